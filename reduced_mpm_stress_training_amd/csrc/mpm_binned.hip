// mpm_binned.hip -- the MI355X substep on particles binned by grid block.
//
// Particles live in bins of 8^3 cells (keyed by the base cell of their
// stencil, mpm_utils.py:197-199), each bin a fixed run of `bin_cap` slots in
// a set of SoA planes, plus an overflow run.  Two plane sets alternate.
//
// One substep (MPM_Simulator_WARP.p2g2p, mpm_solver_warp.py:156-203) is two
// launches:
//   k_grid_bins   per 8^3 grid block: sums the P2G tiles of the (up to 8)
//                 overlapping bins and the spill grid, then
//                 grid_normalization_and_gravity (mpm_utils.py:229-237) and
//                 the colliders (mpm_solver_warp.py:429-449, 505-523).  Its
//                 first workgroup schedules k_g2p2g's work items.
//   k_g2p2g       per work item (a bin, or a run of a large bin's slots):
//                 G2P (mpm_utils.py:241-306) from an LDS tile of grid
//                 velocities, the sand return map and stress in registers,
//                 update_u_from_x, and the P2G of the next substep
//                 (mpm_utils.py:190-227) into an FP64 LDS tile (ds_add_f64;
//                 FP32 LDS atomics run ~8x slower on gfx950), written out as
//                 one plain, coalesced tile per item; every particle is
//                 written to the slot of its new bin in the other plane set.
//                 Workgroup w takes item w of a list sorted largest first.
// Bins' tiles carry an epoch stamp; the grid update uses the tiles of the
// current epoch, so nothing needs clearing between substeps (zero_grid,
// mpm_utils.py:115-123, becomes the epoch bump).  A bin's tiles hold nodes
// b0 .. b0+9 along each axis (b0 its first cell): the stencils of the bases
// inside the bin.  Particles whose next base has left their bin (about one in
// a hundred per substep), and those in the overflow run, scatter with FP64
// atomics into a spill grid that the grid update adds and clears.
// The P2G of the first substep (or after the state is replaced) is
// k_p2g_bins.  So between substeps the tiles hold the next substep's P2G,
// computed with that substep's dt.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_common.h"
#include "mpm_device.h"
#include "mpm_hip.h"
#include "mpm_stencil.h"

using namespace mpm;

namespace {

constexpr int kWG = 512;          // threads per bin workgroup
constexpr int kBinEdge = 8;           // cells per bin edge
constexpr int kTE = kBinEdge + 2;     // tile edge: nodes b0 .. b0+9, the stencils of the bin's bases
constexpr int kTE3 = kTE * kTE * kTE;
constexpr int kSortMax = 6144;        // bins up to this size are processed in bank order
constexpr int kSortPeriod = 16;       // substeps between bank-order sorts of the bins
constexpr int kStageFloats = 12288;   // LDS staging of the sort's copy (48 KB: two bins per CU)
static_assert(kStageFloats >= kSortMax, "a bin's plane must fit the staging buffer");
constexpr int kParts = 4;             // work items (each with its own P2G tile) a bin may be split into

// Plane sets are bin-major: bin b of a scene with `cap` slots per bin is a
// block of MPM_NPLANES planes of cap floats starting at float offset
// bin_off(scene, b), so slot i, plane p is at bin_off + p * cap + i.  After
// all bins comes the overflow run, in blocks of ovf_blk slots laid out the
// same way.  (One scene: bin b at b * MPM_NPLANES * cap, ovf_blk = cap.)
struct BinGeom {
    int n_bins;        // bins of all scenes
    int ovf_cap;       // overflow slots
    int ovf_blk;       // slots per overflow block
    int max_cap;       // largest slots-per-bin of any scene
    int64_t ovf_off;   // float offset of the overflow run
    int64_t stride;
    int n_cu;          // compute units of the device (sizes the fused kernel's work items)
    int pad;
};

// One scene: the mpm_bins_* entry points run one, the mpm_batch_* entry
// points many, through the same kernels.  A batch concatenates the scenes'
// bins (global bin = bin_base + local bin), their grids (node_base) and their
// particles (pid_base); every scene has its own parameters and colliders.
struct SceneDesc {
    MpmParams prm;
    int nb;          // bins per axis = ceil(n_grid / 8)
    int bin_base;    // first global bin (grid blocks are numbered like bins)
    int col_base;    // first collider in the batch's collider array
    int n_cols;
    int cap;         // slots per bin
    int pad;
    int64_t node_base;  // first node of the scene's grids
    int64_t pid_base;   // first particle (flat index) of the scene
    int64_t slot_base;  // float offset of the scene's first bin
};

struct SceneTab {
    SceneDesc one;              // the scene when dev == nullptr
    const SceneDesc* dev;       // device table of n scenes (batch)
    int n;
    const MpmCollider* cols;    // device colliders (batch)
};

__device__ __forceinline__ SceneDesc scene_of_bin(const SceneTab& t, int bin) {
    if (t.dev == nullptr) return t.one;
    int lo = 0, hi = t.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.dev[mid].bin_base <= bin) lo = mid; else hi = mid - 1;
    }
    return t.dev[lo];
}

__device__ __forceinline__ SceneDesc scene_of_pid(const SceneTab& t, int64_t pid) {
    if (t.dev == nullptr) return t.one;
    int lo = 0, hi = t.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.dev[mid].pid_base <= pid) lo = mid; else hi = mid - 1;
    }
    return t.dev[lo];
}

// origin (cells) of local bin lb of a scene with nb bins per axis
__device__ __forceinline__ void bin_origin(int lb, int nb, int b0[3]) {
    b0[0] = (lb / (nb * nb)) * kBinEdge;
    b0[1] = ((lb / nb) % nb) * kBinEdge;
    b0[2] = (lb % nb) * kBinEdge;
}

// float offset of slot 0, plane 0 of global bin b (plane stride sc.cap)
__device__ __forceinline__ int64_t bin_off(const SceneDesc& sc, int b) {
    return sc.slot_base + (int64_t)(b - sc.bin_base) * MPM_NPLANES * sc.cap;
}

// float offset of overflow slot r, plane 0 (plane stride g.ovf_blk)
__device__ __forceinline__ int64_t ovf_off(const BinGeom& g, int r) {
    const int k = r / g.ovf_blk;
    return g.ovf_off + (int64_t)k * MPM_NPLANES * g.ovf_blk + (r - k * g.ovf_blk);
}

__device__ __forceinline__ int bin_of_base(const int b[3], int nb) {
    return ((b[0] >> 3) * nb + (b[1] >> 3)) * nb + (b[2] >> 3);
}

__device__ __forceinline__ bool stencil_ok(const int b[3], int ng) {
    return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] <= ng - 3 && b[1] <= ng - 3 && b[2] <= ng - 3;
}

// workgroup -> bin.  Workgroups are dealt to the 8 XCDs round-robin, so the
// identity mapping (used by k_p2g_bins and when no schedule is current).
__device__ __forceinline__ int xcd_bin(int wg, int n_bins) {
    (void)n_bins;
    return wg;
}

// Schedules of k_g2p2g and of the bank-order sort, built from the bin
// counts by the grid update.  Work is counted in passes: a bin of n
// particles takes ceil(n / 512) passes of a 512-thread workgroup.
//
// Both kernels take one workgroup per list entry, and the command processor
// hands workgroup w to XCD w mod 8, each XCD starting its workgroups in
// order on whichever CU frees up first.  Listing the work by decreasing
// size (longest processing time first) gives every XCD every 8th entry of a
// sorted list -- an even share -- and fills each XCD's CUs greedily from the
// largest down.
//
// For k_g2p2g a whole bin can be several times a CU's share of the work
// (250k particles on a 64^3 grid: 80 bins of up to 6 passes for 256 CUs),
// so bins above q passes, q about half a CU's share of all passes, are split
// into up to kParts work items of consecutive slot ranges; each item
// scatters into its own tile (part p of bin b: tile p * n_bins + b) and the
// grid update sums the parts.
//
// order[0 .. n_bins): bins for the sort, -1 past the last non-empty one;
// order[n_bins + kItemsEpoch]: the epoch of the bin counts it was made from;
// order[n_bins + kItemsFirst + w], w < n_bins: work item w of k_g2p2g,
// (bin << 4) | (part << 2) | (parts - 1), -1 past the last.  Bins are split
// only while the items fit in n_bins slots, so the fused kernel's grid stays
// one workgroup per bin (plus the overflow run's).
constexpr int kSchedKeys = 64;
constexpr int kItemsEpoch = 0, kItemsFirst = 1;

__device__ __forceinline__ int bin_passes(int n) { return (n + kWG - 1) / kWG; }

static_assert(kParts == 4, "split_bin and ceil_div_parts count up to four parts");

// ceil(x / k) for k = 1 .. 4 and 0 <= x < 32766, without a division
__device__ __forceinline__ int ceil_div_parts(int x, int k) {
    return k == 1 ? x : k == 2 ? (x + 1) >> 1 : k == 3 ? ((x + 2) * 21846) >> 16 : (x + 3) >> 2;
}

// parts of a bin of p passes when items should not exceed q passes (split_bin)
__device__ __forceinline__ int split_parts(int p, int q) {
    const int k = 1 + (p > q) + (p > 2 * q) + (p > 3 * q);
    const int pp = ceil_div_parts(p, k);
    return 1 + (p > pp) + (p > 2 * pp) + (p > 3 * pp);
}

// parts of a bin of p passes when items should not exceed q passes, and the
// passes of each (the last one may have fewer): k = min(4, ceil(p / q)),
// per = ceil(p / k), parts = ceil(p / per) <= k
__device__ __forceinline__ void split_bin(int p, int q, int* parts, int* per) {
    const int k = 1 + (p > q) + (p > 2 * q) + (p > 3 * q);
    const int pp = ceil_div_parts(p, k);
    *per = pp;
    *parts = 1 + (p > pp) + (p > 2 * pp) + (p > 3 * pp);
}

// The schedule is built by one workgroup and sits on the grid update's
// critical path: each thread reads the counts of its bins (tid + j * kWG)
// kSchedRegs at a time, all loads of a chunk issued together, and keeps
// their passes in registers; grids of up to kWG * kSchedRegs bins (a 128^3
// grid's 4096) read them once for all passes.
constexpr int kSchedRegs = 8;

__device__ __forceinline__ void load_passes(const int* __restrict__ cnt_cur, int n_bins, int cap, int chunk,
                                            int pv[kSchedRegs]) {
    const int b0 = chunk * kWG * kSchedRegs + (int)threadIdx.x;
#pragma unroll
    for (int j = 0; j < kSchedRegs; ++j) {
        const int b = b0 + j * kWG;
        pv[j] = b < n_bins ? bin_passes(min(cnt_cur[b], cap)) : 0;
    }
}

// prefix sums of s_hist in decreasing key order into s_cursor; s_total = the
// count, s_max = the largest key present (0 if none).  Wave 0, lane l holding
// key kSchedKeys - 1 - l: a 64-lane scan instead of a serial walk.
static_assert(kSchedKeys == 64, "hist_cursors gives one key to each lane of a wave");
__device__ __forceinline__ void hist_cursors(const int* s_hist, int* s_cursor, int* s_total, int* s_max) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x, k = kSchedKeys - 1 - lane;
        const int h = k >= 1 ? s_hist[k] : 0;
        int incl = h;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(incl, d);
            if (lane >= d) incl += o;
        }
        if (k >= 1) s_cursor[k] = incl - h;
        const unsigned long long nz = __ballot(h > 0);
        if (lane == 63) *s_total = incl;
        if (lane == 0) *s_max = nz ? kSchedKeys - __ffsll((long long)nz) : 0;
    }
}

// counter[key] += 1 for every active lane, one LDS atomic per distinct key
// in the wave (most bins share a key, and lanes adding to one LDS address
// are serialized); returns what a per-lane atomicAdd would have returned.
// Called by every lane of the wave.
__device__ __forceinline__ int wave_add(int* counter, int key, bool active) {
    const int lane = __lane_id();
    unsigned long long todo = __ballot(active);
    int old = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int k0 = __shfl(key, leader);
        const unsigned long long same = todo & __ballot(key == k0);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counter[k0], (int)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) old = base + (int)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return old;
}

__device__ void schedule_bins(const int* __restrict__ cnt_cur, const BinGeom& g, int epoch,
                              int* __restrict__ order) {
    __shared__ int s_hist[kSchedKeys];
    __shared__ int s_cursor[kSchedKeys];
    __shared__ int s_total, s_max, s_passes, s_items;
    const int tid = threadIdx.x;
    const int n_bins = g.n_bins, cap = g.max_cap;
    const int n_chunks = (n_bins + kWG * kSchedRegs - 1) / (kWG * kSchedRegs);
    const bool reload = n_chunks > 1;
    int* items = order + n_bins;
    int pv[kSchedRegs];
    if (tid < kSchedKeys) s_hist[tid] = 0;
    if (tid == 0) s_passes = s_items = 0;
    __syncthreads();
    // bins keyed by their passes (their cost), capped
    int passes = 0;
    for (int c = 0; c < n_chunks; ++c) {
        load_passes(cnt_cur, n_bins, cap, c, pv);
#pragma unroll
        for (int j = 0; j < kSchedRegs; ++j) {
            wave_add(s_hist, min(pv[j], kSchedKeys - 1), pv[j] > 0);
            passes += pv[j];
        }
    }
    if (passes > 0) atomicAdd(&s_passes, passes);
    __syncthreads();
    hist_cursors(s_hist, s_cursor, &s_total, &s_max);
    __syncthreads();
    // Split where a bin exceeds a CU's share of all passes (whole bins then
    // leave CUs idle), and where more bins exceed half a share than there are
    // CUs: the ones beyond the n_cu largest would run second on a CU after a
    // large bin, so they are split (those the CUs start with stay whole).
    // Items then take at most half a share each, and all items must fit in
    // n_bins slots.
    const int share = s_passes / max(1, g.n_cu);
    const int q = max(1, share / 2);
    const int n_big = q < kSchedKeys ? s_cursor[q] : 0;   // bins whose key exceeds q (before the cursors move)
    const bool split_all = s_max > max(1, share) && s_max < kSchedKeys - 1;
    const bool split_late = !reload && n_big > g.n_cu && s_max < kSchedKeys - 1;
    bool split = split_all || split_late;
    __syncthreads();   // every thread has read n_big before the cursors move
    int rank[kSchedRegs];   // each bin's place in the sort's list (without reload)
    // the sort's list (and k_g2p2g's when no bin is split)
    for (int c = 0; c < n_chunks; ++c) {
        if (reload) load_passes(cnt_cur, n_bins, cap, c, pv);
#pragma unroll
        for (int j = 0; j < kSchedRegs; ++j) {
            const int pos = wave_add(s_cursor, min(pv[j], kSchedKeys - 1), pv[j] > 0);
            const int b = c * kWG * kSchedRegs + j * kWG + tid;
            rank[j] = pos;
            if (pv[j] > 0) {
                order[pos] = b;
                if (!split) items[kItemsFirst + pos] = b << 4;
            }
        }
    }
    for (int p = s_total + tid; p < n_bins; p += kWG) {
        order[p] = -1;
        if (!split) items[kItemsFirst + p] = -1;
    }
    if (split && !split_all) {
        // Only the bins ranked [n_cu, n_big) split (all of them above q): the
        // items are the n_cu largest bins, then those bins' parts, then the
        // remaining bins in the list's order -- descending but for the parts'
        // place among bins of q passes, without a second sort.
        int my_parts = 0;
#pragma unroll
        for (int j = 0; j < kSchedRegs; ++j)
            if (pv[j] > 0 && rank[j] >= g.n_cu && rank[j] < n_big) my_parts += split_parts(pv[j], q);
        if (my_parts > 0) atomicAdd(&s_items, my_parts);
        __syncthreads();
        const int n_parts = s_items, n_split = n_big - g.n_cu, n_items = s_total + n_parts - n_split;
        const bool fits = n_items <= n_bins;
        if (tid == 0) s_passes = 0;   // from here the next free part slot
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSchedRegs; ++j) {
            const int b = j * kWG + tid, r = rank[j];
            if (pv[j] <= 0) {
            } else if (!fits || r < g.n_cu) {
                items[kItemsFirst + r] = b << 4;
            } else if (r < n_big) {
                const int k = split_parts(pv[j], q);
                const int base = g.n_cu + atomicAdd(&s_passes, k);
#pragma unroll
                for (int i = 0; i < kParts; ++i)
                    if (i < k) items[kItemsFirst + base + i] = (b << 4) | (i << 2) | (k - 1);
            } else {
                items[kItemsFirst + r - n_big + g.n_cu + n_parts] = b << 4;
            }
        }
        for (int p = (fits ? n_items : s_total) + tid; p < n_bins; p += kWG) items[kItemsFirst + p] = -1;
    } else if (split) {
        int my_items = 0;
        for (int c = 0; c < n_chunks; ++c) {
            if (reload) load_passes(cnt_cur, n_bins, cap, c, pv);
#pragma unroll
            for (int j = 0; j < kSchedRegs; ++j)
                if (pv[j] > 0) {
                    int parts = 1, per = pv[j];
                    if (split_all || rank[j] >= g.n_cu) split_bin(pv[j], q, &parts, &per);
                    my_items += parts;
                }
        }
        if (my_items > 0) atomicAdd(&s_items, my_items);
        __syncthreads();   // s_total of the sort's list has been read
        split = s_items <= n_bins;
        if (tid < kSchedKeys) s_hist[tid] = 0;
        __syncthreads();
        for (int c = 0; c < n_chunks; ++c) {
            if (reload) load_passes(cnt_cur, n_bins, cap, c, pv);
#pragma unroll
            for (int j = 0; j < kSchedRegs; ++j) {
                const int p = pv[j];
                int parts = 0, per = 1;
                if (p > 0) {
                    if (split && (split_all || rank[j] >= g.n_cu)) split_bin(p, q, &parts, &per);
                    else { parts = 1; per = p; }
                }
#pragma unroll
                for (int i = 0; i < kParts; ++i)
                    wave_add(s_hist, min(max(min(per, p - i * per), 0), kSchedKeys - 1), i < parts);
            }
        }
        __syncthreads();
        hist_cursors(s_hist, s_cursor, &s_total, &s_max);
        __syncthreads();
        for (int c = 0; c < n_chunks; ++c) {
            if (reload) load_passes(cnt_cur, n_bins, cap, c, pv);
#pragma unroll
            for (int j = 0; j < kSchedRegs; ++j) {
                const int p = pv[j];
                const int b = c * kWG * kSchedRegs + j * kWG + tid;
                int parts = 0, per = 1;
                if (p > 0) {
                    if (split && (split_all || rank[j] >= g.n_cu)) split_bin(p, q, &parts, &per);
                    else { parts = 1; per = p; }
                }
#pragma unroll
                for (int i = 0; i < kParts; ++i) {
                    const int pos =
                        wave_add(s_cursor, min(max(min(per, p - i * per), 0), kSchedKeys - 1), i < parts);
                    if (i < parts) items[kItemsFirst + pos] = (b << 4) | (i << 2) | (parts - 1);
                }
            }
        }
        for (int p = s_total + tid; p < n_bins; p += kWG) items[kItemsFirst + p] = -1;
    }
    if (tid == 0) items[kItemsEpoch] = epoch;
}

// The grid blocks that hold P2G sums of an epoch, listed once each by the
// writer that marks them first (a tile flush or the spill scatter): the grid
// update visits only these.  Two lists, by the epoch's parity: the grid
// update of epoch e reads list e & 1 and empties list (e + 1) & 1 for the
// next epoch's writers.
struct BlockList {
    int* __restrict__ mark;   // per grid block (global index): the newest epoch it was listed for
    int* __restrict__ list;   // [2][n] grid blocks
    int* __restrict__ head;   // [0], [1]: list lengths by epoch parity; [2]: epoch of the last grid update (0: none)
    int n;                    // grid blocks of all scenes (= bins)
};

// The list append of grid block blk once its mark's previous value `old`
// (atomicExch(&bl.mark[blk], epoch)) is known: the first writer of the epoch
// appends.
__device__ __forceinline__ void list_append(const BlockList& bl, int blk, int epoch, int old) {
    if (old != epoch) {
        const int i = atomicAdd(&bl.head[epoch & 1], 1);
        if (i < bl.n) bl.list[(epoch & 1) * bl.n + i] = blk;
    }
}

// Scatter of one particle straight into the spill grid with FP64 atomics
// (overflow particles and particles that left their bin's scatter tile): the
// FP32 contributions sum in FP64 as in the LDS tiles, so the node's value
// does not depend on the order the atomics land in unless the sum needs more
// than 53 bits.
__device__ void scatter_spill(double* __restrict__ acc, int* __restrict__ gflag, int epoch, const SceneDesc& sc,
                              const AxisWeights& A, float m, float vol, const float v[3], const float C[9],
                              const float tau[9], float dt) {
    const int ng = sc.prm.n_grid;
    double* sacc = acc + 4 * sc.node_base;
    p2g_contrib_rolled(A, m, vol, v, C, tau, sc.prm.dx, sc.prm.inv_dx, dt, [&](int c, int i, int j, int k, float val) {
        const int x = A.base[0] + i, y = A.base[1] + j, z = A.base[2] + k;
        unsafeAtomicAdd(sacc + 4 * (((int64_t)x * ng + y) * ng + z) + c, (double)val);
    });
    // the grid blocks of the stencil's nodes (base + 0, 2 per axis) hold
    // spill contributions of this epoch: plain stores, nothing waits on them
    // (the grid update finds blocks no tile listed by scanning these flags)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k)
                gflag[sc.bin_base + (((A.base[0] + 2 * i) >> 3) * sc.nb + ((A.base[1] + 2 * j) >> 3)) * sc.nb +
                      ((A.base[2] + 2 * k) >> 3)] = epoch;
}

// Scatter of one particle into a bin's FP64 LDS tile (SoA: mx, my, mz, m).
__device__ __forceinline__ void scatter_tile(double* tile, int i0, const AxisWeights& A, float m, float vol,
                                             const float v[3], const float C[9], const float tau[9], float dx,
                                             float inv_dx, float dt) {
    p2g_contrib(A, m, vol, v, C, tau, dx, inv_dx, dt, [&](int c, int i, int j, int k, float val) {
        atomicAdd(&tile[c * kTE3 + i0 + (i * kTE + j) * kTE + k], (double)val);
    });
}

// Lists the grid blocks that bin `bin`'s tile overlaps (nodes 8B .. 8B+9
// along each axis: blocks B and B+1, up to 8) for the tile's epoch.  tid:
// the thread's index in the workgroup.
__device__ __forceinline__ void mark_tile_blocks(const BlockList& bl, const SceneDesc& sc, int bin, int epoch,
                                                 int tid) {
    if (tid < 8) {
        const int nb = sc.nb, lb = bin - sc.bin_base;
        const int x = lb / (nb * nb) + (tid >> 2), y = (lb / nb) % nb + ((tid >> 1) & 1), z = lb % nb + (tid & 1);
        if (x < nb && y < nb && z < nb) {
            const int blk = sc.bin_base + (x * nb + y) * nb + z;
            list_append(bl, blk, epoch, atomicExch(&bl.mark[blk], epoch));
        }
    }
}

// Writes a bin's LDS tile (nodes b0 .. b0+9) to its slot of the global
// tile buffer as float4 (mv, m) and stamps it with the epoch (the writer
// lists the grid blocks it overlaps: mark_tile_blocks).  tid: the thread's
// index in the workgroup.
__device__ __forceinline__ void write_tile(const double* s_acc, float4* __restrict__ tiles,
                                           int* __restrict__ stamps, int bin, int epoch, int tid) {
    float4* out = tiles + (int64_t)bin * kTE3;
    for (int t = tid; t < kTE3; t += blockDim.x)
        out[t] = make_float4((float)s_acc[t], (float)s_acc[kTE3 + t], (float)s_acc[2 * kTE3 + t],
                             (float)s_acc[3 * kTE3 + t]);
    if (tid == 0) stamps[bin] = epoch;
}


struct SubstepArgs {
    SceneTab tab;
    BinGeom g;
    float dt;
    int epoch_out;
    const float* __restrict__ src;
    float* __restrict__ dst;
    const int* __restrict__ cnt_cur;
    int* __restrict__ cnt_nxt;
    const float4* __restrict__ vel;
    double* __restrict__ acc;
    int* __restrict__ gflag;
    int* __restrict__ status;
    float4* __restrict__ tiles;
    int* __restrict__ stamps;       // epoch of each bin's tile; -epoch: split, its parts in the spill grid
    BlockList bl;                   // grid blocks written this epoch (tiles, spill)
    int* __restrict__ items;        // work items (schedule_bins): order + n_bins
    int sched_epoch;                // epoch the schedule must carry to be used
    int store_all;                  // 0: write only the planes the next substep reads (see g2p2g_particle)
};

// Destination slot of every active lane.  A bin's workgroup reserves
// n_live slots of its own bin in the next plane set with one atomic before its
// loop (`stay_base`, -1 on the overflow path); lanes that stay in the bin
// take consecutive slots of that run through an LDS counter, so stayers keep
// their order.  Lanes that move take one atomic each on their new bin.
// Returns the float offset of the slot (its plane stride in *stride), or
// -1 when both the bin and the overflow run are full.
__device__ __forceinline__ int64_t reserve_slot(bool act, int new_bin, int own_bin, int stay_base,
                                                int* s_stay, const SubstepArgs& a, const SceneDesc& sc,
                                                int64_t* stride) {
    const bool stay = act && new_bin == own_bin && stay_base >= 0;
    const unsigned long long mask = __ballot(stay);
    const int lane = __lane_id();
    int pos = 0;
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(s_stay, __popcll(mask));
        base = __shfl(base, leader);
        pos = stay_base + base + __popcll(mask & ((1ull << lane) - 1ull));
    }
    if (act && !stay) {
        pos = atomicAdd(&a.cnt_nxt[new_bin], 1);
        if (pos < sc.cap) atomicAdd(&a.cnt_nxt[a.g.n_bins + 1 + new_bin], 1);  // live count
    }
    if (!act) return -1;
    if (pos < sc.cap) {   // new_bin is a bin of the particle's own scene
        *stride = sc.cap;
        return bin_off(sc, new_bin) + pos;
    }
    const int r = atomicAdd(&a.cnt_nxt[a.g.n_bins], 1);
    *stride = a.g.ovf_blk;
    if (r < a.g.ovf_cap) return ovf_off(a.g, r);
    atomicAdd(&a.status[0], 1);
    return -1;
}

// all = false: only the planes the next substep's G2P reads (x0, u, F, m,
// vol, pid, and yield_stress for metal); v, C and tau are then stale in this
// plane set.  Other materials do not carry yield_stress (k_bins_extract then
// leaves the flat plane as it was).
__device__ __forceinline__ void store_particle(float* __restrict__ dst, int64_t off, int64_t stride,
                                               const float x0[3], const float u[3], const float v[3],
                                               const float C[9], const float F[9], const float tau[9],
                                               float m, float vol, float pid, float ys, bool metal, bool all) {
    float* p = dst + off;
    if (metal) p[MPM_P_YIELD * stride] = ys;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        p[(MPM_P_X0 + d) * stride] = x0[d];
        p[(MPM_P_U + d) * stride] = u[d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) p[(MPM_P_F + k) * stride] = F[k];
    p[MPM_P_MASS * stride] = m;
    p[MPM_P_VOL * stride] = vol;
    p[MPM_P_PID * stride] = pid;
    if (all) {
#pragma unroll
        for (int d = 0; d < 3; ++d) p[(MPM_P_V + d) * stride] = v[d];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            p[(MPM_P_C + k) * stride] = C[k];
            p[(MPM_P_TAU + k) * stride] = tau[k];
        }
    }
}

// The per-particle body of k_g2p2g.  `tile_vel`/`tile_acc` are the bin's LDS
// tiles (nullptr on the overflow path); b0 the bin origin in cells.
// The G2P inputs of one particle (x0, u, F, m, vol, pid), loaded ahead of use.
struct ParticleIn {
    float x0[3], u[3], F[9], m, vol, pid, ys;
};

// metal: also the yield_stress plane
__device__ __forceinline__ void load_particle(const float* __restrict__ src, int64_t off, int64_t S, bool act,
                                              bool metal, ParticleIn& q) {
    q.ys = 0.f;
    if (act) {
        if (metal) q.ys = src[off + MPM_P_YIELD * S];
        const float* sp = src + off;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            q.x0[d] = sp[(MPM_P_X0 + d) * S];
            q.u[d] = sp[(MPM_P_U + d) * S];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) q.F[k] = sp[(MPM_P_F + k) * S];
        q.m = sp[MPM_P_MASS * S];
        q.vol = sp[MPM_P_VOL * S];
        q.pid = sp[MPM_P_PID * S];
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) { q.x0[d] = 0.5f; q.u[d] = 0.f; }
#pragma unroll
        for (int k = 0; k < 9; ++k) q.F[k] = (k % 4 == 0) ? 1.f : 0.f;
        q.m = q.vol = 0.f;
        q.pid = __int_as_float(-1);
    }
}

__device__ __forceinline__ void g2p2g_particle(const SubstepArgs& a, const SceneDesc& sc, bool act_slot, int64_t off,
                                               int64_t S, int own_bin, int stay_base, int* s_stay, const int b0[3],
                                               const float* tile_vel, double* tile_acc, const ParticleIn& q) {
    const float* sp = a.src + off;
    const MpmParams& prm = sc.prm;
    const int ng = prm.n_grid;
    const float dt = a.dt;
    // slots left empty by particles that moved to another bin carry pid -1
    const bool act = act_slot && __float_as_int(q.pid) >= 0;
    float x0[3], u[3], F[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) { x0[d] = q.x0[d]; u[d] = q.u[d]; }
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = q.F[k];
    const float m = q.m, vol = q.vol, pid = q.pid;
    const bool metal = prm.material == 1;
    float ys = q.ys;
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = x0[d] + u[d];
    AxisWeights A;
    axis_weights(x, prm.inv_dx, A);
    const bool live = act && stencil_ok(A.base, ng);
    float v[3], Cm[9], G[9], tau[9];
    int new_bin = own_bin;
    int mode = 0;  // next-substep scatter: 0 none, 1 LDS tile, 2 spill grid
    int i0 = 0;
    if (live) {
        const int lx = A.base[0] - b0[0], ly = A.base[1] - b0[1], lz = A.base[2] - b0[2];
        if (tile_vel != nullptr && lx >= 0 && ly >= 0 && lz >= 0 && lx < kBinEdge && ly < kBinEdge &&
            lz < kBinEdge) {
            // the velocity tile holds nodes b0 .. b0+9
            gather_tile<kTE * kTE, kTE>(SoANodes{tile_vel, kTE3}, (lx * kTE + ly) * kTE + lz, A, prm.inv_dx, v,
                                        Cm, G);
        } else {
            // global gather: same contraction, nodes read from the grid
            const float4* vb = a.vel + sc.node_base + (((int64_t)A.base[0] * ng + A.base[1]) * ng + A.base[2]);
            gather_global_rolled(vb, ng, A, prm.inv_dx, v, Cm, G);
        }
        constitutive_update(G, dt, F, prm.material, prm.mu, prm.lam, prm.alpha, prm.hardening, prm.xi, ys, tau);
        // x += dt v ; u = x - x0 ; the next substep recomputes x = x0 + u
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float xn = x[d] + dt * v[d];
            u[d] = xn - x0[d];
            x[d] = x0[d] + u[d];
        }
        axis_weights(x, prm.inv_dx, A);     // weights of the next substep's P2G
        if (stencil_ok(A.base, ng)) {
            new_bin = sc.bin_base + bin_of_base(A.base, sc.nb);
            // the tile takes the stencils of bases inside the bin; a particle
            // that has moved into another bin scatters into the spill grid
            const int sx = A.base[0] - b0[0], sy = A.base[1] - b0[1], sz = A.base[2] - b0[2];
            const bool in_tile = tile_acc != nullptr && sx >= 0 && sy >= 0 && sz >= 0 && sx < kBinEdge &&
                                 sy < kBinEdge && sz < kBinEdge;
            if (in_tile) {
                mode = 1;
                i0 = (sx * kTE + sy) * kTE + sz;
            } else {
                mode = 2;
            }
            // count the bin's particles that left their tile, one atomic per wave
            const unsigned long long out = __ballot(tile_acc != nullptr && !in_tile);
            if (out && __lane_id() == __ffsll((long long)out) - 1) atomicAdd(&a.status[1], __popcll(out));
        } else {
            atomicAdd(&a.status[2], 1);  // left the grid: frozen from here on
            int bc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) bc[d] = min(max(A.base[d], 0), ng - 3);
            new_bin = sc.bin_base + bin_of_base(bc, sc.nb);
        }
    } else if (act) {
        // frozen particle (stencil outside the grid): copied unchanged
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = sp[(MPM_P_V + d) * S];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Cm[k] = sp[(MPM_P_C + k) * S];
            tau[k] = sp[(MPM_P_TAU + k) * S];
        }
        int bc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) bc[d] = min(max(A.base[d], 0), ng - 3);
        new_bin = sc.bin_base + bin_of_base(bc, sc.nb);
    }
    int64_t dstride = 0;
    const int64_t dslot = reserve_slot(act, new_bin, own_bin, stay_base, s_stay, a, sc, &dstride);
    // v, C and tau are no input of the next substep's G2P, and this kernel
    // scatters the next P2G from registers: between the substeps of one run
    // only x0, u, F, m, vol, pid are written.  The last substep of a run
    // (store_all) writes every plane, and so do frozen particles and particles
    // that left the grid (mode 0: they are copied from the source planes in
    // the substeps after, so those planes must hold their values).
    if (dslot >= 0)
        store_particle(a.dst, dslot, dstride, x0, u, v, Cm, F, tau, m, vol, pid, ys, metal, a.store_all || mode == 0);
    // P2G of the next substep, after the store so that x0, u, F are dead
    if (mode == 1) {
        scatter_tile(tile_acc, i0, A, m, vol, v, Cm, tau, prm.dx, prm.inv_dx, dt);
    } else if (mode == 2) {
        scatter_spill(a.acc, a.gflag, a.epoch_out, sc, A, m, vol, v, Cm, tau, dt);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// The tile entries a grid block sums, by source c = (ix, iy, iz) = the tile
// of bin (bx - ix, by - iy, bz - iz): the tile of bin B spans nodes 8B ..
// 8B+9, so source c covers the block's nodes whose local coordinate is < 2
// along each axis it steps down -- 8 or 2 nodes per axis, 1000 entries in
// all.  Source 0 (the block's own bin: 512 entries) is read straight into
// registers; the halo sources 1..7 are staged in LDS from halo_off(c) on.
constexpr int halo_size(int c) { return ((c & 4) ? 2 : 8) * ((c & 2) ? 2 : 8) * ((c & 1) ? 2 : 8); }
constexpr int halo_off(int c) { return c == 0 ? 0 : halo_off(c - 1) + halo_size(c - 1); }
constexpr int kHaloEnd = halo_off(8);
constexpr int kHalo = kHaloEnd - 512;   // halo entries of a block
static_assert(kHaloEnd == 1000, "the block's tile entries: 10^3");
// Each half of a grid-update workgroup (kBlockThreads threads) takes one
// block, two nodes per thread: four waves per block leave each thread the
// registers of two nodes' loads in flight at four waves per SIMD, so that the
// blocks of a 1M-particle column are all resident at once.
constexpr int kBlockThreads = kWG / 2;
static_assert(kWG == 512 && kBlockThreads == 256, "two 8^3 blocks per workgroup, two nodes per thread");

// Stage halo entry e (kHaloEnd > e >= 512) of block (bx, by, bz) of a scene
// with nb^3 bins whose tiles start at btiles.  Entries of sources outside the
// grid read as zero; those of sources without a current tile are loaded all
// the same (in bounds, dropped at the sum).
__device__ __forceinline__ float4 halo_entry(const float4* __restrict__ btiles, int nb, int bx, int by, int bz,
                                             int e) {
    int c = 1, off = halo_off(1);
#pragma unroll
    for (int k = 2; k < 8; ++k)
        if (e >= halo_off(k)) {
            c = k;
            off = halo_off(k);
        }
    const int ix = c >> 2, iy = (c >> 1) & 1, iz = c & 1;
    const int ley = iy ? 1 : 3, lez = iz ? 1 : 3;
    const int r = e - off;
    const int u = r >> (ley + lez), v = (r >> lez) & ((1 << ley) - 1), q = r & ((1 << lez) - 1);
    const int sbx = bx - ix, sby = by - iy, sbz = bz - iz;
    if (sbx < 0 || sby < 0 || sbz < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    return btiles[(int64_t)((sbx * nb + sby) * nb + sbz) * kTE3 + ((u + 8 * ix) * kTE + (v + 8 * iy)) * kTE +
                  (q + 8 * iz)];
}

// Grid update over the grid blocks that hold P2G sums of this epoch: those
// listed by the tile flushes (BlockList, this epoch's parity) and those only
// the spill scatter wrote (spill flag gflag == epoch and not listed; each
// workgroup scans a slice of the flags for them).  For each block: sums the
// P2G tiles of the current epoch from the bins overlapping it and the spill
// grid in FP64, then normalization, gravity and colliders.  Blocks holding
// no P2G sum are not visited: G2P reads none of their nodes, and their
// velocities are left as they are (k_settle_vel zeroes them for readers of
// the whole grid).
// Also resets the next plane set's bin counters and empties the next
// epoch's list.  clear_only: only drains the spill grid.  Workgroup 0
// builds the schedule of k_g2p2g from the current bin counts (when `order`
// is given); workgroups 1 .. gridDim.x - 1 take the blocks two at a time.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWG)
    k_grid_bins(double4* __restrict__ acc, float4* __restrict__ vel, int* __restrict__ gflag,
                const int* __restrict__ stamps, const float4* __restrict__ tiles, int* __restrict__ cnt_nxt,
                const int* __restrict__ cnt_cur, int* __restrict__ order, BinGeom g, SceneTab tab,
                ColliderSet cs, float time, float dt, int epoch, int clear_only, const double* __restrict__ t_base,
                float4* __restrict__ rec, int* __restrict__ status, BlockList bl, int* __restrict__ bvel) {
    // per half: neighbour bin c has a whole-bin tile of this epoch; flags;
    // halo entries (halo_off).  s_extra: the spill-only blocks of a slice.
    __shared__ int s_live[2][27];
    __shared__ int s_any[2], s_bad[2];
    __shared__ float4 s_halo[2][kHalo];
    __shared__ int s_extra[kWG], s_nextra;
    const int tid = threadIdx.x;
    if (blockIdx.x == 0) {
        if (tid == 0) {
            bl.head[(epoch + 1) & 1] = 0;        // the next epoch's writers list from scratch
            bl.head[2] = clear_only ? 0 : epoch;
        }
        if (order) schedule_bins(cnt_cur, g, epoch, order);
        return;
    }
    const int nw = gridDim.x - 1, w = blockIdx.x - 1;
    // half (whole waves: wave-uniform), thread in it
    const int hf = __builtin_amdgcn_readfirstlane(tid / kBlockThreads), ht = tid % kBlockThreads;
    // a captured run reads its start time from the device: `time` is then the offset
    const float tnow = t_base ? (float)(t_base[0] + (double)time) : time;
    const int* lst = bl.list + (epoch & 1) * bl.n;
    // one round of loads: the list's length, this workgroup's first entry and
    // the first chunk of its slice of spill flags and list marks
    const int n_list = min(bl.head[epoch & 1], bl.n);
    const int first = 2 * w + hf < bl.n ? lst[2 * w + hf] : 0;
    const int slice = (g.n_bins + nw - 1) / nw, s0 = w * slice, s1 = min(g.n_bins, s0 + slice);
    int sflag = 0, smark = 0;
    if (s0 + tid < s1) {
        sflag = gflag[s0 + tid];
        smark = bl.mark[s0 + tid];
    }
    if (ht == 0) s_any[hf] = s_bad[hf] = 0;
    if (tid == 0) s_nextra = 0;
    // thread ht: nodes (lx, ly, lz) and (lx + 4, ly, lz) of its half's block
    const int lx = ht >> 6, ly = (ht >> 3) & 7, lz = ht & 7;

    // Blocks blk (has) of the two halves: called by every thread.
    auto update_pair = [&](bool has, int blk) {
        __syncthreads();   // the previous blocks are done with the shared arrays; the flags are reset
        blk = has ? __builtin_amdgcn_readfirstlane(blk) : 0;   // global block (= global bin index)
        const SceneDesc sc = scene_of_bin(tab, blk);
        const int nb = sc.nb, ng = sc.prm.n_grid;
        const int lblk = blk - sc.bin_base;
        const int bz = lblk % nb, by = (lblk / nb) % nb, bx = lblk / (nb * nb);
        double4* const bacc = acc + sc.node_base;
        const float4* const btiles = tiles + (int64_t)sc.bin_base * kTE3;
        // one round of loads: the neighbours' stamps, the spill flag, the
        // block's tile entries and its spill sums (the parts of split bins
        // are in the spill grid)
        if (has && ht < 27) {
            const int c = ht;
            const int x = bx + c / 9 - 1, y = by + (c / 3) % 3 - 1, z = bz + c % 3 - 1;
            const bool in = !clear_only && x >= 0 && y >= 0 && z >= 0 && x < nb && y < nb && z < nb;
            const int st = in ? stamps[sc.bin_base + (x * nb + y) * nb + z] : 0;
            const int live = in && st == epoch;
            s_live[hf][c] = live;
            if (live) atomicOr(&s_any[hf], 1);
        }
        // the spill grid holds contributions of this epoch in this block
        const bool spill = has && gflag[blk] == epoch;
        float4 own[2] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
        double4 sp[2] = {make_double4(0.0, 0.0, 0.0, 0.0), make_double4(0.0, 0.0, 0.0, 0.0)};
        if (has && !clear_only) {
#pragma unroll
            for (int k = 0; k < 2; ++k)
                own[k] = btiles[(int64_t)lblk * kTE3 + ((lx + 4 * k) * kTE + ly) * kTE + lz];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int e = 512 + k * kBlockThreads + ht;
                if (e < kHaloEnd) s_halo[hf][e - 512] = halo_entry(btiles, nb, bx, by, bz, e);
            }
        }
        if (spill && !clear_only) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int x = bx * 8 + lx + 4 * k, y = by * 8 + ly, z = bz * 8 + lz;
                if (x < ng && y < ng && z < ng) sp[k] = bacc[((int64_t)x * ng + y) * ng + z];
            }
        }
        __syncthreads();   // stamps and halo entries in LDS
        const bool work = has && !clear_only && (spill || s_any[hf]);
        if (work) {
            // the two nodes' sums first, then their collider passes: the
            // colliders' registers are not held beside the sums'
            float4 a4[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int nx = lx + 4 * k;
                double sx = sp[k].x, sy = sp[k].y, sz = sp[k].z, sm = sp[k].w;
                // whole-bin tiles in (ix, iy, iz) order
                if (s_any[hf]) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int ix = c >> 2, iy = (c >> 1) & 1, iz = c & 1;
                        const bool cov = (!ix || nx < 2) && (!iy || ly < 2) && (!iz || lz < 2);
                        if (!cov || !s_live[hf][13 - 9 * ix - 3 * iy - iz]) continue;
                        const float4 t4 = c == 0 ? own[k]
                                                 : s_halo[hf][halo_off(c) - 512 +
                                                              (nx * (iy ? 2 : 8) + ly) * (iz ? 2 : 8) + lz];
                        sx += t4.x; sy += t4.y; sz += t4.z; sm += t4.w;
                    }
                }
                a4[k] = make_float4((float)sx, (float)sy, (float)sz, (float)sm);
            }
            float4* const bvelg = vel + sc.node_base;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int x = bx * 8 + lx + 4 * k, y = by * 8 + ly, z = bz * 8 + lz;
                if (x >= ng || y >= ng || z >= ng) continue;
                const int64_t node = ((int64_t)x * ng + y) * ng + z;
                if (rec) rec[sc.node_base + node] = a4[k];   // grid_v_in and grid_m of the reference
                // a NaN or Inf in any particle's state reaches its nodes' P2G sums
                if (!isfinite(a4[k].x + a4[k].y + a4[k].z + a4[k].w)) atomicAdd(&s_bad[hf], 1);
                // the colliders are wave-uniform: read through the scalar cache, from
                // the scene table or from the kernel's arguments
                bvelg[node] = tab.dev ? grid_node_update(a4[k], x, y, z, ng, sc.prm.dx, dt, sc.prm.gravity,
                                                         tab.cols + sc.col_base, sc.n_cols, tnow)
                                      : grid_node_update(a4[k], x, y, z, ng, sc.prm.dx, dt, sc.prm.gravity, cs, tnow);
            }
        }
        if (spill) {   // drained (after the loads above, off their path)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int x = bx * 8 + lx + 4 * k, y = by * 8 + ly, z = bz * 8 + lz;
                if (x < ng && y < ng && z < ng) bacc[((int64_t)x * ng + y) * ng + z] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
        }
        __syncthreads();   // every thread has read the flags and the spill flag
        if (ht == 0 && has) {
            if (spill) gflag[blk] = 0;
            if (work) bvel[blk] = epoch;   // written by this epoch's update (k_settle_vel)
            if (s_bad[hf]) atomicAdd(&status[3], s_bad[hf]);
            s_any[hf] = s_bad[hf] = 0;
        }
    };

    // the listed blocks: pairs w, w + nw, ... of the list
    const int n_pairs = (n_list + 1) / 2;
    for (int pi = w; pi < n_pairs; pi += nw) {
        const int li = 2 * pi + hf;
        update_pair(li < n_list, li < n_list ? (pi == w ? first : lst[li]) : 0);
    }
    // the blocks of this workgroup's slice that only the spill scatter wrote
    // (rare: material moving into bins without a tile), a chunk of the slice
    // at a time
    __syncthreads();   // s_nextra is reset
    for (int c0 = s0; c0 < s1; c0 += kWG) {
        if (c0 > s0 && c0 + tid < s1) {
            sflag = gflag[c0 + tid];
            smark = bl.mark[c0 + tid];
        }
        if (c0 + tid < s1 && sflag == epoch && smark != epoch) s_extra[atomicAdd(&s_nextra, 1)] = c0 + tid;
        __syncthreads();
        const int n_extra = s_nextra;
        for (int k = 0; k < n_extra; k += 2) update_pair(k + hf < n_extra, s_extra[min(k + hf, n_extra - 1)]);
        __syncthreads();   // every thread has read s_nextra and s_extra
        if (tid == 0) s_nextra = 0;
    }
    // last, off the blocks' path: slots and live particles of every bin, and
    // the overflow run's count, for the next fused kernel
    if (cnt_nxt)
        for (int i = w * kWG + tid; i < 2 * g.n_bins + 1; i += nw * kWG) cnt_nxt[i] = 0;
}

// ---------------------------------------------------------------------------
// Fused G2P (this substep) + P2G (next substep), one workgroup per bin plus
// workgroups over the overflow run.
// ---------------------------------------------------------------------------
// LDS bank class of a stencil base node in the 10^3 tiles: t mod 32, t the
// tile index of the base (origin b0).  gfx950's LDS serves an 8-byte access
// of 16 consecutive lanes at once from 32 four-byte banks: the scatter's
// ds_add_f64 of such a group runs without conflicts when its 16 nodes differ
// mod 16 (the offset of a stencil node from the base is the same for all
// lanes of one atomic); lanes 16 apart do not conflict.  The gather's 4-byte
// reads are served 32 lanes at a time and want the nodes to differ mod 32.
// The bank-order sort lays bins out in rounds of the 32 classes in order, so
// that aligned groups of 32 slots hold distinct classes and each aligned
// half of them distinct classes mod 16.
__device__ __forceinline__ int bank_class(const int base[3], const int b0[3]) {
    return ((((base[0] - b0[0]) * kTE + (base[1] - b0[1])) * kTE + (base[2] - b0[2])) & 31);
}

// Launch shapes of k_g2p2g held at three waves per SIMD (~165 VGPRs, no
// next-particle prefetch): one 768-thread workgroup per CU, two of 384 or
// three of 256 (3 x 44 KB of LDS tiles fit the CU's 160 KB).  512 threads
// keep two waves per SIMD and load the next particle ahead.
template <int WG>
constexpr bool kThreeWaves = WG == 768 || WG == 384 || WG == 256;

// One work item of k_g2p2g: slots [lo, hi) of `bin`, part `part` of
// `parts` (the bin's passes dealt out in runs of ceil(passes / parts)), with
// its own velocity and P2G tiles.  Called by every thread of the workgroup.
template <int WG>
__device__ __forceinline__ void g2p2g_item(const SubstepArgs& a, int bin, int part, int parts, float* s_vel,
                                           double* s_acc, int* s_stay, int* s_base, int* s_live) {
    const int tid = threadIdx.x;
    const BinGeom& g = a.g;
    const SceneDesc sc = scene_of_bin(a.tab, bin);
    const int cap = sc.cap;
    const int n_all = min(a.cnt_cur[bin], cap);
    const int per = ceil_div_parts(bin_passes(n_all), parts) * kWG;
    const int lo = part * per, hi = min(n_all, lo + per);
    if (lo >= hi) return;
    const int ng = sc.prm.n_grid;
    int b0[3];
    bin_origin(bin - sc.bin_base, sc.nb, b0);
    const float4* svel = a.vel + sc.node_base;
    const int64_t off0 = bin_off(sc, bin);
    // the grid blocks the item's tile (or its bin's parts' tiles) will
    // overlap are listed now: the exchanges are issued before the velocity
    // tile's loads and their results used after them (lanes 0..7 of wave 1)
    const int mt = tid - 64;
    int mark_blk = -1, mark_old = 0;
    if (part == 0 && mt >= 0 && mt < 8) {
        const int nb = sc.nb, lb = bin - sc.bin_base;
        const int x = lb / (nb * nb) + (mt >> 2), y = (lb / nb) % nb + ((mt >> 1) & 1), z = lb % nb + (mt & 1);
        if (x < nb && y < nb && z < nb) {
            mark_blk = sc.bin_base + (x * nb + y) * nb + z;
            mark_old = atomicExch(&a.bl.mark[mark_blk], a.epoch_out);
        }
    }
    // the run stayers will fill: as many slots as the item has live
    // particles (a whole bin: the live count kept with the bin counts; a
    // part: counted here)
    if (tid == 0) {
        *s_stay = 0;
        *s_live = 0;
        if (parts == 1) *s_base = atomicAdd(&a.cnt_nxt[bin], min(a.cnt_cur[g.n_bins + 1 + bin], n_all));
    }
    if (parts > 1) {
        __syncthreads();
        int live = 0;
        for (int i = lo + tid; i < hi; i += WG) live += __float_as_int(a.src[off0 + (int64_t)MPM_P_PID * cap + i]) >= 0;
        if (live) atomicAdd(s_live, live);
    }
    for (int t = tid; t < kTE3; t += WG) {
        const int x = b0[0] + t / (kTE * kTE), y = b0[1] + (t / kTE) % kTE, z = b0[2] + t % kTE;
        float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
        if (x < ng && y < ng && z < ng) val = svel[((int64_t)x * ng + y) * ng + z];
        s_vel[t] = val.x;
        s_vel[kTE3 + t] = val.y;
        s_vel[2 * kTE3 + t] = val.z;
    }
    for (int t = tid; t < 4 * kTE3; t += WG) s_acc[t] = 0.0;
    if (mark_blk >= 0) list_append(a.bl, mark_blk, a.epoch_out, mark_old);
    __syncthreads();
    if (parts > 1) {
        if (tid == 0) *s_base = atomicAdd(&a.cnt_nxt[bin], *s_live);
        __syncthreads();
    }
    const int n_live =
        parts > 1 ? __builtin_amdgcn_readfirstlane(*s_live) : min(a.cnt_cur[g.n_bins + 1 + bin], n_all);
    // Particles are processed in storage order.  Storage is kept in bank order
    // (k_bins_interleave after insertion and every kSortPeriod substeps;
    // stayers are written back in processing order), so each 32-lane half
    // mostly sees distinct bank classes of its base node.
    const int stay_base = __builtin_amdgcn_readfirstlane(*s_base);
    {
        const int64_t offp = off0 + lo;   // the item's first slot
        const int n_in = hi - lo;
        const bool metal = sc.prm.material == 1;
        if constexpr (kThreeWaves<WG>) {
            // three waves per SIMD leave no registers for loading the next
            // particle ahead
            for (int it = 0; it < n_in; it += WG) {
                const int i = it + tid;
                ParticleIn cur;
                load_particle(a.src, offp + i, cap, i < n_in, metal, cur);
                g2p2g_particle(a, sc, i < n_in, offp + i, cap, bin, stay_base, s_stay, b0, s_vel, s_acc, cur);
            }
        } else {
            // two waves per SIMD: the next particle's inputs load during this one
            ParticleIn q;
            load_particle(a.src, offp + tid, cap, tid < n_in, metal, q);
            for (int it = 0; it < n_in; it += WG) {
                const int i = it + tid;
                const ParticleIn cur = q;
                if (it + WG < n_in) load_particle(a.src, offp + i + WG, cap, i + WG < n_in, metal, q);
                g2p2g_particle(a, sc, i < n_in, offp + i, cap, bin, stay_base, s_stay, b0, s_vel, s_acc, cur);
            }
        }
    }
    __syncthreads();
    // reserved slots nobody stayed in become holes (pid -1)
    const int hole_lo = stay_base + *s_stay, hole_hi = min(stay_base + n_live, cap);
    for (int h = hole_lo + tid; h < hole_hi; h += WG)
        a.dst[off0 + (int64_t)MPM_P_PID * cap + h] = __int_as_float(-1);
    if (tid == 0) atomicAdd(&a.cnt_nxt[g.n_bins + 1 + bin], max(0, min(*s_stay, cap - stay_base)));
    if (parts == 1) {
        write_tile(s_acc, a.tiles, a.stamps, bin, a.epoch_out, tid);
    } else {
        // a part adds its FP64 sums into the spill grid with FP64 atomics (the
        // parts of a bin land in any order; FP64 sums of FP32 contributions do
        // not depend on it unless they need more than 53 bits) and flags the
        // tile's blocks; its bin has no whole tile this epoch (stamp -epoch)
        double* sacc = a.acc + 4 * sc.node_base;
        for (int t = tid; t < kTE3; t += WG) {
            const int x = b0[0] + t / (kTE * kTE), y = b0[1] + (t / kTE) % kTE, z = b0[2] + t % kTE;
            if (x >= ng || y >= ng || z >= ng) continue;
            double* nd = sacc + 4 * (((int64_t)x * ng + y) * ng + z);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double v = s_acc[c * kTE3 + t];
                if (v != 0.0) unsafeAtomicAdd(nd + c, v);
            }
        }
        if (tid < 8) {
            const int x = (b0[0] >> 3) + (tid >> 2), y = (b0[1] >> 3) + ((tid >> 1) & 1), z = (b0[2] >> 3) + (tid & 1);
            if (x < sc.nb && y < sc.nb && z < sc.nb) a.gflag[sc.bin_base + (x * sc.nb + y) * sc.nb + z] = a.epoch_out;
        }
        if (tid == 0 && part == 0) a.stamps[bin] = -a.epoch_out;
    }
}

// Block `blk` (WG particles) of the overflow run: global gather and spill
// scatter, no tiles.
template <int WG>
__device__ __forceinline__ void g2p2g_overflow(const SubstepArgs& a, int blk) {
    const BinGeom& g = a.g;
    const int n_ovf = min(a.cnt_cur[g.n_bins], g.ovf_cap);
    const int i = blk * WG + (int)threadIdx.x;
    if (blk * WG >= n_ovf) return;
    const int b0[3] = {0, 0, 0};
    const int64_t off = ovf_off(g, i);
    ParticleIn q;
    load_particle(a.src, off, g.ovf_blk, i < n_ovf, true, q);
    const SceneDesc sc = scene_of_pid(a.tab, __float_as_int(q.pid));   // the particle's own scene
    g2p2g_particle(a, sc, i < n_ovf, off, g.ovf_blk, g.n_bins, -1, nullptr, b0, nullptr, nullptr, q);
}

// WG: threads per workgroup (256, 384, 512 or 768).  The 76 KB of LDS tiles
// allow one 768- or 512-thread workgroup per CU, or two of 256 or 384
// threads.  The kernel needs ~161 VGPRs (its rarely taken paths -- the spill
// scatter and the overflow run's global gather -- are written with small live
// sets so that they do not set the budget), so 768 threads, or two
// workgroups of 384, give three waves per SIMD; 256 and 512 threads two.
// Workgroup w < n_bins runs work item w of
// the schedule of this epoch (schedule_bins: at most n_bins items, largest
// first; -1 past the last) or, without one, bin w whole; the workgroups
// after them run the overflow run in blocks of WG particles.
template <int WG>
__global__ void __launch_bounds__(WG)
    __attribute__((amdgpu_waves_per_eu(kThreeWaves<WG> ? 3 : 2, kThreeWaves<WG> ? 3 : 2))) k_g2p2g(SubstepArgs a) {
    __shared__ float s_vel[3 * kTE3];
    __shared__ double s_acc[4 * kTE3];
    __shared__ int s_stay, s_base, s_live;
    const BinGeom& g = a.g;
    const int wg = blockIdx.x;
    if (wg >= g.n_bins) {
        g2p2g_overflow<WG>(a, wg - g.n_bins);
        return;
    }
    const int it = a.items[kItemsEpoch] == a.sched_epoch ? a.items[kItemsFirst + wg] : (wg << 4);
    if (it < 0) return;
    g2p2g_item<WG>(a, it >> 4, (it >> 2) & 3, (it & 3) + 1, s_vel, s_acc, &s_stay, &s_base, &s_live);
}

// ---------------------------------------------------------------------------
// P2G of the first substep (p2g_apic_stress_given, mpm_utils.py:190-227) from
// the binned state: same tiles as k_g2p2g.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWG)
    k_p2g_bins(SceneTab tab, BinGeom g, float dt, const float* __restrict__ planes, const int* __restrict__ cnt,
               double* __restrict__ acc, int* __restrict__ gflag, float4* __restrict__ tiles,
               int* __restrict__ stamps, BlockList bl, int epoch_out) {
    __shared__ double s_acc[4 * kTE3];
    const int tid = threadIdx.x;
    const int wg = blockIdx.x;
    const bool ovf = wg >= g.n_bins;
    int n_in, bin;
    SceneDesc sc_bin = tab.one;
    if (ovf) {
        n_in = min(cnt[g.n_bins], g.ovf_cap) - (wg - g.n_bins) * kWG;
        n_in = min(max(n_in, 0), kWG);
        bin = g.n_bins;
    } else {
        bin = xcd_bin(wg, g.n_bins);
        sc_bin = scene_of_bin(tab, bin);
        n_in = min(cnt[bin], sc_bin.cap);
    }
    if (n_in == 0) return;
    const int64_t S = ovf ? g.ovf_blk : sc_bin.cap;
    int b0[3] = {0, 0, 0};
    if (!ovf) {
        bin_origin(bin - sc_bin.bin_base, sc_bin.nb, b0);
        for (int t = tid; t < 4 * kTE3; t += kWG) s_acc[t] = 0.0;
        __syncthreads();
    }
    for (int it = 0; it < n_in; it += kWG) {
        const int i = it + tid;
        if (i >= n_in) continue;
        const float* sp = planes + (ovf ? ovf_off(g, (wg - g.n_bins) * kWG + i) : bin_off(sc_bin, bin) + i);
        float x[3], v[3], Cm[9], tau[9];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[d] = sp[(MPM_P_X0 + d) * S] + sp[(MPM_P_U + d) * S];
            v[d] = sp[(MPM_P_V + d) * S];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Cm[k] = sp[(MPM_P_C + k) * S];
            tau[k] = sp[(MPM_P_TAU + k) * S];
        }
        const float m = sp[MPM_P_MASS * S], vol = sp[MPM_P_VOL * S];
        const int pid = __float_as_int(sp[MPM_P_PID * S]);
        if (pid < 0) continue;  // hole
        const SceneDesc sc = ovf ? scene_of_pid(tab, pid) : sc_bin;
        AxisWeights A;
        axis_weights(x, sc.prm.inv_dx, A);
        if (!stencil_ok(A.base, sc.prm.n_grid)) continue;
        const int sx = A.base[0] - b0[0], sy = A.base[1] - b0[1], sz = A.base[2] - b0[2];
        if (!ovf && sx >= 0 && sy >= 0 && sz >= 0 && sx < kBinEdge && sy < kBinEdge && sz < kBinEdge) {
            scatter_tile(s_acc, (sx * kTE + sy) * kTE + sz, A, m, vol, v, Cm, tau, sc.prm.dx, sc.prm.inv_dx, dt);
        } else {
            scatter_spill(acc, gflag, epoch_out, sc, A, m, vol, v, Cm, tau, dt);
        }
    }
    if (!ovf) {
        __syncthreads();
        write_tile(s_acc, tiles, stamps, bin, epoch_out, threadIdx.x);
        mark_tile_blocks(bl, sc_bin, bin, epoch_out, threadIdx.x);
    }
}

// Copies every bin of plane set `src` into plane set `dst` in bank order and
// without holes: the live particles are ranked within their bank_class
// (rank r, class c), and laid out round by round -- round r holds the r-th
// particle of every class that has one, in class order, so the particle goes
// to slot sum_c' min(count_c', r) + #{c' < c : count_c' > r}.  While every
// class still has particles a round is 32 slots, aligned to the 32-lane
// halves of the fused kernel's waves, with 32 distinct classes; once classes
// run out the rounds shorten and a group may repeat classes (few particles in
// a full bin, whose classes hold near-equal counts).  Bins larger than
// kSortMax and the overflow run are copied as they are.
// Plane j of the planes a fused substep reads from a plane set: x0, u, F, m,
// vol, pid, and yield_stress for metal (18 or 19).
__device__ __forceinline__ int read_plane(int j) {
    return j < 6 ? MPM_P_X0 + j
                 : (j < 15 ? MPM_P_F + j - 6
                           : (j == 15 ? MPM_P_MASS : (j == 16 ? MPM_P_VOL : (j == 17 ? MPM_P_PID : MPM_P_YIELD))));
}

// all_planes false (the periodic sort inside a run): only the planes the next
// fused substep reads are copied -- v, C and tau of live particles are
// rewritten by it before anything reads them -- plus v, C and tau of frozen
// particles (stencil outside the grid), which the fused kernel copies forward.
__device__ void interleave_bin(const float* __restrict__ src, const int* __restrict__ cnt_src,
                               float* __restrict__ dst, int* __restrict__ cnt_dst, const BinGeom& g,
                               const SceneTab& tab, int item, bool all_planes) {
    __shared__ float s_stage[kStageFloats];      // planes of the bin being copied
    __shared__ unsigned short s_from[kSortMax];  // source slot of each destination slot
    __shared__ unsigned short s_rank[kSortMax];
    __shared__ unsigned char s_cls[kSortMax];
    __shared__ int s_ccnt[32];
    __shared__ int s_n, s_frozen;
    const int tid = threadIdx.x;
    int64_t off0;
    int n_in;
    if (item >= g.n_bins) {  // overflow blocks are copied as they are
        const int k = item - g.n_bins;
        const int n_ovf = min(cnt_src[g.n_bins], g.ovf_cap);
        const int B = g.ovf_blk;
        off0 = ovf_off(g, k * B);
        n_in = min(max(n_ovf - k * B, 0), B);
        for (int i = tid; i < n_in; i += kWG)
            for (int pl = 0; pl < MPM_NPLANES; ++pl) dst[off0 + (int64_t)pl * B + i] = src[off0 + (int64_t)pl * B + i];
        if (k == 0 && tid == 0) cnt_dst[g.n_bins] = n_ovf;
        return;
    }
    const int bin = item;
    const SceneDesc sc = scene_of_bin(tab, bin);
    const int cap = sc.cap;
    n_in = min(cnt_src[bin], cap);
    off0 = bin_off(sc, bin);
    const float inv_dx = sc.prm.inv_dx;
    int b0[3];
    bin_origin(bin - sc.bin_base, sc.nb, b0);
    if (n_in > kSortMax) {
        // kept in place, holes included (the live count is then an upper bound)
        for (int i = tid; i < n_in; i += kWG)
            for (int pl = 0; pl < MPM_NPLANES; ++pl) dst[off0 + (int64_t)pl * cap + i] = src[off0 + (int64_t)pl * cap + i];
        if (tid == 0) { cnt_dst[bin] = n_in; cnt_dst[g.n_bins + 1 + bin] = n_in; }
        return;
    }
    if (tid < 32) s_ccnt[tid] = 0;
    if (tid == 0) s_frozen = 0;
    __syncthreads();
    for (int i = tid; i < n_in; i += kWG) {
        const float* sp = src + off0 + i;
        if (__float_as_int(sp[(int64_t)MPM_P_PID * cap]) < 0) { s_cls[i] = 255; continue; }  // hole
        int raw[3], base[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float x = sp[(int64_t)(MPM_P_X0 + d) * cap] + sp[(int64_t)(MPM_P_U + d) * cap];
            raw[d] = (int)(x * inv_dx - 0.5f);
            base[d] = min(max(raw[d], b0[d]), b0[d] + kBinEdge - 1);
        }
        const int c = bank_class(base, b0);
        // frozen, or within a cell of it (a margin that keeps this test on the
        // safe side of the fused kernel's own at cell boundaries)
        const int ng = sc.prm.n_grid;
        const bool frozen = !(raw[0] >= 1 && raw[1] >= 1 && raw[2] >= 1 && raw[0] <= ng - 4 && raw[1] <= ng - 4 &&
                              raw[2] <= ng - 4);
        if (frozen) s_frozen = 1;
        s_cls[i] = (unsigned char)(c | (frozen ? 64 : 0));   // class 0..31; bit 6: frozen
        s_rank[i] = (unsigned short)atomicAdd(&s_ccnt[c], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int total = 0;
        for (int c = 0; c < 32; ++c) total += s_ccnt[c];
        s_n = total;
        cnt_dst[bin] = total;
        cnt_dst[g.n_bins + 1 + bin] = total;
    }
    __syncthreads();
    const unsigned n = (unsigned)s_n;
    for (int i = tid; i < n_in; i += kWG) {
        if (s_cls[i] == 255) continue;
        const int c = s_cls[i] & 31, r = s_rank[i];
        int pos = 0;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int ck = s_ccnt[k];
            pos += min(ck, r) + ((k < c && ck > r) ? 1 : 0);
        }
        s_from[pos] = (unsigned short)i;
    }
    // Permuted copy staged through LDS, as many planes at a time as fit:
    // coalesced reads of whole planes (many loads in flight per thread), then
    // coalesced writes gathering from LDS.
    const int n_planes = all_planes ? MPM_NPLANES : (sc.prm.material == 1 ? 19 : 18);
    const int per = max(1, kStageFloats / max(n_in, 1));
    for (int pl0 = 0; pl0 < n_planes; pl0 += per) {
        const int np = min(per, n_planes - pl0);
        __syncthreads();  // s_from ready / the previous stage is written out
        for (int j = 0; j < np; ++j) {
            const int pl = all_planes ? pl0 + j : read_plane(pl0 + j);
            const float* sp = src + off0 + (int64_t)pl * cap;
            for (int i = tid; i < n_in; i += kWG) s_stage[j * n_in + i] = sp[i];
        }
        __syncthreads();
        for (int j = 0; j < np; ++j) {
            const int pl = all_planes ? pl0 + j : read_plane(pl0 + j);
            float* dp = dst + off0 + (int64_t)pl * cap;
            for (int q = tid; q < (int)n; q += kWG) dp[q] = s_stage[j * n_in + s_from[q]];
        }
    }
    if (!all_planes && s_frozen) {
        // v, C and tau of the frozen particles
        for (int q = tid; q < (int)n; q += kWG) {
            const int i = s_from[q];
            if (!(s_cls[i] & 64)) continue;
            for (int pl = MPM_P_V; pl < MPM_P_TAU + 9; ++pl)
                dst[off0 + (int64_t)pl * cap + q] = src[off0 + (int64_t)pl * cap + i];
        }
    }
}

// One workgroup per bin, then one per block of ovf_blk slots of the overflow
// run.  With a current schedule (schedule_bins, order[n_bins] == epoch) the
// bins are taken in its order, which spreads them over the XCDs and skips
// the empty ones (their counts in `cnt_dst` are then already zero);
// otherwise workgroup b takes bin b.
__global__ void __launch_bounds__(kWG)
    k_bins_interleave(const float* __restrict__ src, const int* __restrict__ cnt_src, float* __restrict__ dst,
                      int* __restrict__ cnt_dst, BinGeom g, SceneTab tab, const int* __restrict__ order,
                      int sched_epoch, int all_planes) {
    int item = blockIdx.x;
    if (item < g.n_bins && order != nullptr && order[g.n_bins] == sched_epoch) {
        item = order[item];
        if (item < 0) return;
    }
    interleave_bin(src, cnt_src, dst, cnt_dst, g, tab, item, all_planes != 0);
}

// fill n ints with v (bin counters; a kernel rather than a memset so that the
// sequence can be captured in a graph)
// Tile stamps at the start of a captured run: the pending tiles (stamp
// `from`, or -from for a split bin) become the run's first epoch `to` (-to),
// every other stamp is voided, so that no tile of an earlier replay matches
// an epoch the run reuses.  The block list of `from` becomes the list of
// `to`, and velocity blocks written before are marked stale for
// k_settle_vel (an epoch of an earlier replay could equal a later one's).
__global__ void k_stamp_rebase(int* __restrict__ stamps, int* __restrict__ gflag, BlockList bl,
                               int* __restrict__ bvel, int n, int from, int to) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int st = stamps[i];
        stamps[i] = st == from ? to : st == -from ? -to : (int)0x80000000;
        bl.mark[i] = bl.mark[i] == from ? to : 0;
        gflag[i] = gflag[i] == from ? to : 0;
        if (bvel[i] != 0) bvel[i] = -1;
        if ((from & 1) != (to & 1)) {
            // the two parities' lengths are different words: read one, write the other
            const int nf = min(bl.head[from & 1], n);
            if (i < nf) bl.list[(to & 1) * n + i] = bl.list[(from & 1) * n + i];
            if (i == 0) bl.head[to & 1] = nf;
        }
    }
}

// Zeroes the velocity nodes of every grid block that the last grid update
// did not write (its particles have left it; zero_grid, mpm_utils.py:115-123,
// clears every node each substep).  The substeps never read such nodes, so
// they leave them as they were; this runs before the grid is read from
// outside.  One workgroup per grid block.
__global__ void __launch_bounds__(kWG) k_settle_vel(float4* __restrict__ vel, int* __restrict__ bvel,
                                                    const int* __restrict__ head, BinGeom g, SceneTab tab) {
    const int blk = blockIdx.x;
    if (blk >= g.n_bins) return;
    const int e = bvel[blk];
    if (e == 0 || e == head[2]) return;
    const SceneDesc sc = scene_of_bin(tab, blk);
    const int nb = sc.nb, ng = sc.prm.n_grid;
    const int lblk = blk - sc.bin_base;
    const int bz = lblk % nb, by = (lblk / nb) % nb, bx = lblk / (nb * nb);
    for (int t = threadIdx.x; t < 512; t += kWG) {
        const int x = bx * 8 + (t >> 6), y = by * 8 + ((t >> 3) & 7), z = bz * 8 + (t & 7);
        if (x >= ng || y >= ng || z >= ng) continue;
        vel[sc.node_base + ((int64_t)x * ng + y) * ng + z] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();   // every thread has read bvel[blk]
    if (threadIdx.x == 0) bvel[blk] = 0;
}

__global__ void k_time_set(double* __restrict__ t, double v, int add) {
    if (threadIdx.x == 0 && blockIdx.x == 0) t[0] = add ? t[0] + v : v;
}

__global__ void __launch_bounds__(kWG) k_zero4(float4* __restrict__ g, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (i < n) g[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i64(int64_t* __restrict__ p, int n, int64_t v) {
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i < n) p[i] = v;
}

// Moving a state between the caller's order (flat, slot = caller index) and
// the bins is a permutation of 40 planes; bins hold their particles in no
// pid order, so one side of it is random.
//
// Insert: every particle of a flat state (in a batch the scenes' particles
// follow each other) takes a slot in the bin of its base cell (x = x0 + u)
// in plane set `planes` and scatters its planes there.  (A gather instead --
// slots first, then every slot reading its particle's planes -- measured
// slower at 1M: 1.09 ms against 0.69 ms, its per-bin workgroups too few to
// hide the random reads; the per-bin slot counters' atomics take ~0.24 ms
// of either.)
__global__ void __launch_bounds__(kWG)
    k_bins_insert(const float* __restrict__ flat, int64_t fstride, int n, SceneTab tab, BinGeom g,
                  float* __restrict__ planes, int* __restrict__ cnt, int* __restrict__ status) {
    const int p = blockIdx.x * kWG + threadIdx.x;
    if (p >= n) return;
    const SceneDesc sc = scene_of_pid(tab, p);
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = flat[(MPM_P_X0 + d) * fstride + p] + flat[(MPM_P_U + d) * fstride + p];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = min(max((int)(x[d] * sc.prm.inv_dx - 0.5f), 0), sc.prm.n_grid - 3);
    const int bin = sc.bin_base + bin_of_base(b, sc.nb);
    const int pos = atomicAdd(&cnt[bin], 1);
    int64_t off, S;
    if (pos < sc.cap) {
        off = bin_off(sc, bin) + pos;
        S = sc.cap;
    } else {
        const int r = atomicAdd(&cnt[g.n_bins], 1);
        if (r >= g.ovf_cap) {
            atomicAdd(&status[0], 1);
            return;
        }
        off = ovf_off(g, r);
        S = g.ovf_blk;
    }
#pragma unroll
    for (int k0 = 0; k0 < MPM_NPLANES; k0 += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = flat[(int64_t)(k0 + j) * fstride + p];
#pragma unroll
        for (int j = 0; j < 8; ++j) planes[off + (int64_t)(k0 + j) * S] = v[j];
    }
}

// The filled slots of workgroup wg: slots [0, n_in) of bin wg, or, for
// wg >= n_bins, overflow slots r0 + [0, n_in) with r0 = (wg - n_bins) * kWG
// (they may straddle two overflow blocks).  slot_at gives slot i's float
// offset (plane 0); S is the plane stride.
struct BlockSlots {
    int n_in;
    bool ovf;
    int64_t base;   // bin: offset of slot 0; overflow: r0
    int64_t S;
    __device__ __forceinline__ int64_t slot_at(const BinGeom& g, int i) const {
        return ovf ? ovf_off(g, (int)base + i) : base + i;
    }
};

__device__ __forceinline__ BlockSlots block_slots(const int* __restrict__ cnt, const BinGeom& g,
                                                  const SceneTab& tab, int wg) {
    BlockSlots b;
    b.ovf = wg >= g.n_bins;
    if (b.ovf) {
        const int r0 = (wg - g.n_bins) * kWG;
        b.base = r0;
        b.S = g.ovf_blk;
        b.n_in = min(max(min(cnt[g.n_bins], g.ovf_cap) - r0, 0), kWG);
    } else {
        const SceneDesc sc = scene_of_bin(tab, wg);
        b.base = bin_off(sc, wg);
        b.S = sc.cap;
        b.n_in = min(cnt[wg], sc.cap);
    }
    return b;
}

// Extract, pass 1: the inverse slot map of plane set `planes`: map[pid] =
// float offset (plane 0) of the particle's slot, for every filled slot
// (the caller fills map with -1: particles not in the bins keep it).
__global__ void __launch_bounds__(kWG)
    k_bins_slotmap(const float* __restrict__ planes, const int* __restrict__ cnt, BinGeom g, SceneTab tab,
                   int64_t* __restrict__ map, int n) {
    const BlockSlots bs = block_slots(cnt, g, tab, blockIdx.x);
    for (int i = threadIdx.x; i < bs.n_in; i += kWG) {
        const int64_t off = bs.slot_at(g, i);
        const int pid = __float_as_int(planes[off + (int64_t)MPM_P_PID * bs.S]);
        if (pid >= 0 && pid < n) map[pid] = off;
    }
}

// Extract, pass 2: flat[k][pid] <- plane k of the particle's slot, for the
// planes of plane_mask (yield_stress only in metal scenes: the other
// materials do not carry it).  blockIdx.y takes planes 8y .. 8y + 7, so the
// workgroups in flight read a few planes' lines, which the caches then hold
// while the other particles of those lines come by.
__global__ void __launch_bounds__(kWG)
    k_bins_gather(const float* __restrict__ planes, const int64_t* __restrict__ map, BinGeom g, SceneTab tab,
                  float* __restrict__ flat, int64_t fstride, int n, uint64_t plane_mask) {
    const int pid = blockIdx.x * kWG + threadIdx.x;
    const int k0 = 8 * blockIdx.y;
    if (pid >= n || ((plane_mask >> k0) & 0xffull) == 0) return;
    const int64_t off = map[pid];
    if (off < 0) return;
    const SceneDesc sc = scene_of_pid(tab, pid);
    const int64_t S = off >= g.ovf_off ? g.ovf_blk : sc.cap;
    const uint64_t want = plane_mask & (sc.prm.material == 1 ? ~0ull : ~(1ull << MPM_P_YIELD));
    {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = ((want >> (k0 + j)) & 1) ? planes[off + (int64_t)(k0 + j) * S] : 0.f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((want >> (k0 + j)) & 1) flat[(int64_t)(k0 + j) * fstride + pid] = v[j];
    }
}

// ===========================================================================
// C-ABI of the binned path (one scene) and of batches of scenes
// ===========================================================================
namespace {

struct WsLayout {
    int64_t off_status;
    int64_t off_vel;
    double4* acc;      // spill grid (all scenes' nodes), FP64
    float4* vel;       // grid velocities (v, m)
    int* gflag;        // spill flags per grid block
    int* status;       // 64 counters
    int* stamps;       // epoch of each bin's tile (-epoch: the bin was split, its parts are in the spill grid)
    int* blive;        // per grid block: the newest epoch it was listed for (BlockList::mark)
    int* bhead;        // BlockList::head (64 ints)
    int* blist;        // BlockList::list, [2][nbins]
    int* bvel;         // per grid block: epoch of the grid update that last wrote its velocities (0: zero, -1: stale)
    int* order;        // schedules of the sort and of k_g2p2g (schedule_bins)
    float4* tiles;     // one tile of kTE^3 nodes per bin
    SceneDesc* table;  // batch: scene table
    MpmCollider* cols; // batch: colliders
    int64_t bytes;
};

int64_t align256(int64_t b) { return ((b + 255) / 256) * 256; }

// Compute units of the current device (sizes the fused kernel's work items),
// queried once per device.  A failed query is not cached and reads as 256:
// the count only steers how finely large bins are split.
int device_cus() {
    static int cache[64];
    int d = 0;
    const bool dev_ok = hipGetDevice(&d) == hipSuccess && d >= 0;
    if (dev_ok && d < 64 && cache[d] > 0) return cache[d];
    int n = 0;
    if (!dev_ok || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || n < 1)
        return 256;
    if (d < 64) cache[d] = n;
    return n;
}

WsLayout ws_layout(int64_t nodes, int64_t nbins, int n_scenes, int n_cols, void* base) {
    char* p = (char*)base;
    WsLayout w;
    int64_t off = 0;
    w.acc = (double4*)(p + off);
    off += align256(nodes * 32);
    w.off_vel = off;
    w.vel = (float4*)(p + off);
    off += align256(nodes * 16);
    w.gflag = (int*)(p + off);
    off += align256(nbins * 4);
    w.status = (int*)(p + off);
    w.off_status = off;
    off += 256;
    w.stamps = (int*)(p + off);
    off += align256(nbins * 4);
    w.blive = (int*)(p + off);
    off += align256(nbins * 4);
    w.bhead = (int*)(p + off);
    off += 256;
    w.blist = (int*)(p + off);
    off += align256(2 * nbins * 4);
    w.bvel = (int*)(p + off);
    off += align256(nbins * 4);
    w.order = (int*)(p + off);
    off += align256((2 * nbins + kItemsFirst) * 4);
    w.tiles = (float4*)(p + off);
    off += align256(nbins * kTE3 * 16);
    w.table = (SceneDesc*)(p + off);
    off += align256((int64_t)n_scenes * (int64_t)sizeof(SceneDesc));
    w.cols = (MpmCollider*)(p + off);
    off += align256((int64_t)n_cols * (int64_t)sizeof(MpmCollider));
    w.bytes = off;
    return w;
}

int64_t bins_of_grid(int ng) {
    const int64_t nb = (ng + kBinEdge - 1) / kBinEdge;
    return nb * nb * nb;
}

// everything a launch needs
struct Ctx {
    BinGeom g;
    SceneTab tab;
    WsLayout w;
    ColliderSet cs;   // the single scene's colliders
};

// n_slots: slots of all bins; max_cap: the largest slots per bin
int check_bins(const MpmBins* b, int64_t n_bins, int64_t n_slots, int max_cap, BinGeom* g) {
    if (!b || !b->planes[0] || !b->planes[1] || !b->count) return set_err(-20, "bins: NULL pointer");
    if (b->bin_cap < 1 || b->overflow_cap < 0 || b->overflow_cap % b->bin_cap != 0)
        return set_err(-22, "bins: bin_cap < 1 or overflow_cap not a multiple of bin_cap");
    if (b->cur != 0 && b->cur != 1) return set_err(-23, "bins: cur must be 0 or 1");
    if (b->wg_threads != 0 && b->wg_threads != 256 && b->wg_threads != 384 && b->wg_threads != 512 &&
        b->wg_threads != 768)
        return set_err(-29, "bins: wg_threads must be 0, 256, 384, 512 or 768");
    if (n_slots + b->overflow_cap > 0x7fffffffLL) return set_err(-28, "bins: more than 2^31-1 slots");
    if (n_bins >= (1LL << 27)) return set_err(-28, "bins: 2^27 bins or more");
    if (max_cap > 16 * 1000 * 1000) return set_err(-22, "bins: more than 16e6 slots per bin");
    const int64_t need = (n_slots + b->overflow_cap) * MPM_NPLANES;
    if (b->stride < need) return set_err(-24, "bins: stride < (bin slots + overflow_cap) * MPM_NPLANES");
    g->n_bins = (int)n_bins;
    g->ovf_cap = b->overflow_cap;
    g->ovf_blk = b->bin_cap;
    g->max_cap = max_cap;
    g->ovf_off = n_slots * MPM_NPLANES;
    g->stride = b->stride;
    g->n_cu = device_cus();
    g->pad = 0;
    return 0;
}

int ctx_single(const MpmParams* prm, const MpmBins* bins, void* workspace, const MpmCollider* colliders,
               int n_colliders, bool need_material, Ctx* c) {
    int rc = check_grid(prm);
    if (rc) return rc;
    if (need_material && (rc = check_params(prm))) return rc;
    const int nb = (prm->n_grid + kBinEdge - 1) / kBinEdge;
    if (bins && bins->nb != nb) return set_err(-21, "bins: nb must be ceil(n_grid / 8)");
    const int64_t n_bins = bins_of_grid(prm->n_grid);
    if (bins && (rc = check_bins(bins, n_bins, n_bins * bins->bin_cap, bins->bin_cap, &c->g))) return rc;
    if ((rc = make_collider_set(colliders, n_colliders, &c->cs))) return rc;
    c->tab.one.prm = *prm;
    c->tab.one.nb = nb;
    c->tab.one.bin_base = 0;
    c->tab.one.col_base = 0;
    c->tab.one.n_cols = 0;
    c->tab.one.cap = bins ? bins->bin_cap : 0;
    c->tab.one.pad = 0;
    c->tab.one.node_base = 0;
    c->tab.one.pid_base = 0;
    c->tab.one.slot_base = 0;
    c->tab.dev = nullptr;
    c->tab.n = 1;
    c->tab.cols = nullptr;
    const int64_t ng = prm->n_grid;
    c->w = ws_layout(ng * ng * ng, bins_of_grid(prm->n_grid), 0, 0, workspace);
    return 0;
}

int ctx_batch(const MpmBatch* batch, const MpmBins* bins, Ctx* c) {
    if (!batch || !batch->workspace || batch->n_scenes < 1) return set_err(-40, "batch: not set up");
    int rc;
    if (bins && (rc = check_bins(bins, batch->n_bins, batch->n_slots, batch->max_cap, &c->g))) return rc;
    make_collider_set(nullptr, 0, &c->cs);
    c->w = ws_layout(batch->n_nodes, batch->n_bins, batch->n_scenes, batch->n_colliders, batch->workspace);
    c->tab.dev = c->w.table;
    c->tab.n = batch->n_scenes;
    c->tab.cols = c->w.cols;
    memset(&c->tab.one, 0, sizeof(SceneDesc));
    return 0;
}

unsigned ovf_wgs(const BinGeom& g, int wg = kWG) { return (unsigned)((g.ovf_cap + wg - 1) / wg); }

int* counts(const MpmBins* b, const BinGeom& g, int set) { return b->count + (int64_t)set * (2 * g.n_bins + 1); }

// k_bins_interleave over every bin and overflow block (order: a schedule or
// nullptr)
int launch_sort(const Ctx& c, const float* src, const int* cnt_src, float* dst, int* cnt_dst, const int* order,
                int sched_epoch, hipStream_t s, bool all_planes = true) {
    const unsigned n_items = (unsigned)(c.g.n_bins + c.g.ovf_cap / c.g.ovf_blk);
    hipLaunchKernelGGL(k_bins_interleave, dim3(n_items), dim3(kWG), 0, s, src, cnt_src, dst, cnt_dst, c.g, c.tab,
                       order, sched_epoch, all_planes ? 1 : 0);
    return check_launch("k_bins_interleave");
}

int op_insert(const Ctx& c, const MpmState* flat, const MpmBins* bins, hipStream_t s) {
    if (!flat || !flat->planes || flat->n < 0 || flat->stride < flat->n) return set_err(-25, "flat state invalid");
    if (flat->n > 0x7fffffffLL) return set_err(-6, "more than 2^31-1 particles");
    int rc;
    // insert into the spare plane set, then copy into `cur` in bank order
    const int spare = 1 - bins->cur;
    int* cnt_spare = counts(bins, c.g, spare);
    int* cnt = counts(bins, c.g, bins->cur);
    const int cb = 2 * c.g.n_bins + 1;
    hipLaunchKernelGGL(k_fill_int, dim3((cb + kWG - 1) / kWG), dim3(kWG), 0, s, cnt_spare, cb, 0);
    if ((rc = check_launch("k_fill_int"))) return rc;
    hipLaunchKernelGGL(k_fill_int, dim3((cb + kWG - 1) / kWG), dim3(kWG), 0, s, cnt, cb, 0);
    if ((rc = check_launch("k_fill_int"))) return rc;
    // new counts: the current k_g2p2g schedule no longer applies
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kWG), 0, s, c.w.order + c.g.n_bins, 1, (int)0x80000000);
    if ((rc = check_launch("k_fill_int"))) return rc;
    if (flat->n == 0) return 0;
    hipLaunchKernelGGL(k_bins_insert, dim3((unsigned)((flat->n + kWG - 1) / kWG)), dim3(kWG), 0, s, flat->planes,
                       flat->stride, (int)flat->n, c.tab, c.g, bins->planes[spare], cnt_spare, c.w.status);
    if ((rc = check_launch("k_bins_insert"))) return rc;
    return launch_sort(c, bins->planes[spare], cnt_spare, bins->planes[bins->cur], cnt, nullptr, 0, s);
}

int op_extract(const Ctx& c, const MpmBins* bins, const MpmState* flat, hipStream_t s,
               uint64_t plane_mask = ~0ull) {
    if (!flat || !flat->planes || flat->n < 0 || flat->stride < flat->n) return set_err(-25, "flat state invalid");
    if (flat->n > 0x7fffffffLL) return set_err(-6, "more than 2^31-1 particles");
    if (flat->n == 0 || plane_mask == 0) return 0;
    // the inverse slot map lives in the other plane set, whose contents are
    // dead between operations (every operation that reads it writes it first)
    if (2 * flat->n > bins->stride) return set_err(-28, "bins: plane set too small for the slot map");
    int64_t* map = (int64_t*)bins->planes[1 - bins->cur];
    const int n = (int)flat->n;
    hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)((n + kWG - 1) / kWG)), dim3(kWG), 0, s, map, n, (int64_t)-1);
    int rc = check_launch("k_fill_i64");
    if (rc) return rc;
    hipLaunchKernelGGL(k_bins_slotmap, dim3(c.g.n_bins + ovf_wgs(c.g)), dim3(kWG), 0, s, bins->planes[bins->cur],
                       (const int*)counts(bins, c.g, bins->cur), c.g, c.tab, map, n);
    if ((rc = check_launch("k_bins_slotmap"))) return rc;
    hipLaunchKernelGGL(k_bins_gather, dim3((unsigned)((n + kWG - 1) / kWG), MPM_NPLANES / 8), dim3(kWG), 0, s,
                       (const float*)bins->planes[bins->cur], (const int64_t*)map, c.g, c.tab, flat->planes,
                       flat->stride, n, plane_mask);
    return check_launch("k_bins_gather");
}

BlockList block_list(const Ctx& c) {
    BlockList bl;
    bl.mark = c.w.blive;
    bl.list = c.w.blist;
    bl.head = c.w.bhead;
    bl.n = (int)c.g.n_bins;
    return bl;
}

// Workgroups of the grid update over the listed blocks, two blocks each (plus
// the schedule's): as many as are resident at once (two per CU at four waves
// per SIMD), which covers every listed block of a 1M-particle column; larger
// lists are taken in turns.
constexpr int kGridWGs = 512;

unsigned grid_wgs(const Ctx& c) { return 1u + (unsigned)(c.g.n_bins < kGridWGs ? c.g.n_bins : kGridWGs); }

int op_clear(const Ctx& c, MpmBins* bins, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_bins, dim3(grid_wgs(c)), dim3(kWG), 0, s, c.w.acc, c.w.vel, c.w.gflag,
                       (const int*)c.w.stamps, (const float4*)c.w.tiles, (int*)nullptr, (const int*)nullptr, (int*)nullptr, c.g, c.tab,
                       c.cs, 0.f, 0.f, bins->epoch, 1, (const double*)nullptr, (float4*)nullptr, c.w.status,
                       block_list(c), c.w.bvel);
    int rc = check_launch("k_grid_bins(clear)");
    if (rc) return rc;
    bins->epoch += 1;  // the pending tiles are no longer current
    return 0;
}

int op_p2g(const Ctx& c, MpmBins* bins, float dt, hipStream_t s) {
    const int epoch_out = bins->epoch + 1;
    // the list of epoch_out starts empty (no grid update of epoch_out - 1 need have run)
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c.w.bhead + (epoch_out & 1), 1, 0);
    int rc = check_launch("k_fill_int");
    if (rc) return rc;
    hipLaunchKernelGGL(k_p2g_bins, dim3(c.g.n_bins + ovf_wgs(c.g)), dim3(kWG), 0, s, c.tab, c.g, dt,
                       bins->planes[bins->cur], (const int*)counts(bins, c.g, bins->cur), (double*)c.w.acc, c.w.gflag,
                       c.w.tiles, c.w.stamps, block_list(c), epoch_out);
    rc = check_launch("k_p2g_bins");
    if (rc) return rc;
    bins->epoch = epoch_out;
    return 0;
}

int op_grid(const Ctx& c, const MpmBins* bins, float time, float dt, hipStream_t s,
            const double* t_base = nullptr) {
    float4* rec = (float4*)bins->record_acc;
    if (rec && c.tab.dev) return set_err(-27, "bins: record_acc is for a single scene, not a batch");
    if (rec) {
        // the record holds this substep's P2G at every node: cleared first
        const int64_t nn = c.tab.dev ? 0 : (int64_t)c.tab.one.prm.n_grid * c.tab.one.prm.n_grid * c.tab.one.prm.n_grid;
        hipLaunchKernelGGL(k_zero4, dim3((unsigned)((nn + kWG - 1) / kWG)), dim3(kWG), 0, s, rec, nn);
        int rc = check_launch("k_zero4");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_grid_bins, dim3(grid_wgs(c)), dim3(kWG), 0, s, c.w.acc, c.w.vel, c.w.gflag,
                       (const int*)c.w.stamps, (const float4*)c.w.tiles, counts(bins, c.g, 1 - bins->cur),
                       (const int*)counts(bins, c.g, bins->cur), c.w.order, c.g, c.tab, c.cs, time, dt, bins->epoch,
                       0, t_base, rec, c.w.status, block_list(c), c.w.bvel);
    return check_launch("k_grid_bins");
}

// store_all = false only for substeps followed by another one of the same
// run (no caller sees the plane set in between)
int op_g2p2g(const Ctx& c, MpmBins* bins, float dt, hipStream_t s, bool store_all = true) {
    const int cur = bins->cur, nxt = 1 - cur;
    SubstepArgs a;
    a.tab = c.tab;
    a.g = c.g;
    a.dt = dt;
    a.epoch_out = bins->epoch + 1;
    a.src = bins->planes[cur];
    a.dst = bins->planes[nxt];
    a.cnt_cur = counts(bins, c.g, cur);
    a.cnt_nxt = counts(bins, c.g, nxt);
    a.vel = c.w.vel;
    a.acc = (double*)c.w.acc;
    a.gflag = c.w.gflag;
    a.status = c.w.status;
    a.tiles = c.w.tiles;
    a.stamps = c.w.stamps;
    a.bl = block_list(c);
    a.items = c.w.order + c.g.n_bins;
    a.sched_epoch = bins->epoch;
    a.store_all = store_all ? 1 : 0;
    if (bins->wg_threads == 256)
        hipLaunchKernelGGL(k_g2p2g<256>, dim3(c.g.n_bins + ovf_wgs(c.g, 256)), dim3(256), 0, s, a);
    else if (bins->wg_threads == 384)
        hipLaunchKernelGGL(k_g2p2g<384>, dim3(c.g.n_bins + ovf_wgs(c.g, 384)), dim3(384), 0, s, a);
    else if (bins->wg_threads == 512)
        hipLaunchKernelGGL(k_g2p2g<512>, dim3(c.g.n_bins + ovf_wgs(c.g, 512)), dim3(512), 0, s, a);
    else
        hipLaunchKernelGGL(k_g2p2g<768>, dim3(c.g.n_bins + ovf_wgs(c.g, 768)), dim3(768), 0, s, a);
    int rc = check_launch("k_g2p2g");
    if (rc) return rc;
    bins->epoch = a.epoch_out;
    bins->cur = nxt;
    return 0;
}

// all_planes false: between a grid update and the fused kernel that follows
// it (the planes that kernel rewrites are not copied, see interleave_bin).
int op_sort(const Ctx& c, MpmBins* bins, hipStream_t s, bool all_planes = true) {
    const int cur = bins->cur, nxt = 1 - cur;
    int* cnt_cur = counts(bins, c.g, cur);
    // With the schedule of this epoch's grid update (checked on the device),
    // only scheduled bins are copied; that grid update has zeroed cnt_nxt, so
    // the others read as empty.  Without one, every bin is copied.
    int rc = launch_sort(c, bins->planes[cur], cnt_cur, bins->planes[nxt], counts(bins, c.g, nxt), c.w.order,
                         bins->epoch, s, all_planes);
    if (rc) return rc;
    // the old set becomes the fused kernel's output: its counters start at zero
    const int cb = 2 * c.g.n_bins + 1;
    hipLaunchKernelGGL(k_fill_int, dim3((cb + kWG - 1) / kWG), dim3(kWG), 0, s, cnt_cur, cb, 0);
    if ((rc = check_launch("k_fill_int"))) return rc;
    bins->cur = nxt;
    bins->since_sort = 0;
    return 0;
}

int op_substep(const Ctx& c, MpmBins* bins, float time, float dt, hipStream_t s, const double* t_base = nullptr,
               bool store_all = true) {
    int rc = op_grid(c, bins, time, dt, s, t_base);
    if (rc) return rc;
    // particles drift across cells and out of bank order; restore it now and
    // then, between the grid update (which schedules the bins) and the fused
    // kernel, which reads the sorted set
    const int period = bins->sort_period == 0 ? kSortPeriod : bins->sort_period;
    if (++bins->since_sort >= period && period > 0 && (rc = op_sort(c, bins, s, false))) return rc;
    return op_g2p2g(c, bins, dt, s, store_all);
}

int op_run(const Ctx& c, MpmBins* bins, double time, float dt, int n_substeps, hipStream_t s) {
    if (n_substeps < 0) return set_err(-26, "n_substeps < 0");
    double t = time;
    for (int k = 0; k < n_substeps; ++k) {
        int rc = op_substep(c, bins, (float)t, dt, s, nullptr, k == n_substeps - 1);
        if (rc) return rc;
        t = t + (double)dt;
    }
    return 0;
}

// the device time base of captured runs: two int32 status slots (48, 49)
double* time_base(const Ctx& c) { return (double*)(c.w.status + 48); }

}  // namespace

// A captured run of n_substeps (opaque to callers, include/mpm_hip.h).
struct MpmGraph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    const void* workspace;   // the batch it was captured for
    int n_bins, n_substeps;
    int cur0;                // plane set a replay starts (and ends) on
    int epoch0, epoch_end;   // epochs of its first grid update and of its pending tiles
    int since_end;           // since_sort after a replay
    float dt;
};

namespace {

int op_graph_create(const Ctx& c, const void* ws_id, const MpmBins* bins, float dt, int n_substeps, MpmGraph** out) {
    if (!out) return set_err(-50, "graph: output pointer is NULL");
    *out = nullptr;
    if (n_substeps < 2) return set_err(-51, "graph: n_substeps < 2");
    MpmBins b = *bins;           // the capture's own bookkeeping; *bins is not advanced
    const int e0 = b.epoch, e_end = b.epoch + n_substeps;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return set_err(-52, "graph: stream");
    int rc = 0;
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        hipStreamDestroy(cs);
        return set_err(-53, "graph: begin capture failed");
    }
    hipLaunchKernelGGL(k_stamp_rebase, dim3((c.g.n_bins + 255) / 256), dim3(256), 0, cs, c.w.stamps, c.w.gflag, block_list(c), c.w.bvel,
                       c.g.n_bins, e_end, e0);
    rc = check_launch("k_stamp_rebase");
    double off = 0.0;
    for (int k = 0; k < n_substeps && !rc; ++k) {
        rc = op_substep(c, &b, (float)off, dt, cs, time_base(c), k == n_substeps - 1);
        off = off + (double)dt;
    }
    if (!rc) {
        hipLaunchKernelGGL(k_time_set, dim3(1), dim3(64), 0, cs, time_base(c), off, 1);
        rc = check_launch("k_time_set");
    }
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (!rc && ec != hipSuccess) rc = set_err(-54, "graph: end capture failed");
    if (!rc && b.cur != bins->cur)
        rc = set_err(-55, "graph: a replay would end on the other plane set; n_substeps plus the sorts it "
                          "contains must be even (e.g. twice the sort period)");
    hipGraphExec_t exec = nullptr;
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        rc = set_err(-56, "graph: instantiate failed");
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    MpmGraph* g = (MpmGraph*)malloc(sizeof(MpmGraph));
    if (!g) {
        hipGraphExecDestroy(exec);
        hipGraphDestroy(graph);
        return set_err(-57, "graph: out of host memory");
    }
    g->graph = graph;
    g->exec = exec;
    g->workspace = ws_id;
    g->n_bins = c.g.n_bins;
    g->n_substeps = n_substeps;
    g->cur0 = bins->cur;
    g->epoch0 = e0;
    g->epoch_end = e_end;
    g->since_end = b.since_sort;
    g->dt = dt;
    *out = g;
    return 0;
}

int op_graph_launch(const Ctx& c, const void* ws_id, MpmGraph* g, MpmBins* bins, double time, int n_replays,
                    hipStream_t s) {
    if (!g) return set_err(-50, "graph is NULL");
    if (g->workspace != ws_id || g->n_bins != c.g.n_bins) return set_err(-58, "graph: captured for another batch");
    if (bins->cur != g->cur0) return set_err(-59, "graph: the state is on the other plane set than at capture");
    if (n_replays < 0) return set_err(-26, "n_replays < 0");
    if (n_replays == 0) return 0;
    // the pending tiles take the stamp a replay's rebase expects; the time base
    hipLaunchKernelGGL(k_stamp_rebase, dim3((c.g.n_bins + 255) / 256), dim3(256), 0, s, c.w.stamps, c.w.gflag, block_list(c), c.w.bvel,
                       c.g.n_bins, bins->epoch, g->epoch_end);
    int rc = check_launch("k_stamp_rebase");
    if (rc) return rc;
    hipLaunchKernelGGL(k_time_set, dim3(1), dim3(64), 0, s, time_base(c), time, 0);
    if ((rc = check_launch("k_time_set"))) return rc;
    for (int r = 0; r < n_replays; ++r)
        if (hipGraphLaunch(g->exec, s) != hipSuccess) return set_err(-60, "graph: launch failed");
    bins->epoch = g->epoch_end;
    bins->since_sort = g->since_end;
    return 0;
}

}  // namespace

extern "C" {

int64_t mpm_bins_workspace_bytes(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    const int64_t ng = prm->n_grid;
    return ws_layout(ng * ng * ng, bins_of_grid(prm->n_grid), 0, 0, nullptr).bytes;
}

int64_t mpm_bins_status_offset(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    const int64_t ng = prm->n_grid;
    return ws_layout(ng * ng * ng, bins_of_grid(prm->n_grid), 0, 0, nullptr).off_status;
}

int64_t mpm_bins_velocity_offset(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    const int64_t ng = prm->n_grid;
    return ws_layout(ng * ng * ng, bins_of_grid(prm->n_grid), 0, 0, nullptr).off_vel;
}

int mpm_bins_insert(const MpmParams* prm, const MpmState* flat, const MpmBins* bins, void* workspace,
                    void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, nullptr, 0, false, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_insert(c, flat, bins, (hipStream_t)stream);
}

int mpm_bins_extract(const MpmParams* prm, const MpmBins* bins, const MpmState* flat, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, nullptr, nullptr, 0, false, &c);
    if (rc) return rc;
    return op_extract(c, bins, flat, (hipStream_t)stream);
}

int mpm_bins_extract_planes(const MpmParams* prm, const MpmBins* bins, const MpmState* flat, uint64_t plane_mask,
                            void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, nullptr, nullptr, 0, false, &c);
    if (rc) return rc;
    if (plane_mask >> MPM_NPLANES) return set_err(-26, "plane_mask names planes past MPM_NPLANES");
    return op_extract(c, bins, flat, (hipStream_t)stream, plane_mask);
}

int mpm_bins_settle_grid(const MpmParams* prm, const MpmBins* bins, void* workspace, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, nullptr, 0, false, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    if (c.g.n_bins < 1) return 0;
    hipLaunchKernelGGL(k_settle_vel, dim3((unsigned)c.g.n_bins), dim3(kWG), 0, (hipStream_t)stream, c.w.vel, c.w.bvel,
                       (const int*)c.w.bhead, c.g, c.tab);
    return check_launch("k_settle_vel");
}

int mpm_bins_clear_grid(const MpmParams* prm, MpmBins* bins, void* workspace, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, nullptr, 0, false, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_clear(c, bins, (hipStream_t)stream);
}

int mpm_bins_p2g(const MpmParams* prm, MpmBins* bins, void* workspace, float dt, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, nullptr, 0, true, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_p2g(c, bins, dt, (hipStream_t)stream);
}

int mpm_bins_grid_update(const MpmParams* prm, const MpmBins* bins, const MpmCollider* colliders,
                         int n_colliders, void* workspace, float time, float dt, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, colliders, n_colliders, true, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_grid(c, bins, time, dt, (hipStream_t)stream);
}

int mpm_bins_g2p2g(const MpmParams* prm, MpmBins* bins, void* workspace, float dt, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, nullptr, 0, true, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_g2p2g(c, bins, dt, (hipStream_t)stream);
}

int mpm_bins_sort(const MpmParams* prm, MpmBins* bins, void* workspace, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, nullptr, 0, false, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_sort(c, bins, (hipStream_t)stream);
}

int mpm_bins_substep(const MpmParams* prm, MpmBins* bins, const MpmCollider* colliders, int n_colliders,
                     void* workspace, float time, float dt, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, colliders, n_colliders, true, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_substep(c, bins, time, dt, (hipStream_t)stream);
}

int mpm_bins_profile(const MpmParams* prm, MpmBins* bins, const MpmCollider* colliders, int n_colliders,
                     void* workspace, double time, float dt, int n_substeps, void* stream, float* grid_ms,
                     float* g2p2g_ms, float* sort_ms) {
    if (!bins || !grid_ms || !g2p2g_ms || !sort_ms) return set_err(-20, "profile: NULL pointer");
    if (n_substeps < 1) return set_err(-26, "n_substeps < 1");
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, colliders, n_colliders, true, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[4];
    for (int k = 0; k < 4; ++k)
        if (hipEventCreate(&ev[k]) != hipSuccess) return set_err(-27, "hipEventCreate failed");
    double t = time;
    float acc_grid = 0.f, acc_g2p2g = 0.f, acc_sort = 0.f;
    const int period = bins->sort_period == 0 ? kSortPeriod : bins->sort_period;
    for (int k = 0; k < n_substeps && rc == 0; ++k) {
        (void)hipEventRecord(ev[0], s);
        rc = op_grid(c, bins, (float)t, dt, s);
        (void)hipEventRecord(ev[1], s);
        const bool sorted = rc == 0 && ++bins->since_sort >= period && period > 0;
        if (sorted) rc = op_sort(c, bins, s, false);
        (void)hipEventRecord(ev[2], s);
        if (rc == 0) rc = op_g2p2g(c, bins, dt, s, k == n_substeps - 1);
        (void)hipEventRecord(ev[3], s);
        (void)hipEventSynchronize(ev[3]);
        float a = 0.f, b = 0.f, cc = 0.f;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&cc, ev[1], ev[2]);
        (void)hipEventElapsedTime(&b, ev[2], ev[3]);
        acc_grid += a;
        acc_g2p2g += b;
        // the window of a substep without a sort holds only the two events'
        // own overhead (~5 us), which is not the sort's
        if (sorted) acc_sort += cc;
        t = t + (double)dt;
    }
    for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ev[k]);
    *grid_ms = acc_grid / n_substeps;
    *g2p2g_ms = acc_g2p2g / n_substeps;
    *sort_ms = acc_sort / n_substeps;
    return rc;
}

int mpm_bins_run(const MpmParams* prm, MpmBins* bins, const MpmCollider* colliders, int n_colliders,
                 void* workspace, double time, float dt, int n_substeps, void* stream) {
    Ctx c;
    int rc = ctx_single(prm, bins, workspace, colliders, n_colliders, true, &c);
    if (rc) return rc;
    if (!workspace) return set_err(-10, "workspace is NULL");
    return op_run(c, bins, time, dt, n_substeps, (hipStream_t)stream);
}

// ---- batches ----------------------------------------------------------------

int64_t mpm_batch_workspace_bytes(const MpmScene* scenes, int n_scenes) {
    if (!scenes || n_scenes < 1) return -1;
    int64_t nodes = 0, nbins = 0, ncols = 0;
    for (int i = 0; i < n_scenes; ++i) {
        if (check_grid(&scenes[i].params)) return -1;
        const int64_t ng = scenes[i].params.n_grid;
        nodes += ng * ng * ng;
        nbins += bins_of_grid(scenes[i].params.n_grid);
        ncols += scenes[i].n_colliders;
    }
    return ws_layout(nodes, nbins, n_scenes, (int)ncols, nullptr).bytes;
}

int mpm_batch_setup(const MpmScene* scenes, int n_scenes, const MpmCollider* colliders, void* workspace,
                    MpmBatch* batch) {
    if (!scenes || n_scenes < 1) return set_err(-41, "batch: no scenes");
    if (!workspace || !batch) return set_err(-42, "batch: NULL workspace or batch");
    SceneDesc* tab = (SceneDesc*)malloc(sizeof(SceneDesc) * (size_t)n_scenes);
    if (!tab) return set_err(-43, "batch: host allocation failed");
    int64_t nodes = 0, nbins = 0, pids = 0, slots = 0;
    int ncols = 0, max_cap = 0;
    for (int i = 0; i < n_scenes; ++i) {
        const MpmScene& sc = scenes[i];
        int rc = check_params(&sc.params);
        if (!rc && (sc.n_particles < 0 || sc.n_colliders < 0 || sc.n_colliders > MPM_MAX_COLLIDERS ||
                    sc.bin_cap < 1))
            rc = set_err(-44, "batch: scene with n_particles < 0, bin_cap < 1 or n_colliders out of [0, 16]");
        if (rc) {
            free(tab);
            return rc;
        }
        const int64_t ng = sc.params.n_grid;
        tab[i].prm = sc.params;
        tab[i].nb = (int)((ng + kBinEdge - 1) / kBinEdge);
        tab[i].bin_base = (int)nbins;
        tab[i].col_base = ncols;
        tab[i].n_cols = sc.n_colliders;
        tab[i].cap = sc.bin_cap;
        tab[i].pad = 0;
        tab[i].node_base = nodes;
        tab[i].pid_base = pids;
        tab[i].slot_base = slots * MPM_NPLANES;
        nodes += ng * ng * ng;
        nbins += bins_of_grid(sc.params.n_grid);
        pids += sc.n_particles;
        slots += bins_of_grid(sc.params.n_grid) * sc.bin_cap;
        ncols += sc.n_colliders;
        max_cap = sc.bin_cap > max_cap ? sc.bin_cap : max_cap;
    }
    if (ncols > 0 && !colliders) {
        free(tab);
        return set_err(-45, "batch: colliders is NULL");
    }
    if (nbins > 0x3fffffff || pids > 0x7fffffffLL || slots > 0x7fffffffLL) {
        free(tab);
        return set_err(-46, "batch: too many bins or particles");
    }
    WsLayout w = ws_layout(nodes, nbins, n_scenes, ncols, workspace);
    hipError_t e = hipMemcpy(w.table, tab, sizeof(SceneDesc) * (size_t)n_scenes, hipMemcpyHostToDevice);
    free(tab);
    if (e == hipSuccess && ncols > 0)
        e = hipMemcpy(w.cols, colliders, sizeof(MpmCollider) * (size_t)ncols, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_err(-47, "batch: copying the scene table failed");
    batch->n_scenes = n_scenes;
    batch->n_bins = (int32_t)nbins;
    batch->n_colliders = ncols;
    batch->n_particles = pids;
    batch->n_nodes = nodes;
    batch->n_slots = slots;
    batch->max_cap = max_cap;
    batch->workspace = workspace;
    return 0;
}

int64_t mpm_batch_status_offset(const MpmBatch* batch) {
    if (!batch || batch->n_scenes < 1) return -1;
    return ws_layout(batch->n_nodes, batch->n_bins, batch->n_scenes, batch->n_colliders, nullptr).off_status;
}

int mpm_batch_insert(const MpmBatch* batch, const MpmState* flat, const MpmBins* bins, void* stream) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    if (flat && flat->n != batch->n_particles) return set_err(-48, "batch: flat state size != batch particles");
    return op_insert(c, flat, bins, (hipStream_t)stream);
}

int mpm_batch_extract(const MpmBatch* batch, const MpmBins* bins, const MpmState* flat, void* stream) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    return op_extract(c, bins, flat, (hipStream_t)stream);
}

int mpm_batch_clear_grid(const MpmBatch* batch, MpmBins* bins, void* stream) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    return op_clear(c, bins, (hipStream_t)stream);
}

int mpm_batch_p2g(const MpmBatch* batch, MpmBins* bins, float dt, void* stream) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    return op_p2g(c, bins, dt, (hipStream_t)stream);
}

int mpm_batch_run(const MpmBatch* batch, MpmBins* bins, double time, float dt, int n_substeps, void* stream) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    return op_run(c, bins, time, dt, n_substeps, (hipStream_t)stream);
}

int mpm_batch_graph_create(const MpmBatch* batch, const MpmBins* bins, float dt, int n_substeps, MpmGraph** graph) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    return op_graph_create(c, batch->workspace, bins, dt, n_substeps, graph);
}

int mpm_batch_graph_launch(MpmGraph* graph, const MpmBatch* batch, MpmBins* bins, double time, int n_replays,
                           void* stream) {
    Ctx c;
    int rc = ctx_batch(batch, bins, &c);
    if (rc) return rc;
    return op_graph_launch(c, batch->workspace, graph, bins, time, n_replays, (hipStream_t)stream);
}

int mpm_graph_destroy(MpmGraph* graph) {
    if (!graph) return 0;
    hipGraphExecDestroy(graph->exec);
    hipGraphDestroy(graph->graph);
    free(graph);
    return 0;
}

}  // extern "C"
