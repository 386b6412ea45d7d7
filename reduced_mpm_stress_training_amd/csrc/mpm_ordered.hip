// mpm_ordered.hip -- the flat state in bin order (C-ABI: mpm_order_bytes,
// mpm_sort_by_bin, mpm_step_ordered, mpm_permute in include/mpm_hip.h).
//
// The stage path's P2G (k_p2g_atomic) issues 27 x 4 float atomics per
// particle into the global grid; with every atomic resolved far from the
// CU, it is the cost of a flat substep.  Here the flat state is first
// counting-sorted by bin and sub-block of its stencil base (k_order_hist,
// k_order_scan, k_order_scatter), which also yields a table of chunks --
// runs of at most 256, 512 or 1024 particles inside one bin.  k_p2g_chunks then gives
// each chunk one workgroup that scatters into an FP64 LDS tile placed one
// node below its bin (chunk_origin; ds_add_f64) and flushes the tile with one
// float atomic per touched node and component.  The order stays valid while
// particles drift (a particle whose stencil leaves its chunk's tile scatters
// straight to the grid), so a caller sorts once and steps many substeps.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <limits.h>

#include "mpm_common.h"
#include "mpm_device.h"
#include "mpm_hip.h"
#include "mpm_ordered.h"
#include "mpm_stencil.h"

using namespace mpm;

namespace {

constexpr int kB = 256;       // threads of the per-particle kernels
constexpr int kScan = 1024;   // threads of the single-workgroup scan

__device__ __forceinline__ void stencil_base(const float* __restrict__ sp, int64_t S, float inv_dx, int base[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float x = sp[(MPM_P_X0 + d) * S] + sp[(MPM_P_U + d) * S];
        base[d] = (int)(x * inv_dx - 0.5f);   // mpm_utils.py:197-199
    }
}

__device__ __forceinline__ bool base_ok(const int b[3], int ng) {
    return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] <= ng - 3 && b[1] <= ng - 3 && b[2] <= ng - 3;
}

__global__ void k_order_fill(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}

// Runs of equal keys in consecutive lanes of a wave: *first = the run's
// first lane, returns the run's length.  Lane 0 always starts a run.  The
// sort's counters take one atomic per run instead of one per particle --
// a state that is already in order has one or two runs per wave.
__device__ __forceinline__ int key_run(int key, int* first) {
    const int lane = __lane_id();
    const int left = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || left != key);
    const unsigned long long upto = (2ull << lane) - 1ull;   // lanes 0 .. lane
    *first = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    return next - *first;
}

// Workgroup aggregation of the sort's counters.  The keys of a workgroup's
// particles mostly lie in a narrow range (a state sorted an earlier segment
// ago has moved little), so each workgroup counts its keys in LDS over a
// window of kWin keys from its smallest key and takes one global atomic per
// key present, instead of one per run of equal keys in a wave (bank order
// inside chunks leaves runs of one or two particles, and thousands of waves
// hitting the same few counters serialize).  Keys outside the window (an
// unsorted state) take one atomic per run of equal keys in a wave.
constexpr int kWin = 512;

// smallest valid key of the workgroup (INT_MAX if none) in *s_kmin, which
// thread 0 set to INT_MAX before a barrier; a barrier must follow
__device__ __forceinline__ void wg_min_key(int key, int* s_kmin) {
    int k = key >= 0 ? key : INT_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k = min(k, __shfl_xor(k, o));
    if (__lane_id() == 0) atomicMin(s_kmin, k);
}

__global__ void __launch_bounds__(kB)
    k_order_hist(const float* __restrict__ planes, int64_t S, int n, float inv_dx, int ng, int nb,
                 int* __restrict__ hist) {
    __shared__ int s_cnt[kWin];
    __shared__ int s_kmin;
    const int tid = threadIdx.x;
    const int p = blockIdx.x * kB + tid;
    int key = -1;
    if (p < n) {
        int b[3];
        stencil_base(planes + p, S, inv_dx, b);
        key = order_key(b, ng, nb);
    }
    if (tid == 0) s_kmin = INT_MAX;
    for (int t = tid; t < kWin; t += kB) s_cnt[t] = 0;
    __syncthreads();
    wg_min_key(key, &s_kmin);
    __syncthreads();
    const int kmin = s_kmin;
    const bool in = key >= 0 && key - kmin < kWin;
    if (in) atomicAdd(&s_cnt[key - kmin], 1);
    int first;
    const int len = key_run(in ? -2 : key, &first);
    if (!in && key >= 0 && __lane_id() == first) atomicAdd(&hist[key], len);
    __syncthreads();
    for (int t = tid; t < kWin; t += kB)
        if (s_cnt[t] > 0) atomicAdd(&hist[kmin + t], s_cnt[t]);
}

// One workgroup: exclusive offsets of the key counts (in place, and copied
// to the cursors) and the chunk table.  Thread t owns a contiguous range of
// bins, so every bin's keys and chunks are handled by one thread.
__global__ void __launch_bounds__(kScan)
    k_order_scan(int* __restrict__ hist, int* __restrict__ cursor, int4* __restrict__ chunks, int* __restrict__ meta,
                 int nbins, int max_chunks, int chunk) {
    __shared__ int s_p[kScan], s_c[kScan];
    const int tid = threadIdx.x;
    const int per = (nbins + kScan - 1) / kScan;
    const int b_lo = min(tid * per, nbins), b_hi = min(b_lo + per, nbins);
    int np = 0, nc = 0;
    for (int b = b_lo; b < b_hi; ++b) {
        int c = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) c += hist[8 * b + k];
        np += c;
        nc += (c + chunk - 1) / chunk;
    }
    s_p[tid] = np;
    s_c[tid] = nc;
    __syncthreads();
    for (int o = 1; o < kScan; o <<= 1) {   // inclusive scan
        const int a = tid >= o ? s_p[tid - o] : 0, ac = tid >= o ? s_c[tid - o] : 0;
        __syncthreads();
        s_p[tid] += a;
        s_c[tid] += ac;
        __syncthreads();
    }
    int run = s_p[tid] - np, crun = s_c[tid] - nc;
    for (int b = b_lo; b < b_hi; ++b) {
        const int first = run;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int c = hist[8 * b + k];
            hist[8 * b + k] = run;
            cursor[8 * b + k] = run;
            run += c;
        }
        const int cnt = run - first;
        for (int s = 0; s < cnt; s += chunk, ++crun)
            if (crun < max_chunks) chunks[crun] = make_int4(first + s, min(chunk, cnt - s), b, 0);
    }
    if (tid == kScan - 1) {
        hist[8 * nbins] = s_p[kScan - 1];
        meta[0] = min(s_c[kScan - 1], max_chunks);
    }
    if (tid < 63) meta[1 + tid] = 0;   // status counters of the substeps on this order
}

// Each particle's sorted slot: perm[slot] = source index (the planes move in
// k_order_interleave, which gathers them straight into bank order).
__global__ void __launch_bounds__(kB)
    k_order_scatter(const float* __restrict__ src, int64_t SS, int n, float inv_dx, int ng, int nb,
                    int* __restrict__ cursor, int* __restrict__ perm) {
    __shared__ int s_cnt[kWin], s_base[kWin];
    __shared__ int s_kmin;
    const int tid = threadIdx.x;
    const int p = blockIdx.x * kB + tid;
    int key = -1;
    if (p < n) {
        int b[3];
        stencil_base(src + p, SS, inv_dx, b);
        key = order_key(b, ng, nb);
    }
    if (tid == 0) s_kmin = INT_MAX;
    for (int t = tid; t < kWin; t += kB) s_cnt[t] = 0;
    __syncthreads();
    wg_min_key(key, &s_kmin);
    __syncthreads();
    const int kmin = s_kmin;
    const bool in = key >= 0 && key - kmin < kWin;
    const int local = in ? atomicAdd(&s_cnt[key - kmin], 1) : 0;   // rank among the workgroup's equal keys
    int first;
    const int len = key_run(in ? -2 : key, &first);
    int base = 0;
    if (!in && key >= 0 && __lane_id() == first) base = atomicAdd(&cursor[key], len);
    base = __shfl(base, first);
    __syncthreads();
    for (int t = tid; t < kWin; t += kB)
        if (s_cnt[t] > 0) s_base[t] = atomicAdd(&cursor[kmin + t], s_cnt[t]);
    __syncthreads();
    if (key < 0) return;
    const int pos = in ? s_base[key - kmin] + local : base + (__lane_id() - first);
    if (pos < 0 || pos >= n) return;   // cannot happen: same keys as the histogram
    perm[pos] = p;
}

// Bank order inside each chunk, in memory: one workgroup per chunk gathers
// the chunk's particles (all planes) from the unsorted source straight into
// bank order (chunk_bank_order) of their stencil bases in the chunk's tile
// (chunk_origin), and reorders the chunk's permutation entries to match.
// The chunk kernels then read consecutive slots (coalesced) whose 32-lane
// halves gather and scatter through distinct LDS banks.
__global__ void __launch_bounds__(kChunkWG)
    k_order_interleave(const float* __restrict__ src, int64_t SS, float* __restrict__ dst, int64_t DS, int n,
                       const int4* __restrict__ chunks, const int* __restrict__ meta, float inv_dx, int ng,
                       int* __restrict__ perm) {
    __shared__ unsigned short s_ord[kChunkMax];
    __shared__ int s_cnt[32];
    const int c = blockIdx.x;
    if (c >= meta[0]) return;
    const int4 ch = chunks[c];
    const int tid = threadIdx.x;
    const int n_per = min(ch.y, kChunkMax);
    if (ch.x < 0 || ch.x + n_per > n || n_per < 1) return;
    int o[3];
    chunk_origin(ch.z, (ng + 7) / 8, o);
    chunk_bank_order(src, SS, perm, ch.x, n_per, n, inv_dx, o, s_ord, s_cnt);   // barriers inside
    constexpr int R = kChunkMax / kChunkWG;
    int from[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int q = tid + r * kChunkWG;
        from[r] = q < n_per ? perm[ch.x + s_ord[q]] : -1;
    }
#pragma unroll 3
    for (int k = 0; k < MPM_NPLANES; ++k)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int q = tid + r * kChunkWG;
            if (from[r] >= 0 && from[r] < n) dst[(int64_t)k * DS + ch.x + q] = src[(int64_t)k * SS + from[r]];
        }
    __syncthreads();   // every thread has read the chunk's old permutation entries
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int q = tid + r * kChunkWG;
        if (q < n_per) perm[ch.x + q] = from[r];
    }
}

// ---------------------------------------------------------------------------
// p2g_apic_stress_given (mpm_utils.py:190-227) over chunks: one workgroup per
// chunk, its particles kChunkWG at a time.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kChunkWG)
    k_p2g_chunks(const float* __restrict__ planes, int64_t S, int n, const int4* __restrict__ chunks,
                 const int* __restrict__ meta, MpmParams prm, float dt, float* __restrict__ grid,
                 int* __restrict__ flags) {
    __shared__ double s_t[4 * kCT3];
    const int c = blockIdx.x;
    if (c >= meta[0]) return;
    const int4 ch = chunks[c];
    const int tid = threadIdx.x;
    const int ng = prm.n_grid;
    const int n_per = min(ch.y, kChunkMax);   // a chunk's particles, kChunkWG at a time
    for (int t = tid; t < 4 * kCT3; t += kChunkWG) s_t[t] = 0.0;
    int o[3];
    chunk_origin(ch.z, (ng + 7) / 8, o);
    __syncthreads();
    for (int i = tid; i < n_per; i += kChunkWG) {
        const int p = ch.x + i;
        if (p < 0 || p >= n) continue;
        const float* sp = planes + p;
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
        AxisWeights A;
        axis_weights(x, prm.inv_dx, A);
        if (!base_ok(A.base, ng)) continue;
        const int ti = chunk_tile_index(A.base, o);
        if (ti >= 0) {
            p2g_contrib(A, m, vol, v, Cm, tau, prm.dx, prm.inv_dx, dt, [&](int cc, int i3, int j3, int k3, float val) {
                atomicAdd(&s_t[cc * kCT3 + ti + (i3 * kCT + j3) * kCT + k3], (double)val);
            });
        } else {
            p2g_contrib_rolled(A, m, vol, v, Cm, tau, prm.dx, prm.inv_dx, dt,
                               [&](int cc, int i3, int j3, int k3, float val) {
                                   const int64_t node =
                                       ((int64_t)(A.base[0] + i3) * ng + (A.base[1] + j3)) * ng + (A.base[2] + k3);
                                   unsafeAtomicAdd(grid + 4 * node + cc, val);
                               });
            if (flags) mark_stencil_blocks(flags, A.base, (ng + 7) / 8);
        }
    }
    __syncthreads();
    chunk_flush<4>(s_t, o, ng, grid);
    if (flags) mark_chunk_blocks(flags, ch.z, (ng + 7) / 8, tid);
}

// ---------------------------------------------------------------------------
// Fused G2P (this substep, g2p_apic_and_stress, mpm_utils.py:241-306) and
// P2G (the next substep, p2g_apic_stress_given, :190-227) over chunks.  The
// chunk's grid velocities are staged in an LDS tile at its origin o
// (chunk_origin); each particle gathers, runs F update / return map / stress
// in registers, writes its new state (u = x - x0, v, C, F, tau; in place, or
// the whole state into `dst`), and scatters its next-substep P2G into an
// FP64 LDS tile at the same origin (the tile holds a cell of drift either
// way), flushed once per node into acc_next.
// Stencils outside the tiles use the global grids.  Frozen particles (stencil
// outside the grid) are left unchanged and do not scatter, as in the stage
// kernels.
// ---------------------------------------------------------------------------
struct OrdNodesTile {
    const float* s_v;
    int ti;
    __device__ __forceinline__ float comp(int c, int t) const {
        const int i = t / 9, j = (t / 3) % 3, k = t % 3;
        return s_v[c * kCT3 + ti + (i * kCT + j) * kCT + k];
    }
};
struct OrdNodesGrid {
    const float4* gv;
    int64_t gb;
    int ng;
    __device__ __forceinline__ float comp(int c, int t) const {
        const int i = t / 9, j = (t / 3) % 3, k = t % 3;
        const float4 g = gv[gb + ((int64_t)i * ng + j) * ng + k];
        return c == 0 ? g.x : (c == 1 ? g.y : g.z);
    }
};

// A particle's inputs to the fused substep; the chunk loop loads the next
// particle's while it processes the current one.
struct FlatIn {
    float x0[3], u[3], F[9], m, vol, pid;
};

// cst: where the constant planes (x0, m, vol, id) live -- src itself, or the
// first state of a recorded segment, whose later states hold only the planes
// a substep changes.
__device__ __forceinline__ bool const_plane(int k) {
    return k < MPM_P_U || k == MPM_P_MASS || k == MPM_P_VOL || k == MPM_P_PID;
}

__device__ __forceinline__ void load_flat(const float* src, const float* cst, int64_t S, int p, FlatIn& q) {
    const float* sp = src + p;
    const float* cp = cst + p;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q.x0[d] = cp[(MPM_P_X0 + d) * S];
        q.u[d] = sp[(MPM_P_U + d) * S];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) q.F[k] = sp[(MPM_P_F + k) * S];
    q.m = cp[MPM_P_MASS * S];
    q.vol = cp[MPM_P_VOL * S];
    q.pid = cp[MPM_P_PID * S];
}

template <typename Nodes>
__device__ __forceinline__ void g2p2g_flat_particle(const FlatIn& q, bool copy, float* dst, int64_t S, int p,
                                                    const float x[3], const AxisWeights& A, const Nodes& V,
                                                    const MpmParams& prm, float dt, double* s_acc, const int o2[3],
                                                    float* __restrict__ acc_next, int* __restrict__ flags_next,
                                                    int* __restrict__ status) {
    const int ng = prm.n_grid;
    float* dp = dst + p;
    const float* x0 = q.x0;
    float v[3], Cm[9], G[9], F[9], tau[9];
    if constexpr (std::is_same<Nodes, OrdNodesGrid>::value)
        gather_global_rolled(V.gv + V.gb, V.ng, A, prm.inv_dx, v, Cm, G);   // few particles: small live set
    else
        gather_tile<9, 3>(V, 0, A, prm.inv_dx, v, Cm, G);
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = q.F[k];
    const float m = q.m, vol = q.vol;
    if (copy) {   // the constant planes, into a state that must hold them
#pragma unroll
        for (int d = 0; d < 3; ++d) dp[(MPM_P_X0 + d) * S] = x0[d];
        dp[MPM_P_MASS * S] = m;
        dp[MPM_P_VOL * S] = vol;
        dp[MPM_P_PID * S] = q.pid;
    }
    sand_update(G, dt, F, prm.mu, prm.lam, prm.alpha, tau);
    float xn[3], u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        u[d] = (x[d] + dt * v[d]) - x0[d];     // update_u_from_x (mpm_utils.py:12-15)
        xn[d] = x0[d] + u[d];                  // the next substep's update_x_from_u
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dp[(MPM_P_U + d) * S] = u[d];
        dp[(MPM_P_V + d) * S] = v[d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        dp[(MPM_P_C + k) * S] = Cm[k];
        dp[(MPM_P_F + k) * S] = F[k];
        dp[(MPM_P_TAU + k) * S] = tau[k];
    }
    AxisWeights B;
    axis_weights(xn, prm.inv_dx, B);
    if (!base_ok(B.base, ng)) {
        atomicAdd(&status[2], 1);   // left the grid: frozen from here on
        return;
    }
    const int t2 = chunk_tile_index(B.base, o2);
    const int64_t gb = ((int64_t)B.base[0] * ng + B.base[1]) * ng + B.base[2];
    if (t2 >= 0) {
        p2g_contrib(B, m, vol, v, Cm, tau, prm.dx, prm.inv_dx, dt, [&](int cc, int i3, int j3, int k3, float val) {
            atomicAdd(&s_acc[cc * kCT3 + t2 + (i3 * kCT + j3) * kCT + k3], (double)val);
        });
    } else {
        // outside the tile (few particles): the same values with a small live set
        p2g_contrib_rolled(B, m, vol, v, Cm, tau, prm.dx, prm.inv_dx, dt,
                           [&](int cc, int i3, int j3, int k3, float val) {
                               unsafeAtomicAdd(acc_next + 4 * (gb + ((int64_t)i3 * ng + j3) * ng + k3) + cc, val);
                           });
        if (flags_next) mark_stencil_blocks(flags_next, B.base, (ng + 7) / 8);
    }
}

__global__ void __launch_bounds__(kChunkWG) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_g2p2g_chunks(const float* planes, const float* cst, float* dst, int copy_const, int64_t S, int n,
                   const int4* __restrict__ chunks, const int* __restrict__ meta, MpmParams prm, float dt,
                   const float4* __restrict__ gv, float* __restrict__ acc_next, int* __restrict__ flags_next,
                   int* __restrict__ status) {
    __shared__ float s_v[3 * kCT3];
    __shared__ double s_acc[4 * kCT3];
    const int c = blockIdx.x;
    if (c >= meta[0]) return;
    const int4 ch = chunks[c];
    const int tid = threadIdx.x;
    const int ng = prm.n_grid;
    const int n_per = min(ch.y, kChunkMax);
    for (int t = tid; t < 4 * kCT3; t += kChunkWG) s_acc[t] = 0.0;
    int o[3];
    chunk_origin(ch.z, (ng + 7) / 8, o);
    const int* o2 = o;   // the next substep's P2G tile: the same origin
    for (int t = tid; t < kCT3; t += kChunkWG) {
        const int x = o[0] + t / (kCT * kCT), y = o[1] + (t / kCT) % kCT, z = o[2] + t % kCT;
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        if (x >= 0 && y >= 0 && z >= 0 && x < ng && y < ng && z < ng) g = gv[((int64_t)x * ng + y) * ng + z];
        s_v[t] = g.x;
        s_v[kCT3 + t] = g.y;
        s_v[2 * kCT3 + t] = g.z;
    }
    __syncthreads();
    // stencils inside the velocity tile first; the others (bit r of the
    // thread's r-th particle) gather from the global grid afterwards.  The
    // next particle's inputs are loaded while the current one is processed.
    const bool copy = copy_const != 0;
    unsigned outside = 0;
    FlatIn q;
    if (tid < n_per && ch.x + tid < n) load_flat(planes, cst, S, ch.x + tid, q);
    for (int i = tid, r = 0; i < n_per; i += kChunkWG, ++r) {
        const int p = ch.x + i;
        const FlatIn cur = q;
        if (i + kChunkWG < n_per && p + kChunkWG < n) load_flat(planes, cst, S, p + kChunkWG, q);
        if (p < 0 || p >= n) continue;
        float x[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = cur.x0[d] + cur.u[d];
        AxisWeights A;
        axis_weights(x, prm.inv_dx, A);
        if (!base_ok(A.base, ng)) {   // frozen: unchanged
            for (int k = 0; k < MPM_NPLANES; ++k) {
                if (const_plane(k)) {
                    if (copy) dst[(int64_t)k * S + p] = cst[(int64_t)k * S + p];
                } else if (dst != planes) {
                    dst[(int64_t)k * S + p] = planes[(int64_t)k * S + p];
                }
            }
            continue;
        }
        const int ti = chunk_tile_index(A.base, o);
        if (ti < 0) {
            outside |= 1u << r;
            continue;
        }
        g2p2g_flat_particle(cur, copy, dst, S, p, x, A, OrdNodesTile{s_v, ti}, prm, dt, s_acc, o2, acc_next, flags_next,
                            status);
    }
    for (int i = tid, r = 0; outside && i < n_per; i += kChunkWG, ++r) {
        if (!(outside & (1u << r))) continue;
        const int p = ch.x + i;
        load_flat(planes, cst, S, p, q);
        float x[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = q.x0[d] + q.u[d];
        AxisWeights A;
        axis_weights(x, prm.inv_dx, A);
        const OrdNodesGrid V{gv, ((int64_t)A.base[0] * ng + A.base[1]) * ng + A.base[2], ng};
        g2p2g_flat_particle(q, copy, dst, S, p, x, A, V, prm, dt, s_acc, o2, acc_next, flags_next, status);
    }
    __syncthreads();
    chunk_flush<4>(s_acc, o2, ng, acc_next);
    if (flags_next) mark_chunk_blocks(flags_next, ch.z, (ng + 7) / 8, tid);
}

__global__ void __launch_bounds__(kB)
    k_permute(const float* __restrict__ src, int64_t ss, float* __restrict__ dst, int64_t ds, int nplanes,
              const int* __restrict__ idx, int n, int scatter) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int q = idx[j];
    if (q < 0 || q >= n) return;
    const int from = scatter ? j : q, to = scatter ? q : j;
    for (int p = 0; p < nplanes; ++p) dst[(int64_t)p * ds + to] = src[(int64_t)p * ss + from];
}

unsigned nblk(int64_t n, int b = kB) { return (unsigned)((n + b - 1) / b); }

int check_flat(const MpmState* st) {
    if (!st || !st->planes) return set_err(-4, "state or planes is NULL");
    if (st->n < 0 || st->stride < st->n) return set_err(-5, "state n/stride invalid");
    if (st->n > 0x7fffffffLL) return set_err(-6, "more than 2^31-1 particles");
    return 0;
}

}  // namespace

namespace mpm {

int p2g_ordered(const MpmParams* prm, const MpmState* st, const void* order, float* grid_acc, float dt,
                void* stream, int* flags) {
    int rc = check_sand(prm);
    if (rc) return rc;
    if ((rc = check_flat(st))) return rc;
    if (!order || ((uintptr_t)order & 255)) return set_err(-40, "order buffer NULL or not 256-byte aligned");
    if (!grid_acc) return set_err(-7, "grid is NULL");
    if (st->n == 0) return 0;
    const OrderLayout L = order_layout(*prm, st->n);
    const char* ob = (const char*)order;
    hipLaunchKernelGGL(k_p2g_chunks, dim3((unsigned)L.max_chunks), dim3(kChunkWG), 0, (hipStream_t)stream,
                       (const float*)st->planes, st->stride, (int)st->n, (const int4*)(ob + L.off_chunks),
                       (const int*)(ob + L.off_meta), *prm, dt, grid_acc, flags);
    return check_launch("k_p2g_chunks");
}

int g2p2g_ordered(const MpmParams* prm, const MpmState* src, const MpmState* dst, const void* order,
                  const float* grid_v, float* acc_next, float dt, const float* cst, bool copy_const,
                  void* stream, int* flags_next) {
    int rc = check_sand(prm);
    if (rc) return rc;
    if ((rc = check_flat(src)) || (rc = check_flat(dst))) return rc;
    if (src->n != dst->n || src->stride != dst->stride) return set_err(-41, "src and dst differ in n or stride");
    if (!order || ((uintptr_t)order & 255)) return set_err(-40, "order buffer NULL or not 256-byte aligned");
    if (!grid_v || !acc_next) return set_err(-7, "grid is NULL");
    if (src->n == 0) return 0;
    const OrderLayout L = order_layout(*prm, src->n);
    char* ob = (char*)order;
    int* status = (int*)(ob + L.off_meta) + kOrderStatus;
    hipLaunchKernelGGL(k_g2p2g_chunks, dim3((unsigned)L.max_chunks), dim3(kChunkWG), 0, (hipStream_t)stream,
                       (const float*)src->planes, cst ? cst : (const float*)src->planes, (float*)dst->planes,
                       copy_const ? 1 : 0, src->stride, (int)src->n, (const int4*)(ob + L.off_chunks),
                       (const int*)(ob + L.off_meta), *prm, dt, (const float4*)grid_v, acc_next, flags_next, status);
    return check_launch("k_g2p2g_chunks");
}

}  // namespace mpm

extern "C" {

int64_t mpm_order_bytes(const MpmParams* prm, int64_t n_particles) {
    if (check_grid(prm) || n_particles < 0) return -1;
    return order_layout(*prm, n_particles).bytes;
}

int64_t mpm_order_perm_offset(const MpmParams* prm, int64_t n_particles) {
    if (check_grid(prm) || n_particles < 0) return -1;
    return order_layout(*prm, n_particles).off_perm;
}

int mpm_sort_by_bin(const MpmParams* prm, const MpmState* src, const MpmState* dst, void* order, void* stream) {
    int rc = check_grid(prm);
    if (rc) return rc;
    if ((rc = check_flat(src)) || (rc = check_flat(dst))) return rc;
    if (src->n != dst->n) return set_err(-41, "sort: src and dst hold different particle counts");
    if (src->planes == dst->planes) return set_err(-42, "sort: src and dst must not alias");
    if (!order || ((uintptr_t)order & 255)) return set_err(-40, "order buffer NULL or not 256-byte aligned");
    const int64_t n = src->n;
    const OrderLayout L = order_layout(*prm, n);
    if (L.nbins > INT_MAX / 8) return set_err(-43, "sort: too many bins");
    char* ob = (char*)order;
    int* hist = (int*)(ob + L.off_hist);
    int* cursor = (int*)(ob + L.off_cursor);
    int* perm = (int*)(ob + L.off_perm);
    int4* chunks = (int4*)(ob + L.off_chunks);
    int* meta = (int*)(ob + L.off_meta);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_order_fill, dim3(nblk(L.nkeys + 1)), dim3(kB), 0, s, hist, L.nkeys + 1, 0);
    if ((rc = check_launch("k_order_fill"))) return rc;
    if (n > 0) {
        hipLaunchKernelGGL(k_order_hist, dim3(nblk(n)), dim3(kB), 0, s, (const float*)src->planes, src->stride,
                           (int)n, prm->inv_dx, prm->n_grid, (int)L.nb, hist);
        if ((rc = check_launch("k_order_hist"))) return rc;
    }
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(kScan), 0, s, hist, cursor, chunks, meta, (int)L.nbins,
                       (int)L.max_chunks, (int)L.chunk);
    if ((rc = check_launch("k_order_scan"))) return rc;
    if (n > 0) {
        hipLaunchKernelGGL(k_order_scatter, dim3(nblk(n)), dim3(kB), 0, s, (const float*)src->planes, src->stride,
                           (int)n, prm->inv_dx, prm->n_grid, (int)L.nb, cursor, perm);
        if ((rc = check_launch("k_order_scatter"))) return rc;
        hipLaunchKernelGGL(k_order_interleave, dim3((unsigned)L.max_chunks), dim3(kChunkWG), 0, s,
                           (const float*)src->planes, src->stride, (float*)dst->planes, dst->stride, (int)n,
                           (const int4*)chunks, (const int*)meta, prm->inv_dx, prm->n_grid, perm);
        if ((rc = check_launch("k_order_interleave"))) return rc;
    }
    return 0;
}

int mpm_step_ordered(const MpmParams* prm, const MpmState* st, const void* order, const MpmCollider* colliders,
                     int n_colliders, void* workspace, int64_t workspace_bytes, float time, float dt, void* stream) {
    int rc = check_sand(prm);
    if (rc) return rc;
    if ((rc = check_flat(st))) return rc;
    const int64_t need = mpm_workspace_bytes(prm, st->n);
    if (!workspace || workspace_bytes < need) return set_err(-10, "workspace too small");
    if (((uintptr_t)workspace) & 255) return set_err(-11, "workspace not 256-byte aligned");
    const int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    float* acc = (float*)workspace;
    float* gv = acc + 4 * nn;
    if ((rc = mpm_zero_grid(prm, acc, stream))) return rc;
    if ((rc = p2g_ordered(prm, st, order, acc, dt, stream))) return rc;
    if ((rc = mpm_grid_update(prm, acc, gv, colliders, n_colliders, time, dt, stream))) return rc;
    return mpm_g2p(prm, st, gv, dt, stream);
}

int64_t mpm_run_ordered_workspace_bytes(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    return 2 * mpm_step_grids_bytes(prm);   // two grid sets (this substep's, the next's)
}

int mpm_run_ordered(const MpmParams* prm, const MpmState* st, const void* order, const MpmCollider* colliders,
                    int n_colliders, void* workspace, int64_t workspace_bytes, int k0, double dt, int n_sub,
                    void* stream) {
    int rc = check_sand(prm);
    if (rc) return rc;
    if ((rc = check_flat(st))) return rc;
    if (n_sub < 0) return set_err(-31, "n_sub < 0");
    if (!workspace || workspace_bytes < mpm_run_ordered_workspace_bytes(prm))
        return set_err(-10, "workspace too small (mpm_run_ordered_workspace_bytes)");
    if (((uintptr_t)workspace) & 255) return set_err(-11, "workspace not 256-byte aligned");
    if (n_sub == 0) return 0;
    const int ng = prm->n_grid;
    const int64_t nn = (int64_t)ng * ng * ng;
    float* set = (float*)workspace;
    float* set2 = set + mpm_step_grids_bytes(prm) / 4;
    // both sets cleared whole (marks included); from then on each grid
    // update clears the blocks the other set's earlier P2G marked
    if ((rc = clear_grid_set(prm, set, stream)) || (rc = clear_grid_set(prm, set2, stream))) return rc;
    if ((rc = p2g_ordered(prm, st, order, set, (float)dt, stream, grid_set_flags(set, ng)))) return rc;
    for (int j = 0; j < n_sub; ++j) {
        if ((rc = grid_update_blocks(prm, set, colliders, n_colliders, (float)((double)(k0 + j) * dt), (float)dt, set2,
                                     stream)))
            return rc;
        if ((rc = g2p2g_ordered(prm, st, st, order, set + 4 * nn, set2, (float)dt, nullptr, false, stream,
                                grid_set_flags(set2, ng))))
            return rc;
        float* t = set; set = set2; set2 = t;
    }
    return 0;
}

int mpm_permute(const float* src, int64_t src_stride, float* dst, int64_t dst_stride, int n_planes, const int* idx,
                int64_t n, int scatter, void* stream) {
    if (!src || !dst || !idx) return set_err(-44, "permute: NULL pointer");
    if (n < 0 || n > 0x7fffffffLL || src_stride < n || dst_stride < n || n_planes < 1)
        return set_err(-45, "permute: bad sizes");
    if (src == dst) return set_err(-46, "permute: src and dst must not alias");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_permute, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, src, src_stride, dst, dst_stride,
                       n_planes, idx, (int)n, scatter ? 1 : 0);
    return check_launch("k_permute");
}

}  // extern "C"
