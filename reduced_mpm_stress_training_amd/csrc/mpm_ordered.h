// mpm_ordered.h -- the flat particle state in bin order and the chunked,
// LDS-privatized scatter shared by the ordered P2G (mpm_ordered.hip) and the
// ordered G2P adjoint (mpm_adjoint.hip).
//
// mpm_sort_by_bin counting-sorts a flat state by the 8^3-cell bin of each
// particle's stencil base (mpm_utils.py:197-199) and, inside a bin, by its
// 4^3-cell sub-block, and writes a chunk table: runs of at most
// chunk_size(n) consecutive particles that never straddle two bins.  A
// chunk's particles then have stencil bases in their bin's 8 cells per axis
// at sort time (plus what they drift in the substeps after), so one
// workgroup accumulates a whole chunk's scatter into a 12^3-node FP64 LDS
// tile placed one node below the bin (chunk_origin: bases b0 - 1 .. b0 + 8
// fit, one cell of drift either way), and flushes each touched node to the
// global grid with one float atomic per component -- instead of 27 atomics
// per particle per component.  Particles whose stencil leaves the tile go
// straight to the global grid.
#pragma once
#include <hip/hip_runtime.h>
#include <limits.h>

#include "mpm_hip.h"
#include "mpm_stencil.h"

namespace mpm {

constexpr int kChunkWG = 256;   // threads per chunk workgroup
constexpr int kChunkMax = 1024; // the largest chunk

// Particles per chunk: the largest of 1024, 512 and 256 that still gives
// about 900 chunks (three or more workgroups per CU).  Measured on the sand
// column: at 1M particles on 128^3, 1024 beats 512 by 5% in the forward
// substep and 256 by 27%; at 250k on 64^3, 256 beats 512 and 1024 by 10% in
// both the forward and the adjoint substep.
inline int chunk_size(int64_t n) { return n >= 900 * 1024 ? 1024 : (n >= 900 * 512 ? 512 : 256); }
constexpr int kCT = 12;         // LDS tile edge in nodes: bases o .. o+9 fit
constexpr int kCT3 = kCT * kCT * kCT;

// Byte layout of the order buffer (256-byte aligned regions).
struct OrderLayout {
    int64_t nb, nbins, nkeys, max_chunks, chunk;
    int64_t off_hist;     // int[nkeys + 1]: key counts, then exclusive offsets
    int64_t off_cursor;   // int[nkeys]: scatter cursors
    int64_t off_perm;     // int[n]: sorted slot -> source slot
    int64_t off_chunks;   // int4[max_chunks]: (first slot, count, bin, 0)
    int64_t off_meta;     // int[64]: [0] = number of chunks; [32..47] status counters
    int64_t bytes;
};

inline OrderLayout order_layout(const MpmParams& prm, int64_t n) {
    auto up = [](int64_t b) { return (b + 255) & ~int64_t(255); };
    OrderLayout L;
    L.nb = (prm.n_grid + 7) / 8;
    L.nbins = L.nb * L.nb * L.nb;
    L.nkeys = 8 * L.nbins;
    // every non-empty bin ends in at most one partial chunk
    L.chunk = chunk_size(n);
    L.max_chunks = (n + L.chunk - 1) / L.chunk + (L.nbins < n ? L.nbins : n);
    int64_t o = 0;
    L.off_hist = o;   o += up(4 * (L.nkeys + 1));
    L.off_cursor = o; o += up(4 * L.nkeys);
    L.off_perm = o;   o += up(4 * (n > 0 ? n : 1));
    L.off_chunks = o; o += up(16 * (L.max_chunks > 0 ? L.max_chunks : 1));
    L.off_meta = o;   o += 256;
    L.bytes = o;
    return L;
}

// status counters of the fused substep in the order buffer's meta block
// (int index from off_meta; +2: particles that left the grid)
constexpr int kOrderStatus = 32;

// sort key of a stencil base: bin (8^3 cells, C order) then sub-block (4^3)
__device__ __forceinline__ int order_key(const int base[3], int ng, int nb) {
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = min(max(base[d], 0), ng - 3);
    const int bin = ((b[0] >> 3) * nb + (b[1] >> 3)) * nb + (b[2] >> 3);
    const int sub = ((((b[0] >> 2) & 1) * 2 + ((b[1] >> 2) & 1)) * 2) + ((b[2] >> 2) & 1);
    return bin * 8 + sub;
}

// Origin (node) of the LDS tiles of a chunk in bin `bin` (C order, nb bins
// per axis): one node below the bin's first cell, so the tile's bases o ..
// o + 9 cover the bin's bases and one cell of drift on each side.  The
// origin may be -1; tile nodes outside the grid are neither read nor
// flushed.
__device__ __forceinline__ void chunk_origin(int bin, int nb, int o[3]) {
    o[0] = 8 * (bin / (nb * nb)) - 1;
    o[1] = 8 * ((bin / nb) % nb) - 1;
    o[2] = 8 * (bin % nb) - 1;
}

// tile index of stencil node (0,0,0) of a particle with base b in the tile at
// origin o, or -1 when its 3^3 stencil leaves the tile
__device__ __forceinline__ int chunk_tile_index(const int b[3], const int o[3]) {
    const int x = b[0] - o[0], y = b[1] - o[1], z = b[2] - o[2];
    if (x < 0 || y < 0 || z < 0 || x > kCT - 3 || y > kCT - 3 || z > kCT - 3) return -1;
    return (x * kCT + y) * kCT + z;
}

// Processing order of a chunk's particles (sorted slots first .. first+n_per,
// whose particles sit at src[idx[slot]]): s_ord[q] = chunk index of the q-th
// particle to process, in bank order of the LDS bank class t mod 32 of each
// stencil base's tile index t in the tile at origin o: particle rank r of
// class c goes to sum_c' min(count_c', r) + #{c' < c : count_c' > r}, so
// that while every class has particles each aligned group of 32 holds the 32
// classes in order -- its 4-byte gathers hit distinct banks, and each 16-lane
// half, which gfx950 serves at once for 8-byte accesses, distinct banks for
// the FP64 scatter.  Particles outside the tile or the grid take class 0.
// Scratch: s_cnt[32]; barriers inside (call from every thread of the
// workgroup).
__device__ __forceinline__ void chunk_bank_order(const float* __restrict__ planes, int64_t S,
                                                 const int* __restrict__ idx, int first, int n_per, int n,
                                                 float inv_dx, const int o[3], unsigned short* s_ord, int* s_cnt) {
    const int tid = threadIdx.x;
    if (tid < 32) s_cnt[tid] = 0;
    __syncthreads();
    int cls[kChunkMax / kChunkWG], rank[kChunkMax / kChunkWG];
#pragma unroll
    for (int r = 0; r < kChunkMax / kChunkWG; ++r) {
        const int i = tid + r * kChunkWG;
        cls[r] = rank[r] = 0;
        if (i >= n_per) continue;
        const int p = idx[first + i];
        int c = 0;
        if (p >= 0 && p < n) {
            int b[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float x = planes[(MPM_P_X0 + d) * S + p] + planes[(MPM_P_U + d) * S + p];
                b[d] = (int)(x * inv_dx - 0.5f);
            }
            const int t = chunk_tile_index(b, o);
            c = t >= 0 ? (t & 31) : 0;
        }
        cls[r] = c;
        rank[r] = atomicAdd(&s_cnt[c], 1);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kChunkMax / kChunkWG; ++r) {
        const int i = tid + r * kChunkWG;
        if (i >= n_per) continue;
        const int c = cls[r], rk = rank[r];
        int pos = 0;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int ck = s_cnt[k];
            pos += min(ck, rk) + ((k < c && ck > rk) ? 1 : 0);
        }
        s_ord[pos] = (unsigned short)i;
    }
    __syncthreads();
}

// Adds NC components of the LDS tile (SoA, kCT3 doubles each) at origin o to
// the float4 grid (components 0 .. NC-1 of each node), skipping zeros.
template <int NC>
__device__ __forceinline__ void chunk_flush(const double* s_t, const int o[3], int ng, float* __restrict__ grid) {
    for (int t = threadIdx.x; t < kCT3; t += blockDim.x) {
        const int x = o[0] + t / (kCT * kCT), y = o[1] + (t / kCT) % kCT, z = o[2] + t % kCT;
        if (x < 0 || y < 0 || z < 0 || x >= ng || y >= ng || z >= ng) continue;
        float* g = grid + 4 * (((int64_t)x * ng + y) * ng + z);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double v = s_t[c * kCT3 + t];
            if (v != 0.0) unsafeAtomicAdd(g + c, (float)v);
        }
    }
}

// Grid sets of the ordered path (mpm_step_grids_bytes): the P2G accumulator
// (n^3 float4), the grid velocities (n^3 float4), then one int per 8^3 grid
// block, nonzero where a P2G into this accumulator may have written.  The
// P2G kernels mark the blocks they touch; the grid update and the grid
// adjoint visit only marked blocks, and an accumulator is cleared block by
// block from its marks before it takes the next P2G.
inline int64_t grid_blocks(int ng) {
    const int64_t nb = (ng + 7) / 8;
    return nb * nb * nb;
}
inline int* grid_set_flags(float* set, int ng) { return (int*)(set + 8 * (int64_t)ng * ng * ng); }

// Marks the grid blocks a chunk's P2G tile (nodes 8B-1 .. 8B+10 of its bin B
// along each axis: blocks B-1 .. B+1) may write.  tid: index in the workgroup.
__device__ __forceinline__ void mark_chunk_blocks(int* __restrict__ flags, int bin, int nb, int tid) {
    if (tid < 27) {
        const int x = bin / (nb * nb) + tid / 9 - 1, y = (bin / nb) % nb + (tid / 3) % 3 - 1,
                  z = bin % nb + tid % 3 - 1;
        if (x >= 0 && y >= 0 && z >= 0 && x < nb && y < nb && z < nb) flags[(x * nb + y) * nb + z] = 1;
    }
}

// Marks the blocks of one particle's 3^3 stencil at `base` (its P2G outside
// the chunk's tile).
__device__ __forceinline__ void mark_stencil_blocks(int* __restrict__ flags, const int base[3], int nb) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int x = (base[0] + 2 * (c >> 2)) >> 3, y = (base[1] + 2 * ((c >> 1) & 1)) >> 3, z = (base[2] + 2 * (c & 1)) >> 3;
        flags[(x * nb + y) * nb + z] = 1;
    }
}

// host: the chunked P2G (p2g_apic_stress_given) of an ordered state into a
// zeroed float4 accumulator grid (mpm_ordered.hip); flags: the grid set's
// block marks, or NULL
int p2g_ordered(const MpmParams* prm, const MpmState* st, const void* order, float* grid_acc, float dt,
                void* stream, int* flags = nullptr);

// host: the fused G2P (grid velocities grid_v) + next P2G (into the zeroed
// acc_next) over the chunks; dst may be src (in place) or another state of
// the same size, which receives the planes the substep changes (u, v, C, F,
// tau) and, when copy_const, the constant ones (x0, m, vol, id).  cst: the
// planes base the constant planes are read from (NULL: src) -- the first
// state of a recorded segment (mpm_ordered.hip)
int g2p2g_ordered(const MpmParams* prm, const MpmState* src, const MpmState* dst, const void* order,
                  const float* grid_v, float* acc_next, float dt, const float* cst, bool copy_const,
                  void* stream, int* flags_next = nullptr);

// host: the grid update of the marked blocks of grid set `set` (its
// velocities from its accumulator), and the clearing of the blocks of
// `next` (a grid set, or NULL) that its earlier P2G marked, marks included
// (mpm_kernels.hip)
int grid_update_blocks(const MpmParams* prm, float* set, const MpmCollider* colliders, int n_colliders, float time,
                       float dt, float* next, void* stream);

// host: the whole accumulator and the marks of grid set `set` to zero
int clear_grid_set(const MpmParams* prm, float* set, void* stream);

}  // namespace mpm
