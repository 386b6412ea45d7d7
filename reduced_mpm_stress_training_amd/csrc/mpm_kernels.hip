// mpm_kernels.hip -- MI355X (gfx950) kernels of the explicit MLS/APIC-MPM
// sand substep and the C-ABI declared in include/mpm_hip.h.
//
// Stage kernels (one reference launch each):
//   k_zero_grid     zero_grid                      mpm_utils.py:115-123
//   k_p2g_atomic    update_x_from_u + p2g_apic_stress_given
//                                                  mpm_utils.py:4-7, 190-227
//   k_grid_update   grid_normalization_and_gravity + colliders
//                                                  mpm_utils.py:229-237,
//                                                  mpm_solver_warp.py:429-449, 505-523
//   k_g2p           g2p_apic_and_stress + update_u_from_x
//                                                  mpm_utils.py:241-306, 12-15
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "mpm_common.h"
#include "mpm_device.h"
#include "mpm_hip.h"
#include "mpm_ordered.h"

using namespace mpm;

namespace {
thread_local char g_err[512] = "";
constexpr int kBlock = 256;
inline unsigned grid_1d(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// The reference indexes the grid without bounds checks (mpm_utils.py:215-217,
// 264-266) and relies on the bounding box to keep particles inside.  Here a
// particle whose 3x3x3 stencil leaves the grid is skipped (P2G) and left
// unchanged (G2P) instead of touching memory outside the grid.
__device__ __forceinline__ bool stencil_inside(const mpm::Weights& W, int ng) {
    return W.base[0] >= 0 && W.base[1] >= 0 && W.base[2] >= 0 && W.base[0] <= ng - 3 &&
           W.base[1] <= ng - 3 && W.base[2] <= ng - 3;
}
}  // namespace

namespace mpm {

int set_err(int code, const char* msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
        return -(int)e - 1000;
    }
    return 0;
}

int check_grid(const MpmParams* prm) {
    if (!prm) return set_err(-1, "params is NULL");
    if (prm->n_grid < 4 || prm->n_grid > 1024) return set_err(-2, "n_grid out of range [4, 1024]");
    return 0;
}

int check_params(const MpmParams* prm) {
    int rc = check_grid(prm);
    if (rc) return rc;
    if (prm->material < 0 || prm->material > 2) return set_err(-3, "material must be 0 (jelly), 1 (metal) or 2 (sand)");
    if (prm->material == 1 && prm->hardening != 0 && prm->hardening != 1) return set_err(-3, "hardening must be 0 or 1");
    return 0;
}

int check_sand(const MpmParams* prm) {
    int rc = check_params(prm);
    if (rc) return rc;
    if (prm->material != 2) return set_err(-3, "this entry point implements material 2 (sand) only");
    return 0;
}

int make_collider_set(const MpmCollider* cols, int n, ColliderSet* cs) {
    if (n < 0 || n > MPM_MAX_COLLIDERS) return set_err(-8, "n_colliders out of range [0, 16]");
    if (n > 0 && !cols) return set_err(-9, "colliders is NULL");
    memset(cs, 0, sizeof(*cs));
    cs->n = n;
    for (int i = 0; i < n; ++i) cs->c[i] = cols[i];
    return 0;
}

int check_state(const MpmState* st) {
    if (!st || !st->planes) return set_err(-4, "state or planes is NULL");
    if (st->n < 0 || st->stride < st->n) return set_err(-5, "state n/stride invalid");
    if (st->n > 0x7fffffffLL) return set_err(-6, "more than 2^31-1 particles");
    return 0;
}
}  // namespace mpm

// --------------------------------------------------------------------------
// zero_grid (mpm_utils.py:115-123)
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_zero_grid(float4* __restrict__ g, int64_t nn) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nn) g[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// --------------------------------------------------------------------------
// update_x_from_u + p2g_apic_stress_given (mpm_utils.py:4-7, 190-227).
// One thread per particle, hardware float atomics into the float4 grid.
// Per node: v_in += w m (v + C dpos) + dt (-vol tau grad_w); m += w m, with
// dpos = ((i,j,k) - fx) dx.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
    k_p2g_atomic(const float* __restrict__ planes, int64_t stride, int n, MpmParams prm, float dt,
                 float* __restrict__ grid) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const float* pl = planes + p;
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = pl[(MPM_P_X0 + d) * stride] + pl[(MPM_P_U + d) * stride];
    float v[3], Cm[9], tau[9];
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = pl[(MPM_P_V + d) * stride];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        Cm[k] = pl[(MPM_P_C + k) * stride];
        tau[k] = pl[(MPM_P_TAU + k) * stride];
    }
    const float m = pl[MPM_P_MASS * stride];
    const float nvol = -pl[MPM_P_VOL * stride];
    Weights W;
    bspline_weights(x, prm.inv_dx, W);
    const int ng = prm.n_grid;
    if (!stencil_inside(W, ng)) return;
    const float dx = prm.dx, inv_dx = prm.inv_dx;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float dpos0 = ((float)i - W.fx[0]) * dx;
                float dpos1 = ((float)j - W.fx[1]) * dx;
                float dpos2 = ((float)k - W.fx[2]) * dx;
                float weight = W.w[0][i] * W.w[1][j] * W.w[2][k];
                float g0 = W.dw[0][i] * W.w[1][j] * W.w[2][k] * inv_dx;
                float g1 = W.w[0][i] * W.dw[1][j] * W.w[2][k] * inv_dx;
                float g2 = W.w[0][i] * W.w[1][j] * W.dw[2][k] * inv_dx;
                float wm = weight * m;
                float add[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float cd = Cm[a * 3 + 0] * dpos0 + Cm[a * 3 + 1] * dpos1 + Cm[a * 3 + 2] * dpos2;
                    float f = (nvol * tau[a * 3 + 0]) * g0 + (nvol * tau[a * 3 + 1]) * g1 +
                              (nvol * tau[a * 3 + 2]) * g2;
                    add[a] = wm * (v[a] + cd) + dt * f;
                }
                int64_t node = ((int64_t)(W.base[0] + i) * ng + (W.base[1] + j)) * ng + (W.base[2] + k);
                float* gn = grid + node * 4;
                unsafeAtomicAdd(gn + 0, add[0]);
                unsafeAtomicAdd(gn + 1, add[1]);
                unsafeAtomicAdd(gn + 2, add[2]);
                unsafeAtomicAdd(gn + 3, wm);
            }
}

__global__ void __launch_bounds__(kBlock)
    k_grid_update(const float4* __restrict__ acc, float4* __restrict__ gv, MpmParams prm,
                  ColliderSet cs, float time, float dt, float4* __restrict__ zero_next) {
    const int ng = prm.n_grid;
    int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int64_t nn = (int64_t)ng * ng * ng;
    if (node >= nn) return;
    int gz = (int)(node % ng);
    int gy = (int)((node / ng) % ng);
    int gx = (int)(node / ((int64_t)ng * ng));
    gv[node] = grid_node_update(acc[node], gx, gy, gz, ng, prm.dx, dt, prm.gravity, cs, time);
    // the next substep's accumulator, cleared here instead of by its own launch
    if (zero_next) zero_next[node] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The grid update on grid sets (mpm_ordered.h): one workgroup per 8^3 grid
// block, one thread per node.  A block the P2G marked gets its velocities;
// a block the next accumulator's earlier P2G marked is cleared and unmarked.
// Blocks marked in neither leave after two loads.
__global__ void __launch_bounds__(512)
    k_grid_update_blocks(const float4* __restrict__ acc, const int* __restrict__ flags, float4* __restrict__ gv,
                         MpmParams prm, ColliderSet cs, float time, float dt, float4* __restrict__ next,
                         int* __restrict__ next_flags) {
    const int ng = prm.n_grid, nb = (ng + 7) / 8;
    const int blk = blockIdx.x;
    const int mark = flags[blk];
    const int dirty = next_flags ? next_flags[blk] : 0;
    if (!mark && !dirty) return;
    const int t = threadIdx.x;
    const int bz = blk % nb, by = (blk / nb) % nb, bx = blk / (nb * nb);
    const int gx = bx * 8 + (t >> 6), gy = by * 8 + ((t >> 3) & 7), gz = bz * 8 + (t & 7);
    if (gx < ng && gy < ng && gz < ng) {
        const int64_t node = ((int64_t)gx * ng + gy) * ng + gz;
        if (mark) gv[node] = grid_node_update(acc[node], gx, gy, gz, ng, prm.dx, dt, prm.gravity, cs, time);
        if (dirty) next[node] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (dirty) {
        __syncthreads();   // every thread has read the mark
        if (t == 0) next_flags[blk] = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_zero_ints(int* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0;
}

// --------------------------------------------------------------------------
// g2p_apic_and_stress (mpm_utils.py:241-306) + update_x_from_u / _u_from_x.
// dpos = (i,j,k) - fx in grid units; C += v_i (x) dpos * (w 4 inv_dx);
// grad v += v_i (x) grad w; x += dt v; F_tr = (I + dt grad v) F; the
// material's return map and stress (constitutive_update).
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
    k_g2p(float* __restrict__ planes, int64_t stride, int n, MpmParams prm, float dt,
          const float4* __restrict__ gv) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    float* pl = planes + p;
    float x0[3], x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x0[d] = pl[(MPM_P_X0 + d) * stride];
        x[d] = x0[d] + pl[(MPM_P_U + d) * stride];
    }
    Weights W;
    bspline_weights(x, prm.inv_dx, W);
    const int ng = prm.n_grid;
    if (!stencil_inside(W, ng)) return;
    const float inv_dx = prm.inv_dx;
    float nv[3] = {0.f, 0.f, 0.f}, nC[9], gradv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) { nC[k] = 0.f; gradv[k] = 0.f; }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                int64_t node = ((int64_t)(W.base[0] + i) * ng + (W.base[1] + j)) * ng + (W.base[2] + k);
                float4 g = gv[node];
                float gvv[3] = {g.x, g.y, g.z};
                float dpos[3] = {(float)i - W.fx[0], (float)j - W.fx[1], (float)k - W.fx[2]};
                float weight = W.w[0][i] * W.w[1][j] * W.w[2][k];
                float sC = weight * inv_dx * 4.0f;
                float gw[3] = {W.dw[0][i] * W.w[1][j] * W.w[2][k] * inv_dx,
                               W.w[0][i] * W.dw[1][j] * W.w[2][k] * inv_dx,
                               W.w[0][i] * W.w[1][j] * W.dw[2][k] * inv_dx};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    nv[a] += gvv[a] * weight;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        nC[a * 3 + b] += (gvv[a] * dpos[b]) * sC;
                        gradv[a * 3 + b] += gvv[a] * gw[b];
                    }
                }
            }
    float F[9], tau[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = pl[(MPM_P_F + k) * stride];
    float ys = prm.material == 1 ? pl[MPM_P_YIELD * stride] : 0.0f;
    constitutive_update(gradv, dt, F, prm.material, prm.mu, prm.lam, prm.alpha, prm.hardening, prm.xi, ys, tau);
    if (prm.material == 1) pl[MPM_P_YIELD * stride] = ys;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float xn = x[d] + dt * nv[d];
        pl[(MPM_P_U + d) * stride] = xn - x0[d];
        pl[(MPM_P_V + d) * stride] = nv[d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        pl[(MPM_P_C + k) * stride] = nC[k];
        pl[(MPM_P_F + k) * stride] = F[k];
        pl[(MPM_P_TAU + k) * stride] = tau[k];
    }
}

// --------------------------------------------------------------------------
// import / export between caller-order (n, ncomp) arrays and solver planes
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
    k_import(float* __restrict__ planes, int64_t stride, int n, int first, int ncomp,
             const float* __restrict__ src, int mode) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int pid = __float_as_int(planes[MPM_P_PID * stride + s]);
    for (int c = 0; c < ncomp; ++c) {
        float val = src[(int64_t)pid * ncomp + c];
        if (mode == 2 && (c % 4) == 0) val += 1.0f;
        planes[(int64_t)(first + c) * stride + s] = val;
    }
}

__global__ void __launch_bounds__(kBlock)
    k_export(const float* __restrict__ planes, int64_t stride, int n, int first, int ncomp,
             float* __restrict__ dst, int mode) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int pid = __float_as_int(planes[MPM_P_PID * stride + s]);
    for (int c = 0; c < ncomp; ++c) {
        float val = planes[(int64_t)(first + c) * stride + s];
        if (mode == 1 && (c % 4) == 0) val -= 1.0f;
        if (mode == 3) val = planes[(int64_t)(MPM_P_X0 + c) * stride + s] + planes[(int64_t)(MPM_P_U + c) * stride + s];
        dst[(int64_t)pid * ncomp + c] = val;
    }
}

// ==========================================================================
// C-ABI
// ==========================================================================
namespace mpm {

int grid_update_zero_next(const MpmParams* prm, const float* grid_acc, float* grid_v, const MpmCollider* colliders,
                          int n_colliders, float time, float dt, float* zero_next, void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if (!grid_acc || !grid_v) return set_err(-7, "grid is NULL");
    if (zero_next && (zero_next == grid_acc || zero_next == grid_v)) return set_err(-7, "grid aliases the next one");
    ColliderSet cs;
    if ((rc = make_collider_set(colliders, n_colliders, &cs))) return rc;
    int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    hipLaunchKernelGGL(k_grid_update, dim3(grid_1d(nn)), dim3(kBlock), 0, (hipStream_t)stream,
                       (const float4*)grid_acc, (float4*)grid_v, *prm, cs, time, dt, (float4*)zero_next);
    return check_launch("k_grid_update");
}

int grid_update_blocks(const MpmParams* prm, float* set, const MpmCollider* colliders, int n_colliders, float time,
                       float dt, float* next, void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if (!set) return set_err(-7, "grid set is NULL");
    if (next == set) return set_err(-7, "grid set aliases the next one");
    ColliderSet cs;
    if ((rc = make_collider_set(colliders, n_colliders, &cs))) return rc;
    const int ng = prm->n_grid;
    const int64_t nn = (int64_t)ng * ng * ng;
    hipLaunchKernelGGL(k_grid_update_blocks, dim3((unsigned)grid_blocks(ng)), dim3(512), 0, (hipStream_t)stream,
                       (const float4*)set, (const int*)grid_set_flags(set, ng), (float4*)(set + 4 * nn), *prm, cs, time,
                       dt, (float4*)next, next ? grid_set_flags(next, ng) : (int*)nullptr);
    return check_launch("k_grid_update_blocks");
}

int clear_grid_set(const MpmParams* prm, float* set, void* stream) {
    int rc = mpm_zero_grid(prm, set, stream);
    if (rc) return rc;
    const int64_t nbk = grid_blocks(prm->n_grid);
    hipLaunchKernelGGL(k_zero_ints, dim3(grid_1d(nbk)), dim3(kBlock), 0, (hipStream_t)stream,
                       grid_set_flags(set, prm->n_grid), nbk);
    return check_launch("k_zero_ints");
}

}  // namespace mpm

extern "C" {

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

const char* mpm_last_error(void) { return g_err; }

int mpm_zero_grid(const MpmParams* prm, float* grid, void* stream) {
    int rc = check_grid(prm);
    if (rc) return rc;
    if (!grid) return set_err(-7, "grid is NULL");
    int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    hipLaunchKernelGGL(k_zero_grid, dim3(grid_1d(nn)), dim3(kBlock), 0, (hipStream_t)stream,
                       (float4*)grid, nn);
    return check_launch("k_zero_grid");
}

int mpm_p2g(const MpmParams* prm, const MpmState* st, float* grid_acc, float dt, void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if ((rc = check_state(st))) return rc;
    if (!grid_acc) return set_err(-7, "grid is NULL");
    if (st->n == 0) return 0;
    hipLaunchKernelGGL(k_p2g_atomic, dim3(grid_1d(st->n)), dim3(kBlock), 0, (hipStream_t)stream,
                       st->planes, st->stride, (int)st->n, *prm, dt, grid_acc);
    return check_launch("k_p2g_atomic");
}


int mpm_grid_update(const MpmParams* prm, const float* grid_acc, float* grid_v,
                    const MpmCollider* colliders, int n_colliders, float time, float dt,
                    void* stream) {
    return mpm::grid_update_zero_next(prm, grid_acc, grid_v, colliders, n_colliders, time, dt, nullptr, stream);
}

int mpm_g2p(const MpmParams* prm, const MpmState* st, const float* grid_v, float dt,
            void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if ((rc = check_state(st))) return rc;
    if (!grid_v) return set_err(-7, "grid is NULL");
    if (st->n == 0) return 0;
    hipLaunchKernelGGL(k_g2p, dim3(grid_1d(st->n)), dim3(kBlock), 0, (hipStream_t)stream,
                       st->planes, st->stride, (int)st->n, *prm, dt, (const float4*)grid_v);
    return check_launch("k_g2p");
}

int64_t mpm_workspace_bytes(const MpmParams* prm, int64_t n_particles) {
    if (check_grid(prm)) return -1;
    (void)n_particles;
    int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    return 2 * nn * 16;  // accumulator + velocity grid
}

int64_t mpm_step_grids_bytes(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    const int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    return 2 * nn * 16 + ((grid_blocks(prm->n_grid) * 4 + 255) / 256) * 256;   // accumulator, velocities, marks
}

int mpm_step(const MpmParams* prm, const MpmState* st, const MpmCollider* colliders,
             int n_colliders, void* workspace, int64_t workspace_bytes, float time, float dt,
             void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if ((rc = check_state(st))) return rc;
    int64_t need = mpm_workspace_bytes(prm, st->n);
    if (!workspace || workspace_bytes < need) return set_err(-10, "workspace too small");
    if (((uintptr_t)workspace) & 255) return set_err(-11, "workspace not 256-byte aligned");
    int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    float* acc = (float*)workspace;
    float* gv = acc + 4 * nn;
    if ((rc = mpm_zero_grid(prm, acc, stream))) return rc;
    if ((rc = mpm_p2g(prm, st, acc, dt, stream))) return rc;
    if ((rc = mpm_grid_update(prm, acc, gv, colliders, n_colliders, time, dt, stream))) return rc;
    return mpm_g2p(prm, st, gv, dt, stream);
}

int mpm_import(const MpmState* st, int first_plane, int ncomp, const float* src, int mode,
               void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (first_plane < 0 || ncomp < 1 || first_plane + ncomp > MPM_P_PID)
        return set_err(-12, "plane range invalid");
    if (!src) return set_err(-13, "src is NULL");
    if (st->n == 0) return 0;
    hipLaunchKernelGGL(k_import, dim3(grid_1d(st->n)), dim3(kBlock), 0, (hipStream_t)stream,
                       st->planes, st->stride, (int)st->n, first_plane, ncomp, src, mode);
    return check_launch("k_import");
}

int mpm_export(const MpmState* st, int first_plane, int ncomp, float* dst, int mode,
               void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (first_plane < 0 || ncomp < 1 || first_plane + ncomp > MPM_P_PID)
        return set_err(-12, "plane range invalid");
    if (!dst) return set_err(-13, "dst is NULL");
    if (st->n == 0) return 0;
    hipLaunchKernelGGL(k_export, dim3(grid_1d(st->n)), dim3(kBlock), 0, (hipStream_t)stream,
                       st->planes, st->stride, (int)st->n, first_plane, ncomp, dst, mode);
    return check_launch("k_export");
}

}  // extern "C"
