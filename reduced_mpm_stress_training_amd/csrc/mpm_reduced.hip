// mpm_reduced.hip -- the reduced-order sampled substep on the flat state:
// MPM_Simulator_WARP.p2g2p_reduced (mpm_solver_warp.py:213-227) and the
// sample G2P its caller launches after it, the time stepper a CROM decoder
// drives (only sample particles are advanced; the decoder fills in the rest).
//
//   k_mark_cells        update_neighboring_cells            mpm_utils.py:318-335
//                       (and ..._for_grid_v, :337-354, with the range -2..4)
//   k_p2g_sample_all    p2g_sample_all                      mpm_utils.py:356-410
//   k_grid_sample       zero_grid + grid_normalization_and_gravity_sample_all
//                       + the colliders                      mpm_utils.py:115-123,
//                                                            413-421;
//                                                            mpm_solver_warp.py:222-225
//   k_g2p_sample        g2p_sample                          mpm_utils.py:423-471
//
// Workspace (mpm_reduced_workspace_bytes): accumulator grid (mv, m) and
// velocity grid (v, m) as float4 per node, then the int32 flag grids
// neighboring_cells and neighboring_cells_grid_v (n_grid^3 each).
//
// Unlike the full substep the scatter here is sparse (only particles in
// cells next to a sample contribute, a few percent of them for a CROM
// sample), so it uses one thread per particle and hardware float atomics
// straight into the global grid, as the stage P2G does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_common.h"
#include "mpm_device.h"
#include "mpm_hip.h"

using namespace mpm;

namespace {
constexpr int kBlock = 256;
inline unsigned grid_1d(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct ReducedWs {
    float4* acc;
    float4* vel;
    int* cells;
    int* cells_v;
};

ReducedWs reduced_ws(const MpmParams& prm, void* ws) {
    const int64_t nn = (int64_t)prm.n_grid * prm.n_grid * prm.n_grid;
    char* p = (char*)ws;
    ReducedWs w;
    w.acc = (float4*)p;
    w.vel = (float4*)(p + nn * 16);
    w.cells = (int*)(p + 2 * nn * 16);
    w.cells_v = (int*)(p + 2 * nn * 16 + nn * 4);
    return w;
}

__device__ __forceinline__ void particle_x(const float* __restrict__ planes, int64_t stride, int64_t p, float x[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = planes[(MPM_P_X0 + d) * stride + p] + planes[(MPM_P_U + d) * stride + p];
}

// zero_grid (mpm_utils.py:115-123): both grids and both flag grids
__global__ void __launch_bounds__(kBlock) k_zero_reduced(ReducedWs w, int64_t nn) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nn) return;
    w.acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    w.vel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    w.cells[i] = 0;
    w.cells_v[i] = 0;
}

// update_neighboring_cells (mpm_utils.py:318-335; lo = -2, hi = 2) and
// update_neighboring_cells_for_grid_v (:337-354; lo = -2, hi = 4): every
// cell base + (i, j, k), i, j, k in lo..hi, inside the grid is flagged.
// The flag is a plain store of 1, so concurrent writers agree.
__global__ void __launch_bounds__(kBlock)
    k_mark_cells(const float* __restrict__ planes, int64_t stride, int64_t n, const int* __restrict__ samples,
                 int n_samples, MpmParams prm, int lo, int hi, int* __restrict__ flags) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n_samples) return;
    const int64_t p = samples[s];
    if (p < 0 || p >= n) return;
    float x[3];
    particle_x(planes, stride, p, x);
    const int ng = prm.n_grid;
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = (int)(x[d] * prm.inv_dx - 0.5f);
    for (int i = lo; i <= hi; ++i) {
        const int gx = b[0] + i;
        if (gx < 0 || gx >= ng) continue;
        for (int j = lo; j <= hi; ++j) {
            const int gy = b[1] + j;
            if (gy < 0 || gy >= ng) continue;
            for (int k = lo; k <= hi; ++k) {
                const int gz = b[2] + k;
                if (gz < 0 || gz >= ng) continue;
                flags[((int64_t)gx * ng + gy) * ng + gz] = 1;
            }
        }
    }
}

// p2g_sample_all (mpm_utils.py:356-410): a particle whose base cell is
// flagged sets F = F_disp + I, forms its stress from the SVD of F (FCR with
// J = s0 s1 s2 for jelly and metal, Drucker-Prager for sand; not stored),
// and scatters the PIC momentum w m v (no affine C term), the force
// dt (-(m / rho) tau grad w) and the mass w m.  F_disp is F - I in this
// state, so F_disp + I is the round trip (F - I) + I the reference's arrays
// hold.
__global__ void __launch_bounds__(kBlock)
    k_p2g_sample_all(float* __restrict__ planes, int64_t stride, int n, const float* __restrict__ density,
                     MpmParams prm, float dt, const int* __restrict__ cells, float* __restrict__ grid) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    float x[3];
    particle_x(planes, stride, p, x);
    Weights W;
    bspline_weights(x, prm.inv_dx, W);
    const int ng = prm.n_grid;
    if (W.base[0] < 0 || W.base[1] < 0 || W.base[2] < 0 || W.base[0] >= ng || W.base[1] >= ng || W.base[2] >= ng)
        return;
    if (cells[((int64_t)W.base[0] * ng + W.base[1]) * ng + W.base[2]] <= 0) return;
    // the stencil must stay inside the grid (the reference has no check)
    if (W.base[0] > ng - 3 || W.base[1] > ng - 3 || W.base[2] > ng - 3) return;
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float fd = planes[(MPM_P_F + k) * stride + p] - ((k % 4) == 0 ? 1.0f : 0.0f);
        F[k] = fd + ((k % 4) == 0 ? 1.0f : 0.0f);
        planes[(MPM_P_F + k) * stride + p] = F[k];
    }
    float U[9], s[3], V[9], tau[9];
    svd3_rot(F, U, s, V);
    if (prm.material == 2) {
        float ls[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ls[d] = logf(s[d]);
        dp_stress_sym(U, ls, prm.mu, prm.lam, tau);
    } else {
        fcr_stress(F, U, V, s[0] * s[1] * s[2], prm.mu, prm.lam, tau);
    }
    float v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = planes[(MPM_P_V + d) * stride + p];
    const float m = planes[MPM_P_MASS * stride + p];
    const float nvol = -(m / density[p]);
    const float inv_dx = prm.inv_dx;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float weight = W.w[0][i] * W.w[1][j] * W.w[2][k];
                const float g0 = W.dw[0][i] * W.w[1][j] * W.w[2][k] * inv_dx;
                const float g1 = W.w[0][i] * W.dw[1][j] * W.w[2][k] * inv_dx;
                const float g2 = W.w[0][i] * W.w[1][j] * W.dw[2][k] * inv_dx;
                const float wm = weight * m;
                float* gn = grid + 4 * (((int64_t)(W.base[0] + i) * ng + (W.base[1] + j)) * ng + (W.base[2] + k));
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float f = (nvol * tau[a * 3 + 0]) * g0 + (nvol * tau[a * 3 + 1]) * g1 +
                                    (nvol * tau[a * 3 + 2]) * g2;
                    unsafeAtomicAdd(gn + a, wm * v[a]);
                    unsafeAtomicAdd(gn + a, dt * f);
                }
                unsafeAtomicAdd(gn + 3, wm);
            }
}

// grid_normalization_and_gravity_sample_all (mpm_utils.py:413-421) on the
// flagged nodes (m > 1e-10, v = mv / m + dt g; every other node keeps the
// zero of zero_grid), then the colliders in registration order on every
// node (mpm_solver_warp.py:222-225).
__global__ void __launch_bounds__(kBlock)
    k_grid_sample(const float4* __restrict__ acc, float4* __restrict__ vel, const int* __restrict__ cells,
                  MpmParams prm, ColliderSet cs, float time, float dt) {
    const int ng = prm.n_grid;
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)ng * ng * ng) return;
    const int gz = (int)(node % ng), gy = (int)((node / ng) % ng), gx = (int)(node / ((int64_t)ng * ng));
    const float4 a = acc[node];
    float o[3] = {0.f, 0.f, 0.f};
    if (cells[node] > 0 && a.w > 1e-10f) {
        o[0] = a.x / a.w + dt * prm.gravity[0];
        o[1] = a.y / a.w + dt * prm.gravity[1];
        o[2] = a.z / a.w + dt * prm.gravity[2];
    }
    float mask[3];
    for (int c = 0; c < cs.n; ++c) apply_collider(cs.c[c], gx, gy, gz, ng, prm.dx, dt, time, o, mask);
    vel[node] = make_float4(o[0], o[1], o[2], a.w);
}

// g2p_sample (mpm_utils.py:423-471) for sample particle q = samples[s]:
// v = sum w v_i (PIC), grad v = sum v_i (x) grad w, x += dt v,
// F_trial = (I + dt grad v) F, the material's return map (no stress), then
// u = x - x0 (F_disp = F - I is F itself here).  Metal reads and hardens the
// yield stress of particle s -- the sample's slot in the list, as the
// reference passes its thread index p to von_mises_return_mapping (:464).
__global__ void __launch_bounds__(kBlock)
    k_g2p_sample(float* __restrict__ planes, int64_t stride, int64_t n, const int* __restrict__ samples,
                 int n_samples, MpmParams prm, float dt, const float4* __restrict__ vel) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n_samples) return;
    const int64_t q = samples[s];
    if (q < 0 || q >= n) return;
    float x0[3], x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x0[d] = planes[(MPM_P_X0 + d) * stride + q];
        x[d] = x0[d] + planes[(MPM_P_U + d) * stride + q];
    }
    Weights W;
    bspline_weights(x, prm.inv_dx, W);
    const int ng = prm.n_grid;
    if (W.base[0] < 0 || W.base[1] < 0 || W.base[2] < 0 || W.base[0] > ng - 3 || W.base[1] > ng - 3 ||
        W.base[2] > ng - 3)
        return;
    const float inv_dx = prm.inv_dx;
    float nv[3] = {0.f, 0.f, 0.f}, G[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) G[k] = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float4 g = vel[((int64_t)(W.base[0] + i) * ng + (W.base[1] + j)) * ng + (W.base[2] + k)];
                const float gv[3] = {g.x, g.y, g.z};
                const float weight = W.w[0][i] * W.w[1][j] * W.w[2][k];
                const float gw[3] = {W.dw[0][i] * W.w[1][j] * W.w[2][k] * inv_dx,
                                     W.w[0][i] * W.dw[1][j] * W.w[2][k] * inv_dx,
                                     W.w[0][i] * W.w[1][j] * W.dw[2][k] * inv_dx};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    nv[a] += gv[a] * weight;
#pragma unroll
                    for (int b = 0; b < 3; ++b) G[a * 3 + b] += gv[a] * gw[b];
                }
            }
    float F[9], A[9], Ftr[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        F[k] = planes[(MPM_P_F + k) * stride + q];
        A[k] = dt * G[k] + ((k % 4) == 0 ? 1.0f : 0.0f);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Ftr[i * 3 + j] = A[i * 3 + 0] * F[0 * 3 + j] + A[i * 3 + 1] * F[1 * 3 + j] + A[i * 3 + 2] * F[2 * 3 + j];
    if (prm.material == 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = Ftr[k];
    } else {
        float U[9], sv[3], V[9], ls[3];
        svd3_rot(Ftr, U, sv, V);
        if (prm.material == 2) {
            sand_return_map(Ftr, U, sv, V, prm.mu, prm.lam, prm.alpha, F, ls);
        } else {
            float ys = planes[MPM_P_YIELD * stride + s];
            von_mises_return_map(Ftr, U, sv, V, prm.mu, prm.lam, ys, prm.hardening, prm.xi, F, ls);
            if (prm.hardening == 1) planes[MPM_P_YIELD * stride + s] = ys;
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        planes[(MPM_P_V + d) * stride + q] = nv[d];
        planes[(MPM_P_U + d) * stride + q] = (x[d] + dt * nv[d]) - x0[d];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) planes[(MPM_P_F + k) * stride + q] = F[k];
}

int check_samples(const MpmState* st, const int* samples, int64_t n_samples) {
    if (n_samples < 0 || n_samples > 0x7fffffffLL) return set_err(-50, "n_samples out of range");
    if (n_samples > 0 && !samples) return set_err(-51, "samples is NULL");
    (void)st;
    return 0;
}
}  // namespace

extern "C" {

int64_t mpm_reduced_workspace_bytes(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    const int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    return 2 * nn * 16 + 2 * nn * 4;
}

int mpm_step_reduced(const MpmParams* prm, const MpmState* st, const int* samples, int64_t n_samples,
                     const float* density, const MpmCollider* colliders, int n_colliders, void* workspace,
                     int64_t workspace_bytes, float time, float dt, void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if ((rc = check_state(st))) return rc;
    if ((rc = check_samples(st, samples, n_samples))) return rc;
    if (!density && st->n > 0) return set_err(-52, "density is NULL");
    if (!workspace || ((uintptr_t)workspace & 255)) return set_err(-11, "workspace NULL or not 256-byte aligned");
    if (workspace_bytes < mpm_reduced_workspace_bytes(prm)) return set_err(-10, "workspace too small");
    ColliderSet cs;
    if ((rc = make_collider_set(colliders, n_colliders, &cs))) return rc;
    const ReducedWs w = reduced_ws(*prm, workspace);
    const int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_reduced, dim3(grid_1d(nn)), dim3(kBlock), 0, s, w, nn);
    if ((rc = check_launch("k_zero_reduced"))) return rc;
    if (n_samples > 0) {
        hipLaunchKernelGGL(k_mark_cells, dim3(grid_1d(n_samples)), dim3(kBlock), 0, s, st->planes, st->stride, st->n,
                           samples, (int)n_samples, *prm, -2, 2, w.cells);
        if ((rc = check_launch("k_mark_cells"))) return rc;
    }
    if (st->n > 0) {
        hipLaunchKernelGGL(k_p2g_sample_all, dim3(grid_1d(st->n)), dim3(kBlock), 0, s, st->planes, st->stride,
                           (int)st->n, density, *prm, dt, (const int*)w.cells, (float*)w.acc);
        if ((rc = check_launch("k_p2g_sample_all"))) return rc;
    }
    hipLaunchKernelGGL(k_grid_sample, dim3(grid_1d(nn)), dim3(kBlock), 0, s, (const float4*)w.acc, w.vel,
                       (const int*)w.cells, *prm, cs, time, dt);
    return check_launch("k_grid_sample");
}

int mpm_g2p_sample(const MpmParams* prm, const MpmState* st, const int* samples, int64_t n_samples,
                   const void* workspace, float dt, void* stream) {
    int rc = check_params(prm);
    if (rc) return rc;
    if ((rc = check_state(st))) return rc;
    if ((rc = check_samples(st, samples, n_samples))) return rc;
    if (!workspace || ((uintptr_t)workspace & 255)) return set_err(-11, "workspace NULL or not 256-byte aligned");
    if (prm->material == 1 && n_samples > st->n)
        return set_err(-53, "metal: the yield stress of sample slot s is read, so n_samples must not exceed n");
    if (n_samples == 0) return 0;
    const ReducedWs w = reduced_ws(*prm, (void*)workspace);
    hipLaunchKernelGGL(k_g2p_sample, dim3(grid_1d(n_samples)), dim3(kBlock), 0, (hipStream_t)stream, st->planes,
                       st->stride, st->n, samples, (int)n_samples, *prm, dt, (const float4*)w.vel);
    return check_launch("k_g2p_sample");
}

int mpm_mark_cells_grid_v(const MpmParams* prm, const MpmState* st, const int* samples, int64_t n_samples,
                          void* workspace, void* stream) {
    int rc = check_grid(prm);
    if (rc) return rc;
    if ((rc = check_state(st))) return rc;
    if ((rc = check_samples(st, samples, n_samples))) return rc;
    if (!workspace || ((uintptr_t)workspace & 255)) return set_err(-11, "workspace NULL or not 256-byte aligned");
    if (n_samples == 0) return 0;
    const ReducedWs w = reduced_ws(*prm, workspace);
    hipLaunchKernelGGL(k_mark_cells, dim3(grid_1d(n_samples)), dim3(kBlock), 0, (hipStream_t)stream, st->planes,
                       st->stride, st->n, samples, (int)n_samples, *prm, -2, 4, w.cells_v);
    return check_launch("k_mark_cells");
}

}  // extern "C"
