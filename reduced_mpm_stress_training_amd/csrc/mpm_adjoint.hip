// mpm_adjoint.hip -- reverse-mode derivative (adjoint) of one explicit
// MLS/APIC-MPM sand substep (MPM_Simulator_WARP.p2g2p,
// mpm_solver_warp.py:156-203) on the flat particle state, and its C-ABI
// (mpm_adjoint_* in include/mpm_hip.h).
//
// The reference never differentiates its solver (no wp.Tape is recorded);
// this is the derivative of the same forward arithmetic.  With the state
// S_n = (x, v, C, F, tau) of every particle, one substep is
//   P2G     m_i = sum_p w m_p,  mv_i = sum_p w m_p (v_p + C_p (x_i - x_p))
//                                      - dt vol_p tau_p grad w   (mpm_utils.py:190-227)
//   grid    v_i = BC(mv_i / m_i + dt g)                         (:229-237, colliders)
//   G2P     v' = sum w v_i,  C' = 4 inv_dx^2 sum w v_i (x) (x_i - x_p),
//           G = sum v_i (x) grad w,  x' = x + dt v',
//           F' = RM((I + dt G) F),  tau' = stress(F')           (:241-306)
// and its adjoint runs the three stages backwards:
//   k_g2p_adj_chunks  per particle: stress and return-map adjoints through
//               the SVD (closed forms of the derivative of isotropic matrix
//               functions, divided differences where singular values meet),
//               the position / F adjoints of the weights and F_tr, and the
//               grid-velocity adjoint, scattered per chunk of the ordered
//               state into an FP64 LDS tile (mpm_ordered.h);
//   k_grid_adj  per node: through the collider masks and the normalization;
//   k_p2g_adj   per particle: gathers v, C, tau and position adjoints.
// The weights' base cell is piecewise constant and the return-map branch is
// held fixed (derivative of the branch taken).  Material parameters collect
// their adjoints (mu, lam, alpha) in double atomics.
#include <hip/hip_runtime.h>
#include <math.h>

#include "mpm_common.h"
#include "mpm_device.h"
#include "mpm_hip.h"
#include "mpm_ordered.h"
#include "mpm_stencil.h"

using namespace mpm;

namespace {

constexpr int kBlk = 256;

// adjoint planes: x (3), v (3), C (9), F (9), tau (9)
constexpr int kAX = 0, kAV = 3, kAC = 6, kAF = 15, kAT = 24;

// The adjoint's own quotients (not those of the recomputed forward return
// map, whose branch must match the forward's) multiply by the hardware
// reciprocal (1 ulp) instead of the correctly rounded division sequence
// (~11 instructions; ~45 of them per particle).
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Divided differences of the logarithm and of a power, robust when the
// arguments meet, from the arguments s_k, s_l > 0 and their logarithms
// lk = ln s_k, ll = ln s_l (which the return map already has: no further
// logarithm or power is evaluated).
// (ln s_l - ln s_k) / (s_l - s_k), rsk = 1 / s_k:
__device__ __forceinline__ float dlog_div(float sk, float sl, float lk, float ll, float rsk) {
    const float r = sl * rsk - 1.0f;
    if (fabsf(r) < 1e-4f) return (1.0f - 0.5f * r + r * r * (1.0f / 3.0f)) * rsk;
    return (ll - lk) * rcp(sl - sk);
}

// (s_l^B - s_k^B) / (s_l - s_k) = s_k^(B-1) expm1(B d) / expm1(d), d = ll - lk:
__device__ __forceinline__ float dpow_div(float lk, float ll, float B) {
    const float d = ll - lk;
    const float pk = expf((B - 1.0f) * lk);
    if (fabsf(d) < 1e-5f) return B * pk * (1.0f + 0.5f * (B - 1.0f) * d);
    return pk * expm1f(B * d) * rcp(expm1f(d));
}

__device__ __forceinline__ void matmul(const float A[9], const float B[9], float C[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}
__device__ __forceinline__ void matmul_tn(const float A[9], const float B[9], float C[9]) {  // A^T B
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
__device__ __forceinline__ void matmul_nt(const float A[9], const float B[9], float C[9]) {  // A B^T
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i * 3 + j] = A[i * 3] * B[j * 3] + A[i * 3 + 1] * B[j * 3 + 1] + A[i * 3 + 2] * B[j * 3 + 2];
}

// Adjoint of  F_tr = (I + dt G) F,  F' = RM(F_tr),  tau' = stress(F')
// (sand_update).  In: G, F, dt, parameters, the adjoints Fb' of F' and Tb of
// tau'.  Out: Gb, Fb (adjoint of the input F), and the parameters' adjoints.
__device__ void sand_update_adj(const float G[9], const float F[9], float dt, float mu, float lam, float alpha,
                                const float Fbn[9], const float Tb[9], float Gb[9], float Fb[9], float& mub,
                                float& lamb, float& alb) {
    float A[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = dt * G[k];
    A[0] += 1.f; A[4] += 1.f; A[8] += 1.f;
    float Ftr[9];
    matmul(A, F, Ftr);
    float U[9], s[3], V[9];
    svd3_rot(Ftr, U, s, V);
    // forward return map quantities (sand_return_map)
    float eps[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) eps[d] = logf(fmaxf(fabsf(s[d]), 1e-14f));
    const float tr = eps[0] + eps[1] + eps[2];
    float eh[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) eh[d] = eps[d] - tr / 3.0f;
    const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    const float kap = (3.0f * lam + 2.0f * mu) / (2.0f * mu);
    const float dg = ehn + kap * tr * alpha;
    int br = 1;
    float se[3], ls[3];
    float beta = 0.f;
    if (dg <= 0.0f) {
#pragma unroll
        for (int d = 0; d < 3; ++d) { se[d] = s[d]; ls[d] = (s[d] >= 1e-14f) ? eps[d] : logf(s[d]); }
    } else if (tr > 0.0f) {
        br = 2;
#pragma unroll
        for (int d = 0; d < 3; ++d) { se[d] = 1.f; ls[d] = 0.f; }
    } else {
        br = 3;
        beta = kap * alpha * tr * rcp(ehn);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ls[d] = eps[d] - eh[d] * (dg / ehn);
            se[d] = expf(ls[d]);
        }
    }
    // stress adjoint in the principal frame of F' = U diag(se) V^T
    float T1[9], Tp[9];
    matmul_tn(U, Tb, T1);
    matmul(T1, U, Tp);          // U^T Tb U
    const float lsum = ls[0] + ls[1] + ls[2];
    const float tsum = Tp[0] + Tp[4] + Tp[8];
    mub += 2.0f * (Tp[0] * ls[0] + Tp[4] * ls[1] + Tp[8] * ls[2]);
    lamb += tsum * lsum;
    float Q[9];
    {
        float Q1[9];
        matmul_tn(U, Fbn, Q1);
        matmul(Q1, V, Q);       // U^T Fb' V
    }
    float rse[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) rse[k] = rcp(se[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) Q[k * 4] += (2.0f * mu * Tp[k * 4] + lam * tsum) * rse[k];
    // off-diagonal terms: the divided differences are symmetric in (k, l),
    // so each is evaluated once per unordered pair
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
        const int k = pr == 2 ? 1 : 0, l = pr == 0 ? 1 : 2;
        const float c = 2.0f * mu * dlog_div(se[k], se[l], ls[k], ls[l], rse[k]) * rcp(se[k] + se[l]);
        const float tt = Tp[k * 3 + l] + Tp[l * 3 + k];
        Q[k * 3 + l] += c * se[l] * tt;
        Q[l * 3 + k] += c * se[k] * tt;
    }
    // return-map adjoint: P = principal-frame adjoint of F_tr
    float P[9];
    if (br == 1) {
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = Q[k];
    } else {
        float dsd[3][3];   // d se_i / d sigma_k
        if (br == 2) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) dsd[i][k] = 0.f;
        } else {
            // H_i = -beta eps_i + (1 + beta) tr/3, beta = kap alpha tr / |eh|
            const float rehn = rcp(ehn);
            float dbeta[3], rs[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dbeta[k] = kap * alpha * (rehn - tr * eh[k] * (rehn * rehn * rehn));
                rs[k] = rcp(s[k]);
            }
            const float b3 = (1.0f + beta) * (1.0f / 3.0f);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float dH = (i == k ? -beta : 0.f) + b3 - eh[i] * dbeta[k];
                    dsd[i][k] = se[i] * dH * rs[k];
                }
            // parameters: dH_i/dalpha = -eh_i kap tr/|eh|, dH_i/dkap = -eh_i alpha tr/|eh|
            float qa = 0.f, qk = 0.f;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                qa += Q[i * 4] * se[i] * (-eh[i] * kap * tr * rehn);
                qk += Q[i * 4] * se[i] * (-eh[i] * alpha * tr * rehn);
            }
            alb += qa;
            // kap = 1 + 3 lam / (2 mu)
            const float r2mu = rcp(2.0f * mu);
            lamb += qk * 3.0f * r2mu;
            mub += qk * (-3.0f * lam * r2mu * rcp(mu));
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float v = 0.f;
#pragma unroll
            for (int i = 0; i < 3; ++i) v += Q[i * 4] * dsd[i][k];
            P[k * 4] = v;
        }
        const float Bexp = -beta;
        const float eA = br == 3 ? expf((1.0f + beta) * tr / 3.0f) : 0.f;
#pragma unroll
        for (int pr = 0; pr < 3; ++pr) {   // symmetric in (k, l): once per unordered pair
            const int k = pr == 2 ? 1 : 0, l = pr == 0 ? 1 : 2;
            const float D = br == 3 ? eA * dpow_div(eps[k], eps[l], Bexp) : 0.f;
            const float S = (se[l] + se[k]) * rcp(s[l] + s[k]);
            const float a = 0.5f * (D + S), b = 0.5f * (D - S);
            const float qkl = Q[k * 3 + l], qlk = Q[l * 3 + k];
            P[k * 3 + l] = a * qkl + b * qlk;
            P[l * 3 + k] = a * qlk + b * qkl;
        }
    }
    // Ftr adjoint = U P V^T
    float Ftb[9];
    {
        float T2[9];
        matmul(U, P, T2);
        matmul_nt(T2, V, Ftb);
    }
    // F_tr = A F: Ab = Ftb F^T, Fb = A^T Ftb; A = I + dt G
    float Ab[9];
    matmul_nt(Ftb, F, Ab);
#pragma unroll
    for (int k = 0; k < 9; ++k) Gb[k] = dt * Ab[k];
    matmul_tn(A, Ftb, Fb);
}

__device__ __forceinline__ void load9(const float* p, int64_t S, int plane, float o[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = p[(int64_t)(plane + k) * S];
}

// The material parameters' adjoints of a workgroup: a wave reduction, the
// waves' sums in s_par (3 per wave), and after a barrier one double atomic per
// parameter from thread 0 (block_flush_params).  Every wave of every
// workgroup adds to the same three doubles, so per-wave atomics would
// serialize thousands of same-address atomics at the end of the kernel.
__device__ __forceinline__ void block_stage_params(float mub, float lamb, float alb, float* s_par) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mub += __shfl_xor(mub, o);
        lamb += __shfl_xor(lamb, o);
        alb += __shfl_xor(alb, o);
    }
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        s_par[3 * w] = mub;
        s_par[3 * w + 1] = lamb;
        s_par[3 * w + 2] = alb;
    }
}

__device__ __forceinline__ void block_flush_params(const float* s_par, double* out) {
    if (threadIdx.x < 3) {
        double acc = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) acc += (double)s_par[3 * w + threadIdx.x];
        atomicAdd(&out[threadIdx.x], acc);
    }
}

// ---------------------------------------------------------------------------
// G2P adjoint of one particle.
//
// Per axis a the quadratic B-spline factors at fx_a: w, e = w (n - fx),
// d = dw/dfx, and their x-derivatives w' = inv_dx dw, e' = inv_dx (dw (n - fx)
// - w), d' = inv_dx ddw (ddw = 1, -2, 1).  The grid-velocity adjoint the
// particle scatters is
//   vbn_a(ijk) = w vt_a + 4 inv_dx w sum_d Cb_ad (n_d - fx_d)
//                + sum_d Gb_ad grad_d w
//              = A_a[i] w_y w_z + w_x B_a[j] w_z + w_x w_y C_a[k]
// with A_a = vt_a w_x + c4 Cb_a0 e_x + inv_dx Gb_a0 d_x, B_a = c4 Cb_a1 e_y +
// inv_dx Gb_a1 d_y, C_a = c4 Cb_a2 e_z + inv_dx Gb_a2 d_z, and the position
// adjoint it adds is sum_nodes sum_a v_a d(vbn_a)/dx_e: the same form with
// the axis-e factor differentiated.  Both are evaluated as tensor-product
// contractions (z, then y, then x) instead of per node.
// ---------------------------------------------------------------------------
// per-axis B-spline factors at fx: w, e = w (n - fx), d = dw/dfx and their
// x-derivatives (n = 0, 1, 2)
struct AxisFactors {
    float w[3], e[3], d[3], wp[3], ep[3], dp[3];
};

__device__ __forceinline__ AxisFactors axis_factors(float fx, float inv_dx) {
    AxisFactors f;
    const float wa = 1.5f - fx, wb = fx - 1.0f, wc = fx - 0.5f;
    f.w[0] = 0.5f * wa * wa;
    f.w[1] = 0.75f - wb * wb;
    f.w[2] = 0.5f * wc * wc;
    f.d[0] = fx - 1.5f;
    f.d[1] = -2.0f * (fx - 1.0f);
    f.d[2] = fx - 0.5f;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const float off = (float)n - fx;
        f.e[n] = f.w[n] * off;
        f.wp[n] = inv_dx * f.d[n];
        f.ep[n] = inv_dx * (f.d[n] * off - f.w[n]);
        f.dp[n] = inv_dx * (n == 1 ? -2.0f : 1.0f);
    }
    return f;
}

// The 27 stencil nodes (t = i*9 + j*3 + k, gather_tile's strides 9, 3) of
// a particle: in the chunk's LDS tile (SoA, kCT^3 nodes; ti = tile index of
// node 0), or in the global grid (node 0 at index gb).
struct TileNodes {
    const float* s_v;
    int ti;
    __device__ __forceinline__ float comp(int c, int t) const {
        const int i = t / 9, j = (t / 3) % 3, k = t % 3;
        return s_v[c * kCT3 + ti + (i * kCT + j) * kCT + k];
    }
};
struct GridNodes {
    const float4* gv;
    int64_t gb;
    int ng;
    __device__ __forceinline__ float comp(int c, int t) const {
        const int i = t / 9, j = (t / 3) % 3, k = t % 3;
        const float4 g = gv[gb + ((int64_t)i * ng + j) * ng + k];
        return c == 0 ? g.x : (c == 1 ? g.y : g.z);
    }
};

// vbn scatter (emit(a, i, j, k, value)) and position adjoint (added to xg).
template <typename Nodes, typename Emit>
__device__ __forceinline__ void g2p_adj_nodes(const Nodes& V, const float fx[3], const float vt[3], const float Cb[9],
                                              const float Gb[9], float inv_dx, float xg[3], Emit&& emit) {
    const float c4 = 4.0f * inv_dx;
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        const AxisFactors X0 = axis_factors(fx[0], inv_dx);
        const AxisFactors X1 = axis_factors(fx[1], inv_dx);
        const AxisFactors X2 = axis_factors(fx[2], inv_dx);
        const float ca0 = c4 * Cb[a * 3], ca1 = c4 * Cb[a * 3 + 1], ca2 = c4 * Cb[a * 3 + 2];
        const float ga0 = inv_dx * Gb[a * 3], ga1 = inv_dx * Gb[a * 3 + 1], ga2 = inv_dx * Gb[a * 3 + 2];
        float A[3], Ap[3], B[3], Bp[3], Cc[3], Ccp[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            A[n] = fmaf(vt[a], X0.w[n], fmaf(ca0, X0.e[n], ga0 * X0.d[n]));
            Ap[n] = fmaf(vt[a], X0.wp[n], fmaf(ca0, X0.ep[n], ga0 * X0.dp[n]));
            B[n] = fmaf(ca1, X1.e[n], ga1 * X1.d[n]);
            Bp[n] = fmaf(ca1, X1.ep[n], ga1 * X1.dp[n]);
            Cc[n] = fmaf(ca2, X2.e[n], ga2 * X2.d[n]);
            Ccp[n] = fmaf(ca2, X2.ep[n], ga2 * X2.dp[n]);
        }
        float t0 = 0.f, t1 = 0.f, t2 = 0.f;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float y0 = 0.f, y1 = 0.f, y2 = 0.f, y3 = 0.f, y4 = 0.f, y5 = 0.f, y6 = 0.f, y7 = 0.f, y8 = 0.f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float ab = fmaf(A[i], X1.w[j], X0.w[i] * B[j]);
                const float w01 = X0.w[i] * X1.w[j];
                float zw = 0.f, zc = 0.f, zwp = 0.f, zcp = 0.f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float v = V.comp(a, i * 9 + j * 3 + k);
                    zw = fmaf(X2.w[k], v, zw);
                    zc = fmaf(Cc[k], v, zc);
                    zwp = fmaf(X2.wp[k], v, zwp);
                    zcp = fmaf(Ccp[k], v, zcp);
                    emit(a, i, j, k, fmaf(ab, X2.w[k], w01 * Cc[k]));
                }
                y0 = fmaf(X1.w[j], zw, y0);
                y1 = fmaf(B[j], zw, y1);
                y2 = fmaf(X1.w[j], zc, y2);
                y3 = fmaf(X1.wp[j], zw, y3);
                y4 = fmaf(Bp[j], zw, y4);
                y5 = fmaf(X1.wp[j], zc, y5);
                y6 = fmaf(X1.w[j], zwp, y6);
                y7 = fmaf(B[j], zwp, y7);
                y8 = fmaf(X1.w[j], zcp, y8);
            }
            t0 = fmaf(Ap[i], y0, fmaf(X0.wp[i], y1 + y2, t0));
            t1 = fmaf(A[i], y3, fmaf(X0.w[i], y4 + y5, t1));
            t2 = fmaf(A[i], y6, fmaf(X0.w[i], y7 + y8, t2));
        }
        xg[0] += t0;
        xg[1] += t1;
        xg[2] += t2;
    }
}

// G2P adjoint of particle p (stencil inside the grid).  st: S_n; an: adjoint
// of S_{n+1}; ac: adjoint of S_n (x and F written here, v, C, tau by
// k_p2g_adj).  V: the particle's 27 grid velocities; emit(a, i, j, k, value)
// receives its grid-velocity adjoint.
// A particle's first inputs to its G2P adjoint: position, F, and the
// adjoint of the next F; the chunk loop loads the next particle's while it
// processes the current one.
struct AdjIn {
    float x[3], F[9], Fbn[9];
};

__device__ __forceinline__ void load_adj_in(const float* __restrict__ st, const float* __restrict__ cst, int64_t S,
                                            const float* __restrict__ an, int64_t SA, int p, AdjIn& q) {
#pragma unroll
    for (int d = 0; d < 3; ++d) q.x[d] = cst[(MPM_P_X0 + d) * S + p] + st[(MPM_P_U + d) * S + p];
    load9(st + p, S, MPM_P_F, q.F);
    load9(an + p, SA, kAF, q.Fbn);
}

// The x, v, C, tau adjoints of S_{n+1} as the G2P adjoint reads them: from
// the adjoint planes (AdjPlanes: ap = an + p), or from registers (AdjRegs:
// the fused kernel has just completed them with the P2G adjoint of
// substep n+1).
struct AdjPlanes {
    const float* ap;
    int64_t SA;
    __device__ __forceinline__ void tau(float o[9]) const { load9(ap, SA, kAT, o); }
    __device__ __forceinline__ void C(float o[9]) const { load9(ap, SA, kAC, o); }
    __device__ __forceinline__ float x(int d) const { return ap[(kAX + d) * SA]; }
    __device__ __forceinline__ float v(int d) const { return ap[(kAV + d) * SA]; }
};
struct AdjNext {
    float x[3], v[3], C[9], tau[9];
};
struct AdjRegs {
    const AdjNext& a;
    __device__ __forceinline__ void tau(float o[9]) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = a.tau[k];
    }
    __device__ __forceinline__ void C(float o[9]) const {
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = a.C[k];
    }
    __device__ __forceinline__ float x(int d) const { return a.x[d]; }
    __device__ __forceinline__ float v(int d) const { return a.v[d]; }
};

template <typename Nodes, typename Src, typename Emit>
__device__ __forceinline__ void g2p_adj_particle(const AdjIn& q, int p, const MpmParams& prm, float dt,
                                                 const AxisWeights& Aw, const Nodes& V, const Src& an, int64_t SA,
                                                 float* __restrict__ ac, float& mub, float& lamb, float& alb,
                                                 Emit&& emit) {
    const float inv_dx = prm.inv_dx;
    const float* x = q.x;
    float G[9];
    {
        float v[3], Cm[9];
        gather_tile<9, 3>(V, 0, Aw, inv_dx, v, Cm, G);
    }
    float xb[3], vt[3], Cb[9], Gb[9], Fb[9];
    {
        float Tb[9];
        an.tau(Tb);
        sand_update_adj(G, q.F, dt, prm.mu, prm.lam, prm.alpha, q.Fbn, Tb, Gb, Fb, mub, lamb, alb);
    }
    an.C(Cb);
    // x' = x + dt v'
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        xb[d] = an.x(d);
        vt[d] = an.v(d) + dt * xb[d];
    }
    float fx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) fx[d] = x[d] * inv_dx - (float)Aw.base[d];
    g2p_adj_nodes(V, fx, vt, Cb, Gb, inv_dx, xb, emit);
    float* cp = ac + p;
#pragma unroll
    for (int d = 0; d < 3; ++d) cp[(kAX + d) * SA] = xb[d];
#pragma unroll
    for (int k = 0; k < 9; ++k) cp[(kAF + k) * SA] = Fb[k];
}

__device__ __forceinline__ bool base_inside(const int b[3], int ng) {
    return b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[0] <= ng - 3 && b[1] <= ng - 3 && b[2] <= ng - 3;
}

// A frozen particle (stencil outside the grid) is copied unchanged by the
// forward substep: its whole adjoint passes through.
__device__ __forceinline__ void frozen_adj(const float* __restrict__ an, int64_t SA, int p, float* __restrict__ ac) {
#pragma unroll 3
    for (int k = 0; k < MPM_ADJ_NPLANES; ++k) ac[k * SA + p] = an[k * SA + p];
}

// P2G adjoint of one particle of S_n whose stencil base lies inside the
// grid (defined with k_p2g_adj below): the v, C, tau adjoints of S_n and the
// P2G part of its position adjoint.
struct P2GAdjOut {
    float v[3], C[9], tau[9], xg[3];
};
__device__ __forceinline__ void p2g_adj_particle(const float* __restrict__ sp, const float* __restrict__ kp,
                                                 int64_t S, const float fx[3], int64_t b0, const MpmParams& prm,
                                                 float dt, const float4* __restrict__ gb, P2GAdjOut& o);

// position, stencil base and fractional offsets of particle p of S_n as the
// P2G (and its adjoint) takes them; false when the stencil leaves the grid
// (a frozen particle)
__device__ __forceinline__ bool p2g_base(const float* __restrict__ st, const float* __restrict__ cst, int64_t S,
                                         int p, const MpmParams& prm, float fx[3], int64_t* b0) {
    const int ng = prm.n_grid;
    int base[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float gp = (cst[(MPM_P_X0 + d) * S + p] + st[(MPM_P_U + d) * S + p]) * prm.inv_dx;
        base[d] = (int)(gp - 0.5f);
        fx[d] = gp - (float)base[d];
    }
    *b0 = ((int64_t)base[0] * ng + base[1]) * ng + base[2];
    return base[0] >= 0 && base[1] >= 0 && base[2] >= 0 && base[0] <= ng - 3 && base[1] <= ng - 3 &&
           base[2] <= ng - 3;
}

// The fused adjoint of one particle: the P2G adjoint of substep n+1 (on
// S_{n+1} = st_next, grid adjoint gb_next) completes the adjoint of S_{n+1},
// of which the planes an hold the part the G2P adjoint of substep n+1 wrote
// (x, F; every plane for a particle frozen at n+1); the G2P adjoint of
// substep n then takes it from registers.
template <typename Nodes, typename Emit>
__device__ __forceinline__ void fused_adj_particle(const AdjIn& q, int p, const float* __restrict__ st_next,
                                                   const float* __restrict__ cst, int64_t S,
                                                   const float4* __restrict__ gb_next, const MpmParams& prm, float dt,
                                                   const AxisWeights& Aw, const Nodes& V,
                                                   const float* __restrict__ an, int64_t SA, float* __restrict__ ac,
                                                   float& mub, float& lamb, float& alb, Emit&& emit) {
    const float* ap = an + p;
    AdjNext a;
    float fx[3];
    int64_t b0;
    if (p2g_base(st_next, cst, S, p, prm, fx, &b0)) {
        P2GAdjOut o;
        p2g_adj_particle(st_next + p, cst + p, S, fx, b0, prm, dt, gb_next, o);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.x[d] = ap[(kAX + d) * SA] + o.xg[d];
            a.v[d] = o.v[d];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            a.C[k] = o.C[k];
            a.tau[k] = o.tau[k];
        }
    } else {
        // frozen at substep n+1: its G2P adjoint passed every plane through
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.x[d] = ap[(kAX + d) * SA];
            a.v[d] = ap[(kAV + d) * SA];
        }
        load9(ap, SA, kAC, a.C);
        load9(ap, SA, kAT, a.tau);
    }
    g2p_adj_particle(q, p, prm, dt, Aw, V, AdjRegs{a}, SA, ac, mub, lamb, alb, emit);
}

// G2P adjoint over the chunks of an ordered state (mpm_sort_by_bin): the
// chunk's grid velocities are staged in an LDS tile at its origin
// (chunk_origin); the grid-velocity adjoint is accumulated in an FP64 LDS tile there and
// flushed once per node (chunk_flush); particles whose stencil leaves the
// tile gather from and add to the global grids directly.  kFused: each
// particle first runs the P2G adjoint of substep n+1 (fused_adj_particle;
// st_next, gb_next), which spares k_p2g_adj's pass over the adjoint planes
// between the substeps of a segment.
template <bool kFused>
__global__ void __launch_bounds__(kChunkWG)
    k_g2p_adj_chunks(const float* __restrict__ st, const float* __restrict__ cst, int64_t S, int n,
                     const int4* __restrict__ chunks,
                     const int* __restrict__ meta, MpmParams prm, float dt, const float4* __restrict__ gv,
                     const float* __restrict__ an, int64_t SA, float* __restrict__ vb, float* __restrict__ ac,
                     double* __restrict__ pb, const float* __restrict__ st_next,
                     const float4* __restrict__ gb_next) {
    __shared__ double s_t[3 * kCT3];
    __shared__ float s_v[3 * kCT3];
    __shared__ float s_par[3 * (kChunkWG / 64)];
    const int c = blockIdx.x;
    if (c >= meta[0]) return;
    const int4 ch = chunks[c];
    const int tid = threadIdx.x;
    const int ng = prm.n_grid;
    const int n_per = min(ch.y, kChunkMax);
    for (int t = tid; t < 3 * kCT3; t += kChunkWG) s_t[t] = 0.0;
    int o[3];
    chunk_origin(ch.z, (ng + 7) / 8, o);
    // the chunk's grid velocities (nodes outside the grid read as 0)
    for (int t = tid; t < kCT3; t += kChunkWG) {
        const int x = o[0] + t / (kCT * kCT), y = o[1] + (t / kCT) % kCT, z = o[2] + t % kCT;
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        if (x >= 0 && y >= 0 && z >= 0 && x < ng && y < ng && z < ng) g = gv[((int64_t)x * ng + y) * ng + z];
        s_v[t] = g.x;
        s_v[kCT3 + t] = g.y;
        s_v[2 * kCT3 + t] = g.z;
    }
    __syncthreads();
    float mub = 0.f, lamb = 0.f, alb = 0.f;
    // particles whose stencil lies in the chunk's tile; the others are
    // marked (bit r for the thread's r-th particle) for the second loop
    unsigned outside = 0;
    AdjIn q;
    if (tid < n_per && ch.x + tid < n) load_adj_in(st, cst, S, an, SA, ch.x + tid, q);
    for (int i = tid, r = 0; i < n_per; i += kChunkWG, ++r) {
        const int p = ch.x + i;
        const AdjIn cur = q;
        if (i + kChunkWG < n_per && p + kChunkWG < n) load_adj_in(st, cst, S, an, SA, p + kChunkWG, q);
        if (p < 0 || p >= n) continue;
        AxisWeights Aw;
        axis_weights(cur.x, prm.inv_dx, Aw);
        if (!base_inside(Aw.base, ng)) {
            frozen_adj(an, SA, p, ac);
            continue;
        }
        const int ti = chunk_tile_index(Aw.base, o);
        if (ti < 0) {
            outside |= 1u << r;
            continue;
        }
        auto emit = [&](int a, int i3, int j3, int k3, float val) {
            atomicAdd(&s_t[a * kCT3 + ti + (i3 * kCT + j3) * kCT + k3], (double)val);
        };
        if constexpr (kFused)
            fused_adj_particle(cur, p, st_next, cst, S, gb_next, prm, dt, Aw, TileNodes{s_v, ti}, an, SA, ac, mub,
                               lamb, alb, emit);
        else
            g2p_adj_particle(cur, p, prm, dt, Aw, TileNodes{s_v, ti}, AdjPlanes{an + p, SA}, SA, ac, mub, lamb, alb,
                             emit);
    }
    for (int i = tid, r = 0; outside && i < n_per; i += kChunkWG, ++r) {
        if (!(outside & (1u << r))) continue;
        const int p = ch.x + i;
        load_adj_in(st, cst, S, an, SA, p, q);
        AxisWeights Aw;
        axis_weights(q.x, prm.inv_dx, Aw);
        const int64_t gb = ((int64_t)Aw.base[0] * ng + Aw.base[1]) * ng + Aw.base[2];
        auto emit = [&](int a, int i3, int j3, int k3, float val) {
            unsafeAtomicAdd(vb + 4 * (gb + ((int64_t)i3 * ng + j3) * ng + k3) + a, val);
        };
        if constexpr (kFused)
            fused_adj_particle(q, p, st_next, cst, S, gb_next, prm, dt, Aw, GridNodes{gv, gb, ng}, an, SA, ac, mub,
                               lamb, alb, emit);
        else
            g2p_adj_particle(q, p, prm, dt, Aw, GridNodes{gv, gb, ng}, AdjPlanes{an + p, SA}, SA, ac, mub, lamb, alb,
                             emit);
    }
    block_stage_params(mub, lamb, alb, s_par);
    __syncthreads();
    block_flush_params(s_par, pb);
    chunk_flush<3>(s_t, o, ng, vb);
}

// ---------------------------------------------------------------------------
// Grid adjoint: gb (mv_b, m_b) from vb through the colliders and the
// normalization, on the recomputed P2G grid acc (mv, m).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlk)
    k_grid_adj(const float4* __restrict__ acc, float* __restrict__ vb, float4* __restrict__ gb, MpmParams prm,
               ColliderSet cs, float time, float dt, int clear_vb) {
    const int64_t nn = (int64_t)prm.n_grid * prm.n_grid * prm.n_grid;
    const int64_t t = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (t >= nn) return;
    const int ng = prm.n_grid;
    const int gx = (int)(t / ((int64_t)ng * ng)), gy = (int)((t / ng) % ng), gz = (int)(t % ng);
    const float4 a4 = acc[t];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a4.w > 1e-15f) {
        const float im = 1.0f / a4.w;
        float o[3] = {a4.x * im + dt * prm.gravity[0], a4.y * im + dt * prm.gravity[1], a4.z * im + dt * prm.gravity[2]};
        float mask[3] = {1.f, 1.f, 1.f};
        for (int c = 0; c < cs.n; ++c)   // the colliders of grid_node_update, recording replaced components
            apply_collider(cs.c[c], gx, gy, gz, ng, prm.dx, dt, time, o, mask);
        const float b0 = vb[4 * t] * mask[0], b1 = vb[4 * t + 1] * mask[1], b2 = vb[4 * t + 2] * mask[2];
        out.x = b0 * im;
        out.y = b1 * im;
        out.z = b2 * im;
        out.w = -(b0 * a4.x + b1 * a4.y + b2 * a4.z) * im * im;
    }
    gb[t] = out;
    // the velocity adjoint the next (earlier) substep scatters into, cleared
    // here instead of by its own launch
    if (clear_vb) reinterpret_cast<float4*>(vb)[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// k_grid_adj over the 8^3 grid blocks the substep's P2G marked (one
// workgroup per block, one thread per node): the G2P adjoint scatters vb
// only into the stencils of the substep's particles, which its P2G marked,
// and the P2G adjoint reads gb only there; the other blocks leave after one
// load (their vb is zero, their gb never read).
__global__ void __launch_bounds__(512)
    k_grid_adj_blocks(const float4* __restrict__ acc, const int* __restrict__ flags, float* __restrict__ vb,
                      float4* __restrict__ gb, MpmParams prm, ColliderSet cs, float time, float dt, int clear_vb) {
    const int ng = prm.n_grid, nb = (ng + 7) / 8;
    const int blk = blockIdx.x;
    if (!flags[blk]) return;
    const int tt = threadIdx.x;
    const int bz = blk % nb, by = (blk / nb) % nb, bx = blk / (nb * nb);
    const int gx = bx * 8 + (tt >> 6), gy = by * 8 + ((tt >> 3) & 7), gz = bz * 8 + (tt & 7);
    if (gx >= ng || gy >= ng || gz >= ng) return;
    const int64_t t = ((int64_t)gx * ng + gy) * ng + gz;
    const float4 a4 = acc[t];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a4.w > 1e-15f) {
        const float im = 1.0f / a4.w;
        float o[3] = {a4.x * im + dt * prm.gravity[0], a4.y * im + dt * prm.gravity[1], a4.z * im + dt * prm.gravity[2]};
        float mask[3] = {1.f, 1.f, 1.f};
        for (int c = 0; c < cs.n; ++c)
            apply_collider(cs.c[c], gx, gy, gz, ng, prm.dx, dt, time, o, mask);
        const float b0 = vb[4 * t] * mask[0], b1 = vb[4 * t + 1] * mask[1], b2 = vb[4 * t + 2] * mask[2];
        out.x = b0 * im;
        out.y = b1 * im;
        out.z = b2 * im;
        out.w = -(b0 * a4.x + b1 * a4.y + b2 * a4.z) * im * im;
    }
    gb[t] = out;
    if (clear_vb) reinterpret_cast<float4*>(vb)[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// P2G adjoint: v, C, tau adjoints of S_n and the position adjoint's P2G part
// (added to what k_g2p_adj_chunks wrote).  For each grid-adjoint field f
// (mv_b x, y, z and m_b) the kernel needs 22 weighted sums over the 27
// nodes, every weight a product of one factor per axis from w, e = w (n - fx),
// d = dw, p = dw (n - fx), dd = d2w (grid units):
//   sum w f (v, C via e; the C^T term of x), sum grad w f (tau, and the v
//   and m_b terms of x), sum grad_e w (x_d - x_p) f (the C term of x), and
//   sum Hessian_de w f (the tau term of x);
// they are contracted z, then y, then x (tensor-product form) instead of
// node by node.
// ---------------------------------------------------------------------------
struct AxisF5 {
    float w[3], e[3], d[3], p[3], dd[3];
};

__device__ __forceinline__ AxisF5 axis_f5(float fx) {
    AxisF5 f;
    const float wa = 1.5f - fx, wb = fx - 1.0f, wc = fx - 0.5f;
    f.w[0] = 0.5f * wa * wa;
    f.w[1] = 0.75f - wb * wb;
    f.w[2] = 0.5f * wc * wc;
    f.d[0] = fx - 1.5f;
    f.d[1] = -2.0f * (fx - 1.0f);
    f.d[2] = fx - 0.5f;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const float off = (float)n - fx;
        f.e[n] = f.w[n] * off;
        f.p[n] = f.d[n] * off;
        f.dd[n] = n == 1 ? -2.0f : 1.0f;
    }
    return f;
}

// indices of the 22 sums: (x, y, z) factors
enum { kS0, kEx, kEy, kEz, kDx, kDy, kDz, kMxy, kMyx, kMxz, kMzx, kMyz, kMzy, kPx, kPy, kPz, kHxy, kHxz, kHyz,
       kDDx, kDDy, kDDz, kNSum };

// The 22 sums of the three momentum-adjoint fields and the three of the mass
// adjoint in one pass over the 27 nodes (one float4 load per node).
__device__ __forceinline__ void p2g_adj_sums(const float4* __restrict__ gb, int64_t b0, int ng, const AxisF5& X,
                                             const AxisF5& Y, const AxisF5& Z, float r[3][kNSum], float rm[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = 0; q < kNSum; ++q) r[a][q] = 0.f;
    rm[0] = rm[1] = rm[2] = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float y[3][12], ym[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 12; ++q) y[a][q] = 0.f;
        ym[0] = ym[1] = ym[2] = 0.f;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float zw[4] = {0.f, 0.f, 0.f, 0.f}, zd[4] = {0.f, 0.f, 0.f, 0.f};
            float ze[3] = {0.f, 0.f, 0.f}, zp[3] = {0.f, 0.f, 0.f}, zdd[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float4 g = gb[b0 + ((int64_t)i * ng + j) * ng + k];
                const float f[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    zw[a] = fmaf(Z.w[k], f[a], zw[a]);
                    zd[a] = fmaf(Z.d[k], f[a], zd[a]);
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    ze[a] = fmaf(Z.e[k], f[a], ze[a]);
                    zp[a] = fmaf(Z.p[k], f[a], zp[a]);
                    zdd[a] = fmaf(Z.dd[k], f[a], zdd[a]);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                y[a][0] = fmaf(Y.w[j], zw[a], y[a][0]);     // (., w, w)
                y[a][1] = fmaf(Y.e[j], zw[a], y[a][1]);     // (., e, w)
                y[a][2] = fmaf(Y.w[j], ze[a], y[a][2]);     // (., w, e)
                y[a][3] = fmaf(Y.d[j], zw[a], y[a][3]);     // (., d, w)
                y[a][4] = fmaf(Y.w[j], zd[a], y[a][4]);     // (., w, d)
                y[a][5] = fmaf(Y.d[j], ze[a], y[a][5]);     // (., d, e)
                y[a][6] = fmaf(Y.e[j], zd[a], y[a][6]);     // (., e, d)
                y[a][7] = fmaf(Y.p[j], zw[a], y[a][7]);     // (., p, w)
                y[a][8] = fmaf(Y.w[j], zp[a], y[a][8]);     // (., w, p)
                y[a][9] = fmaf(Y.d[j], zd[a], y[a][9]);     // (., d, d)
                y[a][10] = fmaf(Y.dd[j], zw[a], y[a][10]);  // (., dd, w)
                y[a][11] = fmaf(Y.w[j], zdd[a], y[a][11]);  // (., w, dd)
            }
            ym[0] = fmaf(Y.w[j], zw[3], ym[0]);
            ym[1] = fmaf(Y.d[j], zw[3], ym[1]);
            ym[2] = fmaf(Y.w[j], zd[3], ym[2]);
        }
        const float wx = X.w[i], ex = X.e[i], dx = X.d[i], px = X.p[i], ddx = X.dd[i];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float* ra = r[a];
            const float* ya = y[a];
            ra[kS0] = fmaf(wx, ya[0], ra[kS0]);
            ra[kEx] = fmaf(ex, ya[0], ra[kEx]);
            ra[kEy] = fmaf(wx, ya[1], ra[kEy]);
            ra[kEz] = fmaf(wx, ya[2], ra[kEz]);
            ra[kDx] = fmaf(dx, ya[0], ra[kDx]);
            ra[kDy] = fmaf(wx, ya[3], ra[kDy]);
            ra[kDz] = fmaf(wx, ya[4], ra[kDz]);
            ra[kMxy] = fmaf(dx, ya[1], ra[kMxy]);
            ra[kMyx] = fmaf(ex, ya[3], ra[kMyx]);
            ra[kMxz] = fmaf(dx, ya[2], ra[kMxz]);
            ra[kMzx] = fmaf(ex, ya[4], ra[kMzx]);
            ra[kMyz] = fmaf(wx, ya[5], ra[kMyz]);
            ra[kMzy] = fmaf(wx, ya[6], ra[kMzy]);
            ra[kPx] = fmaf(px, ya[0], ra[kPx]);
            ra[kPy] = fmaf(wx, ya[7], ra[kPy]);
            ra[kPz] = fmaf(wx, ya[8], ra[kPz]);
            ra[kHxy] = fmaf(dx, ya[3], ra[kHxy]);
            ra[kHxz] = fmaf(dx, ya[4], ra[kHxz]);
            ra[kHyz] = fmaf(wx, ya[9], ra[kHyz]);
            ra[kDDx] = fmaf(ddx, ya[0], ra[kDDx]);
            ra[kDDy] = fmaf(wx, ya[10], ra[kDDy]);
            ra[kDDz] = fmaf(wx, ya[11], ra[kDDz]);
        }
        rm[0] = fmaf(dx, ym[0], rm[0]);   // mass adjoint: D_x
        rm[1] = fmaf(wx, ym[1], rm[1]);   // D_y
        rm[2] = fmaf(wx, ym[2], rm[2]);   // D_z
    }
}

__device__ __forceinline__ void p2g_adj_particle(const float* __restrict__ sp, const float* __restrict__ kp,
                                                 int64_t S, const float fx[3], int64_t b0, const MpmParams& prm,
                                                 float dt, const float4* __restrict__ gb, P2GAdjOut& o) {
    const int ng = prm.n_grid;
    const float inv_dx = prm.inv_dx, dxp = prm.dx;
    const AxisF5 X = axis_f5(fx[0]), Y = axis_f5(fx[1]), Z = axis_f5(fx[2]);
    float r[3][kNSum], rm[3];
    p2g_adj_sums(gb, b0, ng, X, Y, Z, r, rm);
    const float m = kp[MPM_P_MASS * S], vol = kp[MPM_P_VOL * S];
    const float ct = dt * vol * inv_dx * inv_dx;
    float* xg = o.xg;
    xg[0] = m * inv_dx * rm[0];   // the mass adjoint m_b
    xg[1] = m * inv_dx * rm[1];
    xg[2] = m * inv_dx * rm[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float* ra = r[a];
        const float va = sp[(MPM_P_V + a) * S];
        float C[3], tau[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            C[d] = sp[(MPM_P_C + a * 3 + d) * S];
            tau[d] = sp[(MPM_P_TAU + a * 3 + d) * S];
        }
        o.v[a] = m * ra[kS0];
        o.C[a * 3 + 0] = m * dxp * ra[kEx];
        o.C[a * 3 + 1] = m * dxp * ra[kEy];
        o.C[a * 3 + 2] = m * dxp * ra[kEz];
        o.tau[a * 3 + 0] = -dt * vol * inv_dx * ra[kDx];
        o.tau[a * 3 + 1] = -dt * vol * inv_dx * ra[kDy];
        o.tau[a * 3 + 2] = -dt * vol * inv_dx * ra[kDz];
        const float mv = m * va * inv_dx;
        xg[0] += mv * ra[kDx] + m * (C[0] * ra[kPx] + C[1] * ra[kMxy] + C[2] * ra[kMxz] - C[0] * ra[kS0]) -
                 ct * (tau[0] * ra[kDDx] + tau[1] * ra[kHxy] + tau[2] * ra[kHxz]);
        xg[1] += mv * ra[kDy] + m * (C[0] * ra[kMyx] + C[1] * ra[kPy] + C[2] * ra[kMyz] - C[1] * ra[kS0]) -
                 ct * (tau[0] * ra[kHxy] + tau[1] * ra[kDDy] + tau[2] * ra[kHyz]);
        xg[2] += mv * ra[kDz] + m * (C[0] * ra[kMzx] + C[1] * ra[kMzy] + C[2] * ra[kPz] - C[2] * ra[kS0]) -
                 ct * (tau[0] * ra[kHxz] + tau[1] * ra[kHyz] + tau[2] * ra[kDDz]);
    }
}

__global__ void __launch_bounds__(kBlk)
    k_p2g_adj(const float* __restrict__ st, const float* __restrict__ cst, int64_t S, int n, MpmParams prm, float dt,
              const float4* __restrict__ gb, float* __restrict__ ac, int64_t SA) {
    const int p = blockIdx.x * kBlk + threadIdx.x;
    if (p >= n) return;
    float fx[3];
    int64_t b0;
    if (!p2g_base(st, cst, S, p, prm, fx, &b0)) return;   // frozen: k_g2p_adj_chunks passed v, C, tau through
    P2GAdjOut o;
    p2g_adj_particle(st + p, cst + p, S, fx, b0, prm, dt, gb, o);
    float* cp = ac + p;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cp[(kAV + a) * SA] = o.v[a];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            cp[(kAC + a * 3 + d) * SA] = o.C[a * 3 + d];
            cp[(kAT + a * 3 + d) * SA] = o.tau[a * 3 + d];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) cp[(kAX + d) * SA] += o.xg[d];
}

__global__ void k_zero4(float4* __restrict__ g, int64_t nn) {
    const int64_t t = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (t < nn) g[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

unsigned blocks(int64_t n) { return (unsigned)((n + kBlk - 1) / kBlk); }

}  // namespace

extern "C" {

int64_t mpm_adjoint_workspace_bytes(const MpmParams* prm) {
    if (check_grid(prm)) return -1;
    const int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    // a grid set (acc (mv, m), grid v, block marks), then the v adjoint and
    // the (mv, m) adjoint
    return mpm_step_grids_bytes(prm) + 2 * nn * 16;
}

// mpm_adjoint_substep; vb_clear: the workspace's velocity-adjoint grid is
// already zero (cleared by the previous call's k_grid_adj); clear_after: leave
// it zero for the next call.  step_flags: the block marks of step_grids when
// it is a grid set (a recorded segment), NULL for the two dense grids
// mpm_step_ordered leaves.
static int adjoint_substep(const MpmParams* prm, const MpmState* st, const void* order, const void* step_grids,
                           const int* step_flags, const MpmCollider* colliders, int n_colliders, void* workspace,
                           float time, float dt, const float* adj_next, float* adj_cur, int64_t adj_stride,
                           double* param_adj, bool vb_clear, bool clear_after, const float* cst, void* stream) {
    int rc = check_sand(prm);
    if (rc) return rc;
    if (!st || !st->planes || st->n < 0 || st->stride < st->n) return set_err(-4, "state invalid");
    if (st->n > 0x7fffffffLL) return set_err(-6, "more than 2^31-1 particles");
    if (!workspace || ((uintptr_t)workspace & 255)) return set_err(-11, "workspace NULL or not 256-byte aligned");
    if (!order || ((uintptr_t)order & 255)) return set_err(-40, "order buffer NULL or not 256-byte aligned");
    if (!adj_next || !adj_cur || !param_adj || adj_stride < st->n) return set_err(-30, "adjoint buffers invalid");
    ColliderSet cs;
    if ((rc = make_collider_set(colliders, n_colliders, &cs))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int64_t nn = (int64_t)prm->n_grid * prm->n_grid * prm->n_grid;
    if (step_grids && ((uintptr_t)step_grids & 255)) return set_err(-11, "step grids not 256-byte aligned");
    // workspace: a grid set (accumulator, velocities, block marks), then the
    // velocity adjoint and the (mv, m) adjoint
    float* wset = (float*)workspace;
    float4* acc = (float4*)wset;
    float4* gv = acc + nn;
    float4* vbg = (float4*)(wset + mpm_step_grids_bytes(prm) / 4);
    float4* gbg = vbg + nn;
    const int* flags = step_flags;
    if (step_grids) {
        // the grids of this substep recorded by the forward
        acc = (float4*)step_grids;
        gv = acc + nn;
    } else {
        // forward grid of this substep (P2G over the chunks, normalization, colliders)
        if ((rc = clear_grid_set(prm, wset, stream))) return rc;
        int* wflags = grid_set_flags(wset, prm->n_grid);
        if ((rc = p2g_ordered(prm, st, order, wset, dt, stream, wflags))) return rc;
        if ((rc = grid_update_blocks(prm, wset, colliders, n_colliders, time, dt, nullptr, stream))) return rc;
        flags = wflags;
    }
    if (!vb_clear) {
        hipLaunchKernelGGL(k_zero4, dim3(blocks(nn)), dim3(kBlk), 0, s, vbg, nn);
        if ((rc = check_launch("k_zero4"))) return rc;
    }
    const int n = (int)st->n;
    if (n > 0) {
        const OrderLayout L = order_layout(*prm, n);
        const char* ob = (const char*)order;
        hipLaunchKernelGGL(k_g2p_adj_chunks<false>, dim3((unsigned)L.max_chunks), dim3(kChunkWG), 0, s,
                           (const float*)st->planes, cst ? cst : (const float*)st->planes, st->stride, n,
                           (const int4*)(ob + L.off_chunks),
                           (const int*)(ob + L.off_meta), *prm, dt, (const float4*)gv, adj_next, adj_stride,
                           (float*)vbg, adj_cur, param_adj, (const float*)nullptr, (const float4*)nullptr);
        if ((rc = check_launch("k_g2p_adj_chunks"))) return rc;
    }
    if (flags) {
        // only the blocks this substep's P2G marked hold its particles' stencils
        hipLaunchKernelGGL(k_grid_adj_blocks, dim3((unsigned)grid_blocks(prm->n_grid)), dim3(512), 0, s,
                           (const float4*)acc, flags, (float*)vbg, gbg, *prm, cs, time, dt, clear_after ? 1 : 0);
        if ((rc = check_launch("k_grid_adj_blocks"))) return rc;
    } else {
        hipLaunchKernelGGL(k_grid_adj, dim3(blocks(nn)), dim3(kBlk), 0, s, (const float4*)acc, (float*)vbg, gbg,
                           *prm, cs, time, dt, clear_after ? 1 : 0);
        if ((rc = check_launch("k_grid_adj"))) return rc;
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_p2g_adj, dim3(blocks(n)), dim3(kBlk), 0, s, (const float*)st->planes,
                           cst ? cst : (const float*)st->planes, st->stride, n, *prm, dt, (const float4*)gbg, adj_cur,
                           adj_stride);
        if ((rc = check_launch("k_p2g_adj"))) return rc;
    }
    return 0;
}

int mpm_adjoint_substep(const MpmParams* prm, const MpmState* st, const void* order, const void* step_grids,
                        const MpmCollider* colliders, int n_colliders, void* workspace, float time, float dt,
                        const float* adj_next, float* adj_cur, int64_t adj_stride, double* param_adj, void* stream) {
    return adjoint_substep(prm, st, order, step_grids, nullptr, colliders, n_colliders, workspace, time, dt, adj_next,
                           adj_cur, adj_stride, param_adj, false, false, nullptr, stream);
}

// shared checks of a segment's states and per-substep workspaces
static int check_segment(const MpmParams* prm, const MpmState* states, void* const* step_grids, int n_sub) {
    if (!states || !step_grids || n_sub < 1) return set_err(-31, "segment: states, grids or n_sub invalid");
    const int64_t stride = states[0].stride, n = states[0].n;
    for (int j = 0; j < n_sub; ++j)
        if (!states[j].planes || states[j].stride != stride || states[j].n != n)
            return set_err(-32, "segment: states must share n and stride");
    if (check_grid(prm)) return set_err(-2, "n_grid out of range");
    for (int j = 0; j < n_sub; ++j)
        if (!step_grids[j] || ((uintptr_t)step_grids[j] & 255))
            return set_err(-11, "segment: step workspaces NULL or not 256-byte aligned");
    return 0;
}

// The forward half of a segment: the P2G of S_k0 (states[0]; skipped when
// p2g_done says step_grids[0] holds it already), then per substep j its grid
// update into step_grids[j] and the fused G2P + next P2G states[j] ->
// states[j+1] into step_grids[j+1].  When `last` is given, the last substep
// writes S_{k0+n_sub} there and its P2G to last_acc.  Every grid is a grid
// set whose accumulator is zero outside its marked blocks (the caller
// zeroes it before its first use); a grid update clears the marked blocks
// of the set the next P2G writes, so no grid is swept whole.
static int record_segment(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                          const MpmCollider* colliders, int n_colliders, int k0, double dt, int n_sub, int p2g_done,
                          const MpmState* last, float* last_acc, void* stream) {
    const int ng = prm->n_grid;
    const int64_t nn = (int64_t)ng * ng * ng;
    const float fdt = (float)dt;
    int rc;
    float* acc0 = (float*)step_grids[0];
    if (!p2g_done) {
        if ((rc = clear_grid_set(prm, acc0, stream))) return rc;
        if ((rc = p2g_ordered(prm, &states[0], order, acc0, fdt, stream, grid_set_flags(acc0, ng)))) return rc;
    }
    for (int j = 0; j < n_sub; ++j) {
        float* acc = (float*)step_grids[j];
        float* gvj = acc + 4 * nn;
        const bool more = j + 1 < n_sub;
        float* acc_next = more ? (float*)step_grids[j + 1] : (last ? last_acc : nullptr);   // NULL: none after
        if ((rc = grid_update_blocks(prm, acc, colliders, n_colliders, (float)((double)(k0 + j) * dt), fdt, acc_next,
                                     stream)))
            return rc;
        if (!acc_next) break;
        // the segment's later states hold only the planes a substep changes;
        // x0, m, vol, id stay in states[0] (and are written to `last`)
        if ((rc = g2p2g_ordered(prm, &states[j], more ? &states[j + 1] : last, order, gvj, acc_next, fdt,
                                (const float*)states[0].planes, !more, stream, grid_set_flags(acc_next, ng))))
            return rc;
    }
    return 0;
}

// The reverse half: adjoint substeps k0+n_sub-1 .. k0 on the recorded states
// and grids; adj_a holds the adjoint of S_{k0+n_sub} on entry and that of
// S_k0 on return.  Per substep j: the G2P adjoint (scattering the velocity
// adjoint of grid j), the grid adjoint over the blocks grid j's P2G marked,
// then the P2G adjoint of substep j -- which, for j > 0, runs inside the
// G2P adjoint of substep j-1 (k_g2p_adj_chunks<true>): the adjoint of S_j
// passes between the two in registers instead of through the adjoint planes.
static int adjoint_recorded(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                            const MpmCollider* colliders, int n_colliders, void* workspace, int k0, double dt,
                            int n_sub, float* adj_a, float* adj_b, int64_t adj_stride, double* param_adj,
                            void* stream) {
    int rc = check_sand(prm);
    if (rc) return rc;
    const int64_t n64 = states[0].n;
    if (n64 < 0 || states[0].stride < n64) return set_err(-4, "state invalid");
    if (n64 > 0x7fffffffLL) return set_err(-6, "more than 2^31-1 particles");
    if (!workspace || ((uintptr_t)workspace & 255)) return set_err(-11, "workspace NULL or not 256-byte aligned");
    if (!order || ((uintptr_t)order & 255)) return set_err(-40, "order buffer NULL or not 256-byte aligned");
    if (!param_adj || adj_stride < n64) return set_err(-30, "adjoint buffers invalid");
    ColliderSet cs;
    if ((rc = make_collider_set(colliders, n_colliders, &cs))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int ng = prm->n_grid;
    const int64_t nn = (int64_t)ng * ng * ng;
    const int n = (int)n64;
    const float fdt = (float)dt;
    const float* cst = (const float*)states[0].planes;   // x0, m, vol, id of every state of the segment
    float4* vbg = (float4*)((float*)workspace + mpm_step_grids_bytes(prm) / 4);
    float4* gbg = vbg + nn;
    const OrderLayout L = order_layout(*prm, n);
    const char* ob = (const char*)order;
    hipLaunchKernelGGL(k_zero4, dim3(blocks(nn)), dim3(kBlk), 0, s, vbg, nn);
    if ((rc = check_launch("k_zero4"))) return rc;
    float* cur = adj_a;
    float* nxt = adj_b;
    for (int j = n_sub - 1; j >= 0; --j) {
        float* set = (float*)step_grids[j];
        const MpmState& st = states[j];
        if (n > 0) {
            // G2P adjoint of substep j (for j < n_sub-1 with the P2G adjoint of substep j+1 in front)
            const bool fused = j < n_sub - 1;
            auto kern = fused ? k_g2p_adj_chunks<true> : k_g2p_adj_chunks<false>;
            hipLaunchKernelGGL(kern, dim3((unsigned)L.max_chunks), dim3(kChunkWG), 0, s, (const float*)st.planes,
                               cst, st.stride, n, (const int4*)(ob + L.off_chunks), (const int*)(ob + L.off_meta),
                               *prm, fdt, (const float4*)(set + 4 * nn), (const float*)cur, adj_stride, (float*)vbg,
                               nxt, param_adj, fused ? (const float*)states[j + 1].planes : (const float*)nullptr,
                               fused ? (const float4*)gbg : (const float4*)nullptr);
            if ((rc = check_launch("k_g2p_adj_chunks"))) return rc;
        }
        // grid adjoint of substep j over its marked blocks; clears the velocity adjoint behind it
        hipLaunchKernelGGL(k_grid_adj_blocks, dim3((unsigned)grid_blocks(ng)), dim3(512), 0, s, (const float4*)set,
                           (const int*)grid_set_flags(set, ng), (float*)vbg, gbg, *prm, cs,
                           (float)((double)(k0 + j) * dt), fdt, 1);
        if ((rc = check_launch("k_grid_adj_blocks"))) return rc;
        float* t = cur; cur = nxt; nxt = t;
    }
    // P2G adjoint of the segment's first substep
    if (n > 0) {
        hipLaunchKernelGGL(k_p2g_adj, dim3(blocks(n)), dim3(kBlk), 0, s, (const float*)states[0].planes, cst,
                           states[0].stride, n, *prm, fdt, (const float4*)gbg, cur, adj_stride);
        if ((rc = check_launch("k_p2g_adj"))) return rc;
    }
    if (cur != adj_a && hipMemcpyAsync(adj_a, cur, (size_t)MPM_ADJ_NPLANES * adj_stride * sizeof(float),
                                       hipMemcpyDeviceToDevice, (hipStream_t)stream) != hipSuccess)
        return set_err(-33, "segment: adjoint copy failed");
    return 0;
}

int mpm_adjoint_segment(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                        const MpmCollider* colliders, int n_colliders, void* workspace, int k0, double dt, int n_sub,
                        float* adj_a, float* adj_b, int64_t adj_stride, double* param_adj, void* stream) {
    int rc;
    if ((rc = check_segment(prm, states, step_grids, n_sub))) return rc;
    if (!adj_a || !adj_b || adj_a == adj_b) return set_err(-30, "adjoint buffers invalid");
    if ((rc = record_segment(prm, states, order, step_grids, colliders, n_colliders, k0, dt, n_sub, 0, nullptr,
                             nullptr, stream)))
        return rc;
    return adjoint_recorded(prm, states, order, step_grids, colliders, n_colliders, workspace, k0, dt, n_sub, adj_a,
                            adj_b, adj_stride, param_adj, stream);
}

int mpm_record_segment(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                       const MpmCollider* colliders, int n_colliders, int k0, double dt, int n_sub, int p2g_done,
                       const MpmState* last, void* next_grid, int64_t next_grid_bytes, void* stream) {
    int rc;
    if ((rc = check_segment(prm, states, step_grids, n_sub))) return rc;
    if (states[0].n < 0 || states[0].stride < states[0].n || states[0].n > 0x7fffffffLL)
        return set_err(-5, "segment: state n/stride invalid");
    if (!order || ((uintptr_t)order & 255)) return set_err(-40, "order buffer NULL or not 256-byte aligned");
    if (last) {
        if (!last->planes || last->n != states[0].n || last->stride != states[0].stride)
            return set_err(-32, "segment: the last state must share n and stride");
        for (int j = 0; j < n_sub; ++j)
            if (states[j].planes == last->planes) return set_err(-32, "segment: the last state aliases a state");
        if (!next_grid || ((uintptr_t)next_grid & 255) || next_grid_bytes < mpm_step_grids_bytes(prm))
            return set_err(-10, "segment: next_grid must be one 256-byte aligned grid set (mpm_step_grids_bytes)");
        for (int j = 0; j < n_sub; ++j)
            if (step_grids[j] == next_grid) return set_err(-10, "segment: next_grid aliases a step grid");
    }
    return record_segment(prm, states, order, step_grids, colliders, n_colliders, k0, dt, n_sub, p2g_done, last,
                          (float*)next_grid, stream);
}

int mpm_adjoint_recorded(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                         const MpmCollider* colliders, int n_colliders, void* workspace, int k0, double dt, int n_sub,
                         float* adj_a, float* adj_b, int64_t adj_stride, double* param_adj, void* stream) {
    int rc;
    if ((rc = check_segment(prm, states, step_grids, n_sub))) return rc;
    if (!adj_a || !adj_b || adj_a == adj_b) return set_err(-30, "adjoint buffers invalid");
    return adjoint_recorded(prm, states, order, step_grids, colliders, n_colliders, workspace, k0, dt, n_sub, adj_a,
                            adj_b, adj_stride, param_adj, stream);
}


}  // extern "C"
