// mpm_device.h -- per-particle FP32 math of the sand substep, in registers.
//
//   bspline_weights   quadratic B-spline w, dw     mpm_utils.py:196-209, 244-257
//   svd3_rot          3x3 SVD (rotation U, V)       wp.svd3 at mpm_utils.py:92,297
//   sand_return_map   Drucker-Prager projection     mpm_utils.py:87-112
//   dp_stress_sym     Kirchhoff stress of F_e       mpm_utils.py:45-52
//   sand_update_matfn the two above as functions of F_tr F_tr^T (no SVD)
//   von_mises_return_map  metal projection          mpm_utils.py:59-84
//   stvk_stress       StVK (Hencky) stress          mpm_utils.py:36-43
//   fcr_stress        fixed-corotated stress        mpm_utils.py:29-34
//   constitutive_update  the G2P material switch    mpm_utils.py:279-306
//
// All matrices are row-major float[9].  Loops have constant trip counts and
// are fully unrolled so that every array lives in VGPRs.
#pragma once
#include <hip/hip_runtime.h>

namespace mpm {

struct Weights {
    int base[3];
    float fx[3];
    float w[3][3];   // w[axis][node]
    float dw[3][3];  // dw[axis][node]
};

// mpm_utils.py:196-209: base = int(x*inv_dx - 0.5) (truncation), fx in
// [0.5, 1.5); w = (0.5(1.5-fx)^2, 0.75-(fx-1)^2, 0.5(fx-0.5)^2),
// dw = (fx-1.5, -2(fx-1), fx-0.5).
__device__ __forceinline__ void bspline_weights(const float x[3], float inv_dx, Weights& W) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float gp = x[d] * inv_dx;
        int b = (int)(gp - 0.5f);
        W.base[d] = b;
        float fx = gp - (float)b;
        W.fx[d] = fx;
        float wa = 1.5f - fx, wb = fx - 1.0f, wc = fx - 0.5f;
        W.w[d][0] = wa * wa * 0.5f;
        W.w[d][1] = (0.0f - wb * wb) + 0.75f;
        W.w[d][2] = wc * wc * 0.5f;
        W.dw[d][0] = fx - 1.5f;
        W.dw[d][1] = -2.0f * (fx - 1.0f);
        W.dw[d][2] = fx - 0.5f;
    }
}

// 1/sqrt(x) refined by one Newton step from the hardware estimate, so that
// rotation cosines/sines built from it satisfy c^2 + s^2 = 1 to ~1 ulp.
__device__ __forceinline__ float rsqrt_nr(float x) {
    float y = rsqrtf(x);
    float h = 0.5f * x * y;
    return y * fmaf(-h, y, 1.5f);
}

// One Jacobi rotation zeroing a_pq of the symmetric matrix (app, aqq, apq,
// arp, arq); V's columns p, q rotate with it.  t = tan of the rotation
// angle, the smaller root of t^2 + 2 theta t - 1 = 0, theta = (aqq-app)/(2apq).
template <int P, int Q>
__device__ __forceinline__ void jacobi_rotate(float& app, float& aqq, float& apq, float& arp,
                                              float& arq, float V[9]) {
    // t only needs to define a rotation (c, s below are normalised from it),
    // so the hardware sqrt and reciprocal are enough here
    float d = aqq - app;
    float den = fabsf(d) + __builtin_amdgcn_sqrtf(d * d + 4.0f * apq * apq);
    float t = (den > 0.0f) ? ((d >= 0.0f ? 2.0f : -2.0f) * apq) * __builtin_amdgcn_rcpf(den) : 0.0f;
    float c = rsqrt_nr(1.0f + t * t);
    float s = t * c;
    app = app - t * apq;
    aqq = aqq + t * apq;
    apq = 0.0f;
    float rp = arp, rq = arq;
    arp = c * rp - s * rq;
    arq = s * rp + c * rq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float vp = V[k * 3 + P], vq = V[k * 3 + Q];
        V[k * 3 + P] = c * vp - s * vq;
        V[k * 3 + Q] = s * vp + c * vq;
    }
}

// Givens rotation on rows (P, Q) of B zeroing B[Q][COL]; U <- U G^T.
template <int P, int Q, int COL>
__device__ __forceinline__ void givens_qr(float B[9], float U[9]) {
    float a = B[P * 3 + COL], b = B[Q * 3 + COL];
    float r2 = a * a + b * b;
    bool ok = r2 > 1e-36f;
    float ir = ok ? rsqrt_nr(r2) : 0.0f;
    float c = ok ? a * ir : 1.0f;
    float s = ok ? b * ir : 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float bp = B[P * 3 + k], bq = B[Q * 3 + k];
        B[P * 3 + k] = c * bp + s * bq;
        B[Q * 3 + k] = -s * bp + c * bq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float up = U[k * 3 + P], uq = U[k * 3 + Q];
        U[k * 3 + P] = c * up + s * uq;
        U[k * 3 + Q] = -s * up + c * uq;
    }
}

template <int I, int J>
__device__ __forceinline__ void sort_swap(float B[9], float V[9]) {
    float ni = B[I] * B[I] + B[3 + I] * B[3 + I] + B[6 + I] * B[6 + I];
    float nj = B[J] * B[J] + B[3 + J] * B[3 + J] + B[6 + J] * B[6 + J];
    bool sw = ni < nj;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float bi = B[k * 3 + I], bj = B[k * 3 + J];
        B[k * 3 + I] = sw ? bj : bi;
        B[k * 3 + J] = sw ? -bi : bj;
        float vi = V[k * 3 + I], vj = V[k * 3 + J];
        V[k * 3 + I] = sw ? vj : vi;
        V[k * 3 + J] = sw ? -vi : vj;
    }
}

// SVD A = U diag(s) V^T after the structure of McAdams et al. 2011 (Jacobi
// eigenanalysis of A^T A, column sort of AV by decreasing norm, Givens QR of
// AV), with exact Jacobi rotations.  U and V are rotations; s0 >= s1 >= |s2|
// and s2 < 0 iff det A < 0.
__device__ __forceinline__ void svd3_rot(const float A[9], float U[9], float s[3], float V[9]) {
    float s00 = A[0] * A[0] + A[3] * A[3] + A[6] * A[6];
    float s11 = A[1] * A[1] + A[4] * A[4] + A[7] * A[7];
    float s22 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8];
    float s01 = A[0] * A[1] + A[3] * A[4] + A[6] * A[7];
    float s02 = A[0] * A[2] + A[3] * A[5] + A[6] * A[8];
    float s12 = A[1] * A[2] + A[4] * A[5] + A[7] * A[8];
#pragma unroll
    for (int k = 0; k < 9; ++k) V[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    for (int sweep = 0; sweep < 6; ++sweep) {
        float off = s01 * s01 + s02 * s02 + s12 * s12;
        float dia = s00 * s00 + s11 * s11 + s22 * s22;
        if (!(off > dia * 1e-15f)) break;
        jacobi_rotate<0, 1>(s00, s11, s01, s02, s12, V);
        jacobi_rotate<0, 2>(s00, s22, s02, s01, s12, V);
        jacobi_rotate<1, 2>(s11, s22, s12, s01, s02, V);
    }
    float B[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            B[i * 3 + j] = A[i * 3 + 0] * V[0 * 3 + j] + A[i * 3 + 1] * V[1 * 3 + j] +
                           A[i * 3 + 2] * V[2 * 3 + j];
    sort_swap<0, 1>(B, V);
    sort_swap<1, 2>(B, V);
    sort_swap<0, 1>(B, V);
#pragma unroll
    for (int k = 0; k < 9; ++k) U[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    givens_qr<0, 1, 0>(B, U);
    givens_qr<0, 2, 0>(B, U);
    givens_qr<1, 2, 1>(B, U);
    s[0] = B[0];
    s[1] = B[4];
    s[2] = B[8];
}

// sand_return_mapping (mpm_utils.py:87-112) on F_tr given its SVD.  Writes
// F_e and ls = ln of the singular values of F_e = U diag(exp ls) V^T, the
// logarithms the stress needs:
//   branch 1 (dg <= 0): F_e = F_tr bitwise, ls = ln s (no abs/clamp, as the
//                       stress at mpm_utils.py:47-50 takes it);
//   branch 2 (dg > 0, tr > 0): F_e = U V^T, ls = 0;
//   branch 3 (dg > 0, tr <= 0): ls = H = eps - eps_hat dg/|eps_hat|,
//                       F_e = U diag(exp H) V^T.
// Returns the branch.
__device__ __forceinline__ int sand_return_map(const float Ftr[9], const float U[9],
                                               const float s[3], const float V[9], float mu,
                                               float lam, float alpha, float Fe[9],
                                               float ls[3]) {
    float eps[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) eps[d] = logf(fmaxf(fabsf(s[d]), 1e-14f));
    float tr = eps[0] + eps[1] + eps[2];
    float eh0 = eps[0] - tr / 3.0f, eh1 = eps[1] - tr / 3.0f, eh2 = eps[2] - tr / 3.0f;
    float ehn = sqrtf(eh0 * eh0 + eh1 * eh1 + eh2 * eh2);
    float dg = ehn + (3.0f * lam + 2.0f * mu) / (2.0f * mu) * tr * alpha;
    if (dg <= 0.0f) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Fe[k] = Ftr[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) ls[d] = eps[d];
        // ln of a singular value below 1e-14 (NaN when negative), as the
        // stress takes it; a branch so that the common case skips the logf
        if (fminf(fminf(s[0], s[1]), s[2]) < 1e-14f) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (s[d] < 1e-14f) ls[d] = logf(s[d]);
        }
        return 1;
    }
    float se[3];
    int br;
    if (tr > 0.0f) {
        se[0] = se[1] = se[2] = 1.0f;
        ls[0] = ls[1] = ls[2] = 0.0f;
        br = 2;
    } else {
        float r = dg / ehn;
        ls[0] = eps[0] - eh0 * r;
        ls[1] = eps[1] - eh1 * r;
        ls[2] = eps[2] - eh2 * r;
        se[0] = expf(ls[0]);
        se[1] = expf(ls[1]);
        se[2] = expf(ls[2]);
        br = 3;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Fe[i * 3 + j] = U[i * 3 + 0] * se[0] * V[j * 3 + 0] + U[i * 3 + 1] * se[1] * V[j * 3 + 1] +
                            U[i * 3 + 2] * se[2] * V[j * 3 + 2];
    return br;
}

// kirchoff_stress_drucker_prager (mpm_utils.py:45-52):
//   tau = U diag((2 mu ln s_i + lam sum ln s)/s_i) V^T F_e^T.
// With F_e = U diag(s) V^T this is U diag(2 mu ln s_i + lam sum ln s) U^T,
// evaluated in that symmetric form from the SVD the return map already
// holds (the reference recomputes the SVD of F_e, mpm_utils.py:297).
// ln s_i is taken without abs or clamp, as the reference does.
__device__ __forceinline__ void dp_stress_sym(const float U[9], const float ls[3], float mu,
                                              float lam, float tau[9]) {
    float lsum = ls[0] + ls[1] + ls[2];
    float c0 = 2.0f * mu * ls[0] + lam * lsum;
    float c1 = 2.0f * mu * ls[1] + lam * lsum;
    float c2 = 2.0f * mu * ls[2] + lam * lsum;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            float t = U[i * 3 + 0] * c0 * U[j * 3 + 0] + U[i * 3 + 1] * c1 * U[j * 3 + 1] +
                      U[i * 3 + 2] * c2 * U[j * 3 + 2];
            tau[i * 3 + j] = t;
            tau[j * 3 + i] = t;
        }
}

// ---------------------------------------------------------------------------
// The sand update without a decomposition.  With F_tr = U diag(s) V^T and
// B = F_tr F_tr^T = U diag(s^2) U^T, every quantity the return map and the
// stress take is a function of B alone:
//   L = 1/2 ln B = U diag(ln s) U^T       (eps_i = ln s_i; tr L = sum eps)
//   D = L - tr(L)/3 I = U diag(eps_hat) U^T, |eps_hat| = |D|_F
//   branch 1: F_e = F_tr,                    H = L
//   branch 2: F_e = U V^T = exp(-L) F_tr,    H = 0
//   branch 3: F_e = U diag(e^H) V^T = exp(-r D) F_tr, H = L - r D, r = dg/|D|
//   tau = U diag(2 mu H_i + lam sum H) U^T = 2 mu H + lam tr(H) I
// (sand_return_mapping mpm_utils.py:87-112, kirchoff_stress_drucker_prager
// :45-52).  L comes from the series 1/2 ln B = atanh(Z) = Z (1 + Z^2/3 +
// Z^4/5 + ...), Z = X (2I + X)^-1, X = B - I, and exp from its Taylor series;
// all products are of commuting symmetric matrices.  Each particle takes the
// fewest terms that keep the truncation below 1/4 ulp for it (the bounds
// below), so its result does not depend on the other particles of its wave.  Particles whose B is far from I (an
// eigenvalue of B outside [1/3, 3]) or whose F_tr is inverted or singular
// (det <= 0, where the reference's ln takes a negative or zero singular
// value) return false and take the SVD path.
// ---------------------------------------------------------------------------
struct Sym3 {
    float xx, xy, xz, yy, yz, zz;
};

// product of two commuting symmetric matrices (the product is symmetric)
__device__ __forceinline__ Sym3 sym_mul(const Sym3& a, const Sym3& b) {
    Sym3 c;
    c.xx = a.xx * b.xx + a.xy * b.xy + a.xz * b.xz;
    c.xy = a.xx * b.xy + a.xy * b.yy + a.xz * b.yz;
    c.xz = a.xx * b.xz + a.xy * b.yz + a.xz * b.zz;
    c.yy = a.xy * b.xy + a.yy * b.yy + a.yz * b.yz;
    c.yz = a.xy * b.xz + a.yy * b.yz + a.yz * b.zz;
    c.zz = a.xz * b.xz + a.yz * b.yz + a.zz * b.zz;
    return c;
}

// c I + k M
__device__ __forceinline__ Sym3 sym_axpi(float k, const Sym3& m, float c) {
    return Sym3{fmaf(k, m.xx, c), k * m.xy, k * m.xz, fmaf(k, m.yy, c), k * m.yz, fmaf(k, m.zz, c)};
}

__device__ __forceinline__ float sym_norm2(const Sym3& m) {
    return m.xx * m.xx + m.yy * m.yy + m.zz * m.zz + 2.0f * (m.xy * m.xy + m.xz * m.xz + m.yz * m.yz);
}

// Truncation bounds.  atanh: with K terms the first omitted one is
// z^(2K+1)/(2K+1) for an eigenvalue z of Z, |z| <= 1/2; relative to z it is
// below |Z|_F^(2K) / (0.75 (2K+1)); bound k-1 is the largest |Z|_F^2 at
// which k terms keep that below 3e-8.  Taylor exp: with n terms after I the
// first omitted one is |Y|^(n+1)/(n+1)!; bound n-1 is the largest |Y|_F^2
// keeping it below 3e-8.
constexpr int kLogTerms = 12;
constexpr float kLogBound[kLogTerms - 1] = {6.75e-8f, 3.35e-4f, 5.40e-3f, 2.12e-2f, 4.75e-2f, 8.14e-2f,
                                            1.19e-1f, 1.58e-1f, 1.96e-1f, 2.32e-1f, 2.68e-1f};
constexpr float kLogMax = 0.25f;   // |Z|_F^2 beyond this takes the SVD path
constexpr int kExpTerms = 10;
constexpr float kExpBound[kExpTerms - 1] = {6.0e-8f, 3.19e-5f, 8.49e-4f, 6.62e-3f, 2.79e-2f,
                                            8.06e-2f, 1.865e-1f, 3.66e-1f, 6.41e-1f};

// Returns false (F, tau unspecified) when the SVD path must be taken.
// mu, lam, alpha as sand_return_map.  F and tau receive F_e and tau.
__device__ __forceinline__ bool sand_update_matfn(const float Ftr[9], float mu, float lam, float alpha, float F[9],
                                                  float tau[9]) {
    // X = F_tr F_tr^T - I; the diagonal as (F_ii - 1)(F_ii + 1) + ..., which
    // keeps the small difference from I exact where F_ii is near 1
    Sym3 X;
    X.xx = fmaf(Ftr[1], Ftr[1], fmaf(Ftr[2], Ftr[2], (Ftr[0] - 1.0f) * (Ftr[0] + 1.0f)));
    X.yy = fmaf(Ftr[3], Ftr[3], fmaf(Ftr[5], Ftr[5], (Ftr[4] - 1.0f) * (Ftr[4] + 1.0f)));
    X.zz = fmaf(Ftr[6], Ftr[6], fmaf(Ftr[7], Ftr[7], (Ftr[8] - 1.0f) * (Ftr[8] + 1.0f)));
    X.xy = Ftr[0] * Ftr[3] + Ftr[1] * Ftr[4] + Ftr[2] * Ftr[5];
    X.xz = Ftr[0] * Ftr[6] + Ftr[1] * Ftr[7] + Ftr[2] * Ftr[8];
    X.yz = Ftr[3] * Ftr[6] + Ftr[4] * Ftr[7] + Ftr[5] * Ftr[8];
    const float det = Ftr[0] * (Ftr[4] * Ftr[8] - Ftr[5] * Ftr[7]) - Ftr[1] * (Ftr[3] * Ftr[8] - Ftr[5] * Ftr[6]) +
                      Ftr[2] * (Ftr[3] * Ftr[7] - Ftr[4] * Ftr[6]);
    // Z = X adj(M) / det(M), M = 2I + X
    const float m00 = X.xx + 2.0f, m11 = X.yy + 2.0f, m22 = X.zz + 2.0f;
    Sym3 Ad;
    Ad.xx = m11 * m22 - X.yz * X.yz;
    Ad.xy = X.xz * X.yz - X.xy * m22;
    Ad.xz = X.xy * X.yz - X.xz * m11;
    Ad.yy = m00 * m22 - X.xz * X.xz;
    Ad.yz = X.xy * X.xz - m00 * X.yz;
    Ad.zz = m00 * m11 - X.xy * X.xy;
    const float idm = __builtin_amdgcn_rcpf(m00 * Ad.xx + X.xy * Ad.xy + X.xz * Ad.xz);
    Sym3 Z = sym_mul(X, Ad);
    Z = Sym3{Z.xx * idm, Z.xy * idm, Z.xz * idm, Z.yy * idm, Z.yz * idm, Z.zz * idm};
    const Sym3 Z2 = sym_mul(Z, Z);
    const float n2 = Z2.xx + Z2.yy + Z2.zz;   // |Z|_F^2
    const bool ok = det > 0.0f && n2 <= kLogMax;
    // terms of the series: each particle's own count (its result must not
    // depend on the other particles of its wave); the loops run to the most
    // any lane of the wave needs
    int K = 1;
    bool more = ok;
#pragma unroll
    for (int k = 0; k < kLogTerms - 1; ++k) {
        more = more && n2 > kLogBound[k];
        if (!__any(more)) break;
        K += more ? 1 : 0;
    }
    // P = sum_j Z^(2j) / (2j+1), j < K, by Horner
    Sym3 P = Sym3{1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 1.0f};
#pragma unroll
    for (int j = kLogTerms - 1; j >= 1; --j) {
        if (!__any(j < K)) continue;
        const float c = 1.0f / (float)(2 * j + 1);
        if (j == K - 1) P = Sym3{c, 0.0f, 0.0f, c, 0.0f, c};
        else if (j < K - 1) P = sym_axpi(1.0f, sym_mul(Z2, P), c);
    }
    if (K > 1) P = sym_axpi(1.0f, sym_mul(Z2, P), 1.0f);
    const Sym3 L = sym_mul(Z, P);
    const float tr = L.xx + L.yy + L.zz;
    const float t3 = tr * (1.0f / 3.0f);
    const Sym3 D = Sym3{L.xx - t3, L.xy, L.xz, L.yy - t3, L.yz, L.zz - t3};
    const float ehn = __builtin_amdgcn_sqrtf(sym_norm2(D));
    const float dg = ehn + (3.0f * lam + 2.0f * mu) / (2.0f * mu) * tr * alpha;
    Sym3 H = L;
    if (dg > 0.0f) {
        // branch 2 (tr > 0): H = 0, F_e = exp(-L) F_tr; branch 3: H = L - r D, F_e = exp(-r D) F_tr
        Sym3 Y;
        if (tr > 0.0f) {
            H = Sym3{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            Y = Sym3{-L.xx, -L.xy, -L.xz, -L.yy, -L.yz, -L.zz};
        } else {
            const float r = dg / ehn;
            Y = Sym3{-r * D.xx, -r * D.xy, -r * D.xz, -r * D.yy, -r * D.yz, -r * D.zz};
            H = Sym3{L.xx + Y.xx, L.xy + Y.xy, L.xz + Y.xz, L.yy + Y.yy, L.yz + Y.yz, L.zz + Y.zz};
        }
        const float ny2 = sym_norm2(Y);
        int n = 1;   // this particle's terms, as K above
        bool up = ok;
#pragma unroll
        for (int k = 0; k < kExpTerms - 1; ++k) {
            up = up && ny2 > kExpBound[k];
            if (!__any(up)) break;
            n += up ? 1 : 0;
        }
        // E = I + Y (I + Y/2 (I + ... (I + Y/n)))
        Sym3 E = Sym3{1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 1.0f};
#pragma unroll
        for (int k = kExpTerms; k >= 1; --k) {
            if (!__any(k <= n)) continue;
            const float ik = 1.0f / (float)k;
            if (k == n) E = sym_axpi(ik, Y, 1.0f);
            else if (k < n) E = sym_axpi(ik, sym_mul(Y, E), 1.0f);
        }
        const float Ef[9] = {E.xx, E.xy, E.xz, E.xy, E.yy, E.yz, E.xz, E.yz, E.zz};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                F[i * 3 + j] = Ef[i * 3 + 0] * Ftr[0 * 3 + j] + Ef[i * 3 + 1] * Ftr[1 * 3 + j] +
                               Ef[i * 3 + 2] * Ftr[2 * 3 + j];
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = Ftr[k];
    }
    const float c2 = 2.0f * mu, cl = lam * (H.xx + H.yy + H.zz);
    tau[0] = fmaf(c2, H.xx, cl);
    tau[4] = fmaf(c2, H.yy, cl);
    tau[8] = fmaf(c2, H.zz, cl);
    tau[1] = tau[3] = c2 * H.xy;
    tau[2] = tau[6] = c2 * H.xz;
    tau[5] = tau[7] = c2 * H.yz;
    return ok;
}

// F_tr = (I + dt gradv) F  (mpm_utils.py:279-282)
__device__ __forceinline__ void trial_F(const float gradv[9], float dt, const float F[9], float Ftr[9]) {
    float A[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = gradv[k] * dt;
    A[0] += 1.0f; A[4] += 1.0f; A[8] += 1.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Ftr[i * 3 + j] = A[i * 3 + 0] * F[0 * 3 + j] + A[i * 3 + 1] * F[1 * 3 + j] +
                             A[i * 3 + 2] * F[2 * 3 + j];
}

// Whole per-particle constitutive update of g2p_apic_and_stress
// (mpm_utils.py:279-306), sand: F_tr = (I + dt gradv) F, return map, stress
// (sand_update_matfn, the SVD form where it declines).
__device__ __forceinline__ void sand_update(const float gradv[9], float dt, float F[9],
                                            float mu, float lam, float alpha, float tau[9]) {
    float Ftr[9];
    trial_F(gradv, dt, F, Ftr);
    if (!sand_update_matfn(Ftr, mu, lam, alpha, F, tau)) {
        float U[9], s[3], V[9];
        svd3_rot(Ftr, U, s, V);
        float ls[3];
        sand_return_map(Ftr, U, s, V, mu, lam, alpha, F, ls);
        dp_stress_sym(U, ls, mu, lam, tau);
    }
}

// von_mises_return_mapping (mpm_utils.py:59-84) on F_tr given its SVD, for
// metal.  sigma = max(s, 0.01); eps = ln sigma; the trial stress
// t = 2 mu eps + lam (sum eps) and its deviator cond = t - sum(t)/3.  If
// |cond| > ys: eps_hat = eps - mean(eps), n = |eps_hat| + 1e-6,
// dg = n - ys/(2 mu), eps -= (dg/n) eps_hat, F_e = U diag(exp eps) V^T and,
// when hardening == 1, ys += 2 mu xi dg.  Otherwise F_e = F_tr.  Writes ls,
// the logs of the clamped singular values of F_e that the StVK stress takes
// (the reference recomputes the SVD of F_e, mpm_utils.py:297: its singular
// values are exp(eps) in the same order, and max(., 0.01) leaves them).
// Returns true when the particle was projected.
__device__ __forceinline__ bool von_mises_return_map(const float Ftr[9], const float U[9], const float s[3],
                                                     const float V[9], float mu, float lam, float& ys,
                                                     int hardening, float xi, float Fe[9], float ls[3]) {
    float eps[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) eps[d] = logf(fmaxf(s[d], 0.01f));
    const float esum = eps[0] + eps[1] + eps[2];
    const float temp = esum / 3.0f;
    float t[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) t[d] = 2.0f * mu * eps[d] + lam * esum;
    const float st3 = (t[0] + t[1] + t[2]) / 3.0f;
    const float c0 = t[0] - st3, c1 = t[1] - st3, c2 = t[2] - st3;
    const float cn = sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
    if (!(cn > ys)) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Fe[k] = Ftr[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) ls[d] = eps[d];
        return false;
    }
    float eh[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) eh[d] = eps[d] - temp;
    const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]) + 1e-6f;
    const float dg = ehn - ys / (2.0f * mu);
    const float r = dg / ehn;
    float se[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ls[d] = eps[d] - r * eh[d];
        se[d] = expf(ls[d]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Fe[i * 3 + j] = U[i * 3 + 0] * se[0] * V[j * 3 + 0] + U[i * 3 + 1] * se[1] * V[j * 3 + 1] +
                            U[i * 3 + 2] * se[2] * V[j * 3 + 2];
    if (hardening == 1) ys = ys + 2.0f * mu * xi * dg;
    return true;
}

// kirchoff_stress_StVK (mpm_utils.py:36-43) from the SVD of F and
// ls = ln max(s, 0.01): tau = U diag(2 mu ls_i + lam sum ls) V^T -- with V^T,
// as the reference writes it (so tau is not symmetric unless U = V).
__device__ __forceinline__ void stvk_stress(const float U[9], const float V[9], const float ls[3], float mu,
                                            float lam, float tau[9]) {
    const float lsum = ls[0] + ls[1] + ls[2];
    const float c0 = 2.0f * mu * ls[0] + lam * lsum;
    const float c1 = 2.0f * mu * ls[1] + lam * lsum;
    const float c2 = 2.0f * mu * ls[2] + lam * lsum;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tau[i * 3 + j] = U[i * 3 + 0] * c0 * V[j * 3 + 0] + U[i * 3 + 1] * c1 * V[j * 3 + 1] +
                             U[i * 3 + 2] * c2 * V[j * 3 + 2];
}

__device__ __forceinline__ float det3(const float F[9]) {
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
           F[2] * (F[3] * F[7] - F[4] * F[6]);
}

// kirchoff_stress_FCR (mpm_utils.py:29-34) with R = U V^T from the SVD of F:
//   tau = 2 mu (F - R) F^T + lam J (J - 1) I.
// The full G2P passes J = det F (wp.determinant, mpm_utils.py:292), the
// sampled P2G the product of the singular values (mpm_utils.py:372-374).
__device__ __forceinline__ void fcr_stress(const float F[9], const float U[9], const float V[9], float J, float mu,
                                           float lam, float tau[9]) {
    float D[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            D[i * 3 + j] = F[i * 3 + j] - (U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] +
                                           U[i * 3 + 2] * V[j * 3 + 2]);
    const float lj = lam * J * (J - 1.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tau[i * 3 + j] = 2.0f * mu * (D[i * 3 + 0] * F[j * 3 + 0] + D[i * 3 + 1] * F[j * 3 + 1] +
                                          D[i * 3 + 2] * F[j * 3 + 2]) +
                             (i == j ? lj : 0.0f);
}

// The per-particle constitutive part of g2p_apic_and_stress
// (mpm_utils.py:279-306) for every material: F_tr = (I + dt gradv) F, then
//   material 1 (metal): F = von Mises(F_tr) on *ys, StVK stress;
//   material 2 (sand):  F = Drucker-Prager(F_tr), stress of F_e;
//   otherwise (jelly):  F = F_tr, FCR stress.
// One SVD of F_tr serves the return map and the stress (the reference
// recomputes it on F_e, mpm_utils.py:297); sand takes it only where
// sand_update_matfn declines.  `material` is uniform across a workgroup, so
// the material branches do not diverge.
__device__ __forceinline__ void constitutive_update(const float gradv[9], float dt, float F[9], int material,
                                                    float mu, float lam, float alpha, int hardening, float xi,
                                                    float& ys, float tau[9]) {
    float Ftr[9];
    trial_F(gradv, dt, F, Ftr);
    // sand without a decomposition where it can (sand_update_matfn); one SVD
    // serves the other materials and the sand particles it declines
    const bool svd = material != 2 || !sand_update_matfn(Ftr, mu, lam, alpha, F, tau);
    if (svd) {
        float U[9], s[3], V[9];
        svd3_rot(Ftr, U, s, V);
        float ls[3];
        if (material == 2) {
            sand_return_map(Ftr, U, s, V, mu, lam, alpha, F, ls);
            dp_stress_sym(U, ls, mu, lam, tau);
        } else if (material == 1) {
            von_mises_return_map(Ftr, U, s, V, mu, lam, ys, hardening, xi, F, ls);
            stvk_stress(U, V, ls, mu, lam, tau);
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = Ftr[k];
            fcr_stress(F, U, V, det3(F), mu, lam, tau);
        }
    }
}

}  // namespace mpm
