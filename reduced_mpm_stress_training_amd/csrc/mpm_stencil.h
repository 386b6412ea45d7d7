// mpm_stencil.h -- tensor-product forms of the quadratic B-spline P2G
// scatter and G2P gather used by the binned kernels.
//
// Per axis d and stencil node n in {0,1,2} the reference uses
//   w[d][n]  (mpm_utils.py:204-208)          weight factor
//   dw[d][n] (mpm_utils.py:209)              derivative factor (grid units)
//   (n - fx[d])                              offset in grid units
// and every per-node quantity it forms is a product of one factor per axis:
//   weight          = w_x w_y w_z
//   grad w (d = x)  = dw_x w_y w_z * inv_dx            (compute_dweight :54-57)
//   weight * dpos_x = (w_x (n - fx_x)) w_y w_z [* dx in P2G]
// so the 27-node sums and scatters factor axis by axis.  The arithmetic is
// the reference's up to the order of FP32 roundings.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace mpm {

struct AxisWeights {
    int base[3];
    float w[3][3];  // weight factors
    float e[3][3];  // w * (n - fx)
    float d[3][3];  // dw
};

__device__ __forceinline__ void axis_weights(const float x[3], float inv_dx, AxisWeights& A) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float gp = x[a] * inv_dx;
        int b = (int)(gp - 0.5f);
        A.base[a] = b;
        float fx = gp - (float)b;
        float wa = 1.5f - fx, wb = fx - 1.0f, wc = fx - 0.5f;
        A.w[a][0] = wa * wa * 0.5f;
        A.w[a][1] = (0.0f - wb * wb) + 0.75f;
        A.w[a][2] = wc * wc * 0.5f;
        A.d[a][0] = fx - 1.5f;
        A.d[a][1] = -2.0f * (fx - 1.0f);
        A.d[a][2] = fx - 0.5f;
#pragma unroll
        for (int n = 0; n < 3; ++n) A.e[a][n] = A.w[a][n] * ((float)n - fx);
    }
}

// G2P gather (mpm_utils.py:261-273) from a node-velocity tile: node
// (i,j,k) of the stencil is entry idx0 + i*SX + j*SY + k of the tile.
// Returns v, C = 4 inv_dx sum v_i (x) (x_i - x_p) w (grid units) and
// grad v = sum v_i (x) grad w.
// A gather source: component c of stencil node (idx0 + i*SX + j*SY + k).
struct Float4Nodes {  // float4 array, .xyz = v
    const float4* p;
    __device__ __forceinline__ float comp(int c, int t) const {
        const float4 g = p[t];
        return c == 0 ? g.x : (c == 1 ? g.y : g.z);
    }
};
struct SoANodes {     // three float arrays of stride `n` (bank-conflict friendly in LDS)
    const float* p;
    int n;
    __device__ __forceinline__ float comp(int c, int t) const { return p[c * n + t]; }
};

template <int SX, int SY, typename Nodes>
__device__ __forceinline__ void gather_tile(const Nodes& tile, int idx0, const AxisWeights& A,
                                            float inv_dx, float v[3], float C[9], float G[9]) {
    const float c4 = 4.0f * inv_dx;
    // one velocity component at a time keeps the live set small
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float sv = 0.f, cx = 0.f, gx = 0.f, cy = 0.f, gy = 0.f, cz = 0.f, gz = 0.f;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float yww = 0.f, yew = 0.f, ydw = 0.f, ywe = 0.f, ywd = 0.f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float zw = 0.f, ze = 0.f, zd = 0.f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float g = tile.comp(c, idx0 + i * SX + j * SY + k);
                    zw = fmaf(A.w[2][k], g, zw);
                    ze = fmaf(A.e[2][k], g, ze);
                    zd = fmaf(A.d[2][k], g, zd);
                }
                yww = fmaf(A.w[1][j], zw, yww);
                yew = fmaf(A.e[1][j], zw, yew);
                ydw = fmaf(A.d[1][j], zw, ydw);
                ywe = fmaf(A.w[1][j], ze, ywe);
                ywd = fmaf(A.w[1][j], zd, ywd);
            }
            sv = fmaf(A.w[0][i], yww, sv);
            cx = fmaf(A.e[0][i], yww, cx);
            gx = fmaf(A.d[0][i], yww, gx);
            cy = fmaf(A.w[0][i], yew, cy);
            gy = fmaf(A.w[0][i], ydw, gy);
            cz = fmaf(A.w[0][i], ywe, cz);
            gz = fmaf(A.w[0][i], ywd, gz);
        }
        v[c] = sv;
        C[c * 3 + 0] = cx * c4;
        C[c * 3 + 1] = cy * c4;
        C[c * 3 + 2] = cz * c4;
        G[c * 3 + 0] = gx * inv_dx;
        G[c * 3 + 1] = gy * inv_dx;
        G[c * 3 + 2] = gz * inv_dx;
    }
}

// P2G contributions (mpm_utils.py:218-227) of one particle to its 27 nodes:
//   mv_a += w m (v_a + sum_b C_ab dpos_b) + dt sum_b (-vol tau_ab) grad_b w
//   m    += w m
// factored as mv_a(i,j,k) = P_a[i] w_y w_z + Q_a[j] w_x w_z + R_a[k] w_x w_y.
// `emit(c, i, j, k, val)` receives component c (0..2 momentum, 3 mass) of
// node (i,j,k)'s contribution.
template <typename Emit>
__device__ __forceinline__ void p2g_contrib(const AxisWeights& A, float m, float vol, const float v[3],
                                            const float C[9], const float tau[9], float dx, float inv_dx,
                                            float dt, Emit&& emit) {
    const float ca = m * dx;                 // C dpos in physical units
    const float cb = -dt * vol * inv_dx;     // dt (-vol tau) grad w
    float wyz[3][3], wxz[3][3], wxy[3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            wyz[p][q] = A.w[1][p] * A.w[2][q];
            wxz[p][q] = A.w[0][p] * A.w[2][q];
            wxy[p][q] = A.w[0][p] * A.w[1][q];
        }
    // mass, then one momentum component at a time (component index 3 = mass)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float mwx = m * A.w[0][i];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) emit(3, i, j, k, mwx * wyz[j][k]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ax = ca * C[a * 3 + 0], ay = ca * C[a * 3 + 1], az = ca * C[a * 3 + 2];
        const float bx = cb * tau[a * 3 + 0], by = cb * tau[a * 3 + 1], bz = cb * tau[a * 3 + 2];
        const float mv = m * v[a];
        float P[3], Q[3], R[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            P[n] = fmaf(mv, A.w[0][n], fmaf(ax, A.e[0][n], bx * A.d[0][n]));
            Q[n] = fmaf(ay, A.e[1][n], by * A.d[1][n]);
            R[n] = fmaf(az, A.e[2][n], bz * A.d[2][n]);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    emit(a, i, j, k, fmaf(P[i], wyz[j][k], fmaf(Q[j], wxz[i][k], R[k] * wxy[i][j])));
    }
}

// Factor n of a per-axis triple for a run-time n (selects, no dynamic
// register indexing).
__device__ __forceinline__ float pick3(const float a[3], int n) { return n == 0 ? a[0] : (n == 1 ? a[1] : a[2]); }

// p2g_contrib with the x loop kept rolled: every emitted value is the same
// expression with the same FP32 roundings, but the live set stays small.
// For paths few particles take (the spill scatter), so that they do not set
// the register budget of the kernels that contain them.
template <typename Emit>
__device__ __forceinline__ void p2g_contrib_rolled(const AxisWeights& A, float m, float vol, const float v[3],
                                                   const float C[9], const float tau[9], float dx, float inv_dx,
                                                   float dt, Emit&& emit) {
    const float ca = m * dx;
    const float cb = -dt * vol * inv_dx;
    float P[3][3], Q[3][3], R[3][3];   // [component][node]
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ax = ca * C[a * 3 + 0], ay = ca * C[a * 3 + 1], az = ca * C[a * 3 + 2];
        const float bx = cb * tau[a * 3 + 0], by = cb * tau[a * 3 + 1], bz = cb * tau[a * 3 + 2];
        const float mv = m * v[a];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            P[a][n] = fmaf(mv, A.w[0][n], fmaf(ax, A.e[0][n], bx * A.d[0][n]));
            Q[a][n] = fmaf(ay, A.e[1][n], by * A.d[1][n]);
            R[a][n] = fmaf(az, A.e[2][n], bz * A.d[2][n]);
        }
    }
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        const float wx = pick3(A.w[0], i);
        const float mwx = m * wx;
        const float Pi[3] = {pick3(P[0], i), pick3(P[1], i), pick3(P[2], i)};
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float wyz = A.w[1][j] * A.w[2][k];
                const float wxz = wx * A.w[2][k];
                const float wxy = wx * A.w[1][j];
                emit(3, i, j, k, mwx * wyz);
#pragma unroll
                for (int a = 0; a < 3; ++a) emit(a, i, j, k, fmaf(Pi[a], wyz, fmaf(Q[a][j], wxz, R[a][k] * wxy)));
            }
    }
}

// G2P gather from the 27 global float4 nodes vb[(i ng + j) ng + k]: the
// contraction of gather_tile (each accumulator receives the same FMAs in
// the same order), one velocity component at a time with the x and y loops
// rolled, reading only that component of three nodes at a time: a small live
// set for the paths few particles take (the overflow run).
__device__ __forceinline__ void gather_global_rolled(const float4* __restrict__ vb, int64_t ng,
                                                     const AxisWeights& A, float inv_dx, float v[3], float C[9],
                                                     float G[9]) {
    const float c4 = 4.0f * inv_dx;
    const float* vf = reinterpret_cast<const float*>(vb);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float sv = 0.f, cx = 0.f, gx = 0.f, cy = 0.f, gy = 0.f, cz = 0.f, gz = 0.f;
#pragma unroll 1
        for (int i = 0; i < 3; ++i) {
            float yww = 0.f, yew = 0.f, ydw = 0.f, ywe = 0.f, ywd = 0.f;
#pragma unroll 1
            for (int j = 0; j < 3; ++j) {
                const float* row = vf + 4 * (((int64_t)i * ng + j) * ng) + c;
                const float g[3] = {row[0], row[4], row[8]};
                float zw = 0.f, ze = 0.f, zd = 0.f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    zw = fmaf(A.w[2][k], g[k], zw);
                    ze = fmaf(A.e[2][k], g[k], ze);
                    zd = fmaf(A.d[2][k], g[k], zd);
                }
                const float wj = pick3(A.w[1], j), ej = pick3(A.e[1], j), dj = pick3(A.d[1], j);
                yww = fmaf(wj, zw, yww);
                yew = fmaf(ej, zw, yew);
                ydw = fmaf(dj, zw, ydw);
                ywe = fmaf(wj, ze, ywe);
                ywd = fmaf(wj, zd, ywd);
            }
            const float wi = pick3(A.w[0], i), ei = pick3(A.e[0], i), di = pick3(A.d[0], i);
            sv = fmaf(wi, yww, sv);
            cx = fmaf(ei, yww, cx);
            gx = fmaf(di, yww, gx);
            cy = fmaf(wi, yew, cy);
            gy = fmaf(wi, ydw, gy);
            cz = fmaf(wi, ywe, cz);
            gz = fmaf(wi, ywd, gz);
        }
        v[c] = sv;
        C[c * 3 + 0] = cx * c4;
        C[c * 3 + 1] = cy * c4;
        C[c * 3 + 2] = cz * c4;
        G[c * 3 + 0] = gx * inv_dx;
        G[c * 3 + 1] = gy * inv_dx;
        G[c * 3 + 2] = gz * inv_dx;
    }
}

}  // namespace mpm
