/*
 * mpm_oracle.c -- CPU restatement of the explicit MLS/APIC-MPM sand substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * reduced_mpm_stress_training_amd/csrc: it is compiled into oracle/_build/
 * liboracle.so and may be loaded only by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py.  The product path never links,
 * loads or calls it.
 *
 * One source, two precisions: compiled once with -DORACLE_F64 (every value
 * double) and once without (every value float, the reference's Warp `float`).
 * Exported symbols carry the suffix _f64 / _f32.
 *
 * Every function cites the reference file:line it restates
 * (zeshunzong/reduced_mpm_stress_training, engine_stress/).
 *
 * Pinning (tests/test_oracle_*.py): weights against partition-of-unity and
 * moment identities, the SVD against LAPACK (numpy.linalg.svd), the return map
 * and the stress against closed forms on diagonal F and rotation invariance,
 * P2G against mass/momentum conservation, G2P against affine-field
 * reproduction, the whole step against free fall and against the independent
 * numpy restatement oracle/mpm_oracle_np.py (LAPACK SVD, np.add.at scatter).
 * Warp's own svd3 rounding and its sign convention for det F <= 0 are not
 * available here (warp-lang absent): parity unpinned for those two points.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_F64
typedef double real;
#define SFX(n) n##_f64
#define R_LOG log
#define R_EXP exp
#define R_SQRT sqrt
#define R_FABS fabs
#define R_ACOS acos
#define R_SIN sin
#define R_COS cos
#define ORACLE_JACOBI_TOL 1e-30
#else
typedef float real;
#define SFX(n) n##_f32
#define R_LOG logf
#define R_EXP expf
#define R_SQRT sqrtf
#define R_FABS fabsf
#define R_ACOS acosf
#define R_SIN sinf
#define R_COS cosf
#define ORACLE_JACOBI_TOL 1e-14f
#endif

/* Model scalars; the reference keeps them in the Warp struct MPMModelStruct
 * (warp_utils.py:61-81) as 32-bit floats.  They travel here as double and
 * are rounded to `real` on entry, as Warp rounds the Python floats. */
typedef struct {
    int32_t n_grid;
    int32_t material;     /* 0 jelly, 1 metal, 2 sand (mpm_solver_warp.py:90-94) */
    double dx, inv_dx;
    double mu, lam, alpha;
    double gravity[3];    /* acceleration vector; reference: (0, -g, 0) */
    int32_t hardening;    /* metal: MPMModelStruct.hardening (warp_utils.py:80) */
    double xi;            /* metal: MPMModelStruct.xi (warp_utils.py:81) */
} OracleParams;

/* Grid collider (warp_utils.py:126-151, Dirichlet_collider).
 * kind 0 surface half-space (mpm_solver_warp.py:407-497), 1 bounding box
 * (:499-525), 2 velocity on an infinite plane (:527-566), 3 column moving on
 * a circle (:568-605), 4 rotating disk (:607-664), 5 velocity on a cube
 * (:666-693), 6 sphere leader (:695-720). */
typedef struct {
    int32_t kind;
    int32_t surface_type; /* 0 sticky, 1 slip, 2 frictional */
    double point[3];
    double normal[3];     /* already normalised (mpm_solver_warp.py:411-412) */
    double start_time, end_time, friction;
    double velocity[3];
    double threshold, radius, height, width, v_scale;
    double x_unit[3], y_unit[3];
    int32_t reset;
} OracleCollider;

/* ------------------------------------------------------------------ */
/* 3x3 helpers (row-major, as Warp mat33 and torch (N,3,3))             */
/* ------------------------------------------------------------------ */
static void mm3(const real* A, const real* B, real* C) {
    real T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = A[i * 3 + 0] * B[0 * 3 + j];
            s = s + A[i * 3 + 1] * B[1 * 3 + j];
            s = s + A[i * 3 + 2] * B[2 * 3 + j];
            T[i * 3 + j] = s;
        }
    memcpy(C, T, sizeof(T));
}
static void mmT3(const real* A, const real* B, real* C) { /* C = A * B^T */
    real T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = A[i * 3 + 0] * B[j * 3 + 0];
            s = s + A[i * 3 + 1] * B[j * 3 + 1];
            s = s + A[i * 3 + 2] * B[j * 3 + 2];
            T[i * 3 + j] = s;
        }
    memcpy(C, T, sizeof(T));
}

/* ------------------------------------------------------------------ */
/* R11: 3x3 SVD.  The reference calls Warp's builtin wp.svd3             */
/* (mpm_utils.py:92,297), itself after McAdams et al. 2011, "Computing   */
/* the SVD of 3x3 matrices with minimal branching and elementary         */
/* floating point operations" (UW-Madison CS TR1690): Jacobi             */
/* eigenanalysis of A^T A, columns of AV sorted by decreasing norm, then */
/* a Givens QR of AV.  Restated here with exact Jacobi rotations run to  */
/* convergence.  Convention: U, V rotations (det +1), s0 >= s1 >= |s2|,  */
/* s2 carries the sign of det A.                                         */
/* ------------------------------------------------------------------ */
static void jacobi_rotate(real* S, real* V, int p, int q) {
    real apq = S[p * 3 + q];
    if (apq == (real)0) return;
    real app = S[p * 3 + p], aqq = S[q * 3 + q];
    real theta = (aqq - app) / ((real)2 * apq);
    real t = (theta >= 0 ? (real)1 : (real)-1) /
             (R_FABS(theta) + R_SQRT(theta * theta + (real)1));
    real c = (real)1 / R_SQRT(t * t + (real)1);
    real s = t * c;
    /* S <- J^T S J with J = I except J[p][p]=c, J[p][q]=s, J[q][p]=-s, J[q][q]=c */
    for (int k = 0; k < 3; ++k) { /* columns p,q */
        real skp = S[k * 3 + p], skq = S[k * 3 + q];
        S[k * 3 + p] = c * skp - s * skq;
        S[k * 3 + q] = s * skp + c * skq;
    }
    for (int k = 0; k < 3; ++k) { /* rows p,q */
        real spk = S[p * 3 + k], sqk = S[q * 3 + k];
        S[p * 3 + k] = c * spk - s * sqk;
        S[q * 3 + k] = s * spk + c * sqk;
    }
    S[p * 3 + q] = S[q * 3 + p] = 0;
    for (int k = 0; k < 3; ++k) {
        real vkp = V[k * 3 + p], vkq = V[k * 3 + q];
        V[k * 3 + p] = c * vkp - s * vkq;
        V[k * 3 + q] = s * vkp + c * vkq;
    }
}

static void givens_rows(real* B, real* U, int p, int q, int col) {
    /* zero B[q][col] against B[p][col]; accumulate U <- U * G^T */
    real a = B[p * 3 + col], b = B[q * 3 + col];
    real r = R_SQRT(a * a + b * b);
    real c = 1, s = 0;
    if (r > (real)0) { c = a / r; s = b / r; }
    for (int k = 0; k < 3; ++k) {
        real bp = B[p * 3 + k], bq = B[q * 3 + k];
        B[p * 3 + k] = c * bp + s * bq;
        B[q * 3 + k] = -s * bp + c * bq;
    }
    for (int k = 0; k < 3; ++k) {
        real up = U[k * 3 + p], uq = U[k * 3 + q];
        U[k * 3 + p] = c * up + s * uq;
        U[k * 3 + q] = -s * up + c * uq;
    }
}

void SFX(oracle_svd3)(const real* A, real* U, real* sig, real* V) {
    real S[9];
    mm3((real[9]){A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]}, A, S);
    real Vm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 30; ++sweep) {
        real off = S[1] * S[1] + S[2] * S[2] + S[5] * S[5];
        real dia = S[0] * S[0] + S[4] * S[4] + S[8] * S[8];
        if (off <= dia * ORACLE_JACOBI_TOL || off == (real)0) break;
        jacobi_rotate(S, Vm, 0, 1);
        jacobi_rotate(S, Vm, 0, 2);
        jacobi_rotate(S, Vm, 1, 2);
    }
    real B[9];
    mm3(A, Vm, B);
    /* sort columns of B (and V) by decreasing squared norm; a swap negates
     * one column so that V stays a rotation */
    for (int pass = 0; pass < 3; ++pass) {
        int i = (pass == 1) ? 1 : 0, j = i + 1;
        real ni = B[i] * B[i] + B[3 + i] * B[3 + i] + B[6 + i] * B[6 + i];
        real nj = B[j] * B[j] + B[3 + j] * B[3 + j] + B[6 + j] * B[6 + j];
        if (ni < nj) {
            for (int k = 0; k < 3; ++k) {
                real t = B[k * 3 + i]; B[k * 3 + i] = B[k * 3 + j]; B[k * 3 + j] = -t;
                t = Vm[k * 3 + i]; Vm[k * 3 + i] = Vm[k * 3 + j]; Vm[k * 3 + j] = -t;
            }
        }
    }
    real Um[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    givens_rows(B, Um, 0, 1, 0);
    givens_rows(B, Um, 0, 2, 0);
    givens_rows(B, Um, 1, 2, 1);
    sig[0] = B[0]; sig[1] = B[4]; sig[2] = B[8];
    memcpy(U, Um, sizeof(Um));
    memcpy(V, Vm, sizeof(Vm));
}

/* ------------------------------------------------------------------ */
/* R5: quadratic B-spline weights, mpm_utils.py:196-209 / :244-257.      */
/* w[d][node], dw[d][node] -- Warp's mat33(v0,v1,v2) takes the vectors   */
/* as columns, so w[d,i] is the weight of node i along axis d.           */
/* ------------------------------------------------------------------ */
void SFX(oracle_weights)(const real* x, real inv_dx, int* base, real* fx,
                         real* w, real* dw) {
    for (int d = 0; d < 3; ++d) {
        real gp = x[d] * inv_dx;
        base[d] = (int)(gp - (real)0.5);  /* truncation, mpm_utils.py:197 */
        fx[d] = gp - (real)base[d];
        real wa = (real)1.5 - fx[d], wb = fx[d] - (real)1.0, wc = fx[d] - (real)0.5;
        w[d * 3 + 0] = wa * wa * (real)0.5;
        w[d * 3 + 1] = ((real)0 - wb * wb) + (real)0.75;
        w[d * 3 + 2] = wc * wc * (real)0.5;
        dw[d * 3 + 0] = fx[d] - (real)1.5;
        dw[d * 3 + 1] = (real)-2.0 * (fx[d] - (real)1.0);
        dw[d * 3 + 2] = fx[d] - (real)0.5;
    }
}

/* compute_dweight, mpm_utils.py:54-57 */
static void dweight(const real* w, const real* dw, int i, int j, int k,
                    real inv_dx, real* g) {
    g[0] = dw[0 * 3 + i] * w[1 * 3 + j] * w[2 * 3 + k] * inv_dx;
    g[1] = w[0 * 3 + i] * dw[1 * 3 + j] * w[2 * 3 + k] * inv_dx;
    g[2] = w[0 * 3 + i] * w[1 * 3 + j] * dw[2 * 3 + k] * inv_dx;
}

/* ------------------------------------------------------------------ */
/* R10: sand_return_mapping, mpm_utils.py:87-112                         */
/* ------------------------------------------------------------------ */
void SFX(oracle_sand_return_map)(const real* Ftr, real mu, real lam,
                                 real alpha, real* Fe) {
    real U[9], V[9], s[3];
    SFX(oracle_svd3)(Ftr, U, s, V);
    real eps[3];
    for (int d = 0; d < 3; ++d) {
        real a = R_FABS(s[d]);
        eps[d] = R_LOG(a > (real)1e-14 ? a : (real)1e-14);
    }
    real tr = eps[0] + eps[1] + eps[2];
    real eh[3] = {eps[0] - tr / (real)3, eps[1] - tr / (real)3, eps[2] - tr / (real)3};
    real ehn = R_SQRT(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
    real dg = ehn + ((real)3 * lam + (real)2 * mu) / ((real)2 * mu) * tr * alpha;
    if (dg <= (real)0) { memcpy(Fe, Ftr, 9 * sizeof(real)); return; }
    real sn[3];
    if (tr > (real)0) {
        sn[0] = sn[1] = sn[2] = (real)1;   /* F_e = U V^T, :104-105 */
    } else {
        for (int d = 0; d < 3; ++d) sn[d] = R_EXP(eps[d] - eh[d] * (dg / ehn));
    }
    real US[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) US[i * 3 + j] = U[i * 3 + j] * sn[j];
    mmT3(US, V, Fe);
}

/* ------------------------------------------------------------------ */
/* R12: kirchoff_stress_drucker_prager, mpm_utils.py:45-52, applied with */
/* a fresh SVD of the elastic F as at mpm_utils.py:292-306.              */
/* ------------------------------------------------------------------ */
void SFX(oracle_dp_stress)(const real* F, real mu, real lam, real* tau) {
    real U[9], V[9], s[3];
    SFX(oracle_svd3)(F, U, s, V);
    real ls = R_LOG(s[0]) + R_LOG(s[1]) + R_LOG(s[2]);
    real c[3];
    for (int d = 0; d < 3; ++d)
        c[d] = (real)2 * mu * R_LOG(s[d]) * ((real)1 / s[d]) + lam * ls * ((real)1 / s[d]);
    real UC[9], UCV[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) UC[i * 3 + j] = U[i * 3 + j] * c[j];
    mmT3(UC, V, UCV);       /* U diag(c) V^T   */
    mmT3(UCV, F, tau);      /* ... F^T          */
}

/* ------------------------------------------------------------------ */
/* Metal: von_mises_return_mapping, mpm_utils.py:59-84.  Updates *ys      */
/* (model.yield_stress[p]) when hardening == 1 and the particle yields.  */
/* ------------------------------------------------------------------ */
void SFX(oracle_von_mises)(const real* Ftr, real mu, real lam, real* ys,
                           int hardening, real xi, real* Fe) {
    real U[9], V[9], so[3];
    SFX(oracle_svd3)(Ftr, U, so, V);
    real eps[3];
    for (int d = 0; d < 3; ++d) eps[d] = R_LOG(so[d] > (real)0.01 ? so[d] : (real)0.01);
    real temp = (eps[0] + eps[1] + eps[2]) / (real)3;
    real t[3];
    for (int d = 0; d < 3; ++d) t[d] = (real)2 * mu * eps[d] + lam * (eps[0] + eps[1] + eps[2]);
    real sum_tau = t[0] + t[1] + t[2];
    real cond[3] = {t[0] - sum_tau / (real)3, t[1] - sum_tau / (real)3, t[2] - sum_tau / (real)3};
    real cn = R_SQRT(cond[0] * cond[0] + cond[1] * cond[1] + cond[2] * cond[2]);
    if (cn > *ys) {
        real eh[3] = {eps[0] - temp, eps[1] - temp, eps[2] - temp};
        real ehn = R_SQRT(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]) + (real)1e-6;
        real dg = ehn - *ys / ((real)2 * mu);
        real sn[3];
        for (int d = 0; d < 3; ++d) sn[d] = R_EXP(eps[d] - (dg / ehn) * eh[d]);
        real US[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) US[i * 3 + j] = U[i * 3 + j] * sn[j];
        mmT3(US, V, Fe);
        if (hardening == 1) *ys = *ys + (real)2 * mu * xi * dg;
    } else {
        memcpy(Fe, Ftr, 9 * sizeof(real));
    }
}

/* Metal stress: kirchoff_stress_StVK, mpm_utils.py:36-43, on a fresh SVD */
/* of F (mpm_utils.py:297, 301-302): tau = U diag(2 mu ln s + lam sum) V^T */
void SFX(oracle_stvk_stress)(const real* F, real mu, real lam, real* tau) {
    real U[9], V[9], s[3];
    SFX(oracle_svd3)(F, U, s, V);
    for (int d = 0; d < 3; ++d) s[d] = s[d] > (real)0.01 ? s[d] : (real)0.01;
    real ls = R_LOG(s[0]) + R_LOG(s[1]) + R_LOG(s[2]);
    real c[3];
    for (int d = 0; d < 3; ++d) c[d] = (real)2 * mu * R_LOG(s[d]) + lam * ls;
    real UC[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) UC[i * 3 + j] = U[i * 3 + j] * c[j];
    mmT3(UC, V, tau);
}

/* Jelly stress: kirchoff_stress_FCR, mpm_utils.py:29-34, with R = U V^T  */
/* of a fresh SVD of F and J = det F (mpm_utils.py:292, 297-299).         */
void SFX(oracle_fcr_stress)(const real* F, real mu, real lam, real* tau) {
    real U[9], V[9], s[3];
    SFX(oracle_svd3)(F, U, s, V);
    real R[9];
    mmT3(U, V, R);
    real D[9];
    for (int k = 0; k < 9; ++k) D[k] = F[k] - R[k];
    real J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
             F[2] * (F[3] * F[7] - F[4] * F[6]);
    mmT3(D, F, tau);
    for (int k = 0; k < 9; ++k) tau[k] = (real)2 * mu * tau[k];
    for (int d = 0; d < 3; ++d) tau[d * 4] = tau[d * 4] + lam * J * (J - (real)1);
}

/* ------------------------------------------------------------------ */
/* R6: p2g_apic_stress_given, mpm_utils.py:190-227.  Sequential in       */
/* particle order (the reference's float atomics have no fixed order).   */
/* grid_mv is v_in (n^3 x 3), grid_m is m (n^3).                          */
/* ------------------------------------------------------------------ */
void SFX(oracle_p2g)(const OracleParams* P, int64_t n, const real* x,
                     const real* v, const real* C, const real* tau,
                     const real* mass, const real* vol, real* grid_m,
                     real* grid_mv, double dt_in) {
    const int ng = P->n_grid;
    const real dx = (real)P->dx, inv_dx = (real)P->inv_dx, dt = (real)dt_in;
    for (int64_t p = 0; p < n; ++p) {
        int base[3]; real fx[3], w[9], dw[9];
        SFX(oracle_weights)(x + 3 * p, inv_dx, base, fx, w, dw);
        const real* Cp = C + 9 * p;
        const real* tp = tau + 9 * p;
        const real mp = mass[p], vp = vol[p];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    real dpos[3] = {((real)i - fx[0]) * dx, ((real)j - fx[1]) * dx,
                                    ((real)k - fx[2]) * dx};
                    int64_t node = ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k);
                    real weight = w[0 * 3 + i] * w[1 * 3 + j] * w[2 * 3 + k];
                    real g[3];
                    dweight(w, dw, i, j, k, inv_dx, g);
                    real wm = weight * mp;
                    for (int a = 0; a < 3; ++a) {
                        real cd = Cp[a * 3 + 0] * dpos[0] + Cp[a * 3 + 1] * dpos[1] + Cp[a * 3 + 2] * dpos[2];
                        grid_mv[node * 3 + a] += wm * (v[3 * p + a] + cd);
                    }
                    for (int a = 0; a < 3; ++a) {
                        real f = ((-vp) * tp[a * 3 + 0]) * g[0] + ((-vp) * tp[a * 3 + 1]) * g[1] +
                                 ((-vp) * tp[a * 3 + 2]) * g[2];
                        grid_mv[node * 3 + a] += dt * f;
                    }
                    grid_m[node] += wm;
                }
    }
}

/* ------------------------------------------------------------------ */
/* R7: grid_normalization_and_gravity, mpm_utils.py:229-237 (with the    */
/* gravity vector generalising (0,-g,0)); then R8/R9 colliders in        */
/* registration order (mpm_solver_warp.py:182-185):                      */
/*   surface collider  mpm_solver_warp.py:429-449 (every type writes 0), */
/*   bounding box      mpm_solver_warp.py:505-523.                       */
/* grid_v (v_out) must be zero on entry (zero_grid, mpm_utils.py:120).   */
/* ------------------------------------------------------------------ */
/* The active branch of the velocity colliders (kinds 2-6) on one node:    */
/* off = x_node - point, dot = off . normal.                               */
static void moving_collider(const OracleCollider* cl, const real* off, real dot, real* vo) {
    const real vs = (real)cl->v_scale;
    if (cl->kind == 2) {                       /* mpm_solver_warp.py:550-551 */
        if (R_FABS(dot) <= (real)cl->threshold) {
            for (int d = 0; d < 3; ++d) vo[d] = (real)cl->velocity[d];
        }
    } else if (cl->kind == 3) {                /* :584-592 */
        if (off[0] * off[0] + off[2] * off[2] < (real)cl->radius * (real)cl->radius &&
            R_FABS(off[1]) < (real)cl->height) {
            real Z[3] = {(real)cl->point[0] - (real)0.5, (real)cl->point[1], (real)cl->point[2] - (real)0.5};
            real dp = Z[0] / R_SQRT(Z[0] * Z[0] + Z[1] * Z[1] + Z[2] * Z[2]);
            real th = R_ACOS(dp);
            if (Z[2] < (real)0) th = -th;
            vo[0] = -vs * R_SIN(th);
            vo[1] = 0;
            vo[2] = vs * R_COS(th);
        }
    } else if (cl->kind == 4) {                /* :642-657 */
        if (R_FABS(dot) < (real)cl->width) {
            real r[3] = {0, off[1], off[2]};
            real rl = R_SQRT(r[1] * r[1] + r[2] * r[2]);
            if (rl < (real)cl->radius) {
                if (rl > (real)0) {
                    real dp = (r[1] * (real)cl->x_unit[1] + r[2] * (real)cl->x_unit[2]) / rl;
                    real rdoty = r[1] * (real)cl->y_unit[1] + r[2] * (real)cl->y_unit[2];
                    real th = R_ACOS(dp);
                    if (rdoty < (real)0) th = -th;
                    real sx = -rl * R_SIN(th) * vs, sy = rl * R_COS(th) * vs;
                    for (int d = 0; d < 3; ++d) vo[d] = sx * (real)cl->x_unit[d] + sy * (real)cl->y_unit[d];
                } else {
                    /* on the axis: the reference divides 0 by 0; the limit of the rotation is 0 */
                    vo[0] = vo[1] = vo[2] = 0;
                }
            }
        }
    } else if (cl->kind == 5) {                /* :683-684 */
        if (R_FABS(off[0]) < (real)cl->width && R_FABS(off[1]) < (real)cl->width && R_FABS(off[2]) < (real)cl->width) {
            for (int d = 0; d < 3; ++d) vo[d] = (real)cl->velocity[d];
        }
    } else if (cl->kind == 6) {                /* :712-713 */
        if (R_SQRT(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]) < (real)cl->radius) {
            for (int d = 0; d < 3; ++d) vo[d] = (real)cl->velocity[d];
        }
    }
}

/* The colliders' grid kernels over every node, in registration order */
/* (mpm_solver_warp.py:182-185, :222-225).                             */
static void apply_colliders(const OracleParams* P, real* grid_v, const OracleCollider* cols, int ncol,
                            real time, real dt) {
    const int ng = P->n_grid;
    const real dx = (real)P->dx;
    /* each collider's kernel over every node, in registration order */
    for (int c = 0; c < ncol; ++c) {
        const OracleCollider* cl = cols + c;
        const int active = time >= (real)cl->start_time && time < (real)cl->end_time;
        /* outside the window only the reset branches of kinds 2, 4, 5 act */
        int reset = 0;
        if (!active && cl->reset == 1) {
            if (cl->kind == 2) reset = time < (real)cl->end_time + (real)1.5 * dt * (real)10.0;
            if (cl->kind == 4 || cl->kind == 5) reset = time < (real)cl->end_time + (real)1.5 * dt;
        }
        if (!active && !reset) continue;
        for (int gx = 0; gx < ng; ++gx)
            for (int gy = 0; gy < ng; ++gy)
                for (int gz = 0; gz < ng; ++gz) {
                    int64_t node = ((int64_t)gx * ng + gy) * ng + gz;
                    real* vo = grid_v + node * 3;
                    real off[3] = {(real)gx * dx - (real)cl->point[0],
                                   (real)gy * dx - (real)cl->point[1],
                                   (real)gz * dx - (real)cl->point[2]};
                    real dot = off[0] * (real)cl->normal[0] + off[1] * (real)cl->normal[1] +
                               off[2] * (real)cl->normal[2];
                    if (reset) {
                        /* kind 2 zeroes its slab (:557-558), kinds 4, 5 every node (:661, :687) */
                        if (cl->kind != 2 || R_FABS(dot) <= (real)cl->threshold) vo[0] = vo[1] = vo[2] = 0;
                    } else if (cl->kind == 0) {
                        if (dot < (real)0) { vo[0] = vo[1] = vo[2] = 0; }
                    } else if (cl->kind == 1) {
                        const int pad = 3;
                        const int idx[3] = {gx, gy, gz};
                        for (int d = 0; d < 3; ++d) {
                            if (idx[d] < pad && vo[d] < (real)0) vo[d] = 0;
                            if (idx[d] >= ng - pad && vo[d] > (real)0) vo[d] = 0;
                        }
                    } else {
                        moving_collider(cl, off, dot, vo);
                    }
                }
    }
}

void SFX(oracle_grid_update)(const OracleParams* P, const real* grid_m,
                             const real* grid_mv, real* grid_v,
                             const OracleCollider* cols, int ncol,
                             double time_in, double dt_in) {
    const int ng = P->n_grid;
    const real dt = (real)dt_in, time = (real)time_in;
    const real gv[3] = {(real)P->gravity[0], (real)P->gravity[1], (real)P->gravity[2]};
    const int64_t nn = (int64_t)ng * ng * ng;
    for (int64_t node = 0; node < nn; ++node) {
        real m = grid_m[node];
        if (m > (real)1e-15) {
            real im = (real)1.0 / m;
            for (int a = 0; a < 3; ++a)
                grid_v[node * 3 + a] = grid_mv[node * 3 + a] * im + dt * gv[a];
        }
    }
    apply_colliders(P, grid_v, cols, ncol, time, dt);
}

/* ------------------------------------------------------------------ */
/* R13: g2p_apic_and_stress, mpm_utils.py:241-306, with the material     */
/* switch of :284-304 (sand R10/R12, metal von Mises + StVK, jelly FCR). */
/* x is updated in place; F holds F_e on exit; ys (metal) is the         */
/* particle's yield stress.                                              */
/* ------------------------------------------------------------------ */
static void g2p_one(const OracleParams* P, const real* grid_v, real* x, real* v,
                    real* C, real* F, real* tau, real* ys, real dt) {
    const int ng = P->n_grid;
    const real inv_dx = (real)P->inv_dx, mu = (real)P->mu, lam = (real)P->lam,
               alpha = (real)P->alpha;
    int base[3]; real fx[3], w[9], dw[9];
    SFX(oracle_weights)(x, inv_dx, base, fx, w, dw);
    real nv[3] = {0, 0, 0}, nC[9] = {0}, nF[9] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                int64_t node = ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k);
                real dpos[3] = {(real)i - fx[0], (real)j - fx[1], (real)k - fx[2]};
                real weight = w[0 * 3 + i] * w[1 * 3 + j] * w[2 * 3 + k];
                const real* g = grid_v + node * 3;
                real sC = weight * inv_dx * (real)4.0;
                real gw[3];
                dweight(w, dw, i, j, k, inv_dx, gw);
                for (int a = 0; a < 3; ++a) {
                    nv[a] = nv[a] + g[a] * weight;
                    for (int b = 0; b < 3; ++b) {
                        nC[a * 3 + b] = nC[a * 3 + b] + (g[a] * dpos[b]) * sC;
                        nF[a * 3 + b] = nF[a * 3 + b] + g[a] * gw[b];
                    }
                }
            }
    for (int a = 0; a < 3; ++a) {
        v[a] = nv[a];
        x[a] = x[a] + dt * nv[a];
    }
    memcpy(C, nC, sizeof(nC));
    real A[9];
    for (int a = 0; a < 9; ++a) A[a] = nF[a] * dt;
    A[0] = A[0] + (real)1; A[4] = A[4] + (real)1; A[8] = A[8] + (real)1;
    real Ftr[9];
    mm3(A, F, Ftr);
    if (P->material == 1) {
        SFX(oracle_von_mises)(Ftr, mu, lam, ys, P->hardening, (real)P->xi, F);
        SFX(oracle_stvk_stress)(F, mu, lam, tau);
    } else if (P->material == 2) {
        SFX(oracle_sand_return_map)(Ftr, mu, lam, alpha, F);
        SFX(oracle_dp_stress)(F, mu, lam, tau);
    } else {
        memcpy(F, Ftr, sizeof(Ftr));
        SFX(oracle_fcr_stress)(F, mu, lam, tau);
    }
}

void SFX(oracle_g2p)(const OracleParams* P, int64_t n, const real* grid_v,
                     real* x, real* v, real* C, real* F, real* tau, real* yield,
                     double dt_in, int nthreads) {
    const real dt = (real)dt_in;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t p = 0; p < n; ++p) {
        real ys0 = 0;
        real* ys = yield ? yield + p : &ys0;
        g2p_one(P, grid_v, x + 3 * p, v + 3 * p, C + 9 * p, F + 9 * p, tau + 9 * p, ys, dt);
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------ */
/* R15: one substep p2g2p, mpm_solver_warp.py:156-203, in its order:     */
/* (yield: the per-particle yield stress array, metal only; may be NULL) */
/*   update_x_from_u (mpm_utils.py:4-7), zero_grid (:115-123),           */
/*   P2G, grid update + colliders, G2P, update_u_from_x (:12-15).        */
/* F_disp = F - I (:16-19) is a derived output and is not kept here.     */
/* x is scratch (n x 3); grid arrays are scratch of n_grid^3 entries.    */
/* nthreads > 1 runs P2G with OpenMP atomics (order then varies, as in   */
/* the reference) and G2P in parallel.                                   */
/* ------------------------------------------------------------------ */
int SFX(oracle_step)(const OracleParams* P, int64_t n, const real* x0, real* u,
                     real* v, real* C, real* F, real* tau, const real* mass,
                     const real* vol, real* yield, real* x, real* grid_m, real* grid_mv,
                     real* grid_v, const OracleCollider* cols, int ncol,
                     double time, double dt, int nthreads) {
    const int ng = P->n_grid;
    const int64_t nn = (int64_t)ng * ng * ng;
    const real dtr = (real)dt;
    for (int64_t p = 0; p < 3 * n; ++p) x[p] = x0[p] + u[p];
    memset(grid_m, 0, nn * sizeof(real));
    memset(grid_mv, 0, 3 * nn * sizeof(real));
    memset(grid_v, 0, 3 * nn * sizeof(real));
    if (nthreads <= 1) {
        SFX(oracle_p2g)(P, n, x, v, C, tau, mass, vol, grid_m, grid_mv, dt);
    } else {
#ifdef _OPENMP
        const real dx = (real)P->dx, inv_dx = (real)P->inv_dx;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            int base[3]; real fx[3], w[9], dw[9];
            SFX(oracle_weights)(x + 3 * p, inv_dx, base, fx, w, dw);
            const real* Cp = C + 9 * p;
            const real* tp = tau + 9 * p;
            const real mp = mass[p], vp = vol[p];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        real dpos[3] = {((real)i - fx[0]) * dx, ((real)j - fx[1]) * dx,
                                        ((real)k - fx[2]) * dx};
                        int64_t node = ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k);
                        real weight = w[0 * 3 + i] * w[1 * 3 + j] * w[2 * 3 + k];
                        real g[3];
                        dweight(w, dw, i, j, k, inv_dx, g);
                        real wm = weight * mp;
                        for (int a = 0; a < 3; ++a) {
                            real cd = Cp[a * 3 + 0] * dpos[0] + Cp[a * 3 + 1] * dpos[1] + Cp[a * 3 + 2] * dpos[2];
                            real f = ((-vp) * tp[a * 3 + 0]) * g[0] + ((-vp) * tp[a * 3 + 1]) * g[1] +
                                     ((-vp) * tp[a * 3 + 2]) * g[2];
                            real add = wm * (v[3 * p + a] + cd) + dtr * f;
#pragma omp atomic
                            grid_mv[node * 3 + a] += add;
                        }
#pragma omp atomic
                        grid_m[node] += wm;
                    }
        }
#else
        SFX(oracle_p2g)(P, n, x, v, C, tau, mass, vol, grid_m, grid_mv, dt);
#endif
    }
    SFX(oracle_grid_update)(P, grid_m, grid_mv, grid_v, cols, ncol, time, dt);
    SFX(oracle_g2p)(P, n, grid_v, x, v, C, F, tau, yield, dt, nthreads);
    for (int64_t p = 0; p < 3 * n; ++p) u[p] = x[p] - x0[p];
    return 0;
}

/* ------------------------------------------------------------------ */
/* Reduced-order sampled substep (mpm_solver_warp.py:207-227).           */
/* ------------------------------------------------------------------ */

/* update_neighboring_cells (mpm_utils.py:318-335, lo = -2, hi = 2) and   */
/* update_neighboring_cells_for_grid_v (:337-354, lo = -2, hi = 4).      */
void SFX(oracle_mark_cells)(const OracleParams* P, const real* x, const int32_t* samples, int64_t ns, int lo,
                            int hi, int32_t* flags) {
    const int ng = P->n_grid;
    const real inv_dx = (real)P->inv_dx;
    for (int64_t s = 0; s < ns; ++s) {
        const real* xp = x + 3 * (int64_t)samples[s];
        int b[3];
        for (int d = 0; d < 3; ++d) b[d] = (int)(xp[d] * inv_dx - (real)0.5);
        for (int i = lo; i <= hi; ++i)
            for (int j = lo; j <= hi; ++j)
                for (int k = lo; k <= hi; ++k) {
                    int gx = b[0] + i, gy = b[1] + j, gz = b[2] + k;
                    if (0 <= gx && gx < ng && 0 <= gy && gy < ng && 0 <= gz && gz < ng)
                        flags[((int64_t)gx * ng + gy) * ng + gz] = 1;
                }
    }
}

/* p2g_sample_all (mpm_utils.py:356-410).  F holds F_disp + I on entry    */
/* (F_disp = F - I is formed and added back as the reference's arrays     */
/* hold it) and is written for the particles in flagged cells.           */
void SFX(oracle_p2g_sample_all)(const OracleParams* P, int64_t n, const real* x, const real* v, real* F,
                                const real* mass, const real* density, const int32_t* flags, real* grid_m,
                                real* grid_mv, double dt_in) {
    const int ng = P->n_grid;
    const real inv_dx = (real)P->inv_dx, dt = (real)dt_in, mu = (real)P->mu, lam = (real)P->lam;
    for (int64_t p = 0; p < n; ++p) {
        int base[3]; real fx[3], w[9], dw[9];
        SFX(oracle_weights)(x + 3 * p, inv_dx, base, fx, w, dw);
        if (base[0] < 0 || base[1] < 0 || base[2] < 0 || base[0] > ng - 3 || base[1] > ng - 3 || base[2] > ng - 3)
            continue;
        if (flags[((int64_t)base[0] * ng + base[1]) * ng + base[2]] <= 0) continue;
        real* Fp = F + 9 * p;
        for (int k = 0; k < 9; ++k) {
            real id = (k % 4 == 0) ? (real)1 : (real)0;
            Fp[k] = (Fp[k] - id) + id;
        }
        real U[9], V[9], sig[3], tau[9];
        SFX(oracle_svd3)(Fp, U, sig, V);
        real J = (real)1;
        for (int d = 0; d < 3; ++d) J = J * sig[d];
        if (P->material == 2) {
            real ls = R_LOG(sig[0]) + R_LOG(sig[1]) + R_LOG(sig[2]);
            real c[3], UC[9], UCV[9];
            for (int d = 0; d < 3; ++d)
                c[d] = (real)2 * mu * R_LOG(sig[d]) * ((real)1 / sig[d]) + lam * ls * ((real)1 / sig[d]);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) UC[i * 3 + j] = U[i * 3 + j] * c[j];
            mmT3(UC, V, UCV);
            mmT3(UCV, Fp, tau);
        } else {
            real R[9], D[9];
            mmT3(U, V, R);
            for (int k = 0; k < 9; ++k) D[k] = Fp[k] - R[k];
            mmT3(D, Fp, tau);
            for (int k = 0; k < 9; ++k) tau[k] = (real)2 * mu * tau[k];
            for (int d = 0; d < 3; ++d) tau[d * 4] = tau[d * 4] + lam * J * (J - (real)1);
        }
        const real mp = mass[p], nvol = -(mass[p] / density[p]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int64_t node = ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k);
                    real weight = w[0 * 3 + i] * w[1 * 3 + j] * w[2 * 3 + k];
                    real g[3];
                    dweight(w, dw, i, j, k, inv_dx, g);
                    for (int a = 0; a < 3; ++a) {
                        grid_mv[node * 3 + a] += weight * mp * v[3 * p + a];
                        real f = (nvol * tau[a * 3 + 0]) * g[0] + (nvol * tau[a * 3 + 1]) * g[1] +
                                 (nvol * tau[a * 3 + 2]) * g[2];
                        grid_mv[node * 3 + a] += dt * f;
                    }
                    grid_m[node] += weight * mp;
                }
    }
}

/* grid_normalization_and_gravity_sample_all (mpm_utils.py:413-421) on the */
/* flagged nodes, then the colliders; grid_v is zero on entry (zero_grid). */
void SFX(oracle_grid_update_sample)(const OracleParams* P, const int32_t* flags, const real* grid_m,
                                    const real* grid_mv, real* grid_v, const OracleCollider* cols, int ncol,
                                    double time_in, double dt_in) {
    const int ng = P->n_grid;
    const real dt = (real)dt_in;
    const int64_t nn = (int64_t)ng * ng * ng;
    for (int64_t node = 0; node < nn; ++node) {
        if (flags[node] <= 0) continue;
        real m = grid_m[node];
        if (m > (real)1e-10) {
            for (int a = 0; a < 3; ++a) grid_v[node * 3 + a] = grid_mv[node * 3 + a] / m + dt * (real)P->gravity[a];
        }
    }
    apply_colliders(P, grid_v, cols, ncol, (real)time_in, dt);
}

/* g2p_sample (mpm_utils.py:423-471): PIC gather, x += dt v, return map of */
/* (I + dt grad v) F, u = x - x0.  Metal uses yield[s] of the sample slot  */
/* s (the reference passes its thread index, :464).                        */
void SFX(oracle_g2p_sample)(const OracleParams* P, const real* grid_v, const real* x0, real* u, real* v, real* F,
                            real* yield, const int32_t* samples, int64_t ns, double dt_in) {
    const int ng = P->n_grid;
    const real inv_dx = (real)P->inv_dx, dt = (real)dt_in, mu = (real)P->mu, lam = (real)P->lam;
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t q = samples[s];
        real x[3];
        for (int d = 0; d < 3; ++d) x[d] = x0[3 * q + d] + u[3 * q + d];
        int base[3]; real fx[3], w[9], dw[9];
        SFX(oracle_weights)(x, inv_dx, base, fx, w, dw);
        if (base[0] < 0 || base[1] < 0 || base[2] < 0 || base[0] > ng - 3 || base[1] > ng - 3 || base[2] > ng - 3)
            continue;
        real nv[3] = {0, 0, 0}, nF[9] = {0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    int64_t node = ((int64_t)(base[0] + i) * ng + (base[1] + j)) * ng + (base[2] + k);
                    real weight = w[0 * 3 + i] * w[1 * 3 + j] * w[2 * 3 + k];
                    const real* g = grid_v + node * 3;
                    real gw[3];
                    dweight(w, dw, i, j, k, inv_dx, gw);
                    for (int a = 0; a < 3; ++a) {
                        nv[a] = nv[a] + g[a] * weight;
                        for (int b = 0; b < 3; ++b) nF[a * 3 + b] = nF[a * 3 + b] + g[a] * gw[b];
                    }
                }
        real A[9], Ftr[9];
        for (int k = 0; k < 9; ++k) A[k] = dt * nF[k];
        A[0] = A[0] + (real)1; A[4] = A[4] + (real)1; A[8] = A[8] + (real)1;
        mm3(A, F + 9 * q, Ftr);
        if (P->material == 1) {
            SFX(oracle_von_mises)(Ftr, mu, lam, yield + s, P->hardening, (real)P->xi, F + 9 * q);
        } else if (P->material == 2) {
            SFX(oracle_sand_return_map)(Ftr, mu, lam, (real)P->alpha, F + 9 * q);
        } else {
            memcpy(F + 9 * q, Ftr, sizeof(Ftr));
        }
        for (int d = 0; d < 3; ++d) {
            v[3 * q + d] = nv[d];
            u[3 * q + d] = (x[d] + dt * nv[d]) - x0[3 * q + d];
        }
    }
}
