// crom_encoder.hip -- the CROM encoder NetEnc (offline_training_u/
// CROMnet.py:361-438) and the per-frame latent gradient, FP32, and their
// C-ABI (include/crom_hip.h).
//
// NetEnc maps a frame's displacement field q [P][3] to its latent:
// standardizeQ (:440-455), Conv1d(3 -> 3, kernel ks, stride s) + ELU
// repeated while 3 * L_out >= 32 (:377-384, util.py:201-203), flatten, Linear
// -> 32 + ELU, Linear -> lbl + ELU (:385-386, :429-436).  It is a few MFLOP
// per frame -- bandwidth-trivial, so plain FP32 kernels: one thread per
// output element, reductions over frames and positions per weight.
#include <hip/hip_runtime.h>
#include <math.h>

#include "crom_hip.h"
#include "mpm_common.h"

using namespace mpm;

namespace {

constexpr int kB = 256;

__device__ __forceinline__ float elu(float z) { return z > 0.f ? z : expm1f(z); }
__device__ __forceinline__ float elu_grad_of_act(float a) { return a > 0.f ? 1.f : a + 1.f; }

// input element (frame b, channel c, position t): a layer's input [B][3][L],
// or for the first layer the standardized field q [B][P][3]
__device__ __forceinline__ float conv_in(const float* in, int b, int c, int t, int L, const float* qm,
                                         const float* qis) {
    if (qm) return (in[((int64_t)b * L + t) * 3 + c] - qm[c]) * qis[c];
    return in[((int64_t)b * 3 + c) * L + t];
}

__global__ void k_conv_fwd(const float* __restrict__ in, int B, int L_in, int ks, int st, const float* __restrict__ W,
                           const float* __restrict__ bias, float* __restrict__ out, int L_out,
                           const float* __restrict__ qm, const float* __restrict__ qis) {
    const int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (g >= (int64_t)B * 3 * L_out) return;
    const int t = (int)(g % L_out), co = (int)((g / L_out) % 3), b = (int)(g / (3 * (int64_t)L_out));
    float z = bias[co];
    for (int ci = 0; ci < 3; ++ci)
        for (int j = 0; j < ks; ++j) z += W[(co * 3 + ci) * ks + j] * conv_in(in, b, ci, t * st + j, L_in, qm, qis);
    out[g] = elu(z);
}

// dz = dout * ELU'(out); din (position t'): sum over co, j with t' = t s + j
__global__ void k_conv_bwd_in(const float* __restrict__ dout, const float* __restrict__ out, int B, int L_in, int ks,
                              int st, const float* __restrict__ W, int L_out, float* __restrict__ din) {
    const int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (g >= (int64_t)B * 3 * L_in) return;
    const int tp = (int)(g % L_in), ci = (int)((g / L_in) % 3), b = (int)(g / (3 * (int64_t)L_in));
    float s = 0.f;
    for (int j = 0; j < ks; ++j) {
        const int num = tp - j;
        if (num < 0 || num % st) continue;
        const int t = num / st;
        if (t >= L_out) continue;
        for (int co = 0; co < 3; ++co) {
            const int64_t o = ((int64_t)b * 3 + co) * L_out + t;
            s += W[(co * 3 + ci) * ks + j] * dout[o] * elu_grad_of_act(out[o]);
        }
    }
    din[g] = s;
}

// Weight and bias gradients: reductions over frames and output positions.
// Workgroup (w, s) sums slice s of the B * L_out products of weight w (or of
// bias w - 9 ks), then adds its partial with one float atomic into the
// zeroed gW / gb (k_zero_f32 first), so that the reduction spreads over
// kConvSplit workgroups per weight instead of one.
constexpr int kConvSplit = 64;

__global__ void k_zero_f32(float* __restrict__ a, int na, float* __restrict__ b, int nb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0.f;
    else if (i < na + nb) b[i - na] = 0.f;
}

__global__ void __launch_bounds__(kB)
    k_conv_bwd_w(const float* __restrict__ in, const float* __restrict__ dout, const float* __restrict__ out, int B,
                 int L_in, int ks, int st, int L_out, const float* __restrict__ qm, const float* __restrict__ qis,
                 float* __restrict__ gW, float* __restrict__ gb) {
    __shared__ float red[kB];
    const int w = blockIdx.x;
    const int n_w = 9 * ks;
    const int64_t total = (int64_t)B * L_out;
    const int64_t per = (total + kConvSplit - 1) / kConvSplit;
    const int64_t e0 = (int64_t)blockIdx.y * per, e1 = min(total, e0 + per);
    float s = 0.f;
    if (w < n_w) {
        const int co = w / (3 * ks), ci = (w / ks) % 3, j = w % ks;
        for (int64_t e = e0 + threadIdx.x; e < e1; e += kB) {
            const int b = (int)(e / L_out), t = (int)(e % L_out);
            const int64_t o = ((int64_t)b * 3 + co) * L_out + t;
            s += dout[o] * elu_grad_of_act(out[o]) * conv_in(in, b, ci, t * st + j, L_in, qm, qis);
        }
    } else {
        const int co = w - n_w;
        for (int64_t e = e0 + threadIdx.x; e < e1; e += kB) {
            const int b = (int)(e / L_out), t = (int)(e % L_out);
            const int64_t o = ((int64_t)b * 3 + co) * L_out + t;
            s += dout[o] * elu_grad_of_act(out[o]);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && e0 < e1) {
        if (w < n_w) atomicAdd(&gW[w], red[0]);
        else atomicAdd(&gb[w - n_w], red[0]);
    }
}

// y[b][o] = ELU(b[o] + sum_i W[o][i] x[b][i])
__global__ void k_dense_fwd(const float* __restrict__ x, int B, int n_in, int n_out, const float* __restrict__ W,
                            const float* __restrict__ bias, float* __restrict__ y) {
    const int g = blockIdx.x * kB + threadIdx.x;
    if (g >= B * n_out) return;
    const int o = g % n_out, b = g / n_out;
    float z = bias[o];
    for (int i = 0; i < n_in; ++i) z += W[o * n_in + i] * x[b * n_in + i];
    y[g] = elu(z);
}

// dx = W^T dz, dz = dy * ELU'(y)
__global__ void k_dense_bwd_in(const float* __restrict__ dy, const float* __restrict__ y, int B, int n_in, int n_out,
                               const float* __restrict__ W, float* __restrict__ dx) {
    const int g = blockIdx.x * kB + threadIdx.x;
    if (g >= B * n_in) return;
    const int i = g % n_in, b = g / n_in;
    float s = 0.f;
    for (int o = 0; o < n_out; ++o) s += W[o * n_in + i] * dy[b * n_out + o] * elu_grad_of_act(y[b * n_out + o]);
    dx[g] = s;
}

// dW[o][i] = sum_b dz[b][o] x[b][i], db[o] = sum_b dz[b][o]
__global__ void k_dense_bwd_w(const float* __restrict__ dy, const float* __restrict__ y, const float* __restrict__ x,
                              int B, int n_in, int n_out, float* __restrict__ gW, float* __restrict__ gb) {
    const int g = blockIdx.x * kB + threadIdx.x;
    if (g >= n_out * (n_in + 1)) return;
    const int o = g / (n_in + 1), i = g % (n_in + 1);
    float s = 0.f;
    for (int b = 0; b < B; ++b) {
        const float dz = dy[b * n_out + o] * elu_grad_of_act(y[b * n_out + o]);
        s += dz * (i < n_in ? x[b * n_in + i] : 1.f);
    }
    if (i < n_in) gW[o * n_in + i] = s;
    else gb[o] = s;
}

// g[f][c] = sum over the P points of frame f of dx[p][c], c < lbl.
// Workgroup (f, s) reads slice s of the frame's rows (each thread whole rows,
// coalesced across the workgroup), keeps the lbl column sums in registers,
// reduces them through LDS and adds them to the zeroed g with float atomics.
constexpr int kLatSplit = 64;
constexpr int kLatMax = 32;   // lbl <= in_dim <= 32

__global__ void __launch_bounds__(kB) k_latent_grad(const float* __restrict__ dx, int P, int in_dim, int lbl,
                                                    float* __restrict__ g) {
    __shared__ float red[kLatMax][kB / 64];
    const int f = blockIdx.x;
    const int per = (P + kLatSplit - 1) / kLatSplit;
    const int p0 = blockIdx.y * per, p1 = min(P, p0 + per);
    float s[kLatMax];
#pragma unroll
    for (int c = 0; c < kLatMax; ++c) s[c] = 0.f;
    for (int p = p0 + threadIdx.x; p < p1; p += kB) {
        const float* row = dx + ((int64_t)f * P + p) * in_dim;
#pragma unroll
        for (int c = 0; c < kLatMax; ++c)
            if (c < lbl) s[c] += row[c];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < kLatMax; ++c) {
        if (c >= lbl) break;
        float v = s[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[c][wave] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < lbl && p0 < p1) {
        float v = 0.f;
#pragma unroll
        for (int w = 0; w < kB / 64; ++w) v += red[threadIdx.x][w];
        atomicAdd(&g[f * lbl + threadIdx.x], v);
    }
}

unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace

extern "C" {

int crom_conv_forward(const float* in, int B, int L_in, int ks, int stride, const float* W, const float* bias,
                      float* out, const float* q_mean, const float* q_std_inv, void* stream) {
    if (!in || !W || !bias || !out || B < 1 || ks < 1 || stride < 1 || L_in < ks)
        return set_err(-80, "crom conv: bad arguments");
    const int L_out = (L_in - ks) / stride + 1;
    hipLaunchKernelGGL(k_conv_fwd, dim3(nblk((int64_t)B * 3 * L_out)), dim3(kB), 0, (hipStream_t)stream, in, B, L_in,
                       ks, stride, W, bias, out, L_out, q_mean, q_std_inv);
    return check_launch("k_conv_fwd");
}

int crom_conv_backward(const float* in, const float* out, const float* dout, int B, int L_in, int ks, int stride,
                       const float* W, float* din, float* gW, float* gb, const float* q_mean, const float* q_std_inv,
                       void* stream) {
    if (!in || !out || !dout || !W || !gW || !gb || B < 1 || ks < 1 || stride < 1 || L_in < ks)
        return set_err(-80, "crom conv: bad arguments");
    const int L_out = (L_in - ks) / stride + 1;
    int rc;
    if (din) {
        hipLaunchKernelGGL(k_conv_bwd_in, dim3(nblk((int64_t)B * 3 * L_in)), dim3(kB), 0, (hipStream_t)stream, dout,
                           out, B, L_in, ks, stride, W, L_out, din);
        if ((rc = check_launch("k_conv_bwd_in"))) return rc;
    }
    hipLaunchKernelGGL(k_zero_f32, dim3(nblk(9 * ks + 3)), dim3(kB), 0, (hipStream_t)stream, gW, 9 * ks, gb, 3);
    if ((rc = check_launch("k_zero_f32"))) return rc;
    hipLaunchKernelGGL(k_conv_bwd_w, dim3(9 * ks + 3, kConvSplit), dim3(kB), 0, (hipStream_t)stream, in, dout, out, B,
                       L_in, ks, stride, L_out, q_mean, q_std_inv, gW, gb);
    return check_launch("k_conv_bwd_w");
}

int crom_dense_forward(const float* x, int B, int n_in, int n_out, const float* W, const float* bias, float* y,
                       void* stream) {
    if (!x || !W || !bias || !y || B < 1 || n_in < 1 || n_out < 1) return set_err(-81, "crom dense: bad arguments");
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((int64_t)B * n_out)), dim3(kB), 0, (hipStream_t)stream, x, B, n_in,
                       n_out, W, bias, y);
    return check_launch("k_dense_fwd");
}

int crom_dense_backward(const float* x, const float* y, const float* dy, int B, int n_in, int n_out, const float* W,
                        float* dx, float* gW, float* gb, void* stream) {
    if (!x || !y || !dy || !W || !gW || !gb || B < 1 || n_in < 1 || n_out < 1)
        return set_err(-81, "crom dense: bad arguments");
    int rc;
    if (dx) {
        hipLaunchKernelGGL(k_dense_bwd_in, dim3(nblk((int64_t)B * n_in)), dim3(kB), 0, (hipStream_t)stream, dy, y, B,
                           n_in, n_out, W, dx);
        if ((rc = check_launch("k_dense_bwd_in"))) return rc;
    }
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((int64_t)n_out * (n_in + 1))), dim3(kB), 0, (hipStream_t)stream, dy,
                       y, x, B, n_in, n_out, gW, gb);
    return check_launch("k_dense_bwd_w");
}

int crom_latent_grad(const float* dx, int n_frames, int points_per_frame, int in_dim, int lbl, float* g,
                     void* stream) {
    if (!dx || !g || n_frames < 1 || points_per_frame < 1 || lbl < 1 || lbl > in_dim || lbl > kLatMax)
        return set_err(-82, "crom latent grad: bad arguments");
    hipLaunchKernelGGL(k_zero_f32, dim3(nblk((int64_t)n_frames * lbl)), dim3(kB), 0, (hipStream_t)stream, g,
                       n_frames * lbl, g, 0);
    int rc = check_launch("k_zero_f32");
    if (rc) return rc;
    hipLaunchKernelGGL(k_latent_grad, dim3(n_frames, kLatSplit), dim3(kB), 0, (hipStream_t)stream, dx,
                       points_per_frame, in_dim, lbl, g);
    return check_launch("k_latent_grad");
}

}  // extern "C"
