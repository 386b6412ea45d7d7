// crom_mlp.hip -- the CROM decoder MLP (NetDec, offline_training_u/
// CROMnet.py:218-301) on gfx950 matrix cores, and its C-ABI
// (include/crom_hip.h).
//
// Orientation.  Every layer is computed transposed, Z^T = W A^T, with the
// weights as the MFMA A operand (rows = output features) and the activations
// as the B operand (columns = points): one wave owns 32 points (the 32
// columns of v_mfma_f32_32x32x16_bf16) and all 128 features of a layer
// (4 row tiles, 64 accumulator VGPRs).  A 32x32 accumulator leaves lane l
// (column l mod 32, half h = l / 32) holding rows 8i + 4h + j (i, j < 4);
// the B operand of the next layer wants lane l to hold 8 consecutive k of
// its column at k = 16 kb + 8h + e.  Instead of exchanging lanes, the next
// layer's weight columns are packed in the order in which the accumulators
// hold the features: B element (kb, h, e) is accumulator element
// (t = kb / 2, i = 2 (kb mod 2) + e / 4, j = e mod 4), i.e. feature
//   f(kb, h, e) = 32 (kb/2) + 8 (2 (kb mod 2) + e/4) + 4h + (e mod 4)
// so activations go from accumulator to operand registers with only bias,
// ELU and the BF16 conversion in between.  All weights of the network
// (147 KB of BF16 fragments, biases, standardization constants) sit in LDS
// for the life of a persistent workgroup.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "crom_hip.h"
#include "mpm_common.h"

using namespace mpm;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kH = CROM_HIDDEN;          // hidden width
constexpr int kT = kH / 32;              // row tiles of a hidden layer
constexpr int kKB = kH / 16;             // K blocks of a hidden layer
constexpr int kWaves = 8;                // waves per workgroup
constexpr int kThreads = 64 * kWaves;

// Packed blob (crom_pack).  Fragments are 64 lanes x 16 B = 1 KB; fragment
// (t, kb) of layer L lies at frag_off[L] + (t * nkb[L] + kb) KB.
struct Header {
    int32_t in_dim, out_dim, x_col, pad;
    float mean_x[3], inv_std_x[3];
    float mean_q[CROM_MAX_OUT], std_q[CROM_MAX_OUT];
};
constexpr int kHeaderBytes = 512;
constexpr int kBiasFloats = 5 * kH + 32;
constexpr int kBiasBytes = kBiasFloats * 4;
// layer geometry: row tiles and K blocks
constexpr int kRowTiles[CROM_N_LINEAR] = {kT, kT, kT, kT, kT, 1};
constexpr int kKBlocks[CROM_N_LINEAR] = {2, kKB, kKB, kKB, kKB, kKB};
constexpr int frag_count(int L) { return kRowTiles[L] * kKBlocks[L]; }
constexpr int kFragTotal = frag_count(0) + 4 * frag_count(1) + frag_count(5);
constexpr int kFwdBytes = kHeaderBytes + kBiasBytes + kFragTotal * 1024;
// transposed fragments for the backward pass: rows = input features of the
// layer, K = its output features in accumulator order (layer 0: one row
// tile of input features)
constexpr int kRowTilesT[CROM_N_LINEAR] = {1, kT, kT, kT, kT, kT};
constexpr int kKBlocksT[CROM_N_LINEAR] = {kKB, kKB, kKB, kKB, kKB, 2};
constexpr int kFragTotalT = 4 * kT * kKB + kT * 2 + kKB;
constexpr int kBwdBytes = kHeaderBytes + kFragTotalT * 1024;   // backward LDS: header + transposed fragments
constexpr int kPackedBytes = kFwdBytes + kFragTotalT * 1024;
// activation / gradient buffers: per 32-point tile, section 0 holds 2 K
// blocks (the prepared input, resp. the enc layer's output gradient) and
// sections 1..5 hold 8 K blocks (hidden layers)
constexpr int kSec0Bytes = 2 * 1024;
constexpr int kSecBytes = kKB * 1024;
constexpr int kTileBufBytes = kSec0Bytes + 5 * kSecBytes;

static_assert(sizeof(Header) <= kHeaderBytes, "header");
static_assert(kFwdBytes <= 160 * 1024, "forward operands must fit in LDS");

__host__ __device__ constexpr int frag_base(int L) {
    return L == 0 ? 0 : frag_count(0) + (L - 1) * frag_count(1);
}
__host__ __device__ constexpr int frag_base_t(int L) {  // 1..4, then 5, then 0
    return L == 0 ? 4 * kT * kKB + kT * 2 : (L == 5 ? 4 * kT * kKB : (L - 1) * kT * kKB);
}
static_assert(kBwdBytes <= 160 * 1024, "backward operands must fit in LDS");

// byte offset of section `sec` of point tile `tile` in an activation or
// gradient buffer of n_tiles tiles
__host__ __device__ inline int64_t sec_off(int64_t n_tiles, int64_t tile, int sec) {
    return sec == 0 ? tile * kSec0Bytes : n_tiles * kSec0Bytes + ((int64_t)(sec - 1) * n_tiles + tile) * kSecBytes;
}

// feature held by B element (kb, h, e) of an activation operand
__host__ __device__ inline int feat_of(int kb, int h, int e) {
    return 32 * (kb >> 1) + 8 * (2 * (kb & 1) + (e >> 2)) + 4 * h + (e & 3);
}

// ELU (alpha = 1, CROMnet.py:471-472) of an activation that is rounded to
// BF16 right after: exp(z) - 1 with the hardware exponential (absolute error
// ~1e-7, far below BF16's 2^-8 relative step) instead of expm1f, which costs
// several times as many VALU instructions between the layers' MFMAs.
__device__ __forceinline__ float elu(float z) { return z > 0.f ? z : __expf(z) - 1.0f; }

// ---------------------------------------------------------------------------
// Packing: FP32 [out][in] weights -> BF16 A-operand fragments.
// Forward fragment (L, t, kb), lane (r, h): 8 values of row n = 32 t + r at
// k = 16 kb + 8 h + e, k being the input feature (layer 0) or feat_of (L > 0).
// Transposed fragment (L, t, kb): row m = 32 t + r is an input feature of
// layer L, k = feat_of(kb, h, e) an output feature of layer L.
// ---------------------------------------------------------------------------
struct PackArgs {
    const float* w[CROM_N_LINEAR];
    const float* b[CROM_N_LINEAR];
    const float *mean_x, *std_x, *mean_q, *std_q;
    int in_dim, out_dim, x_col;
};

__global__ void k_crom_pack(PackArgs a, uint8_t* __restrict__ packed) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = g & 63, r = lane & 31, h = lane >> 5;
    const int frag = g >> 6;
    if (frag < kFragTotal) {
        int L = 0, f = frag;
        while (L < 5 && f >= frag_count(L)) { f -= frag_count(L); ++L; }
        const int t = f / kKBlocks[L], kb = f % kKBlocks[L];
        const int n = 32 * t + r;
        const int n_in = L == 0 ? a.in_dim : kH;
        const int n_out = L == 5 ? a.out_dim : kH;
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = L == 0 ? 16 * kb + 8 * h + e : feat_of(kb, h, e);
            const float w = (n < n_out && k < n_in) ? a.w[L][(int64_t)n * n_in + k] : 0.f;
            v[e] = (__bf16)w;
        }
        *(bf16x8*)(packed + kHeaderBytes + kBiasBytes + (int64_t)(frag_base(L) + f) * 1024 + lane * 16) = v;
        return;
    }
    const int ft = frag - kFragTotal;
    if (ft < kFragTotalT) {
        // transposed: layers 1..4, then 5, then 0
        int L, f;
        if (ft < 4 * kT * kKB) { L = 1 + ft / (kT * kKB); f = ft % (kT * kKB); }
        else if (ft < 4 * kT * kKB + kT * 2) { L = 5; f = ft - 4 * kT * kKB; }
        else { L = 0; f = ft - 4 * kT * kKB - kT * 2; }
        const int t = f / kKBlocksT[L], kb = f % kKBlocksT[L];
        const int m = 32 * t + r;               // input feature of layer L
        const int n_in = L == 0 ? a.in_dim : kH;
        const int n_out = L == 5 ? a.out_dim : kH;
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int n = feat_of(kb, h, e);     // output feature of layer L
            const float w = (n < n_out && m < n_in) ? a.w[L][(int64_t)n * n_in + m] : 0.f;
            v[e] = (__bf16)w;
        }
        *(bf16x8*)(packed + kFwdBytes + (int64_t)(frag_base_t(L) + f) * 1024 + lane * 16) = v;
        return;
    }
    // header and biases
    const int s = g - 64 * (kFragTotal + kFragTotalT);
    float* bias = (float*)(packed + kHeaderBytes);
    if (s < kBiasFloats) {
        const int L = s < 5 * kH ? s / kH : 5;
        const int n = s < 5 * kH ? s % kH : s - 5 * kH;
        const int n_out = L == 5 ? a.out_dim : kH;
        bias[s] = n < n_out ? a.b[L][n] : 0.f;
    } else if (s == kBiasFloats) {
        Header* hd = (Header*)packed;
        hd->in_dim = a.in_dim;
        hd->out_dim = a.out_dim;
        hd->x_col = a.x_col;
        hd->pad = 0;
        for (int d = 0; d < 3; ++d) {
            hd->mean_x[d] = a.mean_x[d];
            hd->inv_std_x[d] = 1.0f / a.std_x[d];
        }
        for (int d = 0; d < CROM_MAX_OUT; ++d) {
            hd->mean_q[d] = d < a.out_dim ? a.mean_q[d] : 0.f;
            hd->std_q[d] = d < a.out_dim ? a.std_q[d] : 0.f;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused forward.  Persistent workgroups: the forward operands are copied to
// LDS once, then each wave takes 32-point tiles until none is left.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bf16x8 frag_at(const uint8_t* lds, int idx, int lane) {
    return *(const bf16x8*)(lds + kHeaderBytes + kBiasBytes + idx * 1024 + lane * 16);
}

// bias + ELU of accumulator tile t, written as the next layer's B operands
// kb = 2t, 2t+1
__device__ __forceinline__ void act_to_operand(const f32x16& acc, const float* bias, int t, int h, bf16x8& b0,
                                               bf16x8& b1) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i0 = e >> 2, j = e & 3;
        const int n0 = 32 * t + 8 * i0 + 4 * h + j;          // i = i0      -> kb = 2t
        const int n1 = 32 * t + 8 * (2 + i0) + 4 * h + j;    // i = 2 + i0  -> kb = 2t + 1
        b0[e] = (__bf16)elu(acc[4 * i0 + j] + bias[n0]);
        b1[e] = (__bf16)elu(acc[4 * (2 + i0) + j] + bias[n1]);
    }
}

__global__ void __launch_bounds__(kThreads) k_crom_forward(const uint8_t* __restrict__ packed,
                                                           const float* __restrict__ x, int64_t n,
                                                           float* __restrict__ y, uint8_t* __restrict__ acts) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    for (int o = threadIdx.x * 16; o < kFwdBytes; o += kThreads * 16)
        *(uint4*)(lds + o) = *(const uint4*)(packed + o);
    __syncthreads();
    const Header* hd = (const Header*)lds;
    const float* bias = (const float*)(lds + kHeaderBytes);
    const int in_dim = hd->in_dim, out_dim = hd->out_dim, x_col = hd->x_col;
    const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
    const int wave = threadIdx.x >> 6;
    const int64_t n_tiles = (n + 31) / 32;
    for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < n_tiles; tile += (int64_t)gridDim.x * kWaves) {
        const int64_t p = tile * 32 + col;
        const bool valid = p < n;
        // input operand: k = 16 kb + 8 h + e, Prepare on the position columns
        bf16x8 bin[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = 16 * kb + 8 * h + e;
                float v = 0.f;
                if (valid && k < in_dim) {
                    v = x[p * in_dim + k];
                    const int d = k - x_col;
                    if (d >= 0 && d < 3) v = (v - hd->mean_x[d]) * hd->inv_std_x[d];
                }
                bin[kb][e] = (__bf16)v;
            }
        if (acts) {
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) *(bf16x8*)(acts + sec_off(n_tiles, tile, 0) + kb * 1024 + lane * 16) = bin[kb];
        }
        f32x16 acc[kT];
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            acc[t] = (f32x16)0.f;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(frag_at(lds, frag_base(0) + t * 2 + kb, lane),
                                                                 bin[kb], acc[t], 0, 0, 0);
        }
        bf16x8 b[kKB];
#pragma unroll
        for (int t = 0; t < kT; ++t) act_to_operand(acc[t], bias, t, h, b[2 * t], b[2 * t + 1]);
        if (acts) {
#pragma unroll
            for (int kb = 0; kb < kKB; ++kb) *(bf16x8*)(acts + sec_off(n_tiles, tile, 1) + kb * 1024 + lane * 16) = b[kb];
        }
#pragma unroll 1
        for (int L = 1; L <= 4; ++L) {
#pragma unroll
            for (int t = 0; t < kT; ++t) {
                acc[t] = (f32x16)0.f;
#pragma unroll
                for (int kb = 0; kb < kKB; ++kb)
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        frag_at(lds, frag_base(L) + t * kKB + kb, lane), b[kb], acc[t], 0, 0, 0);
            }
#pragma unroll
            for (int t = 0; t < kT; ++t) act_to_operand(acc[t], bias + L * kH, t, h, b[2 * t], b[2 * t + 1]);
            if (acts) {
#pragma unroll
                for (int kb = 0; kb < kKB; ++kb)
                    *(bf16x8*)(acts + sec_off(n_tiles, tile, L + 1) + kb * 1024 + lane * 16) = b[kb];
            }
        }
        // enc: one row tile (out_dim <= 32), then invStandardizeQ
        f32x16 o = (f32x16)0.f;
#pragma unroll
        for (int kb = 0; kb < kKB; ++kb)
            o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(frag_at(lds, frag_base(5) + kb, lane), b[kb], o, 0, 0, 0);
        if (valid) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = 8 * i + 4 * h + j;
                    if (r < out_dim) y[p * out_dim + r] = hd->mean_q[r] + (o[4 * i + j] + bias[5 * kH + r]) * hd->std_q[r];
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Backward through the network for the loss  scale/2 * sum (y - q)^2
// (CROMnet.py:76: 1000 * MSELoss, so scale = 2000 / numel).  Per 32-point
// tile: dz_enc = scale (y - q) std_q (invStandardizeQ), then for each layer
// g = W^T dz with the transposed fragments (same accumulator-order trick,
// now over output features) and dz_prev = g * ELU'(a), ELU'(z) = 1 for
// a > 0 and a + 1 otherwise.  Every dz is stored (BF16 operand fragments,
// the activation buffers' layout) for the weight gradient; the input
// gradient (Prepare's chain rule on the position columns) is optional.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bf16x8 tfrag_at(const uint8_t* lds, int idx, int lane) {
    return *(const bf16x8*)(lds + kHeaderBytes + idx * 1024 + lane * 16);
}

__device__ __forceinline__ float elu_grad_of_act(float a) { return a > 0.f ? 1.f : a + 1.f; }

// dz of accumulator tile t (g * ELU'(a)) as operands kb = 2t, 2t+1; a given
// as the layer's stored operands
__device__ __forceinline__ void grad_to_operand(const f32x16& g, const bf16x8& a0, const bf16x8& a1, bf16x8& d0,
                                                bf16x8& d1) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i0 = e >> 2, j = e & 3;
        d0[e] = (__bf16)(g[4 * i0 + j] * elu_grad_of_act((float)a0[e]));
        d1[e] = (__bf16)(g[4 * (2 + i0) + j] * elu_grad_of_act((float)a1[e]));
    }
}

__global__ void __launch_bounds__(kThreads) k_crom_backward(const uint8_t* __restrict__ packed,
                                                            const uint8_t* __restrict__ acts,
                                                            const float* __restrict__ y, const float* __restrict__ q,
                                                            int64_t n, float scale, uint8_t* __restrict__ dlt,
                                                            float* __restrict__ dx) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    for (int o = threadIdx.x * 16; o < kHeaderBytes; o += kThreads * 16)
        *(uint4*)(lds + o) = *(const uint4*)(packed + o);
    for (int o = threadIdx.x * 16; o < kFragTotalT * 1024; o += kThreads * 16)
        *(uint4*)(lds + kHeaderBytes + o) = *(const uint4*)(packed + kFwdBytes + o);
    __syncthreads();
    const Header* hd = (const Header*)lds;
    const int in_dim = hd->in_dim, out_dim = hd->out_dim, x_col = hd->x_col;
    const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
    const int wave = threadIdx.x >> 6;
    const int64_t n_tiles = (n + 31) / 32;
    for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < n_tiles; tile += (int64_t)gridDim.x * kWaves) {
        const int64_t p = tile * 32 + col;
        const bool valid = p < n;
        // output gradient as operands kb = 0, 1 (rows r = feat_of(kb, h, e) < 32)
        bf16x8 d[kKB];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int r = feat_of(kb, h, e);
                float v = 0.f;
                if (valid && r < out_dim) v = scale * (y[p * out_dim + r] - q[p * out_dim + r]) * hd->std_q[r];
                d[kb][e] = (__bf16)v;
            }
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) *(bf16x8*)(dlt + sec_off(n_tiles, tile, 0) + kb * 1024 + lane * 16) = d[kb];
        // enc: g = W5^T dz (K = 2 blocks), dz4 = g * ELU'(a4)
        f32x16 g[kT];
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            g[t] = (f32x16)0.f;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                g[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(tfrag_at(lds, frag_base_t(5) + t * 2 + kb, lane),
                                                               d[kb], g[t], 0, 0, 0);
        }
#pragma unroll 1
        for (int L = 4; L >= 0; --L) {
            // stored activation of layer L (section L + 1) and dz_L
            const uint8_t* ap = acts + sec_off(n_tiles, tile, L + 1) + lane * 16;
#pragma unroll
            for (int t = 0; t < kT; ++t)
                grad_to_operand(g[t], *(const bf16x8*)(ap + (2 * t) * 1024), *(const bf16x8*)(ap + (2 * t + 1) * 1024),
                                d[2 * t], d[2 * t + 1]);
            uint8_t* dp = dlt + sec_off(n_tiles, tile, L + 1) + lane * 16;
#pragma unroll
            for (int kb = 0; kb < kKB; ++kb) *(bf16x8*)(dp + kb * 1024) = d[kb];
            if (L == 0) break;
            // g = W_L^T dz_L
#pragma unroll
            for (int t = 0; t < kT; ++t) {
                g[t] = (f32x16)0.f;
#pragma unroll
                for (int kb = 0; kb < kKB; ++kb)
                    g[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        tfrag_at(lds, frag_base_t(L) + t * kKB + kb, lane), d[kb], g[t], 0, 0, 0);
            }
        }
        if (dx) {
            // input gradient: W0^T dz0, one row tile of input features
            f32x16 gi = (f32x16)0.f;
#pragma unroll
            for (int kb = 0; kb < kKB; ++kb)
                gi = __builtin_amdgcn_mfma_f32_32x32x16_bf16(tfrag_at(lds, frag_base_t(0) + kb, lane), d[kb], gi, 0, 0, 0);
            if (valid) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int r = 8 * i + 4 * h + j;
                        if (r >= in_dim) continue;
                        float v = gi[4 * i + j];
                        const int dd = r - x_col;
                        if (dd >= 0 && dd < 3) v *= hd->inv_std_x[dd];
                        dx[p * in_dim + r] = v;
                    }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradients dW_L = sum_p dz_L[p] (x) a_{L-1}[p] and db_L = sum_p dz_L[p]
// as MFMA products over the points (K = points).  A workgroup stages 128
// points of dz_L and a_{L-1} in LDS transposed to [feature][point] (logical
// feature order), its 8 waves share the layer's 32x32 output tiles (plus one
// tile per row block against a column of ones for db), and it writes its
// partial sums over its share of the points; k_crom_wgrad_reduce adds them.
// ---------------------------------------------------------------------------
constexpr int kStagePts = 128;
constexpr int kStageRow = kStagePts + 8;   // bf16 per LDS row (16 B pad against bank conflicts)
constexpr int kPartialFloats = kH * kH + kH;

struct LayerGeom {
    int n_rows, n_cols;        // logical dW shape
    int row_tiles, col_tiles;
    int d_sec, d_kb;           // dz section and its K blocks
    int a_sec, a_kb;           // a section and its K blocks
    bool a_natural;            // section 0: features in natural order
};

__device__ __forceinline__ LayerGeom layer_geom(int L, int in_dim, int out_dim) {
    LayerGeom g;
    g.n_rows = L == 5 ? out_dim : kH;
    g.n_cols = L == 0 ? in_dim : kH;
    g.row_tiles = L == 5 ? 1 : kT;
    g.col_tiles = L == 0 ? 1 : kT;
    g.d_sec = L == 5 ? 0 : L + 1;
    g.d_kb = L == 5 ? 2 : kKB;
    g.a_sec = L;
    g.a_kb = L == 0 ? 2 : kKB;
    g.a_natural = L == 0;
    return g;
}

__device__ __forceinline__ void stage_transposed(const uint8_t* __restrict__ buf, int64_t n_tiles, int64_t tile0,
                                                 int sec, int nkb, bool natural, __bf16* s) {
    // 4 point tiles x nkb K blocks x 64 lanes chunks of 8 values
    const int chunks = 4 * nkb * 64;
    for (int c = threadIdx.x; c < chunks; c += kThreads) {
        const int lane = c & 63, kb = (c >> 6) % nkb, tl = (c >> 6) / nkb;
        const int64_t tile = tile0 + tl;
        bf16x8 v = (bf16x8)(__bf16)0.f;
        if (tile < n_tiles) v = *(const bf16x8*)(buf + sec_off(n_tiles, tile, sec) + kb * 1024 + lane * 16);
        const int pt = tl * 32 + (lane & 31), hh = lane >> 5;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int f = natural ? 16 * kb + 8 * hh + e : feat_of(kb, hh, e);
            s[f * kStageRow + pt] = v[e];
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_crom_wgrad(const uint8_t* __restrict__ acts,
                                                         const uint8_t* __restrict__ dlt, int64_t n, int in_dim,
                                                         int out_dim, float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) __bf16 sD[kH * kStageRow];
    __shared__ __attribute__((aligned(16))) __bf16 sA[kH * kStageRow];
    const int L = blockIdx.y;
    const LayerGeom G = layer_geom(L, in_dim, out_dim);
    const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5, wave = threadIdx.x >> 6;
    const int64_t n_tiles = (n + 31) / 32;
    const int64_t n_stages = (n_tiles + 3) / 4;
    // output tiles of this wave: w, w + 8, w + 16 over [row_tiles x col_tiles] then the db tiles
    const int n_wt = G.row_tiles * G.col_tiles, n_all = n_wt + G.row_tiles;
    f32x16 acc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] = (f32x16)0.f;
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.f;
    // zero the rows that a section does not fill (padded features)
    for (int i = threadIdx.x; i < kH * kStageRow; i += kThreads) { sD[i] = (__bf16)0.f; sA[i] = (__bf16)0.f; }
    for (int64_t st = blockIdx.x; st < n_stages; st += gridDim.x) {
        __syncthreads();
        stage_transposed(dlt, n_tiles, st * 4, G.d_sec, G.d_kb, false, sD);
        stage_transposed(acts, n_tiles, st * 4, G.a_sec, G.a_kb, G.a_natural, sA);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int ot = wave + 8 * k;
            if (ot >= n_all) break;
            const bool bias_tile = ot >= n_wt;
            const int rt = bias_tile ? ot - n_wt : ot / G.col_tiles;
            const int ct = bias_tile ? 0 : ot % G.col_tiles;
#pragma unroll
            for (int kp = 0; kp < kStagePts / 16; ++kp) {
                const bf16x8 av = *(const bf16x8*)(sD + (32 * rt + col) * kStageRow + 16 * kp + 8 * h);
                const bf16x8 bv = bias_tile ? ones : *(const bf16x8*)(sA + (32 * ct + col) * kStageRow + 16 * kp + 8 * h);
                acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[k], 0, 0, 0);
            }
        }
    }
    // partial sums of this workgroup: [row][col] then db[row]
    float* out = partial + ((int64_t)L * gridDim.x + blockIdx.x) * kPartialFloats;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ot = wave + 8 * k;
        if (ot >= n_all) break;
        const bool bias_tile = ot >= n_wt;
        const int rt = bias_tile ? ot - n_wt : ot / G.col_tiles;
        const int ct = bias_tile ? 0 : ot % G.col_tiles;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = 32 * rt + 8 * i + 4 * h + j;
                const int c = 32 * ct + col;
                if (bias_tile) {
                    if (col == 0) out[kH * kH + r] = acc[k][4 * i + j];
                } else {
                    out[r * kH + c] = acc[k][4 * i + j];
                }
            }
    }
}

// grads (flat, per layer W [n_rows][n_cols] then b [n_rows]) = sum of the
// workgroups' partials
__global__ void k_crom_wgrad_reduce(const float* __restrict__ partial, int n_parts, int in_dim, int out_dim,
                                    float* __restrict__ grads) {
    const int64_t gidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t off = 0;
    for (int L = 0; L < CROM_N_LINEAR; ++L) {
        const int nr = L == 5 ? out_dim : kH, nc = L == 0 ? in_dim : kH;
        const int64_t cnt = (int64_t)nr * nc + nr;
        if (gidx < off + cnt) {
            const int64_t k = gidx - off;
            const int64_t src = k < (int64_t)nr * nc ? (k / nc) * kH + (k % nc) : kH * kH + (k - (int64_t)nr * nc);
            float sum = 0.f;
            const float* pp = partial + (int64_t)L * n_parts * kPartialFloats + src;
            for (int w = 0; w < n_parts; ++w) sum += pp[(int64_t)w * kPartialFloats];
            grads[gidx] = sum;
            return;
        }
        off += cnt;
    }
}

// Adam (torch.optim.Adam defaults as configured at CROMnet.py:193-196:
// no weight decay, no amsgrad), one thread per parameter
__global__ void k_crom_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                            float* __restrict__ v, int64_t count, float lr, float b1, float b2, float eps,
                            float bc1, float bc2_sqrt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float denom = sqrtf(vi) / bc2_sqrt + eps;
    w[i] = w[i] - (lr / bc1) * (mi / denom);
}

int check_net(const CromNet* net) {
    if (!net) return set_err(-60, "crom: net is NULL");
    if (net->in_dim < 4 || net->in_dim > CROM_MAX_IN) return set_err(-61, "crom: in_dim out of range [4, 32]");
    if (net->out_dim < 1 || net->out_dim > CROM_MAX_OUT) return set_err(-62, "crom: out_dim out of range [1, 32]");
    if (net->x_col < 0 || net->x_col + 3 > net->in_dim) return set_err(-63, "crom: x_col + 3 > in_dim");
    for (int L = 0; L < CROM_N_LINEAR; ++L)
        if (!net->weight[L] || !net->bias[L]) return set_err(-64, "crom: NULL weight or bias");
    if (!net->mean_x || !net->std_x || !net->mean_q || !net->std_q)
        return set_err(-65, "crom: NULL standardization constants");
    return 0;
}

int g_cu_count = 0;
int cu_count() {
    if (g_cu_count == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            g_cu_count = prop.multiProcessorCount;
        if (g_cu_count <= 0) g_cu_count = 256;
    }
    return g_cu_count;
}

}  // namespace

extern "C" {

int64_t crom_packed_bytes(void) { return kPackedBytes; }

int64_t crom_acts_bytes(int64_t n) {
    if (n < 0) return -1;
    return ((n + 31) / 32) * kTileBufBytes;
}

int crom_pack(const CromNet* net, void* packed, void* stream) {
    int rc = check_net(net);
    if (rc) return rc;
    if (!packed || ((uintptr_t)packed & 255)) return set_err(-66, "crom: packed is NULL or not 256-byte aligned");
    PackArgs a;
    for (int L = 0; L < CROM_N_LINEAR; ++L) {
        a.w[L] = net->weight[L];
        a.b[L] = net->bias[L];
    }
    a.mean_x = net->mean_x;
    a.std_x = net->std_x;
    a.mean_q = net->mean_q;
    a.std_q = net->std_q;
    a.in_dim = net->in_dim;
    a.out_dim = net->out_dim;
    a.x_col = net->x_col;
    const int threads = 64 * (kFragTotal + kFragTotalT) + kBiasFloats + 1;
    hipLaunchKernelGGL(k_crom_pack, dim3((threads + 255) / 256), dim3(256), 0, (hipStream_t)stream, a,
                       (uint8_t*)packed);
    return check_launch("k_crom_pack");
}

int crom_forward(const void* packed, const float* x, int64_t n, float* y, void* acts, void* stream) {
    if (!packed || ((uintptr_t)packed & 255)) return set_err(-66, "crom: packed is NULL or not 256-byte aligned");
    if (n < 0) return set_err(-67, "crom: n < 0");
    if (n == 0) return 0;
    if (!x || !y) return set_err(-68, "crom: x or y is NULL");
    if (((uintptr_t)acts & 15)) return set_err(-69, "crom: acts not 16-byte aligned");
    const int64_t n_tiles = (n + 31) / 32;
    const int64_t want = (n_tiles + kWaves - 1) / kWaves;
    const unsigned grid = (unsigned)(want < cu_count() ? want : cu_count());
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void*)k_crom_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                kFwdBytes) != hipSuccess)
            return set_err(-70, "crom: cannot enable 147 KB of dynamic LDS");
        attr_set = true;
    }
    hipLaunchKernelGGL(k_crom_forward, dim3(grid), dim3(kThreads), kFwdBytes, (hipStream_t)stream,
                       (const uint8_t*)packed, x, n, y, (uint8_t*)acts);
    return check_launch("k_crom_forward");
}

int crom_backward(const void* packed, const void* acts, const float* y, const float* q, int64_t n, float scale,
                  void* dlt, float* dx, void* stream) {
    if (!packed || ((uintptr_t)packed & 255)) return set_err(-66, "crom: packed is NULL or not 256-byte aligned");
    if (n < 0) return set_err(-67, "crom: n < 0");
    if (n == 0) return 0;
    if (!acts || !y || !q || !dlt) return set_err(-68, "crom: acts, y, q or grads is NULL");
    if (((uintptr_t)acts & 15) || ((uintptr_t)dlt & 15)) return set_err(-69, "crom: buffers not 16-byte aligned");
    const int64_t n_tiles = (n + 31) / 32;
    const int64_t want = (n_tiles + kWaves - 1) / kWaves;
    const unsigned grid = (unsigned)(want < cu_count() ? want : cu_count());
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void*)k_crom_backward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                kBwdBytes) != hipSuccess)
            return set_err(-70, "crom: cannot enable the backward pass's dynamic LDS");
        attr_set = true;
    }
    hipLaunchKernelGGL(k_crom_backward, dim3(grid), dim3(kThreads), kBwdBytes, (hipStream_t)stream,
                       (const uint8_t*)packed, (const uint8_t*)acts, y, q, n, scale, (uint8_t*)dlt, dx);
    return check_launch("k_crom_backward");
}

int64_t crom_param_count(int32_t in_dim, int32_t out_dim) {
    if (in_dim < 4 || in_dim > CROM_MAX_IN || out_dim < 1 || out_dim > CROM_MAX_OUT) return -1;
    return (int64_t)kH * in_dim + kH + 4 * ((int64_t)kH * kH + kH) + (int64_t)out_dim * kH + out_dim;
}

static unsigned wgrad_parts(int64_t n) {
    const int64_t stages = ((n + 31) / 32 + 3) / 4;
    return (unsigned)(stages < 256 ? (stages > 0 ? stages : 1) : 256);
}

int64_t crom_wgrad_workspace_bytes(int64_t n) {
    if (n < 0) return -1;
    return (int64_t)CROM_N_LINEAR * wgrad_parts(n) * kPartialFloats * 4;
}

int crom_wgrad(const void* acts, const void* dlt, int64_t n, int32_t in_dim, int32_t out_dim, void* workspace,
               float* grads, void* stream) {
    if (crom_param_count(in_dim, out_dim) < 0) return set_err(-61, "crom: in_dim / out_dim out of range");
    if (n < 1) return set_err(-67, "crom: n < 1");
    if (!acts || !dlt || !workspace || !grads) return set_err(-68, "crom: NULL buffer");
    if (((uintptr_t)acts & 15) || ((uintptr_t)dlt & 15)) return set_err(-69, "crom: buffers not 16-byte aligned");
    const unsigned parts = wgrad_parts(n);
    hipLaunchKernelGGL(k_crom_wgrad, dim3(parts, CROM_N_LINEAR), dim3(kThreads), 0, (hipStream_t)stream,
                       (const uint8_t*)acts, (const uint8_t*)dlt, n, in_dim, out_dim, (float*)workspace);
    int rc = check_launch("k_crom_wgrad");
    if (rc) return rc;
    const int64_t count = crom_param_count(in_dim, out_dim);
    hipLaunchKernelGGL(k_crom_wgrad_reduce, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)workspace, (int)parts, in_dim, out_dim, grads);
    return check_launch("k_crom_wgrad_reduce");
}

int crom_adam(float* params, const float* grads, float* m, float* v, int64_t count, float lr, float beta1,
              float beta2, float eps, int32_t step, void* stream) {
    if (!params || !grads || !m || !v) return set_err(-68, "crom: NULL buffer");
    if (count < 0 || step < 1) return set_err(-71, "crom: count < 0 or step < 1");
    if (count == 0) return 0;
    const float bc1 = 1.0f - powf(beta1, (float)step);
    const float bc2_sqrt = sqrtf(1.0f - powf(beta2, (float)step));
    hipLaunchKernelGGL(k_crom_adam, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                       grads, m, v, count, lr, beta1, beta2, eps, bc1, bc2_sqrt);
    return check_launch("k_crom_adam");
}

}  // extern "C"
