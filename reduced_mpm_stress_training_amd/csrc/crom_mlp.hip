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
// transposed fragments for the backward pass (layers 5..1): rows = input
// features of the layer, K = its output features in accumulator order
constexpr int kRowTilesT[CROM_N_LINEAR] = {0, kT, kT, kT, kT, kT};
constexpr int kKBlocksT[CROM_N_LINEAR] = {0, kKB, kKB, kKB, kKB, 2};
constexpr int kFragTotalT = 4 * kT * kKB + kT * 2;
constexpr int kPackedBytes = kFwdBytes + kFragTotalT * 1024;

static_assert(sizeof(Header) <= kHeaderBytes, "header");
static_assert(kFwdBytes <= 160 * 1024, "forward operands must fit in LDS");

__host__ __device__ constexpr int frag_base(int L) {
    return L == 0 ? 0 : frag_count(0) + (L - 1) * frag_count(1);
}
__host__ __device__ constexpr int frag_base_t(int L) {  // L = 1..5
    return L == 5 ? 4 * kT * kKB : (L - 1) * kT * kKB;
}

// feature held by B element (kb, h, e) of an activation operand
__host__ __device__ inline int feat_of(int kb, int h, int e) {
    return 32 * (kb >> 1) + 8 * (2 * (kb & 1) + (e >> 2)) + 4 * h + (e & 3);
}

__device__ __forceinline__ float elu(float z) { return z > 0.f ? z : expm1f(z); }

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
        // transposed: layers 1..4 then 5
        int L, f;
        if (ft < 4 * kT * kKB) { L = 1 + ft / (kT * kKB); f = ft % (kT * kKB); }
        else { L = 5; f = ft - 4 * kT * kKB; }
        const int t = f / kKBlocksT[L], kb = f % kKBlocksT[L];
        const int m = 32 * t + r;               // input feature of layer L (< kH)
        const int n_out = L == 5 ? a.out_dim : kH;
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int n = feat_of(kb, h, e);     // output feature of layer L
            const float w = n < n_out ? a.w[L][(int64_t)n * kH + m] : 0.f;
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
        uint8_t* ap = acts ? acts + (tile * kKB) * 1024 : nullptr;   // layer stride: n_tiles * 8 KB
        const int64_t layer_stride = n_tiles * kKB * 1024;
        if (ap) {
#pragma unroll
            for (int kb = 0; kb < kKB; ++kb) *(bf16x8*)(ap + kb * 1024 + lane * 16) = b[kb];
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
            if (ap) {
#pragma unroll
                for (int kb = 0; kb < kKB; ++kb)
                    *(bf16x8*)(ap + L * layer_stride + kb * 1024 + lane * 16) = b[kb];
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
    return 5 * ((n + 31) / 32) * kKB * 1024;
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

}  // extern "C"
