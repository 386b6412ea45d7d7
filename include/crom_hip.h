/* crom_hip.h -- C-ABI of the CROM decoder MLP on MI355X matrix cores.
 *
 * The network is the reference's decoder NetDec (offline_training_u/
 * CROMnet.py:218-301): Prepare (standardize the 3 reference-position columns,
 * :497-534) -> dec0..dec4 (Linear + ELU(alpha=1), :226-237, :457-473) -> enc
 * (Linear) -> invStandardizeQ (mean_q + y * std_q, :475-487).  The stress head
 * of offline_training_xhat_and_stress_and_mu/CROMnet_secondary.py:203-307 is
 * the same trunk with input [label, X, mu] and 6 outputs; it uses the same
 * entry points with in_dim = lbl + 4, out_dim = 6.
 *
 * Hidden width is CROM_HIDDEN (the configuration's 128); in_dim and out_dim
 * are at most 32.  Matrix products run on BF16 MFMA with FP32 accumulation;
 * weights are kept by the caller in FP32 (torch nn.Linear layout [out][in])
 * and packed to BF16 operand fragments by crom_pack.
 *
 * All pointers are device pointers; `stream` is a hipStream_t.  Status 0 on
 * success, negative on error (message in mpm_last_error(), include/mpm_hip.h).
 * The library allocates nothing.
 */
#ifndef CROM_HIP_H
#define CROM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CROM_HIDDEN 128
#define CROM_N_LINEAR 6   /* dec0..dec4, enc */
#define CROM_MAX_IN 32
#define CROM_MAX_OUT 32

typedef struct CromNet {
    int32_t in_dim;           /* lbl + 3 (+ mu columns), <= 32 */
    int32_t out_dim;          /* 3 (displacement) or 6 (stress), <= 32 */
    int32_t x_col;            /* first of the 3 position columns Prepare standardizes (= lbl) */
    int32_t reserved;
    const float* weight[CROM_N_LINEAR];  /* [out][in] FP32: dec0 [128][in_dim], dec1..4 [128][128], enc [out_dim][128] */
    const float* bias[CROM_N_LINEAR];
    const float* mean_x;      /* 3: Prepare.standardizeX, CROMnet.py:527-528 */
    const float* std_x;       /* 3 */
    const float* mean_q;      /* out_dim: invStandardizeQ, CROMnet.py:486-487 */
    const float* std_q;       /* out_dim */
} CromNet;

/* Bytes of the packed forward operands (BF16 weight fragments, biases,
 * standardization constants). */
int64_t crom_packed_bytes(void);

/* Packs `net` for crom_forward (and the transposed fragments crom_backward
 * uses) into `packed` (crom_packed_bytes() bytes, 256-byte aligned). */
int crom_pack(const CromNet* net, void* packed, void* stream);

/* NetDec.forward (CROMnet.py:293-296) on n points: x [n][in_dim] FP32
 * (decoder input [label, X, (mu)]), y [n][out_dim] FP32.  `acts` may be NULL;
 * otherwise it receives the prepared input and the five hidden activations
 * in BF16 for crom_backward / crom_wgrad (crom_acts_bytes(n) bytes). */
int crom_forward(const void* packed, const float* x, int64_t n, float* y, void* acts, void* stream);

/* Bytes of the activation / gradient buffers of n points. */
int64_t crom_acts_bytes(int64_t n);

/* Backward of the training loss (CROMnet.py:70-83: 1000 * MSELoss(y, q))
 * through the decoder, written with the loss as  scale/2 * sum (y - q)^2
 * (scale = 2000 / (n * out_dim) for that loss).  acts: from crom_forward of
 * the same x; y, q [n][out_dim].  Writes the layers' pre-activation
 * gradients to `grads_buf` (crom_acts_bytes(n) bytes) and, when dx is not
 * NULL, the input gradient dx [n][in_dim] (through Prepare on the position
 * columns) -- the latent columns' gradient is what the encoder needs. */
int crom_backward(const void* packed, const void* acts, const float* y, const float* q, int64_t n, float scale,
                  void* grads_buf, float* dx, void* stream);

/* Number of parameters (weights [out][in] then bias, layer by layer: dec0,
 * dec1..dec4, enc), the layout of the flat parameter / gradient / Adam
 * moment buffers. */
int64_t crom_param_count(int32_t in_dim, int32_t out_dim);

/* Workspace of crom_wgrad for n points. */
int64_t crom_wgrad_workspace_bytes(int64_t n);

/* Parameter gradients (flat, crom_param_count floats) from the activations
 * of crom_forward and the gradients of crom_backward over the same n points:
 * dW = sum_p dz (x) a_prev, db = sum_p dz, on matrix cores over the points. */
int crom_wgrad(const void* acts, const void* grads_buf, int64_t n, int32_t in_dim, int32_t out_dim, void* workspace,
               float* grads, void* stream);

/* One Adam step (torch.optim.Adam as CROMnet.py:193-196 configures it: lr
 * from the LambdaLR schedule, betas, eps, no weight decay) over `count`
 * FP32 parameters in place; step counts from 1. */
int crom_adam(float* params, const float* grads, float* m, float* v, int64_t count, float lr, float beta1,
              float beta2, float eps, int32_t step, void* stream);

/* ---- encoder NetEnc (CROMnet.py:361-438), FP32 ------------------------
 * Conv1d(3 -> 3, kernel ks, stride) + ELU on [B][3][L_in] -> [B][3][L_out],
 * L_out = (L_in - ks) / stride + 1, weights [3][3][ks] (torch Conv1d
 * layout).  For the first layer pass the frame field q [B][L_in][3] with
 * q_mean / q_std_inv (standardizeQ, CROMnet.py:451-452) non-NULL; they are
 * NULL for later layers.  `out` holds post-ELU values, which the backward
 * uses. */
int crom_conv_forward(const float* in, int B, int L_in, int ks, int stride, const float* W, const float* bias,
                      float* out, const float* q_mean, const float* q_std_inv, void* stream);
/* dout: gradient of `out`; din (may be NULL): gradient of the layer input;
 * gW [3][3][ks], gb [3] are overwritten. */
int crom_conv_backward(const float* in, const float* out, const float* dout, int B, int L_in, int ks, int stride,
                       const float* W, float* din, float* gW, float* gb, const float* q_mean, const float* q_std_inv,
                       void* stream);
/* Linear + ELU: y [B][n_out] = ELU(x W^T + b), W [n_out][n_in]. */
int crom_dense_forward(const float* x, int B, int n_in, int n_out, const float* W, const float* bias, float* y,
                       void* stream);
int crom_dense_backward(const float* x, const float* y, const float* dy, int B, int n_in, int n_out, const float* W,
                        float* dx, float* gW, float* gb, void* stream);
/* Gradient of the frames' latents from the decoder's input gradient:
 * g [n_frames][lbl] = sum over the frame's points of dx[p][0 .. lbl) (the
 * latent is broadcast to every point, CROMnet.py:141-142). */
int crom_latent_grad(const float* dx, int n_frames, int points_per_frame, int in_dim, int lbl, float* g,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CROM_HIP_H */
