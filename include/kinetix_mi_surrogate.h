/*
 * kinetix_mi_surrogate.h -- C ABI of the learned transport surrogate (no counterpart in the
 * reference; it sits beside kinetix::mixtureAvgTransportProps, benchmark/src/kinetix.cpp:815-841).
 *
 * Model: y = b(T, p, Y) + out_mean + out_std * MLP((x - in_mean) / in_std), with
 * y = (ln mu, ln lambda, ln D_1m .. ln D_Nm), x = (normalised T, normalised ln p, Y_1..Y_N)
 * padded to 64, b an FP32 physics baseline, MLP 64 -> 3x256 GELU -> 64. All buffers are HIP
 * device pointers; all calls are asynchronous on `stream` (NULL = default stream).
 */
#ifndef KINETIX_MI_SURROGATE_H
#define KINETIX_MI_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#include "kinetix_mi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mechanism-specific (generated) kernels ---- */

/* Standardised MLP inputs xin[n][64], baselines base[n][64] and out-of-hull flags[n]
 * (1 = an input lies outside [hull_lo, hull_hi] or is NaN). d_p: p [Pa] per state.
 * norm = {in_mean[64], in_std[64], hull_lo[64], hull_hi[64]} (FP32). */
int kx_sur_features(kx_mech_t m, int n_states, int offsetT, int offset, const double *d_p,
                    const double *d_state, const float *d_norm, float *d_xin, float *d_base,
                    int *d_flags, void *stream);

/* Training targets tgt[n][64] = (ln label - base - out_mean) / out_std from the plain
 * kernel's viscosity, conductivity and D_km (rows k*offset). ostat = {out_mean[64], out_std[64]}. */
int kx_sur_targets(kx_mech_t m, int n_states, int offset, const double *d_visc,
                   const double *d_cond, const double *d_Dkm, const float *d_base,
                   const float *d_ostat, float *d_tgt, void *stream);

/* Outputs from base + MLP residual for every state whose flag is 0 (flagged states are left
 * for the plain-kernel fallback). d_Dkm may be NULL. */
int kx_sur_outputs(kx_mech_t m, int n_states, int offsetT, int offset, const double *d_p,
                   const double *d_state, const float *d_base, const float *d_resid,
                   const float *d_ostat, const int *d_flags, double *d_viscosity,
                   double *d_conductivity, double *d_rhoD, double *d_Dkm, void *stream);

/* The whole inference path in one kernel: inputs, baselines and hull flags (flags[n], as in
 * kx_sur_features), the MLP on BF16 MFMA (weights from kx_mlp_pack_bf16, biases from the
 * FP32 parameter block; both 16-byte aligned) and the outputs of kx_sur_outputs, for every
 * state whose flag is 0. Flagged states are left untouched; if d_fb_queue is not NULL their
 * indices are appended to it at d_fb_queue[atomicAdd(d_fb_count, 1)]. d_Dkm may be NULL. */
int kx_sur_infer(kx_mech_t m, int n_states, int offsetT, int offset, const double *d_p,
                 const double *d_state, const float *d_norm, const float *d_ostat,
                 const uint16_t *d_weights_bf16, const float *d_params, int *d_flags,
                 double *d_viscosity, double *d_conductivity, double *d_rhoD, double *d_Dkm,
                 int *d_fb_queue, int *d_fb_count, void *stream);

/* Fallback statistics of one kx_sur_transport call. */
typedef struct {
  int n_states;
  int out_of_hull;          /* inputs outside the training box: plain kernel */
  int audited;              /* audit samples (computed by the plain kernel, which they keep) */
  int audit_failed;         /* audit samples off by more than the audit tolerances */
  int whole_call_fallback;  /* 1: audit failures above max_audit_fail, plain kernel everywhere */
  int fallback;             /* out_of_hull + audit_failed, or n_states after a whole-call fallback */
} kx_sur_stats;

/* Device workspace (bytes) kx_sur_transport needs for n_states states and this audit stride:
 * counters, per-state hull flags and the queue of state ids, about 8 bytes per state
 * (with_dkm does not change it). No allocation happens in the call. */
size_t kx_sur_workspace_bytes(kx_mech_t m, int n_states, int audit_stride, int with_dkm);

/* Audited surrogate transport, the drop-in beside kx_mixture_avg_transport_pstate:
 *  1. kx_sur_infer on every state (out-of-hull states queued);
 *  2. the plain FP64 kernel on every audit_stride-th state (from audit_offset % audit_stride;
 *     audit_stride <= 0: no audit) and on the queued states, reading them in place, and its
 *     results written over theirs. Audit samples are compared first; a sample fails if
 *       |mu/mu_plain - 1| > audit_tol, or |lambda/lambda_plain - 1| > audit_tol, or
 *       max_k |rhoD_k/rhoD_k,plain - 1| > audit_tol_rhoD, or any of them is NaN
 *     (the reference's own transport tolerance is 1e-3 for all three, bk.cpp:244-258);
 *  3. if failures exceed max_audit_fail * audited, the plain kernel for all states (decided
 *     on the device).
 * The audit is a sample: it bounds how often a state is off, not every state. Measured with
 * the shipped GRI-3.0 model (kinetix_amd/mechs/gri30/surrogate.npz) on 4,194,304 states over
 * T in [300, 2500] K, p in [0.5, 40] atm: see the tests of this call for the stated bounds.
 * stats == NULL: every step is enqueued on the stream and the call returns at once (no host
 * wait, no host read, no memset: the form for tight loops and stream capture into a graph).
 * stats != NULL: the call reads the counters and returns after the stream has drained. The
 * device work is the same.
 * Outputs as kx_mixture_avg_transport_pstate (rows k * offset); d_Dkm may be NULL.
 * Calls that overlap on one workspace race. */
int kx_sur_transport(kx_mech_t m, int n_states, int offsetT, int offset, const double *d_p,
                     const double *d_state, const float *d_norm, const float *d_ostat,
                     const uint16_t *d_weights_bf16, const float *d_params, int audit_stride,
                     int audit_offset, double audit_tol, double audit_tol_rhoD,
                     double max_audit_fail, double *d_viscosity, double *d_conductivity,
                     double *d_rhoD, double *d_Dkm, void *d_workspace, size_t workspace_bytes,
                     kx_sur_stats *stats, void *stream);

/* Size of the device scratch array behind kx_sur_transport_noalloc (64 MiB per GPU, held
 * from the first kx_init on that GPU until the process ends). kx_sur_workspace_bytes(n,
 * audit_stride) must not exceed it: up to 5,592,384 states at audit_stride 1 and 8,346,816
 * at the default stride of 100. */
#define KX_SUR_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). kx_sur_transport
 * with the workspace taken from a library-owned device array (KX_SUR_SCRATCH_BYTES per GPU)
 * instead of an argument; it calls no hipMalloc/hipFree and allocates nothing. Calls beyond
 * the scratch size return KX_ERR_TOO_LARGE. Every call on one GPU shares that GPU's scratch,
 * which keeps the last call's contents (each call initialises what it reads): calls that
 * overlap on one GPU (other streams, other threads) race -- serialise them, or use
 * kx_sur_transport with a workspace each. Same arguments and results otherwise. */
int kx_sur_transport_noalloc(kx_mech_t m, int n_states, int offsetT, int offset, const double *d_p,
                             const double *d_state, const float *d_norm, const float *d_ostat,
                             const uint16_t *d_weights_bf16, const float *d_params,
                             int audit_stride, int audit_offset, double audit_tol,
                             double audit_tol_rhoD, double max_audit_fail, double *d_viscosity,
                             double *d_conductivity, double *d_rhoD, double *d_Dkm,
                             kx_sur_stats *stats, void *stream);

/* ---- mechanism-independent MLP kernels ---- */

int kx_mlp_width(void);             /* hidden width H (256) */
int kx_mlp_n_params(void);          /* FP32 parameter block length */
int kx_mlp_n_weights_bf16(void);    /* BF16 weight block length */

/* Training forward pass, batch B (multiple of 64): acts receives Z1 H1 Z2 H2 Z3 H3 (B x H
 * each) and Y (B x 64). */
int kx_mlp_forward(const float *params, int B, const float *X, float *acts, void *stream);

/* Gradient of mean((Y - T)^2) over the first 55 outputs into grads (params layout).
 * work: 2 x B x H floats; loss_part: ceil(B * 64 / 256) floats of partial squared errors. */
int kx_mlp_backward(const float *params, int B, const float *X, const float *acts,
                    const float *T, float *grads, float *work, float *loss_part, void *stream);

/* Adam step `step` (1-based) over the whole parameter block. */
int kx_mlp_adam(float *params, const float *grads, float *m, float *v, float lr, float beta1,
                float beta2, float eps, int step, void *stream);

/* BF16 copy of the weight matrices (round to nearest even) in the MFMA fragment order of
 * the inference kernels (kx_mlp_n_weights_bf16() elements). */
int kx_mlp_pack_bf16(const float *params, uint16_t *weights_bf16, void *stream);

/* Inference Y[n][64] = MLP(X[n][64]) on BF16 MFMA with FP32 accumulation. */
int kx_mlp_infer_bf16(const uint16_t *weights_bf16, const float *params, int n_states,
                      const float *X, float *Y, void *stream);

/* Device self-tests of the v_mfma_f32_16x16x16_bf16 and v_mfma_f32_16x16x32_bf16 lane
 * layouts (256 floats each: D[i][j] at i * 16 + j). */
int kx_selftest_mfma_bf16_layout(float *d_out, void *stream);
int kx_selftest_mfma_bf16_k32_layout(float *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* KINETIX_MI_SURROGATE_H */
