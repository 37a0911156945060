/*
 * rsx_gae.h — generalized advantage estimation on the device (librsx.so,
 * gfx950), the step after the env in a PPO trainer: the backward GAE scan
 * over an experience buffer.
 *
 * Replaces rsc::ComputeGAEOnDevice (RocketSimCuda/include/RocketSimCudaTraining.h:
 * 21-34, src/GAECuda.cu:247-303) with the same arguments and the arithmetic
 * of the trainer's CPU loop GGL::GAE::Compute (GigaLearnCPP/src/private/
 * GigaLearnCPP/PPO/GAE.cpp:128-195), per element bit for bit.
 *
 * All pointers are device memory on the current HIP device; `stream` is a
 * hipStream_t (NULL: the default stream); the call is asynchronous on it.
 */
#ifndef RSX_GAE_H
#define RSX_GAE_H

#include <stdint.h>

#include "rsx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* DeviceGAEParams (RocketSimCudaTraining.h:7-12). */
typedef struct rsx_gae_params {
    float gamma;       /* discount */
    float lambda;      /* GAE lambda */
    float return_std;  /* != 0: rewards are divided by it (and then clipped) */
    float clip_range;  /* > 0 with return_std != 0: clip the standardized rewards to +-clip_range */
} rsx_gae_params;

/* ComputeGAEOnDevice.  Inputs, T steps in buffer order: rewards, terminals
 * (int8: 0 none, 1 normal episode end, 2 truncation), value predictions; and
 * the num_truncs bootstrap values of the truncated steps in buffer order
 * (GAE.cpp reads them from the back while walking backwards).  num_truncs
 * must equal the number of truncated steps (the CPU loop raises otherwise;
 * here a missing value reads as 0).  A last step that is not a terminal
 * bootstraps from 0 (GAE.cpp:163-168).  Outputs (T each): advantages,
 * target values (value + advantage), returns (the discounted reward sum
 * within an episode); d_rew_clip_portion (nullable, one float): the share of
 * |standardized reward| removed by the clip.  Uses a grow-only scratch buffer
 * per device (24 bytes per 8192 steps), allocated on first use or
 * growth (then the device is synchronized). */
rsx_status rsx_gae(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                   const float* d_trunc_val_preds, int32_t num_truncs, float* d_advantages, float* d_target_values,
                   float* d_returns, float* d_rew_clip_portion, int32_t T, const rsx_gae_params* params,
                   void* stream);

/* Largest T that rsx_gae_noalloc takes (its fixed scratch: 24 bytes per
 * 8192 steps, 192 KiB per GPU). */
#define RSX_GAE_NOALLOC_MAX_STEPS (1 << 26)

/* No counterpart in the reference: rsx_gae for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as rsx_gae, but no
 * hipMalloc or hipFree and no device synchronization: the scratch is a fixed
 * array in the library's device memory, one copy per GPU held until the
 * process ends, whose address is looked up once per device (by rsx_create
 * or by the first call on that device; make one call before a capture on a
 * device without an engine).  T above RSX_GAE_NOALLOC_MAX_STEPS is refused
 * with RSX_INVALID_ARG before any device work.  Every call on one GPU shares
 * that GPU's array and each call writes all of the array it reads, so calls
 * in order on one stream are safe; calls that overlap on one GPU (other
 * streams or threads) race. */
rsx_status rsx_gae_noalloc(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                           const float* d_trunc_val_preds, int32_t num_truncs, float* d_advantages,
                           float* d_target_values, float* d_returns, float* d_rew_clip_portion, int32_t T,
                           const rsx_gae_params* params, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RSX_GAE_H */
