// rsx_gae.hip — generalized advantage estimation on the device: the
// trainer's backward GAE scan over an experience buffer (GGL::GAE::Compute,
// GigaLearnCPP/src/private/GigaLearnCPP/PPO/GAE.cpp:128-195), the job of the
// reference's ComputeGAEOnDevice (RocketSimCuda/include/RocketSimCudaTraining.h:
// 21-34, src/GAECuda.cu:247-303).
//
// The recurrence restarts at every terminal (normal or truncated), so the
// buffer splits into independent segments, each ending at a terminal or at the
// buffer's last step.  Each segment runs the reference's per-step arithmetic
// in its own lane, backwards, so every element is bit-identical to the CPU
// loop; the segments run in parallel.  What the CPU loop carries across
// segments is only the count of truncated steps already passed (which
// bootstrap value a truncation reads, GAE.cpp:152-162); here that is an
// exclusive prefix count of truncations, per 1024-step block and then across
// blocks.  The reward-clip metric sums |reward| per segment and then over
// segments in a fixed order (deterministic; not the CPU's single running sum,
// so equal to it within float rounding).
//
// Launches, all on the caller's stream, no host synchronisation and no
// memset (the scratch is fully written before it is read):
//   k_gae_block_counts   truncations per block                 (nb blocks)
//   k_gae_block_scan     exclusive scan over the blocks        (1 block)
//   k_gae_segments       the backward scans, per-block sums    (nb blocks)
//   k_gae_finalize       reward clip portion                   (1 block)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/rsx_gae.h"
#include "rsx_internal.h"

namespace {

constexpr int kBlock = 1024;           // steps per block (16 wavefronts)
constexpr int kWaves = kBlock / 64;
constexpr int8_t kNormal = 1, kTruncated = 2;

// Scratch layout (int32 words): [0] truncations in the buffer, [1] unused,
// [2, 2 + nb) truncations per block, then their exclusive prefix,
// then 2 * nb float partial sums (|reward|, |clipped reward|) per block.
constexpr int kScratchHead = 2;
size_t scratch_words(int nb) { return (size_t)kScratchHead + 4 * (size_t)nb; }

__device__ inline int lane_id() { return threadIdx.x & 63; }

// Exclusive prefix of `flag` over the block's threads, and the block total.
__device__ inline int block_prefix(bool flag, int* lds_waves, int& total) {
    const unsigned long long m = __ballot(flag);
    const int wave = threadIdx.x >> 6;
    const int below = __popcll(m & ((1ull << lane_id()) - 1ull));
    if (lane_id() == 0) lds_waves[wave] = __popcll(m);
    __syncthreads();
    int before = 0, sum = 0;
    for (int w = 0; w < kWaves; w++) {
        const int c = lds_waves[w];
        before += w < wave ? c : 0;
        sum += c;
    }
    __syncthreads();
    total = sum;
    return before + below;
}

__global__ void __launch_bounds__(kBlock) k_gae_block_counts(const int8_t* __restrict__ terms, int T,
                                                             int* __restrict__ scratch) {
    __shared__ int waves[kWaves];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool tr = i < T && terms[i] == kTruncated;
    int total = 0;
    block_prefix(tr, waves, total);
    if (threadIdx.x == 0) scratch[kScratchHead + blockIdx.x] = total;
}

// One workgroup: exclusive scan of the per-block counts, in chunks of 1024.
__global__ void __launch_bounds__(kBlock) k_gae_block_scan(int nb, int* __restrict__ scratch) {
    __shared__ int part[kBlock];
    const int* counts = scratch + kScratchHead;
    int* before = scratch + kScratchHead + nb;
    int carry = 0;
    for (int base = 0; base < nb; base += kBlock) {
        const int b = base + threadIdx.x;
        const int v = b < nb ? counts[b] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        // Hillis-Steele inclusive scan in LDS
        for (int off = 1; off < kBlock; off <<= 1) {
            const int add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (b < nb) before[b] = carry + part[threadIdx.x] - v;
        carry += part[kBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) scratch[0] = carry;  // truncations in the buffer
}

__global__ void __launch_bounds__(kBlock) k_gae_segments(const float* __restrict__ rews,
                                                         const int8_t* __restrict__ terms,
                                                         const float* __restrict__ vals,
                                                         const float* __restrict__ trunc_vals, int num_truncs,
                                                         float* __restrict__ adv, float* __restrict__ target,
                                                         float* __restrict__ rets, int T, int nb, rsx_gae_params p,
                                                         int* __restrict__ scratch) {
    __shared__ int waves[kWaves];
    __shared__ float sums[2][kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int8_t ti = i < T ? terms[i] : 0;
    int total = 0;
    const int trunc_in_block = block_prefix(i < T && ti == kTruncated, waves, total);
    float tot_rew = 0.f, tot_clip = 0.f;
    if (i < T && (ti != 0 || i == T - 1)) {
        // This lane's segment ends at step i.  GAE.cpp's loop reads
        // truncValPreds[numTruncs - 1 - truncCount] with truncCount the
        // truncations at later steps.
        const int trunc_before = scratch[kScratchHead + nb + blockIdx.x] + trunc_in_block;
        const int trunc_idx = num_truncs - scratch[0] + trunc_before;
        const float gamma = p.gamma, lambda = p.lambda, rstd = p.return_std, clip = p.clip_range;
        float prev_lambda = 0.f, prev_ret = 0.f;
        for (int step = i; step >= 0; step--) {
            const int8_t term = terms[step];
            if (step < i && term != 0) break;  // the previous segment's end
            const float done = term == kNormal ? 1.f : 0.f;
            float trunc = term == kTruncated ? 1.f : 0.f;
            float cur;
            if (rstd != 0.f) {
                cur = rews[step] / rstd;
                tot_rew += fabsf(cur);
                if (clip > 0.f) cur = fminf(fmaxf(cur, -clip), clip);
                tot_clip += fabsf(cur);
            } else {
                cur = rews[step];
                tot_rew += fabsf(cur);
            }
            float next;
            if (term == kTruncated) {
                // with fewer bootstrap values than truncated steps (the CPU
                // loop raises) the missing ones read as 0
                next = (trunc_idx >= 0 && trunc_idx < num_truncs) ? trunc_vals[trunc_idx] : 0.f;
            } else if (step >= T - 1) {
                next = 0.f;  // a last step without a terminal: no bootstrap, end of the return chain
                trunc = 1.f;
            } else {
                next = vals[step + 1];
            }
            const float pred_return = cur + gamma * next * (1.f - done);
            const float delta = pred_return - vals[step];
            const float cur_ret = rews[step] + prev_ret * gamma * (1.f - done) * (1.f - trunc);
            rets[step] = cur_ret;
            prev_lambda = delta + gamma * lambda * (1.f - done) * (1.f - trunc) * prev_lambda;
            adv[step] = prev_lambda;
            target[step] = vals[step] + prev_lambda;
            prev_ret = cur_ret;
        }
    }
    // per-block sums of the segments that end in this block, in thread order
    sums[0][threadIdx.x] = tot_rew;
    sums[1][threadIdx.x] = tot_clip;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sums[0][threadIdx.x] += sums[0][threadIdx.x + off];
            sums[1][threadIdx.x] += sums[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float* bs = (float*)(scratch + kScratchHead + 2 * nb);
        bs[2 * blockIdx.x] = sums[0][0];
        bs[2 * blockIdx.x + 1] = sums[1][0];
    }
    (void)total;
}

__global__ void __launch_bounds__(kBlock) k_gae_finalize(int nb, const int* __restrict__ scratch,
                                                         float* __restrict__ clip_portion) {
    __shared__ float sums[2][kBlock];
    const float* bs = (const float*)(scratch + kScratchHead + 2 * nb);
    float a = 0.f, c = 0.f;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        a += bs[2 * b];
        c += bs[2 * b + 1];
    }
    sums[0][threadIdx.x] = a;
    sums[1][threadIdx.x] = c;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sums[0][threadIdx.x] += sums[0][threadIdx.x + off];
            sums[1][threadIdx.x] += sums[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && clip_portion) {
        const float tr = sums[0][0], tc = sums[1][0];
        *clip_portion = (tr - tc) / fmaxf(tr, 1e-7f);
    }
}

rsx_status gae_launch(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                      const float* d_trunc_val_preds, int32_t num_truncs, float* d_advantages,
                      float* d_target_values, float* d_returns, float* d_rew_clip_portion, int32_t T,
                      const rsx_gae_params* params, int* scratch, hipStream_t st) {
    const int nb = (T + kBlock - 1) / kBlock;
    const rsx_gae_params p = *params;
    hipLaunchKernelGGL(k_gae_block_counts, dim3(nb), dim3(kBlock), 0, st, d_terminals, T, scratch);
    hipLaunchKernelGGL(k_gae_block_scan, dim3(1), dim3(kBlock), 0, st, nb, scratch);
    hipLaunchKernelGGL(k_gae_segments, dim3(nb), dim3(kBlock), 0, st, d_rews, d_terminals, d_val_preds,
                       d_trunc_val_preds, num_truncs, d_advantages, d_target_values, d_returns, T, nb, p, scratch);
    hipLaunchKernelGGL(k_gae_finalize, dim3(1), dim3(kBlock), 0, st, nb, (const int*)scratch, d_rew_clip_portion);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return rsx::record_failure(RSX_HIP_ERROR, hipGetErrorString(e));
    return RSX_OK;
}

rsx_status gae_check_args(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                          const float* d_trunc_val_preds, int32_t num_truncs, const float* d_advantages,
                          const float* d_target_values, const float* d_returns, int32_t T,
                          const rsx_gae_params* params) {
    if (!params) return rsx::record_failure(RSX_INVALID_ARG, "null GAE parameters");
    if (T < 0 || num_truncs < 0) return rsx::record_failure(RSX_INVALID_ARG, "negative step or truncation count");
    if (T > 0 && (!d_rews || !d_terminals || !d_val_preds || !d_advantages || !d_target_values || !d_returns))
        return rsx::record_failure(RSX_INVALID_ARG, "null device buffer");
    if (num_truncs > 0 && !d_trunc_val_preds)
        return rsx::record_failure(RSX_INVALID_ARG, "null truncation bootstrap values");
    return RSX_OK;
}

// Grow-only scratch of rsx_gae, per device.
struct GaeScratch {
    std::mutex m;
    std::vector<int*> buf;
    std::vector<size_t> words;
};
GaeScratch g_scratch;

constexpr int kNoAllocBlocks = (RSX_GAE_NOALLOC_MAX_STEPS + kBlock - 1) / kBlock;
constexpr int kMaxDevices = 64;

}  // namespace

// Scratch of rsx_gae_noalloc: the layout above for RSX_GAE_NOALLOC_MAX_STEPS
// steps, one copy per GPU for the life of the process.  In host code this
// name is only a stand-in for the symbol: kernels are given the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) int scratch[kScratchHead + 4 * kNoAllocBlocks];

namespace {

// Address of `scratch` on each device, looked up once per device.
std::mutex g_noalloc_m;
int* g_noalloc_ptr[kMaxDevices];

}  // namespace

namespace rsx {

// Looks up the address of rsx_gae_noalloc's scratch on device `dev`, which
// must be the current device; later calls return at once.
int32_t gae_noalloc_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices)
        return record_failure(RSX_INVALID_ARG, "device ordinal beyond the no-alloc GAE's device table");
    std::lock_guard<std::mutex> g(g_noalloc_m);
    if (g_noalloc_ptr[dev]) return RSX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return record_failure(RSX_HIP_ERROR, hipGetErrorString(e));
    if (!p) return record_failure(RSX_HIP_ERROR, "no device address for the no-alloc GAE scratch");
    g_noalloc_ptr[dev] = (int*)p;
    return RSX_OK;
}

}  // namespace rsx

extern "C" {

rsx_status rsx_gae(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                   const float* d_trunc_val_preds, int32_t num_truncs, float* d_advantages, float* d_target_values,
                   float* d_returns, float* d_rew_clip_portion, int32_t T, const rsx_gae_params* params,
                   void* stream) {
    rsx_status s = gae_check_args(d_rews, d_terminals, d_val_preds, d_trunc_val_preds, num_truncs, d_advantages,
                                  d_target_values, d_returns, T, params);
    if (s != RSX_OK || T == 0) return s;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return rsx::record_failure(RSX_HIP_ERROR, hipGetErrorString(e));
    const size_t need = scratch_words((T + kBlock - 1) / kBlock);
    int* buf = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch.m);
        if ((int)g_scratch.buf.size() <= dev) {
            g_scratch.buf.resize(dev + 1, nullptr);
            g_scratch.words.resize(dev + 1, 0);
        }
        if (g_scratch.words[dev] < need) {
            // the old buffer may still be read by work queued on other streams
            e = hipDeviceSynchronize();
            if (e == hipSuccess && g_scratch.buf[dev]) e = hipFree(g_scratch.buf[dev]);
            g_scratch.buf[dev] = nullptr;
            g_scratch.words[dev] = 0;
            if (e == hipSuccess) e = hipMalloc((void**)&g_scratch.buf[dev], need * sizeof(int));
            if (e != hipSuccess) return rsx::record_failure(RSX_HIP_ERROR, hipGetErrorString(e));
            g_scratch.words[dev] = need;
        }
        buf = g_scratch.buf[dev];
    }
    return gae_launch(d_rews, d_terminals, d_val_preds, d_trunc_val_preds, num_truncs, d_advantages, d_target_values,
                      d_returns, d_rew_clip_portion, T, params, buf, (hipStream_t)stream);
}

rsx_status rsx_gae_noalloc(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                           const float* d_trunc_val_preds, int32_t num_truncs, float* d_advantages,
                           float* d_target_values, float* d_returns, float* d_rew_clip_portion, int32_t T,
                           const rsx_gae_params* params, void* stream) {
    if (T > RSX_GAE_NOALLOC_MAX_STEPS)
        return rsx::record_failure(RSX_INVALID_ARG, "T exceeds RSX_GAE_NOALLOC_MAX_STEPS (use rsx_gae)");
    rsx_status s = gae_check_args(d_rews, d_terminals, d_val_preds, d_trunc_val_preds, num_truncs, d_advantages,
                                  d_target_values, d_returns, T, params);
    if (s != RSX_OK || T == 0) return s;
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return rsx::record_failure(RSX_HIP_ERROR, hipGetErrorString(e));
    if ((s = rsx::gae_noalloc_setup(dev)) != RSX_OK) return s;
    int* buf;
    {
        std::lock_guard<std::mutex> g(g_noalloc_m);
        buf = g_noalloc_ptr[dev];
    }
    // every scratch word the kernels read is written earlier in the same
    // call (block counts, their scan, the block sums): nothing to clear
    return gae_launch(d_rews, d_terminals, d_val_preds, d_trunc_val_preds, num_truncs, d_advantages, d_target_values,
                      d_returns, d_rew_clip_portion, T, params, buf, (hipStream_t)stream);
}

}  // extern "C"
