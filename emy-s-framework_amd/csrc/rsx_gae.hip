// rsx_gae.hip — generalized advantage estimation on the device: the
// trainer's backward GAE scan over an experience buffer (GGL::GAE::Compute,
// GigaLearnCPP/src/private/GigaLearnCPP/PPO/GAE.cpp:128-195), the job of the
// reference's ComputeGAEOnDevice (RocketSimCuda/include/RocketSimCudaTraining.h:
// 21-34, src/GAECuda.cu:247-303).
//
// The recurrence restarts at every terminal (normal or truncated), so the
// buffer splits into independent segments, each ending at a terminal or at the
// buffer's last step.  Each segment runs the reference's per-step arithmetic
// in one lane, backwards, so every element is bit-identical to the CPU loop;
// the segments run in parallel.  What the CPU loop carries across segments is
// only the count of truncated steps already passed (which bootstrap value a
// truncation reads, GAE.cpp:152-162); here that is an exclusive prefix count
// of truncations, per 8192-step region and then across regions.
//
// A workgroup owns a region: it lists the region's segment ends in LDS (the
// dense mapping of the reference's segment list, without a global list), and
// its lanes take those segments in turn, so a region with one long segment
// and one with thousands of short ones both keep their lanes busy.  A
// segment starts after the previous end (the list's previous entry, or the
// last terminal before the region, from a max-scan over regions), so a lane
// knows its whole segment before reading it and issues the loads of 16 steps
// at a time ahead of their arithmetic (inside a segment every factor
// (1 - done), (1 - trunc) is exactly 1).  The reward-clip metric sums
// |reward| per lane and then over lanes and regions in a fixed order
// (deterministic; not the CPU's single running sum, so equal to it within
// float rounding).
//
// Launches, all on the caller's stream, no host synchronisation and no
// memset (the scratch is fully written before it is read):
//   k_gae_region_counts   truncations and last terminal per region  (nb workgroups)
//   k_gae_region_scan     their exclusive scans over the regions    (1 workgroup)
//   k_gae_segments        the backward scans, per-region sums       (nb workgroups)
//   k_gae_finalize        reward clip portion                       (1 workgroup)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/rsx_gae.h"
#include "rsx_internal.h"

namespace {

// A workgroup owns a region of kRegion consecutive steps: kThreads lanes,
// kPer consecutive steps each.
constexpr int kThreads = 512;
constexpr int kPer = 16;
constexpr int kRegion = kThreads * kPer;  // 8192 steps
constexpr int kWaves = kThreads / 64;
constexpr int8_t kNormal = 1, kTruncated = 2;

// Scratch layout (int32 words): [0] truncations in the buffer, [1] unused,
// then per region (nb of them): truncations, their exclusive prefix, two
// float partial sums (|reward|, |clipped reward|), the region's last terminal
// step (-1: none), the last terminal step before the region (-1: none).
constexpr int kScratchHead = 2;
constexpr size_t kWordsPerRegion = 6;
size_t scratch_words(int nb) { return (size_t)kScratchHead + kWordsPerRegion * (size_t)nb; }
__device__ inline int* trunc_count(int* s) { return s + kScratchHead; }
__device__ inline int* trunc_prefix(int* s, int nb) { return s + kScratchHead + nb; }
__device__ inline float* region_sums(int* s, int nb) { return (float*)(s + kScratchHead + 2 * nb); }
__device__ inline int* region_last(int* s, int nb) { return s + kScratchHead + 4 * nb; }
__device__ inline int* last_before(int* s, int nb) { return s + kScratchHead + 5 * nb; }
constexpr int kChunk = 16;  // steps whose loads a lane issues together

__device__ inline int lane_id() { return threadIdx.x & 63; }

// Exclusive prefix sum of v over the workgroup's lanes, and the total.
__device__ inline int wg_exscan_sum(int v, int* lds, int& total) {
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane_id() >= off) x += y;
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 63) lds[wave] = x;
    __syncthreads();
    int before = 0, sum = 0;
    for (int w = 0; w < kWaves; w++) {
        const int c = lds[w];
        before += w < wave ? c : 0;
        sum += c;
    }
    __syncthreads();
    total = sum;
    return before + x - v;
}

// Exclusive prefix max of v over the workgroup's lanes (-1 before the first),
// and the overall max.
__device__ inline int wg_exscan_max(int v, int* lds, int& all) {
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane_id() >= off) x = max(x, y);
    }
    int ex = __shfl_up(x, 1);
    if (lane_id() == 0) ex = -1;
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 63) lds[wave] = x;
    __syncthreads();
    int before = -1, m = -1;
    for (int w = 0; w < kWaves; w++) {
        const int c = lds[w];
        before = w < wave ? max(before, c) : before;
        m = max(m, c);
    }
    __syncthreads();
    all = m;
    return max(before, ex);
}

// This lane's kPer steps: which end a segment (a terminal, or the buffer's
// last step), which are truncations, and the last terminal among them.
struct LaneSteps {
    uint32_t ends, truncs;
    int last_term;
};

__device__ inline LaneSteps lane_steps(const int8_t* __restrict__ terms, int T, int first) {
    LaneSteps r{0u, 0u, -1};
    int8_t t[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) t[k] = first + k < T ? terms[first + k] : 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int step = first + k;
        if (step >= T) break;
        if (t[k] != 0 || step == T - 1) r.ends |= 1u << k;
        if (t[k] == kTruncated) r.truncs |= 1u << k;
        if (t[k] != 0) r.last_term = step;
    }
    return r;
}

__global__ void __launch_bounds__(kThreads) k_gae_region_counts(const int8_t* __restrict__ terms, int T, int nb,
                                                               int* __restrict__ scratch) {
    __shared__ int lds[kWaves];
    const LaneSteps ls = lane_steps(terms, T, blockIdx.x * kRegion + threadIdx.x * kPer);
    int total = 0, last = -1;
    wg_exscan_sum(__popc(ls.truncs), lds, total);
    wg_exscan_max(ls.last_term, lds, last);
    if (threadIdx.x == 0) {
        trunc_count(scratch)[blockIdx.x] = total;
        region_last(scratch, nb)[blockIdx.x] = last;
    }
}

// One workgroup: exclusive scans over the regions, kThreads at a time (a sum
// of the truncation counts, a max of the last terminals).
__global__ void __launch_bounds__(kThreads) k_gae_region_scan(int nb, int* __restrict__ scratch) {
    __shared__ int lds[kWaves];
    const int* counts = trunc_count(scratch);
    int* before = trunc_prefix(scratch, nb);
    const int* lasts = region_last(scratch, nb);
    int* lb = last_before(scratch, nb);
    int carry = 0, carry_max = -1;
    for (int base = 0; base < nb; base += kThreads) {
        const int b = base + threadIdx.x;
        int tot = 0, mx = -1;
        const int ex = wg_exscan_sum(b < nb ? counts[b] : 0, lds, tot);
        const int exm = wg_exscan_max(b < nb ? lasts[b] : -1, lds, mx);
        if (b < nb) {
            before[b] = carry + ex;
            lb[b] = max(carry_max, exm);
        }
        carry += tot;
        carry_max = max(carry_max, mx);
    }
    if (threadIdx.x == 0) scratch[0] = carry;  // truncations in the buffer
}

// The reward as the step's arithmetic uses it (standardized and clipped
// when return_std is set); adds its contributions to the clip metric (the
// clipped sum stays 0 without return_std, as in GAE.cpp).  Both sums are
// updated on every path so that they stay in registers.
__device__ inline float gae_reward(float r, float rstd, float clip, float& tot_rew, float& tot_clip) {
    float cur = r, a_rew = fabsf(r), a_clip = 0.f;
    if (rstd != 0.f) {
        cur = r / rstd;
        a_rew = fabsf(cur);
        if (clip > 0.f) cur = fminf(fmaxf(cur, -clip), clip);
        a_clip = fabsf(cur);
    }
    tot_rew += a_rew;
    tot_clip += a_clip;
    return cur;
}

// One segment, steps start .. end, backwards: the CPU loop's arithmetic.
// trunc_idx: the bootstrap value a truncation at `end` reads.
__device__ inline void gae_segment(const float* __restrict__ rews, const float* __restrict__ vals,
                                   const float* __restrict__ trunc_vals, int num_truncs, float* __restrict__ adv,
                                   float* __restrict__ target, float* __restrict__ rets, int T, int8_t te, int start,
                                   int end, int trunc_idx, const rsx_gae_params& p, float& tot_rew,
                                   float& tot_clip) {
    const float gamma = p.gamma, lambda = p.lambda, rstd = p.return_std, clip = p.clip_range;
    float prev_lambda, prev_ret, next_val;
    {
        // the segment's last step: the general arithmetic of GAE.cpp:128-195
        const float done = te == kNormal ? 1.f : 0.f;
        float trunc = te == kTruncated ? 1.f : 0.f;
        const float r = rews[end], v = vals[end];
        const float cur = gae_reward(r, rstd, clip, tot_rew, tot_clip);
        float next;
        if (te == kTruncated) {
            // with fewer bootstrap values than truncated steps (the CPU loop
            // raises) the missing ones read as 0
            next = (trunc_idx >= 0 && trunc_idx < num_truncs) ? trunc_vals[trunc_idx] : 0.f;
        } else if (end >= T - 1) {
            next = 0.f;  // a last step without a terminal: no bootstrap, end of the return chain
            trunc = 1.f;
        } else {
            next = vals[end + 1];
        }
        const float pred_return = cur + gamma * next * (1.f - done);
        const float delta = pred_return - v;
        const float zero = 0.f;  // the later segment's carry times (1 - done) (1 - trunc) = 0
        prev_ret = r + zero * gamma * (1.f - done) * (1.f - trunc);
        prev_lambda = delta + gamma * lambda * (1.f - done) * (1.f - trunc) * zero;
        rets[end] = prev_ret;
        adv[end] = prev_lambda;
        target[end] = v + prev_lambda;
        next_val = v;
    }
    // the steps before it carry no terminal: done = trunc = 0, so every
    // (1 - done) and (1 - trunc) factor is exactly 1 and drops out
    const float gl = gamma * lambda;
    for (int top = end - 1; top >= start; top -= kChunk) {
        const int n = min(kChunk, top - start + 1);
        float rr[kChunk], vv[kChunk];
        // every load issued (past the segment's start they re-read its first
        // step, which is in bounds), none waits on a branch
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            const int s = max(top - k, start);
            rr[k] = rews[s];
            vv[k] = vals[s];
        }
#pragma unroll
        for (int k = 0; k < kChunk; k++) {
            if (k < n) {
                const int step = top - k;
                const float cur = gae_reward(rr[k], rstd, clip, tot_rew, tot_clip);
                const float delta = (cur + gamma * next_val) - vv[k];
                prev_ret = rr[k] + prev_ret * gamma;
                prev_lambda = delta + gl * prev_lambda;
                rets[step] = prev_ret;
                adv[step] = prev_lambda;
                target[step] = vv[k] + prev_lambda;
                next_val = vv[k];
            }
        }
    }
}

// The segments that end in this workgroup's region: the region's segment
// ends are listed in LDS in step order, then lane l takes ends l, l +
// kThreads, ...  A segment starts after the previous end (the list's
// previous entry, or the last terminal before the region).
__global__ void __launch_bounds__(kThreads) k_gae_segments(const float* __restrict__ rews,
                                                           const int8_t* __restrict__ terms,
                                                           const float* __restrict__ vals,
                                                           const float* __restrict__ trunc_vals, int num_truncs,
                                                           float* __restrict__ adv, float* __restrict__ target,
                                                           float* __restrict__ rets, int T, int nb, rsx_gae_params p,
                                                           int* __restrict__ scratch) {
    __shared__ int lds[kWaves];
    __shared__ uint16_t ends[kRegion];  // region offset of each segment end
    __shared__ uint16_t tpre[kRegion];  // truncations among the region's earlier ends
    __shared__ float sums[2][kThreads];
    const int base = blockIdx.x * kRegion;
    const int first = base + threadIdx.x * kPer;
    const LaneSteps ls = lane_steps(terms, T, first);
    int n_ends = 0, n_tr = 0;
    int pos = wg_exscan_sum(__popc(ls.ends), lds, n_ends);
    int tr = wg_exscan_sum(__popc(ls.truncs), lds, n_tr);
    for (int k = 0; k < kPer; k++) {
        if (ls.ends & (1u << k)) {
            ends[pos] = (uint16_t)(threadIdx.x * kPer + k);
            tpre[pos] = (uint16_t)tr;
            pos++;
        }
        if (ls.truncs & (1u << k)) tr++;
    }
    __syncthreads();
    const int region_before = last_before(scratch, nb)[blockIdx.x];
    const int trunc_base = num_truncs - scratch[0] + trunc_prefix(scratch, nb)[blockIdx.x];
    float tot_rew = 0.f, tot_clip = 0.f;
    for (int e = threadIdx.x; e < n_ends; e += kThreads) {
        const int end = base + ends[e];
        const int start = e > 0 ? base + ends[e - 1] + 1 : region_before + 1;
        // GAE.cpp reads truncValPreds[numTruncs - 1 - truncCount], truncCount
        // the truncations at later steps
        gae_segment(rews, vals, trunc_vals, num_truncs, adv, target, rets, T, terms[end], start, end,
                    trunc_base + tpre[e], p, tot_rew, tot_clip);
    }
    // per-region sums, reduced in lane order
    sums[0][threadIdx.x] = tot_rew;
    sums[1][threadIdx.x] = tot_clip;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sums[0][threadIdx.x] += sums[0][threadIdx.x + off];
            sums[1][threadIdx.x] += sums[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float* rs = region_sums(scratch, nb);
        rs[2 * blockIdx.x] = sums[0][0];
        rs[2 * blockIdx.x + 1] = sums[1][0];
    }
}

__global__ void __launch_bounds__(kThreads) k_gae_finalize(int nb, int* __restrict__ scratch,
                                                           float* __restrict__ clip_portion) {
    __shared__ float sums[2][kThreads];
    const float* rs = region_sums(scratch, nb);
    float a = 0.f, c = 0.f;
    for (int b = threadIdx.x; b < nb; b += kThreads) {
        a += rs[2 * b];
        c += rs[2 * b + 1];
    }
    sums[0][threadIdx.x] = a;
    sums[1][threadIdx.x] = c;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
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

int regions_of(int T) { return (T + kRegion - 1) / kRegion; }

rsx_status gae_launch(const float* d_rews, const int8_t* d_terminals, const float* d_val_preds,
                      const float* d_trunc_val_preds, int32_t num_truncs, float* d_advantages,
                      float* d_target_values, float* d_returns, float* d_rew_clip_portion, int32_t T,
                      const rsx_gae_params* params, int* scratch, hipStream_t st) {
    const int nb = regions_of(T);
    const rsx_gae_params p = *params;
    hipLaunchKernelGGL(k_gae_region_counts, dim3(nb), dim3(kThreads), 0, st, d_terminals, T, nb, scratch);
    hipLaunchKernelGGL(k_gae_region_scan, dim3(1), dim3(kThreads), 0, st, nb, scratch);
    hipLaunchKernelGGL(k_gae_segments, dim3(nb), dim3(kThreads), 0, st, d_rews, d_terminals, d_val_preds,
                       d_trunc_val_preds, num_truncs, d_advantages, d_target_values, d_returns, T, nb, p, scratch);
    hipLaunchKernelGGL(k_gae_finalize, dim3(1), dim3(kThreads), 0, st, nb, scratch, d_rew_clip_portion);
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

constexpr int kNoAllocRegions = (RSX_GAE_NOALLOC_MAX_STEPS + kRegion - 1) / kRegion;
constexpr int kMaxDevices = 64;

}  // namespace

// Scratch of rsx_gae_noalloc: the layout above for RSX_GAE_NOALLOC_MAX_STEPS
// steps, one copy per GPU for the life of the process.  In host code this
// name is only a stand-in for the symbol: kernels are given the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) int scratch[kScratchHead + kWordsPerRegion * kNoAllocRegions];

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
    const size_t need = scratch_words(regions_of(T));
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
