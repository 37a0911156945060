// rsx_vanta.hip — the VantaObs kernels (BuildVantaObs, RocketSimCudaApi.cuh:
// 597-640): per arena the ball-prediction slices (k_vanta_pred), then the
// pad and demo timers, every car's coyote timers and its 366-float row
// (k_vanta).
//
// One lane per arena, 64-lane workgroups.  The arena and its Vanta words are
// read field-major (coalesced over the wavefront), the 360-tick ball-only
// rollout runs in the lane with the engine's own tick compiled for an arena
// without cars, and each row is written by the lane that owns the arena.
// The reference splits the rows over a player kernel
// (RocketSimCudaObs.cuh:125-186); one lane per arena keeps the per-arena
// state in registers between the arena update and the rows.
#include "rsx_kern_body.h"

namespace rsx {

// The tick context of a ball-only arena: no cars, so no simplex work, no
// pooled queries and no cull.
__device__ inline void ball_only_ctx(TickCtx& T, const VantaArgs& V, gjk::Simplex* sx) {
    T.cg = V.cg;
    T.bg = V.bg;
    T.mv = V.mv;
    T.pads = V.pads;
    T.dt = 1.f / kTickRate;
    T.sx = sx;
    T.cull = sdf::Cull{};
    T.lds_act = nullptr;
    T.own = true;
    T.pool = nullptr;
}

// The ball-prediction rollouts (vantaArenaUpdateKernel's first half,
// RocketSimCudaObs.cuh:92-121), in passes so that no wavefront runs one
// lane's rollout while its others idle.  k_vanta_pred_step, one lane per
// arena: clears the Vanta words of an arena that was reset; keeps the slices
// when neither the tick nor the ball moved since the last rollout; when the
// ball is, bit for bit, the one the last rollout passed ts ticks after its
// start, lists the arena for k_vanta_pred_cont, which continues that rollout
// (rollout_continue: the same slices as a new rollout) with its four chain
// states on four lanes; otherwise lists it for k_vanta_pred_full, which runs
// the full rollouts, one per lane.  Both run in dense wavefronts and read the
// ball from the arena words.  `lists`: 2n + 2 ints, the new-rollout list
// [0, n), the continue list [n, 2n), then their two lengths.
__device__ inline vanta::BallQ arena_ball_q(const Col& col, int ball_base) {
    return vanta::BallQ{col.getv(ball_base + 0), col.getv(ball_base + 3), col.getv(ball_base + 6)};
}

template <int MAXC>
__device__ void store_rollout(const Col& vc, uint64_t tick, const vanta::BallQ& start, const V3* pos, const V3* vel,
                              const vanta::BallQ* chain, bool chains) {
    for (int k = 0; k < 3; k++) {
        vc.setv(vanta::kVaPredPos + 3 * k, pos[k]);
        vc.setv(vanta::kVaPredVel + 3 * k, vel[k]);
    }
    vc.set64(vanta::kVaT0, tick);
    vanta::set_q(vc, vanta::kVaStart, start);
    for (int k = 0; k < 4; k++) vanta::set_q(vc, vanta::kVaChain + 9 * k, chain[k]);
    const uint32_t hi = vc.get(vanta::kVaAvailHi) & ~vanta::kVaChainsBit;
    vc.set(vanta::kVaAvailHi, hi | (chains ? vanta::kVaChainsBit : 0u));
}

__global__ void k_vanta_count_zero(int* count) {
    if (threadIdx.x < 2) count[threadIdx.x] = 0;
}

// Appends `value` of every lane with `take` to list (length *count), one
// atomic per wavefront: the lanes' slots follow their order in the
// wavefront.  Every lane of the wavefront must call it.
__device__ inline void wave_append(int* list, int* count, bool take, int value) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

template <int MAXC>
__global__ void __launch_bounds__(64) k_vanta_pred_step(BucketArgs B, uint32_t* vwords, const uint8_t* reset,
                                                        VantaArgs V, int* fresh, int* cont, int* count) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    bool go_cont = false, go_fresh = false;
    if (a < B.n) {
        const Col col{B.words + a, B.n};
        const Col vc{vwords + a, B.n};
        if (reset && reset[B.global[a]]) vanta::clear<MAXC>(vc);
        const vanta::BallQ now = arena_ball_q(col, Layout<MAXC>::kBallBase);
        const uint64_t tick = col.get64(Layout<MAXC>::kArenaBase + kArTickLo);
        const bool chains = (vc.get(vanta::kVaAvailHi) & vanta::kVaChainsBit) != 0;
        const uint64_t t0 = vc.get64(vanta::kVaT0);
        const int ts = V.Q.ticks_between;
        const bool moved = !(chains && tick == t0 && vanta::same_ball(now, vanta::get_q(vc, vanta::kVaStart)));
        go_cont = moved && chains && vanta::chains_fit(ts) && tick == t0 + (uint64_t)ts &&
                  vanta::same_ball(now, vanta::get_q(vc, vanta::kVaChain));
        go_fresh = moved && !go_cont;
    }
    wave_append(cont, &count[1], go_cont, a);
    wave_append(fresh, &count[0], go_fresh, a);
}

// A ball-only arena's tick context with the pooled ball gather in `pool`.
__device__ inline void pooled_ball_ctx(TickCtx& T, const VantaArgs& V, gjk::Simplex* sx, uint8_t* pool, bool own) {
    ball_only_ctx(T, V, sx);
    T.pool = pool;
    T.own = own;
}

// The continued rollouts (rollout_continue), four lanes per listed arena:
// lane k advances the rollout's chain state k (its state at ts, 120 + ts,
// 240 + ts and 360 ticks) by ts ticks.  Lane 1's input is the first slice
// (the ball 120 ticks after the new start), lane 2's the second, and lane
// 3's output the third, so each lane stores what it read or computed.  One
// wavefront per workgroup; the ball's mesh gather is pooled over its lanes
// (body_mesh_pool), so lanes past the list's end run a copy of its last lane
// and store nothing.
template <int MAXC>
__global__ void __launch_bounds__(64) k_vanta_pred_cont(BucketArgs B, uint32_t* vwords, VantaArgs V, const int* cont,
                                                        const int* count) {
    __shared__ __align__(16) uint8_t pool[kGatEnd];
    const int n = 4 * count[1];
    if ((int)blockIdx.x * 64 >= n) return;  // the whole wavefront
    const int i = blockIdx.x * 64 + (int)threadIdx.x;
    const bool own = i < n;
    const int ii = own ? i : n - 1;
    const int a = cont[ii >> 2], k = ii & 3;
    const Col col{B.words + a, B.n};
    const Col vc{vwords + a, B.n};
    const int ts = V.Q.ticks_between;
    const vanta::BallQ in = vanta::get_q(vc, vanta::kVaChain + 9 * k);
    gjk::Simplex sx;
    TickCtx T;
    pooled_ball_ctx(T, V, &sx, pool, own);
    TickScratch<0> W;
    W.cull_age = 1 << 30;
    W.cull_counts[0] = W.cull_counts[1] = W.cull_counts[2] = 0;
    W.L.overflow = false;
    Arena<0> P;
    vanta::Roller R{T, W, P};
    const vanta::BallQ out = R.advance(in, ts);
    if (!own) return;
    vanta::set_q(vc, vanta::kVaChain + 9 * k, out);
    if (k == 0) {
        const uint64_t tick = col.get64(Layout<MAXC>::kArenaBase + kArTickLo);
        vc.set64(vanta::kVaT0, tick);
        vanta::set_q(vc, vanta::kVaStart, arena_ball_q(col, Layout<MAXC>::kBallBase));
        vc.set(vanta::kVaAvailHi, vc.get(vanta::kVaAvailHi) | vanta::kVaChainsBit);
    } else if (k < 3) {
        vc.setv(vanta::kVaPredPos + 3 * (k - 1), in.p * kBT2UU);
        vc.setv(vanta::kVaPredVel + 3 * (k - 1), in.v * kBT2UU);
    } else {
        vc.setv(vanta::kVaPredPos + 6, out.p * kBT2UU);
        vc.setv(vanta::kVaPredVel + 6, out.v * kBT2UU);
    }
}

// Arenas per wavefront of the new rollouts: enough wavefronts for about two
// per SIMD of the device (the kernel's occupancy), at least kMinRolloutApw
// arenas each.  A list of a few thousand arenas thus spreads each ball's
// pooled mesh work over 16 lanes instead of leaving most SIMDs idle.
constexpr int kMinRolloutApw = 4;
__device__ inline int rollout_apw(int n) {
    int k = kMinRolloutApw;
    while (k < 64 && (n + k - 1) / k > 2048) k *= 2;
    return k;
}

// One wavefront per workgroup, rollout_apw(n) listed arenas per wavefront.
// The ball's mesh gather and sphere-triangle tests are pooled over the
// wavefront's lanes (body_mesh_pool, in the workgroup's static LDS): the
// lanes that own no listed arena (past the wavefront's share or the list's
// end) take only their part of the pools and store nothing.
template <int MAXC>
__global__ void __launch_bounds__(64) k_vanta_pred_full(BucketArgs B, uint32_t* vwords, VantaArgs V, const int* work,
                                                        const int* count) {
    __shared__ __align__(16) uint8_t pool[kGatEnd];
    const int n = count[0];
    const int k = rollout_apw(n);
    if ((int)blockIdx.x * k >= n) return;  // the whole wavefront
    const int l = (int)threadIdx.x;
    const int i = blockIdx.x * k + (l & (k - 1));
    const bool own = l < k && i < n;
    const int a = work[i < n ? i : n - 1];
    const Col col{B.words + a, B.n};
    const Col vc{vwords + a, B.n};
    const vanta::BallQ now = arena_ball_q(col, Layout<MAXC>::kBallBase);
    const uint64_t tick = col.get64(Layout<MAXC>::kArenaBase + kArTickLo);
    const int ts = V.Q.ticks_between;
    gjk::Simplex sx;
    TickCtx T;
    pooled_ball_ctx(T, V, &sx, pool, own);
    TickScratch<0> W;
    W.cull_age = 1 << 30;
    W.cull_counts[0] = W.cull_counts[1] = W.cull_counts[2] = 0;
    W.L.overflow = false;
    Arena<0> P;
    vanta::Roller R{T, W, P};
    V3 pos[3], vel[3];
    vanta::BallQ chain[4];
    const bool fit = vanta::chains_fit(ts);
    vanta::rollout_full(R, now, fit ? ts : 0, pos, vel, chain);
    if (own) store_rollout<MAXC>(vc, tick, now, pos, vel, chain, fit);
}

// The rows: the arena update (pads, demo timers), then every car's coyote
// timers and row, reading the slices k_vanta_pred stored when V.Q.predict.
template <int MAXC>
__global__ void __launch_bounds__(64) k_vanta(BucketArgs B, uint32_t* vwords, const uint8_t* reset, VantaArgs V) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= B.n) return;
    const Col col{B.words + a, B.n};
    const Col vc{vwords + a, B.n};
    if (reset && reset[B.global[a]]) vanta::clear<MAXC>(vc);
    Arena<MAXC> A;
    unpack_arena<MAXC>(col, A, V.cg);
    vanta::State<MAXC> S;
    vanta::load<MAXC>(vc, S);
    S.has_pred = V.Q.predict != 0;
    vanta::update_arena<MAXC>(S, A, V.P, V.Q.dt);
    const uint32_t lt = vanta::last_touch_id<MAXC>(A);
    float* o = V.obs + B.obs_off[a];
    for (int i = 0; i < A.ncars; i++) vanta::build_row<MAXC>(A, S, V.P, V.Q, i, lt, o + (int64_t)i * vanta::kRow);
    vanta::store<MAXC>(vc, S, A.ncars);
}

template <int MAXC>
__global__ void __launch_bounds__(256) k_vanta_clear(uint32_t* vwords, int n, int only) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n || (only >= 0 && a != only)) return;
    vanta::clear<MAXC>(Col{vwords + a, n});
}

hipError_t launch_vanta(int maxc, const BucketArgs& b, uint32_t* vwords, const uint8_t* reset, const VantaArgs& v,
                        int* lists, hipStream_t st) {
    if (!b.n) return hipSuccess;
    dim3 grid((b.n + 63) / 64), grid4((4 * b.n + 63) / 64), block(64);
    dim3 grid_full((b.n + kMinRolloutApw - 1) / kMinRolloutApw);  // rollout_apw(n) >= kMinRolloutApw
    int* const fresh = lists;
    int* const cont = lists + b.n;
    int* const count = lists + 2 * (size_t)b.n;
    if (v.Q.predict) {
        // the rollouts clear the reset arenas' words first
        hipLaunchKernelGGL(k_vanta_count_zero, dim3(1), dim3(64), 0, st, count);
#define RSX_LAUNCH_PRED(M)                                                                                       \
    hipLaunchKernelGGL(k_vanta_pred_step<M>, grid, block, 0, st, b, vwords, reset, v, fresh, cont, count);     \
    hipLaunchKernelGGL(k_vanta_pred_cont<M>, grid4, block, 0, st, b, vwords, v, (const int*)cont,              \
                       (const int*)count);                                                                   \
    hipLaunchKernelGGL(k_vanta_pred_full<M>, grid_full, block, 0, st, b, vwords, v, (const int*)fresh,          \
                       (const int*)count)
        if (maxc == 2) {
            RSX_LAUNCH_PRED(2);
        } else if (maxc == 4) {
            RSX_LAUNCH_PRED(4);
        } else {
            RSX_LAUNCH_PRED(6);
        }
#undef RSX_LAUNCH_PRED
        const hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
        reset = nullptr;
    }
    if (maxc == 2) hipLaunchKernelGGL(k_vanta<2>, grid, block, 0, st, b, vwords, reset, v);
    else if (maxc == 4) hipLaunchKernelGGL(k_vanta<4>, grid, block, 0, st, b, vwords, reset, v);
    else hipLaunchKernelGGL(k_vanta<6>, grid, block, 0, st, b, vwords, reset, v);
    return hipGetLastError();
}

hipError_t launch_vanta_clear(int maxc, uint32_t* vwords, int n, int only, hipStream_t st) {
    if (!n) return hipSuccess;
    dim3 grid((n + 255) / 256), block(256);
    if (maxc == 2) hipLaunchKernelGGL(k_vanta_clear<2>, grid, block, 0, st, vwords, n, only);
    else if (maxc == 4) hipLaunchKernelGGL(k_vanta_clear<4>, grid, block, 0, st, vwords, n, only);
    else hipLaunchKernelGGL(k_vanta_clear<6>, grid, block, 0, st, vwords, n, only);
    return hipGetLastError();
}

}  // namespace rsx
