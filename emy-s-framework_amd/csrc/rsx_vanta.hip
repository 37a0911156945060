// rsx_vanta.hip — the VantaObs kernel (BuildVantaObs, RocketSimCudaApi.cuh:
// 597-640): per arena the ball-prediction slices, the pad and demo timers,
// then every car's coyote timers and its 366-float row.
//
// One lane per arena, 64-lane workgroups.  The arena and its Vanta words are
// read field-major (coalesced over the wavefront), the 360-tick ball-only
// rollout runs in the lane with the engine's own tick, and each row is
// written by the lane that owns the arena.  The reference splits the same
// work over an arena kernel and a player kernel (RocketSimCudaObs.cuh:70-186);
// one lane per arena keeps the per-arena state in registers between the two.
#include "rsx_kern_body.h"

namespace rsx {

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
    if (V.Q.predict) {
        // the tick context of a ball-only arena: no cars, so no simplex work,
        // no pooled queries and no cull
        gjk::Simplex sx;
        TickCtx T;
        T.cg = V.cg;
        T.bg = V.bg;
        T.mv = V.mv;
        T.pads = V.pads;
        T.dt = 1.f / kTickRate;
        T.sx = &sx;
        T.cull = sdf::Cull{};
        T.lds_act = nullptr;
        T.own = true;
        T.pool = nullptr;
        TickScratch<2> W;
        W.cull_age = 1 << 30;
        W.cull_counts[0] = W.cull_counts[1] = W.cull_counts[2] = 0;
        W.L.overflow = false;
        Arena<2> P;
        vanta::predict_ball(A.ball, T, W, P, S.pred_pos, S.pred_vel);
        S.has_pred = true;
    } else {
        S.has_pred = false;
    }
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
                        hipStream_t st) {
    if (!b.n) return hipSuccess;
    dim3 grid((b.n + 63) / 64), block(64);
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
