// rsx_k_env.hip — the fused env step of one team-size bucket
// (rsx_step_env when every arena has the same team size).
#include "rsx_kern_body.h"

namespace rsx {

template <int MAXC>
__global__ void __launch_bounds__(256) k_env_step(BucketArgs B, WorldArgs G, EnvArgs V, int apw, int lds_wave) {
    extern __shared__ __align__(16) uint8_t lds[];
    env_wave<MAXC>(B, G, V, wave_index(), apw, lds + (threadIdx.x >> 6) * lds_wave);
}

namespace {
size_t g_lds_done[3] = {65536, 65536, 65536};
}

hipError_t launch_env(int maxc, const BucketArgs& b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                      hipStream_t st) {
    if (!b.n) return hipSuccess;
    const size_t per_wave = arena_lds_per_wave(w.cull.enabled != 0, w.pool != 0);
    const size_t lds = per_wave * (s.block / 64);
    dim3 grid(blocks_for_waves(arena_waves(b.n, s), s)), block(s.block);
    if (s.block == 512) return launch_env_wide(maxc, b, w, v, s, st);
    if (maxc == 2)
        hipLaunchKernelGGL(k_env_step<2>, grid, block, lds, st, b, w, v, s.arenas_per_wave, (int)per_wave);
    else if (maxc == 4)
        hipLaunchKernelGGL(k_env_step<4>, grid, block, lds, st, b, w, v, s.arenas_per_wave, (int)per_wave);
    else
        hipLaunchKernelGGL(k_env_step<6>, grid, block, lds, st, b, w, v, s.arenas_per_wave, (int)per_wave);
    return hipGetLastError();
}

hipError_t prepare_env_kernels(size_t bytes) {
    hipError_t e = set_lds_limit(k_env_step<2>, bytes, g_lds_done[0]);
    if (e == hipSuccess) e = set_lds_limit(k_env_step<4>, bytes, g_lds_done[1]);
    if (e == hipSuccess) e = set_lds_limit(k_env_step<6>, bytes, g_lds_done[2]);
    if (e == hipSuccess) e = prepare_env_wide_kernels(bytes);
    return e;
}

}  // namespace rsx
