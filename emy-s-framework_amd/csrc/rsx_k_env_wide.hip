// rsx_k_env_wide.hip — the fused env step of one team-size bucket in
// 512-lane workgroups: eight wavefronts share a CU, two per SIMD, so each
// lane gets at most 256 registers and the compiler keeps the rest in private
// memory.  Only the unpooled shape fits its LDS.  A translation unit of its
// own so that it compiles beside the 256-lane kernels.
#include "rsx_kern_body.h"

namespace rsx {

template <int MAXC>
__global__ void __launch_bounds__(512) k_env_step_wide(BucketArgs B, WorldArgs G, EnvArgs V, int apw, int lds_wave) {
    extern __shared__ __align__(16) uint8_t lds[];
    env_wave<MAXC>(B, G, V, wave_index(), apw, lds + (threadIdx.x >> 6) * lds_wave);
}

namespace {
size_t g_lds_done[3] = {65536, 65536, 65536};
}

hipError_t launch_env_wide(int maxc, const BucketArgs& b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                           hipStream_t st) {
    const size_t per_wave = arena_lds_per_wave(w.cull.enabled != 0, w.pool != 0);
    const size_t lds = per_wave * (s.block / 64);
    dim3 grid(blocks_for_waves(arena_waves(b.n, s), s)), block(s.block);
    if (maxc == 2)
        hipLaunchKernelGGL(k_env_step_wide<2>, grid, block, lds, st, b, w, v, s.arenas_per_wave, (int)per_wave);
    else if (maxc == 4)
        hipLaunchKernelGGL(k_env_step_wide<4>, grid, block, lds, st, b, w, v, s.arenas_per_wave, (int)per_wave);
    else
        hipLaunchKernelGGL(k_env_step_wide<6>, grid, block, lds, st, b, w, v, s.arenas_per_wave, (int)per_wave);
    return hipGetLastError();
}

hipError_t prepare_env_wide_kernels(size_t bytes) {
    hipError_t e = set_lds_limit(k_env_step_wide<2>, bytes, g_lds_done[0]);
    if (e == hipSuccess) e = set_lds_limit(k_env_step_wide<4>, bytes, g_lds_done[1]);
    if (e == hipSuccess) e = set_lds_limit(k_env_step_wide<6>, bytes, g_lds_done[2]);
    return e;
}

}  // namespace rsx
