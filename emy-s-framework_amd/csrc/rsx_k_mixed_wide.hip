// rsx_k_mixed_wide.hip — the single launch over every team-size bucket in
// 512-lane workgroups (two wavefronts per SIMD, at most 256 registers a
// lane); a translation unit of its own so that it compiles beside the
// 256-lane kernels.
#include "rsx_kern_body.h"

namespace rsx {

__global__ void __launch_bounds__(512) k_env_mixed_wide(BucketArgs b2, BucketArgs b4, BucketArgs b6, int w6, int w4,
                                                        WorldArgs G, EnvArgs V, int apw, int lds_wave) {
    mixed_wave(b2, b4, b6, w6, w4, G, V, apw, lds_wave);
}

namespace {
size_t g_lds_done = 65536;
}

hipError_t launch_env_mixed_wide(const BucketArgs* b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                                 hipStream_t st) {
    const int w2 = arena_waves(b[0].n, s), w4 = arena_waves(b[1].n, s), w6 = arena_waves(b[2].n, s);
    if (!(w2 + w4 + w6)) return hipSuccess;
    const size_t per_wave = arena_lds_per_wave(w.cull.enabled != 0, w.pool != 0);
    const size_t lds = per_wave * (s.block / 64);
    dim3 grid(blocks_for_waves(w2 + w4 + w6, s)), block(s.block);
    hipLaunchKernelGGL(k_env_mixed_wide, grid, block, lds, st, b[0], b[1], b[2], w6, w4, w, v, s.arenas_per_wave,
                       (int)per_wave);
    return hipGetLastError();
}

hipError_t prepare_mixed_wide_kernel(size_t bytes) { return set_lds_limit(k_env_mixed_wide, bytes, g_lds_done); }

}  // namespace rsx
