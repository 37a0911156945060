// rsx_k_mixed.hip — the fused env step of every team-size bucket in one
// launch (rsx_step_env when team sizes differ between arenas).  Each
// wavefront picks its bucket from its index (mixed_wave, rsx_kern_body.h):
// the 3v3 wavefronts first, then 2v2, then 1v1, so the longest work is
// dispatched first.
#include "rsx_kern_body.h"

namespace rsx {

__global__ void __launch_bounds__(256) k_env_mixed(BucketArgs b2, BucketArgs b4, BucketArgs b6, int w6, int w4,
                                                   WorldArgs G, EnvArgs V, int apw, int lds_wave) {
    mixed_wave(b2, b4, b6, w6, w4, G, V, apw, lds_wave);
}

namespace {
size_t g_lds_done = 65536;
}

hipError_t launch_env_mixed(const BucketArgs* b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                            hipStream_t st) {
    if (s.block == 512) return launch_env_mixed_wide(b, w, v, s, st);
    const int w2 = arena_waves(b[0].n, s), w4 = arena_waves(b[1].n, s), w6 = arena_waves(b[2].n, s);
    if (!(w2 + w4 + w6)) return hipSuccess;
    const size_t per_wave = arena_lds_per_wave(w.cull.enabled != 0, w.pool != 0);
    const size_t lds = per_wave * (s.block / 64);
    dim3 grid(blocks_for_waves(w2 + w4 + w6, s)), block(s.block);
    hipLaunchKernelGGL(k_env_mixed, grid, block, lds, st, b[0], b[1], b[2], w6, w4, w, v, s.arenas_per_wave,
                       (int)per_wave);
    return hipGetLastError();
}

hipError_t prepare_mixed_kernel(size_t bytes) {
    hipError_t e = set_lds_limit(k_env_mixed, bytes, g_lds_done);
    if (e == hipSuccess) e = prepare_mixed_wide_kernel(bytes);
    return e;
}

}  // namespace rsx
