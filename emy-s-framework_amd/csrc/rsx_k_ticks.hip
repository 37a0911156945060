// rsx_k_ticks.hip — `ticks` physics ticks of one team-size bucket under the
// stored controls (rsx_step_ticks).
#include "rsx_kern_body.h"

namespace rsx {

template <int MAXC>
__global__ void __launch_bounds__(256) k_step(BucketArgs B, WorldArgs G, int ticks, int apw, int lds_wave) {
    extern __shared__ __align__(16) uint8_t lds[];
    ticks_wave<MAXC>(B, G, ticks, wave_index(), apw, lds + (threadIdx.x >> 6) * lds_wave);
}

namespace {
size_t g_lds_done[3] = {65536, 65536, 65536};
}

hipError_t launch_ticks(int maxc, const BucketArgs& b, const WorldArgs& w, int ticks, const LaunchShape& shape,
                        hipStream_t st) {
    if (!b.n) return hipSuccess;
    // the tick path (parity, savestates) runs 512-lane shapes as 256-lane
    // workgroups: the same wavefronts with the same arenas each
    LaunchShape s = shape;
    if (s.block > 256) s.block = 256;
    const size_t per_wave = arena_lds_per_wave(w.cull.enabled != 0, w.pool != 0);
    const size_t lds = per_wave * (s.block / 64);
    dim3 grid(blocks_for_waves(arena_waves(b.n, s), s)), block(s.block);
    if (maxc == 2)
        hipLaunchKernelGGL(k_step<2>, grid, block, lds, st, b, w, ticks, s.arenas_per_wave, (int)per_wave);
    else if (maxc == 4)
        hipLaunchKernelGGL(k_step<4>, grid, block, lds, st, b, w, ticks, s.arenas_per_wave, (int)per_wave);
    else
        hipLaunchKernelGGL(k_step<6>, grid, block, lds, st, b, w, ticks, s.arenas_per_wave, (int)per_wave);
    return hipGetLastError();
}

hipError_t prepare_tick_kernels(size_t bytes) {
    hipError_t e = set_lds_limit(k_step<2>, bytes, g_lds_done[0]);
    if (e == hipSuccess) e = set_lds_limit(k_step<4>, bytes, g_lds_done[1]);
    if (e == hipSuccess) e = set_lds_limit(k_step<6>, bytes, g_lds_done[2]);
    return e;
}

hipError_t prepare_arena_kernels(size_t bytes) {
    hipError_t e = prepare_tick_kernels(bytes);
    if (e == hipSuccess) e = prepare_env_kernels(bytes);
    if (e == hipSuccess) e = prepare_mixed_kernel(bytes);
    return e;
}

// Shape arithmetic shared by every launcher.
size_t arena_lds_per_wave(bool cull, bool pool) {
    size_t b = (size_t)kSimplexBytes + (pool ? (size_t)kPoolBytes : 0);
    if (cull && (size_t)sdf::kLdsActBytes > b) b = sdf::kLdsActBytes;
    return (b + 15) & ~(size_t)15;
}

size_t arena_pool_offset() { return kPoolOffset; }

int arena_waves(int n, const LaunchShape& s) { return (n + s.arenas_per_wave - 1) / s.arenas_per_wave; }

bool shape_valid(const LaunchShape& s) {
    const bool block_ok = s.block == 64 || s.block == 128 || s.block == 256 || s.block == 512;
    const int a = s.arenas_per_wave;
    return block_ok && a >= 1 && a <= 64 && (a & (a - 1)) == 0 && (s.pool == 0 || s.pool == 1);
}

}  // namespace rsx
