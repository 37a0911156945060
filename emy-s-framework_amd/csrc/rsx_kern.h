// rsx_kern.h — launch interface of the arena kernels.
//
// Work mapping: one lane owns one arena for a whole launch.  A launch loads
// the arena's words from the field-major SoA buffer (coalesced across the
// lanes of a wavefront), runs the requested ticks in registers and private
// memory, and stores the words back.  The arenas of one team size form a
// bucket; each bucket has its own instantiation of the tick (MAXC = 2, 4, 6
// car slots), so no wavefront mixes car counts.
//
// Launch shape (chosen by rsx_set_tuning / rsx_autotune):
//   block            lanes per workgroup: 64, 128 or 256 (1, 2 or 4 wavefronts);
//   arenas_per_wave  arenas per 64-lane wavefront: 64, 32, 16 or 8.  With
//                    fewer than 64, lane l runs the arena of lane
//                    l % arenas_per_wave and stores nothing: the wavefront
//                    then follows fewer arenas' branches, and more wavefronts
//                    are in flight to hide memory latency.
// Each wavefront owns arena_lds_per_wave() bytes of LDS; no code waits for
// another wavefront of its workgroup.
//
// The instantiations live in separate translation units (rsx_k_ticks.hip,
// rsx_k_env.hip, rsx_k_mixed.hip) so they compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rsx_types.h"
#include "rsx_train.h"
#include "rsx_sdf.h"
#include "rsx_vanta.h"

namespace rsx {

// Arena flags word (kArFlags): a goal has been scored this episode, and
// whether it went into the orange goal's side (ball y <= 0); a tick ran out of
// contact-manifold storage (MfList::overflow) since the arena was last reset
// by rsx_reset_arena / rsx_reset_all_arenas (kept across auto-resets).
constexpr uint32_t kFlagGoal = 1u, kFlagGoalOrange = 2u, kFlagContactOverflow = 4u;

// One team-size bucket.
struct BucketArgs {
    uint32_t* words = nullptr;           // field-major SoA words, n arenas
    int n = 0;
    const rsx_controls* ctrl = nullptr;  // n * MAXC controls (tick path)
    const int* car_off = nullptr;        // global index of each arena's first car
    const int64_t* obs_off = nullptr;    // float offset of each arena's first obs row
    const int* global = nullptr;         // global arena index
};

// The world every bucket shares.
struct WorldArgs {
    CarGeom cg;
    BallGeom bg;
    MeshView mv;
    const PadGeom* pads;
    sdf::Cull cull;
    int pool;                  // pool each wavefront's car mesh queries and wheel rays (body_mesh_pool, car_mesh_rays_pool)
    int lds_bytes;             // each wavefront's LDS share (arena_lds_per_wave)
};

// The env half of a fused env step.
struct EnvArgs {
    const int32_t* act;        // device action index per car (global car order)
    const rsx_controls* lut;   // action table
    int lut_n;
    int ticks;                 // tick skip
    EnvConfig E;
    float* obs;
    float* rew;
    uint8_t* term;
    uint8_t* masks;               // DefaultAction mask rows (kNumDefaultActions bytes per car), or null
    const uint8_t* mask_tables;   // ground, air, boost, jump masks of the default table
    const RewardEntry* rt;        // training reward table (device), used when rt_n > 0
    int rt_n;                     // 0: the four stub rewards of EnvConfig
};

struct LaunchShape {
    int block = 64;
    int arenas_per_wave = 64;
    int pool = 1;              // pooled box-triangle queries
};

// LDS bytes one wavefront needs: the lanes' GJK / convex-cast simplices and
// (with pooling) the pooled queries, or -- sharing the space, at another time
// of the tick -- the cull's activation buffer when the cull is on.
size_t arena_lds_per_wave(bool cull, bool pool);
// Offset of the pooled queries' LDS within a wavefront's share.
size_t arena_pool_offset();
// Wavefronts of a launch over n arenas.
int arena_waves(int n, const LaunchShape& s);
// Whether s is one of the supported shapes.
bool shape_valid(const LaunchShape& s);

hipError_t launch_ticks(int maxc, const BucketArgs& b, const WorldArgs& w, int ticks, const LaunchShape& s,
                        hipStream_t st);
hipError_t launch_env(int maxc, const BucketArgs& b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                      hipStream_t st);
// One launch over every non-empty bucket (b[0] 1v1, b[1] 2v2, b[2] 3v3); the
// 3v3 wavefronts come first in the grid, so the longest work starts first.
hipError_t launch_env_mixed(const BucketArgs* b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                            hipStream_t st);
// VantaObs rows of one bucket (k_vanta, rsx_vanta.hip): one lane per arena
// clears the arena's Vanta words when reset[global] is non-zero (reset may
// be null), runs the ball prediction when asked, updates the pad / demo and
// coyote timers and writes every car's 366-float row at obs + obs_off.
struct VantaArgs {
    CarGeom cg;
    BallGeom bg;
    MeshView mv;
    const PadGeom* pads;
    vanta::PadTables P;
    vanta::Params Q;
    float* obs;
};
// lists (2 b.n + 2 ints): the rollouts' arena lists and their lengths.
hipError_t launch_vanta(int maxc, const BucketArgs& b, uint32_t* vwords, const uint8_t* reset, const VantaArgs& v,
                        int* lists, hipStream_t st);
// Clears the Vanta words of every arena of a bucket (only < 0) or of arena `only`.
hipError_t launch_vanta_clear(int maxc, uint32_t* vwords, int n, int only, hipStream_t st);

// The same launches in 512-lane workgroups (rsx_k_env_wide.hip, rsx_k_mixed_wide.hip).
hipError_t launch_env_wide(int maxc, const BucketArgs& b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                           hipStream_t st);
hipError_t launch_env_mixed_wide(const BucketArgs* b, const WorldArgs& w, const EnvArgs& v, const LaunchShape& s,
                                 hipStream_t st);
hipError_t prepare_env_wide_kernels(size_t bytes);
hipError_t prepare_mixed_wide_kernel(size_t bytes);

// Sets each arena kernel's dynamic-LDS limit to `bytes` (needed above 64 KiB);
// call outside stream capture.
hipError_t prepare_arena_kernels(size_t bytes);
hipError_t prepare_env_kernels(size_t bytes);
hipError_t prepare_tick_kernels(size_t bytes);
hipError_t prepare_mixed_kernel(size_t bytes);

}  // namespace rsx
