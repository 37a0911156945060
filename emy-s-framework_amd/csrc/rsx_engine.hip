// rsx_engine.hip — the C-ABI of include/rsx.h: engine set-up, state access,
// resets, action expansion and obs kernels, and the launches of the arena
// kernels (rsx_kern.h), with their launch-shape autotuner and hipGraph
// capture of the env step.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rsx.h"
#include "rsx_host.h"
#include "rsx_env.h"
#include "rsx_internal.h"
#include "rsx_kern.h"
#include "rsx_soa.h"

using namespace rsx;

namespace {

thread_local std::string g_last_error;

rsx_status fail(rsx_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

}  // namespace

int32_t rsx::record_failure(int32_t status, const char* msg) { return fail(status, msg); }

namespace {

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) return fail(RSX_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct DevWorld {
    float* tri = nullptr;
    int* tri_mesh = nullptr;
    int* tri_rank = nullptr;
    int* cell_start = nullptr;
    int* cell_tris = nullptr;
    GridEntry* cell_ent = nullptr;
    int* blk_start = nullptr;
    GridEntry* blk_ent = nullptr;
    uint16_t* bp_meshes = nullptr;
    float* mesh_aabb = nullptr;
    float* tri_info = nullptr;
    PadGeom* pads = nullptr;
};

}  // namespace

struct Bucket {
    int maxc = 0;            // car capacity (2, 4, 6)
    int n = 0;               // arenas in this bucket
    int words = 0;           // SoA words per arena
    uint32_t* d_words = nullptr;
    rsx_controls* d_ctrl = nullptr;  // n * maxc
    int* d_car_off = nullptr;        // n: global index of each arena's first car
    int64_t* d_obs_off = nullptr;    // n: float offset of each arena's first obs row
    int* d_global = nullptr;         // n: global arena index
    std::vector<int> global;         // bucket-local -> global arena index
    uint32_t* d_vanta = nullptr;     // VantaObs state words (field-major), allocated with the first Vanta config
    int* d_vwork = nullptr;          // 2n + 2 ints: the rollouts' two arena lists and their lengths
};

struct rsx_engine {
    int device = 0;
    int num_arenas = 0;
    int tick_skip = 8;
    float no_touch_timeout = 0.f;
    HostWorld hw;
    DevWorld dw;
    MeshView dview;
    Bucket bucket[3];
    std::vector<int> arena_bucket, arena_local, arena_team, arena_car_off;
    int num_cars = 0;
    rsx_controls* d_lut = nullptr;
    int lut_n = 0;
    int lut_cap = 0;
    int32_t* d_idx_stage = nullptr;  // staging for host action indices (num_cars)
    // env outputs (device, stable for the engine's lifetime)
    EnvConfig env;
    float* d_obs = nullptr;          // rows of every car, arena-major, row length 51 + 19 * cars in arena
    float* d_rew = nullptr;          // num_cars
    uint8_t* d_term = nullptr;       // num_arenas
    int64_t obs_floats = 0;
    std::vector<int64_t> car_obs_off;  // num_cars
    std::vector<int32_t> car_obs_len;  // num_cars
    sdf::Cull cull{};                  // learned broadphase cull (disabled: exact collision)
    LaunchShape shape;                 // launch shape of the arena kernels
    uint64_t gen = 1;                  // bumped by every change that alters the env step's launches
    size_t lds_per_block_max = 65536;  // device limit of a workgroup's LDS
    // DefaultAction masks: tables (ground, air, boost, jump) and one row per car
    uint8_t* d_mask_tables = nullptr;
    uint8_t* d_masks = nullptr;
    bool masks_on = false;
    // training reward table (rsx_configure_training_rewards); 0 entries: stub rewards
    RewardEntry* d_rt = nullptr;
    int rt_n = 0;
    std::vector<rsx_training_reward_entry> rt_host;
    // VantaObs (rsx_configure_vanta): parameters, pad tables, state allocated
    vanta::Params vq{8.f / 120.f, 0.5f, 0, 0, 8};
    vanta::PadTables vpads{};
    bool vanta_alloc = false;
    // captured env steps (rsx_capture_env_graph)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_gen = 0;
};

namespace rsx {
const HostWorld& engine_host_world(const rsx_engine* e) { return e->hw; }
int engine_device(const rsx_engine* e) { return e->device; }
void engine_set_cull(rsx_engine* e, const sdf::Cull& c) {
    e->cull = c;
    e->gen++;
}
}  // namespace rsx

namespace {

// Reset of one arena column in place under the env's reset curriculum (the
// plain kickoff reset when E is null): keeps car ids, teams and the carried
// wheel values (Car::SetState does not touch them), advances the arena's RNG
// stream, bumps the episode counter.
template <int MAXC>
__host__ __device__ void reset_column(const Col& col, const CarGeom& cg, const EnvConfig* E) {
    using L = Layout<MAXC>;
    Arena<MAXC> A;
    unpack_arena<MAXC>(col, A, cg);
    if (E) arena_reset_curriculum<MAXC>(A, cg, E->reset_w);
    else arena_reset_kickoff<MAXC>(A, cg);
    A.tick = 0;
    const float no_touch = E ? reset_no_touch(A.rng, *E) : 0.f;
    pack_arena<MAXC>(col, A);
    col.set(L::kArenaBase + kArFlags, 0u);
    col.setf(L::kArenaBase + kArNoTouch, no_touch);
    col.set(L::kArenaBase + kArEpisode, col.get(L::kArenaBase + kArEpisode) + 1u);
}

// Reset of the arenas whose mask byte is non-zero (all when mask is null).
template <int MAXC>
__global__ void __launch_bounds__(64) k_reset(uint32_t* words, int n, const uint8_t* mask, CarGeom cg, EnvConfig E) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    if (mask && !mask[a]) return;
    Col col{words + a, n};
    reset_column<MAXC>(col, cg, &E);
}

// DefaultObs rows of the current state with zero previous actions (the
// observation after a reset).
template <int MAXC>
__global__ void __launch_bounds__(64) k_env_obs(uint32_t* words, int n, CarGeom cg, EnvConfig E, float* obs,
                                                const int64_t* obs_off, const int* car_off, uint8_t* masks,
                                                const uint8_t* mask_tables) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    Col col{words + a, n};
    Arena<MAXC> A;
    unpack_arena<MAXC>(col, A, cg);
    float prev[MAXC][kActionFloats];
    for (int i = 0; i < MAXC; i++)
        for (int k = 0; k < kActionFloats; k++) prev[i][k] = 0.f;
    const int row = obs_row_len_of(E.obs_kind, A.ncars);
    float* o = obs + obs_off[a];
    for (int i = 0; i < A.ncars; i++) env_obs_row<MAXC>(A, i, prev, E, o + (int64_t)i * row);
    if (masks)
        for (int i = 0; i < A.ncars; i++)
            action_mask_row(A.car[i], mask_tables, masks + (int64_t)(car_off[a] + i) * kNumDefaultActions);
}

// Discrete action expansion: one lane per car slot of the bucket.
template <int MAXC>
__global__ void __launch_bounds__(256) k_expand_actions(const int32_t* idx, const int* car_off, int n,
                                                        const rsx_controls* lut, int lut_n, rsx_controls* ctrl) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * MAXC) return;
    const int a = t / MAXC, i = t - a * MAXC;
    int k = idx[car_off[a] + i];
    k = k < 0 ? 0 : (k >= lut_n ? lut_n - 1 : k);
    ctrl[t] = lut[k];
}

template <typename T>
rsx_status upload(T** dst, const std::vector<T>& src) {
    HIP_TRY(hipMalloc((void**)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
    if (!src.empty()) HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return RSX_OK;
}

int bucket_of_team(int t) { return t - 1; }

// Obs rows arena-major, one per car, for observation builder `kind`.
void obs_layout_for(rsx_engine* e, int kind) {
    e->car_obs_off.resize(e->num_cars);
    e->car_obs_len.resize(e->num_cars);
    int64_t off = 0;
    for (int a = 0; a < e->num_arenas; a++) {
        const int nc = 2 * e->arena_team[a];
        for (int i = 0; i < nc; i++) {
            e->car_obs_off[e->arena_car_off[a] + i] = off;
            e->car_obs_len[e->arena_car_off[a] + i] = obs_row_len_of(kind, nc);
            off += obs_row_len_of(kind, nc);
        }
    }
    e->obs_floats = off;
}

// Host-side pack/unpack of one arena through a contiguous staging column.
template <int MAXC>
void host_reset_arena(std::vector<uint32_t>& colbuf, const CarGeom& g, const BallGeom& bg, const MeshView& mv, int team,
                      uint32_t seed, int slot_base_id, int car_order) {
    Arena<MAXC> A;
    memset(&A, 0, sizeof(A));
    // broadphase registrations where Bullet adds the bodies: cars at the
    // identity transform (Car.cpp:211-228), the ball resting at the origin
    // (Ball.cpp:90-91)
    V3 mn, mx;
    car_shape_aabb(g, xf_identity(), mn, mx);
    const int car_reg = bp_cell_index(mv, mn);
    ball_shape_aabb(bg, V3{0, 0, bg.radius}, mn, mx);
    A.ball.bp_reg = bp_cell_index(mv, mn);
    A.ncars = 2 * team;
    A.rng = seed ? seed : 0x9E3779B9u;
    A.tick = 0;
    for (int i = 0; i < A.ncars; i++) {
        A.car[i].team = car_order == RSX_CARS_ALTERNATING ? (i & 1) : (i < team ? 0 : 1);
        A.car[i].id = (uint32_t)(slot_base_id + i + 1);
        A.car[i].active = true;
        A.car[i].bp_reg = car_reg;
    }
    for (int i = 0; i < A.ncars; i++) car_set_spawn<MAXC>(A.car[i], g, 0, 0, kCarSpawnRestZ, 0, kBoostSpawnAmount);
    colbuf.assign(Layout<MAXC>::kWords, 0);
    Col c{colbuf.data(), 1};
    pack_arena<MAXC>(c, A);
    c.set(Layout<MAXC>::kArenaBase + kArNcars, (uint32_t)A.ncars);
    reset_column<MAXC>(c, g, nullptr);
    c.set(Layout<MAXC>::kArenaBase + kArEpisode, 0u);
}

#define RSX_TRY(expr)                    \
    do {                                 \
        rsx_status _s = (expr);          \
        if (_s != RSX_OK) return _s;     \
    } while (0)

BucketArgs bucket_args(const Bucket& b) {
    BucketArgs a;
    a.words = b.d_words;
    a.n = b.n;
    a.ctrl = b.d_ctrl;
    a.car_off = b.d_car_off;
    a.obs_off = b.d_obs_off;
    a.global = b.d_global;
    return a;
}

WorldArgs world_args(const rsx_engine* e, const LaunchShape& s) {
    return WorldArgs{e->hw.cg, e->hw.bg, e->dview, e->dw.pads, e->cull, s.pool,
                     (int)arena_lds_per_wave(e->cull.enabled != 0, s.pool != 0)};
}

// Checks that the shape's LDS fits a workgroup and raises the kernels'
// dynamic-LDS limit when it needs more than 64 KiB.
rsx_status prepare_shape(rsx_engine* e, const LaunchShape& s) {
    const size_t lds = arena_lds_per_wave(e->cull.enabled != 0, s.pool != 0) * (s.block / 64);
    if (lds > e->lds_per_block_max)
        return fail(RSX_INVALID_ARG, "launch shape needs " + std::to_string(lds) + " bytes of LDS per workgroup; the "
                                         "device allows " + std::to_string(e->lds_per_block_max));
    HIP_TRY(prepare_arena_kernels(lds));
    return RSX_OK;
}

// One fused env step of every arena: one launch (the mixed kernel when team
// sizes differ).
hipError_t env_launch(rsx_engine* e, const int32_t* act, const LaunchShape& s, hipStream_t st) {
    const WorldArgs w = world_args(e, s);
    EnvArgs v;
    v.act = act;
    v.lut = e->d_lut;
    v.lut_n = e->lut_n;
    v.ticks = e->tick_skip;
    v.E = e->env;
    v.obs = e->d_obs;
    v.rew = e->d_rew;
    v.term = e->d_term;
    v.masks = e->masks_on ? e->d_masks : nullptr;
    v.mask_tables = e->d_mask_tables;
    v.rt = e->d_rt;
    v.rt_n = e->rt_n;
    int nonempty = 0;
    for (const Bucket& b : e->bucket) nonempty += b.n > 0;
    if (nonempty > 1) {
        const BucketArgs b[3] = {bucket_args(e->bucket[0]), bucket_args(e->bucket[1]), bucket_args(e->bucket[2])};
        return launch_env_mixed(b, w, v, s, st);
    }
    for (const Bucket& b : e->bucket)
        if (b.n) return launch_env(b.maxc, bucket_args(b), w, v, s, st);
    return hipSuccess;
}

// VantaObs rows of every arena (BuildVantaObs); `reset`: per-arena bytes
// whose non-zero entries mark arenas reset since the last build (their
// Vanta state is cleared first), or null.
hipError_t vanta_launch(rsx_engine* e, const uint8_t* reset, hipStream_t st) {
    VantaArgs v;
    v.cg = e->hw.cg;
    v.bg = e->hw.bg;
    v.mv = e->dview;
    v.pads = e->dw.pads;
    v.P = e->vpads;
    v.Q = e->vq;
    v.obs = e->d_obs;
    for (const Bucket& b : e->bucket) {
        if (!b.n) continue;
        const hipError_t r = launch_vanta(b.maxc, bucket_args(b), b.d_vanta, reset, v, b.d_vwork, st);
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

// One env step as rsx_step_env runs it: the fused launch, then (with
// VantaObs) the Vanta rows, clearing the state of the arenas the step reset.
hipError_t env_step_launch(rsx_engine* e, const int32_t* act, const LaunchShape& s, hipStream_t st) {
    hipError_t r = env_launch(e, act, s, st);
    if (r == hipSuccess && e->env.obs_kind == kObsVanta)
        r = vanta_launch(e, e->env.auto_reset ? e->d_term : nullptr, st);
    return r;
}

// Pad tables of VantaObs (RocketSimCudaHostUtils.cuh:68-86): each RLGym pad
// slot's arena pad and location, and the slot of its x-mirror.
vanta::PadTables vanta_pad_tables(const PadGeom* pads, const int8_t* pad_map) {
    vanta::PadTables t{};
    for (int i = 0; i < vanta::kPads; i++) {
        t.map[i] = pad_map[i];
        t.loc[i] = pads[pad_map[i]].pos_uu;
    }
    for (int i = 0; i < vanta::kPads; i++) {
        int match = i;
        for (int j = 0; j < vanta::kPads; j++)
            if (fabsf(t.loc[j].x + t.loc[i].x) < 1.f && fabsf(t.loc[j].y - t.loc[i].y) < 1.f &&
                fabsf(t.loc[j].z - t.loc[i].z) < 1.f) {
                match = j;
                break;
            }
        t.mirror[i] = (int8_t)match;
    }
    return t;
}

int vanta_words(int maxc) {
    return maxc == 2 ? vanta::VLayout<2>::kWords : (maxc == 4 ? vanta::VLayout<4>::kWords : vanta::VLayout<6>::kWords);
}

// Allocates every bucket's Vanta words, all zero (no state yet).
rsx_status vanta_alloc(rsx_engine* e) {
    if (e->vanta_alloc) return RSX_OK;
    for (Bucket& b : e->bucket) {
        if (!b.n) continue;
        const size_t bytes = sizeof(uint32_t) * (size_t)vanta_words(b.maxc) * b.n;
        HIP_TRY(hipMalloc((void**)&b.d_vanta, bytes));
        HIP_TRY(hipMemset(b.d_vanta, 0, bytes));
        HIP_TRY(hipMalloc((void**)&b.d_vwork, sizeof(int) * (2 * (size_t)b.n + 2)));
        HIP_TRY(hipMemset(b.d_vwork, 0, sizeof(int) * (2 * (size_t)b.n + 2)));
    }
    e->vanta_alloc = true;
    return RSX_OK;
}

rsx_status column_get(rsx_engine* e, int arena, std::vector<uint32_t>& out, Bucket*& b, int& local) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (arena < 0 || arena >= e->num_arenas) return fail(RSX_INVALID_ARG, "arena index out of range");
    b = &e->bucket[e->arena_bucket[arena]];
    local = e->arena_local[arena];
    out.resize(b->words);
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpy2D(out.data(), sizeof(uint32_t), b->d_words + local, sizeof(uint32_t) * b->n, sizeof(uint32_t),
                        b->words, hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status column_put(rsx_engine* e, Bucket* b, int local, const std::vector<uint32_t>& in) {
    HIP_TRY(hipMemcpy2D(b->d_words + local, sizeof(uint32_t) * b->n, in.data(), sizeof(uint32_t), sizeof(uint32_t),
                        b->words, hipMemcpyHostToDevice));
    return RSX_OK;
}

int ball_base(int maxc) { return maxc * kCarWords; }
int pad_base(int maxc) { return ball_base(maxc) + kBallWords; }
int arena_base(int maxc) { return pad_base(maxc) + kNumPads * kPadWords; }

void arena_info_from(const Col& c, int maxc, int team, rsx_arena_info& out) {
    int o = arena_base(maxc);
    memset(&out, 0, sizeof(out));
    out.tick_count = c.get64(o + kArTickLo);
    out.num_cars = (int32_t)c.get(o + kArNcars);
    out.team_size = team;
    out.no_touch_time = c.getf(o + kArNoTouch);
    out.episode = c.get(o + kArEpisode);
    uint32_t f = c.get(o + kArFlags);
    out.goal_scored = (f & kFlagGoal) ? 1 : 0;
    out.contact_overflow = (f & kFlagContactOverflow) ? 1 : 0;
    out.goal_team = (f & kFlagGoal) ? ((f & kFlagGoalOrange) ? 1 : 0) : -1;
}

}  // namespace

extern "C" {

const char* rsx_last_error(void) { return g_last_error.c_str(); }

rsx_status rsx_create(const rsx_config* cfg, rsx_engine** out) {
    if (!cfg || !out) return fail(RSX_INVALID_ARG, "null config or output pointer");
    if (cfg->num_arenas <= 0) return fail(RSX_INVALID_ARG, "num_arenas must be positive");
    if (!cfg->mesh_dir) return fail(RSX_INVALID_ARG, "mesh_dir is required");
    if (cfg->car_order != RSX_CARS_BLUE_FIRST && cfg->car_order != RSX_CARS_ALTERNATING)
        return fail(RSX_INVALID_ARG, "car_order must be RSX_CARS_BLUE_FIRST or RSX_CARS_ALTERNATING");
    std::unique_ptr<rsx_engine> e(new rsx_engine());
    e->device = cfg->device;
    e->num_arenas = cfg->num_arenas;
    e->tick_skip = cfg->tick_skip > 0 ? cfg->tick_skip : 8;
    e->no_touch_timeout = cfg->no_touch_timeout_s;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return fail(RSX_HIP_ERROR, "no HIP device is available");
    if (e->device < 0 || e->device >= nd) return fail(RSX_INVALID_ARG, "device ordinal out of range");
    HIP_TRY(hipSetDevice(e->device));
    {
        // the GAE entry point that allocates nothing finds its scratch here,
        // before a caller can capture it in a graph
        const rsx_status gs = gae_noalloc_setup(e->device);
        if (gs != RSX_OK) return gs;
    }
    {
        int a = 0, b = 0;
        HIP_TRY(hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device));
        if (hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, e->device) != hipSuccess) b = 0;
        e->lds_per_block_max = (size_t)std::max(a, b);
    }
    std::string err;
    if (!load_world(cfg->mesh_dir, e->hw, err)) return fail(RSX_MESH_LOAD_FAILED, err);
    rsx_status s;
    if ((s = upload(&e->dw.tri, e->hw.tri)) != RSX_OK) return s;
    if ((s = upload(&e->dw.tri_mesh, e->hw.tri_mesh)) != RSX_OK) return s;
    if ((s = upload(&e->dw.tri_rank, e->hw.tri_rank)) != RSX_OK) return s;
    if ((s = upload(&e->dw.cell_start, e->hw.cell_start)) != RSX_OK) return s;
    if ((s = upload(&e->dw.cell_tris, e->hw.cell_tris)) != RSX_OK) return s;
    if ((s = upload(&e->dw.cell_ent, e->hw.cell_ent)) != RSX_OK) return s;
    if ((s = upload(&e->dw.blk_start, e->hw.blk_start)) != RSX_OK) return s;
    if ((s = upload(&e->dw.blk_ent, e->hw.blk_ent)) != RSX_OK) return s;
    if ((s = upload(&e->dw.bp_meshes, e->hw.bp_meshes)) != RSX_OK) return s;
    if ((s = upload(&e->dw.mesh_aabb, e->hw.mesh_aabb)) != RSX_OK) return s;
    if ((s = upload(&e->dw.tri_info, e->hw.tri_info)) != RSX_OK) return s;
    std::vector<PadGeom> pads(e->hw.pads, e->hw.pads + kNumPads);
    if ((s = upload(&e->dw.pads, pads)) != RSX_OK) return s;
    e->dview = e->hw.view();
    e->dview.tri = e->dw.tri;
    e->dview.tri_mesh = e->dw.tri_mesh;
    e->dview.tri_rank = e->dw.tri_rank;
    e->dview.cell_start = e->dw.cell_start;
    e->dview.cell_tris = e->dw.cell_tris;
    e->dview.cell_ent = e->dw.cell_ent;
    e->dview.blk_start = e->dw.blk_start;
    e->dview.blk_ent = e->dw.blk_ent;
    e->dview.bp_meshes = e->dw.bp_meshes;
    e->dview.mesh_aabb = e->dw.mesh_aabb;
    e->dview.tri_info = e->dw.tri_info;

    // assign arenas to team-size buckets
    e->arena_bucket.resize(e->num_arenas);
    e->arena_local.resize(e->num_arenas);
    e->arena_team.resize(e->num_arenas);
    for (int a = 0; a < e->num_arenas; a++) {
        int t = cfg->team_size ? cfg->team_size[a] : cfg->default_team_size;
        if (t < 1 || t > 3) return fail(RSX_INVALID_ARG, "team size must be 1, 2 or 3");
        Bucket& b = e->bucket[bucket_of_team(t)];
        e->arena_bucket[a] = bucket_of_team(t);
        e->arena_local[a] = (int)b.global.size();
        e->arena_team[a] = t;
        b.global.push_back(a);
    }
    e->arena_car_off.resize(e->num_arenas);
    for (int a = 0; a < e->num_arenas; a++) {
        e->arena_car_off[a] = e->num_cars;
        e->num_cars += 2 * e->arena_team[a];
    }
    {
        std::vector<rsx_controls> lut = default_action_lut();
        e->lut_cap = 1024;
        HIP_TRY(hipMalloc((void**)&e->d_lut, sizeof(rsx_controls) * e->lut_cap));
        HIP_TRY(hipMemcpy(e->d_lut, lut.data(), lut.size() * sizeof(rsx_controls), hipMemcpyHostToDevice));
        e->lut_n = (int)lut.size();
        HIP_TRY(hipMalloc((void**)&e->d_idx_stage, sizeof(int32_t) * (size_t)e->num_cars));
    }
    // env outputs: obs rows arena-major, one row per car
    {
        EnvConfig& E = e->env;
        E.goal_w = 1.f;
        E.concede_scale = -1.f;
        E.touch_w = 1.f;
        E.vbtg_w = 1.f;
        E.vptb_w = 1.f;
        E.no_touch_timeout = cfg->no_touch_timeout_s;
        E.goal_terminal = 1;
        E.auto_reset = 1;
        E.obs_kind = kObsDefault;
        E.reset_w[0] = 1.f;
        E.reset_w[1] = E.reset_w[2] = E.reset_w[3] = 0.f;
        E.no_touch_stagger = 0;
        HIP_TRY(hipMalloc((void**)&e->d_rt, sizeof(RewardEntry) * kMaxRewardEntries));
        if (!obs_pad_map(e->hw.pads, E.pad_map)) return fail(RSX_MESH_LOAD_FAILED, "boost pad layout does not match");
        // the obs buffer holds the longest (VantaObs) rows, so its address
        // survives a change of observation builder
        obs_layout_for(e.get(), kObsVanta);
        const int64_t off = e->obs_floats;
        obs_layout_for(e.get(), kObsDefault);
        HIP_TRY(hipMalloc((void**)&e->d_obs, sizeof(float) * std::max<int64_t>(1, off)));
        HIP_TRY(hipMalloc((void**)&e->d_rew, sizeof(float) * (size_t)e->num_cars));
        HIP_TRY(hipMalloc((void**)&e->d_term, (size_t)e->num_arenas));
        HIP_TRY(hipMemset(e->d_obs, 0, sizeof(float) * std::max<int64_t>(1, off)));
        HIP_TRY(hipMemset(e->d_rew, 0, sizeof(float) * (size_t)e->num_cars));
        HIP_TRY(hipMemset(e->d_term, 0, (size_t)e->num_arenas));
        if ((s = upload(&e->d_mask_tables, default_action_masks())) != RSX_OK) return s;
        e->vpads = vanta_pad_tables(e->hw.pads, E.pad_map);
        e->vq.dt = (float)e->tick_skip / 120.f;
        e->vq.ticks_between = e->tick_skip;
    }
    for (int k = 0; k < 3; k++) {
        Bucket& b = e->bucket[k];
        b.maxc = 2 * (k + 1);
        b.n = (int)b.global.size();
        b.words = arena_base(b.maxc) + kArenaWords;
        if (b.n == 0) continue;
        {
            std::vector<int> off(b.n);
            std::vector<int64_t> ooff(b.n);
            for (int l = 0; l < b.n; l++) {
                off[l] = e->arena_car_off[b.global[l]];
                ooff[l] = e->car_obs_off[off[l]];
            }
            rsx_status st = upload(&b.d_car_off, off);
            if (st != RSX_OK) return st;
            st = upload(&b.d_obs_off, ooff);
            if (st != RSX_OK) return st;
            st = upload(&b.d_global, b.global);
            if (st != RSX_OK) return st;
        }
        HIP_TRY(hipMalloc((void**)&b.d_words, sizeof(uint32_t) * (size_t)b.words * b.n));
        HIP_TRY(hipMalloc((void**)&b.d_ctrl, sizeof(rsx_controls) * (size_t)b.maxc * b.n));
        HIP_TRY(hipMemset(b.d_ctrl, 0, sizeof(rsx_controls) * (size_t)b.maxc * b.n));
        // initial kickoff states, built on the host and uploaded field-major
        std::vector<uint32_t> soa((size_t)b.words * b.n), col;
        const MeshView hv = e->hw.view();
        for (int l = 0; l < b.n; l++) {
            uint32_t seed = cfg->seed * 2654435761u + (uint32_t)b.global[l] * 0x9E3779B9u + 1u;
            if (b.maxc == 2) host_reset_arena<2>(col, e->hw.cg, e->hw.bg, hv, 1, seed, 0, cfg->car_order);
            else if (b.maxc == 4) host_reset_arena<4>(col, e->hw.cg, e->hw.bg, hv, 2, seed, 0, cfg->car_order);
            else host_reset_arena<6>(col, e->hw.cg, e->hw.bg, hv, 3, seed, 0, cfg->car_order);
            for (int w = 0; w < b.words; w++) soa[(size_t)w * b.n + l] = col[w];
        }
        HIP_TRY(hipMemcpy(b.d_words, soa.data(), soa.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    *out = e.release();
    return RSX_OK;
}

rsx_status rsx_destroy(rsx_engine* e) {
    if (!e) return RSX_OK;
    (void)hipSetDevice(e->device);
    for (auto& b : e->bucket) {
        (void)hipFree(b.d_words);
        (void)hipFree(b.d_ctrl);
        (void)hipFree(b.d_car_off);
        (void)hipFree(b.d_obs_off);
        (void)hipFree(b.d_global);
        (void)hipFree(b.d_vanta);
        (void)hipFree(b.d_vwork);
    }
    (void)hipFree(e->d_obs);
    (void)hipFree(e->d_rew);
    (void)hipFree(e->d_term);
    (void)hipFree(e->d_lut);
    (void)hipFree(e->d_mask_tables);
    (void)hipFree(e->d_masks);
    (void)hipFree(e->d_idx_stage);
    (void)hipFree(e->d_rt);
    void* ptrs[] = {e->dw.tri, e->dw.tri_mesh, e->dw.tri_rank, e->dw.cell_start, e->dw.cell_tris, e->dw.cell_ent,
                    e->dw.blk_start, e->dw.blk_ent,
                    e->dw.bp_meshes, e->dw.mesh_aabb, e->dw.tri_info, e->dw.pads};
    for (void* p : ptrs) (void)hipFree(p);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    delete e;
    return RSX_OK;
}

rsx_status rsx_num_arenas(const rsx_engine* e, int32_t* out) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    *out = e->num_arenas;
    return RSX_OK;
}

rsx_status rsx_num_cars(const rsx_engine* e, int32_t* out) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    *out = e->num_cars;
    return RSX_OK;
}

rsx_status rsx_team_size(const rsx_engine* e, int32_t arena, int32_t* out) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (arena < 0 || arena >= e->num_arenas) return fail(RSX_INVALID_ARG, "arena index out of range");
    *out = e->arena_team[arena];
    return RSX_OK;
}

rsx_status rsx_get_car_state(rsx_engine* e, int32_t arena, int32_t car, rsx_car_state* out) {
    if (!out) return fail(RSX_INVALID_ARG, "null output");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    if (car < 0 || car >= 2 * e->arena_team[arena]) return fail(RSX_INVALID_ARG, "car index out of range");
    Car k;
    memset(&k, 0, sizeof(k));
    unpack_car(Col{col.data(), 1}, car * kCarWords, k, e->hw.cg);
    car_to_public(k, *out);
    return RSX_OK;
}

rsx_status rsx_set_car_state(rsx_engine* e, int32_t arena, int32_t car, const rsx_car_state* in) {
    if (!in) return fail(RSX_INVALID_ARG, "null input");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    if (car < 0 || car >= 2 * e->arena_team[arena]) return fail(RSX_INVALID_ARG, "car index out of range");
    Col c{col.data(), 1};
    Car k;
    memset(&k, 0, sizeof(k));
    unpack_car(c, car * kCarWords, k, e->hw.cg);  // keeps id, team and wheel values
    uint32_t id = k.id;
    int team = k.team;
    car_from_public(k, *in, e->hw.cg);
    k.id = id;
    k.team = team;
    pack_car(c, car * kCarWords, k);
    return column_put(e, b, local, col);
}

rsx_status rsx_get_ball_state(rsx_engine* e, int32_t arena, rsx_ball_state* out) {
    if (!out) return fail(RSX_INVALID_ARG, "null output");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    Col c{col.data(), 1};
    int o = ball_base(b->maxc);
    out->pos = rsx_vec3{c.getf(o) * kBT2UU, c.getf(o + 1) * kBT2UU, c.getf(o + 2) * kBT2UU};
    out->vel = rsx_vec3{c.getf(o + 3) * kBT2UU, c.getf(o + 4) * kBT2UU, c.getf(o + 5) * kBT2UU};
    out->ang_vel = rsx_vec3{c.getf(o + 6), c.getf(o + 7), c.getf(o + 8)};
    return RSX_OK;
}

rsx_status rsx_set_ball_state(rsx_engine* e, int32_t arena, const rsx_ball_state* in) {
    if (!in) return fail(RSX_INVALID_ARG, "null input");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    Col c{col.data(), 1};
    int o = ball_base(b->maxc);
    c.setv(o, V3{in->pos.x, in->pos.y, in->pos.z} * kUU2BT);
    c.setv(o + 3, V3{in->vel.x, in->vel.y, in->vel.z} * kUU2BT);
    c.setv(o + 6, V3{in->ang_vel.x, in->ang_vel.y, in->ang_vel.z});
    return column_put(e, b, local, col);
}

rsx_status rsx_get_pad_states(rsx_engine* e, int32_t arena, rsx_pad_state* out) {
    if (!out) return fail(RSX_INVALID_ARG, "null output");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    Col c{col.data(), 1};
    for (int p = 0; p < kNumPads; p++) {
        int o = pad_base(b->maxc) + p * kPadWords;
        memset(&out[p], 0, sizeof(out[p]));
        out[p].is_active = c.get(o) != 0;
        out[p].cooldown = c.getf(o + 1);
        out[p].prev_locked_car_id = c.get(o + 2);
    }
    return RSX_OK;
}

rsx_status rsx_set_pad_states(rsx_engine* e, int32_t arena, const rsx_pad_state* in) {
    if (!in) return fail(RSX_INVALID_ARG, "null input");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    Col c{col.data(), 1};
    for (int p = 0; p < kNumPads; p++) {
        int o = pad_base(b->maxc) + p * kPadWords;
        c.set(o, in[p].is_active ? 1u : 0u);
        c.setf(o + 1, in[p].cooldown);
        c.set(o + 2, in[p].prev_locked_car_id);
    }
    return column_put(e, b, local, col);
}

rsx_status rsx_get_arena_info(rsx_engine* e, int32_t arena, rsx_arena_info* out) {
    if (!out) return fail(RSX_INVALID_ARG, "null output");
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    arena_info_from(Col{col.data(), 1}, b->maxc, e->arena_team[arena], *out);
    return RSX_OK;
}

rsx_status rsx_reset_arena(rsx_engine* e, int32_t arena) {
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    Col c{col.data(), 1};
    if (b->maxc == 2) reset_column<2>(c, e->hw.cg, &e->env);
    else if (b->maxc == 4) reset_column<4>(c, e->hw.cg, &e->env);
    else reset_column<6>(c, e->hw.cg, &e->env);
    RSX_TRY(column_put(e, b, local, col));
    if (e->vanta_alloc) {
        // ObsBuilder::Reset: the arena's VantaObs state starts over
        HIP_TRY(launch_vanta_clear(b->maxc, b->d_vanta, b->n, local, nullptr));
        HIP_TRY(hipStreamSynchronize(nullptr));
    }
    return RSX_OK;
}

rsx_status rsx_reset_all_arenas(rsx_engine* e, void* stream) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    for (int k = 0; k < 3; k++) {
        Bucket& b = e->bucket[k];
        if (!b.n) continue;
        dim3 grid((b.n + 63) / 64), block(64);
        if (b.maxc == 2)
            hipLaunchKernelGGL(k_reset<2>, grid, block, 0, st, b.d_words, b.n, nullptr, e->hw.cg, e->env);
        else if (b.maxc == 4)
            hipLaunchKernelGGL(k_reset<4>, grid, block, 0, st, b.d_words, b.n, nullptr, e->hw.cg, e->env);
        else hipLaunchKernelGGL(k_reset<6>, grid, block, 0, st, b.d_words, b.n, nullptr, e->hw.cg, e->env);
        HIP_TRY(hipGetLastError());
        if (e->vanta_alloc) HIP_TRY(launch_vanta_clear(b.maxc, b.d_vanta, b.n, -1, st));
    }
    return RSX_OK;
}

rsx_status rsx_set_tick_count(rsx_engine* e, int32_t arena, uint64_t tick) {
    std::vector<uint32_t> col;
    Bucket* b;
    int local;
    rsx_status s = column_get(e, arena, col, b, local);
    if (s != RSX_OK) return s;
    Col c{col.data(), 1};
    c.set64(arena_base(b->maxc) + kArTickLo, tick);
    return column_put(e, b, local, col);
}

rsx_status rsx_set_all_controls(rsx_engine* e, const rsx_controls* controls, int32_t count) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!controls) return fail(RSX_INVALID_ARG, "null controls");
    int total = 0;
    for (int a = 0; a < e->num_arenas; a++) total += 2 * e->arena_team[a];
    if (count != total) return fail(RSX_INVALID_ARG, "controls count must equal the number of cars");
    HIP_TRY(hipSetDevice(e->device));
    std::vector<rsx_controls> per[3];
    for (int k = 0; k < 3; k++) per[k].assign((size_t)e->bucket[k].n * e->bucket[k].maxc, rsx_controls{});
    int o = 0;
    for (int a = 0; a < e->num_arenas; a++) {
        int k = e->arena_bucket[a], l = e->arena_local[a], nc = 2 * e->arena_team[a];
        for (int i = 0; i < nc; i++) per[k][(size_t)l * e->bucket[k].maxc + i] = controls[o + i];
        o += nc;
    }
    for (int k = 0; k < 3; k++)
        if (e->bucket[k].n)
            HIP_TRY(hipMemcpy(e->bucket[k].d_ctrl, per[k].data(), per[k].size() * sizeof(rsx_controls),
                              hipMemcpyHostToDevice));
    return RSX_OK;
}

rsx_status rsx_set_car_controls(rsx_engine* e, int32_t arena, int32_t car, const rsx_controls* c) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!c) return fail(RSX_INVALID_ARG, "null controls");
    if (arena < 0 || arena >= e->num_arenas) return fail(RSX_INVALID_ARG, "arena index out of range");
    if (car < 0 || car >= 2 * e->arena_team[arena]) return fail(RSX_INVALID_ARG, "car index out of range");
    const Bucket& b = e->bucket[e->arena_bucket[arena]];
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpy(b.d_ctrl + (size_t)e->arena_local[arena] * b.maxc + car, c, sizeof(rsx_controls),
                      hipMemcpyHostToDevice));
    return RSX_OK;
}

rsx_status rsx_upload_action_lut(rsx_engine* e, const rsx_controls* lut, int32_t count) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!lut || count <= 0 || count > e->lut_cap) return fail(RSX_INVALID_ARG, "action table must hold 1..1024 entries");
    if (e->masks_on) {
        // the DefaultAction masks index the default table's entries
        const std::vector<rsx_controls> def = default_action_lut();
        if (count != (int32_t)def.size() || memcmp(lut, def.data(), def.size() * sizeof(rsx_controls)) != 0)
            return fail(RSX_INVALID_ARG, "action masks are on: only the default 90-entry table can be uploaded "
                                         "(turn the masks off first)");
    }
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpy(e->d_lut, lut, sizeof(rsx_controls) * count, hipMemcpyHostToDevice));
    e->lut_n = count;
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_default_action_lut(rsx_controls* out, int32_t capacity, int32_t* count) {
    std::vector<rsx_controls> lut = default_action_lut();
    if (count) *count = (int32_t)lut.size();
    if (!out) return RSX_OK;
    if (capacity < (int32_t)lut.size()) return fail(RSX_INVALID_ARG, "capacity below the table size");
    memcpy(out, lut.data(), lut.size() * sizeof(rsx_controls));
    return RSX_OK;
}

rsx_status rsx_set_actions_device(rsx_engine* e, const int32_t* device_indices, int32_t count, void* stream) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!device_indices) return fail(RSX_INVALID_ARG, "null indices");
    if (count != e->num_cars) return fail(RSX_INVALID_ARG, "index count must equal the number of cars");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    for (int k = 0; k < 3; k++) {
        Bucket& b = e->bucket[k];
        if (!b.n) continue;
        const int threads = b.n * b.maxc;
        dim3 grid((threads + 255) / 256), block(256);
        if (b.maxc == 2)
            hipLaunchKernelGGL(k_expand_actions<2>, grid, block, 0, st, device_indices, b.d_car_off, b.n, e->d_lut,
                               e->lut_n, b.d_ctrl);
        else if (b.maxc == 4)
            hipLaunchKernelGGL(k_expand_actions<4>, grid, block, 0, st, device_indices, b.d_car_off, b.n, e->d_lut,
                               e->lut_n, b.d_ctrl);
        else
            hipLaunchKernelGGL(k_expand_actions<6>, grid, block, 0, st, device_indices, b.d_car_off, b.n, e->d_lut,
                               e->lut_n, b.d_ctrl);
        HIP_TRY(hipGetLastError());
    }
    return RSX_OK;
}

rsx_status rsx_set_actions_from_indices(rsx_engine* e, const int32_t* indices, int32_t count) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!indices) return fail(RSX_INVALID_ARG, "null indices");
    if (count != e->num_cars) return fail(RSX_INVALID_ARG, "index count must equal the number of cars");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpy(e->d_idx_stage, indices, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    rsx_status s = rsx_set_actions_device(e, e->d_idx_stage, count, nullptr);
    if (s != RSX_OK) return s;
    HIP_TRY(hipDeviceSynchronize());
    return RSX_OK;
}

}  // extern "C"

namespace {

// Whole-bucket copies between the device SoA and a host mirror.
rsx_status bucket_download(rsx_engine* e, int k, std::vector<uint32_t>& host) {
    Bucket& b = e->bucket[k];
    host.resize((size_t)b.words * b.n);
    if (b.n) HIP_TRY(hipMemcpy(host.data(), b.d_words, host.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status bucket_upload(rsx_engine* e, int k, const std::vector<uint32_t>& host) {
    Bucket& b = e->bucket[k];
    if (b.n) HIP_TRY(hipMemcpy(b.d_words, host.data(), host.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    return RSX_OK;
}

// Visits every arena with a host-side column over its bucket's mirror.
// fn(arena, Col, maxc) ; returns the first error.
template <typename Fn>
rsx_status for_all_arenas(rsx_engine* e, bool write_back, Fn fn) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    for (int k = 0; k < 3; k++) {
        Bucket& b = e->bucket[k];
        if (!b.n) continue;
        std::vector<uint32_t> host;
        rsx_status s = bucket_download(e, k, host);
        if (s != RSX_OK) return s;
        for (int l = 0; l < b.n; l++) fn(b.global[l], Col{host.data() + l, b.n}, b.maxc);
        if (write_back) {
            s = bucket_upload(e, k, host);
            if (s != RSX_OK) return s;
        }
    }
    return RSX_OK;
}

}  // namespace

extern "C" {

rsx_status rsx_get_all_car_states(rsx_engine* e, rsx_car_state* out, int32_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_cars) return fail(RSX_INVALID_ARG, "count must equal the number of cars");
    return for_all_arenas(e, false, [&](int a, const Col& c, int) {
        for (int i = 0; i < 2 * e->arena_team[a]; i++) {
            Car k;
            memset(&k, 0, sizeof(k));
            unpack_car(c, i * kCarWords, k, e->hw.cg);
            car_to_public(k, out[e->arena_car_off[a] + i]);
        }
    });
}

rsx_status rsx_set_all_car_states(rsx_engine* e, const rsx_car_state* in, int32_t count) {
    if (!e || !in) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_cars) return fail(RSX_INVALID_ARG, "count must equal the number of cars");
    return for_all_arenas(e, true, [&](int a, const Col& c, int) {
        for (int i = 0; i < 2 * e->arena_team[a]; i++) {
            Car k;
            memset(&k, 0, sizeof(k));
            unpack_car(c, i * kCarWords, k, e->hw.cg);
            uint32_t id = k.id;
            int team = k.team;
            car_from_public(k, in[e->arena_car_off[a] + i], e->hw.cg);
            k.id = id;
            k.team = team;
            pack_car(c, i * kCarWords, k);
        }
    });
}

rsx_status rsx_get_all_ball_states(rsx_engine* e, rsx_ball_state* out, int32_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_arenas) return fail(RSX_INVALID_ARG, "count must equal the number of arenas");
    return for_all_arenas(e, false, [&](int a, const Col& c, int maxc) {
        Ball b;
        memset(&b, 0, sizeof(b));
        int o = ball_base(maxc);
        b.xf.o = c.getv(o);
        b.vel = c.getv(o + 3);
        b.ang = c.getv(o + 6);
        ball_to_public(b, out[a]);
    });
}

rsx_status rsx_set_all_ball_states(rsx_engine* e, const rsx_ball_state* in, int32_t count) {
    if (!e || !in) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_arenas) return fail(RSX_INVALID_ARG, "count must equal the number of arenas");
    return for_all_arenas(e, true, [&](int a, const Col& c, int maxc) {
        Ball b;
        memset(&b, 0, sizeof(b));
        ball_from_public(b, in[a]);
        int o = ball_base(maxc);
        c.setv(o, b.xf.o);
        c.setv(o + 3, b.vel);
        c.setv(o + 6, b.ang);
    });
}

rsx_status rsx_get_all_pad_states(rsx_engine* e, rsx_pad_state* out, int32_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != kNumPads * e->num_arenas) return fail(RSX_INVALID_ARG, "count must be 34 x the number of arenas");
    return for_all_arenas(e, false, [&](int a, const Col& c, int maxc) {
        for (int p = 0; p < kNumPads; p++) {
            int o = pad_base(maxc) + p * kPadWords;
            rsx_pad_state& r = out[(size_t)a * kNumPads + p];
            memset(&r, 0, sizeof(r));
            r.is_active = c.get(o) != 0;
            r.cooldown = c.getf(o + 1);
            r.prev_locked_car_id = c.get(o + 2);
        }
    });
}

rsx_status rsx_set_all_pad_states(rsx_engine* e, const rsx_pad_state* in, int32_t count) {
    if (!e || !in) return fail(RSX_INVALID_ARG, "null argument");
    if (count != kNumPads * e->num_arenas) return fail(RSX_INVALID_ARG, "count must be 34 x the number of arenas");
    return for_all_arenas(e, true, [&](int a, const Col& c, int maxc) {
        for (int p = 0; p < kNumPads; p++) {
            int o = pad_base(maxc) + p * kPadWords;
            const rsx_pad_state& r = in[(size_t)a * kNumPads + p];
            c.set(o, r.is_active ? 1u : 0u);
            c.setf(o + 1, r.cooldown);
            c.set(o + 2, r.prev_locked_car_id);
        }
    });
}

rsx_status rsx_get_all_arena_infos(rsx_engine* e, rsx_arena_info* out, int32_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_arenas) return fail(RSX_INVALID_ARG, "count must equal the number of arenas");
    return for_all_arenas(e, false, [&](int a, const Col& c, int maxc) { arena_info_from(c, maxc, e->arena_team[a], out[a]); });
}

rsx_status rsx_configure_env(rsx_engine* e, const rsx_env_config* c) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!c) return fail(RSX_INVALID_ARG, "null env config");
    EnvConfig& E = e->env;
    E.goal_w = c->goal_weight;
    E.concede_scale = c->concede_scale;
    E.touch_w = c->touch_weight;
    E.vbtg_w = c->vel_ball_to_goal_weight;
    E.vptb_w = c->vel_player_to_ball_weight;
    E.no_touch_timeout = c->no_touch_timeout_s;
    E.goal_terminal = c->goal_terminal != 0;
    E.auto_reset = c->auto_reset != 0;
    E.no_touch_stagger = c->no_touch_stagger != 0;
    if (c->obs_builder != RSX_OBS_DEFAULT && c->obs_builder != RSX_OBS_ADVANCED && c->obs_builder != RSX_OBS_VANTA)
        return fail(RSX_INVALID_ARG, "obs_builder must be RSX_OBS_DEFAULT, RSX_OBS_ADVANCED or RSX_OBS_VANTA");
    if (c->obs_builder == RSX_OBS_VANTA) {
        HIP_TRY(hipSetDevice(e->device));
        RSX_TRY(vanta_alloc(e));
    }
    if (c->obs_builder != E.obs_kind) {
        HIP_TRY(hipSetDevice(e->device));
        HIP_TRY(hipDeviceSynchronize());  // no env step may still read the old offsets
        E.obs_kind = c->obs_builder;
        obs_layout_for(e, E.obs_kind);
        for (int k = 0; k < 3; k++) {
            Bucket& b = e->bucket[k];
            if (!b.n) continue;
            std::vector<int64_t> ooff(b.n);
            for (int l = 0; l < b.n; l++) ooff[l] = e->car_obs_off[e->arena_car_off[b.global[l]]];
            HIP_TRY(hipMemcpy(b.d_obs_off, ooff.data(), sizeof(int64_t) * b.n, hipMemcpyHostToDevice));
        }
    }
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_get_env_config(const rsx_engine* e, rsx_env_config* c) {
    if (!e || !c) return fail(RSX_INVALID_ARG, "null argument");
    const EnvConfig& E = e->env;
    c->goal_weight = E.goal_w;
    c->concede_scale = E.concede_scale;
    c->touch_weight = E.touch_w;
    c->vel_ball_to_goal_weight = E.vbtg_w;
    c->vel_player_to_ball_weight = E.vptb_w;
    c->no_touch_timeout_s = E.no_touch_timeout;
    c->goal_terminal = E.goal_terminal;
    c->auto_reset = E.auto_reset;
    c->obs_builder = E.obs_kind;
    c->no_touch_stagger = E.no_touch_stagger;
    return RSX_OK;
}

rsx_status rsx_configure_training_rewards(rsx_engine* e, const rsx_training_reward_entry* entries, int32_t count) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (count < 0 || count > RSX_MAX_TRAINING_REWARDS)
        return fail(RSX_INVALID_ARG, "reward table must hold 0..RSX_MAX_TRAINING_REWARDS entries");
    if (count > 0 && !entries) return fail(RSX_INVALID_ARG, "null reward table");
    RewardEntry t[kMaxRewardEntries];
    for (int k = 0; k < count; k++) {
        const rsx_training_reward_entry& s = entries[k];
        if (s.id >= kRwCount) return fail(RSX_INVALID_ARG, "unknown reward id " + std::to_string((int)s.id));
        t[k].id = s.id;
        t[k].weight = s.weight;
        for (int p = 0; p < kRewardParams; p++) t[k].params[p] = s.params[p];
        t[k].zero_sum = s.is_zero_sum != 0;
        t[k].team_spirit = s.team_spirit;
        t[k].opp_scale = s.opponent_scale;
    }
    HIP_TRY(hipSetDevice(e->device));
    if (count > 0) HIP_TRY(hipMemcpy(e->d_rt, t, sizeof(RewardEntry) * count, hipMemcpyHostToDevice));
    e->rt_n = count;
    e->rt_host.assign(entries, entries + count);
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_get_training_rewards(const rsx_engine* e, rsx_training_reward_entry* out, int32_t capacity,
                                    int32_t* count) {
    if (!e) return fail(RSX_INVALID_ARG, "engine is null");
    if (count) *count = e->rt_n;
    if (!out) return RSX_OK;
    if (capacity < e->rt_n) return fail(RSX_INVALID_ARG, "capacity below the table size");
    for (int k = 0; k < e->rt_n; k++) out[k] = e->rt_host[k];
    return RSX_OK;
}

rsx_status rsx_set_reset_curriculum(rsx_engine* e, float kickoff_w, float fuzzed_w, float aerial_w, float chase_w) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    auto clean = [](float w) { return (!(w >= 0.f) || w > 1e6f) ? 0.f : w; };
    float* w = e->env.reset_w;
    w[0] = clean(kickoff_w);
    w[1] = clean(fuzzed_w);
    w[2] = clean(aerial_w);
    w[3] = clean(chase_w);
    if (w[0] + w[1] + w[2] + w[3] <= 0.f) w[0] = 1.f;
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_get_reset_curriculum(const rsx_engine* e, float* w4) {
    if (!e || !w4) return fail(RSX_INVALID_ARG, "null argument");
    for (int k = 0; k < 4; k++) w4[k] = e->env.reset_w[k];
    return RSX_OK;
}

rsx_status rsx_step_env(rsx_engine* e, const int32_t* device_actions, void* stream) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!device_actions) return fail(RSX_INVALID_ARG, "null actions");
    HIP_TRY(hipSetDevice(e->device));
    RSX_TRY(prepare_shape(e, e->shape));
    HIP_TRY(env_step_launch(e, device_actions, e->shape, (hipStream_t)stream));
    return RSX_OK;
}

rsx_status rsx_build_obs(rsx_engine* e, void* stream) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    HIP_TRY(hipSetDevice(e->device));
    hipStream_t st = (hipStream_t)stream;
    const bool vanta_on = e->env.obs_kind == kObsVanta;
    for (int k = 0; k < 3; k++) {
        Bucket& b = e->bucket[k];
        if (!b.n) continue;
        if (vanta_on && !e->masks_on) continue;  // the Vanta kernel writes the rows
        dim3 grid((b.n + 63) / 64), block(64);
        uint8_t* masks = e->masks_on ? e->d_masks : nullptr;
#define RSX_LAUNCH_OBS(M)                                                                                        \
    hipLaunchKernelGGL(k_env_obs<M>, grid, block, 0, st, b.d_words, b.n, e->hw.cg, e->env, e->d_obs, b.d_obs_off, \
                       b.d_car_off, masks, e->d_mask_tables)
        if (b.maxc == 2) RSX_LAUNCH_OBS(2);
        else if (b.maxc == 4) RSX_LAUNCH_OBS(4);
        else RSX_LAUNCH_OBS(6);
#undef RSX_LAUNCH_OBS
        HIP_TRY(hipGetLastError());
    }
    if (vanta_on) HIP_TRY(vanta_launch(e, nullptr, st));
    return RSX_OK;
}

rsx_status rsx_env_buffers(rsx_engine* e, float** obs, float** rewards, uint8_t** terminals, int64_t* obs_floats) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (obs) *obs = e->d_obs;
    if (rewards) *rewards = e->d_rew;
    if (terminals) *terminals = e->d_term;
    if (obs_floats) *obs_floats = e->obs_floats;
    return RSX_OK;
}

rsx_status rsx_obs_layout(const rsx_engine* e, int64_t* row_offset, int32_t* row_len, int32_t count) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (count != e->num_cars) return fail(RSX_INVALID_ARG, "count must equal the number of cars");
    for (int i = 0; i < count; i++) {
        if (row_offset) row_offset[i] = e->car_obs_off[i];
        if (row_len) row_len[i] = e->car_obs_len[i];
    }
    return RSX_OK;
}

rsx_status rsx_copy_obs(rsx_engine* e, float* out, int64_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->obs_floats) return fail(RSX_INVALID_ARG, "count must equal the obs buffer size");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, e->d_obs, sizeof(float) * count, hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status rsx_copy_rewards(rsx_engine* e, float* out, int32_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_cars) return fail(RSX_INVALID_ARG, "count must equal the number of cars");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, e->d_rew, sizeof(float) * count, hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status rsx_set_action_masks(rsx_engine* e, int32_t enabled) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (enabled && e->lut_n != kNumDefaultActions)
        return fail(RSX_INVALID_ARG, "action masks follow the default 90-entry action table");
    HIP_TRY(hipSetDevice(e->device));
    if (enabled && !e->d_masks) {
        const size_t bytes = (size_t)kNumDefaultActions * std::max(1, e->num_cars);
        HIP_TRY(hipMalloc((void**)&e->d_masks, bytes));
        HIP_TRY(hipMemset(e->d_masks, 0, bytes));
    }
    e->masks_on = enabled != 0;
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_action_masks_buffer(rsx_engine* e, uint8_t** d_masks) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!d_masks) return fail(RSX_INVALID_ARG, "null output");
    *d_masks = e->d_masks;
    return RSX_OK;
}

rsx_status rsx_copy_action_masks(rsx_engine* e, uint8_t* out, int64_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (!e->d_masks) return fail(RSX_INVALID_ARG, "action masks are off");
    if (count != (int64_t)kNumDefaultActions * e->num_cars)
        return fail(RSX_INVALID_ARG, "count must equal 90 x the number of cars");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, e->d_masks, (size_t)count, hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status rsx_copy_terminals(rsx_engine* e, uint8_t* out, int32_t count) {
    if (!e || !out) return fail(RSX_INVALID_ARG, "null argument");
    if (count != e->num_arenas) return fail(RSX_INVALID_ARG, "count must equal the number of arenas");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, e->d_term, count, hipMemcpyDeviceToHost));
    return RSX_OK;
}

rsx_status rsx_step_ticks(rsx_engine* e, int32_t ticks, void* stream) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (ticks < 0) return fail(RSX_INVALID_ARG, "ticks must be >= 0");
    HIP_TRY(hipSetDevice(e->device));
    RSX_TRY(prepare_shape(e, e->shape));
    const WorldArgs w = world_args(e, e->shape);
    for (int k = 0; k < 3; k++) {
        const Bucket& b = e->bucket[k];
        if (b.n) HIP_TRY(launch_ticks(b.maxc, bucket_args(b), w, ticks, e->shape, (hipStream_t)stream));
    }
    return RSX_OK;
}

rsx_status rsx_set_tuning(rsx_engine* e, const rsx_tuning* t) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!t) return fail(RSX_INVALID_ARG, "null tuning");
    const LaunchShape s{t->block_lanes, t->arenas_per_wave, t->pool_narrowphase};
    if (!shape_valid(s))
        return fail(RSX_INVALID_ARG,
                    "block_lanes must be 64, 128, 256 or 512, arenas_per_wave 64, 32, 16 or 8, pool_narrowphase 0 or 1");
    HIP_TRY(hipSetDevice(e->device));
    RSX_TRY(prepare_shape(e, s));
    e->shape = s;
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_get_tuning(const rsx_engine* e, rsx_tuning* t) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!t) return fail(RSX_INVALID_ARG, "null tuning");
    t->block_lanes = e->shape.block;
    t->arenas_per_wave = e->shape.arenas_per_wave;
    t->pool_narrowphase = e->shape.pool;
    return RSX_OK;
}

rsx_status rsx_autotune(rsx_engine* e, const int32_t* device_actions, int32_t env_steps, rsx_tune_result* results,
                        int32_t capacity, int32_t* count) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!device_actions || env_steps < 1) return fail(RSX_INVALID_ARG, "need device actions and env_steps >= 1");
    HIP_TRY(hipSetDevice(e->device));
    // Fewer than 8 arenas per wavefront only with pooled queries: the spare
    // lanes then take their share of the pools; unpooled they would only
    // replay arenas.
    static const int kBlocks[] = {64, 128, 256, 512};
    static const int kApw[] = {64, 32, 16, 8, 4, 2, 1};
    std::vector<rsx_tune_result> res;
    for (int pool = 1; pool >= 0; pool--)
        for (int bl : kBlocks)
            for (int apw : kApw)
                if (pool || apw >= 8) res.push_back(rsx_tune_result{{bl, apw, pool}, -1.f, 0});
    if (results && capacity < (int32_t)res.size()) return fail(RSX_INVALID_ARG, "results capacity below 44");
    // a shape needing more than 8 wavefronts per SIMD for these arenas is
    // reported infeasible without a trial
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
    const int64_t max_waves = 8LL * 4 * std::max(cus, 1);
    // Snapshot of everything a trial overwrites, restored after each trial
    // and at the end.
    std::vector<std::pair<void*, size_t>> bufs;
    for (auto& b : e->bucket)
        if (b.n) bufs.push_back({b.d_words, (size_t)b.n * b.words * sizeof(uint32_t)});
    bufs.push_back({e->d_obs, sizeof(float) * std::max<int64_t>(1, e->obs_floats)});
    bufs.push_back({e->d_rew, sizeof(float) * std::max(1, e->num_cars)});
    bufs.push_back({e->d_term, std::max<size_t>(1, (size_t)e->num_arenas)});
    if (e->cull.enabled) bufs.push_back({e->cull.stats, 3 * sizeof(unsigned long long)});
    if (e->masks_on) bufs.push_back({e->d_masks, (size_t)e->num_cars * kNumDefaultActions});
    std::vector<void*> saved(bufs.size(), nullptr);
    bool saved_all = false;  // every snapshot taken: restore() is safe on any exit
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    rsx_status status = RSX_OK;
    auto cleanup = [&]() {
        for (void* p : saved) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    };
    auto restore = [&]() -> hipError_t {
        for (size_t i = 0; i < bufs.size(); i++) {
            hipError_t r = hipMemcpyAsync(bufs[i].first, saved[i], bufs[i].second, hipMemcpyDeviceToDevice, st);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    };
    auto run = [&]() -> rsx_status {
        // the trials run on the null stream: a queue of their own would need
        // its own private-segment (scratch) allocation for these kernels
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        for (size_t i = 0; i < bufs.size(); i++) {
            HIP_TRY(hipMalloc(&saved[i], bufs[i].second));
            HIP_TRY(hipMemcpyAsync(saved[i], bufs[i].first, bufs[i].second, hipMemcpyDeviceToDevice, st));
        }
        saved_all = true;
        int best = -1;
        for (size_t k = 0; k < res.size(); k++) {
            const LaunchShape s{res[k].tuning.block_lanes, res[k].tuning.arenas_per_wave,
                                res[k].tuning.pool_narrowphase};
            int64_t waves = 0;
            for (const auto& b : e->bucket) waves += arena_waves(b.n, s);
            if (waves > max_waves) continue;  // more than 8 wavefronts per SIMD: not timed
            const size_t lds = arena_lds_per_wave(e->cull.enabled != 0, s.pool != 0) * (s.block / 64);
            if (lds > e->lds_per_block_max || prepare_arena_kernels(lds) != hipSuccess) {
                (void)hipGetLastError();
                continue;  // does not fit the workgroup's LDS
            }
            HIP_TRY(restore());
            HIP_TRY(env_launch(e, device_actions, s, st));  // warm-up (and code-object load)
            HIP_TRY(restore());
            HIP_TRY(hipEventRecord(ev0, st));
            for (int i = 0; i < env_steps; i++) HIP_TRY(env_launch(e, device_actions, s, st));
            HIP_TRY(hipEventRecord(ev1, st));
            HIP_TRY(hipEventSynchronize(ev1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
            res[k].ms_per_env_step = ms / (float)env_steps;
            res[k].feasible = 1;
            if (best < 0 || res[k].ms_per_env_step < res[best].ms_per_env_step) best = (int)k;
        }
        HIP_TRY(restore());
        HIP_TRY(hipStreamSynchronize(st));
        if (best < 0) return fail(RSX_HIP_ERROR, "no launch shape fits this device");
        e->shape = LaunchShape{res[best].tuning.block_lanes, res[best].tuning.arenas_per_wave,
                               res[best].tuning.pool_narrowphase};
        e->gen++;
        return RSX_OK;
    };
    status = run();
    if (status != RSX_OK && saved_all) {
        // a failed trial leaves the arenas advanced: put them back before
        // reporting the failure (best effort; the first error is the one kept)
        (void)restore();
        (void)hipStreamSynchronize(st);
    }
    cleanup();
    if (status != RSX_OK) return status;
    if (results) std::copy(res.begin(), res.end(), results);
    if (count) *count = (int32_t)res.size();
    return RSX_OK;
}

rsx_status rsx_capture_env_graph(rsx_engine* e, const int32_t* device_actions, int32_t env_steps) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!device_actions || env_steps < 1) return fail(RSX_INVALID_ARG, "need device actions and env_steps >= 1");
    HIP_TRY(hipSetDevice(e->device));
    if (e->graph_exec) HIP_TRY(hipGraphExecDestroy(e->graph_exec));
    if (e->graph) HIP_TRY(hipGraphDestroy(e->graph));
    e->graph_exec = nullptr;
    e->graph = nullptr;
    RSX_TRY(prepare_shape(e, e->shape));
    hipStream_t cs = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t r = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    for (int i = 0; i < env_steps && r == hipSuccess; i++) r = env_step_launch(e, device_actions, e->shape, cs);
    hipGraph_t g = nullptr;
    const hipError_t r_end = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (r == hipSuccess) r = r_end;
    if (r == hipSuccess) r = hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0);
    if (r != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        e->graph_exec = nullptr;
        return fail(RSX_HIP_ERROR, std::string("env-step capture: ") + hipGetErrorString(r));
    }
    e->graph = g;
    e->graph_gen = e->gen;
    return RSX_OK;
}

rsx_status rsx_launch_env_graph(rsx_engine* e, void* stream) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!e->graph_exec) return fail(RSX_INVALID_ARG, "no env step captured");
    if (e->graph_gen != e->gen)
        return fail(RSX_INVALID_ARG, "the launch shape, env config, action table or cull changed since the capture");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipGraphLaunch(e->graph_exec, (hipStream_t)stream));
    return RSX_OK;
}

rsx_status rsx_configure_vanta(rsx_engine* e, const rsx_vanta_config* c) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!c) return fail(RSX_INVALID_ARG, "null Vanta config");
    if (c->tick_skip < 0) return fail(RSX_INVALID_ARG, "tick_skip must be >= 0");
    if (!(c->dodge_deadzone >= 0.f && c->dodge_deadzone <= 1.f))
        return fail(RSX_INVALID_ARG, "dodge_deadzone must lie in [0, 1]");
    e->vq.predict = c->predict_ball != 0;
    e->vq.only_closest_opp = c->only_closest_opp != 0;
    e->vq.deadzone = c->dodge_deadzone;
    e->vq.dt = (float)(c->tick_skip > 0 ? c->tick_skip : e->tick_skip) / 120.f;
    e->gen++;
    return RSX_OK;
}

rsx_status rsx_get_vanta_config(const rsx_engine* e, rsx_vanta_config* c) {
    if (!e || !c) return fail(RSX_INVALID_ARG, "null argument");
    c->predict_ball = e->vq.predict;
    c->only_closest_opp = e->vq.only_closest_opp;
    c->dodge_deadzone = e->vq.deadzone;
    c->tick_skip = (int32_t)lroundf(e->vq.dt * 120.f);
    return RSX_OK;
}

rsx_status rsx_get_vanta_prediction(rsx_engine* e, int32_t arena, rsx_ball_state* out3, int32_t* has_prediction) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    if (!out3 || !has_prediction) return fail(RSX_INVALID_ARG, "null argument");
    if (arena < 0 || arena >= e->num_arenas) return fail(RSX_INVALID_ARG, "arena index out of range");
    *has_prediction = 0;
    memset(out3, 0, 3 * sizeof(rsx_ball_state));
    if (!e->vanta_alloc) return RSX_OK;
    Bucket& b = e->bucket[e->arena_bucket[arena]];
    const int local = e->arena_local[arena];
    std::vector<uint32_t> col(vanta::kVaWords);
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpy2D(col.data(), sizeof(uint32_t), b.d_vanta + local, sizeof(uint32_t) * b.n, sizeof(uint32_t),
                        vanta::kVaWords, hipMemcpyDeviceToHost));
    const Col c{col.data(), 1};
    *has_prediction = (c.get(vanta::kVaAvailHi) >> 9) & 1;
    for (int k = 0; k < 3; k++) {
        const V3 p = c.getv(vanta::kVaPredPos + 3 * k), v = c.getv(vanta::kVaPredVel + 3 * k);
        out3[k].pos = rsx_vec3{p.x, p.y, p.z};
        out3[k].vel = rsx_vec3{v.x, v.y, v.z};
        out3[k].ang_vel = rsx_vec3{0.f, 0.f, 0.f};
    }
    return RSX_OK;
}

rsx_status rsx_synchronize(rsx_engine* e) {
    if (!e) return fail(RSX_NOT_INITIALIZED, "engine is null");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    return RSX_OK;
}

}  // extern "C"
