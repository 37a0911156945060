// rsx_step.h — one physics tick of one arena: the order of Arena::Step
// (Arena.cpp:716-812) with Bullet's world step in the middle
// (btDiscreteDynamicsWorld.cpp:343-428: gravity, damping + prediction,
// broadphase + narrowphase, islands, solve, integrate), then car post-tick,
// boost pads, ball limits and the goal test.
#pragma once

#include <type_traits>

#include "rsx_boxbox.h"
#include "rsx_gjk.h"
#include "rsx_solver.h"

namespace rsx {

// ---- Bullet's unstable quicksort (btAlignedObjectArray::quickSortInternal) ----
template <int CAP>
RSX_HD void bt_quicksort(int* key, int* val, int lo, int hi) {
    // explicit stack: every pending range is a distinct, non-empty slice, so
    // CAP (the array length) entries always suffice
    int stk_lo[CAP], stk_hi[CAP];
    int sp = 0;
    stk_lo[sp] = lo;
    stk_hi[sp] = hi;
    sp++;
    while (sp > 0) {
        sp--;
        int L = stk_lo[sp], H = stk_hi[sp];
        int i = L, j = H;
        int x = key[(L + H) / 2];
        do {
            while (key[i] < x) i++;
            while (x < key[j]) j--;
            if (i <= j) {
                int tk = key[i]; key[i] = key[j]; key[j] = tk;
                int tv = val[i]; val[i] = val[j]; val[j] = tv;
                i++;
                j--;
            }
        } while (i <= j);
        // recursion order of the reference: (L, j) first, then (i, H); a stack
        // reverses, so push (i, H) first.
        if (i < H && sp < CAP) { stk_lo[sp] = i; stk_hi[sp] = H; sp++; }
        if (L < j && sp < CAP) { stk_lo[sp] = L; stk_hi[sp] = j; sp++; }
    }
}

RSX_HD int uf_find(int* id, int x) {
    while (x != id[x]) {
        int p = id[id[x]];
        id[x] = p;
        x = p;
    }
    return x;
}

// A ball-only tick has no box queries, so it keeps no GJK / EPA working set.
struct NoGjk {};

// Cars in use.  An arena of MAXC = 0 car slots is the ball alone (the
// ball-prediction rollouts), for which the count is a compile-time zero and
// every car branch of the tick drops out.
template <int MAXC>
RSX_HD int ncars_of(const Arena<MAXC>& A) {
    return MAXC == 0 ? 0 : A.ncars;
}

template <int MAXC>
struct TickScratch {
    MfList<MAXC> L;
    Solver<MAXC> S;
    V3 bmin[MAXC + 1], bmax[MAXC + 1];   // broadphase AABBs of dynamic bodies
    int cand[kMaxCand];                  // a body's candidate keys (pooled car queries: the kept ones), key order
    int pool_base;                       // the lane's first slot in the wavefront's pooled results
    bool pool_local;                     // this car slot's queries run in the lane (its box spans > 2 cells)
    std::conditional_t<(MAXC > 0), gjk::GjkEpa, NoGjk> gj;  // box-vs-triangle GJK / EPA working set
    float cull_pred[MAXC + 1];           // learned-cull distances (uu), -1e30: not usable
    int cull_age;                        // ticks since the network last ran (uniform over the wavefront)
    unsigned long long cull_counts[3];   // the wavefront's queries, culled body-ticks, audit defects
};

template <int MAXC>
RSX_HD Xf body_xf(const Arena<MAXC>& A, int code) {
    if (code == 0) return A.ball.xf;
    if (code >= 1 && code <= MAXC) return A.car[code - 1].xf;
    Xf t = xf_identity();
    int s = code - kBodyStatic;
    if (s >= kNumMeshes) t.o = plane_origin(s - kNumMeshes);
    return t;
}

// Contact-added callback (Arena::_BulletContactAddedCallback, Arena.cpp:218-427)
template <int MAXC>
RSX_HD void contact_added(Arena<MAXC>& A, const TickCtx& T, Manifold& m, int idx, int tri = -1) {
    Contact& cp = m.pt[idx];
    int b0 = m.body0, b1 = m.body1;
    bool st1 = b1 >= kBodyStatic;
    if (b0 == 0 && st1) {
        // ball + world: special resolution, then internal-edge correction
        cp.special = true;
        if (tri >= 0) adjust_internal_edge(T.mv, tri, cp);
        return;
    }
    if (b0 >= 1 && b0 <= MAXC && st1) {
        Car& c = A.car[b0 - 1];
        c.world_contact = true;
        c.world_contact_normal = cp.normal;
        cp.friction = kCarWorldFriction;
        cp.restitution = kCarWorldRestitution;
        if (tri >= 0) adjust_internal_edge(T.mv, tri, cp);
        return;
    }
    if (b0 == 0 && b1 >= 1 && b1 <= MAXC) {
        // car + ball (Arena.cpp:283-334); the ball is manifold body A
        Car& c = A.car[b1 - 1];
        Ball& ball = A.ball;
        cp.friction = kCarBallFriction;
        cp.restitution = kCarBallRestitution;
        c.ball_hit_valid = true;
        c.ball_hit_rel_pos = cp.localA * kBT2UU;
        c.ball_hit_tick = A.tick;
        V3 ball_pos = ball.xf.o * kBT2UU;
        c.ball_hit_ball_pos = ball_pos;
        c.ball_hit_extra_vel = V3{0, 0, 0};
        if ((A.tick > c.ball_hit_extra_tick + 1) || (c.ball_hit_extra_tick > A.tick)) {
            c.ball_hit_extra_tick = A.tick;
        } else {
            return;
        }
        V3 car_fwd_v = col(c.xf.b, 0);
        V3 rel_pos = ball_pos - c.xf.o * kBT2UU;
        V3 rel_vel = ball.vel * kBT2UU - c.vel * kBT2UU;
        float rel_speed = minf(len(rel_vel), kBallCarExtraImpulseMaxDeltaVel);
        if (rel_speed > 0) {
            V3 hit_dir = safe_normalized(rel_pos * V3{1, 1, kBallCarExtraImpulseZScale});
            V3 adj = car_fwd_v * dot(hit_dir, car_fwd_v) * (1 - kBallCarExtraImpulseForwardScale);
            hit_dir = safe_normalized(hit_dir - adj);
            V3 added = (hit_dir * rel_speed) * curve_eval(curve_ball_car_extra(), rel_speed) * 1.f;
            c.ball_hit_extra_vel = added;
            ball.vel_cache += added * kUU2BT;
        }
        return;
    }
    if (b0 >= 1 && b0 <= MAXC && b1 >= 1 && b1 <= MAXC) {
        // car + car (Arena.cpp:336-418): manifold body A is "car1"
        cp.friction = kCarCarFriction;
        cp.restitution = kCarCarRestitution;
        int i1 = b0 - 1, i2 = b1 - 1;
        for (int pass = 0; pass < 2; pass++) {
            const bool swapped = pass == 1;
            if (swapped) {
                int t = i1;
                i1 = i2;
                i2 = t;
            }
            Car& c1 = A.car[i1];
            Car& c2 = A.car[i2];
            if (c1.is_demoed || c2.is_demoed) return;
            if (c1.car_contact_other_id == c2.id && c1.car_contact_cooldown > 0) continue;
            const V3 pos1 = c1.xf.o * kBT2UU, pos2 = c2.xf.o * kBT2UU;
            const V3 vel1 = c1.vel * kBT2UU, vel2 = c2.vel * kBT2UU;
            const V3 delta = pos2 - pos1;
            if (dot(vel1, delta) > 0) {
                const V3 vel_dir = rl_norm(vel1);
                const V3 dir_other = rl_norm(delta);
                const float toward = dot(vel1, dir_other);
                const float away = dot(vel2, vel_dir);
                if (toward > away) {
                    const V3 lp = swapped ? cp.localB : cp.localA;
                    if ((lp.x * kBT2UU) > kBumpMinForwardDist) {
                        bool demo = c1.is_supersonic;
                        if (demo) demo = c1.team != c2.team;  // team demos off
                        if (demo) {
                            c2.is_demoed = true;
                            c2.demo_respawn_timer = kDemoRespawnTime;
                        } else {
                            const bool ground = c2.on_ground;
                            const float base = curve_eval(ground ? curve_bump_ground() : curve_bump_air(), toward);
                            const V3 up = c2.on_ground ? col(c2.xf.b, 2) : V3{0, 0, 1};
                            const V3 imp = vel_dir * base + up * curve_eval(curve_bump_up(), toward) * 1.f;
                            c2.vel_cache += imp * kUU2BT;
                        }
                        c1.car_contact_other_id = c2.id;
                        c1.car_contact_cooldown = kBumpCooldownTime;
                    }
                }
            }
        }
        return;
    }
    (void)T;
}

// Result of one pooled contact query (a car's gjk::closest_points, the
// ball's sphere-triangle test), with the internal-edge correction of the
// contact it adds (adjust_internal_edge).
struct PooledContact {
    V3 normal_on_b, point_on_b;
    float depth;
    int hit;          // bit 0: a point within reach; bit 1: its normal is corrected to `normal_fix`
    V3 normal_fix;
};

// LDS of a wavefront's pooled contact queries: the 64 lanes' body
// transforms, item counts, the item list (owner lane, triangle) and results.
// Queries past kPoolCap run in their owner lane.
constexpr int kPoolCap = 448;
constexpr int kPoolXfOff = 0;
constexpr int kPoolCntOff = kPoolXfOff + 64 * (int)sizeof(Xf);
constexpr int kPoolItemOff = kPoolCntOff + 64 * 4;
constexpr int kPoolResOff = kPoolItemOff + kPoolCap * 4;
constexpr int kPoolBytes = kPoolResOff + kPoolCap * (int)sizeof(PooledContact);

// Narrowphase of the ball against mesh `mesh` (convex-concave, swapped).
// The candidates are sorted by mesh and the meshes visited in increasing
// order, so the scan starts at `pos` (where the previous mesh's run ended)
// and leaves it at the end of this mesh's run.  With pooling the candidates
// are the triangles the ball's sphere test kept, whose contacts (and their
// internal-edge corrections) ran in body_mesh_pool: only their results are
// read here, in the same order.
template <int MAXC>
RSX_HD void ball_vs_mesh(Arena<MAXC>& A, const TickCtx& T, const TickScratch<MAXC>& W, Manifold& m, int mesh,
                         int ncand, int& pos) {
    const Ball& b = A.ball;
    const int* cand = W.cand;
    Xf trB = xf_identity();
    const PooledContact* pooled =
        T.pool && !W.pool_local ? (const PooledContact*)(T.pool + kPoolResOff) : nullptr;
    int mn = m.n;  // the manifold's point count and threshold in registers, m.n stored before the refresh
    const float brk = m.breaking;
    int i = pos;
    int pk_next = i < ncand ? cand[i] : 0;
    for (; i < ncand; i++) {
        const int pk = pk_next;
        if (i + 1 < ncand) pk_next = cand[i + 1];  // the next key's load overlaps this one's contact
        if (cand_mesh(pk) < mesh) continue;
        if (cand_mesh(pk) > mesh) break;
        const int t = cand_tri(pk);
        if (pooled && W.pool_base + i < kPoolCap) {
            const PooledContact& r = pooled[W.pool_base + i];
            const int hit = r.hit;
            if (hit) {
                int idx = manifold_add_n(m, mn, brk, b.xf, trB, r.normal_on_b, r.point_on_b, r.depth,
                                         minf(kBallFriction, 0.6f), maxf(kBallRestitution, 0.3f));
                if (idx >= 0) {
                    contact_added(A, T, m, idx);  // the edge correction ran with the query
                    if (hit & 2) contact_renormal(m.pt[idx], r.normal_fix);
                }
            }
            continue;
        }
        V3 v0, v1, v2;
        load_tri(T.mv, t, v0, v1, v2);
        if (sphere_tri_early_out(b.xf.o, T.bg.radius, v0, v1, v2, brk)) continue;
        V3 pt, nrm;
        float depth;
        if (sphere_triangle(b.xf.o, T.bg.radius, brk, v0, v1, v2, pt, nrm, depth)) {
            int idx = manifold_add_n(m, mn, brk, b.xf, trB, nrm, pt, depth, minf(kBallFriction, 0.6f),
                                     maxf(kBallRestitution, 0.3f));
            if (idx >= 0) contact_added(A, T, m, idx, t);
        }
    }
    pos = i;
    m.n = mn;
    manifold_refresh(m, b.xf, trB);
}

// Early out of a box-triangle query: the box (world transform `child`,
// half extents with margin `hm`) lies wholly on one side of the triangle's
// plane, farther than `breaking` from it.
RSX_HD bool box_tri_separated(const Xf& child, const M3& inv_basis, V3 hm, const V3* tri, float breaking) {
    V3 n = normalized(cross(tri[1] - tri[0], tri[2] - tri[0]));
    for (int side = 0; side < 2; side++) {
        V3 lp = gjk::box_support(hm, inv_basis * n);
        V3 wp = xform(child, lp);
        float dist = dot(n, tri[0]) - dot(n, wp);
        if (dist > breaking) return true;
        n = n * -1.f;
    }
    return false;
}


// LDS of the gather stage of a wavefront's pooled car queries, laid over the
// results (written only after it): per lane its block's first entry, entry
// count and exclusive prefix, its broadphase mesh mask, the meshes its
// candidates belong to, its box, the inverse of its child basis, and a
// 256-bit mask over its block's entries that the filter keeps.
constexpr int kGatFirstOff = kPoolResOff;
constexpr int kGatCntOff = kGatFirstOff + 64 * 4;
constexpr int kGatPreOff = kGatCntOff + 64 * 4;
constexpr int kGatBpOff = kGatPreOff + 64 * 4;
constexpr int kGatMeshOff = kGatBpOff + 64 * 4;
constexpr int kGatBoxOff = kGatMeshOff + 64 * 4;
constexpr int kGatInvOff = kGatBoxOff + 64 * 2 * (int)sizeof(V3);
constexpr int kGatKeptOff = kGatInvOff + 64 * (int)sizeof(M3);
constexpr int kGatEnd = kGatKeptOff + 64 * 4 * 8;
static_assert(kGatKeptOff % 8 == 0 && kGatEnd <= kPoolBytes && kMaxCand <= 256, "car pool gather layout");

#if defined(__HIP_DEVICE_COMPILE__)
// Body `d`'s mesh queries (0 the ball, 1 + i car i) on every lane of the
// wavefront, spread over its 64 lanes: the candidate gather (gather_tris's
// block scan under the body's box) and the filter (meshes outside the
// broadphase mask `bp`; for a car box_tri_separated, for the ball the
// sphere-triangle early-out and test of ball_vs_mesh), then the contact
// queries (a car's gjk::closest_points, the ball's sphere-triangle contact,
// each with the contact's internal-edge correction).  Each lane puts its
// block range, box and transform in LDS; the wavefront tests the block
// entries 64 at a time, marking in the owner's masks the meshes of its
// candidates and the entries the filter keeps; each owner lists its kept
// triangles in key order (W.cand, `ncand` of them) and as query items, which
// every lane then runs lane, lane + 64, ..., the results waiting in LDS for
// car_vs_mesh / ball_vs_mesh.  One lane's gather and tests would otherwise
// run one after another while the rest of its wavefront idles.  Returns the
// meshes with candidates.  A lane whose box spans more than two grid cells
// along some axis gathers with gather_tris and runs its tests in its
// narrowphase (W.pool_local); so does the ball of an arena without car
// slots, over the kept list.  `want` false: no candidates.  Every lane of the
// wavefront must call it, with the same d.
template <int MAXC>
__device__ uint32_t body_mesh_pool(const Arena<MAXC>& A, const TickCtx& T, TickScratch<MAXC>& W, int d, bool want,
                                   uint32_t bp, int& ncand) {
    uint32_t cmask = 0;
    ncand = 0;
    W.pool_local = false;
    const MeshView& w = T.mv;
    const int l = threadIdx.x & 63;
    const CarGeom& g = T.cg;
    Xf* xfs = (Xf*)(T.pool + kPoolXfOff);
    uint32_t* items = (uint32_t*)(T.pool + kPoolItemOff);
    PooledContact* res = (PooledContact*)(T.pool + kPoolResOff);
    int* gfirst = (int*)(T.pool + kGatFirstOff);
    int* gcnt = (int*)(T.pool + kGatCntOff);
    int* gpre = (int*)(T.pool + kGatPreOff);
    uint32_t* gbp = (uint32_t*)(T.pool + kGatBpOff);
    uint32_t* gmesh = (uint32_t*)(T.pool + kGatMeshOff);
    V3* gbox = (V3*)(T.pool + kGatBoxOff);
    M3* ginv = (M3*)(T.pool + kGatInvOff);
    unsigned long long* gkept = (unsigned long long*)(T.pool + kGatKeptOff);
    Xf child = xf_identity();
    V3 cmn = V3{0, 0, 0}, cmx = V3{0, 0, 0};
    int s = 0, n = 0;
    if (want && d == 0) {
        const V3 r = V3{T.bg.radius + 0.08f, T.bg.radius + 0.08f, T.bg.radius + 0.08f};
        child.o = A.ball.xf.o;
        cmn = A.ball.xf.o - r;
        cmx = A.ball.xf.o + r;
    } else if (want) {
        child = car_child_xf(g, A.car[d - 1].xf);
        box_child_aabb(g, child, cmn, cmx);
    }
    if (want) {
        int blk;
        if (box_block(w, cmn, cmx, blk)) {
            s = w.blk_start[blk];
            n = w.blk_start[blk + 1] - s;
        } else {
            W.pool_local = true;
            ncand = gather_tris(w, cmn, cmx, W.cand, kMaxCand);
            for (int i = 0; i < ncand; i++) cmask |= 1u << cand_mesh(W.cand[i]);
        }
    }
    int incl = n;  // inclusive prefix of the entry counts over the lanes
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (l >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    sdf::wave_sync();  // the previous readers of the pool are done
    xfs[l] = child;
    gfirst[l] = s;
    gcnt[l] = n;
    gpre[l] = incl - n;
    gbp[l] = bp;
    gmesh[l] = 0;
    gbox[2 * l] = cmn;
    gbox[2 * l + 1] = cmx;
    if (n && d >= 1) ginv[l] = inverse(child.b);
    for (int k = 0; k < 4; k++) gkept[l * 4 + k] = 0;
    sdf::wave_sync();
    const V3 hm = g.box_implicit + V3{g.box_margin, g.box_margin, g.box_margin};
    for (int t0 = 0; t0 < total; t0 += 64) {
        const int t = t0 + l;
        if (t >= total) continue;
        // the owner: the last lane whose exclusive prefix is <= t (it has entries)
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (gpre[j + step] <= t) j += step;
        const int off = t - gpre[j];
        const GridEntry e = w.blk_ent[gfirst[j] + off];
        const V3 mn = gbox[2 * j], mx = gbox[2 * j + 1];
        // gather_tris's AABB test: z, then x, then y
        if (e.mn[2] > mx.z || e.mx[2] < mn.z) continue;
        if (e.mn[0] > mx.x || e.mx[0] < mn.x) continue;
        if (e.mn[1] > mx.y || e.mx[1] < mn.y) continue;
        const int mesh = cand_mesh(e.key);
        atomicOr(&gmesh[j], 1u << mesh);
        if (!((gbp[j] >> mesh) & 1u)) continue;
        V3 tri[3];
        load_tri(w, cand_tri(e.key), tri[0], tri[1], tri[2]);
        if (d == 0) {
            const V3 c = xfs[j].o;
            if (sphere_tri_early_out(c, T.bg.radius, tri[0], tri[1], tri[2], T.bg.breaking)) continue;
            V3 pt, nrm;
            float depth;
            if (!sphere_triangle(c, T.bg.radius, T.bg.breaking, tri[0], tri[1], tri[2], pt, nrm, depth)) continue;
        } else if (box_tri_separated(xfs[j], ginv[j], hm, tri, g.breaking)) {
            continue;
        }
        atomicOr(&gkept[j * 4 + (off >> 6)], 1ull << (off & 63));
    }
    sdf::wave_sync();
    if (n) {
        // the kept entries in block order, which is key order
        cmask = gmesh[l];
        for (int k = 0; k < 4; k++) {
            unsigned long long b = gkept[l * 4 + k];
            while (b) {
                const int bit = __builtin_ctzll(b);
                b &= b - 1;
                W.cand[ncand++] = w.blk_ent[s + k * 64 + bit].key;
            }
        }
    }
    if constexpr (MAXC == 0) {
        // an arena without car slots (its kernels hold no results array):
        // ball_vs_mesh runs the kept triangles' tests in the lane
        W.pool_local = true;
        return cmask;
    } else {
    const int kept = W.pool_local ? 0 : ncand;
    int kincl = kept;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(kincl, o);
        if (l >= o) kincl += v;
    }
    const int ktotal = __shfl(kincl, 63);
    const int base = kincl - kept;
    W.pool_base = base;
    for (int r = 0; r < kept && base + r < kPoolCap; r++)
        items[base + r] = ((uint32_t)l << 16) | (uint32_t)cand_tri(W.cand[r]);
    sdf::wave_sync();
    const int nq = ktotal < kPoolCap ? ktotal : kPoolCap;
    const float max_d = g.box_margin + g.breaking;
    for (int it0 = 0; it0 < nq; it0 += 64) {
        const int it = it0 + l;
        if (it < nq && d == 0) {
            // ball_vs_mesh's sphere-triangle contact, then contact_added's
            // internal-edge correction
            const uint32_t e = items[it];
            V3 v0, v1, v2, pt, nrm;
            float depth;
            load_tri(T.mv, (int)(e & 0xffffu), v0, v1, v2);
            PooledContact r;
            r.hit = sphere_triangle(xfs[e >> 16].o, T.bg.radius, T.bg.breaking, v0, v1, v2, pt, nrm, depth) ? 1 : 0;
            r.normal_on_b = nrm;
            r.point_on_b = pt;
            r.depth = depth;
            if (r.hit && !(depth > T.bg.breaking)) {
                Contact cp;
                cp.posB = pt;
                cp.posA = pt + nrm * depth;
                cp.localB = inv_xform(xf_identity(), pt);
                cp.normal = nrm;
                cp.dist = depth;
                if (adjust_internal_edge(T.mv, (int)(e & 0xffffu), cp)) {
                    r.hit |= 2;
                    r.normal_fix = cp.normal;
                }
            }
            res[it] = r;
        } else if (it < nq) {
            const uint32_t e = items[it];
            gjk::Shapes sh;
            load_tri(T.mv, (int)(e & 0xffffu), sh.tri[0], sh.tri[1], sh.tri[2]);
            sh.h = g.box_implicit;
            sh.margin = g.box_margin;
            gjk::Contact1 ct{};
            PooledContact r;
            r.hit = gjk::closest_points(W.gj, *T.sx, sh, xfs[e >> 16], xf_identity(), max_d * max_d, ct) ? 1 : 0;
            r.normal_on_b = ct.normal_on_b;
            r.point_on_b = ct.point_on_b;
            r.depth = ct.depth;
            if (r.hit && !(ct.depth > g.breaking)) {
                // the point as manifold_add stores it, then contact_added's
                // internal-edge correction, here instead of in the owner lane
                Contact cp;
                cp.posB = ct.point_on_b;
                cp.posA = ct.point_on_b + ct.normal_on_b * ct.depth;
                cp.localB = inv_xform(xf_identity(), ct.point_on_b);
                cp.normal = ct.normal_on_b;
                cp.dist = ct.depth;
                if (adjust_internal_edge(T.mv, (int)(e & 0xffffu), cp)) {
                    r.hit |= 2;
                    r.normal_fix = cp.normal;
                }
            }
            res[it] = r;
        }
    }
    sdf::wave_sync();
  }
    return cmask;
}

// LDS of a wavefront's pooled wheel-ray tests, in the pool's space (free
// during the pre-tick): each lane's four ray segments, its block's first
// entry and entry count, the lanes' exclusive prefix counts, and per ray the
// smallest (fraction bits, packed key) found so far.
constexpr int kRayPoolSegOff = 0;
constexpr int kRayPoolFirstOff = kRayPoolSegOff + 64 * 8 * (int)sizeof(V3);
constexpr int kRayPoolCntOff = kRayPoolFirstOff + 64 * 4;
constexpr int kRayPoolPreOff = kRayPoolCntOff + 64 * 4;
constexpr int kRayPoolBestOff = kRayPoolPreOff + 64 * 4;
constexpr int kRayPoolBytes = kRayPoolBestOff + 64 * 4 * 8;
static_assert(kRayPoolBestOff % 8 == 0 && kRayPoolBytes <= kPoolBytes, "ray pool layout");

// The mesh part of car `ci`'s four suspension rays on every lane of the
// wavefront (ray_mesh4's results), spread over its 64 lanes: each lane puts
// its rays and its block of grid entries in LDS, and the wavefront's entries
// are tested 64 at a time, each against every ray of its owner whose box it
// meets.  A ray's hit is the smallest (fraction, packed key) among the
// triangles it crosses below fraction 1; a crossing fraction is da / (da - db)
// with da and db of opposite signs, never negative, so its bits order as the
// float does and one 64-bit minimum of (fraction bits, key) keeps it.  The
// owner then recomputes that triangle's crossing for the normal.  A lane
// whose rays' union spans more than two cells per axis walks the grid per ray
// afterwards, as ray_mesh4 does.  `want` false: the lane casts nothing (hit
// stays object -1).  Every lane of the wavefront must call it.
template <int MAXC>
__device__ void car_mesh_rays_pool(const Arena<MAXC>& A, const TickCtx& T, int ci, bool want, RayHit* hit) {
  if constexpr (MAXC > 0) {
    const MeshView& w = T.mv;
    const int l = threadIdx.x & 63;
    V3* seg = (V3*)(T.pool + kRayPoolSegOff);
    int* first = (int*)(T.pool + kRayPoolFirstOff);
    int* cnt = (int*)(T.pool + kRayPoolCntOff);
    int* pre = (int*)(T.pool + kRayPoolPreOff);
    unsigned long long* best = (unsigned long long*)(T.pool + kRayPoolBestOff);
    V3 from[4], to[4];
    for (int k = 0; k < 4; k++) {
        hit[k].object = -1;
        hit[k].fraction = 1.0f;
    }
    bool walk = false;
    int s = 0, n = 0;
    if (want) {
        V3 mn[4], mx[4];
        for (int k = 0; k < 4; k++) {
            wheel_ray_segment(A.car[ci], T.cg, k, from[k], to[k]);
            mn[k] = vmin(from[k], to[k]);
            mx[k] = vmax(from[k], to[k]);
        }
        int blk;
        if (ray4_block(w, mn, mx, blk)) {
            s = w.blk_start[blk];
            n = w.blk_start[blk + 1] - s;
        } else {
            walk = true;
        }
    }
    int incl = n;  // inclusive prefix of the entry counts over the lanes
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (l >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    sdf::wave_sync();  // the previous readers of the pool are done
    if (n) {
        for (int k = 0; k < 4; k++) {
            seg[l * 8 + 2 * k] = from[k];
            seg[l * 8 + 2 * k + 1] = to[k];
        }
    }
    for (int k = 0; k < 4; k++) best[l * 4 + k] = ~0ull;
    first[l] = s;
    cnt[l] = n;
    pre[l] = incl - n;
    sdf::wave_sync();
    for (int t0 = 0; t0 < total; t0 += 64) {
        const int t = t0 + l;
        if (t >= total) continue;
        // the owner: the last lane whose exclusive prefix is <= t (it has entries)
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (pre[j + step] <= t) j += step;
        const GridEntry g = w.blk_ent[first[j] + (t - pre[j])];
        V3 f[4], e[4];
        bool ov[4];
        bool any = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f[k] = seg[j * 8 + 2 * k];
            e[k] = seg[j * 8 + 2 * k + 1];
            ov[k] = entry_meets_segment(g, vmin(f[k], e[k]), vmax(f[k], e[k]));
            any = any || ov[k];
        }
        if (!any) continue;
        V3 a, b, c;
        load_tri(w, cand_tri(g.key), a, b, c);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!ov[k]) continue;
            V3 nrm;
            float d;
            if (!ray_triangle_hit(f[k], e[k], a, b, c, d, nrm) || !(d < 1.0f)) continue;
            atomicMin(&best[j * 4 + k], ((unsigned long long)__float_as_uint(d) << 32) | (unsigned)g.key);
        }
    }
    sdf::wave_sync();
    if (n) {
        for (int k = 0; k < 4; k++) {
            const unsigned long long v = best[l * 4 + k];
            if (v == ~0ull) continue;
            V3 a, b, c, nrm;
            float d;
            load_tri(w, cand_tri((int)(unsigned)v), a, b, c);
            ray_triangle_hit(from[k], to[k], a, b, c, d, nrm);
            hit[k].fraction = d;
            hit[k].normal = nrm;
            hit[k].object = 0;
        }
    }
    if (walk)
        for (int k = 0; k < 4; k++) ray_mesh(w, from[k], to[k], hit[k]);
  }
}
#endif

// Narrowphase of car `ci`'s box against mesh `mesh`
// (btConvexConcaveCollisionAlgorithm with the box as the convex body: the
// triangle early-out, then btConvexConvexAlgorithm's GJK / EPA per triangle;
// one manifold per car and mesh, refreshed after all triangles).  With
// pooling the candidates are the triangles the filter kept, whose queries
// ran in body_mesh_pool: only their results are read here, in the same order.
template <int MAXC>
RSX_HD void car_vs_mesh(Arena<MAXC>& A, const TickCtx& T, TickScratch<MAXC>& W, Manifold& m, int ci, int mesh,
                        int ncand, int& pos) {
  if constexpr (MAXC > 0) {
    gjk::GjkEpa& G = W.gj;
    const int* cand = W.cand;
    const CarGeom& g = T.cg;
    Car& c = A.car[ci];
    const Xf child = car_child_xf(g, c.xf);
    const Xf trB = xf_identity();
    const V3 hm = g.box_implicit + V3{g.box_margin, g.box_margin, g.box_margin};
    const float brk = m.breaking;  // with the point count, in registers (m.n stored before the refresh)
    int mn = m.n;
    const float max_d = g.box_margin + 0.f + brk;
    const PooledContact* pooled =
        T.pool && !W.pool_local ? (const PooledContact*)(T.pool + kPoolResOff) : nullptr;
    M3 inv_basis{};  // for the triangle early-out, which only the lane path runs
    if (!pooled) inv_basis = inverse(child.b);
    int i = pos;
    int pk_next = i < ncand ? cand[i] : 0;
    for (; i < ncand; i++) {
        const int pk = pk_next;
        if (i + 1 < ncand) pk_next = cand[i + 1];  // the next key's load overlaps this one's contact
        if (cand_mesh(pk) < mesh) continue;
        if (cand_mesh(pk) > mesh) break;
        const int t = cand_tri(pk);
        if (pooled) {
            // the wavefront ran this lane's queries already (body_mesh_pool)
            const int it = W.pool_base + i;
            if (it < kPoolCap) {
                const PooledContact& r = pooled[it];
                const int hit = r.hit;
                if (hit) {
                    int idx = manifold_add_n(m, mn, brk, c.xf, trB, r.normal_on_b, r.point_on_b, r.depth, 0.6f, 0.3f);
                    if (idx >= 0) {
                        contact_added(A, T, m, idx);  // the edge correction ran with the query
                        if (hit & 2) contact_renormal(m.pt[idx], r.normal_fix);
                    }
                }
                continue;
            }
        }
        gjk::Shapes sh;
        load_tri(T.mv, t, sh.tri[0], sh.tri[1], sh.tri[2]);
        if (!pooled && box_tri_separated(child, inv_basis, hm, sh.tri, brk)) continue;
        sh.h = g.box_implicit;
        sh.margin = g.box_margin;
        gjk::Contact1 ct;
        if (gjk::closest_points(G, *T.sx, sh, child, trB, max_d * max_d, ct)) {
            int idx = manifold_add_n(m, mn, brk, c.xf, trB, ct.normal_on_b, ct.point_on_b, ct.depth, 0.6f, 0.3f);
            if (idx >= 0) contact_added(A, T, m, idx, t);
        }
    }
    pos = i;
    m.n = mn;
    if (mn) manifold_refresh(m, c.xf, trB);
  }
}

// GJK result for the ball's centre point against the car's core box (margins
// added back as btGjkPairDetector does).
RSX_HD bool sphere_box(V3 c, float r, const Xf& box, V3 h, float margin, float breaking, V3& normal, V3& point_b,
                       float& dist) {
    V3 p = inv_xform(box, c);
    V3 q = V3{clampf(p.x, -h.x, h.x), clampf(p.y, -h.y, h.y), clampf(p.z, -h.z, h.z)};
    bool inside = (q == p);
    if (!inside) {
        V3 qw = xform(box, q);
        V3 axis = c - qw;
        float s2 = len2(axis);
        float s = sqrtf(s2);
        float marg = r + margin;
        // early exit of GJK when clearly separated
        float maxd = r + margin + breaking;
        if (s2 > maxd * maxd) return false;
        normal = axis * (1.f / s);
        point_b = qw + axis * (margin / s);
        dist = s - marg;
        return true;
    }
    // centre inside the core box: push out along the nearest face
    float dx = h.x - fabsf(p.x), dy = h.y - fabsf(p.y), dz = h.z - fabsf(p.z);
    V3 ln;
    float d;
    V3 face = p;
    if (dx <= dy && dx <= dz) { ln = V3{p.x >= 0 ? 1.f : -1.f, 0, 0}; d = dx; face.x = ln.x * h.x; }
    else if (dy <= dz) { ln = V3{0, p.y >= 0 ? 1.f : -1.f, 0}; d = dy; face.y = ln.y * h.y; }
    else { ln = V3{0, 0, p.z >= 0 ? 1.f : -1.f}; d = dz; face.z = ln.z * h.z; }
    normal = box.b * ln;
    point_b = xform(box, face) + normal * margin;
    dist = -d - r - margin;
    return true;
}

// Whether the lane steps its arena.  With pooled queries a lane that stores
// nothing (it mirrors another lane's arena, or pads the grid past the
// bucket's end) takes part only in the wavefront's collective steps, with no
// work of its own, so that the pools spread the other lanes' work over it.
RSX_HD bool tick_live(const TickCtx& T) { return T.own || !T.pool; }

template <int MAXC>
RSX_HD void arena_world_step(Arena<MAXC>& A, const TickCtx& T, TickScratch<MAXC>& W, const bool* far) {
    const float dt = T.dt;
    const int nb = 1 + ncars_of(A);
    const CarGeom& g = T.cg;
    const bool live = tick_live(T);  // not live: the pools' call sites only
    if (live) {
    // gravity on active bodies (applyGravity): m_gravity = acc * (1 / invMass)
    const V3 gacc = V3{0, 0, kGravityZ} * kUU2BT;
    if (!A.ball.sleeping) A.ball.force += gacc * (1.0f / T.bg.inv_mass);
    for (int i = 0; i < ncars_of(A); i++)
        if (A.car[i].active) A.car[i].force += gacc * (1.0f / g.inv_mass);

    // predictUnconstraintMotion: damping + predicted transforms
    {
        Ball& b = A.ball;
        b.vel *= powf(1.f - kBallDrag, dt);
        b.ang *= powf(1.f - 0.f, dt);
        b.interp = Xf{b.xf.b, b.xf.o + b.vel * dt};
    }
    for (int i = 0; i < ncars_of(A); i++) {
        Car& c = A.car[i];
        c.vel *= powf(1.f - 0.f, dt);
        c.ang *= powf(1.f - 0.f, dt);
        c.interp = integrate_transform(c.xf, c.vel, c.ang, dt);
    }

    // broadphase AABBs (updateSingleAabb, continuous union)
    {
        V3 a0, a1, b0, b1;
        ball_shape_aabb(T.bg, A.ball.xf.o, a0, a1);
        ball_shape_aabb(T.bg, A.ball.interp.o, b0, b1);
        broad_aabb(a0, a1, b0, b1, W.bmin[0], W.bmax[0]);
        for (int i = 0; i < ncars_of(A); i++) {
            car_shape_aabb(g, A.car[i].xf, a0, a1);
            car_shape_aabb(g, A.car[i].interp, b0, b1);
            broad_aabb(a0, a1, b0, b1, W.bmin[1 + i], W.bmax[1 + i]);
        }
        // btRSBroadphase::setAabb: the registrations follow the new AABBs
        A.ball.bp_reg = bp_cell_index(T.mv, W.bmin[0]);
        for (int i = 0; i < ncars_of(A); i++) A.car[i].bp_reg = bp_cell_index(T.mv, W.bmin[1 + i]);
    }
    }  // live

    // ---- pairs, manifolds, contacts (dispatch order) ----
    MfList<MAXC>& L = W.L;
    L.count = 0;
    L.nlive = 0;
    MfCursor C;
    int uf[MAXC + 1];
    for (int i = 0; i < nb; i++) uf[i] = i;
    bool active[MAXC + 1];
    active[0] = !A.ball.sleeping;
    for (int i = 0; i < ncars_of(A); i++) active[1 + i] = A.car[i].active;

    for (int d = 0; d < nb; d++) {
        // a car demolished at the start of the tick has simulation and contact
        // response disabled; one demolished during this tick's collision
        // callbacks keeps both until the next tick (Car::_PreTickUpdate)
        bool demoed = (d >= 1) && !A.car[d - 1].active;
        uint32_t mask = live ? T.mv.bp_meshes[bp_cell_index(T.mv, W.bmin[d])] : 0u;
        // statics: meshes then planes.  Every mesh pair that passes the AABB
        // tests opens its manifold, in this order, even when no triangle can
        // reach the body: Bullet hands every pair's manifold, empty or not, to
        // the island sort (btSimulationIslandManager::buildIslands), whose
        // unstable quicksort orders the manifolds that do hold contacts -- and
        // so the solver's rows -- by where the empty ones sit.  The narrowphase
        // runs only for meshes with candidate triangles (`mask`), gathered once
        // per body; a culled body opens its manifolds and skips the rest.
        const uint32_t open = demoed ? 0u : mask;
        int ncand = 0;
        const bool want = !demoed && mask && active[d] && !far[d];
        bool pooled_body = false;
#if defined(__HIP_DEVICE_COMPILE__)
        if (T.pool) {
            // the body's gather, filter and (cars) box-triangle queries,
            // pooled over the wavefront (every lane calls it)
            pooled_body = true;
            const uint32_t cmask = body_mesh_pool(A, T, W, d, want, mask, ncand);
            mask = want ? (mask & cmask) : 0u;
        }
#endif
        if (!live) continue;
        if (!pooled_body && want) {
            if (d == 0) {
                V3 r = V3{T.bg.radius + 0.08f, T.bg.radius + 0.08f, T.bg.radius + 0.08f};
                ncand = gather_tris(T.mv, A.ball.xf.o - r, A.ball.xf.o + r, W.cand, kMaxCand);
            } else {
                Xf child = car_child_xf(g, A.car[d - 1].xf);
                V3 cmn, cmx;
                box_child_aabb(g, child, cmn, cmx);
                ncand = gather_tris(T.mv, cmn, cmx, W.cand, kMaxCand);
            }
            uint32_t cmask = 0;
            for (int i = 0; i < ncand; i++) cmask |= 1u << cand_mesh(W.cand[i]);
            mask &= cmask;
        } else if (!pooled_body) {
            mask = 0;
        }
        if (demoed) continue;  // no contact response: never solved
        // the body's broadphase AABB and (a car) its compound child's AABB,
        // which no contact of this body changes
        const V3 bmn = W.bmin[d], bmx = W.bmax[d];
        V3 cmn = V3{0, 0, 0}, cmx = V3{0, 0, 0};
        if (d >= 1) box_child_aabb(g, car_child_xf(g, A.car[d - 1].xf), cmn, cmx);
        int cpos = 0;  // the candidates not yet consumed by an earlier mesh
        for (int ms = 0; ms < kNumMeshes; ms++) {
            if (!((open >> ms) & 1)) continue;
            const float* ma = T.mv.mesh_aabb + ms * 6;
            const float th = kContactBreakingThreshold;
            V3 smn = V3{ma[0], ma[1], ma[2]}, smx = V3{ma[3], ma[4], ma[5]};  // shape AABB
            if (!aabb_overlap(bmn, bmx, smn - V3{th, th, th}, smx + V3{th, th, th})) continue;
            if (!active[d]) continue;  // needsCollision: both inactive
            float breaking = (d == 0) ? T.bg.breaking : g.breaking;
            // a car: compound child AABB against the mesh's shape AABB (no threshold)
            if (d >= 1 && !aabb_overlap(cmn, cmx, smn, smx)) continue;
            int s;
            int k = mf_open(L, C, d, -1, breaking, s);
            if (s >= 0) {
                Manifold& m = L.live[s];
                m.body1 = kBodyStatic + ms;
                if ((mask >> ms) & 1) {
                    if (d == 0) ball_vs_mesh(A, T, W, m, ms, ncand, cpos);
                    else car_vs_mesh(A, T, W, m, d - 1, ms, ncand, cpos);
                }
            }
            mf_close(L, C, k, s);
        }
        for (int p = 0; p < kNumPlanes; p++) {
            V3 pmn, pmx;
            plane_shape_aabb(p, pmn, pmx);
            {
                const float th = kContactBreakingThreshold;
                if (!aabb_overlap(bmn, bmx, pmn - V3{th, th, th}, pmx + V3{th, th, th})) continue;
            }
            if (!active[d]) continue;
            float breaking = (d == 0) ? T.bg.breaking : g.breaking;
            // a car: compound child AABB test against the plane's own AABB (no
            // threshold); a child whose AABB misses is dropped with its manifold
            if (d >= 1 && !aabb_overlap(cmn, cmx, pmn, pmx)) continue;
            int s;
            int k = mf_open(L, C, d, -1, breaking, s);
            if (s >= 0) {
                Manifold& m = L.live[s];
                m.body1 = kBodyStatic + kNumMeshes + p;
                Xf trB = Xf{m3_identity(), plane_origin(p)};
                if (d == 0) {
                    V3 nrm, pt;
                    float dist;
                    if (sphere_plane(T.bg, A.ball.xf.o, p, m.breaking, nrm, pt, dist)) {
                        int idx = manifold_add(m, A.ball.xf, trB, nrm, pt, dist, minf(kBallFriction, 0.6f),
                                               maxf(kBallRestitution, 0.3f));
                        if (idx >= 0) contact_added(A, T, m, idx);
                    }
                    if (m.n) manifold_refresh(m, A.ball.xf, trB);
                } else {
                    Car& c = A.car[d - 1];
                    Xf child = car_child_xf(g, c.xf);
                    V3 pn = plane_normal(p);
                    Xf plane_in_convex = xf_mul(xf_inverse(child), trB);
                    Xf convex_in_plane = xf_mul(xf_inverse(trB), child);
                    V3 dir = plane_in_convex.b * (-pn);
                    V3 he = g.box_implicit + V3{g.box_margin, g.box_margin, g.box_margin};
                    V3 vtx = V3{dir.x >= 0 ? he.x : -he.x, dir.y >= 0 ? he.y : -he.y, dir.z >= 0 ? he.z : -he.z};
                    V3 vip = xform(convex_in_plane, vtx);
                    float dist = dot(pn, vip) - 0.f;
                    if (dist < m.breaking) {
                        V3 proj = vip - dist * pn;
                        V3 pt = xform(trB, proj);
                        V3 nrm = trB.b * pn;
                        int idx = manifold_add(m, c.xf, trB, nrm, pt, dist, 0.6f, 0.3f);
                        if (idx >= 0) contact_added(A, T, m, idx);
                    }
                    if (m.n) manifold_refresh(m, c.xf, trB);
                }
            }
            mf_close(L, C, k, s);
        }
        // dynamic pairs with later bodies
        for (int e = d + 1; e < nb; e++) {
            bool edemo = (e >= 1) && !A.car[e - 1].active;
            if (edemo) continue;
            if (!aabb_overlap(W.bmin[d], W.bmax[d], W.bmin[e], W.bmax[e])) continue;
            // islands: unite(d, e)
            {
                int i = uf_find(uf, d), j = uf_find(uf, e);
                if (i != j) uf[i] = j;
            }
            if (!active[d] && !active[e]) continue;
            if (d == 0) {
                // ball vs car e-1: compound child AABB test against the ball's shape AABB
                Car& c = A.car[e - 1];
                Xf child = car_child_xf(g, c.xf);
                V3 cmn, cmx, smn, smx;
                box_child_aabb(g, child, cmn, cmx);
                ball_shape_aabb(T.bg, A.ball.xf.o, smn, smx);
                if (!aabb_overlap(cmn, cmx, smn, smx)) continue;
                float breaking = minf(T.bg.breaking, g.breaking);
                int s;
                int k = mf_open(L, C, 0, e, breaking, s);
                if (s >= 0) {
                    Manifold& m = L.live[s];
                    V3 nrm, ptb;
                    float dist;
                    if (sphere_box(A.ball.xf.o, T.bg.radius, child, g.box_implicit, g.box_margin, m.breaking, nrm, ptb,
                                   dist)) {
                        int idx = manifold_add(m, A.ball.xf, c.xf, nrm, ptb, dist, kBallFriction * kCarCollisionFriction,
                                               kBallRestitution * kCarCollisionRestitution);
                        if (idx >= 0) contact_added(A, T, m, idx);
                    }
                    if (m.n) manifold_refresh(m, A.ball.xf, c.xf);
                }
                mf_close(L, C, k, s);
            }
            if (d >= 1) {
                // car d-1 vs car e-1: compound (car d) -> swapped compound
                // (car e) -> box-box; the manifold's body A is the later car
                const Car& cd = A.car[d - 1];
                const Car& ce = A.car[e - 1];
                Xf xd = car_child_xf(g, cd.xf), xe = car_child_xf(g, ce.xf);
                V3 dmn, dmx, emn, emx;
                box_child_aabb(g, xd, dmn, dmx);
                car_shape_aabb(g, ce.xf, emn, emx);
                if (!aabb_overlap(dmn, dmx, emn, emx)) continue;
                box_child_aabb(g, xe, emn, emx);
                if (!aabb_overlap(emn, emx, dmn, dmx)) continue;
                int s;
                int k = mf_open(L, C, e, d, g.breaking, s);
                if (s >= 0) {
                    Manifold& m = L.live[s];
                    const V3 side = (g.box_implicit + V3{g.box_margin, g.box_margin, g.box_margin}) * 2.f;
                    boxbox::Out bo;
                    bo.n = 0;
                    boxbox::box_box(xe.o, xe.b, side, xd.o, xd.b, side, bo);
                    for (int q = 0; q < bo.n; q++) {
                        int idx = manifold_add(m, ce.xf, cd.xf, bo.normal_on_b[q], bo.point[q], bo.depth[q],
                                               kCarCollisionFriction * kCarCollisionFriction,
                                               kCarCollisionRestitution * kCarCollisionRestitution);
                        if (idx >= 0) contact_added(A, T, m, idx);
                    }
                    manifold_refresh(m, ce.xf, cd.xf);
                }
                mf_close(L, C, k, s);
            }
        }
    }

    mf_finish(L, C);
    if (!live) return;
    // ---- islands and manifold order (btSimulationIslandManager) ----
    int tag[MAXC + 1];
    for (int i = 0; i < nb; i++) tag[i] = uf_find(uf, i);
    // a sleeping ball joined to an awake island wakes for this tick
    bool island_awake[MAXC + 1];
    for (int i = 0; i < nb; i++) island_awake[i] = false;
    for (int i = 0; i < nb; i++)
        if (active[i]) island_awake[tag[i]] = true;
    bool in_solver[MAXC + 1];
    for (int i = 0; i < nb; i++) in_solver[i] = island_awake[tag[i]] && !((i >= 1) && !A.car[i - 1].active);
    bool ball_woken = A.ball.sleeping && in_solver[0];

    constexpr int kCap = MfList<MAXC>::kCap;
    int keys[kCap], order[kCap];
    int nm = 0;
    for (int k = 0; k < L.count; k++) {
        int b0 = L.body0[k];
        keys[nm] = tag[b0];
        order[nm] = k;
        nm++;
    }
    if (nm > 1) bt_quicksort<kCap>(keys, order, 0, nm - 1);

    // ---- solver setup ----
    Solver<MAXC>& S = W.S;
    S.nrow = 0;
    S.nfric = 0;
    const int FIX = MAXC + 1;
    for (int i = 0; i < MAXC + 2; i++) S.body[i].used = false;
    sbody_init_fixed(S.body[FIX]);
    if (in_solver[0]) {
        Ball& b = A.ball;
        M3 ii = m3(T.bg.inv_inertia, 0, 0, 0, T.bg.inv_inertia, 0, 0, 0, T.bg.inv_inertia);
        sbody_init(S.body[0], b.xf, b.vel, b.ang, b.force, V3{0, 0, 0}, T.bg.inv_mass, ii, dt);
    }
    for (int i = 0; i < ncars_of(A); i++) {
        if (!in_solver[1 + i]) continue;
        Car& c = A.car[i];
        sbody_init(S.body[1 + i], c.xf, c.vel, c.ang, c.force, c.torque, g.inv_mass, c.inv_inertia_world, dt);
    }
    for (int i = 0; i <= MAXC; i++) {
        S.nspecial[i] = 0;
        S.sp_normal[i] = V3{0, 0, 0};
        S.sp_dist[i] = 0;
        S.sp_friction[i] = 0;
        S.sp_restitution[i] = 0;
    }
    // the row counts in registers, stored once (the rows' stores would
    // otherwise hold back each reread)
    int nrow = 0, nfric = 0;
    for (int oi = 0; oi < nm; oi++) {
        int k = order[oi];
        int s = k >= 0 ? L.slot[k] : -1;
        if (s < 0) continue;
        const Manifold& m = L.live[s];
        int ia = m.body0;
        int ib = (m.body1 >= kBodyStatic) ? FIX : m.body1;
        if (!in_solver[m.body0] && !(m.body1 < kBodyStatic && in_solver[m.body1])) continue;
        Xf xa = body_xf(A, m.body0);
        Xf xb = body_xf(A, m.body1);
        for (int j = 0; j < m.n; j++) {
            if (nrow >= kMaxRows) break;
            const Contact& cp = m.pt[j];
            V3 rel1 = cp.posA - xa.o;
            V3 rel2 = cp.posB - xb.o;
            int ridx = nrow++;
            Row& r = S.row[ridx];
            setup_contact_row(S, r, ia, ib, cp.normal, cp.dist, cp.friction, cp.restitution, cp.special, rel1, rel2, dt);
            r.fidx = (int16_t)nfric;
            if (cp.special) {
                // only non-static bodies accumulate (the world side is static here)
                int bi = m.body0;
                S.nspecial[bi]++;
                S.sp_friction[bi] = cp.friction;
                S.sp_restitution[bi] = cp.restitution;
                S.sp_normal[bi] += cp.normal;
                S.sp_dist[bi] += len(rel1);
            }
            add_contact_friction(S, nfric, ia, ib, cp.normal, cp.friction, rel1, rel2, ridx);
        }
    }
    // special rows, in body order (only the ball produces them)
    for (int bi = 0; bi < nb; bi++) {
        if (S.nspecial[bi] <= 0 || nrow >= kMaxRows) continue;
        float distance = S.sp_dist[bi] / (float)S.nspecial[bi];
        V3 normal = S.sp_normal[bi] / (float)S.nspecial[bi];
        int ridx = nrow++;
        Row& r = S.row[ridx];
        V3 rel1 = normal * -distance;
        V3 rel2 = V3{0, 0, 0};
        setup_contact_row(S, r, bi, FIX, normal, distance, S.sp_friction[bi], S.sp_restitution[bi], false, rel1, rel2,
                          dt);
        add_contact_friction(S, nfric, bi, FIX, normal, S.sp_friction[bi], rel1, rel2, ridx);
    }

    S.nrow = nrow;
    S.nfric = nfric;

    if (T.slv) solver_iterate_lds(S, T.slv);
    else solver_iterate(S);

    // write-back
    if (in_solver[0]) {
        Ball& b = A.ball;
        sbody_writeback(S.body[0], true, dt, b.xf, b.vel, b.ang);
    }
    for (int i = 0; i < ncars_of(A); i++) {
        if (!in_solver[1 + i]) continue;
        Car& c = A.car[i];
        sbody_writeback(S.body[1 + i], false, dt, c.xf, c.vel, c.ang);
    }

    // integrateTransforms (active bodies)
    if (!A.ball.sleeping || ball_woken) {
        Ball& b = A.ball;
        b.xf.o = b.xf.o + b.vel * dt;
    }
    for (int i = 0; i < ncars_of(A); i++) {
        Car& c = A.car[i];
        if (!c.active) continue;
        c.xf = integrate_transform(c.xf, c.vel, c.ang, dt);
        car_update_inertia(c, g);
    }
    // clearForces
    A.ball.force = V3{0, 0, 0};
    for (int i = 0; i < ncars_of(A); i++) {
        A.car[i].force = V3{0, 0, 0};
        A.car[i].torque = V3{0, 0, 0};
    }
}

// ---- boost pads (BoostPad.cpp:51-105, BoostPadGrid.cpp:5-25) -----------------
template <int MAXC>
RSX_HD void pad_check_car(Arena<MAXC>& A, const TickCtx& T, int ci) {
    Car& c = A.car[ci];
    if (c.is_demoed || c.boost >= 100) return;
    V3 pos = c.xf.o * kBT2UU;
    if (pos.z > kPadGridExtentZ) return;
    int ix = (int)(pos.x / kPadCellSizeX + (kPadCellsX / 2));
    int iy = (int)(pos.y / kPadCellSizeY + (kPadCellsY / 2));
    V3 cmin, cmax;
    car_shape_aabb(T.cg, c.xf, cmin, cmax);
    for (int p = 0; p < kNumPads; p++) {
        const PadGeom& pg = T.pads[p];
        int px = (int)(pg.pos_uu.x / kPadCellSizeX + (kPadCellsX / 2));
        int py = (int)(pg.pos_uu.y / kPadCellSizeY + (kPadCellsY / 2));
        if (px < (ix - 1 > 0 ? ix - 1 : 0) || px > (ix + 1 < kPadCellsX - 1 ? ix + 1 : kPadCellsX - 1)) continue;
        if (py < (iy - 1 > 0 ? iy - 1 : 0) || py > (iy + 1 < kPadCellsY - 1 ? iy + 1 : kPadCellsY - 1)) continue;
        Pad& pad = A.pad[p];
        bool colliding = false;
        V3 cp = c.xf.o;
        if (pad.prev_locked == c.id) {
            colliding = (pg.box_max.x > cmin.x && pg.box_max.y > cmin.y && pg.box_max.z > cmin.z) &&
                        (pg.box_min.x < cmax.x && pg.box_min.y < cmax.y && pg.box_min.z < cmax.z);
        } else {
            float dx = cp.x - pg.pos_bt.x, dy = cp.y - pg.pos_bt.y;
            float r = pg.cyl_rad_bt;
            if (dx * dx + dy * dy < (r * r)) colliding = fabsf(cp.z - pg.pos_bt.z) < (kPadCylHeight * kUU2BT);
        }
        if (colliding) pad.cur_locked = c.id;
    }
}

template <int MAXC>
RSX_HD Car* car_by_id(Arena<MAXC>& A, uint32_t id) {
    for (int i = 0; i < ncars_of(A); i++)
        if (A.car[i].id == id) return &A.car[i];
    return nullptr;
}

RSX_HD uint32_t rng_next(uint32_t& s) {
    // xorshift32
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

// Car::Respawn with the engine's own per-arena draw of the respawn slot.
template <int MAXC>
RSX_HD void car_respawn(Arena<MAXC>& A, const TickCtx& T, int ci);

#if defined(__HIP_DEVICE_COMPILE__)
// Whether a mesh triangle lies within r of q (Bullet units): the exact check
// the learned cull is audited against.
__device__ inline bool mesh_within(const MeshView& w, V3 q, float r) {
    const float inv = 1.0f / w.grid_cell;
    const V3 mn = q - V3{r, r, r}, mx = q + V3{r, r, r};
    const int x0 = clampi((int)floorf((mn.x - w.grid_min.x) * inv), 0, w.gx - 1);
    const int y0 = clampi((int)floorf((mn.y - w.grid_min.y) * inv), 0, w.gy - 1);
    const int z0 = clampi((int)floorf((mn.z - w.grid_min.z) * inv), 0, w.gz - 1);
    const int x1 = clampi((int)floorf((mx.x - w.grid_min.x) * inv), 0, w.gx - 1);
    const int y1 = clampi((int)floorf((mx.y - w.grid_min.y) * inv), 0, w.gy - 1);
    const int z1 = clampi((int)floorf((mx.z - w.grid_min.z) * inv), 0, w.gz - 1);
    for (int ix = x0; ix <= x1; ix++)
        for (int iy = y0; iy <= y1; iy++)
            for (int iz = z0; iz <= z1; iz++) {
                const int cell = (ix * w.gy + iy) * w.gz + iz;
                for (int k = w.cell_start[cell]; k < w.cell_start[cell + 1]; k++) {
                    V3 a, b, c;
                    load_tri(w, w.cell_tris[k], a, b, c);
                    if (sdf::point_tri_dist2(q, a, b, c) <= r * r) return true;
                }
            }
    return false;
}

// The learned broadphase cull of this tick: far[0] for the ball, far[1 + i]
// for car i.  Every cull.every ticks the network predicts each body's
// distance to the mesh; until the next evaluation a body is far when the
// prediction minus the margin exceeds its reach plus the distance it can
// have moved since.  A far body skips the mesh queries (narrowphase and, for
// a car, the mesh part of its wheel rays).  The network is asked only for
// body slots where some lane's body has a mesh listed in its broadphase cell
// (other bodies take the exact path, which is then nearly free).  In audit
// mode every culled body is checked against the exact mesh distance at its
// current position.  Every lane of the wavefront must call it.
template <int MAXC>
__device__ void cull_far(const Arena<MAXC>& A, const TickCtx& T, TickScratch<MAXC>& W, bool* far) {
    const sdf::Cull& C = T.cull;
    const bool eval = W.cull_age >= C.every;
    if (eval) W.cull_age = 0;
    unsigned long long nq = 0, nc = 0, nd = 0;
    for (int b = 0; b <= MAXC; b++) {
        const bool have = (b == 0) || (b - 1 < ncars_of(A));
        V3 q = A.ball.xf.o;
        if (b > 0 && have) q = car_child_xf(T.cg, A.car[b - 1].xf).o;
        if (eval) {
            W.cull_pred[b] = -1e30f;
            const bool need = have && T.mv.bp_meshes[bp_cell_index(T.mv, q)] != 0;
            const unsigned long long any = __ballot(need);
            if (any) {
                const float d = sdf::eval_wave(C.net, q * kBT2UU, T.lds_act);
                if (need) W.cull_pred[b] = d;
                nq += __popcll(__ballot(need && T.own));
            }
        }
        const float reach = b == 0 ? C.reach_ball_uu : C.reach_car_uu;
        const float step = b == 0 ? C.step_ball_uu : C.step_car_uu;
        const float margin = b == 0 ? C.margin_ball_uu : C.margin_car_uu;
        const bool cull = have && (W.cull_pred[b] - margin > reach + step * (float)W.cull_age);
        far[b] = cull;
        const bool defect = cull && C.audit && mesh_within(T.mv, q, reach * kUU2BT);
        nc += __popcll(__ballot(cull && T.own));
        nd += __popcll(__ballot(defect && T.own));
    }
    W.cull_age++;
    W.cull_counts[0] += nq;
    W.cull_counts[1] += nc;
    W.cull_counts[2] += nd;
}

#endif

#if defined(__HIPCC__)
// Adds the wavefront's cull counters to the network's totals, once per
// launch (per-tick atomics from every wavefront would contend on three
// addresses).  Every lane of the wavefront must call it.
template <int MAXC>
__device__ void cull_flush(const TickCtx& T, const TickScratch<MAXC>& W) {
    if (!T.cull.enabled || (threadIdx.x & 63) != 0) return;
    for (int k = 0; k < 3; k++)
        if (W.cull_counts[k]) atomicAdd(&T.cull.stats[k], W.cull_counts[k]);
}
#endif

// One full tick (Arena::Step body).  Returns true if the ball is in a goal
// after the tick (Arena::IsBallScored).  With the learned cull enabled or
// pooled queries every lane of the wavefront must call it, and the
// collective steps (the cull's network, the pools) are reached by all its
// lanes at the same point: the work of a lane that is not live is skipped
// around them.
template <int MAXC>
RSX_HD bool arena_tick(Arena<MAXC>& A, const TickCtx& T, TickScratch<MAXC>& W) {
    const float dt = T.dt;
    const bool live = tick_live(T);
    bool far[MAXC + 1];
    for (int b = 0; b <= MAXC; b++) far[b] = false;
#if defined(__HIP_DEVICE_COMPILE__)
    if (T.cull.enabled) cull_far<MAXC>(A, T, W, far);
#endif
    // ball zero-velocity sleep
    if (live) A.ball.sleeping = (len2(A.ball.vel) == 0 && len2(A.ball.ang) == 0);
    // Cars are visited last-added first, as Arena::Step's loops over
    // Arena::_cars (a std::unordered_set<Car*>) visit them: libstdc++ links a
    // node inserted into an empty bucket at the head of the set's list, so
    // iteration runs in reverse insertion order unless two car pointers share
    // a bucket (the oracle's arenas: every 1v1, 199 of 200 2v2 and 3v3).  The
    // order decides which car's wheel rays see another car's response flag
    // as of this tick, whose impulses a car on another car reads, and which
    // of two cars on one pad locks it.
    for (int i = ncars_of(A) - 1; i >= 0; i--) {
        Car& c = A.car[i];
        if (live) {
            Controls& k = c.controls;
            k.throttle = clampf(k.throttle, -1, 1);
            k.steer = clampf(k.steer, -1, 1);
            k.pitch = clampf(k.pitch, -1, 1);
            k.yaw = clampf(k.yaw, -1, 1);
            k.roll = clampf(k.roll, -1, 1);
            c.active = true;
            if (c.is_demoed) {
                c.demo_respawn_timer = maxf(c.demo_respawn_timer - dt, 0);
                if (c.demo_respawn_timer == 0) {
                    car_respawn(A, T, i);
                    far[1 + i] = false;  // moved after the cull was evaluated
                    W.cull_pred[1 + i] = -1e30f;
                }
                c.active = false;  // body disabled for this tick
            }
        }
#if defined(__HIP_DEVICE_COMPILE__)
        if (T.pool) {
            // the mesh part of this car slot's wheel rays on every lane,
            // spread over the wavefront (a demolished car casts none); the
            // one call site every lane reaches together
            RayHit cast[4];
            car_mesh_rays_pool(A, T, i, live && !c.is_demoed && !far[1 + i], cast);
            if (!live || c.is_demoed) continue;
            car_pre_tick(A, i, T.cg, T.bg, T.mv, dt, dt, *T.sx, !far[1 + i], cast);
            continue;
        }
#endif
        if (c.is_demoed) continue;
        car_pre_tick(A, i, T.cg, T.bg, T.mv, dt, dt, *T.sx, !far[1 + i]);
    }
    if (live && ncars_of(A) > 0) {
        for (int p = 0; p < kNumPads; p++) {
            Pad& pad = A.pad[p];
            if (pad.cooldown > 0) pad.cooldown = maxf(pad.cooldown - dt, 0);
            pad.is_active = (pad.cooldown == 0);
            pad.cur_locked = 0;
        }
    }
    M3 rot0[MAXC > 0 ? MAXC : 1];
    for (int i = 0; i < ncars_of(A); i++) rot0[i] = A.car[i].xf.b;
    arena_world_step(A, T, W, far);
    if (!live) return false;
    for (int i = ncars_of(A) - 1; i >= 0; i--) {  // post-tick and pads, last-added car first
        Car& c = A.car[i];
        // A car demolished during this step skips Car::_PostTickUpdate, which
        // is where its public rotation would have followed the body; the body
        // is disabled from the next tick on, so only that rotation is kept.
        if (c.is_demoed && c.active) c.xf.b = rot0[i];
        if (c.is_demoed) continue;
        car_post_tick(c, dt);
        car_finish_tick(c);
        pad_check_car(A, T, i);
    }
    if (ncars_of(A) > 0) {
        for (int p = 0; p < kNumPads; p++) {
            Pad& pad = A.pad[p];
            uint32_t locked = 0;
            if (pad.cur_locked) {
                locked = pad.cur_locked;
                if (pad.is_active) {
                    Car* c = car_by_id(A, pad.cur_locked);
                    float add = T.pads[p].big ? kPadAmountBig : kPadAmountSmall;
                    if (c) c->boost = minf(c->boost + add, kBoostMax);
                    pad.is_active = false;
                    pad.cooldown = T.pads[p].big ? kPadCooldownBig : kPadCooldownSmall;
                }
            }
            pad.prev_locked = locked;
        }
    }
    {
        Ball& b = A.ball;
        if (!is_zero(b.vel_cache)) {
            b.vel += b.vel_cache;
            b.vel_cache = V3{0, 0, 0};
        }
        V3 v = b.vel, a = b.ang;
        float ms = kBallMaxSpeed * kUU2BT;
        if (len2(v) > ms * ms) v = normalized(v) * ms;
        if (len2(a) > (kBallMaxAngSpeed * kBallMaxAngSpeed)) a = normalized(a) * kBallMaxAngSpeed;
        b.vel = v;
        b.ang = a;
    }
    bool scored = fabsf(A.ball.xf.o.y * kBT2UU) > (kGoalThresholdY + kBallRadius);
    A.tick++;
    return scored;
}

// Respawn spots (RLConst.h CAR_RESPAWN_LOCATIONS_SOCCAR), blue side.
RSX_HD void respawn_spot(int idx, float& x, float& y, float& yaw) {
    const float xs[4] = {-2304, -2688, 2304, 2688};
    x = xs[idx & 3];
    y = -4608;
    yaw = (float)(3.14159265358979323846 / 2);
}

RSX_HD M3 yaw_basis(float yaw) {
    // Angle(yaw, 0, 0).ToRotMat(): btMatrix3x3::setEulerYPR(yaw, -0, -0)
    return m3_euler_zyx(-0.f, -0.f, yaw);
}

template <int MAXC>
RSX_HD void car_set_spawn(Car& c, const CarGeom& g, float x, float y, float z, float yaw, float boost) {
    V3 zero = V3{0, 0, 0};
    c.xf.o = V3{x, y, z} * kUU2BT;
    c.xf.b = yaw_basis(yaw);
    c.vel = c.ang = zero;
    car_update_inertia(c, g);
    c.vel_cache = zero;
    c.force = c.torque = zero;
    c.on_ground = true;
    c.has_jumped = c.has_double_jumped = c.has_flipped = c.is_flipping = c.is_jumping = false;
    c.is_supersonic = c.is_auto_flipping = c.is_demoed = c.world_contact = c.ball_hit_valid = false;
    for (int i = 0; i < 4; i++) c.wheel_contact[i] = false;
    c.boost = boost;
    c.time_spent_boosting = c.jump_time = c.flip_time = c.air_time = c.air_time_since_jump = 0;
    c.supersonic_time = c.handbrake_val = c.auto_flip_timer = c.auto_flip_torque_scale = 0;
    c.demo_respawn_timer = c.car_contact_cooldown = 0;
    c.flip_rel_torque = c.world_contact_normal = zero;
    c.car_contact_other_id = 0;
    c.ball_hit_tick = ~0ULL;
    c.ball_hit_extra_tick = ~0ULL;
    c.ball_hit_rel_pos = c.ball_hit_ball_pos = c.ball_hit_extra_vel = zero;
    c.last_controls = Controls{0, 0, 0, 0, 0, false, false, false};
}

template <int MAXC>
RSX_HD void car_respawn(Arena<MAXC>& A, const TickCtx& T, int ci) {
    Car& c = A.car[ci];
    int idx = (int)(rng_next(A.rng) % 4u);
    float x, y, yaw;
    respawn_spot(idx, x, y, yaw);
    float sgn = c.team == 0 ? 1.f : -1.f;
    float yw = yaw + (c.team == 0 ? 0.f : 3.14159265358979323846f);
    car_set_spawn<MAXC>(c, T.cg, x, y * sgn, kCarRespawnZ, yw, kBoostSpawnAmount);
}

}  // namespace rsx

namespace rsx {

// Kickoff reset (Arena::ResetToRandomKickoff, Arena.cpp:112-216) with the
// engine's own per-arena RNG for the order of the five spawn spots; cars take
// the CarState defaults (Car.h:14-115) and keep their carried wheel values,
// as Car::SetState does.
template <int MAXC>
RSX_HD void arena_reset_kickoff(Arena<MAXC>& A, const CarGeom& g) {
    const float sx[5] = {-2048, 2048, -256, 256, 0};
    const float sy[5] = {-2560, -2560, -3840, -3840, -4608};
    const float syaw[5] = {(float)(0.78539816339744830962 * 1), (float)(0.78539816339744830962 * 3),
                           (float)(0.78539816339744830962 * 2), (float)(0.78539816339744830962 * 2),
                           (float)(0.78539816339744830962 * 2)};
    int order[5] = {0, 1, 2, 3, 4};
    for (int i = 4; i > 0; i--) {
        int j = (int)(rng_next(A.rng) % (uint32_t)(i + 1));
        int t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
    int nb = 0, no = 0;
    for (int i = 0; i < ncars_of(A); i++) {
        Car& c = A.car[i];
        int slot = (c.team == 0) ? nb++ : no++;
        float x, y, yaw;
        if (slot < 5) {
            x = sx[order[slot]];
            y = sy[order[slot]];
            yaw = syaw[order[slot]];
        } else {
            respawn_spot((slot - 5) % 4, x, y, yaw);
        }
        if (c.team != 0) {
            x = -x;
            y = -y;
            yaw = (float)((double)yaw + 3.14159265358979323846);
        }
        car_set_spawn<MAXC>(c, g, x, y, kCarSpawnRestZ, yaw, kBoostSpawnAmount);
    }
    A.ball.xf = xf_identity();
    A.ball.xf.o = V3{0, 0, kBallRestZ} * kUU2BT;
    A.ball.vel = A.ball.ang = A.ball.force = A.ball.vel_cache = V3{0, 0, 0};
    A.ball.interp = A.ball.xf;
    for (int p = 0; p < kNumPads; p++) A.pad[p] = Pad{true, 0.f, 0, 0};
}

}  // namespace rsx
