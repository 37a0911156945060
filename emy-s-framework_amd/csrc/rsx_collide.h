// rsx_collide.h — contact generation of one tick, in the order and with the
// per-point arithmetic of RocketSim's Bullet world.
//
// RocketSim's broadphase (btRSBroadphase.cpp:300-360) removes and re-adds every
// overlapping pair each tick, which destroys and recreates every persistent
// manifold; so the contact set of a tick depends only on that tick's state.
// Pair order: dynamic bodies in creation order (ball, then cars), each paired
// with the static bodies listed in its broadphase cell (meshes 0..15, then the
// four planes, all AABB-tested), then with later dynamic bodies.
//
// Reference behaviour restated here:
//   btManifoldResult::addContactPoint        btManifoldResult.cpp:120-207
//   btPersistentManifold add/sort/refresh    btPersistentManifold.cpp:100-304
//   btConvexPlaneCollisionAlgorithm          btConvexPlaneCollisionAlgorithm.cpp:105-165
//   btConvexConcaveCollisionAlgorithm        btConvexConcaveCollisionAlgorithm.cpp:70-200
//   SphereTriangleDetector                   SphereTriangleDetector.cpp:23-230
//   btCompoundCollisionAlgorithm child AABB  btCompoundCollisionAlgorithm.cpp (ProcessChildShape)
//   btGjkPairDetector result (sphere vs box) btGjkPairDetector.cpp (margins, normal on B)
//   Arena::_BulletContactAddedCallback       Arena.cpp:218-427
//   btCollisionWorld::updateSingleAabb       btCollisionWorld.cpp:143-176
#pragma once

#include "rsx_car.h"
#include "rsx_sdf.h"

namespace rsx {

constexpr int kBodyStatic = 100;  // body code base of static bodies

RSX_HD uint32_t __float_as_uint_rsx(float f) {
    union {
        float f;
        uint32_t u;
    } c;
    c.f = f;
    return c.u;
}

struct TickCtx {
    CarGeom cg;
    BallGeom bg;
    MeshView mv;
    const PadGeom* pads;
    float dt;
    gjk::Simplex* sx;  // this lane's Voronoi simplex working set (LDS in the kernels)
    sdf::Cull cull;    // learned broadphase cull (cull.enabled == 0: exact only)
    uint16_t* lds_act; // the wavefront's LDS activation buffer for the cull's network
    bool own;          // this lane stores its arena (false: it mirrors another lane, or pads the grid)
    uint8_t* pool;     // the wavefront's LDS for pooled box-triangle queries (null: each lane runs its own)
    float* slv = nullptr;  // this lane's column of the wavefront's LDS for the solver's body state (null: private)
};

// --- manifold bookkeeping ----------------------------------------------------
RSX_HD float calc_area4(V3 p0, V3 p1, V3 p2, V3 p3) {
    V3 a0 = p0 - p1, b0 = p2 - p3;
    V3 a1 = p0 - p2, b1 = p1 - p3;
    V3 a2 = p0 - p3, b2 = p1 - p2;
    V3 t0 = cross(a0, b0), t1 = cross(a1, b1), t2 = cross(a2, b2);
    return maxf(maxf(len2(t0), len2(t1)), len2(t2));
}

// btPersistentManifold::sortCachedPoints (gContactCalcArea3Points = true)
RSX_HD int sort_cached_points(const Manifold& m, const Contact& pt) {
    int maxi = -1;
    float maxpen = pt.dist;
    for (int i = 0; i < 4; i++)
        if (m.pt[i].dist < maxpen) {
            maxi = i;
            maxpen = m.pt[i].dist;
        }
    float r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    if (maxi != 0) r0 = len2(cross(pt.localA - m.pt[1].localA, m.pt[3].localA - m.pt[2].localA));
    if (maxi != 1) r1 = len2(cross(pt.localA - m.pt[0].localA, m.pt[3].localA - m.pt[2].localA));
    if (maxi != 2) r2 = len2(cross(pt.localA - m.pt[0].localA, m.pt[3].localA - m.pt[1].localA));
    if (maxi != 3) r3 = len2(cross(pt.localA - m.pt[0].localA, m.pt[2].localA - m.pt[1].localA));
    // btVector4::closestAxis4 = index of max |value|, first wins ties
    float v[4] = {fabsf(r0), fabsf(r1), fabsf(r2), fabsf(r3)};
    int best = -1;
    float bv = -kLargeFloat;
    for (int i = 0; i < 4; i++)
        if (v[i] > bv) {
            best = i;
            bv = v[i];
        }
    return best;
}

// Add a point (normal on manifold body1 pointing to body0, point on body1,
// signed distance); returns the index it landed at or -1 if rejected.  The
// _n form keeps the point count in `n` and the breaking threshold in `brk`
// (registers of a caller adding several points; m.n is stored by it).
RSX_HD int manifold_add_n(Manifold& m, int& n, float brk, const Xf& trA, const Xf& trB, V3 normal, V3 point_b,
                          float depth, float friction, float restitution);
RSX_HD int manifold_add(Manifold& m, const Xf& trA, const Xf& trB, V3 normal, V3 point_b, float depth, float friction,
                        float restitution) {
    int n = m.n;
    const int idx = manifold_add_n(m, n, m.breaking, trA, trB, normal, point_b, depth, friction, restitution);
    m.n = n;
    return idx;
}
RSX_HD int manifold_add_n(Manifold& m, int& n, float brk, const Xf& trA, const Xf& trB, V3 normal, V3 point_b,
                          float depth, float friction, float restitution) {
    if (depth > brk) return -1;
    V3 pa = point_b + normal * depth;
    Contact c;
    c.localA = inv_xform(trA, pa);
    c.localB = inv_xform(trB, point_b);
    c.normal = normal;
    c.dist = depth;
    c.posA = pa;
    c.posB = point_b;
    c.friction = friction;
    c.restitution = restitution;
    c.special = false;
    int idx = n;
    if (idx == kMaxManifoldPoints) {
        idx = sort_cached_points(m, c);
        if (idx < 0) idx = 0;
    } else {
        n++;
    }
    m.pt[idx] = c;
    return idx;
}

RSX_HD void manifold_remove(Manifold& m, int i) {
    int last = m.n - 1;
    if (i != last) m.pt[i] = m.pt[last];
    m.n--;
}

// btPersistentManifold::refreshContactPoints.  The points' fields are read
// once into registers (the per-point stores would otherwise hold back the
// next point's reads) and the removal tests use the refreshed values there;
// the stored points are updated and removed exactly as the tests decide.
RSX_HD void manifold_refresh(Manifold& m, const Xf& trA, const Xf& trB) {
    const int n = m.n;
    const float brk = m.breaking;
    V3 pa[kMaxManifoldPoints], pb[kMaxManifoldPoints], nr[kMaxManifoldPoints];
    float d[kMaxManifoldPoints];
#pragma unroll
    for (int i = 0; i < kMaxManifoldPoints; i++) {
        if (i >= n) break;
        const Contact& c = m.pt[i];
        pa[i] = xform(trA, c.localA);
        pb[i] = xform(trB, c.localB);
        nr[i] = c.normal;
        d[i] = dot(pa[i] - pb[i], nr[i]);
    }
#pragma unroll
    for (int i = 0; i < kMaxManifoldPoints; i++) {
        if (i >= n) break;
        Contact& c = m.pt[i];
        c.posA = pa[i];
        c.posB = pb[i];
        c.dist = d[i];
    }
#pragma unroll
    for (int i = kMaxManifoldPoints - 1; i >= 0; i--) {
        if (i >= m.n) continue;
        bool drop;
        if (!(d[i] <= brk)) {
            drop = true;
        } else {
            V3 proj = pa[i] - nr[i] * d[i];
            V3 diff = pb[i] - proj;
            float d2 = dot(diff, diff);
            drop = d2 > brk * brk;
        }
        if (drop) {
            // manifold_remove: the last point takes slot i (mirrored in registers)
            const int last = m.n - 1;
            manifold_remove(m, i);
            if (i != last) {
                pa[i] = pa[last];
                pb[i] = pb[last];
                nr[i] = nr[last];
                d[i] = d[last];
            }
        }
    }
}

// --- manifold list of the tick ------------------------------------------------
// Every pair the dispatch opens gets an entry, in dispatch order (empty ones
// too: they take part in the island sort); only pairs that hold contacts keep
// a manifold's storage.  kCap is the most pairs a tick can open: each body
// against every mesh and plane, and every pair of dynamic bodies.  kLive, the
// manifolds with contacts a tick can hold, is far below the theoretical bound
// (the contact-heavy 1v1 workload peaks at 8).  Once kLive manifolds hold
// contacts, a further pair gets no storage and is skipped; opening one sets
// `overflow` (conservatively: the pair may have had no contact), which the
// engine reports per arena.
template <int MAXC>
struct MfList {
    static constexpr int kCap = (MAXC + 1) * (kNumMeshes + kNumPlanes) + (MAXC + 1) * MAXC / 2;
    static constexpr int kLive = MAXC >= 2 ? 12 + 2 * (MAXC - 2) : 12;
    int count;
    int8_t body0[kCap], body1[kCap];  // 0 ball, 1+i car i, -1 static
    int8_t slot[kCap];                // index into live, -1 empty
    int nlive;
    bool overflow;                    // some pair found no storage since the flag was cleared
    Manifold live[kLive];
};

// The list's two counters while a tick's pairs are opened, kept in
// registers (the pair records' stores would otherwise hold back each
// reread); mf_finish stores them.
struct MfCursor {
    int count = 0, nlive = 0;
};

// Returns the pair's index, or -1 when the pair list is full (the pair is
// then skipped for this tick and the overflow flag set); `slot` receives the
// manifold storage it got (-1: none).
template <int MAXC>
RSX_HD int mf_open(MfList<MAXC>& L, MfCursor& C, int b0, int b1, float breaking, int& slot) {
    slot = -1;
    if (C.count >= MfList<MAXC>::kCap) {
        L.overflow = true;
        return -1;
    }
    int k = C.count++;
    L.body0[k] = (int8_t)b0;
    L.body1[k] = (int8_t)b1;
    L.slot[k] = -1;
    if (C.nlive < MfList<MAXC>::kLive) {
        int s = C.nlive;
        Manifold& m = L.live[s];
        m.body0 = b0;
        m.body1 = b1;
        m.n = 0;
        m.breaking = breaking;
        L.slot[k] = (int8_t)s;
        slot = s;
    } else {
        L.overflow = true;
    }
    return k;
}
// Keep the manifold's storage only if it ended the tick with contacts.
template <int MAXC>
RSX_HD void mf_close(MfList<MAXC>& L, MfCursor& C, int k, int s) {
    if (k < 0 || s < 0) return;
    const int n = L.live[s].n;
    if (n > 0 && s == C.nlive) C.nlive++;
    else if (n == 0) L.slot[k] = -1;
}
template <int MAXC>
RSX_HD void mf_finish(MfList<MAXC>& L, const MfCursor& C) {
    L.count = C.count;
    L.nlive = C.nlive;
}

// --- AABBs (btCollisionWorld::updateSingleAabb) --------------------------------
RSX_HD void ball_shape_aabb(const BallGeom& bg, V3 c, V3& mn, V3& mx) {
    float m = bg.radius + 0.08f;
    V3 e = V3{m, m, m};
    mn = c - e;
    mx = c + e;
}
RSX_HD V3 abs_dot3(V3 h, const M3& b) {
    M3 a = M3{{vabs(b.r[0]), vabs(b.r[1]), vabs(b.r[2])}};
    return V3{dot(h, a.r[0]), dot(h, a.r[1]), dot(h, a.r[2])};
}
RSX_HD void car_shape_aabb(const CarGeom& g, const Xf& t, V3& mn, V3& mx) {
    V3 half = g.compound_half;  // compound margin is 0
    V3 center = xform(t, g.compound_center);
    V3 ext = abs_dot3(half, t.b);
    mn = center - ext;
    mx = center + ext;
}
RSX_HD void box_child_aabb(const CarGeom& g, const Xf& child, V3& mn, V3& mx) {
    V3 h = g.box_implicit + V3{g.box_margin, g.box_margin, g.box_margin};
    V3 ext = abs_dot3(h, child.b);
    mn = child.o - ext;
    mx = child.o + ext;
}

RSX_HD void broad_aabb(V3 mn1, V3 mx1, V3 mn2, V3 mx2, V3& mn, V3& mx) {
    const float t = kContactBreakingThreshold;
    V3 e = V3{t, t, t};
    mn = vmin(mn1 - e, mn2 - e);
    mx = vmax(mx1 + e, mx2 + e);
}

// Plane shape AABB (btStaticPlaneShape::getAabb with the RocketSim slab).
RSX_HD void plane_shape_aabb(int p, V3& mn, V3& mx) {
    const float L = kLargeFloat, off = 0.2f;
    mn = V3{-L, -L, -L};
    mx = V3{L, L, L};
    V3 o = plane_origin(p);
    if (p == 0) { mn.z = o.z + (0 - off); mx.z = o.z + (0 + off); mn.z = -L; }
    if (p == 1) { mn.z = o.z + (0 - off); mx.z = o.z + (0 + off); mx.z = L; }
    if (p == 2) { mn.x = o.x + (0 - off); mx.x = o.x + (0 + off); mn.x = -L; }
    if (p == 3) { mn.x = o.x + (0 - off); mx.x = o.x + (0 + off); mx.x = L; }
}

// btAdjustInternalEdgeContacts (btInternalEdgeUtility.cpp:330-690) with
// normalAdjustFlags = 0; the mesh transform is the identity.
RSX_HD void nearest_on_segment(V3 p, V3 a, V3 b, V3& out) {
    V3 d = b - a;
    if (fuzzy_zero(d)) {
        out = a;
    } else {
        float t = dot(p - a, d) / dot(d, d);
        if (t < 0) t = 0;
        else if (t > 1) t = 1;
        out = a + d * t;
    }
}
RSX_HD float get_angle(V3 edge_a, V3 normal_a, V3 normal_b) {
    return atan2f(dot(normal_b, edge_a), dot(normal_b, normal_a));
}
RSX_HD V3 quat_rotate_v(Q4 r, V3 v) {
    Q4 q = Q4{r.w * v.x + r.y * v.z - r.z * v.y, r.w * v.y + r.z * v.x - r.x * v.z, r.w * v.z + r.x * v.y - r.y * v.x,
              -r.x * v.x - r.y * v.y - r.z * v.z};
    q = qmul(q, qinverse(r));
    return V3{q.x, q.y, q.z};
}
RSX_HD bool clamp_normal(V3 edge, V3 tri_normal, V3 n, float corrected, V3& out) {
    V3 edge_cross = normalized(cross(edge, tri_normal));
    float cur = get_angle(edge_cross, tri_normal, n);
    if (corrected < 0) {
        if (cur < corrected) {
            out = m3_from_quat(q_axis_angle(edge, corrected - cur)) * n;
            return true;
        }
    }
    if (corrected >= 0) {
        if (cur > corrected) {
            out = m3_from_quat(q_axis_angle(edge, corrected - cur)) * n;
            return true;
        }
    }
    return false;
}

// A corrected contact normal: the point on the mesh moves with it
// (btAdjustInternalEdgeContacts; the mesh transform is the identity).
RSX_HD void contact_renormal(Contact& cp, V3 n) {
    cp.normal = n;
    cp.posB = cp.posA - cp.normal * cp.dist;
    cp.localB = inv_xform(xf_identity(), cp.posB);
}

// Returns whether the contact's normal was corrected (contact_renormal ran).
RSX_HD bool adjust_internal_edge(const MeshView& mv, int tri, Contact& cp) {
    const float* info = mv.tri_info + (size_t)tri * 4;
    uint32_t flags = __float_as_uint_rsx(info[3]);
    const float angles[3] = {info[0], info[1], info[2]};
    const float kMaxEdgeAngle = 2.0f * kPi, kEdgeDist = 0.1f, kConvexEps = 0.f;
    if (!(fabsf(angles[0]) < kMaxEdgeAngle) && !(fabsf(angles[1]) < kMaxEdgeAngle) && !(fabsf(angles[2]) < kMaxEdgeAngle))
        return false;
    bool renormal = false;
    V3 v[3];
    load_tri(mv, tri, v[0], v[1], v[2]);
    V3 tri_normal = normalized(cross(v[1] - v[0], v[2] - v[0]));
    V3 contact = cp.localB;
    V3 local_n = normalized(transpose(m3_identity()) * cp.normal);
    int best = -1;
    float best_d = kLargeFloat;
    // edges: 0 = v0v1, 1 = v1v2, 2 = v2v0
    const int ea[3] = {0, 1, 2}, eb[3] = {1, 2, 0};
    for (int e = 0; e < 3; e++) {
        if (fabsf(angles[e]) < kMaxEdgeAngle) {
            V3 nr;
            nearest_on_segment(cp.localB, v[ea[e]], v[eb[e]], nr);
            float l = len(contact - nr);
            if (l < best_d) {
                best = e;
                best_d = l;
            }
        }
    }
    bool near_edge = false;
    int concave_hits = 0;
    const uint32_t convex_bit[3] = {1, 2, 4}, swap_bit[3] = {8, 16, 32};
    for (int e = 0; e < 3; e++) {
        V3 nr;
        nearest_on_segment(contact, v[ea[e]], v[eb[e]], nr);
        if (!(fabsf(angles[e]) < kMaxEdgeAngle)) continue;
        float l = len(contact - nr);
        if (!(l < kEdgeDist) || best != e) continue;
        V3 edge = v[ea[e]] - v[eb[e]];
        near_edge = true;
        if (angles[e] == 0.f) {
            concave_hits++;
            continue;
        }
        bool convex = (flags & convex_bit[e]) != 0;
        float sf = convex ? 1.f : -1.f;
        V3 nA = sf * tri_normal;
        V3 cnb = quat_rotate_v(q_axis_angle(edge, angles[e]), tri_normal);
        if (flags & swap_bit[e]) cnb = cnb * -1.f;
        V3 nB = sf * cnb;
        float da = dot(local_n, nA), db = dot(local_n, nB);
        bool back = (da < kConvexEps) && (db < kConvexEps);
        if (back) {
            concave_hits++;
        } else {
            V3 ln = (e == 0) ? local_n : (transpose(m3_identity()) * cp.normal);
            V3 clamped;
            if (clamp_normal(edge, sf * tri_normal, ln, angles[e], clamped)) {
                if (dot(clamped, 1.f * tri_normal) > 0) {
                    contact_renormal(cp, m3_identity() * clamped);
                    renormal = true;
                }
            }
        }
    }
    if (near_edge && concave_hits > 0) {
        V3 nn = tri_normal * 1.f;
        float d = dot(nn, local_n);
        if (d < 0) return renormal;
        contact_renormal(cp, m3_identity() * nn);
        renormal = true;
    }
    return renormal;
}

// --- sphere vs plane / triangle ------------------------------------------------
// btConvexPlaneCollisionAlgorithm for the ball (convex) against plane p.
RSX_HD bool sphere_plane(const BallGeom& bg, V3 center, int p, float breaking, V3& normal, V3& pt, float& dist) {
    V3 pn = plane_normal(p);
    Xf pt_xf = Xf{m3_identity(), plane_origin(p)};
    Xf ball_xf = Xf{m3_identity(), center};
    Xf plane_in_convex = xf_mul(xf_inverse(ball_xf), pt_xf);
    Xf convex_in_plane = xf_mul(xf_inverse(pt_xf), ball_xf);
    V3 dir = plane_in_convex.b * (-pn);
    // btConvexInternalShape::localGetSupportingVertex on a sphere: 0 + margin*normalize(dir)
    V3 dn = dir;
    if (len2(dn) < kEps * kEps) dn = V3{-1, -1, -1};
    dn = normalized(dn);
    V3 vtx = V3{0, 0, 0} + dn * bg.radius;
    V3 vip = xform(convex_in_plane, vtx);
    dist = dot(pn, vip) - 0.f;
    if (!(dist < breaking)) return false;
    V3 proj = vip - pn * dist;
    pt = xform(pt_xf, proj);
    normal = pt_xf.b * pn;
    return true;
}

// closestPointTriangle (SphereTriangleDetector.cpp, "faster method")
RSX_HD V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return a;
    const V3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return b;
    const V3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return c;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float v = d1 / (d1 - d3);
        return a + v * ab;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float v = d2 / (d2 - d6);
        return a + v * ac;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + v * (c - b);
    }
    const float denom = 1.f / (va + vb + vc);
    const float v = vb * denom;
    const float w = vc * denom;
    return a + v * ab + w * ac;
}

RSX_HD bool point_in_triangle(V3 p0, V3 p1, V3 p2, V3 p) {
    V3 u = p1 - p0, v = p2 - p0;
    V3 n = cross(u, v);
    float nl = dot(n, n);
    V3 w = p - p0;
    float gamma = dot(cross(u, w), n) / nl;
    float beta = dot(cross(w, v), n) / nl;
    float alpha = 1 - gamma - beta;
    return (0 <= alpha) && (alpha <= 1) && (0 <= beta) && (beta <= 1) && (0 <= gamma) && (gamma <= 1);
}

// SphereTriangleDetector::collide (sphere centre in triangle space = world)
RSX_HD bool sphere_triangle(V3 center, float radius, float threshold, V3 v0, V3 v1, V3 v2, V3& point, V3& rnormal,
                            float& depth) {
    float rwt = radius + threshold;
    V3 normal = cross(v1 - v0, v2 - v0);
    float l2 = len2(normal);
    bool has = false;
    V3 cpnt;
    if (l2 >= kEps * kEps) {
        normal = normal / sqrtf(l2);
        V3 p1c = center - v0;
        float dfp = dot(p1c, normal);
        if (dfp < 0.f) {
            dfp *= -1.f;
            normal = normal * -1.f;
        }
        if (dfp < rwt) {
            if (point_in_triangle(v0, v1, v2, center)) {
                has = true;
                cpnt = center - normal * dfp;
            } else {
                float min_d2 = rwt * rwt;
                V3 ne = closest_point_triangle(center, v0, v1, v2);
                float d2 = dist2(ne, center);
                if (d2 < min_d2) {
                    has = true;
                    cpnt = ne;
                }
            }
        }
    }
    if (has) {
        V3 c2c = center - cpnt;
        float d2 = len2(c2c);
        if (d2 < rwt * rwt) {
            if (d2 > kEps) {
                float d = sqrtf(d2);
                rnormal = normalized(c2c);
                point = cpnt;
                depth = -(radius - d);
            } else {
                rnormal = normal;
                point = cpnt;
                depth = -radius;
            }
            return true;
        }
    }
    return false;
}

// btConvexTriangleCallback early-out for a sphere (support points along
// +-triangle normal): true if the triangle can be skipped.
RSX_HD bool sphere_tri_early_out(V3 center, float radius, V3 v0, V3 v1, V3 v2, float threshold) {
    V3 tn = normalized(cross(v1 - v0, v2 - v0));
    for (int s = 0; s < 2; s++) {
        V3 dir = tn;  // ball basis is identity: inverse() * n == n
        V3 dn = dir;
        if (len2(dn) < kEps * kEps) dn = V3{-1, -1, -1};
        dn = normalized(dn);
        V3 world = center + dn * radius;
        float pd = dot(tn, world);
        float pt = dot(tn, v0);
        if (pt - pd > threshold) return true;
        tn = tn * -1.f;
    }
    return false;
}

}  // namespace rsx
