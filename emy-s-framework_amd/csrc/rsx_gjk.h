// rsx_gjk.h — car box vs arena triangle contact: the convex-vs-triangle path
// of Bullet 3.24 that RocketSim uses for car bodies against the soccar mesh.
//
// Reference behaviour restated here (bullet3-3.24/BulletCollision/):
//   btConvexTriangleCallback::processTriangle (early out)  CollisionDispatch/btConvexConcaveCollisionAlgorithm.cpp:72-180
//   btConvexConvexAlgorithm::processCollision (no polyhedral
//     features, no perturbation for box vs triangle)         CollisionDispatch/btConvexConvexAlgorithm.cpp:268-560
//   btGjkPairDetector::getClosestPointsNonVirtual           NarrowPhaseCollision/btGjkPairDetector.cpp:686-959
//   btVoronoiSimplexSolver                                   NarrowPhaseCollision/btVoronoiSimplexSolver.cpp
//   btGjkEpaPenetrationDepthSolver::calcPenDepth            NarrowPhaseCollision/btGjkEpaPenetrationDepthSolver.cpp:22-79
//   btGjkEpaSolver2 (GJK, EPA, Penetration, Distance)       NarrowPhaseCollision/btGjkEpa2.cpp
//   box / triangle support mappings                          CollisionShapes/btConvexShape.cpp:126-193, btBoxShape.h:47-56
//
// The reference's recursive routines (GJK::EncloseOrigin, EPA::expand) are
// written here as bounded loops with explicit stacks, and its pointer-linked
// face and vertex pools as index-linked arrays, so the code runs in one GPU
// lane with a fixed private-memory frame.
//
// This file is an altered restatement of zlib-licensed upstream code, kept
// with its notice as that licence requires (the alterations: one GPU lane per
// query, explicit stacks and index-linked pools instead of recursion and
// pointers, the engine's own types):
//
//   Bullet Continuous Collision Detection and Physics Library
//   Copyright (c) 2003-2008 Erwin Coumans  https://bulletphysics.org
//   GJK-EPA collision solver by Nathanael Presson, 2008
//
//   This software is provided 'as-is', without any express or implied
//   warranty.  In no event will the authors be held liable for any damages
//   arising from the use of this software.  Permission is granted to anyone
//   to use this software for any purpose, including commercial applications,
//   and to alter it and redistribute it freely, subject to the following
//   restrictions:
//   1. The origin of this software must not be misrepresented; you must not
//      claim that you wrote the original software.  If you use this software
//      in a product, an acknowledgment in the product documentation would be
//      appreciated but is not required.
//   2. Altered source versions must be plainly marked as such, and must not
//      be misrepresented as being the original software.
//   3. This notice may not be removed or altered from any source distribution.
#pragma once

#include "rsx_math.h"

namespace rsx {
namespace gjk {

constexpr float kEps = 1.1920928955078125e-07f;  // SIMD_EPSILON (FLT_EPSILON)
constexpr float kLargeFloat = 1e18f;             // BT_LARGE_FLOAT (single precision)
constexpr float kRelError2 = 1.0e-6f;            // REL_ERROR2
constexpr float kVoronoiEqualThreshold = 0.0001f;

// ---- shapes -------------------------------------------------------------------
// Shape A is the car's box (implicit half extents h, margin), shape B one
// triangle (margin 0), both in their own local frames.
struct Shapes {
    V3 h;
    float margin;
    V3 tri[3];
};

RSX_HD float fsel(float a, float b, float c) { return a >= 0 ? b : c; }  // btFsels

RSX_HD V3 box_support(const V3& h, V3 d) { return V3{fsel(d.x, h.x, -h.x), fsel(d.y, h.y, -h.y), fsel(d.z, h.z, -h.z)}; }

RSX_HD int max_axis(V3 v) { return v.x < v.y ? (v.y < v.z ? 2 : 1) : (v.x < v.z ? 2 : 0); }

RSX_HD V3 tri_support(const V3* t, V3 d) {
    V3 dots = V3{dot(d, t[0]), dot(d, t[1]), dot(d, t[2])};
    return t[max_axis(dots)];
}

// btConvexShape::localGetSupportVertexNonVirtual: normalised direction, then
// the margin along it.
RSX_HD V3 norm_dir(V3 d) {
    if (len2(d) < kEps * kEps) d = V3{-1.f, -1.f, -1.f};
    return normalized(d);
}

// ---- Voronoi simplex solver (btVoronoiSimplexSolver) -------------------------
struct SubResult {
    V3 closest;
    bool used[4];
    float bc[4];
    bool degenerate;
};

RSX_HD void sub_reset(SubResult& r) {
    r.degenerate = false;
    for (int i = 0; i < 4; i++) {
        r.bc[i] = 0.f;
        r.used[i] = false;
    }
}
RSX_HD void sub_bc(SubResult& r, float a, float b, float c, float d) {
    r.bc[0] = a;
    r.bc[1] = b;
    r.bc[2] = c;
    r.bc[3] = d;
}
RSX_HD bool sub_valid(const SubResult& r) { return r.bc[0] >= 0.f && r.bc[1] >= 0.f && r.bc[2] >= 0.f && r.bc[3] >= 0.f; }

RSX_HD void closest_pt_triangle(V3 p, V3 a, V3 b, V3 c, SubResult& r) {
    for (int i = 0; i < 4; i++) r.used[i] = false;
    V3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) {
        r.closest = a;
        r.used[0] = true;
        sub_bc(r, 1, 0, 0, 0);
        return;
    }
    V3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) {
        r.closest = b;
        r.used[1] = true;
        sub_bc(r, 0, 1, 0, 0);
        return;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        float v = d1 / (d1 - d3);
        r.closest = a + v * ab;
        r.used[0] = r.used[1] = true;
        sub_bc(r, 1 - v, v, 0, 0);
        return;
    }
    V3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) {
        r.closest = c;
        r.used[2] = true;
        sub_bc(r, 0, 0, 1, 0);
        return;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        float w = d2 / (d2 - d6);
        r.closest = a + w * ac;
        r.used[0] = r.used[2] = true;
        sub_bc(r, 1 - w, 0, w, 0);
        return;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        r.closest = b + w * (c - b);
        r.used[1] = r.used[2] = true;
        sub_bc(r, 0, 1 - w, w, 0);
        return;
    }
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    r.closest = a + ab * v + ac * w;
    r.used[0] = r.used[1] = r.used[2] = true;
    sub_bc(r, 1 - v - w, v, w, 0);
}

RSX_HD int point_outside_plane(V3 p, V3 a, V3 b, V3 c, V3 d) {
    V3 n = cross(b - a, c - a);
    float sp = dot(p - a, n), sd = dot(d - a, n);
    if (sd * sd < (1e-4f * 1e-4f)) return -1;
    return sp * sd < 0.f ? 1 : 0;
}

RSX_HD bool closest_pt_tetra(V3 p, V3 a, V3 b, V3 c, V3 d, SubResult& fin) {
    SubResult t;
    sub_reset(t);
    fin.closest = p;
    for (int i = 0; i < 4; i++) fin.used[i] = true;
    int oABC = point_outside_plane(p, a, b, c, d);
    int oACD = point_outside_plane(p, a, c, d, b);
    int oADB = point_outside_plane(p, a, d, b, c);
    int oBDC = point_outside_plane(p, b, d, c, a);
    if (oABC < 0 || oACD < 0 || oADB < 0 || oBDC < 0) {
        fin.degenerate = true;
        return false;
    }
    if (!oABC && !oACD && !oADB && !oBDC) return false;
    float best = 3.40282347e+38f;
    if (oABC) {
        closest_pt_triangle(p, a, b, c, t);
        V3 q = t.closest;
        float s = dot(q - p, q - p);
        if (s < best) {
            best = s;
            fin.closest = q;
            for (int i = 0; i < 4; i++) fin.used[i] = false;
            fin.used[0] = t.used[0];
            fin.used[1] = t.used[1];
            fin.used[2] = t.used[2];
            sub_bc(fin, t.bc[0], t.bc[1], t.bc[2], 0);
        }
    }
    if (oACD) {
        closest_pt_triangle(p, a, c, d, t);
        V3 q = t.closest;
        float s = dot(q - p, q - p);
        if (s < best) {
            best = s;
            fin.closest = q;
            for (int i = 0; i < 4; i++) fin.used[i] = false;
            fin.used[0] = t.used[0];
            fin.used[2] = t.used[1];
            fin.used[3] = t.used[2];
            sub_bc(fin, t.bc[0], 0, t.bc[1], t.bc[2]);
        }
    }
    if (oADB) {
        closest_pt_triangle(p, a, d, b, t);
        V3 q = t.closest;
        float s = dot(q - p, q - p);
        if (s < best) {
            best = s;
            fin.closest = q;
            for (int i = 0; i < 4; i++) fin.used[i] = false;
            fin.used[0] = t.used[0];
            fin.used[1] = t.used[2];
            fin.used[3] = t.used[1];
            sub_bc(fin, t.bc[0], t.bc[2], 0, t.bc[1]);
        }
    }
    if (oBDC) {
        closest_pt_triangle(p, b, d, c, t);
        V3 q = t.closest;
        float s = dot(q - p, q - p);
        if (s < best) {
            best = s;
            fin.closest = q;
            for (int i = 0; i < 4; i++) fin.used[i] = false;
            fin.used[1] = t.used[0];
            fin.used[2] = t.used[2];
            fin.used[3] = t.used[1];
            sub_bc(fin, 0, t.bc[0], t.bc[2], t.bc[1]);
        }
    }
    return true;
}

struct Simplex {
    V3 W[4], P[4], Q[4];
    int n;
    V3 lastW;
    SubResult bc;
    V3 cP1, cP2, cV;
    bool valid_closest, needs_update;
};

RSX_HD void sx_reset(Simplex& s) {
    s.valid_closest = false;
    s.n = 0;
    s.needs_update = true;
    s.lastW = V3{kLargeFloat, kLargeFloat, kLargeFloat};
    sub_reset(s.bc);
}

RSX_HD void sx_remove(Simplex& s, int i) {
    s.n--;
    s.W[i] = s.W[s.n];
    s.P[i] = s.P[s.n];
    s.Q[i] = s.Q[s.n];
}

RSX_HD void sx_reduce(Simplex& s, const bool* used) {
    if (s.n >= 4 && !used[3]) sx_remove(s, 3);
    if (s.n >= 3 && !used[2]) sx_remove(s, 2);
    if (s.n >= 2 && !used[1]) sx_remove(s, 1);
    if (s.n >= 1 && !used[0]) sx_remove(s, 0);
}

RSX_HD void sx_add(Simplex& s, V3 w, V3 p, V3 q) {
    s.lastW = w;
    s.needs_update = true;
    s.W[s.n] = w;
    s.P[s.n] = p;
    s.Q[s.n] = q;
    s.n++;
}

RSX_HD bool sx_update(Simplex& s) {
    if (!s.needs_update) return s.valid_closest;
    sub_reset(s.bc);
    s.needs_update = false;
    const V3 O = V3{0, 0, 0};
    switch (s.n) {
        case 0:
            s.valid_closest = false;
            break;
        case 1:
            s.cP1 = s.P[0];
            s.cP2 = s.Q[0];
            s.cV = s.cP1 - s.cP2;
            sub_reset(s.bc);
            sub_bc(s.bc, 1, 0, 0, 0);
            s.valid_closest = sub_valid(s.bc);
            break;
        case 2: {
            V3 from = s.W[0], to = s.W[1];
            V3 diff = O - from;
            V3 v = to - from;
            float t = dot(v, diff);
            if (t > 0) {
                float vv = dot(v, v);
                if (t < vv) {
                    t /= vv;
                    diff -= t * v;
                    s.bc.used[0] = s.bc.used[1] = true;
                } else {
                    t = 1;
                    diff -= v;
                    s.bc.used[1] = true;
                }
            } else {
                t = 0;
                s.bc.used[0] = true;
            }
            sub_bc(s.bc, 1 - t, t, 0, 0);
            s.cP1 = s.P[0] + t * (s.P[1] - s.P[0]);
            s.cP2 = s.Q[0] + t * (s.Q[1] - s.Q[0]);
            s.cV = s.cP1 - s.cP2;
            bool used[4] = {s.bc.used[0], s.bc.used[1], s.bc.used[2], s.bc.used[3]};
            sx_reduce(s, used);
            s.valid_closest = sub_valid(s.bc);
            break;
        }
        case 3: {
            closest_pt_triangle(O, s.W[0], s.W[1], s.W[2], s.bc);
            s.cP1 = s.P[0] * s.bc.bc[0] + s.P[1] * s.bc.bc[1] + s.P[2] * s.bc.bc[2];
            s.cP2 = s.Q[0] * s.bc.bc[0] + s.Q[1] * s.bc.bc[1] + s.Q[2] * s.bc.bc[2];
            s.cV = s.cP1 - s.cP2;
            bool used[4] = {s.bc.used[0], s.bc.used[1], s.bc.used[2], s.bc.used[3]};
            sx_reduce(s, used);
            s.valid_closest = sub_valid(s.bc);
            break;
        }
        case 4: {
            bool sep = closest_pt_tetra(O, s.W[0], s.W[1], s.W[2], s.W[3], s.bc);
            if (sep) {
                s.cP1 = s.P[0] * s.bc.bc[0] + s.P[1] * s.bc.bc[1] + s.P[2] * s.bc.bc[2] + s.P[3] * s.bc.bc[3];
                s.cP2 = s.Q[0] * s.bc.bc[0] + s.Q[1] * s.bc.bc[1] + s.Q[2] * s.bc.bc[2] + s.Q[3] * s.bc.bc[3];
                s.cV = s.cP1 - s.cP2;
                bool used[4] = {s.bc.used[0], s.bc.used[1], s.bc.used[2], s.bc.used[3]};
                sx_reduce(s, used);
            } else {
                if (s.bc.degenerate) {
                    s.valid_closest = false;
                } else {
                    s.valid_closest = true;
                    s.cV = V3{0, 0, 0};
                }
                break;
            }
            s.valid_closest = sub_valid(s.bc);
            break;
        }
        default:
            s.valid_closest = false;
    }
    return s.valid_closest;
}

RSX_HD bool sx_in(const Simplex& s, V3 w) {
    bool found = false;
    for (int i = 0; i < s.n; i++) {
        V3 d = s.W[i] - w;
        if (len2(d) <= kVoronoiEqualThreshold) {
            found = true;
            break;
        }
    }
    if (w.x == s.lastW.x && w.y == s.lastW.y && w.z == s.lastW.z) return true;
    return found;
}

// ---- btGjkEpaSolver2 ---------------------------------------------------------
constexpr int kGjkMaxIter = 128;
constexpr float kGjkAccuracy = 0.0001f, kGjkMinDistance = 0.0001f, kGjkDuplicatedEps = 0.0001f;
constexpr int kEpaMaxVertices = 128;
constexpr int kEpaMaxFaces = kEpaMaxVertices * 2;
constexpr int kEpaMaxIter = 255;
constexpr float kEpaAccuracy = 0.0001f, kEpaPlaneEps = 0.00001f;
constexpr int kNoFace = -1;

struct Mink {
    const Shapes* s;
    M3 to1;        // m_toshape1
    Xf to0;        // m_toshape0
    bool margins;  // support with margins (Penetration) or without (Distance)
    RSX_HD V3 support0(V3 d) const {
        if (margins) {
            V3 n = norm_dir(d);
            return box_support(s->h, n) + s->margin * n;
        }
        return box_support(s->h, d);
    }
    RSX_HD V3 support1(V3 d) const {
        V3 ld = to1 * d;
        V3 p = margins ? tri_support(s->tri, norm_dir(ld)) : tri_support(s->tri, ld);
        return xform(to0, p);
    }
    RSX_HD V3 support(V3 d) const { return support0(d) - support1(-d); }
};

// Support-vertex pool shared by GJK (entries 0-3) and EPA (4 and up).
struct SV {
    V3 d, w;
};

struct GjkSimplex {
    uint8_t c[4];
    float p[4];
    int rank;
};

enum GjkStatus { kGjkValid, kGjkInside, kGjkFailed };

struct EpaFace {
    V3 n;
    float d;
    uint8_t c[3];
    int16_t f[3];
    int16_t l[2];
    uint8_t e[3];
    uint8_t pass;
};

struct EpaList {
    int16_t root;
    int count;
};

struct GjkEpa {
    Mink sh;
    SV sv[4 + kEpaMaxVertices];
    // GJK
    V3 ray;
    float distance;
    GjkSimplex simplices[2];
    uint8_t freel[4];
    int nfree;
    int current;
    int status;
    // EPA
    EpaFace fc[kEpaMaxFaces];
    EpaList hull, stock;
    int nextsv;
    int epa_status;
    V3 epa_normal;
    float epa_depth;
    GjkSimplex epa_result;
    int16_t stk_f[kEpaMaxFaces];
    uint8_t stk_e[kEpaMaxFaces], stk_stage[kEpaMaxFaces];
};

RSX_HD void getsupport(GjkEpa& G, V3 d, int i) {
    G.sv[i].d = d * (1.f / len(d));
    G.sv[i].w = G.sh.support(G.sv[i].d);
}
RSX_HD void removevertice(GjkEpa& G, GjkSimplex& s) { G.freel[G.nfree++] = s.c[--s.rank]; }
RSX_HD void appendvertice(GjkEpa& G, GjkSimplex& s, V3 v) {
    s.p[s.rank] = 0;
    s.c[s.rank] = G.freel[--G.nfree];
    getsupport(G, v, s.c[s.rank++]);
}

RSX_HD float det3(V3 a, V3 b, V3 c) {
    return a.y * b.z * c.x + a.z * b.x * c.y - a.x * b.z * c.y - a.y * b.x * c.z + a.x * b.y * c.z - a.z * b.y * c.x;
}

RSX_HD float project2(V3 a, V3 b, float* w, unsigned& m) {
    V3 d = b - a;
    float l = len2(d);
    if (l > 0.f) {
        float t = l > 0 ? -dot(a, d) / l : 0;
        if (t >= 1) {
            w[0] = 0;
            w[1] = 1;
            m = 2;
            return len2(b);
        } else if (t <= 0) {
            w[0] = 1;
            w[1] = 0;
            m = 1;
            return len2(a);
        } else {
            w[1] = t;
            w[0] = 1 - t;
            m = 3;
            return len2(a + d * t);
        }
    }
    return -1;
}

RSX_HD float project3(V3 a, V3 b, V3 c, float* w, unsigned& m) {
    const unsigned imd3[3] = {1, 2, 0};
    const V3 vt[3] = {a, b, c};
    const V3 dl[3] = {a - b, b - c, c - a};
    const V3 n = cross(dl[0], dl[1]);
    const float l = len2(n);
    if (l > 0.f) {
        float mindist = -1;
        float subw[2] = {0.f, 0.f};
        unsigned subm = 0;
        for (unsigned i = 0; i < 3; ++i) {
            if (dot(vt[i], cross(dl[i], n)) > 0) {
                const unsigned j = imd3[i];
                const float subd = project2(vt[i], vt[j], subw, subm);
                if ((mindist < 0) || (subd < mindist)) {
                    mindist = subd;
                    m = ((subm & 1) ? 1u << i : 0u) + ((subm & 2) ? 1u << j : 0u);
                    w[i] = subw[0];
                    w[j] = subw[1];
                    w[imd3[j]] = 0;
                }
            }
        }
        if (mindist < 0) {
            const float d = dot(a, n);
            const float s = sqrtf(l);
            const V3 p = n * (d / l);
            mindist = len2(p);
            m = 7;
            w[0] = len(cross(dl[1], b - p)) / s;
            w[1] = len(cross(dl[2], c - p)) / s;
            w[2] = 1 - (w[0] + w[1]);
        }
        return mindist;
    }
    return -1;
}

RSX_HD float project4(V3 a, V3 b, V3 c, V3 d, float* w, unsigned& m) {
    const unsigned imd3[3] = {1, 2, 0};
    const V3 vt[4] = {a, b, c, d};
    const V3 dl[3] = {a - d, b - d, c - d};
    const float vl = det3(dl[0], dl[1], dl[2]);
    const bool ng = (vl * dot(a, cross(b - c, a - b))) <= 0;
    if (ng && (fabsf(vl) > 0.f)) {
        float mindist = -1;
        float subw[3] = {0.f, 0.f, 0.f};
        unsigned subm = 0;
        for (unsigned i = 0; i < 3; ++i) {
            const unsigned j = imd3[i];
            const float s = vl * dot(d, cross(dl[i], dl[j]));
            if (s > 0) {
                const float subd = project3(vt[i], vt[j], d, subw, subm);
                if ((mindist < 0) || (subd < mindist)) {
                    mindist = subd;
                    m = ((subm & 1) ? 1u << i : 0u) + ((subm & 2) ? 1u << j : 0u) + ((subm & 4) ? 8u : 0u);
                    w[i] = subw[0];
                    w[j] = subw[1];
                    w[imd3[j]] = 0;
                    w[3] = subw[2];
                }
            }
        }
        if (mindist < 0) {
            mindist = 0;
            m = 15;
            w[0] = det3(c, b, d) / vl;
            w[1] = det3(a, c, d) / vl;
            w[2] = det3(b, a, d) / vl;
            w[3] = 1 - (w[0] + w[1] + w[2]);
        }
        return mindist;
    }
    return -1;
}

RSX_HD int gjk_evaluate(GjkEpa& G, V3 guess) {
    unsigned iterations = 0;
    float sqdist = 0, alpha = 0;
    V3 lastw[4];
    unsigned clastw = 0;
    for (int i = 0; i < 4; i++) G.freel[i] = (uint8_t)i;
    G.nfree = 4;
    G.current = 0;
    G.status = kGjkValid;
    G.distance = 0;
    G.simplices[0].rank = 0;
    G.ray = guess;
    const float sqrl = len2(G.ray);
    appendvertice(G, G.simplices[0], sqrl > 0 ? -G.ray : V3{1, 0, 0});
    G.simplices[0].p[0] = 1;
    G.ray = G.sv[G.simplices[0].c[0]].w;
    sqdist = sqrl;
    lastw[0] = lastw[1] = lastw[2] = lastw[3] = G.ray;
    do {
        const int next = 1 - G.current;
        GjkSimplex& cs = G.simplices[G.current];
        GjkSimplex& ns = G.simplices[next];
        const float rl = len(G.ray);
        if (rl < kGjkMinDistance) {
            G.status = kGjkInside;
            break;
        }
        appendvertice(G, cs, -G.ray);
        const V3 w = G.sv[cs.c[cs.rank - 1]].w;
        bool found = false;
        for (int i = 0; i < 4; ++i) {
            if (len2(w - lastw[i]) < kGjkDuplicatedEps) {
                found = true;
                break;
            }
        }
        if (found) {
            removevertice(G, G.simplices[G.current]);
            break;
        } else {
            clastw = (clastw + 1) & 3;
            lastw[clastw] = w;
        }
        const float omega = dot(G.ray, w) / rl;
        alpha = omega > alpha ? omega : alpha;
        if (((rl - alpha) - (kGjkAccuracy * rl)) <= 0) {
            removevertice(G, G.simplices[G.current]);
            break;
        }
        float weights[4];
        unsigned mask = 0;
        switch (cs.rank) {
            case 2:
                sqdist = project2(G.sv[cs.c[0]].w, G.sv[cs.c[1]].w, weights, mask);
                break;
            case 3:
                sqdist = project3(G.sv[cs.c[0]].w, G.sv[cs.c[1]].w, G.sv[cs.c[2]].w, weights, mask);
                break;
            case 4:
                sqdist = project4(G.sv[cs.c[0]].w, G.sv[cs.c[1]].w, G.sv[cs.c[2]].w, G.sv[cs.c[3]].w, weights, mask);
                break;
        }
        if (sqdist >= 0) {
            ns.rank = 0;
            G.ray = V3{0, 0, 0};
            G.current = next;
            for (int i = 0, ni = cs.rank; i < ni; ++i) {
                if (mask & (1u << i)) {
                    ns.c[ns.rank] = cs.c[i];
                    ns.p[ns.rank++] = weights[i];
                    G.ray += G.sv[cs.c[i]].w * weights[i];
                } else {
                    G.freel[G.nfree++] = cs.c[i];
                }
            }
            if (mask == 15) G.status = kGjkInside;
        } else {
            removevertice(G, G.simplices[G.current]);
            break;
        }
        G.status = ((++iterations) < (unsigned)kGjkMaxIter) ? G.status : kGjkFailed;
    } while (G.status == kGjkValid);
    switch (G.status) {
        case kGjkValid:
            G.distance = len(G.ray);
            break;
        case kGjkInside:
            G.distance = 0;
            break;
        default:
            break;
    }
    return G.status;
}

// GJK::EncloseOrigin without recursion: rank 1 -> 2 -> 3 -> 4 search over
// the reference's candidate directions in its order.
RSX_HD bool enclose4(GjkEpa& G) {
    GjkSimplex& s = G.simplices[G.current];
    return fabsf(det3(G.sv[s.c[0]].w - G.sv[s.c[3]].w, G.sv[s.c[1]].w - G.sv[s.c[3]].w,
                      G.sv[s.c[2]].w - G.sv[s.c[3]].w)) > 0;
}
RSX_HD bool enclose3(GjkEpa& G) {
    GjkSimplex& s = G.simplices[G.current];
    const V3 n = cross(G.sv[s.c[1]].w - G.sv[s.c[0]].w, G.sv[s.c[2]].w - G.sv[s.c[0]].w);
    if (len2(n) > 0) {
        appendvertice(G, s, n);
        if (enclose4(G)) return true;
        removevertice(G, s);
        appendvertice(G, s, -n);
        if (enclose4(G)) return true;
        removevertice(G, s);
    }
    return false;
}
RSX_HD bool enclose2(GjkEpa& G) {
    GjkSimplex& s = G.simplices[G.current];
    const V3 d = G.sv[s.c[1]].w - G.sv[s.c[0]].w;
    for (int i = 0; i < 3; ++i) {
        V3 axis = V3{i == 0 ? 1.f : 0.f, i == 1 ? 1.f : 0.f, i == 2 ? 1.f : 0.f};
        const V3 p = cross(d, axis);
        if (len2(p) > 0) {
            appendvertice(G, s, p);
            if (enclose3(G)) return true;
            removevertice(G, s);
            appendvertice(G, s, -p);
            if (enclose3(G)) return true;
            removevertice(G, s);
        }
    }
    return false;
}
RSX_HD bool enclose1(GjkEpa& G) {
    GjkSimplex& s = G.simplices[G.current];
    for (int i = 0; i < 3; ++i) {
        V3 axis = V3{i == 0 ? 1.f : 0.f, i == 1 ? 1.f : 0.f, i == 2 ? 1.f : 0.f};
        appendvertice(G, s, axis);
        if (enclose2(G)) return true;
        removevertice(G, s);
        appendvertice(G, s, -axis);
        if (enclose2(G)) return true;
        removevertice(G, s);
    }
    return false;
}
RSX_HD bool enclose_origin(GjkEpa& G) {
    switch (G.simplices[G.current].rank) {
        case 1:
            return enclose1(G);
        case 2:
            return enclose2(G);
        case 3:
            return enclose3(G);
        case 4:
            return enclose4(G);
    }
    return false;
}

enum EpaStatus {
    kEpaValid, kEpaTouching, kEpaDegenerated, kEpaNonConvex, kEpaInvalidHull, kEpaOutOfFaces, kEpaOutOfVertices,
    kEpaAccuracyReached, kEpaFallBack, kEpaFailed
};

RSX_HD void list_append(GjkEpa& G, EpaList& L, int f) {
    G.fc[f].l[0] = kNoFace;
    G.fc[f].l[1] = L.root;
    if (L.root != kNoFace) G.fc[L.root].l[0] = (int16_t)f;
    L.root = (int16_t)f;
    ++L.count;
}
RSX_HD void list_remove(GjkEpa& G, EpaList& L, int f) {
    EpaFace& F = G.fc[f];
    if (F.l[1] != kNoFace) G.fc[F.l[1]].l[0] = F.l[0];
    if (F.l[0] != kNoFace) G.fc[F.l[0]].l[1] = F.l[1];
    if (f == L.root) L.root = F.l[1];
    --L.count;
}
RSX_HD void bind(GjkEpa& G, int fa, int ea, int fb, int eb) {
    G.fc[fa].e[ea] = (uint8_t)eb;
    G.fc[fa].f[ea] = (int16_t)fb;
    G.fc[fb].e[eb] = (uint8_t)ea;
    G.fc[fb].f[eb] = (int16_t)fa;
}

RSX_HD bool getedgedist(GjkEpa& G, int face, int a, int b, float& dist) {
    const V3 aw = G.sv[a].w, bw = G.sv[b].w;
    const V3 ba = bw - aw;
    const V3 n_ab = cross(ba, G.fc[face].n);
    const float a_dot_nab = dot(aw, n_ab);
    if (a_dot_nab < 0) {
        const float ba_l2 = len2(ba);
        const float a_dot_ba = dot(aw, ba);
        const float b_dot_ba = dot(bw, ba);
        if (a_dot_ba > 0) {
            dist = len(aw);
        } else if (b_dot_ba < 0) {
            dist = len(bw);
        } else {
            const float a_dot_b = dot(aw, bw);
            float q = (len2(aw) * len2(bw) - a_dot_b * a_dot_b) / ba_l2;
            dist = sqrtf(q > 0.f ? q : 0.f);
        }
        return true;
    }
    return false;
}

RSX_HD int newface(GjkEpa& G, int a, int b, int c, bool forced) {
    if (G.stock.root != kNoFace) {
        int face = G.stock.root;
        list_remove(G, G.stock, face);
        list_append(G, G.hull, face);
        EpaFace& F = G.fc[face];
        F.pass = 0;
        F.c[0] = (uint8_t)a;
        F.c[1] = (uint8_t)b;
        F.c[2] = (uint8_t)c;
        F.n = cross(G.sv[b].w - G.sv[a].w, G.sv[c].w - G.sv[a].w);
        const float l = len(F.n);
        const bool v = l > kEpaAccuracy;
        if (v) {
            if (!(getedgedist(G, face, a, b, F.d) || getedgedist(G, face, b, c, F.d) ||
                  getedgedist(G, face, c, a, F.d))) {
                F.d = dot(G.sv[a].w, F.n) / l;
            }
            F.n = F.n * (1.f / l);
            if (forced || (F.d >= -kEpaPlaneEps)) return face;
            G.epa_status = kEpaNonConvex;
        } else {
            G.epa_status = kEpaDegenerated;
        }
        list_remove(G, G.hull, face);
        list_append(G, G.stock, face);
        return kNoFace;
    }
    G.epa_status = G.stock.root != kNoFace ? kEpaOutOfVertices : kEpaOutOfFaces;
    return kNoFace;
}

RSX_HD int findbest(GjkEpa& G) {
    int minf = G.hull.root;
    float mind = G.fc[minf].d * G.fc[minf].d;
    for (int f = G.fc[minf].l[1]; f != kNoFace; f = G.fc[f].l[1]) {
        const float sqd = G.fc[f].d * G.fc[f].d;
        if (sqd < mind) {
            minf = f;
            mind = sqd;
        }
    }
    return minf;
}

struct Horizon {
    int cf, ff, nf;
};

// EPA::expand with an explicit stack (frames: face, edge, stage).
RSX_HD bool expand(GjkEpa& G, unsigned pass, int w, int f0, int e0, Horizon& hz) {
    const unsigned i1m3[3] = {1, 2, 0};
    const unsigned i2m3[3] = {2, 0, 1};
    int sp = 0;
    G.stk_f[0] = (int16_t)f0;
    G.stk_e[0] = (uint8_t)e0;
    G.stk_stage[0] = 0;
    sp = 1;
    bool ret = false;
    while (sp > 0) {
        const int top = sp - 1;
        const int f = G.stk_f[top];
        const unsigned e = G.stk_e[top];
        EpaFace& F = G.fc[f];
        if (G.stk_stage[top] == 0) {
            if (F.pass == pass) {
                ret = false;
                sp--;
                continue;
            }
            const unsigned e1 = i1m3[e];
            if ((dot(F.n, G.sv[w].w) - F.d) < -kEpaPlaneEps) {
                const int nf = newface(G, F.c[e1], F.c[e], w, false);
                if (nf != kNoFace) {
                    bind(G, nf, 0, f, (int)e);
                    if (hz.cf != kNoFace)
                        bind(G, hz.cf, 1, nf, 2);
                    else
                        hz.ff = nf;
                    hz.cf = nf;
                    ++hz.nf;
                    ret = true;
                } else {
                    ret = false;
                }
                sp--;
                continue;
            }
            F.pass = (uint8_t)pass;
            G.stk_stage[top] = 1;
            if (sp >= kEpaMaxFaces) return false;
            G.stk_f[sp] = F.f[e1];
            G.stk_e[sp] = F.e[e1];
            G.stk_stage[sp] = 0;
            sp++;
            continue;
        }
        if (G.stk_stage[top] == 1) {
            if (!ret) {
                sp--;
                continue;
            }
            const unsigned e2 = i2m3[e];
            G.stk_stage[top] = 2;
            if (sp >= kEpaMaxFaces) return false;
            G.stk_f[sp] = F.f[e2];
            G.stk_e[sp] = F.e[e2];
            G.stk_stage[sp] = 0;
            sp++;
            continue;
        }
        // stage 2: both neighbours expanded
        if (ret) {
            list_remove(G, G.hull, f);
            list_append(G, G.stock, f);
            ret = true;
        }
        sp--;
    }
    return ret;
}

RSX_HD int epa_evaluate(GjkEpa& G, V3 guess) {
    GjkSimplex& simplex = G.simplices[G.current];
    // EPA::Initialize
    G.epa_status = kEpaFailed;
    G.epa_normal = V3{0, 0, 0};
    G.epa_depth = 0;
    G.nextsv = 0;
    G.hull.root = kNoFace;
    G.hull.count = 0;
    G.stock.root = kNoFace;
    G.stock.count = 0;
    for (int i = 0; i < kEpaMaxFaces; ++i) list_append(G, G.stock, kEpaMaxFaces - i - 1);
    if ((simplex.rank > 1) && enclose_origin(G)) {
        while (G.hull.root != kNoFace) {
            int f = G.hull.root;
            list_remove(G, G.hull, f);
            list_append(G, G.stock, f);
        }
        G.epa_status = kEpaValid;
        G.nextsv = 0;
        if (det3(G.sv[simplex.c[0]].w - G.sv[simplex.c[3]].w, G.sv[simplex.c[1]].w - G.sv[simplex.c[3]].w,
                 G.sv[simplex.c[2]].w - G.sv[simplex.c[3]].w) < 0) {
            uint8_t tc = simplex.c[0];
            simplex.c[0] = simplex.c[1];
            simplex.c[1] = tc;
            float tp = simplex.p[0];
            simplex.p[0] = simplex.p[1];
            simplex.p[1] = tp;
        }
        const int c0 = simplex.c[0], c1 = simplex.c[1], c2 = simplex.c[2], c3 = simplex.c[3];
        const int tetra[4] = {newface(G, c0, c1, c2, true), newface(G, c1, c0, c3, true), newface(G, c2, c1, c3, true),
                              newface(G, c0, c2, c3, true)};
        if (G.hull.count == 4) {
            int best = findbest(G);
            EpaFace outer = G.fc[best];
            unsigned pass = 0;
            bind(G, tetra[0], 0, tetra[1], 0);
            bind(G, tetra[0], 1, tetra[2], 0);
            bind(G, tetra[0], 2, tetra[3], 0);
            bind(G, tetra[1], 1, tetra[3], 2);
            bind(G, tetra[1], 2, tetra[2], 1);
            bind(G, tetra[2], 2, tetra[3], 1);
            G.epa_status = kEpaValid;
            for (int iterations = 0; iterations < kEpaMaxIter; ++iterations) {
                if (G.nextsv < kEpaMaxVertices) {
                    Horizon hz{kNoFace, kNoFace, 0};
                    const int w = 4 + G.nextsv++;
                    bool valid = true;
                    G.fc[best].pass = (uint8_t)(++pass);
                    getsupport(G, G.fc[best].n, w);
                    const float wdist = dot(G.fc[best].n, G.sv[w].w) - G.fc[best].d;
                    if (wdist > kEpaAccuracy) {
                        for (int j = 0; (j < 3) && valid; ++j)
                            valid &= expand(G, pass, w, G.fc[best].f[j], G.fc[best].e[j], hz);
                        if (valid && (hz.nf >= 3)) {
                            bind(G, hz.cf, 1, hz.ff, 2);
                            list_remove(G, G.hull, best);
                            list_append(G, G.stock, best);
                            best = findbest(G);
                            outer = G.fc[best];
                        } else {
                            G.epa_status = kEpaInvalidHull;
                            break;
                        }
                    } else {
                        G.epa_status = kEpaAccuracyReached;
                        break;
                    }
                } else {
                    G.epa_status = kEpaOutOfVertices;
                    break;
                }
            }
            const V3 projection = outer.n * outer.d;
            G.epa_normal = outer.n;
            G.epa_depth = outer.d;
            G.epa_result.rank = 3;
            G.epa_result.c[0] = outer.c[0];
            G.epa_result.c[1] = outer.c[1];
            G.epa_result.c[2] = outer.c[2];
            G.epa_result.p[0] = len(cross(G.sv[outer.c[1]].w - projection, G.sv[outer.c[2]].w - projection));
            G.epa_result.p[1] = len(cross(G.sv[outer.c[2]].w - projection, G.sv[outer.c[0]].w - projection));
            G.epa_result.p[2] = len(cross(G.sv[outer.c[0]].w - projection, G.sv[outer.c[1]].w - projection));
            const float sum = G.epa_result.p[0] + G.epa_result.p[1] + G.epa_result.p[2];
            G.epa_result.p[0] /= sum;
            G.epa_result.p[1] /= sum;
            G.epa_result.p[2] /= sum;
            return G.epa_status;
        }
    }
    G.epa_status = kEpaFallBack;
    G.epa_normal = -guess;
    const float nl = len(G.epa_normal);
    if (nl > 0)
        G.epa_normal = G.epa_normal * (1.f / nl);
    else
        G.epa_normal = V3{1, 0, 0};
    G.epa_depth = 0;
    G.epa_result.rank = 1;
    G.epa_result.c[0] = simplex.c[0];
    G.epa_result.p[0] = 1;
    return G.epa_status;
}

// MinkowskiDiff set-up (gjkepa2_impl::Initialize).
RSX_HD void mink_init(GjkEpa& G, const Shapes& s, const Xf& w0, const Xf& w1, bool margins) {
    G.sh.s = &s;
    G.sh.to1 = transpose_times(w1.b, w0.b);
    G.sh.to0 = xf_inverse_times(w0, w1);
    G.sh.margins = margins;
}

struct Witness {
    V3 a, b, normal;
    float distance;
};

// btGjkEpaSolver2::Penetration (with margins)
RSX_HD bool penetration(GjkEpa& G, const Shapes& s, const Xf& w0, const Xf& w1, V3 guess, Witness& r) {
    mink_init(G, s, w0, w1, true);
    int st = gjk_evaluate(G, -guess);
    if (st == kGjkInside) {
        int es = epa_evaluate(G, -guess);
        if (es != kEpaFailed) {
            V3 w0v = V3{0, 0, 0};
            for (int i = 0; i < G.epa_result.rank; ++i)
                w0v += G.sh.support0(G.sv[G.epa_result.c[i]].d) * G.epa_result.p[i];
            r.a = xform(w0, w0v);
            r.b = xform(w0, w0v - G.epa_normal * G.epa_depth);
            r.normal = -G.epa_normal;
            r.distance = -G.epa_depth;
            return true;
        }
    }
    return false;
}

// btGjkEpaSolver2::Distance (without margins)
RSX_HD bool distance(GjkEpa& G, const Shapes& s, const Xf& w0, const Xf& w1, V3 guess, Witness& r) {
    mink_init(G, s, w0, w1, false);
    int st = gjk_evaluate(G, guess);
    if (st == kGjkValid) {
        V3 a = V3{0, 0, 0}, b = V3{0, 0, 0};
        const GjkSimplex& sx = G.simplices[G.current];
        for (int i = 0; i < sx.rank; ++i) {
            const float p = sx.p[i];
            a += G.sh.support0(G.sv[sx.c[i]].d) * p;
            b += G.sh.support1(-G.sv[sx.c[i]].d) * p;
        }
        r.a = xform(w0, a);
        r.b = xform(w0, b);
        r.normal = a - b;
        r.distance = len(r.normal);
        r.normal = r.normal * (1.f / (r.distance > kGjkMinDistance ? r.distance : 1.f));
        return true;
    }
    return false;
}

// btGjkEpaPenetrationDepthSolver::calcPenDepth
RSX_HD bool calc_pen_depth(GjkEpa& G, Simplex& sxs, const Shapes& s, const Xf& ta, const Xf& tb, V3& v, V3& wa, V3& wb) {
    V3 guesses[9] = {safe_normalized(tb.o - ta.o), safe_normalized(ta.o - tb.o), V3{0, 0, 1}, V3{0, 1, 0},
                     V3{1, 0, 0}, V3{1, 1, 0}, V3{1, 1, 1}, V3{0, 1, 1}, V3{1, 0, 1}};
    for (int i = 0; i < 9; i++) {
        sx_reset(sxs);
        Witness r;
        if (penetration(G, s, ta, tb, guesses[i], r)) {
            wa = r.a;
            wb = r.b;
            v = r.normal;
            return true;
        }
        if (distance(G, s, ta, tb, guesses[i], r)) {
            wa = r.a;
            wb = r.b;
            v = r.normal;
            return false;
        }
    }
    wa = wb = v = V3{0, 0, 0};
    return false;
}

struct Contact1 {
    V3 normal_on_b, point_on_b;
    float depth;
};

// btGjkPairDetector::getClosestPointsNonVirtual for box (A) vs triangle (B).
// Returns true with one contact when the reference would report one.
// sx is the caller's Voronoi simplex working set (reset here).
RSX_HD bool closest_points(GjkEpa& G, Simplex& sx, const Shapes& s, const Xf& trA, const Xf& trB, float max_dist2,
                           Contact1& out) {
    float dist = 0.f;
    V3 normalInB = V3{0, 0, 0};
    V3 pointOnA = V3{0, 0, 0}, pointOnB = V3{0, 0, 0};
    Xf lA = trA, lB = trB;
    const V3 off = (trA.o + trB.o) * 0.5f;
    lA.o -= off;
    lB.o -= off;
    const float marginA = s.margin, marginB = 0.f;
    int cur_iter = 0;
    const int max_iter = 1000;
    V3 sep = V3{0, 1, 0};
    bool isValid = false, checkSimplex = false;
    int degenerate = 0;
    float squaredDistance = kLargeFloat;
    float delta = 0.f;
    const float margin = marginA + marginB;
    sx_reset(sx);
    while (true) {
        V3 sepA = (-sep) * trA.b;  // (-v) * basisA
        V3 sepB = sep * trB.b;
        V3 pInA = box_support(s.h, sepA);
        V3 qInB = tri_support(s.tri, sepB);
        V3 pW = xform(lA, pInA);
        V3 qW = xform(lB, qInB);
        V3 w = pW - qW;
        delta = dot(sep, w);
        if ((delta > 0.f) && (delta * delta > squaredDistance * max_dist2)) {
            degenerate = 10;
            checkSimplex = true;
            break;
        }
        if (sx_in(sx, w)) {
            degenerate = 1;
            checkSimplex = true;
            break;
        }
        float f0 = squaredDistance - delta;
        float f1 = squaredDistance * kRelError2;
        if (f0 <= f1) {
            degenerate = (f0 <= 0.f) ? 2 : 11;
            checkSimplex = true;
            break;
        }
        sx_add(sx, w, pW, qW);
        V3 nsep;
        bool ok = sx_update(sx);
        nsep = sx.cV;
        if (!ok) {
            degenerate = 3;
            checkSimplex = true;
            break;
        }
        if (len2(nsep) < kRelError2) {
            sep = nsep;
            degenerate = 6;
            checkSimplex = true;
            break;
        }
        float prev = squaredDistance;
        squaredDistance = len2(nsep);
        if (prev - squaredDistance <= kEps * prev) {
            checkSimplex = true;
            degenerate = 12;
            break;
        }
        sep = nsep;
        if (cur_iter++ > max_iter) break;
        if (sx.n == 4) {
            degenerate = 13;
            break;
        }
    }
    if (checkSimplex) {
        sx_update(sx);
        pointOnA = sx.cP1;
        pointOnB = sx.cP2;
        normalInB = sep;
        float lenSqr = len2(sep);
        if (lenSqr < 0.0001f) degenerate = 5;
        if (lenSqr > kEps * kEps) {
            float rlen = 1.f / sqrtf(lenSqr);
            normalInB = normalInB * rlen;
            float sq = sqrtf(squaredDistance);
            pointOnA -= sep * (marginA / sq);
            pointOnB += sep * (marginB / sq);
            dist = (1.f / rlen) - margin;
            isValid = true;
        }
    }
    const bool catchDegenerate = degenerate && ((dist + margin) < 0.01f);
    if (!isValid || catchDegenerate) {
        V3 tA, tB;
        sep = V3{0, 0, 0};
        bool valid2 = calc_pen_depth(G, sx, s, lA, lB, sep, tA, tB);
        if (valid2) {
            V3 tn = tB - tA;
            float lenSqr = len2(tn);
            if (lenSqr <= kEps * kEps) {
                tn = sep;
                lenSqr = len2(sep);
            }
            if (lenSqr > kEps * kEps) {
                tn = tn * (1.f / sqrtf(lenSqr));
                float d2 = -len(tA - tB);
                if (!isValid || (d2 < dist)) {
                    dist = d2;
                    pointOnA = tA;
                    pointOnB = tB;
                    normalInB = tn;
                    isValid = true;
                }
            }
        } else {
            if (len2(sep) > 0.f) {
                float d2 = len(tA - tB) - margin;
                if (!isValid || (d2 < dist)) {
                    dist = d2;
                    pointOnA = tA;
                    pointOnB = tB;
                    pointOnA -= sep * marginA;
                    pointOnB += sep * marginB;
                    normalInB = normalized(sep);
                    isValid = true;
                }
            }
        }
    }
    if (isValid && ((dist < 0) || (dist * dist < max_dist2))) {
        // fix the normal direction against the AABB centres (box: transformed
        // half extents + margin; triangle: support extremes along the axes)
        V3 posA;
        {
            V3 he = s.h + V3{s.margin, s.margin, s.margin};
            V3 ext = V3{dot(he, vabs(lA.b.r[0])), dot(he, vabs(lA.b.r[1])), dot(he, vabs(lA.b.r[2]))};
            V3 mn = lA.o - ext, mx = lA.o + ext;
            posA = (mx + mn) * 0.5f;
        }
        V3 posB;
        {
            float mn[3], mx[3];
            for (int i = 0; i < 3; i++) {
                V3 e = V3{i == 0 ? 1.f : 0.f, i == 1 ? 1.f : 0.f, i == 2 ? 1.f : 0.f};
                V3 sv = tri_support(s.tri, e * lB.b);
                mx[i] = idx(xform(lB, sv), i);
                sv = tri_support(s.tri, (-e) * lB.b);
                mn[i] = idx(xform(lB, sv), i);
            }
            posB = (V3{mn[0], mn[1], mn[2]} + V3{mx[0], mx[1], mx[2]}) * 0.5f;
        }
        V3 diff = posA - posB;
        if (dot(diff, normalInB) < 0.f) normalInB = normalInB * -1.f;
        out.normal_on_b = normalInB;
        out.point_on_b = pointOnB + off;
        out.depth = dist;
        return true;
    }
    return false;
}

// ---- ray vs convex (btSubsimplexConvexCast) ----------------------------------
// btCollisionWorld::rayTestSingleInternal casts a zero-radius sphere from
// `from` to `to` against a convex body with btSubsimplexConvexCast
// (NarrowPhaseCollision/btSubSimplexConvexCast.cpp:38-140; single precision:
// 32 iterations, epsilon 1e-4).  Shape: a box (kind 0, support of the box
// grown by its margin, btBoxShape::localGetSupportingVertex) or a sphere
// (kind 1, radius along the normalised direction).
struct RayShape {
    int kind;
    V3 h;          // box: implicit half extents + margin; sphere: radius in h.x
    Xf xf;         // world transform of the shape
};

RSX_HD V3 ray_shape_support(const RayShape& s, V3 d) {
    if (s.kind == 0) return box_support(s.h, d);
    V3 n = d;
    if (len2(n) < kEps * kEps) n = V3{-1.f, -1.f, -1.f};
    n = normalized(n);
    return V3{0, 0, 0} + s.h.x * n;
}

// Returns true with the hit fraction and the (unnormalised) normal; sx is
// the caller's Voronoi simplex working set.
RSX_HD bool ray_cast(const RayShape& s, V3 from, V3 to, float& fraction, V3& normal, Simplex& sx) {
    sx_reset(sx);
    const V3 r = to - from;  // linVelA - linVelB, the shape does not move
    float lambda = 0.f;
    V3 ia = from;
    V3 supA = ia;
    V3 supB = xform(s.xf, ray_shape_support(s, r * s.xf.b));
    V3 v = supA - supB;
    int max_iter = 32;
    V3 n = V3{0, 0, 0};
    float dist2 = len2(v);
    while ((dist2 > 0.0001f) && max_iter--) {
        supA = ia;
        supB = xform(s.xf, ray_shape_support(s, v * s.xf.b));
        V3 w = supA - supB;
        float vdotw = dot(v, w);
        if (lambda > 1.0f) return false;
        if (vdotw > 0.f) {
            float vdotr = dot(v, r);
            if (vdotr >= -(kEps * kEps)) return false;
            lambda = lambda - vdotw / vdotr;
            float t = 1.0f - lambda;
            ia = V3{t * from.x + lambda * to.x, t * from.y + lambda * to.y, t * from.z + lambda * to.z};
            w = supA - supB;
            n = v;
        }
        if (!sx_in(sx, w)) sx_add(sx, w, supA, supB);
        if (sx_update(sx))
            dist2 = len2(sx.cV), v = sx.cV;
        else
            dist2 = 0.f, v = sx.cV;
    }
    fraction = lambda;
    V3 nn = (len2(n) >= kEps * kEps) ? normalized(n) : V3{0, 0, 0};
    if (dot(nn, r) >= -0.f) return false;
    normal = nn;
    return true;
}

}  // namespace gjk
}  // namespace rsx
