// rsx_query.h — geometric queries against the soccar arena: the static
// planes, the triangle mesh (uniform-grid index, visited in the order of the
// reference's per-mesh tree so that contact insertion order matches), and the
// suspension collision grid's fast path.
//
// Reference behaviour restated here:
//   SuspensionCollisionGrid::CastSuspensionRay   SuspensionCollisionGrid.cpp:124-175
//   btTriangleRaycastCallback::processTriangle   btRaycastCallback.cpp:35-128
//   btCollisionWorld::ClosestRayResultCallback   btCollisionWorld.h:211-250
//   btDefaultVehicleRaycaster::castRay           btDefaultVehicleRaycaster.cpp:32-53
//   TestTriangleAgainstAabb2 / TestAabbAgainstAabb2  btAabbUtil2.h:43-70
#pragma once

#include "rsx_world.h"

namespace rsx {

// Static planes of the soccar arena (Arena.cpp:1060-1101), Bullet units.
RSX_HD V3 plane_normal(int p) {
    return p == 0 ? V3{0, 0, 1} : (p == 1 ? V3{0, 0, -1} : (p == 2 ? V3{1, 0, 0} : V3{-1, 0, 0}));
}
RSX_HD V3 plane_origin(int p) {
    return p == 0 ? V3{0, 0, 0}
                  : (p == 1 ? V3{0, 0, kArenaHeight * kUU2BT}
                            : (p == 2 ? V3{-kArenaExtentX * kUU2BT, 0, kArenaHeight / 2 * kUU2BT}
                                      : V3{kArenaExtentX * kUU2BT, 0, kArenaHeight / 2 * kUU2BT}));
}

RSX_HD bool aabb_overlap(V3 mn1, V3 mx1, V3 mn2, V3 mx2) {
    bool o = true;
    o = (mn1.x > mx2.x || mx1.x < mn2.x) ? false : o;
    o = (mn1.z > mx2.z || mx1.z < mn2.z) ? false : o;
    o = (mn1.y > mx2.y || mx1.y < mn2.y) ? false : o;
    return o;
}

RSX_HD bool tri_aabb_overlap(V3 a, V3 b, V3 c, V3 mn, V3 mx) {
    // z, then x, then y (btAabbUtil2.h:54-70)
    if (minf(minf(a.z, b.z), c.z) > mx.z) return false;
    if (maxf(maxf(a.z, b.z), c.z) < mn.z) return false;
    if (minf(minf(a.x, b.x), c.x) > mx.x) return false;
    if (maxf(maxf(a.x, b.x), c.x) < mn.x) return false;
    if (minf(minf(a.y, b.y), c.y) > mx.y) return false;
    if (maxf(maxf(a.y, b.y), c.y) < mn.y) return false;
    return true;
}

RSX_HD void load_tri(const MeshView& w, int t, V3& a, V3& b, V3& c) {
    const float* p = w.tri + (size_t)t * 9;
    a = V3{p[0], p[1], p[2]};
    b = V3{p[3], p[4], p[5]};
    c = V3{p[6], p[7], p[8]};
}

RSX_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
RSX_HD int maxi(int a, int b) { return a > b ? a : b; }

// A gathered candidate packs the triangle's mesh (bits 26-30), its rank in
// that mesh's tree traversal (bits 13-25) and its index (bits 0-12), so that
// sorting the packed values orders candidates by (mesh, rank).
constexpr int kCandTriBits = 13, kCandRankShift = 13, kCandMeshShift = 26;
constexpr int kMaxTris = 1 << kCandTriBits;
RSX_HD int cand_tri(int p) { return p & (kMaxTris - 1); }
RSX_HD int cand_mesh(int p) { return p >> kCandMeshShift; }

// Candidate triangles overlapping [mn, mx], each once, sorted by (mesh index,
// tree traversal rank); see cand_tri / cand_mesh.  Returns the count (capped
// at cap).  A triangle is listed in every grid cell its AABB touches, so it is
// taken only from the first of those cells that the query visits.
// Capacity of a candidate list.  Every query box is smaller than a grid cell
// (4 BT) along each axis -- the ball's AABB 3.8 BT, a car box's at most
// 3.1 BT, a wheel ray's well under 1 BT -- so it meets at most 2 x 2 x 2
// cells, and build_world rejects meshes where such a block lists more
// triangles than this (the soccar meshes: 220).
constexpr int kMaxCand = 224;

RSX_HD int gather_tris(const MeshView& w, V3 mn, V3 mx, int* out, int cap) {
    const float inv = 1.0f / w.grid_cell;
    int x0 = clampi((int)floorf((mn.x - w.grid_min.x) * inv), 0, w.gx - 1);
    int y0 = clampi((int)floorf((mn.y - w.grid_min.y) * inv), 0, w.gy - 1);
    int z0 = clampi((int)floorf((mn.z - w.grid_min.z) * inv), 0, w.gz - 1);
    int x1 = clampi((int)floorf((mx.x - w.grid_min.x) * inv), 0, w.gx - 1);
    int y1 = clampi((int)floorf((mx.y - w.grid_min.y) * inv), 0, w.gy - 1);
    int z1 = clampi((int)floorf((mx.z - w.grid_min.z) * inv), 0, w.gz - 1);
    int n = 0;
    if (x1 - x0 <= 1 && y1 - y0 <= 1 && z1 - z0 <= 1) {
        // the box's cells lie in the block at its lowest cell: a triangle
        // whose AABB meets the box is listed in one of them (same cell
        // arithmetic), the block lists it once, in key order
        const int blk = (x0 * w.gy + y0) * w.gz + z0;
        const int s = w.blk_start[blk], e = w.blk_start[blk + 1];
        // four entries per round, their loads issued together
        for (int k = s; k < e; k += 4) {
            GridEntry g[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (k + q < e) g[q] = w.blk_ent[k + q];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (k + q >= e) break;
                if (g[q].mn[2] > mx.z || g[q].mx[2] < mn.z) continue;
                if (g[q].mn[0] > mx.x || g[q].mx[0] < mn.x) continue;
                if (g[q].mn[1] > mx.y || g[q].mx[1] < mn.y) continue;
                if (n < cap) out[n++] = g[q].key;
            }
        }
        return n;
    }
    for (int ix = x0; ix <= x1; ix++)
        for (int iy = y0; iy <= y1; iy++)
            for (int iz = z0; iz <= z1; iz++) {
                int cell = (ix * w.gy + iy) * w.gz + iz;
                int s = w.cell_start[cell], e = w.cell_start[cell + 1];
                for (int k = s; k < e; k++) {
                    const GridEntry g = w.cell_ent[k];
                    // first cell of the triangle's own range inside the query range
                    const int fx = maxi(x0, g.first & 1023);
                    const int fy = maxi(y0, (g.first >> 10) & 1023);
                    const int fz = maxi(z0, g.first >> 20);
                    if (fx != ix || fy != iy || fz != iz) continue;
                    // tri_aabb_overlap on the stored AABB: z, then x, then y
                    if (g.mn[2] > mx.z || g.mx[2] < mn.z) continue;
                    if (g.mn[0] > mx.x || g.mx[0] < mn.x) continue;
                    if (g.mn[1] > mx.y || g.mx[1] < mn.y) continue;
                    if (n < cap) out[n++] = g.key;
                }
            }
    // insertion sort of the packed keys
    for (int i = 1; i < n; i++) {
        int key = out[i];
        int j = i - 1;
        while (j >= 0 && out[j] > key) {
            out[j + 1] = out[j];
            j--;
        }
        out[j + 1] = key;
    }
    return n;
}

struct RayHit {
    V3 point, normal;
    float fraction;
    int object;     // -1 none, 0 static world, 1 + b dynamic body b (0 ball, 1+i car i)
};

// btTriangleRaycastCallback::processTriangle on one triangle; updates hit if closer.
RSX_HD bool ray_triangle(V3 from, V3 to, V3 v0, V3 v1, V3 v2, float& hit_fraction, V3& hit_normal) {
    V3 v10 = v1 - v0;
    V3 v20 = v2 - v0;
    V3 tn = cross(v10, v20);
    const float d = dot(v0, tn);
    float da = dot(tn, from);
    da -= d;
    float db = dot(tn, to);
    db -= d;
    if (da * db >= 0.0f) return false;
    const float proj = da - db;
    const float distance = da / proj;
    if (!(distance < hit_fraction)) return false;
    float tol = len2(tn);
    tol *= -0.0001f;
    float s = 1.0f - distance;
    V3 point = V3{s * from.x + distance * to.x, s * from.y + distance * to.y, s * from.z + distance * to.z};
    V3 v0p = v0 - point;
    V3 v1p = v1 - point;
    V3 cp0 = cross(v0p, v1p);
    if (!(dot(cp0, tn) >= tol)) return false;
    V3 v2p = v2 - point;
    V3 cp1 = cross(v1p, v2p);
    if (!(dot(cp1, tn) >= tol)) return false;
    V3 cp2 = cross(v2p, v0p);
    if (!(dot(cp2, tn) >= tol)) return false;
    tn = normalized(tn);
    hit_normal = (da <= 0.0f) ? -tn : tn;
    hit_fraction = distance;
    return true;
}

// The ray test of btTriangleRaycastCallback::processTriangle without the
// closest-hit comparison: the crossing fraction and hit normal of the
// triangle, or false when the ray misses it.
RSX_HD bool ray_triangle_hit(V3 from, V3 to, V3 v0, V3 v1, V3 v2, float& distance, V3& hit_normal) {
    V3 v10 = v1 - v0;
    V3 v20 = v2 - v0;
    V3 tn = cross(v10, v20);
    const float d = dot(v0, tn);
    float da = dot(tn, from);
    da -= d;
    float db = dot(tn, to);
    db -= d;
    if (da * db >= 0.0f) return false;
    const float proj = da - db;
    distance = da / proj;
    float tol = len2(tn);
    tol *= -0.0001f;
    float s = 1.0f - distance;
    V3 point = V3{s * from.x + distance * to.x, s * from.y + distance * to.y, s * from.z + distance * to.z};
    V3 v0p = v0 - point;
    V3 v1p = v1 - point;
    V3 cp0 = cross(v0p, v1p);
    if (!(dot(cp0, tn) >= tol)) return false;
    V3 v2p = v2 - point;
    V3 cp1 = cross(v1p, v2p);
    if (!(dot(cp1, tn) >= tol)) return false;
    V3 cp2 = cross(v2p, v0p);
    if (!(dot(cp2, tn) >= tol)) return false;
    tn = normalized(tn);
    hit_normal = (da <= 0.0f) ? -tn : tn;
    return true;
}

// The mesh triangles near a ray, tested as the grid walk meets them.  Bullet
// visits them in tree order and keeps a hit only when it is strictly closer,
// so the hit it reports is the one with the smallest fraction, the first in
// tree order among equals; keeping the lexicographically smallest
// (fraction, packed key) as the walk finds them gives the same hit without
// listing and sorting the candidates.  Triangles listed in several cells are
// tested once, from the first such cell the walk visits.
RSX_HD void ray_mesh(const MeshView& w, V3 from, V3 to, RayHit& hit) {
    const V3 mn = vmin(from, to), mx = vmax(from, to);
    const float inv = 1.0f / w.grid_cell;
    const int x0 = clampi((int)floorf((mn.x - w.grid_min.x) * inv), 0, w.gx - 1);
    const int y0 = clampi((int)floorf((mn.y - w.grid_min.y) * inv), 0, w.gy - 1);
    const int z0 = clampi((int)floorf((mn.z - w.grid_min.z) * inv), 0, w.gz - 1);
    const int x1 = clampi((int)floorf((mx.x - w.grid_min.x) * inv), 0, w.gx - 1);
    const int y1 = clampi((int)floorf((mx.y - w.grid_min.y) * inv), 0, w.gy - 1);
    const int z1 = clampi((int)floorf((mx.z - w.grid_min.z) * inv), 0, w.gz - 1);
    float frac = hit.object >= 0 ? hit.fraction : 1.0f;
    int best = -1;  // packed key of the hit kept so far (-1: none from the mesh)
    for (int ix = x0; ix <= x1; ix++)
        for (int iy = y0; iy <= y1; iy++)
            for (int iz = z0; iz <= z1; iz++) {
                const int cell = (ix * w.gy + iy) * w.gz + iz;
                const int s = w.cell_start[cell], e = w.cell_start[cell + 1];
                for (int k = s; k < e; k++) {
                    const GridEntry g = w.cell_ent[k];
                    const int fx = maxi(x0, g.first & 1023);
                    const int fy = maxi(y0, (g.first >> 10) & 1023);
                    const int fz = maxi(z0, g.first >> 20);
                    if (fx != ix || fy != iy || fz != iz) continue;
                    if (g.mn[2] > mx.z || g.mx[2] < mn.z) continue;
                    if (g.mn[0] > mx.x || g.mx[0] < mn.x) continue;
                    if (g.mn[1] > mx.y || g.mx[1] < mn.y) continue;
                    V3 a, b, c, nrm;
                    load_tri(w, cand_tri(g.key), a, b, c);
                    float d;
                    if (!ray_triangle_hit(from, to, a, b, c, d, nrm)) continue;
                    if (d < frac || (d == frac && best >= 0 && g.key < best)) {
                        frac = d;
                        best = g.key;
                        hit.fraction = d;
                        hit.normal = nrm;
                        hit.object = 0;
                    }
                }
            }
}

// The mesh part of four rays at once (a car's suspension rays): their
// segments' union spans at most two grid cells per axis, so the 2 x 2 x 2
// block at its lowest cell lists every triangle any of them can meet, once.
// One scan of that list tests each entry against every ray whose box it
// meets, keeping per ray the smallest (fraction, packed key) -- the same set
// of triangles and the same order as ray_mesh for each ray alone, so the
// same hits; a car over an empty block costs one lookup instead of four grid
// walks.  hit[k] holds what ray k hit before (or object -1).
// The 2 x 2 x 2 block (the grid cell at its lowest corner) listing every
// triangle whose AABB can meet the box [mn, mx] (gather_tris's cell
// arithmetic); false when the box spans more than two grid cells along some
// axis.
RSX_HD bool box_block(const MeshView& w, V3 umn, V3 umx, int& blk) {
    const float inv = 1.0f / w.grid_cell;
    const int x0 = clampi((int)floorf((umn.x - w.grid_min.x) * inv), 0, w.gx - 1);
    const int y0 = clampi((int)floorf((umn.y - w.grid_min.y) * inv), 0, w.gy - 1);
    const int z0 = clampi((int)floorf((umn.z - w.grid_min.z) * inv), 0, w.gz - 1);
    const int x1 = clampi((int)floorf((umx.x - w.grid_min.x) * inv), 0, w.gx - 1);
    const int y1 = clampi((int)floorf((umx.y - w.grid_min.y) * inv), 0, w.gy - 1);
    const int z1 = clampi((int)floorf((umx.z - w.grid_min.z) * inv), 0, w.gz - 1);
    blk = (x0 * w.gy + y0) * w.gz + z0;
    return x1 - x0 <= 1 && y1 - y0 <= 1 && z1 - z0 <= 1;
}

// The block for four ray segments (boxes mn[k], mx[k]): that of their union.
RSX_HD bool ray4_block(const MeshView& w, const V3* mn, const V3* mx, int& blk) {
    const V3 umn = vmin(vmin(mn[0], mn[1]), vmin(mn[2], mn[3]));
    const V3 umx = vmax(vmax(mx[0], mx[1]), vmax(mx[2], mx[3]));
    return box_block(w, umn, umx, blk);
}

// Whether a grid entry's AABB meets a segment's AABB (mn, mx).
RSX_HD bool entry_meets_segment(const GridEntry& g, V3 mn, V3 mx) {
    return !(g.mn[2] > mx.z || g.mx[2] < mn.z) && !(g.mn[0] > mx.x || g.mx[0] < mn.x) &&
           !(g.mn[1] > mx.y || g.mx[1] < mn.y);
}

RSX_HD void ray_mesh4(const MeshView& w, const V3* from, const V3* to, RayHit* hit) {
    V3 mn[4], mx[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        mn[k] = vmin(from[k], to[k]);
        mx[k] = vmax(from[k], to[k]);
    }
    int blk;
    if (!ray4_block(w, mn, mx, blk)) {
        for (int k = 0; k < 4; k++) ray_mesh(w, from[k], to[k], hit[k]);
        return;
    }
    float frac[4];
    int best[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        frac[k] = hit[k].object >= 0 ? hit[k].fraction : 1.0f;
        best[k] = -1;
    }
    const int s = w.blk_start[blk], e = w.blk_start[blk + 1];
    for (int j0 = s; j0 < e; j0 += 2) {
        // two entries per round, their loads issued together
        GridEntry gg[2];
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (j0 + q < e) gg[q] = w.blk_ent[j0 + q];
#pragma unroll
        for (int q = 0; q < 2; q++) {
        if (j0 + q >= e) break;
        const GridEntry& g = gg[q];
        bool ov[4];
        bool any = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            ov[k] = entry_meets_segment(g, mn[k], mx[k]);
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
            if (!ray_triangle_hit(from[k], to[k], a, b, c, d, nrm)) continue;
            if (d < frac[k] || (d == frac[k] && best[k] >= 0 && g.key < best[k])) {
                frac[k] = d;
                best[k] = g.key;
                hit[k].fraction = d;
                hit[k].normal = nrm;
                hit[k].object = 0;
            }
        }
        }
    }
}

// The four static planes' part of the Bullet ray test (after the mesh);
// a plane hit replaces only a strictly closer one.
RSX_HD void ray_planes(V3 from, V3 to, RayHit& hit);

// Full Bullet ray test against the static world (mesh triangles near the ray,
// then the four planes), closest hit wins (strictly closer replaces).
// with_mesh = false leaves out the mesh triangles, for a ray the learned cull
// has shown to be out of their reach.
RSX_HD void ray_static_world(const MeshView& w, V3 from, V3 to, RayHit& hit, bool with_mesh = true) {
    if (with_mesh) ray_mesh(w, from, to, hit);
    ray_planes(from, to, hit);
}

RSX_HD void ray_planes(V3 from, V3 to, RayHit& hit) {
    float frac = hit.object >= 0 ? hit.fraction : 1.0f;
    for (int p = 0; p < kNumPlanes; p++) {
        // btStaticPlaneShape::processAllTriangles around the ray's local AABB,
        // then the triangle ray test on its two triangles.
        V3 pn = plane_normal(p);
        V3 po = plane_origin(p);
        Xf inv = xf_inverse(Xf{m3_identity(), po});
        V3 lf = xform(inv, from), lt = xform(inv, to);
        V3 amin = vmin(lf, lt), amax = vmax(lf, lt);
        V3 half = (amax - amin) * 0.5f;
        float radius = len(half);
        V3 center = (amax + amin) * 0.5f;
        V3 t0, t1;
        plane_space(pn, t0, t1);
        V3 pc = center - (dot(pn, center) - 0.f) * pn;
        V3 ta = pc + t0 * radius + t1 * radius;
        V3 tb = pc + t0 * radius - t1 * radius;
        V3 tc = pc - t0 * radius - t1 * radius;
        V3 nrm;
        if (ray_triangle(lf, lt, ta, tb, tc, frac, nrm)) {
            hit.fraction = frac;
            hit.normal = m3_identity() * nrm;
            hit.object = 0;
        }
        V3 ua = pc - t0 * radius - t1 * radius;
        V3 ub = pc - t0 * radius + t1 * radius;
        V3 uc = pc + t0 * radius + t1 * radius;
        if (ray_triangle(lf, lt, ua, ub, uc, frac, nrm)) {
            hit.fraction = frac;
            hit.normal = m3_identity() * nrm;
            hit.object = 0;
        }
    }
}

RSX_HD void susp_cell_indices(V3 pos_uu, int& i, int& j, int& k) {
    i = (int)clampf(pos_uu.x / kSuspCellSizeX + (kSuspCellsX / 2), 0, kSuspCellsX - 1);
    j = (int)clampf(pos_uu.y / kSuspCellSizeY + (kSuspCellsY / 2), 0, kSuspCellsY - 1);
    k = (int)clampf(pos_uu.z / kSuspCellSizeZ, 0, kSuspCellsZ - 1);
}
RSX_HD int susp_cell(int i, int j, int k) { return (i * kSuspCellsY * kSuspCellsZ) + (j * kSuspCellsZ) + k; }

// Broadphase cell of a point (btRSBroadphase::GetCellIdx, clamped to the grid).
RSX_HD int bp_cell_index(const MeshView& w, V3 pos) {
    V3 f = (pos - V3{kBpMinX, kBpMinY, kBpMinZ}) / kBpCellSize;
    int i = (int)f.x, j = (int)f.y, k = (int)f.z;
    i = i < 0 ? 0 : (i > w.bp_nx - 1 ? w.bp_nx - 1 : i);
    j = j < 0 ? 0 : (j > w.bp_ny - 1 ? w.bp_ny - 1 : j);
    k = k < 0 ? 0 : (k > w.bp_nz - 1 ? w.bp_nz - 1 : k);
    return i * w.bp_ny * w.bp_nz + j * w.bp_nz + k;
}

// Whether a body registered at cell `reg` is listed in cell `cell`
// (btRSBroadphase _UpdateCellsDynamic: the 3x3x3 block around reg).
RSX_HD bool bp_listed(const MeshView& w, int reg, int cell) {
    if (reg < 0) return false;
    const int yz = w.bp_ny * w.bp_nz;
    const int ri = reg / yz, rj = (reg / w.bp_nz) % w.bp_ny, rk = reg % w.bp_nz;
    const int ci = cell / yz, cj = (cell / w.bp_nz) % w.bp_ny, ck = cell % w.bp_nz;
    const int di = ri - ci, dj = rj - cj, dk = rk - ck;
    return di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1;
}

}  // namespace rsx
