// rsx_host.cpp — see rsx_host.h.
//
// Shape constants are evaluated with the expression order Bullet uses when the
// reference builds its bodies, so the engine starts from bit-identical values:
//   btBoxShape ctor + setSafeMargin   btBoxShape.cpp:17-26, btConvexInternalShape.h:63-92
//   btBoxShape::calculateLocalInertia btBoxShape.cpp:33-45
//   btCompoundShape local AABB        btCompoundShape.cpp:40-82, 159-179
//   btSphereShape inertia             btSphereShape.cpp:61-66
//   angular motion disc (breaking threshold)  btCollisionShape.cpp:126-150
//   Car::_BulletSetup wheels           Car.cpp:195-283
//   BoostPad::_Setup                   BoostPad.cpp:37-49
// CMF parsing follows CollisionMeshFile.cpp:11-61 and the hash :75-99.
#include "rsx_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>

namespace rsx {

MeshView HostWorld::view() const {
    MeshView v;
    v.tri = tri.data();
    v.tri_mesh = tri_mesh.data();
    v.tri_rank = tri_rank.data();
    v.cell_start = cell_start.data();
    v.cell_tris = cell_tris.data();
    v.cell_ent = cell_ent.data();
    v.blk_start = blk_start.data();
    v.blk_ent = blk_ent.data();
    v.susp_world = nullptr;
    v.bp_meshes = bp_meshes.data();
    v.mesh_aabb = mesh_aabb.data();
    v.tri_info = tri_info.data();
    v.grid_min = grid_min;
    v.grid_cell = grid_cell;
    v.gx = gx;
    v.gy = gy;
    v.gz = gz;
    v.num_tris = (int)tri_mesh.size();
    v.bp_nx = bp_nx;
    v.bp_ny = bp_ny;
    v.bp_nz = bp_nz;
    return v;
}

uint32_t cmf_hash(const std::vector<int32_t>& idx, const std::vector<float>& verts) {
    uint32_t ntri = (uint32_t)(idx.size() / 3), nv = (uint32_t)(verts.size() / 3);
    uint32_t hash = nv + (ntri * nv);
    const uint32_t kMueller = 0x45D9F3B, kShift = 0x9E3779B9;
    for (size_t t = 0; t < ntri; t++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                float f = verts[(size_t)idx[t * 3 + i] * 3 + j];
                // float -> uint32 as x86-64 compilers emit it (truncate through int64)
                uint32_t cur = (uint32_t)(int64_t)f;
                for (int k = 0; k < 2; k++) cur = ((cur >> 16) ^ cur) * kMueller;
                cur = (cur >> 16) ^ cur;
                hash ^= cur + kShift + (hash << 6) + (hash >> 2);
            }
    return hash;
}

static bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

bool load_world(const std::string& dir, HostWorld& w, std::string& err) {
    std::vector<std::vector<uint8_t>> files(kNumMeshes);
    for (int i = 0; i < kNumMeshes; i++) {
        std::string p = dir + "/mesh_" + std::to_string(i) + ".cmf";
        if (!read_file(p, files[i])) {
            err = "cannot read " + p;
            return false;
        }
    }
    return build_world(files, w, err);
}

static void car_geometry(CarGeom& g) {
    // Octane (CarConfig.cpp:19-60)
    const V3 hitbox = V3{120.507f, 86.6994f, 38.6591f};
    const V3 offset = V3{13.87566f, 0, 20.755f};
    const float front_rad = 12.50f, back_rad = 15.00f;
    const float front_rest = 38.755f, back_rest = 37.055f;
    const V3 front_off = V3{51.25f, 25.90f, 20.755f};
    const V3 back_off = V3{-33.75f, 29.50f, 20.755f};

    V3 half = (hitbox * kUU2BT) * 0.5f;
    float margin = kConvexMargin;
    V3 implicit = half - V3{margin, margin, margin};
    float min_dim = minf(minf(half.x, half.y), half.z);
    // btConvexInternalShape::setSafeMargin calls its own (non-virtual)
    // setMargin, so the implicit dimensions keep the default-margin value and
    // only the margin shrinks.
    float safe = 0.1f * min_dim;
    if (safe < margin) margin = safe;
    g.box_implicit = implicit;
    g.box_margin = margin;
    g.box_offset = offset * kUU2BT;
    // local inertia
    V3 hm = implicit + V3{margin, margin, margin};
    float lx = 2.f * hm.x, ly = 2.f * hm.y, lz = 2.f * hm.z;
    float mass = kCarMassBT;
    V3 inertia = V3{mass / 12.f * (ly * ly + lz * lz), mass / 12.f * (lx * lx + lz * lz), mass / 12.f * (lx * lx + ly * ly)};
    g.inv_inertia_local = V3{inertia.x != 0 ? 1.f / inertia.x : 0.f, inertia.y != 0 ? 1.f / inertia.y : 0.f,
                             inertia.z != 0 ? 1.f / inertia.z : 0.f};
    g.inv_mass = 1.f / mass;
    // compound local AABB from the child box at its offset
    V3 child_ext = V3{hm.x, hm.y, hm.z};  // identity basis: dot3 with |I| rows returns hm exactly
    V3 lmin = g.box_offset - child_ext, lmax = g.box_offset + child_ext;
    g.compound_half = (lmax - lmin) * 0.5f;
    g.compound_center = (lmax + lmin) * 0.5f;
    // angular motion disc of the compound: bounding sphere of its identity AABB
    {
        V3 amin = g.compound_center - g.compound_half, amax = g.compound_center + g.compound_half;
        float radius = len(amax - amin) * 0.5f;
        V3 center = (amin + amax) * 0.5f;
        float disc = radius + len(center);
        g.breaking = disc * kContactBreakingThreshold;
    }
    for (int i = 0; i < 4; i++) {
        bool front = i < 2;
        bool left = i % 2;
        V3 off = front ? front_off : back_off;
        if (left) off.y *= -1;
        g.wheel_cs[i] = off * kUU2BT;
        float rest = (front ? front_rest : back_rest) - kMaxSuspensionTravel;
        g.rest_len[i] = rest * kUU2BT;
        g.wheel_radius[i] = (front ? front_rad : back_rad) * kUU2BT;
        g.susp_force_scale[i] = front ? kSuspForceScaleFront : kSuspForceScaleBack;
    }
}

static void ball_geometry(BallGeom& b) {
    b.radius = kBallRadius * kUU2BT;
    float mass = kBallMassBT;
    b.inv_mass = 1.f / mass;
    float elem = 0.4f * mass * b.radius * b.radius;
    b.inv_inertia = 1.f / elem;
    // RocketSim: bounding sphere radius of a sphere = radius + 0.08, centre 0
    float disc = (b.radius + 0.08f) + 0.f;
    b.breaking = disc * kContactBreakingThreshold;
}

static void pad_geometry(PadGeom* pads) {
    const float small[28][2] = {{0.f, -4240.f}, {-1792.f, -4184.f}, {1792.f, -4184.f}, {-940.f, -3308.f}, {940.f, -3308.f},
                                {0.f, -2816.f}, {-3584.f, -2484.f}, {3584.f, -2484.f}, {-1788.f, -2300.f}, {1788.f, -2300.f},
                                {-2048.f, -1036.f}, {0.f, -1024.f}, {2048.f, -1036.f}, {-1024.f, 0.f}, {1024.f, 0.f},
                                {-2048.f, 1036.f}, {0.f, 1024.f}, {2048.f, 1036.f}, {-1788.f, 2300.f}, {1788.f, 2300.f},
                                {-3584.f, 2484.f}, {3584.f, 2484.f}, {0.f, 2816.f}, {-940.f, 3308.f}, {940.f, 3308.f},
                                {-1792.f, 4184.f}, {1792.f, 4184.f}, {0.f, 4240.f}};
    const float big[6][2] = {{-3584.f, 0.f}, {3584.f, 0.f}, {-3072.f, 4096.f}, {3072.f, 4096.f}, {-3072.f, -4096.f}, {3072.f, -4096.f}};
    for (int i = 0; i < kNumPads; i++) {
        PadGeom& p = pads[i];
        p.big = i < kNumBigPads;
        p.pos_uu = p.big ? V3{big[i][0], big[i][1], 73.f} : V3{small[i - 6][0], small[i - 6][1], 70.f};
        p.pos_bt = p.pos_uu * kUU2BT;
        float box_rad = (p.big ? kPadBoxRadBig : kPadBoxRadSmall) * kUU2BT;
        p.box_min = p.pos_bt - V3{box_rad, box_rad, 0};
        p.box_max = p.pos_bt + V3{box_rad, box_rad, kPadBoxHeight * kUU2BT};
        p.cyl_rad_bt = (p.big ? kPadCylRadBig : kPadCylRadSmall) * kUU2BT;
    }
}

static void internal_edge_info(HostWorld& w);

// Traversal order of the quantized BVH that btBvhTriangleMeshShape builds over
// one mesh (btOptimizedBvh::build, btQuantizedBvh::setQuantizationValues /
// quantize / unQuantize / calcSplittingAxis / sortAndCalcSplittingIndex /
// buildTree / updateSubtreeHeaders, bullet3-3.24
// BroadphaseCollision/btQuantizedBvh.cpp).  RocketSim's Bullet always walks
// the tree subtree header by subtree header
// (reportAabbOverlappingNodex -> walkStacklessQuantizedTreeCacheFriendly,
// btQuantizedBvh.cpp:307-325), and node-array order inside each header's
// subtree; node-array order is the final order of the leaf array after the
// in-place partitions.  A header is recorded for each child of at most 128
// nodes (2048 bytes) of a node of more than 128 nodes, when that node's
// build finishes -- so a small left child is visited after the headers
// inside its large right sibling -- or for the whole tree when no node is
// that large.  The returned rank[t] is triangle t's position in that order.
namespace {
struct QBvh {
    V3 amin, amax, quant;
    struct Leaf {
        uint16_t qmin[3], qmax[3];
        int tri;
    };
    std::vector<Leaf> leaves;

    void quantize(uint16_t* out, V3 p, bool is_max) const {
        V3 v = (p - amin) * quant;
        if (is_max) {
            out[0] = (uint16_t)(((uint16_t)(v.x + 1.f)) | 1);
            out[1] = (uint16_t)(((uint16_t)(v.y + 1.f)) | 1);
            out[2] = (uint16_t)(((uint16_t)(v.z + 1.f)) | 1);
        } else {
            out[0] = (uint16_t)(((uint16_t)(v.x)) & 0xfffe);
            out[1] = (uint16_t)(((uint16_t)(v.y)) & 0xfffe);
            out[2] = (uint16_t)(((uint16_t)(v.z)) & 0xfffe);
        }
    }
    V3 unquantize(const uint16_t* q) const {
        V3 o = V3{(float)q[0] / quant.x, (float)q[1] / quant.y, (float)q[2] / quant.z};
        return o + amin;
    }
    void set_quantization(V3 mn, V3 mx) {
        const V3 clamp = V3{1.f, 1.f, 1.f};
        const V3 q65533 = V3{65533.f, 65533.f, 65533.f};
        amin = mn - clamp;
        amax = mx + clamp;
        V3 size = amax - amin;
        quant = V3{q65533.x / size.x, q65533.y / size.y, q65533.z / size.z};
        uint16_t vin[3];
        quantize(vin, amin, false);
        V3 v = unquantize(vin);
        amin = vmin(amin, v - clamp);
        size = amax - amin;
        quant = V3{q65533.x / size.x, q65533.y / size.y, q65533.z / size.z};
        quantize(vin, amax, true);
        v = unquantize(vin);
        amax = vmax(amax, v + clamp);
        size = amax - amin;
        quant = V3{q65533.x / size.x, q65533.y / size.y, q65533.z / size.z};
    }
    V3 center(int i) const { return 0.5f * (unquantize(leaves[i].qmax) + unquantize(leaves[i].qmin)); }
    int split_axis(int s, int e) const {
        V3 means = V3{0, 0, 0}, var = V3{0, 0, 0};
        int n = e - s;
        for (int i = s; i < e; i++) means += center(i);
        means *= (1.f / (float)n);
        for (int i = s; i < e; i++) {
            V3 d = center(i) - means;
            var += d * d;
        }
        var *= (1.f / ((float)n - 1));
        return gjk::max_axis(var);
    }
    int split_index(int s, int e, int axis) {
        int split = s, n = e - s;
        V3 means = V3{0, 0, 0};
        for (int i = s; i < e; i++) means += center(i);
        means *= (1.f / (float)n);
        float sv = idx(means, axis);
        for (int i = s; i < e; i++) {
            if (idx(center(i), axis) > sv) {
                std::swap(leaves[i], leaves[split]);
                split++;
            }
        }
        int bal = n / 3;
        bool unbalanced = (split <= (s + bal)) || (split >= (e - 1 - bal));
        if (unbalanced) split = s + (n >> 1);
        return split;
    }
    // Subtree headers as leaf ranges [first, second), in the order Bullet
    // appends them.  A subtree over k leaves has 2k - 1 nodes.
    static constexpr int kMaxSubtreeNodes = 2048 / 16;  // MAX_SUBTREE_SIZE_IN_BYTES / sizeof(btQuantizedBvhNode)
    std::vector<std::pair<int, int>> headers;
    void build(int s, int e) {
        if (e - s == 1) return;
        int axis = split_axis(s, e);
        int split = split_index(s, e, axis);
        build(s, split);
        build(split, e);
        if (2 * (e - s) - 1 > kMaxSubtreeNodes) {
            if (2 * (split - s) - 1 <= kMaxSubtreeNodes) headers.push_back({s, split});
            if (2 * (e - split) - 1 <= kMaxSubtreeNodes) headers.push_back({split, e});
        }
    }
};
}  // namespace

static std::vector<int> bvh_leaf_rank(const float* tri, int nt, V3 mn, V3 mx) {
    QBvh q;
    q.set_quantization(mn, mx);
    q.leaves.resize(nt);
    for (int t = 0; t < nt; t++) {
        const float* p = tri + (size_t)t * 9;
        V3 lo = V3{kLargeFloat, kLargeFloat, kLargeFloat}, hi = -lo;
        for (int k = 0; k < 3; k++) {
            V3 v = V3{p[k * 3], p[k * 3 + 1], p[k * 3 + 2]};
            lo = vmin(lo, v);
            hi = vmax(hi, v);
        }
        const float kMinDim = 0.002f, kMinHalf = 0.001f;
        if (hi.x - lo.x < kMinDim) { hi.x += kMinHalf; lo.x -= kMinHalf; }
        if (hi.y - lo.y < kMinDim) { hi.y += kMinHalf; lo.y -= kMinHalf; }
        if (hi.z - lo.z < kMinDim) { hi.z += kMinHalf; lo.z -= kMinHalf; }
        q.quantize(q.leaves[t].qmin, lo, false);
        q.quantize(q.leaves[t].qmax, hi, true);
        q.leaves[t].tri = t;
    }
    q.build(0, nt);
    if (q.headers.empty()) q.headers.push_back({0, nt});
    std::vector<int> rank(nt, -1);
    int r = 0;
    for (const auto& h : q.headers)
        for (int i = h.first; i < h.second; i++) rank[q.leaves[i].tri] = r++;
    return rank;
}

bool build_world(const std::vector<std::vector<uint8_t>>& files, HostWorld& w, std::string& err) {
    if (files.size() != (size_t)kNumMeshes) {
        err = "expected 16 collision mesh files";
        return false;
    }
    w.tri.clear();
    w.tri_mesh.clear();
    w.tri_rank.clear();
    w.mesh_aabb.assign(kNumMeshes * 6, 0.f);
    w.mesh_hash.assign(kNumMeshes, 0);
    for (int m = 0; m < kNumMeshes; m++) {
        const std::vector<uint8_t>& f = files[m];
        if (f.size() < 8) {
            err = "collision mesh " + std::to_string(m) + " is truncated";
            return false;
        }
        int32_t nt, nv;
        memcpy(&nt, f.data(), 4);
        memcpy(&nv, f.data() + 4, 4);
        if (nt <= 0 || nv <= 0 || nt > 1000000 || nv > 1000000 || f.size() < 8 + (size_t)nt * 12 + (size_t)nv * 12) {
            err = "collision mesh " + std::to_string(m) + " has a bad triangle/vertex count";
            return false;
        }
        std::vector<int32_t> idx((size_t)nt * 3);
        std::vector<float> v((size_t)nv * 3);
        memcpy(idx.data(), f.data() + 8, (size_t)nt * 12);
        memcpy(v.data(), f.data() + 8 + (size_t)nt * 12, (size_t)nv * 12);
        for (int32_t i : idx)
            if (i < 0 || i >= nv) {
                err = "collision mesh " + std::to_string(m) + " has a bad vertex index";
                return false;
            }
        w.mesh_hash[m] = cmf_hash(idx, v);
        V3 mn = V3{kLargeFloat, kLargeFloat, kLargeFloat}, mx = V3{-kLargeFloat, -kLargeFloat, -kLargeFloat};
        for (int t = 0; t < nt; t++) {
            for (int k = 0; k < 3; k++) {
                const float* p = &v[(size_t)idx[t * 3 + k] * 3];
                w.tri.push_back(p[0]);
                w.tri.push_back(p[1]);
                w.tri.push_back(p[2]);
                mn = vmin(mn, V3{p[0], p[1], p[2]});
                mx = vmax(mx, V3{p[0], p[1], p[2]});
            }
            w.tri_mesh.push_back(m);
            w.tri_rank.push_back(t);
        }
        {
            const size_t base = w.tri.size() - (size_t)nt * 9;
            std::vector<int> rank = bvh_leaf_rank(&w.tri[base], nt, mn, mx);
            const size_t tbase = w.tri_rank.size() - (size_t)nt;
            for (int t = 0; t < nt; t++) w.tri_rank[tbase + t] = rank[t];
        }
        // btTriangleMeshShape::getAabb at identity: centre +- half extents
        V3 half = (mx - mn) * 0.5f;
        V3 center = (mx + mn) * 0.5f;
        V3 amin = center - half, amax = center + half;
        float* a = &w.mesh_aabb[m * 6];
        a[0] = amin.x; a[1] = amin.y; a[2] = amin.z;
        a[3] = amax.x; a[4] = amax.y; a[5] = amax.z;
    }
    const int ntri = (int)w.tri_mesh.size();
    if (ntri > kMaxTris) {
        err = "collision meshes hold " + std::to_string(ntri) + " triangles; the candidate packing allows " +
              std::to_string(kMaxTris);
        return false;
    }

    // Uniform grid over all triangles (4 BT cells).
    V3 gmin = V3{kLargeFloat, kLargeFloat, kLargeFloat}, gmax = -gmin;
    for (size_t i = 0; i < w.tri.size(); i += 3) {
        gmin = vmin(gmin, V3{w.tri[i], w.tri[i + 1], w.tri[i + 2]});
        gmax = vmax(gmax, V3{w.tri[i], w.tri[i + 1], w.tri[i + 2]});
    }
    w.grid_cell = 4.0f;
    w.grid_min = gmin - V3{1, 1, 1};
    w.gx = (int)std::ceil((gmax.x + 1 - w.grid_min.x) / w.grid_cell);
    w.gy = (int)std::ceil((gmax.y + 1 - w.grid_min.y) / w.grid_cell);
    w.gz = (int)std::ceil((gmax.z + 1 - w.grid_min.z) / w.grid_cell);
    const int ncell = w.gx * w.gy * w.gz;
    std::vector<std::vector<int>> cells(ncell);
    MeshView tmp;
    // Cells are found as gather_tris finds them (a multiplication by the
    // inverse cell size), so that a triangle is listed in exactly the cells a
    // query meeting its AABB visits, whatever the cell size.
    const float cell_inv = 1.0f / w.grid_cell;
    for (int t = 0; t < ntri; t++) {
        const float* p = &w.tri[(size_t)t * 9];
        V3 a = V3{p[0], p[1], p[2]}, b = V3{p[3], p[4], p[5]}, c = V3{p[6], p[7], p[8]};
        V3 mn = vmin(vmin(a, b), c), mx = vmax(vmax(a, b), c);
        int x0 = (int)std::floor((mn.x - w.grid_min.x) * cell_inv), x1 = (int)std::floor((mx.x - w.grid_min.x) * cell_inv);
        int y0 = (int)std::floor((mn.y - w.grid_min.y) * cell_inv), y1 = (int)std::floor((mx.y - w.grid_min.y) * cell_inv);
        int z0 = (int)std::floor((mn.z - w.grid_min.z) * cell_inv), z1 = (int)std::floor((mx.z - w.grid_min.z) * cell_inv);
        for (int ix = std::max(x0, 0); ix <= std::min(x1, w.gx - 1); ix++)
            for (int iy = std::max(y0, 0); iy <= std::min(y1, w.gy - 1); iy++)
                for (int iz = std::max(z0, 0); iz <= std::min(z1, w.gz - 1); iz++) cells[(ix * w.gy + iy) * w.gz + iz].push_back(t);
    }
    (void)tmp;
    w.cell_start.assign(ncell + 1, 0);
    w.cell_tris.clear();
    for (int c = 0; c < ncell; c++) {
        w.cell_start[c] = (int)w.cell_tris.size();
        w.cell_tris.insert(w.cell_tris.end(), cells[c].begin(), cells[c].end());
    }
    w.cell_start[ncell] = (int)w.cell_tris.size();
    // the candidate lists' bound: distinct triangles listed in any 2 x 2 x 2
    // block of cells
    {
        std::vector<int> mark(ntri, -1);
        int worst = 0, stamp = 0;
        for (int x = 0; x < w.gx; x++)
            for (int y = 0; y < w.gy; y++)
                for (int z = 0; z < w.gz; z++) {
                    int cnt = 0;
                    stamp++;
                    for (int ix = x; ix <= std::min(x + 1, w.gx - 1); ix++)
                        for (int iy = y; iy <= std::min(y + 1, w.gy - 1); iy++)
                            for (int iz = z; iz <= std::min(z + 1, w.gz - 1); iz++) {
                                const int c = (ix * w.gy + iy) * w.gz + iz;
                                for (int k = w.cell_start[c]; k < w.cell_start[c + 1]; k++)
                                    if (mark[w.cell_tris[k]] != stamp) {
                                        mark[w.cell_tris[k]] = stamp;
                                        cnt++;
                                    }
                            }
                    worst = std::max(worst, cnt);
                }
        if (worst > kMaxCand) {
            err = "a 2x2x2 block of grid cells lists " + std::to_string(worst) +
                  " triangles; the candidate lists hold " + std::to_string(kMaxCand);
            return false;
        }
    }
    // gather entries (after the tree ranks are known)
    w.cell_ent.resize(w.cell_tris.size());
    {
        const float inv = 1.0f / w.grid_cell;
        for (size_t k = 0; k < w.cell_tris.size(); k++) {
            const int t = w.cell_tris[k];
            const float* p = &w.tri[(size_t)t * 9];
            V3 a = V3{p[0], p[1], p[2]}, b = V3{p[3], p[4], p[5]}, c = V3{p[6], p[7], p[8]};
            V3 mn = vmin(vmin(a, b), c), mx = vmax(vmax(a, b), c);
            GridEntry& g = w.cell_ent[k];
            g.key = (w.tri_mesh[t] << kCandMeshShift) | (w.tri_rank[t] << kCandRankShift) | t;
            const int fx = clampi((int)std::floor((mn.x - w.grid_min.x) * inv), 0, w.gx - 1);
            const int fy = clampi((int)std::floor((mn.y - w.grid_min.y) * inv), 0, w.gy - 1);
            const int fz = clampi((int)std::floor((mn.z - w.grid_min.z) * inv), 0, w.gz - 1);
            g.first = fx | (fy << 10) | (fz << 20);
            g.mn[0] = mn.x, g.mn[1] = mn.y, g.mn[2] = mn.z;
            g.mx[0] = mx.x, g.mx[1] = mx.y, g.mx[2] = mx.z;
        }
    }

    // Block lists: for every cell, the distinct triangles of the 2 x 2 x 2
    // block of cells whose lowest corner it is (clamped at the grid's far
    // faces), sorted by packed key.  A body query box spans at most two cells
    // per axis, so the block at its lowest cell holds every triangle a walk
    // over its cells would find, once each and already in key order.
    {
        std::vector<GridEntry> ent_of(ntri);
        for (size_t k = 0; k < w.cell_tris.size(); k++) ent_of[w.cell_tris[k]] = w.cell_ent[k];
        std::vector<int> mark(ntri, -1), list;
        w.blk_start.assign(ncell + 1, 0);
        w.blk_ent.clear();
        for (int x = 0; x < w.gx; x++)
            for (int y = 0; y < w.gy; y++)
                for (int z = 0; z < w.gz; z++) {
                    const int blk = (x * w.gy + y) * w.gz + z;
                    w.blk_start[blk] = (int)w.blk_ent.size();
                    list.clear();
                    for (int ix = x; ix <= std::min(x + 1, w.gx - 1); ix++)
                        for (int iy = y; iy <= std::min(y + 1, w.gy - 1); iy++)
                            for (int iz = z; iz <= std::min(z + 1, w.gz - 1); iz++) {
                                const int c = (ix * w.gy + iy) * w.gz + iz;
                                for (int k = w.cell_start[c]; k < w.cell_start[c + 1]; k++)
                                    if (mark[w.cell_tris[k]] != blk) {
                                        mark[w.cell_tris[k]] = blk;
                                        list.push_back(w.cell_tris[k]);
                                    }
                            }
                    std::sort(list.begin(), list.end(),
                              [&](int a, int b) { return ent_of[a].key < ent_of[b].key; });
                    for (int t : list) w.blk_ent.push_back(ent_of[t]);
                }
        w.blk_start[ncell] = (int)w.blk_ent.size();
    }

    // RocketSim broadphase static cell lists (btRSBroadphase.cpp:72-136): a
    // mesh is listed in the 27 neighbours of every cell its triangles touch.
    const V3 bpmin = V3{kBpMinX, kBpMinY, kBpMinZ}, bpmax = V3{kBpMaxX, kBpMaxY, kBpMaxZ};
    V3 range = bpmax - bpmin;
    w.bp_nx = std::max(1, (int)std::ceil(range.x / kBpCellSize));
    w.bp_ny = std::max(1, (int)std::ceil(range.y / kBpCellSize));
    w.bp_nz = std::max(1, (int)std::ceil(range.z / kBpCellSize));
    w.bp_meshes.assign((size_t)w.bp_nx * w.bp_ny * w.bp_nz, 0);
    auto cell_idx = [&](V3 pos, int& i, int& j, int& k) {
        V3 f = (pos - bpmin) / kBpCellSize;
        i = std::min(std::max((int)f.x, 0), w.bp_nx - 1);
        j = std::min(std::max((int)f.y, 0), w.bp_ny - 1);
        k = std::min(std::max((int)f.z, 0), w.bp_nz - 1);
    };
    for (int m = 0; m < kNumMeshes; m++) {
        const float* a = &w.mesh_aabb[m * 6];
        const float th = kContactBreakingThreshold;
        V3 amin = V3{a[0], a[1], a[2]} - V3{th, th, th};
        V3 amax = vmin(V3{a[3], a[4], a[5]} + V3{th, th, th}, bpmax);
        int i0, j0, k0, i1, j1, k1;
        cell_idx(amin, i0, j0, k0);
        cell_idx(amax, i1, j1, k1);
        for (int i = i0; i <= i1; i++)
            for (int j = j0; j <= j1; j++)
                for (int k = k0; k <= k1; k++) {
                    V3 cmin = bpmin + V3{(float)i, (float)j, (float)k} * kBpCellSize;
                    V3 cmax = cmin + V3{kBpCellSize, kBpCellSize, kBpCellSize};
                    bool hit = false;
                    for (int t = 0; t < ntri && !hit; t++) {
                        if (w.tri_mesh[t] != m) continue;
                        const float* p = &w.tri[(size_t)t * 9];
                        if (tri_aabb_overlap(V3{p[0], p[1], p[2]}, V3{p[3], p[4], p[5]}, V3{p[6], p[7], p[8]}, cmin, cmax)) hit = true;
                    }
                    if (!hit) continue;
                    for (int di = -1; di <= 1; di++)
                        for (int dj = -1; dj <= 1; dj++)
                            for (int dk = -1; dk <= 1; dk++) {
                                int ci = i + di, cj = j + dj, ck = k + dk;
                                if (ci < 0 || cj < 0 || ck < 0 || ci >= w.bp_nx || cj >= w.bp_ny || ck >= w.bp_nz) continue;
                                w.bp_meshes[(size_t)ci * w.bp_ny * w.bp_nz + cj * w.bp_nz + ck] |= (uint16_t)(1u << m);
                            }
                }
    }
    internal_edge_info(w);
    car_geometry(w.cg);
    ball_geometry(w.bg);
    pad_geometry(w.pads);
    return true;
}


// --- internal edge information (btGenerateInternalEdgeInfo) --------------------
// For every triangle and each of its edges shared with another triangle of the
// same mesh, the signed angle to the neighbour and convexity flags, computed
// with the steps of btConnectivityProcessor::processTriangle
// (btInternalEdgeUtility.cpp:46-250).
static float bt_get_angle(V3 edge_a, V3 normal_a, V3 normal_b) {
    return atan2f(dot(normal_b, edge_a), dot(normal_b, normal_a));
}
static V3 quat_rotate(Q4 r, V3 v) {
    // q = r * v (pure quaternion), q *= r^-1
    Q4 q = Q4{r.w * v.x + r.y * v.z - r.z * v.y, r.w * v.y + r.z * v.x - r.x * v.z, r.w * v.z + r.x * v.y - r.y * v.x,
              -r.x * v.x - r.y * v.y - r.z * v.z};
    q = qmul(q, qinverse(r));
    return V3{q.x, q.y, q.z};
}

static void connectivity(const V3* A, const V3* B, float* info) {
    const float equal_thr = 0.0001f * 0.0001f;
    const float planar_eps = 0.0001f;
    int numshared = 0;
    int sa[3] = {-1, -1, -1}, sb[3] = {-1, -1, -1};
    if (len2(cross(B[1] - B[0], B[2] - B[0])) < equal_thr) return;
    if (len2(cross(A[1] - A[0], A[2] - A[0])) < equal_thr) return;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            if (len2(A[i] - B[j]) < equal_thr) {
                sa[numshared] = i;
                sb[numshared] = j;
                numshared++;
                if (numshared >= 3) return;
            }
        }
        if (numshared >= 3) return;
    }
    if (numshared != 2) return;
    if (sa[0] == 0 && sa[1] == 2) {
        sa[0] = 2;
        sa[1] = 0;
        int t = sb[1];
        sb[1] = sb[0];
        sb[0] = t;
    }
    int sum_a = sa[0] + sa[1];
    int other_a = 3 - sum_a;
    V3 edge = A[sa[1]] - A[sa[0]];
    int other_b = 3 - (sb[0] + sb[1]);
    V3 tb[3] = {B[sb[1]], B[sb[0]], B[other_b]};
    V3 normal_a = normalized(cross(A[1] - A[0], A[2] - A[0]));
    V3 normal_b = normalized(cross(tb[1] - tb[0], tb[2] - tb[0]));
    edge = normalized(edge);
    V3 ecA = normalized(cross(edge, normal_a));
    if (dot(ecA, A[other_a] - A[sa[0]]) < 0) ecA = ecA * -1.f;
    V3 ecB = normalized(cross(edge, normal_b));
    if (dot(ecB, B[other_b] - B[sb[0]]) < 0) ecB = ecB * -1.f;
    float ang4 = 0.f, corrected = 0.f;
    bool is_convex = false;
    V3 calc_edge = cross(ecA, ecB);
    float l2 = len2(calc_edge);
    if (!(l2 < planar_eps)) {
        calc_edge = normalized(calc_edge);
        V3 calc_na = normalized(cross(calc_edge, ecA));
        float angle2 = bt_get_angle(calc_na, ecA, ecB);
        ang4 = kPi - angle2;
        float dot_a = dot(normal_a, ecB);
        is_convex = (dot_a < 0.f);
        corrected = is_convex ? ang4 : -ang4;
    }
    uint32_t flags;
    memcpy(&flags, &info[3], 4);
    int slot;
    V3 e;
    uint32_t swap_bit, convex_bit;
    if (sum_a == 1) { e = A[0] - A[1]; slot = 0; swap_bit = 8; convex_bit = 1; }
    else if (sum_a == 2) { e = A[2] - A[0]; slot = 2; swap_bit = 32; convex_bit = 4; }
    else { e = A[1] - A[2]; slot = 1; swap_bit = 16; convex_bit = 2; }
    Q4 orn = q_axis_angle(e, -corrected);
    V3 cnb = quat_rotate(orn, normal_a);
    if (dot(cnb, normal_b) < 0) flags |= swap_bit;
    info[slot] = -corrected;
    if (is_convex) flags |= convex_bit;
    memcpy(&info[3], &flags, 4);
}

static void internal_edge_info(HostWorld& w) {
    const int ntri = (int)w.tri_mesh.size();
    w.tri_info.assign((size_t)ntri * 4, 0.f);
    const float two_pi = 2.0f * kPi;
    for (int t = 0; t < ntri; t++) {
        w.tri_info[(size_t)t * 4 + 0] = two_pi;
        w.tri_info[(size_t)t * 4 + 1] = two_pi;
        w.tri_info[(size_t)t * 4 + 2] = two_pi;
    }
    std::vector<uint8_t> has(ntri, 0);
    MeshView mv = w.view();
    std::vector<int> cand(4096);
    for (int t = 0; t < ntri; t++) {
        V3 A[3];
        load_tri(mv, t, A[0], A[1], A[2]);
        V3 mn = vmin(vmin(A[0], A[1]), A[2]), mx = vmax(vmax(A[0], A[1]), A[2]);
        int n = gather_tris(mv, mn, mx, cand.data(), (int)cand.size());
        float info[4] = {two_pi, two_pi, two_pi, 0.f};
        bool any = false;
        for (int k = 0; k < n; k++) {
            int u = cand_tri(cand[k]);
            if (u == t || w.tri_mesh[u] != w.tri_mesh[t]) continue;
            V3 B[3];
            load_tri(mv, u, B[0], B[1], B[2]);
            float before[4] = {info[0], info[1], info[2], info[3]};
            connectivity(A, B, info);
            if (memcmp(before, info, sizeof(info)) != 0) any = true;
        }
        if (any) memcpy(&w.tri_info[(size_t)t * 4], info, sizeof(info));
    }
}

// --- record conversion -----------------------------------------------------------
static V3 v_in(const rsx_vec3& v) { return V3{v.x, v.y, v.z}; }
static rsx_vec3 v_out(V3 v) { return rsx_vec3{v.x, v.y, v.z}; }

Controls controls_from_public(const rsx_controls& c) {
    return Controls{c.throttle, c.steer, c.pitch, c.yaw, c.roll, c.jump != 0, c.boost != 0, c.handbrake != 0};
}
rsx_controls controls_to_public(const Controls& c) {
    rsx_controls r{};
    r.throttle = c.throttle;
    r.steer = c.steer;
    r.pitch = c.pitch;
    r.yaw = c.yaw;
    r.roll = c.roll;
    r.jump = c.jump;
    r.boost = c.boost;
    r.handbrake = c.handbrake;
    return r;
}

// RLGym's observation pad order (CommonValues.h:45-80 BOOST_LOCATIONS) mapped
// onto the engine's pad indices by 2D position, as GameState's
// boostPadIndexMap does (GameState.cpp:10-50).  Returns false if a location
// has no unique match.
bool obs_pad_map(const PadGeom* pads, int8_t* out) {
    static const float loc[kNumPads][2] = {
        {0.f, -4240.f}, {-1792.f, -4184.f}, {1792.f, -4184.f}, {-3072.f, -4096.f}, {3072.f, -4096.f},
        {-940.f, -3308.f}, {940.f, -3308.f}, {0.f, -2816.f}, {-3584.f, -2484.f}, {3584.f, -2484.f},
        {-1788.f, -2300.f}, {1788.f, -2300.f}, {-2048.f, -1036.f}, {0.f, -1024.f}, {2048.f, -1036.f},
        {-3584.f, 0.f}, {-1024.f, 0.f}, {1024.f, 0.f}, {3584.f, 0.f}, {-2048.f, 1036.f},
        {0.f, 1024.f}, {2048.f, 1036.f}, {-1788.f, 2300.f}, {1788.f, 2300.f}, {-3584.f, 2484.f},
        {3584.f, 2484.f}, {0.f, 2816.f}, {-940.f, 3310.f}, {940.f, 3308.f}, {-3072.f, 4096.f},
        {3072.f, 4096.f}, {-1792.f, 4184.f}, {1792.f, 4184.f}, {0.f, 4240.f}};
    for (int i = 0; i < kNumPads; i++) {
        int found = -1;
        for (int j = 0; j < kNumPads; j++) {
            float dx = pads[j].pos_uu.x - loc[i][0], dy = pads[j].pos_uu.y - loc[i][1];
            if (dx * dx + dy * dy < 10.f) {
                if (found >= 0) return false;
                found = j;
            }
        }
        if (found < 0) return false;
        out[i] = (int8_t)found;
    }
    return true;
}

// RLGym's DefaultAction table (DefaultAction.cpp:3-59): 24 ground actions
// (throttle x steer x boost x handbrake, boosting only at full throttle; yaw
// follows steer), then 66 air actions (pitch x yaw x roll x jump x boost, no
// yaw with jump, no duplicate of the all-zero ground rows; throttle follows
// boost, steer follows yaw, handbrake on for jumps with a direction).
std::vector<rsx_controls> default_action_lut(int* ground_actions) {
    std::vector<rsx_controls> out;
    const float tri[3] = {-1, 0, 1};
    auto push = [&](float thr, float steer, float pitch, float yaw, float roll, int jump, int boost, int hb) {
        rsx_controls c{};
        c.throttle = thr;
        c.steer = steer;
        c.pitch = pitch;
        c.yaw = yaw;
        c.roll = roll;
        c.jump = (uint8_t)jump;
        c.boost = (uint8_t)boost;
        c.handbrake = (uint8_t)hb;
        out.push_back(c);
    };
    for (float thr : tri)
        for (float steer : tri)
            for (int boost = 0; boost < 2; boost++)
                for (int hb = 0; hb < 2; hb++) {
                    if (boost && thr != 1) continue;
                    push(thr, steer, 0, steer, 0, 0, boost, hb);
                }
    if (ground_actions) *ground_actions = (int)out.size();
    for (float pitch : tri)
        for (float yaw : tri)
            for (float roll : tri)
                for (int jump = 0; jump < 2; jump++)
                    for (int boost = 0; boost < 2; boost++) {
                        if (jump && yaw != 0) continue;
                        if (pitch == 0 && roll == 0 && jump == 0) continue;
                        int hb = (jump && (pitch != 0 || yaw != 0 || roll != 0)) ? 1 : 0;
                        push((float)boost, yaw, pitch, yaw, roll, jump, boost, hb);
                    }
    return out;
}

// DefaultAction's constructor (DefaultAction.cpp:60-89): which entries a
// grounded car, an airborne car, boost and jump make available.  The air
// mask starts after the first aerial entry (strict comparison in the
// reference) and takes in the ground entries whose throttle equals their
// boost and whose handbrake equals (yaw != 0).
std::vector<uint8_t> default_action_masks() {
    int ng = 0;
    const std::vector<rsx_controls> lut = default_action_lut(&ng);
    const int n = (int)lut.size();
    std::vector<uint8_t> m(4 * (size_t)n, 0);
    uint8_t *ground = m.data(), *air = ground + n, *boost = air + n, *jump = boost + n;
    for (int i = 0; i < n; i++) {
        const rsx_controls& a = lut[i];
        jump[i] = a.jump != 0;
        boost[i] = a.boost != 0;
        ground[i] = i < ng;
        air[i] = i > ng && !a.jump;
        if (i < ng && a.throttle == (float)a.boost && (a.yaw != 0.f) == (a.handbrake != 0)) air[i] = 1;
    }
    return m;
}

// Car::SetState (Car.cpp:23-36): transform, velocities (uu -> BT), inertia.
void car_from_public(Car& c, const rsx_car_state& s, const CarGeom& g) {
    c.xf.o = v_in(s.pos) * kUU2BT;
    c.xf.b = m3_from_axes(v_in(s.rot.forward), v_in(s.rot.right), v_in(s.rot.up));
    c.vel = v_in(s.vel) * kUU2BT;
    c.ang = v_in(s.ang_vel);
    car_update_inertia(c, g);
    c.vel_cache = V3{0, 0, 0};
    c.force = c.torque = V3{0, 0, 0};
    c.boost = s.boost;
    c.time_spent_boosting = s.time_spent_boosting;
    c.jump_time = s.jump_time;
    c.flip_time = s.flip_time;
    c.air_time = s.air_time;
    c.air_time_since_jump = s.air_time_since_jump;
    c.supersonic_time = s.supersonic_time;
    c.handbrake_val = s.handbrake_val;
    c.auto_flip_timer = s.auto_flip_timer;
    c.auto_flip_torque_scale = s.auto_flip_torque_scale;
    c.demo_respawn_timer = s.demo_respawn_timer;
    c.car_contact_cooldown = s.car_contact_cooldown;
    c.flip_rel_torque = v_in(s.flip_rel_torque);
    c.world_contact_normal = v_in(s.world_contact_normal);
    c.on_ground = s.is_on_ground;
    c.has_jumped = s.has_jumped;
    c.has_double_jumped = s.has_double_jumped;
    c.has_flipped = s.has_flipped;
    c.is_flipping = s.is_flipping;
    c.is_jumping = s.is_jumping;
    c.is_supersonic = s.is_supersonic;
    c.is_auto_flipping = s.is_auto_flipping;
    c.is_demoed = s.is_demoed;
    c.world_contact = s.world_contact;
    c.ball_hit_valid = s.ball_hit_valid;
    c.team = s.team;
    for (int i = 0; i < 4; i++) c.wheel_contact[i] = s.wheels_with_contact[i];
    c.car_contact_other_id = s.car_contact_other_id;
    c.ball_hit_tick = s.ball_hit_tick;
    c.ball_hit_extra_tick = s.ball_hit_extra_impulse_tick;
    c.ball_hit_rel_pos = v_in(s.ball_hit_rel_pos);
    c.ball_hit_ball_pos = v_in(s.ball_hit_ball_pos);
    c.ball_hit_extra_vel = v_in(s.ball_hit_extra_vel);
    c.last_controls = controls_from_public(s.last_controls);
}

// Car::GetState (Car.cpp:10-20)
void car_to_public(const Car& c, rsx_car_state& s) {
    memset(&s, 0, sizeof(s));
    s.pos = v_out(c.xf.o * kBT2UU);
    s.rot.forward = v_out(col(c.xf.b, 0));
    s.rot.right = v_out(col(c.xf.b, 1));
    s.rot.up = v_out(col(c.xf.b, 2));
    s.vel = v_out(c.vel * kBT2UU);
    s.ang_vel = v_out(c.ang);
    s.boost = c.boost;
    s.time_spent_boosting = c.time_spent_boosting;
    s.jump_time = c.jump_time;
    s.flip_time = c.flip_time;
    s.air_time = c.air_time;
    s.air_time_since_jump = c.air_time_since_jump;
    s.supersonic_time = c.supersonic_time;
    s.handbrake_val = c.handbrake_val;
    s.auto_flip_timer = c.auto_flip_timer;
    s.auto_flip_torque_scale = c.auto_flip_torque_scale;
    s.demo_respawn_timer = c.demo_respawn_timer;
    s.car_contact_cooldown = c.car_contact_cooldown;
    s.flip_rel_torque = v_out(c.flip_rel_torque);
    s.world_contact_normal = v_out(c.world_contact_normal);
    s.is_on_ground = c.on_ground;
    s.has_jumped = c.has_jumped;
    s.has_double_jumped = c.has_double_jumped;
    s.has_flipped = c.has_flipped;
    s.is_flipping = c.is_flipping;
    s.is_jumping = c.is_jumping;
    s.is_supersonic = c.is_supersonic;
    s.is_auto_flipping = c.is_auto_flipping;
    s.is_demoed = c.is_demoed;
    s.world_contact = c.world_contact;
    s.ball_hit_valid = c.ball_hit_valid;
    s.team = (uint8_t)c.team;
    for (int i = 0; i < 4; i++) s.wheels_with_contact[i] = c.wheel_contact[i];
    s.id = c.id;
    s.car_contact_other_id = c.car_contact_other_id;
    s.ball_hit_tick = c.ball_hit_tick;
    s.ball_hit_extra_impulse_tick = c.ball_hit_extra_tick;
    s.ball_hit_rel_pos = v_out(c.ball_hit_rel_pos);
    s.ball_hit_ball_pos = v_out(c.ball_hit_ball_pos);
    s.ball_hit_extra_vel = v_out(c.ball_hit_extra_vel);
    s.last_controls = controls_to_public(c.last_controls);
}

void ball_from_public(Ball& b, const rsx_ball_state& s) {
    b.xf.b = m3_identity();
    b.xf.o = v_in(s.pos) * kUU2BT;
    b.vel = v_in(s.vel) * kUU2BT;
    b.ang = v_in(s.ang_vel);
    b.vel_cache = V3{0, 0, 0};
    b.force = V3{0, 0, 0};
    b.interp = b.xf;
}

void ball_to_public(const Ball& b, rsx_ball_state& s) {
    s.pos = v_out(b.xf.o * kBT2UU);
    s.vel = v_out(b.vel * kBT2UU);
    s.ang_vel = v_out(b.ang);
}

}  // namespace rsx
