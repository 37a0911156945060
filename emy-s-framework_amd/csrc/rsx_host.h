// rsx_host.h — host-side construction of the read-only world: the soccar
// collision meshes (CMF files), the spatial indexes over them, the Octane and
// ball shape constants, and conversion between the public C records
// (include/rsx_types.h) and the engine's working state.
#pragma once

#include <string>
#include <vector>

#include "../../include/rsx_types.h"
#include "rsx_step.h"

namespace rsx {

struct HostWorld {
    std::vector<float> tri;        // num_tris * 9
    std::vector<int> tri_mesh, tri_rank;
    std::vector<int> cell_start, cell_tris;
    std::vector<GridEntry> cell_ent;
    std::vector<int> blk_start;
    std::vector<GridEntry> blk_ent;
    std::vector<uint16_t> bp_meshes;
    std::vector<float> mesh_aabb;  // 16 * 6
    std::vector<float> tri_info;   // num_tris * 4
    std::vector<uint32_t> mesh_hash;
    V3 grid_min;
    float grid_cell;
    int gx, gy, gz;
    int bp_nx, bp_ny, bp_nz;
    CarGeom cg;
    BallGeom bg;
    PadGeom pads[kNumPads];

    MeshView view() const;
};

// Loads mesh_0.cmf .. mesh_15.cmf from dir; returns false (and sets err) on failure.
bool load_world(const std::string& dir, HostWorld& w, std::string& err);
// Same from 16 in-memory CMF blobs.
bool build_world(const std::vector<std::vector<uint8_t>>& files, HostWorld& w, std::string& err);

// RocketSim's CMF hash (CollisionMeshFile.cpp:75-99).
uint32_t cmf_hash(const std::vector<int32_t>& idx, const std::vector<float>& verts);

// Public record <-> working state
void car_from_public(Car& c, const rsx_car_state& s, const CarGeom& g);
void car_to_public(const Car& c, rsx_car_state& s);
void ball_from_public(Ball& b, const rsx_ball_state& s);
void ball_to_public(const Ball& b, rsx_ball_state& s);
Controls controls_from_public(const rsx_controls& c);
rsx_controls controls_to_public(const Controls& c);

// Observation pad slot -> engine pad index (RLGym BOOST_LOCATIONS order).
bool obs_pad_map(const PadGeom* pads, int8_t* out34);

// RLGym DefaultAction's 90-entry discrete action table, in its order.
std::vector<rsx_controls> default_action_lut(int* ground_actions = nullptr);
// DefaultAction's masks over the default table, kNumDefaultActions bytes each,
// in the order ground, air, boost, jump.
std::vector<uint8_t> default_action_masks();

}  // namespace rsx
