// rsx_internal.h — hooks between the engine (rsx_engine.hip), the
// learned-cull module (rsx_sdf.hip) and the device GAE (rsx_gae.hip) inside
// librsx.so; not part of the C-ABI.
#pragma once

#include "rsx_host.h"
#include "rsx_sdf.h"

struct rsx_engine;

namespace rsx {

// The engine's host-side world (mesh triangles in Bullet units, grids).
const HostWorld& engine_host_world(const rsx_engine* e);
// The device the engine's buffers live on.
int engine_device(const rsx_engine* e);
// Installs (or, with c.enabled == 0, removes) the cull the tick kernels use.
void engine_set_cull(rsx_engine* e, const sdf::Cull& c);
// Records a failure message for rsx_last_error and returns the status.
int32_t record_failure(int32_t status, const char* msg);
// Looks up the device address of rsx_gae_noalloc's fixed scratch on `dev`
// (the current device), once per device.
int32_t gae_noalloc_setup(int dev);

}  // namespace rsx
