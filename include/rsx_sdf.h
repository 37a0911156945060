/*
 * rsx_sdf.h — C-ABI of the learned broadphase cull (librsx.so, gfx950).
 *
 * A small MLP (3 -> 128 -> 128 -> 128 -> 1, SiLU, BF16 weights, FP32
 * accumulation) predicts the unsigned distance from a point to the soccar
 * collision mesh.  It is trained on the GPU by the library's own forward,
 * backward and Adam kernels on distances from the exact mesh query below, and
 * attached to an engine as a conservative cull: a body whose predicted
 * distance minus the margin exceeds its reach skips the mesh queries of the
 * tick; the exact triangle tests run for every other body.
 *
 * The reference has no learned component (its broadphase is the uniform grid
 * of RocketSimCuda/src/RocketSimCudaMesh.cuh:9-125 and Bullet's BVH in the CPU
 * path); these entry points have no counterpart there.  They serve the
 * learned-cull workload and its audit.
 *
 * Points and distances are in Unreal units.  Device pointers ("d_") are
 * hipMalloc'd memory on the engine's device; the stream argument is a
 * hipStream_t (NULL: the default stream).
 */
#ifndef RSX_SDF_H
#define RSX_SDF_H

#include <stdint.h>

#include "rsx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsx_sdf rsx_sdf;

/* Counters of the cull since the last reset. */
typedef struct rsx_cull_stats {
    uint64_t queries;  /* body queries that reached the network */
    uint64_t culled;   /* body-ticks dropped from the mesh queries */
    uint64_t defects;  /* audited culled body-ticks whose exact distance was within the reach */
} rsx_cull_stats;

/* Number of FP32 parameters (33665) and their order:
 * W1[128][3], b1[128], W2[128][128] (out-major), b2, W3[128][128], b3, w4[128], b4. */
int32_t rsx_sdf_num_params(void);

/* Network for the engine's mesh and device, weights drawn from `seed`. */
rsx_status rsx_sdf_create(rsx_engine* e, uint64_t seed, rsx_sdf** out);
rsx_status rsx_sdf_destroy(rsx_sdf* s);

/* Mesh triangles (9 floats each, uu) into out (cap floats); *n = triangles. */
rsx_status rsx_sdf_mesh_triangles(rsx_sdf* s, float* out, int64_t cap, int32_t* n);

/* Exact unsigned distance from each point to the mesh triangles. */
rsx_status rsx_sdf_exact_distance(rsx_sdf* s, const float* d_points, int32_t n, float* d_dist, void* stream);

/* Network prediction for each point (BF16 matrix-core path of the tick). */
rsx_status rsx_sdf_eval(rsx_sdf* s, const float* d_points, int32_t n, float* d_dist, void* stream);

/* One Adam step on the batch: L1 loss between the prediction and
 * min(label, clamp_uu), both in units of 100 uu, with over-estimates weighed
 * over_weight (> 1 pushes the network below the true distance, which is the
 * side the cull's safety margin pays for).  d_loss (nullable) receives the
 * batch's mean loss. */
rsx_status rsx_sdf_train_step(rsx_sdf* s, const float* d_points, const float* d_labels, int32_t n, float lr,
                              float clamp_uu, float over_weight, float* d_loss, void* stream);

/* Loss and gradients of the batch (host arrays, rsx_sdf_num_params floats),
 * without an update; for checking the backward kernels. */
rsx_status rsx_sdf_gradients(rsx_sdf* s, const float* d_points, const float* d_labels, int32_t n, float clamp_uu,
                             float over_weight, float* h_grads, float* h_loss);

/* FP32 master parameters, host arrays of rsx_sdf_num_params floats. */
rsx_status rsx_sdf_get_params(rsx_sdf* s, float* h_params);
rsx_status rsx_sdf_set_params(rsx_sdf* s, const float* h_params);

/* Attach the network to the engine's tick as a cull with the given safety
 * margins (uu) for the ball and for cars.  The network runs every
 * `every_ticks` ticks (>= 1); between runs a body stays culled only while its
 * prediction still exceeds its margin + reach + the distance it can have
 * moved.  audit != 0 also checks every culled body against the exact mesh
 * distance each tick and counts defects.  s == NULL detaches (exact
 * collision only).  The network must outlive its attachment. */
rsx_status rsx_attach_cull(rsx_engine* e, rsx_sdf* s, float margin_ball_uu, float margin_car_uu, int32_t every_ticks,
                           int32_t audit);
rsx_status rsx_get_cull_stats(rsx_sdf* s, rsx_cull_stats* out);
rsx_status rsx_reset_cull_stats(rsx_sdf* s);

/* The reaches the cull compares against (uu): ball, car (box and wheel rays). */
rsx_status rsx_cull_reach(rsx_sdf* s, float* ball_uu, float* car_uu);

#ifdef __cplusplus
}
#endif

#endif /* RSX_SDF_H */
