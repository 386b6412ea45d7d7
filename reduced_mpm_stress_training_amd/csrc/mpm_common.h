// mpm_common.h -- pieces shared by the stage kernels and the binned kernels:
// host-side error reporting and argument checks, and the per-node grid
// update (grid_normalization_and_gravity + colliders).
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_hip.h"

namespace mpm {

// host side (mpm_kernels.hip)
int set_err(int code, const char* msg);
int check_launch(const char* what);
int check_grid(const MpmParams* prm);
int check_params(const MpmParams* prm);
int check_sand(const MpmParams* prm);  // check_params and material 2

struct ColliderSet {
    MpmCollider c[MPM_MAX_COLLIDERS];
    int n;
};
int make_collider_set(const MpmCollider* cols, int n, ColliderSet* cs);
// mpm_grid_update that also clears `zero_next` (a float4 grid of the same
// size, or NULL): the accumulator the next substep's P2G scatters into
int grid_update_zero_next(const MpmParams* prm, const float* grid_acc, float* grid_v, const MpmCollider* colliders,
                          int n_colliders, float time, float dt, float* zero_next, void* stream);

// grid_normalization_and_gravity (mpm_utils.py:229-237) then each active
// collider in registration order (mpm_solver_warp.py:182-185):
//   surface collider (mpm_solver_warp.py:429-449): v = 0 where
//     (x_node - point) . n < 0, for every surface type (the reference writes
//     0 after computing the slip/friction projection, :449);
//   bounding box (mpm_solver_warp.py:505-523): padding 3, outward components
//     zeroed axis by axis.
// Returns (v, m); v = 0 where m <= 1e-15.
__device__ __forceinline__ float4 grid_node_update(float4 acc, int gx, int gy, int gz, int ng, float dx,
                                                   float dt, const float grav[3], const MpmCollider* cols,
                                                   int n_cols, float time) {
    float4 o = make_float4(0.f, 0.f, 0.f, acc.w);
    if (acc.w > 1e-15f) {
        float im = 1.0f / acc.w;
        o.x = acc.x * im + dt * grav[0];
        o.y = acc.y * im + dt * grav[1];
        o.z = acc.z * im + dt * grav[2];
    }
    for (int c = 0; c < n_cols; ++c) {
        const MpmCollider& cl = cols[c];
        if (!(time >= cl.start_time && time < cl.end_time)) continue;
        if (cl.kind == 0) {
            float o0 = (float)gx * dx - cl.point[0];
            float o1 = (float)gy * dx - cl.point[1];
            float o2 = (float)gz * dx - cl.point[2];
            float dot = o0 * cl.normal[0] + o1 * cl.normal[1] + o2 * cl.normal[2];
            if (dot < 0.0f) { o.x = 0.f; o.y = 0.f; o.z = 0.f; }
        } else {
            const int pad = 3;
            if (gx < pad && o.x < 0.f) o.x = 0.f;
            if (gx >= ng - pad && o.x > 0.f) o.x = 0.f;
            if (gy < pad && o.y < 0.f) o.y = 0.f;
            if (gy >= ng - pad && o.y > 0.f) o.y = 0.f;
            if (gz < pad && o.z < 0.f) o.z = 0.f;
            if (gz >= ng - pad && o.z > 0.f) o.z = 0.f;
        }
    }
    return o;
}

__device__ __forceinline__ float4 grid_node_update(float4 acc, int gx, int gy, int gz, int ng, float dx,
                                                   float dt, const float grav[3], const ColliderSet& cs,
                                                   float time) {
    return grid_node_update(acc, gx, gy, gz, ng, dx, dt, grav, cs.c, cs.n, time);
}

}  // namespace mpm
