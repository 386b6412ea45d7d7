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
int check_state(const MpmState* st);

struct ColliderSet {
    MpmCollider c[MPM_MAX_COLLIDERS];
    int n;
};
int make_collider_set(const MpmCollider* cols, int n, ColliderSet* cs);
// mpm_grid_update that also clears `zero_next` (a float4 grid of the same
// size, or NULL): the accumulator the next substep's P2G scatters into
int grid_update_zero_next(const MpmParams* prm, const float* grid_acc, float* grid_v, const MpmCollider* colliders,
                          int n_colliders, float time, float dt, float* zero_next, void* stream);

// One collider (a grid post-process kernel of the reference, the MpmCollider
// kinds of include/mpm_hip.h) applied to the node velocity o of node
// (gx, gy, gz).  `time` is the solver time before the substep's increment.
// mask[d] is cleared for every component the collider replaced by a value
// that does not depend on the grid (the adjoint uses it).
__device__ __forceinline__ void apply_collider(const MpmCollider& cl, int gx, int gy, int gz, int ng, float dx,
                                               float dt, float time, float o[3], float mask[3]) {
    const bool active = time >= cl.start_time && time < cl.end_time;
    const float o0 = (float)gx * dx - cl.point[0];
    const float o1 = (float)gy * dx - cl.point[1];
    const float o2 = (float)gz * dx - cl.point[2];
    auto set = [&](float a, float b, float c) {
        o[0] = a; o[1] = b; o[2] = c;
        mask[0] = mask[1] = mask[2] = 0.f;
    };
    switch (cl.kind) {
    case 0: {  // surface collider, every surface type writes 0 (mpm_solver_warp.py:432-449)
        if (!active) break;
        const float dot = o0 * cl.normal[0] + o1 * cl.normal[1] + o2 * cl.normal[2];
        if (dot < 0.0f) set(0.f, 0.f, 0.f);
        break;
    }
    case 1: {  // bounding box, padding 3 (mpm_solver_warp.py:508-523)
        if (!active) break;
        const int pad = 3;
        const int gi[3] = {gx, gy, gz};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if ((gi[d] < pad && o[d] < 0.f) || (gi[d] >= ng - pad && o[d] > 0.f)) {
                o[d] = 0.f;
                mask[d] = 0.f;
            }
        }
        break;
    }
    case 2: {  // velocity on an infinite plane (mpm_solver_warp.py:546-558)
        const float dot = o0 * cl.normal[0] + o1 * cl.normal[1] + o2 * cl.normal[2];
        if (active) {
            if (fabsf(dot) <= cl.threshold) set(cl.velocity[0], cl.velocity[1], cl.velocity[2]);
        } else if (cl.reset == 1) {
            if (time < cl.end_time + 1.5f * dt * 10.0f && fabsf(dot) <= cl.threshold) set(0.f, 0.f, 0.f);
        }
        break;
    }
    case 3: {  // column moving on a circle (mpm_solver_warp.py:581-592)
        if (!active) break;
        if (o0 * o0 + o2 * o2 < cl.radius * cl.radius && fabsf(o1) < cl.height) {
            const float z0 = cl.point[0] - 0.5f, z1 = cl.point[1], z2 = cl.point[2] - 0.5f;
            const float dp = z0 / sqrtf(z0 * z0 + z1 * z1 + z2 * z2);
            float th = acosf(dp);
            if (z2 < 0.f) th = -th;
            set(-cl.v_scale * sinf(th), 0.f, cl.v_scale * cosf(th));
        }
        break;
    }
    case 4: {  // rotating disk (mpm_solver_warp.py:638-661)
        const float dot = o0 * cl.normal[0] + o1 * cl.normal[1] + o2 * cl.normal[2];
        if (active) {
            if (fabsf(dot) < cl.width) {
                // r = (0, dy, dz) as the reference takes it (not the in-plane part)
                const float rl = sqrtf(o1 * o1 + o2 * o2);
                if (rl < cl.radius) {
                    if (rl > 0.f) {
                        const float dp = (o1 * cl.x_unit[1] + o2 * cl.x_unit[2]) / rl;
                        const float rdoty = o1 * cl.y_unit[1] + o2 * cl.y_unit[2];
                        float th = acosf(dp);
                        if (rdoty < 0.f) th = -th;
                        const float sx = -rl * sinf(th) * cl.v_scale * 1.0f;
                        const float sy = rl * cosf(th) * cl.v_scale * 1.0f;
                        set(sx * cl.x_unit[0] + sy * cl.y_unit[0], sx * cl.x_unit[1] + sy * cl.y_unit[1],
                            sx * cl.x_unit[2] + sy * cl.y_unit[2]);
                    } else {
                        set(0.f, 0.f, 0.f);   // on the axis (the reference divides 0 by 0 here)
                    }
                }
            }
        } else if (cl.reset == 1) {
            if (time < cl.end_time + 1.5f * dt) set(0.f, 0.f, 0.f);
        }
        break;
    }
    case 5: {  // velocity on a cube (mpm_solver_warp.py:681-687)
        if (active) {
            if (fabsf(o0) < cl.width && fabsf(o1) < cl.width && fabsf(o2) < cl.width)
                set(cl.velocity[0], cl.velocity[1], cl.velocity[2]);
        } else if (cl.reset == 1) {
            if (time < cl.end_time + 1.5f * dt) set(0.f, 0.f, 0.f);
        }
        break;
    }
    case 6: {  // sphere leader (mpm_solver_warp.py:710-713)
        if (!active) break;
        if (sqrtf(o0 * o0 + o1 * o1 + o2 * o2) < cl.radius) set(cl.velocity[0], cl.velocity[1], cl.velocity[2]);
        break;
    }
    default:
        break;
    }
}

// grid_normalization_and_gravity (mpm_utils.py:229-237) then each collider
// in registration order (mpm_solver_warp.py:182-185, apply_collider).
// Returns (v, m); v = 0 where m <= 1e-15 unless a collider sets it.
__device__ __forceinline__ float4 grid_node_update(float4 acc, int gx, int gy, int gz, int ng, float dx,
                                                   float dt, const float grav[3], const MpmCollider* cols,
                                                   int n_cols, float time) {
    float o[3] = {0.f, 0.f, 0.f};
    if (acc.w > 1e-15f) {
        float im = 1.0f / acc.w;
        o[0] = acc.x * im + dt * grav[0];
        o[1] = acc.y * im + dt * grav[1];
        o[2] = acc.z * im + dt * grav[2];
    }
    float mask[3];
    for (int c = 0; c < n_cols; ++c) apply_collider(cols[c], gx, gy, gz, ng, dx, dt, time, o, mask);
    return make_float4(o[0], o[1], o[2], acc.w);
}

__device__ __forceinline__ float4 grid_node_update(float4 acc, int gx, int gy, int gz, int ng, float dx,
                                                   float dt, const float grav[3], const ColliderSet& cs,
                                                   float time) {
    return grid_node_update(acc, gx, gy, gz, ng, dx, dt, grav, cs.c, cs.n, time);
}

}  // namespace mpm
