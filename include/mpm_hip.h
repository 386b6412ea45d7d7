/*
 * mpm_hip.h -- C-ABI of libmpm_hip.so, the MI355X (gfx950) explicit
 * MLS/APIC-MPM sand substep.
 *
 * The reference (zeshunzong/reduced_mpm_stress_training) is Python + Warp:
 * its "FFI" is `wp.launch(kernel, dim, inputs=[MPMStateStruct,
 * MPMModelStruct, dt])` (engine_stress/mpm_solver_warp.py:156-203).  Each
 * entry point below replaces one of those launches (or a fused run of them)
 * and cites it.  The Python solver class in
 * reduced_mpm_stress_training_amd/mpm_solver.py binds them with ctypes.
 *
 * Conventions
 *  - All pointers are device pointers into caller-owned memory (torch
 *    tensors); the library never allocates behind the caller's back except
 *    where an entry point says so.
 *  - Every call is stream-ordered on `stream` (a hipStream_t; NULL = the
 *    legacy default stream) and does not synchronise the host.
 *  - Return value: 0 = OK, negative = error; mpm_last_error() gives the
 *    message of the last failing call on the calling host thread.
 *
 * Particle state layout ("planes"): one float32 buffer of MPM_NPLANES planes,
 * plane k starting at planes + k*stride (stride >= n, in elements), indexed by
 * solver slot.  Plane order: x0(3) u(3) v(3) C(9) F(9) tau(9) mass vol
 * yield_stress pid, with 3x3 matrices row-major (component a*3+b) like Warp
 * mat33 and torch (N,3,3).  The pid plane holds the caller's particle index of
 * each slot as int32 bits.  yield_stress is the reference's per-particle
 * model.yield_stress array (warp_utils.py:75, mpm_solver_warp.py:43); only
 * metal (material 1) reads or writes it.
 *
 * Grid layout: n_grid^3 nodes in C order [ix][iy][iz] (as the reference's
 * wp.array(ndim=3)), node i at position i*dx, one float4 per node:
 *   accumulator  (mv.x, mv.y, mv.z, m)   -- grid_v_in and grid_m
 *   velocity     (v.x,  v.y,  v.z,  m)   -- grid_v_out
 */
#ifndef MPM_HIP_H
#define MPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 15

/* plane indices */
#define MPM_P_X0 0
#define MPM_P_U 3
#define MPM_P_V 6
#define MPM_P_C 9
#define MPM_P_F 18
#define MPM_P_TAU 27
#define MPM_P_MASS 36
#define MPM_P_VOL 37
#define MPM_P_YIELD 38
#define MPM_P_PID 39
#define MPM_NPLANES 40

#define MPM_MAX_COLLIDERS 16

/* Model scalars: MPMModelStruct (engine_stress/warp_utils.py:61-81).
 * gravity generalises the reference's (0, -g, 0) (mpm_utils.py:236).
 * material (the G2P switch of mpm_utils.py:284-304):
 *   0 jelly: F = F_trial, FCR stress (mpm_utils.py:29-34, 298-299);
 *   1 metal: von Mises return map on the yield_stress plane, hardening
 *            when hardening == 1 (mpm_utils.py:59-84), StVK stress
 *            (mpm_utils.py:36-43, 301-302);
 *   2 sand:  Drucker-Prager return map (mpm_utils.py:87-112), stress
 *            (mpm_utils.py:45-52).
 * The adjoint and ordered-state entry points implement sand only. */
typedef struct MpmParams {
    int32_t n_grid;
    int32_t material;
    float dx, inv_dx;
    float mu, lam, alpha;
    float gravity[3];
    int32_t hardening; /* metal: 1 = yield_stress += 2 mu xi dgamma (mpm_utils.py:80-81) */
    float xi;
} MpmParams;

/* Dirichlet_collider (engine_stress/warp_utils.py:126-151).  Each kind is
 * one of the reference's grid post-process kernels, applied to a node's v_out
 * in registration order; "active" means start_time <= time < end_time with
 * the time before the substep's increment (mpm_solver_warp.py:183, 199).
 * kind 0: add_surface_collider / add_surface_collider2 (mpm_solver_warp.py:
 *         407-497) -- every surface type writes v = 0 where
 *         (x_node - point) . normal < 0, as the reference does (:449, :495);
 * kind 1: add_bounding_box (:499-525), padding 3;
 * kind 2: set_velocity_on_infinite_plane (:527-566): v = velocity where
 *         |(x_node - point) . normal| <= threshold; when inactive and reset,
 *         v = 0 there while time < end_time + 15 dt (:552-558);
 * kind 3: set_column_moving_as_circle (:568-605): nodes with
 *         dx^2 + dz^2 < radius^2 and |dy| < height get
 *         v_scale (-sin t, 0, cos t), cos t = Z.x / |Z|, Z = point - (0.5, 0, 0.5);
 * kind 4: set_rotation_on_disk (:607-664): nodes with |offset . normal| <
 *         width and |r| < radius, r = (0, dy, dz), get the rotation
 *         v_scale (-(r.y_unit) x_unit + (r.x_unit) y_unit); when inactive
 *         and reset, every node gets 0 while time < end_time + 1.5 dt;
 * kind 5: set_velocity_on_cube (:666-693): v = velocity where every
 *         |offset_d| < width; inactive + reset: 0 everywhere while
 *         time < end_time + 1.5 dt;
 * kind 6: add_sphere_leader (:695-720): v = velocity where |offset| < radius.
 * The moving kinds (2, 3, 5, 6) advance `point` on the host between substeps
 * (the reference's `modify` callbacks); the library reads it as given. */
typedef struct MpmCollider {
    int32_t kind;
    int32_t surface_type;
    float point[3];
    float normal[3];
    float start_time, end_time;
    float friction;
    float velocity[3];
    float threshold, radius, height, width, v_scale;
    float x_unit[3], y_unit[3];
    int32_t reset;
} MpmCollider;

/* Particle state: MPMStateStruct particle arrays (warp_utils.py:84-124). */
typedef struct MpmState {
    int64_t n;      /* particles */
    int64_t stride; /* plane stride in elements */
    float* planes;  /* MPM_NPLANES planes */
} MpmState;

int mpm_abi_version(void);
const char* mpm_last_error(void);

/* ---- stage entry points (one reference launch each) ------------------- */

/* zero_grid (mpm_utils.py:115-123): clears n_grid^3 float4 nodes. */
int mpm_zero_grid(const MpmParams* prm, float* grid, void* stream);

/* update_x_from_u (mpm_utils.py:4-7) fused into p2g_apic_stress_given
 * (mpm_utils.py:190-227): scatters m*(v + C dpos) + dt*(-vol tau grad w) and
 * w*m of every particle into the accumulator grid (float atomics). */
int mpm_p2g(const MpmParams* prm, const MpmState* st, float* grid_acc, float dt,
            void* stream);

/* grid_normalization_and_gravity (mpm_utils.py:229-237) followed by the
 * colliders in registration order (mpm_solver_warp.py:182-185); `time` is the
 * solver time before the step's increment (mpm_solver_warp.py:183,199).
 * grid_v receives (v, m); nodes with m <= 1e-15 get v = 0. */
int mpm_grid_update(const MpmParams* prm, const float* grid_acc, float* grid_v,
                    const MpmCollider* colliders, int n_colliders, float time,
                    float dt, void* stream);

/* g2p_apic_and_stress (mpm_utils.py:241-306) with sand_return_mapping
 * (:87-112) and kirchoff_stress_drucker_prager (:45-52), fused with
 * update_x_from_u (:4-7) and update_u_from_x (:12-15): writes u, v, C, F,
 * tau of every particle. */
int mpm_g2p(const MpmParams* prm, const MpmState* st, const float* grid_v, float dt,
            void* stream);

/* ---- whole substep ----------------------------------------------------- */

/* Bytes of device workspace mpm_step needs for this grid and particle count. */
int64_t mpm_workspace_bytes(const MpmParams* prm, int64_t n_particles);

/* One substep, MPM_Simulator_WARP.p2g2p (mpm_solver_warp.py:156-203):
 * x = x0 + u, zero grid, P2G, grid update + colliders, G2P, u = x - x0.
 * workspace: mpm_workspace_bytes() bytes, 256-byte aligned. */
int mpm_step(const MpmParams* prm, const MpmState* st, const MpmCollider* colliders,
             int n_colliders, void* workspace, int64_t workspace_bytes, float time,
             float dt, void* stream);

/* ---- no-allocation substep ------------------------------------------------
 *
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  mpm_step_noalloc is mpm_step with its workspace (the
 * accumulator and velocity grids) in a static device array inside the
 * library, sized for n_grid <= mpm_noalloc_max_grid() (64: 2 x 64^3 float4,
 * 8 MiB per GPU, held until the process ends).  Larger grids are rejected
 * with an error (use mpm_step).  It allocates nothing.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int mpm_noalloc_max_grid(void);
int mpm_step_noalloc(const MpmParams* prm, const MpmState* st, const MpmCollider* colliders, int n_colliders,
                     float time, float dt, void* stream);

/* ---- reduced-order sampled substep (CROM time stepper) ------------------
 *
 * MPM_Simulator_WARP.p2g2p_reduced (mpm_solver_warp.py:213-227) and the
 * sample G2P its caller launches after it, on the flat state.  `samples`
 * holds n_samples particle indices (int32, the reference's
 * mpm_state.sample_particle_index_short); `density` is the per-particle
 * density (n floats, mpm_state.particle_density).  workspace:
 * mpm_reduced_workspace_bytes() bytes, 256-byte aligned, laid out as the
 * accumulator grid (mv, m), the velocity grid (v, m) (float4 per node, C
 * order as the stage grids), then the int32 grids neighboring_cells and
 * neighboring_cells_grid_v (n_grid^3 each). */
int64_t mpm_reduced_workspace_bytes(const MpmParams* prm);

/* p2g2p_reduced (mpm_solver_warp.py:214-227) without the time increment:
 * update_x_from_u, zero_grid (both grids and flag grids), update_neighboring_cells
 * (mpm_utils.py:318-335: cells base-2 .. base+2 of every sample),
 * p2g_sample_all (:356-410: particles in flagged cells set F = F_disp + I and
 * scatter PIC momentum, mass and dt (-(m/rho) tau grad w), tau from F: FCR
 * for jelly and metal, Drucker-Prager for sand), grid_normalization_and_
 * gravity_sample_all (:413-421) and the colliders on every node. */
int mpm_step_reduced(const MpmParams* prm, const MpmState* st, const int* samples, int64_t n_samples,
                     const float* density, const MpmCollider* colliders, int n_colliders, void* workspace,
                     int64_t workspace_bytes, float time, float dt, void* stream);

/* g2p_sample (mpm_utils.py:423-471) on the velocity grid of the last
 * mpm_step_reduced: each sample particle gets v, x += dt v, F from the
 * material's return map of (I + dt grad v) F, u = x - x0.  C and tau are not
 * written.  Metal reads and hardens yield_stress[s] of the sample's slot s
 * in `samples`, as the reference does (:464). */
int mpm_g2p_sample(const MpmParams* prm, const MpmState* st, const int* samples, int64_t n_samples,
                   const void* workspace, float dt, void* stream);

/* update_neighboring_cells_for_grid_v (mpm_utils.py:337-354), behind the
 * reference's experiment_for_C (mpm_solver_warp.py:207-209): flags cells
 * base-2 .. base+4 of every sample in the workspace's neighboring_cells_grid_v
 * (cleared by mpm_step_reduced). */
int mpm_mark_cells_grid_v(const MpmParams* prm, const MpmState* st, const int* samples, int64_t n_samples,
                          void* workspace, void* stream);

/* ---- binned path: the solver's own particle layout --------------------
 *
 * Particles are kept in bins of 8^3 grid cells keyed by the base cell of
 * their B-spline stencil (int(x*inv_dx - 0.5), mpm_utils.py:197-199).  A
 * plane set is bin-major: bin b is a block of MPM_NPLANES planes of bin_cap
 * floats (plane order as MpmState), slot i of bin b / plane p at
 * (b*MPM_NPLANES + p)*bin_cap + i; bin b holds count[b] particles (slots
 * with pid -1 are holes left by particles that moved).  The blocks after the
 * nb^3 bins form an overflow run of overflow_cap slots for bins that are
 * full.  Two plane sets alternate between substeps.
 * A substep is the grid update of this substep, then one fused kernel doing
 * G2P of this substep and P2G of the next.  The P2G lands in per-bin tiles
 * of the workspace stamped with `epoch`, so between calls the workspace
 * carries the pending P2G; mpm_bins_p2g computes it for a freshly inserted
 * state.  Entry points that advance the state update `cur`, `epoch` and
 * `since_sort` in place.  Every sort_period substeps mpm_bins_substep
 * re-sorts the bins (mpm_bins_sort): the order of particles inside a bin
 * decides how many LDS bank conflicts the fused kernel's gathers and
 * scatters see.  When one bin holds more than a CU's share of all
 * particles, the fused kernel splits it into up to four slot ranges, each
 * scattering into its own FP64 tile that the grid update adds before the one
 * rounding to FP32.  wg_threads, sort_period and the splitting change speed,
 * not results beyond the order of FP64 tile sums
 * (reduced_mpm_stress_training_amd/tune.py picks the first two by timing). */
typedef struct MpmBins {
    int32_t nb;           /* bins per axis = ceil(n_grid / 8) */
    int32_t bin_cap;      /* slots per bin */
    int32_t overflow_cap; /* slots of the overflow run, a multiple of bin_cap */
    int32_t cur;          /* plane set holding the current state (0 or 1) */
    int32_t epoch;        /* stamp of the pending P2G tiles; start at 0 */
    int32_t since_sort;   /* substeps since the last bank-order sort; start at 0 */
    int64_t stride;       /* floats per plane set >= (nb^3*bin_cap + overflow_cap) * MPM_NPLANES */
    float* planes[2];     /* two sets of MPM_NPLANES planes */
    int32_t* count;       /* [2][2*nb^3 + 1]: per plane set, slots used per bin (nb^3),
                             slots used in the overflow run (1), live particles per bin (nb^3) */
    int32_t wg_threads;   /* threads per bin workgroup of the fused kernel: 256, 384, 512 or 768 (0 = 768) */
    int32_t sort_period;  /* substeps between bank-order sorts (0 = 16, < 0 = never) */
    float* record_acc;    /* NULL, or an n_grid^3 float4 grid that each grid update fills with the
                             substep's summed P2G (mv, m) -- the reference's grid_v_in and grid_m after
                             p2g2p (mpm_solver_warp.py:162); single scene only */
} MpmBins;

/* Workspace of the binned path (zeroed once before first use, 256-byte
 * aligned): the spill grid (n_grid^3 x 4 doubles, FP64 atomics), the
 * velocity grid (n_grid^3 float4: v, m after the grid update), spill flags,
 * 64 int32 status counters (0: particles dropped because bins and overflow
 * were full, 1: particles scattered outside their bin's tile, 2: particles
 * that left the grid (held at their last position), 3: grid nodes whose
 * P2G sum was NaN or Inf (a particle's state was not finite); all four
 * accumulate until the caller clears them; 48-49: the start time, as a double, of a captured run,
 * mpm_batch_graph_launch), tile stamps, the lists of grid blocks each epoch's
 * P2G wrote (the grid update visits only those) and per block the epoch of
 * its last velocity update, the fused kernel's schedule and nb^3 P2G tiles
 * of 10^3 float4 (the parts of a split bin add their sums into the spill
 * grid). */
int64_t mpm_bins_workspace_bytes(const MpmParams* prm);

/* Byte offset of the status counters in the workspace. */
int64_t mpm_bins_status_offset(const MpmParams* prm);

/* Byte offset of the velocity grid (n_grid^3 float4: v after the grid
 * update, m) in the workspace: the reference's grid_v_out and grid_m
 * (mpm_solver_warp.py:65-67) of the last substep once mpm_bins_settle_grid
 * has run.  The grid update writes only the 8^3 blocks that hold P2G sums;
 * blocks the particles have left keep their last values until then. */
int64_t mpm_bins_velocity_offset(const MpmParams* prm);

/* Zeroes the velocity grid's blocks that the last grid update did not write
 * (zero_grid, mpm_utils.py:115-123, clears every node each substep; the
 * substeps here never read those nodes): call before reading the velocity
 * grid as the reference's grid_v_out / grid_m.  No effect on the particles. */
int mpm_bins_settle_grid(const MpmParams* prm, const MpmBins* bins, void* workspace, void* stream);

/* Clears count[cur] and inserts every particle of a flat state (slot =
 * caller index, as MpmState) into the bins of plane set `cur`. */
int mpm_bins_insert(const MpmParams* prm, const MpmState* flat, const MpmBins* bins, void* workspace,
                    void* stream);

/* flat[pid] <- every particle of plane set `cur` (all planes): an inverse
 * slot map (8 bytes per particle, written into the other plane set, whose
 * contents are dead between the bins' operations), then a gather per
 * particle with coalesced stores into flat. */
int mpm_bins_extract(const MpmParams* prm, const MpmBins* bins, const MpmState* flat, void* stream);

/* mpm_bins_extract of the planes whose bits are set in plane_mask (bit k =
 * plane k): what export_uvFstress_from_warp (mpm_solver_warp.py:305-310) or a
 * snapshot needs, without moving the other planes. */
int mpm_bins_extract_planes(const MpmParams* prm, const MpmBins* bins, const MpmState* flat, uint64_t plane_mask,
                            void* stream);

/* P2G (p2g_apic_stress_given, mpm_utils.py:190-227) of plane set `cur` into
 * the tiles of epoch+1, with the given dt; advances epoch. */
int mpm_bins_p2g(const MpmParams* prm, MpmBins* bins, void* workspace, float dt, void* stream);

/* Discards a pending P2G (drains the spill grid, advances epoch). */
int mpm_bins_clear_grid(const MpmParams* prm, MpmBins* bins, void* workspace, void* stream);

/* The two launches of one substep, for per-kernel timing: the grid update
 * (grid_normalization_and_gravity + colliders, mpm_utils.py:229-237 and
 * mpm_solver_warp.py:429-449, 505-523, with `time` the solver time before
 * the increment), which also resets the bin counters of plane set 1-cur; and
 * the fused G2P (g2p_apic_and_stress with the sand return map and stress,
 * mpm_utils.py:241-306) + next-substep P2G, which writes plane set 1-cur,
 * then sets cur = 1-cur and advances epoch. */
int mpm_bins_grid_update(const MpmParams* prm, const MpmBins* bins, const MpmCollider* colliders,
                         int n_colliders, void* workspace, float time, float dt, void* stream);
int mpm_bins_g2p2g(const MpmParams* prm, MpmBins* bins, void* workspace, float dt, void* stream);

/* Copies plane set cur into 1-cur in bank order (each bin counting-sorted
 * by the LDS bank class of its particles' stencil base, holes dropped),
 * zeroes the counters of the old set, then sets cur = 1-cur and
 * since_sort = 0.  No effect on the physics; may sit between
 * mpm_bins_grid_update and mpm_bins_g2p2g. */
int mpm_bins_sort(const MpmParams* prm, MpmBins* bins, void* workspace, void* stream);

/* One substep, MPM_Simulator_WARP.p2g2p (mpm_solver_warp.py:156-203), on the
 * binned state: mpm_bins_grid_update, mpm_bins_sort when since_sort reaches
 * 32, then mpm_bins_g2p2g. */
int mpm_bins_substep(const MpmParams* prm, MpmBins* bins, const MpmCollider* colliders, int n_colliders,
                     void* workspace, float time, float dt, void* stream);

/* n_substeps substeps like mpm_bins_run, each kernel bracketed by HIP events
 * on `stream`; waits for each substep and returns the average duration per
 * substep of the grid update, of the fused kernel and of the periodic sort
 * in milliseconds (benchmarking). */
int mpm_bins_profile(const MpmParams* prm, MpmBins* bins, const MpmCollider* colliders, int n_colliders,
                     void* workspace, double time, float dt, int n_substeps, void* stream, float* grid_ms,
                     float* g2p2g_ms, float* sort_ms);

/* n_substeps consecutive mpm_bins_substep calls; the solver time of substep
 * k is time + k*dt accumulated in double, as the Python solver accumulates
 * self.time (mpm_solver_warp.py:199). */
int mpm_bins_run(const MpmParams* prm, MpmBins* bins, const MpmCollider* colliders, int n_colliders,
                 void* workspace, double time, float dt, int n_substeps, void* stream);

/* ---- adjoint of a substep -------------------------------------------------
 *
 * No counterpart in the reference (it never records a wp.Tape).  The
 * reverse-mode derivative of one substep as mpm_step computes it (flat
 * state): given the adjoint of the state after the substep, the adjoint of
 * the state before it, and the material parameters' adjoints accumulated.
 * Adjoint planes (adj_stride floats each, particle order as MpmState):
 * x (3), v (3), C (9), F (9), tau (9) -- the 33 MPM_ADJ_* planes.  The
 * weights' base cells and the return-map branch are held fixed. */
#define MPM_ADJ_X 0
#define MPM_ADJ_V 3
#define MPM_ADJ_C 6
#define MPM_ADJ_F 15
#define MPM_ADJ_TAU 24
#define MPM_ADJ_NPLANES 33

/* A grid set of the ordered segment path (256-byte aligned): the P2G
 * accumulator and the grid velocities (n_grid^3 float4 each), then one int32
 * per 8^3 grid block, nonzero where a P2G into the accumulator wrote.  The
 * grid update, the next substep's clearing and the grid adjoint visit only
 * marked blocks.  A grid set's accumulator must be zero outside its marked
 * blocks: allocate it zeroed, and afterwards leave it to these calls. */
int64_t mpm_step_grids_bytes(const MpmParams* prm);

/* Workspace of mpm_adjoint_substep (256-byte aligned): a grid set, then two
 * grids of n_grid^3 float4 (the velocity adjoint and the (mv, m) adjoint). */
int64_t mpm_adjoint_workspace_bytes(const MpmParams* prm);

/* st: the state S_n before the substep, in the order of `order` (see
 * mpm_sort_by_bin); time, dt: that substep's.  step_grids: the workspace
 * mpm_step_ordered used for this substep (its P2G accumulator and grid
 * velocities, untouched since), or NULL to recompute them here the same way.
 * adj_next: adjoint of S_{n+1}; adj_cur: receives the adjoint of S_n (may
 * not alias adj_next).  param_adj[0..2] += adjoints of mu, lam, alpha
 * (double). */
int mpm_adjoint_substep(const MpmParams* prm, const MpmState* st, const void* order, const void* step_grids,
                        const MpmCollider* colliders, int n_colliders, void* workspace, float time, float dt,
                        const float* adj_next, float* adj_cur, int64_t adj_stride, double* param_adj, void* stream);

/* The adjoint of the n_sub substeps k0 .. k0+n_sub-1 of one ordered segment
 * in one call: replays the segment from its first state (the substep of
 * mpm_run_ordered, each substep's P2G accumulator and grid velocities kept in
 * its own workspace), then applies the adjoint substeps in reverse on those
 * grids.  states: n_sub states sharing n and stride; states[0] holds S_k0 on
 * entry, states[1..] receive S_{k0+1} .. S_{k0+n_sub-1} -- the planes a
 * substep changes (u, v, C, F, tau); their constant planes (x0, m, vol, id)
 * are left as they are, those of states[0] being used.  step_grids: n_sub
 * grid sets of mpm_step_grids_bytes (256-byte aligned, zeroed before their
 * first use).  Substep k runs at
 * time (float)(k * dt) with step (float)dt.  adj_a holds the adjoint of
 * S_{k0+n_sub} on entry and the adjoint of S_k0 on return; adj_b is scratch of
 * the same size.  workspace: mpm_adjoint_workspace_bytes. */
int mpm_adjoint_segment(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                        const MpmCollider* colliders, int n_colliders, void* workspace, int k0, double dt, int n_sub,
                        float* adj_a, float* adj_b, int64_t adj_stride, double* param_adj, void* stream);

/* The two halves of mpm_adjoint_segment as separate calls, for callers that
 * keep every substep's state and grids from the forward pass (no replay on
 * the way back; 288 GB of HBM holds all 200 substeps of a 1M-particle rollout).
 * mpm_record_segment: the forward substeps k0 .. k0+n_sub-1 of an ordered
 * segment, states and step_grids as in mpm_adjoint_segment.  p2g_done != 0:
 * step_grids[0] already holds the P2G accumulator of states[0] (the
 * next_grid of the segment before, whose last state this segment sorted), so
 * it is not recomputed.  When `last` is not NULL (same n and stride, no
 * alias of a state) it receives S_{k0+n_sub} with every plane, and next_grid (a
 * grid set of mpm_step_grids_bytes, not one of step_grids) that state's P2G
 * accumulator and marks.  mpm_adjoint_recorded: the adjoint substeps in reverse on what
 * mpm_record_segment left, arguments as in mpm_adjoint_segment.  No
 * counterpart in the reference (it records no tape); with p2g_done = 0 the
 * two together equal mpm_adjoint_segment. */
int mpm_record_segment(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                       const MpmCollider* colliders, int n_colliders, int k0, double dt, int n_sub, int p2g_done,
                       const MpmState* last, void* next_grid, int64_t next_grid_bytes, void* stream);
int mpm_adjoint_recorded(const MpmParams* prm, const MpmState* states, const void* order, void* const* step_grids,
                         const MpmCollider* colliders, int n_colliders, void* workspace, int k0, double dt, int n_sub,
                         float* adj_a, float* adj_b, int64_t adj_stride, double* param_adj, void* stream);

/* ---- the flat state in bin order -------------------------------------
 *
 * The same substep as mpm_step (p2g2p, mpm_solver_warp.py:156-203) with
 * the P2G scatter (p2g_apic_stress_given, mpm_utils.py:190-227) privatized
 * per chunk of particles in LDS instead of 27 global float atomics per
 * particle and component.  mpm_sort_by_bin counting-sorts a flat state by
 * the 8^3-cell bin and 4^3-cell sub-block of each particle's stencil base
 * (mpm_utils.py:197-199), then puts each chunk in LDS bank order, into `dst`
 * and fills `order` (mpm_order_bytes
 * bytes, 256-byte aligned): the chunk table (runs of at most 256, 512 or
 * 1024 particles -- the largest that still gives ~900 chunks -- inside one
 * bin) and the permutation
 * perm[sorted slot] = source slot (int32, at byte offset
 * mpm_order_perm_offset).  The order stays valid for any state
 * in the same slot order -- e.g. the substeps after the sort -- as particles
 * drift; it only loses speed.  No counterpart in the reference (Warp's P2G
 * runs one thread per particle in storage order). */
int64_t mpm_order_bytes(const MpmParams* prm, int64_t n_particles);
int64_t mpm_order_perm_offset(const MpmParams* prm, int64_t n_particles);
int mpm_sort_by_bin(const MpmParams* prm, const MpmState* src, const MpmState* dst, void* order, void* stream);

/* p2g2p on a state in the slot order of `order`; workspace as mpm_step. */
int mpm_step_ordered(const MpmParams* prm, const MpmState* st, const void* order, const MpmCollider* colliders,
                     int n_colliders, void* workspace, int64_t workspace_bytes, float time, float dt, void* stream);

/* n_sub substeps k0 .. k0+n_sub-1 on an ordered state, substep k at time
 * (float)(k * dt) with step (float)dt: the P2G of the first, then per
 * substep the grid update and one fused kernel doing this substep's G2P and
 * the next substep's P2G over the chunks (LDS tiles of grid velocities and
 * of the next P2G; the grid work covers only the 8^3 blocks the P2G marked).
 * workspace: mpm_run_ordered_workspace_bytes (two grid sets). */
int64_t mpm_run_ordered_workspace_bytes(const MpmParams* prm);
int mpm_run_ordered(const MpmParams* prm, const MpmState* st, const void* order, const MpmCollider* colliders,
                    int n_colliders, void* workspace, int64_t workspace_bytes, int k0, double dt, int n_sub,
                    void* stream);

/* Plane-wise permutation of n columns: scatter = 0: dst[:, j] = src[:, idx[j]];
 * scatter = 1: dst[:, idx[j]] = src[:, j].  idx: n int32 (a sort's perm, or a
 * state's pid plane); entries outside [0, n) are skipped.  Moves states and
 * adjoints between slot orders (the adjoint of a sort is its inverse
 * permutation). */
int mpm_permute(const float* src, int64_t src_stride, float* dst, int64_t dst_stride, int n_planes, const int* idx,
                int64_t n, int scatter, void* stream);

/* ---- batches of independent scenes -----------------------------------
 *
 * No counterpart in the reference: its generator script steps one scene per process
 * (generate_training_data/generate_sand.py:17-32).  A batch runs many
 * scenes -- each with its own grid size, material and colliders -- through
 * the binned kernels above in the same launches: the scenes' bins are
 * numbered one after the other (one MpmBins with nb = 0 holds them all),
 * their grids likewise, and the batch's flat state (MpmState) holds the
 * scenes' particles one scene after the other, PID plane = flat index.  All
 * scenes advance with the same dt and time. */
typedef struct MpmScene {
    MpmParams params;     /* the scene's grid and material */
    int64_t n_particles;  /* its particles, consecutive in the batch's flat state */
    int32_t n_colliders;  /* its colliders: the next n_colliders entries of the batch's collider array */
    int32_t bin_cap;      /* slots per bin of this scene */
} MpmScene;

typedef struct MpmBatch { /* filled by mpm_batch_setup */
    int32_t n_scenes;
    int32_t n_bins;       /* bins of all scenes: sum of ceil(n_grid/8)^3 */
    int32_t n_colliders;
    int32_t max_cap;      /* largest bin_cap of the scenes */
    int64_t n_slots;      /* bin slots of all scenes: sum of bins * bin_cap */
    int64_t n_particles;  /* particles of all scenes */
    int64_t n_nodes;      /* grid nodes of all scenes */
    void* workspace;      /* the batch's workspace; holds its scene table */
} MpmBatch;

/* Workspace of a batch (zeroed once before use, 256-byte aligned). */
int64_t mpm_batch_workspace_bytes(const MpmScene* scenes, int n_scenes);

/* Checks the scenes, lays the batch out and copies its scene table and
 * colliders (host arrays) into the workspace (a blocking copy).  Bins of
 * scene i have scenes[i].bin_cap slots; the batch's MpmBins has nb = 0,
 * bin_cap = the overflow run's block size, count arrays of
 * [2][2 * n_bins + 1] and stride >= (n_slots + overflow_cap) * MPM_NPLANES. */
int mpm_batch_setup(const MpmScene* scenes, int n_scenes, const MpmCollider* colliders, void* workspace,
                    MpmBatch* batch);

/* Byte offset of the status counters in the batch's workspace. */
int64_t mpm_batch_status_offset(const MpmBatch* batch);

/* The mpm_bins_* operations on every scene of a batch at once. */
int mpm_batch_insert(const MpmBatch* batch, const MpmState* flat, const MpmBins* bins, void* stream);
int mpm_batch_extract(const MpmBatch* batch, const MpmBins* bins, const MpmState* flat, void* stream);
int mpm_batch_clear_grid(const MpmBatch* batch, MpmBins* bins, void* stream);
int mpm_batch_p2g(const MpmBatch* batch, MpmBins* bins, float dt, void* stream);
int mpm_batch_run(const MpmBatch* batch, MpmBins* bins, double time, float dt, int n_substeps, void* stream);

/* A batched run of n_substeps captured once as a HIP graph and replayed:
 * one graph launch per n_substeps instead of two kernel launches per substep
 * (plus the sorts).  Capture does not advance the state.  A replay starts on
 * the plane set current at capture (n_substeps plus the sorts it contains
 * must be even -- twice the sort period is), voids every tile stamp except
 * the pending tiles', and reads its start time from the device, so replays
 * continue one another and the colliders' time windows see the true time.
 * Launch: n_replays replays from time `time` (the caller advances its clock
 * by n_replays * n_substeps * dt); the pending P2G must have been computed
 * with the graph's dt.  Eager substeps may run between launches as long as the
 * state is back on the capture's plane set.  The graph holds the batch's
 * pointers: destroy it before the batch's buffers. */
typedef struct MpmGraph MpmGraph;
int mpm_batch_graph_create(const MpmBatch* batch, const MpmBins* bins, float dt, int n_substeps, MpmGraph** graph);
int mpm_batch_graph_launch(MpmGraph* graph, const MpmBatch* batch, MpmBins* bins, double time, int n_replays,
                           void* stream);
int mpm_graph_destroy(MpmGraph* graph);

/* ---- state import / export (torch loaders and export_uvFstress_from_warp,
 *      mpm_solver_warp.py:229-266 and :305-310) --------------------------- */

/* Copies `ncomp` components of an (n, ncomp) row-major array indexed by the
 * caller's particle index into planes [first_plane, first_plane+ncomp) of
 * every slot (slot <- src[pid[slot]]).  mode 0: copy; mode 2: add the
 * identity (F = F_disp + I, ncomp 9). */
int mpm_import(const MpmState* st, int first_plane, int ncomp, const float* src,
               int mode, void* stream);

/* Inverse of mpm_import: dst[pid[slot]] <- planes (mode 0 copy; mode 1
 * subtract the identity, F_disp = F - I, update_F_disp_from_F
 * mpm_utils.py:16-19; mode 3: x = x0 + u, update_x_from_u, ncomp 3). */
int mpm_export(const MpmState* st, int first_plane, int ncomp, float* dst, int mode,
               void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MPM_HIP_H */
