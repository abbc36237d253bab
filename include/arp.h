/*
 * arp.h -- C ABI of the MI355X particle step (SPH fluid, contact, elastic/plastic solids,
 * explicit Runge-Kutta stages) that drops in for the GLSL compute dispatch of
 * jsmith628/ar-physics.
 *
 * Conventions
 *   - Every entry point returns int: 0 = ARP_OK, negative = error class (ARP_E_*);
 *     arp_last_error(ctx) gives the message.  No C++ exception crosses this boundary.
 *   - Host pointers stay owned by the caller and are copied before the call returns.
 *   - Device pointers (the d_* arguments) are raw HIP device addresses of std430 arrays,
 *     e.g. torch tensors on the same device; the call is ordered on the context's stream.
 *   - A context is used by one thread at a time (the reference's GL-context contract);
 *     separate contexts may live on separate threads.
 *   - Only dim <= 3 is supported: the 4th component of every vec4 must be 0.
 *
 * Wire structs are byte-identical to the reference's std430 GLSL structs, so the Rust host
 * can pass its existing soft_body types unchanged.
 */
#ifndef ARP_H
#define ARP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARP_OK 0
#define ARP_E_ARG (-1)      /* invalid argument / unsupported configuration        */
#define ARP_E_HIP (-2)      /* HIP runtime or module error                         */
#define ARP_E_STATE (-3)    /* call made in the wrong order (e.g. no scene)        */
#define ARP_E_NOMEM (-4)    /* device or host allocation failed                    */
#define ARP_E_LIMIT (-5)    /* a documented size limit was exceeded                */

#define ARP_INVALID_INDEX 0xFFFFFFFFu

/* Particle: src/soft_body/particle_state/mod.rs:24-29 (std430, 48 B) */
typedef struct ArpParticle {
    float den;
    uint32_t mat;
    uint32_t solid_id;
    uint32_t _pad0;
    float pos[4];
    float vel[4];
} ArpParticle;

/* SolidParticle: src/soft_body/particle_state/mod.rs:31-35 (std430, 96 B);
 * stress is a column-major mat4 (stress[4*col + row]); it holds the plastic strain. */
typedef struct ArpSolidParticle {
    uint32_t part_id;
    uint32_t _pad0[3];
    float ref_pos[4];
    float stress[16];
} ArpSolidParticle;

/* Material: src/soft_body/material_region/material.rs:11-28 (std430, 80 B; bool = 4 B) */
typedef struct ArpMaterial {
    uint32_t immobile;
    float mass, start_den;
    uint32_t state_eq; /* 0 zero, 1 constant, 2 ideal gas, 3 Tait (material.rs:53-59) */
    float sound_speed, target_den, visc, bulk_visc;
    int32_t strain_order;
    float normal_stiffness, shear_stiffness, normal_damp, shear_damp;
    uint32_t plastic;
    float yield_strength, work_hardening, work_hardening_exp, kinematic_hardening,
        thermal_softening;
    float relaxation_time;
} ArpMaterial;

/* MatInteraction: src/soft_body/material_region/interaction.rs:10-15 (std430, 20 B);
 * potential: 0 zero, 1 constant, 2 linear, 3 Tait, 4 Lennard-Jones, 5 Hertz */
typedef struct ArpMatInteraction {
    uint32_t potential;
    float strength, radius;
    float dampening, friction;
} ArpMatInteraction;

/* AABB: src/soft_body/material_region/region.rs:11-14 (std430, 32 B) */
typedef struct ArpAABB {
    float min[4];
    float dim[4];
} ArpAABB;

/* Explicit Runge-Kutta tableau (the role of the numerical-integration crate's
 * VelIntegrates implementors, examples/test.rs:165-180).  a is strictly lower triangular. */
#define ARP_MAX_STAGES 6
/* stages == ARP_VERLET selects velocity Verlet (examples/test.rs:167, the TOML default
 * config.rs:451): one force evaluation per step.  With D = (a, rho', eps') the last
 * evaluation, Z = (x + dt (v + dt/2 a), v + dt/2 a, rho + dt/2 rho', eps + dt/2 eps'), the
 * forces D' are evaluated at the predicted X = (Z.x, Z.v + dt/2 a, Z.rho + dt/2 rho',
 * Z.eps + dt/2 eps') and y' = Z + dt/2 D' in v, rho and eps (second order; D' is kept for
 * the next step; the first step of a state evaluates D first, as init_with_vel does,
 * soft_body/mod.rs:220-232).  a, b, c are unused. */
#define ARP_VERLET (-1)
typedef struct ArpTableau {
    int32_t stages;
    float a[ARP_MAX_STAGES][ARP_MAX_STAGES];
    float b[ARP_MAX_STAGES];
    float c[ARP_MAX_STAGES];
} ArpTableau;

/* Neighbour-structure statistics of the last grid build (per kind: boundary, reference,
 * position), as stored by the reference's Bucket.count[3] (neighbor_list.rs:12-16). */
typedef struct ArpGridInfo {
    uint32_t sub[4];          /* subdivisions, neighbor_list.rs:201-206 */
    uint64_t ncells;
    uint32_t n_valid[3];      /* particles that got a bucket slot, per kind */
    uint32_t n_overflow[3];   /* particles refused by the 128-slot bucket cap, per kind */
    uint32_t n_nonfinite[3];  /* NaN/inf positions dropped from the grid, per kind */
    uint32_t occupied_rows;   /* (y,z) cell rows holding at least one mobile particle */
} ArpGridInfo;

typedef struct arp_ctx arp_ctx;

/* ---- context ------------------------------------------------------------------------- */
/* Replaces the GL context + program compilation (soft_body/mod.rs:134-137). */
int arp_create(int device, uint32_t flags, arp_ctx** out);
void arp_destroy(arp_ctx* ctx);
const char* arp_last_error(const arp_ctx* ctx);
/* Order all work of this context on an existing hipStream_t (NULL: the context's own). */
int arp_set_stream(arp_ctx* ctx, void* hip_stream);
int arp_synchronize(arp_ctx* ctx);

/* ---- host-side derivations the reference performs in FluidSim::new -------------------- */
/* poly6 normalisation, computed on the host in f32 (kernel.rs:5-11, 60-63). */
float arp_norm_const(uint32_t dim, float h);
/* MatInteraction::default_between (interaction.rs:42-50). */
void arp_default_interaction(const ArpMaterial* a, const ArpMaterial* b, float h,
                             ArpMatInteraction* out);
/* Interaction-matrix resolution of FluidSim::new (soft_body/mod.rs:93-111): cfg[i*nm+j]
 * is used when present[i*nm+j] != 0, else cfg[j*nm+i] when present, else default_between
 * for i != j, else the zero interaction.  out has nm*nm entries. */
void arp_resolve_interactions(uint32_t nm, const ArpMaterial* mats,
                              const ArpMatInteraction* cfg, const uint8_t* present, float h,
                              ArpMatInteraction* out);
/* Named integrators of examples/test.rs:165-180 ("verlet", "euler", "rk1", "midpoint", "rk2",
 * "heun2", "heun", "ralston", "rk3", "heun3", "rk4", "rk_3_8"); case-insensitive. */
int arp_tableau_by_name(const char* name, ArpTableau* out);

/* ---- scene (FluidSim::new, soft_body/mod.rs:58-166) ------------------------------------ */
/* particles/solids/boundary in the reference's wire format; inter is the resolved nm x nm
 * matrix; bounds is the neighbour-grid AABB.  Builds the static boundary and reference
 * grids, renumbers nothing visible to the caller. */
int arp_scene_upload(arp_ctx* ctx, const ArpParticle* particles, uint32_t n,
                     const ArpSolidParticle* solids, uint32_t ns, const ArpParticle* boundary,
                     uint32_t nb, const ArpMaterial* mats, uint32_t nm,
                     const ArpMatInteraction* inter, const ArpAABB* bounds, float kernel_radius,
                     float gravity, float artificial_viscosity);
/* Integrator + FluidSim's timestep/subticks (soft_body/mod.rs:47-66, 240). */
int arp_set_integrator(arp_ctx* ctx, const ArpTableau* tableau, float timestep,
                       uint32_t subticks);
/* FluidSim::update (soft_body/mod.rs:235-274), called nframes times: each frame advances
 * subticks steps of timestep/subticks.  The fused fast path (sort, gather, RK stages).
 * With neighbour lists kept, a method of several stages and at least 8 steps in the call,
 * steps are issued speculatively (each stage as its list walk alone, a few steps in flight;
 * a step whose lists turn out invalid is run again with the rebuild chain), and the call
 * returns once they have run; shorter calls return without waiting. */
int arp_update(arp_ctx* ctx, uint32_t nframes);
/* Advance nsteps single RK steps of length dt (no subtick division). */
int arp_step(arp_ctx* ctx, float dt, uint32_t nsteps);
/* Simulation time accumulated by arp_update/arp_step (FluidSim::time). */
double arp_time(const arp_ctx* ctx);
/* State transfer in the caller's original particle/solid order (FluidSim::particles). */
int arp_download(arp_ctx* ctx, ArpParticle* particles, uint32_t n, ArpSolidParticle* solids,
                 uint32_t ns);
int arp_upload_state(arp_ctx* ctx, const ArpParticle* particles, uint32_t n,
                     const ArpSolidParticle* solids, uint32_t ns);

/* ---- neighbour lists kept across RK stages (the fused path's replacement for the
 * per-stage rebuild of NeighborList::update_contents, neighbor_list.rs:224-294) ----------
 * For fluid-only scenes arp_update/arp_step/arp_batch_step keep, per particle, the
 * candidates within h * (1 + skin) found by a grid search, and later stages walk those
 * lists instead of sorting and searching again, for as long as every particle stays within
 * skin*h/2 of its position at the search and no bucket would exceed its 128 slots
 * (checked on the device at every stage; otherwise the stage sorts and searches as the
 * reference does).  The neighbours within h, and so the forces, are the reference's; only
 * the order of the per-particle sums follows the list.  skin = 0 disables the lists
 * (default 0.2); cap = entries per particle (0: 48; a list that overflows is not kept).
 * ARP_NL_ALWAYS_BUILD rebuilds at every stage (testing).  Resets the statistics. */
#define ARP_NL_ALWAYS_BUILD 1u
/* Issue the rebuild chain at every stage instead of speculative list-reusing steps (the
 * results are the same; for comparison and debugging). */
#define ARP_NL_NO_SPECULATION 2u
int arp_set_neighbor_skin(arp_ctx* ctx, float skin, uint32_t cap, uint32_t flags);
/* 1 when lists of cap entries (0: 48) can be kept for n particles: every entry must lie
 * below 4 GiB (cap * (n + 16) * 4 bytes); otherwise the steps search at every stage.  No
 * device needed; no counterpart in the reference. */
int arp_neighbor_lists_fit(uint32_t cap, uint32_t n);
/* Rebuild lists older than this many RK stages even when they are still valid (0: only when
 * invalid).  A build re-sorts the particles into bucket order, which keeps the neighbour
 * gathers local as the particles move. */
int arp_set_neighbor_max_age(arp_ctx* ctx, uint32_t stages);
/* Stages that searched and built lists, and stages that reused them, since the last
 * arp_set_neighbor_skin. */
int arp_neighbor_stats(arp_ctx* ctx, uint32_t* builds, uint32_t* reused);
/* Two-level lists (no counterpart in the reference, whose stages always search): the list
 * walks use a short list of the candidates within h * (1 + short_skin), taken from the long
 * list (skin above) while every particle stays within short_skin*h/2 of its position at the
 * short list's refresh; a stage whose short list is stale refreshes it from the long list
 * (no sort, no grid search), and only a stale long list costs a search.  short_skin = 0
 * (default) walks the long list; otherwise 0.01 <= short_skin < skin.  The neighbours within
 * h are the same.  Resets the statistics. */
int arp_set_neighbor_short_skin(arp_ctx* ctx, float short_skin);
/* Short-list refreshes since the last arp_set_neighbor_skin or arp_set_neighbor_short_skin
 * (each is a stage counted as reused by arp_neighbor_stats). */
int arp_neighbor_refreshes(arp_ctx* ctx, uint32_t* refreshes);
/* Autotuner of the lists (no counterpart in the reference, whose step always searches): times
 * trial_steps steps (0: 100) from the current state for each (skin, short skin) pair of a
 * fixed table -- one level at 0.1, 0.2 and 0.3 h; two levels at 0.2/0.04, 0.3/0.03,
 * 0.3/0.05, 0.4/0.04 and 0.5/0.05 h -- in two rounds, restores the state and the time after
 * every trial, and sets the fastest pair (written to *skin and *short_skin, either may be
 * NULL); capacity and flags stay.  The pairs give the same neighbours within h, so results
 * differ between them only in the rounding of the sums.  Fluid-only scenes. */
int arp_tune_lists(arp_ctx* ctx, float dt, uint32_t trial_steps, float* skin, float* short_skin);
/* Steps issued speculatively (each stage as its list walk alone) that completed, and those
 * whose lists turned out invalid and ran again with the rebuild chain, since the context was
 * created.  No counterpart in the reference (whose step always searches); diagnostics. */
int arp_speculation_stats(const arp_ctx* ctx, uint64_t* steps, uint64_t* aborted);

/* ---- per-stage drop-ins on device std430 buffers ------------------------------------- */
/* compute_forces (shaders/mod.rs:11-86): neighbour update at the input positions, fluid
 * forces, and (when the scene has more than one solid) strain + solid forces, summed. */
int arp_compute_forces(arp_ctx* ctx, const ArpParticle* d_state,
                       const ArpSolidParticle* d_solids, ArpParticle* d_deriv,
                       ArpSolidParticle* d_solid_deriv);
/* ArithShaders::linear_combination (particles.rs:381-436, GLSL lin_comb_N/s_lin_comb_N):
 * out = sum_t coef[t] * term[t] over den/pos/vel (+ ref_pos/stress); mat and solid_id come
 * from the first term.  d_solid_terms/d_solid_out may be NULL when the scene has no solids. */
int arp_lincomb(arp_ctx* ctx, uint32_t nterms, const float* coefs,
                const ArpParticle* const* d_terms, const ArpSolidParticle* const* d_solid_terms,
                ArpParticle* d_out, ArpSolidParticle* d_solid_out);
/* arp_lincomb that allocates nothing on the device.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same results as arp_lincomb, bit for bit.
 * With more than 5 terms the chunk partials live in a fixed 128 MiB scratch array that each
 * GPU holds from the library's first use until the process ends: nchunks * (48 B * n
 * [+ 96 B * ns with solids]), each part rounded up to 256 B, must fit (ARP_E_LIMIT
 * otherwise), where nchunks = ceil(nterms / 5).  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams or threads) race. */
int arp_lincomb_noalloc(arp_ctx* ctx, uint32_t nterms, const float* coefs,
                        const ArpParticle* const* d_terms,
                        const ArpSolidParticle* const* d_solid_terms, ArpParticle* d_out,
                        ArpSolidParticle* d_solid_out);
/* ArithShaders::velocity (particles.rs:438-446, GLSL vel/vel_solid). */
int arp_velocity(arp_ctx* ctx, float r, const ArpParticle* d_in,
                 const ArpSolidParticle* d_solid_in, ArpParticle* d_out,
                 ArpSolidParticle* d_solid_out);
/* NeighborList::update_contents (neighbor_list.rs:224-294) at the given device state. */
int arp_neighbors_update(arp_ctx* ctx, const ArpParticle* d_state);

/* ---- diagnostics ------------------------------------------------------------------------ */
int arp_grid_info(arp_ctx* ctx, ArpGridInfo* out);
/* Neighbour structure of the last build, in original particle ids:
 *   cell[i]       bucket id of particle i (ARP_INVALID_INDEX if dropped or refused);
 *   cand_count[i] number of boundary + position candidates in its 27-cell stencil;
 *   cand (optional, size >= sum of cand_count): the candidates of particle i, boundary ids
 *   tagged with bit 31, in ascending (tag, id) order, concatenated over i. */
int arp_neighbor_dump(arp_ctx* ctx, uint32_t* cell, uint32_t* cand_count, uint32_t* cand,
                      uint64_t cand_capacity);
/* Per-cell counts (boundary, reference, position) of the last build for cells
 * [first, first+count): out has 3*count entries (Bucket.count semantics incl. the cap). */
int arp_cell_counts(arp_ctx* ctx, uint64_t first, uint64_t count, uint32_t* out);

/* ---- measurement ----------------------------------------------------------------------- */
/* No counterpart in the reference (its Profiler times host enqueue, src/profiler.rs:10-81).
 * enable != 0: every kernel launch of the context is bracketed by hipEvents recorded on the
 * launch stream; arp_profile_query synchronises and returns, per kernel name, the summed
 * event time and the number of launches since the last reset. */
int arp_profile(arp_ctx* ctx, int enable);
int arp_profile_query(arp_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches);
/* The event times (ms) of each launch of one kernel since the last reset, in launch order:
 * up to capacity into ms (may be NULL with capacity 0), the number of launches into count.
 * (A list walk's launch returns at once when its stage searched instead, or when an earlier
 * speculative stage found the lists stale: these launches are short.) */
int arp_profile_durations(arp_ctx* ctx, const char* kernel, float* ms, uint64_t capacity, uint64_t* count);
int arp_profile_reset(arp_ctx* ctx);
/* Time only the named kernels (comma-separated, e.g. "k_fluid,k_fluid_adj"); NULL or ""
 * selects all.  Untimed kernels launch without events. */
int arp_profile_select(arp_ctx* ctx, const char* kernels);

/* ---- reverse mode (no counterpart in the reference, which has no adjoint) ------------- */
/* Advance nsteps fused steps of length dt from the current state (fluid scenes: no solid
 * path), keeping a checkpoint of the state every checkpoint_every steps and the stage
 * inputs and neighbour structures of the last segment for arp_adjoint_backward.  Memory:
 * one checkpoint (40 B/particle) per segment plus checkpoint_every recorded steps.  A later
 * arp_step/arp_update/arp_upload_state drops the recording. */
int arp_adjoint_record(arp_ctx* ctx, float dt, uint32_t nsteps, uint32_t checkpoint_every);
/* Reverse-mode sweep over the recorded trajectory (segments are replayed from their
 * checkpoints as needed).  d_grad_out: n device ArpParticle records holding dL/d(den,
 * pos.xyz, vel.xyz) of the state after the recorded steps, indexed like arp_download's
 * output; d_grad_in receives dL/d of the state before them (mat/solid_id fields zero).
 * grad_params (host, may be NULL): 6*nmat floats per material m: [6m] bulk_viscosity,
 * [6m+1] viscosity, [6m+2] normal_stiffness, [6m+3] shear_stiffness, [6m+4] normal_damp,
 * [6m+5] shear_damp (at most 16 materials).  Solids: every strain order, plastic flow
 * included (the plastic strain is carried as state; dL/d of the final plastic strain is
 * taken as zero and that of the initial one is not returned).
 * The live state is left at the end of the recorded trajectory. */
int arp_adjoint_backward(arp_ctx* ctx, const void* d_grad_out, void* d_grad_in, float* grad_params);
/* Per-solid stiffness scale k_s multiplying the solid's (normal, shear) stiffness in the
 * stress (no counterpart in the reference, whose stiffness is per material; NULL restores
 * k_s = 1, the reference behaviour).  ns = number of solids. */
int arp_set_stiffness_field(arp_ctx* ctx, const float* scale, uint32_t ns);
/* After arp_adjoint_backward: dL/dk_s of the stiffness scale of every solid (ns floats). */
int arp_adjoint_stiffness_grad(arp_ctx* ctx, float* out, uint32_t ns);

/* ---- batched ensembles (no counterpart in the reference: one FluidSim per scene) ------- */
/* Independent scenes (fluid, or with solids: the fluid scenes take no part in the solid
 * path's launches), one context each (uploaded as usual, same device, same
 * integrator), stepped together: every kernel of an RK stage runs once for the whole batch
 * (each workgroup looks up its scene), and with ARP_BATCH_GRAPH the two-step launch cycle is
 * captured in a hipGraph and replayed.  The contexts keep working on their own between
 * batch calls (download, upload, ...).  The batch runs on the first context's stream and
 * is ordered after / before the other contexts' streams. */
typedef struct arp_batch arp_batch;
#define ARP_BATCH_GRAPH 1u
int arp_batch_create(arp_ctx* const* ctxs, uint32_t nscenes, uint32_t flags, arp_batch** out);
void arp_batch_destroy(arp_batch* batch);
const char* arp_batch_last_error(const arp_batch* batch);
/* nsteps RK steps of length dt on every scene (the time of each context advances). */
int arp_batch_step(arp_batch* batch, float dt, uint32_t nsteps);
/* Reverse mode over a batch (no counterpart in the reference): arp_adjoint_record and
 * arp_adjoint_backward of every scene, scene after scene, each on its context's stream.
 * d_grad_out[s] / d_grad_in[s] are scene s's device cotangent arrays (n_s Particle records),
 * grad_params (may be NULL) holds one host pointer per scene, each NULL or 6 * nmat_s floats
 * (arp_adjoint_backward's layout). */
int arp_batch_adjoint_record(arp_batch* batch, float dt, uint32_t nsteps, uint32_t checkpoint_every);
int arp_batch_adjoint_backward(arp_batch* batch, const void* const* d_grad_out, void* const* d_grad_in,
                               float* const* grad_params);
/* Autotuner: times trial_steps batch steps (after two untimed ones) for each candidate pair
 * of row-sort workgroup size (particles per workgroup: 16 .. 2048) and list-walk form
 * (registers only, or the neighbours within h buffered in LDS), in two rounds, and keeps
 * the fastest (*chosen_div: the row-sort size); the scenes' states are restored after
 * every trial.  The results are the same for every choice. */
int arp_batch_tune(arp_batch* batch, float dt, uint32_t trial_steps, uint32_t* chosen_div);
/* The batch's current choices: particles per row-sort workgroup, and 1 when the list walks
 * buffer the neighbours within h in LDS. */
int arp_batch_tuning(const arp_batch* batch, uint32_t* rowsort_div, uint32_t* walk_buffered);
/* The list autotuner (arp_tune_lists) over the batch: one pair for every scene, each trial
 * trial_steps batched steps (0: 100). */
int arp_batch_tune_lists(arp_batch* batch, float dt, uint32_t trial_steps, float* skin, float* short_skin);

/* ---- particle insertion (FluidSim::add_particles, soft_body/mod.rs:180-206) ----------- */
/* FluidSim::kernel_radius (soft_body/mod.rs:166) and num_materials (mod.rs:176-178). */
float arp_kernel_radius(const arp_ctx* ctx);
uint32_t arp_num_materials(const arp_ctx* ctx);
/* Lengths of the particle and solid-particle buffers (Particles::particles / solids). */
uint32_t arp_num_particles(const arp_ctx* ctx);
uint32_t arp_num_solids(const arp_ctx* ctx);
/* Particles::add_particles (particle_state/particles.rs:547-631): appends count particles of
 * `material` (their den, pos, vel are used; mat and solid_id are assigned).  A stiff material
 * is always added as a new material; otherwise the first equal material is reused.  A new
 * material's interactions are default_between with every material, itself included unless
 * it is stiff (particles.rs:573-598).  New particles of a stiff material become solids with
 * their current position as reference position.  *material_added (may be NULL) is 1 when a
 * material was added.  The live state, time, integrator, neighbour-list settings and
 * stiffness field (new solids: 1) carry over; recordings are dropped.  Deviation: a scene
 * without solids holds one placeholder solid (part_id ARP_INVALID_INDEX) which new solids
 * replace; the reference appends after it. */
int arp_add_particles(arp_ctx* ctx, const ArpMaterial* material, const ArpParticle* particles,
                      uint32_t count, uint32_t* material_added);

/* ---- scene files: the example runner's TOML scenes (host only, no device needed) ------- */
/* Config::parse (examples/test_mods/config.rs:298-474) and the particle generation of
 * FluidSim::new (soft_body/mod.rs:58-111: MaterialRegion::gen_particles per object, the
 * boundary split, init_solid_particles, the interaction matrix).  Objects are numbered as
 * the reference numbers materials: [boundaries.*] first, then [objects.*] without on_click,
 * each group in byte order of the names.  Keys, defaults and type rules are the reference's
 * (a float key given as an integer is an error, as its as_float().unwrap() is); STL models
 * ("model") are read relative to base_dir, or the working directory when it is NULL. */
typedef struct arp_scene_file arp_scene_file;

typedef struct ArpSceneSettings {
    char title[128];            /* name (default "Fluid Test") */
    char integrator[32];        /* integrator as written (default "verlet") */
    float kernel_radius;        /* kernel_radius (1/64) */
    float time_step;            /* time_step (0.01) */
    float artificial_viscosity; /* artificial_viscocity (50) */
    float gravity;              /* gravity (1) */
    float render_radius;        /* kernel_radius * particle_render_factor */
    uint32_t subticks;          /* subticks (1) */
    uint32_t lighting;          /* lighting (default: light_pos given) */
    ArpAABB bounds;             /* min ([-1,-1,0,0]), dim ([2,2,0,0]) */
    uint32_t n_objects;         /* simulation objects = materials */
    uint32_t n_on_click;        /* interactive objects (indices n_objects..) */
    uint32_t n_particles, n_boundary, n_solids; /* after arp_scene_generate */
} ArpSceneSettings;

typedef struct ArpSceneObject {
    char name[64];
    float colors[3][4];       /* color, color_low_density, color_high_density */
    float packing;            /* lattice spacing / kernel radius */
    uint32_t immobile;        /* a [boundaries.*] object */
    uint32_t on_click;        /* 0 simulated, 1 on_click, 2 on_click = "Relative" */
    ArpAABB bound;            /* the region's bounding box */
    ArpMaterial material;     /* with the lattice mass once generated */
    uint32_t first, count;    /* its range in the particle or boundary array */
    uint32_t lattice_points;  /* lattice points tested (the mass divisor) */
} ArpSceneObject;

/* Parses a scene; on error returns ARP_E_ARG with the message in err (may be NULL). */
int arp_scene_parse(const char* toml_text, size_t len, const char* base_dir, arp_scene_file** out,
                    char* err, size_t err_len);
void arp_scene_free(arp_scene_file* f);
const char* arp_scene_error(const arp_scene_file* f);
int arp_scene_settings(const arp_scene_file* f, ArpSceneSettings* out);
/* i < n_objects: simulation objects (material i); n_objects <= i: on_click objects. */
int arp_scene_object(const arp_scene_file* f, uint32_t i, ArpSceneObject* out);
/* Generates every object's particles (regions tested on `threads` host threads, 0 = all
 * cores); the result does not depend on the thread count. */
int arp_scene_generate(arp_scene_file* f, uint32_t threads);
/* Copies the generated arrays (each may be NULL; counts must match the settings):
 * mobile particles, solid particles, boundary particles, materials and the resolved
 * n_objects x n_objects interaction matrix. */
int arp_scene_arrays(const arp_scene_file* f, ArpParticle* particles, uint32_t n,
                     ArpSolidParticle* solids, uint32_t ns, ArpParticle* boundary, uint32_t nb,
                     ArpMaterial* mats, uint32_t nm, ArpMatInteraction* inter);
/* Region::contains of object i for n points (4 floats each): out[k] = 0 or 1. */
int arp_scene_contains(const arp_scene_file* f, uint32_t object, const float* points,
                       uint32_t n, uint8_t* out);
/* FluidSim::new + its integrator (examples/test.rs:160-205): arp_scene_upload of the
 * generated scene with the file's kernel radius, gravity, artificial viscosity and bounds,
 * then arp_set_integrator with the named integrator, time_step and subticks. */
int arp_scene_apply(arp_ctx* ctx, const arp_scene_file* f);
/* FluidSim::add_particles of object i of a scene file (examples/test_mods/renderer.rs:
 * 386-404 spawns the on_click objects this way): its lattice is generated with the context's
 * kernel radius, offset by offset[0..3] when offset is not NULL (positions only), and added
 * with arp_add_particles. */
int arp_scene_spawn(arp_ctx* ctx, const arp_scene_file* f, uint32_t object, const float* offset,
                    uint32_t* material_added);
/* The parsed TOML document as JSON (no counterpart in the reference: it checks the TOML
 * reader).  *needed = bytes including the terminator; returns ARP_E_ARG with the parse
 * message in out when the text is not valid TOML. */
int arp_toml_to_json(const char* text, size_t len, char* out, size_t cap, size_t* needed);

#ifdef __cplusplus
}
#endif
#endif /* ARP_H */
