/*
 * oracle.h -- CPU restatement of ar-physics' particle step.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only to check results against (or, in bench.py, to time the CPU); the product path never
 * calls it.
 *
 * Every function restates the reference GLSL/Rust it cites (paths relative to the
 * reference repository).  Three arithmetics are built from one source (oracle_impl.h): a
 * forward-mode derivative build (oracle_dual.cpp, liboracle_dual.so: REAL = an FP64 dual
 * number, orc_jvp_step gives exact directional derivatives of the step) and two precisions:
 *   *_f  FP32, with the reference's operation order inside each expression and the
 *        81-lane sweep + base-3 LDS reduction order of the compute shaders;
 *   *_d  FP64, the same algorithm in double, used to state FP32 rounding error.
 * Deterministic choices where the reference is not: bucket contents are in ascending id
 * order per kind (the reference appends with atomicAdd, neighbor_list.rs:101), so the
 * particles refused by the 128-slot cap are the highest ids of their cell.
 *
 * Parity status: every function is checked by tests/test_oracle.py and
 * tests/test_oracle_solid.py against closed forms, invariants, brute force or the FP64
 * build (the reference holds no golden vectors); orc_solid_forces_* by rigid motion, pair
 * antisymmetry, the St Venant-Kirchhoff energy gradient, the Hooke traction of a uniform
 * stretch and shear, the hourglass restoring force, the plastic threshold and the
 * self-contact closed form.  Marked "parity unpinned":
 *   the tableau taken for the name "rk2" (the explicit midpoint rule) and velocity Verlet
 *     (numerical-integration crate absent from the reference): parity unpinned.
 */
#ifndef ARP_ORACLE_H
#define ARP_ORACLE_H
#include <stdint.h>
#include "../include/arp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Static scene description (what FluidSim::new hands the shaders). */
typedef struct orc_scene {
    uint32_t n, ns, nb, nm;
    const uint32_t* mat;  /* n */
    const uint32_t* sid;  /* n  (solid id or ARP_INVALID_INDEX) */
    const uint32_t* part; /* ns (particle id of each solid) */
    const float* bden;    /* nb */
    const float* bpos;    /* nb*4 */
    const float* bvel;    /* nb*4 */
    const uint32_t* bmat; /* nb */
    const ArpMaterial* mats;
    const ArpMatInteraction* inter; /* nm*nm */
    ArpAABB bounds;
    float h, g;
    int threads; /* OpenMP threads for particle loops (<=0: library default) */
    const float* kscale; /* ns per-solid stiffness scale of (lambda, mu), or NULL (an extension
                            used to check gradients w.r.t. a stiffness field; NULL = reference) */
} orc_scene;

typedef struct orc_state_f {
    float *den, *pos, *vel;  /* n, n*4, n*4 */
    float *ref, *stress;     /* ns*4, ns*16 */
} orc_state_f;

typedef struct orc_state_d {
    double *den, *pos, *vel;
    double *ref, *stress;
} orc_state_d;

uint32_t orc_dim(const ArpAABB* b);
void orc_grid_sub(const ArpAABB* b, float h, uint32_t sub[4]);
float orc_norm_const_f(uint32_t dim, float h);
double orc_norm_const_d(uint32_t dim, double h);
float orc_pressure_f(uint32_t eq, float density, float c, float d0);
float orc_contact_potential_f(uint32_t eq, float r, float strength, float r0);
float orc_kernel_f(const float r[4], float h, float k);
void orc_grad_w_f(const float r[4], float h, float k, float out[4]);

/* Cell of a position: floor(clamp((p-min)/dim,0,1)*sub) (neighbor_list.rs:95), 3 coords. */
void orc_cell_coords(const ArpAABB* b, const uint32_t sub[4], const float p[4], uint32_t c[3]);

/* Neighbour structure (neighbor_list.rs:37-143, 224-294) at positions pos (n*4) and the
 * solids' ref positions.  Outputs (any may be NULL):
 *   cell[n], ref_cell[n]  Index.pos_index / Index.ref_index;
 *   counts[ncells*3]      Bucket.count;
 *   cand_count[n], cand   boundary+position candidates of each particle in its 27-cell
 *                         stencil (fluid.rs:137-142), boundary ids | 0x80000000, sorted. */
int orc_grid_dump(const orc_scene* sc, const float* pos, const float* ref, uint32_t* cell,
                  uint32_t* ref_cell, uint32_t* counts, uint32_t* cand_count, uint32_t* cand,
                  uint64_t cand_capacity);

/* compute_forces (shaders/mod.rs:11-86) on one state: derivative in d (same layout). */
int orc_compute_forces_f(const orc_scene* sc, const orc_state_f* in, orc_state_f* d);
int orc_compute_forces_d(const orc_scene* sc, const orc_state_d* in, orc_state_d* d);
/* Only fluid_forces (fluid.rs) / only solid_forces (solid.rs, includes compute_strain). */
int orc_fluid_forces_f(const orc_scene* sc, const orc_state_f* in, orc_state_f* d);
int orc_solid_forces_f(const orc_scene* sc, const orc_state_f* in, orc_state_f* d);
int orc_fluid_forces_d(const orc_scene* sc, const orc_state_d* in, orc_state_d* d);
int orc_solid_forces_d(const orc_scene* sc, const orc_state_d* in, orc_state_d* d);
/* compute_strain (strain.rs:71-181): strains[s*48 + 16*k + 4*col + row], k = L^-1, F, F'. */
int orc_strain_f(const orc_scene* sc, const orc_state_f* in, float* strains);
int orc_strain_d(const orc_scene* sc, const orc_state_d* in, double* strains);

/* lin_comb_N / s_lin_comb_N (particles.rs:8-154) chained as reduce_owned (:381-401). */
int orc_lincomb_f(const orc_scene* sc, int nterms, const float* coef,
                  const orc_state_f* const* terms, orc_state_f* out);
int orc_lincomb_d(const orc_scene* sc, int nterms, const double* coef,
                  const orc_state_d* const* terms, orc_state_d* out);
/* vel / vel_solid (particles.rs:156-250). */
int orc_velocity_f(const orc_scene* sc, float r, const orc_state_f* in, orc_state_f* out);

/* Explicit RK driver: stage inputs y + dt*sum_j a_ij k_j and the final y + dt*sum b_i k_i
 * are formed with the lin-comb above, y first then k_1..k_s (one fixed term order). */
int orc_step_f(const orc_scene* sc, const ArpTableau* tab, float dt, int nsteps,
               orc_state_f* y);
int orc_step_d(const orc_scene* sc, const ArpTableau* tab, double dt, int nsteps,
               orc_state_d* y);
/* The same driver on the scalar ODE y' = lambda*y (one step). */
double orc_rk_scalar(const ArpTableau* tab, double lambda, double dt, double y0);

/* Constitutive helpers (solid.rs:75-215) on column-major 4x4 matrices, FP32. */
void orc_seth_hill_f(const float F[16], int two_m, int dim, float out[16]);
void orc_seth_hill_rate_f(const float F[16], const float Fd[16], int two_m, int dim,
                          float out[16]);
void orc_inverse_f(const float A[16], float out[16]);
void orc_hooke_f(const float S[16], float lambda, float mu, int dim, float out[16]);

#ifdef __cplusplus
}
#endif
#endif
