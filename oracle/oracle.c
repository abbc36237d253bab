/*
 * oracle.c -- CPU restatement of ar-physics' particle step.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h): loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg as the checker, never by the product path.
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile) so every FP32
 * expression rounds in source order and isnan/isinf survive.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- precision-independent pieces ---------------------------------------- */

/* dim = number of axes with bounds.dim > 0 (soft_body/mod.rs:113-117) */
uint32_t orc_dim(const ArpAABB* b) {
    uint32_t d = 0;
    for (int i = 0; i < 4; i++)
        if (b->dim[i] > 0.0f) d++;
    return d;
}

/* sub[i] = max(1, dim[i]/h) as uint (neighbor_list.rs:201-206); Rust's f32->u32 cast
 * truncates and saturates. */
void orc_grid_sub(const ArpAABB* b, float h, uint32_t sub[4]) {
    for (int i = 0; i < 4; i++) {
        float q = b->dim[i] / h;
        float m = 1.0f > q ? 1.0f : q;
        if (!(m == m)) m = 1.0f;
        sub[i] = m >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)m;
    }
}

/* volume_nball + poly6 norm_const, f32 on the host (kernel.rs:5-11, 60-63) */
static float vol_nball_f(uint32_t n, float r) {
    if (n == 0) return 1.0f;
    if (n == 1) return 2.0f * r;
    const float pi = 3.14159265358979323846f;
    return 2.0f * pi * r * r * vol_nball_f(n - 2, r) / (float)n;
}
float orc_norm_const_f(uint32_t dim, float h) {
    const float pi = 3.14159265358979323846f;
    return (pi * pi * pi) / (6.0f * vol_nball_f(dim + 6, h));
}
static double vol_nball_d(uint32_t n, double r) {
    if (n == 0) return 1.0;
    if (n == 1) return 2.0 * r;
    return 2.0 * M_PI * r * r * vol_nball_d(n - 2, r) / (double)n;
}
double orc_norm_const_d(uint32_t dim, double h) {
    return (M_PI * M_PI * M_PI) / (6.0 * vol_nball_d(dim + 6, h));
}

/* ---------------- the two precisions -------------------------------------------------- */

#define REAL float
#define SFX _f
#define SQRT sqrtf
#define POW powf
#define LOG logf
#define FMAX fmaxf
#define FMIN fminf
#define FLOOR floorf
#define ISINF isinf
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef LOG
#undef FMAX
#undef FMIN
#undef FLOOR

#define REAL double
#define SFX _d
#define SQRT sqrt
#define POW pow
#define LOG log
#define FMAX fmax
#define FMIN fmin
#define FLOOR floor
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef LOG
#undef FMAX
#undef FMIN
#undef FLOOR

/* ---------------- small exported helpers for known-answer tests ----------------------- */

float orc_pressure_f(uint32_t eq, float density, float c, float d0) {
    return pressure_f(eq, density, c, d0);
}
float orc_contact_potential_f(uint32_t eq, float r, float strength, float r0) {
    return contact_f(eq, r, strength, r0);
}
float orc_kernel_f(const float r[4], float h, float k) { return kern_f(r, h, k); }
void orc_grad_w_f(const float r[4], float h, float k, float out[4]) { gradw_f(r, h, k, out); }

void orc_cell_coords(const ArpAABB* b, const uint32_t sub[4], const float p[4], uint32_t c[3]) {
    (void)cell_of_f(b, sub, p, c);
}

void orc_seth_hill_f(const float F[16], int two_m, int dim, float out[16]) {
    mseth_hill_f(F, two_m, (uint32_t)dim, out);
}
void orc_seth_hill_rate_f(const float F[16], const float Fd[16], int two_m, int dim,
                          float out[16]) {
    mseth_hill_rate_f(F, Fd, two_m, (uint32_t)dim, out);
}
void orc_inverse_f(const float A[16], float out[16]) { minv_f(A, out); }
void orc_hooke_f(const float S[16], float lambda, float mu, int dim, float out[16]) {
    mhooke_f(S, lambda, mu, (uint32_t)dim, out);
}

int orc_velocity_f(const orc_scene* sc, float r, const orc_state_f* in, orc_state_f* out) {
    /* vel (particles.rs:194-203): pos <- r*vel, vel <- 0, den <- 0; vel_solid: 0 */
    for (size_t i = 0; i < sc->n; i++) {
        out->den[i] = 0.0f;
        for (int c = 0; c < 4; c++) {
            out->pos[4 * i + c] = r * in->vel[4 * i + c];
            out->vel[4 * i + c] = 0.0f;
        }
    }
    for (size_t s = 0; s < sc->ns; s++) {
        for (int c = 0; c < 4; c++) out->ref[4 * s + c] = 0.0f;
        for (int c = 0; c < 16; c++) out->stress[16 * s + c] = 0.0f;
    }
    return 0;
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

int orc_grid_dump(const orc_scene* sc, const float* pos, const float* ref, uint32_t* cell,
                  uint32_t* ref_cell, uint32_t* counts, uint32_t* cand_count, uint32_t* cand,
                  uint64_t cand_capacity) {
    orc_grid_f g;
    if (grid_build_f(sc, pos, sc->ns ? ref : NULL, &g)) return -1;
    if (cell) memcpy(cell, g.pos_index, sc->n * sizeof(uint32_t));
    if (ref_cell) memcpy(ref_cell, g.ref_index, sc->n * sizeof(uint32_t));
    if (counts) memcpy(counts, g.cnt, g.ncells * 3 * sizeof(uint32_t));
    uint64_t off = 0;
    int rc = 0;
    for (uint32_t i = 0; i < sc->n; i++) {
        uint32_t b0 = g.pos_index[i];
        uint64_t start = off;
        if (b0 != ARP_INVALID_INDEX) {
            for (int lane = 0; lane < 81; lane++) {
                if ((lane % 9) / 3 != 0) continue; /* one lane per stencil cell */
                uint64_t b;
                if (!lane_cell_f(&g, b0, lane, &b)) continue;
                uint32_t bc = g.cnt[3 * b], rcn = g.cnt[3 * b + 1], pc = g.cnt[3 * b + 2];
                for (uint32_t j = 0; j < bc + rcn + pc; j++) {
                    if (j >= bc && j < bc + rcn) continue;
                    uint32_t id = pidx_f(&g, b, j);
                    if (j < bc) id |= 0x80000000u;
                    if (cand && off < cand_capacity) cand[off] = id;
                    else if (cand) rc = -2;
                    off++;
                }
            }
        }
        if (cand_count) cand_count[i] = (uint32_t)(off - start);
        if (cand && off <= cand_capacity) qsort(cand + start, off - start, sizeof(uint32_t), cmp_u32);
    }
    grid_free_f(&g);
    return rc;
}

double orc_rk_scalar(const ArpTableau* tab, double lambda, double dt, double y0) {
    /* the orc_step structure on y' = lambda*y: stage input y + sum_j (dt*a_ij)*k_j, then
     * y + sum_i (dt*b_i)*k_i, terms summed left to right */
    double k[ARP_MAX_STAGES];
    int S = tab->stages;
    for (int s = 0; s < S; s++) {
        double x = 1.0 * y0;
        for (int j = 0; j < s; j++)
            if (tab->a[s][j] != 0.0f) x = x + (dt * (double)tab->a[s][j]) * k[j];
        k[s] = lambda * x;
    }
    double y = 1.0 * y0;
    for (int j = 0; j < S; j++)
        if (tab->b[j] != 0.0f) y = y + (dt * (double)tab->b[j]) * k[j];
    return y;
}
