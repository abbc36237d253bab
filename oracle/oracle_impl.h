/*
 * oracle_impl.h -- precision-generic body of the CPU restatement (see oracle.h).
 * TEST INFRASTRUCTURE ONLY.  Included twice by oracle.c with REAL = float / double and
 * SFX = _f / _d.  F(name) pastes the suffix.
 *
 * Expression order follows the GLSL source text (left-associative products, dot() as
 * x*x'+y*y'+z*z'+w*w' in order) and the file is built with -ffp-contract=off, so the FP32
 * variant rounds like a non-fusing GLSL implementation would.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)

typedef struct F(orc_mat) {
    REAL mass, start_den, sound_speed, target_den, visc, bulk_visc;
    REAL normal_stiffness, shear_stiffness, normal_damp, shear_damp;
    REAL yield_strength, work_hardening, work_hardening_exp, kinematic_hardening,
        relaxation_time;
    uint32_t state_eq, plastic, immobile;
    int32_t strain_order;
    int elastic; /* normal_stiffness != 0 || shear_stiffness != 0 (material.rs:33) */
} F(orc_mat);

typedef struct F(orc_inter) {
    uint32_t potential;
    REAL strength, radius, dampening, friction;
} F(orc_inter);

typedef struct F(orc_ctx) {
    const orc_scene* sc;
    F(orc_mat) * mats;
    F(orc_inter) * inter;
    REAL h, g, k; /* kernel radius, gravity, poly6 norm_const */
    uint32_t dim;
} F(orc_ctx);

static int F(ctx_init)(F(orc_ctx) * c, const orc_scene* sc) {
    c->sc = sc;
    c->mats = (F(orc_mat)*)calloc(sc->nm ? sc->nm : 1, sizeof(F(orc_mat)));
    c->inter = (F(orc_inter)*)calloc(sc->nm ? (size_t)sc->nm * sc->nm : 1, sizeof(F(orc_inter)));
    if (!c->mats || !c->inter) return -1;
    for (uint32_t i = 0; i < sc->nm; i++) {
        const ArpMaterial* m = &sc->mats[i];
        F(orc_mat)* o = &c->mats[i];
        o->mass = m->mass; o->start_den = m->start_den; o->sound_speed = m->sound_speed;
        o->target_den = m->target_den; o->visc = m->visc; o->bulk_visc = m->bulk_visc;
        o->normal_stiffness = m->normal_stiffness; o->shear_stiffness = m->shear_stiffness;
        o->normal_damp = m->normal_damp; o->shear_damp = m->shear_damp;
        o->yield_strength = m->yield_strength; o->work_hardening = m->work_hardening;
        o->work_hardening_exp = m->work_hardening_exp;
        o->kinematic_hardening = m->kinematic_hardening;
        o->relaxation_time = m->relaxation_time;
        o->state_eq = m->state_eq; o->plastic = m->plastic; o->immobile = m->immobile;
        o->strain_order = m->strain_order;
        o->elastic = (m->normal_stiffness != 0.0f || m->shear_stiffness != 0.0f);
    }
    for (size_t i = 0; i < (size_t)sc->nm * sc->nm; i++) {
        c->inter[i].potential = sc->inter[i].potential;
        c->inter[i].strength = sc->inter[i].strength;
        c->inter[i].radius = sc->inter[i].radius;
        c->inter[i].dampening = sc->inter[i].dampening;
        c->inter[i].friction = sc->inter[i].friction;
    }
    c->h = sc->h;
    c->g = sc->g;
    c->dim = orc_dim(&sc->bounds);
    c->k = F(orc_norm_const)(c->dim, (REAL)sc->h);
#ifdef ORC_CTX_HOOK
    ORC_CTX_HOOK(c); /* the derivative build seeds a material parameter's tangent here */
#endif
    return 0;
}

static void F(ctx_free)(F(orc_ctx) * c) {
    free(c->mats);
    free(c->inter);
}

/* ---------------- vec4 / mat4 (GLSL conventions, column-major m[4*col+row]) ------------ */

static inline REAL F(dot4)(const REAL* a, const REAL* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

static inline REAL F(len4)(const REAL* a) { return SQRT(F(dot4)(a, a)); }

/* poly6 (kernel.rs:81-89) */
static inline REAL F(kern)(const REAL* r, REAL h, REAL k) {
    REAL diff = FMAX(h * h - F(dot4)(r, r), (REAL)0);
    return k * diff * diff * diff;
}

static inline void F(gradw)(const REAL* r, REAL h, REAL k, REAL* out) {
    REAL diff = FMAX(h * h - F(dot4)(r, r), (REAL)0);
    REAL s = (REAL)-6 * k * diff * diff;
    for (int i = 0; i < 4; i++) out[i] = s * r[i];
}

static void F(mzero)(REAL* m) { memset(m, 0, 16 * sizeof(REAL)); }
static void F(mident)(REAL* m) {
    F(mzero)(m);
    m[0] = m[5] = m[10] = m[15] = 1;
}
static void F(mmul)(const REAL* A, const REAL* B, REAL* out) {
    REAL t[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            t[4 * c + r] = A[r] * B[4 * c] + A[4 + r] * B[4 * c + 1] + A[8 + r] * B[4 * c + 2] +
                           A[12 + r] * B[4 * c + 3];
    memcpy(out, t, sizeof t);
}
static void F(mvec)(const REAL* M, const REAL* v, REAL* out) {
    REAL t[4];
    for (int r = 0; r < 4; r++)
        t[r] = M[r] * v[0] + M[4 + r] * v[1] + M[8 + r] * v[2] + M[12 + r] * v[3];
    memcpy(out, t, sizeof t);
}
static void F(mtrans)(const REAL* A, REAL* out) {
    REAL t[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) t[4 * c + r] = A[4 * r + c];
    memcpy(out, t, sizeof t);
}
static void F(madd)(const REAL* A, const REAL* B, REAL* out) {
    for (int i = 0; i < 16; i++) out[i] = A[i] + B[i];
}
static void F(msub)(const REAL* A, const REAL* B, REAL* out) {
    for (int i = 0; i < 16; i++) out[i] = A[i] - B[i];
}
static void F(mscale)(REAL s, const REAL* A, REAL* out) {
    for (int i = 0; i < 16; i++) out[i] = s * A[i];
}
/* outerProduct(c, r): column j = c * r[j] */
static void F(mouter)(const REAL* c, const REAL* r, REAL* out) {
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) out[4 * j + i] = c[i] * r[j];
}
/* inverse(mat4) by cofactors (adjugate / determinant) */
static void F(minv)(const REAL* m, REAL* out) {
    REAL a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3];
    REAL a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
    REAL a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11];
    REAL a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
    REAL b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10;
    REAL b02 = a00 * a13 - a03 * a10, b03 = a01 * a12 - a02 * a11;
    REAL b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    REAL b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30;
    REAL b08 = a20 * a33 - a23 * a30, b09 = a21 * a32 - a22 * a31;
    REAL b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    REAL det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
    REAL id = (REAL)1 / det;
    REAL t[16];
    t[0] = (a11 * b11 - a12 * b10 + a13 * b09) * id;
    t[1] = (a02 * b10 - a01 * b11 - a03 * b09) * id;
    t[2] = (a31 * b05 - a32 * b04 + a33 * b03) * id;
    t[3] = (a22 * b04 - a21 * b05 - a23 * b03) * id;
    t[4] = (a12 * b08 - a10 * b11 - a13 * b07) * id;
    t[5] = (a00 * b11 - a02 * b08 + a03 * b07) * id;
    t[6] = (a32 * b02 - a30 * b05 - a33 * b01) * id;
    t[7] = (a20 * b05 - a22 * b02 + a23 * b01) * id;
    t[8] = (a10 * b10 - a11 * b08 + a13 * b06) * id;
    t[9] = (a01 * b08 - a00 * b10 - a03 * b06) * id;
    t[10] = (a30 * b04 - a31 * b02 + a33 * b00) * id;
    t[11] = (a21 * b02 - a20 * b04 - a23 * b00) * id;
    t[12] = (a11 * b07 - a10 * b09 - a12 * b06) * id;
    t[13] = (a00 * b09 - a01 * b07 + a02 * b06) * id;
    t[14] = (a31 * b01 - a30 * b03 - a32 * b00) * id;
    t[15] = (a20 * b03 - a21 * b01 + a22 * b00) * id;
    memcpy(out, t, sizeof t);
}

/* ---------------- constitutive helpers (solid.rs:75-215) ------------------------------- */

static REAL F(mtrace)(const REAL* A, uint32_t dim) {
    REAL t = A[0];
    for (uint32_t i = 1; i < dim; i++) t += A[5 * i];
    return t;
}
static REAL F(mdotm)(const REAL* x, const REAL* y) {
    return F(dot4)(x, y) + F(dot4)(x + 4, y + 4) + F(dot4)(x + 8, y + 8) + F(dot4)(x + 12, y + 12);
}
/* hooke: normal*trace(strain)*I + 2*shear*strain (solid.rs:139-141) */
static void F(mhooke)(const REAL* S, REAL lambda, REAL mu, uint32_t dim, REAL* out) {
    REAL tr = lambda * F(mtrace)(S, dim);
    REAL s2 = (REAL)2 * mu;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) out[4 * c + r] = (c == r ? tr : (REAL)0) + s2 * S[4 * c + r];
}
/* mat_log: 5-term series (solid.rs:105-116) */
static void F(mlog)(const REAL* A, REAL* out) {
    REAL sum[16], term[16], I[16], AmI[16];
    F(mzero)(sum);
    F(mident)(I);
    F(mscale)((REAL)-1, I, term);
    F(msub)(A, I, AmI);
    F(mscale)((REAL)-1, AmI, AmI);
    for (int i = 1; i <= 5; i++) {
        F(mmul)(term, AmI, term);
        for (int j = 0; j < 16; j++) sum[j] += term[j] / (REAL)i;
    }
    memcpy(out, sum, sizeof sum);
}
/* mat_sqrt: Denman-Beavers, 15 iterations (solid.rs:118-133) */
static void F(msqrt)(const REAL* X, REAL* sq, REAL* isq) {
    REAL Y[16], Z[16], iY[16], iZ[16];
    memcpy(Y, X, sizeof Y);
    F(mident)(Z);
    for (int i = 0; i < 15; i++) {
        F(minv)(Z, iZ);
        F(minv)(Y, iY);
        for (int j = 0; j < 16; j++) {
            REAL y = (REAL)0.5 * (Y[j] + iZ[j]);
            REAL z = (REAL)0.5 * (Z[j] + iY[j]);
            Y[j] = y;
            Z[j] = z;
        }
    }
    memcpy(sq, Y, sizeof Y);
    memcpy(isq, Z, sizeof Z);
}
/* C = F^T F with unused diagonal padded to 1 */
static void F(mcauchy)(const REAL* def, uint32_t dim, REAL* C) {
    REAL Ft[16];
    F(mtrans)(def, Ft);
    F(mmul)(Ft, def, C);
    for (uint32_t i = dim; i < 4; i++) C[5 * i] = 1;
}
/* seth_hill (solid.rs:147-168) */
static void F(mseth_hill)(const REAL* def, int two_m, uint32_t dim, REAL* out) {
    REAL C[16];
    F(mcauchy)(def, dim, C);
    if (two_m == 0) {
        REAL L[16];
        F(mlog)(C, L);
        F(mscale)((REAL)0.5, L, out);
        return;
    }
    uint32_t m = (uint32_t)abs(two_m) >> 1;
    REAL Cm[16], I[16];
    F(mident)(Cm);
    F(mident)(I);
    for (uint32_t i = 0; i < m; i++) F(mmul)(Cm, C, Cm);
    if ((abs(two_m) & 1) == 1) {
        REAL s[16], inv[16];
        F(msqrt)(C, s, inv);
        F(mmul)(Cm, s, Cm);
    }
    if (two_m < 0) F(minv)(Cm, Cm);
    REAL d[16];
    F(msub)(Cm, I, d);
    F(mscale)((REAL)1.0 / (REAL)two_m, d, out);
}
/* seth_hill_rate (solid.rs:170-204) */
static void F(mseth_hill_rate)(const REAL* def, const REAL* rate, int two_m, uint32_t dim,
                               REAL* out) {
    REAL dC[16], t[16], C[16];
    F(mtrans)(rate, t);
    F(mmul)(t, def, dC);
    F(mtrans)(dC, t);
    F(madd)(dC, t, dC);
    F(mcauchy)(def, dim, C);
    if (two_m == 0) {
        REAL iC[16], h[16];
        F(minv)(C, iC);
        F(mscale)((REAL)0.5, dC, h);
        F(mmul)(h, iC, out);
        return;
    }
    uint32_t m = (uint32_t)abs(two_m) >> 1;
    REAL sr[16];
    F(mzero)(sr);
    for (uint32_t i = 0; i < m; i++) {
        REAL term[16];
        F(mident)(term);
        for (uint32_t j = 0; j < m; j++) F(mmul)(term, j == i ? dC : C, term);
        F(madd)(sr, term, sr);
    }
    F(mscale)((REAL)1.0 / (REAL)two_m, sr, sr);
    if (two_m < 0) {
        REAL iC[16], n[16];
        F(minv)(C, iC);
        F(mscale)((REAL)-1, iC, n);
        F(mmul)(n, sr, n);
        F(mmul)(n, iC, out);
    } else {
        memcpy(out, sr, sizeof sr);
    }
}

/* ---------------- EOS and contact potentials (fluid.rs:45-67) ------------------------- */

static inline REAL F(pressure)(uint32_t eq, REAL density, REAL c, REAL d0) {
    switch (eq) {
        case 1: return c * c;
        case 2: return c * c * (density - d0);
        case 3: return (c * c * d0 / (REAL)7) * (POW(density / d0, (REAL)7) - (REAL)1);
        default: return 0;
    }
}

static inline REAL F(contact)(uint32_t eq, REAL r, REAL strength, REAL r0) {
    switch (eq) {
        case 1: return r <= r0 ? strength : (REAL)0;
        case 2: return strength * (r0 - r);
        case 3: return (strength * r / (REAL)7) * (POW(r0 / r, (REAL)7) - (REAL)1);
        case 4: {
            REAL factor = r0 / r;
            factor *= factor * factor;
            factor *= factor;
            return strength * (factor * factor - (REAL)2 * factor);
        }
        case 5: return (strength * r0 / (REAL)7) * (POW(r / r0, (REAL)7) - (REAL)1);
        default: return 0;
    }
}

/* ---------------- neighbour grid (neighbor_list.rs) ----------------------------------- */

typedef struct F(orc_grid) {
    uint32_t sub[4];
    uint64_t ncells;
    uint32_t* cnt;      /* ncells*3 capped counts (Bucket.count) */
    uint32_t* start[3]; /* ncells+1 per kind: CSR offsets over every particle of the kind */
    uint32_t* ids[3];   /* per kind, ids sorted by (cell, id) */
    uint32_t* pos_index;
    uint32_t* ref_index;
} F(orc_grid);

static int F(isfinite4)(const REAL* p) {
    for (int i = 0; i < 4; i++)
        if (!(p[i] == p[i]) || ISINF(p[i])) return 0;
    return 1;
}

/* floor(clamp((p-min)/dim, 0, 1) * sub) (neighbor_list.rs:95); a NaN quotient (an axis
 * with dim == 0) clamps to 0; an index equal to sub (p exactly at bounds.max) is clamped to
 * sub-1 instead of aliasing past the row/plane as the reference's store does. */
static uint64_t F(cell_of)(const ArpAABB* b, const uint32_t* sub, const REAL* p, uint32_t* cc) {
    for (int i = 0; i < 3; i++) {
        REAL t = (p[i] - (REAL)b->min[i]) / (REAL)b->dim[i];
        t = FMIN(FMAX(t, (REAL)0), (REAL)1);
        REAL f = FLOOR(t * (REAL)sub[i]);
        uint32_t c = (uint32_t)f;
        if (c >= sub[i]) c = sub[i] - 1;
        cc[i] = c;
    }
    return (uint64_t)cc[0] + (uint64_t)cc[1] * sub[0] + (uint64_t)cc[2] * sub[0] * sub[1];
}

static void F(grid_free)(F(orc_grid) * g) {
    free(g->cnt);
    for (int k = 0; k < 3; k++) {
        free(g->start[k]);
        free(g->ids[k]);
    }
    free(g->pos_index);
    free(g->ref_index);
}

/* Build the three bucket segments: boundary (mode 0), reference (mode 1), position (2). */
static int F(grid_build)(const orc_scene* sc, const REAL* pos, const REAL* ref,
                         F(orc_grid) * g) {
    memset(g, 0, sizeof *g);
    orc_grid_sub(&sc->bounds, sc->h, g->sub);
    g->ncells = (uint64_t)g->sub[0] * g->sub[1] * g->sub[2];
    uint64_t nc = g->ncells;
    g->cnt = (uint32_t*)calloc(nc * 3, sizeof(uint32_t));
    g->pos_index = (uint32_t*)malloc((sc->n ? sc->n : 1) * sizeof(uint32_t));
    g->ref_index = (uint32_t*)malloc((sc->n ? sc->n : 1) * sizeof(uint32_t));
    if (!g->cnt || !g->pos_index || !g->ref_index) return -1;
    uint32_t nk[3] = {sc->nb, sc->n, sc->n};
    uint64_t* key[3];
    for (int k = 0; k < 3; k++) {
        key[k] = (uint64_t*)malloc((nk[k] ? nk[k] : 1) * sizeof(uint64_t));
        g->start[k] = (uint32_t*)calloc(nc + 1, sizeof(uint32_t));
        g->ids[k] = (uint32_t*)malloc((nk[k] ? nk[k] : 1) * sizeof(uint32_t));
        if (!key[k] || !g->start[k] || !g->ids[k]) return -1;
    }
    /* keys: UINT64_MAX = not in the grid */
    for (uint32_t i = 0; i < sc->nb; i++) {
        REAL p[4];
        for (int c = 0; c < 4; c++) p[c] = sc->bpos[4 * (size_t)i + c];
        uint32_t cc[3];
        key[0][i] = F(isfinite4)(p) ? F(cell_of)(&sc->bounds, g->sub, p, cc) : UINT64_MAX;
    }
    for (uint32_t i = 0; i < sc->n; i++) {
        uint32_t s = sc->sid[i];
        uint32_t cc[3];
        if (s == ARP_INVALID_INDEX || ref == NULL) {
            key[1][i] = UINT64_MAX;
        } else {
            const REAL* p = ref + 4 * (size_t)s;
            key[1][i] = F(isfinite4)(p) ? F(cell_of)(&sc->bounds, g->sub, p, cc) : UINT64_MAX;
        }
        const REAL* q = pos + 4 * (size_t)i;
        key[2][i] = F(isfinite4)(q) ? F(cell_of)(&sc->bounds, g->sub, q, cc) : UINT64_MAX;
    }
    /* stable counting sort per kind */
    for (int k = 0; k < 3; k++) {
        uint32_t* st = g->start[k];
        for (uint32_t i = 0; i < nk[k]; i++)
            if (key[k][i] != UINT64_MAX) st[key[k][i] + 1]++;
        for (uint64_t c = 0; c < nc; c++) st[c + 1] += st[c];
        uint32_t* fill = (uint32_t*)calloc(nc, sizeof(uint32_t));
        if (!fill) return -1;
        for (uint32_t i = 0; i < nk[k]; i++)
            if (key[k][i] != UINT64_MAX) {
                uint64_t c = key[k][i];
                g->ids[k][st[c] + fill[c]++] = i;
            }
        free(fill);
    }
    /* caps: spot = start + atomicAdd(count[mode]); refused when spot >= 128
     * (neighbor_list.rs:98-110), so count[mode] = min(n, 128 - start) */
    for (uint64_t c = 0; c < nc; c++) {
        uint32_t start = 0;
        for (int k = 0; k < 3; k++) {
            uint32_t n = g->start[k][c + 1] - g->start[k][c];
            uint32_t room = 128u - start;
            g->cnt[3 * c + k] = n < room ? n : room;
            start += g->cnt[3 * c + k];
        }
    }
    for (uint32_t i = 0; i < sc->n; i++) {
        g->pos_index[i] = ARP_INVALID_INDEX;
        g->ref_index[i] = ARP_INVALID_INDEX;
    }
    for (int k = 1; k < 3; k++)
        for (uint64_t c = 0; c < nc; c++) {
            uint32_t a = g->start[k][c], n = g->cnt[3 * c + k];
            for (uint32_t j = 0; j < n; j++) {
                uint32_t id = g->ids[k][a + j];
                if (k == 1)
                    g->ref_index[id] = (uint32_t)c;
                else
                    g->pos_index[id] = (uint32_t)c;
            }
        }
    for (int k = 0; k < 3; k++) free(key[k]);
    return 0;
}

/* particle_index(bucket, i) over the three segments (neighbor_list.rs:30-32) */
static inline uint32_t F(pidx)(const F(orc_grid) * g, uint64_t c, uint32_t j) {
    uint32_t bc = g->cnt[3 * c], rc = g->cnt[3 * c + 1];
    if (j < bc) return g->ids[0][g->start[0][c] + j];
    if (j < bc + rc) return g->ids[1][g->start[1][c] + (j - bc)];
    return g->ids[2][g->start[2][c] + (j - bc - rc)];
}

/* local_bucket_pos: lane (x,y,z) of the 9x3x3 group looks at cell + (x%3, y, z) - 1 */
static inline int F(lane_cell)(const F(orc_grid) * g, uint32_t bucket, int lane, uint64_t* cell) {
    int lx = lane % 9, ly = (lane / 9) % 3, lz = lane / 27;
    int64_t x = (int64_t)(bucket % g->sub[0]) + lx % 3 - 1;
    int64_t y = (int64_t)((bucket / g->sub[0]) % g->sub[1]) + ly - 1;
    int64_t z = (int64_t)(bucket / ((uint64_t)g->sub[0] * g->sub[1])) + lz - 1;
    if (x < 0 || y < 0 || z < 0 || x >= g->sub[0] || y >= g->sub[1] || z >= g->sub[2]) return 0;
    *cell = (uint64_t)x + (uint64_t)y * g->sub[0] + (uint64_t)z * g->sub[0] * g->sub[1];
    return 1;
}

/* base-3 LDS tree: s[g] += s[g+k] + s[g+2k], k = 27, 9, 3, 1 (fluid.rs:226-234) */
static void F(reduce81)(REAL* s, int width) {
    for (int k = 27; k >= 1; k /= 3)
        for (int g = 0; g < k; g++)
            for (int c = 0; c < width; c++)
                s[g * width + c] = s[g * width + c] + (s[(g + k) * width + c] + s[(g + 2 * k) * width + c]);
}

/* ---------------- fluid_forces (fluid.rs:95-251) ------------------------------------- */

static void F(fluid_one)(const F(orc_ctx) * cx, const F(orc_grid) * g, const F(orc_state) * in,
                         uint32_t p, F(orc_state) * d) {
    const orc_scene* sc = cx->sc;
    const REAL h = cx->h, k = cx->k;
    const REAL EPS = (REAL)0.1;
    uint32_t s_id = sc->sid[p], mat_id = sc->mat[p];
    const F(orc_mat)* M1 = &cx->mats[mat_id];
    REAL m1 = M1->mass, d1 = in->den[p];
    int elastic = s_id != ARP_INVALID_INDEX && M1->elastic;
    REAL sden[81], sf[81 * 4];
    memset(sden, 0, sizeof sden);
    memset(sf, 0, sizeof sf);
    const REAL* xi = in->pos + 4 * (size_t)p;
    const REAL* vi = in->vel + 4 * (size_t)p;
    uint32_t bucket = g->pos_index[p];
    /* a particle without a bucket (non-finite or refused by the cap) makes the reference
     * read buckets[0xFFFFFFFF] (fluid.rs:117); here it sees no neighbours */
    if (bucket != ARP_INVALID_INDEX) {
        REAL c1 = M1->sound_speed, f1 = M1->visc, bf1 = M1->bulk_visc, d0 = M1->target_den;
        REAL p1 = F(pressure)(M1->state_eq, d1, c1, d0);
        for (int lane = 0; lane < 81; lane++) {
            uint64_t b;
            if (!F(lane_cell)(g, bucket, lane, &b)) continue;
            uint32_t sub = (uint32_t)((lane % 9) / 3);
            uint32_t bc = g->cnt[3 * b], rc = g->cnt[3 * b + 1], pc = g->cnt[3 * b + 2];
            REAL den = 0, force[4] = {0, 0, 0, 0};
            for (uint32_t j = sub; j < pc + rc + bc; j += 3) {
                if (j >= bc && j < rc + bc) continue;
                uint32_t p2 = F(pidx)(g, b, j);
                int bnd = j < bc;
                uint32_t mat2 = bnd ? sc->bmat[p2] : sc->mat[p2];
                /* particles[p_id2].solid_id with a boundary index (fluid.rs:141) only feeds
                 * elastic2, which a non-stiff boundary material makes false anyway */
                uint32_t s2 = (p2 < sc->n) ? sc->sid[p2] : ARP_INVALID_INDEX;
                const F(orc_mat)* M2 = &cx->mats[mat2];
                REAL m2 = M2->mass, f2 = M2->visc, bf2 = M2->bulk_visc, c2 = M2->sound_speed;
                REAL r[4], v[4], d2;
                if (bnd) {
                    for (int c = 0; c < 4; c++) {
                        r[c] = (REAL)sc->bpos[4 * (size_t)p2 + c] - xi[c];
                        v[c] = (REAL)sc->bvel[4 * (size_t)p2 + c] - vi[c];
                    }
                    d2 = sc->bden[p2];
                } else {
                    for (int c = 0; c < 4; c++) {
                        r[c] = in->pos[4 * (size_t)p2 + c] - xi[c];
                        v[c] = in->vel[4 * (size_t)p2 + c] - vi[c];
                    }
                    d2 = in->den[p2];
                }
                REAL p2v = F(pressure)(M2->state_eq, d2, c2, M2->target_den);
                REAL gw[4];
                F(gradw)(r, h, k, gw);
                if (!bnd) den += m2 * F(dot4)(v, gw);
                REAL cf[4] = {0, 0, 0, 0};
                int elastic2 = s2 != ARP_INVALID_INDEX && M2->elastic;
                const F(orc_inter)* it = &cx->inter[(size_t)mat_id * sc->nm + mat2];
                REAL cp = F(contact)(it->potential, F(len4)(r), it->radius, it->strength);
                REAL pf = m1 * m2 * (p1 / (d1 * d1) + p2v / (d2 * d2));
                for (int c = 0; c < 4; c++) force[c] += pf * gw[c];
                if (cp != 0) {
                    REAL cs = m2 * cp / d2;
                    for (int c = 0; c < 4; c++) cf[c] += cs * gw[c];
                    if ((elastic || elastic2 || bnd) && mat_id != mat2) {
                        REAL rr = F(dot4)(r, r), rinv[4], nf[4], tv[4];
                        for (int c = 0; c < 4; c++) rinv[c] = r[c] / rr;
                        REAL cfr = F(dot4)(cf, r), vr = F(dot4)(v, r);
                        for (int c = 0; c < 4; c++) {
                            nf[c] = rinv[c] * cfr;
                            tv[c] = v[c] - rinv[c] * vr;
                        }
                        REAL tl = F(len4)(tv);
                        for (int c = 0; c < 4; c++) tv[c] = tv[c] / tl;
                        int bad = 0;
                        for (int c = 0; c < 4; c++)
                            if (!(tv[c] == tv[c]) || ISINF(tv[c])) bad = 1;
                        if (!bad) {
                            REAL fl = it->friction * F(len4)(nf);
                            REAL tl2 = F(len4)(tv);
                            for (int c = 0; c < 4; c++) force[c] += fl * (tv[c] / tl2);
                        }
                    }
                    for (int c = 0; c < 4; c++) force[c] += cf[c];
                }
                REAL rr = F(dot4)(r, r);
                if (!elastic && !bnd) {
                    REAL s = m1 * m2 * (f1 + f2) * F(dot4)(r, gw);
                    REAL den_ = d1 * d2 * (rr + EPS * h * h);
                    for (int c = 0; c < 4; c++) force[c] -= s * v[c] / den_;
                }
                if (!bnd) {
                    REAL ff = mat_id == mat2 ? bf1 + bf2 : it->dampening;
                    REAL s = m1 * m2 * (ff * F(dot4)(v, r) / ((d1 + d2) * (rr + EPS * h * h)));
                    for (int c = 0; c < 4; c++) force[c] -= s * gw[c];
                }
            }
            sden[lane] = den;
            for (int c = 0; c < 4; c++) sf[4 * lane + c] = force[c];
        }
        F(reduce81)(sden, 1);
        F(reduce81)(sf, 4);
    }
    d->den[p] = sden[0];
    for (int c = 0; c < 4; c++) {
        d->vel[4 * (size_t)p + c] = sf[c] / in->den[p] + (c == 1 ? -cx->g : (REAL)0);
        d->pos[4 * (size_t)p + c] = vi[c];
    }
}

/* ---------------- compute_strain (strain.rs:71-181) ---------------------------------- */

static void F(strain_one)(const F(orc_ctx) * cx, const F(orc_grid) * g, const F(orc_state) * in,
                          uint32_t s, REAL* out /* 48 */) {
    const orc_scene* sc = cx->sc;
    const REAL h = cx->h, k = cx->k;
    memset(out, 0, 48 * sizeof(REAL));
    uint32_t p = sc->part[s];
    if (p == ARP_INVALID_INDEX) return;
    if (!cx->mats[sc->mat[p]].elastic) return;
    uint32_t bucket = g->ref_index[p];
    static const int lanes = 81;
    REAL* corr = (REAL*)calloc(lanes * 16, sizeof(REAL));
    REAL* def = (REAL*)calloc(lanes * 16, sizeof(REAL));
    REAL* rate = (REAL*)calloc(lanes * 16, sizeof(REAL));
    const REAL* X1 = in->ref + 4 * (size_t)s;
    if (bucket != ARP_INVALID_INDEX) {
        for (int lane = 0; lane < lanes; lane++) {
            uint64_t b;
            if (!F(lane_cell)(g, bucket, lane, &b)) continue;
            uint32_t sub = (uint32_t)((lane % 9) / 3);
            uint32_t start = g->cnt[3 * b], end = start + g->cnt[3 * b + 1];
            for (uint32_t i = start + sub; i < end; i += 3) {
                uint32_t p2 = F(pidx)(g, b, i);
                uint32_t s2 = sc->sid[p2];
                if (s2 == ARP_INVALID_INDEX) continue;
                const F(orc_mat)* M2 = &cx->mats[sc->mat[p2]];
                if (!M2->elastic) continue;
                REAL V2 = M2->mass / M2->start_den;
                REAL dX[4], gw[4], o[16];
                for (int c = 0; c < 4; c++) dX[c] = in->ref[4 * (size_t)s2 + c] - X1[c];
                F(gradw)(dX, h, k, gw);
                F(mouter)(gw, dX, o);
                for (int j = 0; j < 16; j++) corr[16 * lane + j] += V2 * o[j];
            }
        }
    }
    F(reduce81)(corr, 16);
    REAL Linv[16];
    memcpy(Linv, corr, sizeof Linv);
    for (uint32_t i = cx->dim; i < 4; i++) Linv[5 * i] = 1;
    F(minv)(Linv, Linv);
    if (bucket != ARP_INVALID_INDEX) {
        const REAL* x1 = in->pos + 4 * (size_t)p;
        const REAL* v1 = in->vel + 4 * (size_t)p;
        for (int lane = 0; lane < lanes; lane++) {
            uint64_t b;
            if (!F(lane_cell)(g, bucket, lane, &b)) continue;
            uint32_t sub = (uint32_t)((lane % 9) / 3);
            uint32_t start = g->cnt[3 * b], end = start + g->cnt[3 * b + 1];
            for (uint32_t i = start + sub; i < end; i += 3) {
                uint32_t p2 = F(pidx)(g, b, i);
                uint32_t s2 = sc->sid[p2];
                if (s2 == ARP_INVALID_INDEX) continue;
                const F(orc_mat)* M2 = &cx->mats[sc->mat[p2]];
                if (!M2->elastic) continue;
                REAL V2 = M2->mass / M2->start_den;
                REAL dX[4], dx[4], dv[4], gw[4], gr[4], o[16];
                for (int c = 0; c < 4; c++) {
                    dX[c] = in->ref[4 * (size_t)s2 + c] - X1[c];
                    dx[c] = in->pos[4 * (size_t)p2 + c] - x1[c];
                    dv[c] = in->vel[4 * (size_t)p2 + c] - v1[c];
                }
                F(gradw)(dX, h, k, gw);
                F(mvec)(Linv, gw, gr);
                F(mouter)(dx, gr, o);
                for (int j = 0; j < 16; j++) def[16 * lane + j] += V2 * o[j];
                F(mouter)(dv, gr, o);
                for (int j = 0; j < 16; j++) rate[16 * lane + j] += V2 * o[j];
            }
        }
    }
    F(reduce81)(def, 16);
    F(reduce81)(rate, 16);
    memcpy(out, Linv, 16 * sizeof(REAL));
    memcpy(out + 16, def, 16 * sizeof(REAL));
    memcpy(out + 32, rate, 16 * sizeof(REAL));
    free(corr);
    free(def);
    free(rate);
}

/* PK2 stress of a solid: hooke(E(F)+eps_p) + hooke(E'(F,F')) (solid.rs:334-339) */
static void F(pk2)(const F(orc_ctx) * cx, const F(orc_mat) * M, const REAL* strains,
                   const REAL* eps_p, uint32_t sid, REAL* S) {
    REAL E[16], Ed[16], a[16], b[16];
    /* optional per-solid stiffness scale (an extension of the reference, 1 when absent) */
    const REAL ks = cx->sc->kscale ? (REAL)cx->sc->kscale[sid] : (REAL)1;
    F(mseth_hill)(strains + 16, M->strain_order, cx->dim, E);
    F(madd)(E, eps_p, E);
    F(mseth_hill_rate)(strains + 16, strains + 32, M->strain_order, cx->dim, Ed);
    F(mhooke)(E, ks * M->normal_stiffness, ks * M->shear_stiffness, cx->dim, a);
    F(mhooke)(Ed, M->normal_damp, M->shear_damp, cx->dim, b);
    F(madd)(a, b, S);
}

/* ---------------- solid_forces (solid.rs:276-524) ------------------------------------ */

static void F(solid_one)(const F(orc_ctx) * cx, const F(orc_grid) * g, const F(orc_state) * in,
                         const REAL* strains, uint32_t p, F(orc_state) * d) {
    const orc_scene* sc = cx->sc;
    const REAL h = cx->h, k = cx->k, EPS = (REAL)0.1;
    uint32_t s_id = sc->sid[p], mat_id = sc->mat[p];
    d->den[p] = 0;
    for (int c = 0; c < 4; c++) d->pos[4 * (size_t)p + c] = d->vel[4 * (size_t)p + c] = 0;
    if (s_id == ARP_INVALID_INDEX) return;
    for (int c = 0; c < 4; c++) d->ref[4 * (size_t)s_id + c] = 0;
    for (int c = 0; c < 16; c++) d->stress[16 * (size_t)s_id + c] = 0;
    const F(orc_mat)* M1 = &cx->mats[mat_id];
    REAL m1 = M1->mass, d1 = in->den[p];
    if (!M1->elastic) return;
    REAL sf[81 * 4], sden[81];
    memset(sf, 0, sizeof sf);
    memset(sden, 0, sizeof sden);
    const REAL* S1 = strains + 48 * (size_t)s_id;
    REAL pk2[16], stress[16], t[16];
    F(pk2)(cx, M1, S1, in->stress + 16 * (size_t)s_id, s_id, pk2);
    F(mtrans)(pk2, t);
    F(mmul)(S1 + 16, t, stress);
    const REAL* X1 = in->ref + 4 * (size_t)s_id;
    uint32_t rb = g->ref_index[p];
    if (rb != ARP_INVALID_INDEX) {
        const REAL* corr = S1;
        REAL V1 = m1 / d1;
        REAL A1[16];
        F(mmul)(stress, corr, A1);
        for (int lane = 0; lane < 81; lane++) {
            uint64_t b;
            if (!F(lane_cell)(g, rb, lane, &b)) continue;
            uint32_t sub = (uint32_t)((lane % 9) / 3);
            uint32_t start = g->cnt[3 * b], rc = g->cnt[3 * b + 1];
            for (uint32_t j = start + sub; j < start + rc; j += 3) {
                uint32_t p2 = F(pidx)(g, b, j);
                uint32_t s2 = sc->sid[p2];
                if (sc->mat[p2] != mat_id) continue;
                if (s2 == ARP_INVALID_INDEX) continue;
                REAL V2 = m1 / in->den[p2];
                const REAL* S2 = strains + 48 * (size_t)s2;
                REAL pk2b[16], st2[16], A[16], r[4], gw[4], el[4];
                F(pk2)(cx, M1, S2, in->stress + 16 * (size_t)s2, s2, pk2b);
                F(mtrans)(pk2b, t);
                F(mmul)(S2 + 16, t, st2);
                F(mmul)(st2, S2, A);
                F(madd)(A1, A, A);
                for (int c = 0; c < 4; c++) r[c] = in->ref[4 * (size_t)s2 + c] - X1[c];
                F(gradw)(r, h, k, gw);
                F(mscale)(-V1 * V2, A, A);
                F(mvec)(A, gw, el);
                for (uint32_t c = cx->dim; c < 4; c++) el[c] = 0;
                for (int c = 0; c < 4; c++) sf[4 * lane + c] += el[c];
            }
        }
    }
    uint32_t pb = g->pos_index[p];
    if (pb != ARP_INVALID_INDEX) {
        const REAL* xi = in->pos + 4 * (size_t)p;
        for (int lane = 0; lane < 81; lane++) {
            uint64_t b;
            if (!F(lane_cell)(g, pb, lane, &b)) continue;
            uint32_t sub = (uint32_t)((lane % 9) / 3);
            uint32_t bc = g->cnt[3 * b], rc = g->cnt[3 * b + 1], pc = g->cnt[3 * b + 2];
            /* boundary candidates (j < bc) never reach a force term: the boundary material
             * differs from mat_id, which both the hourglass and the contact term require
             * (solid.rs:430-450), and the friction block adds nothing (:456-461) */
            for (uint32_t j = sub; j < pc + rc + bc; j += 3) {
                if (j < bc + rc) continue;
                uint32_t p2 = F(pidx)(g, b, j);
                uint32_t s2 = sc->sid[p2];
                uint32_t mat2 = sc->mat[p2];
                if (s2 == ARP_INVALID_INDEX) continue;
                const F(orc_mat)* M2 = &cx->mats[mat2];
                REAL m2 = M2->mass;
                REAL r[4];
                for (int c = 0; c < 4; c++) r[c] = in->pos[4 * (size_t)p2 + c] - xi[c];
                REAL d2 = in->den[p2];
                int elastic2 = M2->elastic;
                REAL dX[4];
                for (int c = 0; c < 4; c++) dX[c] = in->ref[4 * (size_t)s2 + c] - X1[c];
                REAL contact = h, r_cut = (REAL)8 * contact;
                int in_contact = (!elastic2 || F(dot4)(dX, dX) > r_cut * r_cut) &&
                                 F(dot4)(r, r) < r_cut * r_cut;
                if (mat_id == mat2 || in_contact) {
                    REAL l = F(len4)(r);
                    if (mat_id == mat2) {
                        const REAL* F1 = S1 + 16;
                        const REAL* F2 = strains + 48 * (size_t)s2 + 16;
                        REAL Fa[16], err[4];
                        F(madd)(F1, F2, Fa);
                        F(mscale)((REAL)0.5, Fa, Fa);
                        F(mvec)(Fa, dX, err);
                        for (int c = 0; c < 4; c++) err[c] -= r[c];
                        REAL s = ((REAL)100000000 * m1 * m2 / (d1 * d2)) *
                                 ((REAL)0.5 * F(dot4)(err, r) / (l + (REAL)10 * EPS * h)) *
                                 ((REAL)1 / (l * l + EPS * h * h)) * F(kern)(dX, h, k);
                        for (int c = 0; c < 4; c++) sf[4 * lane + c] -= s * r[c] / (l + EPS * h);
                    }
                    if (mat_id == mat2 && in_contact && !M1->plastic) {
                        REAL r_geom = contact * contact / r_cut;
                        REAL dr = FMAX(r_cut - l, (REAL)0);
                        REAL s = (REAL)-5000 * SQRT(dr * r_geom);
                        for (int c = 0; c < 4; c++) sf[4 * lane + c] += s * r[c] / l;
                    }
                }
            }
        }
    }
    F(reduce81)(sf, 4);
    const REAL* vi = in->vel + 4 * (size_t)p;
    for (int c = 0; c < 4; c++) {
        d->vel[4 * (size_t)p + c] = sf[c] / in->den[p] + (c == 1 ? -cx->g : (REAL)0);
        d->pos[4 * (size_t)p + c] = vi[c];
    }
    if (M1->plastic) {
        REAL A = M1->yield_strength, B = M1->work_hardening, C = M1->kinematic_hardening;
        REAL n = M1->work_hardening_exp, T = M1->relaxation_time;
        /* eq_plastic_strain_sq multiplies by the integer quotient (dim-1)/dim == 0
         * (solid.rs:212-215) */
        const REAL* ep = in->stress + 16 * (size_t)s_id;
        REAL dev[16];
        REAL tr = F(mtrace)(ep, cx->dim);
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++)
                dev[4 * c + r] = ep[4 * c + r] - (c == r ? tr : (REAL)0) / (REAL)cx->dim;
        REAL eq = (REAL)((cx->dim - 1) / cx->dim) * F(mdotm)(dev, dev);
        REAL sn = F(mdotm)(pk2, pk2);
        REAL ys = (A + B * POW(eq, n / (REAL)2)) * ((REAL)1 + C * LOG((REAL)1));
        if (ys * ys <= sn) {
            sn = SQRT(sn);
            REAL s = (sn - ys) / T;
            for (int c = 0; c < 16; c++) d->stress[16 * (size_t)s_id + c] = s * (pk2[c] / sn);
        }
    }
}

/* ---------------- orchestration (shaders/mod.rs:11-86) ------------------------------- */

static int F(run_fluid)(const F(orc_ctx) * cx, const F(orc_grid) * g, const F(orc_state) * in,
                        F(orc_state) * d) {
    const orc_scene* sc = cx->sc;
    int nt = sc->threads > 0 ? sc->threads : 1;
    (void)nt;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t p = 0; p < (int64_t)sc->n; p++) F(fluid_one)(cx, g, in, (uint32_t)p, d);
    /* clear_solids (solid.rs:5-37): dest.solids = {part_id, 0, 0} */
    for (uint32_t s = 0; s < sc->ns; s++) {
        for (int c = 0; c < 4; c++) d->ref[4 * (size_t)s + c] = 0;
        for (int c = 0; c < 16; c++) d->stress[16 * (size_t)s + c] = 0;
    }
    return 0;
}

static int F(run_strain)(const F(orc_ctx) * cx, const F(orc_grid) * g, const F(orc_state) * in,
                         REAL* strains) {
    const orc_scene* sc = cx->sc;
    int nt = sc->threads > 0 ? sc->threads : 1;
    (void)nt;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t s = 0; s < (int64_t)sc->ns; s++)
        F(strain_one)(cx, g, in, (uint32_t)s, strains + 48 * (size_t)s);
    return 0;
}

static int F(run_solid)(const F(orc_ctx) * cx, const F(orc_grid) * g, const F(orc_state) * in,
                        const REAL* strains, F(orc_state) * d) {
    const orc_scene* sc = cx->sc;
    int nt = sc->threads > 0 ? sc->threads : 1;
    (void)nt;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t p = 0; p < (int64_t)sc->n; p++) F(solid_one)(cx, g, in, strains, (uint32_t)p, d);
    return 0;
}

static int F(alloc_state)(const orc_scene* sc, F(orc_state) * s) {
    size_t n = sc->n ? sc->n : 1, ns = sc->ns ? sc->ns : 1;
    s->den = (REAL*)calloc(n, sizeof(REAL));
    s->pos = (REAL*)calloc(4 * n, sizeof(REAL));
    s->vel = (REAL*)calloc(4 * n, sizeof(REAL));
    s->ref = (REAL*)calloc(4 * ns, sizeof(REAL));
    s->stress = (REAL*)calloc(16 * ns, sizeof(REAL));
    return (s->den && s->pos && s->vel && s->ref && s->stress) ? 0 : -1;
}

static void F(free_state)(F(orc_state) * s) {
    free(s->den);
    free(s->pos);
    free(s->vel);
    free(s->ref);
    free(s->stress);
}

int F(orc_fluid_forces)(const orc_scene* sc, const F(orc_state) * in, F(orc_state) * d) {
    F(orc_ctx) cx;
    F(orc_grid) g;
    if (F(ctx_init)(&cx, sc)) return -1;
    if (F(grid_build)(sc, in->pos, sc->ns ? in->ref : NULL, &g)) return -1;
    F(run_fluid)(&cx, &g, in, d);
    F(grid_free)(&g);
    F(ctx_free)(&cx);
    return 0;
}

int F(orc_strain)(const orc_scene* sc, const F(orc_state) * in, REAL* strains) {
    F(orc_ctx) cx;
    F(orc_grid) g;
    if (F(ctx_init)(&cx, sc)) return -1;
    if (F(grid_build)(sc, in->pos, in->ref, &g)) return -1;
    F(run_strain)(&cx, &g, in, strains);
    F(grid_free)(&g);
    F(ctx_free)(&cx);
    return 0;
}

int F(orc_solid_forces)(const orc_scene* sc, const F(orc_state) * in, F(orc_state) * d) {
    F(orc_ctx) cx;
    F(orc_grid) g;
    if (F(ctx_init)(&cx, sc)) return -1;
    if (F(grid_build)(sc, in->pos, in->ref, &g)) return -1;
    REAL* strains = (REAL*)calloc(48 * (size_t)(sc->ns ? sc->ns : 1), sizeof(REAL));
    F(run_strain)(&cx, &g, in, strains);
    F(run_solid)(&cx, &g, in, strains, d);
    free(strains);
    F(grid_free)(&g);
    F(ctx_free)(&cx);
    return 0;
}

/* compute_forces: fluid dest, then dest2 from the solid path when solids.len() > 1,
 * returned as the sum dest + dest2 (shaders/mod.rs:51-79, particle_state/mod.rs:129) */
int F(orc_compute_forces)(const orc_scene* sc, const F(orc_state) * in, F(orc_state) * d) {
    F(orc_ctx) cx;
    F(orc_grid) g;
    if (F(ctx_init)(&cx, sc)) return -1;
    if (F(grid_build)(sc, in->pos, sc->ns ? in->ref : NULL, &g)) return -1;
    F(run_fluid)(&cx, &g, in, d);
    if (sc->ns > 1) {
        F(orc_state) d2;
        if (F(alloc_state)(sc, &d2)) return -1;
        REAL* strains = (REAL*)calloc(48 * (size_t)sc->ns, sizeof(REAL));
        F(run_strain)(&cx, &g, in, strains);
        F(run_solid)(&cx, &g, in, strains, &d2);
        for (size_t i = 0; i < sc->n; i++) {
            d->den[i] = (REAL)1 * d->den[i] + (REAL)1 * d2.den[i];
            for (int c = 0; c < 4; c++) {
                d->pos[4 * i + c] = (REAL)1 * d->pos[4 * i + c] + (REAL)1 * d2.pos[4 * i + c];
                d->vel[4 * i + c] = (REAL)1 * d->vel[4 * i + c] + (REAL)1 * d2.vel[4 * i + c];
            }
        }
        for (size_t s = 0; s < sc->ns; s++) {
            for (int c = 0; c < 4; c++) d->ref[4 * s + c] += d2.ref[4 * s + c];
            for (int c = 0; c < 16; c++) d->stress[16 * s + c] += d2.stress[16 * s + c];
        }
        free(strains);
        F(free_state)(&d2);
    }
    F(grid_free)(&g);
    F(ctx_free)(&cx);
    return 0;
}

/* lin_comb_N over <= 5 terms: dst = r0*p0 + r1*p1 + ... (particles.rs:81-84, 146-147) */
static void F(lc_chunk)(const orc_scene* sc, int nt, const REAL* coef,
                        const F(orc_state) * const* t, F(orc_state) * o) {
    for (size_t i = 0; i < sc->n; i++) {
        REAL a = coef[0] * t[0]->den[i];
        for (int k = 1; k < nt; k++) a = a + coef[k] * t[k]->den[i];
        o->den[i] = a;
        for (int c = 0; c < 4; c++) {
            REAL x = coef[0] * t[0]->pos[4 * i + c], v = coef[0] * t[0]->vel[4 * i + c];
            for (int k = 1; k < nt; k++) {
                x = x + coef[k] * t[k]->pos[4 * i + c];
                v = v + coef[k] * t[k]->vel[4 * i + c];
            }
            o->pos[4 * i + c] = x;
            o->vel[4 * i + c] = v;
        }
    }
    for (size_t s = 0; s < sc->ns; s++) {
        for (int c = 0; c < 4; c++) {
            REAL x = coef[0] * t[0]->ref[4 * s + c];
            for (int k = 1; k < nt; k++) x = x + coef[k] * t[k]->ref[4 * s + c];
            o->ref[4 * s + c] = x;
        }
        for (int c = 0; c < 16; c++) {
            REAL x = coef[0] * t[0]->stress[16 * s + c];
            for (int k = 1; k < nt; k++) x = x + coef[k] * t[k]->stress[16 * s + c];
            o->stress[16 * s + c] = x;
        }
    }
}

static void F(copy_state)(const orc_scene* sc, const F(orc_state) * a, F(orc_state) * b) {
    memcpy(b->den, a->den, sc->n * sizeof(REAL));
    memcpy(b->pos, a->pos, 4 * sc->n * sizeof(REAL));
    memcpy(b->vel, a->vel, 4 * sc->n * sizeof(REAL));
    memcpy(b->ref, a->ref, 4 * sc->ns * sizeof(REAL));
    memcpy(b->stress, a->stress, 16 * sc->ns * sizeof(REAL));
}

/* reduce_ref / reduce_owned (particles.rs:381-417): chunks of 5, recursively combined with
 * coefficient 1 until one term with coefficient 1 remains. */
int F(orc_lincomb)(const orc_scene* sc, int nterms, const REAL* coef,
                   const F(orc_state) * const* terms, F(orc_state) * out) {
    if (nterms <= 0) return -1;
    if (nterms <= 5) {
        if (nterms == 1 && coef[0] == (REAL)1) {
            F(lc_chunk)(sc, 1, coef, terms, out);
            return 0;
        }
        F(lc_chunk)(sc, nterms, coef, terms, out);
        return 0;
    }
    int nch = (nterms + 4) / 5;
    F(orc_state)* tmp = (F(orc_state)*)calloc(nch, sizeof(F(orc_state)));
    const F(orc_state)** tp = (const F(orc_state)**)calloc(nch, sizeof(void*));
    REAL* ones = (REAL*)calloc(nch, sizeof(REAL));
    for (int c = 0; c < nch; c++) {
        F(alloc_state)(sc, &tmp[c]);
        int k0 = 5 * c, nk = nterms - k0 < 5 ? nterms - k0 : 5;
        F(lc_chunk)(sc, nk, coef + k0, terms + k0, &tmp[c]);
        tp[c] = &tmp[c];
        ones[c] = 1;
    }
    int rc = F(orc_lincomb)(sc, nch, ones, tp, out);
    for (int c = 0; c < nch; c++) F(free_state)(&tmp[c]);
    free(tmp);
    free(tp);
    free(ones);
    return rc;
}

/* One explicit RK step of the tableau on y (in place). */
int F(orc_step)(const orc_scene* sc, const ArpTableau* tab, REAL dt, int nsteps,
                F(orc_state) * y) {
    int S = tab->stages;
    if (S < 1 || S > ARP_MAX_STAGES) return -1;
    F(orc_state) k[ARP_MAX_STAGES], x;
    for (int s = 0; s < S; s++)
        if (F(alloc_state)(sc, &k[s])) return -1;
    if (F(alloc_state)(sc, &x)) return -1;
    for (int step = 0; step < nsteps; step++) {
        for (int s = 0; s < S; s++) {
            const F(orc_state)* terms[ARP_MAX_STAGES + 1];
            REAL co[ARP_MAX_STAGES + 1];
            int nt = 0;
            terms[nt] = y;
            co[nt++] = 1;
            for (int j = 0; j < s; j++)
                if (tab->a[s][j] != 0.0f) {
                    terms[nt] = &k[j];
                    co[nt++] = dt * (REAL)tab->a[s][j];
                }
            if (nt == 1)
                F(copy_state)(sc, y, &x);
            else
                F(orc_lincomb)(sc, nt, co, terms, &x);
            if (F(orc_compute_forces)(sc, &x, &k[s])) return -1;
        }
        const F(orc_state)* terms[ARP_MAX_STAGES + 1];
        REAL co[ARP_MAX_STAGES + 1];
        int nt = 0;
        terms[nt] = y;
        co[nt++] = 1;
        for (int j = 0; j < S; j++)
            if (tab->b[j] != 0.0f) {
                terms[nt] = &k[j];
                co[nt++] = dt * (REAL)tab->b[j];
            }
        F(orc_lincomb)(sc, nt, co, terms, &x);
        F(copy_state)(sc, &x, y);
    }
    for (int s = 0; s < S; s++) F(free_state)(&k[s]);
    F(free_state)(&x);
    return 0;
}

#undef CAT_
#undef CAT
#undef F
