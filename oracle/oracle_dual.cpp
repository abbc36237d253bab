/*
 * oracle_dual.cpp -- forward-mode derivative build of the CPU restatement.  TEST
 * INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/ as the checker of the GPU's
 * reverse mode, never by the product path.
 *
 * oracle_impl.h is compiled a third time with REAL = a dual number (an FP64 value and one
 * FP64 tangent) whose arithmetic carries the exact derivative of every operation.  One run
 * from a state y with tangent u (and, optionally, a tangent on one material parameter)
 * gives the state after nsteps together with J u (+ dy/dp): exact directional derivatives
 * of the FP64 step, with no finite-difference truncation.  A reverse-mode gradient G =
 * J^T g is then checked through <G, u> = <g, J u> for any number of directions.
 *
 * Non-differentiable points follow the FP64 oracle's branches (comparisons use the value):
 * the grid, the kernel's support cut-off max(h^2 - r^2, 0), the contact potentials'
 * cases and max(), with the one-sided derivative of the branch taken.
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "oracle.h"

struct dual {
    double v, d;
    dual() = default;
    constexpr dual(double x) : v(x), d(0.0) {}
    constexpr dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    explicit operator uint32_t() const { return (uint32_t)v; }
    dual& operator+=(const dual& o) {
        v += o.v;
        d += o.d;
        return *this;
    }
    dual& operator-=(const dual& o) {
        v -= o.v;
        d -= o.d;
        return *this;
    }
    dual& operator*=(const dual& o) {
        d = d * o.v + v * o.d;
        v *= o.v;
        return *this;
    }
    dual& operator/=(const dual& o) {
        const double q = v / o.v;
        d = (d - q * o.d) / o.v;
        v = q;
        return *this;
    }
};
inline dual operator+(dual a, const dual& b) { return a += b; }
inline dual operator-(dual a, const dual& b) { return a -= b; }
inline dual operator*(dual a, const dual& b) { return a *= b; }
inline dual operator/(dual a, const dual& b) { return a /= b; }
inline dual operator-(const dual& a) { return dual(-a.v, -a.d); }
inline dual operator+(const dual& a) { return a; }
inline bool operator<(const dual& a, const dual& b) { return a.v < b.v; }
inline bool operator<=(const dual& a, const dual& b) { return a.v <= b.v; }
inline bool operator>(const dual& a, const dual& b) { return a.v > b.v; }
inline bool operator>=(const dual& a, const dual& b) { return a.v >= b.v; }
inline bool operator==(const dual& a, const dual& b) { return a.v == b.v; }
inline bool operator!=(const dual& a, const dual& b) { return a.v != b.v; }

inline dual dsqrt(const dual& a) {
    const double s = std::sqrt(a.v);
    // at 0 with a zero tangent (the length of the self pair's r = 0): derivative 0
    return dual(s, a.d == 0.0 ? 0.0 : a.d / (2.0 * s));
}
inline dual dpow(const dual& a, const dual& b) {
    const double p = std::pow(a.v, b.v);
    // d(a^b) = b a^(b-1) da + a^b ln(a) db (the second term only when b varies)
    double d = b.v == 0.0 ? 0.0 : b.v * std::pow(a.v, b.v - 1.0) * a.d;
    if (b.d != 0.0) d += p * std::log(a.v) * b.d;
    return dual(p, d);
}
inline dual dlog(const dual& a) { return dual(std::log(a.v), a.d / a.v); }
inline dual dfmax(const dual& a, const dual& b) { return a.v >= b.v ? a : b; }
inline dual dfmin(const dual& a, const dual& b) { return a.v <= b.v ? a : b; }
inline dual dfloor(const dual& a) { return dual(std::floor(a.v)); }
inline bool disinf(const dual& a) { return std::isinf(a.v); }

typedef struct orc_state_x {
    dual *den, *pos, *vel, *ref, *stress;
} orc_state_x;

inline dual orc_norm_const_x(uint32_t dim, dual h) { return dual(orc_norm_const_d(dim, h.v)); }

// the material parameter whose tangent a run seeds (set by orc_jvp_step, read by ctx_init)
static int g_seed_mat = -1, g_seed_field = 0;

#define REAL dual
#define SFX _x
#define SQRT dsqrt
#define POW dpow
#define LOG dlog
#define FMAX dfmax
#define FMIN dfmin
#define FLOOR dfloor
#define ISINF disinf
#define ORC_CTX_HOOK(c) orc_seed_params(c)
struct orc_ctx_x;
static void orc_seed_params(struct orc_ctx_x* c);
#include "oracle_impl.h"

/* field codes: 1 bulk_visc, 2 visc, 3 normal_stiffness, 4 shear_stiffness, 5 normal_damp,
 * 6 shear_damp (the order of arp_adjoint_backward's parameter gradients) */
static void orc_seed_params(orc_ctx_x* c) {
    if (g_seed_mat < 0 || (uint32_t)g_seed_mat >= c->sc->nm) return;
    orc_mat_x* m = &c->mats[g_seed_mat];
    dual* f = g_seed_field == 1 ? &m->bulk_visc : g_seed_field == 2 ? &m->visc
            : g_seed_field == 3 ? &m->normal_stiffness : g_seed_field == 4 ? &m->shear_stiffness
            : g_seed_field == 5 ? &m->normal_damp : g_seed_field == 6 ? &m->shear_damp : nullptr;
    if (f) f->d = 1.0;
}

/* nsteps steps of the tableau from y (FP64 values, in place) with tangent ydot (in place: on
 * return J ydot, plus dy/dp when param_mat >= 0 seeds material param_mat's field
 * param_field).  Returns 0, or -1 on a bad argument or allocation failure. */
extern "C" int orc_jvp_step(const orc_scene* sc, const ArpTableau* tab, double dt, int nsteps,
                            orc_state_d* y, orc_state_d* ydot, int param_mat, int param_field) {
    if (!sc || !tab || !y || !ydot || tab->stages < 1 || tab->stages > ARP_MAX_STAGES) return -1;
    orc_state_x s;
    if (alloc_state_x(sc, &s)) return -1;
    const size_t n = sc->n, ns = sc->ns;
    auto in = [](dual* dst, const double* v, const double* t, size_t k) {
        for (size_t i = 0; i < k; i++) dst[i] = dual(v[i], t[i]);
    };
    auto out = [](const dual* src, double* v, double* t, size_t k) {
        for (size_t i = 0; i < k; i++) {
            v[i] = src[i].v;
            t[i] = src[i].d;
        }
    };
    in(s.den, y->den, ydot->den, n);
    in(s.pos, y->pos, ydot->pos, 4 * n);
    in(s.vel, y->vel, ydot->vel, 4 * n);
    in(s.ref, y->ref, ydot->ref, 4 * ns);
    in(s.stress, y->stress, ydot->stress, 16 * ns);
    g_seed_mat = param_mat;
    g_seed_field = param_field;
    const int rc = orc_step_x(sc, tab, dual(dt), nsteps, &s);
    g_seed_mat = -1;
    g_seed_field = 0;
    out(s.den, y->den, ydot->den, n);
    out(s.pos, y->pos, ydot->pos, 4 * n);
    out(s.vel, y->vel, ydot->vel, 4 * n);
    out(s.ref, y->ref, ydot->ref, 4 * ns);
    out(s.stress, y->stress, ydot->stress, 16 * ns);
    free_state_x(&s);
    return rc;
}
