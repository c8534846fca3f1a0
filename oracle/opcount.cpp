/*
 * opcount.cpp -- TEST INFRASTRUCTURE ONLY: the algorithmic operation count
 * of one RK4 stage, for the roofline's FLOP unit.  Never part of the
 * product; bench.py reads the counts this produces (frozen in
 * profiles/algorithmic_counts.json), the product never loads it.
 *
 * The oracle (grhd_oracle.c, a restatement of the reference's numerics) is
 * compiled here as C++ with every `double` replaced by `Real`, a double that
 * counts its own arithmetic: +, -, *, / and sqrt each count one operation
 * (negation, comparisons, fabs and selects count none).  The count is then
 * taken over the de-duplicated form of the reference's call structure
 * (FixedBGFArrayBoxFiniteVolume.hpp:199-363 evaluates each face twice and
 * the metric ~23 times per cell; here each is evaluated once), per
 * interior cell of one stage:
 *   - one full metric evaluation (KerrSchild / CircularBinaryPN
 *     compute_metric_background, then the geometry and derivative
 *     adapters of FixedBGMetric.hpp with one inverse of gamma);
 *   - per direction: the cell's MC-theta slopes and one face (the
 *     cell's low face): the mean face geometry, then compute_muscl_hlle_flux
 *     (FiniteVolume.hpp:63-98);
 *   - the flux divergence, -inverse_dx (F_hi - F_lo) summed over the three
 *     directions;
 *   - the static-metric sources on the floored cell primitives (skipped
 *     for metric-source atmosphere, StaticMetric.hpp:214-275) and, when on,
 *     the volume-time sources (two more metric evaluations);
 *   - the RK4 stage update (accumulator and stage state);
 *   - the post-update recovery of the new state at the iterations it
 *     takes, the floors, and for a reset cell the conserved variables of
 *     the atmosphere state and their recovery.
 * Ghost-cell work (outflow fill, recovery over the halo) is overhead, not
 * counted.  con2prim is counted as the reference writes it (each Newton
 * step evaluates the residual and then its derivative separately).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stddef.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <cstddef>
#include <vector>

namespace opcount
{
struct Counts
{
    unsigned long long add, mul, div, sqrt, other;
};
static Counts g_counts = {0, 0, 0, 0, 0};
static bool g_on = false;
} // namespace opcount

/* A double that counts its arithmetic while opcount::g_on is set. */
struct Real
{
    double v;
    Real() = default;
    constexpr Real(double x) : v(x) {}
    constexpr Real(int x) : v((double)x) {}
    constexpr Real(long x) : v((double)x) {}
    constexpr Real(long long x) : v((double)x) {}
    constexpr Real(unsigned x) : v((double)x) {}
    constexpr Real(unsigned long long x) : v((double)x) {}
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator int() const { return (int)v; }
    explicit constexpr operator long long() const { return (long long)v; }
    Real &operator+=(Real o);
    Real &operator-=(Real o);
    Real &operator*=(Real o);
    Real &operator/=(Real o);
};

#define OPC_TICK(field)                                                       \
    do                                                                        \
    {                                                                         \
        if (opcount::g_on)                                                    \
            ++opcount::g_counts.field;                                        \
    } while (0)

inline Real operator+(Real a, Real b)
{
    OPC_TICK(add);
    return Real(a.v + b.v);
}
inline Real operator-(Real a, Real b)
{
    OPC_TICK(add);
    return Real(a.v - b.v);
}
inline Real operator*(Real a, Real b)
{
    OPC_TICK(mul);
    return Real(a.v * b.v);
}
inline Real operator/(Real a, Real b)
{
    OPC_TICK(div);
    return Real(a.v / b.v);
}
inline Real operator-(Real a) { return Real(-a.v); }
inline Real operator+(Real a) { return a; }
inline Real &Real::operator+=(Real o) { return *this = *this + o; }
inline Real &Real::operator-=(Real o) { return *this = *this - o; }
inline Real &Real::operator*=(Real o) { return *this = *this * o; }
inline Real &Real::operator/=(Real o) { return *this = *this / o; }
inline bool operator<(Real a, Real b) { return a.v < b.v; }
inline bool operator>(Real a, Real b) { return a.v > b.v; }
inline bool operator<=(Real a, Real b) { return a.v <= b.v; }
inline bool operator>=(Real a, Real b) { return a.v >= b.v; }
inline bool operator==(Real a, Real b) { return a.v == b.v; }
inline bool operator!=(Real a, Real b) { return a.v != b.v; }
inline Real sqrt(Real a)
{
    OPC_TICK(sqrt);
    return Real(std::sqrt(a.v));
}
inline Real fabs(Real a) { return Real(std::fabs(a.v)); }
inline bool isfinite(Real a) { return std::isfinite(a.v); }
inline Real frexp(Real a, int *e) { return Real(std::frexp(a.v, e)); }
inline Real cos(Real a)
{
    OPC_TICK(other);
    return Real(std::cos(a.v));
}
inline Real sin(Real a)
{
    OPC_TICK(other);
    return Real(std::sin(a.v));
}
inline Real pow(Real a, Real b)
{
    OPC_TICK(other);
    return Real(std::pow(a.v, b.v));
}

#define double Real
#include "grhd_oracle.c"
#undef double

namespace
{

struct Tally
{
    unsigned long long n[5];
};

Tally now()
{
    const opcount::Counts &c = opcount::g_counts;
    return Tally{{c.add, c.mul, c.div, c.sqrt, c.other}};
}

/* part[k] += now - since */
void charge(Tally since, unsigned long long *part)
{
    const Tally t = now();
    for (int i = 0; i < 5; ++i)
        part[i] += t.n[i] - since.n[i];
}

/* FixedBGMetric.hpp:32-55 with the geometry's gamma^ij (one inverse per
 * metric evaluation, as a de-duplicated evaluation needs). */
void derivs_with_inverse(const adm_t *v, const geom_t *g, derivs_t *d)
{
    for (int i = 0; i < 3; ++i)
    {
        d->lapse[i] = v->d1_lapse[i];
        for (int j = 0; j < 3; ++j)
            d->shift[i][j] = v->d1_shift[i][j];
    }
    for (int dir = 0; dir < 3; ++dir)
    {
        d->log_sqrt_det[dir] = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                d->g[dir][i][j] = v->d1_gamma[i][j][dir];
                d->log_sqrt_det[dir] += 0.5 * g->gu[i][j] *
                                        v->d1_gamma[i][j][dir];
            }
    }
}

} // namespace

/* Parts of the count, in the order of the out[] rows below. */
enum
{
    P_METRIC = 0,
    P_SLOPES = 1,
    P_FACES = 2,
    P_DIVERGENCE = 3,
    P_SOURCES = 4,
    P_UPDATE = 5,
    P_RECOVERY = 6,
    NPART = 7
};

/* One RK4 stage over the interior of box b from `state` (conserved and
 * primitive components current, ghosts included), stage 2's update form
 * (acc += dt/3 k, U = U0 + dt/2 k).  out[part * 5 + kind] with kind = add
 * (incl. subtract), mul, div, sqrt, other (trig, pow); out[NPART * 5] =
 * interior cells; out[NPART * 5 + 1] = Newton iterations of the
 * recoveries; out[NPART * 5 + 2] = metric-source atmosphere cells;
 * out[NPART * 5 + 3] = reset cells.  Returns 0 or an oracle error code. */
extern "C" int grhd_opcount_rk4_stage(const grhd_box *b, const grhd_params *o,
                                      const grhd_background *bgp, Real time,
                                      Real dt, const Real *state, int ncomp,
                                      int cb, unsigned long long *out)
{
    int rc = check_args(b, o, bgp, ncomp, cb);
    if (rc)
        return rc;
    memset(out, 0, sizeof(unsigned long long) * (NPART * 5 + 4));
    unsigned long long *part = out;
    opcount::g_on = false;
    bg_eval_t bg, bgp_plus, bgp_minus;
    background_at(bgp, time, &bg);
    const Real h = std_max(o->metric_time_derivative_step, 1.0e-12);
    if (o->use_metric_volume_time_sources)
    {
        background_at(bgp, time + h, &bgp_plus);
        background_at(bgp, time - h, &bgp_minus);
    }
    const fab_shape_t ss = shape_of(b->glo, b->ghi);
    const Real gamma = o->adiabatic_index;

    /* geometry and primitives of every storage cell (not counted: the
     * interior cells' own metric evaluation is counted below) */
    std::vector<geom_t> geo((size_t)ss.npts);
    std::vector<prim_t> W((size_t)ss.npts);
    for (int k = b->glo[2]; k <= b->ghi[2]; ++k)
        for (int j = b->glo[1]; j <= b->ghi[1]; ++j)
            for (int i = b->glo[0]; i <= b->ghi[0]; ++i)
            {
                const long long off = fab_off(&ss, i, j, k);
                load_geometry(&bg, i, j, k, b->dx, b->center, &geo[off]);
                load_prim(state, &ss, cb, off, &W[off]);
            }
    /* slopes of every cell that a counted face reads (the interior and
     * one layer below it in each direction), counted for interior cells */
    std::vector<prim_t> slope[3];
    for (int d = 0; d < 3; ++d)
        slope[d].assign((size_t)ss.npts, prim_t());

    const Real c_acc = dt / 3.0, c_stage = dt / 2.0;
    unsigned long long ncells = 0, iters = 0, atmos = 0, resets = 0;
    for (int k = b->lo[2] - 1; k <= b->hi[2]; ++k)
        for (int j = b->lo[1] - 1; j <= b->hi[1]; ++j)
            for (int i = b->lo[0] - 1; i <= b->hi[0]; ++i)
            {
                const int iv[3] = {i, j, k};
                const long long off = fab_off(&ss, i, j, k);
                const bool interior = i >= b->lo[0] && j >= b->lo[1] &&
                                      k >= b->lo[2];
                for (int d = 0; d < 3; ++d)
                {
                    if (!o->use_reconstruction)
                        continue;
                    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    lo[d] -= 1;
                    hi[d] += 1;
                    if (lo[d] < b->glo[d] || hi[d] > b->ghi[d])
                        continue;
                    (void)iv;
                    opcount::g_on = interior;
                    const Tally t0 = now();
                    limited_slope_prim(&W[fab_off(&ss, lo[0], lo[1], lo[2])],
                                       &W[off],
                                       &W[fab_off(&ss, hi[0], hi[1], hi[2])],
                                       o->limiter_theta, &slope[d][off]);
                    charge(t0, part + P_SLOPES * 5);
                    opcount::g_on = false;
                }
            }

    for (int k = b->lo[2]; k <= b->hi[2]; ++k)
        for (int j = b->lo[1]; j <= b->hi[1]; ++j)
            for (int i = b->lo[0]; i <= b->hi[0]; ++i)
            {
                const long long off = fab_off(&ss, i, j, k);
                opcount::g_on = true;
                ++ncells;

                /* one full metric evaluation */
                Tally t0 = now();
                Real x[3];
                cell_coords(i, j, k, b->dx, b->center, x);
                adm_t adm;
                background_metric(&bg, x, &adm);
                geom_t g;
                geometry_from_adm(&adm, &g);
                derivs_t dv;
                derivs_with_inverse(&adm, &g, &dv);
                charge(t0, part + P_METRIC * 5);

                /* the low face of each direction, and the divergence of
                 * that face's flux against the neighbour's (the high face
                 * is the next cell's low face: counted there) */
                cons_t rhs;
                rhs.D = 0.0;
                rhs.tau = 0.0;
                rhs.S[0] = rhs.S[1] = rhs.S[2] = 0.0;
                for (int d = 0; d < 3; ++d)
                {
                    int l[3] = {i, j, k};
                    l[d] -= 1;
                    const long long lo = fab_off(&ss, l[0], l[1], l[2]);
                    t0 = now();
                    geom_t face;
                    average_geometry(&geo[lo], &geo[off], &face);
                    cons_t flo;
                    muscl_hlle_flux(&W[lo], &W[off], &slope[d][lo],
                                    &slope[d][off], gamma, &face, d, o, &flo);
                    charge(t0, part + P_FACES * 5);
                    /* the high face's value from the same face of the
                     * next cell, uncounted */
                    int h3[3] = {i, j, k};
                    h3[d] += 1;
                    const long long hi = fab_off(&ss, h3[0], h3[1], h3[2]);
                    opcount::g_on = false;
                    geom_t fh;
                    average_geometry(&geo[off], &geo[hi], &fh);
                    prim_t shi;
                    if (o->use_reconstruction)
                    {
                        int hh[3] = {h3[0], h3[1], h3[2]};
                        hh[d] += 1;
                        limited_slope_prim(&W[off], &W[hi],
                                           &W[fab_off(&ss, hh[0], hh[1],
                                                      hh[2])],
                                           o->limiter_theta, &shi);
                    }
                    cons_t fhi;
                    muscl_hlle_flux(&W[off], &W[hi], &slope[d][off], &shi,
                                    gamma, &fh, d, o, &fhi);
                    const Real inv = inverse_dx_of(b, d); /* per box */
                    opcount::g_on = true;
                    t0 = now();
                    rhs.D += -inv * (fhi.D - flo.D);
                    rhs.tau += -inv * (fhi.tau - flo.tau);
                    for (int m = 0; m < 3; ++m)
                        rhs.S[m] += -inv * (fhi.S[m] - flo.S[m]);
                    charge(t0, part + P_DIVERGENCE * 5);
                }

                /* sources */
                cons_t U0;
                load_cons(state, &ss, cb, off, &U0);
                t0 = now();
                if (o->use_static_metric_sources)
                {
                    prim_t pf = W[off];
                    enforce_primitive_floors(&pf, gamma, g.g, o);
                    if (is_atmosphere_primitive(
                            &pf, o, o->metric_source_atmosphere_factor))
                        ++atmos;
                    else
                    {
                        cons_t src;
                        static_metric_source_terms(&pf, &U0, gamma, &g, &dv,
                                                   &src);
                        rhs.D += src.D;
                        rhs.tau += src.tau;
                        for (int m = 0; m < 3; ++m)
                            rhs.S[m] += src.S[m];
                    }
                }
                if (o->use_metric_volume_time_sources)
                {
                    adm_t vp, vm;
                    geom_t gp, gm;
                    background_metric(&bgp_plus, x, &vp);
                    geometry_from_adm(&vp, &gp);
                    background_metric(&bgp_minus, x, &vm);
                    geometry_from_adm(&vm, &gm);
                    const Real lp = spatial_metric_log_sqrt_det(&gp);
                    const Real lm = spatial_metric_log_sqrt_det(&gm);
                    const Real dlog = (lp - lm) / (2.0 * h);
                    rhs.D += 0.0 - U0.D * dlog;
                    rhs.tau += 0.0 - U0.tau * dlog;
                    for (int m = 0; m < 3; ++m)
                        rhs.S[m] += 0.0 - U0.S[m] * dlog;
                }
                charge(t0, part + P_SOURCES * 5);

                /* stage update */
                t0 = now();
                cons_t acc = U0, un;
                acc.D += c_acc * rhs.D;
                acc.tau += c_acc * rhs.tau;
                un.D = U0.D + c_stage * rhs.D;
                un.tau = U0.tau + c_stage * rhs.tau;
                for (int m = 0; m < 3; ++m)
                {
                    acc.S[m] += c_acc * rhs.S[m];
                    un.S[m] = U0.S[m] + c_stage * rhs.S[m];
                }
                charge(t0, part + P_UPDATE * 5);

                /* post-update recovery */
                t0 = now();
                recovery_t rec;
                recover_primitive(&un, gamma, g.gu, o, &rec);
                iters += (unsigned long long)rec.iterations;
                prim_t prim;
                bool reset = !rec.success;
                if (rec.success)
                {
                    prim = rec.prim;
                    enforce_primitive_floors(&prim, gamma, g.g, o);
                    reset = primitive_differs(&rec.prim, &prim);
                }
                else
                    make_atmosphere_primitive(gamma, o, &prim);
                if (reset)
                {
                    ++resets;
                    cons_t c2;
                    compute_conserved(&prim, gamma, g.g, &c2);
                    recovery_t rec2;
                    recover_primitive(&c2, gamma, g.gu, o, &rec2);
                    iters += (unsigned long long)rec2.iterations;
                }
                charge(t0, part + P_RECOVERY * 5);
                opcount::g_on = false;
            }
    out[NPART * 5] = ncells;
    out[NPART * 5 + 1] = iters;
    out[NPART * 5 + 2] = atmos;
    out[NPART * 5 + 3] = resets;
    return 0;
}
