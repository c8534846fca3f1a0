/*
 * grhd_adjoint.hpp -- device side of the discrete adjoint of the RK4 step.
 *
 * Every kernel of the forward step is a function of a few cells' data (a
 * face flux of its 4-cell stencil, a cell's sources and recovery of that
 * cell).  The adjoint needs the transposed local Jacobians of these
 * functions applied to adjoint vectors.
 * - The face flux's, in its two face states and Gamma, is applied in one
 *   hand-written reverse sweep (face_flux_reverse); the reconstruction's
 *   per-component Jacobian then takes it to the four stencil cells.
 * - The others, and the spin's tangent everywhere, are formed column by
 *   column in forward mode: the same templated physics as the forward
 *   kernels is evaluated on Dual numbers (value + one tangent), once per
 *   local input, and the tangent of the outputs is dotted with the
 *   adjoint of the outputs.
 * Branches (limiter choice, floors, HLLE wave-speed signs, atmosphere
 * skips and resets, the v^2 clamp) compare values only, so they are frozen
 * at the forward pass's decisions.  con2prim is differentiated at its
 * converged root by the implicit-function theorem (con2prim_ift).
 *
 * Scope: one box with outflow (physical) ghosts on a Kerr-Schild
 * background, FP64 faces, no volume-time sources.  The parameters that
 * carry adjoints are the adiabatic index Gamma and the spin a.
 */
#ifndef GRHD_ADJOINT_HPP_
#define GRHD_ADJOINT_HPP_

#include "grhd_stage_kernels.hpp"

namespace grhd
{

/* ------------------------------------------------------------------ */
/* Forward-mode dual number (one tangent).                             */
/* ------------------------------------------------------------------ */
struct Dual
{
    double v, d;
    Dual() = default;
    __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr Dual(double x, double t) : v(x), d(t) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b)
{
    return Dual(a.v + b.v, a.d + b.d);
}
__device__ __forceinline__ Dual operator-(Dual a, Dual b)
{
    return Dual(a.v - b.v, a.d - b.d);
}
__device__ __forceinline__ Dual operator*(Dual a, Dual b)
{
    return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, double b)
{
    return Dual(a.v + b, a.d);
}
__device__ __forceinline__ Dual operator+(double a, Dual b)
{
    return Dual(a + b.v, b.d);
}
__device__ __forceinline__ Dual operator-(Dual a, double b)
{
    return Dual(a.v - b, a.d);
}
__device__ __forceinline__ Dual operator-(double a, Dual b)
{
    return Dual(a - b.v, -b.d);
}
__device__ __forceinline__ Dual operator*(Dual a, double b)
{
    return Dual(a.v * b, a.d * b);
}
__device__ __forceinline__ Dual operator*(double a, Dual b)
{
    return Dual(a * b.v, a * b.d);
}
__device__ __forceinline__ Dual operator/(Dual a, double b)
{
    return Dual(a.v / b, a.d / b);
}
__device__ __forceinline__ Dual operator/(double a, Dual b)
{
    const double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b)
{
    a = a + b;
    return a;
}
__device__ __forceinline__ Dual &operator-=(Dual &a, Dual b)
{
    a = a - b;
    return a;
}
__device__ __forceinline__ Dual &operator*=(Dual &a, Dual b)
{
    a = a * b;
    return a;
}
#define GRHD_DUAL_CMP(op)                                                     \
    __device__ __forceinline__ bool operator op(Dual a, Dual b)               \
    {                                                                         \
        return a.v op b.v;                                                    \
    }                                                                         \
    __device__ __forceinline__ bool operator op(Dual a, double b)             \
    {                                                                         \
        return a.v op b;                                                      \
    }                                                                         \
    __device__ __forceinline__ bool operator op(double a, Dual b)             \
    {                                                                         \
        return a op b.v;                                                      \
    }
GRHD_DUAL_CMP(<)
GRHD_DUAL_CMP(<=)
GRHD_DUAL_CMP(>)
GRHD_DUAL_CMP(>=)
GRHD_DUAL_CMP(==)
GRHD_DUAL_CMP(!=)
#undef GRHD_DUAL_CMP

__device__ __forceinline__ Dual sqrt(Dual x)
{
    const double r = ::sqrt(x.v);
    return Dual(r, r > 0.0 ? x.d / (2.0 * r) : 0.0);
}
__device__ __forceinline__ Dual fabs(Dual x) { return x.v < 0.0 ? -x : x; }
/* the double overloads, which the Dual ones would otherwise hide here */
__device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ double fabs(double x) { return ::fabs(x); }

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }

/* ------------------------------------------------------------------ */
/* con2prim differentiated at its root.                                */
/* ------------------------------------------------------------------ */

/* presid with every argument a Dual and the v^2 clamp frozen. */
__device__ __forceinline__ Dual presid_dual(Dual p, Dual D, Dual tau, Dual S2,
                                            Dual gamma, double max_v2)
{
    const Dual epp = tau + D + p;
    Dual v2 = S2 / (epp * epp);
    if (v2 > max_v2)
        v2 = Dual(max_v2);
    const Dual W = 1.0 / sqrt(1.0 - v2);
    return D * W + gamma * p * W * W / (gamma - 1.0) - epp;
}

/* Primitive recovery with tangents.  The forward solve (double con2prim,
 * the reference's algorithm) fixes the root p* and the branch.  There
 * f(p*; U, gamma^ij, Gamma) = 0, so dp = -(df with p held) / (df/dp); the
 * primitives are then formed from (U, p) exactly as con2prim forms them.
 * A root taken at the recovery pressure floor by the early exit is a
 * constant.  Returns false where the forward solve fails. */
/* The forward solve alone: the root pressure of con2prim (values). */
__device__ __forceinline__ bool con2prim_root(const Cons &cv,
                                              const Options &o, double gamma,
                                              const double guv[3][3],
                                              double &ps)
{
    Options ov = o;
    ov.gamma = gamma;
    Prim pv;
    double res;
    if (!con2prim(cv, ov, guv, pv, res))
        return false;
    ps = pv.pressure;
    return true;
}

/* The primitives with tangents at a given root ps of the forward solve,
 * on a metric of type G (double, or Dual for the spin's tangent). */
template <class G>
__device__ __forceinline__ void con2prim_at_root(const ConsT<Dual> &c,
                                                 const Options &o, Dual gamma,
                                                 const G gu[3][3], double ps,
                                                 PrimT<Dual> &out);

__device__ bool con2prim_ift(const ConsT<Dual> &c, const Options &o,
                             Dual gamma, const Dual gu[3][3],
                             PrimT<Dual> &out)
{
    Cons cv;
    cv.D = c.D.v;
    cv.tau = c.tau.v;
    double guv[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
    {
        cv.S[m] = c.S[m].v;
#pragma unroll
        for (int n = 0; n < 3; ++n)
            guv[m][n] = gu[m][n].v;
    }
    double ps;
    if (!con2prim_root(cv, o, gamma.v, guv, ps))
        return false;
    con2prim_at_root(c, o, gamma, gu, ps, out);
    return true;
}

template <class G>
__device__ __forceinline__ void con2prim_at_root(const ConsT<Dual> &c,
                                                 const Options &o, Dual gamma,
                                                 const G gu[3][3], double ps,
                                                 PrimT<Dual> &out)
{
    Dual S2 = 0.;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n)
            S2 += gu[m][n] * c.S[m] * c.S[n];
    Dual p(ps, 0.0);
    if (ps != o.rec_pressure_floor)
    {
        const Dual f_u =
            presid_dual(Dual(ps, 0.0), c.D, c.tau, S2, gamma, o.rec_max_v2);
        const Dual f_p = presid_dual(Dual(ps, 1.0), Dual(c.D.v), Dual(c.tau.v),
                                     Dual(S2.v), Dual(gamma.v), o.rec_max_v2);
        p = Dual(ps, -f_u.d / f_p.d);
    }
    const Dual epp = c.tau + c.D + p;
    const Dual v2 = S2 / (epp * epp);
    const Dual W = 1.0 / sqrt(1.0 - v2);
    out.rho = c.D / W;
    out.pressure = p;
    out.eps = p / ((gamma - 1.0) * out.rho);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        Dual vi = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vi += gu[i][j] * c.S[j] / epp;
        out.v[i] = vi;
    }
}

/* ------------------------------------------------------------------ */
/* Local functions, templated on double (values) / Dual (tangents).    */
/* ------------------------------------------------------------------ */

/* Parameters with their tangent slots: which local input is seeded. */
struct AdjParams
{
    double M, a, gamma;
};

enum
{
    SEED_GAMMA = -1,
    SEED_SPIN = -2
};

/* ------------------------------------------------------------------ */
/* Reverse mode of one face flux in its two face states.               */
/* side_forward evaluates one state's part of hlle_face_g (floors,      */
/* prim2con, flux_of, char_speeds) in double, keeping the values the    */
/* backward sweep needs and the branches it took; face_flux_reverse     */
/* applies the transposed Jacobian of the HLLE combination to lambda    */
/* and side_reverse that of each state's part.  Every branch (the       */
/* floors, the v^2 clamp, the radicand clamp, the wave-speed            */
/* selections) is taken as the forward evaluation takes it, with the    */
/* same tie rules (std_max keeps its first argument on a tie, std_min   */
/* too).  One sweep gives what one forward-mode pass per face-state     */
/* component and Gamma would.                                           */
/* ------------------------------------------------------------------ */
struct SideFwd
{
    Prim q;                  /* the state after the floors */
    bool rho_kept, p_kept;   /* floors left rho / P as they were */
    bool scaled;             /* the v^2 clamp rescaled v */
    double v_in[3], v2pre, scale;
    double vL[3], v2, W, h, rhW2;
    double c[5], f[5];       /* D, S1, S2, S3, tau */
    double tv;
    double cs2, cs, den, rad, sq, lo, hi, lower, upper;
};

__device__ __forceinline__ void side_forward(const Prim &p, const Options &o,
                                             double gamma, const Geo &g,
                                             int d, SideFwd &s)
{
    /* floors_g */
    Prim q = p;
    s.rho_kept = !(q.rho < o.rho_floor);
    s.p_kept = !(q.pressure < o.pressure_floor);
    if (!s.rho_kept)
        q.rho = o.rho_floor;
    if (!s.p_kept)
        q.pressure = o.pressure_floor;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        s.v_in[i] = q.v[i];
    s.v2pre = vsq(q.v, g.g);
    s.scaled = s.v2pre >= o.max_v2;
    s.scale = 1.0;
    if (s.scaled)
    {
        s.scale = sqrt(o.max_v2 / s.v2pre);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            q.v[i] *= s.scale;
    }
    q.eps = q.pressure / ((gamma - 1.0) * q.rho);
    s.q = q;
    /* prim2con */
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        s.vL[i] = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            s.vL[i] += g.g[i][j] * q.v[j];
    }
    s.v2 = vsq(q.v, g.g);
    s.W = 1.0 / sqrt(1.0 - s.v2);
    s.h = 1.0 + q.eps + q.pressure / q.rho;
    s.rhW2 = q.rho * s.h * s.W * s.W;
    s.c[0] = q.rho * s.W;
    s.c[4] = s.rhW2 - q.pressure - s.c[0];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        s.c[1 + i] = s.rhW2 * s.vL[i];
    /* flux_of */
    const double vel = q.v[d];
    s.tv = g.lapse * vel - g.shift[d];
    s.f[0] = s.c[0] * s.tv;
    s.f[4] = s.c[4] * s.tv + g.lapse * q.pressure * vel;
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        s.f[1 + j] = s.c[1 + j] * s.tv;
        if (j == d)
            s.f[1 + j] += g.lapse * q.pressure;
    }
    /* char_speeds */
    s.cs2 = gamma * q.pressure / (q.rho * s.h);
    s.cs = sqrt(s.cs2);
    s.den = 1.0 - s.v2 * s.cs2;
    s.rad = (1.0 - s.v2) *
            (g.gu[d][d] * (1.0 - s.v2 * s.cs2) - vel * vel * (1.0 - s.cs2));
    s.sq = sqrt(0.0 < s.rad ? s.rad : 0.0);
    const double root = s.cs * s.sq;
    s.lo = (vel * (1.0 - s.cs2) - root) / s.den;
    s.hi = (vel * (1.0 - s.cs2) + root) / s.den;
    s.lower = g.lapse * s.lo - g.shift[d];
    s.upper = g.lapse * s.hi - g.shift[d];
}

/* vb += w d(v^T g v)/dv */
__device__ __forceinline__ void vsq_bar(const double v[3], const double g[3][3],
                                        double w, double vb[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            t += (g[k][j] + g[j][k]) * v[j];
        vb[k] += w * t;
    }
}

/* The spin tangents of the face-geometry components a face flux of
 * direction d reads: gamma_ij, gamma^dd, alpha and beta^d. */
struct FaceTan
{
    double g[9], gudd, lapse, shiftd;
};

/* w d(v^T gamma v)/d gamma_ij contracted with the tangents of gamma_ij */
__device__ __forceinline__ double vsq_metric_dot(const double v[3], double w,
                                                 const FaceTan &t)
{
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            s += v[i] * v[j] * t.g[3 * i + j];
    return w * s;
}

/* The adjoints of one side's conserved state, flux and wave speeds
 * (cb, fb, lob, hib) taken back to the face state before the floors
 * (pb: rho, eps, P, v1..v3), to Gamma (added into gb), and through the
 * face geometry to the spin (added into ab), given the geometry's spin
 * tangents ft. */
__device__ __forceinline__ void side_reverse(const SideFwd &s, double gamma,
                                             const Geo &g, int d,
                                             const double cb[5],
                                             const double fb[5], double lob,
                                             double hib, const FaceTan &ft,
                                             double pb[6], double &gb,
                                             double &ab)
{
    const Prim &q = s.q;
    const double vel = q.v[d];
    double rho_b = 0.0, P_b = 0.0, eps_b = 0.0, gam_b = 0.0, v2_b = 0.0;
    double v_b[3] = {0.0, 0.0, 0.0};
    /* char_speeds */
    const double lo_b = g.lapse * lob, hi_b = g.lapse * hib;
    const double A_b = (lo_b + hi_b) / s.den;
    const double root_b = (hi_b - lo_b) / s.den;
    const double den_b = -(lo_b * s.lo + hi_b * s.hi) / s.den;
    ab += lob * (s.lo * ft.lapse - ft.shiftd) +
          hib * (s.hi * ft.lapse - ft.shiftd);
    double vel_b = A_b * (1.0 - s.cs2);
    double cs2_b = -A_b * vel;
    const double cs_b = root_b * s.sq;
    const double sq_b = root_b * s.cs;
    const double rad_b =
        (0.0 < s.rad && s.sq > 0.0) ? sq_b / (2.0 * s.sq) : 0.0;
    const double B = g.gu[d][d] * (1.0 - s.v2 * s.cs2) - vel * vel * (1.0 - s.cs2);
    const double B_b = rad_b * (1.0 - s.v2);
    ab += B_b * (1.0 - s.v2 * s.cs2) * ft.gudd;
    v2_b -= rad_b * B;
    v2_b -= B_b * g.gu[d][d] * s.cs2;
    cs2_b -= B_b * g.gu[d][d] * s.v2;
    vel_b -= B_b * 2.0 * vel * (1.0 - s.cs2);
    cs2_b += B_b * vel * vel;
    v2_b -= den_b * s.cs2;
    cs2_b -= den_b * s.v2;
    if (s.cs > 0.0)
        cs2_b += cs_b / (2.0 * s.cs);
    gam_b += cs2_b * q.pressure / (q.rho * s.h);
    P_b += cs2_b * gamma / (q.rho * s.h);
    rho_b -= cs2_b * s.cs2 / q.rho;
    double h_b = -cs2_b * s.cs2 / s.h;
    v_b[d] += vel_b;
    /* flux_of */
    double tv_b = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        tv_b += fb[k] * s.c[k];
    double c_b[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        c_b[k] = cb[k] + fb[k] * s.tv;
    P_b += fb[4] * g.lapse * vel + fb[1 + d] * g.lapse;
    v_b[d] += fb[4] * g.lapse * q.pressure + tv_b * g.lapse;
    ab += tv_b * (vel * ft.lapse - ft.shiftd) +
          (fb[4] * vel + fb[1 + d]) * q.pressure * ft.lapse;
    /* prim2con: tau = rhW2 - P - D, S_i = rhW2 vL_i, D = rho W */
    double rhW2_b = c_b[4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        rhW2_b += c_b[1 + i] * s.vL[i];
    P_b -= c_b[4];
    const double D_b = c_b[0] - c_b[4];
    rho_b += D_b * s.W;
    double W_b = D_b * q.rho;
    rho_b += rhW2_b * s.h * s.W * s.W;
    h_b += rhW2_b * q.rho * s.W * s.W;
    W_b += rhW2_b * 2.0 * q.rho * s.h * s.W;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            v_b[j] += c_b[1 + i] * s.rhW2 * g.g[i][j];
            ab += c_b[1 + i] * s.rhW2 * q.v[j] * ft.g[3 * i + j];
        }
    /* h = 1 + eps + P / rho (prim2con's and char_speeds' alike) */
    eps_b += h_b;
    P_b += h_b / q.rho;
    rho_b -= h_b * q.pressure / (q.rho * q.rho);
    /* W = (1 - v^2)^(-1/2), v^2 shared with char_speeds */
    v2_b += W_b * 0.5 * s.W * s.W * s.W;
    vsq_bar(q.v, g.g, v2_b, v_b);
    ab += vsq_metric_dot(q.v, v2_b, ft);
    /* floors: eps = P / ((Gamma - 1) rho), then the v^2 clamp */
    P_b += eps_b / ((gamma - 1.0) * q.rho);
    rho_b -= eps_b * q.eps / q.rho;
    gam_b -= eps_b * q.eps / (gamma - 1.0);
    double vin_b[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        vin_b[i] = v_b[i];
    if (s.scaled)
    {
        double scale_b = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            scale_b += v_b[i] * s.v_in[i];
            vin_b[i] = v_b[i] * s.scale;
        }
        const double v2pre_b = -0.5 * scale_b * s.scale / s.v2pre;
        vsq_bar(s.v_in, g.g, v2pre_b, vin_b);
        ab += vsq_metric_dot(s.v_in, v2pre_b, ft);
    }
    pb[0] = s.rho_kept ? rho_b : 0.0;
    pb[1] = 0.0; /* eps is recomputed by the floors */
    pb[2] = s.p_kept ? P_b : 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        pb[3 + i] = vin_b[i];
    gb += gam_b;
}

/* lambda . F(lp, rp) differentiated in the two face states (lpb, rpb, in
 * the order rho, eps, P, v1..v3), in Gamma (added into gb) and, through
 * the face geometry with spin tangents ft, in the spin (added into ab). */
__device__ __forceinline__ void face_flux_reverse(const Prim &lp,
                                                  const Prim &rp,
                                                  const Options &o,
                                                  double gamma, const Geo &g,
                                                  const FaceTan &ft, int d,
                                                  const double lam[5],
                                                  double lpb[6], double rpb[6],
                                                  double &gb, double &ab)
{
    SideFwd l, r;
    side_forward(lp, o, gamma, g, d, l);
    side_forward(rp, o, gamma, g, d, r);
    double lcb[5] = {0, 0, 0, 0, 0}, lfb[5] = {0, 0, 0, 0, 0};
    double rcb[5] = {0, 0, 0, 0, 0}, rfb[5] = {0, 0, 0, 0, 0};
    double llob = 0.0, lhib = 0.0, rlob = 0.0, rhib = 0.0;
    /* the HLLE combination (hlle_face_g) */
    const double in_lo = (r.lower < l.lower) ? r.lower : l.lower;
    const double sl = (in_lo < 0.0) ? in_lo : 0.0;
    const double in_hi = (l.upper < r.upper) ? r.upper : l.upper;
    const double su = (0.0 < in_hi) ? in_hi : 0.0;
    if (sl >= 0.0)
    {
#pragma unroll
        for (int k = 0; k < 5; ++k)
            lfb[k] = lam[k];
    }
    else if (su <= 0.0)
    {
#pragma unroll
        for (int k = 0; k < 5; ++k)
            rfb[k] = lam[k];
    }
    else
    {
        /* f = (su F_L - sl F_R + sl su (U_R - U_L)) / (su - sl), with sl
         * from the smaller lower speed and su from the larger upper one */
        const double inv = 1.0 / (su - sl);
        double sl_b = 0.0, su_b = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k)
        {
            const double dc = r.c[k] - l.c[k];
            const double fk = (su * l.f[k] - sl * r.f[k] + sl * su * dc) * inv;
            lfb[k] = lam[k] * su * inv;
            rfb[k] = -lam[k] * sl * inv;
            rcb[k] = lam[k] * sl * su * inv;
            lcb[k] = -rcb[k];
            su_b += lam[k] * ((l.f[k] + sl * dc) * inv - fk * inv);
            sl_b += lam[k] * ((-r.f[k] + su * dc) * inv + fk * inv);
        }
        if (r.lower < l.lower)
            rlob = sl_b;
        else
            llob = sl_b;
        if (l.upper < r.upper)
            rhib = su_b;
        else
            lhib = su_b;
    }
    side_reverse(l, gamma, g, d, lcb, lfb, llob, lhib, ft, lpb, gb, ab);
    side_reverse(r, gamma, g, d, rcb, rfb, rlob, rhib, ft, rpb, gb, ab);
}

/* d(c + h * limited_slope(l, c, r, theta)) / d(l, c, r): the reconstruction
 * of one component (slope_step, h = +1/2 for the plus half, -1/2 for the
 * minus half), with the limiter's branch frozen. */
__device__ __forceinline__ void half_step_jacobian(double l, double c,
                                                   double r, double theta,
                                                   double h, double j[3])
{
#pragma unroll
    for (int q = 0; q < 3; ++q)
    {
        const Dual L(l, q == 0 ? 1.0 : 0.0), C(c, q == 1 ? 1.0 : 0.0),
            R(r, q == 2 ? 1.0 : 0.0);
        const Dual out = C + Dual(h) * limited_slope(L, C, R, Dual(theta));
        j[q] = out.d;
    }
}

/* The face geometry: mean of the geometries of cells L and R. */
template <class T>
__device__ __forceinline__ void face_geo(const T xl[3], const T xr[3], T M,
                                         T a, GeoT<T> &face)
{
    GeoT<T> gl, gr;
    ks_geometry(M, a, xl, gl);
    ks_geometry(M, a, xr, gr);
    average_geo(gl, gr, face);
}

/* The full geometry the static-metric sources read. */
template <class T>
__device__ __forceinline__ void source_geo(const T x[3], T M, T a,
                                           FullGeoT<T> &fg)
{
    KSCoreT<T> core;
    GeoT<T> g;
    ks_geometry_core(M, a, x, g, core);
    ks_derivs(M, a, x, core, g, fg);
}

/* Static-metric sources of an interior cell (k_cells) on a given full
 * geometry: floors on W, the atmosphere skip, then the sources.  Returns
 * false where skipped. */
template <class T, class G>
__device__ __forceinline__ bool sources_geo(const PrimT<T> &w,
                                            const ConsT<T> &u,
                                            const FullGeoT<G> &fg, T gamma,
                                            const Options &o, ConsT<T> &src)
{
    PrimT<T> pw = w;
    floors_g(pw, o, gamma, fg.geo.g);
    Prim pv;
    pv.rho = val(pw.rho);
    pv.pressure = val(pw.pressure);
    if (is_atmosphere(pv, o, o.source_atmosphere_factor))
        return false;
    static_sources(pw, u, gamma, fg, src);
    return true;
}

/* lambda . S(w, u), the static sources after the floors and the
 * atmosphere skip (sources_geo on a double geometry), differentiated in
 * one reverse sweep: in w (wb: rho, eps, P, v1..v3), u (ub: D, S1..S3,
 * tau) and Gamma (added into gb).  Branches as the forward evaluation
 * takes them (floors_g's ties keep the state).  Returns false where the
 * cell is skipped: the sources do not depend on it. */
__device__ __forceinline__ bool sources_reverse(const Prim &w, const Cons &u,
                                                const FullGeo &fg,
                                                double gamma, const Options &o,
                                                const double lam[5],
                                                double wb[6], double ub[5],
                                                double &gb)
{
    const Geo &g = fg.geo;
    /* floors_g */
    Prim q = w;
    const bool rho_kept = !(q.rho < o.rho_floor);
    const bool p_kept = !(q.pressure < o.pressure_floor);
    if (!rho_kept)
        q.rho = o.rho_floor;
    if (!p_kept)
        q.pressure = o.pressure_floor;
    double v_in[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        v_in[i] = q.v[i];
    const double v2pre = vsq(q.v, g.g);
    const bool scaled = v2pre >= o.max_v2;
    double scale = 1.0;
    if (scaled)
    {
        scale = sqrt(o.max_v2 / v2pre);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            q.v[i] *= scale;
    }
    q.eps = q.pressure / ((gamma - 1.0) * q.rho);
    if (is_atmosphere(q, o, o.source_atmosphere_factor))
        return false;
    /* static_sources' values */
    const double W = 1.0 / sqrt(1.0 - vsq(q.v, g.g));
    const double h = 1.0 + q.eps + q.pressure / q.rho;
    const double rhW2 = q.rho * h * W * W;
    const double lD = lam[0], lT = lam[4];
    const double lS[3] = {lam[1], lam[2], lam[3]};
    double rho_b = 0.0, P_b = 0.0, eps_b = 0.0, gam_b = 0.0;
    double v_b[3] = {0.0, 0.0, 0.0};
    double D_b = 0.0, tau_b = 0.0, S_b[3] = {0.0, 0.0, 0.0};
    /* E = D + tau enters S_m as -E dalpha_m */
    double E_b = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
        E_b -= lS[m] * fg.dlapse[m];
    D_b += E_b;
    tau_b += E_b;
    /* gamma^ij S_j enters tau as -mU_d dalpha_d; S_s dbeta^s_m enters S_m */
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const double mU_b = -lT * fg.dlapse[d];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            S_b[j] += mU_b * g.gu[d][j];
    }
#pragma unroll
    for (int sdir = 0; sdir < 3; ++sdir)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            S_b[sdir] += lS[m] * fg.dshift[sdir][m];
    /* the fluxes F^d enter as -F^d d_d ln sqrt(gamma) */
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const double tv = g.lapse * q.v[d] - g.shift[d];
        const double cD_b = -lD * fg.dlogsqrtdet[d];
        const double cT_b = -lT * fg.dlogsqrtdet[d];
        double cS_b[3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
            cS_b[m] = -lS[m] * fg.dlogsqrtdet[d];
        double tv_b = cD_b * u.D + cT_b * u.tau;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tv_b += cS_b[j] * u.S[j];
        D_b += cD_b * tv;
        tau_b += cT_b * tv;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            S_b[j] += cS_b[j] * tv;
        P_b += cT_b * g.lapse * q.v[d] + cS_b[d] * g.lapse;
        v_b[d] += cT_b * g.lapse * q.pressure + tv_b * g.lapse;
    }
    /* T^ij = rho h W^2 v^i v^j + P gamma^ij, contracted with alpha K_ij
     * (tau) and alpha dgamma_ij / 2 (S_m) */
    double rhW2_b = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            double t_b = lT * g.lapse * fg.K[i][j];
#pragma unroll
            for (int m = 0; m < 3; ++m)
                t_b += lS[m] * 0.5 * g.lapse * fg.dg[m][i][j];
            rhW2_b += t_b * q.v[i] * q.v[j];
            v_b[i] += t_b * rhW2 * q.v[j];
            v_b[j] += t_b * rhW2 * q.v[i];
            P_b += t_b * g.gu[i][j];
        }
    rho_b += rhW2_b * h * W * W;
    const double h_b = rhW2_b * q.rho * W * W;
    const double W_b = rhW2_b * 2.0 * q.rho * h * W;
    eps_b += h_b;
    P_b += h_b / q.rho;
    rho_b -= h_b * q.pressure / (q.rho * q.rho);
    vsq_bar(q.v, g.g, W_b * 0.5 * W * W * W, v_b);
    /* floors: eps = P / ((Gamma - 1) rho), then the v^2 clamp */
    P_b += eps_b / ((gamma - 1.0) * q.rho);
    rho_b -= eps_b * q.eps / q.rho;
    gam_b -= eps_b * q.eps / (gamma - 1.0);
    double vin_b[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        vin_b[i] = v_b[i];
    if (scaled)
    {
        double scale_b = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            scale_b += v_b[i] * v_in[i];
            vin_b[i] = v_b[i] * scale;
        }
        vsq_bar(v_in, g.g, -0.5 * scale_b * scale / v2pre, vin_b);
    }
    wb[0] = rho_kept ? rho_b : 0.0;
    wb[1] = 0.0; /* eps is recomputed by the floors */
    wb[2] = p_kept ? P_b : 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        wb[3 + i] = vin_b[i];
    ub[0] = D_b;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        ub[1 + i] = S_b[i];
    ub[4] = tau_b;
    gb += gam_b;
    return true;
}

/* ------------------------------------------------------------------ */
/* Seeding helpers.                                                    */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ void seed_params(const AdjParams &pa, int seed,
                                            Dual &M, Dual &a, Dual &gamma)
{
    M = Dual(pa.M);
    a = Dual(pa.a, seed == SEED_SPIN ? 1.0 : 0.0);
    gamma = Dual(pa.gamma, seed == SEED_GAMMA ? 1.0 : 0.0);
}

__device__ __forceinline__ PrimT<Dual> dual_prim(const Prim &p, int slot)
{
    PrimT<Dual> q;
    q.rho = Dual(p.rho, slot == 0 ? 1.0 : 0.0);
    q.eps = Dual(p.eps, slot == 1 ? 1.0 : 0.0);
    q.pressure = Dual(p.pressure, slot == 2 ? 1.0 : 0.0);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        q.v[m] = Dual(p.v[m], slot == 3 + m ? 1.0 : 0.0);
    return q;
}

__device__ __forceinline__ ConsT<Dual> dual_cons(const Cons &c, int slot)
{
    ConsT<Dual> q;
    q.D = Dual(c.D, slot == 0 ? 1.0 : 0.0);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        q.S[m] = Dual(c.S[m], slot == 1 + m ? 1.0 : 0.0);
    q.tau = Dual(c.tau, slot == 4 ? 1.0 : 0.0);
    return q;
}

__device__ __forceinline__ double dot_tangent(const ConsT<Dual> &f,
                                              const double lam[5])
{
    return lam[0] * f.D.d + lam[1] * f.S[0].d + lam[2] * f.S[1].d +
           lam[3] * f.S[2].d + lam[4] * f.tau.d;
}

/* block sum of two doubles, one atomicAdd each by thread 0 */
__device__ inline void block_add_params(double *out, double g0, double g1)
{
    __shared__ double part[2][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        g0 += __shfl_xor(g0, off, 64);
        g1 += __shfl_xor(g1, off, 64);
    }
    if (lane == 0)
    {
        part[0][wave] = g0;
        part[1][wave] = g1;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < kBlock / 64; ++w)
        {
            s0 += part[0][w];
            s1 += part[1][w];
        }
        if (s0 != 0.0)
            atomicAdd(&out[0], s0);
        if (s1 != 0.0)
            atomicAdd(&out[1], s1);
    }
}

/* Interior arrays of the adjoint: 5 components over the interior box,
 * component-major (the rhs FAB layout). */
__device__ __forceinline__ long long iidx(const BoxDev &b, int i, int j, int k)
{
    return ((long long)(k - b.lo[2]) * b.in[1] + (j - b.lo[1])) * b.in[0] +
           (i - b.lo[0]);
}

__device__ __forceinline__ bool interior(const BoxDev &b, int i, int j, int k)
{
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] &&
           k >= b.lo[2] && k <= b.hi[2];
}

/* ------------------------------------------------------------------ */
/* The spin tangent of the Kerr-Schild geometry of every storage cell,  */
/* evaluated once per adjoint run (the background is static) and read   */
/* by the reverse sweeps of every stage, which contract the adjoint of  */
/* the metric they read with it to give the spin's adjoint.  Component- */
/* major: the tangent of component m (gamma_ij at 3i + j, gamma^ij at   */
/* 9 + 3i + j, alpha at 18, beta^i at 19 + i) at [m][cell].             */
/* ------------------------------------------------------------------ */
constexpr int kGeoN = 22; /* g (9), gu (9), lapse, shift (3) */

struct MetricCache
{
    double *geo; /* kGeoN x storage cells */
};

template <class T> __device__ __forceinline__ T &geo_at(GeoT<T> &g, int m)
{
    return m < 9    ? g.g[m / 3][m % 3]
           : m < 18 ? g.gu[(m - 9) / 3][(m - 9) % 3]
           : m < 19 ? g.lapse
                    : g.shift[m - 19];
}

__global__ void __launch_bounds__(kBlock)
    k_adj_metric_cache(BoxDev b, AdjParams pa, MetricCache mc)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= b.gnpts)
        return;
    const int i = b.glo[0] + (int)(it % b.gn[0]);
    const int j = b.glo[1] + (int)((it / b.gn[0]) % b.gn[1]);
    const int k = b.glo[2] + (int)(it / ((long long)b.gn[0] * b.gn[1]));
    const long long n = b.gnpts;
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    const Dual xd[3] = {Dual(x[0]), Dual(x[1]), Dual(x[2])};
    Dual M, a, gamma;
    seed_params(pa, SEED_SPIN, M, a, gamma);
    GeoT<Dual> gd;
    ks_geometry(M, a, xd, gd);
#pragma unroll
    for (int m = 0; m < kGeoN; ++m)
        mc.geo[(long long)m * n + it] = geo_at(gd, m).d;
}

/* ------------------------------------------------------------------ */
/* k_adj_faces<D>: lambda_F of each face of direction D from the rhs    */
/* adjoint of its two cells, then J_F^T lambda_F into lambda_W of the   */
/* four stencil cells (atomics) and into the parameter adjoints, the    */
/* spin's through the face geometry and its cached tangents.            */
/* One thread per face of the interior face box.                       */
/* ------------------------------------------------------------------ */
template <int D>
__global__ void __launch_bounds__(kBlock)
    k_adj_faces(BoxDev b, Options o, AdjParams pa, MetricCache mc,
                const double *W, const double *lam_k, double *lam_W,
                double *lam_par)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = b.in[0] + (D == 0), n1 = b.in[1] + (D == 1);
    const int n2 = b.in[2] + (D == 2);
    const long long nfaces = (long long)n0 * n1 * n2;
    double g0 = 0.0, g1 = 0.0;
    if (t < nfaces)
    {
        int c[3];
        c[0] = b.lo[0] + (int)(t % n0);
        c[1] = b.lo[1] + (int)((t / n0) % n1);
        c[2] = b.lo[2] + (int)(t / ((long long)n0 * n1));
        /* face between L = c - e_D and R = c */
        int cl[3] = {c[0], c[1], c[2]};
        cl[D] -= 1;
        const long long ni = b.inpts;
        double lam[5];
        const bool Lin = interior(b, cl[0], cl[1], cl[2]);
        const bool Rin = interior(b, c[0], c[1], c[2]);
        const long long il = Lin ? iidx(b, cl[0], cl[1], cl[2]) : 0;
        const long long ir = Rin ? iidx(b, c[0], c[1], c[2]) : 0;
        bool any = false;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            const double lr = Rin ? lam_k[m * ni + ir] : 0.0;
            const double ll = Lin ? lam_k[m * ni + il] : 0.0;
            lam[m] = (lr - ll) * b.inv_dx[D];
            any = any || lam[m] != 0.0;
        }
        if (any)
        {
            const long long n = b.gnpts;
            long long so[4];
            Prim wv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                int cc[3] = {c[0], c[1], c[2]};
                cc[D] += q - 2;
                so[q] = goff(b, cc[0], cc[1], cc[2]);
                wv[q] = load_prim(W, n, so[q]);
            }
            /* the face metric in double, and its spin tangent: the mean of
             * the two cells' cached tangents (average_geo) */
            double xl[3], xr[3];
            coords(cl[0], cl[1], cl[2], b.dx, b.center, xl);
            coords(c[0], c[1], c[2], b.dx, b.center, xr);
            Geo facev;
            face_geo(xl, xr, pa.M, pa.a, facev);
            FaceTan ft;
            {
                /* geo component m's tangent at [m] (gamma_ij at
                 * 3i + j, gamma^ij at 9 + 3i + j, alpha 18, beta^i 19 + i) */
                auto tan = [&](int m) {
                    return 0.5 * (mc.geo[(long long)m * n + so[1]] +
                                  mc.geo[(long long)m * n + so[2]]);
                };
#pragma unroll
                for (int m = 0; m < 9; ++m)
                    ft.g[m] = tan(m);
                ft.gudd = tan(9 + 4 * D);
                ft.lapse = tan(18);
                ft.shiftd = tan(19 + D);
            }
            /* the face states (values), as k_faces forms them */
            Prim lpv = wv[1], rpv = wv[2];
            if (o.use_reconstruction)
            {
                Prim unused;
                slope_step(wv[0], wv[1], wv[2], o.theta, unused, lpv);
                slope_step(wv[1], wv[2], wv[3], o.theta, rpv, unused);
            }
            /* lambda . dF/d(face states), d/dGamma and d/da in one reverse
             * sweep */
            double lam_lp[6], lam_rp[6];
            face_flux_reverse(lpv, rpv, o, pa.gamma, facev, ft, D, lam, lam_lp,
                              lam_rp, g0, g1);
            /* back through the reconstruction, component by component:
             * lp = W_L + s_L/2 from (W_{L-1}, W_L, W_R), rp = W_R - s_R/2
             * from (W_L, W_R, W_{R+1}) */
#pragma unroll
            for (int m = 0; m < 6; ++m)
            {
                double gq[4] = {0.0, lam_lp[m], lam_rp[m], 0.0};
                if (o.use_reconstruction)
                {
                    /* the stencil's W read again rather than kept in
                     * registers through the reverse sweep */
                    double wm[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        wm[q] = W[(long long)m * n + so[q]];
                    double jl[3], jr[3];
                    half_step_jacobian(wm[0], wm[1], wm[2], o.theta, 0.5, jl);
                    half_step_jacobian(wm[1], wm[2], wm[3], o.theta, -0.5, jr);
                    gq[0] = lam_lp[m] * jl[0];
                    gq[1] = lam_lp[m] * jl[1] + lam_rp[m] * jr[0];
                    gq[2] = lam_lp[m] * jl[2] + lam_rp[m] * jr[1];
                    gq[3] = lam_rp[m] * jr[2];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (gq[q] != 0.0)
                        atomicAdd(&lam_W[(long long)m * n + so[q]], gq[q]);
            }
        }
    }
    block_add_params(lam_par, g0, g1);
}

/* ------------------------------------------------------------------ */
/* k_adj_sources: sources of each interior cell: lambda_k -> lambda_W   */
/* (6), lambda_U (5, added), parameter adjoints.                        */
/* ------------------------------------------------------------------ */
template <bool SPIN>
__global__ void __launch_bounds__(kBlock)
    k_adj_sources(BoxDev b, Options o, AdjParams pa, const double *U, const double *W, const double *lam_k,
                  double *lam_W, double *lam_U, double *lam_par)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double g0 = 0.0, g1 = 0.0;
    if (it < b.inpts && o.use_static_sources)
    {
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        const long long n = b.gnpts, ni = b.inpts;
        const long long so = goff(b, i, j, k);
        double lam[5];
        bool any = false;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            lam[m] = lam_k[m * ni + it];
            any = any || lam[m] != 0.0;
        }
        if (any)
        {
            const Prim wv = load_prim(W, n, so);
            const Cons uv = load_cons(U, n, so);
            double x[3];
            coords(i, j, k, b.dx, b.center, x);
            if constexpr (SPIN)
            {
                /* the spin's tangent: the full geometry with duals */
                Dual M, a, gamma;
                seed_params(pa, SEED_SPIN, M, a, gamma);
                const Dual xd[3] = {Dual(x[0]), Dual(x[1]), Dual(x[2])};
                FullGeoT<Dual> fg;
                source_geo(xd, M, a, fg);
                ConsT<Dual> src;
                if (sources_geo(dual_prim(wv, -1), dual_cons(uv, -1), fg,
                                gamma, o, src))
                    g1 = dot_tangent(src, lam);
            }
            else
            {
                /* the full geometry in double: no pass here seeds it */
                FullGeo fgv;
                source_geo(x, pa.M, pa.a, fgv);
                /* 6 W inputs, 5 U inputs and Gamma in one reverse sweep */
                double gW[6], gU[5];
                if (sources_reverse(wv, uv, fgv, pa.gamma, o, lam, gW, gU, g0))
                {
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        lam_U[(long long)m * ni + it] += gU[m];
#pragma unroll
                    for (int m = 0; m < 6; ++m)
                        if (gW[m] != 0.0)
                            atomicAdd(&lam_W[m * n + so], gW[m]);
                }
            }
        }
    }
    block_add_params(lam_par, g0, g1);
}

/* The W of a storage cell as a function of the state (k_prims):
 * interior, con2prim on the cell's metric; outflow ghost, the chain
 * con2prim(source) -> floors(ghost metric) -> prim2con -> con2prim. */
/* The values of that chain for one cell: the geometries of the cell and
 * (ghost) of its source, and the roots of the one or two recoveries.  The
 * roots are the pressures k_prims just wrote into W for this state: the
 * cell's own (interior: con2prim of U; ghost: the chain's last recovery)
 * and, for a ghost, its source cell's (the chain's first recovery, the
 * same solve on the same inputs).  The adjoint runs only after a forward
 * sweep in which every recovery succeeded. */
struct WChain
{
    Geo gt, gs;
    double p1, p2;
};

__device__ __forceinline__ void w_chain_values(const BoxDev &b,
                                               const AdjParams &pa,
                                               const double *W, int i, int j,
                                               int k, const int src[3],
                                               bool physical, WChain &v)
{
    const long long n = b.gnpts;
    const long long st = goff(b, i, j, k);
    double xt[3];
    coords(i, j, k, b.dx, b.center, xt);
    ks_geometry(pa.M, pa.a, xt, v.gt);
    const double p_own = W[2 * n + st];
    if (!physical)
    {
        v.p1 = p_own;
        return;
    }
    const long long ss = goff(b, src[0], src[1], src[2]);
    double xs[3];
    coords(src[0], src[1], src[2], b.dx, b.center, xs);
    ks_geometry(pa.M, pa.a, xs, v.gs);
    v.p1 = W[2 * n + ss];
    v.p2 = p_own;
}

/* ------------------------------------------------------------------ */
/* Reverse mode of the recovery chain (k_prims' W as a function of U).  */
/* ------------------------------------------------------------------ */

/* con2prim_at_root in double: the values its reverse sweep needs. */
struct RootFwd
{
    double S2, p, epp, v2, W;
    Prim w;
    bool free_p; /* the root moves with U (not the recovery floor) */
};

__device__ __forceinline__ void root_forward(const Cons &c, const Options &o,
                                             double gamma,
                                             const double gu[3][3], double ps,
                                             RootFwd &r)
{
    r.S2 = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n)
            r.S2 += gu[m][n] * c.S[m] * c.S[n];
    r.p = ps;
    r.free_p = ps != o.rec_pressure_floor;
    r.epp = c.tau + c.D + ps;
    r.v2 = r.S2 / (r.epp * r.epp);
    r.W = 1.0 / sqrt(1.0 - r.v2);
    r.w.rho = c.D / r.W;
    r.w.pressure = ps;
    r.w.eps = ps / ((gamma - 1.0) * r.w.rho);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        double vi = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vi += gu[i][j] * c.S[j] / r.epp;
        r.w.v[i] = vi;
    }
}

/* The adjoint of the recovered primitives (wb: rho, eps, P, v1..v3) back
 * to the conserved state (cb: D, S1..S3, tau), Gamma (added into gb) and,
 * through gamma^ij with spin tangents tgu, the spin (added into ab).
 * The root pressure moves by the implicit-function theorem,
 * dp = -(df with p held) / (df/dp), on presid with the v^2 clamp as the
 * solve left it; a root at the recovery floor is a constant. */
__device__ __forceinline__ void root_reverse(const Cons &c, const Options &o,
                                             double gamma,
                                             const double gu[3][3],
                                             const double tgu[9],
                                             const RootFwd &r,
                                             const double wb[6], double cb[5],
                                             double &gb, double &ab)
{
    const Prim &w = r.w;
    double p_b = wb[2], rho_b = wb[0], D_b = 0.0, tau_b = 0.0, epp_b = 0.0;
    double S_b[3] = {0.0, 0.0, 0.0};
    /* eps = p / ((Gamma - 1) rho) */
    p_b += wb[1] / ((gamma - 1.0) * w.rho);
    rho_b -= wb[1] * w.eps / w.rho;
    gb -= wb[1] * w.eps / (gamma - 1.0);
    /* v^i = gamma^ij S_j / (E + p) */
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            S_b[j] += wb[3 + i] * gu[i][j] / r.epp;
            ab += wb[3 + i] * (c.S[j] / r.epp) * tgu[3 * i + j];
        }
        epp_b -= wb[3 + i] * w.v[i] / r.epp;
    }
    /* rho = D / W, W = (1 - v^2)^(-1/2), v^2 = S^2 / (E + p)^2 */
    D_b += rho_b / r.W;
    const double W_b = -rho_b * w.rho / r.W;
    const double v2_b = W_b * 0.5 * r.W * r.W * r.W;
    double S2_b = v2_b / (r.epp * r.epp);
    epp_b -= 2.0 * v2_b * r.v2 / r.epp;
    D_b += epp_b;
    tau_b += epp_b;
    p_b += epp_b;
    if (r.free_p)
    {
        /* presid(p) = D W' + Gamma p W'^2 / (Gamma - 1) - (E + p), with
         * v'^2 = min-clamped S^2 / (E + p)^2 */
        const double e = r.epp;
        const double v2r = r.S2 / (e * e);
        const bool clamped = v2r > o.rec_max_v2;
        const double v2c = clamped ? o.rec_max_v2 : v2r;
        const double Wc = 1.0 / sqrt(1.0 - v2c);
        const double f_W = c.D + 2.0 * gamma * r.p * Wc / (gamma - 1.0);
        const double dW_dv2 = clamped ? 0.0 : 0.5 * Wc * Wc * Wc;
        const double via_e = f_W * dW_dv2 * (-2.0 * v2r / e);
        const double f_D = Wc + via_e - 1.0;
        const double f_tau = via_e - 1.0;
        const double f_p = gamma * Wc * Wc / (gamma - 1.0) + via_e - 1.0;
        const double f_S2 = f_W * dW_dv2 / (e * e);
        const double f_g = -r.p * Wc * Wc / ((gamma - 1.0) * (gamma - 1.0));
        const double k = -p_b / f_p;
        D_b += k * f_D;
        tau_b += k * f_tau;
        S2_b += k * f_S2;
        gb += k * f_g;
    }
    /* S^2 = gamma^mn S_m S_n */
#pragma unroll
    for (int m = 0; m < 3; ++m)
    {
        double t = 0.0;
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2)
        {
            t += (gu[m][n2] + gu[n2][m]) * c.S[n2];
            ab += S2_b * c.S[m] * c.S[n2] * tgu[3 * m + n2];
        }
        S_b[m] += S2_b * t;
    }
    cb[0] = D_b;
#pragma unroll
    for (int m = 0; m < 3; ++m)
        cb[1 + m] = S_b[m];
    cb[4] = tau_b;
}

/* The spin tangents of the metric components the recovery chain reads:
 * gamma_ij and gamma^ij of the cell, gamma^ij of its source cell. */
struct ChainTan
{
    double gt_g[9], gt_gu[9], gs_gu[9];
};

/* The chain's adjoint: lambda_W of a storage cell (lw) back to the state
 * of the interior cell it was recovered from (ub, D, S1..S3, tau), to
 * Gamma (added into gb) and, through the metric with tangents t, to the
 * spin (added into ab), in one reverse sweep at the roots of `v`. */
__device__ __forceinline__ void w_chain_reverse(const WChain &v, double gamma,
                                                bool physical,
                                                const Options &o,
                                                const Cons &u,
                                                const ChainTan &t,
                                                const double lw[6],
                                                double ub[5], double &gb,
                                                double &ab)
{
    if (!physical)
    {
        RootFwd r;
        root_forward(u, o, gamma, v.gt.gu, v.p1, r);
        root_reverse(u, o, gamma, v.gt.gu, t.gt_gu, r, lw, ub, gb, ab);
        return;
    }
    /* forward: pr = C(u; gs), floors on gt, cg = prim2con(gt), w = C(cg; gt) */
    RootFwd r1;
    root_forward(u, o, gamma, v.gs.gu, v.p1, r1);
    Prim q = r1.w;
    const bool rho_kept = !(q.rho < o.rho_floor);
    const bool p_kept = !(q.pressure < o.pressure_floor);
    if (!rho_kept)
        q.rho = o.rho_floor;
    if (!p_kept)
        q.pressure = o.pressure_floor;
    double v_in[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        v_in[i] = q.v[i];
    const double v2pre = vsq(q.v, v.gt.g);
    const bool scaled = v2pre >= o.max_v2;
    double scale = 1.0;
    if (scaled)
    {
        scale = sqrt(o.max_v2 / v2pre);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            q.v[i] *= scale;
    }
    q.eps = q.pressure / ((gamma - 1.0) * q.rho);
    Cons cg;
    prim2con(q, gamma, v.gt.g, cg);
    RootFwd r2;
    root_forward(cg, o, gamma, v.gt.gu, v.p2, r2);
    /* reverse */
    double cgb[5];
    root_reverse(cg, o, gamma, v.gt.gu, t.gt_gu, r2, lw, cgb, gb, ab);
    /* prim2con on gt */
    double vL[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        vL[i] = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vL[i] += v.gt.g[i][j] * q.v[j];
    }
    const double W = 1.0 / sqrt(1.0 - vsq(q.v, v.gt.g));
    const double h = 1.0 + q.eps + q.pressure / q.rho;
    const double rhW2 = q.rho * h * W * W;
    double rho_b = 0.0, P_b = 0.0, eps_b = 0.0, v_b[3] = {0.0, 0.0, 0.0};
    double rhW2_b = cgb[4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        rhW2_b += cgb[1 + i] * vL[i];
    P_b -= cgb[4];
    const double D_b = cgb[0] - cgb[4];
    rho_b += D_b * W;
    double W_b = D_b * q.rho;
    rho_b += rhW2_b * h * W * W;
    const double h_b = rhW2_b * q.rho * W * W;
    W_b += rhW2_b * 2.0 * q.rho * h * W;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            v_b[j] += cgb[1 + i] * rhW2 * v.gt.g[i][j];
            ab += cgb[1 + i] * rhW2 * q.v[j] * t.gt_g[3 * i + j];
        }
    eps_b += h_b;
    P_b += h_b / q.rho;
    rho_b -= h_b * q.pressure / (q.rho * q.rho);
    const double v2_b = W_b * 0.5 * W * W * W;
    vsq_bar(q.v, v.gt.g, v2_b, v_b);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            ab += v2_b * q.v[i] * q.v[j] * t.gt_g[3 * i + j];
    /* floors on gt */
    P_b += eps_b / ((gamma - 1.0) * q.rho);
    rho_b -= eps_b * q.eps / q.rho;
    gb -= eps_b * q.eps / (gamma - 1.0);
    double prb[6];
    prb[0] = rho_kept ? rho_b : 0.0;
    prb[1] = 0.0; /* eps is recomputed by the floors */
    prb[2] = p_kept ? P_b : 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        prb[3 + i] = v_b[i];
    if (scaled)
    {
        double scale_b = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            scale_b += v_b[i] * v_in[i];
            prb[3 + i] = v_b[i] * scale;
        }
        double vin_b[3] = {prb[3], prb[4], prb[5]};
        const double v2pre_b = -0.5 * scale_b * scale / v2pre;
        vsq_bar(v_in, v.gt.g, v2pre_b, vin_b);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                ab += v2pre_b * v_in[i] * v_in[j] * t.gt_g[3 * i + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            prb[3 + i] = vin_b[i];
    }
    /* the source cell's recovery on gs */
    root_reverse(u, o, gamma, v.gs.gu, t.gs_gu, r1, prb, ub, gb, ab);
}

/* ------------------------------------------------------------------ */
/* k_adj_prims: lambda_W of every storage cell back to the interior     */
/* state it was recovered from (interior: itself; outflow ghost: its    */
/* clamped source, atomics), plus the parameter adjoints, the spin's    */
/* through the cached metric tangents.                                  */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(kBlock)
    k_adj_prims(BoxDev b, Options o, AdjParams pa, MetricCache mc,
                const double *U, const double *W, const double *lam_W,
                double *lam_U, double *lam_par)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double g0 = 0.0, g1 = 0.0;
    if (it < b.gnpts)
    {
        const int i = b.glo[0] + (int)(it % b.gn[0]);
        const int j = b.glo[1] + (int)((it / b.gn[0]) % b.gn[1]);
        const int k = b.glo[2] + (int)(it / ((long long)b.gn[0] * b.gn[1]));
        const long long n = b.gnpts;
        double lw[6];
        bool any = false;
#pragma unroll
        for (int m = 0; m < 6; ++m)
        {
            lw[m] = lam_W[m * n + it];
            any = any || lw[m] != 0.0;
        }
        const int iv[3] = {i, j, k};
        int src[3] = {i, j, k};
        bool physical = false;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            if (iv[d] < b.lo[d])
            {
                physical = true;
                src[d] = b.lo[d];
            }
            else if (iv[d] > b.hi[d])
            {
                physical = true;
                src[d] = b.hi[d];
            }
        }
        if (any)
        {
            const long long ss = goff(b, src[0], src[1], src[2]);
            const Cons uv = load_cons(U, n, ss);
            const long long is = iidx(b, src[0], src[1], src[2]);
            WChain v;
            w_chain_values(b, pa, W, i, j, k, src, physical, v);
            /* spin tangents of gamma_ij (components 0..8 of the cache) and
             * gamma^ij (9..17) of the cell, and of its source's gamma^ij */
            ChainTan t;
            const long long st = goff(b, i, j, k);
#pragma unroll
            for (int m = 0; m < 9; ++m)
            {
                t.gt_gu[m] = mc.geo[(long long)(9 + m) * n + st];
                t.gt_g[m] = physical ? mc.geo[(long long)m * n + st] : 0.0;
                t.gs_gu[m] = physical ? mc.geo[(long long)(9 + m) * n + ss] : 0.0;
            }
            /* 5 U inputs, Gamma and the spin in one reverse sweep */
            double gu5[5];
            w_chain_reverse(v, pa.gamma, physical, o, uv, t, lw, gu5, g0, g1);
#pragma unroll
            for (int m = 0; m < 5; ++m)
                if (gu5[m] != 0.0)
                    atomicAdd(&lam_U[(long long)m * b.inpts + is], gu5[m]);
        }
    }
    block_add_params(lam_par, g0, g1);
}

/* ------------------------------------------------------------------ */
/* Post-update recovery with atmosphere resets (k_cells' stage tail):   */
/* U_pre -> U_post.                                                     */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ void recover_reset_v(const Cons &un,
                                                const Options &o,
                                                const Geo &g, Cons &out,
                                                Prim &prim, bool &ok,
                                                bool &reset)
{
    Prim rec;
    double res;
    ok = con2prim(un, o, g.gu, rec, res);
    reset = !ok;
    if (ok)
    {
        prim = rec;
        floors(prim, o, g.g);
        reset = prim_differs(rec, prim);
    }
    else
    {
        atmosphere_prim(o, prim);
    }
    out = un;
    if (reset)
        prim2con(prim, o.gamma, g.g, out);
}

/* Dual version with the branch frozen at the value path's decision. */
__device__ __forceinline__ void recover_reset_d(const ConsT<Dual> &un,
                                                const Options &o, Dual gamma,
                                                const GeoT<Dual> &g,
                                                bool ok, bool reset,
                                                ConsT<Dual> &out)
{
    out = un;
    if (!reset)
        return;
    PrimT<Dual> prim;
    if (ok)
    {
        con2prim_ift(un, o, gamma, g.gu, prim);
        floors_g(prim, o, gamma, g.g);
    }
    else
    {
        prim.rho = Dual(o.rho_floor);
        prim.pressure = Dual(o.pressure_floor);
        prim.eps = prim.pressure / ((gamma - 1.0) * prim.rho);
        prim.v[0] = prim.v[1] = prim.v[2] = Dual(0.0);
    }
    prim2con(prim, gamma, g.g, out);
}

/* Horizon excision: an excised cell's rhs is the constant zero
 * (ExcisionTest), so its lambda_k carries nothing back through its fluxes
 * or sources.  Zeroes the 5 rows of lambda_k of every excised cell. */
__global__ void __launch_bounds__(kBlock)
    k_adj_excise(BoxDev b, Options o, AdjParams pa, double *lam_k)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= b.inpts)
        return;
    const int i = b.lo[0] + (int)(it % b.in[0]);
    const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
    const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    if (!ks_excised(pa.M, pa.a, x, o.excise_inner, o.excise_outer))
        return;
    const long long ni = b.inpts;
#pragma unroll
    for (int m = 0; m < 5; ++m)
        lam_k[m * ni + it] = 0.0;
}

/* RK4 stage update of the adjoint's forward sweep: the same combination
 * and order as k_cells<MODE_RK4_s> (U_s = U0 + c_stage k_s; the
 * accumulator U0 + dt/6 k1 + dt/3 k2 + dt/3 k3; U_4 = acc + dt/6 k4), then
 * the post-update recovery with resets.  Writes the stage state into the
 * interior of `X` (5 components over the storage box). */
__global__ void __launch_bounds__(kBlock)
    k_fwd_stage(BoxDev b, Options o, AdjParams pa, int stage, double dt,
                const double *U0, const double *k, double *acc, double *X)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= b.inpts)
        return;
    const int i = b.lo[0] + (int)(it % b.in[0]);
    const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
    const int kk = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
    const long long n = b.gnpts, ni = b.inpts;
    const long long so = goff(b, i, j, kk);
    const double c2 = dt / 2.0, c3 = dt / 3.0, c6 = dt / 6.0;
    const Cons u0 = load_cons(U0, n, so);
    const Cons r = load_cons(k, ni, it);
    Cons un;
    if (stage < 4)
    {
        const double cst = stage == 3 ? dt : c2;
        const double cac = stage == 1 ? c6 : c3;
        Cons ac = stage == 1 ? u0 : load_cons(acc, ni, it);
        ac.D += cac * r.D;
        ac.tau += cac * r.tau;
#pragma unroll
        for (int m = 0; m < 3; ++m)
            ac.S[m] += cac * r.S[m];
        store_cons(acc, ni, it, ac);
        un.D = u0.D + cst * r.D;
        un.tau = u0.tau + cst * r.tau;
#pragma unroll
        for (int m = 0; m < 3; ++m)
            un.S[m] = u0.S[m] + cst * r.S[m];
    }
    else
    {
        const Cons ac = load_cons(acc, ni, it);
        un.D = ac.D + c6 * r.D;
        un.tau = ac.tau + c6 * r.tau;
#pragma unroll
        for (int m = 0; m < 3; ++m)
            un.S[m] = ac.S[m] + c6 * r.S[m];
    }
    double x[3];
    coords(i, j, kk, b.dx, b.center, x);
    Geo g;
    ks_geometry(pa.M, pa.a, x, g);
    Cons out;
    Prim prim;
    bool ok, reset;
    recover_reset_v(un, o, g, out, prim, ok, reset);
    store_cons(X, n, so, out);
}

/* k_adj_recover: lambda of the post-update state back through the
 * recovery with resets to the pre-update state (lambda_pre = R'^T
 * lambda_post), plus the parameter adjoints.  U_pre is rebuilt from the
 * stored U0, k and accumulator exactly as k_fwd_stage formed it. */
__global__ void __launch_bounds__(kBlock)
    k_adj_recover(BoxDev b, Options o, AdjParams pa, int stage, double dt,
                  const double *U0, const double *k, const double *acc_prev,
                  const double *lam_post, double *lam_pre, double *lam_par)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double g0 = 0.0, g1 = 0.0;
    if (it < b.inpts)
    {
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int kk = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        const long long n = b.gnpts, ni = b.inpts;
        const long long so = goff(b, i, j, kk);
        const double c2 = dt / 2.0, c6 = dt / 6.0;
        const Cons r = load_cons(k, ni, it);
        Cons un;
        if (stage < 4)
        {
            const Cons u0 = load_cons(U0, n, so);
            const double cst = stage == 3 ? dt : c2;
            un.D = u0.D + cst * r.D;
            un.tau = u0.tau + cst * r.tau;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                un.S[m] = u0.S[m] + cst * r.S[m];
        }
        else
        {
            const Cons ac = load_cons(acc_prev, ni, it);
            un.D = ac.D + c6 * r.D;
            un.tau = ac.tau + c6 * r.tau;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                un.S[m] = ac.S[m] + c6 * r.S[m];
        }
        double lam[5];
#pragma unroll
        for (int m = 0; m < 5; ++m)
            lam[m] = lam_post[m * ni + it];
        double x[3];
        coords(i, j, kk, b.dx, b.center, x);
        Geo g;
        ks_geometry(pa.M, pa.a, x, g);
        Cons out;
        Prim prim;
        bool ok, reset;
        recover_reset_v(un, o, g, out, prim, ok, reset);
        if (!reset)
        {
#pragma unroll
            for (int m = 0; m < 5; ++m)
                lam_pre[m * ni + it] = lam[m];
        }
        else
        {
            const Dual xd[3] = {Dual(x[0]), Dual(x[1]), Dual(x[2])};
            for (int s = 0; s < 7; ++s)
            {
                const int seed = s < 5 ? 0 : (s == 5 ? SEED_GAMMA : SEED_SPIN);
                Dual M, a, gamma;
                seed_params(pa, seed, M, a, gamma);
                GeoT<Dual> gd;
                ks_geometry(M, a, xd, gd);
                const ConsT<Dual> ud = dual_cons(un, s < 5 ? s : -1);
                ConsT<Dual> od;
                recover_reset_d(ud, o, gamma, gd, ok, reset, od);
                const double gg = dot_tangent(od, lam);
                if (s < 5)
                    lam_pre[(long long)s * ni + it] = gg;
                else if (s == 5)
                    g0 += gg;
                else
                    g1 += gg;
            }
        }
    }
    block_add_params(lam_par, g0, g1);
}

/* Accretion rate through the coordinate sphere |x| = r0 (interior cells
 * with ||x| - r0| < dx/2): -dx^2 sqrt(gamma) D (alpha v^i - beta^i)
 * x_i / |x|, v from con2prim of the state. */
template <class T>
__device__ __forceinline__ T mdot_local(const ConsT<T> &u, const T x[3],
                                        T M, T a, T gamma, const Options &o,
                                        double dx, bool &ok)
{
    GeoT<T> g;
    ks_geometry(M, a, x, g);
    PrimT<T> w;
    if constexpr (sizeof(T) == sizeof(double))
    {
        double res;
        ok = con2prim(u, o, g.gu, w, res);
    }
    else
    {
        ok = con2prim_ift(u, o, gamma, g.gu, w);
    }
    if (!ok)
        return T(0.0);
    const T r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    T radial = T(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        radial += (g.lapse * w.v[i] - g.shift[i]) * x[i];
    const T sqrtg = sqrt(det_sym(g.g));
    return -(dx * dx) * sqrtg * u.D * radial / r;
}

__device__ __forceinline__ bool in_shell(const double x[3], double r0,
                                         double dx)
{
    const double r = ::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    return ::fabs(r - r0) < 0.5 * dx;
}

/* Per-cell accretion-rate terms (0 outside the shell) for a deterministic
 * sum. */
__global__ void __launch_bounds__(kBlock)
    k_mdot_terms(BoxDev b, Options o, AdjParams pa, double r0,
                 const double *U, double *terms)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= b.inpts)
        return;
    const int i = b.lo[0] + (int)(it % b.in[0]);
    const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
    const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    double v = 0.0;
    if (in_shell(x, r0, b.dx))
    {
        const Cons u = load_cons(U, b.gnpts, goff(b, i, j, k));
        bool ok;
        v = mdot_local(u, x, pa.M, pa.a, pa.gamma, o, b.dx, ok);
    }
    terms[it] = v;
}

/* lambda_U += w * d(mdot)/dU over the shell; parameter adjoints += w *
 * d(mdot)/d(Gamma, a). */
__global__ void __launch_bounds__(kBlock)
    k_adj_mdot(BoxDev b, Options o, AdjParams pa, double r0, double w,
               const double *U, double *lam_U, double *lam_par)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double g0 = 0.0, g1 = 0.0;
    if (it < b.inpts)
    {
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        double x[3];
        coords(i, j, k, b.dx, b.center, x);
        if (in_shell(x, r0, b.dx))
        {
            const Cons u = load_cons(U, b.gnpts, goff(b, i, j, k));
            const Dual xd[3] = {Dual(x[0]), Dual(x[1]), Dual(x[2])};
            for (int s = 0; s < 7; ++s)
            {
                const int seed = s < 5 ? 0 : (s == 5 ? SEED_GAMMA : SEED_SPIN);
                Dual M, a, gamma;
                seed_params(pa, seed, M, a, gamma);
                bool ok;
                const Dual m = mdot_local(dual_cons(u, s < 5 ? s : -1), xd, M,
                                          a, gamma, o, b.dx, ok);
                if (!ok)
                    break;
                if (s < 5)
                    lam_U[(long long)s * b.inpts + it] += w * m.d;
                else if (s == 5)
                    g0 += w * m.d;
                else
                    g1 += w * m.d;
            }
        }
    }
    block_add_params(lam_par, g0, g1);
}

/* y += a x over n entries */
__global__ void __launch_bounds__(kBlock)
    k_axpy(long long n, double a, const double *x, double *y)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        y[t] += a * x[t];
}

/* Interior 5-component array <-> the interior of a 5-component storage
 * array (copy in either direction). */
__global__ void __launch_bounds__(kBlock)
    k_interior_copy(BoxDev b, const double *src, double *dst, int to_storage)
{
    const long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= b.inpts)
        return;
    const int i = b.lo[0] + (int)(it % b.in[0]);
    const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
    const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
    const long long so = goff(b, i, j, k);
#pragma unroll
    for (int m = 0; m < 5; ++m)
    {
        if (to_storage)
            dst[m * b.gnpts + so] = src[m * b.inpts + it];
        else
            dst[m * b.inpts + it] = src[m * b.gnpts + so];
    }
}

/* Deterministic double-double sum of n terms: per-block partials, then
 * one block over the partials in order (k_diag_* use the same scheme). */
__global__ void __launch_bounds__(kBlock)
    k_sum_partial(const double *x, long long n, double *part)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    DD acc[5];
    double plain[5];
    double bad = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m)
    {
        acc[m] = DD{0.0, 0.0};
        plain[m] = 0.0;
    }
    if (t < n)
    {
        acc[0].hi = x[t];
        plain[0] = x[t];
    }
    block_reduce_diag(acc, plain, bad);
    if (threadIdx.x == 0)
    {
        part[2 * blockIdx.x] = acc[0].hi;
        part[2 * blockIdx.x + 1] = acc[0].lo;
    }
}

__global__ void __launch_bounds__(kBlock)
    k_sum_final(const double *part, long long nblocks, double *out)
{
    DD acc[5];
    double plain[5];
    double bad = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m)
    {
        acc[m] = DD{0.0, 0.0};
        plain[m] = 0.0;
    }
    for (long long q = threadIdx.x; q < nblocks; q += kBlock)
        acc[0] = dd_add(acc[0], DD{part[2 * q], part[2 * q + 1]});
    block_reduce_diag(acc, plain, bad);
    if (threadIdx.x == 0)
        *out = acc[0].hi;
}

} // namespace grhd

#endif /* GRHD_ADJOINT_HPP_ */
