/*
 * grhd_physics.hpp -- device-inline GRHD numerics for CDNA4 (gfx950).
 *
 * Per-cell and per-face arithmetic of GRDzhadzha's fixed-background GRHD
 * path, in binary64, with the reference's operation order so that results
 * (and the threshold decisions behind the atmosphere and failure masks)
 * agree bit for bit with a CPU evaluation of the same formulas.  The whole
 * translation unit is compiled with -ffp-contract=off: no multiply-add is
 * fused, as on the CPU.  pow(x, -0.5), pow(x, 2), pow(x, 3) of the
 * Kerr-Schild lapse are written 1/sqrt(x), x*x, x*x*x, and the natural log
 * is grhd_log below, so only correctly rounded operations (+ - *, and the
 * division and square root of qdiv and qsqrt, IEEE over the ranges stated
 * there) and frexp appear.
 *
 * Citations are to the reference repository (Source/...).
 */
#ifndef GRHD_PHYSICS_HPP_
#define GRHD_PHYSICS_HPP_

#include <hip/hip_runtime.h>

namespace grhd
{

/* The face-side arithmetic (metric for the face geometry, reconstruction,
 * HLLE) is templated on its scalar type T: double for the bit-exact path,
 * float for the mixed-precision faces.  Constants are written T(...), so
 * the double instantiation is the same arithmetic as plain double code. */
template <class T> struct PrimT
{
    T rho, eps, pressure, v[3];
};

template <class T> struct ConsT
{
    T D, S[3], tau;
};

/* CellGeometryData without K (StaticMetric.hpp:19-40); the flux and
 * recovery paths never read K_ij. */
template <class T> struct GeoT
{
    T g[3][3];
    T gu[3][3];
    T lapse;
    T shift[3];
};

using Prim = PrimT<double>;
using Cons = ConsT<double>;
using Geo = GeoT<double>;

/* Geometry plus everything the static-metric sources read
 * (StaticMetric.hpp:79-102, FixedBGMetric.hpp:17-55). */
template <class T> struct FullGeoT
{
    GeoT<T> geo;
    T K[3][3];
    T dlapse[3];
    T dlogsqrtdet[3];
    T dg[3][3][3];    /* [dir][i][j] */
    T dshift[3][3];   /* [i][dir]    */
};
using FullGeo = FullGeoT<double>;

/* Operator options passed by value to kernels. */
struct Options
{
    double gamma;
    double rho_floor, pressure_floor, max_v2, source_atmosphere_factor;
    int max_iterations;
    double tolerance, rec_pressure_floor, rec_max_v2;
    double theta;
    int use_reconstruction, use_static_sources, use_volume_sources;
    int mixed; /* FP32 reconstruction, face metric and HLLE */
    /* mixed only: FP64 faces beside cells near a floor or the v^2 cap
     * (guard_margin > 0), or where the shift is large (guard_shift > 0);
     * see fp64_cell.  Both 0: every face FP32 */
    double guard_margin, guard_shift;
    /* mixed only: Kerr-Schild static-metric sources in FP32 (cells
     * without kWGuard) */
    int fp32_sources;
    int excise; /* zero rhs inside the buffered Kerr-Schild horizons */
    double excise_inner, excise_outer;
};

/* A background evaluated at one time (background_factory(time)); the
 * binary's positions and velocities come from the host (cos/sin there). */
struct Background
{
    int kind; /* 0 = Kerr-Schild, 1 = circular binary PN */
    double M, a;
    double m1, m2, soft;
    double p1[3], p2[3], u1[3], u2[3];
};

template <class T> __device__ __forceinline__ T std_max(T a, T b)
{
    return (a < b) ? b : a;
}
template <class T> __device__ __forceinline__ T std_min(T a, T b)
{
    return (b < a) ? b : a;
}
/* Division and square root of the device numerics.
 *
 * float (the mixed-precision faces and sources): the hardware reciprocal
 * and square root, about one ulp, without the range-scaling steps of the
 * general FP32 sequences (their values lie far from the denormal and
 * overflow ranges those steps guard).
 *
 * double: correctly rounded, as IEEE a / b and sqrt(a).  The compiler's
 * FP64 division is a Newton-Raphson refinement of the hardware reciprocal
 * (v_rcp_f64, two steps), the quotient corrected once by its residual,
 * wrapped in v_div_scale (both operands scaled by 2^+-128 when their
 * exponents would make an intermediate overflow or go subnormal) and
 * v_div_fixup (zeros, infinities, NaN, overflow, underflow).  qdiv keeps
 * the refinement and v_div_fixup and drops the scaling: 9 instructions
 * instead of 11, none of which writes VCC, and the refined reciprocal of
 * a denominator is shared by every division by it.  Where the scaling
 * would do nothing -- a = 0 or |a| >= 2^-969, b normal with |b| < 2^1022,
 * exponent(a) - exponent(b) < 768, a normal quotient -- the instructions
 * are the compiler's, so the result is the IEEE quotient bit for bit;
 * zero, infinite and NaN operands get the IEEE result from v_div_fixup.
 * The FP64 square root likewise drops the 2^256 scaling of arguments
 * below 2^-767: above that bound (and for +-0, +inf, negative and NaN
 * arguments) it is the IEEE square root bit for bit.  Outside these
 * ranges (operands below about 1e-290 or above 1e300) a result may differ
 * from IEEE's by amounts of that size.  The kernels' values (densities
 * and pressures at or above their floors, metric components, velocities
 * that are zero or physical) lie inside them.  The adjoint's dual numbers
 * use their own operators. */
template <class T> __device__ __forceinline__ T qdiv(T a, T b)
{
    return a / b;
}
__device__ __forceinline__ float qdiv(float a, float b)
{
    return a * __builtin_amdgcn_rcpf(b);
}
__device__ __forceinline__ double qdiv(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    const double q = a * r;
    const double rem = fma(-b, q, a);
    return __builtin_amdgcn_div_fixup(fma(rem, r, q), b, a);
}
template <class T> __device__ __forceinline__ T qsqrt(T a)
{
    return sqrt(a);
}
__device__ __forceinline__ float qsqrt(float a)
{
    return __builtin_amdgcn_sqrtf(a);
}
/* Goldschmidt-style refinement of the hardware reciprocal square root
 * y ~ 1/sqrt(x): s ~ sqrt(x) and h ~ y/2 improved together, then two
 * residual corrections of s; +-0 and +inf (class mask 0x260) return x. */
__device__ __forceinline__ double qsqrt(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y;
    double h = y * 0.5;
    const double r = fma(-h, s, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    double d = fma(-s, s, x);
    s = fma(d, h, s);
    d = fma(-s, s, x);
    s = fma(d, h, s);
    return __builtin_amdgcn_class(x, 0x260) ? x : s;
}

template <class T = double> __device__ __forceinline__ T kdelta(int i, int j)
{
    return (i == j) ? T(1) : T(0);
}

/* KerrSchild::check_if_excised (KerrSchild.hpp:247-280): x lies inside the
 * outer horizon's oblate ellipsoid (x^2 + y^2) / (2 M r_+) + z^2 / r_+^2 = 1
 * scaled by outer_buffer, or inside the inner horizon's scaled by
 * inner_buffer.  With a = 0 the inner ratio is inf or NaN: never inside. */
__device__ __forceinline__ bool ks_excised(double M, double a,
                                           const double x[3],
                                           double inner_buffer,
                                           double outer_buffer)
{
    const double root = sqrt(M * M - a * a);
    const double r_out = M + root, r_in = M - root;
    const double cyl2 = x[0] * x[0] + x[1] * x[1];
    const double z2 = x[2] * x[2];
    const double outer = cyl2 / (2.0 * M * r_out) + z2 / r_out / r_out;
    const double inner = cyl2 / (2.0 * M * r_in) + z2 / r_in / r_in;
    return outer < outer_buffer || inner < inner_buffer;
}

/* ---------------- GRChombo TensorAlgebra (restated) ---------------- */
template <class T> __device__ __forceinline__ T det_sym(const T m[3][3])
{
    return m[0][0] * m[1][1] * m[2][2] + 2 * m[0][1] * m[0][2] * m[1][2] -
           m[0][0] * m[1][2] * m[1][2] - m[1][1] * m[0][2] * m[0][2] -
           m[2][2] * m[0][1] * m[0][1];
}

template <class T>
__device__ __forceinline__ void inverse_sym(const T m[3][3], T out[3][3])
{
    const T deth = det_sym(m);
    const T inv = qdiv(T(1), deth);
    out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[1][2]) * inv;
    out[0][1] = (m[0][2] * m[1][2] - m[0][1] * m[2][2]) * inv;
    out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[0][2]) * inv;
    out[1][2] = (m[0][1] * m[0][2] - m[0][0] * m[1][2]) * inv;
    out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[0][1]) * inv;
    out[1][0] = out[0][1];
    out[2][0] = out[0][2];
    out[2][1] = out[1][2];
}

/* IEEE-only natural log (k ln2 + 2 atanh((m-1)/(m+1))), ~1 ulp. */
__device__ __forceinline__ double grhd_log(double x)
{
    if (!(x > 0.0) || x == __builtin_inf())
    {
        if (x == 0.0)
            return -__builtin_inf();
        if (x == __builtin_inf())
            return __builtin_inf();
        return __builtin_nan("");
    }
    int k;
    double m = frexp(x, &k);
    if (m < 0.70710678118654752440)
    {
        m = m * 2.0;
        k = k - 1;
    }
    const double s = qdiv(m - 1.0, m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 25.0;
    p = 1.0 / 23.0 + s2 * p;
    p = 1.0 / 21.0 + s2 * p;
    p = 1.0 / 19.0 + s2 * p;
    p = 1.0 / 17.0 + s2 * p;
    p = 1.0 / 15.0 + s2 * p;
    p = 1.0 / 13.0 + s2 * p;
    p = 1.0 / 11.0 + s2 * p;
    p = 1.0 / 9.0 + s2 * p;
    p = 1.0 / 7.0 + s2 * p;
    p = 1.0 / 5.0 + s2 * p;
    p = 1.0 / 3.0 + s2 * p;
    const double two_s = 2.0 * s;
    const double log_m = two_s + two_s * (s2 * p);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double dk = (double)k;
    return (dk * ln2_hi + (log_m + dk * ln2_lo));
}

/* ---------------- Coordinates ---------------- */
__device__ __forceinline__ void coords(int i, int j, int k, double dx,
                                       const double c[3], double x[3])
{
    x[0] = (i + 0.5) * dx - c[0];
    x[1] = (j + 0.5) * dx - c[1];
    x[2] = (k + 0.5) * dx - c[2];
}

template <class T> __device__ __forceinline__ T get_radius(const T x[3])
{
    const T r = qsqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    return std_max(r, T(1e-6));
}

/* ---------------- Kerr-Schild (KerrSchild.hpp:65-166) ---------------- */

/* Radius, H and l_i shared by the geometry-only and full evaluations. */
template <class T> struct KSCoreT
{
    T rho, rho2, r2, r, cos_theta, H, el[3];
};
using KSCore = KSCoreT<double>;

template <class T>
__device__ __forceinline__ void ks_core(T M, T a, const T x[3], KSCoreT<T> &c)
{
    const T a2 = a * a;
    const T z = x[2];
    c.rho = get_radius(x);
    c.rho2 = c.rho * c.rho;
    c.r2 = T(0.5) * (c.rho2 - a2) +
           qsqrt(T(0.25) * (c.rho2 - a2) * (c.rho2 - a2) + a2 * z * z);
    c.r = qsqrt(c.r2);
    c.cos_theta = qdiv(z, c.r);
    c.H = qdiv(M * c.r, c.r2 + a2 * c.cos_theta * c.cos_theta);
    c.el[0] = qdiv(c.r * x[0] + a * x[1], c.r2 + a2);
    c.el[1] = qdiv(c.r * x[1] - a * x[0], c.r2 + a2);
    c.el[2] = qdiv(z, c.r);
}

/* gamma, gamma^-1, lapse, shift (KerrSchild.hpp:79-116 with
 * FixedBGMetric.hpp:23-24); `c` keeps r, H and l_i for ks_derivs. */
template <class T>
__device__ __forceinline__ void ks_geometry_core(T M, T a, const T x[3],
                                                 GeoT<T> &g, KSCoreT<T> &c)
{
    ks_core(M, a, x, c);
    const T el_t = T(1);
    g.lapse = qdiv(T(1), qsqrt(T(1) + T(2) * c.H * el_t * el_t));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            g.g[i][j] = kdelta<T>(i, j) + T(2) * c.H * c.el[i] * c.el[j];
    if constexpr (sizeof(T) < sizeof(double))
    {
        /* Reduced precision: the cofactor inverse of gamma sums terms of
         * order one that cancel to order H; the Sherman-Morrison form of
         * the inverse of delta + 2H l l^T has no cancellation:
         * gamma^ij = delta^ij - q l^i l^j, beta^i = q l^i, with
         * q = 2H / (1 + 2H l.l). */
        const T ll = c.el[0] * c.el[0] + c.el[1] * c.el[1] + c.el[2] * c.el[2];
        const T q = qdiv(T(2) * c.H, T(1) + T(2) * c.H * ll);
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                g.gu[i][j] = kdelta<T>(i, j) - q * c.el[i] * c.el[j];
            g.shift[i] = q * c.el[i] * el_t;
        }
    }
    else
    {
        /* KerrSchild.hpp:108 inverse; FixedBGMetric.hpp:23-24 inverse
         * again: both are the same function of the same gamma. */
        inverse_sym(g.g, g.gu);
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            g.shift[i] = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
                g.shift[i] += g.gu[i][j] * T(2) * c.H * c.el[j] * el_t;
        }
    }
}

/* gamma, gamma^-1, lapse, shift only (derivatives are not needed by the
 * flux/recovery paths and do not feed these values). */
template <class T>
__device__ __forceinline__ void ks_geometry(T M, T a, const T x[3],
                                            GeoT<T> &g)
{
    KSCoreT<T> c;
    ks_geometry_core(M, a, x, g, c);
}

/* Reduced precision: K_ij of a metric with time derivative kdiag delta_ij
 * (zero for Kerr-Schild) as the Lie derivative of gamma along the shift,
 * K_ij = (beta^l d_l gamma_ij + gamma_lj d_i beta^l + gamma_il d_j beta^l
 *         - kdiag delta_ij) / (2 alpha),
 * which equals the Christoffel form below (KerrSchild.hpp:146-164) in
 * exact arithmetic: the Christoffel terms reduce to beta^l d_l gamma_ij
 * once Gamma^k_ij beta_k is written as beta^l Gamma_lij.  It is symmetric
 * by construction and takes a ninth of the Christoffel form's operations;
 * only the float instantiations use it (the FP64 path keeps the
 * reference's form and rounding). */
template <class T>
__device__ __forceinline__ void lie_extrinsic_curvature(const GeoT<T> &g,
                                                        const T dg[3][3][3],
                                                        const T dshift[3][3],
                                                        T kdiag, T K[3][3])
{
    const T half_over_lapse = qdiv(T(0.5), g.lapse);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j)
        {
            T v = (i == j) ? -kdiag : T(0);
#pragma unroll
            for (int l = 0; l < 3; ++l)
                v += g.shift[l] * dg[l][i][j] + g.g[l][j] * dshift[l][i] +
                     g.g[i][l] * dshift[l][j];
            K[i][j] = v * half_over_lapse;
            K[j][i] = K[i][j];
        }
}

/* The derivative half of the full ADM evaluation, given the geometry half
 * (KerrSchild.hpp:118-242 + FixedBGMetric.hpp:32-55): dgamma, dalpha,
 * dbeta, K_ij and d ln sqrt gamma. */
template <class T>
__device__ __forceinline__ void ks_derivs(T M, T a, const T xc[3],
                                          const KSCoreT<T> &c,
                                          const GeoT<T> &gin, FullGeoT<T> &f)
{
    (void)M;
    const T a2 = a * a;
    const T z = xc[2];
    const T el_t = T(1.0);
    const T H = c.H;
    const T r = c.r, r2 = c.r2, rho = c.rho, rho2 = c.rho2;

    /* get_KS_derivs (KerrSchild.hpp:171-242) */
    const T cos_theta2 = c.cos_theta * c.cos_theta;
    T drdx[3], dHdx[3], dldx[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const T drhodx = qdiv(xc[i], rho);
        drdx[i] = qdiv(T(0.5), r) *
                  (rho * drhodx +
                   qdiv(T(0.5), qsqrt(T(0.25) * (rho2 - a2) * (rho2 - a2) + a2 * z * z)) *
                       (drhodx * rho * (rho2 - a2) +
                        kdelta<T>(i, 2) * T(2.0) * a2 * z));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const T dcosthetadx = qdiv(-z, r2) * drdx[i] + qdiv(kdelta<T>(i, 2), r);
        dHdx[i] = H * (qdiv(drdx[i], r) -
                       qdiv(T(2.0), r2 + a2 * cos_theta2) *
                           (r * drdx[i] + a2 * c.cos_theta * dcosthetadx));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        dldx[0][i] = qdiv(
            xc[0] * drdx[i] + r * kdelta<T>(i, 0) + a * kdelta<T>(i, 1) -
                qdiv(T(2.0) * r * drdx[i] * (r * xc[0] + a * xc[1]), r2 + a2),
            r2 + a2);
        dldx[1][i] = qdiv(
            xc[1] * drdx[i] + r * kdelta<T>(i, 1) - a * kdelta<T>(i, 0) -
                qdiv(T(2.0) * r * drdx[i] * (r * xc[1] - a * xc[0]), r2 + a2),
            r2 + a2);
        dldx[2][i] = qdiv(-xc[2] * drdx[i], r2) + qdiv(kdelta<T>(i, 2), r);
    }
    const T dltdx = T(0.0);

    f.geo = gin;
    const GeoT<T> &g = f.geo;

    /* d1_gamma[i][j][k] stored as dg[k][i][j] */
    if constexpr (sizeof(T) < sizeof(double))
    {
        /* reduced precision: the symmetric half, mirrored */
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = i; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                {
                    f.dg[k][i][j] =
                        T(2.0) * (c.el[i] * c.el[j] * dHdx[k] +
                                  H * c.el[i] * dldx[j][k] +
                                  H * c.el[j] * dldx[i][k]);
                    f.dg[k][j][i] = f.dg[k][i][j];
                }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    f.dg[k][i][j] =
                        T(2.0) * (c.el[i] * c.el[j] * dHdx[k] +
                               H * c.el[i] * dldx[j][k] + H * c.el[j] * dldx[i][k]);
    }

    const T lapse2 = g.lapse * g.lapse;
    const T lapse3 = g.lapse * g.lapse * g.lapse;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        f.dlapse[i] = -lapse3 * el_t * (el_t * dHdx[i] + T(2.0) * H * dltdx);

#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            f.dshift[i][j] = T(2.0) * el_t * dHdx[j] * lapse2 * c.el[i] +
                             T(4.0) * el_t * H * g.lapse * f.dlapse[j] * c.el[i] +
                             T(2.0) * el_t * H * lapse2 * dldx[i][j] +
                             T(2.0) * dltdx * H * lapse2 * c.el[i];

    /* K_ij with the Christoffel symbols formed per (i, j)
     * (GRChombo compute_christoffel, KerrSchild.hpp:146-164) */
    if constexpr (sizeof(T) < sizeof(double))
        lie_extrinsic_curvature(g, f.dg, f.dshift, T(0), f.K);
    else
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            T LLL[3];
#pragma unroll
            for (int l = 0; l < 3; ++l)
                LLL[l] = T(0.5) * (f.dg[j][i][l] + f.dg[i][j][l] - f.dg[l][i][j]);
            T Kij = T(0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                T ULL = 0;
#pragma unroll
                for (int l = 0; l < 3; ++l)
                    ULL += g.gu[k][l] * LLL[l];
                Kij += g.g[k][j] * f.dshift[k][i] + g.g[k][i] * f.dshift[k][j] +
                       (f.dg[j][k][i] + f.dg[i][k][j]) * g.shift[k];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    Kij += -T(2.0) * ULL * g.g[k][m] * g.shift[m];
            }
            Kij *= qdiv(T(0.5), g.lapse);
            f.K[i][j] = Kij;
        }

    /* make_static_metric_derivatives_from_adm (FixedBGMetric.hpp:32-55) */
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T s = T(0.0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                s += T(0.5) * g.gu[i][j] * f.dg[d][i][j];
        f.dlogsqrtdet[d] = s;
    }
}

/* Full ADM evaluation converted to geometry + source derivatives
 * (KerrSchild.hpp:65-242 + FixedBGMetric.hpp:17-55). */
__device__ __forceinline__ void ks_full(double M, double a, const double xc[3],
                                        FullGeo &f)
{
    KSCore c;
    Geo g;
    ks_geometry_core(M, a, xc, g, c);
    ks_derivs(M, a, xc, c, g, f);
}

/* ---------------- Circular binary PN (CircularBinaryPN.hpp) ---------------- */
template <class T> struct PotentialsT
{
    T scalar, scalar_dt, dscalar[3], vec[3], dvec[3][3];
};
using Potentials = PotentialsT<double>;

template <class T>
__device__ __forceinline__ void pn_add_point(PotentialsT<T> &p, T mass, T soft,
                                             const T src[3], const T vel[3],
                                             const T pos[3])
{
    T disp[3];
    T r2 = soft * soft;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        disp[i] = pos[i] - src[i];
        r2 += disp[i] * disp[i];
    }
    const T ir = qdiv(T(1), qsqrt(r2));
    const T ir3 = ir * ir * ir;
    p.scalar += mass * ir;
    T ddv = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        ddv += disp[i] * vel[i];
    p.scalar_dt += mass * ddv * ir3;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        p.dscalar[d] += -mass * disp[d] * ir3;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            p.dvec[c][d] += -mass * vel[c] * disp[d] * ir3;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        p.vec[c] += mass * vel[c] * ir;
}

template <class T>
__device__ __forceinline__ void pn_potentials(const Background &bg,
                                              const T x[3], PotentialsT<T> &p)
{
    p.scalar = T(0);
    p.scalar_dt = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        p.dscalar[i] = T(0);
        p.vec[i] = T(0);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            p.dvec[i][j] = T(0);
    }
    const T p1[3] = {T(bg.p1[0]), T(bg.p1[1]), T(bg.p1[2])};
    const T p2[3] = {T(bg.p2[0]), T(bg.p2[1]), T(bg.p2[2])};
    const T u1[3] = {T(bg.u1[0]), T(bg.u1[1]), T(bg.u1[2])};
    const T u2[3] = {T(bg.u2[0]), T(bg.u2[1]), T(bg.u2[2])};
    pn_add_point(p, T(bg.m1), T(bg.soft), p1, u1, x);
    pn_add_point(p, T(bg.m2), T(bg.soft), p2, u2, x);
}

/* Conformally flat geometry of the binary (CircularBinaryPN.hpp:80-130);
 * `p` keeps the potentials for pn_derivs. */
template <class T>
__device__ __forceinline__ void pn_geometry_pot(const Background &bg,
                                                const T x[3], GeoT<T> &g,
                                                PotentialsT<T> &p)
{
    pn_potentials(bg, x, p);
    const T U = p.scalar;
    const T cf = T(1) + T(2) * U;
    T vsq = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        vsq += p.vec[i] * p.vec[i];
    const T lapse_squared =
        T(1) - T(2) * U + T(2) * U * U + qdiv(T(16) * vsq, cf);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            g.g[i][j] = T(0);
    g.lapse = qsqrt(lapse_squared);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        g.g[i][i] = cf;
        g.shift[i] = qdiv(T(-4) * p.vec[i], cf);
    }
    inverse_sym(g.g, g.gu);
}

template <class T>
__device__ __forceinline__ void pn_geometry(const Background &bg,
                                            const T x[3], GeoT<T> &g)
{
    PotentialsT<T> p;
    pn_geometry_pot(bg, x, g, p);
}

/* The derivative half of the binary's ADM evaluation given its geometry
 * (CircularBinaryPN.hpp:132-188 + FixedBGMetric.hpp:32-55). */
template <class T>
__device__ __forceinline__ void pn_derivs(const PotentialsT<T> &p,
                                          const GeoT<T> &gin, FullGeoT<T> &f)
{
    f.geo = gin;
    const GeoT<T> &g = f.geo;
    const T U = p.scalar;
    const T dcf_dt = T(2.0) * p.scalar_dt;
    const T cf = T(1.0) + T(2.0) * U;
    T vsq = T(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        vsq += p.vec[i] * p.vec[i];
    const T lapse_squared = T(1.0) - T(2.0) * U + T(2.0) * U * U + qdiv(T(16.0) * vsq, cf);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        f.dlapse[i] = T(0.0);
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            f.dshift[i][j] = T(0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                f.dg[k][i][j] = T(0.0);
        }
    }
    (void)lapse_squared;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const T dU = p.dscalar[d];
        const T dcf = T(2.0) * dU;
#pragma unroll
        for (int m = 0; m < 3; ++m)
            f.dg[d][m][m] = dcf;
        T dvsq = T(0.0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            dvsq += T(2.0) * p.vec[c] * p.dvec[c][d];
            f.dshift[c][d] =
                -T(4.0) * (qdiv(p.dvec[c][d], cf) - qdiv(p.vec[c] * dcf, cf * cf));
        }
        const T dls = (-T(2.0) + T(4.0) * U) * dU +
                           T(16.0) * (qdiv(dvsq, cf) - qdiv(vsq * dcf, cf * cf));
        f.dlapse[d] = qdiv(T(0.5) * dls, g.lapse);
    }
    if constexpr (sizeof(T) < sizeof(double))
        lie_extrinsic_curvature(g, f.dg, f.dshift, dcf_dt, f.K);
    else
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            T LLL[3];
#pragma unroll
            for (int l = 0; l < 3; ++l)
                LLL[l] = T(0.5) * (f.dg[j][i][l] + f.dg[i][j][l] - f.dg[l][i][j]);
            T Kij = T(0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                T ULL = T(0);
#pragma unroll
                for (int l = 0; l < 3; ++l)
                    ULL += g.gu[k][l] * LLL[l];
                Kij += g.g[k][j] * f.dshift[k][i] + g.g[k][i] * f.dshift[k][j] +
                       (f.dg[j][k][i] + f.dg[i][k][j]) * g.shift[k];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    Kij += -T(2.0) * ULL * g.g[k][m] * g.shift[m];
            }
            if (i == j)
                Kij -= dcf_dt;
            Kij *= qdiv(T(0.5), g.lapse);
            f.K[i][j] = Kij;
        }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T s = T(0.0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                s += T(0.5) * g.gu[i][j] * f.dg[d][i][j];
        f.dlogsqrtdet[d] = s;
    }
}

__device__ __forceinline__ void pn_full(const Background &bg,
                                        const double x[3], FullGeo &f)
{
    Potentials p;
    Geo g;
    pn_geometry_pot(bg, x, g, p);
    pn_derivs(p, g, f);
}

/* What the derivative half of either background needs from its geometry
 * half. */
struct GeoCore
{
    KSCore ks;
    Potentials pn;
};

__device__ __forceinline__ void cell_geometry_core(const Background &bg,
                                                   const double x[3], Geo &g,
                                                   GeoCore &core)
{
    if (bg.kind == 1)
        pn_geometry_pot(bg, x, g, core.pn);
    else
        ks_geometry_core(bg.M, bg.a, x, g, core.ks);
}

__device__ __forceinline__ void cell_derivs(const Background &bg,
                                            const double x[3],
                                            const GeoCore &core, const Geo &g,
                                            FullGeo &f)
{
    if (bg.kind == 1)
        pn_derivs(core.pn, g, f);
    else
        ks_derivs(bg.M, bg.a, x, core.ks, g, f);
}

template <class T>
__device__ __forceinline__ void cell_geometry(const Background &bg,
                                              const T x[3], GeoT<T> &g)
{
    if (bg.kind == 1)
        pn_geometry(bg, x, g);
    else
        ks_geometry(T(bg.M), T(bg.a), x, g);
}

__device__ __forceinline__ void cell_full(const Background &bg,
                                          const double x[3], FullGeo &f)
{
    if (bg.kind == 1)
        pn_full(bg, x, f);
    else
        ks_full(bg.M, bg.a, x, f);
}

/* log sqrt det gamma (StaticMetric.hpp:293-302) */
__device__ __forceinline__ double log_sqrt_det(const Geo &g)
{
    const double det = det_sym(g.g);
    if (det <= 0.0 || !isfinite(det))
        return 0.0;
    return 0.5 * grhd_log(det);
}

/* ---------------- Valencia (Valencia.hpp) ---------------- */
/* The geometry-taking functions below have a second type G for the metric
 * (default T): the adjoint evaluates fluid tangents on a metric held in
 * double. */
template <class T, class G = T>
__device__ __forceinline__ T vsq(const T v[3], const G g[3][3])
{
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            s += g[i][j] * v[i] * v[j];
    return s;
}

/* compute_conserved (Valencia.hpp:55-73) */
template <class T, class G = T>
__device__ __forceinline__ void prim2con(const PrimT<T> &p, T gamma,
                                         const G g[3][3], ConsT<T> &c)
{
    (void)gamma;
    T vL[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        vL[i] = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vL[i] += g[i][j] * p.v[j];
    }
    if constexpr (sizeof(T) < sizeof(double))
    {
        /* Reduced precision: tau = rho h W^2 - P - D cancels to
         * rho (eps + O(v^2)) and would lose most of its float digits.  The
         * same quantity without cancellation, from x = W^2 - 1 =
         * v^2 / (1 - v^2):  tau = rho W x/(W + 1) + rho eps W^2 + P x. */
        const T v2 = vsq(p.v, g);
        const T x = qdiv(v2, T(1) - v2);
        const T W = qdiv(T(1), qsqrt(T(1) - v2));
        const T h = T(1) + p.eps + qdiv(p.pressure, p.rho);
        const T rhW2 = p.rho * h * W * W;
        c.D = p.rho * W;
        c.tau = p.rho * W * qdiv(x, W + T(1)) + p.rho * p.eps * W * W +
                p.pressure * x;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            c.S[i] = rhW2 * vL[i];
    }
    else
    {
        const T W = qdiv(T(1), qsqrt(T(1) - vsq(p.v, g)));
        const T h = T(1) + p.eps + qdiv(p.pressure, p.rho);
        const T rhW2 = p.rho * h * W * W;
        c.D = p.rho * W;
        c.tau = rhW2 - p.pressure - c.D;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            c.S[i] = rhW2 * vL[i];
    }
}

/* The pressure residual's intermediate values at p: E + p, the raw and
 * clamped v^2 and W.  The Newton derivative at the same p and the final
 * primitives reuse them instead of recomputing the same divisions and
 * square roots (the values are the same operations on the same inputs). */
struct PresidAt
{
    double epp, v2raw, v2, W;
};

__device__ __forceinline__ double presid_at(double p, double D, double tau,
                                            double S2, double gamma,
                                            double max_v2, PresidAt &s)
{
    s.epp = tau + D + p;
    s.v2raw = qdiv(S2, s.epp * s.epp);
    s.v2 = s.v2raw > max_v2 ? max_v2 : s.v2raw;
    s.W = qdiv(1.0, qsqrt(1.0 - s.v2));
    return D * s.W + qdiv(gamma * p * s.W * s.W, gamma - 1.0) - s.epp;
}

__device__ __forceinline__ double presid(double p, double D, double tau,
                                         double S2, double gamma,
                                         double max_v2)
{
    PresidAt s;
    return presid_at(p, D, tau, S2, gamma, max_v2, s);
}

/* d residual / dp at the p of `s` */
__device__ __forceinline__ double presid_deriv_at(double p, double D,
                                                  double gamma,
                                                  const PresidAt &s)
{
    const double W = s.W;
    const double dW_dp = qdiv(-W * W * W * s.v2, s.epp);
    const double dW2_dp = 2.0 * W * dW_dp;
    return D * dW_dp + qdiv(gamma * (W * W + p * dW2_dp), gamma - 1.0) - 1.0;
}

/* recover_primitive (Valencia.hpp:116-228).  Returns success; `residual`
 * receives the final residual (0 for the early D <= 0 / bracket exits). */
__device__ __forceinline__ bool con2prim(const Cons &c, const Options &o,
                                         const double gu[3][3], Prim &out,
                                         double &residual_out)
{
    residual_out = 0.0;
    if (c.D <= 0.0)
        return false;
    const double gamma = o.gamma;
    double S2 = 0.;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n)
            S2 += gu[m][n] * c.S[m] * c.S[n];
    const double mv2 = o.rec_max_v2;
    double lower = o.rec_pressure_floor;
    const double f_lower = presid(lower, c.D, c.tau, S2, gamma, mv2);
    if (f_lower >= 0.0 && fabs(f_lower) > o.tolerance)
        return false;
    double upper = lower;
    if (f_lower < 0.0)
    {
        upper = std_max(1.0, fabs(c.tau) + c.D);
        double f_upper = presid(upper, c.D, c.tau, S2, gamma, mv2);
        int it = 0;
        while (f_upper <= 0.0 && it < o.max_iterations)
        {
            upper *= 2.0;
            f_upper = presid(upper, c.D, c.tau, S2, gamma, mv2);
            ++it;
        }
        if (f_upper <= 0.0)
            return false;
    }
    double p = (fabs(f_lower) <= o.tolerance)
                   ? lower
                   : std_min(std_max(c.tau * (gamma - 1.0), lower), upper);
    PresidAt at;
    double res = presid_at(p, c.D, c.tau, S2, gamma, mv2, at);
    for (int iter = 0; iter < o.max_iterations; ++iter)
    {
        if (fabs(res) < o.tolerance)
            break;
        if (f_lower < 0.0)
        {
            if (res > 0.0)
                upper = p;
            else
                lower = p;
        }
        const double der = presid_deriv_at(p, c.D, gamma, at);
        double np = p - qdiv(res, der);
        if (!isfinite(np) || np <= lower || np >= upper)
            np = 0.5 * (lower + upper);
        p = np;
        res = presid_at(p, c.D, c.tau, S2, gamma, mv2, at);
    }
    residual_out = res;
    if (fabs(res) >= o.tolerance)
        return false;
    /* E + p and v^2 at the final p; below the cap W is the residual's */
    const double epp = at.epp;
    if (at.v2raw >= mv2)
        return false;
    const double W = at.W;
    out.rho = qdiv(c.D, W);
    out.pressure = p;
    out.eps = qdiv(p, (gamma - 1.0) * out.rho);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        double vi = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vi += qdiv(gu[i][j] * c.S[j], epp);
        out.v[i] = vi;
    }
    return true;
}

/* compute_flux (Valencia.hpp:230-249) */
template <class T, class G = T>
__device__ __forceinline__ void flux_of(const PrimT<T> &p, const ConsT<T> &c,
                                        int d, G lapse, const G shift[3],
                                        ConsT<T> &f)
{
    const T tv = lapse * p.v[d] - shift[d];
    f.D = c.D * tv;
    f.tau = c.tau * tv + lapse * p.pressure * p.v[d];
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        f.S[j] = c.S[j] * tv;
        if (j == d)
            f.S[j] += lapse * p.pressure;
    }
}

/* ---------------- Atmosphere (Atmosphere.hpp:24-60) ---------------- */
template <class T, class G = T>
__device__ __forceinline__ void floors_g(PrimT<T> &p, const Options &o,
                                         T gamma, const G g[3][3])
{
    p.rho = std_max(p.rho, T(o.rho_floor));
    p.pressure = std_max(p.pressure, T(o.pressure_floor));
    const T v2 = vsq(p.v, g);
    if (v2 >= T(o.max_v2))
    {
        const T scale = qsqrt(qdiv(T(o.max_v2), v2));
#pragma unroll
        for (int i = 0; i < 3; ++i)
            p.v[i] *= scale;
    }
    p.eps = qdiv(p.pressure, (gamma - T(1)) * p.rho);
}

template <class T>
__device__ __forceinline__ void floors(PrimT<T> &p, const Options &o,
                                       const T g[3][3])
{
    floors_g(p, o, T(o.gamma), g);
}

__device__ __forceinline__ void atmosphere_prim(const Options &o, Prim &p)
{
    p.rho = o.rho_floor;
    p.pressure = o.pressure_floor;
    p.eps = qdiv(p.pressure, (o.gamma - 1.0) * p.rho);
    p.v[0] = p.v[1] = p.v[2] = 0.0;
}

/* Mixed precision: whether the faces beside a cell are computed in FP64.
 * - guard_margin m > 0: the recovered primitives lie within a factor m of
 *   a floor or of the velocity cap, where the floor and reset decisions of
 *   the next recoveries turn on rounding: rho <= m' rho_floor or
 *   P <= m' P_floor with m' = max(m, the metric-source atmosphere factor),
 *   or 1 - v^2 <= m (1 - max_v2);
 * - guard_shift s > 0: the shift is at least s in size,
 *   gamma_ij beta^i beta^j >= s^2.  In horizon-penetrating coordinates
 *   infalling gas moves with the shift, so the flux speed alpha v^i -
 *   beta^i is a small difference of large terms there, which FP32 loses. */
__device__ __forceinline__ bool fp64_cell(const Prim &w, const Options &o,
                                          const Geo &geo)
{
    bool near = false;
    const double m = o.guard_margin;
    if (m > 0.0)
    {
        const double mf = std_max(m, o.source_atmosphere_factor);
        near = w.rho <= mf * o.rho_floor ||
               w.pressure <= mf * o.pressure_floor ||
               1.0 - vsq(w.v, geo.g) <= m * (1.0 - o.max_v2);
    }
    const double s = o.guard_shift;
    if (s > 0.0 && vsq(geo.shift, geo.g) >= s * s)
        near = true;
    return near;
}

/* FArrayBoxFiniteVolume.hpp:118-126 */
__device__ __forceinline__ bool is_atmosphere(const Prim &p, const Options &o,
                                              double threshold_factor)
{
    const double factor = std_max(1.0, threshold_factor);
    return p.rho <= factor * o.rho_floor &&
           p.pressure <= factor * o.pressure_floor;
}

/* FArrayBoxFiniteVolume.hpp:704-711 */
__device__ __forceinline__ bool prim_differs(const Prim &a, const Prim &b)
{
    return a.rho != b.rho || a.eps != b.eps || a.pressure != b.pressure ||
           a.v[0] != b.v[0] || a.v[1] != b.v[1] || a.v[2] != b.v[2];
}

/* ---------------- Reconstruction (Reconstruction.hpp:16-69) ---------------- */
template <class T> __device__ __forceinline__ T minmod2(T l, T r)
{
    if (l * r <= T(0))
        return T(0);
    return l > T(0) ? std_min(l, r) : std_max(l, r);
}

template <class T>
__device__ __forceinline__ T limited_slope(T l, T c, T r, T theta)
{
    return minmod2(theta * (c - l), minmod2(T(0.5) * (r - l), theta * (r - c)));
}

/* ---------------- Riemann solver (RiemannSolvers.hpp:23-118) ---------------- */
template <class T, class G = T>
__device__ __forceinline__ void char_speeds(const PrimT<T> &p, T gamma,
                                            const GeoT<G> &g, int d, T &lower,
                                            T &upper)
{
    const T v2 = vsq(p.v, g.g);
    const T h = T(1) + p.eps + qdiv(p.pressure, p.rho);
    const T cs2 = qdiv(gamma * p.pressure, p.rho * h);
    const T cs = qsqrt(cs2);
    const T vel = p.v[d];
    const T den = T(1) - v2 * cs2;
    const T rad = (T(1) - v2) *
                  (g.gu[d][d] * (T(1) - v2 * cs2) - vel * vel * (T(1) - cs2));
    const T root = cs * qsqrt(std_max(T(0), rad));
    const T lo = qdiv(vel * (T(1) - cs2) - root, den);
    const T hi = qdiv(vel * (T(1) - cs2) + root, den);
    lower = g.lapse * lo - g.shift[d];
    upper = g.lapse * hi - g.shift[d];
}

/* compute_muscl_hlle_flux (FiniteVolume.hpp:63-98) after the MUSCL step:
 * floors, prim2con and HLLE on the face geometry. */
template <class T, class G = T>
__device__ __forceinline__ void hlle_face_g(PrimT<T> lp, PrimT<T> rp,
                                            const Options &o, T gamma,
                                            const GeoT<G> &face, int d,
                                            ConsT<T> &f)
{
    floors_g(lp, o, gamma, face.g);
    floors_g(rp, o, gamma, face.g);
    ConsT<T> lc, rc;
    prim2con(lp, gamma, face.g, lc);
    prim2con(rp, gamma, face.g, rc);
    ConsT<T> fl, fr;
    flux_of(lp, lc, d, face.lapse, face.shift, fl);
    flux_of(rp, rc, d, face.lapse, face.shift, fr);
    T ll, lu, rl, ru;
    char_speeds(lp, gamma, face, d, ll, lu);
    char_speeds(rp, gamma, face, d, rl, ru);
    const T sl = std_min(T(0), std_min(ll, rl));
    const T su = std_max(T(0), std_max(lu, ru));
    /* The left flux if s_lower >= 0, else the right flux if s_upper <= 0,
     * else the HLLE combination.  The choice is made per component with
     * selects rather than three exits: with exits the compiler merges the
     * three stores of the momentum flux into one at a run-time offset and
     * keeps that array in scratch memory.  The combination is evaluated on
     * every lane; where it is not chosen (s_lower = 0 or s_upper = 0) its
     * value, possibly 0 * inf, is discarded. */
    const bool left = sl >= T(0);
    const bool right = su <= T(0);
    const T inv = qdiv(T(1), su - sl);
    const T hD = (su * fl.D - sl * fr.D + sl * su * (rc.D - lc.D)) * inv;
    const T hT =
        (su * fl.tau - sl * fr.tau + sl * su * (rc.tau - lc.tau)) * inv;
    f.D = left ? fl.D : (right ? fr.D : hD);
    f.tau = left ? fl.tau : (right ? fr.tau : hT);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const T hS =
            (su * fl.S[i] - sl * fr.S[i] + sl * su * (rc.S[i] - lc.S[i])) *
            inv;
        f.S[i] = left ? fl.S[i] : (right ? fr.S[i] : hS);
    }
}

template <class T>
__device__ __forceinline__ void hlle_face(const PrimT<T> &lp,
                                          const PrimT<T> &rp,
                                          const Options &o,
                                          const GeoT<T> &face, int d,
                                          ConsT<T> &f)
{
    hlle_face_g(lp, rp, o, T(o.gamma), face, d, f);
}

/* average_geometry (FArrayBoxFiniteVolume.hpp:194-216), K omitted */
template <class T>
__device__ __forceinline__ void average_geo(const GeoT<T> &l, const GeoT<T> &r,
                                            GeoT<T> &out)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            out.g[i][j] = T(0.5) * (l.g[i][j] + r.g[i][j]);
            out.gu[i][j] = T(0.5) * (l.gu[i][j] + r.gu[i][j]);
        }
    out.lapse = T(0.5) * (l.lapse + r.lapse);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        out.shift[i] = T(0.5) * (l.shift[i] + r.shift[i]);
}

/* ---------------- Static-metric sources (StaticMetric.hpp:193-275) ------- */
template <class T, class G = T>
__device__ __forceinline__ void static_sources(const PrimT<T> &p,
                                               const ConsT<T> &c, T gamma,
                                               const FullGeoT<G> &f,
                                               ConsT<T> &src)
{
    (void)gamma; /* the EOS argument of the reference's signature */
    const GeoT<G> &g = f.geo;
    const T W = qdiv(T(1.0), qsqrt(T(1.0) - vsq(p.v, g.g)));
    const T h = T(1.0) + p.eps + qdiv(p.pressure, p.rho);
    const T rhW2 = p.rho * h * W * W;
    T Tuu[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Tuu[i][j] = rhW2 * p.v[i] * p.v[j] + p.pressure * g.gu[i][j];
    const T E = c.D + c.tau;
    T mU[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        mU[i] = T(0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            mU[i] += g.gu[i][j] * c.S[j];
    ConsT<T> cf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        flux_of(p, c, d, g.lapse, g.shift, cf[d]);
    src.D = T(0.0);
    src.tau = T(0.0);
    src.S[0] = src.S[1] = src.S[2] = T(0.0);
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        src.D -= cf[d].D * f.dlogsqrtdet[d];
        src.tau -= cf[d].tau * f.dlogsqrtdet[d];
        src.tau -= mU[d] * f.dlapse[d];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            src.tau += g.lapse * Tuu[i][j] * f.K[i][j];
#pragma unroll
    for (int m = 0; m < 3; ++m)
    {
        T ms = -E * f.dlapse[m];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                ms += T(0.5) * g.lapse * Tuu[i][j] * f.dg[m][i][j];
#pragma unroll
        for (int s = 0; s < 3; ++s)
            ms += c.S[s] * f.dshift[s][m];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            ms -= cf[d].S[m] * f.dlogsqrtdet[d];
        src.S[m] += ms;
    }
}

} // namespace grhd

#endif /* GRHD_PHYSICS_HPP_ */
