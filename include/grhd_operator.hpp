/*
 * grhd_operator.hpp -- C++ host mirror of the reference's level operators
 * over the C-ABI of grhd_abi.h.
 *
 * FixedBGLevelDataFiniteVolumeOperator and
 * TimeDependentFixedBGLevelDataFiniteVolumeOperator
 * (Source/Matter/GRHD/FixedBGFArrayBoxFiniteVolume.hpp:446-614, :677-830)
 * keep their names, constructor arguments, method names, argument meaning
 * and exceptions: std::runtime_error with the reference's messages for
 * primitive-recovery failures, std::invalid_argument for a non-positive
 * CFL number or bad arguments.  A LevelData<FArrayBox> becomes a
 * DeviceLevelData: the DisjointBoxLayout (interior boxes), the ghost width,
 * the component count, and one HBM pool holding each box's FArrayBox
 * (component-major, i fastest) back to back.  Every method is one launch
 * per kernel for the whole level; the pools stay on the device.
 *
 * The overloads taking inverse_dx_U accept any indexable type (GRChombo's
 * Tensor<1, double>, std::array<double, 3>, double[3]); those without it
 * use 1/dx in every direction, as the reference's make_inverse_dx does.
 * A Box argument is an IntBox; a ProblemDomain built from it is
 * non-periodic.
 *
 * FixedBGAMRLevelHooks and TimeDependentFixedBGAMRLevelHooks
 * (FixedBGAMRLevelHooks.hpp:22-312) and the driver views
 * FixedBGLevelDataDriverView / TimeDependentFixedBGLevelDataDriverView
 * (FixedBGLevelDataDriver.hpp:17-141, :229-395) delegate to the operators
 * exactly as the reference's do.
 *
 * Header-only; link against libgrhd_hip.so.
 */
#ifndef GRHD_OPERATOR_HPP_
#define GRHD_OPERATOR_HPP_

#include "grhd_abi.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace GRHD
{

/* Atmosphere.hpp:16-22 */
struct AtmosphereOptions
{
    double rho_floor = 1.0e-12;
    double pressure_floor = 1.0e-12;
    double max_velocity_squared = 1.0 - 1.0e-12;
    double metric_source_atmosphere_factor = 1.0 + 1.0e-12;
};

/* Valencia.hpp:17-23 */
struct RecoveryOptions
{
    int max_iterations = 50;
    double tolerance = 1.0e-12;
    double pressure_floor = 1.0e-12;
    double max_velocity_squared = 1.0 - 1.0e-12;
};

/* IdealGasEOS.hpp:10-53 (the operator needs the adiabatic index only) */
struct IdealGasEOS
{
    double adiabatic_index;
    explicit IdealGasEOS(double gamma) : adiabatic_index(gamma) {}
};

/* FArrayBoxFiniteVolume.hpp:78-92 */
struct AtmosphereResetDiagnostics
{
    long long num_cells = 0;
    long long num_failed_recoveries = 0;
    long long num_floored_primitives = 0;
    long long num_conserved_resets = 0;

    void add(const AtmosphereResetDiagnostics &o)
    {
        num_cells += o.num_cells;
        num_failed_recoveries += o.num_failed_recoveries;
        num_floored_primitives += o.num_floored_primitives;
        num_conserved_resets += o.num_conserved_resets;
    }
};

/* Conserved<double> of a level sum (GRHDVars.hpp:15-37) */
struct ConservedSum
{
    double D = 0.0;
    std::array<double, 3> S_L{{0.0, 0.0, 0.0}};
    double tau = 0.0;
};

namespace amd
{

/* KerrSchild::params_t (KerrSchild.hpp:30-36) */
inline grhd_background kerr_schild(double mass, double spin)
{
    grhd_background b{};
    b.kind = GRHD_BG_KERR_SCHILD;
    b.mass = mass;
    b.spin = spin;
    return b;
}

/* CircularBinaryPN::params_t (CircularBinaryPN.hpp:25-45) */
inline grhd_background circular_binary_pn(double mass_1, double mass_2,
                                          double separation,
                                          double angular_velocity = 0.0,
                                          double phase = 0.0,
                                          double time = 0.0,
                                          double softening_radius = 0.0)
{
    grhd_background b{};
    b.kind = GRHD_BG_CIRCULAR_BINARY_PN;
    b.mass_1 = mass_1;
    b.mass_2 = mass_2;
    b.separation = separation;
    b.angular_velocity = angular_velocity;
    b.phase = phase;
    b.time = time;
    b.softening_radius = softening_radius;
    return b;
}

/* A Box: inclusive integer corners. */
struct IntBox
{
    std::array<int, 3> lo{{0, 0, 0}}, hi{{-1, -1, -1}};

    long long num_pts() const
    {
        long long n = 1;
        for (int d = 0; d < 3; ++d)
            n *= (long long)(hi[d] - lo[d] + 1);
        return n;
    }
    IntBox grown(int g) const
    {
        IntBox b = *this;
        for (int d = 0; d < 3; ++d)
        {
            b.lo[d] -= g;
            b.hi[d] += g;
        }
        return b;
    }
};

/* ProblemDomain: the domain box and its periodic directions. */
struct ProblemDomain
{
    IntBox box;
    std::array<bool, 3> periodic{{false, false, false}};
    explicit ProblemDomain(const IntBox &b) : box(b) {}
    ProblemDomain(const IntBox &b, const std::array<bool, 3> &p)
        : box(b), periodic(p)
    {
    }
    const IntBox &domainBox() const { return box; }
};

/* inverse_dx_U: three per-direction inverse spacings; all zero stands
 * for 1/dx in every direction. */
using InverseDx = std::array<double, 3>;

/* Anything indexable by 0..2 into doubles (Tensor<1, double>, arrays). */
template <class V, class = void> struct is_inverse_dx : std::false_type
{
};
template <class V>
struct is_inverse_dx<
    V, typename std::enable_if<
           !std::is_arithmetic<V>::value &&
           std::is_convertible<decltype(std::declval<const V &>()[0]),
                               double>::value>::type> : std::true_type
{
};

template <class V> InverseDx to_inverse_dx(const V &v)
{
    return InverseDx{{double(v[0]), double(v[1]), double(v[2])}};
}

/* The bounding box of a box list (the domain of a layout given none). */
inline IntBox bounding_box(const std::vector<IntBox> &grids)
{
    IntBox b = grids.empty() ? IntBox() : grids[0];
    for (const IntBox &g : grids)
        for (int d = 0; d < 3; ++d)
        {
            b.lo[d] = std::min(b.lo[d], g.lo[d]);
            b.hi[d] = std::max(b.hi[d], g.hi[d]);
        }
    return b;
}

/* LevelData<FArrayBox> in HBM.  `pool` is a device pointer to
 * pool_size() doubles (the caller owns it).  The layout's ProblemDomain
 * (DisjointBoxLayout::physDomain) is given, or the grids' bounding box,
 * non-periodic.  With the domain given, the grids may be empty: a rank
 * that owns none of the level's boxes, on which every operation does
 * nothing. */
class DeviceLevelData
{
  public:
    DeviceLevelData(std::vector<IntBox> grids, int ghosts, int ncomp,
                    double *pool)
        : DeviceLevelData(grids, ProblemDomain(bounding_box(grids)), ghosts,
                          ncomp, pool)
    {
        if (m_grids.empty())
            throw std::invalid_argument(
                "DeviceLevelData: no grids and no ProblemDomain");
    }

    DeviceLevelData(std::vector<IntBox> grids, const ProblemDomain &domain,
                    int ghosts, int ncomp, double *pool)
        : m_grids(std::move(grids)), m_domain(domain), m_ghosts(ghosts),
          m_ncomp(ncomp), m_pool(pool)
    {
        if (ghosts < 0 || ncomp <= 0)
            throw std::invalid_argument("DeviceLevelData: bad layout");
        long long off = 0;
        for (const IntBox &b : m_grids)
        {
            m_offsets.push_back(off);
            off += (long long)ncomp * b.grown(ghosts).num_pts();
        }
        m_size = off;
    }

    const std::vector<IntBox> &grids() const { return m_grids; }
    const ProblemDomain &problem_domain() const { return m_domain; }
    int ghosts() const { return m_ghosts; }
    int ncomp() const { return m_ncomp; }
    double *data() const { return m_pool; }
    long long pool_size() const { return m_size; }
    long long fab_offset(size_t b) const { return m_offsets[b]; }
    long long num_interior() const
    {
        long long n = 0;
        for (const IntBox &b : m_grids)
            n += b.num_pts();
        return n;
    }

    /* The C-ABI's box table for this layout. */
    std::vector<grhd_level_box> level_boxes(
        const ProblemDomain &domain, double dx,
        const std::array<double, 3> &c,
        const InverseDx &inverse_dx = InverseDx{{0.0, 0.0, 0.0}}) const
    {
        std::vector<grhd_level_box> out(m_grids.size());
        for (size_t n = 0; n < m_grids.size(); ++n)
        {
            grhd_box &b = out[n].box;
            const IntBox g = m_grids[n].grown(m_ghosts);
            for (int d = 0; d < 3; ++d)
            {
                b.lo[d] = m_grids[n].lo[d];
                b.hi[d] = m_grids[n].hi[d];
                b.glo[d] = g.lo[d];
                b.ghi[d] = g.hi[d];
                b.dlo[d] = domain.box.lo[d];
                b.dhi[d] = domain.box.hi[d];
                b.periodic[d] = domain.periodic[d] ? 1 : 0;
                b.center[d] = c[d];
                b.inverse_dx[d] = inverse_dx[d];
            }
            b.dx = dx;
            out[n].state_offset = m_offsets[n];
        }
        return out;
    }

  private:
    std::vector<IntBox> m_grids;
    ProblemDomain m_domain;
    int m_ghosts;
    int m_ncomp;
    double *m_pool;
    std::vector<long long> m_offsets;
    long long m_size = 0;
};

/* Owns one grhd_ctx (device scratch, status block) on the current device. */
class Context
{
  public:
    Context()
    {
        if (grhd_ctx_create(&m_ctx) != GRHD_OK)
            throw std::runtime_error("grhd_ctx_create failed");
    }
    ~Context() { grhd_ctx_destroy(m_ctx); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    grhd_ctx *get() const { return m_ctx; }
    /* time the face launches on each new level and keep the fastest
     * (results are the same either way) */
    void set_autotune(bool on) { (void)grhd_ctx_set_autotune(m_ctx, on); }
    /* pin the face launch of one precision: walk 1, 2, 4 or 8, 0 for the
     * tiles, -1 to unpin (results are the same either way) */
    void set_face_launch(bool mixed, int walk, bool fused)
    {
        if (grhd_ctx_set_face_launch(m_ctx, mixed ? 1 : 0, walk,
                                     fused ? 1 : 0) != GRHD_OK)
            throw std::invalid_argument("GRHD set_face_launch: walk must "
                                        "be -1, 0, 1, 2, 4 or 8");
    }

  private:
    grhd_ctx *m_ctx = nullptr;
};

/* Status code -> the exception the reference throws. */
inline void check(int rc, const grhd_error *err, const char *what)
{
    if (rc == GRHD_OK)
        return;
    if (rc == GRHD_ERR_INVALID_ARGUMENT)
        throw std::invalid_argument(std::string("GRHD ") + what +
                                    ": invalid argument");
    char iv[64] = "?";
    double res = 0.0;
    if (err)
    {
        std::snprintf(iv, sizeof(iv), "(%d,%d,%d)", err->iv[0], err->iv[1],
                      err->iv[2]);
        res = err->residual;
    }
    char msg[256];
    switch (rc)
    {
    case GRHD_ERR_RECOVERY_GHOST_FILL:
        std::snprintf(msg, sizeof(msg),
                      "GRHD primitive recovery failed at outflow source %s "
                      "while filling fixed-background boundary, residual %g",
                      iv, res);
        break;
    case GRHD_ERR_RECOVERY_RHS:
        std::snprintf(msg, sizeof(msg),
                      "GRHD primitive recovery failed at %s with residual %g "
                      "on fixed background",
                      iv, res);
        break;
    case GRHD_ERR_RECOVERY_SOURCES:
        std::snprintf(msg, sizeof(msg),
                      "GRHD primitive recovery failed at %s while computing "
                      "fixed-background metric sources, residual %g",
                      iv, res);
        break;
    case GRHD_ERR_RECOVERY_CFL:
        std::snprintf(msg, sizeof(msg),
                      "GRHD primitive recovery failed at %s while computing "
                      "fixed-background CFL speed, residual %g",
                      iv, res);
        break;
    default:
        std::snprintf(msg, sizeof(msg), "GRHD %s: HIP device error (code %d)",
                      what, rc);
    }
    throw std::runtime_error(msg);
}

inline AtmosphereResetDiagnostics to_diag(const grhd_diagnostics &d)
{
    AtmosphereResetDiagnostics o;
    o.num_cells = d.num_cells;
    o.num_failed_recoveries = d.num_failed_recoveries;
    o.num_floored_primitives = d.num_floored_primitives;
    o.num_conserved_resets = d.num_conserved_resets;
    return o;
}

/* Shared state of both operators. */
class LevelOperatorBase
{
  public:
    LevelOperatorBase(const IdealGasEOS &eos, const grhd_background &bg,
                      double dx, const std::array<double, 3> &center,
                      const AtmosphereOptions &atm, double limiter_theta,
                      bool use_reconstruction, const RecoveryOptions &rec,
                      bool use_static_metric_sources,
                      bool use_metric_volume_time_sources,
                      double metric_time_derivative_step, int comp_base,
                      std::shared_ptr<Context> ctx, void *stream)
        : m_background(bg), m_dx(dx), m_center(center),
          m_comp_base(comp_base),
          m_ctx(ctx ? std::move(ctx) : std::make_shared<Context>()),
          m_stream(stream)
    {
        grhd_params &p = m_params;
        p.adiabatic_index = eos.adiabatic_index;
        p.rho_floor = atm.rho_floor;
        p.pressure_floor = atm.pressure_floor;
        p.max_velocity_squared = atm.max_velocity_squared;
        p.metric_source_atmosphere_factor =
            atm.metric_source_atmosphere_factor;
        p.max_iterations = rec.max_iterations;
        p.tolerance = rec.tolerance;
        p.recovery_pressure_floor = rec.pressure_floor;
        p.recovery_max_velocity_squared = rec.max_velocity_squared;
        p.limiter_theta = limiter_theta;
        p.use_reconstruction = use_reconstruction ? 1 : 0;
        p.use_static_metric_sources = use_static_metric_sources ? 1 : 0;
        p.use_metric_volume_time_sources =
            use_metric_volume_time_sources ? 1 : 0;
        p.metric_time_derivative_step = metric_time_derivative_step;
        p.excision_inner_buffer = 1.05;
        p.excision_outer_buffer = 0.9;
    }

    const grhd_background &background() const { return m_background; }
    grhd_ctx *context() const { return m_ctx->get(); }
    void *stream() const { return m_stream; }
    int comp_base() const { return m_comp_base; }
    double dx() const { return m_dx; }
    const std::array<double, 3> &center() const { return m_center; }

    /* FP32 reconstruction, face metric and HLLE (see grhd_params); off by
     * default, which is the bit-exact FP64 path. */
    void set_mixed_precision(bool on) { m_params.mixed_precision = on ? 1 : 0; }

    /* With mixed precision: FP64 faces beside cells within a factor
     * `margin` of a floor or the velocity cap (see grhd_params); 0 keeps
     * every face FP32. */
    void set_mixed_fp64_margin(double margin)
    {
        m_params.mixed_fp64_margin = margin;
    }

    /* With mixed precision: FP64 faces beside cells whose shift is at
     * least `shift` in size (see grhd_params); 0 disables the test. */
    void set_mixed_fp64_shift(double shift)
    {
        m_params.mixed_fp64_shift = shift;
    }

    /* KerrSchild::check_if_excised with ExcisionTest's zero right-hand side
     * (see grhd_params); Kerr-Schild backgrounds only. */
    void set_horizon_excision(bool on, double inner_buffer = 1.05,
                              double outer_buffer = 0.9)
    {
        m_params.excise_horizon = on ? 1 : 0;
        m_params.excision_inner_buffer = inner_buffer;
        m_params.excision_outer_buffer = outer_buffer;
    }

  protected:
    static constexpr InverseDx uniform() { return InverseDx{{0.0, 0.0, 0.0}}; }

    std::vector<grhd_level_box> table(const DeviceLevelData &l,
                                      const ProblemDomain &domain,
                                      const InverseDx &inv) const
    {
        return l.level_boxes(domain, m_dx, m_center, inv);
    }

    void rhs_at(DeviceLevelData &rhs, DeviceLevelData &state,
                const ProblemDomain &domain, const InverseDx &inv,
                double time, bool accumulate = false) const
    {
        const auto t = table(state, domain, inv);
        if (rhs.ncomp() != state.ncomp() ||
            rhs.pool_size() != rhs.ncomp() * state.num_interior())
            throw std::invalid_argument(
                "rhs must hold ncomp components over the interior boxes");
        grhd_error err{};
        if (accumulate)
            check(grhd_level_add_conserved_rhs(
                      context(), t.data(), (int)t.size(), &m_params,
                      &m_background, time, state.data(), state.ncomp(),
                      m_comp_base, rhs.data(), &err, m_stream),
                  &err, "add_conserved_rhs_to");
        else
            check(grhd_level_compute_rhs(context(), t.data(), (int)t.size(),
                                         &m_params, &m_background, time,
                                         state.data(), state.ncomp(),
                                         m_comp_base, rhs.data(), nullptr,
                                         &err, m_stream),
                  &err, "compute_rhs");
    }

    AtmosphereResetDiagnostics recover_at(DeviceLevelData &state,
                                          const ProblemDomain &domain,
                                          double time) const
    {
        const auto t = table(state, domain, uniform());
        grhd_diagnostics d{};
        check(grhd_level_recover_primitives(
                  context(), t.data(), (int)t.size(), &m_params,
                  &m_background, time, state.data(), state.ncomp(),
                  m_comp_base, nullptr, &d, m_stream),
              nullptr, "recover_primitives");
        return to_diag(d);
    }

    void update_at(DeviceLevelData &state, const DeviceLevelData &rhs,
                   const ProblemDomain &domain, double dt, bool recover,
                   double recovery_time) const
    {
        const auto t = table(state, domain, uniform());
        grhd_diagnostics d{};
        check(grhd_level_update_conserved(
                  context(), t.data(), (int)t.size(), &m_params,
                  &m_background, state.data(), state.ncomp(), m_comp_base,
                  rhs.data(), dt, recover ? 1 : 0, recovery_time, nullptr, &d,
                  m_stream),
              nullptr, "update_conserved");
    }

    /* The domain box does not enter the CFL estimate (the reference's level
     * version ignores it, FixedBGFArrayBoxFiniteVolume.hpp:423); the
     * state's own ProblemDomain keys the launch plan. */
    double max_inverse_dt_at(const DeviceLevelData &state,
                             const InverseDx &inv, double time) const
    {
        const auto t = table(state, state.problem_domain(), inv);
        double out = 0.0;
        grhd_error err{};
        check(grhd_level_compute_max_inverse_dt(
                  context(), t.data(), (int)t.size(), &m_params,
                  &m_background, time, state.data(), state.ncomp(),
                  m_comp_base, &out, &err, m_stream),
              &err, "compute_max_inverse_dt");
        return out;
    }

    /* compute_stable_dt (:544-555, :769-780) */
    double stable_dt_at(const DeviceLevelData &state, const InverseDx &inv,
                        double cfl, double time) const
    {
        if (cfl <= 0.0)
            throw std::invalid_argument("GRHD CFL number must be positive");
        return cfl / std::max(max_inverse_dt_at(state, inv, time), 1.0e-14);
    }

    void ssprk2_at(DeviceLevelData &state, const ProblemDomain &domain,
                   const InverseDx &inv, double dt, double time) const
    {
        const auto t = table(state, domain, inv);
        grhd_diagnostics d{};
        grhd_error err{};
        check(grhd_level_advance_ssprk2(context(), t.data(), (int)t.size(),
                                        &m_params, &m_background, time, dt,
                                        state.data(), state.ncomp(),
                                        m_comp_base, nullptr, &d, &err,
                                        m_stream),
              &err, "advance_ssprk2");
    }

    AtmosphereResetDiagnostics rk4_at(DeviceLevelData &state,
                                      const ProblemDomain &domain,
                                      const InverseDx &inv, double dt,
                                      int nsteps, double time) const
    {
        const auto t = table(state, domain, inv);
        grhd_diagnostics d{};
        grhd_error err{};
        check(grhd_level_advance_rk4(context(), t.data(), (int)t.size(),
                                     &m_params, &m_background, time, dt,
                                     nsteps, state.data(), state.ncomp(),
                                     m_comp_base, nullptr, &d, &err,
                                     m_stream),
              &err, "advance_rk4");
        return to_diag(d);
    }

    grhd_params m_params{};
    grhd_background m_background;
    double m_dx;
    std::array<double, 3> m_center;
    int m_comp_base;
    std::shared_ptr<Context> m_ctx;
    void *m_stream;
};

#define GRHD_IF_INVERSE_DX(V)                                                 \
    template <class V,                                                        \
              typename std::enable_if<is_inverse_dx<V>::value, int>::type = 0>

/* FixedBGLevelDataFiniteVolumeOperator
 * (FixedBGFArrayBoxFiniteVolume.hpp:446-614): a static background. */
class FixedBGLevelDataFiniteVolumeOperator : public LevelOperatorBase
{
  public:
    FixedBGLevelDataFiniteVolumeOperator(
        const IdealGasEOS &a_eos, const grhd_background &a_background,
        double a_dx, const std::array<double, 3> &a_center,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : LevelOperatorBase(a_eos, a_background, a_dx, a_center, a_atmosphere,
                            a_limiter_theta, a_use_reconstruction,
                            a_recovery_options, a_use_static_metric_sources,
                            false, 1.0e-3, a_comp_base, std::move(a_context),
                            a_stream)
    {
    }

    /* :468-477 */
    GRHD_IF_INVERSE_DX(V)
    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const ProblemDomain &domain, const V &inverse_dx_U) const
    {
        rhs_at(rhs, state, domain, to_inverse_dx(inverse_dx_U), 0.0);
    }

    /* :479-484 */
    GRHD_IF_INVERSE_DX(V)
    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const IntBox &domain_box, const V &inverse_dx_U) const
    {
        compute_rhs(rhs, state, ProblemDomain(domain_box), inverse_dx_U);
    }

    /* :486-490 */
    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const ProblemDomain &domain) const
    {
        rhs_at(rhs, state, domain, uniform(), 0.0);
    }

    /* :492-503: rhs[D, S_i, tau] += the hydro rhs */
    GRHD_IF_INVERSE_DX(V)
    void add_conserved_rhs_to(DeviceLevelData &rhs, DeviceLevelData &state,
                              const ProblemDomain &domain,
                              const V &inverse_dx_U) const
    {
        rhs_at(rhs, state, domain, to_inverse_dx(inverse_dx_U), 0.0, true);
    }

    /* :506-520 */
    AtmosphereResetDiagnostics recover_primitives(DeviceLevelData &state) const
    {
        return recover_at(state, state.problem_domain(), 0.0);
    }

    AtmosphereResetDiagnostics recover_primitives(DeviceLevelData &state,
                                                  const ProblemDomain &domain)
        const
    {
        return recover_at(state, domain, 0.0);
    }

    /* :522-533 */
    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          double dt,
                          bool recover_primitives_after_update = true) const
    {
        update_at(state, rhs, state.problem_domain(), dt,
                  recover_primitives_after_update, 0.0);
    }

    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          const ProblemDomain &domain, double dt,
                          bool recover_primitives_after_update = true) const
    {
        update_at(state, rhs, domain, dt, recover_primitives_after_update,
                  0.0);
    }

    /* :535-542 */
    GRHD_IF_INVERSE_DX(V)
    double compute_max_inverse_dt(const DeviceLevelData &state,
                                  const IntBox &domain_box,
                                  const V &inverse_dx_U) const
    {
        (void)domain_box;
        return max_inverse_dt_at(state, to_inverse_dx(inverse_dx_U), 0.0);
    }

    double compute_max_inverse_dt(const DeviceLevelData &state,
                                  const ProblemDomain &domain) const
    {
        (void)domain;
        return max_inverse_dt_at(state, uniform(), 0.0);
    }

    /* :544-555 */
    GRHD_IF_INVERSE_DX(V)
    double compute_stable_dt(const DeviceLevelData &state,
                             const IntBox &domain_box, const V &inverse_dx_U,
                             double cfl) const
    {
        (void)domain_box;
        return stable_dt_at(state, to_inverse_dx(inverse_dx_U), cfl, 0.0);
    }

    /* :557-561 */
    double compute_stable_dt(const DeviceLevelData &state,
                             const IntBox &domain_box, double cfl) const
    {
        (void)domain_box;
        return stable_dt_at(state, uniform(), cfl, 0.0);
    }

    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain, double cfl) const
    {
        (void)domain;
        return stable_dt_at(state, uniform(), cfl, 0.0);
    }

    /* :563-589 */
    GRHD_IF_INVERSE_DX(V)
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        const V &inverse_dx_U, double dt) const
    {
        ssprk2_at(state, domain, to_inverse_dx(inverse_dx_U), dt, 0.0);
    }

    /* :591-595 */
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        double dt) const
    {
        ssprk2_at(state, domain, uniform(), dt, 0.0);
    }

    /* Classical RK4, `nsteps` steps (no counterpart in the reference; see
     * grhd_advance_rk4_box). */
    AtmosphereResetDiagnostics advance_rk4(DeviceLevelData &state,
                                           const ProblemDomain &domain,
                                           double dt, int nsteps = 1) const
    {
        return rk4_at(state, domain, uniform(), dt, nsteps, 0.0);
    }

    GRHD_IF_INVERSE_DX(V)
    AtmosphereResetDiagnostics advance_rk4(DeviceLevelData &state,
                                           const ProblemDomain &domain,
                                           const V &inverse_dx_U, double dt,
                                           int nsteps = 1) const
    {
        return rk4_at(state, domain, to_inverse_dx(inverse_dx_U), dt, nsteps,
                      0.0);
    }
};

/* TimeDependentFixedBGLevelDataFiniteVolumeOperator
 * (FixedBGFArrayBoxFiniteVolume.hpp:677-830): the background at `time`
 * is the binary's parameters with time += `time` (its factory), and the
 * metric volume-time sources may be switched on. */
class TimeDependentFixedBGLevelDataFiniteVolumeOperator
    : public LevelOperatorBase
{
  public:
    TimeDependentFixedBGLevelDataFiniteVolumeOperator(
        const IdealGasEOS &a_eos, const grhd_background &a_background,
        double a_dx, const std::array<double, 3> &a_center,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true,
        bool a_use_metric_volume_time_sources = true,
        double a_metric_time_derivative_step = 1.0e-3, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : LevelOperatorBase(a_eos, a_background, a_dx, a_center, a_atmosphere,
                            a_limiter_theta, a_use_reconstruction,
                            a_recovery_options, a_use_static_metric_sources,
                            a_use_metric_volume_time_sources,
                            a_metric_time_derivative_step, a_comp_base,
                            std::move(a_context), a_stream)
    {
    }

    /* :702-720 */
    GRHD_IF_INVERSE_DX(V)
    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const ProblemDomain &domain, const V &inverse_dx_U,
                     double time) const
    {
        rhs_at(rhs, state, domain, to_inverse_dx(inverse_dx_U), time);
    }

    /* :722-726 */
    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const ProblemDomain &domain, double time) const
    {
        rhs_at(rhs, state, domain, uniform(), time);
    }

    /* :728-743 */
    AtmosphereResetDiagnostics recover_primitives(DeviceLevelData &state,
                                                  double time) const
    {
        return recover_at(state, state.problem_domain(), time);
    }

    AtmosphereResetDiagnostics recover_primitives(DeviceLevelData &state,
                                                  const ProblemDomain &domain,
                                                  double time) const
    {
        return recover_at(state, domain, time);
    }

    /* :745-757 */
    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          double dt, double recovery_time,
                          bool recover_primitives_after_update = true) const
    {
        update_at(state, rhs, state.problem_domain(), dt,
                  recover_primitives_after_update, recovery_time);
    }

    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          const ProblemDomain &domain, double dt,
                          double recovery_time,
                          bool recover_primitives_after_update = true) const
    {
        update_at(state, rhs, domain, dt, recover_primitives_after_update,
                  recovery_time);
    }

    /* :759-767 */
    GRHD_IF_INVERSE_DX(V)
    double compute_max_inverse_dt(const DeviceLevelData &state,
                                  const IntBox &domain_box,
                                  const V &inverse_dx_U, double time) const
    {
        (void)domain_box;
        return max_inverse_dt_at(state, to_inverse_dx(inverse_dx_U), time);
    }

    double compute_max_inverse_dt(const DeviceLevelData &state,
                                  const ProblemDomain &domain,
                                  double time) const
    {
        (void)domain;
        return max_inverse_dt_at(state, uniform(), time);
    }

    /* :769-780 */
    GRHD_IF_INVERSE_DX(V)
    double compute_stable_dt(const DeviceLevelData &state,
                             const IntBox &domain_box, const V &inverse_dx_U,
                             double cfl, double time) const
    {
        (void)domain_box;
        return stable_dt_at(state, to_inverse_dx(inverse_dx_U), cfl, time);
    }

    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain, double cfl,
                             double time) const
    {
        (void)domain;
        return stable_dt_at(state, uniform(), cfl, time);
    }

    /* :782-809 */
    GRHD_IF_INVERSE_DX(V)
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        const V &inverse_dx_U, double time, double dt) const
    {
        ssprk2_at(state, domain, to_inverse_dx(inverse_dx_U), dt, time);
    }

    /* the same with 1/dx in every direction */
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        double time, double dt) const
    {
        ssprk2_at(state, domain, uniform(), dt, time);
    }

    AtmosphereResetDiagnostics advance_rk4(DeviceLevelData &state,
                                           const ProblemDomain &domain,
                                           double dt, int nsteps,
                                           double time) const
    {
        return rk4_at(state, domain, uniform(), dt, nsteps, time);
    }
};

/* FixedBGAMRLevelHooks (FixedBGAMRLevelHooks.hpp:22-167): what a GRAMRLevel
 * delegates to on a static background -- RHS, post-update recovery, CFL and
 * a standalone SSPRK2 step -- forwarded to the level operator. */
class FixedBGAMRLevelHooks
{
  public:
    FixedBGAMRLevelHooks(
        const IdealGasEOS &a_eos, const grhd_background &a_background,
        double a_dx, const std::array<double, 3> &a_center,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : m_level_operator(a_eos, a_background, a_dx, a_center, a_atmosphere,
                           a_limiter_theta, a_use_reconstruction,
                           a_recovery_options, a_use_static_metric_sources,
                           a_comp_base, std::move(a_context), a_stream),
          m_dx(a_dx)
    {
    }

    const FixedBGLevelDataFiniteVolumeOperator &level_operator() const
    {
        return m_level_operator;
    }

    /* :40-46 */
    GRHD_IF_INVERSE_DX(V)
    void specific_eval_rhs(DeviceLevelData &soln, DeviceLevelData &rhs,
                           const ProblemDomain &domain,
                           const V &inverse_dx_U) const
    {
        m_level_operator.compute_rhs(rhs, soln, domain, inverse_dx_U);
    }

    /* :48-53 */
    void specific_eval_rhs(DeviceLevelData &soln, DeviceLevelData &rhs,
                           const ProblemDomain &domain, double dx) const
    {
        specific_eval_rhs(soln, rhs, domain, make_inverse_dx(dx));
    }

    /* :55-63 (the background is static: time is ignored) */
    GRHD_IF_INVERSE_DX(V)
    void specific_eval_rhs(DeviceLevelData &soln, DeviceLevelData &rhs,
                           const ProblemDomain &domain, const V &inverse_dx_U,
                           double time) const
    {
        (void)time;
        specific_eval_rhs(soln, rhs, domain, inverse_dx_U);
    }

    /* :65-71 */
    void specific_eval_rhs(DeviceLevelData &soln, DeviceLevelData &rhs,
                           const ProblemDomain &domain, double dx,
                           double time) const
    {
        specific_eval_rhs(soln, rhs, domain, make_inverse_dx(dx), time);
    }

    /* :73-77 */
    AtmosphereResetDiagnostics specific_update_ode(DeviceLevelData &soln) const
    {
        return m_level_operator.recover_primitives(soln);
    }

    /* :79-86 */
    AtmosphereResetDiagnostics specific_update_ode(DeviceLevelData &soln,
                                                   const DeviceLevelData &rhs,
                                                   double dt) const
    {
        (void)rhs;
        (void)dt;
        return specific_update_ode(soln);
    }

    /* :88-95 */
    GRHD_IF_INVERSE_DX(V)
    double compute_stable_dt(const DeviceLevelData &state,
                             const IntBox &domain_box, const V &inverse_dx_U,
                             double cfl) const
    {
        return m_level_operator.compute_stable_dt(state, domain_box,
                                                  inverse_dx_U, cfl);
    }

    /* :97-104 */
    GRHD_IF_INVERSE_DX(V)
    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain,
                             const V &inverse_dx_U, double cfl) const
    {
        return compute_stable_dt(state, domain.domainBox(), inverse_dx_U,
                                 cfl);
    }

    /* :106-111 */
    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain, double dx,
                             double cfl) const
    {
        return compute_stable_dt(state, domain, make_inverse_dx(dx), cfl);
    }

    /* :113-124 */
    GRHD_IF_INVERSE_DX(V)
    double compute_coupled_dt(const DeviceLevelData &state,
                              const ProblemDomain &domain,
                              const V &inverse_dx_U, double cfl,
                              double metric_dt) const
    {
        if (metric_dt <= 0.0)
            throw std::invalid_argument(
                "GRHD fixed-background coupled timestep must be positive");
        const double hydro_dt =
            compute_stable_dt(state, domain, inverse_dx_U, cfl);
        return std::min(metric_dt, hydro_dt);
    }

    /* :126-132 */
    double compute_coupled_dt(const DeviceLevelData &state,
                              const ProblemDomain &domain, double dx,
                              double cfl, double metric_dt) const
    {
        return compute_coupled_dt(state, domain, make_inverse_dx(dx), cfl,
                                  metric_dt);
    }

    /* :134-140 */
    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          double dt,
                          bool recover_primitives_after_update = true) const
    {
        m_level_operator.update_conserved(state, rhs, dt,
                                          recover_primitives_after_update);
    }

    /* :142-148 */
    GRHD_IF_INVERSE_DX(V)
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        const V &inverse_dx_U, double dt) const
    {
        m_level_operator.advance_ssprk2(state, domain, inverse_dx_U, dt);
    }

    /* :150-154 */
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        double dt) const
    {
        advance_ssprk2(state, domain, make_inverse_dx(m_dx), dt);
    }

  private:
    static InverseDx make_inverse_dx(double dx)
    {
        return InverseDx{{1.0 / dx, 1.0 / dx, 1.0 / dx}};
    }

    FixedBGLevelDataFiniteVolumeOperator m_level_operator;
    double m_dx;
};

/* TimeDependentFixedBGAMRLevelHooks (FixedBGAMRLevelHooks.hpp:170-310). */
class TimeDependentFixedBGAMRLevelHooks
{
  public:
    TimeDependentFixedBGAMRLevelHooks(
        const IdealGasEOS &a_eos, const grhd_background &a_background,
        double a_dx, const std::array<double, 3> &a_center,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true,
        bool a_use_metric_volume_time_sources = true,
        double a_metric_time_derivative_step = 1.0e-3, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : m_level_operator(a_eos, a_background, a_dx, a_center, a_atmosphere,
                           a_limiter_theta, a_use_reconstruction,
                           a_recovery_options, a_use_static_metric_sources,
                           a_use_metric_volume_time_sources,
                           a_metric_time_derivative_step, a_comp_base,
                           std::move(a_context), a_stream),
          m_dx(a_dx)
    {
    }

    const TimeDependentFixedBGLevelDataFiniteVolumeOperator &
    level_operator() const
    {
        return m_level_operator;
    }

    /* :194-201 */
    GRHD_IF_INVERSE_DX(V)
    void specific_eval_rhs(DeviceLevelData &soln, DeviceLevelData &rhs,
                           const ProblemDomain &domain, const V &inverse_dx_U,
                           double time) const
    {
        m_level_operator.compute_rhs(rhs, soln, domain, inverse_dx_U, time);
    }

    /* :203-209 */
    void specific_eval_rhs(DeviceLevelData &soln, DeviceLevelData &rhs,
                           const ProblemDomain &domain, double dx,
                           double time) const
    {
        specific_eval_rhs(soln, rhs, domain, make_inverse_dx(dx), time);
    }

    /* :211-215 */
    AtmosphereResetDiagnostics specific_update_ode(DeviceLevelData &soln,
                                                   double time) const
    {
        return m_level_operator.recover_primitives(soln, time);
    }

    /* :217-224 */
    AtmosphereResetDiagnostics specific_update_ode(DeviceLevelData &soln,
                                                   const DeviceLevelData &rhs,
                                                   double dt,
                                                   double time) const
    {
        (void)rhs;
        (void)dt;
        return specific_update_ode(soln, time);
    }

    /* :226-233 */
    GRHD_IF_INVERSE_DX(V)
    double compute_stable_dt(const DeviceLevelData &state,
                             const IntBox &domain_box, const V &inverse_dx_U,
                             double cfl, double time) const
    {
        return m_level_operator.compute_stable_dt(state, domain_box,
                                                  inverse_dx_U, cfl, time);
    }

    /* :235-242 */
    GRHD_IF_INVERSE_DX(V)
    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain,
                             const V &inverse_dx_U, double cfl,
                             double time) const
    {
        return compute_stable_dt(state, domain.domainBox(), inverse_dx_U, cfl,
                                 time);
    }

    /* :244-250 */
    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain, double dx,
                             double cfl, double time) const
    {
        return compute_stable_dt(state, domain, make_inverse_dx(dx), cfl,
                                 time);
    }

    /* :252-264 */
    GRHD_IF_INVERSE_DX(V)
    double compute_coupled_dt(const DeviceLevelData &state,
                              const ProblemDomain &domain,
                              const V &inverse_dx_U, double cfl,
                              double metric_dt, double time) const
    {
        if (metric_dt <= 0.0)
            throw std::invalid_argument(
                "GRHD time-dependent fixed-background coupled timestep must "
                "be positive");
        const double hydro_dt =
            compute_stable_dt(state, domain, inverse_dx_U, cfl, time);
        return std::min(metric_dt, hydro_dt);
    }

    /* :266-273 */
    double compute_coupled_dt(const DeviceLevelData &state,
                              const ProblemDomain &domain, double dx,
                              double cfl, double metric_dt,
                              double time) const
    {
        return compute_coupled_dt(state, domain, make_inverse_dx(dx), cfl,
                                  metric_dt, time);
    }

    /* :275-282 */
    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          double dt, double recovery_time,
                          bool recover_primitives_after_update = true) const
    {
        m_level_operator.update_conserved(state, rhs, dt, recovery_time,
                                          recover_primitives_after_update);
    }

    /* :284-291 */
    GRHD_IF_INVERSE_DX(V)
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        const V &inverse_dx_U, double time, double dt) const
    {
        m_level_operator.advance_ssprk2(state, domain, inverse_dx_U, time,
                                        dt);
    }

    /* :293-298 */
    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        double time, double dt) const
    {
        advance_ssprk2(state, domain, make_inverse_dx(m_dx), time, dt);
    }

  private:
    static InverseDx make_inverse_dx(double dx)
    {
        return InverseDx{{1.0 / dx, 1.0 / dx, 1.0 / dx}};
    }

    TimeDependentFixedBGLevelDataFiniteVolumeOperator m_level_operator;
    double m_dx;
};

/* compute_leveldata_conserved_sum (LevelDataDriver.hpp:22-57) */
inline ConservedSum compute_leveldata_conserved_sum(
    const DeviceLevelData &state, const ProblemDomain &domain, double dx,
    const std::array<double, 3> &center, int comp_base, grhd_ctx *ctx,
    void *stream = nullptr)
{
    const auto t = state.level_boxes(domain, dx, center);
    double s[5];
    check(grhd_level_conserved_sum(ctx, t.data(), (int)t.size(), state.data(),
                                   state.ncomp(), comp_base, s, stream),
          nullptr, "conserved_sum");
    ConservedSum out;
    out.D = s[0];
    out.S_L = {{s[1], s[2], s[3]}};
    out.tau = s[4];
    return out;
}

/* leveldata_conserved_state_is_finite (LevelDataDriver.hpp:59-92) */
inline bool leveldata_conserved_state_is_finite(
    const DeviceLevelData &state, const ProblemDomain &domain, double dx,
    const std::array<double, 3> &center, int comp_base, grhd_ctx *ctx,
    void *stream = nullptr)
{
    const auto t = state.level_boxes(domain, dx, center);
    int ok = 0;
    check(grhd_level_conserved_state_is_finite(ctx, t.data(), (int)t.size(),
                                               state.data(), state.ncomp(),
                                               comp_base, &ok, stream),
          nullptr, "conserved_state_is_finite");
    return ok != 0;
}

/* FixedBGLevelDataDriverView (FixedBGLevelDataDriver.hpp:17-141): a level
 * state, its domain, 1/dx in every direction, the level operator, and a
 * time and step count advanced by each update. */
class FixedBGLevelDataDriverView
{
  public:
    FixedBGLevelDataDriverView(
        DeviceLevelData &a_state, const ProblemDomain &a_domain, double a_dx,
        const std::array<double, 3> &a_center, const IdealGasEOS &a_eos,
        const grhd_background &a_background,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : m_state(a_state), m_domain(a_domain), m_dx(a_dx),
          m_center(a_center),
          m_inverse_dx_U{{1.0 / a_dx, 1.0 / a_dx, 1.0 / a_dx}},
          m_operator(a_eos, a_background, a_dx, a_center, a_atmosphere,
                     a_limiter_theta, a_use_reconstruction,
                     a_recovery_options, a_use_static_metric_sources,
                     a_comp_base, std::move(a_context), a_stream)
    {
        if (a_dx <= 0.0)
            throw std::invalid_argument(
                "GRHD fixed-background driver dx must be positive");
    }

    FixedBGLevelDataDriverView(
        DeviceLevelData &a_state, const IntBox &a_domain_box, double a_dx,
        const std::array<double, 3> &a_center, const IdealGasEOS &a_eos,
        const grhd_background &a_background,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : FixedBGLevelDataDriverView(
              a_state, ProblemDomain(a_domain_box), a_dx, a_center, a_eos,
              a_background, a_atmosphere, a_limiter_theta,
              a_use_reconstruction, a_recovery_options,
              a_use_static_metric_sources, a_comp_base, std::move(a_context),
              a_stream)
    {
    }

    DeviceLevelData &state() { return m_state; }
    const DeviceLevelData &state() const { return m_state; }
    const std::vector<IntBox> &grids() const { return m_state.grids(); }
    const IntBox &domain_box() const { return m_domain.domainBox(); }
    const ProblemDomain &problem_domain() const { return m_domain; }
    double dx() const { return m_dx; }
    const std::array<double, 3> &center() const { return m_center; }
    const InverseDx &inverse_dx_U() const { return m_inverse_dx_U; }
    double time() const { return m_time; }
    int step() const { return m_step; }

    void compute_rhs(DeviceLevelData &rhs)
    {
        m_operator.compute_rhs(rhs, m_state, m_domain, m_inverse_dx_U);
    }

    void add_conserved_rhs_to(DeviceLevelData &rhs)
    {
        m_operator.add_conserved_rhs_to(rhs, m_state, m_domain,
                                        m_inverse_dx_U);
    }

    double compute_stable_dt(double cfl) const
    {
        return m_operator.compute_stable_dt(m_state, m_domain.domainBox(),
                                            m_inverse_dx_U, cfl);
    }

    double compute_coupled_dt(double cfl, double metric_dt) const
    {
        if (metric_dt <= 0.0)
            throw std::invalid_argument(
                "GRHD fixed-background coupled timestep must be positive");
        return std::min(metric_dt, compute_stable_dt(cfl));
    }

    AtmosphereResetDiagnostics recover_primitives()
    {
        return m_operator.recover_primitives(m_state, m_domain);
    }

    void update_conserved(const DeviceLevelData &rhs, double dt,
                          bool recover_primitives_after_update = true)
    {
        m_operator.update_conserved(m_state, rhs, m_domain, dt,
                                    recover_primitives_after_update);
        m_time += dt;
        ++m_step;
    }

    void advance_ssprk2(double dt)
    {
        m_operator.advance_ssprk2(m_state, m_domain, m_inverse_dx_U, dt);
        m_time += dt;
        ++m_step;
    }

    ConservedSum conserved_sum() const
    {
        return compute_leveldata_conserved_sum(
            m_state, m_domain, m_dx, m_center, m_operator.comp_base(),
            m_operator.context(), m_operator.stream());
    }

    bool conserved_state_is_finite() const
    {
        return leveldata_conserved_state_is_finite(
            m_state, m_domain, m_dx, m_center, m_operator.comp_base(),
            m_operator.context(), m_operator.stream());
    }

  private:
    DeviceLevelData &m_state;
    ProblemDomain m_domain;
    double m_dx;
    std::array<double, 3> m_center;
    InverseDx m_inverse_dx_U;
    FixedBGLevelDataFiniteVolumeOperator m_operator;
    double m_time = 0.0;
    int m_step = 0;
};

/* TimeDependentFixedBGLevelDataDriverView
 * (FixedBGLevelDataDriver.hpp:232-382): as above with the background at the
 * view's time; an update recovers at time + dt. */
class TimeDependentFixedBGLevelDataDriverView
{
  public:
    TimeDependentFixedBGLevelDataDriverView(
        DeviceLevelData &a_state, const ProblemDomain &a_domain, double a_dx,
        const std::array<double, 3> &a_center, const IdealGasEOS &a_eos,
        const grhd_background &a_background,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true,
        bool a_use_metric_volume_time_sources = true,
        double a_metric_time_derivative_step = 1.0e-3, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : m_state(a_state), m_domain(a_domain), m_dx(a_dx),
          m_center(a_center),
          m_inverse_dx_U{{1.0 / a_dx, 1.0 / a_dx, 1.0 / a_dx}},
          m_operator(a_eos, a_background, a_dx, a_center, a_atmosphere,
                     a_limiter_theta, a_use_reconstruction,
                     a_recovery_options, a_use_static_metric_sources,
                     a_use_metric_volume_time_sources,
                     a_metric_time_derivative_step, a_comp_base,
                     std::move(a_context), a_stream)
    {
        if (a_dx <= 0.0)
            throw std::invalid_argument(
                "GRHD fixed-background driver dx must be positive");
    }

    TimeDependentFixedBGLevelDataDriverView(
        DeviceLevelData &a_state, const IntBox &a_domain_box, double a_dx,
        const std::array<double, 3> &a_center, const IdealGasEOS &a_eos,
        const grhd_background &a_background,
        const AtmosphereOptions &a_atmosphere = AtmosphereOptions(),
        double a_limiter_theta = 1.5, bool a_use_reconstruction = true,
        const RecoveryOptions &a_recovery_options = RecoveryOptions(),
        bool a_use_static_metric_sources = true,
        bool a_use_metric_volume_time_sources = true,
        double a_metric_time_derivative_step = 1.0e-3, int a_comp_base = 0,
        std::shared_ptr<Context> a_context = nullptr, void *a_stream = nullptr)
        : TimeDependentFixedBGLevelDataDriverView(
              a_state, ProblemDomain(a_domain_box), a_dx, a_center, a_eos,
              a_background, a_atmosphere, a_limiter_theta,
              a_use_reconstruction, a_recovery_options,
              a_use_static_metric_sources, a_use_metric_volume_time_sources,
              a_metric_time_derivative_step, a_comp_base,
              std::move(a_context), a_stream)
    {
    }

    DeviceLevelData &state() { return m_state; }
    const DeviceLevelData &state() const { return m_state; }
    const std::vector<IntBox> &grids() const { return m_state.grids(); }
    const IntBox &domain_box() const { return m_domain.domainBox(); }
    const ProblemDomain &problem_domain() const { return m_domain; }
    double dx() const { return m_dx; }
    const std::array<double, 3> &center() const { return m_center; }
    const InverseDx &inverse_dx_U() const { return m_inverse_dx_U; }
    double time() const { return m_time; }
    int step() const { return m_step; }
    void set_time(double time) { m_time = time; }

    void compute_rhs(DeviceLevelData &rhs)
    {
        m_operator.compute_rhs(rhs, m_state, m_domain, m_inverse_dx_U,
                               m_time);
    }

    void compute_rhs(DeviceLevelData &rhs, double time)
    {
        m_operator.compute_rhs(rhs, m_state, m_domain, m_inverse_dx_U, time);
    }

    double compute_stable_dt(double cfl) const
    {
        return m_operator.compute_stable_dt(m_state, m_domain.domainBox(),
                                            m_inverse_dx_U, cfl, m_time);
    }

    double compute_stable_dt(double cfl, double time) const
    {
        return m_operator.compute_stable_dt(m_state, m_domain.domainBox(),
                                            m_inverse_dx_U, cfl, time);
    }

    double compute_coupled_dt(double cfl, double metric_dt) const
    {
        if (metric_dt <= 0.0)
            throw std::invalid_argument(
                "GRHD fixed-background coupled timestep must be positive");
        return std::min(metric_dt, compute_stable_dt(cfl));
    }

    AtmosphereResetDiagnostics recover_primitives()
    {
        return m_operator.recover_primitives(m_state, m_domain, m_time);
    }

    AtmosphereResetDiagnostics recover_primitives(double time)
    {
        return m_operator.recover_primitives(m_state, m_domain, time);
    }

    void update_conserved(const DeviceLevelData &rhs, double dt,
                          bool recover_primitives_after_update = true)
    {
        const double recovery_time = m_time + dt;
        m_operator.update_conserved(m_state, rhs, m_domain, dt, recovery_time,
                                    recover_primitives_after_update);
        m_time = recovery_time;
        ++m_step;
    }

    void advance_ssprk2(double dt)
    {
        m_operator.advance_ssprk2(m_state, m_domain, m_inverse_dx_U, m_time,
                                  dt);
        m_time += dt;
        ++m_step;
    }

    ConservedSum conserved_sum() const
    {
        return compute_leveldata_conserved_sum(
            m_state, m_domain, m_dx, m_center, m_operator.comp_base(),
            m_operator.context(), m_operator.stream());
    }

    bool conserved_state_is_finite() const
    {
        return leveldata_conserved_state_is_finite(
            m_state, m_domain, m_dx, m_center, m_operator.comp_base(),
            m_operator.context(), m_operator.stream());
    }

  private:
    DeviceLevelData &m_state;
    ProblemDomain m_domain;
    double m_dx;
    std::array<double, 3> m_center;
    InverseDx m_inverse_dx_U;
    TimeDependentFixedBGLevelDataFiniteVolumeOperator m_operator;
    double m_time = 0.0;
    int m_step = 0;
};

#undef GRHD_IF_INVERSE_DX

} // namespace amd
} // namespace GRHD

#endif /* GRHD_OPERATOR_HPP_ */
