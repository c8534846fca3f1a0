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
 * Header-only; link against libgrhd_hip.so.  Uniform dx only (the
 * reference's inverse_dx_U overloads are 1/dx in every direction).
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
};

/* LevelData<FArrayBox> in HBM.  `pool` is a device pointer to
 * pool_size() doubles (the caller owns it). */
class DeviceLevelData
{
  public:
    DeviceLevelData(std::vector<IntBox> grids, int ghosts, int ncomp,
                    double *pool)
        : m_grids(std::move(grids)), m_ghosts(ghosts), m_ncomp(ncomp),
          m_pool(pool)
    {
        if (ghosts < 0 || ncomp <= 0 || m_grids.empty())
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
    std::vector<grhd_level_box> level_boxes(const ProblemDomain &domain,
                                            double dx,
                                            const std::array<double, 3> &c)
        const
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
            }
            b.dx = dx;
            out[n].state_offset = m_offsets[n];
        }
        return out;
    }

  private:
    std::vector<IntBox> m_grids;
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

    /* FP32 reconstruction, face metric and HLLE (see grhd_params); off by
     * default, which is the bit-exact FP64 path. */
    void set_mixed_precision(bool on) { m_params.mixed_precision = on ? 1 : 0; }

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
    std::vector<grhd_level_box> table(const DeviceLevelData &l,
                                      const ProblemDomain &domain) const
    {
        return l.level_boxes(domain, m_dx, m_center);
    }

    void rhs_at(DeviceLevelData &rhs, DeviceLevelData &state,
                const ProblemDomain &domain, double time) const
    {
        const auto t = table(state, domain);
        if (rhs.ncomp() != state.ncomp() ||
            rhs.pool_size() != rhs.ncomp() * state.num_interior())
            throw std::invalid_argument(
                "rhs must hold ncomp components over the interior boxes");
        grhd_error err{};
        check(grhd_level_compute_rhs(context(), t.data(), (int)t.size(),
                                     &m_params, &m_background, time,
                                     state.data(), state.ncomp(), m_comp_base,
                                     rhs.data(), nullptr, &err, m_stream),
              &err, "compute_rhs");
    }

    AtmosphereResetDiagnostics recover_at(DeviceLevelData &state,
                                          const ProblemDomain &domain,
                                          double time) const
    {
        const auto t = table(state, domain);
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
        const auto t = table(state, domain);
        grhd_diagnostics d{};
        check(grhd_level_update_conserved(
                  context(), t.data(), (int)t.size(), &m_params,
                  &m_background, state.data(), state.ncomp(), m_comp_base,
                  rhs.data(), dt, recover ? 1 : 0, recovery_time, nullptr, &d,
                  m_stream),
              nullptr, "update_conserved");
    }

    double max_inverse_dt_at(const DeviceLevelData &state,
                             const ProblemDomain &domain, double time) const
    {
        const auto t = table(state, domain);
        double out = 0.0;
        grhd_error err{};
        check(grhd_level_compute_max_inverse_dt(
                  context(), t.data(), (int)t.size(), &m_params,
                  &m_background, time, state.data(), state.ncomp(),
                  m_comp_base, &out, &err, m_stream),
              &err, "compute_max_inverse_dt");
        return out;
    }

    void ssprk2_at(DeviceLevelData &state, const ProblemDomain &domain,
                   double dt, double time) const
    {
        const auto t = table(state, domain);
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
                                      const ProblemDomain &domain, double dt,
                                      int nsteps, double time) const
    {
        const auto t = table(state, domain);
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

    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const ProblemDomain &domain) const
    {
        rhs_at(rhs, state, domain, 0.0);
    }

    AtmosphereResetDiagnostics recover_primitives(DeviceLevelData &state,
                                                  const ProblemDomain &domain)
        const
    {
        return recover_at(state, domain, 0.0);
    }

    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          const ProblemDomain &domain, double dt,
                          bool recover_primitives_after_update = true) const
    {
        update_at(state, rhs, domain, dt, recover_primitives_after_update,
                  0.0);
    }

    double compute_max_inverse_dt(const DeviceLevelData &state,
                                  const ProblemDomain &domain) const
    {
        return max_inverse_dt_at(state, domain, 0.0);
    }

    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain, double cfl) const
    {
        if (cfl <= 0.0)
            throw std::invalid_argument("GRHD CFL number must be positive");
        return cfl / std::max(compute_max_inverse_dt(state, domain), 1.0e-14);
    }

    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        double dt) const
    {
        ssprk2_at(state, domain, dt, 0.0);
    }

    /* Classical RK4, `nsteps` steps (no counterpart in the reference; see
     * grhd_advance_rk4_box). */
    AtmosphereResetDiagnostics advance_rk4(DeviceLevelData &state,
                                           const ProblemDomain &domain,
                                           double dt, int nsteps = 1) const
    {
        return rk4_at(state, domain, dt, nsteps, 0.0);
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

    void compute_rhs(DeviceLevelData &rhs, DeviceLevelData &state,
                     const ProblemDomain &domain, double time) const
    {
        rhs_at(rhs, state, domain, time);
    }

    AtmosphereResetDiagnostics recover_primitives(DeviceLevelData &state,
                                                  const ProblemDomain &domain,
                                                  double time) const
    {
        return recover_at(state, domain, time);
    }

    void update_conserved(DeviceLevelData &state, const DeviceLevelData &rhs,
                          const ProblemDomain &domain, double dt,
                          double recovery_time,
                          bool recover_primitives_after_update = true) const
    {
        update_at(state, rhs, domain, dt, recover_primitives_after_update,
                  recovery_time);
    }

    double compute_max_inverse_dt(const DeviceLevelData &state,
                                  const ProblemDomain &domain,
                                  double time) const
    {
        return max_inverse_dt_at(state, domain, time);
    }

    double compute_stable_dt(const DeviceLevelData &state,
                             const ProblemDomain &domain, double cfl,
                             double time) const
    {
        if (cfl <= 0.0)
            throw std::invalid_argument("GRHD CFL number must be positive");
        return cfl /
               std::max(compute_max_inverse_dt(state, domain, time), 1.0e-14);
    }

    void advance_ssprk2(DeviceLevelData &state, const ProblemDomain &domain,
                        double dt, double time) const
    {
        ssprk2_at(state, domain, dt, time);
    }

    AtmosphereResetDiagnostics advance_rk4(DeviceLevelData &state,
                                           const ProblemDomain &domain,
                                           double dt, int nsteps,
                                           double time) const
    {
        return rk4_at(state, domain, dt, nsteps, time);
    }
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

} // namespace amd
} // namespace GRHD

#endif /* GRHD_OPERATOR_HPP_ */
