/*
 * Test driver for the AMR-level hooks and driver views of the C++ host
 * mirror (include/grhd_operator.hpp), after
 * Tests/GRHDFixedBGAMRHookTest/GRHDFixedBGAMRHookTest.cpp:208-589: on a
 * 2-box level the hooks and the driver views must give what the level
 * operator gives directly (the reference allows 1e-12 on the rhs and the
 * SSPRK2 state and 1e-14 on dt and the coupled dt; here they must be
 * equal), the recovery must fail and reset nothing, and the views' time and
 * step must advance.  add_conserved_rhs_to on an rhs already holding the
 * hydro rhs must double its conserved components and leave the others.
 * The hook results are written out for the Python test to compare with the
 * oracle.
 *
 * usage: hooks_driver <dir> <time_dependent 0|1> <time> <idx> <idy> <idz>
 *   <dir>/level.txt   ghosts ncomp nbox dx cx cy cz dlo[3] dhi[3]
 *                     then one line per box: lo[3] hi[3]
 *   <dir>/params.bin  grhd_params, <dir>/background.bin grhd_background
 *   <dir>/state.bin   pool (doubles)
 *   idx idy idz       inverse_dx_U passed to every call that takes one
 * writes <dir>/hook_rhs.bin, <dir>/hook_advanced.bin and <dir>/result.txt:
 * "pass" or the failed checks, then the hook dt and the SSPRK2 dt.
 */
#include "grhd_operator.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace GRHD;
using namespace GRHD::amd;

namespace
{

/* stands in for GRChombo's Tensor<1, double>: anything with operator[] */
struct Tensor1
{
    double v[3];
    double operator[](int i) const { return v[i]; }
};

template <class T> T read_struct(const std::string &path)
{
    T v{};
    std::ifstream in(path, std::ios::binary);
    in.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!in)
        throw std::runtime_error("cannot read " + path);
    return v;
}

/* One device pool (owned here, by the test). */
class Pool
{
  public:
    explicit Pool(long long n) : m_n(n)
    {
        if (hipMalloc((void **)&m_d, sizeof(double) * (size_t)n) !=
                hipSuccess ||
            hipMemset(m_d, 0, sizeof(double) * (size_t)n) != hipSuccess)
            throw std::runtime_error("device allocation failed");
    }
    ~Pool() { (void)hipFree(m_d); }
    Pool(const Pool &) = delete;
    Pool &operator=(const Pool &) = delete;
    double *data() const { return m_d; }
    void put(const std::vector<double> &h)
    {
        if (hipMemcpy(m_d, h.data(), sizeof(double) * h.size(),
                      hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("copy to device failed");
    }
    std::vector<double> get() const
    {
        std::vector<double> h((size_t)m_n);
        if (hipMemcpy(h.data(), m_d, sizeof(double) * (size_t)m_n,
                      hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("copy to host failed");
        return h;
    }

  private:
    double *m_d = nullptr;
    long long m_n;
};

std::vector<std::string> failures;

void expect(bool ok, const std::string &what)
{
    if (!ok)
        failures.push_back(what);
}

void write(const std::string &path, const std::vector<double> &v)
{
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char *>(v.data()),
              (std::streamsize)(sizeof(double) * v.size()));
}

struct Level
{
    std::vector<IntBox> grids;
    ProblemDomain domain{IntBox()};
    int ghosts = 0, ncomp = 0;
    double dx = 0.0;
    std::array<double, 3> center{{0.0, 0.0, 0.0}};
    std::vector<double> state;

    long long state_size() const
    {
        long long n = 0;
        for (const IntBox &b : grids)
            n += (long long)ncomp * b.grown(ghosts).num_pts();
        return n;
    }
    long long rhs_size() const
    {
        long long n = 0;
        for (const IntBox &b : grids)
            n += (long long)ncomp * b.num_pts();
        return n;
    }
};

Level read_level(const std::string &dir)
{
    Level L;
    std::ifstream lv(dir + "/level.txt");
    int nbox = 0;
    lv >> L.ghosts >> L.ncomp >> nbox >> L.dx >> L.center[0] >> L.center[1] >>
        L.center[2];
    IntBox dom;
    for (int d = 0; d < 3; ++d)
        lv >> dom.lo[d];
    for (int d = 0; d < 3; ++d)
        lv >> dom.hi[d];
    L.domain = ProblemDomain(dom);
    L.grids.resize(nbox);
    for (IntBox &b : L.grids)
    {
        for (int d = 0; d < 3; ++d)
            lv >> b.lo[d];
        for (int d = 0; d < 3; ++d)
            lv >> b.hi[d];
    }
    if (!lv)
        throw std::runtime_error("cannot read level.txt");
    L.state.resize((size_t)L.state_size());
    std::ifstream in(dir + "/state.bin", std::ios::binary);
    in.read(reinterpret_cast<char *>(L.state.data()),
            (std::streamsize)(sizeof(double) * L.state.size()));
    if (!in)
        throw std::runtime_error("cannot read state.bin");
    return L;
}

AtmosphereOptions atmosphere_of(const grhd_params &p)
{
    AtmosphereOptions a;
    a.rho_floor = p.rho_floor;
    a.pressure_floor = p.pressure_floor;
    a.max_velocity_squared = p.max_velocity_squared;
    a.metric_source_atmosphere_factor = p.metric_source_atmosphere_factor;
    return a;
}

RecoveryOptions recovery_of(const grhd_params &p)
{
    RecoveryOptions r;
    r.max_iterations = p.max_iterations;
    r.tolerance = p.tolerance;
    r.pressure_floor = p.recovery_pressure_floor;
    r.max_velocity_squared = p.recovery_max_velocity_squared;
    return r;
}

/* check_static_hook / check_static_driver (:208-274, :367-462) */
void check_static(const std::string &dir, const Level &L,
                  const grhd_params &p, const grhd_background &bg,
                  const Tensor1 &inv, double &hook_dt_out, double &dt_out)
{
    const IdealGasEOS eos(p.adiabatic_index);
    const AtmosphereOptions atm = atmosphere_of(p);
    const RecoveryOptions rec = recovery_of(p);
    auto ctx = std::make_shared<Context>();
    Pool hs(L.state_size()), ds(L.state_size()), ha(L.state_size()),
        da(L.state_size()), hr(L.rhs_size()), dr(L.rhs_size());
    for (Pool *q : {&hs, &ds, &ha, &da})
        q->put(L.state);
    auto level = [&](Pool &q, bool rhs) {
        return DeviceLevelData(L.grids, L.domain, rhs ? 0 : L.ghosts,
                               L.ncomp, q.data());
    };
    DeviceLevelData hook_state = level(hs, false),
                    direct_state = level(ds, false),
                    hook_advanced = level(ha, false),
                    direct_advanced = level(da, false),
                    hook_rhs = level(hr, true), direct_rhs = level(dr, true);

    FixedBGAMRLevelHooks hooks(eos, bg, L.dx, L.center, atm,
                               p.limiter_theta, p.use_reconstruction != 0,
                               rec, p.use_static_metric_sources != 0, 0, ctx);
    FixedBGLevelDataFiniteVolumeOperator direct(
        eos, bg, L.dx, L.center, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, p.use_static_metric_sources != 0, 0,
        ctx);

    hooks.specific_eval_rhs(hook_state, hook_rhs, L.domain, inv);
    direct.compute_rhs(direct_rhs, direct_state, L.domain, inv);
    expect(hr.get() == dr.get(), "static hook rhs");
    write(dir + "/hook_rhs.bin", hr.get());

    const double hook_dt =
        hooks.compute_stable_dt(hook_state, L.domain, inv, 0.2);
    const double direct_dt =
        direct.compute_stable_dt(direct_state, L.domain.domainBox(), inv, 0.2);
    expect(hook_dt == direct_dt, "static hook dt");
    expect(hooks.compute_coupled_dt(hook_state, L.domain, inv, 0.2,
                                    0.5 * hook_dt) == 0.5 * hook_dt,
           "static coupled dt");
    const AtmosphereResetDiagnostics r = hooks.specific_update_ode(hook_state);
    expect(r.num_failed_recoveries == 0 && r.num_conserved_resets == 0,
           "static hook recovery reset");

    const double dt = std::min(0.02, 0.25 * hook_dt);
    hooks.advance_ssprk2(hook_advanced, L.domain, inv, dt);
    direct.advance_ssprk2(direct_advanced, L.domain, inv, dt);
    expect(ha.get() == da.get(), "static hook ssprk2");
    write(dir + "/hook_advanced.bin", ha.get());

    /* add_conserved_rhs_to (FixedBGFArrayBoxFiniteVolume.hpp:492-503) */
    {
        Pool s2(L.state_size()), r2(L.rhs_size());
        s2.put(L.state);
        const std::vector<double> seed = dr.get(); /* the hydro rhs */
        r2.put(seed);
        DeviceLevelData st2 = level(s2, false), rh2 = level(r2, true);
        direct.add_conserved_rhs_to(rh2, st2, L.domain, inv);
        const std::vector<double> got = r2.get();
        bool ok = true;
        long long off = 0;
        for (const IntBox &b : L.grids)
        {
            const long long n = b.num_pts();
            for (int c = 0; c < L.ncomp; ++c)
                for (long long i = 0; i < n; ++i)
                {
                    const double s = seed[(size_t)(off + c * n + i)];
                    const double g = got[(size_t)(off + c * n + i)];
                    ok = ok && (c < 5 ? g == s + s : g == s);
                }
            off += (long long)L.ncomp * n;
        }
        expect(ok, "add_conserved_rhs_to");
    }

    /* the driver view (:367-462) */
    Pool vs(L.state_size()), vr(L.rhs_size()), us(L.state_size()),
        as(L.state_size());
    vs.put(L.state);
    DeviceLevelData view_state = level(vs, false), view_rhs = level(vr, true);
    FixedBGLevelDataDriverView driver(
        view_state, L.domain, L.dx, L.center, eos, bg, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, p.use_static_metric_sources != 0, 0,
        ctx);
    driver.compute_rhs(view_rhs);
    /* the view uses 1/dx, which is the direct rhs only when inv is */
    Pool ur(L.rhs_size()), ucheck(L.state_size());
    ucheck.put(L.state);
    DeviceLevelData uniform_state = level(ucheck, false),
                    uniform_rhs = level(ur, true);
    direct.compute_rhs(uniform_rhs, uniform_state, L.domain);
    expect(vr.get() == ur.get(), "static driver rhs");
    const double view_dt = driver.compute_stable_dt(0.2);
    expect(view_dt == direct.compute_stable_dt(uniform_state,
                                               L.domain.domainBox(), 0.2),
           "static driver dt");
    expect(driver.compute_coupled_dt(0.2, 0.5 * view_dt) == 0.5 * view_dt,
           "static driver coupled dt");
    const AtmosphereResetDiagnostics vrec = driver.recover_primitives();
    expect(vrec.num_failed_recoveries == 0 && vrec.num_conserved_resets == 0,
           "static driver recovery reset");
    const double vdt = std::min(0.02, 0.25 * view_dt);
    us.put(L.state);
    DeviceLevelData update_state = level(us, false);
    FixedBGLevelDataDriverView update_driver(
        update_state, L.domain, L.dx, L.center, eos, bg, atm,
        p.limiter_theta, p.use_reconstruction != 0, rec,
        p.use_static_metric_sources != 0, 0, ctx);
    update_driver.update_conserved(view_rhs, vdt);
    Pool ds2(L.state_size());
    ds2.put(L.state);
    DeviceLevelData direct_update = level(ds2, false);
    direct.update_conserved(direct_update, view_rhs, vdt);
    expect(us.get() == ds2.get(), "static driver update");
    expect(update_driver.time() == vdt && update_driver.step() == 1,
           "static driver update time/step");
    as.put(L.state);
    DeviceLevelData adv_state = level(as, false);
    FixedBGLevelDataDriverView advance_driver(
        adv_state, L.domain, L.dx, L.center, eos, bg, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, p.use_static_metric_sources != 0, 0,
        ctx);
    advance_driver.advance_ssprk2(vdt);
    Pool da2(L.state_size());
    da2.put(L.state);
    DeviceLevelData direct_adv = level(da2, false);
    direct.advance_ssprk2(direct_adv, L.domain, vdt);
    expect(as.get() == da2.get(), "static driver ssprk2");
    expect(advance_driver.conserved_state_is_finite(),
           "static driver finite");
    expect(advance_driver.time() == vdt && advance_driver.step() == 1,
           "static driver advance time/step");
    hook_dt_out = hook_dt;
    dt_out = dt;
}

/* check_time_dependent_hook / check_time_dependent_driver (:276-365,
 * :464-589) */
void check_time_dependent(const std::string &dir, const Level &L,
                          const grhd_params &p, const grhd_background &bg,
                          double time, const Tensor1 &inv,
                          double &hook_dt_out, double &dt_out)
{
    const IdealGasEOS eos(p.adiabatic_index);
    const AtmosphereOptions atm = atmosphere_of(p);
    const RecoveryOptions rec = recovery_of(p);
    const bool stat = p.use_static_metric_sources != 0;
    const bool vol = p.use_metric_volume_time_sources != 0;
    const double step = p.metric_time_derivative_step;
    auto ctx = std::make_shared<Context>();
    Pool hs(L.state_size()), ds(L.state_size()), ha(L.state_size()),
        da(L.state_size()), hr(L.rhs_size()), dr(L.rhs_size());
    for (Pool *q : {&hs, &ds, &ha, &da})
        q->put(L.state);
    auto level = [&](Pool &q, bool rhs) {
        return DeviceLevelData(L.grids, L.domain, rhs ? 0 : L.ghosts,
                               L.ncomp, q.data());
    };
    DeviceLevelData hook_state = level(hs, false),
                    direct_state = level(ds, false),
                    hook_advanced = level(ha, false),
                    direct_advanced = level(da, false),
                    hook_rhs = level(hr, true), direct_rhs = level(dr, true);

    TimeDependentFixedBGAMRLevelHooks hooks(
        eos, bg, L.dx, L.center, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, stat, vol, step, 0, ctx);
    TimeDependentFixedBGLevelDataFiniteVolumeOperator direct(
        eos, bg, L.dx, L.center, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, stat, vol, step, 0, ctx);

    hooks.specific_eval_rhs(hook_state, hook_rhs, L.domain, inv, time);
    direct.compute_rhs(direct_rhs, direct_state, L.domain, inv, time);
    expect(hr.get() == dr.get(), "time-dependent hook rhs");
    write(dir + "/hook_rhs.bin", hr.get());

    const double hook_dt =
        hooks.compute_stable_dt(hook_state, L.domain, inv, 0.2, time);
    const double direct_dt = direct.compute_stable_dt(
        direct_state, L.domain.domainBox(), inv, 0.2, time);
    expect(hook_dt == direct_dt, "time-dependent hook dt");
    expect(hooks.compute_coupled_dt(hook_state, L.domain, inv, 0.2,
                                    0.5 * hook_dt, time) == 0.5 * hook_dt,
           "time-dependent coupled dt");
    const AtmosphereResetDiagnostics r =
        hooks.specific_update_ode(hook_state, time);
    expect(r.num_failed_recoveries == 0 && r.num_conserved_resets == 0,
           "time-dependent hook recovery reset");

    const double dt = std::min(0.02, 0.25 * hook_dt);
    hooks.advance_ssprk2(hook_advanced, L.domain, inv, time, dt);
    direct.advance_ssprk2(direct_advanced, L.domain, inv, time, dt);
    expect(ha.get() == da.get(), "time-dependent hook ssprk2");
    write(dir + "/hook_advanced.bin", ha.get());

    /* the driver view at the same time */
    Pool vs(L.state_size()), vr(L.rhs_size());
    vs.put(L.state);
    DeviceLevelData view_state = level(vs, false), view_rhs = level(vr, true);
    TimeDependentFixedBGLevelDataDriverView driver(
        view_state, L.domain, L.dx, L.center, eos, bg, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, stat, vol, step, 0, ctx);
    driver.set_time(time);
    driver.compute_rhs(view_rhs);
    Pool ur(L.rhs_size()), ucheck(L.state_size());
    ucheck.put(L.state);
    DeviceLevelData uniform_state = level(ucheck, false),
                    uniform_rhs = level(ur, true);
    direct.compute_rhs(uniform_rhs, uniform_state, L.domain, time);
    expect(vr.get() == ur.get(), "time-dependent driver rhs");
    const double view_dt = driver.compute_stable_dt(0.2);
    expect(view_dt == direct.compute_stable_dt(uniform_state, L.domain, 0.2,
                                               time),
           "time-dependent driver dt");
    expect(driver.compute_coupled_dt(0.2, 0.5 * view_dt) == 0.5 * view_dt,
           "time-dependent driver coupled dt");
    const AtmosphereResetDiagnostics vrec = driver.recover_primitives();
    expect(vrec.num_failed_recoveries == 0 && vrec.num_conserved_resets == 0,
           "time-dependent driver recovery reset");
    const double vdt = std::min(0.02, 0.25 * view_dt);
    Pool us(L.state_size()), ds2(L.state_size());
    us.put(L.state);
    ds2.put(L.state);
    DeviceLevelData update_state = level(us, false),
                    direct_update = level(ds2, false);
    TimeDependentFixedBGLevelDataDriverView update_driver(
        update_state, L.domain, L.dx, L.center, eos, bg, atm,
        p.limiter_theta, p.use_reconstruction != 0, rec, stat, vol, step, 0,
        ctx);
    update_driver.set_time(time);
    update_driver.update_conserved(view_rhs, vdt);
    direct.update_conserved(direct_update, view_rhs, vdt, time + vdt);
    expect(us.get() == ds2.get(), "time-dependent driver update");
    expect(update_driver.time() == time + vdt && update_driver.step() == 1,
           "time-dependent driver update time/step");
    Pool as(L.state_size()), da2(L.state_size());
    as.put(L.state);
    da2.put(L.state);
    DeviceLevelData adv_state = level(as, false),
                    direct_adv = level(da2, false);
    TimeDependentFixedBGLevelDataDriverView advance_driver(
        adv_state, L.domain, L.dx, L.center, eos, bg, atm, p.limiter_theta,
        p.use_reconstruction != 0, rec, stat, vol, step, 0, ctx);
    advance_driver.set_time(time);
    advance_driver.advance_ssprk2(vdt);
    direct.advance_ssprk2(direct_adv, L.domain, time, vdt);
    expect(as.get() == da2.get(), "time-dependent driver ssprk2");
    expect(advance_driver.conserved_state_is_finite(),
           "time-dependent driver finite");
    expect(advance_driver.time() == time + vdt &&
               advance_driver.step() == 1,
           "time-dependent driver advance time/step");
    hook_dt_out = hook_dt;
    dt_out = dt;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 7)
    {
        std::fprintf(stderr,
                     "usage: %s dir time_dependent time idx idy idz\n",
                     argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const bool td = std::atoi(argv[2]) != 0;
    const double time = std::atof(argv[3]);
    const Tensor1 inv{{std::atof(argv[4]), std::atof(argv[5]),
                       std::atof(argv[6])}};
    try
    {
        const Level L = read_level(dir);
        const grhd_params p = read_struct<grhd_params>(dir + "/params.bin");
        const grhd_background bg =
            read_struct<grhd_background>(dir + "/background.bin");
        double hook_dt = 0.0, dt = 0.0;
        if (td)
            check_time_dependent(dir, L, p, bg, time, inv, hook_dt, dt);
        else
            check_static(dir, L, p, bg, inv, hook_dt, dt);
        std::ofstream out(dir + "/result.txt");
        out.precision(17);
        if (failures.empty())
            out << "pass\n";
        for (const std::string &f : failures)
            out << "FAILED " << f << "\n";
        out << hook_dt << " " << dt << "\n";
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "hooks_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
