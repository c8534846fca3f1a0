/*
 * Test driver for the C++ host mirror (include/grhd_operator.hpp): reads a
 * level (boxes, options, background, FAB pool) written by the Python test,
 * runs the reference's LevelData SSPRK2 loop with
 * FixedBGLevelDataFiniteVolumeOperator -- dt = min(compute_stable_dt(cfl),
 * max_dt, t_end - t) until t_end, as GRHDLevelDataMichelAccretion.cpp does
 * -- and writes the final pool, the step count and the level sums.
 *
 * usage: level_driver <dir> <cfl> <max_dt> <t_end>
 *   <dir>/level.txt   ghosts ncomp nbox dx cx cy cz dlo[3] dhi[3]
 *                     then one line per box: lo[3] hi[3]
 *   <dir>/params.bin  grhd_params, <dir>/background.bin grhd_background
 *   <dir>/state.bin   pool (doubles)
 * writes <dir>/state_out.bin and <dir>/result.txt
 */
#include "grhd_operator.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

using namespace GRHD;
using namespace GRHD::amd;

template <class T> static T read_struct(const std::string &path)
{
    T v{};
    std::ifstream in(path, std::ios::binary);
    in.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!in)
        throw std::runtime_error("cannot read " + path);
    return v;
}

int main(int argc, char **argv)
{
    if (argc != 5 && argc != 7)
    {
        std::fprintf(stderr,
                     "usage: %s dir cfl max_dt t_end [walk fused]\n",
                     argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const double cfl = std::atof(argv[2]), max_dt = std::atof(argv[3]),
                 t_end = std::atof(argv[4]);
    try
    {
        std::ifstream lv(dir + "/level.txt");
        int ghosts, ncomp, nbox;
        double dx;
        std::array<double, 3> c;
        IntBox dom;
        lv >> ghosts >> ncomp >> nbox >> dx >> c[0] >> c[1] >> c[2];
        for (int d = 0; d < 3; ++d)
            lv >> dom.lo[d];
        for (int d = 0; d < 3; ++d)
            lv >> dom.hi[d];
        std::vector<IntBox> grids(nbox);
        for (auto &b : grids)
            lv >> b.lo[0] >> b.lo[1] >> b.lo[2] >> b.hi[0] >> b.hi[1] >>
                b.hi[2];
        if (!lv)
            throw std::runtime_error("bad level.txt");
        const auto p = read_struct<grhd_params>(dir + "/params.bin");
        const auto bg = read_struct<grhd_background>(dir + "/background.bin");

        AtmosphereOptions atm;
        atm.rho_floor = p.rho_floor;
        atm.pressure_floor = p.pressure_floor;
        atm.max_velocity_squared = p.max_velocity_squared;
        atm.metric_source_atmosphere_factor =
            p.metric_source_atmosphere_factor;
        RecoveryOptions rec;
        rec.max_iterations = p.max_iterations;
        rec.tolerance = p.tolerance;
        rec.pressure_floor = p.recovery_pressure_floor;
        rec.max_velocity_squared = p.recovery_max_velocity_squared;

        double *pool = nullptr;
        DeviceLevelData probe(grids, ghosts, ncomp, nullptr);
        const size_t bytes = sizeof(double) * (size_t)probe.pool_size();
        if (hipMalloc(&pool, bytes) != hipSuccess)
            throw std::runtime_error("hipMalloc failed");
        DeviceLevelData state(grids, ghosts, ncomp, pool);
        std::vector<double> host((size_t)state.pool_size());
        {
            std::ifstream in(dir + "/state.bin", std::ios::binary);
            in.read(reinterpret_cast<char *>(host.data()), bytes);
            if (!in)
                throw std::runtime_error("cannot read state.bin");
        }
        if (hipMemcpy(pool, host.data(), bytes, hipMemcpyHostToDevice) !=
            hipSuccess)
            throw std::runtime_error("upload failed");

        const ProblemDomain domain(dom);
        auto ctx = std::make_shared<Context>();
        if (argc == 7) /* a pinned FP64 face launch */
            ctx->set_face_launch(false, std::atoi(argv[5]),
                                 std::atoi(argv[6]) != 0);
        FixedBGLevelDataFiniteVolumeOperator op(
            IdealGasEOS(p.adiabatic_index), bg, dx, c, atm, p.limiter_theta,
            p.use_reconstruction != 0, rec, p.use_static_metric_sources != 0,
            0, ctx);
        const ConservedSum s0 =
            compute_leveldata_conserved_sum(state, domain, dx, c, 0,
                                            ctx->get());
        double t = 0.0;
        int steps = 0;
        while (t < t_end - 1.0e-14)
        {
            const double dt = std::min(
                std::min(op.compute_stable_dt(state, domain, cfl), max_dt),
                t_end - t);
            op.advance_ssprk2(state, domain, dt);
            t += dt;
            ++steps;
        }
        const ConservedSum s1 =
            compute_leveldata_conserved_sum(state, domain, dx, c, 0,
                                            ctx->get());
        const bool finite = leveldata_conserved_state_is_finite(
            state, domain, dx, c, 0, ctx->get());
        if (hipMemcpy(host.data(), pool, bytes, hipMemcpyDeviceToHost) !=
            hipSuccess)
            throw std::runtime_error("download failed");
        (void)hipFree(pool);
        std::ofstream out(dir + "/state_out.bin", std::ios::binary);
        out.write(reinterpret_cast<const char *>(host.data()), bytes);
        /* a rank that owns none of the level's boxes: the same calls on
         * an empty layout of the same domain do nothing, and its stable
         * dt is the reference's cfl / 1e-14 */
        DeviceLevelData none(std::vector<IntBox>{}, domain, ghosts, ncomp,
                             nullptr);
        const double dt_none = op.compute_stable_dt(none, domain, cfl);
        op.advance_ssprk2(none, domain, max_dt);
        const ConservedSum s_none =
            compute_leveldata_conserved_sum(none, domain, dx, c, 0,
                                            ctx->get());
        std::ofstream res(dir + "/result.txt");
        res.precision(17);
        res << steps << ' ' << t << ' ' << s0.D << ' ' << s1.D << ' '
            << (finite ? 1 : 0) << ' ' << dt_none << ' ' << s_none.D
            << '\n';
        std::printf("level_driver: %d steps to t = %.17g\n", steps, t);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "level_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
