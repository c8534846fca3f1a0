/*
 * grhd_abi.h -- C-ABI of the MI355X fixed-background GRHD operator.
 *
 * One call = one FArrayBox (one box of a level).  Every array argument is a
 * DEVICE pointer into HBM laid out exactly like a Chombo FArrayBox:
 * component-major, Fortran order, i fastest:
 *
 *     fab[c * npts + ((k - glo[2]) * ny + (j - glo[1])) * nx + (i - glo[0])]
 *
 * with nx = ghi[0] - glo[0] + 1 (etc.) and npts = nx * ny * nz.  The eleven
 * GRHD components sit at comp_base + {0..10} in the order of the reference's
 * UserVariables (Tests/GRHDFixedBGTest/UserVariables.hpp:24-40):
 *     D, S1, S2, S3, tau, rho, eps, pressure, v1, v2, v3.
 * The rhs FAB covers the interior box only (the reference allocates it with
 * IntVect::Zero ghosts) and uses the same component numbering.
 *
 * Calls are stream-ordered on `stream` (a hipStream_t, NULL = default
 * stream).  Every entry point that returns a status waits for its own work
 * to finish (the reference is synchronous and reports failures by throwing);
 * grhd_advance_rk4_box runs `nsteps` steps and waits once at the end.
 * The caller owns every buffer; scratch space lives in a grhd_ctx and is
 * reused across calls.  A grhd_ctx must not be used from two host threads at
 * once.  All arithmetic is IEEE binary64 without contraction, following the
 * reference's operation order.
 *
 * Return value: GRHD_OK (0) or one of the GRHD_ERR_* codes below; on a
 * recovery error `err` (if non-NULL) names the first failing cell in the
 * reference's BoxIterator order and the con2prim residual there.  After a
 * recovery error the contents of the state, rhs and mask buffers are
 * unspecified (the reference throws part-way through the same update).
 */
#ifndef GRHD_ABI_H_
#define GRHD_ABI_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of GRHD components starting at comp_base. */
#define GRHD_NUM_COMPS 11
enum
{
    GRHD_D = 0,
    GRHD_S1 = 1,
    GRHD_S2 = 2,
    GRHD_S3 = 3,
    GRHD_TAU = 4,
    GRHD_RHO = 5,
    GRHD_EPS = 6,
    GRHD_PRESSURE = 7,
    GRHD_V1 = 8,
    GRHD_V2 = 9,
    GRHD_V3 = 10
};

/* Status codes.  The recovery codes map onto the std::runtime_error the
 * reference throws at FixedBGFArrayBoxFiniteVolume.hpp:157-164 (ghost fill),
 * :48-55 (RHS primitive recovery), :272-279 (metric sources) and :383-390
 * (CFL speed); GRHD_ERR_INVALID_ARGUMENT maps onto std::invalid_argument
 * (:549-550) and MayDay::Error for bad background parameters
 * (KerrSchild.hpp:41-46, CircularBinaryPN.hpp:52-56). */
enum
{
    GRHD_OK = 0,
    GRHD_ERR_RECOVERY_GHOST_FILL = 1,
    GRHD_ERR_RECOVERY_RHS = 2,
    GRHD_ERR_RECOVERY_SOURCES = 3,
    GRHD_ERR_RECOVERY_CFL = 4,
    GRHD_ERR_INVALID_ARGUMENT = 5,
    GRHD_ERR_DEVICE = 6
};

/* Background kinds (Source/Background/KerrSchild.hpp,
 * Source/Background/CircularBinaryPN.hpp). */
enum
{
    GRHD_BG_KERR_SCHILD = 0,
    GRHD_BG_CIRCULAR_BINARY_PN = 1
};

/* Per-cell recovery mask bits (uint8 per interior cell). */
enum
{
    GRHD_MASK_FAILED = 1,          /* con2prim failed -> atmosphere      */
    GRHD_MASK_FLOORED = 2,         /* floors changed the primitive       */
    GRHD_MASK_RESET = 4,           /* conserved variables were rewritten */
    GRHD_MASK_SOURCE_ATMOSPHERE = 8 /* metric sources skipped (RHS)      */
};

/* Geometry of one box of a level: Chombo's grids[dit] (interior),
 * state[dit].box() (FAB storage), ProblemDomain, and the Coordinates
 * convention x = (i + 1/2) dx - center (GRChombo Coordinates). */
typedef struct grhd_box
{
    int lo[3], hi[3];   /* interior (valid) box, inclusive      */
    int glo[3], ghi[3]; /* FAB storage box, inclusive (ghosted) */
    int dlo[3], dhi[3]; /* ProblemDomain box, inclusive         */
    int periodic[3];    /* ProblemDomain::isPeriodic(dir)       */
    double dx;
    double center[3];
    /* inverse_dx_U of the reference's operator overloads
     * (FixedBGFArrayBoxFiniteVolume.hpp:468-484, :535-589, :702-809):
     * the flux divergence of direction d is scaled by inverse_dx[d]
     * (:243) and the CFL speed sums max|lambda_d| inverse_dx[d] (:407);
     * coordinates still use dx.  All three zero means 1/dx in every
     * direction (the overloads without inverse_dx_U, make_inverse_dx);
     * otherwise each must be positive and finite. */
    double inverse_dx[3];
} grhd_box;

/* IdealGasEOS + AtmosphereOptions + RecoveryOptions + operator flags
 * (IdealGasEOS.hpp:10-53, Atmosphere.hpp:16-22, Valencia.hpp:17-23,
 * FixedBGFArrayBoxFiniteVolume.hpp:450-464 and :681-699). */
typedef struct grhd_params
{
    double adiabatic_index;
    /* AtmosphereOptions */
    double rho_floor;
    double pressure_floor;
    double max_velocity_squared;
    double metric_source_atmosphere_factor;
    /* RecoveryOptions */
    int max_iterations;
    double tolerance;
    double recovery_pressure_floor;
    double recovery_max_velocity_squared;
    /* operator */
    double limiter_theta;
    int use_reconstruction;
    int use_static_metric_sources;
    int use_metric_volume_time_sources;
    double metric_time_derivative_step;
    /* No counterpart in the reference: nonzero puts reconstruction, the
     * face metric and HLLE in FP32 (face fluxes stored as float), and the
     * static-metric sources (metric derivatives, K_ij and the source
     * sums); con2prim, W, the flux-divergence sum, the addition of the
     * sources, the volume-time sources and the RK update stay FP64.  Zero
     * (the default) is the bit-exact FP64 path. */
    int mixed_precision;
    /* KerrSchild::check_if_excised (KerrSchild.hpp:247-280) applied the way
     * ExcisionTest applies it to the matter right-hand side
     * (Tests/KerrBHScalarTest/ExcisionTest.hpp:40-63): nonzero sets the
     * conserved right-hand side of every interior cell with
     * (x^2 + y^2) / (2 M r_+) + z^2 / r_+^2 < excision_outer_buffer or
     * (x^2 + y^2) / (2 M r_-) + z^2 / r_-^2 < excision_inner_buffer to zero,
     * after the fluxes and sources; those cells then keep their state.
     * Kerr-Schild backgrounds only.  Zero (the default) leaves the
     * right-hand side untouched. */
    int excise_horizon;
    double excision_inner_buffer; /* the reference's default: 1.05 */
    double excision_outer_buffer; /* the reference's default: 0.9 */
    /* No counterpart in the reference.  With mixed_precision, a margin
     * m >= 1 computes in FP64 every face adjacent to a cell whose recovered
     * primitives lie within a factor m of a floor or of the velocity cap:
     * rho <= m' rho_floor or P <= m' pressure_floor, with m' the larger of
     * m and metric_source_atmosphere_factor (so the metric-source
     * atmosphere test is covered), or 1 - v^2 <= m (1 -
     * max_velocity_squared); faces of cells whose recovery failed are FP64
     * as well.  The other faces stay FP32.  Such cells sit
     * on the floor and cap thresholds, where a rounding difference flips a
     * reset decision; with their faces in FP64, cells that only such cells
     * reach evolve exactly as on the FP64 path.  Fluxes are then stored as
     * double.  Zero (the default) keeps every face in FP32; other values
     * below 1, and non-finite ones, are invalid. */
    double mixed_fp64_margin;
    /* No counterpart in the reference.  With mixed_precision, s > 0 also
     * computes in FP64 every face adjacent to a cell whose shift has
     * gamma_ij beta^i beta^j >= s^2: in horizon-penetrating coordinates
     * infalling gas moves with the shift, and the flux speed
     * alpha v^i - beta^i is a small difference of large terms that FP32
     * loses.  Zero (the default) disables the test; negative or non-finite
     * values are invalid. */
    double mixed_fp64_shift;
} grhd_params;

/* Analytic background.  For the binary, the orbital time at a call is
 * `time` (the field) plus the call's `time` argument, which is what the
 * reference's time factory does (GRHDFixedBGTest.cpp:435-446). */
typedef struct grhd_background
{
    int kind;
    /* KerrSchild::params_t */
    double mass;
    double spin;
    /* CircularBinaryPN::params_t */
    double mass_1;
    double mass_2;
    double separation;
    double angular_velocity;
    double phase;
    double time;
    double softening_radius;
} grhd_background;

/* AtmosphereResetDiagnostics (FArrayBoxFiniteVolume.hpp:78-92). */
typedef struct grhd_diagnostics
{
    long long num_cells;
    long long num_failed_recoveries;
    long long num_floored_primitives;
    long long num_conserved_resets;
} grhd_diagnostics;

/* Where and why a call failed. */
typedef struct grhd_error
{
    int code;
    int iv[3];
    double residual;
} grhd_error;

/* Opaque scratch/stream context: device scratch buffers, status words,
 * and two side streams on which the y and z face kernels of large calls
 * run beside the x kernel (joined back into the caller's stream). */
typedef struct grhd_ctx grhd_ctx;

/* Create / destroy a context on the current HIP device.  No reference
 * counterpart: the reference allocates its temporaries per call
 * (FixedBGFArrayBoxFiniteVolume.hpp:214-217, :569-573). */
int grhd_ctx_create(grhd_ctx **out);
void grhd_ctx_destroy(grhd_ctx *ctx);

/* RHS of one box: flux divergence over the three directional sweeps plus
 * the static-metric sources (and, if enabled, the metric volume-time
 * sources).  Mutates the physical ghost cells of `state` exactly as the
 * reference does; overwrites the interior of `rhs`.
 * Replaces compute_leveldata_flux_rhs_from_fixed_background
 * (FixedBGFArrayBoxFiniteVolume.hpp:322-363) for one box, as called by
 * FixedBGLevelDataFiniteVolumeOperator::compute_rhs (:468-477) and
 * TimeDependentFixedBGLevelDataFiniteVolumeOperator::compute_rhs (:702-720),
 * including add_metric_volume_time_sources_... (:633-657).  `source_mask`
 * (nullable, uint8 per interior cell) receives GRHD_MASK_SOURCE_ATMOSPHERE
 * where the sources were skipped (:284-285). */
int grhd_compute_rhs_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double time,
                         double *state, int ncomp, int comp_base, double *rhs,
                         unsigned char *source_mask, grhd_error *err,
                         void *stream);

/* Post-update primitive recovery with atmosphere resets over the interior.
 * Replaces recover_primitives_and_apply_atmosphere_from_fixed_background
 * (FixedBGFArrayBoxFiniteVolume.hpp:60-110), as called by
 * recover_primitives (:506-520, :728-743).  `mask` (nullable) receives the
 * GRHD_MASK_FAILED/FLOORED/RESET bits per interior cell. */
int grhd_recover_primitives_box(grhd_ctx *ctx, const grhd_box *box,
                                const grhd_params *params,
                                const grhd_background *background,
                                double time, double *state, int ncomp,
                                int comp_base, unsigned char *mask,
                                grhd_diagnostics *diagnostics, void *stream);

/* state += dt * rhs over the interior conserved components, then (if
 * recover != 0) recovery at recovery_time.  Replaces update_conserved
 * (FixedBGFArrayBoxFiniteVolume.hpp:522-533, :745-757) for one box, built on
 * add_scaled_conserved_rhs (FArrayBoxFiniteVolume.hpp:927-940). */
int grhd_update_conserved_box(grhd_ctx *ctx, const grhd_box *box,
                              const grhd_params *params,
                              const grhd_background *background,
                              double *state, int ncomp, int comp_base,
                              const double *rhs, double dt, int recover,
                              double recovery_time, unsigned char *mask,
                              grhd_diagnostics *diagnostics, void *stream);

/* Maximum over non-atmosphere interior cells of sum_d max|lambda|/dx.
 * Replaces compute_max_inverse_dt_from_fixed_background
 * (FixedBGFArrayBoxFiniteVolume.hpp:364-412). */
int grhd_compute_max_inverse_dt_box(grhd_ctx *ctx, const grhd_box *box,
                                    const grhd_params *params,
                                    const grhd_background *background,
                                    double time, const double *state,
                                    int ncomp, int comp_base,
                                    double *max_inverse_dt, grhd_error *err,
                                    void *stream);

/* cfl / max(max_inverse_dt, 1e-14); GRHD_ERR_INVALID_ARGUMENT if cfl <= 0.
 * Replaces compute_stable_dt (FixedBGFArrayBoxFiniteVolume.hpp:544-555,
 * :769-780). */
int grhd_compute_stable_dt_box(grhd_ctx *ctx, const grhd_box *box,
                               const grhd_params *params,
                               const grhd_background *background, double time,
                               const double *state, int ncomp, int comp_base,
                               double cfl, double *stable_dt, grhd_error *err,
                               void *stream);

/* One SSPRK2 step, the reference's integrator:
 * U1 = U + dt L(U); recover(U1, t+dt); U = (U + U1 + dt L(U1)) / 2;
 * recover(U, t+dt).  Replaces advance_ssprk2
 * (FixedBGFArrayBoxFiniteVolume.hpp:563-589, :782-809).  `diagnostics`
 * (nullable) receives the final recovery's counts, `mask` its bits. */
int grhd_advance_ssprk2_box(grhd_ctx *ctx, const grhd_box *box,
                            const grhd_params *params,
                            const grhd_background *background, double time,
                            double dt, double *state, int ncomp,
                            int comp_base, unsigned char *mask,
                            grhd_diagnostics *diagnostics, grhd_error *err,
                            void *stream);

/* `nsteps` classical RK4 steps of size dt starting at `time`.  Stage states
 * U0 + (dt/2) k1, U0 + (dt/2) k2, U0 + dt k3 and the result
 * ((((U0 + dt/6 k1) + dt/3 k2) + dt/3 k3) + dt/6 k4) are each recovered with
 * atmosphere resets.  No counterpart in the reference, which integrates GRHD
 * only with SSPRK2; it drives the per-stage seam of
 * FixedBGAMRLevelHooks::specific_eval_rhs / specific_update_ode
 * (FixedBGAMRLevelHooks.hpp:40-86) the way GRChombo's RK4 drives
 * GRAMRLevel.  `diagnostics` receives the sum over all recoveries, `mask`
 * the last recovery's bits. */
int grhd_advance_rk4_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double time,
                         double dt, int nsteps, double *state, int ncomp,
                         int comp_base, unsigned char *mask,
                         grhd_diagnostics *diagnostics, grhd_error *err,
                         void *stream);

/* Bytes of device scratch grhd_advance_rk4_box_noalloc may use per GPU:
 * enough for a box of up to 72^3 interior cells with 3 ghost layers
 * (64^3 needs 83 MiB), or 68^3 on the guarded mixed path, whose FP64 face
 * list and guarded-cell sources take room too. */
#define GRHD_NOALLOC_SCRATCH_BYTES (128ull << 20)

/* grhd_advance_rk4_box without a context and without allocating: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Arguments, results and errors are those
 * of grhd_advance_rk4_box (bit for bit the same state, mask and counts).
 * Every scratch buffer (primitives, face fluxes, stage states, the RK
 * accumulator, the status block and the launch tables) lives in one static
 * device array of GRHD_NOALLOC_SCRATCH_BYTES, which each GPU the library
 * runs on holds until the process ends; the call makes no hipMalloc,
 * hipFree or other allocating call.  A box whose scratch would not fit
 * returns GRHD_ERR_INVALID_ARGUMENT, as does a box that is not its own
 * non-periodic ProblemDomain.  All calls on one GPU share the array and it
 * keeps the last call's contents: calls that overlap on one GPU (from two
 * streams or two host threads) race and must be serialised by the
 * caller. */
int grhd_advance_rk4_box_noalloc(const grhd_box *box,
                                 const grhd_params *params,
                                 const grhd_background *background,
                                 double time, double dt, int nsteps,
                                 double *state, int ncomp, int comp_base,
                                 unsigned char *mask,
                                 grhd_diagnostics *diagnostics,
                                 grhd_error *err, void *stream);

/* Analytic cell geometry over the FAB storage box, for checking the device
 * metric: out holds 16 components [g11 g12 g13 g22 g23 g33, gu11 gu12 gu13
 * gu22 gu23 gu33, lapse, beta1 beta2 beta3] in the FAB layout of the
 * storage box.  Replaces load_fixed_background_geometry
 * (FixedBGFArrayBoxFiniteVolume.hpp:15-22) evaluated per cell. */
int grhd_cell_geometry_box(grhd_ctx *ctx, const grhd_box *box,
                           const grhd_background *background, double time,
                           double *out, void *stream);

/* The FP64 division and square root every kernel uses, applied
 * elementwise for checking them against IEEE arithmetic (no counterpart
 * in the reference, whose CPU build uses the IEEE operations directly):
 * quotient[i] = a[i] / b[i] and root[i] = sqrt(a[i]) for n doubles in
 * device memory.  They equal the IEEE results bit for bit where the
 * hardware sequences need no range scaling: a = 0 or |a| >= 2^-969, b
 * normal below 2^1022, exponent(a) - exponent(b) < 768, a normal
 * quotient; square roots of arguments of 2^-767 and above, +-0, +inf,
 * negative and NaN arguments.  Synchronises the stream. */
int grhd_fp64_div_sqrt(const double *a, const double *b, double *quotient,
                       double *root, long long n, void *stream);



/* ------------------------------------------------------------------ */
/* Levels: every box of an AMR level in one launch per kernel.        */
/* ------------------------------------------------------------------ */

/* One box of a level and where its FArrayBox starts in the level's state
 * pool (in doubles).  The pool holds each box's FAB (ncomp components
 * over its storage box) at its offset.
 *
 * Every grhd_level_* call takes the boxes this process owns, `nbox` of
 * them, and nbox may be 0 (a rank of a distributed level can own none;
 * the reference's loops over its DataIterator then do nothing).  Such a
 * call checks its parameters, launches nothing and returns GRHD_OK with
 * an empty level's outputs: zero diagnostics, a maximum 1/dt of 0 (the
 * reference's local maximum before MPI_Allreduce), zero sums, a finite
 * state, and a packed checkpoint of zeros; `boxes` and the pools may be
 * NULL. */
typedef struct grhd_level_box
{
    grhd_box box;
    long long state_offset;
} grhd_level_box;

/* Level RHS: LevelData exchange() of internal ghosts (cells of the
 * ProblemDomain -- periodic directions wrapped -- that another box of the
 * level owns), outflow fill of physical ghosts, then the RHS of every box.
 * rhs_pool holds the boxes' interior FABs back to back in box order
 * (ncomp x interior cells each); source_mask (nullable) one byte per
 * interior cell, boxes back to back.  Replaces
 * compute_leveldata_flux_rhs_from_fixed_background
 * (FixedBGFArrayBoxFiniteVolume.hpp:322-363) with its exchange() calls
 * (:310, :355) and the DataIterator loops over the level's boxes.
 * GRHD_ERR_INVALID_ARGUMENT if a ghost inside the domain has no owner (a
 * coarse-fine boundary, which needs interpolation from the coarser level). */
int grhd_level_compute_rhs(grhd_ctx *ctx, const grhd_level_box *boxes,
                           int nbox, const grhd_params *params,
                           const grhd_background *background, double time,
                           double *state_pool, int ncomp, int comp_base,
                           double *rhs_pool, unsigned char *source_mask,
                           grhd_error *err, void *stream);

/* add_conserved_rhs_to (FixedBGFArrayBoxFiniteVolume.hpp:484-496): the
 * level rhs of grhd_level_compute_rhs (exchange, ghost fill -- which
 * mutates the state as there -- fluxes and sources) added to the D, S_i
 * and tau components of rhs_pool, rhs += hydro_rhs cell by cell
 * (add_conserved_rhs, FArrayBoxFiniteVolume.hpp:912-925); the other
 * components of rhs_pool are left as they are.  For a caller (a CCZ4 or
 * matter rhs) that sums several right-hand sides into one. */
int grhd_level_add_conserved_rhs(grhd_ctx *ctx, const grhd_level_box *boxes,
                                 int nbox, const grhd_params *params,
                                 const grhd_background *background,
                                 double time, double *state_pool, int ncomp,
                                 int comp_base, double *rhs_pool,
                                 grhd_error *err, void *stream);

/* recover_primitives over the level (FixedBGFArrayBoxFiniteVolume.hpp:506-520,
 * :728-743); diagnostics are level totals, as after the reference's
 * MPI_Allreduce (FArrayBoxFiniteVolume.hpp:94-116). */
int grhd_level_recover_primitives(grhd_ctx *ctx, const grhd_level_box *boxes,
                                  int nbox, const grhd_params *params,
                                  const grhd_background *background,
                                  double time, double *state_pool, int ncomp,
                                  int comp_base, unsigned char *mask,
                                  grhd_diagnostics *diagnostics, void *stream);

/* update_conserved over the level (FixedBGFArrayBoxFiniteVolume.hpp:522-532,
 * :745-758): add_scaled_conserved_rhs (FArrayBoxFiniteVolume.hpp:927-940)
 * of the rhs pool into every box, then, if `recover`, the post-update
 * recovery at `recovery_time` with level-total diagnostics. */
int grhd_level_update_conserved(grhd_ctx *ctx, const grhd_level_box *boxes,
                                int nbox, const grhd_params *params,
                                const grhd_background *background,
                                double *state_pool, int ncomp, int comp_base,
                                const double *rhs_pool, double dt, int recover,
                                double recovery_time, unsigned char *mask,
                                grhd_diagnostics *diagnostics, void *stream);

/* compute_leveldata_max_inverse_dt_from_fixed_background
 * (FixedBGFArrayBoxFiniteVolume.hpp:414-444): maximum over the level. */
int grhd_level_compute_max_inverse_dt(grhd_ctx *ctx,
                                      const grhd_level_box *boxes, int nbox,
                                      const grhd_params *params,
                                      const grhd_background *background,
                                      double time, const double *state_pool,
                                      int ncomp, int comp_base,
                                      double *max_inverse_dt, grhd_error *err,
                                      void *stream);

/* compute_leveldata_conserved_sum (LevelDataDriver.hpp:22-57): sum[0..4] =
 * sums of D, S_1, S_2, S_3, tau over every interior cell of the level.  The
 * reference adds sequentially in box and BoxIterator order; this adds in a
 * fixed order in double-double arithmetic and rounds once, so the result is
 * deterministic and within (cells - 1) eps sum|x| of the reference's (the
 * bound of its own sequential rounding).  If any value is not finite the
 * plain floating-point sum is returned. */
int grhd_level_conserved_sum(grhd_ctx *ctx, const grhd_level_box *boxes,
                             int nbox, const double *state_pool, int ncomp,
                             int comp_base, double *sum, void *stream);

/* leveldata_conserved_state_is_finite + conserved_is_finite
 * (LevelDataDriver.hpp:59-92):
 * *is_finite = 1 if D, S_i and tau are finite in every interior cell of
 * the level, else 0. */
int grhd_level_conserved_state_is_finite(grhd_ctx *ctx,
                                         const grhd_level_box *boxes,
                                         int nbox, const double *state_pool,
                                         int ncomp, int comp_base,
                                         int *is_finite, void *stream);

/* Discrete adjoint of `nsteps` RK4 steps of size dt of one box (the steps
 * of grhd_advance_rk4_box, starting at time 0) for the objective
 *   J = sum_{n = first_step}^{nsteps} dt * mdot(U_n),
 *   mdot(U) = -dx^2 sum over interior cells with ||x| - r0| < dx/2 of
 *             sqrt(gamma) D (alpha v^i - beta^i) x_i / |x|,
 * the rest-mass flux through the coordinate sphere |x| = r0 = shell_radius
 * (v from con2prim of U_n).  No counterpart in the reference.  Returns J
 * (host), dJ/dU_0 (device, 5 x interior cells: D, S1, S2, S3, tau,
 * component-major over the interior box) and dJ/dGamma, dJ/da (host,
 * 2 doubles) at fixed U_0.  Limiter, floor, HLLE wave-speed, atmosphere and
 * reset branches are frozen at the forward pass; con2prim is differentiated
 * at its converged root by the implicit-function theorem.  The forward
 * state is checkpointed every `checkpoint_every` steps and each segment is
 * recomputed in the reverse sweep.  With dJ_dconserved or dJ_dparams NULL
 * only the forward sweep runs and J is returned.  `state` is not
 * modified.  Horizon excision (grhd_params.excise_horizon) is honoured in
 * both sweeps; near a black hole it keeps J differentiable, since without
 * it recoveries that fail inside the horizon reset cells on either side
 * of any perturbation.  Scope:
 * Kerr-Schild background, box == ProblemDomain (outflow ghosts), no
 * volume-time sources, FP64 faces; otherwise GRHD_ERR_INVALID_ARGUMENT. */
int grhd_rk4_adjoint_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double dt,
                         int nsteps, const double *state, int ncomp,
                         int comp_base, double shell_radius, int first_step,
                         int checkpoint_every, double *objective,
                         double *dJ_dconserved, double *dJ_dparams,
                         grhd_error *err, void *stream);

/* KITPCHK1 checkpoint payload <-> level pool, on the device.  The payload
 * covers the ProblemDomain [domain_lo, domain_hi] grown by `ghosts`, point
 * major with the ncomp components of a cell adjacent (ncomp = NUM_VARS):
 *   values[(((k - lo2) * ny + (j - lo1)) * nx + (i - lo0)) * ncomp + c].
 * Pack zeroes the payload, then writes each storage cell whose index clamped
 * to the domain lies in its box's interior (write_checkpoint,
 * Main_GRHDLevelDataBinaryPNKITPDisk.cpp:1385-1425); unpack fills every
 * storage cell of every box (read_checkpoint_state, :1516-1586), so a
 * checkpoint written with one box layout restores another.  Each box's
 * storage must lie inside the grown domain.  The header and file I/O stay
 * on the host with the caller. */
int grhd_level_checkpoint_pack(grhd_ctx *ctx, const grhd_level_box *boxes,
                               int nbox, const double *state_pool, int ncomp,
                               const int *domain_lo, const int *domain_hi,
                               int ghosts, double *checkpoint_values,
                               void *stream);
int grhd_level_checkpoint_unpack(grhd_ctx *ctx, const grhd_level_box *boxes,
                                 int nbox, const double *checkpoint_values,
                                 int ncomp, const int *domain_lo,
                                 const int *domain_hi, int ghosts,
                                 double *state_pool, void *stream);

/* advance_ssprk2 of a whole level (FixedBGFArrayBoxFiniteVolume.hpp:563-589,
 * :782-809), exchanging internal ghosts before each RHS. */
int grhd_level_advance_ssprk2(grhd_ctx *ctx, const grhd_level_box *boxes,
                              int nbox, const grhd_params *params,
                              const grhd_background *background, double time,
                              double dt, double *state_pool, int ncomp,
                              int comp_base, unsigned char *mask,
                              grhd_diagnostics *diagnostics, grhd_error *err,
                              void *stream);

/* `nsteps` RK4 steps of a whole level, as grhd_advance_rk4_box does for
 * one box, exchanging internal ghosts before each stage's RHS: the level
 * RK driver launches per level, not per box.  No counterpart in the
 * reference (it steps GRHD with SSPRK2 only). */
int grhd_level_advance_rk4(grhd_ctx *ctx, const grhd_level_box *boxes,
                           int nbox, const grhd_params *params,
                           const grhd_background *background, double time,
                           double dt, int nsteps, double *state_pool,
                           int ncomp, int comp_base, unsigned char *mask,
                           grhd_diagnostics *diagnostics, grhd_error *err,
                           void *stream);

/* ------------------------------------------------------------------ */
/* Instrumentation (no counterpart in the reference).                 */
/* Per-kernel-class timing of every call made on `ctx` while enabled: */
/* a HIP event is recorded on the call's stream after each launch and */
/* the gaps are summed per class when the call waits.                 */
/* ------------------------------------------------------------------ */
enum
{
    GRHD_KERNEL_PRIMS = 0, /* ghost fill + con2prim over the storage box */
    GRHD_KERNEL_FACES = 1, /* MUSCL-HLLE faces: the three directions run as
                              one launch or concurrently and are timed
                              together                                   */
    GRHD_KERNEL_CELLS = 2, /* divergence + sources + stage update         */
    GRHD_KERNEL_OTHER = 3,
    GRHD_NUM_KERNEL_CLASSES = 4
};

/* Enable (1) or disable (0) timing; either resets the accumulators. */
int grhd_ctx_set_timing(grhd_ctx *ctx, int enable);

/* Accumulated milliseconds and launch counts per kernel class
 * (arrays of GRHD_NUM_KERNEL_CLASSES). */
int grhd_ctx_get_timing(grhd_ctx *ctx, double *total_ms,
                        long long *launches);

/* ------------------------------------------------------------------ */
/* Face-launch tuning (no counterpart in the reference: its CPU loops */
/* have no launch shape to choose).  The y and z face kernels walk a  */
/* run of 1, 2, 4 or 8 faces per thread, or form faces in tiles of    */
/* 8 x 8 cells (one cell per lane), and the three face directions run */
/* as one fused launch or as three kernels on concurrent streams.  By */
/* default size rules choose both: tiles and one fused launch below   */
/* 3 x 2^20 interior cells, walks of 8 on streams above.  With tuning */
/* on, the first call on a new level (a new box list) or in a new     */
/* face precision times all ten choices on that call's own primitives */
/* and keeps the fastest for the plan.  A launch can also be pinned.  */
/* The fluxes, and so every result, are the same under every choice.  */
/* A tuning call waits on the host once per choice.  The guarded      */
/* mixed path's two face passes keep the size rules.                 */
/* ------------------------------------------------------------------ */

/* Enable (1) or disable (0) tuning; disabling returns to the rules. */
int grhd_ctx_set_autotune(grhd_ctx *ctx, int enable);

/* Pin the face launch of this context's later calls in one face
 * precision (mixed = 0: FP64 faces, 1: FP32 faces): walk 1, 2, 4 or 8
 * faces per thread, or 0 for the tiles, and fused (1) or three launches
 * on streams (0).  A pinned launch comes before tuning; walk = -1 unpins
 * (back to tuning or the rules).  For a launch measured once offline, or
 * for checking that every launch gives the same results. */
int grhd_ctx_set_face_launch(grhd_ctx *ctx, int mixed, int walk,
                             int fused);

/* The face launch the current level uses for FP64 (mixed = 0) or FP32
 * (mixed = 1) faces: walk length (0 for the tiles), fused (1) or not (0),
 * whether it was measured (1) or is the rule's or a pinned one (0), and
 * the measured milliseconds per face stage (0 otherwise).  Any output
 * pointer may be NULL.  Returns GRHD_ERR_INVALID_ARGUMENT before the
 * first level call. */
int grhd_ctx_get_face_launch(grhd_ctx *ctx, int mixed, int *walk,
                             int *fused, int *tuned, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* GRHD_ABI_H_ */
