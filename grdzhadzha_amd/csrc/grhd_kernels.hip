/*
 * grhd_kernels.hip -- CDNA4 kernels and the C-ABI (include/grhd_abi.h) of
 * the fixed-background GRHD operator.
 *
 * Work decomposition of one right-hand side (the reference's three
 * directional sweeps, FixedBGFArrayBoxFiniteVolume.hpp:199-363, with every
 * duplicated quantity computed once):
 *
 *   k_prims      one thread per storage cell: outflow ghost fill
 *                (:113-177) and con2prim over the ghosted box (:33-58)
 *                -> primitives W (SoA, 6 x npts) + failure byte
 *   k_faces<d>   one thread per face of direction d: MC/minmod slopes of
 *                the four cells around the face, face metric = mean of the
 *                two cell metrics (:179-197), floors, prim2con, HLLE
 *                -> face flux F_d (SoA, 5 x nfaces); each face once
 *   k_cells<M>   one thread per interior cell: flux divergence in the
 *                reference's summation order, static-metric sources
 *                (:255-298) from one full analytic metric evaluation and
 *                the cell's W, volume-time sources (:616-657), then per
 *                mode: write rhs, or an RK4 / SSPRK2 stage update fused
 *                with the post-update recovery (:60-110) that also
 *                produces W of the next stage
 *
 * Memory is the FArrayBox layout (component-major, i fastest), so every
 * load and store of a component is coalesced along i.  Reductions (reset
 * counts, max 1/dt) are wavefront-64 shuffles, one LDS step across the
 * waves of a block and one atomic per block.  No MFMA: nothing here is a
 * dense contraction.  Compiled with -ffp-contract=off.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/grhd_abi.h"
#include "grhd_physics.hpp"

using namespace grhd;

namespace
{

constexpr int kBlock = 256;
constexpr unsigned long long kNoFail = ~0ull;
constexpr int kLinBits = 40;

/* Device-side status block of one call. */
struct DevStatus
{
    unsigned long long fail_key; /* (eval << 41) | (phase << 40) | index */
    unsigned long long cfl_key;  /* interior linear index */
    unsigned long long counts[4];
    unsigned long long max_bits;
};

/* Box geometry as seen by the kernels. */
struct BoxDev
{
    int lo[3], hi[3];
    int glo[3], ghi[3];
    int dlo[3], dhi[3];
    int periodic[3];
    int gn[3];
    long long gnpts;
    int in[3];
    long long inpts;
    double dx;
    double center[3];
};

__device__ __forceinline__ long long goff(const BoxDev &b, int i, int j, int k)
{
    return ((long long)(k - b.glo[2]) * b.gn[1] + (j - b.glo[1])) * b.gn[0] +
           (i - b.glo[0]);
}

__device__ __forceinline__ Cons load_cons(const double *U, long long n,
                                          long long o)
{
    Cons c;
    c.D = U[o];
    c.S[0] = U[n + o];
    c.S[1] = U[2 * n + o];
    c.S[2] = U[3 * n + o];
    c.tau = U[4 * n + o];
    return c;
}

__device__ __forceinline__ void store_cons(double *U, long long n, long long o,
                                           const Cons &c)
{
    U[o] = c.D;
    U[n + o] = c.S[0];
    U[2 * n + o] = c.S[1];
    U[3 * n + o] = c.S[2];
    U[4 * n + o] = c.tau;
}

/* primitives: rho, eps, pressure, v1, v2, v3 */
__device__ __forceinline__ Prim load_prim(const double *P, long long n,
                                          long long o)
{
    Prim p;
    p.rho = P[o];
    p.eps = P[n + o];
    p.pressure = P[2 * n + o];
    p.v[0] = P[3 * n + o];
    p.v[1] = P[4 * n + o];
    p.v[2] = P[5 * n + o];
    return p;
}

__device__ __forceinline__ void store_prim(double *P, long long n, long long o,
                                           const Prim &p)
{
    P[o] = p.rho;
    P[n + o] = p.eps;
    P[2 * n + o] = p.pressure;
    P[3 * n + o] = p.v[0];
    P[4 * n + o] = p.v[1];
    P[5 * n + o] = p.v[2];
}

__device__ __forceinline__ unsigned long long
fail_key(unsigned long long eval, unsigned long long phase,
         unsigned long long idx)
{
    return (eval << (kLinBits + 1)) | (phase << kLinBits) | idx;
}

/* wavefront-64 + block reductions */
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

__device__ void block_add_counts(DevStatus *st, unsigned long long c0,
                                 unsigned long long c1, unsigned long long c2,
                                 unsigned long long c3)
{
    __shared__ unsigned long long part[4][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    c0 = wave_sum(c0);
    c1 = wave_sum(c1);
    c2 = wave_sum(c2);
    c3 = wave_sum(c3);
    if (lane == 0)
    {
        part[0][wave] = c0;
        part[1][wave] = c1;
        part[2][wave] = c2;
        part[3][wave] = c3;
    }
    __syncthreads();
    if (threadIdx.x < 4)
    {
        unsigned long long s = 0;
        for (int w = 0; w < kBlock / 64; ++w)
            s += part[threadIdx.x][w];
        if (s)
            atomicAdd(&st->counts[threadIdx.x], s);
    }
}

/* The e-th storage cell outside the interior box: whole z-planes below
 * and above the interior, then per interior plane the rows below and
 * above the interior in y, then the x-ghost cells of interior rows. */
__device__ __forceinline__ void exterior_cell(const BoxDev &b, long long e,
                                              int &i, int &j, int &k)
{
    const long long plane = (long long)b.gn[0] * b.gn[1];
    const long long zlo = (long long)(b.lo[2] - b.glo[2]) * plane;
    const long long zhi = (long long)(b.ghi[2] - b.hi[2]) * plane;
    if (e < zlo + zhi)
    {
        const long long q = (e < zlo) ? e : e - zlo;
        k = ((e < zlo) ? b.glo[2] : b.hi[2] + 1) + (int)(q / plane);
        const long long r = q % plane;
        j = b.glo[1] + (int)(r / b.gn[0]);
        i = b.glo[0] + (int)(r % b.gn[0]);
        return;
    }
    e -= zlo + zhi;
    const long long per_plane = plane - (long long)b.in[0] * b.in[1];
    k = b.lo[2] + (int)(e / per_plane);
    long long r = e % per_plane;
    const long long ylo = (long long)(b.lo[1] - b.glo[1]) * b.gn[0];
    const long long yhi = (long long)(b.ghi[1] - b.hi[1]) * b.gn[0];
    if (r < ylo + yhi)
    {
        const long long q = (r < ylo) ? r : r - ylo;
        j = ((r < ylo) ? b.glo[1] : b.hi[1] + 1) + (int)(q / b.gn[0]);
        i = b.glo[0] + (int)(q % b.gn[0]);
        return;
    }
    r -= ylo + yhi;
    const int gx = b.lo[0] - b.glo[0];
    const int per_row = b.gn[0] - b.in[0];
    j = b.lo[1] + (int)(r / per_row);
    const int q = (int)(r % per_row);
    i = (q < gx) ? b.glo[0] + q : b.hi[0] + 1 + (q - gx);
}

__global__ void k_status_init(DevStatus *st)
{
    if (threadIdx.x == 0)
    {
        st->fail_key = kNoFail;
        st->cfl_key = kNoFail;
        st->counts[0] = st->counts[1] = st->counts[2] = st->counts[3] = 0;
        st->max_bits = 0;
    }
}

/* ------------------------------------------------------------------ */
/* k_prims: ghost fill + con2prim over the storage box.               */
/* EXTERIOR_ONLY: interior W already holds con2prim of the interior   */
/* (written by the previous k_cells); only cells outside the interior */
/* box are processed, and all of them are physical ghosts.            */
/* ------------------------------------------------------------------ */
template <bool EXTERIOR_ONLY>
__global__ void __launch_bounds__(kBlock)
    k_prims(BoxDev b, Options o, Background bg, double *U, double *P,
            double *W, unsigned char *wbad, DevStatus *st,
            unsigned long long eval)
{
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (EXTERIOR_ONLY)
    {
        if (tid >= b.gnpts - b.inpts)
            return;
        exterior_cell(b, tid, i, j, k);
    }
    else
    {
        if (tid >= b.gnpts)
            return;
        i = b.glo[0] + (int)(tid % b.gn[0]);
        j = b.glo[1] + (int)((tid / b.gn[0]) % b.gn[1]);
        k = b.glo[2] + (int)(tid / ((long long)b.gn[0] * b.gn[1]));
    }
    const long long idx = goff(b, i, j, k);
    const int iv[3] = {i, j, k};
    int src[3] = {i, j, k};
    bool physical = false;
    bool interior = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        if (iv[d] < b.lo[d] || iv[d] > b.hi[d])
            interior = false;
        if (iv[d] < b.dlo[d])
        {
            if (!b.periodic[d])
            {
                physical = true;
                src[d] = b.lo[d];
            }
        }
        else if (iv[d] > b.dhi[d])
        {
            if (!b.periodic[d])
            {
                physical = true;
                src[d] = b.hi[d];
            }
        }
    }
    if (EXTERIOR_ONLY && interior)
        return;

    const long long n = b.gnpts;
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    Geo gt;
    cell_geometry(bg, x, gt);
    Cons c;
    if (physical)
    {
        const long long so = goff(b, src[0], src[1], src[2]);
        Prim p;
        bool src_ok;
        double res = 0.0;
        if (EXTERIOR_ONLY)
        {
            src_ok = wbad[so] == 0;
            if (src_ok)
                p = load_prim(W, n, so);
            else
                res = W[so];
        }
        else
        {
            double xs[3];
            coords(src[0], src[1], src[2], b.dx, b.center, xs);
            Geo gs;
            cell_geometry(bg, xs, gs);
            const Cons cs = load_cons(U, n, so);
            src_ok = con2prim(cs, o, gs.gu, p, res);
        }
        if (!src_ok)
        {
            atomicMin(&st->fail_key, fail_key(eval, 0, (unsigned long long)idx));
            W[idx] = res;
            wbad[idx] = 1;
            return;
        }
        floors(p, o, gt.g);
        if (P)
            store_prim(P, n, idx, p);
        prim2con(p, o.gamma, gt.g, c);
        store_cons(U, n, idx, c);
    }
    else
    {
        c = load_cons(U, n, idx);
    }
    Prim w;
    double res;
    if (!con2prim(c, o, gt.gu, w, res))
    {
        atomicMin(&st->fail_key, fail_key(eval, 1, (unsigned long long)idx));
        W[idx] = res;
        wbad[idx] = 1;
        return;
    }
    store_prim(W, n, idx, w);
    wbad[idx] = 0;
}

/* ------------------------------------------------------------------ */
/* k_faces<D>: MUSCL-HLLE fluxes of direction D, marching.            */
/* Face box: interior box with hi[D] + 1; face f sits between cells   */
/* f - e_D (left) and f (right).  Each thread walks kSeg consecutive  */
/* faces along D, carrying the right cell's primitives, the next      */
/* face's reconstructed left state and the right cell's metric, so a  */
/* cell's slope and metric are computed once per walk instead of once */
/* per adjacent face.  Lanes of a wavefront are consecutive along the */
/* fastest other axis (i for D = 1, 2; j for D = 0).                  */
/* ------------------------------------------------------------------ */
constexpr int kSeg = 8;

__device__ __forceinline__ void slope_step(const Prim &l, const Prim &c,
                                           const Prim &r, double theta,
                                           Prim &minus_half, Prim &plus_half)
{
    /* add_scaled(c, slope, -0.5) and add_scaled(c, slope, +0.5)
     * (Reconstruction.hpp:59-69) of the MC-theta slope of c */
    double s = limited_slope(l.rho, c.rho, r.rho, theta);
    minus_half.rho = c.rho + -0.5 * s;
    plus_half.rho = c.rho + 0.5 * s;
    s = limited_slope(l.eps, c.eps, r.eps, theta);
    minus_half.eps = c.eps + -0.5 * s;
    plus_half.eps = c.eps + 0.5 * s;
    s = limited_slope(l.pressure, c.pressure, r.pressure, theta);
    minus_half.pressure = c.pressure + -0.5 * s;
    plus_half.pressure = c.pressure + 0.5 * s;
#pragma unroll
    for (int m = 0; m < 3; ++m)
    {
        s = limited_slope(l.v[m], c.v[m], r.v[m], theta);
        minus_half.v[m] = c.v[m] + -0.5 * s;
        plus_half.v[m] = c.v[m] + 0.5 * s;
    }
}

template <int D>
__global__ void __launch_bounds__(kBlock)
    k_faces(BoxDev b, Options o, Background bg, const double *W, double *F,
            long long nfaces, int nseg)
{
    constexpr int A = (D == 0) ? 1 : 0; /* fastest other axis */
    constexpr int B = (D == 2) ? 1 : 2;
    const int na = b.in[A], nb = b.in[B];
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)na * nb * nseg)
        return;
    const int ia = (int)(idx % na);
    const int ib = (int)((idx / na) % nb);
    const int sg = (int)(idx / ((long long)na * nb));
    const int fnd = b.in[D] + 1;
    const int p0 = sg * kSeg;
    const int p1 = min(p0 + kSeg, fnd);
    const long long n = b.gnpts;
    const long long stride =
        (D == 0) ? 1 : ((D == 1) ? (long long)b.gn[0]
                                 : (long long)b.gn[0] * b.gn[1]);
    int c[3];
    c[A] = b.lo[A] + ia;
    c[B] = b.lo[B] + ib;
    c[D] = b.lo[D] + p0; /* right cell of the first face */
    long long oR = goff(b, c[0], c[1], c[2]);
    const int fn0 = b.in[0] + (D == 0), fn1 = b.in[1] + (D == 1);
    const long long f_stride =
        (D == 0) ? 1 : ((D == 1) ? (long long)fn0 : (long long)fn0 * fn1);
    long long fo = ((long long)(c[2] - b.lo[2]) * fn1 + (c[1] - b.lo[1])) *
                       fn0 + (c[0] - b.lo[0]);

    Prim wL = load_prim(W, n, oR - stride);
    Prim wR = load_prim(W, n, oR);
    Prim lp = wL;
    if (o.use_reconstruction)
    {
        const Prim wLL = load_prim(W, n, oR - 2 * stride);
        Prim unused;
        slope_step(wLL, wL, wR, o.theta, unused, lp);
    }
    double x[3];
    c[D] -= 1;
    coords(c[0], c[1], c[2], b.dx, b.center, x);
    Geo gL;
    cell_geometry(bg, x, gL);
    for (int p = p0; p < p1; ++p)
    {
        c[D] += 1;
        Prim rp = wR, lp_next = wR;
        Prim wRR;
        if (o.use_reconstruction)
        {
            wRR = load_prim(W, n, oR + stride);
            slope_step(wL, wR, wRR, o.theta, rp, lp_next);
        }
        coords(c[0], c[1], c[2], b.dx, b.center, x);
        Geo gR, face;
        cell_geometry(bg, x, gR);
        average_geo(gL, gR, face);
        Cons f;
        hlle_face(lp, rp, o, face, D, f);
        F[fo] = f.D;
        F[nfaces + fo] = f.S[0];
        F[2 * nfaces + fo] = f.S[1];
        F[3 * nfaces + fo] = f.S[2];
        F[4 * nfaces + fo] = f.tau;
        /* advance one cell along D */
        if (p + 1 < p1)
        {
            lp = lp_next;
            wL = wR;
            wR = o.use_reconstruction ? wRR : load_prim(W, n, oR + stride);
            gL = gR;
            oR += stride;
            fo += f_stride;
        }
    }
}

/* ------------------------------------------------------------------ */
/* k_faces_x: MUSCL-HLLE fluxes of direction 0, one lane per cell.    */
/* The cells lo-1 .. hi+1 of every (j, k) row are laid end to end; a  */
/* wavefront takes 64 consecutive ones, overlapping the next wave by  */
/* one.  Each lane evaluates its cell's metric and the two half-cell  */
/* reconstructed states once; the face between a lane and the next is */
/* formed with __shfl_down, so loads are unit-stride along i and no   */
/* cell metric is evaluated twice within a wave.                      */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ double shfl_next(double v)
{
    return __shfl_down(v, 1, 64);
}

__global__ void __launch_bounds__(kBlock)
    k_faces_x(BoxDev b, Options o, Background bg, const double *W, double *F,
              long long nfaces, long long next)
{
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long e = wave * 63 + lane;
    const int E = b.in[0] + 2;
    const bool live = e < next;
    const long long ee = live ? e : next - 1;
    const int x = (int)(ee % E);
    const long long row = ee / E;
    const int j = b.lo[1] + (int)(row % b.in[1]);
    const int k = b.lo[2] + (int)(row / b.in[1]);
    const int i = b.lo[0] - 1 + x;
    const long long n = b.gnpts;
    const long long oc = goff(b, i, j, k);
    const Prim wc = load_prim(W, n, oc);
    Prim mh = wc, ph = wc; /* c - 1/2 (right state of the face below),
                              c + 1/2 (left state of the face above) */
    if (o.use_reconstruction)
    {
        const Prim wl = load_prim(W, n, oc - 1);
        const Prim wr = load_prim(W, n, oc + 1);
        slope_step(wl, wc, wr, o.theta, mh, ph);
    }
    double xc[3];
    coords(i, j, k, b.dx, b.center, xc);
    Geo gc;
    cell_geometry(bg, xc, gc);

    /* neighbour (lane + 1): its minus-half state and metric */
    Prim rp;
    rp.rho = shfl_next(mh.rho);
    rp.eps = shfl_next(mh.eps);
    rp.pressure = shfl_next(mh.pressure);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        rp.v[m] = shfl_next(mh.v[m]);
    Geo gn;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            gn.g[a][c] = shfl_next(gc.g[a][c]);
            gn.gu[a][c] = shfl_next(gc.gu[a][c]);
        }
    gn.lapse = shfl_next(gc.lapse);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        gn.shift[m] = shfl_next(gc.shift[m]);

    /* face between this cell and the next one of the same row */
    if (!live || lane == 63 || x == E - 1 || e + 1 >= next)
        return;
    Geo face;
    average_geo(gc, gn, face);
    Cons f;
    hlle_face(ph, rp, o, face, 0, f);
    const int fn0 = b.in[0] + 1;
    const long long fo =
        ((long long)(k - b.lo[2]) * b.in[1] + (j - b.lo[1])) * fn0 + x;
    F[fo] = f.D;
    F[nfaces + fo] = f.S[0];
    F[2 * nfaces + fo] = f.S[1];
    F[3 * nfaces + fo] = f.S[2];
    F[4 * nfaces + fo] = f.tau;
}

enum CellMode
{
    MODE_RHS = 0,
    MODE_RK4_1 = 1,
    MODE_RK4_2 = 2,
    MODE_RK4_3 = 3,
    MODE_RK4_4 = 4,
    MODE_SSP_1 = 5,
    MODE_SSP_2 = 6
};

struct CellArgs
{
    const double *F0, *F1, *F2;
    long long nf0, nf1, nf2;
    const double *Us;   /* state whose RHS is evaluated (5 comps, storage) */
    const double *U0;   /* RK: step start state / SSP: old state */
    double *acc;        /* RK accumulator (5 comps, interior) */
    double *Unext;      /* output state (5 comps, storage) */
    double *Pout;       /* primitive slots to store (final stage) or null */
    double *rhs;        /* MODE_RHS: rhs FAB comp D (interior layout) */
    unsigned char *mask;
    double *W;
    unsigned char *wbad;
    DevStatus *st;
    double c_stage;     /* U0 coefficient of k */
    double c_acc;       /* accumulator coefficient of k */
    double dt;
    double inv_dx;
    double vol_h;       /* max(step, 1e-12) */
    unsigned long long next_eval;
    int count;          /* add recovery counts */
};

template <int MODE>
__global__ void __launch_bounds__(kBlock)
    k_cells(BoxDev b, Options o, Background bg_s, Background bg_next,
            Background bg_plus, Background bg_minus, CellArgs a)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt[4] = {0, 0, 0, 0};
    if (idx < b.inpts)
    {
        const int ii = (int)(idx % b.in[0]);
        const int jj = (int)((idx / b.in[0]) % b.in[1]);
        const int kk = (int)(idx / ((long long)b.in[0] * b.in[1]));
        const int i = b.lo[0] + ii, j = b.lo[1] + jj, k = b.lo[2] + kk;
        const long long n = b.gnpts;
        const long long so = goff(b, i, j, k);

        /* flux divergence: rhs = ((0 + r_x) + r_y) + r_z */
        Cons rhs;
        rhs.D = 0.0;
        rhs.tau = 0.0;
        rhs.S[0] = rhs.S[1] = rhs.S[2] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double *F = (d == 0) ? a.F0 : ((d == 1) ? a.F1 : a.F2);
            const long long nf = (d == 0) ? a.nf0 : ((d == 1) ? a.nf1 : a.nf2);
            const int fn0 = b.in[0] + (d == 0), fn1 = b.in[1] + (d == 1);
            const long long flo =
                ((long long)kk * fn1 + jj) * fn0 + ii;
            const long long fhi =
                flo + ((d == 0) ? 1
                                : ((d == 1) ? (long long)fn0
                                            : (long long)fn0 * fn1));
            const double mi = -a.inv_dx;
            rhs.D += mi * (F[fhi] - F[flo]);
            rhs.tau += mi * (F[4 * nf + fhi] - F[4 * nf + flo]);
#pragma unroll
            for (int m = 0; m < 3; ++m)
                rhs.S[m] += mi * (F[(1 + m) * nf + fhi] - F[(1 + m) * nf + flo]);
        }

        const Cons us = load_cons(a.Us, n, so);
        double x[3];
        coords(i, j, k, b.dx, b.center, x);
        unsigned char mbits = 0;
        Geo geo_s;
        bool have_geo = false;
        if (o.use_static_sources)
        {
            /* geometry first; the derivative half of the metric only for
             * cells whose sources are not skipped (:284-288) */
            GeoCore core;
            cell_geometry_core(bg_s, x, geo_s, core);
            have_geo = true;
            Prim p = load_prim(a.W, n, so);
            floors(p, o, geo_s.g);
            if (is_atmosphere(p, o, o.source_atmosphere_factor))
            {
                mbits |= GRHD_MASK_SOURCE_ATMOSPHERE;
            }
            else
            {
                FullGeo fg;
                cell_derivs(bg_s, x, core, geo_s, fg);
                Cons src;
                static_sources(p, us, o.gamma, fg, src);
                rhs.D += src.D;
                rhs.tau += src.tau;
                rhs.S[0] += src.S[0];
                rhs.S[1] += src.S[1];
                rhs.S[2] += src.S[2];
            }
        }
        if (o.use_volume_sources)
        {
            Geo gp, gm;
            cell_geometry(bg_plus, x, gp);
            cell_geometry(bg_minus, x, gm);
            const double lp = log_sqrt_det(gp), lm = log_sqrt_det(gm);
            const double dlog = (lp - lm) / (2.0 * a.vol_h);
            rhs.D += 0.0 - us.D * dlog;
            rhs.tau += 0.0 - us.tau * dlog;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                rhs.S[m] += 0.0 - us.S[m] * dlog;
        }

        if (MODE == MODE_RHS)
        {
            const long long ni = b.inpts;
            a.rhs[idx] = rhs.D;
            a.rhs[ni + idx] = rhs.S[0];
            a.rhs[2 * ni + idx] = rhs.S[1];
            a.rhs[3 * ni + idx] = rhs.S[2];
            a.rhs[4 * ni + idx] = rhs.tau;
            if (a.mask)
                a.mask[idx] = mbits;
        }
        else
        {
            /* stage update */
            Cons un;
            if (MODE == MODE_RK4_1 || MODE == MODE_RK4_2 ||
                MODE == MODE_RK4_3)
            {
                const Cons u0 = (MODE == MODE_RK4_1) ? us
                                                     : load_cons(a.U0, n, so);
                const long long ni = b.inpts;
                Cons ac = (MODE == MODE_RK4_1) ? u0
                                               : load_cons(a.acc, ni, idx);
                ac.D += a.c_acc * rhs.D;
                ac.tau += a.c_acc * rhs.tau;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    ac.S[m] += a.c_acc * rhs.S[m];
                store_cons(a.acc, ni, idx, ac);
                un.D = u0.D + a.c_stage * rhs.D;
                un.tau = u0.tau + a.c_stage * rhs.tau;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    un.S[m] = u0.S[m] + a.c_stage * rhs.S[m];
            }
            else if (MODE == MODE_RK4_4)
            {
                const long long ni = b.inpts;
                const Cons ac = load_cons(a.acc, ni, idx);
                un.D = ac.D + a.c_acc * rhs.D;
                un.tau = ac.tau + a.c_acc * rhs.tau;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    un.S[m] = ac.S[m] + a.c_acc * rhs.S[m];
            }
            else if (MODE == MODE_SSP_1)
            {
                un.D = us.D + a.dt * rhs.D;
                un.tau = us.tau + a.dt * rhs.tau;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    un.S[m] = us.S[m] + a.dt * rhs.S[m];
            }
            else /* MODE_SSP_2: 0.5 (old + stage + dt k) */
            {
                const Cons old = load_cons(a.U0, n, so);
                un.D = 0.5 * (old.D + us.D + a.dt * rhs.D);
                un.tau = 0.5 * (old.tau + us.tau + a.dt * rhs.tau);
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    un.S[m] = 0.5 * (old.S[m] + us.S[m] + a.dt * rhs.S[m]);
            }

            /* recovery with atmosphere (FixedBGFArrayBoxFiniteVolume.hpp:60-110) */
            Geo g;
            if (have_geo && bg_next.kind == 0)
                g = geo_s;
            else
                cell_geometry(bg_next, x, g);
            Prim rec;
            double res;
            const bool ok = con2prim(un, o, g.gu, rec, res);
            bool reset = !ok;
            Prim prim;
            unsigned char m = 0;
            cnt[0] = 1;
            if (ok)
            {
                prim = rec;
                floors(prim, o, g.g);
                reset = prim_differs(rec, prim);
                if (reset)
                {
                    cnt[2] = 1;
                    m |= GRHD_MASK_FLOORED;
                }
            }
            else
            {
                cnt[1] = 1;
                m |= GRHD_MASK_FAILED;
                atmosphere_prim(o, prim);
            }
            if (a.Pout)
                store_prim(a.Pout, n, so, prim);
            Prim wn = rec;
            bool wok = ok;
            if (reset)
            {
                cnt[3] = 1;
                m |= GRHD_MASK_RESET;
                prim2con(prim, o.gamma, g.g, un);
                wok = con2prim(un, o, g.gu, wn, res);
            }
            store_cons(a.Unext, n, so, un);
            if (wok)
            {
                store_prim(a.W, n, so, wn);
                a.wbad[so] = 0;
            }
            else
            {
                atomicMin(&a.st->fail_key,
                          fail_key(a.next_eval, 1, (unsigned long long)so));
                a.W[so] = res;
                a.wbad[so] = 1;
            }
            if (a.mask)
                a.mask[idx] = m;
        }
    }
    if (MODE != MODE_RHS && a.count)
        block_add_counts(a.st, cnt[0], cnt[1], cnt[2], cnt[3]);
}

/* Post-update recovery of the interior of `state` (no stage update). */
__global__ void __launch_bounds__(kBlock)
    k_recover(BoxDev b, Options o, Background bg, double *U, double *P,
              unsigned char *mask, DevStatus *st)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt[4] = {0, 0, 0, 0};
    if (idx < b.inpts)
    {
        const int i = b.lo[0] + (int)(idx % b.in[0]);
        const int j = b.lo[1] + (int)((idx / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(idx / ((long long)b.in[0] * b.in[1]));
        const long long n = b.gnpts;
        const long long so = goff(b, i, j, k);
        double x[3];
        coords(i, j, k, b.dx, b.center, x);
        Geo g;
        cell_geometry(bg, x, g);
        Cons c = load_cons(U, n, so);
        Prim rec, prim;
        double res;
        const bool ok = con2prim(c, o, g.gu, rec, res);
        bool reset = !ok;
        unsigned char m = 0;
        cnt[0] = 1;
        if (ok)
        {
            prim = rec;
            floors(prim, o, g.g);
            reset = prim_differs(rec, prim);
            if (reset)
            {
                cnt[2] = 1;
                m |= GRHD_MASK_FLOORED;
            }
        }
        else
        {
            cnt[1] = 1;
            m |= GRHD_MASK_FAILED;
            atmosphere_prim(o, prim);
        }
        store_prim(P, n, so, prim);
        if (reset)
        {
            cnt[3] = 1;
            m |= GRHD_MASK_RESET;
            prim2con(prim, o.gamma, g.g, c);
            store_cons(U, n, so, c);
        }
        if (mask)
            mask[idx] = m;
    }
    block_add_counts(st, cnt[0], cnt[1], cnt[2], cnt[3]);
}

/* add_scaled_conserved_rhs over the interior (rhs in interior layout) */
__global__ void __launch_bounds__(kBlock)
    k_add_scaled(BoxDev b, double *U, const double *rhs, double scale)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= b.inpts)
        return;
    const int i = b.lo[0] + (int)(idx % b.in[0]);
    const int j = b.lo[1] + (int)((idx / b.in[0]) % b.in[1]);
    const int k = b.lo[2] + (int)(idx / ((long long)b.in[0] * b.in[1]));
    const long long n = b.gnpts, ni = b.inpts;
    const long long so = goff(b, i, j, k);
#pragma unroll
    for (int c = 0; c < 5; ++c)
        U[c * n + so] += scale * rhs[c * ni + idx];
}

/* max over non-atmosphere cells of sum_d max|lambda| / dx */
__global__ void __launch_bounds__(kBlock)
    k_max_inv_dt(BoxDev b, Options o, Background bg, const double *U,
                 double inv_dx, DevStatus *st, double *W)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double best = 0.0;
    if (idx < b.inpts)
    {
        const int i = b.lo[0] + (int)(idx % b.in[0]);
        const int j = b.lo[1] + (int)((idx / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(idx / ((long long)b.in[0] * b.in[1]));
        const long long so = goff(b, i, j, k);
        double x[3];
        coords(i, j, k, b.dx, b.center, x);
        Geo g;
        cell_geometry(bg, x, g);
        const Cons c = load_cons(U, b.gnpts, so);
        Prim p;
        double res;
        if (!con2prim(c, o, g.gu, p, res))
        {
            atomicMin(&st->cfl_key, (unsigned long long)idx);
            W[idx] = res;
        }
        else
        {
            floors(p, o, g.g);
            if (!is_atmosphere(p, o, 1.0 + 1.0e-12))
            {
                double cell = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    double lo, hi;
                    char_speeds(p, o.gamma, g, d, lo, hi);
                    const double ms = std_max(fabs(lo), fabs(hi));
                    cell += ms * inv_dx;
                }
                best = std_max(best, cell);
            }
        }
    }
    /* non-negative doubles order like their bit patterns */
    unsigned long long bits = __double_as_longlong(best);
    bits = wave_max(bits);
    __shared__ unsigned long long part[kBlock / 64];
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long m = 0;
        for (int w = 0; w < kBlock / 64; ++w)
            m = part[w] > m ? part[w] : m;
        if (m)
            atomicMax(&st->max_bits, m);
    }
}

/* copy the 5 conserved components of the interior + ghosts */
__global__ void __launch_bounds__(kBlock)
    k_copy5(long long n, const double *src, double *dst)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
#pragma unroll
    for (int c = 0; c < 5; ++c)
        dst[c * n + idx] = src[c * n + idx];
}

__global__ void __launch_bounds__(kBlock)
    k_geometry(BoxDev b, Background bg, double *out)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= b.gnpts)
        return;
    const int i = b.glo[0] + (int)(idx % b.gn[0]);
    const int j = b.glo[1] + (int)((idx / b.gn[0]) % b.gn[1]);
    const int k = b.glo[2] + (int)(idx / ((long long)b.gn[0] * b.gn[1]));
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    Geo g;
    cell_geometry(bg, x, g);
    const long long n = b.gnpts;
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int m = 0; m < 6; ++m)
    {
        out[m * n + idx] = g.g[ii[m]][jj[m]];
        out[(6 + m) * n + idx] = g.gu[ii[m]][jj[m]];
    }
    out[12 * n + idx] = g.lapse;
#pragma unroll
    for (int m = 0; m < 3; ++m)
        out[(13 + m) * n + idx] = g.shift[m];
}

__global__ void k_fill_zero(long long n, double *p)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n)
        p[idx] = 0.0;
}

/* ------------------------------------------------------------------ */
/* Host side                                                          */
/* ------------------------------------------------------------------ */
inline unsigned grid_of(long long n)
{
    return (unsigned)((n + kBlock - 1) / kBlock);
}

BoxDev make_box(const grhd_box *b)
{
    BoxDev d;
    for (int i = 0; i < 3; ++i)
    {
        d.lo[i] = b->lo[i];
        d.hi[i] = b->hi[i];
        d.glo[i] = b->glo[i];
        d.ghi[i] = b->ghi[i];
        d.dlo[i] = b->dlo[i];
        d.dhi[i] = b->dhi[i];
        d.periodic[i] = b->periodic[i];
        d.gn[i] = b->ghi[i] - b->glo[i] + 1;
        d.in[i] = b->hi[i] - b->lo[i] + 1;
        d.center[i] = b->center[i];
    }
    d.gnpts = (long long)d.gn[0] * d.gn[1] * d.gn[2];
    d.inpts = (long long)d.in[0] * d.in[1] * d.in[2];
    d.dx = b->dx;
    return d;
}

Options make_options(const grhd_params *p)
{
    Options o;
    o.gamma = p->adiabatic_index;
    o.rho_floor = p->rho_floor;
    o.pressure_floor = p->pressure_floor;
    o.max_v2 = p->max_velocity_squared;
    o.source_atmosphere_factor = p->metric_source_atmosphere_factor;
    o.max_iterations = p->max_iterations;
    o.tolerance = p->tolerance;
    o.rec_pressure_floor = p->recovery_pressure_floor;
    o.rec_max_v2 = p->recovery_max_velocity_squared;
    o.theta = p->limiter_theta;
    o.use_reconstruction = p->use_reconstruction;
    o.use_static_sources = p->use_static_metric_sources;
    o.use_volume_sources = p->use_metric_volume_time_sources;
    return o;
}

/* background_factory(time): CircularBinaryPN::compute_state
 * (CircularBinaryPN.hpp:209-250) at params.time = bg.time + time. */
Background make_background(const grhd_background *p, double time)
{
    Background b;
    std::memset(&b, 0, sizeof(b));
    b.kind = p->kind;
    b.M = p->mass;
    b.a = p->spin;
    b.m1 = p->mass_1;
    b.m2 = p->mass_2;
    b.soft = p->softening_radius;
    if (p->kind == GRHD_BG_CIRCULAR_BINARY_PN)
    {
        const double params_time = p->time + time;
        const double total_mass = p->mass_1 + p->mass_2;
        const double omega =
            (p->angular_velocity != 0.0)
                ? p->angular_velocity
                : std::sqrt(total_mass / (p->separation * p->separation *
                                          p->separation));
        const double phase = p->phase + omega * params_time;
        const double cp = std::cos(phase);
        const double sp = std::sin(phase);
        const double r1 = p->mass_2 / total_mass * p->separation;
        const double r2 = p->mass_1 / total_mass * p->separation;
        b.p1[0] = r1 * cp;
        b.p1[1] = r1 * sp;
        b.p2[0] = -r2 * cp;
        b.p2[1] = -r2 * sp;
        b.u1[0] = -omega * r1 * sp;
        b.u1[1] = omega * r1 * cp;
        b.u2[0] = omega * r2 * sp;
        b.u2[1] = -omega * r2 * cp;
    }
    return b;
}

int check_args(const grhd_box *b, const grhd_params *o,
               const grhd_background *bg, int ncomp, int cb)
{
    if (!b || !o || !bg)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (cb < 0 || cb + GRHD_NUM_COMPS > ncomp)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (!(b->dx > 0.0))
        return GRHD_ERR_INVALID_ARGUMENT;
    for (int d = 0; d < 3; ++d)
    {
        if (b->lo[d] > b->hi[d])
            return GRHD_ERR_INVALID_ARGUMENT;
        if (b->glo[d] > b->lo[d] - 2 || b->ghi[d] < b->hi[d] + 2)
            return GRHD_ERR_INVALID_ARGUMENT;
    }
    long long gnpts = 1;
    for (int d = 0; d < 3; ++d)
        gnpts *= (long long)(b->ghi[d] - b->glo[d] + 1);
    if (gnpts >= (1ll << kLinBits))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (o->max_iterations < 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (bg->kind == GRHD_BG_KERR_SCHILD)
    {
        if (bg->spin > bg->mass || bg->spin < -bg->mass)
            return GRHD_ERR_INVALID_ARGUMENT;
    }
    else if (bg->kind == GRHD_BG_CIRCULAR_BINARY_PN)
    {
        if (bg->mass_1 <= 0.0 || bg->mass_2 <= 0.0 || bg->separation <= 0.0 ||
            bg->softening_radius < 0.0)
            return GRHD_ERR_INVALID_ARGUMENT;
    }
    else
        return GRHD_ERR_INVALID_ARGUMENT;
    return GRHD_OK;
}

/* The integrators keep their stage states internally, so every storage
 * cell outside the interior must be a physical (outflow) ghost. */
bool all_ghosts_physical(const grhd_box *b)
{
    for (int d = 0; d < 3; ++d)
        if (b->periodic[d] || b->lo[d] != b->dlo[d] || b->hi[d] != b->dhi[d])
            return false;
    return true;
}

} // namespace

struct grhd_ctx
{
    int device = 0;
    double *W = nullptr;
    unsigned char *wbad = nullptr;
    long long cap_g = 0;
    double *F[3] = {nullptr, nullptr, nullptr};
    long long cap_f[3] = {0, 0, 0};
    double *UA = nullptr, *UB = nullptr;
    double *acc = nullptr;
    long long cap_i = 0;
    long long cap_u = 0;
    DevStatus *dstat = nullptr;
    DevStatus *hstat = nullptr;
    /* instrumentation */
    bool timing = false;
    std::vector<hipEvent_t> events;
    std::vector<int> classes; /* class of the launch ending at events[i+1] */
    size_t nev = 0;
    double total_ms[GRHD_NUM_KERNEL_CLASSES] = {0};
    long long launches[GRHD_NUM_KERNEL_CLASSES] = {0};
};

namespace
{

template <class T> int grow(T **p, long long *cap, long long need)
{
    if (*cap >= need)
        return GRHD_OK;
    if (*p)
        (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc((void **)p, sizeof(T) * (size_t)need) != hipSuccess)
        return GRHD_ERR_DEVICE;
    *cap = need;
    return GRHD_OK;
}

int ensure_rhs_scratch(grhd_ctx *c, const BoxDev &b)
{
    if (c->cap_g < b.gnpts)
    {
        (void)hipFree(c->W);
        (void)hipFree(c->wbad);
        c->W = nullptr;
        c->wbad = nullptr;
        c->cap_g = 0;
        if (hipMalloc((void **)&c->W, sizeof(double) * 6 * (size_t)b.gnpts) !=
                hipSuccess ||
            hipMalloc((void **)&c->wbad, (size_t)b.gnpts) != hipSuccess)
            return GRHD_ERR_DEVICE;
        c->cap_g = b.gnpts;
    }
    for (int d = 0; d < 3; ++d)
    {
        const long long nf = b.inpts / b.in[d] * (b.in[d] + 1);
        const int rc = grow(&c->F[d], &c->cap_f[d], 5 * nf);
        if (rc)
            return rc;
    }
    return GRHD_OK;
}

int ensure_stage_scratch(grhd_ctx *c, const BoxDev &b)
{
    if (c->cap_u < 5 * b.gnpts)
    {
        long long ca = 0, cb = 0;
        if (c->UA)
            (void)hipFree(c->UA);
        if (c->UB)
            (void)hipFree(c->UB);
        c->UA = c->UB = nullptr;
        c->cap_u = 0;
        int rc = grow(&c->UA, &ca, 5 * b.gnpts);
        if (rc)
            return rc;
        rc = grow(&c->UB, &cb, 5 * b.gnpts);
        if (rc)
            return rc;
        c->cap_u = 5 * b.gnpts;
    }
    return grow(&c->acc, &c->cap_i, 5 * b.inpts);
}

/* Record an event on `s`; the first call of a timed API call opens the
 * sequence, each later one closes the launch of class `cls`. */
void tick(grhd_ctx *c, int cls, hipStream_t s)
{
    if (!c->timing)
        return;
    if (c->nev == c->events.size())
    {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess)
            return;
        c->events.push_back(e);
        c->classes.push_back(0);
    }
    c->classes[c->nev] = cls;
    (void)hipEventRecord(c->events[c->nev], s);
    ++c->nev;
}

/* after the stream has drained: fold the recorded gaps into the totals */
void collect_ticks(grhd_ctx *c)
{
    if (!c->timing)
        return;
    for (size_t i = 1; i < c->nev; ++i)
    {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->events[i - 1], c->events[i]) ==
            hipSuccess)
        {
            c->total_ms[c->classes[i]] += ms;
            c->launches[c->classes[i]] += 1;
        }
    }
    c->nev = 0;
}

void launch_status_init(grhd_ctx *c, hipStream_t s)
{
    c->nev = 0;
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(64), 0, s, c->dstat);
    tick(c, GRHD_KERNEL_OTHER, s);
}

/* RHS pieces: primitives, faces */
void launch_prims(grhd_ctx *c, const BoxDev &b, const Options &o,
                  const Background &bg, double *U, double *P, bool exterior,
                  unsigned long long eval, hipStream_t s)
{
    if (exterior)
        hipLaunchKernelGGL(k_prims<true>, dim3(grid_of(b.gnpts - b.inpts)),
                           dim3(kBlock), 0, s, b, o, bg, U, P, c->W, c->wbad,
                           c->dstat, eval);
    else
        hipLaunchKernelGGL(k_prims<false>, dim3(grid_of(b.gnpts)),
                           dim3(kBlock), 0, s, b, o, bg, U, P, c->W, c->wbad,
                           c->dstat, eval);
    tick(c, GRHD_KERNEL_PRIMS, s);
}

int face_segments(const BoxDev &b, int d)
{
    return (b.in[d] + 1 + kSeg - 1) / kSeg;
}

long long face_threads(const BoxDev &b, int d)
{
    return b.inpts / b.in[d] * face_segments(b, d);
}

void launch_faces(grhd_ctx *c, const BoxDev &b, const Options &o,
                  const Background &bg, hipStream_t s, long long nf[3])
{
    for (int d = 0; d < 3; ++d)
        nf[d] = b.inpts / b.in[d] * (b.in[d] + 1);
    {
        /* one lane per cell lo-1 .. hi+1 of each row, 63 new cells per
         * 64-lane wave */
        const long long next = (long long)(b.in[0] + 2) * b.in[1] * b.in[2];
        const long long waves = (next - 1 + 62) / 63;
        hipLaunchKernelGGL(k_faces_x, dim3(grid_of(waves * 64)), dim3(kBlock),
                           0, s, b, o, bg, c->W, c->F[0], nf[0], next);
    }
    tick(c, GRHD_KERNEL_FACES_X, s);
    hipLaunchKernelGGL(k_faces<1>, dim3(grid_of(face_threads(b, 1))),
                       dim3(kBlock), 0, s, b, o, bg, c->W, c->F[1], nf[1],
                       face_segments(b, 1));
    tick(c, GRHD_KERNEL_FACES_Y, s);
    hipLaunchKernelGGL(k_faces<2>, dim3(grid_of(face_threads(b, 2))),
                       dim3(kBlock), 0, s, b, o, bg, c->W, c->F[2], nf[2],
                       face_segments(b, 2));
    tick(c, GRHD_KERNEL_FACES_Z, s);
}

CellArgs base_args(grhd_ctx *c, const BoxDev &b, const long long nf[3])
{
    CellArgs a;
    std::memset(&a, 0, sizeof(a));
    a.F0 = c->F[0];
    a.F1 = c->F[1];
    a.F2 = c->F[2];
    a.nf0 = nf[0];
    a.nf1 = nf[1];
    a.nf2 = nf[2];
    a.W = c->W;
    a.wbad = c->wbad;
    a.st = c->dstat;
    a.inv_dx = 1.0 / b.dx;
    return a;
}

void launch_cells_raw(int mode, const BoxDev &b, const Options &o,
                  const Background &bs, const Background &bn,
                  const Background &bp, const Background &bm,
                  const CellArgs &a, hipStream_t s)
{
    const dim3 g(grid_of(b.inpts)), t(kBlock);
    switch (mode)
    {
    case MODE_RHS:
        hipLaunchKernelGGL(k_cells<MODE_RHS>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    case MODE_RK4_1:
        hipLaunchKernelGGL(k_cells<MODE_RK4_1>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    case MODE_RK4_2:
        hipLaunchKernelGGL(k_cells<MODE_RK4_2>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    case MODE_RK4_3:
        hipLaunchKernelGGL(k_cells<MODE_RK4_3>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    case MODE_RK4_4:
        hipLaunchKernelGGL(k_cells<MODE_RK4_4>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    case MODE_SSP_1:
        hipLaunchKernelGGL(k_cells<MODE_SSP_1>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    default:
        hipLaunchKernelGGL(k_cells<MODE_SSP_2>, g, t, 0, s, b, o, bs, bn, bp,
                           bm, a);
        break;
    }
}

void launch_cells(grhd_ctx *c, int mode, const BoxDev &b, const Options &o,
                  const Background &bs, const Background &bn,
                  const Background &bp, const Background &bm,
                  const CellArgs &a, hipStream_t s)
{
    launch_cells_raw(mode, b, o, bs, bn, bp, bm, a, s);
    tick(c, GRHD_KERNEL_CELLS, s);
}

/* Volume-time source backgrounds at t +- max(step, 1e-12). */
void vol_backgrounds(const grhd_background *bg, const grhd_params *p,
                     double t, Background &bp, Background &bm, double &h)
{
    h = std::max(p->metric_time_derivative_step, 1.0e-12);
    bp = make_background(bg, t + h);
    bm = make_background(bg, t - h);
}

/* Wait, then translate the status block into a return code. */
int finish(grhd_ctx *c, const BoxDev &b, hipStream_t s, grhd_error *err,
           grhd_diagnostics *diag, bool want_cfl)
{
    if (hipMemcpyAsync(c->hstat, c->dstat, sizeof(DevStatus),
                       hipMemcpyDeviceToHost, s) != hipSuccess)
        return GRHD_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess)
        return GRHD_ERR_DEVICE;
    if (hipGetLastError() != hipSuccess)
        return GRHD_ERR_DEVICE;
    collect_ticks(c);
    if (diag)
    {
        diag->num_cells = (long long)c->hstat->counts[0];
        diag->num_failed_recoveries = (long long)c->hstat->counts[1];
        diag->num_floored_primitives = (long long)c->hstat->counts[2];
        diag->num_conserved_resets = (long long)c->hstat->counts[3];
    }
    const unsigned long long key = c->hstat->fail_key;
    if (key != kNoFail)
    {
        const unsigned long long phase = (key >> kLinBits) & 1ull;
        const long long lin = (long long)(key & ((1ull << kLinBits) - 1));
        int iv[3] = {b.glo[0] + (int)(lin % b.gn[0]),
                     b.glo[1] + (int)((lin / b.gn[0]) % b.gn[1]),
                     b.glo[2] + (int)(lin / ((long long)b.gn[0] * b.gn[1]))};
        if (phase == 0)
        {
            /* the reference names the outflow source cell */
            for (int d = 0; d < 3; ++d)
            {
                if (iv[d] < b.dlo[d] && !b.periodic[d])
                    iv[d] = b.lo[d];
                else if (iv[d] > b.dhi[d] && !b.periodic[d])
                    iv[d] = b.hi[d];
            }
        }
        if (err)
        {
            err->code = phase == 0 ? GRHD_ERR_RECOVERY_GHOST_FILL
                                   : GRHD_ERR_RECOVERY_RHS;
            err->iv[0] = iv[0];
            err->iv[1] = iv[1];
            err->iv[2] = iv[2];
            double r = 0.0;
            (void)hipMemcpy(&r, c->W + lin, sizeof(double),
                            hipMemcpyDeviceToHost);
            err->residual = r;
        }
        return phase == 0 ? GRHD_ERR_RECOVERY_GHOST_FILL
                          : GRHD_ERR_RECOVERY_RHS;
    }
    if (want_cfl && c->hstat->cfl_key != kNoFail)
    {
        const long long lin = (long long)c->hstat->cfl_key;
        if (err)
        {
            err->code = GRHD_ERR_RECOVERY_CFL;
            err->iv[0] = b.lo[0] + (int)(lin % b.in[0]);
            err->iv[1] = b.lo[1] + (int)((lin / b.in[0]) % b.in[1]);
            err->iv[2] = b.lo[2] + (int)(lin / ((long long)b.in[0] * b.in[1]));
            double r = 0.0;
            (void)hipMemcpy(&r, c->W + lin, sizeof(double),
                            hipMemcpyDeviceToHost);
            err->residual = r;
        }
        return GRHD_ERR_RECOVERY_CFL;
    }
    return GRHD_OK;
}

inline double *state_U(double *state, int cb, const BoxDev &b)
{
    return state + (long long)cb * b.gnpts;
}

} // namespace

extern "C" {

int grhd_ctx_create(grhd_ctx **out)
{
    if (!out)
        return GRHD_ERR_INVALID_ARGUMENT;
    grhd_ctx *c = new grhd_ctx();
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipMalloc((void **)&c->dstat, sizeof(DevStatus)) != hipSuccess ||
        hipHostMalloc((void **)&c->hstat, sizeof(DevStatus)) != hipSuccess)
    {
        delete c;
        return GRHD_ERR_DEVICE;
    }
    *out = c;
    return GRHD_OK;
}

void grhd_ctx_destroy(grhd_ctx *c)
{
    if (!c)
        return;
    (void)hipDeviceSynchronize();
    (void)hipFree(c->W);
    (void)hipFree(c->wbad);
    for (int d = 0; d < 3; ++d)
        (void)hipFree(c->F[d]);
    (void)hipFree(c->UA);
    (void)hipFree(c->UB);
    (void)hipFree(c->acc);
    for (hipEvent_t e : c->events)
        (void)hipEventDestroy(e);
    (void)hipFree(c->dstat);
    (void)hipHostFree(c->hstat);
    delete c;
}

int grhd_compute_rhs_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double time,
                         double *state, int ncomp, int comp_base, double *rhs,
                         unsigned char *source_mask, grhd_error *err,
                         void *stream)
{
    if (!ctx || !state || !rhs)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_args(box, params, background, ncomp, comp_base);
    if (rc)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    rc = ensure_rhs_scratch(ctx, b);
    if (rc)
        return rc;
    const Options o = make_options(params);
    const Background bg = make_background(background, time);
    Background bp = bg, bm = bg;
    double h = 1.0;
    if (o.use_volume_sources)
        vol_backgrounds(background, params, time, bp, bm, h);

    launch_status_init(ctx, s);
    /* rhs.setVal(0.0) over every component */
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_of((long long)ncomp * b.inpts)),
                       dim3(kBlock), 0, s, (long long)ncomp * b.inpts, rhs);
    tick(ctx, GRHD_KERNEL_OTHER, s);
    double *U = state_U(state, comp_base, b);
    launch_prims(ctx, b, o, bg, U, U + 5 * b.gnpts, false, 0, s);
    long long nf[3];
    launch_faces(ctx, b, o, bg, s, nf);
    CellArgs a = base_args(ctx, b, nf);
    a.Us = U;
    a.rhs = rhs + (long long)comp_base * b.inpts;
    a.mask = source_mask;
    a.vol_h = h;
    launch_cells(ctx, MODE_RHS, b, o, bg, bg, bp, bm, a, s);
    return finish(ctx, b, s, err, nullptr, false);
}

int grhd_recover_primitives_box(grhd_ctx *ctx, const grhd_box *box,
                                const grhd_params *params,
                                const grhd_background *background,
                                double time, double *state, int ncomp,
                                int comp_base, unsigned char *mask,
                                grhd_diagnostics *diagnostics, void *stream)
{
    if (!ctx || !state)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_args(box, params, background, ncomp, comp_base);
    if (rc)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    const Options o = make_options(params);
    const Background bg = make_background(background, time);
    launch_status_init(ctx, s);
    double *U = state_U(state, comp_base, b);
    hipLaunchKernelGGL(k_recover, dim3(grid_of(b.inpts)), dim3(kBlock), 0, s,
                       b, o, bg, U, U + 5 * b.gnpts, mask, ctx->dstat);
    tick(ctx, GRHD_KERNEL_OTHER, s);
    return finish(ctx, b, s, nullptr, diagnostics, false);
}

int grhd_update_conserved_box(grhd_ctx *ctx, const grhd_box *box,
                              const grhd_params *params,
                              const grhd_background *background,
                              double *state, int ncomp, int comp_base,
                              const double *rhs, double dt, int recover,
                              double recovery_time, unsigned char *mask,
                              grhd_diagnostics *diagnostics, void *stream)
{
    if (!ctx || !state || !rhs)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_args(box, params, background, ncomp, comp_base);
    if (rc)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    double *U = state_U(state, comp_base, b);
    launch_status_init(ctx, s);
    hipLaunchKernelGGL(k_add_scaled, dim3(grid_of(b.inpts)), dim3(kBlock), 0,
                       s, b, U, rhs + (long long)comp_base * b.inpts, dt);
    tick(ctx, GRHD_KERNEL_OTHER, s);
    if (recover)
    {
        const Options o = make_options(params);
        const Background bg = make_background(background, recovery_time);
        hipLaunchKernelGGL(k_recover, dim3(grid_of(b.inpts)), dim3(kBlock), 0,
                           s, b, o, bg, U, U + 5 * b.gnpts, mask, ctx->dstat);
        tick(ctx, GRHD_KERNEL_OTHER, s);
    }
    grhd_diagnostics d;
    rc = finish(ctx, b, s, nullptr, &d, false);
    if (diagnostics)
    {
        if (recover)
            *diagnostics = d;
        else
            std::memset(diagnostics, 0, sizeof(*diagnostics));
    }
    return rc;
}

int grhd_compute_max_inverse_dt_box(grhd_ctx *ctx, const grhd_box *box,
                                    const grhd_params *params,
                                    const grhd_background *background,
                                    double time, const double *state,
                                    int ncomp, int comp_base,
                                    double *max_inverse_dt, grhd_error *err,
                                    void *stream)
{
    if (!ctx || !state || !max_inverse_dt)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_args(box, params, background, ncomp, comp_base);
    if (rc)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    rc = ensure_rhs_scratch(ctx, b);
    if (rc)
        return rc;
    const Options o = make_options(params);
    const Background bg = make_background(background, time);
    launch_status_init(ctx, s);
    hipLaunchKernelGGL(k_max_inv_dt, dim3(grid_of(b.inpts)), dim3(kBlock), 0,
                       s, b, o, bg, state + (long long)comp_base * b.gnpts,
                       1.0 / b.dx, ctx->dstat, ctx->W);
    tick(ctx, GRHD_KERNEL_OTHER, s);
    rc = finish(ctx, b, s, err, nullptr, true);
    if (rc)
        return rc;
    double m;
    const unsigned long long bits = ctx->hstat->max_bits;
    std::memcpy(&m, &bits, sizeof(m));
    *max_inverse_dt = m;
    return GRHD_OK;
}

int grhd_compute_stable_dt_box(grhd_ctx *ctx, const grhd_box *box,
                               const grhd_params *params,
                               const grhd_background *background, double time,
                               const double *state, int ncomp, int comp_base,
                               double cfl, double *stable_dt, grhd_error *err,
                               void *stream)
{
    if (!stable_dt)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (cfl <= 0.0)
    {
        if (err)
        {
            std::memset(err, 0, sizeof(*err));
            err->code = GRHD_ERR_INVALID_ARGUMENT;
        }
        return GRHD_ERR_INVALID_ARGUMENT;
    }
    double m = 0.0;
    const int rc = grhd_compute_max_inverse_dt_box(ctx, box, params,
                                                   background, time, state,
                                                   ncomp, comp_base, &m, err,
                                                   stream);
    if (rc)
        return rc;
    *stable_dt = cfl / std::max(m, 1.0e-14);
    return GRHD_OK;
}

int grhd_advance_ssprk2_box(grhd_ctx *ctx, const grhd_box *box,
                            const grhd_params *params,
                            const grhd_background *background, double time,
                            double dt, double *state, int ncomp,
                            int comp_base, unsigned char *mask,
                            grhd_diagnostics *diagnostics, grhd_error *err,
                            void *stream)
{
    if (!ctx || !state)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_args(box, params, background, ncomp, comp_base);
    if (rc)
        return rc;
    if (!all_ghosts_physical(box))
        return GRHD_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    if ((rc = ensure_rhs_scratch(ctx, b)) || (rc = ensure_stage_scratch(ctx, b)))
        return rc;
    const Options o = make_options(params);
    double *U = state_U(state, comp_base, b);
    launch_status_init(ctx, s);
    long long nf[3];

    /* k1 = L(U, t); stage = U + dt k1; recover(stage, t + dt) */
    const Background b0 = make_background(background, time);
    const Background b1 = make_background(background, time + dt);
    Background bp, bm;
    double h = 1.0;
    bp = bm = b0;
    if (o.use_volume_sources)
        vol_backgrounds(background, params, time, bp, bm, h);
    launch_prims(ctx, b, o, b0, U, U + 5 * b.gnpts, false, 0, s);
    launch_faces(ctx, b, o, b0, s, nf);
    CellArgs a = base_args(ctx, b, nf);
    a.Us = U;
    a.Unext = ctx->UA;
    a.dt = dt;
    a.vol_h = h;
    a.next_eval = 1;
    launch_cells(ctx, MODE_SSP_1, b, o, b0, b1, bp, bm, a, s);

    /* k2 = L(stage, t + dt); U = (U + stage + dt k2) / 2; recover(U) */
    bp = bm = b1;
    if (o.use_volume_sources)
        vol_backgrounds(background, params, time + dt, bp, bm, h);
    launch_prims(ctx, b, o, b1, ctx->UA, nullptr, true, 1, s);
    launch_faces(ctx, b, o, b1, s, nf);
    a = base_args(ctx, b, nf);
    a.Us = ctx->UA;
    a.U0 = U;
    a.Unext = U;
    a.Pout = U + 5 * b.gnpts;
    a.mask = mask;
    a.dt = dt;
    a.vol_h = h;
    a.next_eval = 2;
    a.count = 1;
    launch_cells(ctx, MODE_SSP_2, b, o, b1, b1, bp, bm, a, s);
    return finish(ctx, b, s, err, diagnostics, false);
}

int grhd_advance_rk4_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double time,
                         double dt, int nsteps, double *state, int ncomp,
                         int comp_base, unsigned char *mask,
                         grhd_diagnostics *diagnostics, grhd_error *err,
                         void *stream)
{
    if (!ctx || !state || nsteps < 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_args(box, params, background, ncomp, comp_base);
    if (rc)
        return rc;
    if (!all_ghosts_physical(box))
        return GRHD_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    if ((rc = ensure_rhs_scratch(ctx, b)) || (rc = ensure_stage_scratch(ctx, b)))
        return rc;
    const Options o = make_options(params);
    double *U = state_U(state, comp_base, b);
    double *P = U + 5 * b.gnpts;
    launch_status_init(ctx, s);
    const double c2 = dt / 2.0, c3 = dt / 3.0, c6 = dt / 6.0;
    unsigned long long eval = 0;
    long long nf[3];
    for (int step = 0; step < nsteps; ++step)
    {
        const double t = time + step * dt;
        const double ts[4] = {t, t + c2, t + c2, t + dt};
        const double tn[4] = {t + c2, t + c2, t + dt, t + dt};
        const double cst[4] = {c2, c2, dt, 0.0};
        const double cac[4] = {c6, c3, c3, c6};
        double *Us[4] = {U, ctx->UA, ctx->UB, ctx->UA};
        double *Un[4] = {ctx->UA, ctx->UB, ctx->UA, U};
        for (int st = 0; st < 4; ++st)
        {
            const Background bs = make_background(background, ts[st]);
            const Background bn = make_background(background, tn[st]);
            Background bp = bs, bm = bs;
            double h = 1.0;
            if (o.use_volume_sources)
                vol_backgrounds(background, params, ts[st], bp, bm, h);
            /* the first evaluation of the call recovers every cell; later
             * ones reuse the interior W the previous stage produced */
            launch_prims(ctx, b, o, bs, Us[st], st == 0 ? P : nullptr,
                         eval != 0, eval, s);
            launch_faces(ctx, b, o, bs, s, nf);
            CellArgs a = base_args(ctx, b, nf);
            a.Us = Us[st];
            a.U0 = U;
            a.acc = ctx->acc;
            a.Unext = Un[st];
            a.Pout = (st == 3) ? P : nullptr;
            a.mask = (st == 3 && step == nsteps - 1) ? mask : nullptr;
            a.c_stage = cst[st];
            a.c_acc = cac[st];
            a.dt = dt;
            a.vol_h = h;
            a.next_eval = eval + 1;
            a.count = 1;
            launch_cells(ctx, MODE_RK4_1 + st, b, o, bs, bn, bp, bm, a, s);
            ++eval;
        }
    }
    return finish(ctx, b, s, err, diagnostics, false);
}

int grhd_cell_geometry_box(grhd_ctx *ctx, const grhd_box *box,
                           const grhd_background *background, double time,
                           double *out, void *stream)
{
    if (!ctx || !box || !background || !out)
        return GRHD_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    const BoxDev b = make_box(box);
    const Background bg = make_background(background, time);
    hipLaunchKernelGGL(k_geometry, dim3(grid_of(b.gnpts)), dim3(kBlock), 0, s,
                       b, bg, out);
    if (hipStreamSynchronize(s) != hipSuccess)
        return GRHD_ERR_DEVICE;
    return hipGetLastError() == hipSuccess ? GRHD_OK : GRHD_ERR_DEVICE;
}

int grhd_ctx_set_timing(grhd_ctx *ctx, int enable)
{
    if (!ctx)
        return GRHD_ERR_INVALID_ARGUMENT;
    ctx->timing = enable != 0;
    ctx->nev = 0;
    for (int i = 0; i < GRHD_NUM_KERNEL_CLASSES; ++i)
    {
        ctx->total_ms[i] = 0.0;
        ctx->launches[i] = 0;
    }
    return GRHD_OK;
}

int grhd_ctx_get_timing(grhd_ctx *ctx, double *total_ms, long long *launches)
{
    if (!ctx)
        return GRHD_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < GRHD_NUM_KERNEL_CLASSES; ++i)
    {
        if (total_ms)
            total_ms[i] = ctx->total_ms[i];
        if (launches)
            launches[i] = ctx->launches[i];
    }
    return GRHD_OK;
}

} /* extern "C" */
