/*
 * grhd_stage_kernels.hpp -- the CDNA4 kernels of one right-hand side /
 * stage over every box of a level.
 *
 * A launch covers all boxes of a level.  The host writes a device box
 * table (BoxDev: geometry of each box and its offsets into the shared
 * pools) and, per kind of work item, a block table: block b processes
 * items first .. first+255 of box `box` (blocks never straddle boxes), so
 * one launch per kernel serves 1 or 540 boxes alike.  Box parameters are
 * block-uniform and load into scalar registers.
 *
 * Work decomposition of one right-hand side (the reference's three
 * directional sweeps, FixedBGFArrayBoxFiniteVolume.hpp:199-363, with every
 * duplicated quantity computed once):
 *   k_exchange   internal ghosts copied from the owning box (LevelData
 *                exchange(), :310, :355)
 *   k_prims      outflow ghost fill (:113-177) and con2prim over the
 *                ghosted box (:33-58) -> primitives W (6 x npts) + a
 *                failure byte per cell
 *   k_faces_x,   MUSCL slopes (FArrayBoxFiniteVolume.hpp:573-589), face
 *   k_faces<1,2> metric = mean of the two cell metrics (:179-197), floors,
 *   (walkers) or prim2con, HLLE -> face fluxes F_d (5 x nfaces), each
 *   k_faces_t<1,2> face once
 *   (tiles)
 *   k_guard_sources  (guarded mixed path only) the FP64 static sources
 *                of the guarded cells the FP32 face pass listed
 *   k_cells<M>   flux divergence in the reference's summation order,
 *                static-metric sources (:255-298), volume-time sources
 *                (:616-657), then per mode: write rhs, or an RK4 / SSPRK2
 *                stage update fused with the post-update recovery (:60-110)
 *                that also produces W of the next stage
 *
 * Memory follows the FArrayBox layout (component-major, i fastest), so a
 * wavefront's loads of one component are unit-stride along i.  Reductions
 * (reset counts, max 1/dt) are wavefront-64 shuffles, one LDS step across
 * the block's waves and one atomic per block.  No MFMA: nothing here is a
 * dense contraction.  Compiled with -ffp-contract=off.
 */
#ifndef GRHD_STAGE_KERNELS_HPP_
#define GRHD_STAGE_KERNELS_HPP_

#include <hip/hip_runtime.h>

#include "../../include/grhd_abi.h"
#include "grhd_box_table.hpp"
#include "grhd_physics.hpp"

namespace grhd
{

/* Pools shared by every box of the level. */
enum PoolId
{
    POOL_STATE = 0,
    POOL_A = 1,
    POOL_B = 2
};

struct Pools
{
    double *state; /* caller's FABs (ncomp components each) */
    int cb;        /* comp_base */
    double *UA, *UB;
    double *acc;
    double *W;
    unsigned char *wbad;
    double *F[3];
    double *rhs; /* rhs FABs: rhs_ncomp x inpts per box */
    int rhs_ncomp;
    unsigned char *mask;
    DevStatus *st;
    /* guarded mixed path: the faces left to FP64, one entry each
     * ((direction << 30) | box, face index), as many as DevStatus::nlist */
    int2 *list;
    /* guarded mixed path: the interior cells that carry kWGuard, one
     * entry each (box, interior index), as many as DevStatus::ncell, and
     * the FP64 static sources of those outside the metric-source
     * atmosphere (k_guard_sources), 5 components over each box's interior,
     * like the rhs pool */
    int2 *clist;
    double *src;
};

/* Per storage cell flags beside W (Pools::wbad): kWFailed when the cell's
 * recovery failed (W then holds the residual), kWGuard when, on the
 * guarded mixed path, the faces beside the cell are computed in FP64. */
constexpr unsigned char kWFailed = 1;
constexpr unsigned char kWGuard = 2;

__device__ __forceinline__ bool guarded_mixed(const Options &o)
{
    return o.guard_margin > 0.0 || o.guard_shift > 0.0;
}

__device__ __forceinline__ unsigned char failed_flags(const Options &o)
{
    return guarded_mixed(o) ? (unsigned char)(kWFailed | kWGuard) : kWFailed;
}

__device__ __forceinline__ unsigned char
recovered_flags(const Prim &w, const Options &o, const Geo &geo)
{
    return (guarded_mixed(o) && fp64_cell(w, o, geo)) ? kWGuard : 0;
}

__device__ __forceinline__ double *pool_cons(const Pools &p, int which,
                                             const BoxDev &b)
{
    if (which == POOL_STATE)
        return p.state + b.off_u + (long long)p.cb * b.gnpts;
    return (which == POOL_A ? p.UA : p.UB) + 5 * b.off_g;
}

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
fail_key(unsigned long long eval, int phase, const BoxDev &b, long long idx)
{
    return (eval << kKeyBits) |
           (unsigned long long)(2 * b.off_g + (phase ? b.gnpts : 0) + idx);
}

/* The block's box and the item this thread processes (-1 past the end).
 * The box is bound by reference, not copied: its fields are block-uniform
 * scalar loads the compiler may issue where they are used, which keeps the
 * scalar register file from spilling in the large kernels. */
__device__ __forceinline__ long long block_item(const int2 *__restrict__ items,
                                                const BoxDev *__restrict__ boxes,
                                                int cls, const BoxDev *&b)
{
    const int2 bt = items[blockIdx.x];
    b = boxes + bt.x;
    const long long it = (long long)bt.y + threadIdx.x;
    return it < class_items(*b, cls) ? it : -1;
}

/* True once a recovery failure of an evaluation before `eval` is
 * recorded.  A kernel writing primitives of evaluation `eval` then drops
 * its stores: the call reports the first failure, and the scratch slot
 * holding that failure's residual must survive later evaluations.  Failures
 * of `eval` itself do not stop it, since a ghost-fill failure of an
 * evaluation precedes its interior failures in the reference's order.  The
 * word is read once per wavefront at kernel start and consulted only at the
 * stores, so its latency hides behind the arithmetic; every lane of the
 * wavefront sees the same value. */
__device__ __forceinline__ bool failed_before(const DevStatus *st,
                                              unsigned long long eval)
{
    return __hip_atomic_load(&st->fail_key, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT) <
           (eval << kKeyBits);
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

__device__ inline void block_add_counts(DevStatus *st, unsigned long long c0,
                                        unsigned long long c1,
                                        unsigned long long c2,
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
/* k_exchange: internal ghost cells of every box copied from the box  */
/* whose interior owns them (host-built list: dst box, dst index, src */
/* box, src index), `ncopy` components of the pool.                   */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(kBlock)
    k_exchange(const BoxDev *__restrict__ boxes, const int4 *list, long long n, Pools p,
               int which, int ncopy)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int4 e = list[t];
    const BoxDev &bd = boxes[e.x];
    const BoxDev &bs = boxes[e.z];
    double *dst = (which == POOL_STATE)
                      ? p.state + bd.off_u
                      : pool_cons(p, which, bd);
    const double *src = (which == POOL_STATE)
                            ? p.state + bs.off_u
                            : pool_cons(p, which, bs);
    for (int c = 0; c < ncopy; ++c)
        dst[c * bd.gnpts + e.y] = src[c * bs.gnpts + e.w];
}

/* ------------------------------------------------------------------ */
/* k_prims: outflow ghost fill + con2prim over the storage box.       */
/* EXTERIOR: interior W already holds con2prim of the interior state  */
/* (written by the previous k_cells); only exterior cells are done.   */
/* ------------------------------------------------------------------ */
template <bool EXTERIOR>
__global__ void __launch_bounds__(kBlock)
    k_prims(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o, Background bg,
            Pools p, int which, int store_prims, unsigned long long eval)
{
    const BoxDev *bp;
    const long long it =
        block_item(items, boxes, EXTERIOR ? C_EXT : C_ALL, bp);
    const BoxDev &b = *bp;
    if (it < 0)
        return;
    const bool dead = failed_before(p.st, eval);
    int i, j, k;
    if (EXTERIOR)
    {
        exterior_cell(b, it, i, j, k);
    }
    else
    {
        i = b.glo[0] + (int)(it % b.gn[0]);
        j = b.glo[1] + (int)((it / b.gn[0]) % b.gn[1]);
        k = b.glo[2] + (int)(it / ((long long)b.gn[0] * b.gn[1]));
    }
    const long long idx = goff(b, i, j, k);
    const int iv[3] = {i, j, k};
    int src[3] = {i, j, k};
    bool physical = false;
    bool src_interior = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
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
        if (src[d] < b.lo[d] || src[d] > b.hi[d])
            src_interior = false;
    }

    const long long n = b.gnpts;
    double *U = pool_cons(p, which, b);
    double *W = p.W + 6 * b.off_g;
    unsigned char *wbad = p.wbad + b.off_g;
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    Geo gt;
    cell_geometry(bg, x, gt);
    Cons c;
    if (physical)
    {
        const long long so = goff(b, src[0], src[1], src[2]);
        Prim pr;
        bool src_ok;
        double res = 0.0;
        if (EXTERIOR && src_interior)
        {
            /* con2prim of the source was done by the previous stage */
            src_ok = (wbad[so] & kWFailed) == 0;
            if (src_ok)
                pr = load_prim(W, n, so);
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
            src_ok = con2prim(cs, o, gs.gu, pr, res);
        }
        if (!src_ok)
        {
            if (!dead)
            {
                atomicMin(&p.st->fail_key, fail_key(eval, 0, b, idx));
                W[idx] = res;
                wbad[idx] = failed_flags(o);
            }
            return;
        }
        floors(pr, o, gt.g);
        prim2con(pr, o.gamma, gt.g, c);
        if (!dead)
        {
            if (store_prims)
                store_prim(U + 5 * n, n, idx, pr);
            store_cons(U, n, idx, c);
        }
    }
    else
    {
        c = load_cons(U, n, idx);
    }
    Prim w;
    double res;
    const bool ok = con2prim(c, o, gt.gu, w, res);
    if (dead)
        return;
    if (!ok)
    {
        atomicMin(&p.st->fail_key, fail_key(eval, 1, b, idx));
        W[idx] = res;
        wbad[idx] = failed_flags(o);
        return;
    }
    store_prim(W, n, idx, w);
    wbad[idx] = recovered_flags(w, o, gt);
}

/* The face kernels compute in T: double on the bit-exact path, float on
 * the mixed-precision path (reconstruction, face metric and HLLE in FP32;
 * W comes from the FP64 recovery and is rounded once on load; fluxes are
 * stored as T and summed in FP64 by k_cells). */
template <class T>
__device__ __forceinline__ PrimT<T> load_prim_as(const double *P, long long n,
                                                 long long o)
{
    PrimT<T> p;
    p.rho = T(P[o]);
    p.eps = T(P[n + o]);
    p.pressure = T(P[2 * n + o]);
    p.v[0] = T(P[3 * n + o]);
    p.v[1] = T(P[4 * n + o]);
    p.v[2] = T(P[5 * n + o]);
    return p;
}

template <class T>
__device__ __forceinline__ void coords_as(int i, int j, int k, const BoxDev &b,
                                          T x[3])
{
    double xd[3];
    coords(i, j, k, b.dx, b.center, xd);
    x[0] = T(xd[0]);
    x[1] = T(xd[1]);
    x[2] = T(xd[2]);
}

/* Face-flux pool of direction d for box b, stored as T. */
template <class T>
__device__ __forceinline__ T *face_pool(const Pools &p, const BoxDev &b, int d)
{
    return reinterpret_cast<T *>(p.F[d]) + 5 * b.off_f[d];
}

/* add_scaled(c, slope, -1/2) and add_scaled(c, slope, +1/2)
 * (Reconstruction.hpp:59-69) of the MC-theta slope of c */
template <class T>
__device__ __forceinline__ void slope_step(const PrimT<T> &l,
                                           const PrimT<T> &c,
                                           const PrimT<T> &r, T theta,
                                           PrimT<T> &minus_half,
                                           PrimT<T> &plus_half)
{
    T s = limited_slope(l.rho, c.rho, r.rho, theta);
    minus_half.rho = c.rho + T(-0.5) * s;
    plus_half.rho = c.rho + T(0.5) * s;
    s = limited_slope(l.eps, c.eps, r.eps, theta);
    minus_half.eps = c.eps + T(-0.5) * s;
    plus_half.eps = c.eps + T(0.5) * s;
    s = limited_slope(l.pressure, c.pressure, r.pressure, theta);
    minus_half.pressure = c.pressure + T(-0.5) * s;
    plus_half.pressure = c.pressure + T(0.5) * s;
#pragma unroll
    for (int m = 0; m < 3; ++m)
    {
        s = limited_slope(l.v[m], c.v[m], r.v[m], theta);
        minus_half.v[m] = c.v[m] + T(-0.5) * s;
        plus_half.v[m] = c.v[m] + T(0.5) * s;
    }
}

template <class T, class O>
__device__ __forceinline__ void store_flux(O *F, long long nf, long long fo,
                                           const ConsT<T> &f)
{
    F[fo] = O(f.D);
    F[nf + fo] = O(f.S[0]);
    F[2 * nf + fo] = O(f.S[1]);
    F[3 * nf + fo] = O(f.S[2]);
    F[4 * nf + fo] = O(f.tau);
}

/* Which faces a face kernel writes, and as what.  FACES_ALL: every face,
 * stored as the compute type.  FACES_FP32, the first pass of the guarded
 * mixed path: the faces neither of whose cells carries kWGuard, stored as
 * double; the others go to the FP64 face list, which k_faces_list then
 * computes into the same double pool.  k_cells<double> reads that pool as
 * on the FP64 path. */
enum FaceSel
{
    FACES_ALL = 0,
    FACES_FP32 = 1
};

template <class T, int SEL> struct FluxOut
{
    using type = double;
};
template <class T> struct FluxOut<T, FACES_ALL>
{
    using type = T;
};

/* Whether the face between two cells with these flags is computed in
 * FP64 on the guarded path. */
__device__ __forceinline__ bool fp64_face(int left, int right)
{
    return ((left | right) & kWGuard) != 0;
}

__device__ __forceinline__ int2 list_entry(int box, int d, long long fo)
{
    return make_int2(box | (d << 30), (int)fo);
}

/* Append one entry per lane that takes one to `list` (length *count): a
 * single atomic per wavefront, entries in lane order from there.  Every
 * lane of the wavefront must reach it. */
__device__ __forceinline__ void list_append(unsigned long long *count,
                                            int2 *list, bool take,
                                            int2 entry)
{
    const unsigned long long bal = __ballot(take);
    if (bal == 0ull)
        return;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)bal) - 1;
    unsigned long long base = 0ull;
    if (lane == leader)
        base = atomicAdd(count, (unsigned long long)__popcll(bal));
    base = __shfl(base, leader, 64);
    if (take)
        list[base + __popcll(bal & ((1ull << lane) - 1ull))] = entry;
}

/* The FP64 face list. */
__device__ __forceinline__ void list_append_lane(const Pools &p, bool take,
                                                 int2 entry)
{
    list_append(&p.st->nlist, p.list, take, entry);
}

/* Append the faces of a walk whose bits are set (bit q: face fo0 +
 * q f_stride): one atomic per wavefront, after a scan of the counts.
 * Every lane of the wavefront must reach it. */
__device__ __forceinline__ void list_append_walk(const Pools &p,
                                                 unsigned bits, int box,
                                                 int d, long long fo0,
                                                 long long f_stride)
{
    if (!__any(bits != 0u))
        return;
    const int lane = (int)__lane_id();
    const int cnt = __popc(bits);
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += y;
    }
    const int total = __shfl(incl, 63, 64);
    if (total == 0)
        return;
    unsigned long long base = 0ull;
    if (lane == 63)
        base = atomicAdd(&p.st->nlist, (unsigned long long)total);
    base = __shfl(base, 63, 64);
    unsigned long long at = base + (unsigned long long)(incl - cnt);
    while (bits)
    {
        const int q = __ffs(bits) - 1;
        bits &= bits - 1u;
        p.list[at++] = list_entry(box, d, fo0 + q * f_stride);
    }
}

/* ------------------------------------------------------------------ */
/* k_faces_x: fluxes of direction 0, one lane per cell.  The cells    */
/* lo-1 .. hi+1 of every (j, k) row are laid end to end; a wavefront  */
/* takes 64 consecutive ones, overlapping the next wave by one.  Each */
/* lane evaluates its cell's metric and two half-cell states once;    */
/* the face between a lane and the next is formed with __shfl_down,   */
/* so loads are unit-stride along i and no metric is evaluated twice. */
/* ------------------------------------------------------------------ */
template <class T> __device__ __forceinline__ T shfl_next(T v)
{
    return __shfl_down(v, 1, 64);
}

template <class T, int SEL = FACES_ALL>
__device__ __forceinline__ void faces_x_block(const BoxDev *__restrict__ boxes,
                                              const int2 bt, const Options &o,
                                              const Background &bg,
                                              const Pools &p)
{
    using O = typename FluxOut<T, SEL>::type;
    const BoxDev &b = boxes[bt.x];
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)bt.y + threadIdx.x) >> 6;
    const long long e = wave * 63 + lane;
    const long long next = fx_cells(b);
    const int E = b.in[0] + 2;
    const bool live = e < next;
    const long long ee = live ? e : next - 1;
    const int x = (int)(ee % E);
    const long long row = ee / E;
    const int j = b.lo[1] + (int)(row % b.in[1]);
    const int k = b.lo[2] + (int)(row / b.in[1]);
    const int i = b.lo[0] - 1 + x;
    const long long n = b.gnpts;
    const double *W = p.W + 6 * b.off_g;
    const long long oc = goff(b, i, j, k);
    const bool makes_face =
        live && lane != 63 && x != E - 1 && e + 1 < next;
    const long long fo =
        ((long long)(k - b.lo[2]) * b.in[1] + (j - b.lo[1])) * (b.in[0] + 1) +
        x;
    bool mine = makes_face;
    if (SEL == FACES_FP32)
    {
        /* FP64 faces go to the list; a wavefront with none of its own
         * leaves at once (the exit is uniform, so the shuffles below see
         * whole wavefronts) */
        const int fc = p.wbad[b.off_g + oc];
        /* every lane shuffles (a lane that skipped it would leave its
         * neighbour an undefined value) */
        const int fn = shfl_next(fc);
        const bool to64 = makes_face && fp64_face(fc, fn);
        list_append_lane(p, to64, list_entry(bt.x, 0, fo));
        /* the guarded interior cells, for k_guard_sources: each interior
         * cell once, from the lane that holds it first */
        if (o.use_static_sources)
        {
            const bool owns = live && lane != 63 && x >= 1 && x <= E - 2;
            const long long ci =
                ((long long)(k - b.lo[2]) * b.in[1] + (j - b.lo[1])) *
                    b.in[0] + (i - b.lo[0]);
            list_append(&p.st->ncell, p.clist, owns && (fc & kWGuard) != 0,
                        make_int2(bt.x, (int)ci));
        }
        mine = makes_face && !to64;
        if (!__any(mine))
            return;
    }
    const PrimT<T> wc = load_prim_as<T>(W, n, oc);
    PrimT<T> mh = wc, ph = wc;
    if (o.use_reconstruction)
    {
        const PrimT<T> wl = load_prim_as<T>(W, n, oc - 1);
        const PrimT<T> wr = load_prim_as<T>(W, n, oc + 1);
        slope_step(wl, wc, wr, T(o.theta), mh, ph);
    }
    T xc[3];
    coords_as(i, j, k, b, xc);
    GeoT<T> gc;
    cell_geometry(bg, xc, gc);

    PrimT<T> rp;
    rp.rho = shfl_next(mh.rho);
    rp.eps = shfl_next(mh.eps);
    rp.pressure = shfl_next(mh.pressure);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        rp.v[m] = shfl_next(mh.v[m]);
    GeoT<T> gn;
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

    if (!mine)
        return;
    GeoT<T> face;
    average_geo(gc, gn, face);
    ConsT<T> f;
    hlle_face(ph, rp, o, face, 0, f);
    store_flux(face_pool<O>(p, b, 0), b.nface[0], fo, f);
}

template <class T, int SEL = FACES_ALL>
__global__ void __launch_bounds__(kBlock)
    k_faces_x(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o,
              Background bg, Pools p)
{
    faces_x_block<T, SEL>(boxes, items[blockIdx.x], o, bg, p);
}

/* ------------------------------------------------------------------ */
/* k_faces<D> (D = 1, 2): each thread walks b.seg consecutive faces    */
/* along D carrying the right cell's primitives, the next face's left  */
/* state and the right cell's metric, so a cell's slope and metric are */
/* computed once per walk; lanes are consecutive along i.              */
/* ------------------------------------------------------------------ */
template <int D, class T, int SEL = FACES_ALL>
__device__ __forceinline__ void faces_march_block(
    const BoxDev *__restrict__ boxes, const int2 bt, const Options &o,
    const Background &bg, const Pools &p)
{
    using O = typename FluxOut<T, SEL>::type;
    const BoxDev &b = boxes[bt.x];
    const long long it0 = (long long)bt.y + threadIdx.x;
    /* the FP32 pass keeps every lane to the list append below */
    const bool valid = it0 < class_items(b, D == 1 ? C_FY : C_FZ);
    if (SEL == FACES_ALL && !valid)
        return;
    const long long it = valid ? it0 : 0;
    constexpr int B = (D == 2) ? 1 : 2; /* the other slow axis */
    const int na = b.in[0], nb = b.in[B];
    const int ia = (int)(it % na);
    const int ib = (int)((it / na) % nb);
    const int sg = (int)(it / ((long long)na * nb));
    const int fnd = b.in[D] + 1;
    const int p0 = sg * b.seg;
    const int p1 = min(p0 + b.seg, fnd);
    const long long n = b.gnpts;
    const double *W = p.W + 6 * b.off_g;
    const long long stride =
        (D == 1) ? (long long)b.gn[0] : (long long)b.gn[0] * b.gn[1];
    int c[3];
    c[0] = b.lo[0] + ia;
    c[B] = b.lo[B] + ib;
    c[D] = b.lo[D] + p0; /* right cell of the first face */
    long long oR = goff(b, c[0], c[1], c[2]);
    const int fn0 = b.in[0], fn1 = b.in[1] + (D == 1);
    const long long f_stride =
        (D == 1) ? (long long)fn0 : (long long)fn0 * fn1;
    long long fo = ((long long)(c[2] - b.lo[2]) * fn1 + (c[1] - b.lo[1])) *
                       fn0 + (c[0] - b.lo[0]);
    const long long nf = b.nface[D];
    O *F = face_pool<O>(p, b, D);
    const T theta = T(o.theta);
    /* bit q - p0: face q is written here (every face when SEL is
     * FACES_ALL); the FP32 pass lists its FP64 faces, and a walk with none
     * of its own leaves at once */
    unsigned mine = ~0u;
    if (SEL == FACES_FP32)
    {
        unsigned to64 = 0u;
        if (valid)
        {
            const unsigned char *fl = p.wbad + b.off_g;
            int left = fl[oR - stride];
            for (int q = p0; q < p1; ++q)
            {
                const int right = fl[oR + (long long)(q - p0) * stride];
                if (fp64_face(left, right))
                    to64 |= 1u << (q - p0);
                left = right;
            }
        }
        list_append_walk(p, to64, bt.x, D, fo, f_stride);
        mine = valid ? ~to64 & ((1u << (p1 - p0)) - 1u) : 0u;
        if (mine == 0u)
            return;
    }

    PrimT<T> wL = load_prim_as<T>(W, n, oR - stride);
    PrimT<T> wR = load_prim_as<T>(W, n, oR);
    PrimT<T> lp = wL;
    if (o.use_reconstruction)
    {
        const PrimT<T> wLL = load_prim_as<T>(W, n, oR - 2 * stride);
        PrimT<T> unused;
        slope_step(wLL, wL, wR, theta, unused, lp);
    }
    T x[3];
    c[D] -= 1;
    coords_as(c[0], c[1], c[2], b, x);
    GeoT<T> gL;
    cell_geometry(bg, x, gL);
    for (int q = p0; q < p1; ++q)
    {
        c[D] += 1;
        PrimT<T> rp = wR, lp_next = wR;
        PrimT<T> wRR;
        if (o.use_reconstruction)
        {
            wRR = load_prim_as<T>(W, n, oR + stride);
            slope_step(wL, wR, wRR, theta, rp, lp_next);
        }
        coords_as(c[0], c[1], c[2], b, x);
        GeoT<T> gR, face;
        cell_geometry(bg, x, gR);
        average_geo(gL, gR, face);
        ConsT<T> f;
        hlle_face(lp, rp, o, face, D, f);
        if (SEL == FACES_ALL || ((mine >> (q - p0)) & 1u))
            store_flux(F, nf, fo, f);
        if (q + 1 < p1)
        {
            lp = lp_next;
            wL = wR;
            wR = o.use_reconstruction ? wRR
                                      : load_prim_as<T>(W, n, oR + stride);
            gL = gR;
            oR += stride;
            fo += f_stride;
        }
    }
}

template <int D, class T, int SEL = FACES_ALL>
__global__ void __launch_bounds__(kBlock)
    k_faces(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o, Background bg,
            Pools p)
{
    faces_march_block<D, T, SEL>(boxes, items[blockIdx.x], o, bg, p);
}

/* ------------------------------------------------------------------ */
/* k_faces_t<D> (D = 1, 2): fluxes of direction D by tiles, one lane   */
/* per cell.  A wavefront holds kTileI consecutive cells along i (lane */
/* bits 0-2) by kTileD consecutive cells along D (bits 3-5), starting  */
/* one cell below the tile's first face; each lane evaluates its       */
/* cell's metric and two half-cell states once, and the face between a */
/* lane and the lane one cell further along D is formed with           */
/* __shfl_down by kTileI.  Loads are unit-stride along i in rows of    */
/* kTileI cells; tiles along D overlap by one cell, so a cell's metric */
/* is evaluated kTileD / (kTileD - 1) times per direction.             */
/* ------------------------------------------------------------------ */
template <int D, class T, int SEL = FACES_ALL>
__device__ __forceinline__ void faces_tile_block(
    const BoxDev *__restrict__ boxes, const int2 bt, const Options &o,
    const Background &bg, const Pools &p)
{
    using O = typename FluxOut<T, SEL>::type;
    const BoxDev &b = boxes[bt.x];
    const long long wave = ((long long)bt.y + threadIdx.x) >> 6;
    /* the block's last waves may lie past the box's tiles (uniform per
     * wavefront, so the shuffles below see whole wavefronts) */
    if (wave >= face_tiles(b, D))
        return;
    constexpr int B = (D == 2) ? 1 : 2; /* the other slow axis */
    const int lane = threadIdx.x & 63;
    const int li = lane % kTileI, ld = lane / kTileI;
    const int nti = (b.in[0] + kTileI - 1) / kTileI;
    const int ntd = (b.in[D] + 1 + kTileD - 2) / (kTileD - 1);
    const int ti = (int)(wave % nti);
    const int td = (int)((wave / nti) % ntd);
    const int ib = (int)(wave / ((long long)nti * ntd));
    /* face f along D lies between cells lo - 1 + f and lo + f; this lane
     * holds cell lo - 1 + f and forms face f with the next row's lane */
    const int f = td * (kTileD - 1) + ld;
    const int iu = ti * kTileI + li; /* interior offset along i */
    const bool makes_face =
        iu < b.in[0] && ld < kTileD - 1 && f <= b.in[D];
    int c[3];
    c[0] = b.lo[0] + min(iu, b.in[0] - 1);
    c[B] = b.lo[B] + ib;
    c[D] = b.lo[D] - 1 + min(f, b.in[D] + 1); /* clamped to cell hi + 1 */
    const long long n = b.gnpts;
    const double *W = p.W + 6 * b.off_g;
    const long long stride =
        (D == 1) ? (long long)b.gn[0] : (long long)b.gn[0] * b.gn[1];
    const long long oc = goff(b, c[0], c[1], c[2]);
    const long long fo =
        (D == 1) ? ((long long)ib * (b.in[1] + 1) + f) * b.in[0] + iu
                 : ((long long)f * b.in[1] + ib) * b.in[0] + iu;
    bool mine = makes_face;
    if (SEL == FACES_FP32)
    {
        /* as in faces_x_block: every lane shuffles, a wavefront with no
         * FP32 face of its own leaves at once */
        const int fc = p.wbad[b.off_g + oc];
        const int fn = __shfl_down(fc, kTileI, 64);
        const bool to64 = makes_face && fp64_face(fc, fn);
        list_append_lane(p, to64, list_entry(bt.x, D, fo));
        mine = makes_face && !to64;
        if (!__any(mine))
            return;
    }
    const PrimT<T> wc = load_prim_as<T>(W, n, oc);
    PrimT<T> mh = wc, ph = wc;
    if (o.use_reconstruction)
    {
        const PrimT<T> wl = load_prim_as<T>(W, n, oc - stride);
        const PrimT<T> wr = load_prim_as<T>(W, n, oc + stride);
        slope_step(wl, wc, wr, T(o.theta), mh, ph);
    }
    T xc[3];
    coords_as(c[0], c[1], c[2], b, xc);
    GeoT<T> gc;
    cell_geometry(bg, xc, gc);

    PrimT<T> rp;
    rp.rho = __shfl_down(mh.rho, kTileI, 64);
    rp.eps = __shfl_down(mh.eps, kTileI, 64);
    rp.pressure = __shfl_down(mh.pressure, kTileI, 64);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        rp.v[m] = __shfl_down(mh.v[m], kTileI, 64);
    GeoT<T> gn;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int e = 0; e < 3; ++e)
        {
            gn.g[a][e] = __shfl_down(gc.g[a][e], kTileI, 64);
            gn.gu[a][e] = __shfl_down(gc.gu[a][e], kTileI, 64);
        }
    gn.lapse = __shfl_down(gc.lapse, kTileI, 64);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        gn.shift[m] = __shfl_down(gc.shift[m], kTileI, 64);

    if (!mine)
        return;
    GeoT<T> face;
    average_geo(gc, gn, face);
    ConsT<T> fl;
    hlle_face(ph, rp, o, face, D, fl);
    store_flux(face_pool<O>(p, b, D), b.nface[D], fo, fl);
}

template <int D, class T, int SEL = FACES_ALL>
__global__ void __launch_bounds__(kBlock)
    k_faces_t(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o,
              Background bg, Pools p)
{
    faces_tile_block<D, T, SEL>(boxes, items[blockIdx.x], o, bg, p);
}

/* The three directions in one launch with the tile kernels for y and z:
 * x blocks, then y tile blocks, then z tile blocks. */
template <class T, int SEL = FACES_ALL>
__global__ void __launch_bounds__(kBlock)
    k_faces_all_t(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, int nbx,
                  int nby, Options o, Background bg, Pools p)
{
    const int2 bt = items[blockIdx.x];
    if ((int)blockIdx.x < nbx)
        faces_x_block<T, SEL>(boxes, bt, o, bg, p);
    else if ((int)blockIdx.x < nbx + nby)
        faces_tile_block<1, T, SEL>(boxes, bt, o, bg, p);
    else
        faces_tile_block<2, T, SEL>(boxes, bt, o, bg, p);
}

/* All three face directions of a call in one launch, for levels too small
 * to fill the device with one direction at a time: the blocks of the
 * C_FALL table are the x blocks, then the y blocks, then the z blocks,
 * so every wavefront runs one direction's code.  The kernel needs the
 * registers of the marching kernels, which costs the x blocks occupancy;
 * on small levels that matters less than the two launch boundaries it
 * removes. */
template <class T, int SEL = FACES_ALL>
__global__ void __launch_bounds__(kBlock)
    k_faces_all(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, int nbx, int nby,
                Options o, Background bg, Pools p)
{
    const int2 bt = items[blockIdx.x];
    if ((int)blockIdx.x < nbx)
        faces_x_block<T, SEL>(boxes, bt, o, bg, p);
    else if ((int)blockIdx.x < nbx + nby)
        faces_march_block<1, T, SEL>(boxes, bt, o, bg, p);
    else
        faces_march_block<2, T, SEL>(boxes, bt, o, bg, p);
}

/* The FP64 faces of the guarded mixed path, one face per list entry: the
 * same loads, slopes, metric evaluations, average and HLLE as a walk of
 * one face (or the x kernel's lane pair), so each flux equals the FP64
 * path's bit for bit for the same W. */
template <int D>
__device__ __forceinline__ void face_one(const BoxDev &b, long long fo,
                                         const Options &o,
                                         const Background &bg,
                                         const Pools &p)
{
    const int fn0 = b.in[0] + (D == 0), fn1 = b.in[1] + (D == 1);
    int c[3]; /* the face's right cell */
    c[0] = b.lo[0] + (int)(fo % fn0);
    c[1] = b.lo[1] + (int)((fo / fn0) % fn1);
    c[2] = b.lo[2] + (int)(fo / ((long long)fn0 * fn1));
    const long long n = b.gnpts;
    const double *W = p.W + 6 * b.off_g;
    const long long stride =
        (D == 0) ? 1ll
                 : ((D == 1) ? (long long)b.gn[0]
                             : (long long)b.gn[0] * b.gn[1]);
    const long long oR = goff(b, c[0], c[1], c[2]);
    const PrimT<double> wL = load_prim_as<double>(W, n, oR - stride);
    const PrimT<double> wR = load_prim_as<double>(W, n, oR);
    PrimT<double> lp = wL, rp = wR;
    if (o.use_reconstruction)
    {
        const PrimT<double> wLL = load_prim_as<double>(W, n, oR - 2 * stride);
        const PrimT<double> wRR = load_prim_as<double>(W, n, oR + stride);
        PrimT<double> unused;
        slope_step(wLL, wL, wR, o.theta, unused, lp);
        slope_step(wL, wR, wRR, o.theta, rp, unused);
    }
    double x[3];
    GeoT<double> gL, gR, face;
    c[D] -= 1;
    coords_as(c[0], c[1], c[2], b, x);
    cell_geometry(bg, x, gL);
    c[D] += 1;
    coords_as(c[0], c[1], c[2], b, x);
    cell_geometry(bg, x, gR);
    average_geo(gL, gR, face);
    ConsT<double> f;
    hlle_face(lp, rp, o, face, D, f);
    store_flux(face_pool<double>(p, b, D), b.nface[D], fo, f);
}

/* Grid-stride over the FP64 face list (DevStatus::nlist entries). */
__global__ void __launch_bounds__(kBlock)
    k_faces_list(const BoxDev *__restrict__ boxes, Options o, Background bg,
                 Pools p)
{
    const unsigned long long cnt = p.st->nlist;
    for (unsigned long long q = (unsigned long long)blockIdx.x * kBlock +
                                threadIdx.x;
         q < cnt; q += (unsigned long long)gridDim.x * kBlock)
    {
        const int2 e = p.list[q];
        const int d = (int)((unsigned)e.x >> 30);
        const BoxDev &b = boxes[e.x & ((1 << 30) - 1)];
        if (d == 0)
            face_one<0>(b, e.y, o, bg, p);
        else if (d == 1)
            face_one<1>(b, e.y, o, bg, p);
        else
            face_one<2>(b, e.y, o, bg, p);
    }
}

/* Empty the FP64 face list and the guarded-cell list before a stage's
 * FP32 pass. */
__global__ void k_list_reset(DevStatus *st)
{
    if (threadIdx.x == 0)
    {
        st->nlist = 0ull;
        st->ncell = 0ull;
    }
}

/* Mixed precision: the static-metric sources of a cell in FP32 (the
 * metric's derivatives, K_ij and the source sums), from the FP64 metric,
 * floored primitives and state rounded once; the result is added to the
 * FP64 right-hand side.  Guarded cells keep FP64 sources, like their faces.
 * The binary's volume-time sources (a difference of log sqrt(gamma) at
 * t +- h over 2h) stay FP64. */
template <bool S32>
__device__ __forceinline__ bool sources_in_fp32(const Options &o,
                                                const Background &bg,
                                                unsigned char wflags)
{
    if constexpr (S32)
        return o.fp32_sources && !(wflags & kWGuard);
    else
        return false;
}

/* The derivative half runs in FP32 on the cell's FP64 geometry half
 * (metric, inverse, lapse, shift, and the Kerr-Schild radius, H and l_i
 * or the binary's potentials), rounded once: the cell has evaluated that
 * half in FP64 already for its floors, so it is not evaluated again. */
__device__ __forceinline__ GeoT<float> round_geo(const Geo &g)
{
    GeoT<float> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            r.g[i][j] = float(g.g[i][j]);
            r.gu[i][j] = float(g.gu[i][j]);
        }
        r.shift[i] = float(g.shift[i]);
    }
    r.lapse = float(g.lapse);
    return r;
}

__device__ __forceinline__ ConsT<float>
static_sources_fp32(const Background &bg, const double x[3],
                    const GeoCore &core, const Geo &geo, const Prim &pw,
                    const Cons &us, double gamma)
{
    const float xf[3] = {float(x[0]), float(x[1]), float(x[2])};
    const GeoT<float> g = round_geo(geo);
    FullGeoT<float> f;
    if (bg.kind == 1)
    {
        const Potentials &pd = core.pn;
        PotentialsT<float> pot;
        pot.scalar = float(pd.scalar);
        pot.scalar_dt = float(pd.scalar_dt);
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            pot.dscalar[i] = float(pd.dscalar[i]);
            pot.vec[i] = float(pd.vec[i]);
#pragma unroll
            for (int j = 0; j < 3; ++j)
                pot.dvec[i][j] = float(pd.dvec[i][j]);
        }
        pn_derivs(pot, g, f);
    }
    else
    {
        const KSCore &kd = core.ks;
        KSCoreT<float> kc;
        kc.rho = float(kd.rho);
        kc.rho2 = float(kd.rho2);
        kc.r2 = float(kd.r2);
        kc.r = float(kd.r);
        kc.cos_theta = float(kd.cos_theta);
        kc.H = float(kd.H);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            kc.el[i] = float(kd.el[i]);
        ks_derivs(float(bg.M), float(bg.a), xf, kc, g, f);
    }
    PrimT<float> pf;
    pf.rho = float(pw.rho);
    pf.eps = float(pw.eps);
    pf.pressure = float(pw.pressure);
    ConsT<float> uf;
    uf.D = float(us.D);
    uf.tau = float(us.tau);
#pragma unroll
    for (int m = 0; m < 3; ++m)
    {
        pf.v[m] = float(pw.v[m]);
        uf.S[m] = float(us.S[m]);
    }
    ConsT<float> src;
    static_sources(pf, uf, float(gamma), f, src);
    return src;
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
    int us;          /* pool whose RHS is evaluated */
    int u0;          /* RK: step start / SSP: old state */
    int unext;       /* pool receiving the stage result */
    int store_prims; /* write floored primitives into the state FAB */
    int write_mask;
    int accumulate;  /* MODE_RHS: add the conserved rhs to the rhs FAB */
    int count;       /* add recovery counts */
    double c_stage;  /* U0 coefficient of k */
    double c_acc;    /* accumulator coefficient of k */
    double dt;
    double vol_h;    /* max(step, 1e-12) */
    unsigned long long next_eval;
};

/* The guarded mixed path's FP64 static sources: for each interior cell
 * that the FP32 face pass listed (its recovery set kWGuard) and that lies
 * outside the metric-source atmosphere, the same geometry, floors, metric
 * derivatives and sources as the FP64 path's k_cells (:255-298), into
 * Pools::src, which k_cells then adds where the FP64 path adds its own.
 * Grid-stride over the DevStatus::ncell entries. */
__global__ void __launch_bounds__(kBlock)
    k_guard_sources(const BoxDev *__restrict__ boxes, Options o,
                    Background bg_s, Pools p, int us_pool)
{
    const unsigned long long cnt = p.st->ncell;
    for (unsigned long long q = (unsigned long long)blockIdx.x * kBlock +
                                threadIdx.x;
         q < cnt; q += (unsigned long long)gridDim.x * kBlock)
    {
        const int2 e = p.clist[q];
        const BoxDev &b = boxes[e.x];
        const long long it = e.y;
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        const long long so = goff(b, i, j, k);
        const long long n = b.gnpts;
        const double *W = p.W + 6 * b.off_g;
        double x[3];
        coords(i, j, k, b.dx, b.center, x);
        Geo geo_s;
        GeoCore core;
        cell_geometry_core(bg_s, x, geo_s, core);
        Prim pw = load_prim(W, n, so);
        floors(pw, o, geo_s.g);
        if (is_atmosphere(pw, o, o.source_atmosphere_factor))
            continue;
        const Cons us = load_cons(pool_cons(p, us_pool, b), n, so);
        FullGeo fg;
        cell_derivs(bg_s, x, core, geo_s, fg);
        Cons src;
        static_sources(pw, us, o.gamma, fg, src);
        store_cons(p.src + 5 * b.off_i, b.inpts, it, src);
    }
}

/* ------------------------------------------------------------------ */
/* k_cells<MODE>: one thread per interior cell.                        */
/* ------------------------------------------------------------------ */
/* FT: the type of the face-flux pool.  S32: the mixed paths' FP32
 * sources may run (sources_in_fp32); false on the FP64 path, whose
 * instantiations carry no FP32 code. */
template <int MODE, class FT, bool S32>
__device__ __forceinline__ void
cells_block(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items,
            Options o, Background bg_s, Background bg_next,
            Background bg_plus, Background bg_minus, Pools p, CellArgs a)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_INT, bp);
    const BoxDev &b = *bp;
    const bool dead = failed_before(p.st, a.next_eval);
    unsigned long long cnt[4] = {0, 0, 0, 0};
    if (it >= 0)
    {
        const int ii = (int)(it % b.in[0]);
        const int jj = (int)((it / b.in[0]) % b.in[1]);
        const int kk = (int)(it / ((long long)b.in[0] * b.in[1]));
        const int i = b.lo[0] + ii, j = b.lo[1] + jj, k = b.lo[2] + kk;
        const long long n = b.gnpts;
        const long long so = goff(b, i, j, k);
        double *W = p.W + 6 * b.off_g;

        /* flux divergence: rhs = ((0 + r_x) + r_y) + r_z */
        Cons rhs;
        rhs.D = 0.0;
        rhs.tau = 0.0;
        rhs.S[0] = rhs.S[1] = rhs.S[2] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double mi = -b.inv_dx[d];
            const FT *F = face_pool<const FT>(p, b, d);
            const long long nf = b.nface[d];
            const int fn0 = b.in[0] + (d == 0), fn1 = b.in[1] + (d == 1);
            const long long flo = ((long long)kk * fn1 + jj) * fn0 + ii;
            const long long fhi =
                flo + ((d == 0) ? 1
                                : ((d == 1) ? (long long)fn0
                                            : (long long)fn0 * fn1));
            rhs.D += mi * ((double)F[fhi] - (double)F[flo]);
            rhs.tau += mi * ((double)F[4 * nf + fhi] - (double)F[4 * nf + flo]);
#pragma unroll
            for (int m = 0; m < 3; ++m)
                rhs.S[m] += mi * ((double)F[(1 + m) * nf + fhi] -
                                  (double)F[(1 + m) * nf + flo]);
        }

        const Cons us = load_cons(pool_cons(p, a.us, b), n, so);
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
            Prim pw = load_prim(W, n, so);
            floors(pw, o, geo_s.g);
            if (is_atmosphere(pw, o, o.source_atmosphere_factor))
            {
                mbits |= GRHD_MASK_SOURCE_ATMOSPHERE;
            }
            else if (sources_in_fp32<S32>(o, bg_s, p.wbad[b.off_g + so]))
            {
                const ConsT<float> src = static_sources_fp32(
                    bg_s, x, core, geo_s, pw, us, o.gamma);
                rhs.D += (double)src.D;
                rhs.tau += (double)src.tau;
                rhs.S[0] += (double)src.S[0];
                rhs.S[1] += (double)src.S[1];
                rhs.S[2] += (double)src.S[2];
            }
            else if constexpr (!S32)
            {
                FullGeo fg;
                cell_derivs(bg_s, x, core, geo_s, fg);
                Cons src;
                static_sources(pw, us, o.gamma, fg, src);
                rhs.D += src.D;
                rhs.tau += src.tau;
                rhs.S[0] += src.S[0];
                rhs.S[1] += src.S[1];
                rhs.S[2] += src.S[2];
            }
            else
            {
                /* a guarded cell of the mixed path: its FP64 sources,
                 * computed the FP64 path's way by k_guard_sources, which
                 * keeps the derivative half of the metric out of this
                 * kernel's registers */
                const Cons src = load_cons(p.src + 5 * b.off_i, b.inpts, it);
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
            const double dlog = qdiv(lp - lm, 2.0 * a.vol_h);
            rhs.D += 0.0 - us.D * dlog;
            rhs.tau += 0.0 - us.tau * dlog;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                rhs.S[m] += 0.0 - us.S[m] * dlog;
        }
        /* ExcisionTest (ExcisionTest.hpp:40-63): after the matter rhs */
        if (o.excise &&
            ks_excised(bg_s.M, bg_s.a, x, o.excise_inner, o.excise_outer))
        {
            rhs.D = 0.0;
            rhs.tau = 0.0;
            rhs.S[0] = rhs.S[1] = rhs.S[2] = 0.0;
        }

        if (MODE == MODE_RHS)
        {
            if (!dead)
            {
                const long long ni = b.inpts;
                double *r =
                    p.rhs + p.rhs_ncomp * b.off_i + (long long)p.cb * ni;
                if (a.accumulate)
                {
                    /* add_conserved_rhs (FArrayBoxFiniteVolume.hpp:912-925) */
                    r[it] += rhs.D;
                    r[ni + it] += rhs.S[0];
                    r[2 * ni + it] += rhs.S[1];
                    r[3 * ni + it] += rhs.S[2];
                    r[4 * ni + it] += rhs.tau;
                }
                else
                {
                    r[it] = rhs.D;
                    r[ni + it] = rhs.S[0];
                    r[2 * ni + it] = rhs.S[1];
                    r[3 * ni + it] = rhs.S[2];
                    r[4 * ni + it] = rhs.tau;
                }
                if (a.write_mask)
                    p.mask[b.off_i + it] = mbits;
            }
        }
        else
        {
            /* stage update */
            Cons un;
            const long long ni = b.inpts;
            double *acc = p.acc + 5 * b.off_i;
            if (MODE == MODE_RK4_1 || MODE == MODE_RK4_2 ||
                MODE == MODE_RK4_3)
            {
                const Cons u0 = (MODE == MODE_RK4_1)
                                    ? us
                                    : load_cons(pool_cons(p, a.u0, b), n, so);
                Cons ac = (MODE == MODE_RK4_1) ? u0 : load_cons(acc, ni, it);
                ac.D += a.c_acc * rhs.D;
                ac.tau += a.c_acc * rhs.tau;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    ac.S[m] += a.c_acc * rhs.S[m];
                store_cons(acc, ni, it, ac);
                un.D = u0.D + a.c_stage * rhs.D;
                un.tau = u0.tau + a.c_stage * rhs.tau;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    un.S[m] = u0.S[m] + a.c_stage * rhs.S[m];
            }
            else if (MODE == MODE_RK4_4)
            {
                const Cons ac = load_cons(acc, ni, it);
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
                const Cons old = load_cons(pool_cons(p, a.u0, b), n, so);
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
            if (dead)
            {
                cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
            }
            else
            {
                double *Un = pool_cons(p, a.unext, b);
                if (a.store_prims)
                    store_prim(Un + 5 * n, n, so, prim);
                Prim wn = rec;
                bool wok = ok;
                if (reset)
                {
                    cnt[3] = 1;
                    m |= GRHD_MASK_RESET;
                    prim2con(prim, o.gamma, g.g, un);
                    wok = con2prim(un, o, g.gu, wn, res);
                }
                store_cons(Un, n, so, un);
                unsigned char *wbad = p.wbad + b.off_g;
                if (wok)
                {
                    store_prim(W, n, so, wn);
                    wbad[so] = recovered_flags(wn, o, g);
                }
                else
                {
                    atomicMin(&p.st->fail_key, fail_key(a.next_eval, 1, b, so));
                    W[so] = res;
                    wbad[so] = failed_flags(o);
                }
                if (a.write_mask)
                    p.mask[b.off_i + it] = m;
            }
        }
    }
    if (MODE != MODE_RHS && a.count)
        block_add_counts(p.st, cnt[0], cnt[1], cnt[2], cnt[3]);
}

/* The FP64 path's cell kernel. */
template <int MODE, class FT, bool S32 = (sizeof(FT) < sizeof(double))>
__global__ void __launch_bounds__(kBlock)
    k_cells(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o,
            Background bg_s, Background bg_next, Background bg_plus,
            Background bg_minus, Pools p, CellArgs a)
{
    cells_block<MODE, FT, S32>(boxes, items, o, bg_s, bg_next, bg_plus,
                               bg_minus, p, a);
}

/* The mixed paths' cell kernel (FP32 sources; plain mixed with a float
 * flux pool, guarded mixed with a double one), held to 168 VGPRs, three
 * waves per SIMD: without its FP64 metric derivatives it needs about that
 * many, and the compiler's own choice of a few more would halve the
 * waves. */
template <int MODE, class FT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
    k_cells_mixed(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items,
                  Options o, Background bg_s, Background bg_next,
                  Background bg_plus, Background bg_minus, Pools p, CellArgs a)
{
    cells_block<MODE, FT, true>(boxes, items, o, bg_s, bg_next, bg_plus,
                                bg_minus, p, a);
}

/* Post-update recovery of the interior of the state pool. */
__global__ void __launch_bounds__(kBlock)
    k_recover(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o,
              Background bg, Pools p, int write_mask)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_INT, bp);
    const BoxDev &b = *bp;
    unsigned long long cnt[4] = {0, 0, 0, 0};
    if (it >= 0)
    {
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        const long long n = b.gnpts;
        const long long so = goff(b, i, j, k);
        double *U = pool_cons(p, POOL_STATE, b);
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
        store_prim(U + 5 * n, n, so, prim);
        if (reset)
        {
            cnt[3] = 1;
            m |= GRHD_MASK_RESET;
            prim2con(prim, o.gamma, g.g, c);
            store_cons(U, n, so, c);
        }
        if (write_mask)
            p.mask[b.off_i + it] = m;
    }
    block_add_counts(p.st, cnt[0], cnt[1], cnt[2], cnt[3]);
}

/* add_scaled_conserved_rhs (FArrayBoxFiniteVolume.hpp:927-940) */
__global__ void __launch_bounds__(kBlock)
    k_add_scaled(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Pools p,
                 double scale)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_INT, bp);
    const BoxDev &b = *bp;
    if (it < 0)
        return;
    const int i = b.lo[0] + (int)(it % b.in[0]);
    const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
    const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
    const long long n = b.gnpts, ni = b.inpts;
    double *U = pool_cons(p, POOL_STATE, b);
    const double *r = p.rhs + p.rhs_ncomp * b.off_i + (long long)p.cb * ni;
    const long long so = goff(b, i, j, k);
#pragma unroll
    for (int c = 0; c < 5; ++c)
        U[c * n + so] += scale * r[c * ni + it];
}

/* rhs.setVal(0.0) over every component of every box */
__global__ void __launch_bounds__(kBlock)
    k_zero_rhs(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Pools p)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_INT, bp);
    const BoxDev &b = *bp;
    if (it < 0)
        return;
    double *r = p.rhs + p.rhs_ncomp * b.off_i;
    for (int c = 0; c < p.rhs_ncomp; ++c)
        r[c * b.inpts + it] = 0.0;
}

/* max over non-atmosphere cells of sum_d max|lambda| / dx */
__global__ void __launch_bounds__(kBlock)
    k_max_inv_dt(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Options o,
                 Background bg, Pools p)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_INT, bp);
    const BoxDev &b = *bp;
    double best = 0.0;
    if (it >= 0)
    {
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        const long long so = goff(b, i, j, k);
        double x[3];
        coords(i, j, k, b.dx, b.center, x);
        Geo g;
        cell_geometry(bg, x, g);
        const Cons c = load_cons(pool_cons(p, POOL_STATE, b), b.gnpts, so);
        Prim pr;
        double res;
        if (!con2prim(c, o, g.gu, pr, res))
        {
            const unsigned long long key =
                (unsigned long long)(b.off_i + it);
            atomicMin(&p.st->cfl_key, key);
            p.W[key] = res;
        }
        else
        {
            floors(pr, o, g.g);
            if (!is_atmosphere(pr, o, 1.0 + 1.0e-12))
            {
                double cell = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    double lo, hi;
                    char_speeds(pr, o.gamma, g, d, lo, hi);
                    const double ms = std_max(fabs(lo), fabs(hi));
                    cell += ms * b.inv_dx[d];
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
            atomicMax(&p.st->max_bits, m);
    }
}

/* analytic cell geometry over the storage box of box 0 */
__global__ void __launch_bounds__(kBlock)
    k_geometry(const BoxDev *__restrict__ boxes, const int2 *__restrict__ items, Background bg,
               double *out)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_ALL, bp);
    const BoxDev &b = *bp;
    if (it < 0)
        return;
    const int i = b.glo[0] + (int)(it % b.gn[0]);
    const int j = b.glo[1] + (int)((it / b.gn[0]) % b.gn[1]);
    const int k = b.glo[2] + (int)(it / ((long long)b.gn[0] * b.gn[1]));
    double x[3];
    coords(i, j, k, b.dx, b.center, x);
    Geo g;
    cell_geometry(bg, x, g);
    const long long n = b.gnpts;
    double *o = out + 16 * b.off_g;
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int m = 0; m < 6; ++m)
    {
        o[m * n + it] = g.g[ii[m]][jj[m]];
        o[(6 + m) * n + it] = g.gu[ii[m]][jj[m]];
    }
    o[12 * n + it] = g.lapse;
#pragma unroll
    for (int m = 0; m < 3; ++m)
        o[(13 + m) * n + it] = g.shift[m];
}

/* ------------------------------------------------------------------ */
/* Level diagnostics: sums of D, S_1, S_2, S_3, tau over every interior */
/* cell and their finiteness (LevelDataDriver.hpp:22-92).  The         */
/* reference adds sequentially; here every cell value enters a         */
/* double-double accumulator (Knuth's TwoSum, the QD library's         */
/* accurate addition) in a fixed order -- wavefront butterfly, waves   */
/* of a block in order, blocks in order -- so the result is the exact  */
/* sum rounded once in all but contrived cases, and the same on every  */
/* run.  A non-finite cell makes the plain sum the result.             */
/* ------------------------------------------------------------------ */
struct DD
{
    double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ DD fast_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

constexpr int kDiagWords = 16; /* per block: hi[5], lo[5], plain[5], bad */

__device__ __forceinline__ DD shfl_xor_dd(DD v, int off)
{
    return {__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
}

/* Reduce (acc[5], plain[5], bad) over the block in a fixed order; thread 0
 * returns the block's values.  dd_add is commutative bit for bit (TwoSum
 * is exact in either order), so both lanes of a butterfly pair agree. */
__device__ inline void block_reduce_diag(DD acc[5], double plain[5],
                                         double &bad)
{
    __shared__ double part[kBlock / 64][kDiagWords];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            acc[m] = dd_add(acc[m], shfl_xor_dd(acc[m], off));
            plain[m] += __shfl_xor(plain[m], off, 64);
        }
        bad += __shfl_xor(bad, off, 64);
    }
    if (lane == 0)
    {
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            part[wave][m] = acc[m].hi;
            part[wave][5 + m] = acc[m].lo;
            part[wave][10 + m] = plain[m];
        }
        part[wave][15] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < kBlock / 64; ++w)
        {
#pragma unroll
            for (int m = 0; m < 5; ++m)
            {
                acc[m] = dd_add(acc[m], DD{part[w][m], part[w][5 + m]});
                plain[m] += part[w][10 + m];
            }
            bad += part[w][15];
        }
    }
}

/* Per block of interior cells: partial sums into diag[block * 16 ..]. */
__global__ void __launch_bounds__(kBlock)
    k_diag_cells(const BoxDev *__restrict__ boxes,
                 const int2 *__restrict__ items, Pools p, double *diag)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_INT, bp);
    const BoxDev &b = *bp;
    DD acc[5];
    double plain[5];
    double bad = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m)
    {
        acc[m] = DD{0.0, 0.0};
        plain[m] = 0.0;
    }
    if (it >= 0)
    {
        const int i = b.lo[0] + (int)(it % b.in[0]);
        const int j = b.lo[1] + (int)((it / b.in[0]) % b.in[1]);
        const int k = b.lo[2] + (int)(it / ((long long)b.in[0] * b.in[1]));
        const Cons c =
            load_cons(pool_cons(p, POOL_STATE, b), b.gnpts, goff(b, i, j, k));
        const double v[5] = {c.D, c.S[0], c.S[1], c.S[2], c.tau};
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            acc[m].hi = v[m];
            plain[m] = v[m];
            if (!isfinite(v[m]))
                bad = 1.0;
        }
    }
    block_reduce_diag(acc, plain, bad);
    if (threadIdx.x == 0)
    {
        double *d = diag + (long long)blockIdx.x * kDiagWords;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            d[m] = acc[m].hi;
            d[5 + m] = acc[m].lo;
            d[10 + m] = plain[m];
        }
        d[15] = bad;
    }
}

/* One block: the partials of `nblocks` blocks, thread t taking blocks
 * t, t + 256, ... in order, then the block reduction. */
__global__ void __launch_bounds__(kBlock)
    k_diag_finish(const double *diag, long long nblocks, DevStatus *st)
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
    {
        const double *d = diag + q * kDiagWords;
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            acc[m] = dd_add(acc[m], DD{d[m], d[5 + m]});
            plain[m] += d[10 + m];
        }
        bad += d[15];
    }
    block_reduce_diag(acc, plain, bad);
    if (threadIdx.x == 0)
    {
        bool finite = true;
#pragma unroll
        for (int m = 0; m < 5; ++m)
            finite = finite && isfinite(acc[m].hi) && isfinite(plain[m]);
#pragma unroll
        for (int m = 0; m < 5; ++m)
            st->sums[m] = finite ? acc[m].hi : plain[m];
        st->nonfinite = (unsigned long long)bad;
    }
}

/* ------------------------------------------------------------------ */
/* KITPCHK1 payload (Main_GRHDLevelDataBinaryPNKITPDisk.cpp:1348-1375,  */
/* :1385-1425, :1516-1586): the ProblemDomain grown by the ghost width, */
/* point-major with the ncomp components of a cell adjacent:            */
/*   global[((k * ny + j) * nx + i) * ncomp + c],                       */
/* indices relative to the grown box.  Pack writes every storage cell   */
/* whose index clamped to the domain lies in the box's interior (so     */
/* interior cells and the physical ghosts beside them, once each);      */
/* unpack fills every storage cell of every box.                        */
/* ------------------------------------------------------------------ */
struct CkptGrid
{
    int lo[3];  /* corner of the grown domain */
    int n[3];   /* its extent                 */
    int dlo[3]; /* ProblemDomain              */
    int dhi[3];
};

__device__ __forceinline__ long long ckpt_cell(const CkptGrid &g, int i,
                                               int j, int k)
{
    return ((long long)(k - g.lo[2]) * g.n[1] + (j - g.lo[1])) * g.n[0] +
           (i - g.lo[0]);
}

__device__ __forceinline__ void storage_cell(const BoxDev &b, long long it,
                                             int &i, int &j, int &k)
{
    i = b.glo[0] + (int)(it % b.gn[0]);
    j = b.glo[1] + (int)((it / b.gn[0]) % b.gn[1]);
    k = b.glo[2] + (int)(it / ((long long)b.gn[0] * b.gn[1]));
}

__global__ void __launch_bounds__(kBlock)
    k_fill(double *x, long long n, double v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        x[t] = v;
}

__global__ void __launch_bounds__(kBlock)
    k_ckpt_pack(const BoxDev *__restrict__ boxes,
                const int2 *__restrict__ items, const double *state,
                int ncomp, CkptGrid g, double *global)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_ALL, bp);
    const BoxDev &b = *bp;
    if (it < 0)
        return;
    int i, j, k;
    storage_cell(b, it, i, j, k);
    const int iv[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const int o = iv[d] < g.dlo[d] ? g.dlo[d]
                                       : (iv[d] > g.dhi[d] ? g.dhi[d] : iv[d]);
        if (o < b.lo[d] || o > b.hi[d])
            return;
    }
    const double *src = state + b.off_u + goff(b, i, j, k);
    double *dst = global + ckpt_cell(g, i, j, k) * ncomp;
    for (int c = 0; c < ncomp; ++c)
        dst[c] = src[(long long)c * b.gnpts];
}

__global__ void __launch_bounds__(kBlock)
    k_ckpt_unpack(const BoxDev *__restrict__ boxes,
                  const int2 *__restrict__ items, const double *global,
                  int ncomp, CkptGrid g, double *state)
{
    const BoxDev *bp;
    const long long it = block_item(items, boxes, C_ALL, bp);
    const BoxDev &b = *bp;
    if (it < 0)
        return;
    int i, j, k;
    storage_cell(b, it, i, j, k);
    const double *src = global + ckpt_cell(g, i, j, k) * ncomp;
    double *dst = state + b.off_u + goff(b, i, j, k);
    for (int c = 0; c < ncomp; ++c)
        dst[(long long)c * b.gnpts] = src[c];
}

} // namespace grhd

#endif /* GRHD_STAGE_KERNELS_HPP_ */
