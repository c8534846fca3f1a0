/*
 * grhd_box_table.hpp -- the level's box table and work-item classes, shared
 * by the kernels (device) and the plan builder (host; grhd_plan.hpp, which
 * a host compiler also builds for the CPU tests).  No device intrinsics.
 */
#ifndef GRHD_BOX_TABLE_HPP_
#define GRHD_BOX_TABLE_HPP_

#include <hip/hip_runtime.h>

namespace grhd
{

constexpr int kBlock = 256;
constexpr unsigned long long kNoFail = ~0ull;
constexpr int kKeyBits = 43; /* 2 x (cells of a level) < 2^43 */

/* Tiles of the transverse face kernels (k_faces_t, BoxDev::seg ==
 * kSegTile): a wavefront holds kTileI cells along i by kTileD cells along
 * the face direction d and forms the kTileD - 1 faces between them in
 * every column; tiles along d overlap by one cell. */
constexpr int kTileI = 8, kTileD = 8;
constexpr int kSegTile = 0;

/* Interior cells of a level below which the y/z faces are formed in tiles
 * (and, in grhd_kernels.hip, the three face directions run as one fused
 * launch). */
constexpr long long kTileCells = 3ll << 20;

/* The y/z face launch of a level of `cells` interior cells: kSegTile for
 * the tile kernels (k_faces_t: every cell its own lane, 8 / 7 metric
 * evaluations per face) below kTileCells, else walks of 8 faces per
 * thread of the marching kernels (k_faces: each cell's metric and slopes
 * shared by the two faces it touches, 9 / 8 evaluations per face, but
 * sequential along the walk).  Measured in RK4 steps of the Kerr torus
 * (HIP events, face launch per stage, FP64): 64^3 0.038 ms tiles against
 * 0.045 for walks of 4, 96^3 0.099 against 0.115 for walks of 8, 128^3
 * 0.223 against 0.234; FP32 tiles equal to walks of 8 at 128^3 and 4%
 * faster over the step at 96^3.  On the 192^3 and 256^3 boxes and the
 * 6M-cell box-set levels the walks of 8 were faster in the steps (the
 * tiles lost 3-4% there).  The fluxes do not depend on the choice. */
__host__ inline int march_segment(long long cells)
{
    return cells < kTileCells ? kSegTile : 8;
}

/* Device-side status block of one call.  The failure key, which every
 * wavefront of the stage kernels reads, has a 128-byte line of its own,
 * apart from the counters that every block updates with atomics. */
struct DevStatus
{
    /* first RHS recovery failure: (evaluation << 43) | position, where a
     * box owns positions [2 off_g, 2 (off_g + gnpts)): its ghost-fill
     * failures first, then its recovery failures, each in BoxIterator
     * order -- the cell the reference would throw at first */
    alignas(128) unsigned long long fail_key;
    alignas(128) unsigned long long cfl_key; /* off_i + interior index */
    unsigned long long counts[4];
    unsigned long long max_bits;
    double sums[5];              /* level conserved sums */
    unsigned long long nonfinite; /* interior cells with a non-finite value */
    /* entries of the guarded mixed path's FP64 face list (Pools::list)
     * and of its guarded-cell list (Pools::clist) */
    alignas(128) unsigned long long nlist;
    unsigned long long ncell;
};

/* One box of a level as the kernels see it. */
struct BoxDev
{
    int lo[3], hi[3];
    int glo[3], ghi[3];
    int dlo[3], dhi[3];
    int periodic[3];
    int gn[3], in[3];
    long long gnpts, inpts;
    double dx;
    double center[3];
    long long off_u;    /* this box's FAB in the state pool (doubles) */
    long long off_g;    /* storage cells of the boxes before it */
    long long off_i;    /* interior cells of the boxes before it */
    long long off_f[3]; /* faces of direction d of the boxes before it */
    long long nface[3]; /* faces of direction d of this box */
    double inv_dx[3];   /* inverse_dx_U: 1 / dx (rounded once on the
                           host) unless the caller gives its own */
    int seg;            /* faces per walker of the marching y/z kernels,
                           or kSegTile for the tile kernels */
};

/* Kinds of work item, one block table each. */
enum ItemClass
{
    C_ALL = 0, /* storage cells                          */
    C_EXT = 1, /* storage cells outside the interior     */
    C_INT = 2, /* interior cells                         */
    C_FX = 3,  /* x-face lanes: 64 per 63 cells of a row */
    C_FY = 4,  /* marching y-face walkers, or tile lanes */
    C_FZ = 5,  /* marching z-face walkers, or tile lanes */
    C_FALL = 6, /* the C_FX, C_FY and C_FZ tables, in that order */
    NCLASS = 7
};

__host__ __device__ inline long long face_tiles(const BoxDev &b, int d)
{
    const int other = (d == 2) ? 1 : 2;
    return (long long)((b.in[0] + kTileI - 1) / kTileI) *
           ((b.in[d] + 1 + kTileD - 2) / (kTileD - 1)) * b.in[other];
}

__host__ __device__ inline int face_segments(const BoxDev &b, int d)
{
    return (b.in[d] + 1 + b.seg - 1) / b.seg;
}

__host__ __device__ inline long long fx_cells(const BoxDev &b)
{
    return (long long)(b.in[0] + 2) * b.in[1] * b.in[2];
}

__host__ __device__ inline long long class_items(const BoxDev &b, int c)
{
    switch (c)
    {
    case C_ALL:
        return b.gnpts;
    case C_EXT:
        return b.gnpts - b.inpts;
    case C_INT:
        return b.inpts;
    case C_FX:
        return (fx_cells(b) - 1 + 62) / 63 * 64;
    case C_FY:
        if (b.seg == kSegTile)
            return 64 * face_tiles(b, 1);
        return (long long)b.in[0] * b.in[2] * face_segments(b, 1);
    case C_FZ:
        if (b.seg == kSegTile)
            return 64 * face_tiles(b, 2);
        return (long long)b.in[0] * b.in[1] * face_segments(b, 2);
    default: /* C_FALL is made of the three face tables */
        return 0;
    }
}

__host__ __device__ inline long long faces(const BoxDev &b, int d)
{
    return b.inpts / b.in[d] * (b.in[d] + 1);
}

} // namespace grhd

#endif /* GRHD_BOX_TABLE_HPP_ */
