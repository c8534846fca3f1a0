/*
 * grhd_plan.hpp -- host-side construction of a level's tables: each box's
 * kernel view (BoxDev), the block table of every work-item class and the
 * internal-ghost exchange list.  Host code only (no HIP runtime calls):
 * the library uploads what these functions return, and the CPU tests
 * build the same code with a host compiler under the address and
 * undefined-behaviour sanitizers.
 */
#ifndef GRHD_PLAN_HPP_
#define GRHD_PLAN_HPP_

#include <algorithm>
#include <vector>

#include "../../include/grhd_abi.h"
#include "grhd_box_table.hpp"

namespace grhd
{

/* Block table of one item class: blocks of kBlock items, box by box. */
inline std::vector<int2> class_table(const std::vector<BoxDev> &hb, int c)
{
    std::vector<int2> items;
    if (c == C_FALL)
    {
        for (int f : {C_FX, C_FY, C_FZ})
        {
            const std::vector<int2> t = class_table(hb, f);
            items.insert(items.end(), t.begin(), t.end());
        }
        return items;
    }
    for (int n = 0; n < (int)hb.size(); ++n)
    {
        const long long m = class_items(hb[n], c);
        for (long long first = 0; first < m; first += kBlock)
            items.push_back(make_int2(n, (int)first));
    }
    return items;
}

/* A coarse bucket grid over the boxes' interiors: each bucket lists, in
 * table order, the boxes overlapping it, so finding a cell's owner costs a
 * handful of checks instead of one per box of the level. */
struct OwnerIndex
{
    int edge = 1;
    int org[3] = {0, 0, 0};
    int nb[3] = {0, 0, 0};
    std::vector<std::vector<int>> bucket;

    void build(const std::vector<BoxDev> &hb)
    {
        int hi[3];
        for (int d = 0; d < 3; ++d)
        {
            org[d] = hb[0].lo[d];
            hi[d] = hb[0].hi[d];
        }
        for (const BoxDev &b : hb)
            for (int d = 0; d < 3; ++d)
            {
                org[d] = std::min(org[d], b.lo[d]);
                hi[d] = std::max(hi[d], b.hi[d]);
            }
        int ext = 1;
        for (int d = 0; d < 3; ++d)
            ext = std::max(ext, hi[d] - org[d] + 1);
        /* at most ~128 buckets per axis, at least 8 cells per bucket */
        edge = std::max(8, (ext + 127) / 128);
        for (int d = 0; d < 3; ++d)
            nb[d] = (hi[d] - org[d]) / edge + 1;
        bucket.assign((size_t)nb[0] * nb[1] * nb[2], std::vector<int>());
        for (int n = 0; n < (int)hb.size(); ++n)
        {
            const BoxDev &b = hb[n];
            int q0[3], q1[3];
            for (int d = 0; d < 3; ++d)
            {
                q0[d] = (b.lo[d] - org[d]) / edge;
                q1[d] = (b.hi[d] - org[d]) / edge;
            }
            for (int z = q0[2]; z <= q1[2]; ++z)
                for (int y = q0[1]; y <= q1[1]; ++y)
                    for (int x = q0[0]; x <= q1[0]; ++x)
                        bucket[((size_t)z * nb[1] + y) * nb[0] + x]
                            .push_back(n);
        }
    }

    const std::vector<int> *at(const int w[3]) const
    {
        int q[3];
        for (int d = 0; d < 3; ++d)
        {
            if (w[d] < org[d])
                return nullptr;
            q[d] = (w[d] - org[d]) / edge;
            if (q[d] >= nb[d])
                return nullptr;
        }
        return &bucket[((size_t)q[2] * nb[1] + q[1]) * nb[0] + q[0]];
    }
};

/* The box whose interior contains cell (i, j, k), wrapping periodic
 * directions of the domain; -1 if none.  The first such box in table
 * order, skipping `self` when the cell was not wrapped. */
inline int owner(const OwnerIndex &ix, const std::vector<BoxDev> &hb,
                 int self, const int c[3], int w[3])
{
    const BoxDev &s = hb[self];
    for (int d = 0; d < 3; ++d)
    {
        w[d] = c[d];
        if (s.periodic[d])
        {
            const int nd = s.dhi[d] - s.dlo[d] + 1;
            w[d] = s.dlo[d] + (((c[d] - s.dlo[d]) % nd) + nd) % nd;
        }
    }
    const std::vector<int> *cand = ix.at(w);
    if (!cand)
        return -1;
    for (int b : *cand)
    {
        const BoxDev &o = hb[b];
        if (w[0] >= o.lo[0] && w[0] <= o.hi[0] && w[1] >= o.lo[1] &&
            w[1] <= o.hi[1] && w[2] >= o.lo[2] && w[2] <= o.hi[2])
        {
            if (b == self && w[0] == c[0] && w[1] == c[1] && w[2] == c[2])
                continue;
            return b;
        }
    }
    return -1;
}

/* The kernels' view of box `lb`, placed after the storage cells, interior
 * cells and faces (tot_g, tot_i, tot_f) of the boxes before it, which it
 * then advances.  Returns whether every exterior cell of the box is an
 * outflow ghost (box = ProblemDomain, no periodic direction). */
inline bool add_boxdev(const grhd_level_box &lb, BoxDev &d,
                       long long &tot_g, long long &tot_i, long long tot_f[3])
{
    const grhd_box *b = &lb.box;
    bool physical = true;
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
        if (b->periodic[i] || b->lo[i] != b->dlo[i] || b->hi[i] != b->dhi[i])
            physical = false;
    }
    d.gnpts = (long long)d.gn[0] * d.gn[1] * d.gn[2];
    d.inpts = (long long)d.in[0] * d.in[1] * d.in[2];
    d.dx = b->dx;
    d.off_u = lb.state_offset;
    d.off_g = tot_g;
    d.off_i = tot_i;
    for (int k = 0; k < 3; ++k)
    {
        d.off_f[k] = tot_f[k];
        d.nface[k] = faces(d, k);
        tot_f[k] += d.nface[k];
    }
    const bool default_inv = b->inverse_dx[0] == 0.0 &&
                             b->inverse_dx[1] == 0.0 &&
                             b->inverse_dx[2] == 0.0;
    for (int k = 0; k < 3; ++k)
        d.inv_dx[k] = default_inv ? 1.0 / d.dx : b->inverse_dx[k];
    tot_g += d.gnpts;
    tot_i += d.inpts;
    return physical;
}

/* Internal-ghost exchange list of a level: every exterior cell that is not
 * an outflow ghost takes its value from the box that owns it, one entry
 * (box, storage index, source box, source storage index) per such cell,
 * box by box in storage order.  Cells that no box owns clear
 * `exchange_ok` and get no entry. */
inline std::vector<int4> exchange_list(const std::vector<BoxDev> &hb,
                                       bool &exchange_ok)
{
    std::vector<int4> xl;
    if (hb.empty())
        return xl;
    OwnerIndex ix;
    ix.build(hb);
    for (int n = 0; n < (int)hb.size(); ++n)
    {
        const BoxDev &b = hb[n];
        for (int k = b.glo[2]; k <= b.ghi[2]; ++k)
            for (int j = b.glo[1]; j <= b.ghi[1]; ++j)
                for (int i = b.glo[0]; i <= b.ghi[0]; ++i)
                {
                    const int c[3] = {i, j, k};
                    bool interior = true, physical = false;
                    for (int d = 0; d < 3; ++d)
                    {
                        if (c[d] < b.lo[d] || c[d] > b.hi[d])
                            interior = false;
                        if ((c[d] < b.dlo[d] || c[d] > b.dhi[d]) &&
                            !b.periodic[d])
                            physical = true;
                    }
                    if (interior || physical)
                        continue;
                    int w[3];
                    const int src = owner(ix, hb, n, c, w);
                    if (src < 0)
                    {
                        exchange_ok = false;
                        continue;
                    }
                    const BoxDev &s = hb[src];
                    const long long di =
                        ((long long)(k - b.glo[2]) * b.gn[1] +
                         (j - b.glo[1])) * b.gn[0] + (i - b.glo[0]);
                    const long long si =
                        ((long long)(w[2] - s.glo[2]) * s.gn[1] +
                         (w[1] - s.glo[1])) * s.gn[0] + (w[0] - s.glo[0]);
                    xl.push_back(make_int4(n, (int)di, src, (int)si));
                }
    }
    return xl;
}

} // namespace grhd

#endif /* GRHD_PLAN_HPP_ */
