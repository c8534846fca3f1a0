/*
 * plan_check.cpp -- builds level tables with the library's own host-side
 * plan code (grdzhadzha_amd/csrc/grhd_plan.hpp) and checks them against
 * brute-force restatements: each box's offsets, every class's block
 * table, and the internal-ghost exchange list (owner, wrapped position,
 * coverage of every non-outflow exterior cell).  No GPU is used; the test
 * suite compiles it with the address and undefined-behaviour sanitizers.
 *
 * Levels:
 *   tiled   4096 boxes: 16 x 16 x 16 boxes of 8^3 cells, 3 ghosts,
 *           periodic in x (LevelData over a DisjointBoxLayout that tiles
 *           the ProblemDomain, as BoxIterator-driven exchange() sees it)
 *   split   4096 boxes of ragged sides from random bisections of a 128^3
 *           domain, periodic in y, 2 ghosts
 *   holed   the split level without one box: the cells of the missing box
 *           have no owner, so the plan reports exchange as impossible
 *   single  one box equal to a non-periodic domain: outflow ghosts only
 *
 * Exit status 0 and a line "plan_check ok ..." when every check holds.
 */
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../grdzhadzha_amd/csrc/grhd_plan.hpp"

using namespace grhd;

namespace
{

int failures = 0;

#define CHECK(cond, ...)                                                  \
    do                                                                    \
    {                                                                     \
        if (!(cond))                                                      \
        {                                                                 \
            if (failures < 20)                                            \
            {                                                             \
                std::fprintf(stderr, "%s:%d: check failed: %s: ",        \
                             __FILE__, __LINE__, #cond);                  \
                std::fprintf(stderr, __VA_ARGS__);                        \
                std::fprintf(stderr, "\n");                               \
            }                                                             \
            ++failures;                                                   \
        }                                                                 \
    } while (0)

struct Cells
{
    int lo[3], hi[3];
};

grhd_level_box make_box(const Cells &c, const Cells &dom, const int per[3],
                        int ghosts, long long &offset)
{
    grhd_level_box lb{};
    for (int d = 0; d < 3; ++d)
    {
        lb.box.lo[d] = c.lo[d];
        lb.box.hi[d] = c.hi[d];
        lb.box.glo[d] = c.lo[d] - ghosts;
        lb.box.ghi[d] = c.hi[d] + ghosts;
        lb.box.dlo[d] = dom.lo[d];
        lb.box.dhi[d] = dom.hi[d];
        lb.box.periodic[d] = per[d];
        lb.box.center[d] = 0.5 * (dom.lo[d] + dom.hi[d]);
    }
    lb.box.dx = 0.25;
    lb.state_offset = offset;
    long long g = 1;
    for (int d = 0; d < 3; ++d)
        g *= c.hi[d] - c.lo[d] + 1 + 2 * ghosts;
    offset += 5 * g;
    return lb;
}

/* Split `c` into `count` boxes (a power of two) by bisections of the
 * longest side at a random point between 30% and 70% of it. */
void bisect(const Cells &c, int count, std::mt19937 &rng,
            std::vector<Cells> &out)
{
    if (count == 1)
    {
        out.push_back(c);
        return;
    }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
        if (c.hi[d] - c.lo[d] > c.hi[ax] - c.lo[ax])
            ax = d;
    const int n = c.hi[ax] - c.lo[ax] + 1;
    std::uniform_int_distribution<int> pick(n * 3 / 10, n * 7 / 10);
    const int left = std::max(1, std::min(n - 1, pick(rng)));
    Cells a = c, b = c;
    a.hi[ax] = c.lo[ax] + left - 1;
    b.lo[ax] = c.lo[ax] + left;
    bisect(a, count / 2, rng, out);
    bisect(b, count / 2, rng, out);
}

bool inside(const int w[3], const BoxDev &b)
{
    return w[0] >= b.lo[0] && w[0] <= b.hi[0] && w[1] >= b.lo[1] &&
           w[1] <= b.hi[1] && w[2] >= b.lo[2] && w[2] <= b.hi[2];
}

void cell_of(const BoxDev &b, long long idx, int c[3])
{
    c[0] = b.glo[0] + (int)(idx % b.gn[0]);
    c[1] = b.glo[1] + (int)(idx / b.gn[0] % b.gn[1]);
    c[2] = b.glo[2] + (int)(idx / ((long long)b.gn[0] * b.gn[1]));
}

struct Built
{
    std::vector<BoxDev> hb;
    long long tot_g = 0, tot_i = 0, tot_f[3] = {0, 0, 0};
    bool all_physical = true;
    bool exchange_ok = true;
    std::vector<int4> xl;
};

Built build(const std::vector<grhd_level_box> &boxes)
{
    Built b;
    b.hb.resize(boxes.size());
    for (size_t n = 0; n < boxes.size(); ++n)
        if (!add_boxdev(boxes[n], b.hb[n], b.tot_g, b.tot_i, b.tot_f))
            b.all_physical = false;
    const int seg = march_segment(b.tot_i);
    for (BoxDev &d : b.hb)
        d.seg = seg;
    if (!b.all_physical)
        b.xl = exchange_list(b.hb, b.exchange_ok);
    return b;
}

void check_boxes(const std::vector<grhd_level_box> &boxes, const Built &b)
{
    long long g = 0, i = 0, f[3] = {0, 0, 0};
    for (size_t n = 0; n < boxes.size(); ++n)
    {
        const BoxDev &d = b.hb[n];
        const grhd_box &x = boxes[n].box;
        long long gn = 1, in = 1;
        for (int k = 0; k < 3; ++k)
        {
            CHECK(d.gn[k] == x.ghi[k] - x.glo[k] + 1, "box %zu", n);
            CHECK(d.in[k] == x.hi[k] - x.lo[k] + 1, "box %zu", n);
            CHECK(d.inv_dx[k] == 1.0 / x.dx, "box %zu", n);
            gn *= d.gn[k];
            in *= d.in[k];
        }
        CHECK(d.gnpts == gn && d.inpts == in, "box %zu", n);
        CHECK(d.off_g == g && d.off_i == i, "box %zu", n);
        CHECK(d.off_u == boxes[n].state_offset, "box %zu", n);
        for (int k = 0; k < 3; ++k)
        {
            const long long nf = in / d.in[k] * (d.in[k] + 1);
            CHECK(d.off_f[k] == f[k] && d.nface[k] == nf, "box %zu dir %d",
                  n, k);
            f[k] += nf;
        }
        g += gn;
        i += in;
    }
    CHECK(b.tot_g == g && b.tot_i == i, "level totals");
    for (int k = 0; k < 3; ++k)
        CHECK(b.tot_f[k] == f[k], "face total %d", k);
}

/* Every class's block table: per box, blocks at first = 0, kBlock, ...
 * up to the class's item count; C_FALL the C_FX, C_FY, C_FZ tables. */
void check_tables(const Built &b)
{
    std::vector<int2> fall;
    for (int c = 0; c < NCLASS; ++c)
    {
        const std::vector<int2> t = class_table(b.hb, c);
        if (c == C_FALL)
        {
            CHECK(t.size() == fall.size(), "C_FALL size");
            for (size_t q = 0; q < t.size() && q < fall.size(); ++q)
                CHECK(t[q].x == fall[q].x && t[q].y == fall[q].y,
                      "C_FALL entry %zu", q);
            continue;
        }
        size_t q = 0;
        for (int n = 0; n < (int)b.hb.size(); ++n)
        {
            const long long m = class_items(b.hb[n], c);
            for (long long first = 0; first < m; first += kBlock, ++q)
                CHECK(q < t.size() && t[q].x == n && t[q].y == first,
                      "class %d box %d first %lld", c, n, first);
        }
        CHECK(q == t.size(), "class %d has %zu extra blocks", c,
              t.size() - q);
        if (c == C_FX || c == C_FY || c == C_FZ)
            fall.insert(fall.end(), t.begin(), t.end());
    }
}

/* Brute-force owner: the first box in table order whose interior holds the
 * wrapped cell, other than `self` when the cell did not wrap. */
int owner_linear(const std::vector<BoxDev> &hb, int self, const int c[3],
                 int w[3])
{
    const BoxDev &s = hb[self];
    for (int d = 0; d < 3; ++d)
    {
        w[d] = c[d];
        if (s.periodic[d])
        {
            const int nd = s.dhi[d] - s.dlo[d] + 1;
            int r = (c[d] - s.dlo[d]) % nd;
            if (r < 0)
                r += nd;
            w[d] = s.dlo[d] + r;
        }
    }
    const bool wrapped = w[0] != c[0] || w[1] != c[1] || w[2] != c[2];
    for (int n = 0; n < (int)hb.size(); ++n)
        if (inside(w, hb[n]) && (n != self || wrapped))
            return n;
    return -1;
}

/* The exchange list against the definition: for every box, every exterior
 * cell that is not an outflow ghost, in storage order, with its owner;
 * owners checked by linear search on every `stride`-th entry and on all
 * entries whose cell wrapped. */
void check_exchange(const Built &b, int stride, long long &orphans)
{
    size_t q = 0;
    orphans = 0;
    long long sampled = 0;
    for (int n = 0; n < (int)b.hb.size(); ++n)
    {
        const BoxDev &d = b.hb[n];
        for (long long idx = 0; idx < d.gnpts; ++idx)
        {
            int c[3];
            cell_of(d, idx, c);
            bool interior = true, physical = false;
            for (int k = 0; k < 3; ++k)
            {
                if (c[k] < d.lo[k] || c[k] > d.hi[k])
                    interior = false;
                if ((c[k] < d.dlo[k] || c[k] > d.dhi[k]) && !d.periodic[k])
                    physical = true;
            }
            if (interior || physical)
                continue;
            int w[3];
            const bool sample = (q % stride) == 0;
            const bool listed =
                q < b.xl.size() && b.xl[q].x == n && b.xl[q].y == idx;
            if (!listed)
            {
                /* an orphan: no box owns it */
                const int o = owner_linear(b.hb, n, c, w);
                CHECK(o < 0, "box %d cell %lld owned by %d but not listed",
                      n, idx, o);
                ++orphans;
                continue;
            }
            const int4 e = b.xl[q++];
            CHECK(e.z >= 0 && e.z < (int)b.hb.size(), "entry %zu", q - 1);
            if (e.z < 0 || e.z >= (int)b.hb.size())
                continue;
            const BoxDev &s = b.hb[e.z];
            CHECK(e.w >= 0 && e.w < s.gnpts, "entry %zu", q - 1);
            int sc[3];
            cell_of(s, e.w, sc);
            CHECK(inside(sc, s), "entry %zu source not interior", q - 1);
            const bool wrapped = sc[0] != c[0] || sc[1] != c[1] ||
                                 sc[2] != c[2];
            if (sample || wrapped)
            {
                const int o = owner_linear(b.hb, n, c, w);
                CHECK(o == e.z && w[0] == sc[0] && w[1] == sc[1] &&
                          w[2] == sc[2],
                      "entry %zu: box %d cell (%d,%d,%d) listed from box %d,"
                      " linear search says %d", q - 1, n, c[0], c[1], c[2],
                      e.z, o);
                ++sampled;
            }
        }
    }
    CHECK(q == b.xl.size(), "%zu exchange entries beyond the exterior cells",
          b.xl.size() - q);
    CHECK(b.exchange_ok == (orphans == 0), "exchange_ok %d orphans %lld",
          (int)b.exchange_ok, orphans);
    CHECK(sampled > 0, "no entry sampled");
}

std::vector<grhd_level_box> tiled_level()
{
    const Cells dom = {{0, 0, 0}, {127, 127, 127}};
    const int per[3] = {1, 0, 0};
    std::vector<grhd_level_box> boxes;
    long long off = 0;
    for (int z = 0; z < 16; ++z)
        for (int y = 0; y < 16; ++y)
            for (int x = 0; x < 16; ++x)
            {
                const Cells c = {{8 * x, 8 * y, 8 * z},
                                 {8 * x + 7, 8 * y + 7, 8 * z + 7}};
                boxes.push_back(make_box(c, dom, per, 3, off));
            }
    return boxes;
}

std::vector<grhd_level_box> split_level()
{
    const Cells dom = {{-64, -64, -64}, {63, 63, 63}};
    const int per[3] = {0, 1, 0};
    std::mt19937 rng(20240607u);
    std::vector<Cells> cells;
    bisect(dom, 4096, rng, cells);
    std::shuffle(cells.begin(), cells.end(), rng);
    std::vector<grhd_level_box> boxes;
    long long off = 0;
    for (const Cells &c : cells)
        boxes.push_back(make_box(c, dom, per, 2, off));
    return boxes;
}

void run(const char *name, const std::vector<grhd_level_box> &boxes,
         int stride, bool expect_ok)
{
    const Built b = build(boxes);
    check_boxes(boxes, b);
    check_tables(b);
    long long orphans = 0;
    if (!b.all_physical)
        check_exchange(b, stride, orphans);
    CHECK(b.exchange_ok == expect_ok, "%s: exchange_ok %d", name,
          (int)b.exchange_ok);
    std::printf("%s: %zu boxes, %lld storage cells, %zu exchange entries, "
                "%lld orphan ghosts, walk length %d\n",
                name, boxes.size(), b.tot_g, b.xl.size(), orphans,
                b.hb.empty() ? 0 : b.hb[0].seg);
}

} // namespace

int main()
{
    run("tiled", tiled_level(), 101, true);
    std::vector<grhd_level_box> split = split_level();
    run("split", split, 53, true);
    split.erase(split.begin() + 1234);
    run("holed", split, 53, false);
    {
        const Cells dom = {{0, 0, 0}, {19, 11, 7}};
        const int per[3] = {0, 0, 0};
        long long off = 0;
        std::vector<grhd_level_box> one = {make_box(dom, dom, per, 3, off)};
        const Built b = build(one);
        CHECK(b.all_physical && b.xl.empty(), "single box");
        check_boxes(one, b);
        check_tables(b);
    }
    {
        /* the checks catch planted mistakes: a wrong source box, a source
         * cell off by one, a dropped entry */
        const std::vector<grhd_level_box> t = tiled_level();
        const std::vector<grhd_level_box> small(t.begin(), t.begin() + 64);
        const Built good = build(small);
        for (int m = 0; m < 3; ++m)
        {
            Built bad = good;
            const size_t q = bad.xl.size() / 3;
            if (m == 0)
                bad.xl[q].z = (bad.xl[q].z + 1) % (int)bad.hb.size();
            else if (m == 1)
                bad.xl[q].w += 1;
            else
                bad.xl.erase(bad.xl.begin() + (long)q);
            const int before = failures;
            long long orphans = 0;
            check_exchange(bad, 1, orphans);
            if (failures == before)
            {
                std::fprintf(stderr, "planted mistake %d not caught\n", m);
                return 1;
            }
            failures = before;
        }
        std::printf("planted exchange mistakes caught\n");
    }
    if (failures)
    {
        std::fprintf(stderr, "plan_check: %d failed checks\n", failures);
        return 1;
    }
    std::printf("plan_check ok\n");
    return 0;
}
