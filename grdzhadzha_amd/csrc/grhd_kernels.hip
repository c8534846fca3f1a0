/*
 * grhd_kernels.hip -- host side of the C-ABI (include/grhd_abi.h): level
 * plans (device box table, block tables, internal-ghost exchange list),
 * scratch pools, the RHS / stage launch sequences and the entry points.
 * The kernels are in grhd_stage_kernels.hpp; a single-box call is a level
 * of one box.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/grhd_abi.h"
#include "grhd_physics.hpp"
#include "grhd_stage_kernels.hpp"
#include "grhd_plan.hpp"
#include "grhd_adjoint.hpp"

using namespace grhd;

/* Scratch of grhd_advance_rk4_box_noalloc: one array per GPU, held from the
 * library's load until the process ends.  Host code never uses its name
 * as an address (on the host it is only a stand-in); it uses what
 * hipGetSymbolAddress returns for the current device. */
static __device__ __align__(256) unsigned char
    scratch[GRHD_NOALLOC_SCRATCH_BYTES];

namespace
{

inline unsigned grid_of(long long n)
{
    return (unsigned)((n + kBlock - 1) / kBlock);
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
    o.mixed = p->mixed_precision != 0;
    o.guard_margin = o.mixed ? p->mixed_fp64_margin : 0.0;
    o.guard_shift = o.mixed ? p->mixed_fp64_shift : 0.0;
    o.fp32_sources = o.mixed;
    o.excise = p->excise_horizon != 0;
    o.excise_inner = p->excision_inner_buffer;
    o.excise_outer = p->excision_outer_buffer;
    return o;
}

/* background_factory(time): CircularBinaryPN::compute_state
 * (CircularBinaryPN.hpp:209-250) at params.time = bg.time + time; the
 * binary's positions are computed here, on the host, with the C library's
 * cos and sin. */
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

int check_box(const grhd_box *b)
{
    if (!(b->dx > 0.0))
        return GRHD_ERR_INVALID_ARGUMENT;
    /* inverse_dx: all zero (1/dx), or positive and finite */
    if (b->inverse_dx[0] != 0.0 || b->inverse_dx[1] != 0.0 ||
        b->inverse_dx[2] != 0.0)
        for (int d = 0; d < 3; ++d)
            if (!(b->inverse_dx[d] > 0.0) || !std::isfinite(b->inverse_dx[d]))
                return GRHD_ERR_INVALID_ARGUMENT;
    for (int d = 0; d < 3; ++d)
    {
        if (b->lo[d] > b->hi[d])
            return GRHD_ERR_INVALID_ARGUMENT;
        /* the sweeps read two cells beyond the interior */
        if (b->glo[d] > b->lo[d] - 2 || b->ghi[d] < b->hi[d] + 2)
            return GRHD_ERR_INVALID_ARGUMENT;
        if (b->dlo[d] > b->dhi[d])
            return GRHD_ERR_INVALID_ARGUMENT;
    }
    return GRHD_OK;
}

/* Mixed precision with FP64 faces near floors, the cap or a large shift
 * (grhd_params::mixed_fp64_margin, mixed_fp64_shift). */
bool guarded(const Options &o)
{
    return o.mixed && (o.guard_margin > 0.0 || o.guard_shift > 0.0);
}

int check_common(const grhd_params *o, const grhd_background *bg, int ncomp,
                 int cb)
{
    if (!o || !bg)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (cb < 0 || cb + GRHD_NUM_COMPS > ncomp)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (o->max_iterations < 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (!(o->mixed_fp64_margin == 0.0 ||
          (o->mixed_fp64_margin >= 1.0 &&
           o->mixed_fp64_margin <= 1.0e300)))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (!(o->mixed_fp64_shift >= 0.0 && o->mixed_fp64_shift <= 1.0e300))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (o->excise_horizon && bg->kind != GRHD_BG_KERR_SCHILD)
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

inline grhd_level_box one_box(const grhd_box *b)
{
    grhd_level_box lb;
    std::memset(&lb, 0, sizeof(lb));
    if (b)
        lb.box = *b;
    lb.state_offset = 0;
    return lb;
}

/* Walk lengths of the marching y/z face kernels, and the tile kernels,
 * that the tuner times. */
constexpr int kNumSegs = 5;
constexpr int kSegs[kNumSegs] = {1, 2, 4, 8, kSegTile};

/* The device copies that depend on the walk length: the box table (each
 * BoxDev carries it) and the y, z and fused face block tables. */
struct SegTables
{
    bool built = false;
    BoxDev *d_boxes = nullptr;
    int2 *d_items[3] = {nullptr, nullptr, nullptr}; /* C_FY, C_FZ, C_FALL */
    int nblocks[3] = {0, 0, 0};
};

/* Device copies of a level's box table, block tables and exchange list.
 * d_boxes and the C_FY, C_FZ and C_FALL entries of d_items point into the
 * SegTables of the walk length in use (seg_cur); the plan owns the other
 * tables itself. */
struct Plan
{
    /* what the plan was built from */
    std::vector<grhd_level_box> key;
    int ncomp = -1;
    /* host view */
    std::vector<BoxDev> hb;
    long long tot_g = 0, tot_i = 0, tot_f[3] = {0, 0, 0};
    int nblocks[NCLASS] = {0};
    bool all_physical = true; /* every exterior cell is an outflow ghost */
    bool exchange_ok = true;  /* every other exterior cell has an owner */
    long long nexchange = 0;
    /* device */
    BoxDev *d_boxes = nullptr;
    int2 *d_items[NCLASS] = {nullptr};
    int4 *d_exchange = nullptr;
    /* walk-length variants of the face tables, and the face launch chosen
     * for FP64 ([0]) and FP32 ([1]) faces: by the size rules, or by timing
     * every variant on the level when the context tunes */
    SegTables segt[kNumSegs];
    int seg_rule = 0; /* index into kSegs */
    int seg_cur = -1;
    bool tuned[2] = {false, false};      /* measured choice in seg/fused_pick */
    bool tune_tried[2] = {false, false}; /* a measurement was started */
    int seg_pick[2] = {0, 0};
    bool fused_pick[2] = {false, false};
    float tuned_ms[2] = {0.0f, 0.0f};
};

/* The face-launch choices measured on one level (box list). */
struct TuneRecord
{
    std::vector<grhd_level_box> key;
    int ncomp = -1;
    bool tuned[2] = {false, false};
    int seg_pick[2] = {0, 0};
    bool fused_pick[2] = {false, false};
    float tuned_ms[2] = {0.0f, 0.0f};
};

/* Levels whose tuned choices a context remembers (oldest dropped first). */
constexpr size_t kMaxTuneRecords = 64;

bool is_seg_class(int c)
{
    return c == C_FY || c == C_FZ || c == C_FALL;
}

int seg_slot(int c)
{
    return c == C_FY ? 0 : (c == C_FZ ? 1 : 2);
}

void free_plan(Plan &p)
{
    for (int c = 0; c < NCLASS; ++c)
        if (!is_seg_class(c))
            (void)hipFree(p.d_items[c]);
    for (SegTables &t : p.segt)
    {
        (void)hipFree(t.d_boxes);
        for (int2 *q : t.d_items)
            (void)hipFree(q);
    }
    (void)hipFree(p.d_exchange);
    p = Plan();
}

int upload_table(const std::vector<int2> &items, int2 **d)
{
    if (items.empty())
        return GRHD_OK;
    if (hipMalloc((void **)d, sizeof(int2) * items.size()) != hipSuccess ||
        hipMemcpy(*d, items.data(), sizeof(int2) * items.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
        return GRHD_ERR_DEVICE;
    return GRHD_OK;
}

/* Build (once) the box table and face tables of walk length kSegs[k]. */
int build_seg_tables(Plan &p, int k)
{
    SegTables &t = p.segt[k];
    if (t.built)
        return GRHD_OK;
    /* what an earlier failed attempt left */
    (void)hipFree(t.d_boxes);
    t.d_boxes = nullptr;
    for (int2 *&q : t.d_items)
    {
        (void)hipFree(q);
        q = nullptr;
    }
    std::vector<BoxDev> hb = p.hb;
    for (BoxDev &d : hb)
        d.seg = kSegs[k];
    for (int c : {C_FY, C_FZ, C_FALL})
    {
        const std::vector<int2> items = class_table(hb, c);
        const int rc = upload_table(items, &t.d_items[seg_slot(c)]);
        if (rc)
            return rc;
        t.nblocks[seg_slot(c)] = (int)items.size();
    }
    if (hipMalloc((void **)&t.d_boxes, sizeof(BoxDev) * hb.size()) !=
            hipSuccess ||
        hipMemcpy(t.d_boxes, hb.data(), sizeof(BoxDev) * hb.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
        return GRHD_ERR_DEVICE;
    t.built = true;
    return GRHD_OK;
}

/* Point the plan's box table and face tables at walk length kSegs[k]
 * (built).  Kernels take these pointers by value at launch, so switching
 * between launches needs no synchronisation. */
void select_seg(Plan &p, int k)
{
    if (p.seg_cur == k)
        return;
    const SegTables &t = p.segt[k];
    p.d_boxes = t.d_boxes;
    for (int c : {C_FY, C_FZ, C_FALL})
    {
        p.d_items[c] = t.d_items[seg_slot(c)];
        p.nblocks[c] = t.nblocks[seg_slot(c)];
    }
    p.seg_cur = k;
}

bool same_key(const Plan &p, const grhd_level_box *boxes, int nbox,
              int ncomp)
{
    if (p.ncomp != ncomp || (int)p.key.size() != nbox)
        return false;
    return std::memcmp(p.key.data(), boxes,
                       sizeof(grhd_level_box) * (size_t)nbox) == 0;
}

int build_plan(Plan &p, const grhd_level_box *boxes, int nbox, int ncomp)
{
    free_plan(p);
    p.key.assign(boxes, boxes + nbox);
    p.ncomp = ncomp;
    p.hb.resize(nbox);
    for (int n = 0; n < nbox; ++n)
        if (!add_boxdev(boxes[n], p.hb[n], p.tot_g, p.tot_i, p.tot_f))
            p.all_physical = false;
    if (2 * p.tot_g >= (1ll << kKeyBits))
        return GRHD_ERR_INVALID_ARGUMENT;
    const int seg = march_segment(p.tot_i);
    for (int k = 0; k < kNumSegs; ++k)
        if (kSegs[k] == seg)
            p.seg_rule = k;
    for (BoxDev &d : p.hb)
        d.seg = seg;

    /* block tables: the classes that do not depend on the walk length
     * here, the face tables of the rule's walk length in segt */
    for (int c = 0; c < NCLASS; ++c)
    {
        if (is_seg_class(c))
            continue;
        const std::vector<int2> items = class_table(p.hb, c);
        const int rc = upload_table(items, &p.d_items[c]);
        if (rc)
            return rc;
        p.nblocks[c] = (int)items.size();
    }
    {
        const int rc = build_seg_tables(p, p.seg_rule);
        if (rc)
            return rc;
        select_seg(p, p.seg_rule);
        p.seg_pick[0] = p.seg_pick[1] = p.seg_rule;
    }

    /* internal-ghost exchange list */
    std::vector<int4> xl;
    if (!p.all_physical)
        xl = exchange_list(p.hb, p.exchange_ok);
    if (!xl.empty())
    {
        if (hipMalloc((void **)&p.d_exchange, sizeof(int4) * xl.size()) !=
                hipSuccess ||
            hipMemcpy(p.d_exchange, xl.data(), sizeof(int4) * xl.size(),
                      hipMemcpyHostToDevice) != hipSuccess)
            return GRHD_ERR_DEVICE;
    }
    p.nexchange = (long long)xl.size();
    return GRHD_OK;
}

} // namespace

struct grhd_ctx
{
    int device = 0;
    Plan plan;
    double *W = nullptr;
    unsigned char *wbad = nullptr;
    long long cap_g = 0;
    double *F[3] = {nullptr, nullptr, nullptr};
    long long cap_f[3] = {0, 0, 0};
    int2 *list = nullptr; /* FP64 face list of the guarded mixed path */
    long long cap_list = 0;
    int2 *clist = nullptr; /* its guarded-cell list */
    long long cap_clist = 0;
    double *src = nullptr; /* FP64 sources of its guarded cells */
    long long cap_src = 0;
    double *UA = nullptr, *UB = nullptr;
    long long cap_u = 0;
    double *acc = nullptr;
    long long cap_i = 0;
    double *diag = nullptr; /* per-block partials of the level diagnostics */
    long long cap_diag = 0;
    DevStatus *dstat = nullptr;
    DevStatus *hstat = nullptr;
    /* side streams of the y and z face kernels, which run beside the x
     * kernel (default priority), and the events that fork them off and
     * join them back into the caller's stream */
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[2] = {nullptr, nullptr};
    /* face-launch tuning (grhd_ctx_set_autotune), its two timing events,
     * and the choices measured on the levels this context has seen (the
     * plan is rebuilt whenever a call names another box list) */
    bool autotune = false;
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    /* a face launch pinned per precision (grhd_ctx_set_face_launch): the
     * kSegs index, -1 for none, and fused or not */
    int pin_seg[2] = {-1, -1};
    bool pin_fused[2] = {false, false};
    std::vector<TuneRecord> tuned_levels;
    /* instrumentation */
    bool timing = false;
    std::vector<hipEvent_t> events;
    std::vector<int> classes; /* class of the launch ending at events[i] */
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
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc((void **)p, sizeof(T) * (size_t)std::max(need, 1ll)) !=
        hipSuccess)
        return GRHD_ERR_DEVICE;
    *cap = need;
    return GRHD_OK;
}

int ensure_scratch(grhd_ctx *c, bool stages, const grhd_params *params)
{
    const Plan &p = c->plan;
    if (guarded(make_options(params)))
    {
        int rc =
            grow(&c->list, &c->cap_list, p.tot_f[0] + p.tot_f[1] + p.tot_f[2]);
        if (rc || (rc = grow(&c->clist, &c->cap_clist, p.tot_i)) ||
            (rc = grow(&c->src, &c->cap_src, 5 * p.tot_i)))
            return rc;
    }
    if (c->cap_g < p.tot_g)
    {
        (void)hipFree(c->W);
        (void)hipFree(c->wbad);
        c->W = nullptr;
        c->wbad = nullptr;
        c->cap_g = 0;
        if (hipMalloc((void **)&c->W, sizeof(double) * 6 * (size_t)p.tot_g) !=
                hipSuccess ||
            hipMalloc((void **)&c->wbad, (size_t)p.tot_g) != hipSuccess)
            return GRHD_ERR_DEVICE;
        c->cap_g = p.tot_g;
    }
    for (int d = 0; d < 3; ++d)
    {
        const int rc = grow(&c->F[d], &c->cap_f[d], 5 * p.tot_f[d]);
        if (rc)
            return rc;
    }
    if (!stages)
        return GRHD_OK;
    if (c->cap_u < 5 * p.tot_g)
    {
        long long ca = 0, cb = 0;
        (void)hipFree(c->UA);
        (void)hipFree(c->UB);
        c->UA = c->UB = nullptr;
        c->cap_u = 0;
        int rc = grow(&c->UA, &ca, 5 * p.tot_g);
        if (rc)
            return rc;
        rc = grow(&c->UB, &cb, 5 * p.tot_g);
        if (rc)
            return rc;
        c->cap_u = 5 * p.tot_g;
    }
    return grow(&c->acc, &c->cap_i, 5 * p.tot_i);
}

int prepare(grhd_ctx *c, const grhd_level_box *boxes, int nbox, int ncomp)
{
    if (!c || !boxes || nbox <= 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    for (int n = 0; n < nbox; ++n)
    {
        const int rc = check_box(&boxes[n].box);
        if (rc)
            return rc;
    }
    if (same_key(c->plan, boxes, nbox, ncomp))
        return GRHD_OK;
    const int rc = build_plan(c->plan, boxes, nbox, ncomp);
    if (rc)
    {
        /* no half-built plan survives: the next call with the same boxes
         * builds it again (and fails the same way) */
        free_plan(c->plan);
        return rc;
    }
    /* a level tuned before keeps what was measured on it */
    Plan &pl = c->plan;
    for (const TuneRecord &t : c->tuned_levels)
        if (t.ncomp == ncomp && t.key.size() == (size_t)nbox &&
            std::memcmp(t.key.data(), boxes,
                        sizeof(grhd_level_box) * (size_t)nbox) == 0)
            for (int q = 0; q < 2; ++q)
            {
                pl.tuned[q] = t.tuned[q];
                pl.tune_tried[q] = t.tuned[q];
                pl.seg_pick[q] = t.seg_pick[q];
                pl.fused_pick[q] = t.fused_pick[q];
                pl.tuned_ms[q] = t.tuned_ms[q];
            }
    return GRHD_OK;
}

/* Remember the plan's tuned choices for its box list. */
void remember_tuning(grhd_ctx *c)
{
    const Plan &pl = c->plan;
    TuneRecord *rec = nullptr;
    for (TuneRecord &t : c->tuned_levels)
        if (t.ncomp == pl.ncomp && t.key.size() == pl.key.size() &&
            std::memcmp(t.key.data(), pl.key.data(),
                        sizeof(grhd_level_box) * pl.key.size()) == 0)
            rec = &t;
    if (!rec)
    {
        if (c->tuned_levels.size() >= kMaxTuneRecords)
            c->tuned_levels.erase(c->tuned_levels.begin());
        c->tuned_levels.emplace_back();
        rec = &c->tuned_levels.back();
        rec->key = pl.key;
        rec->ncomp = pl.ncomp;
    }
    for (int q = 0; q < 2; ++q)
    {
        rec->tuned[q] = pl.tuned[q];
        rec->seg_pick[q] = pl.seg_pick[q];
        rec->fused_pick[q] = pl.fused_pick[q];
        rec->tuned_ms[q] = pl.tuned_ms[q];
    }
}

/* ---------------- instrumentation ---------------- */
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

/* ---------------- launch helpers ---------------- */
Pools make_pools(grhd_ctx *c, double *state, int cb)
{
    Pools p;
    std::memset(&p, 0, sizeof(p));
    p.state = state;
    p.cb = cb;
    p.UA = c->UA;
    p.UB = c->UB;
    p.acc = c->acc;
    p.W = c->W;
    p.wbad = c->wbad;
    for (int d = 0; d < 3; ++d)
        p.F[d] = c->F[d];
    p.list = c->list;
    p.clist = c->clist;
    p.src = c->src;
    p.st = c->dstat;
    return p;
}

void launch_status_init(grhd_ctx *c, hipStream_t s)
{
    c->nev = 0;
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(64), 0, s, c->dstat);
    tick(c, GRHD_KERNEL_OTHER, s);
}

#define GRHD_LAUNCH(c, cls, kern, s, ...)                                     \
    do                                                                        \
    {                                                                         \
        if ((c)->plan.nblocks[cls] > 0)                                       \
            hipLaunchKernelGGL(kern, dim3((c)->plan.nblocks[cls]),            \
                               dim3(kBlock), 0, s, (c)->plan.d_boxes,         \
                               (c)->plan.d_items[cls], __VA_ARGS__);          \
    } while (0)

void launch_exchange(grhd_ctx *c, const Pools &p, int which, int ncopy,
                     hipStream_t s)
{
    if (c->plan.nexchange == 0)
        return;
    hipLaunchKernelGGL(k_exchange, dim3(grid_of(c->plan.nexchange)),
                       dim3(kBlock), 0, s, c->plan.d_boxes,
                       c->plan.d_exchange, c->plan.nexchange, p, which, ncopy);
    tick(c, GRHD_KERNEL_OTHER, s);
}

void launch_prims(grhd_ctx *c, const Options &o, const Background &bg,
                  const Pools &p, int which, bool store_prims, bool exterior,
                  unsigned long long eval, hipStream_t s)
{
    if (exterior)
        GRHD_LAUNCH(c, C_EXT, k_prims<true>, s, o, bg, p, which,
                    (int)store_prims, eval);
    else
        GRHD_LAUNCH(c, C_ALL, k_prims<false>, s, o, bg, p, which,
                    (int)store_prims, eval);
    tick(c, GRHD_KERNEL_PRIMS, s);
}

/* Interior cells of a call from which the three face kernels run as
 * separate launches on concurrent streams; below it one fused launch
 * (k_faces_all, or k_faces_all_t with the tiles, which the size rule
 * takes below the same count) does all three directions.  Measured per
 * RK4 step with walks, fused against streams: 32^3 0.148 / 0.262 ms,
 * 64^3 0.394 / 0.473, 128^3 2.260 / 2.293, 160^3 4.232 / 4.145, 256^3
 * 16.86 / 16.54, the box-set levels (6M cells each) 22.4 / 21.2 ms for
 * all three levels. */
constexpr long long kFusedFacesCells = kTileCells;

/* The three face directions are independent (W in, each its own flux
 * pool out), so they can run as one fused launch or as three kernels.
 * Unfused, y and z run on the context's side streams beside x, forked
 * after everything already queued on s (the ghost fill and recovery of
 * this stage, the previous stage's cell kernel) and joined back into s
 * before the cell kernel. */
template <class T, int SEL = FACES_ALL>
void faces_launches(grhd_ctx *c, const Options &o, const Background &bg,
                    const Pools &p, hipStream_t s, bool fused)
{
    const bool tiles = kSegs[c->plan.seg_cur] == kSegTile;
    if (fused)
    {
        if (c->plan.nblocks[C_FALL] > 0)
        {
            if (tiles)
                hipLaunchKernelGGL((k_faces_all_t<T, SEL>),
                                   dim3(c->plan.nblocks[C_FALL]),
                                   dim3(kBlock), 0, s, c->plan.d_boxes,
                                   c->plan.d_items[C_FALL],
                                   c->plan.nblocks[C_FX],
                                   c->plan.nblocks[C_FY], o, bg, p);
            else
                hipLaunchKernelGGL((k_faces_all<T, SEL>),
                                   dim3(c->plan.nblocks[C_FALL]),
                                   dim3(kBlock), 0, s, c->plan.d_boxes,
                                   c->plan.d_items[C_FALL],
                                   c->plan.nblocks[C_FX],
                                   c->plan.nblocks[C_FY], o, bg, p);
        }
        return;
    }
    /* a failed fork runs the three kernels in order on `s`; the error
     * stays pending for the call's closing hipGetLastError check */
    const bool forked = hipEventRecord(c->fork, s) == hipSuccess &&
                        hipStreamWaitEvent(c->side[0], c->fork, 0) ==
                            hipSuccess &&
                        hipStreamWaitEvent(c->side[1], c->fork, 0) ==
                            hipSuccess;
    hipStream_t sy = forked ? c->side[0] : s;
    hipStream_t sz = forked ? c->side[1] : s;
    if (!forked)
        GRHD_LAUNCH(c, C_FX, (k_faces_x<T, SEL>), s, o, bg, p);
    if (tiles)
    {
        GRHD_LAUNCH(c, C_FY, (k_faces_t<1, T, SEL>), sy, o, bg, p);
        GRHD_LAUNCH(c, C_FZ, (k_faces_t<2, T, SEL>), sz, o, bg, p);
    }
    else
    {
        GRHD_LAUNCH(c, C_FY, (k_faces<1, T, SEL>), sy, o, bg, p);
        GRHD_LAUNCH(c, C_FZ, (k_faces<2, T, SEL>), sz, o, bg, p);
    }
    if (!forked)
        return;
    GRHD_LAUNCH(c, C_FX, (k_faces_x<T, SEL>), s, o, bg, p);
    /* join: `s` waits for both side streams; if an event operation fails,
     * the host waits for the side stream instead, so k_cells never runs
     * before the y and z fluxes are written */
    for (int q = 0; q < 2; ++q)
        if (hipEventRecord(c->join[q], c->side[q]) != hipSuccess ||
            hipStreamWaitEvent(s, c->join[q], 0) != hipSuccess)
            (void)hipStreamSynchronize(c->side[q]);
}

/* Launches per candidate in a tuning measurement (after one untimed). */
constexpr int kTuneReps = 3;

/* Time every face launch of the level -- each walk length of the marching
 * kernels and the tile kernels, fused and unfused -- on the call's own
 * primitives, and keep the
 * fastest for faces of type T.  The faces only read W and write their
 * flux pools, and the fluxes do not depend on the walk length or the
 * launch form, so the extra launches change no result: the call's own
 * face launch follows and writes the same fluxes.  The rule's choice is
 * timed first and kept on ties.  The host waits for each measurement, so
 * a tuning call synchronises once per candidate; it runs once per plan
 * and precision.  If an event operation fails, the rule's choice stays. */
template <class T>
void tune_faces(grhd_ctx *c, const Options &o, const Background &bg,
                const Pools &p, hipStream_t s)
{
    Plan &pl = c->plan;
    const int prec = sizeof(T) < sizeof(double) ? 1 : 0;
    pl.tune_tried[prec] = true;
    int best_k = pl.seg_rule;
    bool best_f = pl.tot_i < kFusedFacesCells;
    float best = 0.0f;
    bool have = false;
    int order[kNumSegs];
    order[0] = pl.seg_rule;
    for (int k = 0, n = 1; k < kNumSegs; ++k)
        if (k != pl.seg_rule)
            order[n++] = k;
    for (int k : order)
    {
        if (build_seg_tables(pl, k) != GRHD_OK)
            continue;
        select_seg(pl, k);
        const bool rule_fused = pl.tot_i < kFusedFacesCells;
        for (bool fused : {rule_fused, !rule_fused})
        {
            faces_launches<T>(c, o, bg, p, s, fused);
            float ms = 0.0f;
            if (hipEventRecord(c->tune_ev[0], s) != hipSuccess)
                return;
            for (int r = 0; r < kTuneReps; ++r)
                faces_launches<T>(c, o, bg, p, s, fused);
            if (hipEventRecord(c->tune_ev[1], s) != hipSuccess ||
                hipEventSynchronize(c->tune_ev[1]) != hipSuccess ||
                hipEventElapsedTime(&ms, c->tune_ev[0], c->tune_ev[1]) !=
                    hipSuccess)
                return;
            if (!have || ms < best)
            {
                best = ms;
                best_k = k;
                best_f = fused;
                have = true;
            }
        }
    }
    if (!have)
        return;
    pl.seg_pick[prec] = best_k;
    pl.fused_pick[prec] = best_f;
    pl.tuned_ms[prec] = best / kTuneReps;
    pl.tuned[prec] = true;
    remember_tuning(c);
}

/* The faces of one stage.  By default the size rules choose: tiles or
 * the walk length from march_segment, and one fused launch below
 * kFusedFacesCells interior cells, where the fused kernel's register
 * count costs the x faces less than the two launch boundaries it
 * removes.  A launch pinned on the context comes first, then what a
 * tuning context measured (tune_faces).  The faces timing class covers
 * all three directions. */
template <class T, int SEL = FACES_ALL>
void launch_faces_t(grhd_ctx *c, const Options &o, const Background &bg,
                    const Pools &p, hipStream_t s)
{
    Plan &pl = c->plan;
    const int prec = sizeof(T) < sizeof(double) ? 1 : 0;
    /* the two passes of the guarded path keep the size rules */
    const bool pinned = c->pin_seg[prec] >= 0 && SEL == FACES_ALL;
    const bool tunes = c->autotune && SEL == FACES_ALL && !pinned;
    if (tunes && !pl.tune_tried[prec])
        tune_faces<T>(c, o, bg, p, s);
    bool fused = pl.tot_i < kFusedFacesCells;
    int k = pl.seg_rule;
    if (pinned && build_seg_tables(pl, c->pin_seg[prec]) == GRHD_OK)
    {
        k = c->pin_seg[prec];
        fused = c->pin_fused[prec];
    }
    else if (tunes && pl.tuned[prec] &&
             build_seg_tables(pl, pl.seg_pick[prec]) == GRHD_OK)
    {
        k = pl.seg_pick[prec];
        fused = pl.fused_pick[prec];
    }
    select_seg(pl, k);
    faces_launches<T, SEL>(c, o, bg, p, s, fused);
    tick(c, GRHD_KERNEL_FACES, s);
}

/* Blocks of the grid-stride FP64 face-list kernel: the list's length is
 * only known on the device, so the grid is sized for the level and idle
 * blocks leave at once. */
constexpr long long kListBlocksMax = 4096;

/* Guarded mixed precision: the FP32 pass writes its faces and lists the
 * others, which k_faces_list computes in FP64; both write the one double
 * flux pool, disjoint faces each. */
void launch_faces(grhd_ctx *c, const Options &o, const Background &bg,
                  const Pools &p, hipStream_t s)
{
    if (guarded(o))
    {
        hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, s, p.st);
        launch_faces_t<float, FACES_FP32>(c, o, bg, p, s);
        const Plan &pl = c->plan;
        const long long nf = pl.tot_f[0] + pl.tot_f[1] + pl.tot_f[2];
        const long long blocks =
            std::min(kListBlocksMax, std::max(1ll, (nf + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_faces_list, dim3((unsigned)blocks), dim3(kBlock),
                           0, s, pl.d_boxes, o, bg, p);
        tick(c, GRHD_KERNEL_FACES, s);
    }
    else if (o.mixed)
        launch_faces_t<float>(c, o, bg, p, s);
    else
        launch_faces_t<double>(c, o, bg, p, s);
}

/* One stage mode's cell kernel: the FP64 path's, or the mixed paths'. */
template <int M, class FT, bool S32>
void launch_cells_mode(grhd_ctx *c, const Options &o, const Background &bs,
                       const Background &bn, const Background &bp,
                       const Background &bm, const Pools &p,
                       const CellArgs &a, hipStream_t s)
{
    if constexpr (S32)
        GRHD_LAUNCH(c, C_INT, (k_cells_mixed<M, FT>), s, o, bs, bn, bp, bm, p,
                    a);
    else
        GRHD_LAUNCH(c, C_INT, (k_cells<M, FT, false>), s, o, bs, bn, bp, bm,
                    p, a);
}

template <class FT, bool S32 = (sizeof(FT) < sizeof(double))>
void launch_cells_t(grhd_ctx *c, int mode, const Options &o,
                    const Background &bs, const Background &bn,
                    const Background &bp, const Background &bm,
                    const Pools &p, const CellArgs &a, hipStream_t s)
{
    switch (mode)
    {
    case MODE_RHS:
        launch_cells_mode<MODE_RHS, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    case MODE_RK4_1:
        launch_cells_mode<MODE_RK4_1, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    case MODE_RK4_2:
        launch_cells_mode<MODE_RK4_2, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    case MODE_RK4_3:
        launch_cells_mode<MODE_RK4_3, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    case MODE_RK4_4:
        launch_cells_mode<MODE_RK4_4, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    case MODE_SSP_1:
        launch_cells_mode<MODE_SSP_1, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    default:
        launch_cells_mode<MODE_SSP_2, FT, S32>(c, o, bs, bn, bp, bm, p, a, s);
        break;
    }
    tick(c, GRHD_KERNEL_CELLS, s);
}

void launch_cells(grhd_ctx *c, int mode, const Options &o,
                  const Background &bs, const Background &bn,
                  const Background &bp, const Background &bm, const Pools &p,
                  const CellArgs &a, hipStream_t s)
{
    if (guarded(o))
        launch_cells_t<double, true>(c, mode, o, bs, bn, bp, bm, p, a, s);
    else if (o.mixed)
        launch_cells_t<float>(c, mode, o, bs, bn, bp, bm, p, a, s);
    else
        launch_cells_t<double>(c, mode, o, bs, bn, bp, bm, p, a, s);
}

/* Volume-time source backgrounds at t +- max(step, 1e-12). */
void vol_backgrounds(const grhd_background *bg, const grhd_params *p,
                     double t, Background &bp, Background &bm, double &h)
{
    h = std::max(p->metric_time_derivative_step, 1.0e-12);
    bp = make_background(bg, t + h);
    bm = make_background(bg, t - h);
}

/* One right-hand-side evaluation of pool `us`: exchange, ghost fill +
 * con2prim (all storage cells, or only the exterior ones when the
 * interior W is already current), faces, then the cell kernel. */
void rhs_sequence(grhd_ctx *c, int mode, const grhd_params *params,
                  const grhd_background *bgp, double t, double t_next,
                  const Pools &p, const CellArgs &a, bool exterior_only,
                  bool store_ghost_prims, bool exchange, int exchange_ncopy,
                  unsigned long long eval, hipStream_t s)
{
    const Options o = make_options(params);
    const Background bs = make_background(bgp, t);
    const Background bn = make_background(bgp, t_next);
    Background bp = bs, bm = bs;
    double h = 1.0;
    if (o.use_volume_sources)
        vol_backgrounds(bgp, params, t, bp, bm, h);
    if (exchange)
        launch_exchange(c, p, a.us, exchange_ncopy, s);
    launch_prims(c, o, bs, p, a.us, store_ghost_prims, exterior_only, eval,
                 s);
    launch_faces(c, o, bs, p, s);
    CellArgs aa = a;
    aa.vol_h = h;
    if (guarded(o) && o.use_static_sources)
    {
        const long long blocks = std::min(
            kListBlocksMax, std::max(1ll, (c->plan.tot_i + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_guard_sources, dim3((unsigned)blocks),
                           dim3(kBlock), 0, s, c->plan.d_boxes, o, bs, p,
                           a.us);
        tick(c, GRHD_KERNEL_CELLS, s);
    }
    launch_cells(c, mode, o, bs, bn, bp, bm, p, aa, s);
}

/* Translate the plan's global failure key into (code, cell, residual). */
int finish(grhd_ctx *c, hipStream_t s, grhd_error *err,
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
    const Plan &pl = c->plan;
    const unsigned long long key = c->hstat->fail_key;
    if (key != kNoFail)
    {
        const long long v = (long long)(key & ((1ull << kKeyBits) - 1));
        int nb = 0;
        while (nb + 1 < (int)pl.hb.size() && 2 * pl.hb[nb + 1].off_g <= v)
            ++nb;
        const BoxDev &b = pl.hb[nb];
        long long r = v - 2 * b.off_g;
        const int phase = r >= b.gnpts ? 1 : 0;
        const long long lin = phase ? r - b.gnpts : r;
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
        const int code = phase == 0 ? GRHD_ERR_RECOVERY_GHOST_FILL
                                    : GRHD_ERR_RECOVERY_RHS;
        if (err)
        {
            err->code = code;
            err->iv[0] = iv[0];
            err->iv[1] = iv[1];
            err->iv[2] = iv[2];
            double res = 0.0;
            (void)hipMemcpy(&res, c->W + 6 * b.off_g + lin, sizeof(double),
                            hipMemcpyDeviceToHost);
            err->residual = res;
        }
        return code;
    }
    if (want_cfl && c->hstat->cfl_key != kNoFail)
    {
        const long long v = (long long)c->hstat->cfl_key;
        int nb = 0;
        while (nb + 1 < (int)pl.hb.size() && pl.hb[nb + 1].off_i <= v)
            ++nb;
        const BoxDev &b = pl.hb[nb];
        const long long lin = v - b.off_i;
        if (err)
        {
            err->code = GRHD_ERR_RECOVERY_CFL;
            err->iv[0] = b.lo[0] + (int)(lin % b.in[0]);
            err->iv[1] = b.lo[1] + (int)((lin / b.in[0]) % b.in[1]);
            err->iv[2] = b.lo[2] + (int)(lin / ((long long)b.in[0] * b.in[1]));
            double res = 0.0;
            (void)hipMemcpy(&res, c->W + v, sizeof(double),
                            hipMemcpyDeviceToHost);
            err->residual = res;
        }
        return GRHD_ERR_RECOVERY_CFL;
    }
    return GRHD_OK;
}

/* ---------------- level operations ---------------- */
/* A level call with no boxes: a process of a distributed level may own
 * none of its boxes, and the reference's loops over the level's boxes
 * then do nothing.  Such a call checks its parameters, launches nothing
 * and succeeds; its outputs are an empty level's (zero counts, max 1/dt
 * and sums, a finite state).  `diag` (nullable) is zeroed. */
int empty_level(const grhd_ctx *c, const grhd_params *params,
                const grhd_background *bg, int ncomp, int cb,
                grhd_diagnostics *diag = nullptr)
{
    if (!c)
        return GRHD_ERR_INVALID_ARGUMENT;
    const int rc = params ? check_common(params, bg, ncomp, cb)
                   : (cb < 0 || cb + GRHD_NUM_COMPS > ncomp)
                       ? GRHD_ERR_INVALID_ARGUMENT
                       : GRHD_OK;
    if (!rc && diag)
        std::memset(diag, 0, sizeof(*diag));
    return rc;
}

int level_rhs(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
              const grhd_params *params, const grhd_background *bg,
              double time, double *state, int ncomp, int cb, double *rhs,
              unsigned char *mask, bool exchange, grhd_error *err, void *st,
              bool accumulate = false)
{
    if (nbox == 0)
        return empty_level(c, params, bg, ncomp, cb);
    if (!state || !rhs)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, bg, ncomp, cb);
    if (rc || (rc = prepare(c, boxes, nbox, ncomp)) ||
        (rc = ensure_scratch(c, false, params)))
        return rc;
    if (exchange && !c->plan.exchange_ok)
        return GRHD_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)st;
    Pools p = make_pools(c, state, cb);
    p.rhs = rhs;
    p.rhs_ncomp = ncomp;
    p.mask = mask;
    launch_status_init(c, s);
    if (!accumulate)
    {
        GRHD_LAUNCH(c, C_INT, k_zero_rhs, s, p);
        tick(c, GRHD_KERNEL_OTHER, s);
    }
    CellArgs a;
    std::memset(&a, 0, sizeof(a));
    a.us = POOL_STATE;
    a.write_mask = mask != nullptr;
    a.accumulate = accumulate ? 1 : 0;
    rhs_sequence(c, MODE_RHS, params, bg, time, time, p, a, false, true,
                 exchange, ncomp, 0, s);
    return finish(c, s, err, nullptr, false);
}

int level_recover(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
                  const grhd_params *params, const grhd_background *bg,
                  double time, double *state, int ncomp, int cb,
                  unsigned char *mask, grhd_diagnostics *diag, void *st)
{
    if (nbox == 0)
        return empty_level(c, params, bg, ncomp, cb, diag);
    if (!state)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, bg, ncomp, cb);
    if (rc || (rc = prepare(c, boxes, nbox, ncomp)))
        return rc;
    hipStream_t s = (hipStream_t)st;
    Pools p = make_pools(c, state, cb);
    p.mask = mask;
    launch_status_init(c, s);
    GRHD_LAUNCH(c, C_INT, k_recover, s, make_options(params),
                make_background(bg, time), p, (int)(mask != nullptr));
    tick(c, GRHD_KERNEL_OTHER, s);
    return finish(c, s, nullptr, diag, false);
}

int level_max_inv_dt(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
                     const grhd_params *params, const grhd_background *bg,
                     double time, const double *state, int ncomp, int cb,
                     double *out, grhd_error *err, void *st)
{
    if (nbox == 0)
    {
        const int rc = out ? empty_level(c, params, bg, ncomp, cb)
                           : GRHD_ERR_INVALID_ARGUMENT;
        if (!rc)
            *out = 0.0;
        return rc;
    }
    if (!state || !out)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, bg, ncomp, cb);
    if (rc || (rc = prepare(c, boxes, nbox, ncomp)) ||
        (rc = ensure_scratch(c, false, params)))
        return rc;
    hipStream_t s = (hipStream_t)st;
    Pools p = make_pools(c, const_cast<double *>(state), cb);
    launch_status_init(c, s);
    GRHD_LAUNCH(c, C_INT, k_max_inv_dt, s, make_options(params),
                make_background(bg, time), p);
    tick(c, GRHD_KERNEL_OTHER, s);
    rc = finish(c, s, err, nullptr, true);
    if (rc)
        return rc;
    double m;
    const unsigned long long bits = c->hstat->max_bits;
    std::memcpy(&m, &bits, sizeof(m));
    *out = m;
    return GRHD_OK;
}

/* Conserved sums and the count of interior cells holding a non-finite
 * conserved value, over the level (LevelDataDriver.hpp:22-92). */
int level_diagnostics(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
                      const double *state, int ncomp, int cb, void *st)
{
    if (!state || cb < 0 || cb + GRHD_NUM_COMPS > ncomp)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = prepare(c, boxes, nbox, ncomp);
    if (rc)
        return rc;
    const long long nb = c->plan.nblocks[C_INT];
    if ((rc = grow(&c->diag, &c->cap_diag, kDiagWords * nb)))
        return rc;
    hipStream_t s = (hipStream_t)st;
    Pools p = make_pools(c, const_cast<double *>(state), cb);
    launch_status_init(c, s);
    GRHD_LAUNCH(c, C_INT, k_diag_cells, s, p, c->diag);
    hipLaunchKernelGGL(k_diag_finish, dim3(1), dim3(kBlock), 0, s, c->diag,
                       nb, c->dstat);
    tick(c, GRHD_KERNEL_OTHER, s);
    return finish(c, s, nullptr, nullptr, false);
}

/* The checkpoint's grid: the ProblemDomain grown by `ghosts`.  Every box's
 * storage must lie inside it ("state ghost box exceeds checkpoint storage",
 * Main_GRHDLevelDataBinaryPNKITPDisk.cpp:1576-1579). */
int ckpt_grid(const Plan &pl, const int *dlo, const int *dhi, int ghosts,
              CkptGrid &g, long long &cells)
{
    if (!dlo || !dhi || ghosts < 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    cells = 1;
    for (int d = 0; d < 3; ++d)
    {
        if (dhi[d] < dlo[d])
            return GRHD_ERR_INVALID_ARGUMENT;
        g.lo[d] = dlo[d] - ghosts;
        g.n[d] = dhi[d] - dlo[d] + 1 + 2 * ghosts;
        g.dlo[d] = dlo[d];
        g.dhi[d] = dhi[d];
        cells *= g.n[d];
    }
    for (const BoxDev &b : pl.hb)
        for (int d = 0; d < 3; ++d)
            if (b.glo[d] < g.lo[d] || b.ghi[d] > g.lo[d] + g.n[d] - 1)
                return GRHD_ERR_INVALID_ARGUMENT;
    return GRHD_OK;
}

int level_checkpoint(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
                     double *state, int ncomp, const int *dlo, const int *dhi,
                     int ghosts, double *global, bool pack, void *st)
{
    if (!global || ncomp < 1 || (!state && nbox != 0))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (nbox == 0)
    {
        /* no boxes: packing gives the zero-filled checkpoint grid,
         * unpacking writes nothing */
        if (!c)
            return GRHD_ERR_INVALID_ARGUMENT;
        Plan none;
        CkptGrid g;
        long long cells = 0;
        const int rc = ckpt_grid(none, dlo, dhi, ghosts, g, cells);
        if (rc || !pack)
            return rc;
        hipStream_t s = (hipStream_t)st;
        const long long n = cells * ncomp;
        hipLaunchKernelGGL(k_fill, dim3(grid_of(n)), dim3(kBlock), 0, s,
                           global, n, 0.0);
        if (hipGetLastError() != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return GRHD_ERR_DEVICE;
        return GRHD_OK;
    }
    int rc = prepare(c, boxes, nbox, ncomp);
    if (rc)
        return rc;
    CkptGrid g;
    long long cells = 0;
    if ((rc = ckpt_grid(c->plan, dlo, dhi, ghosts, g, cells)))
        return rc;
    hipStream_t s = (hipStream_t)st;
    if (pack)
    {
        const long long n = cells * ncomp;
        hipLaunchKernelGGL(k_fill, dim3(grid_of(n)), dim3(kBlock), 0, s,
                           global, n, 0.0);
        GRHD_LAUNCH(c, C_ALL, k_ckpt_pack, s, state, ncomp, g, global);
    }
    else
    {
        GRHD_LAUNCH(c, C_ALL, k_ckpt_unpack, s, global, ncomp, g, state);
    }
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipGetLastError() != hipSuccess)
        return GRHD_ERR_DEVICE;
    return GRHD_OK;
}

/* add_scaled_conserved_rhs over the level, then (optionally) the
 * post-update recovery at `recovery_time`. */
int level_update(grhd_ctx *ctx, const grhd_level_box *boxes, int nbox,
                 const grhd_params *params, const grhd_background *background,
                 double *state, int ncomp, int comp_base, const double *rhs,
                 double dt, int recover, double recovery_time,
                 unsigned char *mask, grhd_diagnostics *diagnostics,
                 void *stream)
{
    if (nbox == 0)
        return empty_level(ctx, params, background, ncomp, comp_base,
                           diagnostics);
    if (!state || !rhs)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, background, ncomp, comp_base);
    if (rc || (rc = prepare(ctx, boxes, nbox, ncomp)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    Pools p = make_pools(ctx, state, comp_base);
    p.rhs = const_cast<double *>(rhs);
    p.rhs_ncomp = ncomp;
    p.mask = mask;
    launch_status_init(ctx, s);
    GRHD_LAUNCH(ctx, C_INT, k_add_scaled, s, p, dt);
    tick(ctx, GRHD_KERNEL_OTHER, s);
    if (recover)
    {
        GRHD_LAUNCH(ctx, C_INT, k_recover, s, make_options(params),
                    make_background(background, recovery_time), p,
                    (int)(mask != nullptr));
        tick(ctx, GRHD_KERNEL_OTHER, s);
    }
    grhd_diagnostics d;
    rc = finish(ctx, s, nullptr, &d, false);
    if (diagnostics)
    {
        if (recover)
            *diagnostics = d;
        else
            std::memset(diagnostics, 0, sizeof(*diagnostics));
    }
    return rc;
}

int level_ssprk2(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
                 const grhd_params *params, const grhd_background *bg,
                 double time, double dt, double *state, int ncomp, int cb,
                 unsigned char *mask, bool exchange, grhd_diagnostics *diag,
                 grhd_error *err, void *st)
{
    if (nbox == 0)
        return empty_level(c, params, bg, ncomp, cb, diag);
    if (!state)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, bg, ncomp, cb);
    if (rc || (rc = prepare(c, boxes, nbox, ncomp)) ||
        (rc = ensure_scratch(c, true, params)))
        return rc;
    /* stage states are internal: every ghost must be an outflow ghost or
     * be filled from another box of the level */
    if (!c->plan.all_physical && !(exchange && c->plan.exchange_ok))
        return GRHD_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)st;
    Pools p = make_pools(c, state, cb);
    p.mask = mask;
    launch_status_init(c, s);
    /* k1 = L(U, t); stage = U + dt k1; recover(stage, t + dt) */
    CellArgs a;
    std::memset(&a, 0, sizeof(a));
    a.us = POOL_STATE;
    a.unext = POOL_A;
    a.dt = dt;
    a.next_eval = 1;
    rhs_sequence(c, MODE_SSP_1, params, bg, time, time + dt, p, a, false, true,
                 exchange, ncomp, 0, s);
    /* k2 = L(stage, t + dt); U = (U + stage + dt k2) / 2; recover(U) */
    std::memset(&a, 0, sizeof(a));
    a.us = POOL_A;
    a.u0 = POOL_STATE;
    a.unext = POOL_STATE;
    a.store_prims = 1;
    a.write_mask = mask != nullptr;
    a.dt = dt;
    a.next_eval = 2;
    a.count = 1;
    /* with exchanged ghosts the interior-only W shortcut does not cover
     * internal ghosts, so every storage cell is recovered */
    const bool ext = c->plan.all_physical;
    rhs_sequence(c, MODE_SSP_2, params, bg, time + dt, time + dt, p, a, ext,
                 false, exchange, 5, 1, s);
    return finish(c, s, err, diag, false);
}

int level_rk4(grhd_ctx *c, const grhd_level_box *boxes, int nbox,
              const grhd_params *params, const grhd_background *bg,
              double time, double dt, int nsteps, double *state, int ncomp,
              int cb, unsigned char *mask, bool exchange,
              grhd_diagnostics *diag, grhd_error *err, void *st)
{
    if (nsteps < 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (nbox == 0)
        return empty_level(c, params, bg, ncomp, cb, diag);
    if (!state)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, bg, ncomp, cb);
    if (rc || (rc = prepare(c, boxes, nbox, ncomp)) ||
        (rc = ensure_scratch(c, true, params)))
        return rc;
    if (!c->plan.all_physical && !(exchange && c->plan.exchange_ok))
        return GRHD_ERR_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)st;
    Pools p = make_pools(c, state, cb);
    p.mask = mask;
    launch_status_init(c, s);
    const double c2 = dt / 2.0, c3 = dt / 3.0, c6 = dt / 6.0;
    const bool ext_ok = c->plan.all_physical;
    unsigned long long eval = 0;
    for (int step = 0; step < nsteps; ++step)
    {
        const double t = time + step * dt;
        const double ts[4] = {t, t + c2, t + c2, t + dt};
        const double tn[4] = {t + c2, t + c2, t + dt, t + dt};
        const double cst[4] = {c2, c2, dt, 0.0};
        const double cac[4] = {c6, c3, c3, c6};
        const int us[4] = {POOL_STATE, POOL_A, POOL_B, POOL_A};
        const int un[4] = {POOL_A, POOL_B, POOL_A, POOL_STATE};
        for (int sg = 0; sg < 4; ++sg)
        {
            CellArgs a;
            std::memset(&a, 0, sizeof(a));
            a.us = us[sg];
            a.u0 = POOL_STATE;
            a.unext = un[sg];
            a.store_prims = sg == 3;
            a.write_mask = (sg == 3 && step == nsteps - 1 && mask);
            a.count = 1;
            a.c_stage = cst[sg];
            a.c_acc = cac[sg];
            a.dt = dt;
            a.next_eval = eval + 1;
            /* the first evaluation of a call recovers every cell; later
             * ones reuse the interior W the previous stage produced */
            rhs_sequence(c, MODE_RK4_1 + sg, params, bg, ts[sg], tn[sg], p, a,
                         eval != 0 && ext_ok, sg == 0, exchange,
                         sg == 0 ? ncomp : 5, eval, s);
            ++eval;
        }
    }
    return finish(c, s, err, diag, false);
}

/* ---------------- RK4 of one box without allocation ---------------- */

constexpr int kMaxScratchDevices = 256;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxScratchDevices] = {};

/* The device address of `scratch` on the current device, looked up on
 * the first call on that device and kept. */
int scratch_on_current_device(unsigned char **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 ||
        dev >= kMaxScratchDevices)
        return GRHD_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev])
    {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess ||
            !ptr)
            return GRHD_ERR_DEVICE;
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return GRHD_OK;
}

/* Where each buffer of one box's RK4 sits in the scratch array (byte
 * offsets, 256-aligned). */
struct NoAllocLayout
{
    size_t status, box, items[NCLASS], W, wbad, F[3], list, clist, src, UA,
        UB, acc, total;
    int nblocks[NCLASS];
};

/* with_list: room for the guarded mixed path's FP64 face list */
NoAllocLayout noalloc_layout(const BoxDev &b, bool with_list)
{
    NoAllocLayout L;
    size_t at = 0;
    auto take = [&at](size_t bytes) {
        const size_t o = at;
        at += (bytes + 255) / 256 * 256;
        return o;
    };
    L.status = take(sizeof(DevStatus));
    L.box = take(sizeof(BoxDev));
    for (int c = 0; c < NCLASS; ++c)
    {
        if (c == C_FALL)
            L.nblocks[c] =
                L.nblocks[C_FX] + L.nblocks[C_FY] + L.nblocks[C_FZ];
        else
            L.nblocks[c] = (int)((class_items(b, c) + kBlock - 1) / kBlock);
        L.items[c] = take(sizeof(int2) * (size_t)L.nblocks[c]);
    }
    const size_t G = (size_t)b.gnpts, I = (size_t)b.inpts;
    L.W = take(sizeof(double) * 6 * G);
    L.wbad = take(G);
    for (int d = 0; d < 3; ++d)
        L.F[d] = take(sizeof(double) * 5 * (size_t)b.nface[d]);
    L.list = take(with_list ? sizeof(int2) * (size_t)(b.nface[0] + b.nface[1] +
                                                      b.nface[2])
                            : 0);
    L.clist = take(with_list ? sizeof(int2) * (size_t)b.inpts : 0);
    L.src = take(with_list ? sizeof(double) * 5 * (size_t)b.inpts : 0);
    L.UA = take(sizeof(double) * 5 * G);
    L.UB = take(sizeof(double) * 5 * G);
    L.acc = take(sizeof(double) * 5 * I);
    L.total = at;
    return L;
}

struct NoAllocTables
{
    BoxDev *box;
    int2 *items[NCLASS];
    int nblocks[NCLASS];
};

/* The box table and block tables of one box, written on the device (block
 * q of a class covers items 256 q .. 256 q + 255 of box 0; the fused face
 * table is the x, y and z tables one after another). */
__global__ void __launch_bounds__(kBlock)
    k_noalloc_tables(BoxDev b, NoAllocTables t)
{
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q == 0)
        *t.box = b;
    for (int c = 0; c < NCLASS; ++c)
    {
        if (q >= t.nblocks[c])
            continue;
        int first = q;
        if (c == C_FALL)
        {
            const int nx = t.nblocks[C_FX], ny = t.nblocks[C_FY];
            first = q < nx ? q : (q < nx + ny ? q - nx : q - nx - ny);
        }
        t.items[c][q] = make_int2(0, first * kBlock);
    }
}

/* grhd_advance_rk4_box with every scratch buffer in `scratch`: a context
 * on the host stack whose plan and pools point into the array, driven by
 * the same launch sequence.  Nothing here allocates device memory: the
 * context's own buffers are never grown (their capacities are set to what
 * the box needs), the tables are written by a kernel, the status block is
 * cleared by a kernel and read back into host stack memory. */
int rk4_box_noalloc(const grhd_box *box, const grhd_params *params,
                    const grhd_background *bg, double time, double dt,
                    int nsteps, double *state, int ncomp, int cb,
                    unsigned char *mask, grhd_diagnostics *diag,
                    grhd_error *err, void *st)
{
    if (!box || !state || nsteps < 0)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_box(box);
    if (rc || (rc = check_common(params, bg, ncomp, cb)))
        return rc;
    const grhd_level_box lb = one_box(box);
    BoxDev b;
    std::memset(&b, 0, sizeof(b));
    long long tot_g = 0, tot_i = 0, tot_f[3] = {0, 0, 0};
    const bool physical = add_boxdev(lb, b, tot_g, tot_i, tot_f);
    /* as grhd_advance_rk4_box: every ghost an outflow ghost */
    if (!physical)
        return GRHD_ERR_INVALID_ARGUMENT;
    b.seg = march_segment(tot_i);
    const bool with_list = guarded(make_options(params));
    const NoAllocLayout L = noalloc_layout(b, with_list);
    if (L.total > (size_t)GRHD_NOALLOC_SCRATCH_BYTES ||
        tot_i >= kFusedFacesCells)
        return GRHD_ERR_INVALID_ARGUMENT;
    unsigned char *base = nullptr;
    if ((rc = scratch_on_current_device(&base)))
        return rc;
    hipStream_t s = (hipStream_t)st;

    NoAllocTables t;
    t.box = reinterpret_cast<BoxDev *>(base + L.box);
    int most = 0;
    for (int c = 0; c < NCLASS; ++c)
    {
        t.items[c] = reinterpret_cast<int2 *>(base + L.items[c]);
        t.nblocks[c] = L.nblocks[c];
        most = std::max(most, L.nblocks[c]);
    }
    hipLaunchKernelGGL(k_noalloc_tables, dim3(grid_of(std::max(most, 1))),
                       dim3(kBlock), 0, s, b, t);

    grhd_ctx c;
    Plan &pl = c.plan;
    pl.key.assign(&lb, &lb + 1);
    pl.ncomp = ncomp;
    pl.hb.assign(&b, &b + 1);
    pl.tot_g = tot_g;
    pl.tot_i = tot_i;
    for (int d = 0; d < 3; ++d)
        pl.tot_f[d] = tot_f[d];
    pl.all_physical = true;
    pl.d_boxes = t.box;
    for (int q = 0; q < NCLASS; ++q)
    {
        pl.d_items[q] = t.items[q];
        pl.nblocks[q] = t.nblocks[q];
    }
    for (int k = 0; k < kNumSegs; ++k)
        if (kSegs[k] == b.seg)
            pl.seg_rule = k;
    pl.seg_cur = pl.seg_rule; /* the face tables above are for b.seg */
    c.W = reinterpret_cast<double *>(base + L.W);
    c.wbad = base + L.wbad;
    c.cap_g = tot_g;
    for (int d = 0; d < 3; ++d)
    {
        c.F[d] = reinterpret_cast<double *>(base + L.F[d]);
        c.cap_f[d] = 5 * tot_f[d];
    }
    if (with_list)
    {
        c.list = reinterpret_cast<int2 *>(base + L.list);
        c.cap_list = tot_f[0] + tot_f[1] + tot_f[2];
        c.clist = reinterpret_cast<int2 *>(base + L.clist);
        c.cap_clist = tot_i;
        c.src = reinterpret_cast<double *>(base + L.src);
        c.cap_src = 5 * tot_i;
    }
    c.UA = reinterpret_cast<double *>(base + L.UA);
    c.UB = reinterpret_cast<double *>(base + L.UB);
    c.cap_u = 5 * tot_g;
    c.acc = reinterpret_cast<double *>(base + L.acc);
    c.cap_i = 5 * tot_i;
    c.dstat = reinterpret_cast<DevStatus *>(base + L.status);
    DevStatus host_status;
    c.hstat = &host_status;
    /* c owns no device memory: it goes out of scope without a free */
    return level_rk4(&c, &lb, 1, params, bg, time, dt, nsteps, state, ncomp,
                     cb, mask, false, diag, err, st);
}

/* ---------------- discrete adjoint of RK4 ---------------- */

/* Device buffers of one adjoint call, freed on every return path. */
struct DevBufs
{
    std::vector<void *> ptrs;
    ~DevBufs()
    {
        for (void *q : ptrs)
            (void)hipFree(q);
    }
    double *get(long long n)
    {
        void *q = nullptr;
        if (n <= 0 || hipMalloc(&q, sizeof(double) * (size_t)n) != hipSuccess)
            return nullptr;
        ptrs.push_back(q);
        return (double *)q;
    }
};

struct AdjRun
{
    grhd_ctx *c;
    BoxDev b;
    Options o;
    AdjParams pa;
    Background bgd;
    double dt;
    hipStream_t s;
    double *work; /* 11-component FAB (comp_base 0) the forward kernels use */
    Pools p;
    MetricCache mc; /* the cells' metric spin tangents (reverse) */
};

void fill0(double *x, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(grid_of(n)), dim3(kBlock), 0, s, x, n,
                       0.0);
}

/* k = L(X): X (5 components over the storage box) into the work FAB, ghost
 * fill + recovery of every storage cell, faces, rhs into k (5 x interior).
 * Leaves W of X in the context's scratch. */
void adj_eval_rhs(AdjRun &r, const double *X, double *k)
{
    const BoxDev &b = r.b;
    (void)hipMemcpyAsync(r.work, X, sizeof(double) * 5 * b.gnpts,
                         hipMemcpyDeviceToDevice, r.s);
    launch_prims(r.c, r.o, r.bgd, r.p, POOL_STATE, false, false, 0, r.s);
    launch_faces(r.c, r.o, r.bgd, r.p, r.s);
    CellArgs a;
    std::memset(&a, 0, sizeof(a));
    a.us = POOL_STATE;
    Pools pp = r.p;
    pp.rhs = k;
    pp.rhs_ncomp = 5;
    launch_cells(r.c, MODE_RHS, r.o, r.bgd, r.bgd, r.bgd, r.bgd, pp, a, r.s);
}

/* One RK4 step from X0 (storage, 5 components): stage states X[1..4]
 * (X[4] = the new state), rhs k[0..3], accumulator after stage 3 in acc3
 * (acc is scratch). */
void adj_fwd_step(AdjRun &r, const double *X0, double *const X[5],
                  double *const k[4], double *acc, double *acc3)
{
    const BoxDev &b = r.b;
    const unsigned gi = grid_of(b.inpts);
    for (int sg = 0; sg < 4; ++sg)
    {
        adj_eval_rhs(r, sg == 0 ? X0 : X[sg], k[sg]);
        /* ghosts of the next stage state are refilled by its own eval */
        if (X[sg + 1] != X0)
            (void)hipMemcpyAsync(X[sg + 1], sg == 0 ? X0 : X[sg],
                                 sizeof(double) * 5 * b.gnpts,
                                 hipMemcpyDeviceToDevice, r.s);
        hipLaunchKernelGGL(k_fwd_stage, dim3(gi), dim3(kBlock), 0, r.s, b, r.o,
                           r.pa, sg + 1, r.dt, X0, k[sg], acc, X[sg + 1]);
        if (sg == 2)
            (void)hipMemcpyAsync(acc3, acc, sizeof(double) * 5 * b.inpts,
                                 hipMemcpyDeviceToDevice, r.s);
    }
}

/* lam_X = L'(X)^T lam_k (5 x interior), parameter adjoints into lam_par.
 * With horizon excision the rows of excised cells in lam_k are zeroed in
 * place first (their rhs is the constant zero). */
void adj_rhs_transpose(AdjRun &r, const double *X, double *lam_k,
                       double *lam_X, double *lam_W, double *lam_par)
{
    const BoxDev &b = r.b;
    if (r.o.excise)
        hipLaunchKernelGGL(k_adj_excise, dim3(grid_of(b.inpts)), dim3(kBlock),
                           0, r.s, b, r.o, r.pa, lam_k);
    (void)hipMemcpyAsync(r.work, X, sizeof(double) * 5 * b.gnpts,
                         hipMemcpyDeviceToDevice, r.s);
    launch_prims(r.c, r.o, r.bgd, r.p, POOL_STATE, false, false, 0, r.s);
    fill0(lam_W, 6 * b.gnpts, r.s);
    fill0(lam_X, 5 * b.inpts, r.s);
    const long long nf[3] = {(long long)(b.in[0] + 1) * b.in[1] * b.in[2],
                             (long long)b.in[0] * (b.in[1] + 1) * b.in[2],
                             (long long)b.in[0] * b.in[1] * (b.in[2] + 1)};
    /* the face Jacobians in one reverse sweep each, the spin's through the
     * cached metric tangents */
    hipLaunchKernelGGL((k_adj_faces<0>), dim3(grid_of(nf[0])), dim3(kBlock),
                       0, r.s, b, r.o, r.pa, r.mc, r.c->W, lam_k, lam_W,
                       lam_par);
    hipLaunchKernelGGL((k_adj_faces<1>), dim3(grid_of(nf[1])), dim3(kBlock),
                       0, r.s, b, r.o, r.pa, r.mc, r.c->W, lam_k, lam_W,
                       lam_par);
    hipLaunchKernelGGL((k_adj_faces<2>), dim3(grid_of(nf[2])), dim3(kBlock),
                       0, r.s, b, r.o, r.pa, r.mc, r.c->W, lam_k, lam_W,
                       lam_par);
    /* sources: the double metric's terms in a reverse sweep, and the
     * spin's forward-mode pass, whose dual metric needs far more registers,
     * in a launch of its own; recoveries: one reverse sweep with the spin
     * through the cached metric tangents */
    hipLaunchKernelGGL(k_adj_sources<false>, dim3(grid_of(b.inpts)),
                       dim3(kBlock), 0, r.s, b, r.o, r.pa, r.work,
                       r.c->W, lam_k, lam_W, lam_X, lam_par);
    hipLaunchKernelGGL(k_adj_sources<true>, dim3(grid_of(b.inpts)),
                       dim3(kBlock), 0, r.s, b, r.o, r.pa, r.work,
                       r.c->W, lam_k, lam_W, lam_X, lam_par);
    hipLaunchKernelGGL(k_adj_prims, dim3(grid_of(b.gnpts)), dim3(kBlock), 0,
                       r.s, b, r.o, r.pa, r.mc, r.work, r.c->W, lam_W, lam_X,
                       lam_par);
}

int rk4_adjoint_box(grhd_ctx *c, const grhd_box *box,
                    const grhd_params *params, const grhd_background *bg,
                    double time, double dt, int nsteps, const double *state,
                    int ncomp, int cb, double r0, int first_step,
                    int ckpt_every, double *J, double *dJ_dU0,
                    double *dJ_dparams, grhd_error *err, void *st)
{
    (void)time;
    if (!box || !state || !J || nsteps < 1 ||
        first_step < 1 || first_step > nsteps || ckpt_every < 1 || dt <= 0.0)
        return GRHD_ERR_INVALID_ARGUMENT;
    int rc = check_common(params, bg, ncomp, cb);
    if (rc)
        return rc;
    if (bg->kind != GRHD_BG_KERR_SCHILD ||
        params->use_metric_volume_time_sources || params->mixed_precision)
        return GRHD_ERR_INVALID_ARGUMENT;
    for (int d = 0; d < 3; ++d)
        if (box->periodic[d] || box->lo[d] != box->dlo[d] ||
            box->hi[d] != box->dhi[d])
            return GRHD_ERR_INVALID_ARGUMENT;
    grhd_level_box lb;
    std::memset(&lb, 0, sizeof(lb));
    lb.box = *box;
    if ((rc = prepare(c, &lb, 1, GRHD_NUM_COMPS)) ||
        (rc = ensure_scratch(c, false, params)))
        return rc;
    AdjRun r;
    r.c = c;
    r.b = c->plan.hb[0];
    r.o = make_options(params);
    r.pa.M = bg->mass;
    r.pa.a = bg->spin;
    r.pa.gamma = params->adiabatic_index;
    r.bgd = make_background(bg, 0.0);
    r.dt = dt;
    r.s = (hipStream_t)st;
    const BoxDev &b = r.b;
    const long long G = b.gnpts, I = b.inpts;
    /* the objective alone needs neither checkpoints nor adjoint buffers */
    const bool grad = dJ_dU0 && dJ_dparams;
    if (!grad)
        ckpt_every = nsteps;
    const int nseg = (nsteps + ckpt_every - 1) / ckpt_every;

    DevBufs mem;
    r.work = mem.get(GRHD_NUM_COMPS * G);
    /* per step of a segment: X[1..4] (5 x G) and k[0..3], acc3 (5 x I) */
    const long long per_step = 4 * 5 * G + 5 * 5 * I;
    const long long nadj = grad ? 1 : 0;
    double *ckpt = mem.get(nadj * nseg * 5 * G + 1);
    double *seg = mem.get(nadj * ckpt_every * per_step + 1);
    double *acc = mem.get(5 * I);
    double *lam = mem.get(nadj * 5 * I + 1);
    double *lam_k[4] = {mem.get(nadj * 5 * I + 1), mem.get(nadj * 5 * I + 1),
                        mem.get(nadj * 5 * I + 1), mem.get(nadj * 5 * I + 1)};
    double *lam_X = mem.get(nadj * 5 * I + 1);
    double *lam_t = mem.get(nadj * 5 * I + 1);
    double *lam_W = mem.get(nadj * 6 * G + 1);
    r.mc.geo = mem.get(nadj * kGeoN * G + 1);
    double *lam_par = mem.get(2);
    double *terms = mem.get(I);
    const long long nb = (I + kBlock - 1) / kBlock;
    double *part = mem.get(2 * nb);
    double *mdot = mem.get(nsteps + 1);
    double *X0 = mem.get(5 * G);
    double *fx[5];
    double *fk[4];
    double *facc3 = mem.get(5 * I);
    for (int q = 1; q < 5; ++q)
        fx[q] = mem.get(5 * G);
    for (int q = 0; q < 4; ++q)
        fk[q] = mem.get(5 * I);
    if (!r.work || !ckpt || !seg || !acc || !lam || !lam_k[3] || !lam_X ||
        !lam_t || !lam_W || !lam_par || !terms || !part || !mdot || !X0 ||
        !fx[4] || !fk[3] || !facc3 || !r.mc.geo)
        return GRHD_ERR_DEVICE;
    r.p = make_pools(c, r.work, 0);
    hipStream_t s = r.s;
    launch_status_init(c, s);
    fill0(r.work, GRHD_NUM_COMPS * G, s);
    fill0(mdot, nsteps + 1, s);
    fill0(lam_par, 2, s);
    if (grad)
        fill0(lam, 5 * I, s);
    (void)hipMemcpyAsync(X0, state + (long long)cb * G,
                         sizeof(double) * 5 * G, hipMemcpyDeviceToDevice, s);

    auto mdot_of = [&](const double *X, int n) {
        hipLaunchKernelGGL(k_mdot_terms, dim3(grid_of(I)), dim3(kBlock), 0, s,
                           b, r.o, r.pa, r0, X, terms);
        hipLaunchKernelGGL(k_sum_partial, dim3((unsigned)nb), dim3(kBlock), 0,
                           s, terms, I, part);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, s, part, nb,
                           mdot + n);
    };

    /* forward sweep: checkpoints and J */
    fx[0] = X0;
    for (int n = 0; n < nsteps; ++n)
    {
        if (grad && n % ckpt_every == 0)
            (void)hipMemcpyAsync(ckpt + (long long)(n / ckpt_every) * 5 * G,
                                 X0, sizeof(double) * 5 * G,
                                 hipMemcpyDeviceToDevice, s);
        adj_fwd_step(r, X0, fx, fk, acc, facc3);
        (void)hipMemcpyAsync(X0, fx[4], sizeof(double) * 5 * G,
                             hipMemcpyDeviceToDevice, s);
        if (n + 1 >= first_step)
            mdot_of(X0, n + 1);
    }
    if ((rc = finish(c, s, err, nullptr, false)))
        return rc;
    std::vector<double> hm(nsteps + 1);
    if (hipMemcpy(hm.data(), mdot, sizeof(double) * (nsteps + 1),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return GRHD_ERR_DEVICE;
    double Jv = 0.0;
    for (int n = first_step; n <= nsteps; ++n)
        Jv += dt * hm[n];
    *J = Jv;
    if (!grad)
        return GRHD_OK;

    /* the cells' metric spin tangents, once for the whole sweep */
    hipLaunchKernelGGL(k_adj_metric_cache, dim3(grid_of(G)), dim3(kBlock), 0,
                       s, b, r.pa, r.mc);

    /* reverse sweep, one checkpoint segment at a time */
    const unsigned gi = grid_of(I);
    for (int sgm = nseg - 1; sgm >= 0; --sgm)
    {
        const int n0 = sgm * ckpt_every;
        const int n1 = std::min(nsteps, n0 + ckpt_every);
        /* recompute the segment, keeping each step's stage data */
        (void)hipMemcpyAsync(X0, ckpt + (long long)sgm * 5 * G,
                             sizeof(double) * 5 * G, hipMemcpyDeviceToDevice,
                             s);
        for (int n = n0; n < n1; ++n)
        {
            double *base = seg + (long long)(n - n0) * per_step;
            double *X[5];
            double *K[4];
            X[0] = n == n0 ? ckpt + (long long)sgm * 5 * G
                           : seg + (long long)(n - n0 - 1) * per_step +
                                 3 * 5 * G;
            for (int q = 1; q < 5; ++q)
                X[q] = base + (q - 1) * 5 * G;
            for (int q = 0; q < 4; ++q)
                K[q] = base + 4 * 5 * G + q * 5 * I;
            double *a3 = base + 4 * 5 * G + 4 * 5 * I;
            adj_fwd_step(r, X[0], X, K, acc, a3);
        }
        for (int n = n1 - 1; n >= n0; --n)
        {
            double *base = seg + (long long)(n - n0) * per_step;
            const double *X[5];
            const double *K[4];
            X[0] = n == n0 ? ckpt + (long long)sgm * 5 * G
                           : seg + (long long)(n - n0 - 1) * per_step +
                                 3 * 5 * G;
            for (int q = 1; q < 5; ++q)
                X[q] = base + (q - 1) * 5 * G;
            for (int q = 0; q < 4; ++q)
                K[q] = base + 4 * 5 * G + q * 5 * I;
            const double *a3 = base + 4 * 5 * G + 4 * 5 * I;
            /* the objective's seed at U_{n+1} */
            if (n + 1 >= first_step)
                hipLaunchKernelGGL(k_adj_mdot, dim3(gi), dim3(kBlock), 0, s,
                                   b, r.o, r.pa, r0, dt, X[4], lam, lam_par);
            const double c2 = dt / 2.0, c3 = dt / 3.0, c6 = dt / 6.0;
            /* U_4 = R(acc3 + dt/6 k4) */
            hipLaunchKernelGGL(k_adj_recover, dim3(gi), dim3(kBlock), 0, s, b,
                               r.o, r.pa, 4, dt, X[0], K[3], a3, lam, lam_t,
                               lam_par);
            /* lam_U0 accumulates in lam_X's partner: keep it in `acc` */
            double *lam_u0 = acc;
            (void)hipMemcpyAsync(lam_u0, lam_t, sizeof(double) * 5 * I,
                                 hipMemcpyDeviceToDevice, s);
            fill0(lam_k[0], 5 * I, s);
            fill0(lam_k[1], 5 * I, s);
            fill0(lam_k[2], 5 * I, s);
            fill0(lam_k[3], 5 * I, s);
            hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock), 0,
                               s, 5 * I, c6, lam_t, lam_k[3]);
            hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock), 0,
                               s, 5 * I, c6, lam_t, lam_k[0]);
            hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock), 0,
                               s, 5 * I, c3, lam_t, lam_k[1]);
            hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock), 0,
                               s, 5 * I, c3, lam_t, lam_k[2]);
            /* stages 3, 2, 1: lam_{X_s} = L'(X_s)^T lam_k[s];
             * X_s = R(U0 + c_s k_s) */
            const double cst[3] = {c2, c2, dt};
            for (int sg = 3; sg >= 1; --sg)
            {
                adj_rhs_transpose(r, X[sg], lam_k[sg], lam_X, lam_W, lam_par);
                hipLaunchKernelGGL(k_adj_recover, dim3(gi), dim3(kBlock), 0, s,
                                   b, r.o, r.pa, sg, dt, X[0], K[sg - 1],
                                   a3, lam_X, lam_t, lam_par);
                hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock),
                                   0, s, 5 * I, 1.0, lam_t, lam_u0);
                hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock),
                                   0, s, 5 * I, cst[sg - 1], lam_t,
                                   lam_k[sg - 1]);
            }
            /* k1 = L(U0) */
            adj_rhs_transpose(r, X[0], lam_k[0], lam_X, lam_W, lam_par);
            hipLaunchKernelGGL(k_axpy, dim3(grid_of(5 * I)), dim3(kBlock), 0,
                               s, 5 * I, 1.0, lam_X, lam_u0);
            (void)hipMemcpyAsync(lam, lam_u0, sizeof(double) * 5 * I,
                                 hipMemcpyDeviceToDevice, s);
        }
    }
    (void)hipMemcpyAsync(dJ_dU0, lam, sizeof(double) * 5 * I,
                         hipMemcpyDeviceToDevice, s);
    double hp[2];
    if (hipMemcpyAsync(hp, lam_par, sizeof(hp), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess ||
        hipGetLastError() != hipSuccess)
        return GRHD_ERR_DEVICE;
    dJ_dparams[0] = hp[0];
    dJ_dparams[1] = hp[1];
    return GRHD_OK;
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
        hipHostMalloc((void **)&c->hstat, sizeof(DevStatus)) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side[0], hipStreamNonBlocking) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&c->side[1], hipStreamNonBlocking) !=
            hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) !=
            hipSuccess ||
        hipEventCreateWithFlags(&c->join[0], hipEventDisableTiming) !=
            hipSuccess ||
        hipEventCreateWithFlags(&c->join[1], hipEventDisableTiming) !=
            hipSuccess ||
        hipEventCreate(&c->tune_ev[0]) != hipSuccess ||
        hipEventCreate(&c->tune_ev[1]) != hipSuccess)
    {
        grhd_ctx_destroy(c);
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
    free_plan(c->plan);
    (void)hipFree(c->W);
    (void)hipFree(c->wbad);
    for (int d = 0; d < 3; ++d)
        (void)hipFree(c->F[d]);
    (void)hipFree(c->list);
    (void)hipFree(c->clist);
    (void)hipFree(c->src);
    (void)hipFree(c->UA);
    (void)hipFree(c->UB);
    (void)hipFree(c->acc);
    (void)hipFree(c->diag);
    for (hipEvent_t e : c->events)
        (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i)
    {
        if (c->side[i])
            (void)hipStreamDestroy(c->side[i]);
        if (c->join[i])
            (void)hipEventDestroy(c->join[i]);
        if (c->tune_ev[i])
            (void)hipEventDestroy(c->tune_ev[i]);
    }
    if (c->fork)
        (void)hipEventDestroy(c->fork);
    (void)hipFree(c->dstat);
    (void)hipHostFree(c->hstat);
    delete c;
}

/* ---------------- one box ---------------- */
int grhd_compute_rhs_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double time,
                         double *state, int ncomp, int comp_base, double *rhs,
                         unsigned char *source_mask, grhd_error *err,
                         void *stream)
{
    if (!box)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    /* per-FArrayBox semantics: ghosts inside the domain are the caller's
     * (as after exchange()); outflow ghosts are filled here */
    return level_rhs(ctx, &lb, 1, params, background, time, state, ncomp,
                     comp_base, rhs, source_mask, false, err, stream);
}

int grhd_recover_primitives_box(grhd_ctx *ctx, const grhd_box *box,
                                const grhd_params *params,
                                const grhd_background *background,
                                double time, double *state, int ncomp,
                                int comp_base, unsigned char *mask,
                                grhd_diagnostics *diagnostics, void *stream)
{
    if (!box)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    return level_recover(ctx, &lb, 1, params, background, time, state, ncomp,
                         comp_base, mask, diagnostics, stream);
}

int grhd_update_conserved_box(grhd_ctx *ctx, const grhd_box *box,
                              const grhd_params *params,
                              const grhd_background *background,
                              double *state, int ncomp, int comp_base,
                              const double *rhs, double dt, int recover,
                              double recovery_time, unsigned char *mask,
                              grhd_diagnostics *diagnostics, void *stream)
{
    if (!box)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    return level_update(ctx, &lb, 1, params, background, state, ncomp,
                        comp_base, rhs, dt, recover, recovery_time, mask,
                        diagnostics, stream);
}

int grhd_compute_max_inverse_dt_box(grhd_ctx *ctx, const grhd_box *box,
                                    const grhd_params *params,
                                    const grhd_background *background,
                                    double time, const double *state,
                                    int ncomp, int comp_base,
                                    double *max_inverse_dt, grhd_error *err,
                                    void *stream)
{
    if (!box)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    return level_max_inv_dt(ctx, &lb, 1, params, background, time, state,
                            ncomp, comp_base, max_inverse_dt, err, stream);
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
    if (!box)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    return level_ssprk2(ctx, &lb, 1, params, background, time, dt, state,
                        ncomp, comp_base, mask, false, diagnostics, err,
                        stream);
}

int grhd_advance_rk4_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double time,
                         double dt, int nsteps, double *state, int ncomp,
                         int comp_base, unsigned char *mask,
                         grhd_diagnostics *diagnostics, grhd_error *err,
                         void *stream)
{
    if (!box)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    return level_rk4(ctx, &lb, 1, params, background, time, dt, nsteps, state,
                     ncomp, comp_base, mask, false, diagnostics, err, stream);
}

int grhd_advance_rk4_box_noalloc(const grhd_box *box,
                                 const grhd_params *params,
                                 const grhd_background *background,
                                 double time, double dt, int nsteps,
                                 double *state, int ncomp, int comp_base,
                                 unsigned char *mask,
                                 grhd_diagnostics *diagnostics,
                                 grhd_error *err, void *stream)
{
    return rk4_box_noalloc(box, params, background, time, dt, nsteps, state,
                           ncomp, comp_base, mask, diagnostics, err, stream);
}

int grhd_cell_geometry_box(grhd_ctx *ctx, const grhd_box *box,
                           const grhd_background *background, double time,
                           double *out, void *stream)
{
    if (!ctx || !box || !background || !out)
        return GRHD_ERR_INVALID_ARGUMENT;
    const grhd_level_box lb = one_box(box);
    int rc = prepare(ctx, &lb, 1, GRHD_NUM_COMPS);
    if (rc)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    GRHD_LAUNCH(ctx, C_ALL, k_geometry, s, make_background(background, time),
                out);
    if (hipStreamSynchronize(s) != hipSuccess)
        return GRHD_ERR_DEVICE;
    return hipGetLastError() == hipSuccess ? GRHD_OK : GRHD_ERR_DEVICE;
}

/* The kernels' FP64 division and square root (qdiv, qsqrt), elementwise,
 * grid-stride so a grid of at most kMaxDivSqrtBlocks covers any n. */
constexpr unsigned kMaxDivSqrtBlocks = 4096;

__global__ void __launch_bounds__(kBlock)
    k_div_sqrt(const double *__restrict__ a, const double *__restrict__ b,
               double *__restrict__ q, double *__restrict__ r, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < n; t += stride)
    {
        q[t] = qdiv(a[t], b[t]);
        r[t] = qsqrt(a[t]);
    }
}

int grhd_fp64_div_sqrt(const double *a, const double *b, double *quotient,
                       double *root, long long n, void *stream)
{
    if (n < 0 || (n > 0 && (!a || !b || !quotient || !root)))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (n == 0)
        return GRHD_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned grid = std::min<long long>(grid_of(n), kMaxDivSqrtBlocks);
    hipLaunchKernelGGL(k_div_sqrt, dim3(grid), dim3(kBlock), 0, s, a, b,
                       quotient, root, n);
    if (hipGetLastError() != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return GRHD_ERR_DEVICE;
    return GRHD_OK;
}

/* ---------------- a level ---------------- */
int grhd_level_compute_rhs(grhd_ctx *ctx, const grhd_level_box *boxes,
                           int nbox, const grhd_params *params,
                           const grhd_background *background, double time,
                           double *state_pool, int ncomp, int comp_base,
                           double *rhs_pool, unsigned char *source_mask,
                           grhd_error *err, void *stream)
{
    return level_rhs(ctx, boxes, nbox, params, background, time, state_pool,
                     ncomp, comp_base, rhs_pool, source_mask, true, err,
                     stream);
}

int grhd_level_add_conserved_rhs(grhd_ctx *ctx, const grhd_level_box *boxes,
                                 int nbox, const grhd_params *params,
                                 const grhd_background *background,
                                 double time, double *state_pool, int ncomp,
                                 int comp_base, double *rhs_pool,
                                 grhd_error *err, void *stream)
{
    return level_rhs(ctx, boxes, nbox, params, background, time, state_pool,
                     ncomp, comp_base, rhs_pool, nullptr, true, err, stream,
                     true);
}

int grhd_level_recover_primitives(grhd_ctx *ctx, const grhd_level_box *boxes,
                                  int nbox, const grhd_params *params,
                                  const grhd_background *background,
                                  double time, double *state_pool, int ncomp,
                                  int comp_base, unsigned char *mask,
                                  grhd_diagnostics *diagnostics, void *stream)
{
    return level_recover(ctx, boxes, nbox, params, background, time,
                         state_pool, ncomp, comp_base, mask, diagnostics,
                         stream);
}

int grhd_level_compute_max_inverse_dt(grhd_ctx *ctx,
                                      const grhd_level_box *boxes, int nbox,
                                      const grhd_params *params,
                                      const grhd_background *background,
                                      double time, const double *state_pool,
                                      int ncomp, int comp_base,
                                      double *max_inverse_dt, grhd_error *err,
                                      void *stream)
{
    return level_max_inv_dt(ctx, boxes, nbox, params, background, time,
                            state_pool, ncomp, comp_base, max_inverse_dt, err,
                            stream);
}

int grhd_level_conserved_sum(grhd_ctx *ctx, const grhd_level_box *boxes,
                             int nbox, const double *state_pool, int ncomp,
                             int comp_base, double *sum, void *stream)
{
    if (!sum)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (nbox == 0)
    {
        const int rc = empty_level(ctx, nullptr, nullptr, ncomp, comp_base);
        if (!rc)
            for (int m = 0; m < 5; ++m)
                sum[m] = 0.0;
        return rc;
    }
    const int rc = level_diagnostics(ctx, boxes, nbox, state_pool, ncomp,
                                     comp_base, stream);
    if (rc)
        return rc;
    for (int m = 0; m < 5; ++m)
        sum[m] = ctx->hstat->sums[m];
    return GRHD_OK;
}

int grhd_level_conserved_state_is_finite(grhd_ctx *ctx,
                                         const grhd_level_box *boxes,
                                         int nbox, const double *state_pool,
                                         int ncomp, int comp_base,
                                         int *is_finite, void *stream)
{
    if (!is_finite)
        return GRHD_ERR_INVALID_ARGUMENT;
    if (nbox == 0)
    {
        const int rc = empty_level(ctx, nullptr, nullptr, ncomp, comp_base);
        if (!rc)
            *is_finite = 1;
        return rc;
    }
    const int rc = level_diagnostics(ctx, boxes, nbox, state_pool, ncomp,
                                     comp_base, stream);
    if (rc)
        return rc;
    *is_finite = ctx->hstat->nonfinite == 0 ? 1 : 0;
    return GRHD_OK;
}

int grhd_level_update_conserved(grhd_ctx *ctx, const grhd_level_box *boxes,
                                int nbox, const grhd_params *params,
                                const grhd_background *background,
                                double *state_pool, int ncomp, int comp_base,
                                const double *rhs_pool, double dt, int recover,
                                double recovery_time, unsigned char *mask,
                                grhd_diagnostics *diagnostics, void *stream)
{
    return level_update(ctx, boxes, nbox, params, background, state_pool,
                        ncomp, comp_base, rhs_pool, dt, recover,
                        recovery_time, mask, diagnostics, stream);
}

int grhd_rk4_adjoint_box(grhd_ctx *ctx, const grhd_box *box,
                         const grhd_params *params,
                         const grhd_background *background, double dt,
                         int nsteps, const double *state, int ncomp,
                         int comp_base, double shell_radius, int first_step,
                         int checkpoint_every, double *objective,
                         double *dJ_dconserved, double *dJ_dparams,
                         grhd_error *err, void *stream)
{
    return rk4_adjoint_box(ctx, box, params, background, 0.0, dt, nsteps,
                           state, ncomp, comp_base, shell_radius, first_step,
                           checkpoint_every, objective, dJ_dconserved,
                           dJ_dparams, err, stream);
}

int grhd_level_checkpoint_pack(grhd_ctx *ctx, const grhd_level_box *boxes,
                               int nbox, const double *state_pool, int ncomp,
                               const int *domain_lo, const int *domain_hi,
                               int ghosts, double *checkpoint_values,
                               void *stream)
{
    return level_checkpoint(ctx, boxes, nbox, const_cast<double *>(state_pool),
                            ncomp, domain_lo, domain_hi, ghosts,
                            checkpoint_values, true, stream);
}

int grhd_level_checkpoint_unpack(grhd_ctx *ctx, const grhd_level_box *boxes,
                                 int nbox, const double *checkpoint_values,
                                 int ncomp, const int *domain_lo,
                                 const int *domain_hi, int ghosts,
                                 double *state_pool, void *stream)
{
    return level_checkpoint(ctx, boxes, nbox, state_pool, ncomp, domain_lo,
                            domain_hi, ghosts,
                            const_cast<double *>(checkpoint_values), false,
                            stream);
}

int grhd_level_advance_ssprk2(grhd_ctx *ctx, const grhd_level_box *boxes,
                              int nbox, const grhd_params *params,
                              const grhd_background *background, double time,
                              double dt, double *state_pool, int ncomp,
                              int comp_base, unsigned char *mask,
                              grhd_diagnostics *diagnostics, grhd_error *err,
                              void *stream)
{
    return level_ssprk2(ctx, boxes, nbox, params, background, time, dt,
                        state_pool, ncomp, comp_base, mask, true, diagnostics,
                        err, stream);
}

int grhd_level_advance_rk4(grhd_ctx *ctx, const grhd_level_box *boxes,
                           int nbox, const grhd_params *params,
                           const grhd_background *background, double time,
                           double dt, int nsteps, double *state_pool,
                           int ncomp, int comp_base, unsigned char *mask,
                           grhd_diagnostics *diagnostics, grhd_error *err,
                           void *stream)
{
    return level_rk4(ctx, boxes, nbox, params, background, time, dt, nsteps,
                     state_pool, ncomp, comp_base, mask, true, diagnostics,
                     err, stream);
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

int grhd_ctx_set_autotune(grhd_ctx *ctx, int enable)
{
    if (!ctx)
        return GRHD_ERR_INVALID_ARGUMENT;
    ctx->autotune = enable != 0;
    return GRHD_OK;
}

int grhd_ctx_set_face_launch(grhd_ctx *ctx, int mixed, int walk, int fused)
{
    if (!ctx || (mixed != 0 && mixed != 1) || (fused != 0 && fused != 1))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (walk == -1)
    {
        ctx->pin_seg[mixed] = -1;
        return GRHD_OK;
    }
    for (int k = 0; k < kNumSegs; ++k)
        if (kSegs[k] == walk)
        {
            ctx->pin_seg[mixed] = k;
            ctx->pin_fused[mixed] = fused != 0;
            return GRHD_OK;
        }
    return GRHD_ERR_INVALID_ARGUMENT;
}

int grhd_ctx_get_face_launch(grhd_ctx *ctx, int mixed, int *walk, int *fused,
                             int *tuned, double *ms)
{
    if (!ctx || ctx->plan.ncomp < 0 || (mixed != 0 && mixed != 1))
        return GRHD_ERR_INVALID_ARGUMENT;
    const Plan &pl = ctx->plan;
    const int pin = ctx->pin_seg[mixed];
    const bool t = pin < 0 && ctx->autotune && pl.tuned[mixed];
    if (walk)
        *walk = kSegs[pin >= 0 ? pin : t ? pl.seg_pick[mixed] : pl.seg_rule];
    if (fused)
        *fused = pin >= 0 ? ctx->pin_fused[mixed]
                 : t      ? pl.fused_pick[mixed]
                          : pl.tot_i < kFusedFacesCells;
    if (tuned)
        *tuned = t;
    if (ms)
        *ms = t ? (double)pl.tuned_ms[mixed] : 0.0;
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
