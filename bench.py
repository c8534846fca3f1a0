"""Benchmark: GRHD cell-updates/s per RK4 step, plain FP64, one MI355X.

Workload (the 128^3 plain-FP64-at-scale configuration): one 128^3 box with 3
ghost cells on [-32, 32]^3 M (dx = 0.5 M), Kerr a = 0.9 M in Kerr-Schild
coordinates, seeded Gaussian torus at R = 12 M with Keplerian rotation,
Gamma = 4/3, minmod limiter, RK4 with dt = 0.25 dx.  A step is one full RK4
step of the box (4 right-hand sides with ghost fill, reconstruction, HLLE,
sources, and 4 recoveries with atmosphere resets); the state is resident in
HBM before timing starts.

Output: one JSON line with the throughput, a roofline record for the
dominant kernel (its average launch time measured with HIP events on the
launch stream during the timed steps) and a CPU baseline (the oracle -- a C
restatement of the reference's CPU path with its call structure -- timed on
this host's cores on one RK4 step of the same box).

Roofline units are algorithmic, not issued: FLOPs per cell from the
de-duplicated operation count of the oracle (oracle/opcount.cpp, frozen in
profiles/algorithmic_counts.json), bytes from the values each kernel must
move once.  The bound of the dominant kernel is the roof its algorithmic
intensity (FLOP/B) meets first against the MI355X ridge of
78.6 TFLOP/s / 8 TB/s = 9.8 FLOP/B; the other roof is reported beside it,
and so is the whole step's (F_alg and 689 B per cell-update).  Nothing here
is a dense contraction, so MFMA peaks do not apply.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("GRHD cell-updates/s per RK4 step, plain FP64 vs mixed/batched "
          "at fixed error")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
RIDGE_FLOP_PER_BYTE = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
# VALU issue model: a wave64 instruction holds a SIMD 4 cycles, an FP64
# transcendental (rcp, rsq) 16; 1024 SIMDs at the 2.4 GHz peak clock
SIMDS = 1024
PEAK_CLOCK_HZ = 2.4e9

# Which parts of the de-duplicated per-cell count (profiles/
# algorithmic_counts.json, flop_per_cell_stage_by_part) each kernel class
# performs: k_cells evaluates the cell's metric with its derivatives, the
# divergence, the sources, the stage update and the recovery; the face
# launch the slopes and the faces.
KERNEL_PARTS = {"cells": ("metric", "divergence", "sources", "update",
                          "recovery"),
                "faces": ("slopes", "faces")}

# Algorithmic HBM bytes per unit of each kernel class: every value the
# kernel must read or write once, in FP64.
#   faces_d: per face, the primitives of one cell (6 x 8 B) read, one flux
#            (5 x 8 B) written.
#   cells:   per interior cell, three face fluxes (each face is shared by
#            two cells: 3 x 5 x 8 B), W (6 x 8), U_s, U0 and the RK
#            accumulator read (3 x 5 x 8), U_next, the accumulator and W of
#            the next stage written (5 x 8 + 5 x 8 + 6 x 8).
#   prims:   per exterior cell: one state read, state + W written.
ALG_BYTES = {"faces": 88,
             "cells": 120 + 48 + 120 + 40 + 40 + 48, "prims": 40 + 40 + 48}
# Whole RK4 step, per cell-update: FP64 state traffic of a fused RK4 (stage
# 1: U0 read, U1 and acc written, 120 B; stages 2 and 3: U_s, U0, acc read,
# U_{s+1}, acc written, 200 B each; stage 4: U3, acc read, U written,
# 120 B), plus the primitives (48 B) and the mask byte.
STEP_ALG_BYTES = 120 + 200 + 200 + 120 + 48 + 1


# The "faces" timing class covers the three face directions of a stage
# (one fused launch, or three concurrent ones), so its unit count, FLOPs and
# traffic are those of all three.


def units_per_launch(cls, n, g):
    if cls == "faces":
        return 3 * (n + 1) * n * n
    if cls == "cells":
        return n ** 3
    if cls == "prims":
        return (n + 2 * g) ** 3 - n ** 3
    return 0


def measured_traffic(cls, n):
    """HBM bytes per launch of kernel class `cls` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-kernel calibrated), for
    the 128^3 box it was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get("bytes_per_launch", {}).get(cls)


def algorithmic_counts(config):
    """The frozen algorithmic count of one workload (None if absent)."""
    path = os.path.join(ROOT, "profiles", "algorithmic_counts.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["configs"].get(config)


def measured_step_bytes(n):
    """Counter HBM bytes of one RK4 step of the n^3 box (4 stages of each
    kernel class, summed from the committed FETCH/WRITE summary); None
    when the summary does not cover this size."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_rk4_step")


def valu_issue(cls, n, avg_launch_s):
    """The kernel class's issue time at peak, from the committed counts of
    VALU (and FP64 transcendental) wavefront instructions per launch on the
    128^3 box, as a fraction of its measured launch time; None otherwise."""
    path = os.path.join(ROOT, "profiles", "fp64_flops.json")
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    insts = data.get("valu_insts_per_launch", {}).get(cls)
    trans = data.get("trans_f64_per_launch", {}).get(cls)
    if insts is None or trans is None:
        return None
    issue_s = (4.0 * insts + 12.0 * trans) / (SIMDS * PEAK_CLOCK_HZ)
    return {"valu_insts_per_launch": insts,
            "trans_f64_per_launch": trans,
            "issue_ms_at_peak": issue_s * 1e3,
            "frac": issue_s / avg_launch_s}


def measured_flops(cls, n):
    """FP64 FLOPs per launch of kernel class `cls` from the committed
    rocprofv3 summary (SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes), for the 128^3
    box it was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "fp64_flops.json")
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get("flops_per_launch", {}).get(cls)


def cpu_baseline(w, steps=3):
    """Oracle RK4 steps of the same box on this host's cores (about 10 s on
    16 cores for the 128^3 box)."""
    import numpy as np

    import oracle
    O = oracle.Oracle(w.background, w.box, w.options)
    st = np.ascontiguousarray(w.state).copy()
    t0 = time.perf_counter()
    O.advance_rk4(st, w.dt, steps)
    t1 = time.perf_counter()
    cells = w.box.num_interior
    return {"value": cells * steps / (t1 - t0), "unit": "cell-updates/s",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": "%d RK4 steps of the same %d^3 box (%d interior cells) "
                      "from the same initial state, %.1f s wall, OpenMP over "
                      "cells, reference call structure"
                      % (steps, w.box.interior_shape[0], cells, t1 - t0)}


def bench_c1(args):
    """The reference's own CPU-runnable case: the 32^3 Michel-like box on
    [-8, 8]^3 (Schwarzschild, MC limiter, dt = 0.25 dx), plain FP64 RK4 on
    the GPU, and the oracle on this host with all its threads and with one.
    At 32,768 cells the GPU is launch-latency bound, not compute bound.
    Returns the JSON record."""
    import numpy as np
    import torch

    import oracle
    from grdzhadzha_amd import workloads
    from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator

    w = workloads.michel_box_32()
    cells = w.box.num_interior
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options)
    op.ctx.set_autotune(args.autotune)
    init = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
    steps = 200
    op.advance_rk4(init.clone(), w.dt, 2)
    st = init.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.advance_rk4(st, w.dt, steps)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    O = oracle.Oracle(w.background, w.box, w.options)
    full = oracle.num_threads()

    def cpu(k):
        s = np.ascontiguousarray(w.state).copy()
        t = time.perf_counter()
        O.advance_rk4(s, w.dt, k)
        return cells * k / (time.perf_counter() - t)

    c_all = cpu(20)
    oracle.set_num_threads(1)
    try:
        c_one = cpu(3)
    finally:
        oracle.set_num_threads(full)
    print("bench: C1 32^3 Michel GPU %.3e cell-updates/s (%.3f ms/step); "
          "oracle %.3e (%d threads), %.3e (1 thread)"
          % (cells * steps / tg, tg / steps * 1e3, c_all, full, c_one),
          file=sys.stderr, flush=True)
    return {"workload": "C1 michel-like 32^3 on [-8,8]^3 Schwarzschild, MC, "
                        "plain FP64 RK4",
            "interior_cells": cells, "steps": steps,
            "value": cells * steps / tg, "ms_per_step": tg / steps * 1e3,
            "face_launch": op.ctx.face_launch(),
            "cpu_oracle": {"value": c_all, "cores": full, "steps": 20},
            "cpu_oracle_1_thread": {"value": c_one, "cores": 1, "steps": 3}}


def bench_sizes(args):
    """Plain FP64 RK4 throughput against box size: the 128^3 torus recipe
    (dx = 0.5, box n/4 M wide, Kerr a = 0.9) at n = 32 .. args.max_n, each
    timed over enough steps for about a second of GPU time.  Shows where
    the launch-latency-bound regime ends and how much of the device the
    largest boxes occupy (device memory in use, the library's scratch and
    the caching allocator included).  Returns the JSON record."""
    import numpy as np
    import torch

    from grdzhadzha_amd import workloads
    from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator

    rows = []
    for n in (32, 64, 96, 128, 192, 256, 384, 512):
        if n > args.max_n:
            break
        print("bench: building the %d^3 torus" % n, file=sys.stderr,
              flush=True)
        w = workloads.torus_box(n=n)
        op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box,
                                                  w.options)
        st = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
        del w.state
        op.advance_rk4(st.clone(), w.dt, 1)
        steps = max(2, min(200, int(2.0e8 / n ** 3)))
        s = st.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.advance_rk4(s, w.dt, steps)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        free, total = torch.cuda.mem_get_info()
        rows.append({"n": n, "steps": steps, "ms_per_step": t / steps * 1e3,
                     "value": n ** 3 * steps / t,
                     "device_used_gb": (total - free) / 1e9})
        print("bench: %d^3 %.3e cell-updates/s (%.3f ms/step)"
              % (n, rows[-1]["value"], rows[-1]["ms_per_step"]),
              file=sys.stderr, flush=True)
        del op, st, s
        torch.cuda.empty_cache()
    return {"workload": "torus recipe, plain FP64 RK4, box size sweep",
            "unit": "cell-updates/s", "sizes": rows}


def calibrate_copy():
    """Five dst.copy_(src) of 64 Mi float64 (512 MiB read + 512 MiB write
    per dispatch): the known-size kernel that converts rocprofv3's
    FETCH_SIZE / WRITE_SIZE readings into bytes."""
    import torch
    n = 64 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copy calibration: 5 x %d bytes read + written" % (8 * n),
          flush=True)


def bench_batched(args):
    """The 3-level x 180-box circular-binary PN box set: every box of a
    level advanced by one level call per RK4 run (box table, one launch per
    kernel per level), against the plain path (one call per box) on the
    same boxes.  Returns the JSON record."""
    import torch

    from grdzhadzha_amd import workloads
    from grdzhadzha_amd.level import LevelOperator
    from grdzhadzha_amd.operator import (
        Context, TimeDependentFixedBGLevelDataFiniteVolumeOperator as TD)

    bg, opts, levels, nsteps = workloads.binary_pn_levels()
    steps = args.steps if args.steps_given else nsteps
    ctx = Context()
    ctx.set_autotune(args.autotune)
    ops = [LevelOperator(bg, boxes, opts, time_dependent=True, context=ctx)
           for boxes, _, _ in levels]
    init = [lev.pack(fabs) for lev, (_, fabs, _) in zip(ops, levels)]
    cells = sum(lev.num_interior for lev in ops)
    nbox = sum(len(b) for b, _, _ in levels)
    print("bench: binary-PN box set, %d levels, %d boxes, %d interior cells"
          % (len(levels), nbox, cells), file=sys.stderr, flush=True)

    def run_batched(k):
        pools = [p.clone() for p in init]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for lev, pool, (_, _, dt) in zip(ops, pools, levels):
            lev.advance_rk4(pool, dt, k, time=0.0)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run_batched(args.warmup or 1)
    tb = run_batched(steps)
    batched = cells * steps / tb

    # batched mixed precision on the same box set, from the same state:
    # throughput and relative L1 / Linf error per conserved variable over
    # every interior cell of every level against batched FP64 (which is
    # bit-exact with the oracle)
    import dataclasses
    o32 = dataclasses.replace(opts, mixed_precision=True)
    ops32 = [LevelOperator(bg, boxes, o32, time_dependent=True, context=ctx)
             for boxes, _, _ in levels]

    def run_levels(lops, k):
        pools = [p.clone() for p in init]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for lev, pool, (_, _, dt) in zip(lops, pools, levels):
            lev.advance_rk4(pool, dt, k, time=0.0)
        torch.cuda.synchronize()
        return pools, time.perf_counter() - t0

    names = ("D", "S1", "S2", "S3", "tau")

    def level_errors(p64, p32):
        num = {k: 0.0 for k in names}
        den = {k: 0.0 for k in names}
        mx = {k: 0.0 for k in names}
        mref = {k: 0.0 for k in names}
        smag = 0.0
        for lev, a64, a32, (boxes, _, _) in zip(ops, p64, p32, levels):
            for b, f64, f32 in zip(boxes, lev.unpack(a64), lev.unpack(a32)):
                sl = b.interior_slices()
                for c, name in enumerate(names):
                    ref = f64[c][sl]
                    diff = abs(f32[c][sl] - ref)
                    num[name] += float(diff.sum())
                    den[name] += float(abs(ref).sum())
                    mx[name] = max(mx[name], float(diff.max()))
                    mref[name] = max(mref[name], float(abs(ref).max()))
                smag += float(((f64[1][sl] ** 2 + f64[2][sl] ** 2 +
                                f64[3][sl] ** 2) ** 0.5).sum())
        out = {k: {"l1_rel": num[k] / den[k], "linf_rel": mx[k] / mref[k]}
               for k in names}
        # S_i normalised by the momentum magnitude sum |S| rather than by
        # sum |S_i|, which is near zero for S3 in a ring about z = 0
        for k in ("S1", "S2", "S3"):
            out[k]["l1_rel_to_abs_S"] = num[k] / smag
        return out

    # one step: the FP32 rounding of a single step, before branch decisions
    # (resets, floors, atmosphere) can come out differently
    err_one = level_errors(run_levels(ops, 1)[0], run_levels(ops32, 1)[0])
    run_levels(ops32, args.warmup or 1)
    p32, tm = run_levels(ops32, steps)
    p64, _ = run_levels(ops, steps)
    mixed = cells * steps / tm
    err_m = level_errors(p64, p32)
    del p32
    # guarded mixed: FP64 faces beside cells whose shift has |beta| >=
    # args.batched_shift (the binary's shift is small: the default 0.05
    # covers the neighbourhood of each mass)
    og = dataclasses.replace(opts, mixed_precision=True,
                             mixed_fp64_shift=args.batched_shift)
    opsg = [LevelOperator(bg, boxes, og, time_dependent=True, context=ctx)
            for boxes, _, _ in levels]
    run_levels(opsg, 1)
    pg, tg = run_levels(opsg, steps)
    guarded = cells * steps / tg
    err_g = level_errors(p64, pg)
    del pg, p64
    print("bench: batched mixed %.3f s (%.3e cell-updates/s), max L1 rel "
          "error %.3e; guarded %.3f s (%.3e), max L1 rel error %.3e"
          % (tm, mixed, max(e["l1_rel"] for e in err_m.values()), tg,
             guarded, max(e["l1_rel"] for e in err_g.values())),
          file=sys.stderr, flush=True)

    # plain path: one call per box, the same boxes and steps; each box has
    # its own context, so its launch plan is built once, not per call
    plain_steps = max(1, min(steps, args.plain_steps))
    per_box = []
    for lev, pool, (boxes, _, dt) in zip(ops, init, levels):
        for n, (b, fab) in enumerate(zip(boxes, lev.unpack(pool))):
            st = torch.from_numpy(fab.copy()).cuda()
            per_box.append((TD(bg, b, opts, context=Context()), st, dt))
    for op, st, dt in per_box:  # warm every per-box plan
        op.advance_rk4(st.clone(), dt, 1, time=0.0)
    sts = [st.clone() for _, st, _ in per_box]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (op, _, dt), st in zip(per_box, sts):
        op.advance_rk4(st, dt, plain_steps, time=0.0)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    plain = cells * plain_steps / tp
    # the face launch each level used (a call of no steps selects the
    # level's plan, with what was tuned on it, and runs nothing)
    face_launch = []
    for lev, pool, (_, _, dt) in zip(ops, init, levels):
        lev.advance_rk4(pool, dt, 0, time=0.0)
        face_launch.append({"fp64": ctx.face_launch(),
                            "fp32": ctx.face_launch(mixed=True)})
    print("bench: batched %.3f s for %d steps (%.3e cell-updates/s); plain "
          "%.3f s for %d steps (%.3e)" % (tb, steps, batched, tp,
                                           plain_steps, plain),
          file=sys.stderr, flush=True)
    return {"metric": METRIC, "value": batched, "unit": "cell-updates/s",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup or 1,
            "ms_per_step": tb / steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "binary-PN box set: 3 levels x 180 "
                       "seeded boxes (sides 8..64), circumbinary ring, "
                       "batched FP64 RK4 per level", "boxes": nbox,
                       "interior_cells": cells, "parallelism": "single GPU"},
            "plain_per_box": {"value": plain, "steps": plain_steps,
                              "ms_per_step": tp / plain_steps * 1e3},
            "speedup_vs_plain": batched / plain,
            "face_launch_per_level": face_launch,
            "batched_mixed": {
                "value": mixed, "ms_per_step": tm / steps * 1e3,
                "speedup_vs_plain_per_box": mixed / plain,
                "speedup_vs_batched_fp64": mixed / batched,
                "error_vs_batched_fp64": err_m,
                "gate_1e-6": gate_1e6(err_m),
                "l1_rel_after_one_step": err_one},
            "batched_guarded_mixed": {
                "fp64_faces": "beside cells with |beta| >= %g"
                              % args.batched_shift,
                "value": guarded, "ms_per_step": tg / steps * 1e3,
                "speedup_vs_plain_per_box": guarded / plain,
                "speedup_vs_batched_fp64": guarded / batched,
                "error_vs_batched_fp64": err_g,
                "gate_1e-6": gate_1e6(err_g)}}


def gate_1e6(err):
    """The precision target: relative L1 error against FP64 <= 1e-6 for
    every conserved variable (and the accretion rate, when reported).  A
    speedup only counts as "at fixed error" where this holds."""
    failing = {k: e["l1_rel"] for k, e in err.items()
               if k != "mdot" and not e["l1_rel"] <= 1e-6}
    if "mdot" in err and not err["mdot"]["rel"] <= 1e-6:
        failing["mdot"] = err["mdot"]["rel"]
    return {"meets": not failing, "failing": failing,
            "speedup_is_at_fixed_error": not failing}


def bench_mixed(args):
    """Mixed precision on the 256^3 torus (the 128^3 recipe scaled 2x in
    radius: [-64, 64]^3, R = 24, sigma = 6, seed 4), 200 RK4 steps, against
    the FP64 path (which is bit-exact with the oracle), all on this GPU from
    the same state:
    - plain mixed: FP32 reconstruction, face metric and HLLE on every face;
    - guarded mixed: the same, except FP64 faces beside cells whose shift
      has |beta| >= args.mixed_shift (grhd_params.mixed_fp64_shift).
    Returns the JSON record: throughput, relative L1 / Linf error per
    conserved variable, the accretion rate through |x| = 2 r_+, and the
    cells whose final recovery mask differs from FP64's."""
    import dataclasses

    import numpy as np
    import torch

    from grdzhadzha_amd import workloads
    from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator

    n = args.mixed_n
    steps = args.mixed_steps
    w = workloads.torus_box(n=n, seed=4, R0=12.0 * n / 128,
                            sigma=3.0 * n / 128, nsteps=steps)
    op64 = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box,
                                                w.options)
    op64.ctx.set_autotune(args.autotune)
    variants = {
        "mixed": dataclasses.replace(w.options, mixed_precision=True),
        "guarded": dataclasses.replace(w.options, mixed_precision=True,
                                       mixed_fp64_shift=args.mixed_shift)}
    ops = {k: FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, o,
                                                   context=op64.ctx)
           for k, o in variants.items()}
    init = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
    cells = w.box.num_interior
    print("bench: mixed precision, %s, %d interior cells, %d steps"
          % (w.name, cells, steps), file=sys.stderr, flush=True)

    def run(op, k):
        st = init.clone()
        mask = torch.zeros(w.box.interior_shape, dtype=torch.uint8,
                           device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.advance_rk4(st, w.dt, k, mask=mask)
        torch.cuda.synchronize()
        return st, mask, time.perf_counter() - t0

    sl = (slice(None),) + w.box.interior_slices()
    names = ("D", "S1", "S2", "S3", "tau")

    def l1(b64, b32):
        return {name: float(np.abs(b32[c][sl[1:]] - b64[c][sl[1:]]).sum() /
                            np.abs(b64[c][sl[1:]]).sum())
                for c, name in enumerate(names)}

    spin = w.background.spin
    r_plus = 1.0 + (1.0 - spin * spin) ** 0.5
    # one step from the same state: the FP32 rounding of a single step,
    # before branch decisions (resets, floors, atmosphere) can diverge
    one64 = run(op64, 1)[0].cpu().numpy()
    s64, m64, t64 = run(op64, steps)
    a64 = s64.cpu().numpy()
    del s64
    md64 = workloads.mass_accretion_rate(w.box, a64, 1.0, spin, 2 * r_plus)
    smag = float(np.sqrt(sum(a64[c][sl[1:]] ** 2 for c in (1, 2, 3))).sum())
    v64 = cells * steps / t64
    rec = {}
    for key, op in ops.items():
        err_one = l1(one64, run(op, 1)[0].cpu().numpy())
        s32, m32, t32 = run(op, steps)
        flips = int((m32 != m64).sum())
        a32 = s32.cpu().numpy()
        del s32, m32
        err = {}
        for c, name in enumerate(names):
            ref = a64[c][sl[1:]]
            diff = np.abs(a32[c][sl[1:]] - ref)
            err[name] = {"l1_rel": float(diff.sum() / np.abs(ref).sum()),
                         "linf_rel": float(diff.max() / np.abs(ref).max())}
            if name.startswith("S"):
                # normalised by the momentum magnitude sum |S|
                err[name]["l1_rel_to_abs_S"] = float(diff.sum() / smag)
        md32 = workloads.mass_accretion_rate(w.box, a32, 1.0, spin,
                                             2 * r_plus)
        del a32
        err["mdot"] = {"fp64": md64, "mixed": md32,
                       "rel": abs(md32 - md64) / max(abs(md64), 1e-300)}
        v32 = cells * steps / t32
        print("bench: %s %.3f s (%.3e cell-updates/s, %.3fx FP64); max L1 "
              "rel error %.3e" % (key, t32, v32, v32 / v64,
                                  max(e["l1_rel"] for k, e in err.items()
                                      if k != "mdot")),
              file=sys.stderr, flush=True)
        rec[key] = {"value": v32, "speedup_vs_fp64": v32 / v64,
                    "gate_1e-6": gate_1e6(err),
                    "ms_per_step": t32 / steps * 1e3,
                    "error_vs_fp64": err,
                    "final_mask_cells_differing_from_fp64": flips,
                    "l1_rel_after_one_step": err_one}
    print("bench: FP64 %.3f s (%.3e cell-updates/s)" % (t64, v64),
          file=sys.stderr, flush=True)
    out = {"workload": "%s mixed FP32 faces / FP64 recovery+update RK4, "
                       "%d steps" % (w.name, steps),
           "interior_cells": cells, "steps": steps, "fp64_value": v64,
           "fp64_ms_per_step": t64 / steps * 1e3,
           "reference": "the FP64 GPU path, bit-exact with the oracle"}
    out.update(rec["mixed"])
    out["guarded"] = dict(rec["guarded"],
                          fp64_faces="beside cells with |beta| >= %g"
                          % args.mixed_shift)
    return out


def bench_adjoint(args):
    """Differentiation: 64^3 Michel-like inflow on [-16, 16]^3 (dx 0.5),
    Kerr a = 0.5, Gamma 4/3, MC limiter, 200 RK4 steps (dt = 0.25 dx).
    J = time-integrated rest-mass flux through r = 4 M over the last 50
    steps.  Gradient of J with respect to the initial density field, Gamma
    and a by the hand-written adjoint (checkpoints every 20 steps), through
    the torch initial-data map; checked against central differences
    (h = 1e-6) along 16 seeded smooth random directions of rho0 and in Gamma
    and a.  The horizon is excised (KerrSchild::check_if_excised, zero
    right-hand side inside) and the recovery tolerance is 1e-16: without
    excision, recoveries that fail near the ring singularity reset cells to
    atmosphere on either side of any perturbation, J jumps, and central
    differences grow like 1/h instead of converging -- the record
    `without_excision` shows it at h = 1e-4 and 1e-6 along 4 directions.
    Returns the JSON record."""
    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import adjoint_harness as ah

    from grdzhadzha_amd.operator import (Context,
                                         FixedBGLevelDataFiniteVolumeOperator)
    from grdzhadzha_amd.options import KerrSchildParams
    from grdzhadzha_amd.workloads import cube_box

    n, steps, window, r0, every = 64, args.adjoint_steps, 50, 4.0, 20
    gamma0, spin0 = 4.0 / 3.0, 0.5
    box = cube_box(n, 3, 16.0)
    dt = 0.25 * box.dx
    first = max(1, steps - window + 1)
    ctx = Context()
    rho0 = torch.from_numpy(ah.seed_rho0(box, 5)).cuda()
    print("bench: adjoint, michel-like inflow %d^3 Kerr a=%.1f, %d steps"
          % (n, spin0, steps), file=sys.stderr, flush=True)

    def f64(x, grad=False):
        return torch.tensor(x, dtype=torch.float64, requires_grad=grad)

    def run(excise, ndir, hs):
        kw = dict(tolerance=1e-16, excise=excise)
        opts = ah.options(gamma0, **kw)
        # plain forward of the same evolution, for the cost ratio
        with torch.no_grad():
            U0 = ah.michel_inflow_U0(box, rho0, f64(gamma0), f64(spin0),
                                     opts)
        op = FixedBGLevelDataFiniteVolumeOperator(
            KerrSchildParams(1.0, spin0), box, opts, context=ctx)
        st = ah.full_state(box, U0)
        op.advance_rk4(st.clone(), dt, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.advance_rk4(st.clone(), dt, steps)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0

        rho = rho0.clone().requires_grad_(True)
        gam, spn = f64(gamma0, True), f64(spin0, True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = ah.objective(box, rho, gam, spn, dt, steps, r0, first, every,
                         ctx, **kw)
        J.backward()
        torch.cuda.synchronize()
        t_adj = time.perf_counter() - t0

        def obj(r, g, a):
            return ah.objective_only(box, r, g, a, dt, steps, r0, first,
                                     ctx, **kw)

        rng = np.random.default_rng(16)
        dirs = [ah.smooth_field(box, rng) * rho0 for _ in range(ndir)]
        fd_out = {}
        if not hs:
            fd_out["skipped"] = True
        for h in hs:
            rels = []
            for d in dirs:
                fd = (obj(rho0 + h * d, gamma0, spin0) -
                      obj(rho0 - h * d, gamma0, spin0)) / (2 * h)
                ad = float((rho.grad * d).sum())
                rels.append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
            fdg = (obj(rho0, gamma0 * (1 + h), spin0) -
                   obj(rho0, gamma0 * (1 - h), spin0)) / (2 * h * gamma0)
            fda = (obj(rho0, gamma0, spin0 * (1 + h)) -
                   obj(rho0, gamma0, spin0 * (1 - h))) / (2 * h * spin0)
            fd_out["h=%g" % h] = {
                "rho0_rel_median": float(np.median(rels)),
                "rho0_rel_max": max(rels),
                "gamma_rel": abs(fdg - float(gam.grad)) / abs(fdg),
                "spin_rel": abs(fda - float(spn.grad)) / abs(fda)}
        print("bench: adjoint (excision %s) %.2f s, forward %.2f s; FD %s"
              % (excise, t_adj, t_fwd, json.dumps(fd_out)), file=sys.stderr,
              flush=True)
        return {"J": float(J.detach()), "dJ_dgamma": float(gam.grad),
                "dJ_dspin": float(spn.grad),
                "dJ_drho0_l2": float(rho.grad.norm()),
                "adjoint_s": t_adj, "forward_s": t_fwd,
                "adjoint_over_forward": t_adj / t_fwd,
                "fd_check": dict(directions=ndir, **fd_out)}

    rec = {"workload": "michel-like inflow %d^3 on [-16,16]^3 Kerr a=%.1f, "
                       "%d RK4 steps, J = mass flux through r = 4M over the "
                       "last %d steps, horizon excised" % (n, spin0, steps,
                                                          window),
           "checkpoint_every": every, "recovery_tolerance": 1e-16,
           "forward_model": "horizon excision on (KerrSchild::"
                            "check_if_excised, zero rhs inside) and recovery "
                            "tolerance 1e-16, not the reference's default "
                            "1e-12 without excision: with those, recoveries "
                            "that fail inside the horizon reset cells on "
                            "either side of any perturbation and J is not "
                            "differentiable (see without_excision)"}
    if args.no_adjoint_checks:
        rec.update(run(True, 0, ()))
        return rec
    rec.update(run(True, 16, (1e-6,)))
    rec["without_excision"] = run(False, 4, (1e-4, 1e-6))
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint record of the default run")
    ap.add_argument("--no-adjoint-checks", action="store_true",
                    help="adjoint record without the finite-difference "
                         "checks and the run without excision (profiling)")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 32^3 Michel (C1) record of the default run")
    ap.add_argument("--max-n", type=int, default=512,
                    help="largest box of --workload sizes")
    ap.add_argument("--workload", choices=["torus", "batched", "mixed",
                                           "adjoint", "c1", "sizes"],
                    default="torus",
                    help="torus: the 128^3 plain FP64 box (default); "
                         "batched: the 3-level binary-PN box set; mixed: "
                         "the 256^3 torus in mixed precision")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision record of the default run")
    ap.add_argument("--mixed-n", type=int, default=256)
    ap.add_argument("--mixed-steps", type=int, default=200)
    ap.add_argument("--batched-shift", type=float, default=0.05,
                    help="guarded batched mixed record: FP64 faces beside "
                         "cells whose shift has |beta| at least this")
    ap.add_argument("--mixed-shift", type=float, default=0.1,
                    help="guarded mixed record: FP64 faces beside cells "
                         "whose shift has |beta| at least this")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the binary-PN box-set record of the default "
                         "run")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="steps of the one-call-per-box comparison run")
    ap.add_argument("--autotune", action="store_true",
                    help="time every face launch (walk length, fused or "
                         "not) on the first call of each level and keep "
                         "the fastest, instead of the size rules")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 512 MiB device copy (x5), the "
                         "reference kernel for scaling FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    args.steps_given = "--steps" in sys.argv
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py: the operator runs on one GPU only; --gpus %d is "
              "not supported (the box would shard by z-slabs across GPUs, "
              "which is not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "mixed":
        print(json.dumps(bench_mixed(args)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(bench_adjoint(args)), flush=True)
        return
    if args.workload == "c1":
        print(json.dumps(bench_c1(args)), flush=True)
        return
    if args.workload == "sizes":
        print(json.dumps(bench_sizes(args)), flush=True)
        return

    import numpy as np
    import torch

    from grdzhadzha_amd import workloads
    from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator

    w = workloads.torus_box(n=args.n, nsteps=args.steps)
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options)
    op.ctx.set_autotune(args.autotune)
    init = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
    print("bench: %s, %d interior cells, dt %.4g; warmup %d, timed %d steps"
          % (w.name, w.box.num_interior, w.dt, args.warmup, args.steps),
          file=sys.stderr, flush=True)

    st = init.clone()
    if args.warmup:
        op.advance_rk4(st, w.dt, args.warmup)
    st = init.clone()
    torch.cuda.synchronize()
    op.ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    diag = op.advance_rk4(st, w.dt, args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timing = op.ctx.timing()
    op.ctx.set_timing(False)
    elapsed = t1 - t0
    cells = w.box.num_interior
    value = cells * args.steps / elapsed
    print("bench: %.3f s for %d steps; recoveries %s" % (elapsed, args.steps,
                                                        diag),
          file=sys.stderr, flush=True)
    for k, (ms, n) in timing.items():
        if n:
            print("bench:   %-8s %9.3f ms total, %5d launches, %.4f ms avg"
                  % (k, ms, n, ms / n), file=sys.stderr, flush=True)

    counts = algorithmic_counts("c2" if args.n == 128 else None)
    dom = max((k for k in timing if k in ALG_BYTES),
              key=lambda k: timing[k][0])
    ms, nl = timing[dom]
    avg_s = ms / nl / 1e3
    units = units_per_launch(dom, args.n, w.box.ghosts)
    gbs = ALG_BYTES[dom] * units / avg_s / 1e9
    hbm = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": gbs / HBM_PEAK_GBS, "alg_bytes_per_unit": ALG_BYTES[dom]}
    common = {"kernel": dom, "traffic": measured_traffic(dom, args.n),
              "avg_launch_ms": ms / nl, "units_per_launch": units,
              "ridge_flop_per_byte": RIDGE_FLOP_PER_BYTE}
    issued = measured_flops(dom, args.n)
    if issued:
        common["issued_fp64_flop_per_unit"] = issued / units
    vi = valu_issue(dom, args.n, avg_s)
    if vi:
        common["valu_issue"] = vi
    if counts and dom in KERNEL_PARTS:
        parts = counts["flop_per_cell_stage_by_part"]
        # faces: one face per direction per cell, so 3 faces per cell
        per_cell = sum(parts[p] for p in KERNEL_PARTS[dom])
        cells_per_launch = args.n ** 3
        flops = per_cell * cells_per_launch
        tf = flops / avg_s / 1e12
        fp64 = {"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / FP64_PEAK_TFLOPS,
                "alg_flop_per_cell": per_cell}
        intensity = flops / (ALG_BYTES[dom] * units)
        common["alg_intensity_flop_per_byte"] = intensity
        if intensity < RIDGE_FLOP_PER_BYTE:
            roofline = dict(bound="hbm", achieved=gbs, peak=HBM_PEAK_GBS,
                            unit="GB/s", frac=gbs / HBM_PEAK_GBS,
                            alg_bytes_per_unit=ALG_BYTES[dom], fp64=fp64,
                            **common)
        else:
            roofline = dict(bound="fp64", achieved=tf,
                            peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                            frac=tf / FP64_PEAK_TFLOPS,
                            alg_flop_per_cell=per_cell, hbm=hbm, **common)
    else:
        roofline = dict(bound="hbm", achieved=gbs, peak=HBM_PEAK_GBS,
                        unit="GB/s", frac=gbs / HBM_PEAK_GBS,
                        alg_bytes_per_unit=ALG_BYTES[dom], **common)

    step = {"alg_bytes_per_cell_update": STEP_ALG_BYTES,
            "achieved_gbs": value * STEP_ALG_BYTES / 1e9,
            "hbm_frac": value * STEP_ALG_BYTES / 1e9 / HBM_PEAK_GBS}
    if counts:
        f_alg = counts["flop_per_cell_update"]
        step.update(alg_flop_per_cell_update=f_alg,
                    achieved_tflops=value * f_alg / 1e12,
                    fp64_frac=value * f_alg / 1e12 / FP64_PEAK_TFLOPS,
                    alg_intensity_flop_per_byte=f_alg / STEP_ALG_BYTES,
                    bound=("fp64" if f_alg / STEP_ALG_BYTES >=
                           RIDGE_FLOP_PER_BYTE else "hbm"))
    sb = measured_step_bytes(args.n)
    if sb:
        step["measured_bytes_per_cell_update"] = sb / cells

    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": elapsed / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": w.name + " plain FP64 RK4 (one box, "
                      "3 ghosts, minmod, dt = 0.25 dx)",
                      "interior_cells": cells, "ghosts": w.box.ghosts,
                      "dt": w.dt, "parallelism": "single GPU"},
           "roofline": roofline,
           "face_launch": op.ctx.face_launch(),
           "step": step}
    if not args.no_batched:
        del op, st, init
        torch.cuda.empty_cache()
        args.steps_given = False
        b = bench_batched(args)
        out["batched"] = {"workload": b["config"]["workload"],
                          "boxes": b["config"]["boxes"],
                          "interior_cells": b["config"]["interior_cells"],
                          "steps": b["steps"], "value": b["value"],
                          "ms_per_step": b["ms_per_step"],
                          "plain_per_box": b["plain_per_box"],
                          "speedup_vs_plain": b["speedup_vs_plain"],
                          "face_launch_per_level":
                              b["face_launch_per_level"],
                          "batched_mixed": b["batched_mixed"],
                          "batched_guarded_mixed":
                              b["batched_guarded_mixed"]}
    if not args.no_mixed:
        torch.cuda.empty_cache()
        out["mixed"] = bench_mixed(args)
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        out["adjoint"] = bench_adjoint(args)
    if not args.no_c1:
        out["c1"] = bench_c1(args)
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle on 3 RK4 steps ...",
              file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
