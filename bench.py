#!/usr/bin/env python3
"""Benchmark of the DIM RK3-TVD hot path on one MI355X.

Workload (the reference's headline configuration that fits one GPU): two vapour
bubbles of radius 0.08 in liquid at T0 = 0.95 with mean flow u = 0.1 on a
4096 x 4096 interior (plus 2 ghost layers), periodic in x, walls in y, FP64,
RK3-TVD with dt from the device-side CFL reduction (CFL 0.1).  Test physics
Re=5, We=10, Pr=20, viscous_r=-1.5, cv_r=2.5.

One step = one full RK3-TVD step (three fused stage kernels + wall ghost
kernels) over the whole field, inputs resident in HBM.  value = interior cells
x steps / wall time (grid-point updates/s).  The same run in the mixed
FP32/FP64 mode is reported beside it with its error against the FP64 run, and
the CPU oracle is timed on a bounded sample on the host.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt updates/s + %HBM peak; FP64 vs mixed FP32 at ≤1e-6 rel err"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# MI355X VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD
# every 4 cycles (16 lanes per cycle; FP64 and non-packed FP32 alike), 2.4 GHz
VALU_PEAK = 256 * 4 * 2.4e9 / 4
# compulsory HBM bytes per interior cell and launch.  FP64: stage 1 reads u^n
# and writes u1 (64 B); stages 2 and 3 read the stage state and u^n and write
# (96 B); the fused stages 1+2 read u^n and write u2 (64 B).  Mixed: u1/u2 are
# FP32 increments -> 48, 64, 80 B, and 48 B for the fused pair.
ALG_BYTES = {"f64": {"stage1": 64, "stage2": 96, "stage3": 96, "pair": 64},
             "mixed": {"stage1": 48, "stage2": 64, "stage3": 80, "pair": 48}}
KERNEL = {"stage1": "stage_kernel<{m}, 1> ({p} stage 1)", "stage2": "stage_kernel<{m}, 2> ({p} stage 2)",
          "stage3": "stage_kernel<{m}, 3> ({p} stage 3, with the CFL reduction)",
          "pair": "stage_pair_kernel<{m}> ({p} stages 1+2 fused)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_errors(a, b):
    """max and L2 relative error of density and total energy (interior)."""
    import numpy as np
    out = {}
    for k, name in ((0, "rho"), (3, "E")):
        x, y = a[k, 2:-2, 2:-2], b[k, 2:-2, 2:-2]
        out[f"{name}_max_rel"] = float(np.abs(x - y).max() / np.abs(y).max())
        out[f"{name}_l2_rel"] = float(np.sqrt(((x - y) ** 2).sum() / (y ** 2).sum()))
    return out


def clock_warmup(field, cfl, dt=0.0, seconds=1.0, chunk=50):
    """Untimed steps until `seconds` of GPU work have run, so the timed steps
    start at the clocks a sustained run reaches (a short run from an idle GPU
    otherwise times the clock ramp).  Returns (steps, seconds)."""
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        field.run(chunk, dt=dt, cfl=cfl)
        field.synchronize()
        n += chunk
    return n, time.perf_counter() - t0


def timed_run(field, steps, warmup, cfl, dt=0.0, clock_s=1.0):
    """The clock warm-up, `warmup` untimed steps, then `steps` timed steps on
    the production path: profiling off, so dim2d_integrate replays pairs of
    steps from captured HIP graphs.  Afterwards a separate window of
    max(steps, 20) eager steps records HIP events on the context's stream
    around every kernel launch, for each kernel's mean duration (events make
    the launches eager and add a few microseconds per launch, so they stay out
    of the timed window).  Returns (elapsed s, dominant kind, its (mean ms,
    launches), per-kind means, clock warm-up (steps, s))."""
    import torch
    clock = clock_warmup(field, cfl, dt, clock_s)
    field.run(warmup, dt=dt, cfl=cfl)
    torch.cuda.synchronize()
    field.synchronize()
    t0 = time.perf_counter()
    field.run(steps, dt=dt, cfl=cfl, sync=False)
    field.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    field.profile(True)
    field.run(max(steps, 20), dt=dt, cfl=cfl)
    kt = field.kernel_times()
    field.profile(False)
    if kt["stage2"][1] == 0 and kt["stage1"][1] > 0:
        # fused stages 1 and 2: their launch is timed in the stage-1 slot
        kt = {"pair": kt["stage1"], "stage3": kt["stage3"], "wall": kt["wall"]}
    stages = [k for k in ("stage1", "stage2", "stage3", "pair") if k in kt]
    dom = max(stages, key=lambda k: kt[k][0] * kt[k][1])
    field.time = field.device_time()[0]
    return el, dom, kt[dom], kt, clock


def cpu_baseline(seconds_hint=15.0):
    """The CPU oracle (one core) on a bounded sample of the workload: the same
    two-bubble physics and boundaries at 1024^2, a few RK3 steps."""
    import numpy as np

    from dim2d_finite_volume_amd import initial_conditions as ic
    from oracle import oracle as orc
    n = 1024
    p, nodes = ic.two_bubbles(n=n)
    m = orc.make_model(p.nx, p.ny, bc=p.bc)
    xm, ym = p.coordinates()
    orc.apply_bc(m, nodes, xm, ym)
    dt = ic.cfl_dt(p, nodes, 0.1)
    # one step to size the sample, then enough steps for ~seconds_hint
    t0 = time.perf_counter()
    orc.run(m, nodes, xm, ym, np.array([dt]))
    one = time.perf_counter() - t0
    steps = max(1, min(8, int(seconds_hint / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    bad = orc.run(m, nodes, xm, ym, np.full(steps, dt))
    el = time.perf_counter() - t0
    total_steps = steps + 1
    return {"value": n * n * total_steps / (el + one), "unit": "grid-pt updates/s", "cores": 1,
            "kind": "port",
            "sample": f"CPU oracle (C restatement of the reference's serial Fortran, gcc -O2, "
                      f"one thread) on the same two-bubble case at {n}x{n}, periodic x / wall y, "
                      f"FP64, {total_steps} RK3 steps, {el + one:.1f} s"
                      + (" (diverged)" if bad else "")}


def batch_bench(steps, nsims=256, seed=4):
    """The batch workload: 256 independent drop/bubble simulations, n^2 with n
    log-uniform in [64, 512], reflection edges, each with its own fixed dt
    (CFL 0.1 of its initial state); persistent kernels under a HIP graph, RK
    stages 1 and 2 fused in one (the stage pair) and stage 3 in another, with
    tiles chosen by dim2d_batch_autotune; FP64 and mixed, each in its default
    mode (FP64 pipelined: two groups half a step apart, each launch running
    one group's stage pairs beside the other's stage 3)."""
    import dataclasses

    import numpy as np

    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    sims = ic.batch_configs(nsims=nsims, seed=seed)
    dts = np.array([ic.cfl_dt(p, n, 0.1) for p, n in sims])
    cells = sum((p.nx - 4) * (p.ny - 4) for p, _ in sims)
    out = {"nsims": nsims, "interior_cells": cells, "steps": steps,
           "sizes": [min(p.nx - 4 for p, _ in sims), max(p.nx - 4 for p, _ in sims)]}
    for prec, name in ((d.PREC_FP64, "f64"), (d.PREC_MIXED, "mixed")):
        b = d.Batch([dataclasses.replace(p, precision=prec) for p, _ in sims])
        for i, (_, nodes) in enumerate(sims):
            b.upload(i, nodes)
        mass0 = [b.download(i)[0, 2:-2, 2:-2].sum() for i in (0, nsims - 1)]
        rows, wg, tuned_ms = b.autotune(dts, 10)
        tiles = b.tiles()
        b.integrate(dts, 2)
        t0 = time.perf_counter()
        b.integrate(dts, steps)
        el = time.perf_counter() - t0
        last = [b.download(i) for i in (0, nsims - 1)]
        b.close()
        value = cells * steps / el
        unfused = sum(ALG_BYTES[name][k] for k in ("stage1", "stage2", "stage3"))
        fused = ALG_BYTES[name]["pair"] + ALG_BYTES[name]["stage3"]
        out[name] = {"value": value, "unit": "grid-pt updates/s", "ms_per_step": 1e3 * el / steps,
                     "algorithmic_bytes_per_update": fused if tiles["stage_fusion"] else unfused,
                     "hbm_frac_algorithmic": value * (fused if tiles["stage_fusion"] else unfused)
                     / HBM_PEAK,
                     "hbm_frac_unfused_bytes": value * unfused / HBM_PEAK,
                     "tiles": dict(tiles, autotune_ms_per_step=tuned_ms),
                     "finite": bool(all(np.isfinite(x).all() for x in last)),
                     "mass_drift": [float((x[0, 2:-2, 2:-2].sum() - m0) / m0)
                                    for x, m0 in zip(last, mass0)]}
        log(f"[bench] batch {name}: {value:.4e} updates/s, tiles {tiles}")
    return out


def open_boundary_bench(n, steps):
    """The 8192^2 open-edge workload on a fixed domain: a vapour bubble (r = 0.1) in a mean flow
    (0.2, 0.05) with Hedstrom open edges, mixed precision against FP64 with a
    common fixed dt (CFL 0.1 of the initial state)."""
    import dataclasses

    import numpy as np

    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    p, nodes = ic.open_bubble(n=n)
    dt = ic.cfl_dt(p, nodes, 0.1)
    res = {}
    finals = {}
    for prec, name in ((d.PREC_FP64, "f64"), (d.PREC_MIXED, "mixed")):
        f = d.Field(dataclasses.replace(p, precision=prec), nodes)
        f.run(2, dt=dt)
        f.synchronize()
        t0 = time.perf_counter()
        f.run(steps, dt=dt)
        el = time.perf_counter() - t0
        finals[name] = f.download()
        f.close()
        res[name] = {"value": n * n * steps / el, "unit": "grid-pt updates/s",
                     "ms_per_step": 1e3 * el / steps}
        log(f"[bench] open {n}^2 {name}: {res[name]['value']:.4e} updates/s")
    del nodes
    res["mixed"]["error_vs_fp64"] = rel_errors(finals["mixed"], finals["f64"])
    res["config"] = (f"vapour bubble r=0.1 in liquid, T0=0.95, (u,v)=(0.2,0.05), {n}x{n} "
                     f"interior, Hedstrom open edges, fixed dt={dt:.4e}, {steps + 2} steps")
    return res


def open_extension_bench(n, steps, every, gap, band):
    """The 8192^2 open-edge workload with the domain extension: the bubble's interface starts
    `gap` points from the east edge; every `every` steps the FP64 run calls
    adapt_domain (density criterion eps = 0.3 over a band of `band` points,
    one layer per edge and call, new points by the characteristic
    extrapolation from the far field of the initial condition).  The mixed
    run replays the same growth schedule, so both end on the same grid and
    the precision error is taken point by point.  Throughput counts the
    updates of the grid as it grows, over the whole run (checks, growths and
    reallocations included)."""
    import dataclasses

    import numpy as np

    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    radius, flow = 0.1, (0.2, 0.05)
    dx = 1.0 / (n - 1)
    center = (0.5 - radius - gap * dx, 0.0)
    p, nodes = ic.open_bubble(n=n, center=center, radius=radius, flow=flow)
    dt = ic.cfl_dt(p, nodes, 0.1)
    ff = ic.far_field(p, 0.95, flow)
    res, finals, schedule = {}, {}, []
    for prec, name in ((d.PREC_FP64, "f64"), (d.PREC_MIXED, "mixed")):
        f = d.Field(dataclasses.replace(p, precision=prec), nodes)
        f.set_far_field(ff)
        f.synchronize()
        updates, done, k = 0, 0, 0
        t0 = time.perf_counter()
        while done < steps:
            m = min(every, steps - done)
            f.run(m, dt=dt)
            updates += (f.params.nx - 4) * (f.params.ny - 4) * m
            done += m
            if prec == d.PREC_FP64:
                a = f.adapt_domain("WESN", distance=band, grow=1, md_threshold=0.3)
                schedule.append(a.tolist())
            else:
                a = schedule[k]
                if any(a):
                    f.extend(*a)
            k += 1
        f.synchronize()
        el = time.perf_counter() - t0
        finals[name] = f.download()
        grid = [f.params.nx, f.params.ny]
        f.close()
        grown = np.array(schedule).sum(axis=0).tolist()
        res[name] = {"value": updates / el, "unit": "grid-pt updates/s",
                     "ms_per_step": 1e3 * el / steps, "final_grid": grid,
                     "finite": bool(np.isfinite(finals[name]).all())}
        log(f"[bench] open extension {n}^2 {name}: {res[name]['value']:.4e} updates/s, "
            f"grid {grid}, layers added {grown}")
    del nodes
    res["layers_added_wesn"] = grown
    res["growth_calls"] = int(sum(1 for a in schedule if any(a)))
    res["mixed"]["error_vs_fp64"] = rel_errors(finals["mixed"], finals["f64"])
    res["config"] = (f"vapour bubble r=0.1 in liquid, T0=0.95, (u,v)=(0.2,0.05), {n}x{n} "
                     f"interior at the start, interface {gap} points from the east edge, "
                     f"Hedstrom open edges, adapt_domain every {every} steps (band {band}, "
                     f"eps 0.3, one layer per call, characteristic new points), fixed "
                     f"dt={dt:.4e}, {steps} steps")
    return res


def variants_bench(n, steps):
    """The other integrator and operator choices on the headline's two-bubble
    state at n^2, CFL 0.1 device-side dt, FP64: the low-storage RK3 (periodic
    x / wall y, as the headline; also in mixed precision) and the
    Cockburn-Gau stencils (periodic edges: they take reflection and periodic
    edges only; FP64 and mixed precision), with the Mattsson RK3-TVD on the
    same periodic case beside them."""
    import dataclasses

    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    p, nodes = ic.two_bubbles(n=n)
    per = (d.BC_PERIODIC,) * 4
    cases = (("lsrk3", dict(integrator=d.INT_LSRK3)),
             ("lsrk3_mixed", dict(integrator=d.INT_LSRK3, precision=d.PREC_MIXED)),
             ("rk3tvd_periodic", dict(bc=per)),
             ("cockburn_gau_periodic", dict(bc=per, operators=d.OPS_COCKBURN_GAU)),
             ("cockburn_gau_periodic_mixed", dict(bc=per, operators=d.OPS_COCKBURN_GAU,
                                                  precision=d.PREC_MIXED)))
    out = {}
    for name, kw in cases:
        f = d.Field(dataclasses.replace(p, **kw), nodes)
        f.run(5, cfl=0.1)
        f.synchronize()
        t0 = time.perf_counter()
        f.run(steps, cfl=0.1)
        el = time.perf_counter() - t0
        f.close()
        out[name] = {"value": n * n * steps / el, "unit": "grid-pt updates/s",
                     "ms_per_step": 1e3 * el / steps}
        log(f"[bench] {name} {n}^2: {out[name]['value']:.4e} updates/s")
    out["config"] = (f"two-bubble state, {n}x{n} interior, FP64 (*_mixed: mixed precision), "
                     f"CFL 0.1, {steps} timed steps")
    return out


def read_profile(name, key):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except Exception:
        return None


def kernel_table(kt, prec, cells, n):
    """per kernel kind: mean launch time, algorithmic HBM rate and its fraction
    of 8 TB/s, counted HBM traffic (FETCH_SIZE + WRITE_SIZE profile) and VALU
    issue rate against its peak (SQ_INSTS_VALU profile)"""
    out = {}
    for k, (ms, cnt) in kt.items():
        if k == "wall" or cnt == 0:
            continue
        key = f"{'stage' + k[-1] if k.startswith('stage') else 'pair'}_{prec}_two_bubbles_{n}"
        alg = ALG_BYTES["f64" if prec == "fp64" else "mixed"][k] * cells
        e = {"kernel": KERNEL[k].format(m=0 if prec == "fp64" else 1,
                                        p="FP64" if prec == "fp64" else "mixed"),
             "mean_launch_ms": ms, "launches": cnt, "algorithmic_bytes_per_launch": alg,
             "achieved_GBps": alg / (ms * 1e-3) / 1e9, "hbm_frac": alg / (ms * 1e-3) / HBM_PEAK,
             "traffic": read_profile(key + "_traffic.json", "hbm_bytes_per_launch")}
        pd = read_profile(key + "_issue.json", "per_dispatch")
        if pd and pd.get("SQ_INSTS_VALU"):
            v = pd["SQ_INSTS_VALU"]
            e["issue"] = {"valu_instructions_per_launch": v, "achieved": v / (ms * 1e-3),
                          "peak": VALU_PEAK, "unit": "wave64 VALU instructions/s",
                          "frac": v / (ms * 1e-3) / VALU_PEAK}
        out[k] = e
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096, help="interior points per axis")
    ap.add_argument("--block-x", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=500,
                    help="steps of the 256-simulation batch (0: skip)")
    ap.add_argument("--open-n", type=int, default=8192)
    ap.add_argument("--open-steps", type=int, default=2000,
                    help="steps of the open-boundary case (0: skip)")
    ap.add_argument("--ext-steps", type=int, default=2000,
                    help="steps of the open-boundary case with domain extension (0: skip)")
    ap.add_argument("--ext-every", type=int, default=50)
    ap.add_argument("--ext-gap", type=int, default=300)
    ap.add_argument("--ext-band", type=int, default=340)
    ap.add_argument("--variant-steps", type=int, default=200,
                    help="steps of the LSRK3 / Cockburn-Gau variants (0: skip)")
    ap.add_argument("--error-steps", default="100,1000",
                    help="steps after which the mixed run is compared with FP64 "
                         "(comma-separated, increasing; empty: no comparison)")
    ap.add_argument("--clock-warmup-s", type=float, default=1.0,
                    help="seconds of untimed steps before the warmup steps")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run an elementwise kernel reading and writing 1 GiB "
                         "(for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic

    torch.cuda.init()
    dev = torch.cuda.get_device_name(0)
    if args.calibrate_copy:
        # a kernel of known traffic for HBM counter scaling: an elementwise
        # product that reads 1 GiB and writes 1 GiB per launch with 16-byte
        # vector accesses, like the stage kernels
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    n = args.n
    cells = n * n
    log(f"[bench] device {dev}; building the {n}^2 two-bubble state")
    p, nodes = ic.two_bubbles(n=n)

    # ---- FP64, CFL time step (the headline) ----
    f = d.Field(p, nodes)
    if args.block_x or args.rows:
        f.set_tiles(args.block_x, args.rows)
    f.apply_bc_on_nodes()
    diag0 = f.diagnostics()
    log(f"[bench] FP64: {args.clock_warmup_s} s clock warm-up, {args.warmup} warmup + "
        f"{args.steps} timed steps")
    el, dom, dom_kt, kt, clock = timed_run(f, args.steps, args.warmup, cfl=0.1,
                                           clock_s=args.clock_warmup_s)
    diag1 = f.diagnostics()
    t_end, dt_last = f.device_time()
    value = cells * args.steps / el
    ms_step = 1e3 * el / args.steps
    log(f"[bench] FP64 {value:.4e} updates/s, {ms_step:.4f} ms/step, dominant {dom} "
        f"{dom_kt}, profiled-window kernel means {kt}")
    kernels = kernel_table(kt, "fp64", cells, n)
    dk = kernels[dom]
    roof = {"bound": "hbm", "kernel": dk["kernel"], "achieved": dk["achieved_GBps"],
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": dk["hbm_frac"],
            "traffic": dk["traffic"], "algorithmic_bytes_per_launch": dk["algorithmic_bytes_per_launch"],
            "mean_launch_ms": dom_kt[0], "profiled_launches": dom_kt[1],
            "timing": "HIP events on the context's stream around every launch of a "
                      "window of eager steps run right after the timed (graph-replay) steps"}
    if "issue" in dk:
        roof["issue"] = dk["issue"]
    step_bytes = sum(ALG_BYTES["f64"][k] for k in kt if k != "wall")
    f.close()
    del f

    # ---- mixed FP32/FP64 ----
    mixed = None
    if not args.no_mixed:
        pm, _ = ic.two_bubbles(n=n, precision=d.PREC_MIXED)
        fm = d.Field(pm, nodes)
        if args.block_x or args.rows:
            fm.set_tiles(args.block_x, args.rows)
        fm.apply_bc_on_nodes()
        log(f"[bench] mixed: {args.warmup} warmup + {args.steps} timed steps")
        elm, domm, domm_kt, ktm, clockm = timed_run(fm, args.steps, args.warmup, cfl=0.1,
                                                    clock_s=args.clock_warmup_s)
        vm = cells * args.steps / elm
        fm.close()
        del fm
        # the mixed-precision criterion: error against FP64 with a common fixed dt (CFL 0.1 of the
        # initial state) at steps 100 and 1000, whatever --steps is
        dt0 = ic.cfl_dt(p, nodes, 0.1)
        marks = tuple(int(x) for x in args.error_steps.split(",") if x)
        outs = {}
        for prec, pp in (("f64", p), ("mixed", pm)):
            a = d.Field(pp, nodes)
            a.apply_bc_on_nodes()
            done = 0
            for nst in marks:
                a.run(nst - done, dt=dt0)
                done = nst
                outs[(prec, nst)] = (a.download(), a.diagnostics())
            a.close()
        errs = {}
        for nst in marks:
            (ra, da), (rb, db) = outs[("f64", nst)], outs[("mixed", nst)]
            e = rel_errors(rb, ra)
            e["mass_drift_mixed"] = float((db[0] - da[0]) / da[0])
            errs[f"step_{nst}"] = e
            log(f"[bench] mixed vs FP64 after {nst} steps: {e}")
        del outs
        mixed = {"value": vm, "unit": "grid-pt updates/s", "ms_per_step": 1e3 * elm / args.steps,
                 "speedup_vs_fp64": vm / value,
                 "kernel_ms": {k: ktm[k][0] for k in ktm},
                 "clock_warmup": {"steps": clockm[0], "seconds": clockm[1]},
                 "dominant_kernel_frac": None,
                 "error_vs_fp64": errs,
                 "error_config": f"same initial state, fixed dt={dt0:.6e} (CFL 0.1 of the "
                                 f"initial state) in both precisions"}
        mk = kernel_table(ktm, "mixed", cells, n)
        mixed["kernels"] = mk
        mixed["dominant_kernel"] = mk[domm]["kernel"]
        mixed["dominant_kernel_frac"] = mk[domm]["hbm_frac"]
        step_bytes_m = sum(ALG_BYTES["mixed"][k] for k in ktm if k != "wall")
        mixed["algorithmic_bytes_per_update"] = step_bytes_m
        mixed["hbm_frac_algorithmic"] = vm * step_bytes_m / HBM_PEAK
        mixed["hbm_frac_unfused_bytes"] = vm * 192 / HBM_PEAK

    batch = batch_bench(args.batch_steps) if args.batch_steps > 0 else None
    open_bc = open_boundary_bench(args.open_n, args.open_steps) if args.open_steps > 0 else None
    open_ext = (open_extension_bench(args.open_n, args.ext_steps, args.ext_every, args.ext_gap,
                                     args.ext_band) if args.ext_steps > 0 else None)
    variants = variants_bench(n, args.variant_steps) if args.variant_steps > 0 else None
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": METRIC, "value": value, "unit": "grid-pt updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"two vapour bubbles r=0.08 in liquid, T0=0.95, u=0.1, "
                               f"{n}x{n} interior + 2 ghost layers, periodic x / wall y, "
                               f"RK3-TVD, CFL 0.1 device-side dt, FP64",
                   "grid": [n + 4, n + 4], "physics": "Re=5 We=10 Pr=20 viscous_r=-1.5 cv_r=2.5",
                   "device": dev},
        "algorithmic_bytes_per_update": step_bytes,
        "hbm_frac_algorithmic": value * step_bytes / HBM_PEAK,
        "hbm_frac_unfused_bytes": value * 256 / HBM_PEAK,
        "kernel_ms": {k: kt[k][0] for k in kt},
        "kernels": kernels,
        "clock_warmup": {"steps": clock[0], "seconds": clock[1],
                         "note": "untimed steps before the warmup steps, until the GPU "
                                 "clocks reach their sustained level"},
        "timed_path": "graph replay (dim2d_integrate, profiling off)",
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "batch": batch,
        "open_boundary": open_bc,
        "open_extension": open_ext,
        "variants": variants,
        "checks": {"mass_drift": float((diag1[0] - diag0[0]) / diag0[0]),
                   "nonfinite": int(diag1[6]), "t_end": t_end, "dt_last": dt_last},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
