"""Summaries of rocprofv3 CSV output for the stage kernels.

Three reports, each written as a short Markdown file (and, for traffic and
issue counts, one JSON file per stage kernel that bench.py reads for the
roofline's `traffic` and `issue` entries):

  stats:    per-kernel call count and durations from `--kernel-trace --stats`
            (<prefix>_kernel_stats.csv) and, with --trace, the same per kernel
            and launch grid from <prefix>_kernel_trace.csv, so that one kernel
            run on two grid sizes reads as two rows
  traffic:  HBM bytes per dispatch from two `--pmc` runs (FETCH_SIZE, then
            WRITE_SIZE, KiB per dispatch), each scaled by what the same counter
            reads for a calibration kernel of known size in the same run
  issue:    instructions per dispatch (SQ_INSTS_VALU, SQ_INSTS_LDS, ...) and
            wave count from one `--pmc` run, summed over the device
  gaps:     the time of a step split into kernels and the idle gaps between
            them, from a kernel trace (<prefix>_kernel_trace.csv), for steps
            [skip, skip + steps) of each precision (a step starts at its
            stage-1 or stage-pair launch)

    python -m dim2d_finite_volume_amd.rocprof_summary stats STATS.csv OUT.md [--trace TRACE.csv] --command "..."
    python -m dim2d_finite_volume_amd.rocprof_summary traffic FETCH.csv WRITE.csv OUTDIR \
        --calib SUBSTRING:READ_BYTES:WRITE_BYTES --workload two_bubbles_4096 --command "..."
    python -m dim2d_finite_volume_amd.rocprof_summary issue COUNTERS.csv OUTDIR \
        --workload two_bubbles_4096 --command "..."
    python -m dim2d_finite_volume_amd.rocprof_summary gaps TRACE.csv OUT.md --skip 2 \
        --steps 20 --command "..."
"""
import argparse
import csv
import json
import os
import re
from collections import defaultdict

STAGE_RE = re.compile(r"dim2d::stage_kernel<(\d+),\s*(\d+)>")
PAIR_RE = re.compile(r"dim2d::stage_pair_kernel<(\d+)(?:,\s*(true|false))?(?:,\s*(true|false))?>")
BSTAGE_RE = re.compile(r"dim2d::stage_batch_kernel<(\d+),\s*(\d+)>")
BPAIR_RE = re.compile(r"dim2d::stage_pair_batch_kernel<(\d+)>")
CG_RE = re.compile(r"dim2d::cg_stage_kernel<(\d+),\s*(\d+)>")
BMIX_RE = re.compile(r"dim2d::stage_mix_batch_kernel<(\d+)>")


def kernel_key(name):
    """file-name key of a stage kernel: [batch_]stage<s>_<fp64|mixed> or
    [batch_]pair_<fp64|mixed>[_open] or batch_mix_<fp64|mixed>"""
    m = BMIX_RE.search(name)
    if m:
        return f"batch_mix_{'mixed' if m.group(1) == '1' else 'fp64'}"
    m = BSTAGE_RE.search(name)
    if m:
        return f"batch_stage{m.group(2)}_{'mixed' if m.group(1) == '1' else 'fp64'}"
    m = BPAIR_RE.search(name)
    if m:
        return f"batch_pair_{'mixed' if m.group(1) == '1' else 'fp64'}"
    m = STAGE_RE.search(name)
    if m:
        return f"stage{m.group(2)}_{'mixed' if m.group(1) == '1' else 'fp64'}"
    m = CG_RE.search(name)
    if m:
        return f"cg_stage{m.group(2)}_{'mixed' if m.group(1) == '1' else 'fp64'}"
    m = PAIR_RE.search(name)
    if m:
        return (f"pair_{'mixed' if m.group(1) == '1' else 'fp64'}"
                f"{'_open' if m.group(2) == 'true' else ''}{'_lsrk3' if m.group(3) == 'true' else ''}")
    return None


def short_name(name):
    m = CG_RE.search(name)
    if m:
        prec = "mixed" if m.group(1) == "1" else "FP64"
        return f"cg_stage_kernel<{m.group(1)}, {m.group(2)}> (Cockburn-Gau, {prec} stage {m.group(2)})"
    m = BMIX_RE.search(name)
    if m:
        prec = "mixed" if m.group(1) == "1" else "FP64"
        return (f"stage_mix_batch_kernel<{m.group(1)}> (pipelined batch, {prec} stages 1+2 "
                f"of one group beside stage 3 of the other)")
    m = BPAIR_RE.search(name)
    if m:
        prec = "mixed" if m.group(1) == "1" else "FP64"
        return f"stage_pair_batch_kernel<{m.group(1)}> (batch, {prec} stages 1+2)"
    m = BSTAGE_RE.search(name)
    if m:
        prec = "mixed" if m.group(1) == "1" else "FP64"
        return f"stage_batch_kernel<{m.group(1)}, {m.group(2)}> (batch, {prec} stage {m.group(2)})"
    m = PAIR_RE.search(name)
    if m:
        prec = "mixed" if m.group(1) == "1" else "FP64"
        if m.group(2) == "true":
            return f"stage_pair_kernel<{m.group(1)}, true> ({prec} stages 1+2, open edges)"
        if m.group(3) == "true":
            return (f"stage_pair_kernel<{m.group(1)}, false, true> "
                    f"({prec} low-storage RK3 stages 1+2)")
        return f"stage_pair_kernel<{m.group(1)}> ({prec} stages 1+2)"
    m = STAGE_RE.search(name)
    if m:
        prec = "mixed" if m.group(1) == "1" else "FP64"
        st = int(m.group(2))
        what = ("time derivative" if st == 0 else
                f"stage {st}" if st <= 3 else f"low-storage RK3 stage {st - 3}")
        return f"stage_kernel<{m.group(1)}, {m.group(2)}> ({prec} {what})"
    name = name.replace("void ", "")
    return name.split("(")[0][:90]


def stats(args):
    rows = []
    with open(args.csv) as f:
        for r in csv.DictReader(f):
            rows.append(r)
    lines = ["# Kernel durations (rocprofv3 --kernel-trace --stats)", "",
             f"Command: `{args.command}`", "",
             "| kernel | calls | mean us | min us | max us | % of GPU time |",
             "|---|---:|---:|---:|---:|---:|"]
    for r in rows:
        if args.all or "dim2d::" in r["Name"]:
            lines.append(f"| {short_name(r['Name'])} | {r['Calls']} | "
                         f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                         f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    if args.trace:
        lines += ["", "Per launch grid (from the kernel trace):", "",
                  "| kernel | grid (work-items x, y) | workgroup | VGPRs | calls | mean us | min us | max us |",
                  "|---|---|---|---:|---:|---:|---:|---:|"]
        groups = defaultdict(list)
        meta = {}
        with open(args.trace) as f:
            for r in csv.DictReader(f):
                if not (args.all or "dim2d::" in r["Kernel_Name"]):
                    continue
                key = (short_name(r["Kernel_Name"]), int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]))
                groups[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
                meta[key] = (f"{r['Workgroup_Size_X']}x{r['Workgroup_Size_Y']}",
                             int(r["VGPR_Count"]) + int(r.get("Accum_VGPR_Count", 0) or 0))
        grand = sum(sum(d) for d in groups.values()) or 1.0
        keys = sorted(groups, key=lambda k: -sum(groups[k]))
        small = [k for k in keys if sum(groups[k]) < 0.001 * grand]
        for key in keys:
            if key in small:
                continue
            d = groups[key]
            lines.append(f"| {key[0]} | {key[1]} x {key[2]} | {meta[key][0]} | {meta[key][1]} | "
                         f"{len(d)} | {sum(d) / len(d):.1f} | {min(d):.1f} | {max(d):.1f} |")
        if small:
            lines += ["", f"{len(small)} smaller launch groups (each under 0.1% of the time above, "
                          f"{sum(sum(groups[k]) for k in small) / grand * 100:.2f}% together) "
                          f"are left out."]
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def per_dispatch(path, counter):
    """mean counter value (KiB) per dispatch, per kernel name"""
    acc = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def traffic(args):
    fetch = per_dispatch(args.fetch, "FETCH_SIZE")
    write = per_dispatch(args.write, "WRITE_SIZE")
    sub, rb, wb = args.calib.split(":")
    rb, wb = float(rb), float(wb)

    def calib(d, nbytes):
        cands = [(k, v) for k, v in d.items() if sub in k]
        if not cands or nbytes <= 0:
            return None, None
        k, (kib, _) = max(cands, key=lambda kv: kv[1][0])
        return nbytes / (kib * 1024.0), kib

    fr, fk = calib(fetch, rb)
    fw, wk = calib(write, wb)
    lines = ["# HBM traffic per dispatch (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", "",
             f"Command: `{args.command}`", "",
             f"Calibration kernel: name contains `{sub}`, reads {rb:.0f} B and writes {wb:.0f} B "
             f"per dispatch; the counters read {fk} KiB and {wk} KiB, so FETCH_SIZE is scaled by "
             f"{fr:.4f} and WRITE_SIZE by {fw:.4f}.", "",
             "| kernel | dispatches | FETCH KiB | WRITE KiB | scaled HBM bytes / dispatch |",
             "|---|---:|---:|---:|---:|"]
    os.makedirs(args.outdir, exist_ok=True)
    for name in sorted(set(fetch) | set(write)):
        if "dim2d::" not in name:
            continue
        f_kib, n = fetch.get(name, (0.0, 0))
        w_kib, _ = write.get(name, (0.0, 0))
        total = f_kib * 1024.0 * (fr or 1.0) + w_kib * 1024.0 * (fw or 1.0)
        lines.append(f"| {short_name(name)} | {n} | {f_kib:.1f} | {w_kib:.1f} | {total:.4e} |")
        key = kernel_key(name)
        if key:
            js = {"kernel": short_name(name), "workload": args.workload,
                  "fetch_kib_per_dispatch": f_kib, "write_kib_per_dispatch": w_kib,
                  "fetch_scale": fr, "write_scale": fw, "hbm_bytes_per_launch": total,
                  "command": args.command}
            with open(os.path.join(args.outdir, f"{key}_{args.workload}_traffic.json"), "w") as f:
                json.dump(js, f, indent=1)
    with open(os.path.join(args.outdir, f"stage_kernels_{args.workload}_traffic.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def issue(args):
    """instructions per dispatch of the stage kernels (device totals)"""
    acc = defaultdict(lambda: defaultdict(list))
    with open(args.csv) as f:
        for r in csv.DictReader(f):
            acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted(k for k in acc if kernel_key(k))
    counters = sorted({c for k in names for c in acc[k]})
    lines = ["# Instructions per dispatch (rocprofv3 --pmc, device totals)", "",
             f"Command: `{args.command}`", "",
             "| kernel | dispatches | " + " | ".join(counters) + " |",
             "|---|---:|" + "---:|" * len(counters)]
    os.makedirs(args.outdir, exist_ok=True)
    for name in names:
        d = acc[name]
        n = max(len(v) for v in d.values())
        mean = {c: sum(v) / len(v) for c, v in d.items()}
        lines.append(f"| {short_name(name)} | {n} | " +
                     " | ".join(f"{mean.get(c, 0):.4g}" for c in counters) + " |")
        js = {"kernel": short_name(name), "workload": args.workload, "dispatches": n,
              "per_dispatch": mean, "command": args.command}
        with open(os.path.join(args.outdir, f"{kernel_key(name)}_{args.workload}_issue.json"),
                  "w") as f:
            json.dump(js, f, indent=1)
    with open(os.path.join(args.outdir, f"stage_kernels_{args.workload}_issue.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def step_kind(name):
    """(precision, role) of a kernel of the stepping, or None"""
    m = PAIR_RE.search(name)
    if m:
        return int(m.group(1)), "pair"
    m = STAGE_RE.search(name)
    if m and m.group(2) in ("1", "2", "3"):
        return int(m.group(1)), f"stage{m.group(2)}"
    if "dim2d::wall_kernel" in name:
        return None, "wall"
    if "dim2d::frame_kernel" in name:
        return None, "frame"
    return None


def gaps(args):
    rows = []
    with open(args.trace) as f:
        for r in csv.DictReader(f):
            k = step_kind(r["Kernel_Name"])
            if k is not None:
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), k))
    rows.sort()
    lines = ["# Step time: kernels and idle gaps (rocprofv3 --kernel-trace)", "",
             f"Command: `{args.command}`", "",
             f"Steps {args.skip} to {args.skip + args.steps - 1} of each precision; a step runs "
             "from its first launch (stage pair or stage 1) to the next step's.  Busy: the "
             "kernels' own durations (overlapping launches counted once); idle: the rest.", ""]
    for prec, label in ((0, "FP64"), (1, "mixed")):
        # the wall and frame kernels carry no precision: they belong to the
        # precision of the stage launch before them
        seq, cur = [], None
        for r in rows:
            if r[2][0] is not None:
                cur = r[2][0]
            if cur == prec:
                seq.append(r)
        starts = [i for i, r in enumerate(seq) if r[2][1] in ("pair", "stage1")]
        sel = starts[args.skip:args.skip + args.steps + 1]
        if len(sel) < 2:
            continue
        periods, busy, by_role, trans = [], [], defaultdict(float), defaultdict(list)
        for a, b in zip(sel, sel[1:]):
            t0, t1 = seq[a][0], seq[b][0]
            periods.append((t1 - t0) / 1e3)
            spans = sorted((r[0], min(r[1], t1)) for r in seq[a:b])
            tot, end = 0, t0
            for s0, s1 in spans:
                if s1 > end:
                    tot += s1 - max(s0, end)
                    end = s1
            busy.append(tot / 1e3)
            for r in seq[a:b]:
                by_role[r[2][1]] += (r[1] - r[0]) / 1e3
            for x, y in zip(seq[a:b + 1], seq[a + 1:b + 1]):
                trans[(x[2][1], y[2][1])].append(max(0, y[0] - x[1]) / 1e3)
        n = len(periods)
        per = sorted(periods)[n // 2]
        mean_p, mean_b = sum(periods) / n, sum(busy) / n
        lines += [f"## {label}: {n} steps", "",
                  f"- step period: median {per:.1f} us, mean {mean_p:.1f} us",
                  f"- kernels busy: {mean_b:.1f} us per step ({100 * mean_b / mean_p:.1f} %); "
                  f"idle {mean_p - mean_b:.1f} us ({100 * (1 - mean_b / mean_p):.1f} %)", "",
                  "| kernel | us per step | % of step |", "|---|---:|---:|"]
        for role, t in sorted(by_role.items(), key=lambda kv: -kv[1]):
            lines.append(f"| {role} | {t / n:.1f} | {100 * t / n / mean_p:.1f} |")
        lines += ["", "| gap after -> before | count per step | mean us | total us per step |",
                  "|---|---:|---:|---:|"]
        for (x, y), v in sorted(trans.items(), key=lambda kv: -sum(kv[1])):
            lines.append(f"| {x} -> {y} | {len(v) / n:.1f} | {sum(v) / len(v):.2f} | "
                         f"{sum(v) / n:.2f} |")
        lines.append("")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    sp = ap.add_subparsers(dest="mode", required=True)
    a = sp.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("out")
    a.add_argument("--command", required=True)
    a.add_argument("--trace", help="the run's kernel_trace.csv: adds a table per launch grid")
    a.add_argument("--all", action="store_true", help="list every kernel, not only dim2d's")
    b = sp.add_parser("traffic")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("outdir")
    b.add_argument("--calib", required=True)
    b.add_argument("--workload", required=True)
    b.add_argument("--command", required=True)
    c = sp.add_parser("issue")
    c.add_argument("csv")
    c.add_argument("outdir")
    c.add_argument("--workload", required=True)
    c.add_argument("--command", required=True)
    g = sp.add_parser("gaps")
    g.add_argument("trace")
    g.add_argument("out")
    g.add_argument("--skip", type=int, default=2)
    g.add_argument("--steps", type=int, default=20)
    g.add_argument("--command", required=True)
    args = ap.parse_args()
    {"stats": stats, "traffic": traffic, "issue": issue, "gaps": gaps}[args.mode](args)


if __name__ == "__main__":
    main()
