"""Text summaries of rocprofv3 output for the repository's profiles folder.

    python -m kunquant_amd.profsum trace  KERNEL_TRACE_CSV  [--launches N] [--title T] [--command C]
    python -m kunquant_amd.profsum valu   COUNTER_CSV       [--title T] [--command C] [--json OUT]
    python -m kunquant_amd.profsum traffic FETCH_CSV WRITE_CSV COPY_FETCH_CSV COPY_WRITE_CSV
                                    --module M --stocks S --days T --json OUT [--command C]
    python -m kunquant_amd.profsum traffic-md TRAFFIC_JSON [--top N]

`trace` reads `rocprofv3 --kernel-trace --output-format csv` and prints a
markdown table of kernels (calls, total and average duration) and, per batch
run (the library's kernels in dispatch order, N launches a run, or kernels
separated by gaps over 2 ms), the wall time, the GPU-busy
fraction (union of kernel intervals over the wall time) and the average
number of kernels in flight.

`valu` reads a `--pmc` run with SQ_WAVES, SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU,
SQ_WAVE_CYCLES (and optionally SQ_INSTS_LDS, SQ_INSTS_SALU, SQ_BUSY_CYCLES):
per kernel, the VALU instructions per dispatch (64-lane wavefront
instructions), their rate over the dispatch's duration and that rate as a
fraction of MI355X's VALU issue peak, 256 CUs x one wavefront instruction
per clock x 2.4 GHz = 614.4 G/s (non-packed fp32: 78.6 TFLOP/s of FMA).

`traffic` reads FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch) of a
run and of a plain 1 GiB device copy, scales each counter by what it reads
for the copy (`bench.py --copy-probe`) and writes per-kernel HBM bytes per
dispatch as JSON, stamped with the commit and the sha256 of the runtime and
module libraries, which `bench.py` checks before quoting it."""
import argparse
import collections
import csv
import hashlib
import json
import os
import subprocess
import sys

VALU_PEAK = 256 * 2.4e9  # wavefront VALU instructions per second, whole GPU
GAP_NS = 2_000_000


def short_name(n: str) -> str:
    n = n.replace("void (anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    if n.startswith(("rank_rows<", "rank_segs<", "scale_")):
        return n.split(">(")[0] + ">"
    return n.split("(")[0]


def kernel_class(n: str) -> str:
    if "rank_" in n:
        return "rank"
    if "scale_" in n:
        return "scale"
    if n.startswith("kun_"):
        return "time-series stage"
    return "other"


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def trace(args):
    rows = _rows(args.csv)
    ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short_name(r["Kernel_Name"]))
                for r in rows)
    out = []
    if args.title:
        out += [f"# {args.title}", ""]
    if args.command:
        out += ["Command (from the repository root):", "", "    " + args.command, ""]
    by = collections.defaultdict(list)
    for s, e, n in ks:
        by[n].append(e - s)
    total = sum(sum(v) for v in by.values())
    out += [f"All kernels: {len(ks)} dispatches, {total / 1e6:.1f} ms summed.", "",
            "| kernel | calls | total ms | avg us | % of summed |", "|---|---|---|---|---|"]
    for n, v in sorted(by.items(), key=lambda kv: -sum(kv[1]))[:args.top]:
        out.append(f"| {n} | {len(v)} | {sum(v) / 1e6:.2f} | {sum(v) / len(v) / 1e3:.1f} | "
                   f"{100 * sum(v) / total:.1f} |")
    if args.launches or args.runs:
        # runs submitted back to back: the library's kernels in dispatch
        # order, `launches` per run
        own = sorted(((int(r["Dispatch_Id"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                       short_name(r["Kernel_Name"])) for r in rows), key=lambda x: x[0])
        own = [x[1:] for x in own if kernel_class(x[3]) != "other"]
        n = args.launches or len(own) // args.runs
        runs = [sorted(own[i:i + n]) for i in range(0, len(own) - n + 1, n)]
        how = f"runs of {n} launches in dispatch order"
    else:
        runs, cur, end = [], [], 0
        for k in ks:
            if cur and k[0] - end > GAP_NS:
                runs.append(cur)
                cur = []
            end = k[1] if not cur else max(end, k[1])
            cur.append(k)
        if cur:
            runs.append(cur)
        how = "runs separated by gaps over 2 ms"
    out += ["", f"Timeline per run ({how}): wall time from the first "
            "kernel's start to the last kernel's end, GPU-busy = union of kernel intervals / wall, "
            "in flight = summed kernel time / busy time.", "",
            "| run | kernels | wall ms | GPU-busy | idle ms | in flight | summed ms by class |",
            "|---|---|---|---|---|---|---|"]
    if len(runs) > args.max_runs:
        out.append(f"(first {args.max_runs} of {len(runs)} runs)")
    for i, run in enumerate(runs[:args.max_runs]):
        t0, t1 = run[0][0], max(e for _, e, _ in run)
        busy, hi = 0, t0
        for s, e, _ in run:
            if e > hi:
                busy += e - max(s, hi)
                hi = e
        tot = sum(e - s for s, e, _ in run)
        cls = collections.Counter()
        for s, e, n in run:
            cls[kernel_class(n)] += e - s
        out.append(f"| {i} | {len(run)} | {(t1 - t0) / 1e6:.2f} | {busy / (t1 - t0):.3f} | "
                   f"{(t1 - t0 - busy) / 1e6:.2f} | {tot / max(busy, 1):.2f} | "
                   + ", ".join(f"{k} {v / 1e6:.1f}" for k, v in cls.most_common()) + " |")
    print("\n".join(out))


def valu(args):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = collections.defaultdict(list)
    seen = set()
    for r in _rows(args.csv):
        n = short_name(r["Kernel_Name"])
        agg[n][r["Counter_Name"]].append(float(r["Counter_Value"]))
        key = (r.get("Dispatch_Id"), n)
        if key not in seen:
            seen.add(key)
            dur[n].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows = []
    for n, c in agg.items():
        avg = {k: sum(v) / len(v) for k, v in c.items()}
        d = sum(dur[n]) / len(dur[n]) if dur[n] else 0
        insts = avg.get("SQ_INSTS_VALU", 0.0)
        rate = insts / (d * 1e-9) if d else 0.0
        act = avg.get("SQ_ACTIVE_INST_VALU", 0.0)
        wc = avg.get("SQ_WAVE_CYCLES", 0.0)
        rows.append({"kernel": n, "dispatches": len(dur[n]), "avg_us": d / 1e3, "valu_insts": insts,
                     "valu_rate": rate, "valu_frac": rate / VALU_PEAK,
                     "valu_active_per_wave_cycle": act / wc if wc else None,
                     "waves": avg.get("SQ_WAVES"), "lds_insts": avg.get("SQ_INSTS_LDS"),
                     "salu_insts": avg.get("SQ_INSTS_SALU"), "total_us": d * len(dur[n]) / 1e3})
    rows.sort(key=lambda r: -r["total_us"])
    out = []
    if args.title:
        out += [f"# {args.title}", ""]
    if args.command:
        out += ["Command (from the repository root):", "", "    " + args.command, ""]
    tot_i = sum(r["valu_insts"] * r["dispatches"] for r in rows)
    tot_t = sum(r["total_us"] for r in rows)
    out += ["Counter runs serialise dispatches, so durations here are of each kernel alone. VALU peak: "
            "256 CUs x 1 wavefront instruction / clock x 2.4 GHz = 614.4 G instr/s "
            "(non-packed fp32 FMA: 78.6 TFLOP/s).", "",
            f"All kernels: {tot_i / 1e9:.2f} G VALU instructions in {tot_t / 1e3:.1f} ms of kernel time; "
            f"at peak issue they would take {tot_i / VALU_PEAK * 1e3:.1f} ms.", "",
            "| kernel | dispatches | avg us | VALU instr / dispatch | G instr/s | fraction of VALU peak | "
            "VALU-active / wave-cycle | LDS instr / dispatch |", "|---|---|---|---|---|---|---|---|"]
    for r in rows[:args.top]:
        act = "" if r["valu_active_per_wave_cycle"] is None else f"{r['valu_active_per_wave_cycle']:.3f}"
        lds = "" if r["lds_insts"] is None else f"{r['lds_insts']:.3g}"
        out.append(f"| {r['kernel']} | {r['dispatches']} | {r['avg_us']:.1f} | {r['valu_insts']:.4g} | "
                   f"{r['valu_rate'] / 1e9:.1f} | {r['valu_frac']:.3f} | {act} | {lds} |")
    print("\n".join(out))
    if args.json:
        keep = ("dispatches", "avg_us", "valu_insts", "valu_frac")
        rec = {"command": args.command, "peak_valu_insts_per_s": VALU_PEAK, "commit": args.commit or _commit(),
               "runtime_sha256": _sha(_runtime_lib()),
               "kernels": {r["kernel"]: {k: (round(r[k], 4) if isinstance(r[k], float) else r[k]) for k in keep}
                           for r in rows}}
        if args.module:
            rec["module_sha256"] = _sha(_module_lib(args.module))
        with open(args.json, "w") as f:
            json.dump(rec, f, indent=1, sort_keys=True)


def _per_dispatch(path, counter):
    acc = collections.defaultdict(list)
    for r in _rows(path):
        if r["Counter_Name"] == counter:
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def _copy_kib(path, counter):
    vals = [v for k, v in _per_dispatch(path, counter).items() if "copy" in k.lower()]
    if not vals:
        raise SystemExit(f"{path}: no copy kernel found for {counter}")
    return max(vals)


def _repo():
    return os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _commit():
    try:
        return subprocess.check_output(["git", "rev-parse", "--short=12", "HEAD"], cwd=_repo(), text=True,
                                       stderr=subprocess.DEVNULL).strip()
    except Exception:
        return None


def _sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def _runtime_lib():
    from .build import RUNTIME_LIB
    return RUNTIME_LIB


def _module_lib(name):
    from .workloads import catalog
    return catalog.library_path(name)


def traffic(args):
    gib_kib = (1 << 30) / 1024
    sf = gib_kib / _copy_kib(args.copy_fetch, "FETCH_SIZE")
    sw = gib_kib / _copy_kib(args.copy_write, "WRITE_SIZE")
    f, w = _per_dispatch(args.fetch, "FETCH_SIZE"), _per_dispatch(args.write, "WRITE_SIZE")
    kern = {k: {"fetch_bytes": round(f[k] * 1024 * sf), "write_bytes": round(w.get(k, 0.0) * 1024 * sw)} for k in f}
    sums = [k for k in kern if k.startswith("scale_sums<")]
    divs = [k for k in kern if k.startswith("scale_div<")]
    if sums and divs:  # one Scale stage = row sums + division, as bench.py groups them
        kern["scale_sums + scale_div"] = {x: kern[sums[0]][x] + kern[divs[0]][x] for x in ("fetch_bytes", "write_bytes")}
    rec = {"module": args.module, "stocks": args.stocks, "days": args.days, "commit": args.commit or _commit(),
           "module_sha256": _sha(_module_lib(args.module)), "runtime_sha256": _sha(_runtime_lib()),
           "scale_fetch": sf, "scale_write": sw, "command": args.command,
           "method": "FETCH_SIZE and WRITE_SIZE (KiB per dispatch) x 1024 x (1 GiB / the counter's reading "
                     "for a 1 GiB device copy)", "kernels": kern}
    with open(args.json, "w") as fo:
        json.dump(rec, fo, indent=1, sort_keys=True)
    print(f"{args.json}: {len(kern)} kernels; copy scale fetch {sf:.3f}, write {sw:.3f}")


def traffic_md(args):
    with open(args.json) as f:
        rec = json.load(f)
    k = rec["kernels"]
    rows = sorted(k.items(), key=lambda kv: -(kv[1]["fetch_bytes"] + kv[1]["write_bytes"]))
    out = [f"# HBM traffic per launch, module {rec['module']}, {rec['stocks']} stocks x {rec['days']} days", "",
           "Command (from the repository root):", "", "    " + (rec.get("command") or ""), "",
           f"Recorded at commit {rec.get('commit')}; runtime library sha256 {rec['runtime_sha256'][:16]}..., "
           f"module library sha256 {rec['module_sha256'][:16]}... (`bench.py` quotes a kernel's figure only "
           f"when the library it runs has that hash).", "",
           f"Calibration: FETCH_SIZE x {rec['scale_fetch']:.3f}, WRITE_SIZE x {rec['scale_write']:.3f} "
           f"(a 1 GiB device copy read back through the same counters).", "",
           "| kernel | MB read | MB written |", "|---|---|---|"]
    for n, v in rows[:args.top]:
        out.append(f"| {n} | {v['fetch_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} |")
    print("\n".join(out))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m kunquant_amd.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--launches", type=int, default=0, help="kernel launches per run (runs back to back)")
    t.add_argument("--runs", type=int, default=0, help="number of runs the library's launches divide into")
    t.add_argument("--max-runs", type=int, default=12, help="rows of the timeline table")
    v = sub.add_parser("valu")
    v.add_argument("csv")
    v.add_argument("--json")
    v.add_argument("--module")
    v.add_argument("--commit", help="commit the libraries were built from (default: git HEAD)")
    for p in (t, v):
        p.add_argument("--title", default="")
        p.add_argument("--command", default="")
        p.add_argument("--top", type=int, default=20)
    r = sub.add_parser("traffic")
    for x in ("fetch", "write", "copy_fetch", "copy_write"):
        r.add_argument(x)
    r.add_argument("--module", required=True)
    r.add_argument("--stocks", type=int, required=True)
    r.add_argument("--days", type=int, required=True)
    r.add_argument("--json", required=True)
    r.add_argument("--command", default="")
    r.add_argument("--commit", help="commit the libraries were built from (default: git HEAD)")
    tm = sub.add_parser("traffic-md")
    tm.add_argument("json")
    tm.add_argument("--top", type=int, default=20)
    a = ap.parse_args(argv)
    {"trace": trace, "valu": valu, "traffic": traffic, "traffic-md": traffic_md}[a.cmd](a)


if __name__ == "__main__":
    main(sys.argv[1:])
