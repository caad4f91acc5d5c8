"""Benchmark of the factor pipeline on one MI355X.

Workload (default): the seeded 100-factor random graph over synthetic OHLCV,
8192 stocks x 1024 days, fp32, 10% NaN input cells, TS layout. A step is one
batch run of the whole graph (every stage kernel) with inputs and outputs
already resident in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--module cfg2_f32|cfg2_f64]

Batched uneven jobs (--workload jobs): 64 independent jobs (stocks uniform
in 50..6000, days in 250..2500, graphs of 20..60 factors), each job's
chunking picked by the autotuner, enqueued together in one call; compared
with the same jobs as sequential plain runs (device-resident, and host
buffers through runGraph).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 256 * 2.4e9  # MI355X: 256 CUs x one wavefront VALU instruction per clock
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


GRAPH_DESC = {"cfg2_f32": "seeded random graph", "cfg2_f64": "seeded random graph",
              "alpha158_f64": "Alpha158 library", "prec_f32": "precision graph", "prec_f64": "precision graph"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--module", default="cfg2_f32")
    ap.add_argument("--stocks", type=int, default=8192)
    ap.add_argument("--days", type=int, default=1024)
    ap.add_argument("--cpu-rows", type=int, default=256, help="rows of the CPU baseline sample")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer (PCIe-inclusive) runGraph timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("factors", "jobs", "serve", "precision"), default="factors")
    ap.add_argument("--sessions", type=int, default=256, help="--workload serve: concurrent sessions")
    ap.add_argument("--rate", type=float, default=20000.0, help="--workload serve: Poisson ticks/s, all sessions")
    ap.add_argument("--rate-sweep", default="5000,10000,15000",
                    help="--workload serve: further offered rates, reported beside --rate")
    ap.add_argument("--seconds", type=float, default=3.0, help="--workload serve: length of the arrival trace")
    ap.add_argument("--state-budget-gb", type=float, default=64.0,
                    help="--workload serve: HBM for session rows and state (admission limit)")
    ap.add_argument("--inflight", type=int, default=4, help="--workload serve: cycles in flight")
    ap.add_argument("--max-computing", type=int, default=3,
                    help="--workload serve: cycles in flight that may still be computing when another starts")
    ap.add_argument("--readback", choices=("all", "none"), default="all",
                    help="--workload serve: copy every output row of each cycle to host memory inside the cycle")
    ap.add_argument("--jobs", type=int, default=64, help="jobs of --workload jobs")
    ap.add_argument("--no-host-baseline", action="store_true",
                    help="--workload jobs: skip the sequential host-buffer runGraph pass")
    ap.add_argument("--no-graph", action="store_true", help="--workload jobs: skip the batch as one HIP graph")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload only: skip the other workloads' lines embedded under other_workloads")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device-to-device three times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(1 << 28, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 << 28, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the factor pipeline "
              f"shards by time or by job across GPUs, which this build does not run.", file=sys.stderr)
        sys.exit(2)
    extra = None
    if args.workload == "factors" and args.module == "cfg2_f32" and not args.no_extra:
        # the other workloads, each in a child process started before this
        # one touches the GPU; their lines are embedded in this one
        extra = other_workloads()
    if args.workload == "jobs":
        run_jobs(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "precision":
        run_precision(args)
        return

    import numpy as np
    import torch
    from kunquant_amd.build import RUNTIME_LIB
    from kunquant_amd.runner import KunRunner as kr
    from kunquant_amd.workloads import catalog, data

    lib, mod = catalog.load(args.module)
    S, T = args.stocks, args.days
    np_dt = mod.dtype
    t_dt = torch.float32 if np_dt == np.float32 else torch.float64
    esz = 4 if np_dt == np.float32 else 8
    d = data.ohlcv(S, T, seed=2, dtype=np_dt, nan_frac=0.1)
    inputs = mod.getInputNames()
    outputs = mod.getOutputNames()
    dev = {k: torch.from_numpy(d[k]).to("cuda:0") for k in inputs}
    outs = {k: torch.empty((T, S), dtype=t_dt, device="cuda:0") for k in outputs}
    ptrs = {k: v.data_ptr() for k, v in list(dev.items()) + list(outs.items())}
    ex = kr.createSingleThreadExecutor()
    torch.cuda.synchronize()

    # runs fork from and join into torch's current stream, so torch events
    # between them time each step without adding any synchronisation
    st = torch.cuda.current_stream()
    for _ in range(args.warmup):
        kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T, stream=st.cuda_stream)
    ex.synchronize()
    torch.cuda.synchronize()
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    marks[0].record(st)
    for k in range(args.steps):
        kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T, stream=st.cuda_stream)
        marks[k + 1].record(st)
    ex.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    step_ms = sorted(marks[k].elapsed_time(marks[k + 1]) for k in range(args.steps))
    cells = len(outputs) * S * T
    value = cells / (ms / 1e3)

    # per-kernel time (HIP events on the executor stream the kernels run on)
    ex.enable_profiling(True)
    prof_steps = max(1, min(3, args.steps))
    for _ in range(prof_steps):
        kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T)
    ex.synchronize()
    stages = mod.stages()
    bufs = mod.buffers

    def buf_bytes(b):
        return S * esz if bufs[b][0].startswith("__init") else S * T * esz

    per_kernel = []  # (avg ms, stage, kernel, kind, algorithmic bytes)
    stage_ms = {}
    for i, (name, kind, ins, outs_) in enumerate(stages):
        kinfo = mod.kernels(i)
        for j in range(max(1, len(kinfo))):
            tot, n = ex.kernel_time(mod, i, j)
            avg = tot / max(n, 1)
            if kind == 0:
                # each distinct HBM value the kernel reads or writes, once per (stock, row)
                nbytes = kinfo[j][4] * S * T * esz
                label = ("chunked" if kinfo[j][0] == 0 else "sequential") + " time-series"
            else:
                # cross-sectional: the input row read once, the output row written once
                nbytes = sum(buf_bytes(b) for b in ins) + sum(buf_bytes(b) for b in outs_)
                label = {1: "rank", 2: "scale", 3: "generic cross-sectional"}[kind]
            per_kernel.append((avg, i, j, label, nbytes))
            stage_ms[i] = stage_ms.get(i, 0.0) + avg
    ex.enable_profiling(False)
    per_kernel.sort(reverse=True)
    kernel_sum = sum(p[0] for p in per_kernel)
    # The dominant kernel: the one kernel function with the most time summed
    # over its launches. Every Rank stage of one row length launches the
    # same radix-sort kernel; every Scale stage the same row-sum + division
    # pair (timed together); each time-series stage kernel is a function of
    # its own.
    groups = {}
    for avg, i, j, label, nbytes in per_kernel:
        key = label if label in ("rank", "scale") else f"kun_{args.module}_s{i}_k{j}"
        g = groups.setdefault(key, {"ms": 0.0, "launches": 0, "bytes": 0, "label": label, "stage": i, "kern": j})
        g["ms"] += avg
        g["launches"] += 1
        g["bytes"] += nbytes
    dom_key = max(groups, key=lambda k: groups[k]["ms"])
    dom = groups[dom_key]
    dom_ms = dom["ms"] / dom["launches"]
    dom_bytes = dom["bytes"] / dom["launches"]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    dom_name = {"rank": rank_kernel_name(np_dt, S), "scale": "scale_sums + scale_div"}.get(dom_key, dom_key)
    # the longest single launch, for comparison
    lo_ms, lo_i, lo_j, lo_label, lo_bytes = per_kernel[0]
    # HBM traffic of the dominant kernel: rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes scaled by the same counters on a plain copy, recorded with the
    # sha256 of the library the kernel lives in; taken only when that library
    # is the one benched
    traffic, traffic_note = None, "no traffic record for this kernel"
    tpath = os.path.join(ROOT, "profiles", f"{args.module}_{S}x{T}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f)
        tk = rec.get("kernels", {}).get(dom_name)
        lib_path = RUNTIME_LIB if dom_key in ("rank", "scale") else catalog.library_path(args.module)
        want = rec.get("runtime_sha256" if dom_key in ("rank", "scale") else "module_sha256")
        if tk and want == sha256_file(lib_path):
            traffic = tk["fetch_bytes"] + tk["write_bytes"]
            traffic_note = f"rocprofv3 counters recorded at commit {rec.get('commit')}"
        elif tk:
            traffic_note = "traffic record is of another build of this kernel's library"
    # VALU instructions of the dominant kernel (rocprofv3 SQ counters,
    # recorded with the sha256 of the library the kernel lives in)
    valu_rec, valu_note = None, "no VALU counter record for this kernel"
    vpath = os.path.join(ROOT, "profiles", f"{args.module}_{S}x{T}_valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            rec = json.load(f)
        vk = rec.get("kernels", {}).get(dom_name)
        lib_path = RUNTIME_LIB if dom_key in ("rank", "scale") else catalog.library_path(args.module)
        want = rec.get("runtime_sha256" if dom_key in ("rank", "scale") else "module_sha256")
        if vk and want == sha256_file(lib_path):
            rate = vk["valu_insts"] / (dom_ms / 1e3)
            valu_rec = {"insts_per_launch": vk["valu_insts"], "alone_us": vk["avg_us"],
                        "alone_frac_of_valu_peak": vk["valu_frac"],
                        "in_graph_rate": rate, "in_graph_frac_of_valu_peak": rate / VALU_PEAK,
                        "peak": VALU_PEAK, "unit": "wavefront VALU instructions/s",
                        "source": f"rocprofv3 SQ_INSTS_VALU recorded at commit {rec.get('commit')}"}
            valu_note = None
        elif vk:
            valu_note = "VALU record is of another build of this kernel's library"
    timeline = dag_timeline(kr, mod, ptrs, S, T, ms)
    graph_bytes = (len(inputs) + len(outputs)) * S * T * esz
    graph_gbs = graph_bytes / (ms / 1e3) / 1e9
    by_kind = {}
    for p in per_kernel:
        by_kind[p[3]] = by_kind.get(p[3], 0.0) + p[0]

    result = {
        "metric": "factor-cells/s",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "step_ms": {"median": step_ms[len(step_ms) // 2], "min": step_ms[0], "max": step_ms[-1],
                    "note": "per step, torch events between the timed runs"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if np_dt == np.float32 else "f64",
        "data": "synthetic (seeded OHLCV, 10% NaN cells)",
        "config": {"workload": f"{len(outputs)}-factor {GRAPH_DESC.get(args.module, 'graph')}, {S} stocks x {T} days, "
                               f"{'fp32' if esz == 4 else 'fp64'}, batch runGraph with HBM-resident inputs",
                   "module": args.module, "stocks": S, "days": T, "factors": len(outputs), "inputs": len(inputs),
                   "stages": len(stages)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": f"{dom_name} ({dom['label']})", "launches_per_step": dom["launches"],
                     "kernel_ms": dom_ms, "total_ms_per_step": dom["ms"], "algorithmic_bytes": dom_bytes,
                     "valu": valu_rec if valu_rec else valu_note},
        "timeline": timeline,
        "longest_launch": {"kernel": f"stage {lo_i} kernel {lo_j} ({lo_label})", "kernel_ms": lo_ms,
                           "algorithmic_bytes": lo_bytes,
                           "achieved_GBs": lo_bytes / (lo_ms / 1e3) / 1e9,
                           "frac": lo_bytes / (lo_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
        "graph": {"algorithmic_bytes": graph_bytes, "achieved_GBs": graph_gbs,
                  "frac_of_hbm_peak": graph_gbs / HBM_PEAK_GBS, "sum_kernel_ms": kernel_sum,
                  "kernel_ms_by_kind": {k: round(v, 3) for k, v in by_kind.items()},
                  "launches_per_step": len(per_kernel),
                  "top_kernels_ms": [[p[1], p[2], p[3], round(p[0], 4)] for p in per_kernel[:8]]},
    }

    # end to end: the reference's API on host buffers (inputs copied to HBM,
    # outputs back, PCIe both ways), same workload; not the headline value
    if not args.no_e2e:
        host_in = {k: d[k] for k in inputs}
        host_out = {k: np.empty((T, S), dtype=np_dt) for k in outputs}
        kr.runGraph(ex, mod, host_in, 0, T, host_out)
        e2e = []
        for _ in range(3):
            t0 = time.perf_counter()
            kr.runGraph(ex, mod, host_in, 0, T, host_out)
            e2e.append((time.perf_counter() - t0) * 1e3)
        e2e_ms = sorted(e2e)[1]
        result["end_to_end"] = {"ms": e2e_ms, "cells_per_s": cells / (e2e_ms / 1e3),
                                "pcie_bytes": graph_bytes, "note": "runGraph on host buffers, median of 3"}
        del host_out

    if not args.no_cpu_baseline:
        rows = min(args.cpu_rows, T)
        cores = len(os.sched_getaffinity(0))
        sample = {k: np.ascontiguousarray(d[k][:rows]) for k in inputs}
        base, ref_outs = reference_cpu_baseline(args.module, sample, outputs, S, rows, cores, np_dt)
        if ref_outs is not None:
            # the reference's results against the device's on the same rows
            same = sum(np.array_equal(outs[k][:rows].cpu().numpy(), ref_outs[k], equal_nan=True) for k in outputs)
            base["device_outputs_identical"] = f"{same}/{len(outputs)} factors bit-identical (NaN payloads aside)"
            del ref_outs
        if base is None:
            from kunquant_amd.codegen.program import program_text
            from oracle.kun_oracle import OracleProgram
            prog = OracleProgram(program_text(catalog.module_ir(args.module)))
            prog.run(sample, S, min(rows, 64), 0, min(rows, 64), nthreads=cores)
            runs = []
            for _ in range(5):
                t0 = time.perf_counter()
                prog.run(sample, S, rows, 0, rows, nthreads=cores)
                runs.append(time.perf_counter() - t0)
            cpu_s = sorted(runs)[2]
            base = {"value": len(outputs) * S * rows / cpu_s, "unit": "cells/s", "cores": cores, "kind": "port",
                    "sample": f"same graph and data, first {rows} of {T} days, {S} stocks "
                              f"(median of 5 runs: {cpu_s:.2f} s on {cores} host threads)"}
        else:
            base["sample"] = f"same graph and data, first {rows} of {T} days, {S} stocks " + base["sample"]
        base["host"] = host_cpu()
        result["cpu_baseline"] = base
    if extra is not None:
        result["other_workloads"] = extra
    print(json.dumps(result), flush=True)


# the other workloads of the benchmark, as run by `python bench.py` with no
# flags: (name, arguments, time limit in seconds)
OTHER_WORKLOADS = [
    ("fp64_100_factors", ["--module", "cfg2_f64", "--no-e2e"], 300),
    ("alpha158_fp64", ["--module", "alpha158_f64", "--no-e2e"], 300),
    ("jobs64_uneven", ["--workload", "jobs"], 400),
    ("serve_256_sessions", ["--workload", "serve"], 400),
    ("precision_150_factors", ["--workload", "precision"], 400),
]


def other_workloads():
    """Each other workload's JSON line (bulky per-factor tables dropped),
    from a child process; a failure is recorded, not raised."""
    import subprocess
    out = {}
    for name, extra_args, limit in OTHER_WORKLOADS:
        cmd = [sys.executable, os.path.abspath(__file__), "--no-extra"] + extra_args
        t0 = time.perf_counter()
        print(f"bench: {name} ...", file=sys.stderr, flush=True)
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=limit)
            line = [x for x in r.stdout.strip().splitlines() if x.startswith("{")]
            if r.returncode != 0 or not line:
                out[name] = {"error": f"exit {r.returncode}: {r.stderr.strip()[-400:]}"}
                continue
            rec = json.loads(line[-1])
            rec.get("precision", {}).pop("per_factor_max_err", None)
            rec["wall_s"] = time.perf_counter() - t0
            rec["command"] = "python bench.py " + " ".join(extra_args)
            out[name] = rec
        except subprocess.TimeoutExpired:
            out[name] = {"error": f"time limit {limit} s"}
        print(f"bench: {name} done in {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    return out


def dag_timeline(kr, mod, ptrs, S, T, wall_ms):
    """Every kernel of one run timed alone (an executor without lanes runs
    them one after another, HIP events around each): their sum is the run's
    serial work, and the longest path of the launch DAG weighted by those
    times is its critical path -- no schedule finishes sooner."""
    old = os.environ.get("KUN_STAGE_LANES")
    os.environ["KUN_STAGE_LANES"] = "0"
    try:
        ser = kr.createSingleThreadExecutor()
    finally:
        if old is None:
            del os.environ["KUN_STAGE_LANES"]
        else:
            os.environ["KUN_STAGE_LANES"] = old
    kr.runGraphDevice(ser, mod, ptrs, S, T, 0, T)
    ser.synchronize()
    ser.enable_profiling(True)
    kr.runGraphDevice(ser, mod, ptrs, S, T, 0, T)
    ser.synchronize()
    alone, kind_of = {}, {}
    for i, (_, kind, _, _) in enumerate(mod.stages()):
        kinfo = mod.kernels(i)
        for j in range(max(1, len(kinfo))):
            tot, n = ser.kernel_time(mod, i, j)
            alone[(i, j)] = tot / max(n, 1)
            kind_of[(i, j)] = (("chunked" if kinfo[j][0] == 0 else "sequential") + " time-series" if kind == 0
                               else {1: "rank", 2: "scale", 3: "generic cross-sectional"}.get(kind, "other"))
    ser.enable_profiling(False)
    finish, key_of, prev = {}, {}, {}
    for line in kr.describeGraphBatch([(mod, S, T, 0, T)]).splitlines():
        if not line.startswith("dag "):
            continue
        head, deps = line.split(":")
        _, _, node, stage, kern, _ = head.split()
        node = int(node)
        preds = [int(d) for d in deps.split()]
        best = max(preds, key=lambda d: finish[d], default=None)
        key_of[node] = (int(stage), int(kern))
        prev[node] = best
        finish[node] = (finish[best] if best is not None else 0.0) + alone.get(key_of[node], 0.0)
    serial = sum(alone.values())
    last = max(finish, key=lambda n: finish[n])
    crit = finish[last]
    by_kind: dict = {}
    node = last
    while node is not None:
        k = kind_of.get(key_of[node], "other")
        by_kind[k] = by_kind.get(k, 0.0) + alone.get(key_of[node], 0.0)
        node = prev[node]
    return {"serial_ms": serial, "critical_path_ms": crit, "dag_nodes": len(finish),
            "critical_path_ms_by_kind": {k: round(v, 3) for k, v in by_kind.items()},
            "wall_ms": wall_ms, "work_over_wall": serial / wall_ms,
            "note": "kernels timed alone; critical path = longest launch-DAG path of those times"}


# modules whose graph and compile options the reference's library was built
# with under oracle/_ref (oracle/build_ref.sh + tests/golden/gen_ref_vectors.py)
REF_MODULES = {"cfg2_f32": "rand_f32", "alpha158_f64": "alpha158_f64", "cfg2_f64": "cfg2_f64",
               "prec_f32": "prec_f32", "prec_f64": "prec_f64"}


def reference_cpu_baseline(module, sample, outputs, S, rows, cores, np_dt):
    """The reference's own CPU library (built from its sources into
    oracle/_ref, AVX2) running the same graph through its C API with a
    `cores`-thread executor: (baseline entry, its outputs), or (None, None)
    when that build is not present."""
    import ctypes
    import numpy as np
    ref = REF_MODULES.get(module)
    rdir = os.path.join(ROOT, "oracle", "_ref")
    rt, so = os.path.join(rdir, "libKunRuntime.so"), os.path.join(rdir, f"{ref}.so") if ref else None
    if not ref or not (os.path.exists(rt) and os.path.exists(so)):
        return None, None
    L = ctypes.CDLL(rt, mode=ctypes.RTLD_GLOBAL)
    vp, sz, cp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p
    for n, res, argt in [("kunCreateMultiThreadExecutor", vp, [ctypes.c_int]), ("kunLoadLibrary", vp, [cp]),
                         ("kunGetModuleFromLibrary", vp, [vp, cp]), ("kunCreateBufferNameMap", vp, []),
                         ("kunSetBufferNameMap", None, [vp, cp, vp]), ("kunDestoryBufferNameMap", None, [vp]),
                         ("kunRunGraph", None, [vp, vp, vp, sz, sz, sz, sz])]:
        f = getattr(L, n)
        f.restype, f.argtypes = res, argt
    ex = L.kunCreateMultiThreadExecutor(cores)
    mod = L.kunGetModuleFromLibrary(L.kunLoadLibrary(so.encode()), ref.encode())
    if not ex or not mod:
        return None, None
    outs = {k: np.empty((rows, S), dtype=np_dt) for k in outputs}
    m = L.kunCreateBufferNameMap()
    for k, v in list(sample.items()) + list(outs.items()):
        L.kunSetBufferNameMap(m, k.encode(), v.ctypes.data)
    warm = min(rows, 64)
    L.kunRunGraph(ex, mod, m, S, rows, 0, warm)
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        L.kunRunGraph(ex, mod, m, S, rows, 0, rows)
        runs.append(time.perf_counter() - t0)
    cpu_s = sorted(runs)[2]
    L.kunDestoryBufferNameMap(m)
    return {"value": len(outputs) * S * rows / cpu_s, "unit": "cells/s", "cores": cores, "kind": "reference",
            "sample": f"(the reference's CPU library built from its sources, AVX2, {cores}-thread executor: "
                      f"median of 5 runs {cpu_s:.2f} s, runs {', '.join(f'{r:.2f}' for r in runs)} s)"}, outs


def host_cpu():
    """Cores and model of the host the CPU baseline ran on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "model": model}


def sha256_file(path):
    import hashlib
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def rank_kernel_name(np_dt, S):
    """Name (as rocprofv3 prints it) of the Rank kernel for rows of S stocks."""
    import numpy as np
    t = "float" if np_dt == np.float32 else "double"
    for lim, block, items in ((256, 256, 1), (512, 256, 2), (1024, 256, 4), (2048, 512, 4), (4096, 1024, 4),
                              (8192, 1024, 8), (16384, 1024, 16)):
        if S <= lim:
            kind = 1 if (items == 16 and t == "double") else 0  # 1: keys-only radix sort
            return f"rank_rows<{t}, {block}, {items}, {kind}>"
    return "rank_rows"


def _checksums(outs):
    """Per output: sum of the value bits as int64 (NaN payloads made one)."""
    import torch
    res = []
    for v in outs.values():
        bits = torch.where(torch.isnan(v), torch.full_like(v, float("nan")), v).view(torch.int32)
        res.append(int(bits.to(torch.int64).sum().item()))
    return res


def run_jobs(args):
    import torch
    from kunquant_amd import autotune
    from kunquant_amd.runner import KunRunner as kr
    from kunquant_amd.workloads import catalog, jobs

    specs = jobs.draw_jobs(args.jobs, seed=4)
    mods = {g: catalog.load(g)[1] for g in jobs.JOB_GRAPHS}
    ex = kr.createSingleThreadExecutor()
    items = []
    for sp in specs:
        mod = mods[sp.graph]
        d = jobs.ohlcv_device(sp.stocks, sp.days, sp.seed)
        ins = {k: d[k] for k in mod.getInputNames()}
        del d
        outs = {k: torch.empty((sp.days, sp.stocks), dtype=torch.float32, device="cuda:0")
                for k in mod.getOutputNames()}
        items.append((sp, mod, ins, outs))
    torch.cuda.synchronize()
    tuples = [(mod, {k: v.data_ptr() for k, v in list(i.items()) + list(o.items())}, sp.stocks, sp.days, 0, sp.days)
              for sp, mod, i, o in items]
    cells = sum(len(o) * sp.stocks * sp.days for sp, _, _, o in items)
    print(f"jobs: {len(items)} jobs, {cells:.3g} output cells", file=sys.stderr, flush=True)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        ex.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ex.synchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    # 1. sequential plain runs, device-resident, default chunking
    def sequential():
        for t in tuples:
            kr.runGraphDevice(ex, *t)
            ex.synchronize()
    seq_ms = timed(sequential, args.steps, args.warmup)
    ref_sums = [_checksums(o) for _, _, _, o in items]
    ex.release_temps()
    print(f"jobs: sequential device {seq_ms:.1f} ms", file=sys.stderr, flush=True)

    # 2. autotune every job's chunking (kept on the executor per module and shape)
    t0 = time.perf_counter()
    picked = {}
    for t in tuples:
        best, _ = autotune.tune(ex, t[0], t[1], t[2], t[3], reps=1)
        picked[best] = picked.get(best, 0) + 1
        ex.release_temps()
    tune_s = time.perf_counter() - t0
    print(f"jobs: autotune {tune_s:.1f} s", file=sys.stderr, flush=True)

    # 3. batched enqueue (one call, kernels of all jobs interleaved over the lanes)
    for _, _, _, o in items:
        for v in o.values():
            v.fill_(0.0)
    torch.cuda.synchronize()
    batch_ms = timed(lambda: kr.runGraphBatchDevice(ex, tuples), args.steps, args.warmup)
    same = all(_checksums(o) == r for (_, _, _, o), r in zip(items, ref_sums))
    ex.release_temps()
    print(f"jobs: batched {batch_ms:.1f} ms, results equal {same}", file=sys.stderr, flush=True)

    # 4. the batch as one HIP graph, replayed with one launch per step
    graph_ms, graph_same, build_s = None, None, None
    if not args.no_graph:
        for _, _, _, o in items:
            for v in o.values():
                v.fill_(0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = kr.BatchGraph(ex, tuples)
        build_s = time.perf_counter() - t0
        graph_ms = timed(g.launch, args.steps, args.warmup)
        graph_same = all(_checksums(o) == r for (_, _, _, o), r in zip(items, ref_sums))
        del g
        torch.cuda.synchronize()
        print(f"jobs: graph {graph_ms:.1f} ms (built in {build_s:.2f} s), results equal {graph_same}",
              file=sys.stderr, flush=True)

    # 5. sequential plain runGraph on host buffers (the reference's API: PCIe both ways)
    host_ms = None
    if not args.no_host_baseline:
        host = [(mod, {k: v.cpu().numpy() for k, v in i.items()}, sp.days) for sp, mod, i, _ in items]
        t0 = time.perf_counter()
        for mod, hin, days in host:
            kr.runGraph(ex, mod, hin, 0, days)
        host_ms = (time.perf_counter() - t0) * 1e3
        ex.release_temps()

    result = {
        "metric": "factor-cells/s",
        "value": cells / (batch_ms / 1e3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded OHLCV generated on the device, 10% NaN cells)",
        "config": {"workload": f"{len(items)} batched uneven jobs (stocks 50..6000, days 250..2500, "
                               f"20..60-factor graphs), one batched enqueue, autotuned chunking",
                   "jobs": len(items), "output_cells": cells,
                   "graphs": {gname: sum(1 for sp in specs if sp.graph == gname) for gname in jobs.JOB_GRAPHS}},
        "batched": {"batched_ms": batch_ms, "sequential_device_ms": seq_ms,
                    "sequential_host_runGraph_ms": host_ms,
                    "speedup_vs_sequential_device": seq_ms / batch_ms,
                    "speedup_vs_sequential_host": host_ms / batch_ms if host_ms else None,
                    "graph_ms": graph_ms, "graph_build_s": build_s, "graph_results_equal_sequential": graph_same,
                    "graph_speedup_vs_sequential_device": seq_ms / graph_ms if graph_ms else None,
                    "graph_speedup_vs_sequential_host": host_ms / graph_ms if (graph_ms and host_ms) else None,
                    "autotune_s": tune_s,
                    "policies": {f"{p[0]}/{p[1]}/{p[2]}": n for p, n in picked.items()},
                    "results_equal_sequential": same},
    }
    print(json.dumps(result), flush=True)


def run_serve(args):
    import numpy as np
    from kunquant_amd.runner import KunRunner as kr
    from kunquant_amd.serving import TickServer
    from kunquant_amd.workloads import catalog, data

    _, mod = catalog.load("serve_f32")
    rng = np.random.Generator(np.random.PCG64(5))
    blk = mod.blocking_len
    sizes = [int(-(-int(rng.integers(500, 8001)) // blk) * blk) for _ in range(args.sessions)]
    bps = kr.StreamPool.bytes_per_stock(mod)
    cap = int(args.state_budget_gb * 1e9 // bps)
    ex = kr.createSingleThreadExecutor()
    pool = kr.StreamPool(ex, mod, cap, max_sessions=args.sessions, max_inflight=args.inflight)
    server = TickServer(pool, max_inflight=args.inflight, max_computing=args.max_computing)
    for j, S in enumerate(sizes):
        server.open(j, S)
    live = sorted(server.sid)
    names = mod.getInputNames()
    NR = 8  # distinct rows per session, cycled
    rows = {}
    for j in live:
        d = data.ohlcv(sizes[j], NR, seed=5000 + j, nan_frac=0.01)
        rows[j] = [pool.rows(server.sid[j], {k: np.ascontiguousarray(d[k][t]) for k in names}) for t in range(NR)]
    print(f"serve: {len(live)}/{len(sizes)} sessions admitted, {sum(sizes[j] for j in live)} lanes, "
          f"{bps} B/lane", file=sys.stderr, flush=True)

    # warm-up: 60 rows into every session, as fast as cycles go (saturated)
    t0 = time.perf_counter()
    for t in range(60):
        for j in live:
            server.push(j, rows[j][t % NR])
    server.drain()
    sat_s = time.perf_counter() - t0
    sat_cycles = server.cycles
    print(f"serve: warm-up {sat_s:.2f} s, {sat_cycles} cycles", file=sys.stderr, flush=True)
    server.latencies.clear()
    server.cycles = server.ticks_in_cycles = 0

    # results to host inside every cycle (latency then includes the readback)
    outs_rb = mod.getOutputNames() if args.readback == "all" else []
    pool.set_readback(outs_rb)
    # per-phase breakdown of full cycles (every session ticks once)
    from kunquant_amd.codegen.hip import stream_phases
    phases = stream_phases(catalog.module_ir("serve_f32"))
    pool.profile(True)
    prof = []
    for t in range(4):
        c = pool.submit([rows[j][t % NR] for j in live])
        pool.done(c, wait=True)
        prof.append(pool.phase_times(c))
    pool.profile(False)
    pt = np.median(np.array(prof[1:]), axis=0)
    kinds = ["time-series" if k == 0 else "cross-sectional" for k, _ in phases]
    ph_rows = [{"phase": i, "kind": kinds[i], "stages": len(phases[i][1]), "ms": float(pt[i + 1])}
               for i in range(len(phases))]
    breakdown = {"input_staging_ms": float(pt[0]), "readback_ms": float(pt[-1]),
                 "time_series_phases_ms": float(sum(r["ms"] for r in ph_rows if r["kind"] == "time-series")),
                 "cross_sectional_phases_ms": float(sum(r["ms"] for r in ph_rows if r["kind"] != "time-series")),
                 "cycle_ms": float(pt.sum()), "phases": len(ph_rows),
                 "slowest_phases": sorted(ph_rows, key=lambda r: -r["ms"])[:8],
                 "what": f"one cycle of all {len(live)} sessions, events between phases, median of 3"}
    print(f"serve: full cycle {breakdown['cycle_ms']:.2f} ms: staging {breakdown['input_staging_ms']:.2f}, "
          f"time-series {breakdown['time_series_phases_ms']:.2f}, cross-sectional "
          f"{breakdown['cross_sectional_phases_ms']:.2f}, readback {breakdown['readback_ms']:.2f}",
          file=sys.stderr, flush=True)

    # open loop: Poisson arrivals over all sessions
    def open_loop(rate):
        server.latencies.clear()
        server.cycles = server.ticks_in_cycles = 0
        server.submit_s = 0.0
        n = int(rate * args.seconds)
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        who = rng.integers(0, len(live), n)
        k = 0
        start = time.perf_counter()
        while k < n or server.inflight or server.ready:
            now = time.perf_counter() - start
            while k < n and arr[k] <= now:
                j = live[who[k]]
                server.push(j, rows[j][k % NR], start + arr[k])
                k += 1
            server.poll()
            server.dispatch()
        end = time.perf_counter()
        lat = np.array(server.latencies) * 1e3
        r = {"offered_ticks_per_s": rate, "achieved_ticks_per_s": len(lat) / (end - start),
             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
             "max_ms": float(lat.max()), "ticks_per_cycle": server.ticks_in_cycles / max(1, server.cycles),
             "cycles": server.cycles, "host_submit_ms_per_cycle": server.submit_s / max(1, server.cycles) * 1e3}
        print(f"serve: offered {rate:.0f}/s: p50 {r['p50_ms']:.3f} ms p99 {r['p99_ms']:.3f} ms, "
              f"{r['achieved_ticks_per_s']:.0f} ticks/s, {r['ticks_per_cycle']:.1f} ticks/cycle, "
              f"submit {r['host_submit_ms_per_cycle']:.3f} ms/cycle", file=sys.stderr, flush=True)
        return r, lat

    sweep = [open_loop(float(x))[0] for x in args.rate_sweep.split(",") if x]
    main_r, lat = open_loop(args.rate)
    p50, p99, achieved = main_r["p50_ms"], main_r["p99_ms"], main_r["achieved_ticks_per_s"]
    hbm_only = None
    if outs_rb:
        # the same offered rate with results left in HBM (no readback)
        pool.set_readback([])
        hbm_only = open_loop(args.rate)[0]
        hbm_only["latency"] = "arrival to the tick's results complete in HBM"

    # plain per-session path: one stream context per session, one tick at a time
    plain = []
    for j in live[:8]:
        st = kr.StreamContext(ex, mod, sizes[j])
        hs = {k_: st.queryBufferHandle(k_) for k_ in names}
        for t in range(40):
            t1 = time.perf_counter()
            for k_, v in zip(names, rows[j][t % NR].arrays):
                st.pushData(hs[k_], v)
            st.run()
            if t >= 10:
                plain.append((time.perf_counter() - t1) * 1e3)
        del st
    plain = np.array(plain)
    result = {
        "metric": "stream tick latency p99",
        "value": p99,
        "unit": "ms",
        "n_gpus": 1,
        "steps": len(lat),
        "warmup": 60,
        "ms_per_step": float(lat.mean()),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded OHLCV rows, 1% NaN cells; Poisson arrival trace)",
        "config": {"workload": f"{args.sessions} stream sessions (universes 500..8000 stocks), 100-factor graph, "
                               f"Poisson {args.rate:.0f} ticks/s over all sessions, continuous batching in one "
                               f"session pool",
                   "sessions": args.sessions, "admitted": len(live), "lanes": int(sum(sizes[j] for j in live)),
                   "state_budget_GB": args.state_budget_gb, "bytes_per_lane": bps},
        "serving": {**main_r,
                    "rate_sweep": sweep,
                    "saturated_ticks_per_s": 60 * len(live) / sat_s, "saturated_ms_per_cycle": sat_s / sat_cycles * 1e3,
                    "latency": ("arrival (scheduled time) to the tick's results in host memory (every output "
                                "row copied to pinned host memory inside the cycle)" if outs_rb else
                                "arrival (scheduled time) to the tick's results complete in HBM"),
                    "readback_bytes_per_tick": len(outs_rb) * float(np.mean([sizes[j] for j in live])) * 4,
                    "cycle_breakdown": breakdown,
                    "same_rate_results_in_hbm": hbm_only,
                    "cycle_share": server.cycle_share, "cycles_in_flight": args.inflight,
                    "max_computing": server.max_computing,
                    "plain_per_session": {"p50_ms": float(np.percentile(plain, 50)),
                                          "p99_ms": float(np.percentile(plain, 99)),
                                          "ticks_per_s": 1e3 / float(plain.mean()),
                                          "what": "StreamContext per session: push 6 rows, run, one tick at a "
                                                  "time (8 sessions x 30 ticks)"},
                    "speedup_ticks_per_s_vs_plain": (60 * len(live) / sat_s) / (1e3 / float(plain.mean()))},
    }
    print(json.dumps(result), flush=True)


def factor_errors(got, ref, floor=1e-6):
    """Per factor (torch tensors [T][S] on the device): relative error
    |x - r| / max(|r|, floor * max|r|) over cells where both are finite --
    its maximum, its median and the fraction of cells within 1e-5 -- and
    the cells where exactly one of the two is NaN."""
    import torch
    res = {}
    for k, r in ref.items():
        x = got[k].to(torch.float64)
        fin = torch.isfinite(r) & torch.isfinite(x)
        mask_diff = int((torch.isnan(r) != torch.isnan(x)).sum().item())
        if not bool(fin.any()):
            res[k] = (0.0, 0.0, 1.0, mask_diff)
            continue
        ra = r.abs()
        scale = float(ra[fin].max().item())
        den = torch.clamp(ra, min=max(scale * floor, 1e-300))
        err = ((x - r).abs() / den)[fin]
        res[k] = (float(err.max().item()), float(err.median().item()), float((err <= 1e-5).double().mean().item()),
                  mask_diff)
    return res


def run_precision(args):
    import numpy as np
    import torch
    from kunquant_amd.runner import KunRunner as kr
    from kunquant_amd.workloads import catalog, data

    S, T = 4096, 2520
    d64 = data.ohlcv(S, T, seed=3, dtype=np.float64, price_scale=(1e3, 1e5), volume_scale=(1e6, 1e9))
    ex = kr.createSingleThreadExecutor()
    outs, times = {}, {}
    for name in ("prec_f64", "prec_f32", "prec_f32x"):
        _, mod = catalog.load(name)
        t_dt = torch.float64 if mod.dtype == np.float64 else torch.float32
        dev = {k: torch.from_numpy(d64[k].astype(mod.dtype)).to("cuda:0") for k in mod.getInputNames()}
        o = {k: torch.empty((T, S), dtype=t_dt, device="cuda:0") for k in mod.getOutputNames()}
        ptrs = {k: v.data_ptr() for k, v in list(dev.items()) + list(o.items())}
        kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T)
        ex.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T)
        ex.synchronize()
        times[name] = (time.perf_counter() - t0) / args.steps * 1e3
        outs[name] = o
        del dev
        ex.release_temps()
        print(f"precision: {name} {times[name]:.1f} ms", file=sys.stderr, flush=True)
    ref = outs["prec_f64"]
    e32 = factor_errors(outs["prec_f32"], ref)
    e32x = factor_errors(outs["prec_f32x"], ref)
    names = sorted(ref)
    a = np.array([e32[k][0] for k in names])
    b = np.array([e32x[k][0] for k in names])
    cells = len(names) * S * T

    def summary(e):
        mx = np.array([e[k][0] for k in names])
        med = np.array([e[k][1] for k in names])
        frac = np.array([e[k][2] for k in names])
        return {"max_err": {"median": float(np.median(mx)), "p90": float(np.percentile(mx, 90)),
                            "max": float(mx.max()), "factors_le_1e-5": int((mx <= 1e-5).sum())},
                "median_cell_err": {"median": float(np.median(med)), "max": float(med.max()),
                                    "factors_le_1e-5": int((med <= 1e-5).sum())},
                "cells_within_1e-5": float(np.mean(frac))}
    result = {
        "metric": "per-factor max relative error vs fp64 (fp32 path, median over factors)",
        "value": float(np.median(a)),
        "unit": "relative",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": times["prec_f32"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded OHLCV, prices scaled to 1e3..1e5, volumes to 1e6..1e9)",
        "config": {"workload": f"150-factor precision graph (quantile, regression slope/R2, 60-row var/corr), "
                               f"{S} stocks x {T} days",
                   "stocks": S, "days": T, "factors": len(names)},
        "precision": {
            "error": "per cell |x - r| / max(|r|, 1e-6 max|r| of the factor), r = fp64 run of the same graph and "
                     "algorithm (bit-exact with the fp64 oracle), over cells where both are finite; per factor its "
                     "max and median, then summarised over the 150 factors",
            "fp32": summary(e32), "fp32_fp64_finalize": summary(e32x),
            "nan_mask_mismatch_cells": {"fp32": int(sum(e32[k][3] for k in names)),
                                        "fp32_fp64_finalize": int(sum(e32x[k][3] for k in names))},
            "cells_per_s": {k: cells / (v / 1e3) for k, v in times.items()},
            "per_factor_max_err": {k: [e32[k][0], e32x[k][0]] for k in names},
        },
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
