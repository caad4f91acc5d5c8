"""Benchmark of the factor pipeline on one MI355X.

Workload (default): the seeded 100-factor random graph over synthetic OHLCV,
8192 stocks x 1024 days, fp32, 10% NaN input cells, TS layout. A step is one
batch run of the whole graph (every stage kernel) with inputs and outputs
already resident in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--module cfg2_f32|cfg2_f64]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--module", default="cfg2_f32")
    ap.add_argument("--stocks", type=int, default=8192)
    ap.add_argument("--days", type=int, default=1024)
    ap.add_argument("--cpu-rows", type=int, default=128, help="rows of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    import numpy as np
    import torch
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

    for _ in range(args.warmup):
        kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T)
    ex.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kr.runGraphDevice(ex, mod, ptrs, S, T, 0, T)
    ex.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
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
                label = {1: "rank", 2: "scale"}[kind]
            per_kernel.append((avg, i, j, label, nbytes))
            stage_ms[i] = stage_ms.get(i, 0.0) + avg
    ex.enable_profiling(False)
    per_kernel.sort(reverse=True)
    kernel_sum = sum(p[0] for p in per_kernel)
    dom_ms, dom_i, dom_j, dom_label, dom_bytes = per_kernel[0]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    # HBM traffic of the dominant kernel, from rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes scaled by the same counters on a plain copy (profiles/*_traffic.json)
    dom_name = f"kun_{args.module}_s{dom_i}_k{dom_j}"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{args.module}_{S}x{T}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tk = json.load(f).get("kernels", {}).get(dom_name)
        if tk:
            traffic = tk["fetch_bytes"] + tk["write_bytes"]
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
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if np_dt == np.float32 else "f64",
        "data": "synthetic (seeded OHLCV, 10% NaN cells)",
        "config": {"workload": f"{len(outputs)}-factor seeded random graph, {S} stocks x {T} days, "
                               f"{'fp32' if esz == 4 else 'fp64'}, batch runGraph with HBM-resident inputs",
                   "module": args.module, "stocks": S, "days": T, "factors": len(outputs), "inputs": len(inputs),
                   "stages": len(stages)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"stage {dom_i} kernel {dom_j} ({dom_label})", "kernel_ms": dom_ms,
                     "algorithmic_bytes": dom_bytes},
        "graph": {"algorithmic_bytes": graph_bytes, "achieved_GBs": graph_gbs,
                  "frac_of_hbm_peak": graph_gbs / HBM_PEAK_GBS, "sum_kernel_ms": kernel_sum,
                  "kernel_ms_by_kind": {k: round(v, 3) for k, v in by_kind.items()},
                  "launches_per_step": len(per_kernel),
                  "top_kernels_ms": [[p[1], p[2], p[3], round(p[0], 4)] for p in per_kernel[:8]]},
    }

    if not args.no_cpu_baseline:
        from kunquant_amd.codegen.program import program_text
        from oracle.kun_oracle import OracleProgram
        rows = min(args.cpu_rows, T)
        cores = min(16, os.cpu_count() or 1)
        prog = OracleProgram(program_text(catalog.module_ir(args.module)))
        sample = {k: np.ascontiguousarray(d[k][:rows]) for k in inputs}
        t0 = time.perf_counter()
        prog.run(sample, S, rows, 0, rows, nthreads=cores)
        cpu_s = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": len(outputs) * S * rows / cpu_s, "unit": "cells/s", "cores": cores,
                                  "kind": "port",
                                  "sample": f"same graph and data, first {rows} of {T} days, {S} stocks "
                                            f"({cpu_s:.1f} s on {cores} host threads)"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
