#!/usr/bin/env python3
"""Throughput of the explicit hex path on one MI355X.

A "step" is one explicit cycle (element forces + assembly + time integration)
of the whole model, inputs resident in HBM.  Workload: the 160x160x160 hex
block (4,096,000 SFORC3 elements, 1 mm edge) of the GPU-filling roofline
model, law-2 steel, initial velocity 10 m/s along z plus N(0, 0.5 m/s) noise
per component from a fixed seed.  value = element-updates/s of the plain FP64
method; the split-precision form is reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 160]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20261203


def build_workload(n: int):
    import numpy as np

    from openradioss_amd import model as M

    x, conn = M.hex_block(n, n, n, 1.0)
    rng = np.random.default_rng(SEED)
    v = 500.0 * rng.standard_normal(x.shape)
    v[:, 2] += 1.0e4
    return M.build_hex_model(x, conn, M.STEEL, v0=v)


def gpu_run(model, precision: str, steps: int, warmup: int):
    import numpy as np

    from openradioss_amd import engine as E
    from openradioss_amd import traffic as T

    g = E.GpuHexModel(model, precision=precision)
    g.run(warmup)
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, sync=False)
    g.sync()
    wall = time.perf_counter() - t0
    # kernel durations with HIP events on the model stream (separate short run)
    g.set_timing(True)
    g.kernel_times(reset=True)
    nt = max(10, min(steps, 50))
    g.run(nt)
    elem_ms, node_ms, ncyc = g.kernel_times(reset=True)
    g.set_timing(False)
    ts = g.time_state()
    x, v, d = g.nodes()
    g.close()
    ne, nn = model.ne, model.nn
    eb = T.hex_element_kernel_bytes(ne, nn, precision)
    nb = T.node_kernel_bytes(ne, nn, precision)
    return dict(wall=wall, elem_ms=elem_ms, node_ms=node_ms, ts=ts, d=d, eb=eb, nb=nb,
                finite=bool(np.isfinite(d).all()))


def cpu_baseline(model, budget_s: float = 12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = max(1, min(16, os.cpu_count() or 1))
    o = O.HexOracle(model, nthreads=threads)
    o.cycle(1)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.cycle(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 200:
            break
    return dict(value=model.ne * n / el, unit="element-updates/s", cores=threads, kind="port",
                sample=f"C oracle (FP64, OpenMP over groups of 128) on the same {model.ne}-element block, "
                       f"{n} cycles after 1 warm-up cycle, {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=160, help="block edge in elements (160 = roofline model)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    t0 = time.perf_counter()
    model = build_workload(args.n)
    print(f"[bench] model built: {model.ne} hex, {model.nn} nodes in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)

    r64 = gpu_run(model, "fp64", args.steps, args.warmup)
    eu64 = model.ne * args.steps / r64["wall"]
    print(f"[bench] fp64: {eu64:.3e} element-updates/s, elem {r64['elem_ms']:.3f} ms, node {r64['node_ms']:.3f} ms",
          file=sys.stderr, flush=True)
    elem_gbs = r64["eb"]["as_implemented"] / (r64["elem_ms"] * 1e-3) / 1e9
    out = {
        "metric": "element-updates/s (explicit hex cycle, plain FP64 method)",
        "value": eu64,
        "unit": "element-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r64["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"hex block {args.n}^3 (1 mm), law-2 steel, v0 = 10 m/s + N(0,0.5) seed {SEED}",
                   "elements": model.ne, "nodes": model.nn, "cycles": args.steps},
        "roofline": {"bound": "hbm", "achieved": elem_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": elem_gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_hex_forces<double>", "kernel_ms": r64["elem_ms"],
                     "bytes_per_launch": r64["eb"]["as_implemented"]},
        "node_kernel": {"ms": r64["node_ms"],
                        "gbs": r64["nb"]["as_implemented"] / (r64["node_ms"] * 1e-3) / 1e9},
        "final_time": r64["ts"].tt,
    }
    if not args.no_split:
        rs = gpu_run(model, "split", args.steps, args.warmup)
        import numpy as np
        eus = model.ne * args.steps / rs["wall"]
        derr = float(np.abs(rs["d"] - r64["d"]).max() / np.abs(r64["d"]).max())
        out["split_form"] = {"value": eus, "speedup_vs_fp64": eus / eu64, "ms_per_step": 1e3 * rs["wall"] / args.steps,
                             "elem_kernel_ms": rs["elem_ms"], "node_kernel_ms": rs["node_ms"],
                             "elem_kernel_gbs": rs["eb"]["as_implemented"] / (rs["elem_ms"] * 1e-3) / 1e9,
                             "disp_rel_err_vs_fp64_gpu": derr, "dtype": "f32 element / f64 node"}
        print(f"[bench] split: {eus:.3e} element-updates/s (x{eus / eu64:.2f}), disp err {derr:.2e}",
              file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(model)
        print(f"[bench] cpu: {out['cpu_baseline']['value']:.3e} element-updates/s", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
