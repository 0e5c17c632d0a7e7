#!/usr/bin/env python3
"""Throughput of the explicit hex + Q4 path on one MI355X.

A "step" is one explicit cycle (hex and shell element forces, assembly, time
integration) of the whole model, inputs resident in HBM.  Workload: the
GPU-filling roofline model -- a 160x160x160 hex block (4,096,000 SFORC3
elements, 1 mm edge) with a 1000x1000 Q4 plate (1,000,000 CFORC3 shells,
t = 2 mm, 5 points) tied to the block's top face at 25,921 shared nodes,
law-2 steel, initial velocity 10 m/s along z plus N(0, 0.5 m/s) per component
from a fixed seed.  value = element-updates/s of the plain FP64 method; the
split-precision form (FP32 element arithmetic and state, FP64 nodes) is
reported beside it with its displacement and energy error against FP64.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 160] [--plate 1000]
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(n: int, plate: int):
    from openradioss_amd import model as M
    return M.block_with_plate(n=n, plate=plate, seed=SEED)


def gpu_run(model, precision: str, steps: int, warmup: int):
    import numpy as np

    from openradioss_amd import engine as E

    g = E.GpuModel(model, precision=precision)
    g.run(warmup)
    variants = g.autotune(5)  # launch variants of the element kernels (results do not depend on them)
    g.sync()
    t0 = time.perf_counter()
    g.run(steps, sync=False)
    g.sync()
    wall = time.perf_counter() - t0
    # kernel durations with HIP events on the model's stream (a separate short run)
    g.set_timing(True)
    g.kernel_times(reset=True)
    g.run(max(10, min(steps, 50)))
    hex_ms, q4_ms, node_ms, _ = g.kernel_times(reset=True)
    g.set_timing(False)
    ts = g.time_state()
    x, v, vr, d = g.nodes6()
    e = g.energy()
    foot = g.footprint()
    plan = g.assembly_stats()
    g.close()
    return dict(wall=wall, hex_ms=hex_ms, q4_ms=q4_ms, node_ms=node_ms, ts=ts, d=d, energy=e, footprint=foot,
                finite=bool(np.isfinite(d).all()), variants=variants, plan=plan)


def kernel_name(family: str, real: str, variant: int) -> str:
    """Name of an element-kernel launch variant as the profiles list it (one 256-element tile per
    workgroup; register budget for 2 waves per SIMD for even variants, 3 for odd ones; the
    instantiation without the time history's EHOUR, the one the bench cycles run)."""
    return f"k_{family}_forces<{real}, {real}, 256, {2 + (variant & 1)}, false>"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on the default workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled with the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "block160_plate1000_hbm_traffic.md")
    col = None
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if "scaled traffic GB" in cells:
                col = cells.index("scaled traffic GB")
            elif len(cells) > 1 and cells[1] == "kernel":
                col = None  # another table
            elif col is not None and len(cells) > col and cells[1] == f"`{kernel}`":
                return float(cells[col]) * 1e9
    except (OSError, ValueError):
        return None
    return None


def cpu_model_name() -> str:
    """The host CPU's model name, from /proc/cpuinfo."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(model, budget_s: float = 12.0):
    """The oracle's C restatement of the FP64 cycle (OpenMP dynamic over groups of 128, static
    over nodes) on every host core, on a bounded sample of cycles of the same model."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = max(1, os.cpu_count() or 1)
    o = O.ModelOracle(model, nthreads=threads)
    o.cycle(1)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.cycle(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 200:
            break
    ne = model.ne_hex + model.ne_q4
    return dict(value=ne * n / el, unit="element-updates/s", cores=threads, kind="port",
                cpu_model=cpu_model_name(), logical_cpus=os.cpu_count(),
                sample=f"C oracle (FP64 restatement of SFORC3/CFORC3 + nodal update, OpenMP dynamic over groups of "
                       f"128, {threads} threads = every logical CPU of the host) on the same {ne}-element model, "
                       f"{n} cycles after 1 warm-up cycle, {el:.1f} s")


def small_leg(cycles: int = 1000) -> tuple[dict, dict]:
    """The small reference workloads (bar 20x20x80, 32,000 hexes; tube 60x300, 72,000 shells), FP64:
    `cycles` timed cycles from the initial state (after 20 warm-up cycles and the launch-variant
    autotuning on a separate instance);
    at these sizes each SIMD holds at most one or two waves of an element kernel, so a cycle is bound
    by latency (one element's chain, and the three dependent launches), not by the kernels' work.
    Then each run continues to the length of its reference fixture (bar 1,000 cycles, tube 8,800)
    and is compared with the state the reference's own compiled routines (glibc libm) reached
    there: the parity block (displacement and energy deviations)."""
    import numpy as np

    from openradioss_amd import engine as E
    from openradioss_amd import model as M
    gold = os.path.join(ROOT, "tests", "golden")
    out, parity = {}, {}
    for name, build, full, fixture in (
            ("bar_20x20x80", lambda: M.hex_model_to_model(M.cantilever_bar(20, 20, 80)), 1000,
             "bar_20x20x80_c1000.npz"),
            ("tube_60x300", M.square_tube, 8800, "tube_60x300_c8800.npz")):
        m = build()
        ne = m.ne_hex + m.ne_q4
        with E.GpuModel(m, precision="fp64") as g:
            g.run(20)
            vh, vq = g.autotune(5)       # every variant gives the same bits
        with E.GpuModel(m, precision="fp64") as g:
            g.set_variant("hex", vh)
            g.set_variant("q4", vq)
            t0 = time.perf_counter()
            g.run(cycles)
            t = time.perf_counter() - t0
            out[name] = dict(elements=ne, cycles=cycles, us_per_cycle=1e6 * t / cycles,
                             element_updates_per_s=ne * cycles / t, kernel_variants={"hex": vh, "q4": vq})
            g.run(full - cycles)
            ts = g.time_state()
            x, v, vr, d = g.nodes6()
            e = g.energy()
        try:
            ref = np.load(os.path.join(gold, fixture))
        except OSError:
            continue
        every = int(ref["d_every"][0])
        dref = ref["d_sub"]
        r = dict(cycles=int(ts.ncycle), reference=f"tests/golden/{fixture} (the reference's Fortran, glibc libm)",
                 d_rel_err=float(np.abs(d[::every] - dref).max() / np.abs(dref).max()),
                 time_rel_err=float(abs(ts.tt - ref["time"][0]) / ref["time"][0]))
        if "energy" in ref.files:
            er = ref["energy"]
            r["energy_rel_err"] = float(abs(e.sum() - er.sum()) / abs(er.sum()))
            r["internal_energy_rel_err"] = float(abs(e[0] - er[0]) / abs(er[0]))
        parity[name] = r
    return out, parity


def sweep_leg(precision: str, cycles: int, warmup: int, n: int = 256):
    """The parametric tube sweep (n tubes, element size 1-4 mm, 4.5k-72k shells each) advanced
    together in one batch; a bounded sample of `cycles` batch cycles after `warmup`."""
    from openradioss_amd import engine as E
    from openradioss_amd import model as M

    t0 = time.perf_counter()
    sw = M.tube_sweep(n, seed=SEED)
    b = E.GpuBatch()
    gms = []
    for m, _ in sw:
        g = E.GpuModel(m, precision=precision)
        b.add(g, 5.0e-3)
        gms.append(g)
    setup = time.perf_counter() - t0
    b.run(warmup)
    variants = b.autotune(2)
    t1 = time.perf_counter()
    b.run(cycles)
    wall = time.perf_counter() - t1
    active = b.active()
    ne = sum(m.ne_q4 for m, _ in sw)
    tts = [g.time_state().tt for g in gms]
    b.close()
    for g in gms:
        g.close()
    return dict(models=n, shells=ne, precision=precision, cycles=cycles, ms_per_cycle=1e3 * wall / cycles,
                element_updates_per_s=ne * cycles / wall, active_after=active, setup_s=setup,
                variants=dict(zip(("fp64_hex", "fp64_q4", "split_hex", "split_q4"), variants)),
                sim_time_range_s=[min(tts), max(tts)])


def sweep_full(precision: str, n: int = 256, t_end: float = 5.0e-3, step: int = 500, max_wall: float = 150.0):
    """The whole tube sweep to its end time: every tube runs its own cycles to t_end in one batch;
    stopped tubes leave the launches.  Reports wall time and element-updates/s, and the tubes
    still running after max_wall seconds (their time, cycle count and time step)."""
    from openradioss_amd import engine as E
    from openradioss_amd import model as M

    sw = M.tube_sweep(n, seed=SEED)
    b = E.GpuBatch()
    gms = []
    for m, _ in sw:
        g = E.GpuModel(m, precision=precision)
        b.add(g, t_end)
        gms.append(g)
    b.run(10)
    variants = b.autotune(2)
    t0 = time.perf_counter()
    active = b.active()
    last = t0
    while active > 0:
        b.run(step, sync=False)
        active = b.active()
        now = time.perf_counter()
        if now - last > 30.0:
            log(f"sweep-full {precision}: {active} tubes running after {now - t0:.0f} s")
            last = now
        if now - t0 > max_wall:
            break
    wall = time.perf_counter() - t0
    states = [g.time_state() for g in gms]
    cyc = [s.ncycle for s in states]
    tts = [s.tt for s in states]
    late = [dict(tube=i, h=p["h"], tt=s.tt, ncycle=s.ncycle, dt2=s.dt2)
            for i, (s, (_, p)) in enumerate(zip(states, sw)) if s.tt < t_end and s.mstop == 0]
    stopped = [dict(tube=i, h=p["h"], tt=s.tt, ncycle=s.ncycle, dt2=s.dt2, mstop=s.mstop)
               for i, (s, (_, p)) in enumerate(zip(states, sw)) if s.mstop != 0]
    eu = sum(m.ne_q4 * c for (m, _), c in zip(sw, cyc))
    b.close()
    for g in gms:
        g.close()
    return dict(models=n, precision=precision, t_end=t_end, wall_s=wall, finished=active == 0, running=late,
                stopped=stopped,
                element_updates=eu, element_updates_per_s=eu / wall, cycles_min=min(cyc), cycles_max=max(cyc),
                sim_time_min=min(tts), variants=variants)


def serve_leg(nreq: int, rate: float, time_scale: float):
    """Served crash solves: Poisson arrivals, bars and tubes of 5k-500k elements, FP64 / split,
    end times U(1, 10) ms x time_scale, admitted by HBM footprint into one running batch."""
    from openradioss_amd import lib as L
    from openradioss_amd import serve as S

    reqs = S.request_stream(nreq, rate=rate, seed=SEED, time_scale=time_scale)
    free, total = L.device_memory(0)
    budget = int(0.8 * free)
    srv = S.Server(hbm_budget=budget, cycles_per_step=200, keep_results=False)
    st = srv.serve(reqs)
    srv.close()
    st.update(rate_per_s=rate, seed=SEED, time_scale=time_scale, hbm_budget_gb=budget / 1e9,
              max_used_gb=srv.max_used / 1e9, elements=[min(r.ne for r in reqs), max(r.ne for r in reqs)],
              trace=[dict(arrival_s=round(r.arrival, 4), kind=r.kind, elements=r.ne, precision=r.precision,
                          t_end_ms=round(1e3 * r.t_end, 4), cycles=r.ncycle, latency_s=round(r.latency, 4))
                     for r in reqs])
    return st


def copy_probe(mib: int):
    """Three plain device-to-device copies of a known size: each reads and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 3 copies of {mib} MiB")


def stream_copy(mib: int = 1024, reps: int = 10) -> dict:
    """Achievable HBM bandwidth: a device-to-device hipMemcpy of mib MiB (reads and writes mib
    MiB), timed with HIP events over reps copies.  Calls the HIP runtime directly, so it runs
    beside the path's library without another runtime on the device."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    size = mib * (1 << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    ms = C.c_float()

    def ok(rc):
        if rc != 0:
            raise RuntimeError(f"HIP error {rc} in the copy measurement")

    ok(hip.hipSetDevice(0))
    ok(hip.hipMalloc(C.byref(src), C.c_size_t(size)))
    ok(hip.hipMalloc(C.byref(dst), C.c_size_t(size)))
    try:
        ok(hip.hipMemset(src, 1, C.c_size_t(size)))
        ok(hip.hipEventCreate(C.byref(e0)))
        ok(hip.hipEventCreate(C.byref(e1)))
        ok(hip.hipMemcpy(dst, src, C.c_size_t(size), 3))  # hipMemcpyDeviceToDevice, warm-up
        ok(hip.hipEventRecord(e0, None))
        for _ in range(reps):
            ok(hip.hipMemcpy(dst, src, C.c_size_t(size), 3))
        ok(hip.hipEventRecord(e1, None))
        ok(hip.hipEventSynchronize(e1))
        ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1))
        ok(hip.hipEventDestroy(e0))
        ok(hip.hipEventDestroy(e1))
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    t = ms.value / reps
    gbs = 2 * size / (t * 1e-3) / 1e9
    return {"gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS, "bytes_per_copy": 2 * size, "ms_per_copy": t,
            "what": f"{mib} MiB device-to-device hipMemcpy, read + write bytes over the event time of {reps} copies"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=160, help="hex block edge in elements (160 = roofline model)")
    ap.add_argument("--plate", type=int, default=1000, help="Q4 plate edge in elements (1000 = roofline model)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision form")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched tube sweep")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving loop")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the small workloads (bar, tube) and their parity block against the reference")
    ap.add_argument("--sweep-full", action="store_true",
                    help="only run the whole 256-tube sweep to 5 ms (FP64 and split) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.sweep_full:
        res = {p: sweep_full(p) for p in ("fp64", "split")}
        print(json.dumps({"sweep_full": res}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; runs on several GPUs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    from openradioss_amd import traffic as T

    t0 = time.perf_counter()
    model = build_workload(args.n, args.plate)
    ne = model.ne_hex + model.ne_q4
    log(f"model built: {model.ne_hex} hex + {model.ne_q4} Q4, {model.nn} nodes in {time.perf_counter() - t0:.1f} s")

    r64 = gpu_run(model, "fp64", args.steps, args.warmup)
    eu64 = ne * args.steps / r64["wall"]
    log(f"fp64: {eu64:.3e} element-updates/s; hex {r64['hex_ms']:.3f} ms, q4 {r64['q4_ms']:.3f} ms, "
        f"node {r64['node_ms']:.3f} ms per cycle")
    tb = T.cycle_bytes(model, "fp64", r64["plan"])
    # achieved bandwidth from the compulsory bytes (element state, connectivity and node data once);
    # the as-implemented figure (with the assembly's rows and row indices) is reported beside it
    hex_gbs = tb["hex"]["compulsory"] / (r64["hex_ms"] * 1e-3) / 1e9
    hex_gbs_impl = tb["hex"]["as_implemented"] / (r64["hex_ms"] * 1e-3) / 1e9
    hex_traffic = (measured_traffic(kernel_name("hex", "double", r64["variants"][0]))
                   if (args.n, args.plate) == (160, 1000) else None)
    cyc_ms = r64["hex_ms"] + r64["q4_ms"] + r64["node_ms"]
    out = {
        "metric": "element-updates/s & %HBM peak vs FP64 plain; p50/p99 latency per served solve",
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
        "config": {"workload": f"hex block {args.n}^3 (1 mm) + Q4 plate {args.plate}^2 (1 mm, t=2 mm, 5 points) "
                               f"tied on the block top, law-2 steel, v0 = 10 m/s + N(0,0.5) seed {SEED}",
                   "hex_elements": model.ne_hex, "q4_elements": model.ne_q4, "nodes": model.nn},
        "roofline": {"bound": "hbm", "achieved": hex_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hex_gbs / HBM_PEAK_GBS,
                     "traffic": hex_traffic,
                     "kernel": kernel_name("hex", "double", r64["variants"][0]), "kernel_ms": r64["hex_ms"],
                     "compulsory_bytes_per_launch": tb["hex"]["compulsory"],
                     "as_implemented_bytes_per_launch": tb["hex"]["as_implemented"],
                     "as_implemented_gbs": hex_gbs_impl,
                     "traffic_over_compulsory": (hex_traffic / tb["hex"]["compulsory"]) if hex_traffic else None},
        "kernels_fp64": {
            k: {"ms": r64[f"{k}_ms"], "compulsory_bytes": tb[k]["compulsory"],
                "as_implemented_bytes": tb[k]["as_implemented"],
                "gbs": tb[k]["compulsory"] / (r64[f"{k}_ms"] * 1e-3) / 1e9,
                "as_implemented_gbs": tb[k]["as_implemented"] / (r64[f"{k}_ms"] * 1e-3) / 1e9}
            for k in ("hex", "q4", "node")},
        "cycle_fp64": {"kernel_ms": cyc_ms, "bytes_as_implemented": tb["total"]["as_implemented"],
                       "bytes_compulsory": tb["total"]["compulsory"],
                       "gbs": tb["total"]["compulsory"] / (cyc_ms * 1e-3) / 1e9,
                       "frac": tb["total"]["compulsory"] / (cyc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "as_implemented_gbs": tb["total"]["as_implemented"] / (cyc_ms * 1e-3) / 1e9,
                       "bytes_per_element_update": tb["total"]["compulsory"] / ne},
        "final_time": r64["ts"].tt,
        "kernel_variants_fp64": {"hex": r64["variants"][0], "q4": r64["variants"][1]},
        "footprint_gb_fp64": r64["footprint"] / 1e9,
        "assembly_plan": r64["plan"],
    }
    out["hbm_copy"] = stream_copy()
    out["roofline"]["frac_of_copy"] = hex_gbs / out["hbm_copy"]["gbs"]
    log(f"copy: {out['hbm_copy']['gbs']:.0f} GB/s ({out['hbm_copy']['frac_of_peak']:.2f} of peak); hex kernel "
        f"{hex_gbs:.0f} GB/s")
    if not args.no_split:
        rs = gpu_run(model, "split", args.steps, args.warmup)
        eus = ne * args.steps / rs["wall"]
        derr = float(np.abs(rs["d"] - r64["d"]).max() / np.abs(r64["d"]).max())
        eerr = float(abs(rs["energy"].sum() - r64["energy"].sum()) / abs(r64["energy"].sum()))
        ts32 = T.cycle_bytes(model, "split", rs["plan"])
        cyc32 = rs["hex_ms"] + rs["q4_ms"] + rs["node_ms"]
        out["split_form"] = {
            "value": eus, "speedup_vs_fp64": eus / eu64, "ms_per_step": 1e3 * rs["wall"] / args.steps,
            "kernels": {k: {"ms": rs[f"{k}_ms"], "gbs": ts32[k]["compulsory"] / (rs[f"{k}_ms"] * 1e-3) / 1e9,
                            "as_implemented_gbs": ts32[k]["as_implemented"] / (rs[f"{k}_ms"] * 1e-3) / 1e9}
                        for k in ("hex", "q4", "node")},
            "cycle_gbs": ts32["total"]["compulsory"] / (cyc32 * 1e-3) / 1e9,
            "cycle_frac": ts32["total"]["compulsory"] / (cyc32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            # bandwidth bound of the split form's speed-up: FP64 cycle bytes over split cycle bytes
            "compulsory_bytes_ratio_fp64_over_split": tb["total"]["compulsory"] / ts32["total"]["compulsory"],
            "disp_rel_err_vs_fp64_gpu": derr, "energy_rel_err_vs_fp64_gpu": eerr,
            "kernel_variants": {"hex": rs["variants"][0], "q4": rs["variants"][1]},
            "dtype": "f32 element / f64 node"}
        log(f"split: {eus:.3e} element-updates/s (x{eus / eu64:.2f}), disp err {derr:.2e}, energy err {eerr:.2e}")
    if not args.no_small:
        out["small_models"], out["parity"] = small_leg()
        for k, r in out["small_models"].items():
            log(f"{k}: {r['element_updates_per_s']:.3e} element-updates/s, {r['us_per_cycle']:.1f} us per cycle "
                f"(variants {r['kernel_variants']})")
        for k, r in out["parity"].items():
            log(f"parity {k} at cycle {r['cycles']}: displacement {r['d_rel_err']:.2e}, "
                f"energy {r.get('energy_rel_err', float('nan')):.2e} against the reference")
    if not args.no_sweep:
        out["batch_sweep"] = {}
        for prec in ("fp64", "split"):
            r = sweep_leg(prec, cycles=100, warmup=10)
            out["batch_sweep"][prec] = r
            log(f"sweep {prec}: {r['element_updates_per_s']:.3e} element-updates/s over {r['models']} tubes "
                f"({r['shells']} shells), {r['ms_per_cycle']:.3f} ms per batch cycle")
    if not args.no_serve:
        # the served workload at its full end times (1-10 ms), at three arrival rates
        out["serving"] = []
        for rate in (2.0, 8.0, 32.0):
            r = serve_leg(16, rate=rate, time_scale=1.0)
            out["serving"].append(r)
            log(f"serve at {rate:g}/s: {r['requests']} requests, p50 {r['p50_s']:.3f} s, p99 {r['p99_s']:.3f} s, "
                f"{r['element_updates_per_s']:.3e} element-updates/s")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(model)
        log(f"cpu: {out['cpu_baseline']['value']:.3e} element-updates/s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
