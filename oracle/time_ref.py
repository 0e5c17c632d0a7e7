#!/usr/bin/env python3
"""CPU timing of the reference's own element routines beside the C restatement (oracle), on the
same element groups: TEST INFRASTRUCTURE ONLY, run in the build container where oracle/_ref
(the reference's Fortran compiled from /root/reference) exists.

Both run one thread on a 20^3 hex block with a 64^2 Q4 plate tied on its top (the bench
workload's families at 8,000 hexes + 4,096 shells), from the same initial state: the reference's
SFORC3 / CFORC3 chains over all groups of 128 (element phase, into FSKY rows), driven per group
through its Fortran harness (oracle/ref_driver.F, oracle/ref_shell.F) with the group's arrays
staged from Python as in the golden-vector runs; the restatement's whole cycle (element phase
and nodal update) in C.

Usage: python oracle/time_ref.py [seconds per leg]
"""
import ctypes as C
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import refchain as R  # noqa: E402
from openradioss_amd import model as M  # noqa: E402


def main():
    budget = float(sys.argv[1]) if len(sys.argv) > 1 else 5.0
    if not R.available():
        raise SystemExit("oracle/_ref/libref.so missing: run oracle/build_ref.py first")
    m = M.block_with_plate(n=20, plate=64, seed=20261203)
    ne_h, ne_q = m.ne_hex, m.ne_q4
    dt1 = 1.0e-7
    # the reference's chains, group by group
    r = R.RefModel(m)
    n, el = 0, 0.0
    while el < budget:
        dt = C.c_double(1.0e6)
        t0 = time.perf_counter()
        r._hex_forces(dt1, dt, ne_h)
        r._q4_forces(dt1, dt, ne_q)
        el += time.perf_counter() - t0
        n += 1
    ref_rate = (ne_h + ne_q) * n / el
    # the restatement (one thread), whole cycles
    o = O.ModelOracle(m, nthreads=1)
    o.cycle(1)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < budget:
        o.cycle(1)
        k += 1
    port_rate = (m.ne_hex + m.ne_q4) * k / (time.perf_counter() - t0)
    print(f"reference SFORC3 + CFORC3 chains (amdflang -O2, one thread, groups of 128, element phase): "
          f"{ref_rate:.3e} element-updates/s over {n} passes of {ne_h} hexes + {ne_q} shells")
    print(f"C restatement (gcc -O2, one thread, whole cycle incl. nodal update): {port_rate:.3e} element-updates/s "
          f"over {k} cycles of {m.ne_hex + m.ne_q4} elements")
    print(f"cpu: {os.cpu_count()} logical CPUs; model {open('/proc/cpuinfo').read().split('model name')[1].split(':')[1].splitlines()[0].strip()}")


if __name__ == "__main__":
    main()
