#!/usr/bin/env python3
"""Golden fixtures of the two reference workloads at their full length, from the reference's
own compiled routines (oracle/_ref, built by oracle/build_ref.py from /root/reference).

* bar_20x20x80_c1000.npz: the 20x20x80 cantilever bar (32,000 hexes, law-2 steel, one end
  fixed, 10 m/s along the axis) after 1,000 cycles: D in full; V and X as SHA-256 digests of
  their bytes plus V at every 7th node; the time scalars.
* tube_60x300_c8800.npz: the 60x60x300 mm square tube (72,000 Q4 shells, 5 points, base fixed,
  top at 15 m/s) after 8,800 cycles, before its first fold: digests of D, V, VR and X, D at
  every 4th node, the time scalars.

Both hold `energy` = [internal, translational kinetic, rotational kinetic] of the final state:
the internal energy as the reference's compiled SBILAN / CBILAN add it (oracle/ref_bilan.F),
the kinetic sums 0.5*MS*|V|^2 and 0.5*IN*|VR|^2 of the reference's velocities, node after node.

Run in the build container:  python tests/golden/make_golden_full.py [bar] [tube]
(about 1 minute for the bar, 20 for the tube).
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refchain as R  # noqa: E402
from openradioss_amd import model as M  # noqa: E402

FULL = {
    "bar_20x20x80_c1000": dict(cycles=1000, build=lambda: M.hex_model_to_model(M.cantilever_bar(20, 20, 80)),
                               d_every=1, v_every=7),
    "tube_60x300_c8800": dict(cycles=8800, build=lambda: M.square_tube(), d_every=4, v_every=0),
}


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libref.so missing: run oracle/build_ref.py first")
    only = set(sys.argv[1:])
    for name, c in FULL.items():
        if only and not any(k in name for k in only):
            continue
        m = c["build"]()
        r = R.RefModel(m)
        t0 = time.time()
        for k in range(0, c["cycles"], 100):
            r.cycle(min(100, c["cycles"] - k))
            print(f"{name}: cycle {r.ncycle}, {time.time() - t0:.0f} s", flush=True)
        assert r.mstop == 0
        out = dict(d_sub=r.d[::c["d_every"]], d_every=np.array([c["d_every"]]),
                   sha_d=np.array(digest(r.d)), sha_v=np.array(digest(r.v)), sha_x=np.array(digest(r.x)),
                   time=np.array([r.tt, r.dt1, r.dt2, r.dt12, r.dt2old]), ncycle=np.array([r.ncycle]))
        if c["v_every"]:
            out.update(v_sub=r.v[::c["v_every"]], v_every=np.array([c["v_every"]]))
        if m.q4 is not None:
            out.update(sha_vr=np.array(digest(r.vr)))
        kw = dict(v=r.v, x=r.x, vr=r.vr)
        if m.hexp is not None:
            kw.update(heint=r.heint, hrho=r.hrho, hoffg=r.hoffg)
        if m.q4 is not None:
            kw.update(q_eint=r.q["eint"], q_offg=r.q["offg"])
        eh, eq = R.bilan_internal_energy(m, **kw)
        ke = 0.0
        kr = 0.0
        for n in range(m.nn):
            ke = ke + 0.5 * float(m.ms[n]) * float(r.v[n] @ r.v[n])
            kr = kr + 0.5 * float(m.inr[n]) * float(r.vr[n] @ r.vr[n])
        out.update(energy=np.array([eh + eq, ke, kr]))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "written, tt", r.tt)


if __name__ == "__main__":
    main()
