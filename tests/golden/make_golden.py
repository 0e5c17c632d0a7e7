#!/usr/bin/env python3
"""Write golden vectors of the hex path from the reference's own compiled routines.

Run in the build container (needs oracle/_ref/libref.so, built by
oracle/build_ref.py from /root/reference).  Each fixture holds the inputs that
define the case (generator arguments, so the model is rebuilt bit-identically
by openradioss_amd.model) and the reference outputs after N cycles.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refchain as R  # noqa: E402
sys.path.insert(0, os.path.dirname(HERE))
from golden_cases import CASES, build_case  # noqa: E402

def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libref.so missing: run oracle/build_ref.py first")
    for name, c in CASES.items():
        m = build_case(c)
        r = R.RefHex(m)
        r.cycle(c["cycles"])
        assert r.mstop == 0
        np.savez_compressed(os.path.join(HERE, name + ".npz"), x=r.x, v=r.v, d=r.d, sig=r.sig, eint=r.eint,
                            rho=r.rho, pla=r.pla, epsd=r.epsd, offg=r.offg, qvis=r.qvis,
                            time=np.array([r.tt, r.dt1, r.dt2, r.dt12, r.dt2old]),
                            ncycle=np.array([r.ncycle]))
        print(name, m.ne, "elements", c["cycles"], "cycles, tt", r.tt)


if __name__ == "__main__":
    main()
