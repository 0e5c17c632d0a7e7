"""The time history of a model run on the oracle, written with openradioss_amd.th: the checker of
GpuModel.run_time_history (same output times, same records)."""
import numpy as np

import oracle as O
from openradioss_amd import th as TH


def oracle_record(o, m, header):
    """One output cycle of the oracle as the record GpuModel.run_th returns."""
    nodes, hexes, shells = TH.group_indices(header)
    r = o.cycle_th()
    r["nodes"] = r["nodes"][nodes]
    if len(hexes):
        r["hex"] = np.stack([o.hoffg[hexes], *o.hsig[:, hexes], o.heint[hexes], o.hrho[hexes], o.hqvis[hexes],
                             o.hvol0[hexes], o.hpla[hexes], o.hepsd[hexes]], axis=1)
    if len(shells):
        q = o.q
        r["q4"] = np.concatenate([q["forc"][:, shells].T, q["mom"][:, shells].T, q["eint"][:, shells].T,
                                  q["offg"][shells, None], q["thk"][shells, None], q["pla"][:, shells].T], axis=1)
    return r


def oracle_time_history(m, header, dthis, tstop, max_cycles, nthreads=16, math="cr"):
    """The time-history file of the oracle run (bytes), the records and the next output time."""
    o = O.ModelOracle(m, nthreads=nthreads, math=math)
    o.track_ehour()
    xmass = float(np.cumsum(np.asarray(m.ms, np.float64))[-1])
    out, recs, this, done = [TH.header_bytes(header)], [], 0.0, 0
    while True:
        tt = o.ts.tt
        if tt >= tstop or o.ts.mstop or done >= max_cycles:
            break
        if tt >= this:
            r = oracle_record(o, m, header)
            out.append(TH.time_bytes(r["tt"], TH.global_values(r, xmass), TH.group_values(header, m, r)))
            recs.append(r)
            this = TH.next_time(this, dthis, r["tt"], tstop)
        else:
            o.cycle(1)
        done += 1
    return b"".join(out), recs, this
