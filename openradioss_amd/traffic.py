"""Algorithmic HBM traffic of the path, implemented once (used by bench.py and DESIGN.md).

Word counts per element-update follow the compulsory-traffic model of the
explicit hex cycle: element state read once and written once, connectivity
read once, node data gathered once per node per cycle, and the nodal update
reading and writing each node once.  The deterministic two-phase assembly
adds the skyline rows (written by the element kernel, read by the node
kernel); it is reported separately as the "as implemented" extra.
"""
from __future__ import annotations

HEX_STATE_READ = 13   # SIG(6), EINT, RHO, VOL0, PLA, EPSD, OFF, QVIS
HEX_STATE_WRITE = 12  # the same minus VOL0
HEX_CONN_BYTES = 8 * 4
HEX_IADS_BYTES = 8 * 4
NODE_GATHER_BYTES = 6 * 8    # X(3), V(3) in FP64
NODE_UPDATE_READ = 6 * 8 + 3 * 8 + 8 + 4 + 4  # X,V + D + MS + BC code + ADSKY
NODE_UPDATE_WRITE = 6 * 8 + 3 * 8             # X,V + D


def state_word(precision: str) -> int:
    return 8 if precision == "fp64" else 4


def row_bytes(precision: str) -> int:
    return 4 * state_word(precision)  # (FX, FY, FZ, STI/4)


def hex_element_kernel_bytes(ne: int, nn: int, precision: str) -> dict:
    """Bytes one launch of the element kernel must move (per launch, for ne elements)."""
    w = state_word(precision)
    compulsory = ne * ((HEX_STATE_READ + HEX_STATE_WRITE) * w + HEX_CONN_BYTES) + nn * NODE_GATHER_BYTES
    skyline = ne * (HEX_IADS_BYTES + 8 * row_bytes(precision))
    return {"compulsory": compulsory, "as_implemented": compulsory + skyline}


def node_kernel_bytes(ne: int, nn: int, precision: str) -> dict:
    compulsory = nn * (NODE_UPDATE_READ + NODE_UPDATE_WRITE)
    skyline = 8 * ne * row_bytes(precision)
    return {"compulsory": compulsory, "as_implemented": compulsory + skyline}


def cycle_bytes_per_element_update(ne: int, nn: int, precision: str) -> dict:
    e = hex_element_kernel_bytes(ne, nn, precision)
    n = node_kernel_bytes(ne, nn, precision)
    return {k: (e[k] + n[k]) / ne for k in ("compulsory", "as_implemented")}
