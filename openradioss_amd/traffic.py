"""Algorithmic HBM traffic of the path, implemented once (used by bench.py).

Word counts per element-update follow the compulsory-traffic model of the
explicit cycle: element state read once and written once, connectivity read
once, node data gathered once per node per cycle, and the nodal update reading
and writing each node once.  The deterministic assembly adds, as implemented:

* per element corner, the index of its stored ("rest") row;
* per node and family, one FP64 prefix written by the element tile that sums the
  node's leading rows in LDS and read by the node kernel, with its list entry
  (node id and eight LDS sources);
* the rest rows themselves (written by the element kernels, read by the node
  kernel: 3 force words per hex row, 3 force + 3 moment words per shell row);
* per node and family, the row count and the first rest row.

"as_implemented" includes them; with no assembly plan at hand (`plan=None`) every
row is counted as a stored row, the bound of an assembly without prefixes.
"""
from __future__ import annotations

HEX_STATE_READ = 13   # SIG(6), EINT, RHO, VOL0, PLA, EPSD, OFF, QVIS
HEX_STATE_WRITE = 12  # the same minus VOL0
HEX_CONN_BYTES = 8 * 4
HEX_REST_IDX_BYTES = 8 * 4

# Q4: GBUF OFF, THK, THKE, FOR(5), MOM(3), STRA(8), HOURG(5), EINT(2), EPSD + per point SIG(5), PLA
Q4_GBUF_READ = 1 + 1 + 1 + 5 + 3 + 8 + 5 + 2 + 1
Q4_GBUF_WRITE = Q4_GBUF_READ - 1      # THKE is read only
Q4_LBUF_WORDS = 5 + 1                 # per point, read and written
Q4_CONN_BYTES = 4 * 4
Q4_REST_IDX_BYTES = 4 * 4

NODE_GATHER_BYTES = 6 * 8      # X(3), V(3) in FP64
NODE_GATHER_ROT_BYTES = 3 * 8  # VR(3)
HEX_ROW_WORDS = 3              # fx, fy, fz
Q4_ROW_WORDS = 6               # fx, fy, fz, mx, my, mz
PREFIX_LIST_BYTES = 4 + 16     # node id + 8 LDS sources of 16 bits


def state_word(precision: str) -> int:
    return 8 if precision == "fp64" else 4


def hex_element_kernel_bytes(ne: int, nn: int, precision: str, prefixes: int | None = None,
                             rest_rows: int | None = None) -> dict:
    """Bytes one launch of the hex element kernel must move (ne elements, nn nodes gathered),
    and as implemented with the given plan (prefixes, rest rows; None: every row stored)."""
    w = state_word(precision)
    compulsory = ne * ((HEX_STATE_READ + HEX_STATE_WRITE) * w + HEX_CONN_BYTES) + nn * NODE_GATHER_BYTES
    if rest_rows is None:
        prefixes, rest_rows = 0, 8 * ne
    asm = ne * HEX_REST_IDX_BYTES + rest_rows * HEX_ROW_WORDS * w + prefixes * (HEX_ROW_WORDS * 8 + PREFIX_LIST_BYTES)
    return {"compulsory": compulsory, "as_implemented": compulsory + asm}


def q4_element_kernel_bytes(ne: int, nn: int, npt: int, precision: str, prefixes: int | None = None,
                            rest_rows: int | None = None) -> dict:
    """Bytes one launch of the Q4 element kernel must move (ne elements, nn nodes gathered)."""
    w = state_word(precision)
    words = Q4_GBUF_READ + Q4_GBUF_WRITE + 2 * npt * Q4_LBUF_WORDS
    compulsory = ne * (words * w + Q4_CONN_BYTES) + nn * (NODE_GATHER_BYTES + NODE_GATHER_ROT_BYTES)
    if rest_rows is None:
        prefixes, rest_rows = 0, 4 * ne
    asm = ne * Q4_REST_IDX_BYTES + rest_rows * Q4_ROW_WORDS * w + prefixes * (Q4_ROW_WORDS * 8 + PREFIX_LIST_BYTES)
    return {"compulsory": compulsory, "as_implemented": compulsory + asm}


def node_kernel_bytes(nn: int, nn_rot: int, precision: str, fams: int, hex_prefixes: int, hex_rest: int,
                      q4_prefixes: int, q4_rest: int) -> dict:
    """Bytes of one launch of the node kernel: X, V, D read and written, MS and the node code
    read; for the nn_rot nodes with rotational work (shell rows or rotational BCs) also VR read
    and written and IN read.  As implemented: the prefixes and rest rows, per node and family
    the row count and the first rest row, and VR padded to 4 words."""
    compulsory = nn * ((6 + 3) * 8 * 2 + 8 + 4) + nn_rot * (3 * 8 * 2 + 8)
    w = state_word(precision)
    asm = (hex_prefixes * HEX_ROW_WORDS * 8 + hex_rest * HEX_ROW_WORDS * w + q4_prefixes * Q4_ROW_WORDS * 8 +
           q4_rest * Q4_ROW_WORDS * w + nn * fams * (4 + 4) + nn_rot * 2 * 8 * 2)
    return {"compulsory": compulsory, "as_implemented": compulsory + asm}


def cycle_bytes(model, precision: str, plan: dict | None = None) -> dict:
    """Per-cycle bytes of every kernel of a Model, and the total.  plan: the assembly plan of the
    uploaded model (GpuModel.assembly_stats()), or None to count every row as stored."""
    import numpy as np
    out = {}
    nn = model.nn
    p = plan or {}
    hp, hr = p.get("hex_prefixes"), p.get("hex_rest_rows")
    qp, qr = p.get("q4_prefixes"), p.get("q4_rest_rows")
    if model.hexp is not None:
        nh = nn_touched(model.hexp.conn, nn)
        out["hex"] = hex_element_kernel_bytes(model.ne_hex, nh, precision, hp, hr)
    if model.q4 is not None:
        nq = nn_touched(model.q4.conn, nn)
        out["q4"] = q4_element_kernel_bytes(model.ne_q4, nq, model.q4.prop.npt, precision, qp, qr)
    nn_rot = 0
    if model.q4 is not None:
        rot = np.zeros(nn, dtype=bool)
        rot[model.q4.conn.reshape(-1)] = True
        rot |= (np.asarray(model.code) & (32 | 16 | 8)) != 0
        nn_rot = int(rot.sum())
    fams = (model.hexp is not None) + (model.q4 is not None)
    out["node"] = node_kernel_bytes(nn, nn_rot, precision, fams,
                                    hp or 0, 8 * model.ne_hex if hr is None else hr,
                                    qp or 0, 4 * model.ne_q4 if qr is None else qr)
    out["total"] = {k: sum(v[k] for v in out.values()) for k in ("compulsory", "as_implemented")}
    return out


def nn_touched(conn, nn: int) -> int:
    import numpy as np
    seen = np.zeros(nn, dtype=bool)
    seen[conn.reshape(-1)] = True
    return int(seen.sum())


def hex_bytes_per_element_update(ne: int, nn: int, precision: str) -> dict:
    e = hex_element_kernel_bytes(ne, nn, precision)
    n = node_kernel_bytes(nn, 0, precision, 1, 0, 8 * ne, 0, 0)
    return {k: (e[k] + n[k]) / ne for k in ("compulsory", "as_implemented")}
