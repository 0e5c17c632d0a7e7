"""Time-history file format (openradioss_amd.th) against the reference's own files and encoder.

Fixtures: tests/golden/LOI14T01.th and DEUXELCT01.th are time-history files of the reference's QA
runs (copied by tests/golden/make_th_fixtures.py); the encoder is compared with the reference's
real_to_IEEE_ASCII compiled from engine/source/output/tools/ieee.cpp (oracle/_ref/libref_ieee.so),
and the reader with every time-history file under qa-tests when the reference tree is present.
"""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from openradioss_amd import th as TH

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_IEEE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_ieee.so")


def _records(path):
    b = open(path, "rb").read()
    out, p = [], 0
    while p < len(b):
        n = int.from_bytes(b[p:p + 4], "big")
        out.append(b[p:p + 8 + n])
        p += 8 + n
    return out


def _loi14_header(nglob=12):
    # LOI14_0000.rad: /PART/1..3 (title no_named_mat_1, property 2, 4, 6, material 1), /MAT/LAW14/1,
    # /PROP/TYPE6/2, 4, 6, /TH/NODE/1 th_nodes DEF on nodes 5, 14, 23, /TH/BRIC/6 th_3Dsolid DEF
    # on bricks 1, 2, 3 named b1, b2, b3 (nodes and bricks are their own internal numbers minus one)
    groups = [TH.ThGroup.make("NODE", 1, "th_nodes", ["DEF"], [4, 13, 22], [5, 14, 23]),
              TH.ThGroup.make("BRIC", 6, "th_3Dsolid", ["DEF"], [0, 1, 2], [1, 2, 3], ["b1", "b2", "b3"])]
    return TH.ThHeader("TEST 1 ELEMENT 3D 8 PT INT",
                       [(1, "no_named_mat_1", 1, 1), (2, "no_named_mat_1", 1, 2), (3, "no_named_mat_1", 1, 3)],
                       [(1, "test des directions d_orthotropie")],
                       [(2, "no_named_mat_1"), (4, "no_named_mat_1"), (6, "no_named_mat_1")], groups,
                       nglob=nglob, date="Wed Nov 23 16:21:42 2022")


def test_header_equals_reference_file():
    ref = _records(os.path.join(GOLD, "LOI14T01.th"))
    ours = _records_of(TH.header_bytes(_loi14_header()))
    assert len(ours) == TH.read(os.path.join(GOLD, "LOI14T01.th"))["header_records"]
    for i, (a, b) in enumerate(zip(ours, ref)):
        if i == 1:  # date, program and machine line
            assert a[4:28] == b[4:28] and len(a) == len(b)
            continue
        assert a == b, f"header record {i}"


def _records_of(buf):
    out, p = [], 0
    while p < len(buf):
        n = int.from_bytes(buf[p:p + 4], "big")
        out.append(buf[p:p + 8 + n])
        p += 8 + n
    return out


def test_time_blocks_equal_reference_file():
    path = os.path.join(GOLD, "LOI14T01.th")
    d = TH.read(path)
    body = b"".join(_records(path)[d["header_records"]:])
    ours = b"".join(TH.time_bytes(t["time"], t["glob"], t["groups"]) for t in d["times"])
    assert len(d["times"]) == 48
    assert ours == body


def test_shell_groups_equal_reference_file():
    path = os.path.join(GOLD, "DEUXELCT01.th")
    d = TH.read(path)
    assert [g["codes"].tolist() for g in d["groups"]] == [[1, 2, 3, 6, 7, 8, 9, 10, 11, 13, 14]] * 2
    h = TH.ThHeader("x", [], [], [], [TH.ThGroup.make("SHEL", 10, "spec", ["DEF"], [1], [1000001], ["compr"]),
                                      TH.ThGroup.make("SHEL", 11, "spect", ["DEF"], [0], [1000000], ["tens"])],
                    nglob=12, date="x" * 24)
    ours = _records_of(TH.header_bytes(h))[-6:]
    ref = _records(path)[16:22]
    assert ours == ref
    # the group values of every output time re-encode to the file's bytes
    recs = _records(path)
    k = d["header_records"]
    for t in d["times"]:
        k += 3  # time, globals, part variables (the subsets have none)
        for v in t["groups"]:
            assert recs[k][4:-4] == TH.encode_reals(v)
            k += 1


def test_structure_of_reference_files():
    d = TH.read(os.path.join(GOLD, "LOI14T01.th"))
    assert d["icode"] == 3040 and d["nglob"] == 12
    assert [m[0] for m in d["materials"]] == [1, 0] and d["materials"][1][1] == "no_title"
    assert d["subsets"][0]["title"] == "GLOBAL MODEL" and d["subsets"][0]["parts"].tolist() == [1, 2, 3]
    g = d["groups"]
    assert [(x["id"], x["ity"]) for x in g] == [(6, 1), (1, 0)]  # bricks before nodes
    assert g[0]["codes"].tolist() == list(TH.var_codes("BRIC", ["DEF"]))
    assert g[1]["codes"].tolist() == list(TH.var_codes("NODE", ["DEF"]))
    t = [r["time"] for r in d["times"]]
    assert t[0] == 0.0 and all(b > a for a, b in zip(t, t[1:]))
    e = TH.read(os.path.join(GOLD, "DEUXELCT01.th"))
    assert [p["codes"].tolist() for p in e["parts"]] == [[1, 2, 3, 4, 5, 6, 7, 28]] * 2
    assert e["times"][1]["parts"].shape == (16,)


def test_var_codes():
    assert TH.var_codes("NODE", ["DEF", "XYZ"]) == (1, 2, 3, 4, 5, 6, 16, 17, 18)
    assert TH.var_codes("NODE", ["A", "VX"]) == (4, 7, 8, 9)
    assert TH.var_codes("BRIC", ["QVIS", "VOL", "PLSR"]) == (10, 11, 14)
    assert TH.var_codes("SHEL", ["PLAS", "THIC"]) == (12, 13, 14)
    with pytest.raises(TH.ThError):
        TH.var_codes("NODE", ["TEMP"])          # needs the thermal solver
    with pytest.raises(TH.ThError):
        TH.var_codes("BRIC", ["DAM1"])          # a variable of other laws
    with pytest.raises(TH.ThError):
        TH.var_codes("SHEL", ["NOPE"])


def test_next_time():
    # written at TT >= THIS; the next time is max(TT, THIS + DTHIS), capped at TSTOP
    assert TH.next_time(0.0, 0.01, 0.0, 1.0) == 0.01
    assert TH.next_time(0.01, 0.01, 0.084, 1.0) == 0.084   # a time step longer than DTHIS
    assert TH.next_time(0.99, 0.05, 0.995, 1.0) == 1.0


@pytest.mark.skipif(not os.path.exists(REF_IEEE), reason="oracle/_ref/libref_ieee.so not built")
def test_encoder_equals_reference_encoder():
    lib = C.CDLL(REF_IEEE)
    lib.real_to_IEEE_ASCII.argtypes = [C.c_float, C.c_char_p]
    lib.real_to_IEEE_ASCII.restype = None
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2 ** 32, 60000, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    f = f[np.isfinite(f)]
    special = np.array([0.0, -0.0, 1.0, -1.0, 1.17e-38, -1.17e-38, 1.1754942e-38, -1.1754942e-38, 1.1e-38,
                        -1.1e-38, 1e-45, -1e-45, 3.4e38, -3.4e38, 3.4028235e38, -3.4028235e38, 3.39e38],
                       dtype=np.float32)
    f = np.concatenate([f, special, (rng.standard_normal(20000) * 10.0 ** rng.integers(-40, 39, 20000))
                        .astype(np.float32)])
    out = C.create_string_buffer(4)
    ref = bytearray()
    for x in f:
        lib.real_to_IEEE_ASCII(C.c_float(float(x)), out)
        ref += out.raw
    assert TH.encode_reals(f.astype(np.float64)) == bytes(ref)
    # REAL*4 = double: round to nearest before encoding
    dd = rng.standard_normal(5000) * 10.0 ** rng.integers(-30, 30, 5000)
    ref = bytearray()
    for x in dd.astype(np.float32):
        lib.real_to_IEEE_ASCII(C.c_float(float(x)), out)
        ref += out.raw
    assert TH.encode_reals(dd) == bytes(ref)


QA = sorted(glob.glob("/root/reference/qa-tests/**/*T01", recursive=True))


@pytest.mark.skipif(not QA, reason="reference QA tree absent")
def test_every_reference_file_parses_and_reencodes():
    n_vals = 0
    for path in QA:
        d = TH.read(path)
        recs = _records(path)[d["header_records"]:]
        k = 0
        for t in d["times"]:
            arrays = [np.array([t["time"]], np.float32)]
            if d["nglob"]:
                arrays.append(t["glob"])
            for a in ("parts", "subsets"):
                if t[a].size:
                    arrays.append(t[a])
            arrays += [g for g in t["groups"] if g.size]
            for a in arrays:
                raw = recs[k][4:-4]
                enc = TH.encode_reals(np.asarray(a, np.float32).astype(np.float64))
                if raw != enc:
                    # the writer's -0 stands for a negative value below 1.17e-38, which reads back as -0.0
                    r, e = np.frombuffer(raw, ">u4"), np.frombuffer(enc, ">u4")
                    assert np.all((r == e) | ((r == 0x80000000) & (e == 0))), path
                n_vals += a.size
                k += 1
        assert k == len(recs), path
    assert n_vals > 3_000_000
