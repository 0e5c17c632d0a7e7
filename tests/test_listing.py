"""The listing cycle line (ecrit.F FORMAT 1100) against the Engine listings kept by the reference's
QA tests (tests/golden/listing_lines.txt, made by tests/golden/make_listing_lines.py): every line
parsed by column and printed again from its fields gives the same text, which pins the I8, G11.4,
F5.1 and A5 editing (F and E forms of G, the dropped E of three-digit exponents, negative zero
errors, the 4 trailing blanks)."""
import os

import pytest

from openradioss_amd import listing as LS

GOLD = os.path.join(os.path.dirname(__file__), "golden", "listing_lines.txt")


def _lines():
    return [s.rstrip("\n") for s in open(GOLD) if not s.startswith("#")]


def test_reference_lines_print_back_identically():
    lines = _lines()
    assert len(lines) > 1000
    bad = []
    for s in lines:
        f = LS.parse_cycle_line(s)
        assert f is not None, s
        ity = LS.ELTYP.index(f["eltyp"])
        t = LS.cycle_line(f["ncycle"], f["tt"], f["dt2"], ity, f["nelts"], f["err1"], f["enint"], f["encin"],
                          f["enrot"], f["wfext"], f["emass"], f["xmass"], f["madd"])
        if t != s:
            bad.append((s, t))
    assert not bad, bad[:5]


def test_lines_read_like_verify_results():
    """qa_tools/verify_results.py:48-56: 13 blank-separated fields, the sixth ending in %."""
    s = LS.cycle_line(1000, 1.234567e-3, 1.1111e-6, LS.ITY_SOLID, 32000, -0.04, 1254.3, 59.26, 0.0, 1313.5, 0.0,
                      0.192e-3, 0.0)
    fields = " ".join(s.split()).split(" ")
    assert len(fields) == 13 and fields[5].endswith("%")
    assert fields[:6] == ["1000", "0.1235E-02", "0.1111E-05", "SOLID", "32000", "-0.0%"]
    assert fields[6:9] == ["1254.", "59.26", "0.000"]


@pytest.mark.parametrize("x,text", [
    (0.0, "  0.000    "), (1.0, "  1.000    "), (9999.4, "  9999.    "), (9999.6, " 0.1000E+05"),
    (0.09999, " 0.9999E-01"), (0.099996, " 0.1000    "), (-0.5, "-0.5000    "), (1.0e-100, " 0.1000E-99"),
    (1.0e-101, " 0.1000-100"), (-1.5e20, "-0.1500E+21")])
def test_g_editing_closed_cases(x, text):
    assert LS.fmt_g(x) == text


def test_energy_balance_error():
    """ERR = ENTOT / (ENTOT0 + WFEXT) - 1 in percent (ecrit.F:508-512): a run whose total energy
    grows with the external work reports 0%, one that loses 1% of it reports -1.0%."""
    L = LS.Listing(xmass=2.0)
    L.add(0, 0.0, 1e-6, 1.0, 10.0, 0.0, 0.0, LS.ITY_SOLID, 0)
    s = L.add(100, 1e-4, 1e-6, 6.0, 10.0, 0.0, 5.0, LS.ITY_SOLID, 4)
    assert " 0.0%" in s and s.split()[4] == "5"
    s = L.add(200, 2e-4, 1e-6, 6.0, 9.84, 0.0, 5.0, LS.ITY_SHELL, 9)
    assert "-1.0%" in s and "SHELL" in s
