"""The Engine listing's cycle line (ecrit.F: FORMAT 1000 header, FORMAT 1100 line) and the
Fortran edit descriptors it uses, written the way a Fortran run prints them.

FORMAT 1100 (ecrit.F:704) is ``(I8,2(1X,G11.4),1X,A5,1X,I10,1X,F5.1,1H%,7(1X,G11.4))`` with the
fields NCYCLE, TT, DT2, ELTYP(ITYPTS), NELTS, ERR1, ENINT, ENCIN, ENROT, WFEXT, EMASS, XMASS and
XMASS-MASS0-DM_INOUT (ecrit.F:651); qa_tools/verify_results.py reads the last such line of a run
(13 blank-separated fields, the sixth ending in ``%``).

G editing (Fortran 2008, 13.7.5): a value whose magnitude, rounded to d significant digits, lies
in [0.1, 10**d) is written in F(w-4).(d-k) editing followed by 4 blanks, with k its number of
digits before the point; other values, in E editing w.d; zero as F(w-4).(d-1).  E editing writes
0.ddddE+ee, dropping the E when the exponent needs three digits.
"""
from __future__ import annotations

from decimal import ROUND_HALF_EVEN, Decimal

# ELTYP(0:12) of ecrit.F:129-131 (the types this path's elements report)
ELTYP = ("FIXED", "SOLID", "QUAD ", "SHELL", "TRUSS", "BEAM ", "SPRIN", "SH_3N", "TRIA ", "AIRBA", "INTER",
         "NODE ", "BLAST")
ITY_SOLID, ITY_SHELL = 1, 3

HEADER = ("   CYCLE    TIME      TIME-STEP  ELEMENT          ERROR  I-ENERGY    K-ENERGY T  K-ENERGY R  "
          "EXT-WORK     MAS.ERR     TOTAL MASS  MASS ADDED")


def _round_sig(x: float, d: int) -> tuple[int, Decimal]:
    """(sign, |x| rounded to d significant digits) for x != 0."""
    v = Decimal(x)  # exact binary value
    sign = -1 if v < 0 else 1
    v = abs(v)
    e = v.adjusted()  # 10**e <= v < 10**(e+1)
    q = v.quantize(Decimal(1).scaleb(e - d + 1), rounding=ROUND_HALF_EVEN)
    return sign, q


def fmt_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d."""
    s = f"{x:.{d}f}" + ("." if d == 0 else "")
    if s.startswith("0.") and len(s) > w:
        s = s[1:]
    elif s.startswith("-0.") and len(s) > w:
        s = "-" + s[2:]
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def fmt_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d with one digit-free leading zero: [-]0.ddddE+ee (E dropped for |ee| > 99)."""
    if x == 0.0:
        body, ex = "0." + "0" * d, 0
        neg = str(x).startswith("-")
    else:
        sign, q = _round_sig(x, d)
        neg = sign < 0
        ex = q.adjusted() + 1                      # 0.1 <= q / 10**ex < 1
        digits = str(q.scaleb(-ex + d).to_integral_value()).rjust(d, "0")
        body = "0." + digits
    if -99 <= ex <= 99:
        tail = f"E{'-' if ex < 0 else '+'}{abs(ex):02d}"
    else:
        tail = f"{'-' if ex < 0 else '+'}{abs(ex):03d}"
    s = ("-" if neg else "") + body + tail
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def fmt_g(x: float, w: int = 11, d: int = 4) -> str:
    """Fortran Gw.d for a real."""
    if x == 0.0:
        return fmt_f(x, w - 4, d - 1) + "    "
    sign, q = _round_sig(x, d)
    if Decimal("0.1") <= q < Decimal(10) ** d:
        k = q.adjusted() + 1  # digits before the point (0 for 0.1 <= q < 1)
        return fmt_f(float(sign * q), w - 4, d - k) + "    "
    return fmt_e(x, w, d)


def fmt_i(n: int, w: int) -> str:
    s = str(int(n))
    return "*" * w if len(s) > w else s.rjust(w)


def cycle_line(ncycle: int, tt: float, dt2: float, ityp: int, nelts: int, err1: float, enint: float,
               encin: float, enrot: float, wfext: float, emass: float, xmass: float, madd: float) -> str:
    """One line of FORMAT 1100 (ecrit.F:651, 704)."""
    out = fmt_i(ncycle, 8) + " " + fmt_g(tt) + " " + fmt_g(dt2) + " " + ELTYP[ityp] + " " + fmt_i(nelts, 10)
    out += " " + fmt_f(err1, 5, 1) + "%"
    for v in (enint, encin, enrot, wfext, emass, xmass, madd):
        out += " " + fmt_g(v)
    return out


def parse_cycle_line(line: str) -> dict | None:
    """The fields of a FORMAT 1100 line, by column, or None for any other line."""
    s = line.rstrip("\n")
    if len(s) < 8 + 12 * 2 + 6 + 11 + 7 + 12 * 7 - 4 or s[8 + 24 + 6 + 11 + 6] != "%":
        return None
    try:
        p = 0
        f = {}
        f["ncycle"] = int(s[p:p + 8]); p += 8
        f["tt"] = float(s[p + 1:p + 12]); p += 12
        f["dt2"] = float(s[p + 1:p + 12]); p += 12
        f["eltyp"] = s[p + 1:p + 6]; p += 6
        f["nelts"] = int(s[p + 1:p + 11]); p += 11
        f["err1"] = float(s[p + 1:p + 6]); p += 7
        for k in ("enint", "encin", "enrot", "wfext", "emass", "xmass", "madd"):
            f[k] = _fortran_float(s[p + 1:p + 12]); p += 12
    except ValueError:
        return None
    return f


def _fortran_float(t: str) -> float:
    t = t.strip()
    if "E" not in t and ("+" in t[1:] or "-" in t[1:]):  # 0.1234-100: exponent without its E
        i = max(t.rfind("+"), t.rfind("-"))
        t = t[:i] + "E" + t[i:]
    return float(t)


class Listing:
    """The cycle lines of one run, built from what the resident model reports at its output
    cycles.  ECRIT's energy balance: ENTOT = ENCIN + ENINT + ENROT; at cycle 0 ENTOT0 = ENTOT -
    WFEXT and MASS0 = XMASS; ERR = ENTOT / (ENTOT0 + WFEXT) - 1 (ecrit.F:501-516), printed in
    percent and clamped to +-99.9; EMASS = (XMASS - MASS0) / MASS0 (ecrit.F:517)."""

    def __init__(self, xmass: float, user_ids_hex=None, user_ids_q4=None):
        self.xmass = float(xmass)
        self.mass0 = None
        self.entot0 = None
        self.ids_h = user_ids_hex
        self.ids_q = user_ids_q4
        self.lines = [HEADER]

    def add(self, ncycle: int, tt: float, dt2: float, enint: float, encin: float, enrot: float, wfext: float,
            ityp: int, elem: int) -> str:
        """Append the line of an output cycle; elem is the controlling element's index in its family
        (its user id is taken from the family's id table when one was given, else index + 1)."""
        entot = encin + enint + enrot
        if self.entot0 is None:
            self.entot0 = entot - wfext
            self.mass0 = self.xmass
        entot1 = self.entot0 + wfext
        err1 = 0.0
        if abs(entot1) > 1.0e-20:
            err = entot / entot1 - 1.0
            err1 = max(-99.9, min(99.9, err * 100.0))
        emass = (self.xmass - self.mass0) / max(self.mass0, 1.0e-20)
        if ityp == ITY_SOLID:
            nelts = int(self.ids_h[elem]) if self.ids_h is not None else elem + 1
        elif ityp == ITY_SHELL:
            nelts = int(self.ids_q[elem]) if self.ids_q is not None else elem + 1
        else:
            nelts = 0
        line = cycle_line(ncycle, tt, dt2, ityp, nelts, err1, enint, encin, enrot, wfext, emass, self.xmass,
                          self.xmass - self.mass0)
        self.lines.append(line)
        return line

    def terminate(self, ncycle: int, mstop: int = 0) -> None:
        """The end of the listing (arret.F:762-768): NORMAL TERMINATION and the cycle count, or
        ERROR TERMINATION for a run stopped by MSTOP (a non-positive Jacobian or time step)."""
        if mstop:
            self.lines += ["", "", "", "     ERROR TERMINATION"]
        else:
            self.lines += ["", "", "", "     NORMAL TERMINATION      ", "", "", "",
                           "     TOTAL NUMBER OF CYCLES  :" + fmt_i(ncycle, 8), ""]

    def text(self) -> str:
        return "\n".join(self.lines) + "\n"
