"""Starter input decks (`*_0000.rad`, block format) for this path: read one into a `Model`, or
write a `Model` as one.

The reader takes the keywords that describe a model on this path -- nodes, 8-node bricks and
4-node shells, their parts, one law-2 (Johnson-Cook) material and one property per element
family, boundary conditions, initial and imposed velocities -- and builds the model with the
Starter-equivalent builder (`model.build_model`: volumes, SMASS3 / CINMAS masses and inertias,
FILLCNE skyline). Anything that would take the model off this path (another material law or
element formulation, skew frames, time-dependent imposed velocities, thermal softening, ...)
raises `DeckError` with the keyword and the reason, instead of being ignored.

Card layouts are those of the reference's Starter input format (HyperMesh card formats under
hm_cfg_files/config/CFG/radioss2025): 10-character integer fields and 20-character real fields,
blank fields read as zero, lines starting with '#' or '$' are comments.

  /NODE                 id, X, Y, Z                                   (I10, 3 x F20)
  /BRICK/part           id, 8 node ids                                (9 x I10)
  /SHELL/part           id, 4 node ids, phi, thick                    (5 x I10, 2 x F20)
  /PART/id              title; prop, mat, subset                      (3 x I10)
  /MAT/PLAS_JOHNS/id    matl2_plas_johns.cfg; read as hm_read_mat02_jc.F90:107-138
  /PROP/SOLID/id        prop_p14_solid.cfg; defaults hm_read_prop14.F:355-369
  /PROP/SHELL/id        prop_p1_shell.cfg; defaults hm_read_prop01.F:156-212
  /BCS/id               title; '   XYZ XYZ' codes, skew, node group
  /GRNOD/NODE/id        title; node ids (I10 fields)
  /INIVEL/TRA/id        inivel.cfg: Vx, Vy, Vz, node group, skew
  /IMPVEL/id            function, direction, skew, sensor, node group, frame; scales, start, stop
  /FUNCT/id             title; (x, y) pairs

Node and element order: nodes are numbered in the order the deck lists them, and the elements
of each family in the order of their /BRICK and /SHELL blocks. The skyline then follows
FILLCNE's rule (solids before shells, ascending element number). The reference Starter also
renumbers elements into groups by part (and, with several domains, by domain); a deck whose
element blocks are already in part order with one part per family numbers them the same way.
"""
from __future__ import annotations

import dataclasses
import io
import re

import numpy as np

from . import model as M

__all__ = ["DeckError", "Deck", "EngineSettings", "parse", "parse_engine", "read", "read_run", "to_model", "write",
           "write_engine"]


class DeckError(ValueError):
    """A deck that is malformed or that describes a model off this path."""


# ----------------------------------------------------------------------------------- cards
_FORTRAN_EXP = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))([+-]\d+)$")


def _real(s: str) -> float:
    s = s.strip()
    if not s:
        return 0.0
    s = s.replace("D", "E").replace("d", "e")
    m = _FORTRAN_EXP.match(s)          # Fortran list style without the exponent letter: 1.5-3
    if m:
        s = m.group(1) + "e" + m.group(2)
    try:
        return float(s)
    except ValueError:
        raise DeckError(f"not a real number: {s!r}") from None


def _int(s: str) -> int:
    s = s.strip()
    if not s:
        return 0
    try:
        return int(s)
    except ValueError:
        raise DeckError(f"not an integer: {s!r}") from None


def _cut(line: str, widths) -> list[str]:
    out, pos = [], 0
    for w in widths:
        out.append(line[pos:pos + w])
        pos += w
    return out


@dataclasses.dataclass
class _Block:
    keyword: str            # e.g. "/PROP/SHELL/2"
    lines: list             # data lines, comments removed
    lineno: int

    @property
    def parts(self) -> list[str]:
        return [p for p in self.keyword.strip().split("/") if p]

    def need(self, n: int) -> None:
        if len(self.lines) < n:
            raise DeckError(f"{self.keyword} (line {self.lineno}): expected at least {n} card(s), "
                            f"found {len(self.lines)}")

    def ident(self) -> int:
        p = self.parts
        if not p or not p[-1].strip().lstrip("-").isdigit():
            raise DeckError(f"{self.keyword} (line {self.lineno}): keyword needs an id")
        return int(p[-1])


def _blocks(text: str) -> list[_Block]:
    blocks: list[_Block] = []
    cur = None
    for no, raw in enumerate(text.splitlines(), 1):
        line = raw.rstrip("\r\n")
        if line.startswith("#") or line.startswith("$"):
            continue
        if line.startswith("/"):
            kw = line.split()[0].upper()
            if kw == "/END":
                break
            cur = _Block(kw, [], no)
            blocks.append(cur)
        elif cur is not None:
            # a blank line inside a keyword is a card whose fields all take their defaults
            cur.lines.append(line.expandtabs())
    for b in blocks:
        while b.lines and not b.lines[-1].strip():
            b.lines.pop()
    return blocks


def _data(lines) -> list[str]:
    """The non-blank lines of a list keyword (nodes, elements, function points)."""
    return [line for line in lines if line.strip()]


# ----------------------------------------------------------------------------- deck content
@dataclasses.dataclass
class Deck:
    """The keywords of a deck, by id, in the units the deck is written in."""

    node_ids: np.ndarray                      # (nn,) user ids in deck order
    x: np.ndarray                             # (nn,3)
    bricks: list                              # [(part, ids (ne,), conn (ne,8) user node ids)]
    shells: list                              # [(part, ids (ne,), conn (ne,4), thick (ne,))]
    parts: dict                               # part id -> (prop id, mat id)
    mats: dict                                # mat id -> Law2 (or a str naming another law)
    solid_props: dict                         # prop id -> SolidProp
    shell_props: dict                         # prop id -> ShellProp
    grnod: dict                               # group id -> user node ids
    bcs: list                                 # [(code bits, group id)]
    inivel: list                              # [(vx, vy, vz, group id)]
    impvel: list                              # [(direction 0..2, value, group id)]
    titles: dict = dataclasses.field(default_factory=dict)  # ('PART'|'MAT'|'PROP', id) -> title; 'RUN' -> run title
    th: list = dataclasses.field(default_factory=list)      # [(kind, id, title, variables, entity ids, names)]
    th_other: list = dataclasses.field(default_factory=list)  # other /TH/ keywords (not written on this path)


def _law2(b: _Block) -> M.Law2:
    """hm_read_mat02_jc.F90:107-138, with the card layout of matl2_plas_johns.cfg."""
    b.need(5)
    c = b.lines[1:]
    rho, rho_o = (_real(x) for x in _cut(c[0], (20, 20)))
    e, nu, iflag, vp, pmin = _cut(c[1], (20, 20, 10, 10, 20))
    e, nu, iflag, vp, pmin = _real(e), _real(nu), _int(iflag), _int(vp), _real(pmin)
    a, bb, n, epsm, sigm = (_real(s) for s in _cut(c[2], (20,) * 5))
    cc, eps0, icc, _fsmooth, fcut, chard = _cut(c[3], (20, 20, 10, 10, 20, 20))
    cc, eps0, icc, fcut, chard = _real(cc), _real(eps0), _int(icc), _real(fcut), _real(chard)
    m_exp, tmelt = (_real(s) for s in _cut(c[4], (20, 20))) if len(c) > 4 else (0.0, 0.0)
    where = f"{b.keyword} (line {b.lineno})"
    if iflag != 0:
        raise DeckError(f"{where}: Iflag = {iflag} (yield and UTS input) is not read here; give a, b, n")
    if vp == 1:
        raise DeckError(f"{where}: VP = 1 (plastic strain rate, filtered after the return) is not on this path")
    if pmin != 0.0:
        raise DeckError(f"{where}: a pressure cut-off Pmin is not on this path")
    if chard != 0.0:
        raise DeckError(f"{where}: kinematic hardening (Chard) is not on this path")
    if tmelt != 0.0:
        raise DeckError(f"{where}: thermal softening (T_melt) is not on this path")
    if rho <= 0.0 or e <= 0.0:
        raise DeckError(f"{where}: density and Young's modulus must be positive")
    if rho_o != 0.0 and rho_o != rho:
        raise DeckError(f"{where}: a reference density RHO_O different from RHO_I is not on this path")
    return M.Law2(rho=rho, young=e, nu=nu, a=a, b=bb, n=n, epsmax=epsm, sigmax=sigm, c=cc, eps0=eps0,
                  icc=icc, vp=vp, fcut=fcut, m_exp=m_exp)


def _solid_prop(b: _Block) -> M.SolidProp:
    """prop_p14_solid.cfg cards; QA/QB defaults of hm_read_prop14.F:355-369 (Isolid=1)."""
    b.need(3)
    f = _cut(b.lines[1], (10,) * 8 + (20,))
    isolid, ismstr, iale, icpre = (_int(s) for s in f[:4])
    iframe, dn = _int(f[7]), _real(f[8])
    where = f"{b.keyword} (line {b.lineno})"
    if isolid != 1:
        raise DeckError(f"{where}: Isolid = {isolid}; this path runs Isolid = 1 (state it explicitly)")
    if ismstr != 0 or iale != 0 or icpre != 0:
        raise DeckError(f"{where}: Ismstr/Iale/Icpre must keep their defaults on this path")
    if iframe not in (0, 1):
        # Iframe = 2 is the corotational frame (JCVT, hm_read_prop14.F:277-282): not SDEFO3's JCVT=0 branch
        raise DeckError(f"{where}: Iframe = {iframe}; this path runs the global frame (Iframe 0 or 1)")
    if dn != 0.0:
        raise DeckError(f"{where}: numerical damping Dn is not on this path")
    if len(b.lines) > 4 and len(_cut(b.lines[4], (10, 10, 10))) > 2 and _int(_cut(b.lines[4], (10, 10, 10))[2]):
        # Icontrol = 1 changes the hourglass default QH from 0.1 to 1 (hm_read_prop14.F:366-367)
        raise DeckError(f"{where}: Icontrol is not on this path")
    qa, qb, qh, lam, mu = (_real(s) for s in _cut(b.lines[2], (20,) * 5))
    if lam != 0.0 or mu != 0.0:
        raise DeckError(f"{where}: numerical viscosities lambda_v / mu_v are not on this path")
    if len(b.lines) > 3 and _real(_cut(b.lines[3], (20,))[0]) != 0.0:
        raise DeckError(f"{where}: an element time-step floor (dt_min) is not on this path")
    qa = qa if qa != 0.0 else 1.1
    qb = qb if qb != 0.0 else 0.05
    qh = qh if qh != 0.0 else 0.1
    return M.SolidProp(qa=qa, qb=qb, qh=qh)


def _shell_prop(b: _Block) -> M.ShellProp:
    """prop_p1_shell.cfg cards; Hm/Hf/Hr defaults of hm_read_prop01.F:205-212."""
    b.need(4)
    ishell, ismstr, ish3, idrill, ipinch = (_int(s) for s in _cut(b.lines[1], (10,) * 5))
    hm, hf, hr, dm, dn = (_real(s) for s in _cut(b.lines[2], (20,) * 5))
    nip, _, thick, ashear, _, ithick, iplas, ipos = _cut(b.lines[3], (10, 10, 20, 20, 10, 10, 10, 10))
    nip, thick, ashear = _int(nip), _real(thick), _real(ashear)
    ithick, iplas, ipos = _int(ithick), _int(iplas), _int(ipos)
    where = f"{b.keyword} (line {b.lineno})"
    if ishell != 1:
        raise DeckError(f"{where}: Ishell = {ishell}; this path runs Ishell = 1 (state it explicitly)")
    if ismstr not in (0, 2) or idrill != 0 or ipinch != 0 or ish3 != 0:
        raise DeckError(f"{where}: Ismstr/Ish3n/Idrill/Ipinch must keep their defaults on this path")
    if dm != 0.0 or dn != 0.0 or ashear != 0.0 or ipos != 0:
        raise DeckError(f"{where}: Dm, Dn, a shear area and Ipos are not on this path")
    if not 2 <= nip <= 10:
        raise DeckError(f"{where}: N = {nip}; this path integrates with 2..10 points through the thickness")
    if ithick == 0 or iplas == 0:
        raise DeckError(f"{where}: Ithick = 0 or Iplas = 0 takes a /DEF_SHELL default; state the flag")
    if thick <= 0.0:
        raise DeckError(f"{where}: shell thickness must be positive")
    return M.ShellProp(thick=thick, npt=nip, ipla=iplas, ithk=ithick,
                       hm=hm if hm != 0.0 else 0.01, hf=hf if hf != 0.0 else 0.01,
                       hr=hr if hr != 0.0 else 0.01)


def _ids(lines) -> list[int]:
    out = []
    for line in lines:
        out.extend(_int(s) for s in _cut(line, (10,) * 10) if s.strip())
    return out


# Keywords that do not change the model on this path: titles, output requests, run controls, and
# defaults that the properties read here must state explicitly anyway.
_IGNORED = {"TITLE", "ANALY", "IOFLAG", "SPMD", "SUBSET", "SHFRA"}
# Global property defaults: a property field left 0 takes their value (hm_read_prop14.F:277-291,
# hm_read_prop01.F:194-273), so they are read only when they leave every default as it is.
_DEFAULTS = {"DEF_SHELL", "DEF_SOLID"}


def _all_default(lines) -> bool:
    """True when every field of the cards is blank or zero."""
    for line in lines:
        for tok in line.split():
            try:
                if float(tok.replace("d", "e").replace("D", "e")) != 0.0:
                    return False
            except ValueError:
                return False
    return True


def _th_group(b: _Block, kind: str, where: str):
    """/TH/NODE, /TH/BRIC, /TH/SHEL (th_node.cfg, th_bric.cfg, th_shel.cfg): title; variable names
    in 10-character fields, on lines up to the first entity line; then per entity its id (I10), a
    skew frame id (I10, nodes and shells) and a name (from column 21)."""
    b.need(1)
    names, ents, labels = [], [], []
    for line in _data(b.lines[1:]):
        first = line[:10].strip()
        if not ents and first and not first.lstrip("-").isdigit():
            names += [line[i:i + 10].strip() for i in range(0, len(line), 10) if line[i:i + 10].strip()]
            continue
        if not first:
            continue
        if kind in ("NODE", "SHEL") and _int(line[10:20]) != 0:
            raise DeckError(f"{where}: time-history output in a skew frame is not on this path")
        ents.append(_int(first))
        labels.append(line[20:].strip()[:100])
    return (kind, b.ident(), b.lines[0].rstrip(), names, np.array(ents, np.int64), labels)


def parse(text: str) -> Deck:
    """Read the keywords of a Starter deck given as text."""
    node_ids, xs = [], []
    bricks, shells = [], []
    parts, mats, sprops, cprops, grnod, funct = {}, {}, {}, {}, {}, {}
    bcs, inivel, impvel_raw = [], [], []
    titles, th, th_other = {}, [], []
    for b in _blocks(text):
        p = [s.upper() for s in b.parts]
        head = p[0]
        where = f"{b.keyword} (line {b.lineno})"
        if head in _IGNORED:
            continue
        if head in _DEFAULTS:
            if not _all_default(b.lines):
                raise DeckError(f"{where}: /{head} changes property defaults; state the flags in the properties")
            continue
        if head == "BEGIN":
            if len(b.lines) >= 4 and b.lines[2].split() != b.lines[3].split():
                raise DeckError(f"{where}: input and work units differ; convert the deck first")
            if b.lines:
                titles["RUN"] = b.lines[0].rstrip()
            continue
        if head == "TH":
            kind = p[1] if len(p) > 2 else "?"
            if kind not in ("NODE", "BRIC", "SHEL"):
                th_other.append(b.keyword)
                continue
            th.append(_th_group(b, kind, where))
            continue
        if head in ("UNIT", "RANDOM"):
            raise DeckError(f"{where}: keyword changes the model ({head.lower()}) and is not applied here")
        if head == "NODE":
            for line in _data(b.lines):
                i, x, y, z = _cut(line, (10, 20, 20, 20))
                node_ids.append(_int(i))
                xs.append((_real(x), _real(y), _real(z)))
        elif head == "BRICK":
            part = b.ident()
            rows = [[_int(s) for s in _cut(line, (10,) * 9)] for line in _data(b.lines)]
            a = np.array(rows, dtype=np.int64).reshape(-1, 9)
            bricks.append((part, a[:, 0], a[:, 1:]))
        elif head == "SHELL":
            part = b.ident()
            ids, conn, thk = [], [], []
            for line in _data(b.lines):
                f = _cut(line, (10,) * 5 + (20, 20))
                ids.append(_int(f[0]))
                conn.append([_int(s) for s in f[1:5]])
                thk.append(_real(f[6]))
            shells.append((part, np.array(ids, np.int64), np.array(conn, np.int64).reshape(-1, 4),
                           np.array(thk, np.float64)))
        elif head == "PART":
            b.need(2)
            prop, mat, sub = (_int(s) for s in _cut(b.lines[1], (10, 10, 10)))
            parts[b.ident()] = (prop, mat)
            titles[("PART", b.ident())] = b.lines[0].rstrip()
            if sub:
                titles[("PART_SUBSET", b.ident())] = sub
        elif head == "MAT":
            law = p[1] if len(p) > 2 else "?"
            mats[b.ident()] = _law2(b) if law in ("PLAS_JOHNS", "LAW2") else law
            titles[("MAT", b.ident())] = b.lines[0].rstrip() if b.lines else ""
        elif head == "PROP":
            kind = p[1] if len(p) > 2 else "?"
            titles[("PROP", b.ident())] = b.lines[0].rstrip() if b.lines else ""
            if kind in ("SOLID", "TYPE14"):
                sprops[b.ident()] = _solid_prop(b)
            elif kind in ("SHELL", "TYPE1"):
                cprops[b.ident()] = _shell_prop(b)
            else:
                sprops[b.ident()] = kind          # reported if a part on the path uses it
        elif head == "GRNOD":
            if len(p) < 3 or p[1] != "NODE":
                raise DeckError(f"{where}: only /GRNOD/NODE node groups are read")
            b.need(1)
            grnod[b.ident()] = _ids(b.lines[1:])
        elif head == "BCS":
            b.need(2)
            line = b.lines[1]
            tra, rot = line[0:6].strip(), line[6:10].strip()
            skew, gid = _int(line[10:20]), _int(line[20:30])
            if skew != 0:
                raise DeckError(f"{where}: boundary conditions in a skew frame are not on this path")
            code = 0
            for ch, bit in zip(tra.rjust(3, "0")[-3:], (M.BC_TX, M.BC_TY, M.BC_TZ)):
                code |= bit if ch == "1" else 0
            for ch, bit in zip(rot.rjust(3, "0")[-3:], (M.BC_RX, M.BC_RY, M.BC_RZ)):
                code |= bit if ch == "1" else 0
            bcs.append((code, gid))
        elif head == "INIVEL":
            if len(p) < 3 or p[1] != "TRA":
                raise DeckError(f"{where}: only /INIVEL/TRA is read")
            b.need(2)
            vx, vy, vz, gid, skew = _cut(b.lines[1], (20, 20, 20, 10, 10))
            if _int(skew) != 0:
                raise DeckError(f"{where}: initial velocities in a skew frame are not on this path")
            if len(b.lines) > 2 and len(_cut(b.lines[2], (20, 10))) > 1 and _int(_cut(b.lines[2], (20, 10))[1]):
                raise DeckError(f"{where}: an initial velocity started by a sensor is not on this path")
            if len(b.lines) > 2 and _real(_cut(b.lines[2], (20,))[0]) != 0.0:
                raise DeckError(f"{where}: a delayed initial velocity (Tstart) is not on this path")
            inivel.append((_real(vx), _real(vy), _real(vz), _int(gid)))
        elif head == "IMPVEL":
            b.need(2)
            fct, d, skew, sens, gid, frame, icoor = _cut(b.lines[1], (10,) * 7)
            scales = _cut(b.lines[2], (20,) * 4) if len(b.lines) > 2 else ("",) * 4
            ascale, fscale, tstart, tstop = (_real(s) for s in scales)
            d = d.strip().upper()
            if d not in ("X", "Y", "Z"):
                raise DeckError(f"{where}: direction {d!r}; this path imposes X, Y or Z velocities")
            if _int(skew) or _int(sens) or _int(frame) or _int(icoor) or tstart != 0.0 or tstop != 0.0:
                raise DeckError(f"{where}: skew, sensor, frame or a time window is not on this path")
            impvel_raw.append((_int(fct), "XYZ".index(d), fscale if fscale != 0.0 else 1.0, _int(gid), where))
        elif head == "FUNCT":
            b.need(1)
            ys = []
            for line in _data(b.lines[1:]):
                _, y = _cut(line, (20, 20))
                ys.append(_real(y))
            funct[b.ident()] = ys
        else:
            raise DeckError(f"{where}: keyword not read by this path's deck reader")
    impvel = []
    for fct, d, fscale, gid, where in impvel_raw:
        ys = funct.get(fct)
        if not ys:
            raise DeckError(f"{where}: function {fct} is not defined")
        if any(y != ys[0] for y in ys):
            raise DeckError(f"{where}: function {fct} varies in time; this path imposes constant velocities")
        impvel.append((d, fscale * ys[0], gid))
    if not node_ids:
        raise DeckError("the deck has no /NODE block")
    return Deck(node_ids=np.array(node_ids, np.int64), x=np.array(xs, np.float64).reshape(-1, 3),
                bricks=bricks, shells=shells, parts=parts, mats=mats, solid_props=sprops,
                shell_props=cprops, grnod=grnod, bcs=bcs, inivel=inivel, impvel=impvel, titles=titles, th=th,
                th_other=th_other)


def _family(deck: Deck, blocks, props: dict, kind: str):
    """The one (material, property) pair of an element family; its blocks concatenated."""
    if not blocks:
        return None
    pairs = set()
    for part in {blk[0] for blk in blocks}:
        if part not in deck.parts:
            raise DeckError(f"/{kind}/{part}: part {part} is not defined")
        prop, mat = deck.parts[part]
        m, pr = deck.mats.get(mat), props.get(prop)
        if not isinstance(m, M.Law2):
            raise DeckError(f"/PART/{part}: material {mat} is {m or 'not defined'}; "
                            "this path runs law 2 (/MAT/PLAS_JOHNS)")
        if pr is None or isinstance(pr, str):
            raise DeckError(f"/PART/{part}: property {prop} is {pr or 'not defined'} for a /{kind} element")
        pairs.add((mat, prop))
    if len(pairs) != 1:
        raise DeckError(f"/{kind}: parts use {len(pairs)} material/property pairs; "
                        "this path takes one per element family")
    mat, prop = pairs.pop()
    return deck.mats[mat], props[prop]


def to_model(deck: Deck, dtfac_hex: float = 0.9, dtfac_q4: float = 0.9, dtmx: float = M.EP20) -> M.Model:
    """Build the post-Starter model of a deck (nodes renumbered 0.. in deck order)."""
    ids = deck.node_ids
    order = np.argsort(ids, kind="stable")
    if np.any(np.diff(ids[order]) == 0):
        raise DeckError("a node id appears twice in /NODE")
    sorted_ids = ids[order]

    def local(user):
        user = np.asarray(user, np.int64)
        pos = np.searchsorted(sorted_ids, user)
        pos = np.minimum(pos, len(sorted_ids) - 1)
        bad = sorted_ids[pos] != user
        if np.any(bad):
            raise DeckError(f"node {int(user[bad][0])} is used but not defined")
        return order[pos].astype(np.int64)

    nn = len(ids)
    hex_conn = hex_mat = hex_prop = None
    fam = _family(deck, deck.bricks, deck.solid_props, "BRICK")
    if fam:
        hex_mat, hex_prop = fam
        hex_conn = local(np.concatenate([blk[2] for blk in deck.bricks])).reshape(-1, 8).astype(np.int32)
    q4_conn = q4_mat = q4_prop = thke = None
    fam = _family(deck, deck.shells, deck.shell_props, "SHELL")
    if fam:
        q4_mat, q4_prop = fam
        conn = np.concatenate([blk[2] for blk in deck.shells])
        conn[:, 3] = np.where(conn[:, 3] == 0, conn[:, 2], conn[:, 3])     # triangle: node 4 = node 3
        q4_conn = local(conn).reshape(-1, 4).astype(np.int32)
        t = np.concatenate([blk[3] for blk in deck.shells])
        thke = np.where(t > 0.0, t, q4_prop.thick)
    if hex_conn is None and q4_conn is None:
        raise DeckError("the deck has no /BRICK or /SHELL elements")

    def group(gid, where):
        if gid not in deck.grnod:
            raise DeckError(f"{where}: node group {gid} is not defined")
        return local(deck.grnod[gid])

    code = np.zeros(nn, np.int32)
    for bits, gid in deck.bcs:
        code[group(gid, "/BCS")] |= bits
    v0 = np.zeros((nn, 3))
    for vx, vy, vz, gid in deck.inivel:
        v0[group(gid, "/INIVEL")] = (vx, vy, vz)
    vimp = np.zeros((nn, 3))
    bit = (M.IMP_X, M.IMP_Y, M.IMP_Z)
    for d, val, gid in deck.impvel:
        n = group(gid, "/IMPVEL")
        code[n] |= bit[d]
        vimp[n, d] = val
    hex_ids = np.concatenate([blk[1] for blk in deck.bricks]).astype(np.int64) if hex_conn is not None else None
    q4_ids = np.concatenate([blk[1] for blk in deck.shells]).astype(np.int64) if q4_conn is not None else None
    for ids, kind in ((hex_ids, "BRICK"), (q4_ids, "SHELL")):
        if ids is not None and np.unique(ids).shape[0] != ids.shape[0]:
            raise DeckError(f"/{kind}: an element id appears twice")
    return M.build_model(deck.x, hex_conn=hex_conn, hex_mat=hex_mat, hex_prop=hex_prop, q4_conn=q4_conn,
                         q4_mat=q4_mat, q4_prop=q4_prop, thke=thke, v0=v0, code=code, vimp=vimp,
                         dtfac_hex=dtfac_hex, dtfac_q4=dtfac_q4, dtmx=dtmx, hex_ids=hex_ids, q4_ids=q4_ids)


def read_run(path: str, engine: str, **kw) -> tuple[M.Model, EngineSettings]:
    """A Starter deck and its Engine file: the model (with the Engine file's /DT/BRICK and /DT/SHELL
    scale factors applied) and the Engine settings, whose `tstop` is the run's end time."""
    with open(engine, "r", encoding="latin-1") as f:
        e = parse_engine(f.read())
    return read(path, engine, **kw), e


def read(path: str, engine: str | None = None, **kw) -> M.Model:
    """Read a Starter deck file into a model. With `engine`, the Engine file's /DT/BRICK and
    /DT/SHELL scale factors are applied; other `to_model` keywords pass through.  The model has
    no end time: take it from `read_run` (or `parse_engine(...).tstop`)."""
    if engine is not None:
        with open(engine, "r", encoding="latin-1") as f:
            e = parse_engine(f.read())
        kw = dict(dict(dtfac_hex=e.dtfac_hex, dtfac_q4=e.dtfac_q4), **kw)
    with open(path, "r", encoding="latin-1") as f:
        return to_model(parse(f.read()), **kw)


# ------------------------------------------------------------------------------ engine file
@dataclasses.dataclass
class EngineSettings:
    """What an Engine input file (`*_0001.rad`) sets for this path's cycle."""

    tstop: float | None = None            # /RUN end time
    dtfac_hex: float = 0.9                # DTFAC1(1), /DT/BRICK
    dtfac_q4: float = 0.9                 # DTFAC1(3), /DT/SHELL
    dthis: float | None = None            # /TFILE time-history interval
    th_vers: int = 40                     # /TH/VERS (freform.F:1543-1549)


# Engine keywords that request output or name the run; they leave the cycle unchanged.
_ENGINE_IGNORED = {"TITLE", "VERS", "ANIM", "PRINT", "RFILE", "MON", "H3D", "OUTP"}


def _free(line: str) -> list[str]:
    """The items of a list-directed (free format) card."""
    return [t for t in re.split(r"[\s,]+", line.strip()) if t]


def parse_engine(text: str) -> EngineSettings:
    """/RUN and the /DT/BRICK, /DT/SHELL scale factors of an Engine file; the factors are read
    as freform.F:1076-1115 reads them (list-directed Tsca, Tmin into DTFAC1 / DTMIN1 of the
    brick and shell types). Keywords that would change the cycle otherwise are refused."""
    out = EngineSettings()
    for b in _blocks(text):
        p = [x.upper() for x in b.parts]
        where = f"{b.keyword} (line {b.lineno})"
        if p[0] in _ENGINE_IGNORED:
            continue
        if p[0] == "RUN":
            b.need(1)
            out.tstop = _real(_free(b.lines[0])[0])
        elif p[0] == "TFILE":
            b.need(1)
            out.dthis = _real(_free(b.lines[0])[0])
        elif p[0] == "TH" and len(p) > 1 and p[1] == "VERS":
            out.th_vers = int(p[2]) if len(p) > 2 and p[2].isdigit() else 40
        elif p[0] == "STOP":
            # /STOP: Emax, Mmax, Nmax (energy / mass error limits that end the run) then output
            # counts (freform.F:770-786); a limit of 0 means none
            if len(p) > 1 and p[1] == "LSENS":
                raise DeckError(f"{where}: a run stopped by sensors is not on this path")
            vals = _free(b.lines[0]) if b.lines else []
            if any(_real(v) != 0.0 for v in vals[:3]):
                raise DeckError(f"{where}: energy or mass error limits that stop the run are not applied here")
        elif p[0] == "DT":
            kind = p[1] if len(p) > 1 else ""
            extra = [x for x in p[2:] if not x.isdigit()]
            if kind not in ("BRICK", "SHELL") or extra:
                raise DeckError(f"{where}: only the /DT/BRICK and /DT/SHELL scale factors apply on this path")
            b.need(1)
            vals = _free(b.lines[0])
            tsca = _real(vals[0]) if vals else 0.0
            tmin = _real(vals[1]) if len(vals) > 1 else 0.0
            if tsca <= 0.0:
                raise DeckError(f"{where}: state a positive time-step scale factor")
            if tmin != 0.0:
                raise DeckError(f"{where}: a minimum element time step is not on this path")
            if kind == "BRICK":
                out.dtfac_hex = tsca
            else:
                out.dtfac_q4 = tsca
        else:
            raise DeckError(f"{where}: Engine keyword not applied by this path")
    return out


# ----------------------------------------------------------------------------------- writer
def _r(v: float) -> str:
    """A 20-character real field; exact (shortest round-trip form) whenever that fits."""
    s = repr(float(v))
    if len(s) > 20:
        s = f"{v:.12e}"
    return s.rjust(20)


def _i10(v: int) -> str:
    return str(int(v)).rjust(10)


def write(m: M.Model, title: str = "model") -> str:
    """A Starter deck of a model (ids are 1-based local numbers). Reading it back with `read`
    gives the same model wherever the reals fit 20 characters exactly."""
    if m.hexp is None and m.q4 is None:
        raise DeckError("the model has no elements")
    if np.any(np.asarray(m.vr) != 0.0):
        raise DeckError("initial rotational velocities are not written (the reader takes /INIVEL/TRA only)")
    if m.q4 is not None and (m.q4.prop.ishear != 0 or m.q4.prop.dtmin != 0.0):
        raise DeckError("a non-default shell Ishear or dt_min is not written")
    if m.dtmx != M.EP20:
        raise DeckError("a maximum time step DTMX is not written")
    out = io.StringIO()
    w = out.write
    w(f"#RADIOSS STARTER\n/BEGIN\n{title}\n      2025         0\n"
      "                   1                   1                   1\n"
      "                   1                   1                   1\n")
    w("/NODE\n")
    for i, (x, y, z) in enumerate(m.x, 1):
        w(f"{_i10(i)}{_r(x)}{_r(y)}{_r(z)}\n")
    mats = []
    if m.hexp is not None:
        mats.append((1, m.hexp.mat))
    if m.q4 is not None:
        mats.append((2, m.q4.mat))
    for mid, mat in mats:
        w(f"/MAT/PLAS_JOHNS/{mid}\nlaw 2\n{_r(mat.rho)}\n{_r(mat.young)}{_r(mat.nu)}{_i10(0)}{_i10(mat.vp)}"
          f"{_r(0.0)}\n{_r(mat.a)}{_r(mat.b)}{_r(mat.n)}{_r(mat.epsmax)}{_r(mat.sigmax)}\n"
          f"{_r(mat.c)}{_r(mat.eps0)}{_i10(mat.icc)}{_i10(0)}{_r(mat.fcut)}{_r(0.0)}\n"
          f"{_r(mat.m_exp)}{_r(0.0)}{_r(0.0)}{_r(0.0)}{_r(0.0)}\n")
    if m.hexp is not None:
        pr = m.hexp.prop
        if pr.vns1 != 0.0 or pr.vns2 != 0.0:
            raise DeckError("numerical viscosities are not written")
        w(f"/PROP/SOLID/1\nsolids\n{_i10(1)}{_i10(0) * 7}{_r(0.0)}\n"
          f"{_r(pr.qa)}{_r(pr.qb)}{_r(pr.qh)}{_r(0.0)}{_r(0.0)}\n{_r(0.0)}\n")
        w(f"/PART/1\nsolids\n{_i10(1)}{_i10(1)}{_i10(0)}\n/BRICK/1\n")
        hids = np.arange(1, m.ne_hex + 1) if m.hexp.ids is None else m.hexp.ids
        for e, row in zip(hids, m.hexp.conn):
            w(_i10(int(e)) + "".join(_i10(n + 1) for n in row) + "\n")
    if m.q4 is not None:
        pr = m.q4.prop
        w(f"/PROP/SHELL/2\nshells\n{_i10(1)}{_i10(2)}{_i10(0) * 3}{' ' * 10}{_r(0.0)}\n"
          f"{_r(pr.hm)}{_r(pr.hf)}{_r(pr.hr)}{_r(0.0)}{_r(0.0)}\n"
          f"{_i10(pr.npt)}{' ' * 10}{_r(pr.thick)}{_r(0.0)}{' ' * 10}{_i10(pr.ithk)}{_i10(pr.ipla)}{_i10(0)}\n")
        w(f"/PART/2\nshells\n{_i10(2)}{_i10(2)}{_i10(0)}\n/SHELL/2\n")
        qids = np.arange(m.ne_hex + 1, m.ne_hex + m.ne_q4 + 1) if m.q4.ids is None else m.q4.ids
        for e, row, t in zip(qids, m.q4.conn, m.q4.thke):
            w(_i10(int(e)) + "".join(_i10(n + 1) for n in row) + _r(0.0) + _r(t) + "\n")
    gid = [0]

    def grnod(nodes, name):
        gid[0] += 1
        w(f"/GRNOD/NODE/{gid[0]}\n{name}\n")
        for k in range(0, len(nodes), 10):
            w("".join(_i10(n + 1) for n in nodes[k:k + 10]) + "\n")
        return gid[0]

    bc = m.code & M.BC_ALL
    for c in np.unique(bc[bc != 0]):
        g = grnod(np.nonzero(bc == c)[0], f"bcs {c}")
        tra = "".join("1" if c & b else "0" for b in (M.BC_TX, M.BC_TY, M.BC_TZ))
        rot = "".join("1" if c & b else "0" for b in (M.BC_RX, M.BC_RY, M.BC_RZ))
        w(f"/BCS/{g}\nbcs\n   {tra} {rot}{_i10(0)}{_i10(g)}\n")
    vel = np.ascontiguousarray(m.v)
    moving = np.nonzero(np.any(vel != 0.0, axis=1))[0]
    uniq, inv = np.unique(vel[moving], axis=0, return_inverse=True)
    for k, (vx, vy, vz) in enumerate(uniq):
        g = grnod(moving[inv.reshape(-1) == k], "inivel")
        w(f"/INIVEL/TRA/{g}\ninivel\n{_r(vx)}{_r(vy)}{_r(vz)}{_i10(g)}{_i10(0)}\n")
    fct = 0
    for d, bit in enumerate((M.IMP_X, M.IMP_Y, M.IMP_Z)):
        nodes = np.nonzero(m.code & bit)[0]
        for val in np.unique(m.vimp[nodes, d]):
            g = grnod(nodes[m.vimp[nodes, d] == val], "impvel")
            fct += 1
            w(f"/FUNCT/{fct}\nconstant\n{_r(0.0)}{_r(val)}\n{_r(1.0e30)}{_r(val)}\n")
            w(f"/IMPVEL/{fct}\nimpvel\n{_i10(fct)}{'XYZ'[d].rjust(10)}{_i10(0) * 2}{_i10(g)}{_i10(0) * 2}\n"
              f"{_r(0.0)}{_r(1.0)}{_r(0.0)}{_r(0.0)}\n")
    w("/END\n")
    return out.getvalue()


def write_engine(m: M.Model, tstop: float, title: str = "model") -> str:
    """An Engine file for a model written by `write`: /RUN with the end time and the /DT/BRICK,
    /DT/SHELL scale factors of the model's families (read back by `parse_engine` / `read_run`)."""
    out = [f"#RADIOSS ENGINE\n/RUN/{title}/1\n{_r(tstop)}\n"]
    if m.hexp is not None:
        out.append(f"/DT/BRICK\n{_r(m.hexp.dtfac)}{_r(0.0)}\n")
    if m.q4 is not None:
        out.append(f"/DT/SHELL\n{_r(m.q4.dtfac)}{_r(0.0)}\n")
    out.append("/END\n")
    return "".join(out)
