"""Model builder: the Starter-side inputs of the explicit hex/Q4 path.

This module produces, from a mesh and material description, exactly the data
the Engine's per-cycle path consumes after the Starter has run:

* nodal arrays X, V, D, MS (SMASS3 lumping, starter/source/elements/solid/
  solide/smass3.F: ``MASS = RHO*VOLU*ONE_OVER_8`` per element node);
* element groups with the ELBUF state of law 2 (SIG(NEL,6) component-major,
  EINT, RHO, VOL, PLA, EPSD, OFF, QVIS);
* the Parith/ON skyline: per-node contribution rows ordered as FILLCNE orders
  them (starter/source/spmd/domdec2.F:2046, 2166-2317): element families in
  the order solids -> quads -> shells -> trusses -> beams -> springs ->
  triangles, inside a family by ascending user element id.  ``ADSKY`` is the
  CSR pointer per node and ``IADS(k, e)`` the row written by local node k of
  element e (scumu3p.F / asspar4.F:164-180);
* boundary conditions (BCS codes as bcs1.F ``LCOD``: 4=x, 2=y, 1=z) and
  constant imposed velocities (fixvel.F global-frame branch).

Everything is numpy on the host; nothing here touches the GPU.
"""
from __future__ import annotations

import dataclasses
import numpy as np

# Johnson-Cook / law-2 defaults applied by the Starter reader
# (starter/source/materials/mat/mat002/hm_read_mat02_jc.F90).
EP20 = 1.0e20


@dataclasses.dataclass
class Law2:
    """Law 2 (Johnson-Cook) material, with the Starter's defaulting applied."""

    rho: float
    young: float
    nu: float
    a: float
    b: float
    n: float
    epsmax: float = 0.0
    sigmax: float = 0.0
    c: float = 0.0
    eps0: float = 0.0
    icc: int = 0
    vp: int = 0
    fcut: float = 0.0
    m_exp: float = 0.0

    def params(self) -> dict:
        """Resolved parameters as M2LAW reads them (mat_param%uparam / %iparam)."""
        cn = self.n if self.n != 0.0 else 1.0
        anu = self.nu
        if anu == 0.5:
            anu = 0.499
        epsm = self.epsmax if self.epsmax != 0.0 else EP20
        sigm = self.sigmax if self.sigmax != 0.0 else EP20
        icc = self.icc if self.icc != 0 else 1
        vp = self.vp if self.vp != 0 else 2
        israte = 1
        eps0 = self.eps0
        if self.c == 0.0:
            eps0 = 1.0
            israte = 0
        fcut = self.fcut
        asrate = EP20 if fcut == 0.0 else 2.0 * np.pi * fcut
        m_exp = self.m_exp if self.m_exp != 0.0 else 1.0
        g = self.young / (2.0 * (1.0 + anu))
        bulk = self.young / (3.0 * (1.0 - 2.0 * anu))
        return dict(rho0=self.rho, bulk=bulk, shear=g, ca=self.a, cb=self.b, cn=cn,
                    epmx=epsm, sigmx=sigm, cc=self.c, epdr=eps0, fisokin=0.0,
                    asrate=asrate, m_exp=m_exp, icc=icc, vp=vp, israte=israte, iform=0,
                    young=self.young, nu=anu)


@dataclasses.dataclass
class SolidProp:
    """/PROP/SOLID with Isolid=1 defaults (hm_read_prop14.F:355-369)."""

    qa: float = 1.1
    qb: float = 0.05
    qh: float = 0.1
    vns1: float = 0.0
    vns2: float = 0.0


@dataclasses.dataclass
class HexModel:
    """A hex-only model in the Engine's post-Starter form (0-based indices)."""

    x: np.ndarray          # (nn,3) float64
    v: np.ndarray          # (nn,3)
    ms: np.ndarray         # (nn,)
    conn: np.ndarray       # (ne,8) int32, 0-based
    vol0: np.ndarray       # (ne,) initial volume (GBUF%VOL)
    iads: np.ndarray       # (ne,8) int64 skyline row per element node
    adsky: np.ndarray      # (nn+1,) int64 CSR pointers into the skyline
    bcs: np.ndarray        # (nn,) int32 BCS code
    imp_node: np.ndarray   # (nimp,) int32
    imp_dof: np.ndarray    # (nimp,) int32
    imp_val: np.ndarray    # (nimp,) float64
    mat: Law2
    prop: SolidProp
    dtfac: float = 0.9
    dtmx: float = EP20

    @property
    def nn(self) -> int:
        return int(self.x.shape[0])

    @property
    def ne(self) -> int:
        return int(self.conn.shape[0])

    @property
    def lsky(self) -> int:
        return int(self.adsky[-1])

    def params(self) -> dict:
        p = self.mat.params()
        p.update(qa=self.prop.qa, qb=self.prop.qb, qh=self.prop.qh, vns1=self.prop.vns1,
                 vns2=self.prop.vns2, dtfac=self.dtfac, dtmx=self.dtmx)
        return p

    def initial_state(self) -> dict:
        """ELBUF initial state of law 2 (zero stress, RHO=RHO0, OFF=1)."""
        ne = self.ne
        p = self.mat.params()
        return dict(sig=np.zeros((6, ne)), eint=np.zeros(ne), rho=np.full(ne, p["rho0"]),
                    pla=np.zeros(ne), epsd=np.zeros(ne), offg=np.ones(ne), qvis=np.zeros(ne))


def hex_volume(xyz: np.ndarray) -> np.ndarray:
    """SDERI3 volume ``ONE_OVER_64*(JAC1*JAC_59_68+JAC2*JAC_67_49+JAC3*JAC_48_57)``.

    xyz: (ne,8,3).  Same operation order as sderi3.F, vectorised over elements.
    """
    x = [xyz[:, k, 0] for k in range(8)]
    y = [xyz[:, k, 1] for k in range(8)]
    z = [xyz[:, k, 2] for k in range(8)]

    def jac(c):
        c17 = c[6] - c[0]
        c28 = c[7] - c[1]
        c35 = c[4] - c[2]
        c46 = c[5] - c[3]
        j1 = c17 + c28 - c35 - c46
        a = c17 + c46
        b = c28 + c35
        return j1, a + b, a - b

    j1, j4, j7 = jac(x)
    j2, j5, j8 = jac(y)
    j3, j6, j9 = jac(z)
    j5968 = j5 * j9 - j6 * j8
    j6749 = j6 * j7 - j4 * j9
    j4857 = j4 * j8 - j5 * j7
    return (1.0 / 64.0) * (j1 * j5968 + j2 * j6749 + j3 * j4857)


def skyline(conn: np.ndarray, nn: int, elem_rank: np.ndarray | None = None):
    """Parith/ON skyline for element connectivity ``conn`` (ne, nnode).

    Rows of node n are ordered by ``elem_rank`` (FILLCNE: family, then user id;
    default = element index).  Returns (iads (ne,nnode) int64, adsky (nn+1,) int64).
    """
    ne, nk = conn.shape
    if elem_rank is None:
        elem_rank = np.arange(ne, dtype=np.int64)
    node = conn.reshape(-1).astype(np.int64)
    rank = np.repeat(elem_rank.astype(np.int64), nk)
    order = np.lexsort((rank, node))          # by node, then element rank
    counts = np.bincount(node, minlength=nn).astype(np.int64)
    adsky = np.zeros(nn + 1, dtype=np.int64)
    np.cumsum(counts, out=adsky[1:])
    iads_flat = np.empty(ne * nk, dtype=np.int64)
    iads_flat[order] = np.arange(ne * nk, dtype=np.int64)
    return iads_flat.reshape(ne, nk), adsky


def skyline_sum(values: np.ndarray, iads: np.ndarray, adsky: np.ndarray) -> np.ndarray:
    """Sum per-element-node values into nodes in skyline order (left to right)."""
    nn = adsky.shape[0] - 1
    lsky = int(adsky[-1])
    rows = np.zeros(lsky)
    rows[iads.reshape(-1)] = values.reshape(-1)
    out = np.zeros(nn)
    counts = np.diff(adsky)
    for j in range(int(counts.max()) if nn else 0):
        has = counts > j
        out[has] = out[has] + rows[adsky[:-1][has] + j]
    return out


def hex_block(nx: int, ny: int, nz: int, h: float, origin=(0.0, 0.0, 0.0)):
    """Structured block of nx*ny*nz hexes of edge h; x fastest.

    Node (i,j,k) -> i+(nx+1)*(j+(ny+1)*k); element (i,j,k) -> i+nx*(j+ny*k);
    local nodes 1-4 on the lower z face counter-clockwise seen from +z, 5-8 above.
    """
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)  # noqa: E731
    x = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    ids = nid(ii, jj, kk).reshape(-1)
    x[ids, 0] = origin[0] + ii.reshape(-1) * h
    x[ids, 1] = origin[1] + jj.reshape(-1) * h
    x[ids, 2] = origin[2] + kk.reshape(-1) * h
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    eid = (ei + nx * (ej + ny * ek)).reshape(-1)
    ei, ej, ek = ei.reshape(-1), ej.reshape(-1), ek.reshape(-1)
    conn = np.empty((nx * ny * nz, 8), dtype=np.int32)
    conn[eid, 0] = nid(ei, ej, ek)
    conn[eid, 1] = nid(ei + 1, ej, ek)
    conn[eid, 2] = nid(ei + 1, ej + 1, ek)
    conn[eid, 3] = nid(ei, ej + 1, ek)
    conn[eid, 4] = nid(ei, ej, ek + 1)
    conn[eid, 5] = nid(ei + 1, ej, ek + 1)
    conn[eid, 6] = nid(ei + 1, ej + 1, ek + 1)
    conn[eid, 7] = nid(ei, ej + 1, ek + 1)
    return x, conn


def build_hex_model(x, conn, mat: Law2, prop: SolidProp | None = None, v0=None,
                    bcs=None, imp=None, dtfac: float = 0.9) -> HexModel:
    """Assemble a HexModel: volumes, SMASS3 lumped masses, skyline, BCs."""
    prop = prop or SolidProp()
    nn = x.shape[0]
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    vol0 = hex_volume(x[conn])
    iads, adsky = skyline(conn, nn)
    rho = mat.params()["rho0"]
    mass = (1.0 * rho) * vol0 * (1.0 / 8.0)   # FILL*RHO*VOLU*ONE_OVER_8
    ms = skyline_sum(np.repeat(mass[:, None], 8, axis=1), iads, adsky)
    v = np.zeros((nn, 3)) if v0 is None else np.array(v0, dtype=np.float64).reshape(nn, 3)
    bcs = np.zeros(nn, dtype=np.int32) if bcs is None else np.asarray(bcs, dtype=np.int32)
    if imp is None:
        imp = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    return HexModel(x=np.ascontiguousarray(x, dtype=np.float64), v=v, ms=ms, conn=conn, vol0=vol0,
                    iads=iads, adsky=adsky, bcs=bcs, imp_node=np.asarray(imp[0], np.int32),
                    imp_dof=np.asarray(imp[1], np.int32), imp_val=np.asarray(imp[2], np.float64),
                    mat=mat, prop=prop, dtfac=dtfac)


STEEL = Law2(rho=7.85e-9, young=210000.0, nu=0.3, a=350.0, b=275.0, n=0.36)


def cantilever_bar(nx: int = 20, ny: int = 20, nz: int = 80, h: float = 1.0, v0: float = 1.0e4,
                   mat: Law2 = STEEL) -> HexModel:
    """Cantilever bar: block along z, z=0 end fully fixed, initial axial velocity v0 (mm/s)."""
    x, conn = hex_block(nx, ny, nz, h)
    nn = x.shape[0]
    v = np.zeros((nn, 3))
    v[:, 2] = v0
    bcs = np.where(x[:, 2] == 0.0, 7, 0).astype(np.int32)
    return build_hex_model(x, conn, mat, v0=v, bcs=bcs)


# ---------------------------------------------------------------------------
# Q4 shells and mixed hex + Q4 models (6-DOF nodes)
# ---------------------------------------------------------------------------

# node code bits (BCS translations / rotations, imposed velocity), see bcs1.F LCOD
BC_TX, BC_TY, BC_TZ = 4, 2, 1
BC_RX, BC_RY, BC_RZ = 32, 16, 8
IMP_X, IMP_Y, IMP_Z = 256, 128, 64
BC_ALL = BC_TX | BC_TY | BC_TZ | BC_RX | BC_RY | BC_RZ


@dataclasses.dataclass
class ShellProp:
    """/PROP/SHELL (IGTYP=1) as the Starter resolves it (hm_read_prop01.F:156-245).

    Ishell=1 (IHBE=1), Ismstr=2, NPT points with the COQINI Newton-Cotes-type
    tables, hourglass coefficients Hm/Hf/Hr (default 0.01), shear factor 5/6.
    """

    thick: float
    npt: int = 5
    ipla: int = 1
    ithk: int = 1
    ishear: int = 0
    hm: float = 0.01
    hf: float = 0.01
    hr: float = 0.01
    dtmin: float = 0.0


def shell_params(mat: Law2, prop: ShellProp, dtfac: float) -> dict:
    """Resolved parameters of the Q4 path (PM words, uparam, GEO words)."""
    p = mat.params()
    young, nu, rho = mat.young, p["nu"], mat.rho
    ssp = float(np.sqrt(young / max(rho, 1.0e-20)))
    a11 = young / (1.0 - nu * nu)
    return dict(rho=rho, young=young, nu=nu, shear=p["shear"], a11=a11, a12=a11 * nu, ssp=ssp,
                ca=p["ca"], cb=p["cb"], cn=p["cn"], epmx=p["epmx"], ymax=p["sigmx"], cc=p["cc"],
                epdr=p["epdr"], asrate=p["asrate"], m_exp=p["m_exp"], icc=p["icc"], vp=p["vp"],
                israte=p["israte"], iform=0, npt=prop.npt, ipla=prop.ipla, ithk=prop.ithk, ishear=prop.ishear,
                hm=prop.hm, hf=prop.hf, hr=prop.hr, fsh=5.0 / 6.0, dtfac=dtfac, dtmin=prop.dtmin)


@dataclasses.dataclass
class HexPart:
    conn: np.ndarray     # (ne,8) int32 0-based
    vol0: np.ndarray     # (ne,)
    iads: np.ndarray     # (ne,8) int64 skyline rows
    mat: Law2
    prop: SolidProp
    dtfac: float = 0.9
    ids: np.ndarray | None = None  # (ne,) element user ids (None: index + 1)

    @property
    def ne(self) -> int:
        return int(self.conn.shape[0])

    def params(self, dtmx: float) -> dict:
        p = self.mat.params()
        p.update(qa=self.prop.qa, qb=self.prop.qb, qh=self.prop.qh, vns1=self.prop.vns1, vns2=self.prop.vns2,
                 dtfac=self.dtfac, dtmx=dtmx)
        return p

    def initial_state(self) -> dict:
        ne = self.ne
        return dict(sig=np.zeros((6, ne)), eint=np.zeros(ne), rho=np.full(ne, self.mat.rho), pla=np.zeros(ne),
                    epsd=np.zeros(ne), offg=np.ones(ne), qvis=np.zeros(ne))


@dataclasses.dataclass
class Q4Part:
    conn: np.ndarray     # (ne,4) int32 0-based (node 4 == node 3 for a triangle)
    thke: np.ndarray     # (ne,) initial thickness
    iadc: np.ndarray     # (ne,4) int64 skyline rows
    mat: Law2
    prop: ShellProp
    dtfac: float = 0.9
    ids: np.ndarray | None = None  # (ne,) element user ids (None: index + 1)

    @property
    def ne(self) -> int:
        return int(self.conn.shape[0])

    def params(self) -> dict:
        return shell_params(self.mat, self.prop, self.dtfac)

    def initial_state(self) -> dict:
        """GBUF/LBUF initial state: THK = THKE, OFF = 1, everything else zero."""
        ne, npt = self.ne, self.prop.npt
        return dict(offg=np.ones(ne), thk=self.thke.astype(np.float64).copy(), forc=np.zeros((5, ne)),
                    mom=np.zeros((3, ne)), gstr=np.zeros((8, ne)), hourg=np.zeros((5, ne)),
                    eint=np.zeros((2, ne)), epsd=np.zeros(ne), sig=np.zeros((npt, 5, ne)),
                    pla=np.zeros((npt, ne)), epsd_ip=np.zeros((npt, ne)))


@dataclasses.dataclass
class Model:
    """A hex and/or Q4 model in the Engine's post-Starter form (0-based indices).

    Skyline rows of a node are ordered as FILLCNE orders them: solids before
    shells, inside a family by ascending element id (domdec2.F:2046, 2166-2317).
    """

    x: np.ndarray
    v: np.ndarray
    vr: np.ndarray
    ms: np.ndarray
    inr: np.ndarray          # nodal rotational inertia IN (zero for hex-only nodes)
    code: np.ndarray         # (nn,) int32 BC / imposed-velocity bits
    vimp: np.ndarray         # (nn,3) imposed velocities
    adsky: np.ndarray        # (nn+1,) int64
    hexp: HexPart | None = None
    q4: Q4Part | None = None
    dtmx: float = EP20

    @property
    def nn(self) -> int:
        return int(self.x.shape[0])

    @property
    def lsky(self) -> int:
        return int(self.adsky[-1])

    @property
    def ne_hex(self) -> int:
        return 0 if self.hexp is None else self.hexp.ne

    @property
    def ne_q4(self) -> int:
        return 0 if self.q4 is None else self.q4.ne

    @property
    def rotational(self) -> bool:
        return self.q4 is not None


def id_rank(ids, ne: int) -> np.ndarray:
    """Position of each element in ascending user-id order (the order MY_ORDERS gives FILLCNE
    and SPMD_MSIN, domdec2.F:2138-2166 and spmd_msin.F:95-99); element index order without ids."""
    if ids is None:
        return np.arange(ne, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.int64)
    if ids.shape != (ne,):
        raise ValueError("one user id per element")
    rank = np.empty(ne, dtype=np.int64)
    rank[np.argsort(ids, kind="stable")] = np.arange(ne, dtype=np.int64)
    return rank


def mixed_skyline(nn: int, hex_conn=None, q4_conn=None, hex_ids=None, q4_ids=None):
    """FILLCNE-ordered skyline over hex then Q4 contributions (solids before shells, each family
    by ascending user id, domdec2.F:2166-2240); returns (iads, iadc, adsky)."""
    nodes, ranks, sizes = [], [], []
    base = 0
    for conn, ids in ((hex_conn, hex_ids), (q4_conn, q4_ids)):
        if conn is None or conn.shape[0] == 0:
            sizes.append(None)
            continue
        ne, nk = conn.shape
        nodes.append(conn.reshape(-1).astype(np.int64))
        ranks.append(np.repeat(base + id_rank(ids, ne), nk))
        sizes.append((ne, nk))
        base += ne
    node = np.concatenate(nodes)
    rank = np.concatenate(ranks)
    order = np.lexsort((rank, node))
    counts = np.bincount(node, minlength=nn).astype(np.int64)
    adsky = np.zeros(nn + 1, dtype=np.int64)
    np.cumsum(counts, out=adsky[1:])
    rows = np.empty(node.shape[0], dtype=np.int64)
    rows[order] = np.arange(node.shape[0], dtype=np.int64)
    out, off = [], 0
    for sz in sizes:
        if sz is None:
            out.append(None)
            continue
        ne, nk = sz
        out.append(rows[off:off + ne * nk].reshape(ne, nk))
        off += ne * nk
    return out[0], out[1], adsky


def q4_area(xyz: np.ndarray) -> np.ndarray:
    """Area of Q4 elements (ne,4,3): half the norm of the diagonal cross product."""
    d13 = xyz[:, 2] - xyz[:, 0]
    d24 = xyz[:, 3] - xyz[:, 1]
    return 0.5 * np.linalg.norm(np.cross(d13, d24), axis=1)


def build_model(x, hex_conn=None, hex_mat: Law2 | None = None, hex_prop: SolidProp | None = None,
                q4_conn=None, q4_mat: Law2 | None = None, q4_prop: ShellProp | None = None, thke=None,
                v0=None, vr0=None, code=None, vimp=None, dtfac_hex: float = 0.9, dtfac_q4: float = 0.9,
                dtmx: float = EP20, hex_ids=None, q4_ids=None) -> Model:
    """Starter-equivalent builder: volumes, SMASS3 / CINMAS lumped masses and
    inertias, FILLCNE skyline, BC codes."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nn = x.shape[0]
    hex_conn = None if hex_conn is None else np.ascontiguousarray(hex_conn, dtype=np.int32)
    q4_conn = None if q4_conn is None else np.ascontiguousarray(q4_conn, dtype=np.int32)
    iads, iadc, adsky = mixed_skyline(nn, hex_conn, q4_conn, hex_ids, q4_ids)
    ms = np.zeros(nn)
    inr = np.zeros(nn)
    hexp = q4p = None
    # nodal masses are summed in skyline order (rows per node, FILLCNE order)
    rows_m = np.zeros(int(adsky[-1]))
    rows_i = np.zeros(int(adsky[-1]))
    if hex_conn is not None and hex_conn.shape[0]:
        vol0 = hex_volume(x[hex_conn])
        rho = hex_mat.params()["rho0"]
        mass = (1.0 * rho) * vol0 * (1.0 / 8.0)          # smass3.F: RHO*VOLU*ONE_OVER_8
        rows_m[iads.reshape(-1)] = np.repeat(mass, 8)
        hexp = HexPart(conn=hex_conn, vol0=vol0, iads=iads, mat=hex_mat, prop=hex_prop or SolidProp(),
                       dtfac=dtfac_hex, ids=None if hex_ids is None else np.asarray(hex_ids, np.int64))
    if q4_conn is not None and q4_conn.shape[0]:
        ne = q4_conn.shape[0]
        thke = np.full(ne, q4_prop.thick) if thke is None else np.asarray(thke, dtype=np.float64)
        area = q4_area(x[q4_conn])
        ems = q4_mat.rho * thke * area * 0.25              # cinmas.F:851
        xi = ems * (area / 9.0 + thke * thke * (1.0 / 12.0))  # cinmas.F:1383 (FAC=9, no offset)
        rows_m[iadc.reshape(-1)] = np.repeat(ems, 4)
        rows_i[iadc.reshape(-1)] = np.repeat(xi, 4)
        q4p = Q4Part(conn=q4_conn, thke=thke, iadc=iadc, mat=q4_mat, prop=q4_prop, dtfac=dtfac_q4,
                     ids=None if q4_ids is None else np.asarray(q4_ids, np.int64))
    counts = np.diff(adsky)
    for j in range(int(counts.max()) if nn else 0):
        has = counts > j
        ms[has] = ms[has] + rows_m[adsky[:-1][has] + j]
        inr[has] = inr[has] + rows_i[adsky[:-1][has] + j]
    v = np.zeros((nn, 3)) if v0 is None else np.array(v0, dtype=np.float64).reshape(nn, 3)
    vr = np.zeros((nn, 3)) if vr0 is None else np.array(vr0, dtype=np.float64).reshape(nn, 3)
    code = np.zeros(nn, np.int32) if code is None else np.asarray(code, dtype=np.int32)
    vimp = np.zeros((nn, 3)) if vimp is None else np.asarray(vimp, dtype=np.float64).reshape(nn, 3)
    return Model(x=x, v=v, vr=vr, ms=ms, inr=inr, code=code, vimp=vimp, adsky=adsky, hexp=hexp, q4=q4p,
                 dtmx=dtmx)


def hex_model_to_model(m: HexModel) -> Model:
    """A HexModel as a general Model (same skyline, masses and BCs)."""
    nn = m.nn
    code = np.asarray(m.bcs, np.int32).copy()
    vimp = np.zeros((nn, 3))
    bit = np.array([IMP_X, IMP_Y, IMP_Z], np.int32)
    for n, dof, val in zip(m.imp_node, m.imp_dof, m.imp_val):
        code[n] |= bit[dof]
        vimp[n, dof] = val
    hexp = HexPart(conn=m.conn, vol0=m.vol0, iads=m.iads, mat=m.mat, prop=m.prop, dtfac=m.dtfac)
    return Model(x=m.x, v=m.v, vr=np.zeros((nn, 3)), ms=m.ms, inr=np.zeros(nn), code=code, vimp=vimp,
                 adsky=m.adsky, hexp=hexp, q4=None, dtmx=m.dtmx)


def q4_grid(nu: int, nv: int, origin, du, dv):
    """Structured nu x nv quad grid spanned by vectors du, dv (per cell); node (i,j) -> i + (nu+1)*j."""
    i, j = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1), indexing="xy")
    x = (np.asarray(origin, float)[None, :] + i.reshape(-1, 1) * np.asarray(du, float)[None, :]
         + j.reshape(-1, 1) * np.asarray(dv, float)[None, :])
    ei, ej = np.meshgrid(np.arange(nu), np.arange(nv), indexing="xy")
    ei, ej = ei.reshape(-1), ej.reshape(-1)
    nid = lambda a, b: a + (nu + 1) * b  # noqa: E731
    conn = np.stack([nid(ei, ej), nid(ei + 1, ej), nid(ei + 1, ej + 1), nid(ei, ej + 1)], axis=1).astype(np.int32)
    return x, conn


ALUMINIUM = Law2(rho=2.7e-9, young=70000.0, nu=0.33, a=200.0, b=150.0, n=0.3)


def square_tube(width: float = 60.0, length: float = 300.0, h: float = 1.0, thick: float = 1.5,
                v_top: float = -1.5e4, mat: Law2 = ALUMINIUM, prop: ShellProp | None = None,
                dtfac: float = 0.9) -> Model:
    """Thin-walled square tube along z: base (z=0) fully fixed, top (z=length)
    given the constant axial velocity v_top (mm/s), Q4 shells of edge h."""
    nside = int(round(width / h))
    nper = 4 * nside
    nlen = int(round(length / h))
    # perimeter points, counter-clockwise seen from +z, starting at (-w/2,-w/2)
    t = np.arange(nper) * h
    px = np.empty(nper)
    py = np.empty(nper)
    w2 = 0.5 * width
    for k in range(nper):
        s = t[k]
        side, r = divmod(s, width)
        side = int(side)
        if side == 0:
            px[k], py[k] = -w2 + r, -w2
        elif side == 1:
            px[k], py[k] = w2, -w2 + r
        elif side == 2:
            px[k], py[k] = w2 - r, w2
        else:
            px[k], py[k] = -w2, w2 - r
    zz = np.arange(nlen + 1) * h
    x = np.zeros(((nlen + 1) * nper, 3))
    for j in range(nlen + 1):
        x[j * nper:(j + 1) * nper, 0] = px
        x[j * nper:(j + 1) * nper, 1] = py
        x[j * nper:(j + 1) * nper, 2] = zz[j]
    ei, ej = np.meshgrid(np.arange(nper), np.arange(nlen), indexing="xy")
    ei, ej = ei.reshape(-1), ej.reshape(-1)
    n0 = ej * nper + ei
    n1 = ej * nper + (ei + 1) % nper
    conn = np.stack([n0, n1, n1 + nper, n0 + nper], axis=1).astype(np.int32)
    nn = x.shape[0]
    code = np.zeros(nn, np.int32)
    vimp = np.zeros((nn, 3))
    code[x[:, 2] == 0.0] = BC_ALL
    top = x[:, 2] == zz[-1]
    code[top] |= IMP_Z
    vimp[top, 2] = v_top
    prop = prop or ShellProp(thick=thick)
    return build_model(x, q4_conn=conn, q4_mat=mat, q4_prop=prop, code=code, vimp=vimp, dtfac_q4=dtfac)


def flat_plate(nx: int, ny: int, h: float = 1.0, thick: float = 2.0, z: float = 0.0, origin=(0.0, 0.0),
               mat: Law2 = STEEL, prop: ShellProp | None = None):
    """(x, conn) of a flat nx x ny Q4 plate in the plane z."""
    x, conn = q4_grid(nx, ny, (origin[0], origin[1], z), (h, 0.0, 0.0), (0.0, h, 0.0))
    return x, conn


def merge_nodes(xa: np.ndarray, xb: np.ndarray, conn_b: np.ndarray, tol: float = 1e-9):
    """Append the nodes of mesh b to mesh a, sharing nodes that coincide (within tol).

    Returns (x, conn_b renumbered into x)."""
    key = {tuple(np.round(p / tol).astype(np.int64)): i for i, p in enumerate(xa)}
    remap = np.array([key.get(tuple(np.round(p / tol).astype(np.int64)), -1) for p in xb], dtype=np.int64)
    keep = remap < 0
    newid = remap.copy()
    newid[keep] = xa.shape[0] + np.arange(int(keep.sum()))
    return np.vstack([xa, xb[keep]]), newid[conn_b].astype(np.int32)


def tie_q4_on_hex(xh, ch, xp, cp, rng, v0: float = -5.0e3, noise: float = 300.0, thick: float = 2.0) -> Model:
    """Hex block with a Q4 plate sharing the block's coincident nodes; base (z=0) fixed;
    initial velocity v0 along z plus Gaussian noise."""
    x, cq = merge_nodes(xh, xp, cp)
    nn = x.shape[0]
    v = np.zeros((nn, 3))
    v[:, 2] = v0
    v += noise * rng.standard_normal((nn, 3))
    code = np.where(x[:, 2] == 0.0, BC_ALL, 0).astype(np.int32)
    return build_model(x, hex_conn=ch, hex_mat=STEEL, q4_conn=cq, q4_mat=STEEL, q4_prop=ShellProp(thick=thick),
                       v0=v, code=code)


def block_with_plate(n: int = 160, plate: int = 1000, h: float = 1.0, thick: float = 2.0, seed: int = 20261203,
                     v0: float = 1.0e4, sigma: float = 500.0, npt: int = 5) -> Model:
    """The GPU-filling roofline model: an n^3 hex block (edge h) with an plate x plate Q4 plate
    (edge h, thickness thick) lying on its top face, the block's top nodes shared with the plate.
    Law-2 steel for both; initial velocity v0 along z plus N(0, sigma) per component (seeded)."""
    xh, ch = hex_block(n, n, n, h)
    nhex_nodes = xh.shape[0]
    off = (plate - n) // 2
    i, j = np.meshgrid(np.arange(plate + 1), np.arange(plate + 1), indexing="xy")
    i, j = i.reshape(-1), j.reshape(-1)
    bi, bj = i - off, j - off
    shared = (bi >= 0) & (bi <= n) & (bj >= 0) & (bj <= n)
    pid = np.empty(i.shape[0], dtype=np.int64)
    # block node (ix,iy,iz) -> ix + (n+1)*(iy + (n+1)*iz); top face iz = n
    pid[shared] = bi[shared] + (n + 1) * (bj[shared] + (n + 1) * n)
    nnew = int((~shared).sum())
    pid[~shared] = nhex_nodes + np.arange(nnew)
    xp = np.zeros((nnew, 3))
    xp[:, 0] = (i[~shared] - off) * h
    xp[:, 1] = (j[~shared] - off) * h
    xp[:, 2] = n * h
    x = np.vstack([xh, xp])
    ei, ej = np.meshgrid(np.arange(plate), np.arange(plate), indexing="xy")
    ei, ej = ei.reshape(-1), ej.reshape(-1)
    g = lambda a, b: pid[a + (plate + 1) * b]  # noqa: E731
    cq = np.stack([g(ei, ej), g(ei + 1, ej), g(ei + 1, ej + 1), g(ei, ej + 1)], axis=1).astype(np.int32)
    rng = np.random.default_rng(seed)
    v = sigma * rng.standard_normal(x.shape)
    v[:, 2] += v0
    return build_model(x, hex_conn=ch, hex_mat=STEEL, q4_conn=cq, q4_mat=STEEL,
                       q4_prop=ShellProp(thick=thick, npt=npt), v0=v)


# ---------------------------------------------------------------------------
# Parametric sweep of tubes (many uneven models advanced together)
# ---------------------------------------------------------------------------
SWEEP_SIZES = (1.0, 1.5, 2.0, 3.0, 4.0)


def tube_sweep(n: int = 256, seed: int = 20261203, width: float = 60.0, length: float = 300.0,
               sizes=SWEEP_SIZES) -> list:
    """n variants of the square-tube crush: element size uniform on `sizes` (mm), thickness
    U(1.0, 2.5) mm, crush velocity U(5, 20) m/s, yield A U(150, 300) MPa (aluminium otherwise),
    drawn in that order per variant from one PCG64 stream.  Returns [(Model, params dict)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        h = float(sizes[int(rng.integers(0, len(sizes)))])
        thick = float(rng.uniform(1.0, 2.5))
        vel = float(rng.uniform(5.0e3, 2.0e4))
        a = float(rng.uniform(150.0, 300.0))
        mat = dataclasses.replace(ALUMINIUM, a=a)
        m = square_tube(width=width, length=length, h=h, thick=thick, v_top=-vel, mat=mat)
        out.append((m, dict(h=h, thick=thick, v=vel, a=a)))
    return out
