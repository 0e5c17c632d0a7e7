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
