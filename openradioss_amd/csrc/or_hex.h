// or_hex.h -- device code of the 8-node one-point hex (SFORC3 chain) for gfx950.
//
// One thread owns one element.  The arithmetic of every step follows the
// reference routine it names, operation for operation, so that with
// -ffp-contract=off the FP64 instantiation rounds exactly like the reference
// CPU engine (pow() is evaluated in double-double and correctly rounded except in
// rare near-tie cases, see or_math.h).
// The split instantiation (R = float) keeps coordinates and their differences
// in FP64 and runs strain, stress and hourglass arithmetic in FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "or_math.h"

namespace orgpu {

constexpr double kEP06 = 1.0e6;
constexpr double kEM15 = 1.0 / 1.0e15;
constexpr double kEM20 = 1.0 / 1.0e20;

// Resolved law-2 / property constants handed to the kernels by value.
struct HexConsts {
  double rho0, bulk, shear, ca, cb, cn, epmx, sigmx, cc, epdr, fisokin, asrate, m_exp;
  int32_t icc, vp, israte, iform;
  double qa, qb, qh, vns1, vns2, dtfac, dtmx;
  double onep333;   // ONEP333 of constant_mod.F, computed on the host
  double dt_growth; // the single-precision literal 1.1 of resol.F:6352, widened
};

// Device-resident cycle scalars, double-buffered by cycle parity.
struct DevTime {
  double tt, dt1, dt2, dt12, dt2old;
  int64_t ncycle;
  int32_t mstop;
  int32_t done; // the model has stopped: end time reached or MSTOP
};

template <typename T> __device__ __forceinline__ T dmaxT(T a, T b) { return (a > b) ? a : b; }
template <typename T> __device__ __forceinline__ T dminT(T a, T b) { return (a < b) ? a : b; }

// FP64: the correctly rounded double-double routines of or_math.h; FP32 (split
// form): the device library.
__device__ __forceinline__ double powR(double a, double b) { return ormath::or_pow(a, b); }
__device__ __forceinline__ float powR(float a, float b) { return powf(a, b); }
// a**b1 and a**b2 (FP64: one shared logarithm, the results of two powR calls)
__device__ __forceinline__ void pow2R(double a, double b1, double b2, double &r1, double &r2) {
  ormath::or_pow2(a, b1, b2, r1, r2);
}
__device__ __forceinline__ void pow2R(float a, float b1, float b2, float &r1, float &r2) {
  r1 = powf(a, b1);
  r2 = powf(a, b2);
}
__device__ __forceinline__ double sqrtR(double a) { return sqrt(a); }
__device__ __forceinline__ float sqrtR(float a) { return sqrtf(a); }
__device__ __forceinline__ double logR(double a) { return ormath::or_log(a); }
__device__ __forceinline__ float logR(float a) { return logf(a); }
__device__ __forceinline__ double fabsR(double a) { return fabs(a); }
__device__ __forceinline__ float fabsR(float a) { return fabsf(a); }

// SLEN (slen.F:30): E*G-F^2 of one face.
template <typename R>
__device__ __forceinline__ R face_measure(R x1, R x2, R x3, R x4, R y1, R y2, R y3, R y4, R z1, R z2, R z3,
                                          R z4) {
  R x13 = x3 - x1, x24 = x4 - x2;
  R y13 = y3 - y1, y24 = y4 - y2;
  R z13 = z3 - z1, z24 = z4 - z2;
  R fs1 = x13 - x24, ft1 = x13 + x24;
  R fs2 = y13 - y24, ft2 = y13 + y24;
  R fs3 = z13 - z24, ft3 = z13 + z24;
  R e = fs1 * fs1 + fs2 * fs2 + fs3 * fs3;
  R f = fs1 * ft1 + fs2 * ft2 + fs3 * ft3;
  R g = ft1 * ft1 + ft2 * ft2 + ft3 * ft3;
  return e * g - f * f;
}

// Per-element state in / out (storage type S).
template <typename S> struct HexState {
  S sig[6];
  S eint, rho, vol0, pla, epsd, offg, qvis;
};

// Scalar output of one element: STI/4, DTX (or +EP06 when excluded), MSTOP, EHOU.  The
// 8 nodal force triples leave through the caller's write(node, fx, fy, fz).
template <typename R> struct HexOut {
  R sti4;
  double dtx;
  int mstop;
  R ehou;
};

// Receives an element's EHOU as soon as it is known (the default ignores it).
struct NoEhou {
  template <typename T> __device__ __forceinline__ void operator()(T) const {}
};

// SHVIS3 hourglass force of node j (0..7) in the three directions, H(d, 1..4) per direction:
// F = -G(j,1)*H1 - G(j,2)*H2 - G(j,3)*H3 -/+ H4 with G(j,m) = +-1 -/+ PXH(m, P(j)) as SHVIS3
// forms them (P = 1,2,3,4,3,4,1,2; minus for nodes 1-4, plus for nodes 5-8).
template <typename R>
__device__ __forceinline__ void hg_node_force3(int j, const R (&PXH)[3][4], const R (&H)[3][4], R (&f)[3]) {
  constexpr signed char kS[3][8] = {{1, -1, 1, -1, 1, -1, 1, -1},
                                    {1, 1, -1, -1, -1, -1, 1, 1},
                                    {1, -1, -1, 1, -1, 1, 1, -1}};
  constexpr int kP[8] = {0, 1, 2, 3, 2, 3, 0, 1};
  constexpr signed char kH4[8] = {-1, 1, -1, 1, 1, -1, 1, -1}; // sign of H4
  R G[3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
    G[m] = (j < 4) ? R(kS[m][j]) - PXH[m][kP[j]] : R(kS[m][j]) + PXH[m][kP[j]];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const R a = -G[0] * H[d][0] - G[1] * H[d][1] - G[2] * H[d][2];
    f[d] = kH4[j] > 0 ? a + H[d][3] : a - H[d][3];
  }
}

// The SFORC3 chain of one element in three phases with the operation order of the
// reference: geometry (coordinates only), kinematics (velocities), then material
// and forces (element state).  A kernel can load each phase's inputs just before
// the phase, which keeps fewer values live at once; hex_element runs all three.

// Geometry: SCOOR3 OFF, SDERI3, SCHKJABT3, the PXiHj projections, SDLEN3.
template <typename R> struct HexGeom {
  R PX1, PX2, PX3, PX4, PY1, PY2, PY3, PY4, PZ1, PZ2, PZ3, PZ4;
  R PXH[3][4];
  R VOLN, DELTAX;
  R off, offg;
  bool zero_vel; // OFFG < 0 on entry: SCOOR3 zeroes the velocities
  int mstop;
};

// Kinematics: SDEFO3 (JCVT=0) and the hourglass velocity projections of SHVIS3.
template <typename R> struct HexKin {
  R DXX, DYY, DZZ, D4, D5, D6, WXX, WYY, WZZ;
  R HG[3][4]; // [x,y,z][mode]
};

// xd/yd/zd: gathered coordinates (FP64, X in the reference); offg_in: GBUF%OFF.
template <typename R>
__device__ __forceinline__ void hex_geometry(const double (&xd)[8], const double (&yd)[8], const double (&zd)[8],
                                             double offg_in, HexGeom<R> &g) {
  const R ZERO = R(0), ONE = R(1), FOUR = R(4);
  const R ONE_OVER_64 = R(1.0 / 64.0);
  const R EM20 = R(kEM20);
  g.mstop = 0;
  // ---- SCOOR3 (scoor3.F:233-258 OFF; 327-340 OFFG<0 zero velocities) ----
  R offg = R(offg_in);
  R off = fabsR(offg);
  g.zero_vel = offg < ZERO;
  // ---- SDERI3 (sderi3.F): differences in FP64, Jacobian terms in R ----
  R JAC1, JAC2, JAC3, JAC4, JAC5, JAC6, JAC7, JAC8, JAC9;
  {
    double X17 = xd[6] - xd[0], X28 = xd[7] - xd[1], X35 = xd[4] - xd[2], X46 = xd[5] - xd[3];
    double Y17 = yd[6] - yd[0], Y28 = yd[7] - yd[1], Y35 = yd[4] - yd[2], Y46 = yd[5] - yd[3];
    double Z17 = zd[6] - zd[0], Z28 = zd[7] - zd[1], Z35 = zd[4] - zd[2], Z46 = zd[5] - zd[3];
    JAC1 = R(X17 + X28 - X35 - X46);
    JAC2 = R(Y17 + Y28 - Y35 - Y46);
    JAC3 = R(Z17 + Z28 - Z35 - Z46);
    R X_17_46 = R(X17 + X46), X_28_35 = R(X28 + X35);
    R Y_17_46 = R(Y17 + Y46), Y_28_35 = R(Y28 + Y35);
    R Z_17_46 = R(Z17 + Z46), Z_28_35 = R(Z28 + Z35);
    JAC4 = X_17_46 + X_28_35; JAC5 = Y_17_46 + Y_28_35; JAC6 = Z_17_46 + Z_28_35;
    JAC7 = X_17_46 - X_28_35; JAC8 = Y_17_46 - Y_28_35; JAC9 = Z_17_46 - Z_28_35;
  }
  R JAC_59_68 = JAC5 * JAC9 - JAC6 * JAC8;
  R JAC_67_49 = JAC6 * JAC7 - JAC4 * JAC9;
  R JAC_48_57 = JAC4 * JAC8 - JAC5 * JAC7;
  R DET = ONE_OVER_64 * (JAC1 * JAC_59_68 + JAC2 * JAC_67_49 + JAC3 * JAC_48_57);
  // SCHKJABT3 (solide4/schkjabt3.F): JLAG, ISMSTR=4, INEG_V=0 -> delete + MSTOP
  if (off == ZERO) {
    DET = ONE;
  } else if (offg > ONE) {
  } else if (DET <= ZERO) {
    DET = ONE; off = ZERO; offg = ZERO; g.mstop = 1;
  }
  R &PX1 = g.PX1, &PX2 = g.PX2, &PX3 = g.PX3, &PX4 = g.PX4;
  R &PY1 = g.PY1, &PY2 = g.PY2, &PY3 = g.PY3, &PY4 = g.PY4;
  R &PZ1 = g.PZ1, &PZ2 = g.PZ2, &PZ3 = g.PZ3, &PZ4 = g.PZ4;
  {
    R DETT = ONE_OVER_64 / DET;
    R JACI1 = DETT * JAC_59_68, JACI4 = DETT * JAC_67_49, JACI7 = DETT * JAC_48_57;
    R JACI2 = DETT * (JAC3 * JAC8 - JAC2 * JAC9);
    R JACI5 = DETT * (JAC1 * JAC9 - JAC3 * JAC7);
    R JACI8 = DETT * (JAC2 * JAC7 - JAC1 * JAC8);
    R JACI3 = DETT * (JAC2 * JAC6 - JAC3 * JAC5);
    R JACI6 = DETT * (JAC3 * JAC4 - JAC1 * JAC6);
    R JACI9 = DETT * (JAC1 * JAC5 - JAC2 * JAC4);
    R JAC12 = JACI1 + JACI2, JAC45 = JACI4 + JACI5, JAC78 = JACI7 + JACI8;
    PX1 = -JAC12 - JACI3; PY1 = -JAC45 - JACI6; PZ1 = -JAC78 - JACI9;
    PX2 = -JAC12 + JACI3; PY2 = -JAC45 + JACI6; PZ2 = -JAC78 + JACI9;
    JAC12 = JACI1 - JACI2; JAC45 = JACI4 - JACI5; JAC78 = JACI7 - JACI8;
    PX3 = JAC12 + JACI3; PY3 = JAC45 + JACI6; PZ3 = JAC78 + JACI9;
    PX4 = JAC12 - JACI3; PY4 = JAC45 - JACI6; PZ4 = JAC78 - JACI9;
  }
  // hourglass projections PXiHj (JHBE/=0); the H vectors are FP64 sums, then rounded to R
  R (&PXH)[3][4] = g.PXH;
  {
    double hx, hy, hz;
    hx = (xd[0] - xd[1] + xd[2] - xd[3] + xd[4] - xd[5] + xd[6] - xd[7]);
    hy = (yd[0] - yd[1] + yd[2] - yd[3] + yd[4] - yd[5] + yd[6] - yd[7]);
    hz = (zd[0] - zd[1] + zd[2] - zd[3] + zd[4] - zd[5] + zd[6] - zd[7]);
    R HX = R(hx), HY = R(hy), HZ = R(hz);
    PXH[0][0] = PX1 * HX + PY1 * HY + PZ1 * HZ;
    PXH[0][1] = PX2 * HX + PY2 * HY + PZ2 * HZ;
    PXH[0][2] = PX3 * HX + PY3 * HY + PZ3 * HZ;
    PXH[0][3] = PX4 * HX + PY4 * HY + PZ4 * HZ;
    hx = (xd[0] + xd[1] - xd[2] - xd[3] - xd[4] - xd[5] + xd[6] + xd[7]);
    hy = (yd[0] + yd[1] - yd[2] - yd[3] - yd[4] - yd[5] + yd[6] + yd[7]);
    hz = (zd[0] + zd[1] - zd[2] - zd[3] - zd[4] - zd[5] + zd[6] + zd[7]);
    HX = R(hx); HY = R(hy); HZ = R(hz);
    PXH[1][0] = PX1 * HX + PY1 * HY + PZ1 * HZ;
    PXH[1][1] = PX2 * HX + PY2 * HY + PZ2 * HZ;
    PXH[1][2] = PX3 * HX + PY3 * HY + PZ3 * HZ;
    PXH[1][3] = PX4 * HX + PY4 * HY + PZ4 * HZ;
    hx = (xd[0] - xd[1] - xd[2] + xd[3] - xd[4] + xd[5] + xd[6] - xd[7]);
    hy = (yd[0] - yd[1] - yd[2] + yd[3] - yd[4] + yd[5] + yd[6] - yd[7]);
    hz = (zd[0] - zd[1] - zd[2] + zd[3] - zd[4] + zd[5] + zd[6] - zd[7]);
    HX = R(hx); HY = R(hy); HZ = R(hz);
    PXH[2][0] = PX1 * HX + PY1 * HY + PZ1 * HZ;
    PXH[2][1] = PX2 * HX + PY2 * HY + PZ2 * HZ;
    PXH[2][2] = PX3 * HX + PY3 * HY + PZ3 * HZ;
    PXH[2][3] = PX4 * HX + PY4 * HY + PZ4 * HZ;
  }
  R VOLN = DET;
  // ---- SDLEN3 / SLEN: characteristic length (coordinates rounded to R like X1..Z8) ----
  R &DELTAX = g.DELTAX;
  {
    R x[8], y[8], z[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { x[k] = R(xd[k]); y[k] = R(yd[k]); z[k] = R(zd[k]); }
    R AREAM = EM20, ar;
    ar = face_measure(x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3], z[0], z[1], z[2], z[3]);
    AREAM = dmaxT(ar, AREAM);
    ar = face_measure(x[4], x[5], x[6], x[7], y[4], y[5], y[6], y[7], z[4], z[5], z[6], z[7]);
    AREAM = dmaxT(ar, AREAM);
    ar = face_measure(x[0], x[1], x[5], x[4], y[0], y[1], y[5], y[4], z[0], z[1], z[5], z[4]);
    AREAM = dmaxT(ar, AREAM);
    ar = face_measure(x[1], x[2], x[6], x[5], y[1], y[2], y[6], y[5], z[1], z[2], z[6], z[5]);
    AREAM = dmaxT(ar, AREAM);
    ar = face_measure(x[2], x[3], x[7], x[6], y[2], y[3], y[7], y[6], z[2], z[3], z[7], z[6]);
    AREAM = dmaxT(ar, AREAM);
    ar = face_measure(x[3], x[0], x[4], x[7], y[3], y[0], y[4], y[7], z[3], z[0], z[4], z[7]);
    AREAM = dmaxT(ar, AREAM);
    DELTAX = FOUR * VOLN * ONE / sqrtR(AREAM);
  }
  g.VOLN = VOLN;
  g.off = off;
  g.offg = offg;
}

// vx/vy/vz: gathered velocities (zeroed here when SCOOR3 asks for it).
template <typename R>
__device__ __forceinline__ void hex_kinematics(R (&vx)[8], R (&vy)[8], R (&vz)[8], const HexGeom<R> &g, const R DT1,
                                               HexKin<R> &kn) {
  const R ZERO = R(0), ONE = R(1), HALF = R(1.0 / 2.0);
  if (g.zero_vel) {
#pragma unroll
    for (int k = 0; k < 8; ++k) { vx[k] = ZERO; vy[k] = ZERO; vz[k] = ZERO; }
  }
  const R PX1 = g.PX1, PX2 = g.PX2, PX3 = g.PX3, PX4 = g.PX4;
  const R PY1 = g.PY1, PY2 = g.PY2, PY3 = g.PY3, PY4 = g.PY4;
  const R PZ1 = g.PZ1, PZ2 = g.PZ2, PZ3 = g.PZ3, PZ4 = g.PZ4;
  const R(&PXH)[3][4] = g.PXH;
  // ---- SDEFO3 (JCVT=0, JHBE=1 branch) ----
  R &DXX = kn.DXX, &DYY = kn.DYY, &DZZ = kn.DZZ, &D4 = kn.D4, &D5 = kn.D5, &D6 = kn.D6;
  R &WXX = kn.WXX, &WYY = kn.WYY, &WZZ = kn.WZZ;
  {
    R VX17 = vx[0] - vx[6], VX28 = vx[1] - vx[7], VX35 = vx[2] - vx[4], VX46 = vx[3] - vx[5];
    R VY17 = vy[0] - vy[6], VY28 = vy[1] - vy[7], VY35 = vy[2] - vy[4], VY46 = vy[3] - vy[5];
    R VZ17 = vz[0] - vz[6], VZ28 = vz[1] - vz[7], VZ35 = vz[2] - vz[4], VZ46 = vz[3] - vz[5];
    DXX = PX1 * VX17 + PX2 * VX28 + PX3 * VX35 + PX4 * VX46;
    DYY = PY1 * VY17 + PY2 * VY28 + PY3 * VY35 + PY4 * VY46;
    DZZ = PZ1 * VZ17 + PZ2 * VZ28 + PZ3 * VZ35 + PZ4 * VZ46;
    R DXY = PY1 * VX17 + PY2 * VX28 + PY3 * VX35 + PY4 * VX46;
    R DXZ = PZ1 * VX17 + PZ2 * VX28 + PZ3 * VX35 + PZ4 * VX46;
    R DYX = PX1 * VY17 + PX2 * VY28 + PX3 * VY35 + PX4 * VY46;
    R DYZ = PZ1 * VY17 + PZ2 * VY28 + PZ3 * VY35 + PZ4 * VY46;
    R DZX = PX1 * VZ17 + PX2 * VZ28 + PX3 * VZ35 + PX4 * VZ46;
    R DZY = PY1 * VZ17 + PY2 * VZ28 + PY3 * VZ35 + PY4 * VZ46;
    R DT1D2 = HALF * DT1;
    D4 = DXY + DYX; D5 = DYZ + DZY; D6 = DXZ + DZX;
    WZZ = DT1D2 * (DYX - DXY);
    WYY = DT1D2 * (DXZ - DZX);
    WXX = DT1D2 * (DZY - DYZ);
  }
  // ---- hourglass velocity projections of SHVIS3 (JHBE>=1); they depend only on
  //      PXiHj and the velocities, so they are formed here to free the velocities ----
  R (&HG)[3][4] = kn.HG;
  {
    R G11 = ONE - PXH[0][0], G21 = -ONE - PXH[0][1], G31 = ONE - PXH[0][2], G41 = -ONE - PXH[0][3];
    R G51 = ONE + PXH[0][2], G61 = -ONE + PXH[0][3], G71 = ONE + PXH[0][0], G81 = -ONE + PXH[0][1];
    R G12 = ONE - PXH[1][0], G22 = ONE - PXH[1][1], G32 = -ONE - PXH[1][2], G42 = -ONE - PXH[1][3];
    R G52 = -ONE + PXH[1][2], G62 = -ONE + PXH[1][3], G72 = ONE + PXH[1][0], G82 = ONE + PXH[1][1];
    R G13 = ONE - PXH[2][0], G23 = -ONE - PXH[2][1], G33 = -ONE - PXH[2][2], G43 = ONE - PXH[2][3];
    R G53 = -ONE + PXH[2][2], G63 = ONE + PXH[2][3], G73 = ONE + PXH[2][0], G83 = -ONE + PXH[2][1];
#define OR_HG(V, out)                                                                                         \
  out[0] = G11 * V[0] + G21 * V[1] + G31 * V[2] + G41 * V[3] + G51 * V[4] + G61 * V[5] + G71 * V[6] + G81 * V[7]; \
  out[1] = G12 * V[0] + G22 * V[1] + G32 * V[2] + G42 * V[3] + G52 * V[4] + G62 * V[5] + G72 * V[6] + G82 * V[7]; \
  out[2] = G13 * V[0] + G23 * V[1] + G33 * V[2] + G43 * V[3] + G53 * V[4] + G63 * V[5] + G73 * V[6] + G83 * V[7]; \
  out[3] = V[0] - V[1] + V[2] - V[3] - V[4] + V[5] - V[6] + V[7];
    OR_HG(vx, HG[0]);
    OR_HG(vy, HG[1]);
    OR_HG(vz, HG[2]);
#undef OR_HG
  }
}

// Material and forces: SRHO3, SROTA3, MMAIN / M2LAW, MSTRAIN_RATE, MQVISCB, SMALLB3,
// SHVIS3 forces, SFINT3, SCUMU3P.  Updates the element state st and fills o.
template <typename R, typename S, typename Write, typename Ehou = NoEhou>
__device__ __forceinline__ void hex_material_forces(const HexGeom<R> &g, const HexKin<R> &kn, HexState<S> &st,
                                                    const HexConsts &c, const R DT1, HexOut<R> &o, Write &&write,
                                                    Ehou ehou_sink = Ehou{}) {
  const R ZERO = R(0), ONE = R(1), TWO = R(2), THREE = R(3);
  const R HALF = R(1.0 / 2.0), THIRD = R(1.0 / 3.0), FOURTH = R(1.0 / 4.0);
  const R TWO_THIRD = R(2.0 / 3.0), HUNDRED = R(100.0);
  const R EM15 = R(kEM15), EM20 = R(kEM20);
  o.mstop = g.mstop;
  const R PX1 = g.PX1, PX2 = g.PX2, PX3 = g.PX3, PX4 = g.PX4;
  const R PY1 = g.PY1, PY2 = g.PY2, PY3 = g.PY3, PY4 = g.PY4;
  const R PZ1 = g.PZ1, PZ2 = g.PZ2, PZ3 = g.PZ3, PZ4 = g.PZ4;
  const R(&PXH)[3][4] = g.PXH;
  const R(&HG)[3][4] = kn.HG;
  const R DXX = kn.DXX, DYY = kn.DYY, DZZ = kn.DZZ, D4 = kn.D4, D5 = kn.D5, D6 = kn.D6;
  const R WXX = kn.WXX, WYY = kn.WYY, WZZ = kn.WZZ;
  const R DELTAX = g.DELTAX;
  R VOLN = g.VOLN;
  const R off = g.off;
  R offg = g.offg;
  // ---- SRHO3 (Lagrangian) ----
  const R RHO0 = R(c.rho0);
  const R volo = R(st.vol0);
  if (offg == ZERO && VOLN == ONE) VOLN = volo;
  R rhon = R(st.rho);
  R DVOL = VOLN - (RHO0 / rhon) * volo;
  rhon = RHO0 * (volo / VOLN);
  R eint = R(st.eint) * volo;
  // ---- SROTA3: Jaumann rotation; S1..S6 keep the stress before rotation ----
  R S1 = R(st.sig[0]), S2 = R(st.sig[1]), S3 = R(st.sig[2]);
  R S4 = R(st.sig[3]), S5 = R(st.sig[4]), S6 = R(st.sig[5]);
  R sg0, sg1, sg2, sg3, sg4, sg5;
  {
    R Q1 = TWO * S4 * WZZ, Q2 = TWO * S6 * WYY, Q3 = TWO * S5 * WXX;
    sg0 = S1 - Q1 + Q2;
    sg1 = S2 + Q1 - Q3;
    sg2 = S3 - Q2 + Q3;
    sg3 = S4 + WZZ * (S1 - S2) + WYY * S5 - WXX * S6;
    sg4 = S5 + WXX * (S2 - S3) + WZZ * S6 - WYY * S4;
    sg5 = S6 + WYY * (S3 - S1) + WXX * S4 - WZZ * S5;
  }
  // ---- MMAIN glue for law 2 (mmain.F90:668-745) ----
  R vol_avg = VOLN - HALF * DVOL;
  R amu = rhon / RHO0 - ONE;
  R rhoref = rhon;
  // ---- M2LAW (m2law.F) ----
  R G = R(c.shear) * off;
  R SIGMX = R(c.sigmx);
  R P = -THIRD * (sg0 + sg1 + sg2);
  R DAV = -THIRD * (DXX + DYY + DZZ);
  R G1 = DT1 * G, G2 = TWO * G1;
  R SSP = sqrtR((R(c.onep333) * G + R(c.bulk)) / RHO0);
  sg0 = sg0 + P + G2 * (DXX + DAV);
  sg1 = sg1 + P + G2 * (DYY + DAV);
  sg2 = sg2 + P + G2 * (DZZ + DAV);
  sg3 = sg3 + G1 * D4;
  sg4 = sg4 + G1 * D5;
  sg5 = sg5 + G1 * D6;
  R AJ2 = HALF * (sg0 * sg0 + sg1 * sg1 + sg2 * sg2) + sg3 * sg3 + sg4 * sg4 + sg5 * sg5;
  AJ2 = sqrtR(THREE * AJ2);
  // MSTRAIN_RATE (mstrain_rate.F), IDEV = VP-2
  R epsd = R(st.epsd);
  {
    R epsdot;
    if (c.vp - 2 == 0) {
      R E1 = DXX, E2 = DYY, E3 = DZZ, E4 = HALF * D4, E5 = HALF * D5, E6 = HALF * D6;
      R epsp = E1 * E1 + E2 * E2 + E3 * E3 + TWO * (E4 * E4 + E5 * E5 + E6 * E6);
      epsdot = sqrtR(epsp);
    } else {
      R dav = (DXX + DYY + DZZ) * THIRD;
      R E1 = DXX - dav, E2 = DYY - dav, E3 = DZZ - dav;
      R E4 = HALF * D4, E5 = HALF * D5, E6 = HALF * D6;
      R epsp = HALF * (E1 * E1 + E2 * E2 + E3 * E3) + E4 * E4 + E5 * E5 + E6 * E6;
      epsdot = sqrtR(THREE * epsp) / R(3.0 / 2.0);
    }
    if (c.israte == 0) {
      epsd = epsdot;
    } else if (c.israte > 0) {
      R asr = dminT(ONE, R(c.asrate) * DT1);
      epsd = asr * epsdot + (ONE - asr) * epsd;
    }
  }
  // rate / thermal factor; TSTAR = 0 on this branch (no thermal), so 1 - TSTAR**MT = 1
  R EPD = ONE;
  if (c.cc != 0.0) {
    if (c.vp == 1) {
      EPD = dmaxT(epsd, R(c.epdr));
      EPD = logR(EPD / R(c.epdr));
    } else {
      EPD = dmaxT(epsd, EM15);
      EPD = logR(EPD / R(c.epdr));
    }
    EPD = dmaxT(ZERO, EPD);
    EPD = (ONE + R(c.cc) * EPD) * ONE;
  }
  if (c.icc == 1) SIGMX = R(c.sigmx) * EPD;
  R epxe = R(st.pla);
  R AK, QH;
  const R CB = R(c.cb), CN = R(c.cn);
  if (CN == ONE) {
    AK = R(c.ca) + CB * epxe;
    QH = CB * EPD;
  } else if (epxe > ZERO) {
    R pn, pq; // EPXE**CN, and EPXE**(CN-1) or EPXE**(1-CN)
    pow2R(epxe, CN, (CN > ONE) ? CN - ONE : ONE - CN, pn, pq);
    AK = R(c.ca) + CB * pn;
    if (CN > ONE) QH = (CB * CN * pq) * EPD;
    else QH = (CB * CN / pq) * EPD;
  } else {
    AK = R(c.ca);
    QH = ZERO;
  }
  AK = AK * EPD;
  if (SIGMX < AK) { AK = SIGMX; QH = ZERO; }
  if (epxe > R(c.epmx)) { AK = ZERO; QH = ZERO; }
  {
    R SCALE = dminT(ONE, AK / dmaxT(AJ2, EM15));
    R DPLA = (ONE - SCALE) * AJ2 / dmaxT(THREE * G + QH, EM15);
    AK = AK + (ONE - R(c.fisokin)) * DPLA * QH;
    SCALE = dminT(ONE, AK / dmaxT(AJ2, EM15));
    sg0 = SCALE * sg0; sg1 = SCALE * sg1; sg2 = SCALE * sg2;
    sg3 = SCALE * sg3; sg4 = SCALE * sg4; sg5 = SCALE * sg5;
    epxe = epxe + DPLA;
  }
  // ---- MQVISCB (IMPL=0, N2D=0, NPG=1, JSMS=0, VIS=VD2=0 on the Lagrangian branch) ----
  R QNEW, STI, vol23;
  double dtx_out;
  {
    R DD = -DXX - DYY - DZZ;
    R AD = ZERO, AL = ZERO;
    R CX = SSP + ZERO; // SSP + SQRT(VD2), VD2 = 0
    // AL = VOLN**(1/3) here and VOLN**(2/3) of SHVIS3 below, from one logarithm of VOLN
    if (off == ONE && VOLN > ZERO) pow2R(VOLN, ONE / THREE, TWO_THIRD, AL, vol23);
    else vol23 = powR(VOLN, TWO_THIRD);
    if (off == ONE) AD = dmaxT(ZERO, DD);
    R NRHO = sqrtR(rhoref * RHO0);
    R QA = ONE * R(c.qa), QB = ONE * R(c.qb);
    R CNS1_0 = ONE * R(c.vns1), CNS2_0 = ONE * R(c.vns2);
    R QAA_0 = QA * QA;
    R CNS1 = CNS1_0 * AL * NRHO * SSP * off;
    R CNS2 = CNS2_0 * AL * NRHO * SSP * off;
    R QAA = QAA_0 * AD;
    R QX = QB * SSP + AL * QAA + ONE * TWO * ZERO / dmaxT(EM20, rhon * DELTAX) +
           (CNS1 + ONE * CNS2) / dmaxT(EM20, rhoref * DELTAX);
    QNEW = rhon * AD * AL * (QAA * AL + QB * SSP);
    R SSP_EQ = dmaxT(EM20, QX + sqrtR(QX * QX + CX * CX));
    R DTX = DELTAX / SSP_EQ;
    STI = ZERO;
    if (!(off == ZERO || offg < ZERO)) {
      R TIDT = ONE / DTX;
      R TRHO = rhon * TIDT;
      R TVOL = VOLN * TIDT;
      STI = TRHO * TVOL;
    }
    DTX = R(c.dtfac) * DTX;
    dtx_out = kEP06;
    if (VOLN > ZERO && (off != ZERO && offg >= ZERO)) dtx_out = double(DTX);
  }
  // pressure and internal energy (m2law.F, IEOS == 0)
  {
    R PNEW = R(c.bulk) * amu;
    sg0 = (sg0 - PNEW) * off;
    sg1 = (sg1 - PNEW) * off;
    sg2 = (sg2 - PNEW) * off;
    sg3 = sg3 * off;
    sg4 = sg4 * off;
    sg5 = sg5 * off;
    R DTA = HALF * DT1;
    R E1 = DXX * (S1 + sg0);
    R E2 = DYY * (S2 + sg1);
    R E3 = DZZ * (S3 + sg2);
    R E4 = D4 * (S4 + sg3);
    R E5 = D5 * (S5 + sg4);
    R E6 = D6 * (S6 + sg5);
    R EINC = vol_avg * (E1 + E2 + E3 + E4 + E5 + E6) * DTA - HALF * DVOL * (R(st.qvis) + QNEW);
    eint = (eint + EINC * off) / dmaxT(EM15, volo);
  }
  st.sig[0] = S(sg0); st.sig[1] = S(sg1); st.sig[2] = S(sg2);
  st.sig[3] = S(sg3); st.sig[4] = S(sg4); st.sig[5] = S(sg5);
  st.eint = S(eint);
  st.rho = S(rhon);
  st.pla = S(epxe);
  st.epsd = S(epsd);
  st.qvis = S(QNEW);
  // ---- SMALLB3 (ISMSTR=4) ----
  if (off < ONE) {
    if (off == ZERO) offg = ZERO;
    else if (offg > ONE) offg = ONE + off;
    else offg = off;
  }
  st.offg = S(offg);
  // ---- SHVIS3 forces (INVSTR>=35: CAQ from GEO(13); FLUID=0; ISMSTR=4; JHBE>=1), SFINT3 and
  //      SCUMU3P, node pair by node pair: each force is the node's SHVIS3 hourglass force,
  //      from which SFINT3 subtracts (nodes 1-4) or to which it adds (nodes 7, 8, 5, 6) the
  //      stress term of the pair, as the two routines do in sequence; deleted elements
  //      (OFFG < 0) contribute zero.  Forces leave through write(node, fx, fy, fz). ----
  o.sti4 = FOURTH * STI;
  o.dtx = dtx_out;
  {
    R CAQ = FOURTH * off * R(c.qh);
    R FCL = CAQ * rhon * vol23;
    R FCQ = FCL * CAQ * HUNDRED;
    FCL = FCL * SSP;
    R H[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = 0; j < 4; ++j) H[d][j] = HG[d][j] * (FCL + fabsR(HG[d][j]) * FCQ);
    // EHOU order: z modes, then x, then y (shvis3.F)
    o.ehou = DT1 * (H[2][0] * HG[2][0] + H[2][1] * HG[2][1] + H[2][2] * HG[2][2] + H[2][3] * HG[2][3] +
                    H[0][0] * HG[0][0] + H[0][1] * HG[0][1] + H[0][2] * HG[0][2] + H[0][3] * HG[0][3] +
                    H[1][0] * HG[1][0] + H[1][1] * HG[1][1] + H[1][2] * HG[1][2] + H[1][3] * HG[1][3]);
    ehou_sink(o.ehou);
    R s1 = (sg0 - QNEW) * VOLN, s2 = (sg1 - QNEW) * VOLN, s3 = (sg2 - QNEW) * VOLN;
    R s4 = sg3 * VOLN, s5 = sg4 * VOLN, s6 = sg5 * VOLN;
    const R PXk[4] = {PX1, PX2, PX3, PX4};
    const R PYk[4] = {PY1, PY2, PY3, PY4};
    const R PZk[4] = {PZ1, PZ2, PZ3, PZ4};
    const bool deleted = offg < ZERO;
    if constexpr (sizeof(R) == 8) {
      // FP64: node pair by node pair, each pair written as soon as it is final (fewest live
      // values; lets the FP64 chain run two waves per SIMD)
      const bool deleted = offg < ZERO;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int op = (k < 2) ? k + 6 : k + 2; // 1-7, 2-8, 3-5, 4-6
        R fk[3], fo[3];
        hg_node_force3<R>(k, PXH, H, fk);
        hg_node_force3<R>(op, PXH, H, fo);
        R FINT;
        FINT = s1 * PXk[k] + s4 * PYk[k] + s6 * PZk[k];
        fk[0] = fk[0] - FINT; fo[0] = fo[0] + FINT;
        FINT = s2 * PYk[k] + s4 * PXk[k] + s5 * PZk[k];
        fk[1] = fk[1] - FINT; fo[1] = fo[1] + FINT;
        FINT = s3 * PZk[k] + s6 * PXk[k] + s5 * PYk[k];
        fk[2] = fk[2] - FINT; fo[2] = fo[2] + FINT;
        if (deleted) {
          fk[0] = ZERO; fk[1] = ZERO; fk[2] = ZERO;
          fo[0] = ZERO; fo[1] = ZERO; fo[2] = ZERO;
        }
        write(k, fk[0], fk[1], fk[2]);
        write(op, fo[0], fo[1], fo[2]);
      }
    } else {
      // FP32: all hourglass forces, then SFINT3 over the pairs, then the writes (registers
      // are not the limit here and this order schedules better)
      R f1[8], f2[8], f3[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        R f[3];
        hg_node_force3<R>(j, PXH, H, f);
        f1[j] = f[0]; f2[j] = f[1]; f3[j] = f[2];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int op = (k < 2) ? k + 6 : k + 2;
        R FINT;
        FINT = s1 * PXk[k] + s4 * PYk[k] + s6 * PZk[k];
        f1[k] = f1[k] - FINT; f1[op] = f1[op] + FINT;
        FINT = s2 * PYk[k] + s4 * PXk[k] + s5 * PZk[k];
        f2[k] = f2[k] - FINT; f2[op] = f2[op] + FINT;
        FINT = s3 * PZk[k] + s6 * PXk[k] + s5 * PYk[k];
        f3[k] = f3[k] - FINT; f3[op] = f3[op] + FINT;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (deleted) { f1[j] = ZERO; f2[j] = ZERO; f3[j] = ZERO; }
        write(j, f1[j], f2[j], f3[j]);
      }
    }
  }
}

// The whole chain for one element.  st.offg and the gathered nodes are read beforehand;
// load_rest() reads the rest of st between the kinematics and the material phase.
template <typename R, typename S, typename Write, typename LoadRest, typename Ehou = NoEhou>
__device__ __forceinline__ void hex_element(const double (&xd)[8], const double (&yd)[8], const double (&zd)[8],
                                            R (&vx)[8], R (&vy)[8], R (&vz)[8], HexState<S> &st,
                                            const HexConsts &c, const R DT1, HexOut<R> &o, Write &&write,
                                            LoadRest &&load_rest, Ehou ehou_sink = Ehou{}) {
  HexGeom<R> g;
  hex_geometry<R>(xd, yd, zd, double(st.offg), g);
  HexKin<R> kn;
  hex_kinematics<R>(vx, vy, vz, g, DT1, kn);
  load_rest();
  hex_material_forces<R, S>(g, kn, st, c, DT1, o, write, ehou_sink);
}

} // namespace orgpu
