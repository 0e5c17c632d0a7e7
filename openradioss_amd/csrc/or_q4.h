// or_q4.h -- device code of the 4-node Q4 shell (CFORC3 chain, Ishell=1, law 2) for gfx950.
//
// One thread owns one element.  Each step follows the reference routine it
// names operation for operation, so that with -ffp-contract=off the FP64
// instantiation rounds like the reference CPU engine (exp/log of the hardening law
// are evaluated in double-double and correctly rounded except in rare near-tie
// cases, see or_math.h).  The split
// instantiation (R = float) computes the convected frame and local coordinates
// from FP64 coordinates, then runs strain, material and hourglass arithmetic in
// FP32 with FP32 element state.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "or_hex.h"

namespace orgpu {

// Resolved law-2 / shell-property constants handed to the kernels by value.
struct Q4Consts {
  double rho, young, nu, shear, ssp, ca, cb, cn, epmx, ymax, cc, epdr, m_exp, asrate;
  int32_t icc, vp, israte, npt, ipla, ithk, ishear, pad;
  double hm, hf, hr, fsh, dtfac, dtmin;
  // plane-stress moduli of M2CPLR, A1 = E / (1 - nu^2) and A2 = A1 * nu, in FP64 on the host
  double a1, a2;
  // parameter constants of constant_mod.F / CHVIS3, computed on the host
  double zep81, zep072169, sr2d2, srh1, srh2, srh3;
  // through-thickness rule of COQINI for npt points (Z0, WF, WM), widened from float
  double z0[10], wf[10], wm[10];
};

// Per-element shell state held in registers for the whole element.
template <typename R> struct Q4Elem {
  R offg, thk, thke, epsd;
};

// GBUF%FOR, MOM, STRA, HOURG and EINT: read together after the through-thickness loop
// through the element's accessor (late_load, one memory round trip, nothing live across
// the loop) and written back as soon as each is final (late_store for FOR, MOM, STRA and
// EINT, hourg_store for HOURG).  late_load fills hourg[0..2] (HOURG 4 and 5 are set,
// not updated).
template <typename R> struct Q4Late {
  R forc[5], mom[3], gstr[8], hourg[5], eint[2];
};

template <typename R> struct Q4Out {
  R f[3][4], m[3][4];
  R sti;
  double dt;    // element time step, or +EP06 when excluded from the minimum
  R ehou;       // CHVIS3's hourglass energy of the cycle (EHOU)
};

__device__ __forceinline__ double expR(double a) { return ormath::or_exp(a); }
__device__ __forceinline__ float expR(float a) { return expf(a); }
template <typename T> __device__ __forceinline__ T fsignT(T a, T b) { return copysign(fabs(a), b); }

// M2CPLR (m2cplr.F) for one integration point, IFORM=0, FISOKIN=0, IMPL_S=0.
// off is OFF_OLD, the element OFF before this cycle's rupture test (sigeps02c.F:135).
template <typename R>
__device__ __forceinline__ void m2cplr_point(const Q4Consts &c, const R a1, const R a2, const R epdr, R off, R gs,
                                             R epsp_in, R (&s)[5], R dxx, R dyy, R dxy, R dyz, R dzx, R &pla,
                                             R &ezz, R &epchk) {
  const R ZERO = R(0), ONE = R(1), TWO = R(2), THREE = R(3), HALF = R(0.5), FOURTH = R(0.25);
  const R EM20 = R(1.0 / 1.0e20), EM15 = R(1.0 / 1.0e15), SMALL = R(1.0 / 1.0e7);
  const R YOUNG = R(c.young), G = R(c.shear), CN = R(c.cn);
  R CA = R(c.ca), CB = R(c.cb), YMAX = R(c.ymax), H = ZERO;
  ezz = ZERO;
  R sxx = s[0] + a1 * dxx + a2 * dyy;
  R syy = s[1] + a2 * dxx + a1 * dyy;
  R sxy = s[2] + G * dxy;
  R syz = s[3] + gs * dyz;
  R szx = s[4] + gs * dzx;
  if (c.cc != 0.0) {
    R epsp = epsp_in;
    if (c.israte == 0 && c.vp == 2) epsp = dmaxT(dmaxT(fabsR(dxx), fabsR(dyy)), HALF * fabsR(dxy));
    epsp = dmaxT(epsp, epdr);
    R logep = logR(epsp / epdr);
    R q = (ONE + R(c.cc) * logep);
    q = dmaxT(q, EM20);
    CA = CA * q;
    CB = CB * q;
    if (c.icc == 1) YMAX = YMAX * q;
  }
  R YLD;
  if (pla == ZERO) YLD = CA;
  else YLD = CA + (CB * (ONE - ZERO)) * expR(CN * logR(pla));
  YLD = dminT(YLD, YMAX);
  if (c.ipla == 0) {
    R svm = sqrtR(sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy);
    R r = dminT(ONE, YLD / (svm + EM15));
    if (r < ONE) {
      sxx = sxx * r; syy = syy * r; sxy = sxy * r;
      R dpla = off * dmaxT(ZERO, (svm - YLD) / YOUNG);
      R s1 = HALF * (sxx + syy);
      ezz = dpla * s1 / YLD;
      pla = pla + dpla;
      epchk = dmaxT(pla, epchk);
    }
  } else if (c.ipla == 1) {
    R s1 = sxx + syy, s2 = sxx - syy, s3 = sxy;
    R A = FOURTH * s1 * s1;
    R B = R(3.0 / 4.0) * s2 * s2 + THREE * s3 * s3;
    R svm = sqrtR(A + B);
    if (svm > YLD && off == ONE) {
      const R nu = R(c.nu);
      R nu1 = ONE / (ONE - nu), nu2 = ONE / (ONE + nu);
      if (YLD >= YMAX) H = ZERO;
      else H = CN * CB * expR((CN - ONE) * logR(pla + SMALL));
      R dpla_j = (svm - YLD) / (THREE * G + H);
      R anu1 = A * nu1, bnu2 = THREE * B * nu2, h2 = TWO * H;
      R dpla_i = ZERO, P = ONE, Q = ONE, DR = ZERO;
#pragma unroll
      for (int it = 0; it < 3; ++it) {
        dpla_i = dpla_j;
        R pla_i = pla + dpla_i;
        R yld_i;
        if (pla_i == ZERO) yld_i = dminT(YMAX, CA);
        else yld_i = dminT(YMAX, CA + CB * expR(CN * logR(pla_i)));
        DR = HALF * YOUNG * dpla_i / yld_i;
        P = ONE / (ONE + DR * nu1);
        Q = ONE / (ONE + THREE * DR * nu2);
        R p2 = P * P, q2 = Q * Q;
        R F = A * p2 + B * q2 - yld_i * yld_i;
        R DF = -(anu1 * p2 * P + bnu2 * q2 * Q) * (YOUNG - DR * h2) / yld_i - h2 * yld_i;
        if (dpla_i > ZERO) dpla_j = dmaxT(ZERO, dpla_i - F / DF);
        else dpla_j = ZERO;
      }
      pla = pla + dpla_i;
      epchk = dmaxT(pla, epchk);
      s1 = (sxx + syy) * P;
      s2 = (sxx - syy) * Q;
      sxx = HALF * (s1 + s2);
      syy = HALF * (s1 - s2);
      sxy = sxy * Q;
      ezz = DR * s1 / YOUNG;
    }
  } else { // IPLA == 2
    R svm2 = sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy;
    R svm = sqrtR(svm2);
    R yld2 = YLD * YLD;
    if (svm2 > yld2 && off == ONE) {
      if (YLD >= YMAX) H = ZERO;
      else H = CN * CB * expR((CN - ONE) * logR(pla + SMALL));
      R aa = (svm2 - yld2) / (R(5) * svm2 + THREE * (-sxx * syy + sxy * sxy));
      R s1 = (ONE - TWO * aa) * sxx + aa * syy;
      R s2 = aa * sxx + (ONE - TWO * aa) * syy;
      R s3 = (ONE - THREE * aa) * sxy;
      sxx = s1; syy = s2; sxy = s3;
      R dpla = off * (svm - YLD) / (THREE * G + H);
      pla = pla + dpla;
      YLD = YLD + H * dpla;
      svm = sqrtR(sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy);
      R r = dminT(ONE, YLD / dmaxT(EM20, svm));
      sxx = sxx * r; syy = syy * r; sxy = sxy * r;
      ezz = dpla * HALF * (sxx + syy) / YLD;
    }
  }
  s[0] = sxx; s[1] = syy; s[2] = sxy; s[3] = syz; s[4] = szx;
}

// The convected frame and local geometry of CNVEC3 (ISHFRAM=0) and CDERI3, in FP64 from
// the FP64 coordinates, rounded to R.
template <typename R> struct Q4Frame {
  R E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y, E3Z;
  R X2, X3, X4, Y2, Y3, Y4, PX1, PX2, PY1, PY2, AREA, VHX, VHY;
};
template <typename R>
__device__ __forceinline__ void q4_frame(const double (&xg)[4], const double (&yg)[4], const double (&zg)[4],
                                         Q4Frame<R> &f) {
  const double DZERO = 0.0, DONE = 1.0, DHALF = 0.5, DTWO = 2.0, DEM20 = 1.0 / 1.0e20;
  double X21 = xg[1] - xg[0], X32 = xg[2] - xg[1], X34 = xg[2] - xg[3], X41 = xg[3] - xg[0];
  double Y21 = yg[1] - yg[0], Y32 = yg[2] - yg[1], Y34 = yg[2] - yg[3], Y41 = yg[3] - yg[0];
  double Z21 = zg[1] - zg[0], Z32 = zg[2] - zg[1], Z34 = zg[2] - zg[3], Z41 = zg[3] - zg[0];
  double e1x = (X21 + X34), e1y = (Y21 + Y34), e1z = (Z21 + Z34);
  double e2x = (X32 + X41), e2y = (Y32 + Y41), e2z = (Z32 + Z41);
  double e3x = e1y * e2z - e1z * e2y;
  double e3y = e1z * e2x - e1x * e2z;
  double e3z = e1x * e2y - e1y * e2x;
  double suma = e3x * e3x + e3y * e3y + e3z * e3z;
  suma = DONE / dmaxT(sqrt(suma), DEM20);
  e3x = e3x * suma; e3y = e3y * suma; e3z = e3z * suma;
  double s1 = e1x * e1x + e1y * e1y + e1z * e1z;
  double s2 = e2x * e2x + e2y * e2y + e2z * e2z;
  suma = sqrt(s1 / s2);
  e1x = e1x + (e2y * e3z - e2z * e3y) * suma;
  e1y = e1y + (e2z * e3x - e2x * e3z) * suma;
  e1z = e1z + (e2x * e3y - e2y * e3x) * suma;
  suma = e1x * e1x + e1y * e1y + e1z * e1z;
  suma = DONE / dmaxT(sqrt(suma), DEM20);
  e1x = e1x * suma; e1y = e1y * suma; e1z = e1z * suma;
  e2x = e3y * e1z - e3z * e1y;
  e2y = e3z * e1x - e3x * e1z;
  e2z = e3x * e1y - e3y * e1x;
  double X31GA = xg[2] - xg[0], Y31GA = yg[2] - yg[0], Z31GA = zg[2] - zg[0];
  double X41GA = xg[3] - xg[0], Y41GA = yg[3] - yg[0], Z41GA = zg[3] - zg[0];
  double x2 = e1x * X21 + e1y * Y21 + e1z * Z21;
  double y2 = e2x * X21 + e2y * Y21 + e2z * Z21;
  double y3 = e2x * X31GA + e2y * Y31GA + e2z * Z31GA;
  double x3 = e1x * X31GA + e1y * Y31GA + e1z * Z31GA;
  double x4 = e1x * X41GA + e1y * Y41GA + e1z * Z41GA;
  double y4 = e2x * X41GA + e2y * Y41GA + e2z * Z41GA;
  double px1 = DHALF * (y2 - y4), py1 = DHALF * (x4 - x2), px2 = DHALF * y3, py2 = -DHALF * x3;
  double area = dmaxT(DTWO * (py2 * px1 - py1 * px2), DEM20);
  double vhx = (-x2 + x3 - x4) / area, vhy = (-y2 + y3 - y4) / area;
  (void)DZERO;
  f.E1X = R(e1x); f.E1Y = R(e1y); f.E1Z = R(e1z); f.E2X = R(e2x); f.E2Y = R(e2y); f.E2Z = R(e2z);
  f.E3X = R(e3x); f.E3Y = R(e3y); f.E3Z = R(e3z);
  f.X2 = R(x2); f.X3 = R(x3); f.X4 = R(x4); f.Y2 = R(y2); f.Y3 = R(y3); f.Y4 = R(y4);
  f.PX1 = R(px1); f.PX2 = R(px2); f.PY1 = R(py1); f.PY2 = R(py2); f.AREA = R(area); f.VHX = R(vhx); f.VHY = R(vhy);
}

// The CFORC3 chain for one element.  xg/yg/zg: gathered coordinates (FP64);
// vl/vrl: gathered velocities and rotational velocities; tri: node 3 == node 4.
// sig/pla/epsd_ip: per-point state, loaded and stored by the caller through the
// accessor ip_io (so that the kernel chooses the memory layout), which also reads
// and writes FOR, MOM, STRA, HOURG and EINT (Q4Late).  nload(k, x, v) gives node k's coordinates and
// velocity again after the through-thickness loop (the values of xg/yg/zg and vl), from
// which the frame and the local velocities that CHVIS3 and CFINT3 use are formed again
// rather than kept live across the loop.
template <typename R, typename IPIO, typename NLOAD, typename Ehou = NoEhou>
__device__ __forceinline__ void q4_element(const double (&xg)[4], const double (&yg)[4], const double (&zg)[4],
                                           R (&vl)[4][3], R (&vrl)[4][3], bool tri, Q4Elem<R> &st,
                                           const Q4Consts &c, const R DT1, IPIO &ip_io, NLOAD &&nload,
                                           Q4Out<R> &o, Ehou ehou_sink = Ehou{}) {
  const R ZERO = R(0), ONE = R(1), TWO = R(2), THREE = R(3), HALF = R(0.5), FOURTH = R(0.25);
  const R EM20 = R(1.0 / 1.0e20), EM01 = R(1.0 / 10.0);
  const R FOUR_OVER_5 = R(4.0 / 5.0);
  // ---- CCOOR3 (ccoor3.F:68-150) ----
  R offg = st.offg;
  R off = dminT(ONE, fabsR(offg));
  const bool zero_vel = offg < ZERO;
  if (zero_vel) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) { vl[k][d] = ZERO; vrl[k][d] = ZERO; }
  }
  R thk0 = st.thke;
  // ---- CNVEC3 (ISHFRAM=0) and CDERI3 ----
  Q4Frame<R> fr;
  q4_frame<R>(xg, yg, zg, fr);
  R E1X = fr.E1X, E1Y = fr.E1Y, E1Z = fr.E1Z, E2X = fr.E2X, E2Y = fr.E2Y, E2Z = fr.E2Z;
  R E3X = fr.E3X, E3Y = fr.E3Y, E3Z = fr.E3Z;
  R PX1 = fr.PX1, PX2 = fr.PX2, PY1 = fr.PY1, PY2 = fr.PY2, AREA = fr.AREA, VHX = fr.VHX, VHY = fr.VHY;
  // ---- CCOEF3 (ccoef3.F:84-170) ----
  if (c.ithk > 0) thk0 = st.thk;
  // SHF of CCOEF3 and ALDT of CDLEN3 are formed from the frame here and again after the
  // through-thickness loop (same inputs, same values), so they are not live across it
  auto shear_factor = [&](R area, R thk02) {
    const R nu = R(c.nu), fsh = R(c.fsh), ish = R(c.ishear);
    R fac1tmp = TWO * (ONE + nu) * thk02;
    return fsh * ((ONE - ish) + ish * fac1tmp / (fsh * area + fac1tmp));
  };
  // CDLEN3 (cdlen3.F:68-124), IHBE /= 0
  auto cdlen3 = [&](const Q4Frame<R> &f) {
    const R X2 = f.X2, X3 = f.X3, X4 = f.X4, Y2 = f.Y2, Y3 = f.Y3, Y4 = f.Y4;
    R AL1 = X2 * X2 + Y2 * Y2;
    R AL2 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
    R AL6 = X3 * X3 + Y3 * Y3;
    R AL3 = (X4 - X3) * (X4 - X3) + (Y4 - Y3) * (Y4 - Y3);
    R AL4 = X4 * X4 + Y4 * Y4;
    R AL5 = (X4 - X2) * (X4 - X2) + (Y4 - Y2) * (Y4 - Y2);
    R ALMIN = dminT(dminT(AL1, AL2), AL4);
    R ALQUAD = dminT(dminT(AL3, AL5), AL6);
    if (AL3 != ZERO) ALMIN = dminT(ALMIN, ALQUAD);
    R DTDYN = f.AREA * f.AREA / dmaxT(dmaxT(AL5, AL6), EM20);
    R aldt = dmaxT(DTDYN, ALMIN);
    R DTHOUR = HALF * (ALMIN + aldt) / dmaxT(R(c.hm), R(c.hf));
    if (DTHOUR < aldt) aldt = DTHOUR;
    return sqrtR(aldt);
  };
  R VOL0 = thk0 * AREA;
  R THK02 = thk0 * thk0;
  R SHF = shear_factor(AREA, THK02);
  const R GS = R(c.shear) * SHF;
  R ALDT = cdlen3(fr);
  // ---- CDEFO3, IHBE<=1 (cdefo3.F:94-140) ----
  R VX[4], VY[4], VZ[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    VX[k] = E1X * vl[k][0] + E1Y * vl[k][1] + E1Z * vl[k][2];
    VY[k] = E2X * vl[k][0] + E2Y * vl[k][1] + E2Z * vl[k][2];
    VZ[k] = E3X * vl[k][0] + E3Y * vl[k][1] + E3Z * vl[k][2];
  }
  R EXX, EYY, EXY, EXZ, EYZ;
  {
    R VZ13 = VZ[0] - VZ[2], VZ24 = VZ[1] - VZ[3];
    EYZ = PY1 * VZ13 + PY2 * VZ24;
    EXZ = PX1 * VZ13 + PX2 * VZ24;
    R DT1V4 = FOURTH * DT1;
    R TMP2A = PY2 + PY1;
    R TMP3A = fsignT(dmaxT(fabsR(TMP2A), EM20), TMP2A);
    R TMP1A = DT1V4 * (VZ13 - VZ24) * (VZ13 - VZ24) / TMP3A;
    R VX13 = VX[0] - VX[2], VX24 = VX[1] - VX[3];
    VX13 = VX13 - TMP1A;
    VX24 = VX24 + TMP1A;
    EXX = PX1 * VX13 + PX2 * VX24;
    EXY = PY1 * VX13 + PY2 * VX24;
    R TMP1B = PX2 - PX1;
    R TMP3B = fsignT(dmaxT(fabsR(TMP1B), EM20), TMP1B);
    R TMP2B = DT1V4 * (VZ13 + VZ24) * (VZ13 + VZ24) / TMP3B;
    R VY13 = VY[0] - VY[2], VY24 = VY[1] - VY[3];
    VY13 = VY13 + TMP2B;
    VY24 = VY24 + TMP2B;
    EXY = EXY + PX1 * VY13 + PX2 * VY24;
    EYY = PY1 * VY13 + PY2 * VY24;
  }
  // ---- CCURV3 (ccurv3.F:64-92) ----
  R RX[4], RY[4], KXX, KYY, KXY;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    RX[k] = E1X * vrl[k][0] + E1Y * vrl[k][1] + E1Z * vrl[k][2];
    RY[k] = E2X * vrl[k][0] + E2Y * vrl[k][1] + E2Z * vrl[k][2];
  }
  {
    R RX13TA = RX[0] - RX[2];
    R RXAVTA = RX[0] + RX[1] + RX[2] + RX[3];
    R RX24TA = RX[1] - RX[3];
    KYY = -PY1 * RX13TA - PY2 * RX24TA;
    KXY = PX1 * RX13TA + PX2 * RX24TA;
    R RY13TA = RY[0] - RY[2];
    R RYAVTA = RY[0] + RY[1] + RY[2] + RY[3];
    R RY24TA = RY[1] - RY[3];
    KXX = PX1 * RY13TA + PX2 * RY24TA;
    KXY = PY1 * RY13TA + PY2 * RY24TA - KXY;
    EXZ = EXZ + RYAVTA * (R(0.25) * AREA);
    EYZ = EYZ - RXAVTA * (R(0.25) * AREA);
  }
  // rotational hourglass rates of CHVIS3 (chvis3.F), formed here from RX / RY
  const R HGRX = RX[0] - RX[1] + RX[2] - RX[3];
  const R HGRY = RY[0] - RY[1] + RY[2] - RY[3];
  // ---- CSTRA3 (cstra3.F:83-93, 168-178), ISTRAIN=1 ----
  {
    R FAC1 = DT1 / AREA;
    EXX = EXX * FAC1; EYY = EYY * FAC1; EXY = EXY * FAC1; EYZ = EYZ * FAC1; EXZ = EXZ * FAC1;
    KXX = KXX * FAC1; KYY = KYY * FAC1; KXY = KXY * FAC1;
    // GSTR = GSTR + increments, applied with the state read after the loop
  }
  // ---- element strain rate of CFORC3 (cforc3.F:537-546) ----
  const R dtinv = DT1 / dmaxT(DT1 * DT1, EM20);
  R epsd_pg;
  {
    R thk = st.thk;
    R eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) * R(1.0 / 9.0) * (thk * thk);
    R eps_m2 = R(4.0 / 3.0) * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
    epsd_pg = sqrtR(eps_k2 + eps_m2) * dtinv;
    st.epsd = ONE * epsd_pg + (ONE - ONE) * st.epsd;
  }
  // ---- MULAWC law-2 branch ----
  VOL0 = AREA * thk0;
  R thkn = st.thk;
  R FOR[5] = {ZERO, ZERO, ZERO, ZERO, ZERO}, MOM[3] = {ZERO, ZERO, ZERO};
  const R off_old = off;
  R epchk = ZERO;
  bool ioff_duct = false;
  const R a1 = R(c.a1);
  const R a2 = (sizeof(R) == 8) ? R(c.a2) : a1 * R(c.nu);
  const R epdr = dmaxT(R(c.epdr) * DT1, EM20);
  const R asrate = (c.israte > 0) ? dminT(ONE, R(c.asrate) * DT1) : ONE;
  const R nu = R(c.nu);
  // each point's state is loaded one point ahead, so that its loads overlap the
  // previous point's arithmetic
  R s_next[5], pla_next, epsd_next;
  ip_io.load(0, s_next, pla_next, epsd_next);
  for (int ip = 0; ip < c.npt; ++ip) {
    const R thkly = R(c.wf[ip]);
    const R wmc = R(c.wm[ip]);
    const R thklyl = thkly * thk0;
    const R zt = (R(c.z0[ip]) + ZERO) * thk0;
    const R dxx = EXX + zt * KXX, dyy = EYY + zt * KYY, dxy = EXY + zt * KXY;
    R s[5] = {s_next[0], s_next[1], s_next[2], s_next[3], s_next[4]};
    R pla = pla_next, epsd_ip = epsd_next;
    if (ip + 1 < c.npt) ip_io.load(ip + 1, s_next, pla_next, epsd_next);
    epsd_ip = asrate * epsd_pg + (ONE - asrate) * epsd_ip;
    const R epsdot = epsd_ip * DT1;
    R ezz;
    m2cplr_point<R>(c, a1, a2, epdr, off_old, GS, epsdot, s, dxx, dyy, dxy, EYZ, EXZ, pla, ezz, epchk);
    if (off == off_old && off > ZERO) {
      if (off == ONE && epchk >= R(c.epmx)) {
        off = FOUR_OVER_5;
        ioff_duct = true;
      } else if (off < ONE) {
        off = off * FOUR_OVER_5;
      }
    }
    ezz = -(dxx + dyy) * nu - (ONE - TWO * nu) * ezz;
    ezz = ezz / (ONE - nu);
    thkn = thkn + ezz * thklyl * off;
    ip_io.store(ip, s, pla, epsd_ip);
    FOR[0] = FOR[0] + thkly * s[0];
    FOR[1] = FOR[1] + thkly * s[1];
    FOR[2] = FOR[2] + thkly * s[2];
    FOR[3] = FOR[3] + thkly * s[3];
    FOR[4] = FOR[4] + thkly * s[4];
    MOM[0] = MOM[0] + wmc * s[0];
    MOM[1] = MOM[1] + wmc * s[1];
    MOM[2] = MOM[2] + wmc * s[2];
  }
  if ((off == FOUR_OVER_5 && !ioff_duct) || (off > ZERO && off_old < EM01)) off = ZERO;
  // the frame of CNVEC3 / CDERI3 again (same inputs and operations, so the same values)
  // and the node velocities of CDEFO3 again, zeroed as SCOOR3 / CCOOR3 zero them; the
  // element state of Q4Late is read in the same batch of loads
  Q4Late<R> lt;
  ip_io.late_load(lt);
  R vn[4][3];
  {
    double xa[4], ya[4], za[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double xk[3];
      nload(k, xk, vn[k]);
      xa[k] = xk[0]; ya[k] = xk[1]; za[k] = xk[2];
      if (zero_vel) { vn[k][0] = ZERO; vn[k][1] = ZERO; vn[k][2] = ZERO; }
    }
    q4_frame<R>(xa, ya, za, fr);
  }
  E1X = fr.E1X; E1Y = fr.E1Y; E1Z = fr.E1Z; E2X = fr.E2X; E2Y = fr.E2Y; E2Z = fr.E2Z;
  E3X = fr.E3X; E3Y = fr.E3Y; E3Z = fr.E3Z;
  PX1 = fr.PX1; PX2 = fr.PX2; PY1 = fr.PY1; PY2 = fr.PY2; AREA = fr.AREA; VHX = fr.VHX; VHY = fr.VHY;
  VOL0 = AREA * thk0;
  THK02 = thk0 * thk0;
  SHF = shear_factor(AREA, THK02);
  ALDT = cdlen3(fr);
  st.thk = dmaxT(thkn, R(1.0 / (1.0e20 * 1.0e10)));
#pragma unroll
  for (int k = 0; k < 5; ++k) FOR[k] = FOR[k] * off * ONE;
#pragma unroll
  for (int k = 0; k < 3; ++k) MOM[k] = MOM[k] * off * ONE;
  {
    const R inc[8] = {EXX, EYY, EXY, EYZ, EXZ, KXX, KYY, KXY};
#pragma unroll
    for (int k = 0; k < 8; ++k) lt.gstr[k] = lt.gstr[k] + inc[k];
  }
  // with the resultants of the previous cycle, then of this one
  R degmb = lt.forc[0] * EXX + lt.forc[1] * EYY + lt.forc[2] * EXY + lt.forc[3] * EYZ + lt.forc[4] * EXZ;
  R degfx = lt.mom[0] * KXX + lt.mom[1] * KYY + lt.mom[2] * KXY;
  degmb = degmb + FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
  degfx = degfx + MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
  {
    R vol2 = HALF * VOL0;
    lt.eint[0] = lt.eint[0] + degmb * vol2;
    lt.eint[1] = lt.eint[1] + degfx * thk0 * vol2;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) lt.forc[k] = FOR[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) lt.mom[k] = MOM[k];
  ip_io.late_store(lt);
  // VISCMX = 1, ALPE = 1 on this branch (DM = 0)
  // ---- CHVIS3, IHBE=1, ISMSTR=2 (chvis3.F:112-321) ----
  R H11, H12, H13, H21, H22, H23, H31, H32, H33, B1k, B2k;
  {
    R PX1V = PX1 * VHX, PX2V = PX2 * VHX, PY1V = PY1 * VHY, PY2V = PY2 * VHY;
    R GAMA1 = off * (ONE - PX1V - PY1V);
    R GAMA3 = off * (ONE + PX1V + PY1V);
    R GAMA2 = off * (-ONE - PX2V - PY2V);
    R GAMA4 = off * (-ONE + PX2V + PY2V);
    const R HVISC = HALF, HELAS = HALF, HVLIN = ZERO;
    R SHFPR3 = SHF / (THREE * (ONE + nu));
    R HVISH1 = HVISC * R(c.hm), HVISH2 = HVISC * R(c.hf);
    R R0 = FOURTH * R(c.rho);
    R R1 = R0 * R(100.0);
    R0 = R0 * HVLIN;
    const R SR2D2 = R(c.sr2d2), ZEP072169 = R(c.zep072169);
    R A1 = R1 * HVISH1;
    R A2 = R0 * SR2D2 * R(c.srh1);
    R SRSHFPR3 = sqrtR(SHFPR3);
    R A3 = R1 * HVISH2 * SRSHFPR3;
    R A4 = R0 * SR2D2 * R(c.srh2) * SRSHFPR3;
    R HH3 = HELAS * R(c.hr);
    R A5 = HH3 * R1 * ZEP072169;
    HH3 = SR2D2 * R(c.srh3);
    R A6 = HH3 * R0 * ZEP072169;
    R0 = FOURTH * R(c.young) * HELAS;
    R A7 = R(c.hm) * R0;
    R A8 = R(c.hf) * R0 * SHFPR3;
    R T2A = THK02 * AREA;
    R TSA = sqrtR(T2A);
    const R SSP = R(c.ssp);
    R H1Q = A1 * TSA, H1L = A2 * SSP * TSA;
    R H2Q = A3 * THK02, H2L = A4 * SSP * THK02;
    R H3Q = A5 * T2A, H3L = A6 * SSP * T2A;
    R TD = thk0 * DT1;
    R HH1 = A7 * TD;
    R B1 = PX1 * PX1 + PY1 * PY1, B2 = PX2 * PX2 + PY2 * PY2;
    R HH2 = A8 * THK02 * TD / (B1 + B2);
    if (tri) { H1Q = ZERO; H1L = ZERO; H2Q = ZERO; H2L = ZERO; H3Q = ZERO; H3L = ZERO; HH1 = ZERO; HH2 = ZERO; }
    // the local velocities of CDEFO3 again
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const R(&v)[3] = vn[k];
      VX[k] = E1X * v[0] + E1Y * v[1] + E1Z * v[2];
      VY[k] = E2X * v[0] + E2Y * v[1] + E2Z * v[2];
      VZ[k] = E3X * v[0] + E3Y * v[1] + E3Z * v[2];
    }
    R HG1 = VX[0] * GAMA1 + VX[1] * GAMA2 + VX[2] * GAMA3 + VX[3] * GAMA4;
    R HG2 = VY[0] * GAMA1 + VY[1] * GAMA2 + VY[2] * GAMA3 + VY[3] * GAMA4;
    R(&hourg)[5] = lt.hourg;
    hourg[0] = hourg[0] + HG1 * HH1;
    hourg[1] = hourg[1] + HG2 * HH1;
    R HOUR1A = hourg[0] + HG1 * (H1L + H1Q * fabsR(HG1));
    H11 = HOUR1A * GAMA1; H12 = HOUR1A * GAMA2; H13 = HOUR1A * GAMA3;
    R HOUR2A = hourg[1] + HG2 * (H1L + H1Q * fabsR(HG2));
    H21 = HOUR2A * GAMA1; H22 = HOUR2A * GAMA2; H23 = HOUR2A * GAMA3;
    R EH = HOUR1A * HG1 + HOUR2A * HG2; // EHOU (chvis3.F:299, 322, 336-338)
    HG1 = VZ[0] * GAMA1 + VZ[1] * GAMA2 + VZ[2] * GAMA3 + VZ[3] * GAMA4;
    hourg[2] = hourg[2] + HG1 * HH2;
    R HOUR3A = hourg[2] + HG1 * (H2L + H2Q * fabsR(HG1));
    H31 = HOUR3A * GAMA1; H32 = HOUR3A * GAMA2; H33 = HOUR3A * GAMA3;
    EH = EH + HOUR3A * HG1;
    HG1 = HGRX;
    HG2 = HGRY;
    hourg[3] = HG1 * (H3L + H3Q * fabsR(HG1));
    hourg[4] = HG2 * (H3L + H3Q * fabsR(HG2));
    EH = EH + hourg[3] * HG1 + hourg[4] * HG2;
    o.ehou = DT1 * EH * off; // DT1C = DT1
    ehou_sink(o.ehou);
    ip_io.hourg_store(lt.hourg);
    B1k = hourg[3] * off;
    B2k = hourg[4] * off;
  }
  // ---- CDT3 (cdt3.F:108-112, 141-165, 200-233), IDTMIN(3)=2 ----
  R STI;
  {
    R DT = R(c.dtfac) * ALDT / R(c.ssp);
    if (!(DT > R(c.dtmin) || off < ONE || offg < ZERO)) off = ZERO;
    o.dt = 1.0e6;
    if (offg > ZERO && off != ZERO) o.dt = double(DT);
    R DIVM = dmaxT(ALDT * ALDT, EM20);
    STI = HALF * VOL0 * R(c.young) / DIVM;
    STI = R(c.zep81) * STI * off;
  }
  // ---- CFINT3, ISIGI=0, IHBE<2 (cfint3.F:139-208) ----
  {
    R F1A = FOR[0] * thk0, F2A = FOR[1] * thk0, F3A = FOR[2] * thk0, F4A = FOR[3] * thk0, F5A = FOR[4] * thk0;
    R M4 = F4A * AREA, M5 = F5A * AREA;
    R F12 = F1A * PX2 + F3A * PY2;
    R F22 = F2A * PY2 + F3A * PX2;
    R F32 = F5A * PX2 + F4A * PY2;
    R F11 = F1A * PX1 + F3A * PY1;
    R F21 = F2A * PY1 + F3A * PX1;
    R F31 = F5A * PX1 + F4A * PY1;
    R G11 = F11 + H11, G13 = H13 - F11;
    R G21 = F21 + H21, G23 = H23 - F21;
    R G31 = F31 + H31, G33 = H33 - F31;
    R G12 = F12 + H12, G22 = F22 + H22, G32 = F32 + H32;
    o.f[0][0] = E1X * G11 + E2X * G21 + E3X * G31;
    o.f[0][1] = E1X * G12 + E2X * G22 + E3X * G32;
    o.f[0][2] = E1X * G13 + E2X * G23 + E3X * G33;
    o.f[1][0] = E1Y * G11 + E2Y * G21 + E3Y * G31;
    o.f[1][1] = E1Y * G12 + E2Y * G22 + E3Y * G32;
    o.f[1][2] = E1Y * G13 + E2Y * G23 + E3Y * G33;
    o.f[2][0] = E1Z * G11 + E2Z * G21 + E3Z * G31;
    o.f[2][1] = E1Z * G12 + E2Z * G22 + E3Z * G32;
    o.f[2][2] = E1Z * G13 + E2Z * G23 + E3Z * G33;
#pragma unroll
    for (int d = 0; d < 3; ++d) o.f[d][3] = -o.f[d][0] - o.f[d][1] - o.f[d][2];
    R M1A = MOM[0] * THK02, M2A = MOM[1] * THK02, M3A = MOM[2] * THK02;
    M4 = M4 * FOURTH;
    M5 = M5 * FOURTH;
    R M11 = -M2A * PY1 - M3A * PX1;
    R M21 = M1A * PX1 + M3A * PY1;
    R M12 = -M2A * PY2 - M3A * PX2;
    R M22 = M1A * PX2 + M3A * PY2;
    R G11T = M11 - M4 + B1k, G13T = -M11 - M4 + B1k;
    R G12T = M12 - M4 + (-B1k), G14T = -M12 - M4 + (-B1k);
    R G21T = M21 + M5 + B2k, G23T = -M21 + M5 + B2k;
    R G22T = M22 + M5 + (-B2k), G24T = -M22 + M5 + (-B2k);
    o.m[0][0] = E1X * G11T + E2X * G21T; o.m[0][1] = E1X * G12T + E2X * G22T;
    o.m[0][2] = E1X * G13T + E2X * G23T; o.m[0][3] = E1X * G14T + E2X * G24T;
    o.m[1][0] = E1Y * G11T + E2Y * G21T; o.m[1][1] = E1Y * G12T + E2Y * G22T;
    o.m[1][2] = E1Y * G13T + E2Y * G23T; o.m[1][3] = E1Y * G14T + E2Y * G24T;
    o.m[2][0] = E1Z * G11T + E2Z * G21T; o.m[2][1] = E1Z * G12T + E2Z * G22T;
    o.m[2][2] = E1Z * G13T + E2Z * G23T; o.m[2][3] = E1Z * G14T + E2Z * G24T;
  }
  // ---- CUPDT3P (cupdt3.F:1000-1105) ----
  if (off < ONE) offg = off;
  if (offg < ZERO) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int k = 0; k < 4; ++k) { o.f[d][k] = ZERO; o.m[d][k] = ZERO; }
    STI = ZERO;
  }
  st.offg = offg;
  o.sti = STI;
}

} // namespace orgpu
