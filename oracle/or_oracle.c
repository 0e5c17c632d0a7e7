/*
 * or_oracle.c -- CPU restatement of the OpenRadioss Engine explicit hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see or_oracle.h).  Built with
 *   gcc -O2 -ffp-contract=off -fno-fast-math
 * so that every expression rounds exactly as the reference Fortran does
 * under its release flags (-ffp-contract=off -frounding-math).
 *
 * Branch restated: Lagrangian (JLAG=1), Isolid=1 (JHBE=1), JCVT=0,
 * ISMSTR=4, law 2 Johnson-Cook (IFORM=0, FISOKIN=0), IPARIT=1 (skyline),
 * INVSTR>=35, IMPL=0, N2D=0, no thermal, no ALE, no failure.
 */
#include "or_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Constants: the reference builds its constants from integer literals by
 * arithmetic in the working precision (common_source/modules/constant_mod.F);
 * the same arithmetic is restated here so the values are bit-identical. */
#define ZERO 0.0
#define ONE 1.0
#define TWO 2.0
#define THREE 3.0
#define FOUR 4.0
#define HALF (1.0 / 2.0)
#define THIRD (1.0 / 3.0)
#define FOURTH (1.0 / 4.0)
#define ONE_OVER_8 (1.0 / 8.0)
#define ONE_OVER_64 (1.0 / 64.0)
#define TWO_THIRD (2.0 / 3.0)
#define HUNDRED 100.0
#define EP06 1.0e6
#define EP15 1.0e15   /* 10^15, exact */
#define EP20 1.0e20   /* 10^20, exact */
#define EM15 (1.0 / EP15)
#define EM20 (1.0 / EP20)
/* ONEP333 = ((ONE + THREE/TEN) + THREE/EP02) + THREE/EP03 (constant_mod.F:675,684,692,865) */
static double c_onep333(void) {
  volatile double zep3 = 3.0 / 10.0, zep03 = 3.0 / 100.0, threeem3 = 3.0 / 1000.0;
  volatile double onep3 = 1.0 + zep3;
  volatile double onep33 = onep3 + zep03;
  return onep33 + threeem3;
}
/* ONEP1 = ONE + ONE/TEN (constant_mod.F:651,673,833) */
static double c_onep1(void) {
  volatile double em01 = 1.0 / 10.0;
  return 1.0 + em01;
}
/* resol.F:6352 writes 1.1*DT2OLD with a default-kind (single precision) literal. */
#define DT_GROWTH ((double)1.1f)

double or_const(int which) {
  switch (which) {
  case 0: return c_onep333();
  case 1: return c_onep1();
  case 2: return DT_GROWTH;
  case 3: return EM15;
  case 4: return EM20;
  default: return 0.0;
  }
}

static inline double dmax(double a, double b) { return (a > b) ? a : b; }
static inline double dmin(double a, double b) { return (a < b) ? a : b; }

/* SLEN (slen.F:30): face area measure E*G-F^2 of one quadrilateral face. */
static inline double slen_face(double x1, double x2, double x3, double x4,
                               double y1, double y2, double y3, double y4,
                               double z1, double z2, double z3, double z4) {
  double x13 = x3 - x1, x24 = x4 - x2;
  double y13 = y3 - y1, y24 = y4 - y2;
  double z13 = z3 - z1, z24 = z4 - z2;
  double fs1 = x13 - x24, ft1 = x13 + x24;
  double fs2 = y13 - y24, ft2 = y13 + y24;
  double fs3 = z13 - z24, ft3 = z13 + z24;
  double e = fs1 * fs1 + fs2 * fs2 + fs3 * fs3;
  double f = fs1 * ft1 + fs2 * ft2 + fs3 * ft3;
  double g = ft1 * ft1 + ft2 * ft2 + ft3 * ft3;
  return e * g - f * f;
}

void or_hex_forces(int64_t e0, int64_t e1, int64_t ne, const int32_t *conn,
                   const double *X, const double *V, double *SIG,
                   double *EINT, double *RHO, const double *VOL0,
                   double *PLA, double *EPSD, double *OFFG, double *QVIS,
                   double *EHOU, const int64_t *IADS, double *FSKY,
                   const or_hex_params *p, double DT1, double *dt2t, int *mstop) {
  or_hex_forces_dt(e0, e1, ne, conn, X, V, SIG, EINT, RHO, VOL0, PLA, EPSD, OFFG, QVIS, EHOU, IADS, FSKY, p, DT1,
                   dt2t, mstop, NULL);
}

void or_hex_forces_dt(int64_t e0, int64_t e1, int64_t ne, const int32_t *conn,
                      const double *X, const double *V, double *SIG,
                      double *EINT, double *RHO, const double *VOL0,
                      double *PLA, double *EPSD, double *OFFG, double *QVIS,
                      double *EHOU, const int64_t *IADS, double *FSKY,
                      const or_hex_params *p, double DT1, double *dt2t, int *mstop, double *dtel) {
  const double ONEP333 = c_onep333();
  double DT2T = *dt2t;
  for (int64_t e = e0; e < e1; ++e) {
    const int32_t *nc = conn + 8 * e;
    /* ---- SCOOR3 (scoor3.F:109-118, 149-172, 233-258, 286-340) ---- */
    double xd[8], yd[8], zd[8], vx[8], vy[8], vz[8];
    for (int k = 0; k < 8; ++k) {
      xd[k] = X[3 * (int64_t)nc[k] + 0];
      yd[k] = X[3 * (int64_t)nc[k] + 1];
      zd[k] = X[3 * (int64_t)nc[k] + 2];
      vx[k] = V[3 * (int64_t)nc[k] + 0];
      vy[k] = V[3 * (int64_t)nc[k] + 1];
      vz[k] = V[3 * (int64_t)nc[k] + 2];
    }
    double offg = OFFG[e];
    double off = fabs(offg);            /* |OFFG| <= 1 on this branch */
    if (offg < ZERO) {
      for (int k = 0; k < 8; ++k) { vx[k] = ZERO; vy[k] = ZERO; vz[k] = ZERO; }
    }
    const double VD2 = ZERO, VIS = ZERO;  /* JLAG /= 0: scoor3.F:343-371 */

    /* ---- SDERI3 (sderi3.F) ---- */
    double X17 = xd[6] - xd[0], X28 = xd[7] - xd[1], X35 = xd[4] - xd[2], X46 = xd[5] - xd[3];
    double Y17 = yd[6] - yd[0], Y28 = yd[7] - yd[1], Y35 = yd[4] - yd[2], Y46 = yd[5] - yd[3];
    double Z17 = zd[6] - zd[0], Z28 = zd[7] - zd[1], Z35 = zd[4] - zd[2], Z46 = zd[5] - zd[3];
    double JAC1 = X17 + X28 - X35 - X46;
    double JAC2 = Y17 + Y28 - Y35 - Y46;
    double JAC3 = Z17 + Z28 - Z35 - Z46;
    double X_17_46 = X17 + X46, X_28_35 = X28 + X35;
    double Y_17_46 = Y17 + Y46, Y_28_35 = Y28 + Y35;
    double Z_17_46 = Z17 + Z46, Z_28_35 = Z28 + Z35;
    double JAC4 = X_17_46 + X_28_35, JAC5 = Y_17_46 + Y_28_35, JAC6 = Z_17_46 + Z_28_35;
    double JAC7 = X_17_46 - X_28_35, JAC8 = Y_17_46 - Y_28_35, JAC9 = Z_17_46 - Z_28_35;
    double JAC_59_68 = JAC5 * JAC9 - JAC6 * JAC8;
    double JAC_67_49 = JAC6 * JAC7 - JAC4 * JAC9;
    double JAC_48_57 = JAC4 * JAC8 - JAC5 * JAC7;
    double DET = ONE_OVER_64 * (JAC1 * JAC_59_68 + JAC2 * JAC_67_49 + JAC3 * JAC_48_57);
    /* SCHKJABT3 (solide4/schkjabt3.F), JLAG/=0, ISMSTR=4, INCONV=1, INEG_V=0 */
    if (off == ZERO) {
      DET = ONE;
    } else if (offg > ONE) {
    } else if (DET <= ZERO) {
      DET = ONE; off = ZERO; offg = ZERO; *mstop = 1;
    }
    double DETT = ONE_OVER_64 / DET;
    double JACI1 = DETT * JAC_59_68, JACI4 = DETT * JAC_67_49, JACI7 = DETT * JAC_48_57;
    double JACI2 = DETT * (JAC3 * JAC8 - JAC2 * JAC9);
    double JACI5 = DETT * (JAC1 * JAC9 - JAC3 * JAC7);
    double JACI8 = DETT * (JAC2 * JAC7 - JAC1 * JAC8);
    double JACI3 = DETT * (JAC2 * JAC6 - JAC3 * JAC5);
    double JACI6 = DETT * (JAC3 * JAC4 - JAC1 * JAC6);
    double JACI9 = DETT * (JAC1 * JAC5 - JAC2 * JAC4);
    double JAC12 = JACI1 + JACI2, JAC45 = JACI4 + JACI5, JAC78 = JACI7 + JACI8;
    double PX1 = -JAC12 - JACI3, PY1 = -JAC45 - JACI6, PZ1 = -JAC78 - JACI9;
    double PX2 = -JAC12 + JACI3, PY2 = -JAC45 + JACI6, PZ2 = -JAC78 + JACI9;
    JAC12 = JACI1 - JACI2; JAC45 = JACI4 - JACI5; JAC78 = JACI7 - JACI8;
    double PX3 = JAC12 + JACI3, PY3 = JAC45 + JACI6, PZ3 = JAC78 + JACI9;
    double PX4 = JAC12 - JACI3, PY4 = JAC45 - JACI6, PZ4 = JAC78 - JACI9;
    double HX, HY, HZ;
    HX = (xd[0] - xd[1] + xd[2] - xd[3] + xd[4] - xd[5] + xd[6] - xd[7]);
    HY = (yd[0] - yd[1] + yd[2] - yd[3] + yd[4] - yd[5] + yd[6] - yd[7]);
    HZ = (zd[0] - zd[1] + zd[2] - zd[3] + zd[4] - zd[5] + zd[6] - zd[7]);
    double PX1H1 = PX1 * HX + PY1 * HY + PZ1 * HZ;
    double PX2H1 = PX2 * HX + PY2 * HY + PZ2 * HZ;
    double PX3H1 = PX3 * HX + PY3 * HY + PZ3 * HZ;
    double PX4H1 = PX4 * HX + PY4 * HY + PZ4 * HZ;
    HX = (xd[0] + xd[1] - xd[2] - xd[3] - xd[4] - xd[5] + xd[6] + xd[7]);
    HY = (yd[0] + yd[1] - yd[2] - yd[3] - yd[4] - yd[5] + yd[6] + yd[7]);
    HZ = (zd[0] + zd[1] - zd[2] - zd[3] - zd[4] - zd[5] + zd[6] + zd[7]);
    double PX1H2 = PX1 * HX + PY1 * HY + PZ1 * HZ;
    double PX2H2 = PX2 * HX + PY2 * HY + PZ2 * HZ;
    double PX3H2 = PX3 * HX + PY3 * HY + PZ3 * HZ;
    double PX4H2 = PX4 * HX + PY4 * HY + PZ4 * HZ;
    HX = (xd[0] - xd[1] - xd[2] + xd[3] - xd[4] + xd[5] + xd[6] - xd[7]);
    HY = (yd[0] - yd[1] - yd[2] + yd[3] - yd[4] + yd[5] + yd[6] - yd[7]);
    HZ = (zd[0] - zd[1] - zd[2] + zd[3] - zd[4] + zd[5] + zd[6] - zd[7]);
    double PX1H3 = PX1 * HX + PY1 * HY + PZ1 * HZ;
    double PX2H3 = PX2 * HX + PY2 * HY + PZ2 * HZ;
    double PX3H3 = PX3 * HX + PY3 * HY + PZ3 * HZ;
    double PX4H3 = PX4 * HX + PY4 * HY + PZ4 * HZ;
    double VOLN = DET;

    /* ---- SDLEN3 / SLEN (sdlen3.F, slen.F) faces 1..6 ---- */
    double AREAM = EM20, ar;
    ar = slen_face(xd[0], xd[1], xd[2], xd[3], yd[0], yd[1], yd[2], yd[3], zd[0], zd[1], zd[2], zd[3]);
    AREAM = dmax(ar, AREAM);
    ar = slen_face(xd[4], xd[5], xd[6], xd[7], yd[4], yd[5], yd[6], yd[7], zd[4], zd[5], zd[6], zd[7]);
    AREAM = dmax(ar, AREAM);
    ar = slen_face(xd[0], xd[1], xd[5], xd[4], yd[0], yd[1], yd[5], yd[4], zd[0], zd[1], zd[5], zd[4]);
    AREAM = dmax(ar, AREAM);
    ar = slen_face(xd[1], xd[2], xd[6], xd[5], yd[1], yd[2], yd[6], yd[5], zd[1], zd[2], zd[6], zd[5]);
    AREAM = dmax(ar, AREAM);
    ar = slen_face(xd[2], xd[3], xd[7], xd[6], yd[2], yd[3], yd[7], yd[6], zd[2], zd[3], zd[7], zd[6]);
    AREAM = dmax(ar, AREAM);
    ar = slen_face(xd[3], xd[0], xd[4], xd[7], yd[3], yd[0], yd[4], yd[7], zd[3], zd[0], zd[4], zd[7]);
    AREAM = dmax(ar, AREAM);
    double DELTAX = FOUR * VOLN * ONE / sqrt(AREAM);

    /* ---- SDEFO3 (sdefo3.F), INTEG8*JEUL=0, JCVT=0, JHBE=1 branch ---- */
    double VX17 = vx[0] - vx[6], VX28 = vx[1] - vx[7], VX35 = vx[2] - vx[4], VX46 = vx[3] - vx[5];
    double VY17 = vy[0] - vy[6], VY28 = vy[1] - vy[7], VY35 = vy[2] - vy[4], VY46 = vy[3] - vy[5];
    double VZ17 = vz[0] - vz[6], VZ28 = vz[1] - vz[7], VZ35 = vz[2] - vz[4], VZ46 = vz[3] - vz[5];
    double DXX = PX1 * VX17 + PX2 * VX28 + PX3 * VX35 + PX4 * VX46;
    double DYY = PY1 * VY17 + PY2 * VY28 + PY3 * VY35 + PY4 * VY46;
    double DZZ = PZ1 * VZ17 + PZ2 * VZ28 + PZ3 * VZ35 + PZ4 * VZ46;
    double DXY = PY1 * VX17 + PY2 * VX28 + PY3 * VX35 + PY4 * VX46;
    double DXZ = PZ1 * VX17 + PZ2 * VX28 + PZ3 * VX35 + PZ4 * VX46;
    double DYX = PX1 * VY17 + PX2 * VY28 + PX3 * VY35 + PX4 * VY46;
    double DYZ = PZ1 * VY17 + PZ2 * VY28 + PZ3 * VY35 + PZ4 * VY46;
    double DZX = PX1 * VZ17 + PX2 * VZ28 + PX3 * VZ35 + PX4 * VZ46;
    double DZY = PY1 * VZ17 + PY2 * VZ28 + PY3 * VZ35 + PY4 * VZ46;
    double DT1D2 = HALF * DT1;
    double D4 = DXY + DYX, D5 = DYZ + DZY, D6 = DXZ + DZX;
    double WZZ = DT1D2 * (DYX - DXY);
    double WYY = DT1D2 * (DXZ - DZX);
    double WXX = DT1D2 * (DZY - DYZ);

    /* ---- SRHO3 (srho3.F), JLAG, MTN/=115, IRESP=0 ---- */
    const double RHO0 = p->rho0;
    double volo = VOL0[e];
    if (offg == ZERO && VOLN == ONE) VOLN = volo;
    double rhon = RHO[e];
    double DVOL = VOLN - (RHO0 / rhon) * volo;
    rhon = RHO0 * (volo / VOLN);
    double eint = EINT[e] * volo;

    /* ---- SROTA3 (srota3.F): Jaumann rotation; S = stress before rotation ---- */
    double S1 = SIG[0 * ne + e], S2 = SIG[1 * ne + e], S3 = SIG[2 * ne + e];
    double S4 = SIG[3 * ne + e], S5 = SIG[4 * ne + e], S6 = SIG[5 * ne + e];
    double sg[6];
    {
      double Q1 = TWO * S4 * WZZ, Q2 = TWO * S6 * WYY, Q3 = TWO * S5 * WXX;
      sg[0] = S1 - Q1 + Q2;
      sg[1] = S2 + Q1 - Q3;
      sg[2] = S3 - Q2 + Q3;
      sg[3] = S4 + WZZ * (S1 - S2) + WYY * S5 - WXX * S6;
      sg[4] = S5 + WXX * (S2 - S3) + WZZ * S6 - WYY * S4;
      sg[5] = S6 + WYY * (S3 - S1) + WXX * S4 - WZZ * S5;
    }

    /* ---- MMAIN law-2 glue (mmain.F90:668-745) ---- */
    double vol_avg = VOLN - HALF * DVOL;
    double amu = rhon / RHO0 - ONE;     /* mtn == 2 */
    double rhoref = rhon;               /* large strain, ismstr == 4 */
    const double tstar = ZERO;          /* no thermal */

    /* ---- M2LAW (m2law.F) ---- */
    double G = p->shear * off;
    double CA = p->ca, SIGMX = p->sigmx;
    double P = -THIRD * (sg[0] + sg[1] + sg[2]);
    double DAV = -THIRD * (DXX + DYY + DZZ);
    double G1 = DT1 * G, G2 = TWO * G1;
    double SSP = sqrt((ONEP333 * G + p->bulk) / RHO0);
    sg[0] = sg[0] + P + G2 * (DXX + DAV);
    sg[1] = sg[1] + P + G2 * (DYY + DAV);
    sg[2] = sg[2] + P + G2 * (DZZ + DAV);
    sg[3] = sg[3] + G1 * D4;
    sg[4] = sg[4] + G1 * D5;
    sg[5] = sg[5] + G1 * D6;
    double AJ2 = HALF * (sg[0] * sg[0] + sg[1] * sg[1] + sg[2] * sg[2]) + sg[3] * sg[3] + sg[4] * sg[4] +
                 sg[5] * sg[5];
    AJ2 = sqrt(THREE * AJ2);
    /* MSTRAIN_RATE (mstrain_rate.F), IDEV = VP-2 */
    double epsd = EPSD[e];
    {
      int idev = p->vp - 2;
      double epsdot;
      if (idev == 0) {
        double E1 = DXX, E2 = DYY, E3 = DZZ, E4 = HALF * D4, E5 = HALF * D5, E6 = HALF * D6;
        double epsp = E1 * E1 + E2 * E2 + E3 * E3 + TWO * (E4 * E4 + E5 * E5 + E6 * E6);
        epsdot = sqrt(epsp);
      } else {
        double dav = (DXX + DYY + DZZ) * THIRD;
        double E1 = DXX - dav, E2 = DYY - dav, E3 = DZZ - dav;
        double E4 = HALF * D4, E5 = HALF * D5, E6 = HALF * D6;
        double epsp = HALF * (E1 * E1 + E2 * E2 + E3 * E3) + E4 * E4 + E5 * E5 + E6 * E6;
        epsdot = sqrt(THREE * epsp) / (THREE / TWO);
      }
      if (p->israte == 0) {
        epsd = epsdot;
      } else if (p->israte > 0) {
        double asr = dmin(ONE, p->asrate * DT1);
        epsd = asr * epsdot + (ONE - asr) * epsd;
      }
    }
    double EPD = ONE;
    if (p->cc != ZERO) {
      if (p->vp == 1) {
        EPD = dmax(epsd, p->epdr);
        EPD = OR_LOGS(OR_SITE_M2LAW_RATE, EPD / p->epdr);
      } else {
        EPD = dmax(epsd, EM15);
        EPD = OR_LOGS(OR_SITE_M2LAW_RATE, EPD / p->epdr);
      }
      double MT = dmax(EM15, p->m_exp);
      EPD = dmax(ZERO, EPD);
      EPD = (ONE + p->cc * EPD) * (ONE - OR_POWS(OR_SITE_M2LAW_RATE, tstar, MT));
      if (p->icc == 1) SIGMX = p->sigmx * EPD;
    } else {
      double MT = dmax(EM15, p->m_exp);
      EPD = ONE - OR_POWS(OR_SITE_M2LAW_RATE, tstar, MT);
      if (p->icc == 1) SIGMX = p->sigmx * EPD;
    }
    double epxe = PLA[e];
    double AK, QH;
    if (p->cn == ONE) {
      AK = CA + p->cb * epxe;
      QH = p->cb * EPD;
    } else if (epxe > ZERO) {
      AK = CA + p->cb * OR_POWS(OR_SITE_M2LAW_AK, epxe, p->cn);
      if (p->cn > ONE) QH = (p->cb * p->cn * OR_POWS(OR_SITE_M2LAW_QH, epxe, p->cn - ONE)) * EPD;
      else QH = (p->cb * p->cn / OR_POWS(OR_SITE_M2LAW_QH, epxe, ONE - p->cn)) * EPD;
    } else {
      AK = CA;
      QH = ZERO;
    }
    AK = AK * EPD;
    if (SIGMX < AK) { AK = SIGMX; QH = ZERO; }
    if (epxe > p->epmx) { AK = ZERO; QH = ZERO; }
    {
      double SCALE = dmin(ONE, AK / dmax(AJ2, EM15));
      double DPLA = (ONE - SCALE) * AJ2 / dmax(THREE * G + QH, EM15);
      AK = AK + (ONE - p->fisokin) * DPLA * QH;
      SCALE = dmin(ONE, AK / dmax(AJ2, EM15));
      for (int c = 0; c < 6; ++c) sg[c] = SCALE * sg[c];
      epxe = epxe + DPLA;
    }
    /* ---- MQVISCB (mqviscb.F), IMPL=0, N2D=0, NPG=1, JSMS=0 ---- */
    double QNEW, STI;
    {
      double DD = -DXX - DYY - DZZ;
      double AD = ZERO, AL = ZERO;
      double CX = SSP + sqrt(VD2);
      if (off == ONE) {
        if (VOLN > ZERO) AL = OR_POWS(OR_SITE_MQVISCB_AL, VOLN, ONE / THREE);
        else AL = ZERO;
        AD = dmax(ZERO, DD);
      }
      double NRHO = sqrt(rhoref * RHO0);
      double QA = ONE * p->qa, QB = ONE * p->qb;
      double CNS1_0 = ONE * p->vns1, CNS2_0 = ONE * p->vns2;
      double QAA_0 = QA * QA;
      double CNS1 = CNS1_0 * AL * NRHO * SSP * off;
      double CNS2 = CNS2_0 * AL * NRHO * SSP * off;
      double QAA = QAA_0 * AD;
      double QX = QB * SSP + AL * QAA + ONE * TWO * VIS / dmax(EM20, rhon * DELTAX) +
                  (CNS1 + ONE * CNS2) / dmax(EM20, rhoref * DELTAX);
      QNEW = rhon * AD * AL * (QAA * AL + QB * SSP);
      double SSP_EQ = dmax(EM20, QX + sqrt(QX * QX + CX * CX));
      double DTX = DELTAX / SSP_EQ;
      STI = ZERO;
      if (!(off == ZERO || offg < ZERO)) {
        double TIDT = ONE / DTX;
        double TRHO = rhon * TIDT;
        double TVOL = VOLN * TIDT;
        STI = TRHO * TVOL;
      }
      DTX = p->dtfac * DTX;
      if (VOLN > ZERO && (off != ZERO && offg >= ZERO)) DT2T = dmin(DTX, DT2T);
      /* the candidate MQVISCB weighs for NELTS (mqviscb.F:622-630): -1 when the element is
         skipped (OFF <= 0, OFFG <= 0 or VOL <= 0) */
      if (dtel) dtel[e] = (VOLN > ZERO && off > ZERO && offg > ZERO) ? DTX : -ONE;
    }
    /* pressure + energy (m2law.F, IEOS == 0) */
    {
      double PNEW = p->bulk * amu;
      sg[0] = (sg[0] - PNEW) * off;
      sg[1] = (sg[1] - PNEW) * off;
      sg[2] = (sg[2] - PNEW) * off;
      sg[3] = sg[3] * off;
      sg[4] = sg[4] * off;
      sg[5] = sg[5] * off;
      double DTA = HALF * DT1;
      double E1 = DXX * (S1 + sg[0]);
      double E2 = DYY * (S2 + sg[1]);
      double E3 = DZZ * (S3 + sg[2]);
      double E4 = D4 * (S4 + sg[3]);
      double E5 = D5 * (S5 + sg[4]);
      double E6 = D6 * (S6 + sg[5]);
      double EINC = vol_avg * (E1 + E2 + E3 + E4 + E5 + E6) * DTA - HALF * DVOL * (QVIS[e] + QNEW);
      eint = (eint + EINC * off) / dmax(EM15, volo);
    }
    /* store the material state (M2LAW tail; lbuf%pla/epsd/qvis) */
    for (int c = 0; c < 6; ++c) SIG[c * ne + e] = sg[c];
    EINT[e] = eint;
    RHO[e] = rhon;
    PLA[e] = epxe;
    EPSD[e] = epsd;
    QVIS[e] = QNEW;

    /* ---- SMALLB3 (smallb3.F), ISMSTR=4 ---- */
    if (off < ONE) {
      if (off == ZERO) offg = ZERO;
      else if (offg > ONE) offg = ONE + off;
      else offg = off;
    }
    OFFG[e] = offg;

    /* ---- SHVIS3 (shvis3.F), INVSTR>=35, FLUID=0, ISMSTR=4, JHBE>=1 ---- */
    double F1[8], F2[8], F3[8];
    double EHOU_e;
    {
      double CAQ = FOURTH * off * p->qh;
      double FCL = CAQ * rhon * OR_POWS(OR_SITE_SHVIS3_FCL, VOLN, TWO_THIRD);
      double FCQ = FCL * CAQ * HUNDRED;
      FCL = FCL * SSP;
      double G11 = ONE - PX1H1, G21 = -ONE - PX2H1, G31 = ONE - PX3H1, G41 = -ONE - PX4H1;
      double G51 = ONE + PX3H1, G61 = -ONE + PX4H1, G71 = ONE + PX1H1, G81 = -ONE + PX2H1;
      double HGX1 = G11 * vx[0] + G21 * vx[1] + G31 * vx[2] + G41 * vx[3] + G51 * vx[4] + G61 * vx[5] + G71 * vx[6] + G81 * vx[7];
      double HGY1 = G11 * vy[0] + G21 * vy[1] + G31 * vy[2] + G41 * vy[3] + G51 * vy[4] + G61 * vy[5] + G71 * vy[6] + G81 * vy[7];
      double HGZ1 = G11 * vz[0] + G21 * vz[1] + G31 * vz[2] + G41 * vz[3] + G51 * vz[4] + G61 * vz[5] + G71 * vz[6] + G81 * vz[7];
      double G12 = ONE - PX1H2, G22 = ONE - PX2H2, G32 = -ONE - PX3H2, G42 = -ONE - PX4H2;
      double G52 = -ONE + PX3H2, G62 = -ONE + PX4H2, G72 = ONE + PX1H2, G82 = ONE + PX2H2;
      double HGX2 = G12 * vx[0] + G22 * vx[1] + G32 * vx[2] + G42 * vx[3] + G52 * vx[4] + G62 * vx[5] + G72 * vx[6] + G82 * vx[7];
      double HGY2 = G12 * vy[0] + G22 * vy[1] + G32 * vy[2] + G42 * vy[3] + G52 * vy[4] + G62 * vy[5] + G72 * vy[6] + G82 * vy[7];
      double HGZ2 = G12 * vz[0] + G22 * vz[1] + G32 * vz[2] + G42 * vz[3] + G52 * vz[4] + G62 * vz[5] + G72 * vz[6] + G82 * vz[7];
      double G13 = ONE - PX1H3, G23 = -ONE - PX2H3, G33 = -ONE - PX3H3, G43 = ONE - PX4H3;
      double G53 = -ONE + PX3H3, G63 = ONE + PX4H3, G73 = ONE + PX1H3, G83 = -ONE + PX2H3;
      double HGX3 = G13 * vx[0] + G23 * vx[1] + G33 * vx[2] + G43 * vx[3] + G53 * vx[4] + G63 * vx[5] + G73 * vx[6] + G83 * vx[7];
      double HGY3 = G13 * vy[0] + G23 * vy[1] + G33 * vy[2] + G43 * vy[3] + G53 * vy[4] + G63 * vy[5] + G73 * vy[6] + G83 * vy[7];
      double HGZ3 = G13 * vz[0] + G23 * vz[1] + G33 * vz[2] + G43 * vz[3] + G53 * vz[4] + G63 * vz[5] + G73 * vz[6] + G83 * vz[7];
      double HGX4 = vx[0] - vx[1] + vx[2] - vx[3] - vx[4] + vx[5] - vx[6] + vx[7];
      double HGY4 = vy[0] - vy[1] + vy[2] - vy[3] - vy[4] + vy[5] - vy[6] + vy[7];
      double HGZ4 = vz[0] - vz[1] + vz[2] - vz[3] - vz[4] + vz[5] - vz[6] + vz[7];
      double HX1 = HGX1 * (FCL + fabs(HGX1) * FCQ), HX2 = HGX2 * (FCL + fabs(HGX2) * FCQ);
      double HX3 = HGX3 * (FCL + fabs(HGX3) * FCQ), HX4 = HGX4 * (FCL + fabs(HGX4) * FCQ);
      double HY1 = HGY1 * (FCL + fabs(HGY1) * FCQ), HY2 = HGY2 * (FCL + fabs(HGY2) * FCQ);
      double HY3 = HGY3 * (FCL + fabs(HGY3) * FCQ), HY4 = HGY4 * (FCL + fabs(HGY4) * FCQ);
      double HZ1 = HGZ1 * (FCL + fabs(HGZ1) * FCQ), HZ2 = HGZ2 * (FCL + fabs(HGZ2) * FCQ);
      double HZ3 = HGZ3 * (FCL + fabs(HGZ3) * FCQ), HZ4 = HGZ4 * (FCL + fabs(HGZ4) * FCQ);
      F1[0] = -G11 * HX1 - G12 * HX2 - G13 * HX3 - HX4;
      F1[1] = -G21 * HX1 - G22 * HX2 - G23 * HX3 + HX4;
      F1[2] = -G31 * HX1 - G32 * HX2 - G33 * HX3 - HX4;
      F1[3] = -G41 * HX1 - G42 * HX2 - G43 * HX3 + HX4;
      F1[4] = -G51 * HX1 - G52 * HX2 - G53 * HX3 + HX4;
      F1[5] = -G61 * HX1 - G62 * HX2 - G63 * HX3 - HX4;
      F1[6] = -G71 * HX1 - G72 * HX2 - G73 * HX3 + HX4;
      F1[7] = -G81 * HX1 - G82 * HX2 - G83 * HX3 - HX4;
      F2[0] = -G11 * HY1 - G12 * HY2 - G13 * HY3 - HY4;
      F2[1] = -G21 * HY1 - G22 * HY2 - G23 * HY3 + HY4;
      F2[2] = -G31 * HY1 - G32 * HY2 - G33 * HY3 - HY4;
      F2[3] = -G41 * HY1 - G42 * HY2 - G43 * HY3 + HY4;
      F2[4] = -G51 * HY1 - G52 * HY2 - G53 * HY3 + HY4;
      F2[5] = -G61 * HY1 - G62 * HY2 - G63 * HY3 - HY4;
      F2[6] = -G71 * HY1 - G72 * HY2 - G73 * HY3 + HY4;
      F2[7] = -G81 * HY1 - G82 * HY2 - G83 * HY3 - HY4;
      F3[0] = -G11 * HZ1 - G12 * HZ2 - G13 * HZ3 - HZ4;
      F3[1] = -G21 * HZ1 - G22 * HZ2 - G23 * HZ3 + HZ4;
      F3[2] = -G31 * HZ1 - G32 * HZ2 - G33 * HZ3 - HZ4;
      F3[3] = -G41 * HZ1 - G42 * HZ2 - G43 * HZ3 + HZ4;
      F3[4] = -G51 * HZ1 - G52 * HZ2 - G53 * HZ3 + HZ4;
      F3[5] = -G61 * HZ1 - G62 * HZ2 - G63 * HZ3 - HZ4;
      F3[6] = -G71 * HZ1 - G72 * HZ2 - G73 * HZ3 + HZ4;
      F3[7] = -G81 * HZ1 - G82 * HZ2 - G83 * HZ3 - HZ4;
      EHOU_e = DT1 * (HZ1 * HGZ1 + HZ2 * HGZ2 + HZ3 * HGZ3 + HZ4 * HGZ4 + HX1 * HGX1 + HX2 * HGX2 + HX3 * HGX3 +
                      HX4 * HGX4 + HY1 * HGY1 + HY2 * HGY2 + HY3 * HGY3 + HY4 * HGY4);
    }
    if (EHOU) EHOU[e] = EHOU[e] + EHOU_e;

    /* ---- SFINT3 (sfint3.F), ISFINT/ICAA inactive, INT22=0, JEUL=0 ---- */
    {
      double s1 = (sg[0] - QNEW) * VOLN;
      double s2 = (sg[1] - QNEW) * VOLN;
      double s3 = (sg[2] - QNEW) * VOLN;
      double s4 = sg[3] * VOLN;
      double s5 = sg[4] * VOLN;
      double s6 = sg[5] * VOLN;
      const double PXk[4] = {PX1, PX2, PX3, PX4};
      const double PYk[4] = {PY1, PY2, PY3, PY4};
      const double PZk[4] = {PZ1, PZ2, PZ3, PZ4};
      const int opp[4] = {6, 7, 4, 5}; /* nodes 1-7, 2-8, 3-5, 4-6 */
      for (int k = 0; k < 4; ++k) {
        double FINT;
        FINT = s1 * PXk[k] + s4 * PYk[k] + s6 * PZk[k];
        F1[k] = F1[k] - FINT; F1[opp[k]] = F1[opp[k]] + FINT;
        FINT = s2 * PYk[k] + s4 * PXk[k] + s5 * PZk[k];
        F2[k] = F2[k] - FINT; F2[opp[k]] = F2[opp[k]] + FINT;
        FINT = s3 * PZk[k] + s6 * PXk[k] + s5 * PYk[k];
        F3[k] = F3[k] - FINT; F3[opp[k]] = F3[opp[k]] + FINT;
      }
    }

    /* ---- SCUMU3P (scumu3p.F), IPARTSPH=0, JTHE>=0 ---- */
    if (offg < ZERO) {
      for (int k = 0; k < 8; ++k) { F1[k] = ZERO; F2[k] = ZERO; F3[k] = ZERO; }
    }
    STI = FOURTH * STI;
    for (int k = 0; k < 8; ++k) {
      int64_t K = IADS[8 * e + k];
      FSKY[8 * K + 0] = F1[k];
      FSKY[8 * K + 1] = F2[k];
      FSKY[8 * K + 2] = F3[k];
      FSKY[8 * K + 6] = STI;
    }
  }
  *dt2t = DT2T;
}

void or_asspar4(int64_t nn, const int64_t *adsky, const double *fsky, double *a, double *ar,
                double *stifn, double *stifr) {
  for (int64_t n = 0; n < nn; ++n) {
    double a1 = ZERO, a2 = ZERO, a3 = ZERO, r1 = ZERO, r2 = ZERO, r3 = ZERO, sn = ZERO, sr = ZERO;
    for (int64_t k = adsky[n]; k < adsky[n + 1]; ++k) {
      const double *f = fsky + 8 * k;
      a1 = a1 + f[0]; a2 = a2 + f[1]; a3 = a3 + f[2];
      r1 = r1 + f[3]; r2 = r2 + f[4]; r3 = r3 + f[5];
      sn = sn + f[6]; sr = sr + f[7];
    }
    a[3 * n + 0] = a1; a[3 * n + 1] = a2; a[3 * n + 2] = a3;
    if (ar) { ar[3 * n + 0] = r1; ar[3 * n + 1] = r2; ar[3 * n + 2] = r3; }
    if (stifn) stifn[n] = sn;
    if (stifr) stifr[n] = sr;
  }
}

/* BCS1V (bcs1.F), global frame (ISK==1). */
static void bcs1v(int32_t code, double *a, double *v) {
  if (code & 4) { a[0] = ZERO; v[0] = ZERO; }
  if (code & 2) { a[1] = ZERO; v[1] = ZERO; }
  if (code & 1) { a[2] = ZERO; v[2] = ZERO; }
}

void or_cycle_hex(int64_t nn, int64_t ne, int64_t lsky, const int32_t *conn, double *x, double *v, double *a,
                  double *d, const double *ms, double *stifn, double *sig, double *eint, double *rho,
                  const double *vol0, double *pla, double *epsd, double *offg, double *qvis, double *ehou,
                  const int64_t *iads, const int64_t *adsky, double *fsky, const int32_t *bcs, int64_t nimp,
                  const int32_t *imp_node, const int32_t *imp_dof, const double *imp_val,
                  const or_hex_params *p, or_time_state *ts, int nthreads) {
  (void)lsky;
  /* resol.F:2721-2722, 2826 */
  ts->dt1 = ts->dt2;
  ts->dt2 = EP06;
  double DT2T = ts->dt2;
  /* element forces: groups of 128 (MVSIZ-1), the reference's OpenMP pattern */
  const int64_t G = 128;
  int64_t ngroups = (ne + G - 1) / G;
  int mstop = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(min : DT2T) \
    reduction(| : mstop)
#endif
  for (int64_t g = 0; g < ngroups; ++g) {
    int64_t e0 = g * G, e1 = e0 + G < ne ? e0 + G : ne;
    double dtl = DT2T;
    int ms = 0;
    or_hex_forces(e0, e1, ne, conn, x, v, sig, eint, rho, vol0, pla, epsd, offg, qvis, ehou, iads, fsky, p,
                  ts->dt1, &dtl, &ms);
    if (dtl < DT2T) DT2T = dtl;
    mstop |= ms;
  }
  (void)nthreads;
  ts->mstop |= mstop;
  /* ASSPAR4 */
  or_asspar4(nn, adsky, fsky, a, NULL, stifn, NULL);
  /* time step: resol.F:6125-6129, 6352, 6494-6497 */
  if (DT2T < ts->dt2) ts->dt2 = DT2T;
  {
    double d = ts->dt2;
    double g = DT_GROWTH * ts->dt2old;
    d = dmin(d, g);
    d = dmin(d, p->dtmx);
    ts->dt2 = d;
  }
  ts->dt2old = ts->dt2;
  ts->dt12 = HALF * (ts->dt1 + ts->dt2);
  if (ts->dt2 <= ZERO) ts->mstop |= 2; /* time step <= 0 stops the run (resol.F:8575, 9599) */
  /* ACCELE (accele.F:65-80) */
  for (int64_t n = 0; n < nn; ++n) {
    if (ms[n] > ZERO) {
      double rtmp = ONE / ms[n];
      a[3 * n + 0] = a[3 * n + 0] * rtmp;
      a[3 * n + 1] = a[3 * n + 1] * rtmp;
      a[3 * n + 2] = a[3 * n + 2] * rtmp;
    } else {
      a[3 * n + 0] = ZERO; a[3 * n + 1] = ZERO; a[3 * n + 2] = ZERO;
    }
  }
  /* BCS10 */
  if (bcs)
    for (int64_t n = 0; n < nn; ++n)
      if (bcs[n] != 0) bcs1v(bcs[n], a + 3 * n, v + 3 * n);
  /* FIXVEL global-frame branch (fixvel.F:366-379), constant imposed velocity */
  for (int64_t i = 0; i < nimp; ++i) {
    int64_t n = imp_node[i];
    int j = imp_dof[i];
    double yc = (imp_val[i] - v[3 * n + j]) / ts->dt12;
    a[3 * n + j] = yc;
  }
  /* VELOCITY (velocity.F:58-66) and DEPLA (displacement.F, IRESP=0) */
  for (int64_t n = 0; n < nn; ++n) {
    for (int c = 0; c < 3; ++c) {
      v[3 * n + c] = v[3 * n + c] + ts->dt12 * a[3 * n + c];
      a[3 * n + c] = ZERO;
    }
    for (int c = 0; c < 3; ++c) {
      double vdt = ts->dt2 * v[3 * n + c];
      d[3 * n + c] = d[3 * n + c] + vdt;
      x[3 * n + c] = x[3 * n + c] + vdt;
    }
  }
  /* resol.F:8605: TT_DOUBLE = TT_DOUBLE + DT2 */
  ts->tt = ts->tt + ts->dt2;
  ts->ncycle += 1;
}

void or_energy(int64_t nn, int64_t ne, const double *eint, const double *vol0, const double *ms, const double *v,
               double out[2]) {
  double ie = 0.0, ke = 0.0;
  for (int64_t e = 0; e < ne; ++e) ie += eint[e] * vol0[e];
  for (int64_t n = 0; n < nn; ++n)
    ke += ms[n] * (v[3 * n] * v[3 * n] + v[3 * n + 1] * v[3 * n + 1] + v[3 * n + 2] * v[3 * n + 2]);
  out[0] = ie;
  out[1] = HALF * ke;
}
