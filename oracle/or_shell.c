/*
 * or_shell.c -- CPU restatement of the OpenRadioss Engine Q4 shell path
 * (CFORC3 chain with law 2 through MULAWC / SIGEPS02C / M2CPLR), of the
 * mixed hex + Q4 cycle (ASSPAR4 with rotational rows, ACCELE with IN,
 * BCS10 on translations and rotations, FIXVEL, VELOCITY with VR, DEPLA)
 * and of its energies.
 *
 * TEST INFRASTRUCTURE ONLY (see or_oracle.h): the checker of tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 * Built with gcc -O2 -ffp-contract=off -fno-fast-math, so that every
 * expression rounds as the reference Fortran does under its release flags.
 *
 * Branch restated (engine defaults for /PROP/SHELL Ishell=1 with law 2):
 *   IHBE=1, ISMSTR=2 with |OFFG| /= 2 (no /DT/SHELL/CST), ISHFRAM=0 (symmetric
 *   convected frame), IGTYP=1 (NPT points with the COQINI tables), IREP=0,
 *   ISTRAIN=1, IEPSDOT=0, IPLA in {0,1,2}, ITHK in {0,1}, VP=2, IFORM=0,
 *   FISOKIN=0, no thermal, no failure model, no non-local, DM=0, ISIGI=0,
 *   NODADT=0, IDT1SH=0, IDTMINS=0, IDTMIN(3)=2, Parith/ON (CUPDT3P).
 */
#include "or_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZERO 0.0
#define ONE 1.0
#define TWO 2.0
#define THREE 3.0
#define FOUR 4.0
#define HALF (1.0 / 2.0)
#define THIRD (1.0 / 3.0)
#define FOURTH (1.0 / 4.0)
#define HUNDRED 100.0
#define EP06 1.0e6
#define EM20 (1.0 / 1.0e20)
#define EM15 (1.0 / 1.0e15)
#define EM30 (1.0 / (1.0e20 * 1.0e10))
#define ONE_OVER_9 (1.0 / 9.0)
#define ONE_OVER_12 (1.0 / 12.0)
#define FOUR_OVER_3 (4.0 / 3.0)
#define FOUR_OVER_5 (4.0 / 5.0)
#define THREE_OVER_4 (3.0 / 4.0)
#define EM01 (1.0 / 10.0)

/* Parameter constants of constant_mod.F, built by the same chains of
 * double-precision operations (constant_mod.F:651-876). */
static double c_em04(void) { volatile double ep04 = ((100.0 * 10.0) * 10.0); return 1.0 / ep04; }
static double c_zep81(void) {
  /* ZEP81 = ONE - ZEP19, ZEP19 = ZEP1 + ZEP09, ZEP09 = ZEP1 - ZEP01 */
  volatile double zep1 = 1.0 / 10.0, zep01 = 1.0 / 100.0;
  volatile double zep09 = zep1 - zep01;
  volatile double zep19 = zep1 + zep09;
  return 1.0 - zep19;
}
static double c_zep072169(void) {
  /* ZEP07 + TWOEM3 + EM04 + SIXEM5 + NINEEM6 (constant_mod.F:820) */
  volatile double zep1 = 1.0 / 10.0, zep03 = 3.0 / 100.0;
  volatile double zep07 = zep1 - zep03;
  volatile double twoem3 = 2.0 / 1000.0;
  volatile double em04 = c_em04();
  volatile double ep05 = (((100.0 * 10.0) * 10.0) * 10.0);
  volatile double fourem5 = 4.0 / ep05;
  volatile double sixem5 = em04 - fourem5;
  volatile double em05 = 1.0 / ep05;
  volatile double ep06 = ep05 * 10.0;
  volatile double em06 = 1.0 / ep06;
  volatile double nineem6 = em05 - em06;
  volatile double s = zep07 + twoem3;
  s = s + em04;
  s = s + sixem5;
  return s + nineem6;
}

double or_shell_const(int which) {
  switch (which) {
  case 0: return c_zep81();
  case 1: return c_zep072169();
  case 2: { volatile double two = 2.0; return sqrt(two) * HALF; } /* SR2D2 of chvis3.F */
  default: return 0.0;
  }
}

/* Through-thickness tables of COQINI (coqini.F:46-152): Z01, WF1 and the
 * WM1 of COQINI_WM.  The DATA constants are default-kind (single precision)
 * literals stored into REAL*8 arrays, hence the float literals widened here.
 * Row n-1 holds the n-point rule. */
static const float Z01T[11][11] = {
    {0.f},
    {-.5f, 0.5f},
    {-.5f, 0.f, 0.5f},
    {-.5f, -.1666667f, 0.1666667f, 0.5f},
    {-.5f, -.25f, 0.f, 0.25f, 0.5f},
    {-.5f, -.3f, -.1f, 0.1f, 0.3f, 0.5f},
    {-.5f, -.3333333f, -.1666667f, 0.0f, 0.1666667f, 0.3333333f, 0.5f},
    {-.5f, -.3571429f, -.2142857f, -.0714286f, 0.0714286f, 0.2142857f, 0.3571429f, 0.5f},
    {-.5f, -.375f, -.25f, -.125f, 0.0f, 0.125f, 0.25f, 0.375f, 0.5f},
    {-.5f, -.3888889f, -.2777778f, -.1666667f, -.0555555f, 0.0555555f, 0.1666667f, 0.2777778f, 0.3888889f, 0.5f},
    {-.5f, -.4f, -.3f, -.2f, -.1f, 0.f, 0.1f, 0.2f, 0.3f, 0.4f, 0.5f}};
static const float WF1T[11][11] = {
    {1.f},
    {0.5f, 0.5f},
    {0.25f, 0.5f, 0.25f},
    {0.1666667f, 0.3333333f, 0.3333333f, 0.1666667f},
    {0.125f, 0.25f, 0.25f, 0.25f, 0.125f},
    {0.1f, 0.2f, 0.2f, 0.2f, 0.2f, 0.1f},
    {0.0833333f, 0.1666667f, 0.1666667f, 0.1666667f, 0.1666667f, 0.1666667f, 0.0833333f},
    {0.0714286f, 0.1428571f, 0.1428571f, 0.1428571f, 0.1428571f, 0.1428571f, 0.1428571f, 0.0714286f},
    {0.0625f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.0625f},
    {0.0555556f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f,
     0.0555556f},
    {0.05f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.05f}};
static const float WM1T[11][11] = {
    {0.f},
    {-.0833333f, 0.0833333f},
    {-.0833333f, 0.f, 0.0833333f},
    {-.0648148f, -.0555556f, 0.0555556f, 0.0648148f},
    {-.0520833f, -.0625f, 0.f, 0.0625f, 0.0520833f},
    {-.0433333f, -.0600000f, -.0200000f, 0.0200000f, 0.0600000f, 0.0433333f},
    {-.0370370f, -.0555556f, -.0277778f, 0.0000000f, 0.0277778f, 0.0555556f, 0.0370370f},
    {-.0323129f, -.0510204f, -.0306122f, -.0102041f, 0.0102041f, 0.0306122f, 0.0510204f, 0.0323129f},
    {-.0286458f, -.0468750f, -.0312500f, -.0156250f, 0.0000000f, 0.0156250f, 0.0312500f, 0.0468750f, 0.0286458f},
    {-.0257202f, -.0432099f, -.0308642f, -.0185185f, -.0061728f, 0.0061728f, 0.0185185f, 0.0308642f, 0.0432099f,
     0.0257202f},
    {-.0233333f, -.0400000f, -.0300000f, -.0200000f, -.0100000f, 0.0000000f, 0.0100000f, 0.0200000f, 0.0300000f,
     0.0400000f, 0.0233333f}};

double or_shell_table(int which, int ipt, int npt) {
  if (npt < 1 || npt > 11 || ipt < 0 || ipt >= npt) return 0.0;
  if (which == 0) return (double)Z01T[npt - 1][ipt];
  if (which == 1) return (double)WF1T[npt - 1][ipt];
  return (double)WM1T[npt - 1][ipt];
}

static inline double dmax(double a, double b) { return (a > b) ? a : b; }
static inline double dmin(double a, double b) { return (a < b) ? a : b; }
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------------------
 * M2CPLR (m2cplr.F) for one integration point, IFORM=0, FISOKIN=0, IMPL_S=0.
 * In: sig (old, 5 comps), strain increments, PLA; out: new sig, PLA, DPLA, EZZ,
 * EPCHK.  off is OFF_OLD, the element OFF before this cycle's rupture test.
 * ---------------------------------------------------------------------- */
typedef struct {
  double young, g, nu, a1, a2, ca, cb, cn, ymax, cc, epdr, m_exp;
  int icc, ipla, israte, vp;
} m2c_t;

static void m2cplr_point(const m2c_t *c, double off, double gs, double epsp_in, const double so[5],
                         double dxx, double dyy, double dxy, double dyz, double dzx, double s[5],
                         double *pla, double *dpla_out, double *ezz_out, double *epchk) {
  const double SMALL = 1.0e-7; /* EM7 = ONE/EP07 */
  double CA = c->ca, CB = c->cb, YMAX = c->ymax, H = ZERO, EZZ = ZERO, DPLA = ZERO;
  double sxx = so[0], syy = so[1], sxy = so[2], syz = so[3], szx = so[4];
  /* elastic predictor */
  sxx = sxx + c->a1 * dxx + c->a2 * dyy;
  syy = syy + c->a2 * dxx + c->a1 * dyy;
  sxy = sxy + c->g * dxy;
  syz = syz + gs * dyz;
  szx = szx + gs * dzx;
  /* strain-rate factor (JOHNSON-COOK, TSTAR=0) */
  if (c->cc != ZERO) {
    double epsp = epsp_in;
    if (c->israte == 0 && c->vp == 2) epsp = dmax(dmax(fabs(dxx), fabs(dyy)), HALF * fabs(dxy));
    epsp = dmax(epsp, c->epdr);
    double logep = OR_LOGS(OR_SITE_M2CPLR, epsp / c->epdr);
    double q = (ONE + c->cc * logep);
    q = dmax(q, EM20);
    CA = CA * q;
    CB = CB * q;
    if (c->icc == 1) YMAX = YMAX * q;
  }
  /* yield */
  double YLD;
  if (*pla == ZERO) {
    YLD = CA;
  } else {
    double beta = CB * (ONE - ZERO);
    YLD = CA + beta * OR_EXPS(OR_SITE_M2CPLR, c->cn * OR_LOGS(OR_SITE_M2CPLR, *pla));
  }
  YLD = dmin(YLD, YMAX);
  if (c->ipla == 0) {
    double svm = sqrt(sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy);
    double r = dmin(ONE, YLD / (svm + EM15));
    if (r < ONE) {
      sxx = sxx * r;
      syy = syy * r;
      sxy = sxy * r;
      DPLA = off * dmax(ZERO, (svm - YLD) / c->young);
      double s1 = HALF * (sxx + syy);
      EZZ = DPLA * s1 / YLD;
      *pla = *pla + DPLA;
      *epchk = dmax(*pla, *epchk);
      if (YLD >= YMAX) H = ZERO;
      else H = c->cn * CB * OR_EXPS(OR_SITE_M2CPLR, (c->cn - ONE) * OR_LOGS(OR_SITE_M2CPLR, *pla + SMALL));
    }
  } else if (c->ipla == 1) {
    double s1 = sxx + syy, s2 = sxx - syy, s3 = sxy;
    double A = FOURTH * s1 * s1;
    double B = THREE_OVER_4 * s2 * s2 + THREE * s3 * s3;
    double svm = sqrt(A + B);
    if (svm > YLD && off == ONE) {
      double nu1 = ONE / (ONE - c->nu);
      double nu2 = ONE / (ONE + c->nu);
      if (YLD >= YMAX) H = ZERO;
      else H = c->cn * CB * OR_EXPS(OR_SITE_M2CPLR, (c->cn - ONE) * OR_LOGS(OR_SITE_M2CPLR, *pla + SMALL));
      double dpla_j = (svm - YLD) / (THREE * c->g + H);
      double anu1 = A * nu1;
      double bnu2 = THREE * B * nu2;
      double h2 = TWO * H;
      double dpla_i = ZERO, P = ONE, Q = ONE, DR = ZERO;
      for (int it = 0; it < 3; ++it) {
        dpla_i = dpla_j;
        double pla_i = *pla + dpla_i;
        DPLA = dpla_j;
        double yld_i;
        if (pla_i == ZERO) yld_i = dmin(YMAX, CA);
        else yld_i = dmin(YMAX, CA + CB * OR_EXPS(OR_SITE_M2CPLR, c->cn * OR_LOGS(OR_SITE_M2CPLR, pla_i)));
        DR = HALF * c->young * dpla_i / yld_i;
        P = ONE / (ONE + DR * nu1);
        Q = ONE / (ONE + THREE * DR * nu2);
        double p2 = P * P, q2 = Q * Q;
        double F = A * p2 + B * q2 - yld_i * yld_i;
        double DF = -(anu1 * p2 * P + bnu2 * q2 * Q) * (c->young - DR * h2) / yld_i - h2 * yld_i;
        if (dpla_i > ZERO) dpla_j = dmax(ZERO, dpla_i - F / DF);
        else dpla_j = ZERO;
      }
      *pla = *pla + dpla_i;
      *epchk = dmax(*pla, *epchk);
      s1 = (sxx + syy) * P;
      s2 = (sxx - syy) * Q;
      sxx = HALF * (s1 + s2);
      syy = HALF * (s1 - s2);
      sxy = sxy * Q;
      EZZ = DR * s1 / c->young;
    }
  } else { /* IPLA == 2 */
    double svm2 = sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy;
    double svm = sqrt(svm2);
    double yld2 = YLD * YLD;
    if (svm2 > yld2 && off == ONE) {
      if (YLD >= YMAX) H = ZERO;
      else H = c->cn * CB * OR_EXPS(OR_SITE_M2CPLR, (c->cn - ONE) * OR_LOGS(OR_SITE_M2CPLR, *pla + SMALL));
      double aa = (svm2 - yld2) / (5.0 * svm2 + THREE * (-sxx * syy + sxy * sxy));
      double s1 = (ONE - TWO * aa) * sxx + aa * syy;
      double s2 = aa * sxx + (ONE - TWO * aa) * syy;
      double s3 = (ONE - THREE * aa) * sxy;
      sxx = s1; syy = s2; sxy = s3;
      DPLA = off * (svm - YLD) / (THREE * c->g + H);
      *pla = *pla + DPLA;
      YLD = YLD + H * DPLA;
      svm = sqrt(sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy);
      double r = dmin(ONE, YLD / dmax(EM20, svm));
      sxx = sxx * r; syy = syy * r; sxy = sxy * r;
      EZZ = DPLA * HALF * (sxx + syy) / YLD;
    }
  }
  s[0] = sxx; s[1] = syy; s[2] = sxy; s[3] = syz; s[4] = szx;
  *dpla_out = DPLA;
  *ezz_out = EZZ;
}

/* ------------------------------------------------------------------------
 * CFORC3 chain for Q4 elements [e0,e1).
 * ---------------------------------------------------------------------- */
void or_q4_forces(int64_t e0, int64_t e1, const int32_t *conn, const double *X, const double *V,
                  const double *VR, or_q4_state *s, const int64_t *IADC, double *FSKY,
                  const or_q4_params *p, double DT1, double *dt2t) {
  or_q4_forces_dt(e0, e1, conn, X, V, VR, s, IADC, FSKY, p, DT1, dt2t, NULL);
}

void or_q4_forces_dt(int64_t e0, int64_t e1, const int32_t *conn, const double *X, const double *V,
                     const double *VR, or_q4_state *s, const int64_t *IADC, double *FSKY,
                     const or_q4_params *p, double DT1, double *dt2t, double *dtel) {
  or_q4_forces_dte(e0, e1, conn, X, V, VR, s, IADC, FSKY, p, DT1, dt2t, dtel, NULL);
}

void or_q4_forces_dte(int64_t e0, int64_t e1, const int32_t *conn, const double *X, const double *V,
                      const double *VR, or_q4_state *s, const int64_t *IADC, double *FSKY,
                      const or_q4_params *p, double DT1, double *dt2t, double *dtel, double *ehou) {
  const int64_t ne = s->ne;
  const int npt = p->npt;
  const double ZEP81 = c_zep81(), ZEP072169 = c_zep072169();
  const double SR2D2 = or_shell_const(2);
  const double HVISC = HALF, HELAS = HALF, HVLIN = ZERO; /* radioss2.F:641-643 */
  const double SRH1 = sqrt(p->hm), SRH2 = sqrt(p->hf), SRH3 = sqrt(p->hr); /* resol_init.F:684-686 */
  m2c_t mc;
  mc.young = p->young; mc.g = p->shear; mc.nu = p->nu;
  mc.a1 = p->young / (ONE - p->nu * p->nu); /* sigeps02c.F: a11 = young/(one-nu**2) */
  mc.a2 = mc.a1 * p->nu;
  mc.ca = p->ca; mc.cb = p->cb; mc.cn = p->cn; mc.ymax = p->ymax; mc.cc = p->cc;
  mc.epdr = dmax(p->epdr * DT1, EM20);
  mc.m_exp = p->m_exp; mc.icc = p->icc; mc.ipla = p->ipla; mc.israte = p->israte; mc.vp = p->vp;
  const double dtinv = DT1 / dmax(DT1 * DT1, EM20);
  const double asrate = (p->israte > 0) ? dmin(ONE, p->asrate * DT1) : ONE;
  double DT2T = *dt2t;
  for (int64_t e = e0; e < e1; ++e) {
    const int32_t *nc = conn + 4 * e;
    /* ---- CCOOR3 (ccoor3.F:68-150) ---- */
    double xg[4], yg[4], zg[4], vl[4][3], vrl[4][3];
    for (int k = 0; k < 4; ++k) {
      const int64_t n = nc[k];
      xg[k] = X[3 * n]; yg[k] = X[3 * n + 1]; zg[k] = X[3 * n + 2];
      for (int c = 0; c < 3; ++c) { vl[k][c] = V[3 * n + c]; vrl[k][c] = VR[3 * n + c]; }
    }
    double offg = s->offg[e];
    double off = dmin(ONE, fabs(offg));
    if (offg < ZERO)
      for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 3; ++c) { vl[k][c] = ZERO; vrl[k][c] = ZERO; }
    double thk0 = s->thke[e];
    /* ---- CNVEC3, ISHFRAM=0 (cnvec3.F:86-138) ---- */
    double E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y, E3Z;
    {
      double X21 = xg[1] - xg[0], X32 = xg[2] - xg[1], X34 = xg[2] - xg[3], X41 = xg[3] - xg[0];
      double Y21 = yg[1] - yg[0], Y32 = yg[2] - yg[1], Y34 = yg[2] - yg[3], Y41 = yg[3] - yg[0];
      double Z21 = zg[1] - zg[0], Z32 = zg[2] - zg[1], Z34 = zg[2] - zg[3], Z41 = zg[3] - zg[0];
      E1X = (X21 + X34); E1Y = (Y21 + Y34); E1Z = (Z21 + Z34);
      E2X = (X32 + X41); E2Y = (Y32 + Y41); E2Z = (Z32 + Z41);
      E3X = E1Y * E2Z - E1Z * E2Y;
      E3Y = E1Z * E2X - E1X * E2Z;
      E3Z = E1X * E2Y - E1Y * E2X;
      double suma = E3X * E3X + E3Y * E3Y + E3Z * E3Z;
      suma = ONE / dmax(sqrt(suma), EM20);
      E3X = E3X * suma; E3Y = E3Y * suma; E3Z = E3Z * suma;
      double S1 = E1X * E1X + E1Y * E1Y + E1Z * E1Z;
      double S2 = E2X * E2X + E2Y * E2Y + E2Z * E2Z;
      suma = sqrt(S1 / S2);
      E1X = E1X + (E2Y * E3Z - E2Z * E3Y) * suma;
      E1Y = E1Y + (E2Z * E3X - E2X * E3Z) * suma;
      E1Z = E1Z + (E2X * E3Y - E2Y * E3X) * suma;
      suma = E1X * E1X + E1Y * E1Y + E1Z * E1Z;
      suma = ONE / dmax(sqrt(suma), EM20);
      E1X = E1X * suma; E1Y = E1Y * suma; E1Z = E1Z * suma;
      E2X = E3Y * E1Z - E3Z * E1Y;
      E2Y = E3Z * E1X - E3X * E1Z;
      E2Z = E3X * E1Y - E3Y * E1X;
    }
    /* ---- CDERI3 (cderi3.F:100-190) ---- */
    double X2, X3, X4, Y2, Y3, Y4, Z2, PX1, PX2, PY1, PY2, AREA, VHX, VHY;
    {
      double X21GA = xg[1] - xg[0], Y21GA = yg[1] - yg[0], Z21GA = zg[1] - zg[0];
      double X31GA = xg[2] - xg[0], Y31GA = yg[2] - yg[0], Z31GA = zg[2] - zg[0];
      double X41GA = xg[3] - xg[0], Y41GA = yg[3] - yg[0], Z41GA = zg[3] - zg[0];
      X2 = E1X * X21GA + E1Y * Y21GA + E1Z * Z21GA;
      Y2 = E2X * X21GA + E2Y * Y21GA + E2Z * Z21GA;
      Y3 = E2X * X31GA + E2Y * Y31GA + E2Z * Z31GA;
      X3 = E1X * X31GA + E1Y * Y31GA + E1Z * Z31GA;
      X4 = E1X * X41GA + E1Y * Y41GA + E1Z * Z41GA;
      Y4 = E2X * X41GA + E2Y * Y41GA + E2Z * Z41GA;
      Z2 = E3X * X21GA + E3Y * Y21GA + E3Z * Z21GA;
      PX1 = HALF * (Y2 - Y4);
      PY1 = HALF * (X4 - X2);
      PX2 = HALF * Y3;
      PY2 = -HALF * X3;
      AREA = dmax(TWO * (PY2 * PX1 - PY1 * PX2), EM20);
      VHX = (-X2 + X3 - X4) / AREA;
      VHY = (-Y2 + Y3 - Y4) / AREA;
    }
    (void)Z2;
    /* ---- CCOEF3 (ccoef3.F:84-170) ---- */
    double VOL0, THK02;
    {
      if (p->ithk > 0) {
        thk0 = s->thk[e];
        VOL0 = thk0 * AREA;
      } else {
        VOL0 = thk0 * AREA;
      }
      THK02 = thk0 * thk0;
    }
    double SHF;
    {
      double fac1tmp = 2.0 * (1.0 + p->nu) * THK02;
      double ish = (double)p->ishear;
      double fsh = p->fsh;
      SHF = fsh * ((1.0 - ish) + ish * fac1tmp / (fsh * AREA + fac1tmp));
    }
    const double GS = p->shear * SHF;
    /* ---- CDLEN3 (cdlen3.F:68-124), IHBE /= 0 ---- */
    double ALDT;
    {
      double AL1 = X2 * X2 + Y2 * Y2;
      double AL2 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
      double AL6 = X3 * X3 + Y3 * Y3;
      double AL3 = (X4 - X3) * (X4 - X3) + (Y4 - Y3) * (Y4 - Y3);
      double AL4 = X4 * X4 + Y4 * Y4;
      double AL5 = (X4 - X2) * (X4 - X2) + (Y4 - Y2) * (Y4 - Y2);
      double ALMIN = dmin(dmin(AL1, AL2), AL4);
      double ALQUAD = dmin(dmin(AL3, AL5), AL6);
      if (AL3 != ZERO) ALMIN = dmin(ALMIN, ALQUAD);
      double DTDYN = AREA * AREA / dmax(dmax(AL5, AL6), EM20);
      ALDT = dmax(DTDYN, ALMIN);
      double DTHOUR = HALF * (ALMIN + ALDT) / dmax(p->hm, p->hf);
      if (DTHOUR < ALDT) ALDT = DTHOUR;
      ALDT = sqrt(ALDT);
    }
    /* ---- CDEFO3, IHBE<=1 (cdefo3.F:94-140) ---- */
    double VX[4], VY[4], VZ[4];
    for (int k = 0; k < 4; ++k) {
      VX[k] = E1X * vl[k][0] + E1Y * vl[k][1] + E1Z * vl[k][2];
      VY[k] = E2X * vl[k][0] + E2Y * vl[k][1] + E2Z * vl[k][2];
      VZ[k] = E3X * vl[k][0] + E3Y * vl[k][1] + E3Z * vl[k][2];
    }
    double EXX, EYY, EXY, EXZ, EYZ;
    {
      double VZ13 = VZ[0] - VZ[2], VZ24 = VZ[1] - VZ[3];
      EYZ = PY1 * VZ13 + PY2 * VZ24;
      EXZ = PX1 * VZ13 + PX2 * VZ24;
      double DT1V4 = FOURTH * DT1;
      double TMP2A = PY2 + PY1;
      double TMP3A = fsign(dmax(fabs(TMP2A), EM20), TMP2A);
      double TMP1A = DT1V4 * (VZ13 - VZ24) * (VZ13 - VZ24) / TMP3A;
      double VX13 = VX[0] - VX[2], VX24 = VX[1] - VX[3];
      VX13 = VX13 - TMP1A;
      VX24 = VX24 + TMP1A;
      EXX = PX1 * VX13 + PX2 * VX24;
      EXY = PY1 * VX13 + PY2 * VX24;
      double TMP1B = PX2 - PX1;
      double TMP3B = fsign(dmax(fabs(TMP1B), EM20), TMP1B);
      double TMP2B = DT1V4 * (VZ13 + VZ24) * (VZ13 + VZ24) / TMP3B;
      double VY13 = VY[0] - VY[2], VY24 = VY[1] - VY[3];
      VY13 = VY13 + TMP2B;
      VY24 = VY24 + TMP2B;
      EXY = EXY + PX1 * VY13 + PX2 * VY24;
      EYY = PY1 * VY13 + PY2 * VY24;
    }
    /* ---- CCURV3 (ccurv3.F:64-92) ---- */
    double RX[4], RY[4], KXX, KYY, KXY;
    for (int k = 0; k < 4; ++k) {
      RX[k] = E1X * vrl[k][0] + E1Y * vrl[k][1] + E1Z * vrl[k][2];
      RY[k] = E2X * vrl[k][0] + E2Y * vrl[k][1] + E2Z * vrl[k][2];
    }
    {
      double RX13TA = RX[0] - RX[2];
      double RXAVTA = RX[0] + RX[1] + RX[2] + RX[3];
      double RX24TA = RX[1] - RX[3];
      KYY = -PY1 * RX13TA - PY2 * RX24TA;
      KXY = PX1 * RX13TA + PX2 * RX24TA;
      double RY13TA = RY[0] - RY[2];
      double RYAVTA = RY[0] + RY[1] + RY[2] + RY[3];
      double RY24TA = RY[1] - RY[3];
      KXX = PX1 * RY13TA + PX2 * RY24TA;
      KXY = PY1 * RY13TA + PY2 * RY24TA - KXY;
      EXZ = EXZ + RYAVTA * (0.25 * AREA);
      EYZ = EYZ - RXAVTA * (0.25 * AREA);
    }
    /* ---- CSTRA3 (cstra3.F:83-93, 168-178), ISTRAIN=1, ISMSTR=2 ---- */
    {
      double FAC1 = DT1 / AREA;
      EXX = EXX * FAC1; EYY = EYY * FAC1; EXY = EXY * FAC1; EYZ = EYZ * FAC1; EXZ = EXZ * FAC1;
      KXX = KXX * FAC1; KYY = KYY * FAC1; KXY = KXY * FAC1;
      double *g = s->gstr;
      g[0 * ne + e] = g[0 * ne + e] + EXX;
      g[1 * ne + e] = g[1 * ne + e] + EYY;
      g[2 * ne + e] = g[2 * ne + e] + EXY;
      g[3 * ne + e] = g[3 * ne + e] + EYZ;
      g[4 * ne + e] = g[4 * ne + e] + EXZ;
      g[5 * ne + e] = g[5 * ne + e] + KXX;
      g[6 * ne + e] = g[6 * ne + e] + KYY;
      g[7 * ne + e] = g[7 * ne + e] + KXY;
    }
    /* ---- element strain rate of CFORC3 (cforc3.F:537-546), asrate = one ---- */
    double epsd_pg;
    {
      double thk = s->thk[e];
      double eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) * ONE_OVER_9 * (thk * thk);
      double eps_m2 = FOUR_OVER_3 * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
      epsd_pg = sqrt(eps_k2 + eps_m2) * dtinv;
      s->epsd[e] = ONE * epsd_pg + (ONE - ONE) * s->epsd[e];
    }
    /* ---- MULAWC law-2 branch (mulawc.F90:542-579, 745-1151, 2628-2660, 3015-3062) ---- */
    double FOR[5], MOM[3];
    for (int c = 0; c < 5; ++c) FOR[c] = s->forc[c * ne + e];
    for (int c = 0; c < 3; ++c) MOM[c] = s->mom[c * ne + e];
    double degmb = FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
    double degfx = MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
    VOL0 = AREA * thk0;
    double thkn = s->thk[e];
    for (int c = 0; c < 5; ++c) FOR[c] = ZERO;
    for (int c = 0; c < 3; ++c) MOM[c] = ZERO;
    const double off_old = off;
    double epchk = ZERO;
    int ioff_duct = 0;
    for (int ip = 0; ip < npt; ++ip) {
      const double thkly = (double)WF1T[npt - 1][ip];
      const double wmc = (double)WM1T[npt - 1][ip];
      const double posly = (double)Z01T[npt - 1][ip] + ZERO;
      const double thklyl = thkly * thk0;
      const double zt = posly * thk0;
      const double dxx = EXX + zt * KXX;
      const double dyy = EYY + zt * KYY;
      const double dxy = EXY + zt * KXY;
      const double dyz = EYZ, dzx = EXZ;
      double *sg = s->sig + (int64_t)ip * 5 * ne;
      double so[5], sn[5];
      for (int c = 0; c < 5; ++c) so[c] = sg[c * ne + e];
      /* SIGEPS02C (sigeps02c.F:120-133, VP=2) */
      double *epsd_ip = s->epsd_ip + (int64_t)ip * ne + e;
      *epsd_ip = asrate * epsd_pg + (ONE - asrate) * (*epsd_ip);
      const double epsdot = (*epsd_ip) * DT1;
      double *pla = s->pla + (int64_t)ip * ne + e;
      double dpla, ezz;
      /* SIGEPS02C hands OFF_OLD to M2CPLR (sigeps02c.F:135) */
      m2cplr_point(&mc, off_old, GS, epsdot, so, dxx, dyy, dxy, dyz, dzx, sn, pla, &dpla, &ezz, &epchk);
      /* progressive rupture (sigeps02c.F:147-156) */
      if (off == off_old && off > ZERO) {
        if (off == ONE && epchk >= p->epmx) {
          off = FOUR_OVER_5;
          ioff_duct = 1;
        } else if (off < ONE) {
          off = off * FOUR_OVER_5;
        }
      }
      /* thickness (sigeps02c.F:166-173), INLOC=0 */
      ezz = -(dxx + dyy) * p->nu - (ONE - TWO * p->nu) * ezz;
      ezz = ezz / (ONE - p->nu);
      thkn = thkn + ezz * thklyl * off;
      for (int c = 0; c < 5; ++c) sg[c * ne + e] = sn[c];
      FOR[0] = FOR[0] + thkly * sn[0];
      FOR[1] = FOR[1] + thkly * sn[1];
      FOR[2] = FOR[2] + thkly * sn[2];
      FOR[3] = FOR[3] + thkly * sn[3];
      FOR[4] = FOR[4] + thkly * sn[4];
      MOM[0] = MOM[0] + wmc * sn[0];
      MOM[1] = MOM[1] + wmc * sn[1];
      MOM[2] = MOM[2] + wmc * sn[2];
    }
    /* element rupture test (mulawc.F90:2898-2912) */
    if ((off == FOUR_OVER_5 && ioff_duct == 0) || (off > ZERO && off_old < EM01)) off = ZERO;
    s->thk[e] = dmax(thkn, EM30);
    for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] * off * ONE;
    for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] * off * ONE;
    degmb = degmb + FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
    degfx = degfx + MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
    {
      double vol2 = HALF * VOL0;
      s->eint[e] = s->eint[e] + degmb * vol2;
      s->eint[ne + e] = s->eint[ne + e] + degfx * thk0 * vol2;
    }
    for (int c = 0; c < 5; ++c) s->forc[c * ne + e] = FOR[c];
    for (int c = 0; c < 3; ++c) s->mom[c * ne + e] = MOM[c];
    /* VISCMX = SQRT(ONE+VISCMX**2)-VISCMX with VISCMX = MAX(DM,0) = 0 (cforc3.F:576-578) */
    const double VISCMX = ONE, ALPE = ONE;
    /* ---- CHVIS3, IHBE=1, ISMSTR=2 (chvis3.F:112-321) ---- */
    double H11, H12, H13, H21, H22, H23, H31, H32, H33, B1k, B2k;
    {
      double PX1V = PX1 * VHX, PX2V = PX2 * VHX, PY1V = PY1 * VHY, PY2V = PY2 * VHY;
      double GAMA1 = off * (ONE - PX1V - PY1V);
      double GAMA3 = off * (ONE + PX1V + PY1V);
      double GAMA2 = off * (-ONE - PX2V - PY2V);
      double GAMA4 = off * (-ONE + PX2V + PY2V);
      double SHFPR3 = SHF / (THREE * (ONE + p->nu));
      double HVISH1 = HVISC * p->hm, HVISH2 = HVISC * p->hf;
      double R0 = FOURTH * p->rho;
      double R1 = R0 * HUNDRED;
      R0 = R0 * HVLIN;
      double A1 = R1 * HVISH1;
      double A2 = R0 * SR2D2 * SRH1;
      double SRSHFPR3 = sqrt(SHFPR3);
      double A3 = R1 * HVISH2 * SRSHFPR3;
      double A4 = R0 * SR2D2 * SRH2 * SRSHFPR3;
      double HH3 = HELAS * p->hr;
      double A5 = HH3 * R1 * ZEP072169;
      HH3 = SR2D2 * SRH3;
      double A6 = HH3 * R0 * ZEP072169;
      R0 = FOURTH * p->young * HELAS;
      double A7 = p->hm * R0;
      double A8 = p->hf * R0 * SHFPR3;
      double T2A = THK02 * AREA;
      double TSA = sqrt(T2A);
      double H1Q = A1 * TSA, H1L = A2 * p->ssp * TSA;
      double H2Q = A3 * THK02, H2L = A4 * p->ssp * THK02;
      double H3Q = A5 * T2A, H3L = A6 * p->ssp * T2A;
      double TD = thk0 * DT1;
      double HH1 = A7 * TD;
      double B1 = PX1 * PX1 + PY1 * PY1, B2 = PX2 * PX2 + PY2 * PY2;
      double HH2 = A8 * THK02 * TD / (B1 + B2);
      if (nc[2] == nc[3]) { /* triangles */
        H1Q = ZERO; H1L = ZERO; H2Q = ZERO; H2L = ZERO; H3Q = ZERO; H3L = ZERO; HH1 = ZERO; HH2 = ZERO;
      }
      double *hour = s->hourg;
      double HG1 = VX[0] * GAMA1 + VX[1] * GAMA2 + VX[2] * GAMA3 + VX[3] * GAMA4;
      double HG2 = VY[0] * GAMA1 + VY[1] * GAMA2 + VY[2] * GAMA3 + VY[3] * GAMA4;
      double hour1 = hour[0 * ne + e] + HG1 * HH1;
      double hour2 = hour[1 * ne + e] + HG2 * HH1;
      hour[0 * ne + e] = hour1;
      hour[1 * ne + e] = hour2;
      double HOUR1A = hour1 + HG1 * (H1L + H1Q * fabs(HG1));
      H11 = HOUR1A * GAMA1; H12 = HOUR1A * GAMA2; H13 = HOUR1A * GAMA3;
      double HOUR2A = hour2 + HG2 * (H1L + H1Q * fabs(HG2));
      H21 = HOUR2A * GAMA1; H22 = HOUR2A * GAMA2; H23 = HOUR2A * GAMA3;
      double EH = HOUR1A * HG1 + HOUR2A * HG2; /* chvis3.F:299 */
      HG1 = VZ[0] * GAMA1 + VZ[1] * GAMA2 + VZ[2] * GAMA3 + VZ[3] * GAMA4;
      double hour3 = hour[2 * ne + e] + HG1 * HH2;
      hour[2 * ne + e] = hour3;
      double HOUR3A = hour3 + HG1 * (H2L + H2Q * fabs(HG1));
      H31 = HOUR3A * GAMA1; H32 = HOUR3A * GAMA2; H33 = HOUR3A * GAMA3;
      EH = EH + HOUR3A * HG1; /* chvis3.F:322 */
      HG1 = RX[0] - RX[1] + RX[2] - RX[3];
      HG2 = RY[0] - RY[1] + RY[2] - RY[3];
      double hour4 = HG1 * (H3L + H3Q * fabs(HG1));
      double hour5 = HG2 * (H3L + H3Q * fabs(HG2));
      hour[3 * ne + e] = hour4;
      hour[4 * ne + e] = hour5;
      EH = EH + hour4 * HG1 + hour5 * HG2; /* chvis3.F:336-338, DT1C = DT1 */
      if (ehou) ehou[e] = DT1 * EH * off;
      B1k = hour4 * off;
      B2k = hour5 * off;
    }
    /* ---- CDT3 (cdt3.F:108-112, 141-165, 200-233), IDTMIN(3)=2 ---- */
    double STI;
    {
      ALDT = ALDT * VISCMX / sqrt(ALPE);
      double DT = p->dtfac * ALDT / p->ssp;
      if (!(DT > p->dtmin || off < ONE || offg < ZERO)) off = ZERO; /* IDTMIN(3)=2 deletion */
      if (offg > ZERO && off != ZERO && DT < DT2T) DT2T = DT;
      /* the candidate CDT3 weighs for NELTS (cdt3.F:205-215): -1 when OFFG <= 0 or OFF = 0 */
      if (dtel) dtel[e] = (offg > ZERO && off != ZERO) ? DT : -ONE;
      double DIVM = dmax(ALDT * ALDT, EM20);
      STI = HALF * VOL0 * p->young / DIVM;
      STI = ZEP81 * STI * off;
    }
    double STIR = ZERO;
    /* ---- CFINT3, ISIGI=0, IHBE<2 (cfint3.F:139-208) ---- */
    double Fk[3][4], Mk[3][4];
    {
      double F1A = FOR[0] * thk0, F2A = FOR[1] * thk0, F3A = FOR[2] * thk0, F4A = FOR[3] * thk0,
             F5A = FOR[4] * thk0;
      double M4 = F4A * AREA, M5 = F5A * AREA;
      double F12 = F1A * PX2 + F3A * PY2;
      double F22 = F2A * PY2 + F3A * PX2;
      double F32 = F5A * PX2 + F4A * PY2;
      double F11 = F1A * PX1 + F3A * PY1;
      double F21 = F2A * PY1 + F3A * PX1;
      double F31 = F5A * PX1 + F4A * PY1;
      double G11 = F11 + H11, G13 = H13 - F11;
      double G21 = F21 + H21, G23 = H23 - F21;
      double G31 = F31 + H31, G33 = H33 - F31;
      double G12 = F12 + H12, G22 = F22 + H22, G32 = F32 + H32;
      Fk[0][0] = E1X * G11 + E2X * G21 + E3X * G31;
      Fk[0][1] = E1X * G12 + E2X * G22 + E3X * G32;
      Fk[0][2] = E1X * G13 + E2X * G23 + E3X * G33;
      Fk[1][0] = E1Y * G11 + E2Y * G21 + E3Y * G31;
      Fk[1][1] = E1Y * G12 + E2Y * G22 + E3Y * G32;
      Fk[1][2] = E1Y * G13 + E2Y * G23 + E3Y * G33;
      Fk[2][0] = E1Z * G11 + E2Z * G21 + E3Z * G31;
      Fk[2][1] = E1Z * G12 + E2Z * G22 + E3Z * G32;
      Fk[2][2] = E1Z * G13 + E2Z * G23 + E3Z * G33;
      for (int c = 0; c < 3; ++c) Fk[c][3] = -Fk[c][0] - Fk[c][1] - Fk[c][2];
      double M1A = MOM[0] * THK02, M2A = MOM[1] * THK02, M3A = MOM[2] * THK02;
      M4 = M4 * FOURTH;
      M5 = M5 * FOURTH;
      double M11 = -M2A * PY1 - M3A * PX1;
      double M21 = M1A * PX1 + M3A * PY1;
      double M12 = -M2A * PY2 - M3A * PX2;
      double M22 = M1A * PX2 + M3A * PY2;
      double G11T = M11 - M4 + B1k;
      double G13T = -M11 - M4 + B1k;
      double G12T = M12 - M4 + (-B1k);
      double G14T = -M12 - M4 + (-B1k);
      double G21T = M21 + M5 + B2k;
      double G23T = -M21 + M5 + B2k;
      double G22T = M22 + M5 + (-B2k);
      double G24T = -M22 + M5 + (-B2k);
      Mk[0][0] = E1X * G11T + E2X * G21T; Mk[0][1] = E1X * G12T + E2X * G22T;
      Mk[0][2] = E1X * G13T + E2X * G23T; Mk[0][3] = E1X * G14T + E2X * G24T;
      Mk[1][0] = E1Y * G11T + E2Y * G21T; Mk[1][1] = E1Y * G12T + E2Y * G22T;
      Mk[1][2] = E1Y * G13T + E2Y * G23T; Mk[1][3] = E1Y * G14T + E2Y * G24T;
      Mk[2][0] = E1Z * G11T + E2Z * G21T; Mk[2][1] = E1Z * G12T + E2Z * G22T;
      Mk[2][2] = E1Z * G13T + E2Z * G23T; Mk[2][3] = E1Z * G14T + E2Z * G24T;
    }
    /* ---- CUPDT3P (cupdt3.F:1000-1105) ---- */
    if (off < ONE) offg = off;
    if (offg < ZERO) {
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 4; ++k) { Fk[c][k] = ZERO; Mk[c][k] = ZERO; }
      STI = ZERO;
      STIR = ZERO;
    }
    s->offg[e] = offg;
    for (int k = 0; k < 4; ++k) {
      double *f = FSKY + 8 * IADC[4 * e + k];
      f[0] = -Fk[0][k]; f[1] = -Fk[1][k]; f[2] = -Fk[2][k];
      f[3] = -Mk[0][k]; f[4] = -Mk[1][k]; f[5] = -Mk[2][k];
      f[6] = STI; f[7] = STIR;
    }
  }
  *dt2t = DT2T;
}

/* ------------------------------------------------------------------------
 * One cycle of a mixed hex + Q4 model (RESOL order, resol.F:2721-9042).
 * ---------------------------------------------------------------------- */
void or_cycle_model(or_model_c *m, or_time_state *ts, int nthreads) {
  const double DT1 = ts->dt2; /* resol.F:2721-2722 */
  double DT2T = EP06;
  int mstop = 0;
  const int64_t G = 128;
  const int64_t ngh = (m->ne_hex + G - 1) / G, ngq = (m->ne_q4 + G - 1) / G;
  const int track = m->ehou_h != NULL || m->ehou_q != NULL;
  double *gsum = track ? (double *)calloc((size_t)(ngh + ngq > 0 ? ngh + ngq : 1), sizeof(double)) : NULL;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(min : DT2T) \
    reduction(| : mstop)
#endif
  for (int64_t g = 0; g < ngh + ngq; ++g) {
    double dtl = DT2T;
    if (g < ngh) {
      int64_t e0 = g * G, e1 = (e0 + G < m->ne_hex) ? e0 + G : m->ne_hex;
      int ms = 0;
      if (m->ehou_h)
        for (int64_t e = e0; e < e1; ++e) m->ehou_h[e] = ZERO;
      or_hex_forces_dt(e0, e1, m->ne_hex, m->hconn, m->x, m->v, m->hsig, m->heint, m->hrho, m->hvol0, m->hpla,
                       m->hepsd, m->hoffg, m->hqvis, m->ehou_h, m->hiads, m->fsky, m->hp, DT1, &dtl, &ms,
                       m->ecrit ? m->dtel_h : NULL);
      mstop |= ms;
      if (m->ehou_h) { /* EHOURT = sum of the group's EHOU (shvis3.F:425-428) */
        double t = ZERO;
        for (int64_t e = e0; e < e1; ++e) t = t + m->ehou_h[e];
        gsum[g] = t;
      }
    } else {
      int64_t e0 = (g - ngh) * G, e1 = (e0 + G < m->ne_q4) ? e0 + G : m->ne_q4;
      or_q4_forces_dte(e0, e1, m->qconn, m->x, m->v, m->vr, &m->q, m->qiadc, m->fsky, m->qp, DT1, &dtl,
                       m->ecrit ? m->dtel_q : NULL, m->ehou_q);
      if (m->ehou_q) { /* chvis3.F:349-351 */
        double t = ZERO;
        for (int64_t e = e0; e < e1; ++e) t = t + m->ehou_q[e];
        gsum[g] = t;
      }
    }
    if (dtl < DT2T) DT2T = dtl;
  }
  (void)nthreads;
  if (gsum) { /* EHOUR = EHOUR + EHOURT, group after group (shvis3.F:464, chvis3.F:397) */
    for (int64_t g = 0; g < ngh + ngq; ++g) m->ehour = m->ehour + gsum[g];
    free(gsum);
  }
  if (m->ecrit && (m->dtel_h || m->dtel_q)) {
    /* NELTS / ITYPTS of the cycle, element after element in group order with DT2T = EP06 on
       entry (resol.F:2826): solids replace it when DTX <= DT2T (mqviscb.F:622-630), shells when
       DT < DT2T (cdt3.F:205-215) */
    double t = EP06;
    m->itypts = 0;
    m->nelts = -1;
    for (int64_t e = 0; m->dtel_h && e < m->ne_hex; ++e) {
      const double d = m->dtel_h[e];
      if (d < ZERO || d > t) continue;
      t = d; m->itypts = 1; m->nelts = e;
    }
    for (int64_t e = 0; m->dtel_q && e < m->ne_q4; ++e) {
      const double d = m->dtel_q[e];
      if (d >= ZERO && d < t) { t = d; m->itypts = 3; m->nelts = e; }
    }
  }
  or_nodes_model(m, ts, DT2T, mstop);
}

/* The nodal half of a cycle whose element half has filled FSKY with DT1 = ts->dt2:
 * ASSPAR4, the DT2 pipeline with the element minimum dt2t, ACCELE, BCS10, FIXVEL,
 * (ECRIT), VELOCITY, DEPLA and the time advance. */
void or_nodes_model(or_model_c *m, or_time_state *ts, double dt2t, int mstop) {
  const int64_t nn = m->nn;
  const double DT2T = dt2t;
  ts->dt1 = ts->dt2;
  ts->dt2 = EP06;
  ts->mstop |= mstop;
  /* ASSPAR4 (asspar4.F:164-180) */
  or_asspar4(nn, m->adsky, m->fsky, m->a, m->ar, m->stifn, m->stifr);
  /* DT2 pipeline (resol.F:6125-6129, 6352, 6494-6497) */
  if (DT2T < ts->dt2) ts->dt2 = DT2T;
  {
    double d = ts->dt2;
    double g = ((double)1.1f) * ts->dt2old;
    d = dmin(d, g);
    d = dmin(d, m->dtmx);
    ts->dt2 = d;
  }
  ts->dt2old = ts->dt2;
  ts->dt12 = HALF * (ts->dt1 + ts->dt2);
  if (ts->dt2 <= ZERO) ts->mstop |= 2; /* time step <= 0 stops the run (resol.F:8575, 9599) */
  const double DT12 = ts->dt12, DT2 = ts->dt2;
  /* ECRIT output point (sortie_main.F:368 from resol.F:8523, after ACCELE/BCS10/FIXVEL and
   * before VELOCITY): DT05 = HALF*DT1 (ecrit.F:193) */
  const double DT05 = HALF * ts->dt1;
  double encin = ZERO, enrot = ZERO, xmom = ZERO, ymom = ZERO, zmom = ZERO;
  for (int64_t n = 0; n < nn; ++n) {
    double *a = m->a + 3 * n, *v = m->v + 3 * n;
    /* ACCELE (accele.F:65-80, 97-109) */
    if (m->ms[n] > ZERO) {
      double r = ONE / m->ms[n];
      a[0] = a[0] * r; a[1] = a[1] * r; a[2] = a[2] * r;
    } else {
      a[0] = ZERO; a[1] = ZERO; a[2] = ZERO;
    }
    double *ar = m->ar + 3 * n, *vr = m->vr + 3 * n;
    if (m->in) {
      if (m->in[n] > ZERO) {
        double r = ONE / m->in[n];
        ar[0] = ar[0] * r; ar[1] = ar[1] * r; ar[2] = ar[2] * r;
      } else {
        ar[0] = ZERO; ar[1] = ZERO; ar[2] = ZERO;
      }
    }
    /* BCS10 / BCS1V global frame: translations bits 4,2,1; rotations bits 32,16,8 */
    const int32_t code = m->code ? m->code[n] : 0;
    if (code) {
      if (code & 4) { a[0] = ZERO; v[0] = ZERO; }
      if (code & 2) { a[1] = ZERO; v[1] = ZERO; }
      if (code & 1) { a[2] = ZERO; v[2] = ZERO; }
      if (code & 32) { ar[0] = ZERO; vr[0] = ZERO; }
      if (code & 16) { ar[1] = ZERO; vr[1] = ZERO; }
      if (code & 8) { ar[2] = ZERO; vr[2] = ZERO; }
      /* FIXVEL global frame (fixvel.F:366-379): bits 256,128,64; each imposed DOF adds its work
         DT1*DW to WFEXT now and DT2*DW at the next cycle (fixvel.F:376-378, 834, 837),
         DW = FOURTH*MS*AXI*WEIGHT*(A*DT12 + TWO*V)*(A - AOLD)
         with AXI = ONE - AX + AX*X(2,I) = ONE (AX = 0 in 3D, fixvel.F:123-127) and WEIGHT = 1 */
      static const int bit[3] = {256, 128, 64};
      for (int j = 0; j < 3; ++j) {
        if (!(code & bit[j])) continue;
        /* the DOF's DT2*DW of the previous cycle, kept in VEL(4,N), enters WFEXT first
           (fixvel.F:342-344 with RW_SMS = ONE without mass scaling, fixvel.F:168-169) */
        m->wfext = m->wfext + m->wpend[3 * n + j];
        m->wpend[3 * n + j] = ZERO;
        const double aold = a[j];
        a[j] = (m->vimp[3 * n + j] - v[j]) / DT12;
        const double DW = FOURTH * m->ms[n] * ONE * ONE * (a[j] * DT12 + TWO * v[j]) * (a[j] - aold);
        m->wfext = m->wfext + ts->dt1 * DW;                       /* fixvel.F:834 */
        m->wpend[3 * n + j] = m->wpend[3 * n + j] + DT2 * DW;    /* fixvel.F:837 */
      }
    }
    if (m->ecrit) {
      /* ecrit.F:221-225 (MAS = MS*WEIGHT_MD, WEIGHT_MD = 1) and 319-324 */
      const double vx = v[0] + DT05 * a[0], vy = v[1] + DT05 * a[1], vz = v[2] + DT05 * a[2];
      encin = encin + (vx * vx + vy * vy + vz * vz) * HALF * m->ms[n];
      xmom = xmom + vx * m->ms[n];
      ymom = ymom + vy * m->ms[n];
      zmom = zmom + vz * m->ms[n];
      if (m->in) {
        const double wx = vr[0] + DT05 * ar[0], wy = vr[1] + DT05 * ar[1], wz = vr[2] + DT05 * ar[2];
        enrot = enrot + (wx * wx + wy * wy + wz * wz) * HALF * m->in[n];
      }
      if (m->th) { /* THNOD's D, V, A, VR, AR, X at HIST2's point (thnod.F:59-97) */
        double *t = m->th + 18 * n;
        const double *xx = m->x + 3 * n, *dd = m->d + 3 * n;
        for (int c = 0; c < 3; ++c) {
          t[c] = dd[c]; t[3 + c] = v[c]; t[6 + c] = a[c];
          t[9 + c] = m->in ? vr[c] : ZERO; t[12 + c] = m->in ? ar[c] : ZERO; t[15 + c] = xx[c];
        }
      }
    }
    /* VELOCITY (velocity.F:58-86) */
    v[0] = v[0] + DT12 * a[0]; v[1] = v[1] + DT12 * a[1]; v[2] = v[2] + DT12 * a[2];
    a[0] = ZERO; a[1] = ZERO; a[2] = ZERO;
    if (m->in) {
      vr[0] = vr[0] + DT12 * ar[0]; vr[1] = vr[1] + DT12 * ar[1]; vr[2] = vr[2] + DT12 * ar[2];
      ar[0] = ZERO; ar[1] = ZERO; ar[2] = ZERO;
    }
    /* DEPLA (displacement.F:108-118), IRESP=0 */
    double *x = m->x + 3 * n, *d = m->d + 3 * n;
    for (int c = 0; c < 3; ++c) {
      double vdt = DT2 * v[c];
      d[c] = d[c] + vdt;
      x[c] = x[c] + vdt;
    }
  }
  if (m->ecrit) {
    /* internal energy: SBILAN EI = EINT*VOL (sbilan.F:140), CBILAN EI = EINT(1)+EINT(2)
     * (cbilan.F:168), summed in element order into the part's PARTSAV(1) (solids part 1,
     * shells part 2), the parts added in part order by ECRIT (ecrit.F:344) */
    double ie_h = ZERO, ie_q = ZERO;
    for (int64_t e = 0; e < m->ne_hex; ++e) ie_h = ie_h + m->heint[e] * m->hvol0[e];
    for (int64_t e = 0; e < m->ne_q4; ++e) ie_q = ie_q + (m->q.eint[e] + m->q.eint[m->ne_q4 + e]);
    m->ecrit[0] = ie_h + ie_q;
    m->ecrit[1] = encin;
    m->ecrit[2] = enrot;
    m->ecrit[3] = m->wfext;
    m->ecrit[4] = xmom;
    m->ecrit[5] = ymom;
    m->ecrit[6] = zmom;
  }
  ts->tt = ts->tt + ts->dt2;
  ts->ncycle += 1;
}

/* Energies of a mixed model: out[0] internal (hex EINT*VOL0 + shell EINT(1)+EINT(2)),
 * out[1] translational kinetic 0.5*sum MS|V|^2, out[2] rotational 0.5*sum IN|VR|^2. */
void or_energy_model(const or_model_c *m, double out[3]) {
  /* internal energy per part as SBILAN / CBILAN accumulate PARTSAV(1,part) (sbilan.F:140,
     cbilan.F:168), element after element; ECRIT adds the parts in part order (ecrit.F:344):
     the solids are part 1, the shells part 2 */
  double ie_h = 0.0, ie_q = 0.0, ke = 0.0, kr = 0.0;
  for (int64_t e = 0; e < m->ne_hex; ++e) ie_h += m->heint[e] * m->hvol0[e];
  for (int64_t e = 0; e < m->ne_q4; ++e) ie_q += m->q.eint[e] + m->q.eint[m->ne_q4 + e];
  const double ie = ie_h + ie_q;
  for (int64_t n = 0; n < m->nn; ++n) {
    const double *v = m->v + 3 * n;
    ke += m->ms[n] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (m->in) {
      const double *w = m->vr + 3 * n;
      kr += m->in[n] * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    }
  }
  out[0] = ie;
  out[1] = HALF * ke;
  out[2] = HALF * kr;
}
