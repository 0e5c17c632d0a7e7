// or_crmath.cpp -- correctly rounded pow / log / exp for the oracle's OR_CR_MATH build.
//
// TEST INFRASTRUCTURE ONLY (part of the oracle).  Host build of the
// double-double routines in openradioss_amd/csrc/or_math.h, so that the oracle
// can be run under the same elementary functions as the FP64 device path.
// Their rounding is checked here against 80-bit long double (64-bit
// significand), independently of the routines themselves.
#include "../openradioss_amd/csrc/or_math.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

extern "C" {
double or_cr_pow(double x, double y) { return ormath::or_pow(x, y); }
double or_cr_log(double x) { return ormath::or_log(x); }
double or_cr_exp(double x) { return ormath::or_exp(x); }

// Counting build of the oracle (OR_COUNT_MATH): the host library's value is returned, and each
// call site counts its calls and those whose correctly rounded value differs.
static int64_t g_site[16][2];
static inline double count_site(int site, double host, double cr) {
  __atomic_fetch_add(&g_site[site][0], 1, __ATOMIC_RELAXED);
  if (memcmp(&host, &cr, sizeof(double)) != 0) __atomic_fetch_add(&g_site[site][1], 1, __ATOMIC_RELAXED);
  return host;
}
double or_count_pow(int site, double x, double y) { return count_site(site, pow(x, y), or_cr_pow(x, y)); }
double or_count_log(int site, double x) { return count_site(site, log(x), or_cr_log(x)); }
double or_count_exp(int site, double x) { return count_site(site, exp(x), or_cr_exp(x)); }
void or_math_site_counts(int64_t *out) {
  for (int s = 0; s < 16; ++s) {
    out[2 * s] = __atomic_load_n(&g_site[s][0], __ATOMIC_RELAXED);
    out[2 * s + 1] = __atomic_load_n(&g_site[s][1], __ATOMIC_RELAXED);
  }
}
void or_math_site_reset(void) { memset(g_site, 0, sizeof(g_site)); }

// Vector form: out[i] = log(x[i]) (fn 0), exp(x[i]) (fn 1) or pow(x[i], y[i]) (fn 2).
void or_cr_eval(int fn, int64_t n, const double *x, const double *y, double *out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = fn == 0 ? or_cr_log(x[i]) : fn == 1 ? or_cr_exp(x[i]) : or_cr_pow(x[i], y[i]);
}
}

namespace {

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double unif(uint64_t &s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

// One argument of function fn drawn over the ranges the path uses:
// fn 0 log: x = e^u, u in [-40, 25] (plastic strain, rate ratios, volumes)
// fn 1 exp: z in [-60, 12] (n*log(pla) and (n-1)*log(pla) of the hardening law)
// fn 2 pow: x = e^u, u in [-30, 14]; y one of 1/3, 2/3 (bulk viscosity, sound
//           speed terms) or uniform in [-1, 1] (hardening exponent n and n-1)
void draw(int fn, uint64_t &s, double &x, double &y) {
  if (fn == 0) { x = exp(-40.0 + 65.0 * unif(s)); y = 0.0; }
  else if (fn == 1) { x = -60.0 + 72.0 * unif(s); y = 0.0; }
  else {
    x = exp(-30.0 + 44.0 * unif(s));
    const uint64_t k = splitmix(s) & 3;
    y = k == 0 ? 1.0 / 3.0 : k == 1 ? 2.0 / 3.0 : -1.0 + 2.0 * unif(s);
  }
}

}  // namespace

extern "C" {
// Rounding check of or_cr_* on n random arguments of function fn (0 log, 1 exp,
// 2 pow).  counts[0] = cases where or_cr_* differs from the host library,
// counts[1] = cases where the host library is strictly closer to the long-double
// value (a rounding error of or_cr_*), counts[2] = cases where or_cr_* is strictly
// closer, counts[3] = results more than 1 ulp from the long-double value.
void or_cr_selftest(int fn, int64_t n, uint64_t seed, int64_t *counts) {
  uint64_t s = seed;
  for (int i = 0; i < 4; ++i) counts[i] = 0;
  for (int64_t k = 0; k < n; ++k) {
    double x, y;
    draw(fn, s, x, y);
    double ours, host;
    long double ref;
    if (fn == 0) { ours = or_cr_log(x); host = log(x); ref = logl((long double)x); }
    else if (fn == 1) { ours = or_cr_exp(x); host = exp(x); ref = expl((long double)x); }
    else { ours = or_cr_pow(x, y); host = pow(x, y); ref = powl((long double)x, (long double)y); }
    const long double ulp = (long double)nextafter(fabs(ours), INFINITY) - (long double)fabs(ours);
    if (fabsl((long double)ours - ref) > ulp) counts[3]++;
    if (ours == host) continue;
    counts[0]++;
    const long double d_ours = fabsl((long double)ours - ref), d_host = fabsl((long double)host - ref);
    if (d_host < d_ours) counts[1]++;
    else if (d_ours < d_host) counts[2]++;
  }
}
}

extern "C" {
// Consistency of the shared-logarithm form used by the device path: or_pow2(x, y1, y2)
// must give exactly or_pow(x, y1) and or_pow(x, y2).  Draws n bases over the pow range
// above plus the special bases (0, 1, subnormal, huge, negative, infinite, NaN) and the
// exponent pairs of the path (1/3 with 2/3; n with n-1 or 1-n).  Returns the number of
// mismatching bits patterns.
int64_t or_cr_pow2_check(int64_t n, uint64_t seed) {
  const double special[] = {0.0, 1.0, 0x1p-1074, 0x1p-1030, 1.0e300, -2.0, INFINITY, NAN};
  uint64_t s = seed;
  int64_t bad = 0;
  for (int64_t k = 0; k < n; ++k) {
    double x = k < 8 ? special[k] : exp(-30.0 + 44.0 * unif(s));
    const double cn = unif(s) * 1.5;
    double y1, y2;
    if (k & 1) { y1 = 1.0 / 3.0; y2 = 2.0 / 3.0; }
    else { y1 = cn; y2 = cn > 1.0 ? cn - 1.0 : 1.0 - cn; }
    double r1, r2;
    ormath::or_pow2(x, y1, y2, r1, r2);
    const double p1 = ormath::or_pow(x, y1), p2 = ormath::or_pow(x, y2);
    uint64_t b[4];
    memcpy(&b[0], &r1, 8); memcpy(&b[1], &p1, 8); memcpy(&b[2], &r2, 8); memcpy(&b[3], &p2, 8);
    if (b[0] != b[1] || b[2] != b[3]) bad++;
  }
  return bad;
}
}
