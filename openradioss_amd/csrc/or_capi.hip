// or_capi.hip -- kernels and C ABI of the MI355X explicit hex path (gfx950).
//
// Per cycle, on one stream, with every scalar kept on the device:
//   k_hex_forces : SFORC3 chain per element -> skyline rows + block min of DTX
//                  folded into a device word with an integer atomicMin (min is exact,
//                  so the result does not depend on block order)
//   k_node_update: ASSPAR4 fixed-order segmented sum per node, DT2 pipeline,
//                  ACCELE, BCS10, FIXVEL, VELOCITY, DEPLA; thread 0 advances the
//                  cycle scalars (double-buffered by cycle parity)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/or_gpu.h"
#include "or_hex.h"

using namespace orgpu;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define OR_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(OR_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
  } while (0)

constexpr int kBlock = 256;

struct alignas(32) RowD { double fx, fy, fz, sti; };
struct alignas(16) RowF { float fx, fy, fz, sti; };

template <typename S> struct HexPtrs {
  const int32_t *conn;  // [8][ne]
  const int32_t *iads;  // [8][ne]
  S *sig;               // [6][ne]
  S *eint, *rho, *pla, *epsd, *offg, *qvis;
  const S *vol0;
  int64_t ne;
};

struct NodePtrs {
  double *xv;          // [nn][6] x,y,z,vx,vy,vz
  double *d;           // [nn][3]
  const double *ms;    // [nn]
  const int32_t *code; // [nn] bits 0-2 BCS (z,y,x), bits 3-5 imposed velocity (z,y,x)
  const double *vimp;  // [nn][3] or null
  double *stifn;       // [nn] or null
  const int32_t *adsky; // [nn+1]
  int64_t nn;
};

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

template <typename T> __device__ __forceinline__ T to_r(double v) { return T(v); }

template <typename Row, typename R> __device__ __forceinline__ void store_row(Row *p, R fx, R fy, R fz, R sti);
template <> __device__ __forceinline__ void store_row<RowD, double>(RowD *p, double fx, double fy, double fz, double sti) {
  double2 *q = reinterpret_cast<double2 *>(p);
  q[0] = make_double2(fx, fy);
  q[1] = make_double2(fz, sti);
}
template <> __device__ __forceinline__ void store_row<RowF, float>(RowF *p, float fx, float fy, float fz, float sti) {
  *reinterpret_cast<float4 *>(p) = make_float4(fx, fy, fz, sti);
}

__device__ __forceinline__ double block_min(double v) {
  __shared__ double red[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) v = (red[w] < v) ? red[w] : v;
  }
  return v;
}

// ---------------------------------------------------------------- element kernel
template <typename R, typename S, typename Row>
__global__ __launch_bounds__(kBlock) void k_hex_forces(HexPtrs<S> e, const double *__restrict__ xv,
                                                       Row *__restrict__ sky, HexConsts c, const DevTime *ts,
                                                       int parity, unsigned long long *dtmin, int *flags,
                                                       S *__restrict__ ehou) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double dtx = kEP06;
  int mstop = 0;
  if (i < e.ne) {
    const R DT1 = R(ts[parity].dt2);
    double xd[8], yd[8], zd[8];
    R vx[8], vy[8], vz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t n = e.conn[(int64_t)k * e.ne + i];
      const double2 *p = reinterpret_cast<const double2 *>(xv + 6 * n);
      double2 a = p[0], b = p[1], q = p[2];
      xd[k] = a.x; yd[k] = a.y; zd[k] = b.x;
      vx[k] = R(b.y); vy[k] = R(q.x); vz[k] = R(q.y);
    }
    HexState<S> st;
#pragma unroll
    for (int k = 0; k < 6; ++k) st.sig[k] = e.sig[(int64_t)k * e.ne + i];
    st.eint = e.eint[i]; st.rho = e.rho[i]; st.vol0 = e.vol0[i]; st.pla = e.pla[i];
    st.epsd = e.epsd[i]; st.offg = e.offg[i]; st.qvis = e.qvis[i];
    HexOut<R> o;
    hex_element<R, S>(xd, yd, zd, vx, vy, vz, st, c, DT1, o);
#pragma unroll
    for (int k = 0; k < 6; ++k) e.sig[(int64_t)k * e.ne + i] = st.sig[k];
    e.eint[i] = st.eint; e.rho[i] = st.rho; e.pla[i] = st.pla;
    e.epsd[i] = st.epsd; e.offg[i] = st.offg; e.qvis[i] = st.qvis;
    if (ehou) ehou[i] = S(R(ehou[i]) + o.ehou);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t K = e.iads[(int64_t)k * e.ne + i];
      store_row<Row, R>(sky + K, o.f1[k], o.f2[k], o.f3[k], o.sti4);
    }
    dtx = o.dtx;
    mstop = o.mstop;
  }
  if (__any(mstop) && mstop) atomicOr(flags, 1);
  const double m = block_min(dtx);
  if (threadIdx.x == 0) atomicMin(dtmin + parity, dbits(m));
}

// ---------------------------------------------------------------- node kernel
__device__ __forceinline__ void accumulate_row(const RowD *r, double &ax, double &ay, double &az, double &sn) {
  const double2 *q = reinterpret_cast<const double2 *>(r);
  double2 a = q[0], b = q[1];
  ax = ax + a.x; ay = ay + a.y; az = az + b.x; sn = sn + b.y;
}
__device__ __forceinline__ void accumulate_row(const RowF *r, double &ax, double &ay, double &az, double &sn) {
  float4 a = *reinterpret_cast<const float4 *>(r);
  ax = ax + double(a.x); ay = ay + double(a.y); az = az + double(a.z); sn = sn + double(a.w);
}

template <typename Row>
__global__ __launch_bounds__(kBlock) void k_node_update(NodePtrs nd, const Row *__restrict__ sky, DevTime *ts,
                                                        int parity, unsigned long long *dtmin, const int *flags,
                                                        double dt_growth, double dtmx) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  // DT2 pipeline (resol.F:2721-2722, 2826, 6125-6129, 6352, 6494-6497)
  const DevTime prev = ts[parity];
  const double DT1 = prev.dt2;
  const double dt2t = bitsd(dtmin[parity]);
  double DT2 = kEP06;
  if (dt2t < DT2) DT2 = dt2t;
  {
    double g = dt_growth * prev.dt2old;
    DT2 = (g < DT2) ? g : DT2;
    DT2 = (dtmx < DT2) ? dtmx : DT2;
  }
  const double DT12 = (1.0 / 2.0) * (DT1 + DT2);
  if (n == 0) {
    DevTime nx;
    nx.tt = prev.tt + DT2;
    nx.dt1 = DT1;
    nx.dt2 = DT2;
    nx.dt12 = DT12;
    nx.dt2old = DT2;
    nx.ncycle = prev.ncycle + 1;
    nx.mstop = prev.mstop | *flags;
    nx.pad = 0;
    ts[parity ^ 1] = nx;
    dtmin[parity ^ 1] = dbits(kEP06);
  }
  if (n >= nd.nn) return;
  // ASSPAR4: rows of this node in skyline order, starting from zero
  double ax = 0.0, ay = 0.0, az = 0.0, sn = 0.0;
  const int32_t k0 = nd.adsky[n], k1 = nd.adsky[n + 1];
  for (int32_t k = k0; k < k1; ++k) accumulate_row(sky + k, ax, ay, az, sn);
  if (nd.stifn) nd.stifn[n] = sn;
  // ACCELE (accele.F:65-80)
  const double msn = nd.ms[n];
  if (msn > 0.0) {
    const double rtmp = 1.0 / msn;
    ax = ax * rtmp; ay = ay * rtmp; az = az * rtmp;
  } else {
    ax = 0.0; ay = 0.0; az = 0.0;
  }
  double2 *p = reinterpret_cast<double2 *>(nd.xv + 6 * n);
  double2 pa = p[0], pb = p[1], pc = p[2];
  double x = pa.x, y = pa.y, z = pb.x, vx = pb.y, vy = pc.x, vz = pc.y;
  const int32_t code = nd.code[n];
  if (code) {
    // BCS1V global frame (bcs1.F): LCOD bit 4 = x, 2 = y, 1 = z
    if (code & 4) { ax = 0.0; vx = 0.0; }
    if (code & 2) { ay = 0.0; vy = 0.0; }
    if (code & 1) { az = 0.0; vz = 0.0; }
    // FIXVEL global frame (fixvel.F:366-379): A = (VIMP - V)/DT12
    if (code & (4 << 3)) ax = (nd.vimp[3 * n + 0] - vx) / DT12;
    if (code & (2 << 3)) ay = (nd.vimp[3 * n + 1] - vy) / DT12;
    if (code & (1 << 3)) az = (nd.vimp[3 * n + 2] - vz) / DT12;
  }
  // VELOCITY (velocity.F:58-66)
  vx = vx + DT12 * ax; vy = vy + DT12 * ay; vz = vz + DT12 * az;
  // DEPLA (displacement.F, IRESP=0 form in FP64)
  double *dd = nd.d + 3 * n;
  double vdt;
  vdt = DT2 * vx; dd[0] = dd[0] + vdt; x = x + vdt;
  vdt = DT2 * vy; dd[1] = dd[1] + vdt; y = y + vdt;
  vdt = DT2 * vz; dd[2] = dd[2] + vdt; z = z + vdt;
  p[0] = make_double2(x, y);
  p[1] = make_double2(z, vx);
  p[2] = make_double2(vy, vz);
}

// ---------------------------------------------------------------- energies (fixed-order)
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_energy_partial(const S *eint, const S *vol0, int64_t ne, const double *xv,
                                                           const double *ms, int64_t nn, int64_t per_block,
                                                           double *partial) {
  double ie = 0.0, ke = 0.0;
  const int64_t b0 = (int64_t)blockIdx.x * per_block;
  for (int64_t j = b0 + threadIdx.x; j < b0 + per_block; j += kBlock) {
    if (j < ne) ie = ie + double(eint[j]) * double(vol0[j]);
    if (j < nn) {
      const double *p = xv + 6 * j;
      ke = ke + ms[j] * (p[3] * p[3] + p[4] * p[4] + p[5] * p[5]);
    }
  }
  ie = block_sum(ie);
  ke = block_sum(ke);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x + 0] = ie;
    partial[2 * blockIdx.x + 1] = ke;
  }
}

__global__ __launch_bounds__(kBlock) void k_energy_final(const double *partial, int nb, double *out) {
  double ie = 0.0, ke = 0.0;
  for (int j = threadIdx.x; j < nb; j += kBlock) {
    ie = ie + partial[2 * j];
    ke = ke + partial[2 * j + 1];
  }
  ie = block_sum(ie);
  ke = block_sum(ke);
  if (threadIdx.x == 0) {
    out[0] = ie;
    out[1] = 0.5 * ke;
  }
}

// ---------------------------------------------------------------- group drop-in kernel
// SFORC3 on one group with the reference host layouts (X(3,*), V(3,*), IXS(11,*),
// SIG(NEL,6), FSKY(8,*), IADS(8,*) 1-based).
__global__ __launch_bounds__(kBlock) void k_sforc3_group(int nel, int64_t nft, const double *x, const double *v,
                                                         const int32_t *ixs, double *sig, double *eint, double *rho,
                                                         const double *vol, double *pla, double *epsd,
                                                         double *off, double *qvis, double *fsky,
                                                         const int32_t *iads, HexConsts c, double dt1,
                                                         double *dtx_out, int *mstop_out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nel) return;
  const int64_t col = nft + i; // 0-based column of IXS / IADS
  double xd[8], yd[8], zd[8], vx[8], vy[8], vz[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int64_t n = (int64_t)ixs[11 * col + 1 + k] - 1;
    xd[k] = x[3 * n]; yd[k] = x[3 * n + 1]; zd[k] = x[3 * n + 2];
    vx[k] = v[3 * n]; vy[k] = v[3 * n + 1]; vz[k] = v[3 * n + 2];
  }
  HexState<double> st;
#pragma unroll
  for (int k = 0; k < 6; ++k) st.sig[k] = sig[(int64_t)k * nel + i];
  st.eint = eint[i]; st.rho = rho[i]; st.vol0 = vol[i]; st.pla = pla[i];
  st.epsd = epsd[i]; st.offg = off[i]; st.qvis = qvis[i];
  HexOut<double> o;
  hex_element<double, double>(xd, yd, zd, vx, vy, vz, st, c, dt1, o);
#pragma unroll
  for (int k = 0; k < 6; ++k) sig[(int64_t)k * nel + i] = st.sig[k];
  eint[i] = st.eint; rho[i] = st.rho; pla[i] = st.pla;
  epsd[i] = st.epsd; off[i] = st.offg; qvis[i] = st.qvis;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int64_t K = (int64_t)iads[8 * col + k] - 1;
    fsky[8 * K + 0] = o.f1[k];
    fsky[8 * K + 1] = o.f2[k];
    fsky[8 * K + 2] = o.f3[k];
    fsky[8 * K + 6] = o.sti4;
  }
  dtx_out[i] = o.dtx;
  mstop_out[i] = o.mstop;
}

// ---------------------------------------------------------------- host helpers
double host_onep333() {
  volatile double zep3 = 3.0 / 10.0, zep03 = 3.0 / 100.0, threeem3 = 3.0 / 1000.0;
  volatile double onep3 = 1.0 + zep3;
  volatile double onep33 = onep3 + zep03;
  return onep33 + threeem3;
}

HexConsts make_consts(const or_hex_params &p) {
  HexConsts c;
  c.rho0 = p.rho0; c.bulk = p.bulk; c.shear = p.shear; c.ca = p.ca; c.cb = p.cb; c.cn = p.cn;
  c.epmx = p.epmx; c.sigmx = p.sigmx; c.cc = p.cc; c.epdr = p.epdr; c.fisokin = p.fisokin;
  c.asrate = p.asrate; c.m_exp = p.m_exp; c.icc = p.icc; c.vp = p.vp; c.israte = p.israte;
  c.iform = p.iform; c.qa = p.qa; c.qb = p.qb; c.qh = p.qh; c.vns1 = p.vns1; c.vns2 = p.vns2;
  c.dtfac = p.dtfac; c.dtmx = p.dtmx;
  c.onep333 = host_onep333();
  c.dt_growth = (double)1.1f;
  return c;
}

int check_params(const or_hex_params &p) {
  if (p.fisokin != 0.0) return fail(OR_ERR_ARG, "law 2 kinematic hardening (FISOKIN>0) is not on this path");
  if (p.iform != 0) return fail(OR_ERR_ARG, "law 2 Zerilli form (IFORM=1) is not on this path");
  if (!(p.rho0 > 0.0) || !(p.dtfac > 0.0)) return fail(OR_ERR_ARG, "rho0 and dtfac must be positive");
  return OR_OK;
}

template <typename T> int dalloc(T **p, int64_t n, int64_t &acc) {
  if (n <= 0) { *p = nullptr; return OR_OK; }
  OR_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)n));
  acc += (int64_t)(sizeof(T) * (size_t)n);
  return OR_OK;
}

} // namespace

// ============================================================== model object
struct or_model {
  int device = 0;
  int prec = OR_PREC_FP64;
  int64_t nn = 0, ne = 0, lsky = 0;
  int64_t bytes = 0;
  hipStream_t stream = nullptr;
  bool have_params = false, have_nodes = false, have_hex = false, have_sky = false;
  HexConsts consts{};
  // nodes
  double *xv = nullptr, *d = nullptr, *ms = nullptr, *vimp = nullptr, *stifn = nullptr;
  int32_t *code = nullptr, *adsky = nullptr;
  // hex (storage type depends on precision)
  int32_t *conn = nullptr, *iads = nullptr;
  void *sig = nullptr, *eint = nullptr, *rho = nullptr, *vol0 = nullptr, *pla = nullptr, *epsd = nullptr,
       *offg = nullptr, *qvis = nullptr;
  void *sky = nullptr;
  // scalars
  DevTime *ts = nullptr;               // [2]
  unsigned long long *dtmin = nullptr; // [2]
  int *flags = nullptr;
  double *energy = nullptr;            // [2 + 2*nb]
  int64_t parity = 0;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;          // ring of 3 events per cycle
  int ring_used = 0;
  double t_elem = 0.0, t_node = 0.0;
  int64_t t_cycles = 0;
};

static size_t state_size(const or_model *m) { return m->prec == OR_PREC_FP64 ? sizeof(double) : sizeof(float); }

// Entry points: C linkage comes from their declarations in or_gpu.h.

int or_version(void) { return 100; }

const char *or_last_error(void) { return g_err.c_str(); }

int or_device_count(int *count) {
  if (!count) return fail(OR_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) { *count = 0; return fail(OR_ERR_NODEV, hipGetErrorString(e)); }
  *count = n;
  return OR_OK;
}

int or_model_create(or_model **out, int device, int64_t nn, int64_t ne_hex, int precision) {
  if (!out) return fail(OR_ERR_ARG, "null out");
  *out = nullptr;
  if (nn <= 0 || ne_hex < 0) return fail(OR_ERR_ARG, "nn must be > 0 and ne_hex >= 0");
  if (precision != OR_PREC_FP64 && precision != OR_PREC_SPLIT) return fail(OR_ERR_ARG, "bad precision");
  if (8 * ne_hex >= (int64_t)1 << 31 || nn >= (int64_t)1 << 31)
    return fail(OR_ERR_ARG, "skyline rows and node ids must fit in int32");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(OR_ERR_NODEV, "no HIP device");
  if (device < 0 || device >= ndev) return fail(OR_ERR_ARG, "device index out of range");
  OR_HIP(hipSetDevice(device));
  or_model *m = new or_model();
  m->device = device;
  m->prec = precision;
  m->nn = nn;
  m->ne = ne_hex;
  m->lsky = 8 * ne_hex;
  int rc = OR_OK;
  auto A = [&](auto **p, int64_t n) { if (rc == OR_OK) rc = dalloc(p, n, m->bytes); };
  A(&m->xv, 6 * nn);
  A(&m->d, 3 * nn);
  A(&m->ms, nn);
  A(&m->vimp, 3 * nn);
  A(&m->code, nn);
  A(&m->adsky, nn + 1);
  A(&m->conn, 8 * ne_hex);
  A(&m->iads, 8 * ne_hex);
  const size_t ss = state_size(m);
  auto AS = [&](void **p, int64_t n) {
    if (rc != OR_OK || n <= 0) return;
    if (hipMalloc(p, ss * (size_t)n) != hipSuccess) { rc = fail(OR_ERR_HIP, "hipMalloc state"); return; }
    m->bytes += (int64_t)(ss * (size_t)n);
  };
  AS(&m->sig, 6 * ne_hex);
  AS(&m->eint, ne_hex); AS(&m->rho, ne_hex); AS(&m->vol0, ne_hex); AS(&m->pla, ne_hex);
  AS(&m->epsd, ne_hex); AS(&m->offg, ne_hex); AS(&m->qvis, ne_hex);
  const size_t row = (precision == OR_PREC_FP64) ? sizeof(RowD) : sizeof(RowF);
  if (rc == OR_OK && m->lsky > 0) {
    if (hipMalloc(&m->sky, row * (size_t)m->lsky) != hipSuccess) rc = fail(OR_ERR_HIP, "hipMalloc skyline");
    else m->bytes += (int64_t)(row * (size_t)m->lsky);
  }
  A(&m->ts, 2);
  A(&m->dtmin, 2);
  A(&m->flags, 1);
  const int64_t nb = 1024;
  A(&m->energy, 2 + 2 * nb);
  if (rc == OR_OK && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(OR_ERR_HIP, "hipStreamCreate");
  if (rc == OR_OK) {
    if (hipMemsetAsync(m->d, 0, sizeof(double) * 3 * nn, m->stream) != hipSuccess ||
        hipMemsetAsync(m->code, 0, sizeof(int32_t) * nn, m->stream) != hipSuccess ||
        hipMemsetAsync(m->vimp, 0, sizeof(double) * 3 * nn, m->stream) != hipSuccess ||
        hipMemsetAsync(m->flags, 0, sizeof(int), m->stream) != hipSuccess)
      rc = fail(OR_ERR_HIP, "hipMemsetAsync init");
    DevTime t0{};
    t0.dt2old = 1.0e6 / (1.0 + 1.0 / 10.0); // DTINI/ONEP1 (starter contrl.F:1125-1127)
    DevTime tsh[2] = {t0, t0};
    unsigned long long dm[2];
    double big = kEP06;
    std::memcpy(&dm[0], &big, 8);
    dm[1] = dm[0];
    if (rc == OR_OK && (hipMemcpy(m->ts, tsh, sizeof(tsh), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(m->dtmin, dm, sizeof(dm), hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(OR_ERR_HIP, "init scalars");
  }
  if (rc != OR_OK) {
    std::string keep = g_err;
    or_model_destroy(m);
    g_err = keep;
    return rc;
  }
  *out = m;
  return OR_OK;
}

int or_model_destroy(or_model *m) {
  if (!m) return OR_OK;
  hipSetDevice(m->device);
  if (m->stream) hipStreamSynchronize(m->stream);
  void *ptrs[] = {m->xv, m->d, m->ms, m->vimp, m->stifn, m->code, m->adsky, m->conn, m->iads, m->sig,
                  m->eint, m->rho, m->vol0, m->pla, m->epsd, m->offg, m->qvis, m->sky, m->ts, m->dtmin,
                  m->flags, m->energy};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (hipEvent_t e : m->ev) hipEventDestroy(e);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return OR_OK;
}

int or_model_set_hex_params(or_model *m, const or_hex_params *p) {
  if (!m || !p) return fail(OR_ERR_ARG, "null argument");
  int rc = check_params(*p);
  if (rc) return rc;
  m->consts = make_consts(*p);
  m->have_params = true;
  return OR_OK;
}

int or_model_upload_nodes(or_model *m, const double *x, const double *v, const double *ms, const int32_t *bcs,
                          const int32_t *impcode, const double *vimp) {
  if (!m || !x || !v || !ms) return fail(OR_ERR_ARG, "null argument");
  if (impcode && !vimp) return fail(OR_ERR_ARG, "impcode given without vimp");
  OR_HIP(hipSetDevice(m->device));
  const int64_t nn = m->nn;
  std::vector<double> xv((size_t)(6 * nn));
  std::vector<int32_t> code((size_t)nn, 0);
  for (int64_t n = 0; n < nn; ++n) {
    xv[6 * n + 0] = x[3 * n]; xv[6 * n + 1] = x[3 * n + 1]; xv[6 * n + 2] = x[3 * n + 2];
    xv[6 * n + 3] = v[3 * n]; xv[6 * n + 4] = v[3 * n + 1]; xv[6 * n + 5] = v[3 * n + 2];
    int32_t cd = 0;
    if (bcs) {
      if (bcs[n] < 0 || bcs[n] > 7) return fail(OR_ERR_ARG, "BCS code must be in 0..7");
      cd |= bcs[n];
    }
    if (impcode) {
      if (impcode[n] < 0 || impcode[n] > 7) return fail(OR_ERR_ARG, "impvel code must be in 0..7");
      cd |= impcode[n] << 3;
    }
    code[n] = cd;
  }
  OR_HIP(hipMemcpyAsync(m->xv, xv.data(), sizeof(double) * 6 * nn, hipMemcpyHostToDevice, m->stream));
  OR_HIP(hipMemcpyAsync(m->ms, ms, sizeof(double) * nn, hipMemcpyHostToDevice, m->stream));
  OR_HIP(hipMemcpyAsync(m->code, code.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, m->stream));
  if (vimp) OR_HIP(hipMemcpyAsync(m->vimp, vimp, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, m->stream));
  OR_HIP(hipMemsetAsync(m->d, 0, sizeof(double) * 3 * nn, m->stream));
  OR_HIP(hipStreamSynchronize(m->stream));
  m->have_nodes = true;
  return OR_OK;
}

template <typename S> static int upload_state(or_model *m, void *dst, const double *src, int64_t n) {
  if (!src) return fail(OR_ERR_ARG, "null state array");
  if (sizeof(S) == sizeof(double)) {
    OR_HIP(hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyHostToDevice));
  } else {
    std::vector<float> t((size_t)n);
    for (int64_t i = 0; i < n; ++i) t[i] = (float)src[i];
    OR_HIP(hipMemcpy(dst, t.data(), sizeof(float) * n, hipMemcpyHostToDevice));
  }
  return OR_OK;
}

template <typename S> static int download_state(or_model *m, double *dst, const void *src, int64_t n) {
  if (!dst) return OR_OK;
  if (sizeof(S) == sizeof(double)) {
    OR_HIP(hipMemcpy(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> t((size_t)n);
    OR_HIP(hipMemcpy(t.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) dst[i] = (double)t[i];
  }
  return OR_OK;
}

int or_model_upload_hex(or_model *m, const int32_t *conn, const double *sig, const double *eint, const double *rho,
                        const double *vol0, const double *pla, const double *epsd, const double *offg,
                        const double *qvis) {
  if (!m || !conn) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  const int64_t ne = m->ne;
  if (ne == 0) { m->have_hex = true; return OR_OK; }
  std::vector<int32_t> soa((size_t)(8 * ne));
  for (int64_t e = 0; e < ne; ++e)
    for (int k = 0; k < 8; ++k) {
      const int32_t n = conn[8 * e + k];
      if (n < 0 || n >= m->nn) return fail(OR_ERR_ARG, "connectivity node id out of range");
      soa[(size_t)(k * ne + e)] = n;
    }
  OR_HIP(hipMemcpy(m->conn, soa.data(), sizeof(int32_t) * 8 * ne, hipMemcpyHostToDevice));
  int rc;
  if (m->prec == OR_PREC_FP64) {
    if ((rc = upload_state<double>(m, m->sig, sig, 6 * ne))) return rc;
    if ((rc = upload_state<double>(m, m->eint, eint, ne))) return rc;
    if ((rc = upload_state<double>(m, m->rho, rho, ne))) return rc;
    if ((rc = upload_state<double>(m, m->vol0, vol0, ne))) return rc;
    if ((rc = upload_state<double>(m, m->pla, pla, ne))) return rc;
    if ((rc = upload_state<double>(m, m->epsd, epsd, ne))) return rc;
    if ((rc = upload_state<double>(m, m->offg, offg, ne))) return rc;
    if ((rc = upload_state<double>(m, m->qvis, qvis, ne))) return rc;
  } else {
    if ((rc = upload_state<float>(m, m->sig, sig, 6 * ne))) return rc;
    if ((rc = upload_state<float>(m, m->eint, eint, ne))) return rc;
    if ((rc = upload_state<float>(m, m->rho, rho, ne))) return rc;
    if ((rc = upload_state<float>(m, m->vol0, vol0, ne))) return rc;
    if ((rc = upload_state<float>(m, m->pla, pla, ne))) return rc;
    if ((rc = upload_state<float>(m, m->epsd, epsd, ne))) return rc;
    if ((rc = upload_state<float>(m, m->offg, offg, ne))) return rc;
    if ((rc = upload_state<float>(m, m->qvis, qvis, ne))) return rc;
  }
  m->have_hex = true;
  return OR_OK;
}

int or_model_upload_skyline(or_model *m, const int64_t *adsky, const int64_t *iads) {
  if (!m || !adsky || (m->ne > 0 && !iads)) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  const int64_t nn = m->nn, ne = m->ne;
  if (adsky[0] != 0 || adsky[nn] != m->lsky)
    return fail(OR_ERR_ARG, "adsky must start at 0 and end at 8*ne");
  std::vector<int32_t> a32((size_t)(nn + 1));
  for (int64_t n = 0; n <= nn; ++n) {
    if (n > 0 && adsky[n] < adsky[n - 1]) return fail(OR_ERR_ARG, "adsky must be non-decreasing");
    a32[n] = (int32_t)adsky[n];
  }
  std::vector<int32_t> soa((size_t)(8 * ne));
  std::vector<uint8_t> seen((size_t)m->lsky, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int k = 0; k < 8; ++k) {
      const int64_t K = iads[8 * e + k];
      if (K < 0 || K >= m->lsky) return fail(OR_ERR_ARG, "iads row out of range");
      if (seen[K]) return fail(OR_ERR_ARG, "iads rows must be distinct");
      seen[K] = 1;
      soa[(size_t)(k * ne + e)] = (int32_t)K;
    }
  OR_HIP(hipMemcpy(m->adsky, a32.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice));
  if (ne) OR_HIP(hipMemcpy(m->iads, soa.data(), sizeof(int32_t) * 8 * ne, hipMemcpyHostToDevice));
  m->have_sky = true;
  return OR_OK;
}

int or_model_set_time(or_model *m, const or_time_state *t) {
  if (!m || !t) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(hipStreamSynchronize(m->stream));
  DevTime dt{};
  dt.tt = t->tt; dt.dt1 = t->dt1; dt.dt2 = t->dt2; dt.dt12 = t->dt12; dt.dt2old = t->dt2old;
  dt.ncycle = t->ncycle; dt.mstop = t->mstop;
  DevTime two[2] = {dt, dt};
  OR_HIP(hipMemcpy(m->ts, two, sizeof(two), hipMemcpyHostToDevice));
  int zero = 0;
  OR_HIP(hipMemcpy(m->flags, &zero, sizeof(int), hipMemcpyHostToDevice));
  m->parity = 0;
  return OR_OK;
}

int or_model_get_time(or_model *m, or_time_state *t) {
  if (!m || !t) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(hipStreamSynchronize(m->stream));
  DevTime dt;
  OR_HIP(hipMemcpy(&dt, m->ts + (m->parity & 1), sizeof(DevTime), hipMemcpyDeviceToHost));
  int flags = 0;
  OR_HIP(hipMemcpy(&flags, m->flags, sizeof(int), hipMemcpyDeviceToHost));
  t->tt = dt.tt; t->dt1 = dt.dt1; t->dt2 = dt.dt2; t->dt12 = dt.dt12; t->dt2old = dt.dt2old;
  t->ncycle = dt.ncycle; t->mstop = dt.mstop | flags;
  return OR_OK;
}

static int flush_timing(or_model *m) {
  if (m->ring_used == 0) return OR_OK;
  OR_HIP(hipEventSynchronize(m->ev[3 * (m->ring_used - 1) + 2]));
  for (int j = 0; j < m->ring_used; ++j) {
    float a = 0.f, b = 0.f;
    OR_HIP(hipEventElapsedTime(&a, m->ev[3 * j], m->ev[3 * j + 1]));
    OR_HIP(hipEventElapsedTime(&b, m->ev[3 * j + 1], m->ev[3 * j + 2]));
    m->t_elem += a;
    m->t_node += b;
    m->t_cycles += 1;
  }
  m->ring_used = 0;
  return OR_OK;
}

template <typename R, typename S, typename Row> static int launch_cycle(or_model *m) {
  const int p = (int)(m->parity & 1);
  NodePtrs nd{m->xv, m->d, m->ms, m->code, m->vimp, m->stifn, m->adsky, m->nn};
  if (m->ne > 0) {
    HexPtrs<S> e{m->conn, m->iads, (S *)m->sig, (S *)m->eint, (S *)m->rho, (S *)m->pla, (S *)m->epsd,
                 (S *)m->offg, (S *)m->qvis, (const S *)m->vol0, m->ne};
    const unsigned gb = (unsigned)((m->ne + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_hex_forces<R, S, Row>), dim3(gb), dim3(kBlock), 0, m->stream, e, m->xv, (Row *)m->sky,
                       m->consts, m->ts, p, m->dtmin, m->flags, (S *)nullptr);
    OR_HIP(hipGetLastError());
  }
  if (m->timing) OR_HIP(hipEventRecord(m->ev[3 * m->ring_used + 1], m->stream));
  const unsigned gn = (unsigned)((m->nn + kBlock - 1) / kBlock);
  hipLaunchKernelGGL((k_node_update<Row>), dim3(gn), dim3(kBlock), 0, m->stream, nd, (const Row *)m->sky, m->ts, p,
                     m->dtmin, m->flags, m->consts.dt_growth, m->consts.dtmx);
  OR_HIP(hipGetLastError());
  m->parity ^= 1;
  return OR_OK;
}

int or_model_run(or_model *m, int64_t ncycles) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  if (ncycles < 0) return fail(OR_ERR_ARG, "ncycles < 0");
  if (!m->have_params || !m->have_nodes || !m->have_hex || !m->have_sky)
    return fail(OR_ERR_STATE, "params, nodes, hex state and skyline must be uploaded before run");
  OR_HIP(hipSetDevice(m->device));
  for (int64_t c = 0; c < ncycles; ++c) {
    if (m->timing) {
      if (m->ring_used * 3 + 3 > (int)m->ev.size()) {
        int rc = flush_timing(m);
        if (rc) return rc;
      }
      OR_HIP(hipEventRecord(m->ev[3 * m->ring_used], m->stream));
    }
    int rc = (m->prec == OR_PREC_FP64) ? launch_cycle<double, double, RowD>(m) : launch_cycle<float, float, RowF>(m);
    if (rc) return rc;
    if (m->timing) {
      OR_HIP(hipEventRecord(m->ev[3 * m->ring_used + 2], m->stream));
      m->ring_used += 1;
    }
  }
  return OR_OK;
}

int or_model_sync(or_model *m) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(hipStreamSynchronize(m->stream));
  int flags = 0;
  OR_HIP(hipMemcpy(&flags, m->flags, sizeof(int), hipMemcpyDeviceToHost));
  if (flags & 1) return fail(OR_ERR_MSTOP, "non-positive Jacobian: element deleted and MSTOP raised");
  return OR_OK;
}

int or_model_set_timing(or_model *m, int enable) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  if (enable && m->ev.empty()) {
    m->ev.resize(3 * 512);
    for (auto &e : m->ev) OR_HIP(hipEventCreate(&e));
  }
  m->timing = enable != 0;
  return OR_OK;
}

int or_model_kernel_times(or_model *m, int reset, double *elem_ms, double *node_ms, int64_t *cycles) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  int rc = flush_timing(m);
  if (rc) return rc;
  const double nc = m->t_cycles > 0 ? (double)m->t_cycles : 1.0;
  if (elem_ms) *elem_ms = m->t_elem / nc;
  if (node_ms) *node_ms = m->t_node / nc;
  if (cycles) *cycles = m->t_cycles;
  if (reset) { m->t_elem = 0.0; m->t_node = 0.0; m->t_cycles = 0; }
  return OR_OK;
}

int or_model_download_nodes(or_model *m, double *x, double *v, double *d, double *stifn) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(hipStreamSynchronize(m->stream));
  const int64_t nn = m->nn;
  std::vector<double> xv((size_t)(6 * nn));
  OR_HIP(hipMemcpy(xv.data(), m->xv, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost));
  for (int64_t n = 0; n < nn; ++n) {
    if (x) { x[3 * n] = xv[6 * n]; x[3 * n + 1] = xv[6 * n + 1]; x[3 * n + 2] = xv[6 * n + 2]; }
    if (v) { v[3 * n] = xv[6 * n + 3]; v[3 * n + 1] = xv[6 * n + 4]; v[3 * n + 2] = xv[6 * n + 5]; }
  }
  if (d) OR_HIP(hipMemcpy(d, m->d, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost));
  if (stifn) {
    if (!m->stifn) return fail(OR_ERR_STATE, "STIFN assembly is not enabled on this model");
    OR_HIP(hipMemcpy(stifn, m->stifn, sizeof(double) * nn, hipMemcpyDeviceToHost));
  }
  return OR_OK;
}

int or_model_download_hex(or_model *m, double *sig, double *eint, double *rho, double *pla, double *epsd,
                          double *offg, double *qvis) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(hipStreamSynchronize(m->stream));
  const int64_t ne = m->ne;
  int rc;
  if (m->prec == OR_PREC_FP64) {
    if ((rc = download_state<double>(m, sig, m->sig, 6 * ne))) return rc;
    if ((rc = download_state<double>(m, eint, m->eint, ne))) return rc;
    if ((rc = download_state<double>(m, rho, m->rho, ne))) return rc;
    if ((rc = download_state<double>(m, pla, m->pla, ne))) return rc;
    if ((rc = download_state<double>(m, epsd, m->epsd, ne))) return rc;
    if ((rc = download_state<double>(m, offg, m->offg, ne))) return rc;
    if ((rc = download_state<double>(m, qvis, m->qvis, ne))) return rc;
  } else {
    if ((rc = download_state<float>(m, sig, m->sig, 6 * ne))) return rc;
    if ((rc = download_state<float>(m, eint, m->eint, ne))) return rc;
    if ((rc = download_state<float>(m, rho, m->rho, ne))) return rc;
    if ((rc = download_state<float>(m, pla, m->pla, ne))) return rc;
    if ((rc = download_state<float>(m, epsd, m->epsd, ne))) return rc;
    if ((rc = download_state<float>(m, offg, m->offg, ne))) return rc;
    if ((rc = download_state<float>(m, qvis, m->qvis, ne))) return rc;
  }
  return OR_OK;
}

int or_model_energy(or_model *m, double out[2]) {
  if (!m || !out) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  const int64_t n = m->nn > m->ne ? m->nn : m->ne;
  const int nb = 1024;
  const int64_t per = (n + nb - 1) / nb;
  const int64_t per_block = ((per + kBlock - 1) / kBlock) * kBlock;
  double *partial = m->energy + 2;
  if (m->prec == OR_PREC_FP64)
    hipLaunchKernelGGL((k_energy_partial<double>), dim3(nb), dim3(kBlock), 0, m->stream, (const double *)m->eint,
                       (const double *)m->vol0, m->ne, m->xv, m->ms, m->nn, per_block, partial);
  else
    hipLaunchKernelGGL((k_energy_partial<float>), dim3(nb), dim3(kBlock), 0, m->stream, (const float *)m->eint,
                       (const float *)m->vol0, m->ne, m->xv, m->ms, m->nn, per_block, partial);
  OR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, m->stream, (const double *)partial, nb, m->energy);
  OR_HIP(hipGetLastError());
  OR_HIP(hipStreamSynchronize(m->stream));
  OR_HIP(hipMemcpy(out, m->energy, 2 * sizeof(double), hipMemcpyDeviceToHost));
  return OR_OK;
}

int or_model_footprint(const or_model *m, int64_t *bytes) {
  if (!m || !bytes) return fail(OR_ERR_ARG, "null argument");
  *bytes = m->bytes;
  return OR_OK;
}

int or_sforc3_group(int device, const or_hex_params *p, int64_t numnod, int32_t nel, int64_t nft, const double *x,
                    const double *v, const int32_t *ixs, double *sig, double *eint, double *rho, const double *vol,
                    double *pla, double *epsd, double *off, double *qvis, int64_t lsky, double *fsky,
                    const int32_t *iads, double dt1, double *dt2t, int32_t *mstop) {
  if (!p || !x || !v || !ixs || !sig || !eint || !rho || !vol || !pla || !epsd || !off || !qvis || !fsky || !iads ||
      !dt2t || !mstop)
    return fail(OR_ERR_ARG, "null argument");
  if (nel <= 0 || nel > 128) return fail(OR_ERR_ARG, "nel must be in 1..128 (MVSIZ-1)");
  if (numnod <= 0 || lsky <= 0 || nft < 0) return fail(OR_ERR_ARG, "bad sizes");
  int rc = check_params(*p);
  if (rc) return rc;
  for (int i = 0; i < nel; ++i) {
    for (int k = 0; k < 8; ++k) {
      const int32_t nid = ixs[11 * (nft + i) + 1 + k];
      if (nid < 1 || nid > numnod) return fail(OR_ERR_ARG, "IXS node id out of range");
      const int32_t K = iads[8 * (nft + i) + k];
      if (K < 1 || K > lsky) return fail(OR_ERR_ARG, "IADS row out of range");
    }
  }
  OR_HIP(hipSetDevice(device));
  HexConsts c = make_consts(*p);
  const int64_t ncol = nft + nel;
  double *dx = nullptr, *dv = nullptr, *dbuf = nullptr, *dfsky = nullptr, *ddtx = nullptr;
  int32_t *dixs = nullptr, *diads = nullptr, *dms = nullptr;
  int ok = 1;
  ok &= hipMalloc(&dx, sizeof(double) * 3 * numnod) == hipSuccess;
  ok &= hipMalloc(&dv, sizeof(double) * 3 * numnod) == hipSuccess;
  ok &= hipMalloc(&dixs, sizeof(int32_t) * 11 * ncol) == hipSuccess;
  ok &= hipMalloc(&diads, sizeof(int32_t) * 8 * ncol) == hipSuccess;
  ok &= hipMalloc(&dbuf, sizeof(double) * 13 * nel) == hipSuccess;
  ok &= hipMalloc(&dfsky, sizeof(double) * 8 * lsky) == hipSuccess;
  ok &= hipMalloc(&ddtx, sizeof(double) * nel) == hipSuccess;
  ok &= hipMalloc(&dms, sizeof(int32_t) * nel) == hipSuccess;
  auto cleanup = [&]() {
    void *ps[] = {dx, dv, dixs, diads, dbuf, dfsky, ddtx, dms};
    for (void *q : ps)
      if (q) hipFree(q);
  };
  if (!ok) { cleanup(); return fail(OR_ERR_HIP, "hipMalloc in or_sforc3_group"); }
  double *dsig = dbuf, *deint = dbuf + 6 * nel, *drho = dbuf + 7 * nel, *dvol = dbuf + 8 * nel,
         *dpla = dbuf + 9 * nel, *depsd = dbuf + 10 * nel, *doff = dbuf + 11 * nel, *dq = dbuf + 12 * nel;
  hipError_t e = hipSuccess;
  auto H2D = [&](void *d, const void *h, size_t b) { if (e == hipSuccess) e = hipMemcpy(d, h, b, hipMemcpyHostToDevice); };
  auto D2H = [&](void *h, const void *d, size_t b) { if (e == hipSuccess) e = hipMemcpy(h, d, b, hipMemcpyDeviceToHost); };
  H2D(dx, x, sizeof(double) * 3 * numnod);
  H2D(dv, v, sizeof(double) * 3 * numnod);
  H2D(dixs, ixs, sizeof(int32_t) * 11 * ncol);
  H2D(diads, iads, sizeof(int32_t) * 8 * ncol);
  H2D(dsig, sig, sizeof(double) * 6 * nel);
  H2D(deint, eint, sizeof(double) * nel);
  H2D(drho, rho, sizeof(double) * nel);
  H2D(dvol, vol, sizeof(double) * nel);
  H2D(dpla, pla, sizeof(double) * nel);
  H2D(depsd, epsd, sizeof(double) * nel);
  H2D(doff, off, sizeof(double) * nel);
  H2D(dq, qvis, sizeof(double) * nel);
  H2D(dfsky, fsky, sizeof(double) * 8 * lsky);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sforc3_group, dim3((nel + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, nel, nft, dx, dv, dixs,
                       dsig, deint, drho, dvol, dpla, depsd, doff, dq, dfsky, diads, c, dt1, ddtx, dms);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  std::vector<double> dtx((size_t)nel);
  std::vector<int32_t> msv((size_t)nel);
  D2H(sig, dsig, sizeof(double) * 6 * nel);
  D2H(eint, deint, sizeof(double) * nel);
  D2H(rho, drho, sizeof(double) * nel);
  D2H(pla, dpla, sizeof(double) * nel);
  D2H(epsd, depsd, sizeof(double) * nel);
  D2H(off, doff, sizeof(double) * nel);
  D2H(qvis, dq, sizeof(double) * nel);
  D2H(fsky, dfsky, sizeof(double) * 8 * lsky);
  D2H(dtx.data(), ddtx, sizeof(double) * nel);
  D2H(msv.data(), dms, sizeof(int32_t) * nel);
  cleanup();
  if (e != hipSuccess) return fail(OR_ERR_HIP, std::string("or_sforc3_group: ") + hipGetErrorString(e));
  for (int i = 0; i < nel; ++i) {
    if (dtx[i] < *dt2t) *dt2t = dtx[i];
    if (msv[i]) *mstop = 1;
  }
  return OR_OK;
}

