// or_capi.hip -- kernels and C ABI of the MI355X explicit hex / Q4 path (gfx950).
//
// Per cycle, on one stream, with every scalar kept on the device:
//   k_hex_forces : SFORC3 chain per 8-node hex -> 4-word skyline rows
//   k_q4_forces  : CFORC3 chain per Q4 shell   -> 8-word skyline rows
//                  (both fold their element time steps into one device word with an
//                  integer atomicMin on the FP64 bit pattern: min is exact, so the
//                  result does not depend on block order)
//   k_node_update: ASSPAR4 fixed-order segmented sum per node (hex rows, then Q4 rows,
//                  the FILLCNE order), DT2 pipeline, ACCELE, BCS10, FIXVEL, VELOCITY,
//                  DEPLA; thread 0 advances the cycle scalars (double-buffered by parity)
//
// The file is compiled twice: once with the compiler's default instruction scheduler
// (everything), and once with -DOR_ILP_TU and the ILP-first scheduler
// (-mllvm -amdgpu-sched-strategy=max-ilp), which only provides the element-kernel launches
// of variants 2 and 3 (or_ilp_elements).  Scheduling reorders instructions and never changes
// an operation or its operands, so every variant gives the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <numeric>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/or_gpu.h"
#include "or_assembly.h"
#include "or_hex.h"
#include "or_q4.h"

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
constexpr double kNoEnd = 1.0e300; // end time of a model advanced by or_model_run

// Launch variants of the element kernels.  Variant v picks the minimum number of waves
// per SIMD the register allocation is built for (two for even v, three for odd v, where
// the FP64 kernels spill to scratch; a wave's register budget is 512 / waves VGPRs and
// AGPRs together) and the instruction scheduler (default for v < 2, ILP-first for v >= 2).
// Every element workgroup is one tile of kTile elements, the unit of the LDS assembly.
// Every variant runs the same arithmetic, so the choice (or_model_set_variant,
// or_model_autotune, or_batch_autotune) changes speed only, never a result bit.
constexpr int kVariants = 4;
constexpr int kTile = orasm::kTile;
static_assert(kTile == kBlock, "element tiles and node blocks share the workgroup size");
template <typename R> constexpr int variant_waves(int v) { return 2 + (v & 1); }

// MSTOP codes (include/or_gpu.h): an element's non-positive Jacobian, a time step <= 0
constexpr int32_t kMstopJacobian = 1, kMstopTimeStep = 2;

// node code bits (include/or_gpu.h)
constexpr int32_t kTX = 4, kTY = 2, kTZ = 1, kRX = 32, kRY = 16, kRZ = 8, kIX = 256, kIY = 128, kIZ = 64;

// Assembly of the resident model (plan in or_assembly.h).  The rows of a node that
// lead its ASSPAR4 order and whose elements share the tile of its first element are
// summed in that tile's LDS from zero and stored once, in FP64, as the node's prefix
// (pre[n][W]); the node kernel starts from the prefix and adds the node's remaining
// ("rest") rows in order.  Rest rows are stored node after node (CSR, rs[n] ..
// rs[n + 1] - 1), W words each in the storage type: W = 3 (fx, fy, fz) for hexes,
// 6 (plus mx, my, mz) for shells.  Every node with hex rows has a hex prefix; a node
// with shell rows has a shell prefix only when it has no hex rows (its sum does not
// start with a shell row otherwise).  The stiffness words (FSKY 7, 8) have no reader
// on this path (element time step, NODADT=0, no interfaces), so the resident rows
// leave them out; the group drop-ins write them.
constexpr int kHexRowWords = 3; // fx, fy, fz
constexpr int kQ4RowWords = 6;  // fx, fy, fz, mx, my, mz
typedef uint32_t v4u __attribute__((ext_vector_type(4))); // a prefix's 8 LDS sources

// Workgroups of a launch are dispatched to the 8 XCDs in turn (workgroup b runs on
// XCD b mod 8), each XCD with its own L2.  Element kernels renumber their workgroups
// so that each XCD works through one contiguous eighth of the element range:
// neighbouring elements, which share nodes, then gather through the same L2.
// The map is a bijection of [0, nb).
__device__ __forceinline__ int64_t xcd_contiguous_block(int64_t b, int64_t nb) {
  constexpr int64_t kXcd = 8;
  const int64_t q = nb / kXcd, r = nb % kXcd;
  const int64_t x = b % kXcd, j = b / kXcd;
  return x * q + (x < r ? x : r) + j;
}

// Pointers into HBM as the kernels use them: typed in the global address space, so
// that their loads and stores compile to global (not flat) memory instructions even
// though the kernels read them from a descriptor in memory.
template <typename T> using gptr = __attribute__((address_space(1))) T *;
typedef double v2d __attribute__((ext_vector_type(2))); // 16-byte load / store unit
template <typename T> __device__ __forceinline__ gptr<T> to_global(T *p) { return (gptr<T>)p; }
template <typename T> __device__ __forceinline__ gptr<T> to_global(const void *p) { return (gptr<T>)(T *)p; }

// A model's cycle scalars, read field by field with global loads: a flat load among
// an element's global loads makes the compiler wait for all of them at its first use.
static_assert(offsetof(DevTime, ncycle) == 40 && offsetof(DevTime, mstop) == 48 && offsetof(DevTime, done) == 52,
              "DevTime layout");
__device__ __forceinline__ DevTime load_time(const DevTime *t) {
  const gptr<const double> d = to_global<const double>(t);
  DevTime r;
  r.tt = d[0]; r.dt1 = d[1]; r.dt2 = d[2]; r.dt12 = d[3]; r.dt2old = d[4];
  r.ncycle = to_global<const int64_t>(t)[5];
  const gptr<const int32_t> w = to_global<const int32_t>(t);
  r.mstop = w[12]; r.done = w[13];
  return r;
}
__device__ __forceinline__ double time_dt2(const DevTime *t) { return to_global<const double>(t)[2]; }
__device__ __forceinline__ int32_t time_done(const DevTime *t) { return to_global<const int32_t>(t)[13]; }

template <typename S> struct HexPtrs {
  gptr<const int32_t> conn;  // [8][ne]
  gptr<const int32_t> rest;  // [8][ne] rest row of the corner, or -1 (summed in LDS)
  gptr<S> sig;               // [6][ne]
  gptr<S> eint, rho, pla, epsd, offg, qvis;
  gptr<const S> vol0;
  int64_t ne;
};

template <typename S> struct Q4Ptrs {
  gptr<const int32_t> conn;  // [4][ne]
  gptr<const int32_t> rest;  // [4][ne] rest row of the corner, or -1 (summed in LDS)
  gptr<S> offg, thk;
  gptr<const S> thke;
  gptr<S> forc, mom, gstr, hourg, eint, epsd; // [5][ne] [3][ne] [8][ne] [5][ne] [2][ne] [ne]
  gptr<S> sig, pla, epsd_ip;                  // [npt][5][ne] [npt][ne] [npt][ne] (epsd_ip null when israte=0)
  int64_t ne;
};

struct NodePtrs {
  double *xv;            // [nn][6] x,y,z,vx,vy,vz
  double *vr;            // [nn][4] rotational velocity (padded) or null
  double *d;             // [nn][3]
  const double *ms;      // [nn]
  const double *in;      // [nn] or null
  const int32_t *code;   // [nn]
  const double *vimp;    // [nn][3]
  const int32_t *deg_h, *deg_q;    // [nn] rows of each node per family (null when the family is absent)
  const int32_t *rs_h, *rs_q;      // [nn + 1] first rest row of each node per family
  const double *pre_h, *pre_q;     // [nn][3], [nn][6] FP64 prefixes
  double *wext;                    // [2][nn] external work of the node's imposed DOFs, and the part
                                   // each cycle defers to the next (FIXVEL's VEL(4,N)), or null
  int64_t nn;
};

// One model as the cycle kernels see it: pointers into its resident arrays, its
// parameters and its cycle scalars.  Held in device memory, one per model (and one
// per slot of a batch); a launch covers any number of models through a tile list.
struct ModelDev {
  // hexes; element state in the model's storage type (double, or float in the split form)
  const int32_t *hconn, *hrest;
  const int32_t *hfr_start, *hfr_node; // prefixes of each hex tile: [ntiles + 1], [nfr]
  const uint16_t *hfr_src;              // [nfr][8] LDS sources of each prefix
  void *hsig, *heint, *hrho, *hpla, *hepsd, *hoffg, *hqvis;
  const void *hvol0;
  int64_t neh;
  // Q4 shells
  const int32_t *qconn, *qrest;
  const int32_t *qfr_start, *qfr_node;
  const uint16_t *qfr_src;
  void *qoffg, *qthk, *qforc, *qmom, *qgstr, *qhourg, *qeint, *qepsd, *qsig, *qpla, *qepsdip;
  const void *qthke;
  int64_t neq;
  NodePtrs nd;
  void *rowh, *rowq;          // rest rows (storage type)
  double *preh, *preq;        // prefixes (written by the element kernels)
  DevTime *ts;                // [2] cycle scalars, by cycle parity
  unsigned long long *dtmin;  // [2] element time-step minimum (FP64 bits), by parity
  int *flags;                 // bit 0: an element reported MSTOP
  double *ecrit;              // [5][nn] per-node kinetic and momentum terms of ECRIT on output cycles, or null
  const int32_t *th_slot;     // [nn] the node's row in th_node, or -1 (time history), or null
  double *th_node;            // [rows][18] D, V, A, VR, AR, X of the time-history nodes on output cycles
  double *ehou_h, *ehou_q;    // per tile: its elements' hourglass energy summed over the cycles, or null
  double *dtel_h, *dtel_q;    // per-element time-step candidates on output cycles (storage order), or null
  double dtmx;                // DTMX
  double tend;                // the model stops after the cycle whose time reaches tend
  HexConsts hc;
  Q4Consts qc;
};

// One block of a cycle launch: consecutive elements (the launch's workgroup size) or
// kBlock consecutive nodes of one model.
struct Tile {
  int32_t model; // index in the launch's ModelDev array
  int32_t first; // first element (node) of the block
};

template <typename S> __device__ __forceinline__ HexPtrs<S> hex_ptrs(const ModelDev &M) {
  return HexPtrs<S>{to_global(M.hconn), to_global(M.hrest), to_global<S>(M.hsig), to_global<S>(M.heint),
                    to_global<S>(M.hrho), to_global<S>(M.hpla), to_global<S>(M.hepsd), to_global<S>(M.hoffg),
                    to_global<S>(M.hqvis), to_global<const S>(M.hvol0), M.neh};
}
template <typename S> __device__ __forceinline__ Q4Ptrs<S> q4_ptrs(const ModelDev &M) {
  return Q4Ptrs<S>{to_global(M.qconn), to_global(M.qrest), to_global<S>(M.qoffg), to_global<S>(M.qthk),
                   to_global<const S>(M.qthke), to_global<S>(M.qforc), to_global<S>(M.qmom), to_global<S>(M.qgstr),
                   to_global<S>(M.qhourg), to_global<S>(M.qeint), to_global<S>(M.qepsd), to_global<S>(M.qsig),
                   to_global<S>(M.qpla), to_global<S>(M.qepsdip), M.neq};
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

template <int TB> __device__ __forceinline__ double block_min(double v) {
  __shared__ double red[TB / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < TB / 64; ++w) v = (red[w] < v) ? red[w] : v;
  }
  return v;
}

__device__ __forceinline__ void load_node(gptr<const double> xv, int64_t n, double &x, double &y, double &z,
                                          double &vx, double &vy, double &vz) {
  gptr<const v2d> p = (gptr<const v2d>)(xv + 6 * n);
  const v2d a = p[0], b = p[1], q = p[2];
  x = a.x; y = a.y; z = b.x; vx = b.y; vy = q.x; vz = q.y;
}

// ---------------------------------------------------------------- tile assembly
// The element chains leave every corner's W row words in LDS, in the storage type (the
// value a row holds): corner k, word w of the tile's element t at lds[(k * W + w) * kTile + t].
// After a barrier the tile writes its rest rows from there and sums its prefixes.
template <typename S, int W>
__device__ __forceinline__ void lds_corner(S *lds, int k, const S (&w)[W]) {
#pragma unroll
  for (int c = 0; c < W; ++c) lds[(k * W + c) * kTile + threadIdx.x] = w[c];
}

// The tile's assembly after its element chains: this thread's rest rows (the corners of its
// element whose rest row is >= 0, written from LDS) and the tile's prefixes (each summed in
// row order from zero in FP64, ASSPAR4's running sum A = A + FSKY, stored in pre[n][W]),
// thread t taking prefixes t, t + kTile, ...  The indices are loaded before the tile's
// barrier (prefetch), so that their latency overlaps the wait for the tile's other waves.
template <typename S, int W, int NK> struct TileAssembly {
  int32_t rest[NK];
  int32_t j, j1, node;
  v4u src;

  __device__ __forceinline__ void prefetch(gptr<const int32_t> restp, int64_t ne, int64_t i,
                                           gptr<const int32_t> fr_start, gptr<const int32_t> fr_node,
                                           gptr<const v4u> fr_src, int64_t tile) {
    __asm__ __volatile__("" ::: "memory"); // after the element chain, not hoisted into it
#pragma unroll
    for (int k = 0; k < NK; ++k) rest[k] = (i < ne) ? restp[(int64_t)k * ne + i] : -1;
    j = fr_start[tile] + (int32_t)threadIdx.x;
    j1 = fr_start[tile + 1];
    if (j < j1) {
      node = fr_node[j];
      src = fr_src[j];
    }
  }

  __device__ __forceinline__ void finish(const S *lds, gptr<S> rows, gptr<const int32_t> fr_node,
                                         gptr<const v4u> fr_src, gptr<double> pre) {
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      if (rest[k] >= 0) {
        const gptr<S> w = rows + (int64_t)rest[k] * W;
#pragma unroll
        for (int c = 0; c < W; ++c) w[c] = lds[(k * W + c) * kTile + threadIdx.x];
      }
    }
    for (; j < j1; j += kTile) {
      double a[W];
#pragma unroll
      for (int c = 0; c < W; ++c) a[c] = 0.0;
#pragma unroll
      for (int r = 0; r < orasm::kRunCap; ++r) {
        const uint32_t u = (src[r >> 1] >> (16 * (r & 1))) & 0xFFFFu;
        if (u == orasm::kNoSrc) break;
        const int base = (int)(u >> 8) * W * kTile + (int)(u & 0xFFu);
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = a[c] + double(lds[base + c * kTile]);
      }
      const gptr<double> p = pre + (int64_t)node * W;
#pragma unroll
      for (int c = 0; c < W; ++c) p[c] = a[c];
      if (j + kTile < j1) { // the tile holds more prefixes than threads
        node = fr_node[j + kTile];
        src = fr_src[j + kTile];
      }
    }
  }
};

// The element time-step minimum of a tile: each wave's minimum goes to red[] before the
// tile's barrier, and after it the first thread folds them into the model's word with an
// integer atomicMin on the FP64 bits (min is exact, so the order does not matter).
__device__ __forceinline__ void wave_min_to(double *red, double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
}
template <int TB> __device__ __forceinline__ void tile_min_commit(const double *red, unsigned long long *word) {
  if (threadIdx.x == 0) {
    double v = red[0];
#pragma unroll
    for (int w = 1; w < TB / 64; ++w) v = (red[w] < v) ? red[w] : v;
    atomicMin(word, dbits(v));
  }
}

// The hourglass energy (EHOU) of a tile's elements: each element stores its EHOU in the tile's
// LDS as soon as it is known (no register holds it through the rest of the chain); after the
// tile's barrier the first wave sums them in a fixed order (per lane its elements in order,
// then a shuffle tree) and adds the sum to the tile's accumulator, which only this tile
// writes, so the sum is the same on every run.
template <int TB, typename R> __device__ __forceinline__ void tile_ehou_commit(const R *eh, int64_t first, int64_t ne,
                                                                               double *acc) {
  if (threadIdx.x < 64) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < TB / 64; ++k) {
      const int t = threadIdx.x + 64 * k;
      if (first + t < ne) v = v + double(eh[t]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    if (threadIdx.x == 0) *to_global(acc) = *to_global(acc) + v;
  }
}

// ---------------------------------------------------------------- hex element kernel
// In the split form, the hex element state read by the material phase (SIG(6), EINT, RHO,
// VOL0, PLA, EPSD, QVIS) is copied from HBM to LDS by the wave itself at the start of the chain, with
// global-to-LDS loads that occupy no registers: the copy is in flight during the gather,
// the geometry and the kinematics, and the material phase reads LDS.  Word a of the
// tile's element t sits at sst[a * kTile + t]; each wave copies its own 64 elements, so
// only the wave's own wait for its loads orders the copy before its reads.
constexpr int kHexStagedWords = 12;
typedef __attribute__((address_space(3))) void *lds_void;

template <typename S>
__device__ __forceinline__ void stage_hex_state(S *sst, const HexPtrs<S> &e, int64_t first) {
  constexpr int D = sizeof(S) / 4; // dwords per word
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t wfirst = first + 64 * w;
  const gptr<const S> src[kHexStagedWords] = {e.sig,          e.sig + e.ne,     e.sig + 2 * e.ne, e.sig + 3 * e.ne,
                                              e.sig + 4 * e.ne, e.sig + 5 * e.ne, e.eint,           e.rho,
                                              e.vol0,         e.pla,            e.epsd,           e.qvis};
#pragma unroll
  for (int a = 0; a < kHexStagedWords; ++a) {
#pragma unroll
    for (int h = 0; h < D; ++h) {
      const int d = h * 64 + l; // dword of the wave's slice
      if (wfirst + d / D < e.ne)
        __builtin_amdgcn_global_load_lds((gptr<void>)((gptr<const char>)(src[a] + wfirst) + 4 * d),
                                         (lds_void)((char *)(sst + a * kTile + 64 * w) + 256 * h), 4, 0, 0);
    }
  }
}

// The wave's global-to-LDS loads have landed (they count as vector memory loads).
__device__ __forceinline__ void wait_staged() { __asm__ __volatile__("s_waitcnt vmcnt(0)" ::: "memory"); }

template <typename R, typename S, int TB, int W, bool EH>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(W))) void k_hex_forces(
    const ModelDev *__restrict__ md, const Tile *__restrict__ tiles, int parity_output) {
  static_assert(TB == kTile, "a hex workgroup is one assembly tile");
  // bit 0: cycle parity; bit 1: an output cycle (each element's time-step candidate is kept)
  const int parity = parity_output & 1;
  const bool output = (parity_output & 2) != 0;
  __shared__ S lds[8 * kHexRowWords * kTile];
  const Tile tl = tiles[xcd_contiguous_block(blockIdx.x, gridDim.x)];
  const ModelDev &M = md[tl.model];
  const DevTime *ts = M.ts;
  // a stopped model computes but stores nothing; the flag is tested only before the
  // stores, so that its load overlaps the element's own loads
  const int stopped = time_done(ts + parity);
  const HexPtrs<S> e = hex_ptrs<S>(M);
  const gptr<const double> xv = to_global<const double>(M.nd.xv);
  const HexConsts &c = M.hc;
  const int64_t i = (int64_t)tl.first + threadIdx.x;
  double dtx = kEP06;
  int mstop = 0;
  __shared__ R seh[EH ? TB : 1];
  // the split form stages its element state through LDS; the FP64 chain, bound by its
  // arithmetic at two waves per SIMD, reads it from HBM after the kinematics
  constexpr bool kStage = sizeof(S) == 4;
  __shared__ S sst[kStage ? kHexStagedWords * kTile : 1];
  if constexpr (kStage) stage_hex_state<S>(sst, e, tl.first);
  if (i < e.ne) {
    const R DT1 = R(time_dt2(ts + parity));
    HexState<S> st;
    HexOut<R> o;
    // EH: the model keeps EHOUR (its own instantiation, so that the other keeps its registers)
    auto ehou_sink = [&](R v) {
      if constexpr (EH) seh[threadIdx.x] = v;
    };
    double xd[8], yd[8], zd[8];
    R vx[8], vy[8], vz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double a, b, cc, u, v, w;
      load_node(xv, e.conn[(int64_t)k * e.ne + i], a, b, cc, u, v, w);
      xd[k] = a; yd[k] = b; zd[k] = cc;
      vx[k] = R(u); vy[k] = R(v); vz[k] = R(w);
    }
    // GBUF%OFF enters the geometry; the rest of the element state is read after the
    // kinematics (the fence keeps the compiler from hoisting the loads), so it is not
    // live through the geometry and the kinematics
    st.offg = e.offg[i];
    auto load_rest = [&]() {
      if constexpr (kStage) {
        wait_staged();
        const S *s = sst + threadIdx.x;
#pragma unroll
        for (int k = 0; k < 6; ++k) st.sig[k] = s[k * kTile];
        st.eint = s[6 * kTile]; st.rho = s[7 * kTile]; st.vol0 = s[8 * kTile]; st.pla = s[9 * kTile];
        st.epsd = s[10 * kTile]; st.qvis = s[11 * kTile];
      } else {
        __asm__ __volatile__("" ::: "memory");
#pragma unroll
        for (int k = 0; k < 6; ++k) st.sig[k] = e.sig[(int64_t)k * e.ne + i];
        st.eint = e.eint[i]; st.rho = e.rho[i]; st.vol0 = e.vol0[i]; st.pla = e.pla[i];
        st.epsd = e.epsd[i]; st.qvis = e.qvis[i];
      }
    };
    if constexpr (sizeof(R) == 8) {
      // FP64: each node's force triple leaves for LDS as soon as it is final, which keeps
      // the chain within two waves' register budget
      auto write_row = [&](int k, R fx, R fy, R fz) {
        const S w[3] = {S(fx), S(fy), S(fz)};
        lds_corner<S, 3>(lds, k, w);
      };
      hex_element<R, S>(xd, yd, zd, vx, vy, vz, st, c, DT1, o, write_row, load_rest, ehou_sink);
    } else {
      // FP32: the forces are gathered first and stored together (registers are not the limit)
      R f1[8], f2[8], f3[8];
      auto keep = [&](int k, R fx, R fy, R fz) { f1[k] = fx; f2[k] = fy; f3[k] = fz; };
      hex_element<R, S>(xd, yd, zd, vx, vy, vz, st, c, DT1, o, keep, load_rest, ehou_sink);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const S w[3] = {S(f1[k]), S(f2[k]), S(f3[k])};
        lds_corner<S, 3>(lds, k, w);
      }
    }
    if (!stopped) {
#pragma unroll
      for (int k = 0; k < 6; ++k) e.sig[(int64_t)k * e.ne + i] = st.sig[k];
      e.eint[i] = st.eint; e.rho[i] = st.rho; e.pla[i] = st.pla;
      e.epsd[i] = st.epsd; e.offg[i] = st.offg; e.qvis[i] = st.qvis;
      if (output && M.dtel_h) to_global(M.dtel_h)[i] = o.dtx;
    }
    dtx = o.dtx;
    mstop = o.mstop;
  }
  if (stopped) return; // uniform over the workgroup
  __shared__ double red[TB / 64];
  TileAssembly<S, kHexRowWords, 8> ta;
  ta.prefetch(e.rest, e.ne, i, to_global(M.hfr_start), to_global(M.hfr_node), to_global<const v4u>(M.hfr_src),
              tl.first / kTile);
  wave_min_to(red, dtx);
  if (mstop) atomicOr(M.flags, kMstopJacobian);
  __syncthreads();
  tile_min_commit<TB>(red, M.dtmin + parity);
  if constexpr (EH) {
    if (M.ehou_h) tile_ehou_commit<TB>(seh, tl.first, e.ne, M.ehou_h + tl.first / kTile);
  }
  ta.finish(lds, to_global<S>(M.rowh), to_global(M.hfr_node), to_global<const v4u>(M.hfr_src), to_global(M.preh));
}

// ---------------------------------------------------------------- Q4 element kernel
template <typename R, typename S> struct Q4IpIO {
  const Q4Ptrs<S> &q;
  int64_t i;
  bool active; // false: the model has stopped, nothing is stored
  __device__ __forceinline__ void load(int ip, R (&s)[5], R &pla, R &epsd_ip) const {
    const gptr<const S> sg = q.sig + (int64_t)ip * 5 * q.ne + i;
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] = R(sg[(int64_t)k * q.ne]);
    pla = R(q.pla[(int64_t)ip * q.ne + i]);
    epsd_ip = q.epsd_ip ? R(q.epsd_ip[(int64_t)ip * q.ne + i]) : R(0);
  }
  __device__ __forceinline__ void store(int ip, const R (&s)[5], R pla, R epsd_ip) const {
    if (!active) return;
    const gptr<S> sg = q.sig + (int64_t)ip * 5 * q.ne + i;
#pragma unroll
    for (int k = 0; k < 5; ++k) sg[(int64_t)k * q.ne] = S(s[k]);
    q.pla[(int64_t)ip * q.ne + i] = S(pla);
    if (q.epsd_ip) q.epsd_ip[(int64_t)ip * q.ne + i] = S(epsd_ip);
  }
  __device__ __forceinline__ void late_load(Q4Late<R> &L) const {
#pragma unroll
    for (int k = 0; k < 5; ++k) L.forc[k] = R(q.forc[(int64_t)k * q.ne + i]);
#pragma unroll
    for (int k = 0; k < 3; ++k) L.mom[k] = R(q.mom[(int64_t)k * q.ne + i]);
#pragma unroll
    for (int k = 0; k < 8; ++k) L.gstr[k] = R(q.gstr[(int64_t)k * q.ne + i]);
#pragma unroll
    for (int k = 0; k < 3; ++k) L.hourg[k] = R(q.hourg[(int64_t)k * q.ne + i]);
    L.eint[0] = R(q.eint[i]);
    L.eint[1] = R(q.eint[q.ne + i]);
  }
  __device__ __forceinline__ void late_store(const Q4Late<R> &L) const {
    if (!active) return;
#pragma unroll
    for (int k = 0; k < 5; ++k) q.forc[(int64_t)k * q.ne + i] = S(L.forc[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) q.mom[(int64_t)k * q.ne + i] = S(L.mom[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) q.gstr[(int64_t)k * q.ne + i] = S(L.gstr[k]);
    q.eint[i] = S(L.eint[0]);
    q.eint[q.ne + i] = S(L.eint[1]);
  }
  __device__ __forceinline__ void hourg_store(const R (&h)[5]) const {
    if (!active) return;
#pragma unroll
    for (int k = 0; k < 5; ++k) q.hourg[(int64_t)k * q.ne + i] = S(h[k]);
  }
};

template <typename R, typename S, int TB, int W, bool EH>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(W))) void k_q4_forces(
    const ModelDev *__restrict__ md, const Tile *__restrict__ tiles, int parity_output) {
  static_assert(TB == kTile, "a shell workgroup is one assembly tile");
  const int parity = parity_output & 1; // bit 1: an output cycle, as for the hex kernel
  const bool output = (parity_output & 2) != 0;
  __shared__ S lds[4 * kQ4RowWords * kTile];
  const Tile tl = tiles[xcd_contiguous_block(blockIdx.x, gridDim.x)];
  const ModelDev &M = md[tl.model];
  const DevTime *ts = M.ts;
  const int stopped = time_done(ts + parity); // tested before the stores only
  const Q4Ptrs<S> q = q4_ptrs<S>(M);
  const gptr<const double> xv = to_global<const double>(M.nd.xv);
  const gptr<const double> vr = to_global<const double>(M.nd.vr);
  const Q4Consts &c = M.qc;
  const int64_t i = (int64_t)tl.first + threadIdx.x;
  double dt = kEP06;
  __shared__ R seh[EH ? TB : 1];
  if (i < q.ne) {
    const R DT1 = R(time_dt2(ts + parity));
    double xg[4], yg[4], zg[4];
    R vl[4][3], vrl[4][3];
    int32_t nc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      nc[k] = q.conn[(int64_t)k * q.ne + i];
      double a, b, cc, u, v, w;
      load_node(xv, nc[k], a, b, cc, u, v, w);
      xg[k] = a; yg[k] = b; zg[k] = cc;
      vl[k][0] = R(u); vl[k][1] = R(v); vl[k][2] = R(w);
      const gptr<const v2d> p = (gptr<const v2d>)(vr + 4 * (int64_t)nc[k]);
      const v2d r0 = p[0], r1 = p[1];
      vrl[k][0] = R(r0.x); vrl[k][1] = R(r0.y); vrl[k][2] = R(r1.x);
    }
    Q4Elem<R> st;
    st.offg = R(q.offg[i]); st.thk = R(q.thk[i]); st.thke = R(q.thke[i]); st.epsd = R(q.epsd[i]);
    Q4IpIO<R, S> io{q, i, !stopped};
    Q4Out<R> o;
    // EH: the model keeps EHOUR (its own instantiation, so that the other keeps its registers)
    auto ehou_sink = [&](R v) {
      if constexpr (EH) seh[threadIdx.x] = v;
    };
    auto nload = [&](int k, double(&x)[3], R(&v)[3]) {
      __asm__ __volatile__("" ::: "memory"); // a fresh load, not the values held since the gather
      double a, b, cc, u, w, z;
      load_node(xv, nc[k], a, b, cc, u, w, z);
      x[0] = a; x[1] = b; x[2] = cc;
      v[0] = R(u); v[1] = R(w); v[2] = R(z);
    };
    q4_element<R>(xg, yg, zg, vl, vrl, nc[2] == nc[3], st, c, DT1, io, nload, o, ehou_sink);
    if (!stopped) {
      q.offg[i] = S(st.offg); q.thk[i] = S(st.thk); q.epsd[i] = S(st.epsd);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const S w[6] = {S(-o.f[0][k]), S(-o.f[1][k]), S(-o.f[2][k]), S(-o.m[0][k]), S(-o.m[1][k]), S(-o.m[2][k])};
        lds_corner<S, 6>(lds, k, w);
      }
      if (output && M.dtel_q) to_global(M.dtel_q)[i] = o.dt;
    }
    dt = o.dt;
  }
  if (stopped) return; // uniform over the workgroup
  __shared__ double red[TB / 64];
  TileAssembly<S, kQ4RowWords, 4> ta;
  ta.prefetch(q.rest, q.ne, i, to_global(M.qfr_start), to_global(M.qfr_node), to_global<const v4u>(M.qfr_src),
              tl.first / kTile);
  wave_min_to(red, dt);
  __syncthreads();
  tile_min_commit<TB>(red, M.dtmin + parity);
  if constexpr (EH) {
    if (M.ehou_q) tile_ehou_commit<TB>(seh, tl.first, q.ne, M.ehou_q + tl.first / kTile);
  }
  ta.finish(lds, to_global<S>(M.rowq), to_global(M.qfr_node), to_global<const v4u>(M.qfr_src), to_global(M.preq));
}

// ---------------------------------------------------------------- nodal update pieces
// ACCELE (accele.F:65-80, 97-109): A = F / MS, zero where the mass (inertia) is not positive.
__device__ __forceinline__ void accele3(double &a0, double &a1, double &a2, double mass) {
  if (mass > 0.0) {
    const double rtmp = 1.0 / mass;
    a0 = a0 * rtmp; a1 = a1 * rtmp; a2 = a2 * rtmp;
  } else {
    a0 = 0.0; a1 = 0.0; a2 = 0.0;
  }
}

// BCS1V (bcs1.F, global frame) and FIXVEL (fixvel.F:366-379, global frame) on one node:
// fixed DOFs get zero acceleration and velocity, imposed ones A = (VIMP - V) / DT12.
template <typename P>
__device__ __forceinline__ void bcs_fixvel(int32_t code, double &ax, double &ay, double &az, double &vx, double &vy,
                                           double &vz, double &rx, double &ry, double &rz, double &wx, double &wy,
                                           double &wz, P vimp, double DT12, double DT1, double DT2, double msn,
                                           double &work, double &pend) {
  if (code & kTX) { ax = 0.0; vx = 0.0; }
  if (code & kTY) { ay = 0.0; vy = 0.0; }
  if (code & kTZ) { az = 0.0; vz = 0.0; }
  if (code & kRX) { rx = 0.0; wx = 0.0; }
  if (code & kRY) { ry = 0.0; wy = 0.0; }
  if (code & kRZ) { rz = 0.0; wz = 0.0; }
  // the imposed DOF's work (fixvel.F:376-378, 834, 837): DW = FOURTH*MS*AXI*WEIGHT*(A*DT12 +
  // TWO*V)*(A - AOLD) with AXI = WEIGHT = 1; DT1*DW enters WFEXT now, DT2*DW (kept per node in
  // pend) at the next cycle, before the DOF's new work (fixvel.F:342-344)
  const double carried = pend;
  work = work + carried;
  pend = 0.0;
  auto imposed = [&](double &a, double v, double vi) {
    const double aold = a;
    a = (vi - v) / DT12;
    const double dw = ((1.0 / 4.0) * msn) * (a * DT12 + 2.0 * v) * (a - aold);
    work = work + DT1 * dw;
    pend = pend + DT2 * dw;
  };
  if (code & kIX) imposed(ax, vx, vimp[0]);
  if (code & kIY) imposed(ay, vy, vimp[1]);
  if (code & kIZ) imposed(az, vz, vimp[2]);
}

// The DT2 pipeline (resol.F:2826, 6125-6129, 6352, 6494-6497): DT2 from the element
// minimum dt2t, the growth limit on DT2OLD and DTMX.
__host__ __device__ __forceinline__ double dt2_pipeline(double dt2t, double dt2old, double dt_growth, double dtmx) {
  double DT2 = kEP06;
  if (dt2t < DT2) DT2 = dt2t;
  const double g = dt_growth * dt2old;
  DT2 = (g < DT2) ? g : DT2;
  DT2 = (dtmx < DT2) ? dtmx : DT2;
  return DT2;
}

// ---------------------------------------------------------------- node kernel
template <typename S>
__global__ __launch_bounds__(kBlock) void k_node_update(const ModelDev *__restrict__ md,
                                                        const Tile *__restrict__ tiles, int parity,
                                                        double dt_growth, int output) {
  const Tile tl = tiles[blockIdx.x];
  const ModelDev &M = md[tl.model];
  const NodePtrs &nd = M.nd;
  const gptr<const S> rowh = to_global<const S>(M.rowh);
  const gptr<const S> rowq = to_global<const S>(M.rowq);
  DevTime *ts = M.ts;
  unsigned long long *dtmin = M.dtmin;
  const double dtmx = M.dtmx;
  const int64_t n = (int64_t)tl.first + threadIdx.x;
  const DevTime prev = load_time(ts + parity);
  const bool in_range = n < nd.nn;
  // ASSPAR4: rows of this node in skyline order (solids, then shells), from zero: the
  // prefix the element tiles summed, then the rest rows in order
  double ax = 0.0, ay = 0.0, az = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
  int32_t dh = 0;
  if (nd.deg_h && in_range) {
    dh = to_global(nd.deg_h)[n];
    if (dh > 0) {
      const gptr<const double> p = to_global(nd.pre_h) + 3 * n;
      ax = p[0]; ay = p[1]; az = p[2];
      const gptr<const int32_t> rs = to_global(nd.rs_h);
      const int32_t r1 = rs[n + 1];
      for (int32_t r = rs[n]; r < r1; ++r) {
        const gptr<const S> w = rowh + (int64_t)r * kHexRowWords;
        ax = ax + double(w[0]); ay = ay + double(w[1]); az = az + double(w[2]);
      }
    }
  }
  bool rot = false; // node carries rotational DOF work this cycle
  if (nd.deg_q && in_range) {
    const int32_t dq = to_global(nd.deg_q)[n];
    rot = dq > 0;
    if (dq > 0) {
      if (dh == 0) { // the node's sum starts with its shell rows: their prefix
        const gptr<const v2d> p = (gptr<const v2d>)(to_global(nd.pre_q) + 6 * n);
        const v2d p0 = p[0], p1 = p[1], p2 = p[2];
        ax = p0.x; ay = p0.y; az = p1.x; rx = p1.y; ry = p2.x; rz = p2.y;
      }
      const gptr<const int32_t> rs = to_global(nd.rs_q);
      const int32_t r1 = rs[n + 1];
      for (int32_t r = rs[n]; r < r1; ++r) {
        const gptr<const S> w = rowq + (int64_t)r * kQ4RowWords;
        ax = ax + double(w[0]); ay = ay + double(w[1]); az = az + double(w[2]);
        rx = rx + double(w[3]); ry = ry + double(w[4]); rz = rz + double(w[5]);
      }
    }
  }
  if (prev.done) {
    // a stopped model keeps its state; its scalars are carried to the other parity
    if (n == 0) {
      ts[parity ^ 1] = prev;
      dtmin[parity ^ 1] = dbits(kEP06);
    }
    return;
  }
  // DT2 pipeline (resol.F:2721-2722, 2826, 6125-6129, 6352, 6494-6497)
  const double DT1 = prev.dt2;
  const double DT2 = dt2_pipeline(bitsd(to_global<const unsigned long long>(dtmin)[parity]), prev.dt2old,
                                  dt_growth, dtmx);
  const double DT12 = (1.0 / 2.0) * (DT1 + DT2);
  if (n == 0) {
    DevTime nx;
    nx.tt = prev.tt + DT2;
    nx.dt1 = DT1;
    nx.dt2 = DT2;
    nx.dt12 = DT12;
    nx.dt2old = DT2;
    nx.ncycle = prev.ncycle + 1;
    // a time step <= 0 stops the run (resol.F:8575, 9599)
    nx.mstop = prev.mstop | *M.flags | (DT2 <= 0.0 ? kMstopTimeStep : 0);
    nx.done = (nx.mstop != 0 || nx.tt >= M.tend) ? 1 : 0;
    ts[parity ^ 1] = nx;
    dtmin[parity ^ 1] = dbits(kEP06);
  }
  if (!in_range) return;
  // ACCELE (accele.F:65-80, 97-109)
  const double msn = to_global(nd.ms)[n];
  accele3(ax, ay, az, msn);
  const int32_t code = to_global(nd.code)[n];
  // A node without shell rows and without rotational BCs has AR = 0, so VELOCITY leaves
  // its VR as it is: its VR and IN are not touched.
  rot = rot || (code & (kRX | kRY | kRZ));
  double wx = 0.0, wy = 0.0, wz = 0.0;
  gptr<v2d> pr = nullptr;
  if (nd.vr && rot) {
    accele3(rx, ry, rz, to_global(nd.in)[n]);
    pr = (gptr<v2d>)to_global(nd.vr + 4 * n);
    const v2d r0 = pr[0], r1 = pr[1];
    wx = r0.x; wy = r0.y; wz = r1.x;
  }
  const gptr<v2d> p = (gptr<v2d>)to_global(nd.xv + 6 * n);
  const v2d pa = p[0], pb = p[1], pc = p[2];
  double x = pa.x, y = pa.y, z = pb.x, vx = pb.y, vy = pc.x, vz = pc.y;
  if (code) {
    // the node's external work (WFEXT) accumulates in wext[n] when its DOFs are imposed
    const bool imp = (code & (kIX | kIY | kIZ)) && nd.wext;
    double work = imp ? to_global(nd.wext)[n] : 0.0, pend = imp ? to_global(nd.wext)[nd.nn + n] : 0.0;
    bcs_fixvel(code, ax, ay, az, vx, vy, vz, rx, ry, rz, wx, wy, wz, to_global(nd.vimp) + 3 * n, DT12, DT1, DT2,
               msn, work, pend);
    if (imp) {
      to_global(nd.wext)[n] = work;
      to_global(nd.wext)[nd.nn + n] = pend;
    }
  }
  if (output && M.ecrit) {
    // ECRIT's kinetic energies at its output point, between FIXVEL and VELOCITY
    // (sortie_main.F:368 from resol.F:8523): V + DT05*A with DT05 = HALF*DT1 (ecrit.F:193,
    // 221-225, 319-324; MAS = MS*WEIGHT_MD, WEIGHT_MD = 1 on one domain)
    const double DT05 = (1.0 / 2.0) * DT1;
    const double ux = vx + DT05 * ax, uy = vy + DT05 * ay, uz = vz + DT05 * az;
    const gptr<double> ke = to_global(M.ecrit);
    ke[n] = (ux * ux + uy * uy + uz * uz) * (1.0 / 2.0) * msn;
    double kr = 0.0;
    if (nd.vr) {
      double qx = wx, qy = wy, qz = wz, inn = 0.0;
      if (!pr) { // rotational work skipped: AR = 0, VR as stored
        const gptr<const v2d> q = (gptr<const v2d>)to_global(nd.vr + 4 * n);
        const v2d q0 = q[0], q1 = q[1];
        qx = q0.x; qy = q0.y; qz = q1.x;
      }
      inn = to_global(nd.in)[n];
      const double sx = qx + DT05 * rx, sy = qy + DT05 * ry, sz = qz + DT05 * rz;
      kr = (sx * sx + sy * sy + sz * sz) * (1.0 / 2.0) * inn;
    }
    ke[nd.nn + n] = kr;
    // XMOMT, YMOMT, ZMOMT (ecrit.F:226-228)
    ke[2 * nd.nn + n] = ux * msn;
    ke[3 * nd.nn + n] = uy * msn;
    ke[4 * nd.nn + n] = uz * msn;
    // the time history's nodal values at the same point (thnod.F:59-97): D and X before
    // DEPLA, V after BCS10, A after FIXVEL, VR and AR (zero without rotational DOFs)
    const int32_t slot = M.th_slot ? to_global(M.th_slot)[n] : -1;
    if (slot >= 0) {
      const gptr<double> t = to_global(M.th_node) + 18 * (int64_t)slot;
      const gptr<const double> dn = to_global(nd.d) + 3 * n;
      double qx = 0.0, qy = 0.0, qz = 0.0;
      if (nd.vr) {
        qx = wx; qy = wy; qz = wz;
        if (!pr) {
          const gptr<const v2d> q = (gptr<const v2d>)to_global(nd.vr + 4 * n);
          const v2d q0 = q[0], q1 = q[1];
          qx = q0.x; qy = q0.y; qz = q1.x;
        }
      }
      t[0] = dn[0]; t[1] = dn[1]; t[2] = dn[2];
      t[3] = vx; t[4] = vy; t[5] = vz;
      t[6] = ax; t[7] = ay; t[8] = az;
      t[9] = qx; t[10] = qy; t[11] = qz;
      t[12] = rx; t[13] = ry; t[14] = rz;
      t[15] = x; t[16] = y; t[17] = z;
    }
  }
  // VELOCITY (velocity.F:58-86)
  vx = vx + DT12 * ax; vy = vy + DT12 * ay; vz = vz + DT12 * az;
  if (pr) {
    wx = wx + DT12 * rx; wy = wy + DT12 * ry; wz = wz + DT12 * rz;
    pr[0] = v2d{wx, wy};
    pr[1] = v2d{wz, 0.0};
  }
  // DEPLA (displacement.F:108-118, IRESP=0 form in FP64)
  const gptr<double> dd = to_global(nd.d) + 3 * n;
  double vdt;
  vdt = DT2 * vx; dd[0] = dd[0] + vdt; x = x + vdt;
  vdt = DT2 * vy; dd[1] = dd[1] + vdt; y = y + vdt;
  vdt = DT2 * vz; dd[2] = dd[2] + vdt; z = z + vdt;
  p[0] = v2d{x, y};
  p[1] = v2d{z, vx};
  p[2] = v2d{vy, vz};
}

// ---------------------------------------------------------------- energies (fixed-order)
// ENINT, ENCIN, ENROT, WFEXT, XMOMT, YMOMT, ZMOMT
constexpr int kEnergySums = 7;
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
__global__ __launch_bounds__(kBlock) void k_energy_partial(const S *heint, const S *hvol0, int64_t neh,
                                                           const S *qeint, int64_t neq, const double *xv,
                                                           const double *vr, const double *ms, const double *in,
                                                           const double *ecrit, const double *wext, int64_t nn,
                                                           int64_t per_block, double *partial) {
  double ie = 0.0, ke = 0.0, kr = 0.0, we = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  const int64_t b0 = (int64_t)blockIdx.x * per_block;
  for (int64_t j = b0 + threadIdx.x; j < b0 + per_block; j += kBlock) {
    if (j < nn && wext) we = we + wext[j];
    if (j < neh) ie = ie + double(heint[j]) * double(hvol0[j]);
    if (j < neq) ie = ie + (double(qeint[j]) + double(qeint[neq + j]));
    if (j < nn && ecrit) {
      // per-node terms written by the node kernel on an output cycle
      ke = ke + ecrit[j];
      kr = kr + ecrit[nn + j];
      mx = mx + ecrit[2 * nn + j];
      my = my + ecrit[3 * nn + j];
      mz = mz + ecrit[4 * nn + j];
    } else if (j < nn) {
      const double *p = xv + 6 * j;
      ke = ke + ms[j] * (p[3] * p[3] + p[4] * p[4] + p[5] * p[5]);
      if (vr) {
        const double *w = vr + 4 * j;
        kr = kr + in[j] * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      }
    }
  }
  const double v[kEnergySums] = {block_sum(ie), block_sum(ke), block_sum(kr), block_sum(we),
                                 block_sum(mx), block_sum(my), block_sum(mz)};
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kEnergySums; ++k) partial[kEnergySums * blockIdx.x + k] = v[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_energy_final(const double *partial, int nb, double kin_factor,
                                                         double *out) {
  double a[kEnergySums];
#pragma unroll
  for (int k = 0; k < kEnergySums; ++k) a[k] = 0.0;
  for (int j = threadIdx.x; j < nb; j += kBlock) {
#pragma unroll
    for (int k = 0; k < kEnergySums; ++k) a[k] = a[k] + partial[kEnergySums * j + k];
  }
#pragma unroll
  for (int k = 0; k < kEnergySums; ++k) a[k] = block_sum(a[k]);
  if (threadIdx.x == 0) {
    out[0] = a[0];
    out[1] = kin_factor * a[1];
    out[2] = kin_factor * a[2];
    for (int k = 3; k < kEnergySums; ++k) out[k] = a[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_fill(double *p, int64_t n, double v) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j < n) p[j] = v;
}

// The element that controls the time step of an output cycle, as MQVISCB and CDT3 name it
// (NELTS, ITYPTS): with DT2T = EP06 on entry (resol.F:2826), the solids in element order
// replace it when DTX <= DT2T (mqviscb.F:622-630), so the last solid with the smallest DTX;
// then the shells when DT < DT2T (cdt3.F:205-215), so the first shell whose DT is below
// every solid's.  Elements that cannot control the step carry EP06.  w[0], w[1]: the
// minimum of each family (FP64 bits, integer atomicMin); w[2]: the largest solid index at
// the solid minimum (atomicMax); w[3]: the smallest shell index at the shell minimum.
__global__ __launch_bounds__(kBlock) void k_ctrl_min(const double *dt, int64_t ne, unsigned long long *w) {
  double v = kEP06;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < ne; j += (int64_t)gridDim.x * kBlock)
    v = (dt[j] < v) ? dt[j] : v;
  const double b = block_min<kBlock>(v);
  if (threadIdx.x == 0 && b < kEP06) atomicMin(w, dbits(b));
}
__global__ __launch_bounds__(kBlock) void k_ctrl_pick(const double *dt, const int32_t *logical, int64_t ne,
                                                      const unsigned long long *wmin, int last,
                                                      unsigned long long *pick) {
  const double m = bitsd(*wmin);
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < ne; j += (int64_t)gridDim.x * kBlock) {
    if (dt[j] == m && m < kEP06) {
      const unsigned long long e = (unsigned long long)logical[j];
      if (last) atomicMax(pick, e);
      else atomicMin(pick, e);
    }
  }
}

// ---------------------------------------------------------------- time history
// The element values of THSOL (thsol.F:189-200, 511-548) and THCOQ (thcoq.F:296-340, 452-520)
// that the time-history file reads from the element buffers, for the listed elements (storage
// slots): hexes OFF, SIG(6), EINT, RHO, QVIS, VOL0, PLA, EPSD; shells FOR(5), MOM(3), EINT(2),
// OFF, THK and the PLA of each through-thickness point.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_th_elements(const ModelDev *__restrict__ md, const int32_t *hex,
                                                        int64_t nh, const int32_t *q4, int64_t nq, int npt,
                                                        double *hout, double *qout) {
  const ModelDev &M = md[0];
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j < nh) {
    const HexPtrs<S> e = hex_ptrs<S>(M);
    const int64_t i = hex[j];
    const gptr<double> o = to_global(hout) + 13 * j;
    o[0] = double(e.offg[i]);
#pragma unroll
    for (int k = 0; k < 6; ++k) o[1 + k] = double(e.sig[(int64_t)k * e.ne + i]);
    o[7] = double(e.eint[i]); o[8] = double(e.rho[i]); o[9] = double(e.qvis[i]); o[10] = double(e.vol0[i]);
    o[11] = double(e.pla[i]); o[12] = double(e.epsd[i]);
  } else if (j < nh + nq) {
    const Q4Ptrs<S> q = q4_ptrs<S>(M);
    const int64_t i = q4[j - nh];
    const gptr<double> o = to_global(qout) + (12 + npt) * (j - nh);
#pragma unroll
    for (int k = 0; k < 5; ++k) o[k] = double(q.forc[(int64_t)k * q.ne + i]);
#pragma unroll
    for (int k = 0; k < 3; ++k) o[5 + k] = double(q.mom[(int64_t)k * q.ne + i]);
    o[8] = double(q.eint[i]); o[9] = double(q.eint[q.ne + i]);
    o[10] = double(q.offg[i]); o[11] = double(q.thk[i]);
    for (int k = 0; k < npt; ++k) o[12 + k] = double(q.pla[(int64_t)k * q.ne + i]);
  }
}

// EHOUR: the tiles' accumulators, solid tiles then shell tiles, summed in a fixed order.
__global__ __launch_bounds__(kBlock) void k_th_ehour(const double *h, int64_t nh, const double *q, int64_t nq,
                                                     double *out) {
  double v = 0.0;
  for (int64_t j = threadIdx.x; j < nh + nq; j += kBlock) v = v + (j < nh ? h[j] : q[j - nh]);
  v = block_sum(v);
  if (threadIdx.x == 0) *out = v;
}

// ---------------------------------------------------------------- group drop-in kernels
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
  auto write_row = [&](int k, double fx, double fy, double fz) {
    const int64_t K = (int64_t)iads[8 * col + k] - 1;
    fsky[8 * K + 0] = fx;
    fsky[8 * K + 1] = fy;
    fsky[8 * K + 2] = fz;
    fsky[8 * K + 6] = o.sti4;
  };
  hex_element<double, double>(xd, yd, zd, vx, vy, vz, st, c, dt1, o, write_row, [] {});
#pragma unroll
  for (int k = 0; k < 6; ++k) sig[(int64_t)k * nel + i] = st.sig[k];
  eint[i] = st.eint; rho[i] = st.rho; pla[i] = st.pla;
  epsd[i] = st.epsd; off[i] = st.offg; qvis[i] = st.qvis;
  dtx_out[i] = o.dtx;
  mstop_out[i] = o.mstop;
}

// Per-point state of the CFORC3 drop-in: LBUF arrays of one group, SIG(NEL,5) per point.
struct GroupIpIO {
  double *sig, *pla, *epsd_ip, *gstr, *hourg, *eint, *forc, *mom;
  int nel, i;
  __device__ __forceinline__ void late_load(Q4Late<double> &L) const {
#pragma unroll
    for (int k = 0; k < 5; ++k) L.forc[k] = forc[(int64_t)k * nel + i];
#pragma unroll
    for (int k = 0; k < 3; ++k) L.mom[k] = mom[(int64_t)k * nel + i];
#pragma unroll
    for (int k = 0; k < 8; ++k) L.gstr[k] = gstr[(int64_t)k * nel + i];
#pragma unroll
    for (int k = 0; k < 3; ++k) L.hourg[k] = hourg[(int64_t)k * nel + i];
    L.eint[0] = eint[i];
    L.eint[1] = eint[nel + i];
  }
  __device__ __forceinline__ void late_store(const Q4Late<double> &L) const {
#pragma unroll
    for (int k = 0; k < 5; ++k) forc[(int64_t)k * nel + i] = L.forc[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) mom[(int64_t)k * nel + i] = L.mom[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) gstr[(int64_t)k * nel + i] = L.gstr[k];
    eint[i] = L.eint[0];
    eint[nel + i] = L.eint[1];
  }
  __device__ __forceinline__ void hourg_store(const double (&h)[5]) const {
#pragma unroll
    for (int k = 0; k < 5; ++k) hourg[(int64_t)k * nel + i] = h[k];
  }
  __device__ __forceinline__ void load(int ip, double (&s)[5], double &p, double &e) const {
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] = sig[((int64_t)ip * 5 + k) * nel + i];
    p = pla[(int64_t)ip * nel + i];
    e = epsd_ip[(int64_t)ip * nel + i];
  }
  __device__ __forceinline__ void store(int ip, const double (&s)[5], double p, double e) const {
#pragma unroll
    for (int k = 0; k < 5; ++k) sig[((int64_t)ip * 5 + k) * nel + i] = s[k];
    pla[(int64_t)ip * nel + i] = p;
    epsd_ip[(int64_t)ip * nel + i] = e;
  }
};

// CFORC3 on one group with the reference host layouts (X,V,VR(3,*), IXC(7,*), GBUF arrays of
// the group, LBUF SIG(NEL,5) / PLA(NEL) / EPSD(NEL) per point, FSKY(8,*), IADC(4,*) 1-based).
__global__ __launch_bounds__(kBlock) void k_cforc3_group(int nel, int64_t nft, const double *x, const double *v,
                                                         const double *vr, const int32_t *ixc, double *offg,
                                                         double *thk, const double *thke, double *forc,
                                                         double *mom, double *gstr, double *hourg, double *eint,
                                                         double *epsd, double *sig, double *pla, double *epsd_ip,
                                                         double *fsky, const int32_t *iadc, Q4Consts c, double dt1,
                                                         double *dt_out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nel) return;
  const int64_t col = nft + i;
  double xg[4], yg[4], zg[4], vl[4][3], vrl[4][3];
  int32_t nc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nc[k] = ixc[7 * col + 1 + k];
    const int64_t n = (int64_t)nc[k] - 1;
    xg[k] = x[3 * n]; yg[k] = x[3 * n + 1]; zg[k] = x[3 * n + 2];
#pragma unroll
    for (int d = 0; d < 3; ++d) { vl[k][d] = v[3 * n + d]; vrl[k][d] = vr[3 * n + d]; }
  }
  Q4Elem<double> st;
  st.offg = offg[i]; st.thk = thk[i]; st.thke = thke[i]; st.epsd = epsd[i];
  GroupIpIO io{sig, pla, epsd_ip, gstr, hourg, eint, forc, mom, nel, i};
  Q4Out<double> o;
  auto nload = [&](int k, double(&p)[3], double(&w)[3]) {
    const int64_t n = (int64_t)nc[k] - 1;
    p[0] = x[3 * n]; p[1] = x[3 * n + 1]; p[2] = x[3 * n + 2];
    w[0] = v[3 * n]; w[1] = v[3 * n + 1]; w[2] = v[3 * n + 2];
  };
  q4_element<double>(xg, yg, zg, vl, vrl, nc[2] == nc[3], st, c, dt1, io, nload, o);
  offg[i] = st.offg; thk[i] = st.thk; epsd[i] = st.epsd;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double *f = fsky + 8 * ((int64_t)iadc[4 * col + k] - 1);
    f[0] = -o.f[0][k]; f[1] = -o.f[1][k]; f[2] = -o.f[2][k];
    f[3] = -o.m[0][k]; f[4] = -o.m[1][k]; f[5] = -o.m[2][k];
    f[6] = o.sti; f[7] = 0.0;
  }
  dt_out[i] = o.dt;
}

// Nodal half of a cycle on the reference host layouts: ASSPAR4 over FSKY(8,LSKY) rows
// listed by ADSKY (1-based CSR, FILLCNE order), ACCELE, BCS10, FIXVEL, VELOCITY, DEPLA on
// X/V/VR/D(3,NUMNOD).  DT12 and DT2 are the cycle's, from the DT2 pipeline on the host.
__global__ __launch_bounds__(kBlock) void k_nodal_dropin(int64_t nn, const int64_t *adsky, const double *fsky,
                                                         const double *ms, const double *in, const int32_t *code,
                                                         const double *vimp, double *x, double *v, double *vr,
                                                         double *d, double *stifn, double *stifr, double DT12,
                                                         double DT2) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= nn) return;
  // ASSPAR4 (asspar4.F:164-180)
  double ax = 0.0, ay = 0.0, az = 0.0, rx = 0.0, ry = 0.0, rz = 0.0, sn = 0.0, sr = 0.0;
  for (int64_t k = adsky[n] - 1; k < adsky[n + 1] - 1; ++k) {
    const double *f = fsky + 8 * k;
    ax = ax + f[0]; ay = ay + f[1]; az = az + f[2];
    rx = rx + f[3]; ry = ry + f[4]; rz = rz + f[5];
    sn = sn + f[6]; sr = sr + f[7];
  }
  if (stifn) stifn[n] = sn;
  if (stifr) stifr[n] = sr;
  accele3(ax, ay, az, ms[n]);
  const bool rot = vr != nullptr;
  double wx = 0.0, wy = 0.0, wz = 0.0;
  if (rot) {
    accele3(rx, ry, rz, in[n]);
    wx = vr[3 * n]; wy = vr[3 * n + 1]; wz = vr[3 * n + 2];
  }
  double vx = v[3 * n], vy = v[3 * n + 1], vz = v[3 * n + 2];
  const int32_t cd = code ? code[n] : 0;
  if (cd) {
    double work = 0.0, pend = 0.0; // the drop-in reports no external work (the caller's FIXVEL keeps WFEXT)
    bcs_fixvel(cd, ax, ay, az, vx, vy, vz, rx, ry, rz, wx, wy, wz, vimp + 3 * n, DT12, 0.0, 0.0, ms[n], work, pend);
  }
  // VELOCITY (velocity.F:58-86)
  vx = vx + DT12 * ax; vy = vy + DT12 * ay; vz = vz + DT12 * az;
  v[3 * n] = vx; v[3 * n + 1] = vy; v[3 * n + 2] = vz;
  if (rot) {
    wx = wx + DT12 * rx; wy = wy + DT12 * ry; wz = wz + DT12 * rz;
    vr[3 * n] = wx; vr[3 * n + 1] = wy; vr[3 * n + 2] = wz;
  }
  // DEPLA (displacement.F:108-118)
  double vdt;
  vdt = DT2 * vx; d[3 * n] = d[3 * n] + vdt; x[3 * n] = x[3 * n] + vdt;
  vdt = DT2 * vy; d[3 * n + 1] = d[3 * n + 1] + vdt; x[3 * n + 1] = x[3 * n + 1] + vdt;
  vdt = DT2 * vz; d[3 * n + 2] = d[3 * n + 2] + vdt; x[3 * n + 2] = x[3 * n + 2] + vdt;
}

// ---------------------------------------------------------------- host helpers
double host_onep333() {
  volatile double zep3 = 3.0 / 10.0, zep03 = 3.0 / 100.0, threeem3 = 3.0 / 1000.0;
  volatile double onep3 = 1.0 + zep3;
  volatile double onep33 = onep3 + zep03;
  return onep33 + threeem3;
}

HexConsts make_hex_consts(const or_hex_params &p) {
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

int check_hex_params(const or_hex_params &p) {
  if (p.fisokin != 0.0) return fail(OR_ERR_ARG, "law 2 kinematic hardening (FISOKIN>0) is not on this path");
  if (p.iform != 0) return fail(OR_ERR_ARG, "law 2 Zerilli form (IFORM=1) is not on this path");
  // VP=1 (plastic strain rate) filters EPSD after the return (m2law.F:536-544) with the Starter's
  // default cut-off frequency; that update is not on this path
  if (p.vp != 2 && p.vp != 3) return fail(OR_ERR_ARG, "law 2 hexes: VP must be 2 (total) or 3 (deviatoric strain rate)");
  if (!(p.rho0 > 0.0) || !(p.dtfac > 0.0)) return fail(OR_ERR_ARG, "rho0 and dtfac must be positive");
  return OR_OK;
}

// Tables of COQINI (coqini.F:46-152), single-precision DATA literals.
const float kZ0[11][11] = {
    {0.f}, {-.5f, 0.5f}, {-.5f, 0.f, 0.5f}, {-.5f, -.1666667f, 0.1666667f, 0.5f}, {-.5f, -.25f, 0.f, 0.25f, 0.5f},
    {-.5f, -.3f, -.1f, 0.1f, 0.3f, 0.5f}, {-.5f, -.3333333f, -.1666667f, 0.0f, 0.1666667f, 0.3333333f, 0.5f},
    {-.5f, -.3571429f, -.2142857f, -.0714286f, 0.0714286f, 0.2142857f, 0.3571429f, 0.5f},
    {-.5f, -.375f, -.25f, -.125f, 0.0f, 0.125f, 0.25f, 0.375f, 0.5f},
    {-.5f, -.3888889f, -.2777778f, -.1666667f, -.0555555f, 0.0555555f, 0.1666667f, 0.2777778f, 0.3888889f, 0.5f},
    {-.5f, -.4f, -.3f, -.2f, -.1f, 0.f, 0.1f, 0.2f, 0.3f, 0.4f, 0.5f}};
const float kWF[11][11] = {
    {1.f}, {0.5f, 0.5f}, {0.25f, 0.5f, 0.25f}, {0.1666667f, 0.3333333f, 0.3333333f, 0.1666667f},
    {0.125f, 0.25f, 0.25f, 0.25f, 0.125f}, {0.1f, 0.2f, 0.2f, 0.2f, 0.2f, 0.1f},
    {0.0833333f, 0.1666667f, 0.1666667f, 0.1666667f, 0.1666667f, 0.1666667f, 0.0833333f},
    {0.0714286f, 0.1428571f, 0.1428571f, 0.1428571f, 0.1428571f, 0.1428571f, 0.1428571f, 0.0714286f},
    {0.0625f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.0625f},
    {0.0555556f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f, 0.1111111f,
     0.0555556f},
    {0.05f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.05f}};
const float kWM[11][11] = {
    {0.f}, {-.0833333f, 0.0833333f}, {-.0833333f, 0.f, 0.0833333f}, {-.0648148f, -.0555556f, 0.0555556f, 0.0648148f},
    {-.0520833f, -.0625f, 0.f, 0.0625f, 0.0520833f},
    {-.0433333f, -.0600000f, -.0200000f, 0.0200000f, 0.0600000f, 0.0433333f},
    {-.0370370f, -.0555556f, -.0277778f, 0.0000000f, 0.0277778f, 0.0555556f, 0.0370370f},
    {-.0323129f, -.0510204f, -.0306122f, -.0102041f, 0.0102041f, 0.0306122f, 0.0510204f, 0.0323129f},
    {-.0286458f, -.0468750f, -.0312500f, -.0156250f, 0.0000000f, 0.0156250f, 0.0312500f, 0.0468750f, 0.0286458f},
    {-.0257202f, -.0432099f, -.0308642f, -.0185185f, -.0061728f, 0.0061728f, 0.0185185f, 0.0308642f, 0.0432099f,
     0.0257202f},
    {-.0233333f, -.0400000f, -.0300000f, -.0200000f, -.0100000f, 0.0000000f, 0.0100000f, 0.0200000f, 0.0300000f,
     0.0400000f, 0.0233333f}};

int make_q4_consts(const or_q4_params &p, Q4Consts &c) {
  if (p.npt < 2 || p.npt > 10) return fail(OR_ERR_ARG, "shell NPT must be in 2..10 (NPT=1 uses MHVIS3, not on this path)");
  if (p.ipla < 0 || p.ipla > 2) return fail(OR_ERR_ARG, "Iplas must be 0, 1 or 2");
  if (p.iform != 0) return fail(OR_ERR_ARG, "law 2 Zerilli form (IFORM=1) is not on this path");
  if (p.vp != 2) return fail(OR_ERR_ARG, "law 2 shells: only VP=2 (total strain rate) is on this path");
  if (!(p.rho > 0.0) || !(p.young > 0.0) || !(p.dtfac > 0.0) || !(p.ssp > 0.0))
    return fail(OR_ERR_ARG, "rho, young, ssp and dtfac must be positive");
  std::memset(&c, 0, sizeof(c));
  c.rho = p.rho; c.young = p.young; c.nu = p.nu; c.shear = p.shear; c.ssp = p.ssp; c.ca = p.ca; c.cb = p.cb;
  c.cn = p.cn; c.epmx = p.epmx; c.ymax = p.ymax; c.cc = p.cc; c.epdr = p.epdr; c.m_exp = p.m_exp;
  c.asrate = p.asrate; c.icc = p.icc; c.vp = p.vp; c.israte = p.israte; c.npt = p.npt; c.ipla = p.ipla;
  c.ithk = p.ithk; c.ishear = p.ishear; c.hm = p.hm; c.hf = p.hf; c.hr = p.hr; c.fsh = p.fsh; c.dtfac = p.dtfac;
  c.dtmin = p.dtmin;
  {
    volatile double a1 = p.young / (1.0 - p.nu * p.nu);
    c.a1 = a1;
    c.a2 = a1 * p.nu;
  }
  {
    volatile double zep1 = 1.0 / 10.0, zep01 = 1.0 / 100.0, zep03 = 3.0 / 100.0;
    volatile double zep09 = zep1 - zep01;
    volatile double zep19 = zep1 + zep09;
    c.zep81 = 1.0 - zep19;
    volatile double zep07 = zep1 - zep03, twoem3 = 2.0 / 1000.0;
    volatile double ep04 = (100.0 * 10.0) * 10.0, ep05 = ep04 * 10.0, ep06 = ep05 * 10.0;
    volatile double em04 = 1.0 / ep04, em05 = 1.0 / ep05, em06 = 1.0 / ep06;
    volatile double sixem5 = em04 - 4.0 / ep05, nineem6 = em05 - em06;
    volatile double s = zep07 + twoem3;
    s = s + em04;
    s = s + sixem5;
    c.zep072169 = s + nineem6;
    volatile double two = 2.0;
    c.sr2d2 = std::sqrt(two) * 0.5;
  }
  c.srh1 = std::sqrt(p.hm); c.srh2 = std::sqrt(p.hf); c.srh3 = std::sqrt(p.hr);
  for (int k = 0; k < p.npt; ++k) {
    c.z0[k] = (double)kZ0[p.npt - 1][k];
    c.wf[k] = (double)kWF[p.npt - 1][k];
    c.wm[k] = (double)kWM[p.npt - 1][k];
  }
  return OR_OK;
}

template <typename T> int dalloc(T **p, int64_t n, int64_t &acc) {
  if (n <= 0) { *p = nullptr; return OR_OK; }
  OR_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)n));
  acc += (int64_t)(sizeof(T) * (size_t)n);
  return OR_OK;
}

} // namespace

#ifndef OR_ILP_TU
// ============================================================== model object
struct or_model {
  int device = 0;
  int prec = OR_PREC_FP64;
  int64_t nn = 0, neh = 0, neq = 0;
  int npt = 0;
  int64_t lsky_h = 0, lsky_q = 0;
  int64_t bytes = 0;
  hipStream_t stream = nullptr;
  bool have_hparams = false, have_qparams = false, have_nodes = false, have_hex = false, have_q4 = false,
       have_sky = false;
  HexConsts hc{};
  Q4Consts qc{};
  // nodes
  double *xv = nullptr, *vr = nullptr, *d = nullptr, *ms = nullptr, *in = nullptr, *vimp = nullptr;
  int32_t *code = nullptr, *deg_h = nullptr, *deg_q = nullptr, *rs_h = nullptr, *rs_q = nullptr;
  double *preh = nullptr, *preq = nullptr;
  int64_t sky_bytes = 0;               // rest rows and prefix lists, sized at upload_skyline
  // element storage order of each family: slot -> element (or_assembly.h tile_order); empty
  // until the first element upload, fixed after it
  std::vector<int32_t> hperm, qperm;
  int64_t asm_stats[4] = {0, 0, 0, 0}; // prefixes and rest rows of the hex and shell plans
  // hex (storage type depends on precision)
  int32_t *hconn = nullptr, *hrest = nullptr, *hfr_start = nullptr, *hfr_node = nullptr;
  uint16_t *hfr_src = nullptr;
  void *hsig = nullptr, *heint = nullptr, *hrho = nullptr, *hvol0 = nullptr, *hpla = nullptr, *hepsd = nullptr,
       *hoffg = nullptr, *hqvis = nullptr, *rowh = nullptr;
  // Q4
  int32_t *qconn = nullptr, *qrest = nullptr, *qfr_start = nullptr, *qfr_node = nullptr;
  uint16_t *qfr_src = nullptr;
  void *qoffg = nullptr, *qthk = nullptr, *qthke = nullptr, *qforc = nullptr, *qmom = nullptr, *qgstr = nullptr,
       *qhourg = nullptr, *qeint = nullptr, *qepsd = nullptr, *qsig = nullptr, *qpla = nullptr, *qepsdip = nullptr,
       *rowq = nullptr;
  // scalars
  DevTime *ts = nullptr;               // [2]
  unsigned long long *dtmin = nullptr; // [2]
  int *flags = nullptr;
  double *energy = nullptr;            // [kEnergySums * (1 + nb)]
  double *ecrit = nullptr;             // [5][nn] per-node ECRIT terms, allocated at the first output cycle
  double momentum[3] = {0.0, 0.0, 0.0}; // XMOMT, YMOMT, ZMOMT of the last output cycle
  // time history: the rows of its nodes (th_slot), their output-cycle values and the sampled
  // elements (th_buf: nodes [n][18], hexes [n][13], shells [n][12 + npt]), EHOUR per tile
  int32_t *th_slot = nullptr, *th_hex = nullptr, *th_q4 = nullptr;
  double *th_buf = nullptr, *ehou_h = nullptr, *ehou_q = nullptr;
  int64_t th_nn = 0, th_nh = 0, th_nq = 0;
  double *wext = nullptr;              // [2][nn] external work per node (and its deferred part), when
                                       // the model has imposed DOFs
  // the controlling element of output cycles: per-element candidates (storage order), the
  // element index of each storage slot, and the scan's words; allocated at the first output cycle
  double *dtel_h = nullptr, *dtel_q = nullptr;
  int32_t *lidx_h = nullptr, *lidx_q = nullptr;
  unsigned long long *ctrl = nullptr;  // [4]
  int64_t parity = 0;
  // what the kernels see (host copy, device copy) and the model's own tile lists
  ModelDev hdev{};
  ModelDev *dev = nullptr;
  bool dev_dirty = true;
  Tile *tiles_h = nullptr, *tiles_q = nullptr, *tiles_n = nullptr;
  int32_t nt_h = 0, nt_q = 0, nt_n = 0;
  int var_h = 0, var_q = 0;            // element kernel variants (tiles_h / tiles_q follow their block)
  or_batch *batch = nullptr;           // set while the model belongs to a batch
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;          // ring of 4 events per cycle (hex | q4 | node)
  int ring_used = 0;
  double t_hex = 0.0, t_q4 = 0.0, t_node = 0.0;
  int64_t t_cycles = 0;
};

// A set of resident models advanced together: every cycle launches each kernel once
// per precision class over the tile lists of all live members (or_batch_*).
struct or_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<or_model *> members; // slot order
  std::vector<double> tend;        // per slot
  std::vector<char> live;          // per slot: still cycling (cleared by or_batch_active)
  ModelDev *md = nullptr;          // device array of the slots' kernel views
  int64_t md_cap = 0;
  Tile *tiles[2][3] = {};          // [precision][hex, q4, node]
  int var[2][2] = {};              // element kernel variants [precision][hex, q4]
  int32_t nt[2][3] = {};
  int64_t tcap[2][3] = {};
  bool dirty = true;
  int64_t parity = 0;
  int64_t cycles = 0;              // cycles launched
};

// Wait for every stream that may hold work of the model (its own, and its batch's).
static hipError_t quiesce(or_model *m) {
  if (m->batch) {
    hipError_t e = hipStreamSynchronize(m->batch->stream);
    if (e != hipSuccess) return e;
  }
  return hipStreamSynchronize(m->stream);
}

// Pointers or parameters of the model changed: its kernel views must be rebuilt.
static void mark_changed(or_model *m) {
  m->dev_dirty = true;
  if (m->batch) m->batch->dirty = true;
}

static size_t ssize(const or_model *m) { return m->prec == OR_PREC_FP64 ? sizeof(double) : sizeof(float); }

// Element tiles of a model alone, for its current variants.
static int put_elem_tiles(or_model *m) {
  for (int f = 0; f < 2; ++f) {
    const int64_t ne = f == 0 ? m->neh : m->neq;
    if (ne == 0) continue;
    std::vector<Tile> tl;
    for (int64_t first = 0; first < ne; first += kTile) tl.push_back(Tile{0, (int32_t)first});
    OR_HIP(hipMemcpy(f == 0 ? m->tiles_h : m->tiles_q, tl.data(), sizeof(Tile) * tl.size(), hipMemcpyHostToDevice));
    (f == 0 ? m->nt_h : m->nt_q) = (int32_t)tl.size();
  }
  return OR_OK;
}

// Entry points: C linkage comes from their declarations in or_gpu.h.

int or_version(void) { return 200; }

const char *or_last_error(void) { return g_err.c_str(); }

int or_device_count(int *count) {
  if (!count) return fail(OR_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) { *count = 0; return fail(OR_ERR_NODEV, hipGetErrorString(e)); }
  *count = n;
  return OR_OK;
}

int or_device_memory(int device, int64_t *free_bytes, int64_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(OR_ERR_ARG, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(OR_ERR_NODEV, "no HIP device");
  if (device < 0 || device >= ndev) return fail(OR_ERR_ARG, "device index out of range");
  OR_HIP(hipSetDevice(device));
  size_t f = 0, t = 0;
  OR_HIP(hipMemGetInfo(&f, &t));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return OR_OK;
}

int or_model_create(or_model **out, int device, int64_t nn, int64_t ne_hex, int64_t ne_q4, int32_t npt,
                    int precision) {
  if (!out) return fail(OR_ERR_ARG, "null out");
  *out = nullptr;
  if (nn <= 0 || ne_hex < 0 || ne_q4 < 0) return fail(OR_ERR_ARG, "nn must be > 0, ne_hex and ne_q4 >= 0");
  if (ne_q4 > 0 && (npt < 2 || npt > 10)) return fail(OR_ERR_ARG, "shell NPT must be in 2..10");
  if (precision != OR_PREC_FP64 && precision != OR_PREC_SPLIT) return fail(OR_ERR_ARG, "bad precision");
  if (8 * ne_hex >= (int64_t)1 << 31 || 4 * ne_q4 >= (int64_t)1 << 31 || nn >= (int64_t)1 << 31)
    return fail(OR_ERR_ARG, "skyline rows and node ids must fit in int32");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(OR_ERR_NODEV, "no HIP device");
  if (device < 0 || device >= ndev) return fail(OR_ERR_ARG, "device index out of range");
  OR_HIP(hipSetDevice(device));
  or_model *m = new or_model();
  m->device = device;
  m->prec = precision;
  m->nn = nn;
  m->neh = ne_hex;
  m->neq = ne_q4;
  m->npt = ne_q4 > 0 ? npt : 0;
  m->lsky_h = 8 * ne_hex;
  m->lsky_q = 4 * ne_q4;
  int rc = OR_OK;
  auto A = [&](auto **p, int64_t n) { if (rc == OR_OK) rc = dalloc(p, n, m->bytes); };
  const size_t ss = ssize(m);
  auto AS = [&](void **p, int64_t n) {
    if (rc != OR_OK || n <= 0) return;
    if (hipMalloc(p, ss * (size_t)n) != hipSuccess) { rc = fail(OR_ERR_HIP, "hipMalloc state"); return; }
    m->bytes += (int64_t)(ss * (size_t)n);
  };
  A(&m->xv, 6 * nn);
  A(&m->d, 3 * nn);
  A(&m->ms, nn);
  A(&m->vimp, 3 * nn);
  A(&m->code, nn);
  if (ne_q4 > 0) {
    A(&m->vr, 4 * nn);
    A(&m->in, nn);
  }
  if (ne_hex > 0) {
    A(&m->deg_h, nn);
    A(&m->rs_h, nn + 1);
    A(&m->preh, 3 * nn);
    A(&m->hfr_start, (ne_hex + kTile - 1) / kTile + 1);
    A(&m->hconn, 8 * ne_hex);
    A(&m->hrest, 8 * ne_hex);
    AS(&m->hsig, 6 * ne_hex);
    AS(&m->heint, ne_hex); AS(&m->hrho, ne_hex); AS(&m->hvol0, ne_hex); AS(&m->hpla, ne_hex);
    AS(&m->hepsd, ne_hex); AS(&m->hoffg, ne_hex); AS(&m->hqvis, ne_hex);
  }
  if (ne_q4 > 0) {
    A(&m->deg_q, nn);
    A(&m->rs_q, nn + 1);
    A(&m->preq, 6 * nn);
    A(&m->qfr_start, (ne_q4 + kTile - 1) / kTile + 1);
    A(&m->qconn, 4 * ne_q4);
    A(&m->qrest, 4 * ne_q4);
    AS(&m->qoffg, ne_q4); AS(&m->qthk, ne_q4); AS(&m->qthke, ne_q4); AS(&m->qepsd, ne_q4);
    AS(&m->qforc, 5 * ne_q4); AS(&m->qmom, 3 * ne_q4); AS(&m->qgstr, 8 * ne_q4); AS(&m->qhourg, 5 * ne_q4);
    AS(&m->qeint, 2 * ne_q4);
    AS(&m->qsig, (int64_t)npt * 5 * ne_q4);
    AS(&m->qpla, (int64_t)npt * ne_q4);
  }
  A(&m->ts, 2);
  A(&m->dtmin, 2);
  A(&m->flags, 1);
  A(&m->dev, 1);
  m->nt_n = (int32_t)((nn + kBlock - 1) / kBlock);
  A(&m->tiles_h, (ne_hex + kTile - 1) / kTile);
  A(&m->tiles_q, (ne_q4 + kTile - 1) / kTile);
  A(&m->tiles_n, m->nt_n);
  const int64_t nb = 1024;
  A(&m->energy, kEnergySums * (1 + nb));
  if (rc == OR_OK && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(OR_ERR_HIP, "hipStreamCreate");
  if (rc == OR_OK) {
    if (hipMemsetAsync(m->d, 0, sizeof(double) * 3 * nn, m->stream) != hipSuccess ||
        hipMemsetAsync(m->code, 0, sizeof(int32_t) * nn, m->stream) != hipSuccess ||
        hipMemsetAsync(m->vimp, 0, sizeof(double) * 3 * nn, m->stream) != hipSuccess ||
        (m->vr && hipMemsetAsync(m->vr, 0, sizeof(double) * 4 * nn, m->stream) != hipSuccess) ||
        hipMemsetAsync(m->flags, 0, sizeof(int), m->stream) != hipSuccess)
      rc = fail(OR_ERR_HIP, "hipMemsetAsync init");
    DevTime t0{};
    t0.dt2old = 1.0e6 / (1.0 + 1.0 / 10.0); // DTINI/ONEP1 (starter contrl.F:1125-1127)
    DevTime tsh[2] = {t0, t0};
    unsigned long long dm[2];
    double big = kEP06;
    std::memcpy(&dm[0], &big, 8);
    dm[1] = dm[0];
    std::vector<Tile> tl;
    auto put_tiles = [&](Tile *dst, int32_t nt) {
      if (rc != OR_OK || nt == 0) return;
      tl.resize((size_t)nt);
      for (int32_t b = 0; b < nt; ++b) tl[(size_t)b] = Tile{0, b * kBlock};
      if (hipMemcpy(dst, tl.data(), sizeof(Tile) * (size_t)nt, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(OR_ERR_HIP, "tile upload");
    };
    put_tiles(m->tiles_n, m->nt_n);
    if (rc == OR_OK) rc = put_elem_tiles(m);
    if (rc == OR_OK && (hipMemcpyAsync(m->ts, tsh, sizeof(tsh), hipMemcpyHostToDevice, m->stream) != hipSuccess ||
                        hipMemcpyAsync(m->dtmin, dm, sizeof(dm), hipMemcpyHostToDevice, m->stream) != hipSuccess ||
                        hipStreamSynchronize(m->stream) != hipSuccess))
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
  if (m->batch) (void)or_batch_remove(m->batch, m);
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void *ptrs[] = {m->xv, m->vr, m->d, m->ms, m->in, m->vimp, m->code, m->deg_h, m->deg_q, m->rs_h, m->rs_q,
                  m->preh, m->preq, m->hconn, m->hrest, m->hfr_start, m->hfr_node, m->hfr_src,
                  m->hsig, m->heint, m->hrho, m->hvol0, m->hpla, m->hepsd, m->hoffg, m->hqvis, m->rowh, m->qconn,
                  m->qrest, m->qfr_start, m->qfr_node, m->qfr_src, m->qoffg, m->qthk, m->qthke, m->qforc, m->qmom,
                  m->qgstr, m->qhourg, m->qeint, m->qepsd,
                  m->qsig, m->qpla, m->qepsdip, m->rowq, m->ts, m->dtmin, m->flags, m->energy, m->dev,
                  m->tiles_h, m->tiles_q, m->tiles_n, m->ecrit, m->wext, m->dtel_h, m->dtel_q, m->lidx_h,
                  m->lidx_q, m->ctrl, m->th_slot, m->th_hex, m->th_q4, m->th_buf, m->ehou_h, m->ehou_q};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : m->ev) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return OR_OK;
}

int or_model_set_hex_params(or_model *m, const or_hex_params *p) {
  if (!m || !p) return fail(OR_ERR_ARG, "null argument");
  int rc = check_hex_params(*p);
  if (rc) return rc;
  m->hc = make_hex_consts(*p);
  m->have_hparams = true;
  mark_changed(m);
  return OR_OK;
}

int or_model_set_q4_params(or_model *m, const or_q4_params *p) {
  if (!m || !p) return fail(OR_ERR_ARG, "null argument");
  if (m->neq > 0 && p->npt != m->npt) return fail(OR_ERR_ARG, "q4 params NPT differs from the model's NPT");
  Q4Consts c;
  int rc = make_q4_consts(*p, c);
  if (rc) return rc;
  // LBUF%EPSD is only distinct from GBUF%EPSD when the strain-rate filter is on (ASRATE < 1)
  if (c.israte > 0 && m->neq > 0 && !m->qepsdip) {
    const size_t ss = ssize(m);
    OR_HIP(hipSetDevice(m->device));
    OR_HIP(hipMalloc(&m->qepsdip, ss * (size_t)(m->npt * m->neq)));
    OR_HIP(hipMemsetAsync(m->qepsdip, 0, ss * (size_t)(m->npt * m->neq), m->stream));
    OR_HIP(quiesce(m));
    m->bytes += (int64_t)(ss * (size_t)(m->npt * m->neq));
  }
  m->qc = c;
  m->have_qparams = true;
  mark_changed(m);
  return OR_OK;
}

int or_model_upload_nodes(or_model *m, const double *x, const double *v, const double *vr, const double *ms,
                          const double *in, const int32_t *code, const double *vimp) {
  if (!m || !x || !v || !ms) return fail(OR_ERR_ARG, "null argument");
  if (m->neq > 0 && !in) return fail(OR_ERR_ARG, "a model with shells needs nodal inertias IN");
  OR_HIP(hipSetDevice(m->device));
  const int64_t nn = m->nn;
  std::vector<double> xv((size_t)(6 * nn));
  for (int64_t n = 0; n < nn; ++n) {
    xv[6 * n + 0] = x[3 * n]; xv[6 * n + 1] = x[3 * n + 1]; xv[6 * n + 2] = x[3 * n + 2];
    xv[6 * n + 3] = v[3 * n]; xv[6 * n + 4] = v[3 * n + 1]; xv[6 * n + 5] = v[3 * n + 2];
    if (code && (code[n] & ~511)) return fail(OR_ERR_ARG, "node code has bits outside 0..511");
    if (code && (code[n] & (kIX | kIY | kIZ)) && !vimp)
      return fail(OR_ERR_ARG, "imposed-velocity bits given without vimp");
  }
  OR_HIP(hipMemcpyAsync(m->xv, xv.data(), sizeof(double) * 6 * nn, hipMemcpyHostToDevice, m->stream));
  OR_HIP(hipMemcpyAsync(m->ms, ms, sizeof(double) * nn, hipMemcpyHostToDevice, m->stream));
  if (code) OR_HIP(hipMemcpyAsync(m->code, code, sizeof(int32_t) * nn, hipMemcpyHostToDevice, m->stream));
  if (vimp) OR_HIP(hipMemcpyAsync(m->vimp, vimp, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, m->stream));
  std::vector<double> vr4;
  if (m->vr) {
    vr4.assign((size_t)(4 * nn), 0.0);
    if (vr)
      for (int64_t n = 0; n < nn; ++n) {
        vr4[4 * n] = vr[3 * n]; vr4[4 * n + 1] = vr[3 * n + 1]; vr4[4 * n + 2] = vr[3 * n + 2];
      }
    OR_HIP(hipMemcpyAsync(m->vr, vr4.data(), sizeof(double) * 4 * nn, hipMemcpyHostToDevice, m->stream));
    OR_HIP(hipMemcpyAsync(m->in, in, sizeof(double) * nn, hipMemcpyHostToDevice, m->stream));
  }
  OR_HIP(hipMemsetAsync(m->d, 0, sizeof(double) * 3 * nn, m->stream));
  // external work of the imposed DOFs (FIXVEL's WFEXT), from zero at this upload
  bool imposed = false;
  for (int64_t n = 0; code && n < nn && !imposed; ++n) imposed = (code[n] & (kIX | kIY | kIZ)) != 0;
  if (imposed && !m->wext) {
    OR_HIP(quiesce(m));
    OR_HIP(hipMalloc((void **)&m->wext, sizeof(double) * 2 * (size_t)nn));
    m->bytes += (int64_t)(sizeof(double) * 2 * nn);
    mark_changed(m);
  }
  if (m->wext) OR_HIP(hipMemsetAsync(m->wext, 0, sizeof(double) * 2 * (size_t)nn, m->stream));
  OR_HIP(quiesce(m));
  m->have_nodes = true;
  return OR_OK;
}

int or_model_upload_displacement(or_model *m, const double *d) {
  if (!m || !d) return fail(OR_ERR_ARG, "null argument");
  if (!m->have_nodes) return fail(OR_ERR_STATE, "upload the nodes before their displacements");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  OR_HIP(hipMemcpyAsync(m->d, d, sizeof(double) * 3 * m->nn, hipMemcpyHostToDevice, m->stream));
  OR_HIP(quiesce(m));
  return OR_OK;
}

// Element arrays [ncomp][ne] move between the caller's element order and the storage
// order of slots (perm: slot -> element; empty: the same order).
static inline int64_t elem_index(const std::vector<int32_t> &perm, int64_t i) {
  if (perm.empty()) return i;
  const int64_t ne = (int64_t)perm.size();
  return (i / ne) * ne + perm[(size_t)(i % ne)];
}
template <typename S>
static int put(or_model *m, void *dst, const double *src, int64_t n, double dflt, const std::vector<int32_t> &perm) {
  std::vector<S> t((size_t)n);
  for (int64_t i = 0; i < n; ++i) t[i] = (S)(src ? src[elem_index(perm, i)] : dflt);
  OR_HIP(hipMemcpyAsync(dst, t.data(), sizeof(S) * n, hipMemcpyHostToDevice, m->stream));
  OR_HIP(quiesce(m));
  return OR_OK;
}
static int put_state(or_model *m, void *dst, const double *src, int64_t n, double dflt,
                     const std::vector<int32_t> &perm) {
  if (!dst || n <= 0) return OR_OK;
  return m->prec == OR_PREC_FP64 ? put<double>(m, dst, src, n, dflt, perm) : put<float>(m, dst, src, n, dflt, perm);
}
template <typename S>
static int get(or_model *m, double *dst, const void *src, int64_t n, const std::vector<int32_t> &perm) {
  std::vector<S> t((size_t)n);
  OR_HIP(hipMemcpyAsync(t.data(), src, sizeof(S) * n, hipMemcpyDeviceToHost, m->stream));
  OR_HIP(quiesce(m));
  for (int64_t i = 0; i < n; ++i) dst[elem_index(perm, i)] = (double)t[i];
  return OR_OK;
}
static int get_state(or_model *m, double *dst, const void *src, int64_t n, const std::vector<int32_t> &perm) {
  if (!dst || n <= 0) return OR_OK;
  if (!src) return fail(OR_ERR_STATE, "state array not held by this model");
  return m->prec == OR_PREC_FP64 ? get<double>(m, dst, src, n, perm) : get<float>(m, dst, src, n, perm);
}

// Node coordinates on the host ([nn][3]), or empty before the nodes are uploaded.
static int host_coords(or_model *m, std::vector<double> &x) {
  x.clear();
  if (!m->have_nodes) return OR_OK;
  const int64_t nn = m->nn;
  std::vector<double> xv((size_t)(6 * nn));
  OR_HIP(quiesce(m));
  OR_HIP(hipMemcpy(xv.data(), m->xv, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost));
  x.resize((size_t)(3 * nn));
  for (int64_t n = 0; n < nn; ++n)
    for (int d = 0; d < 3; ++d) x[(size_t)(3 * n + d)] = xv[(size_t)(6 * n + d)];
  return OR_OK;
}

// The storage order of a family is chosen at its first element upload: tiles from the
// node coordinates when the nodes are already uploaded, the caller's order otherwise.
static int choose_order(or_model *m, int nk, int64_t ne, const int32_t *conn, std::vector<int32_t> &perm) {
  if (!perm.empty() || ne == 0) return OR_OK;
  std::vector<double> x;
  int rc = host_coords(m, x);
  if (rc) return rc;
  perm = orasm::tile_order(nk, ne, conn, x.empty() ? nullptr : x.data());
  return OR_OK;
}

int or_model_upload_hex(or_model *m, const int32_t *conn, const double *sig, const double *eint, const double *rho,
                        const double *vol0, const double *pla, const double *epsd, const double *offg,
                        const double *qvis) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  const int64_t ne = m->neh;
  if (ne == 0) { m->have_hex = true; return OR_OK; }
  if (!conn || !vol0 || !rho) return fail(OR_ERR_ARG, "conn, rho and vol0 are required");
  OR_HIP(hipSetDevice(m->device));
  for (int64_t j = 0; j < 8 * ne; ++j)
    if (conn[j] < 0 || conn[j] >= m->nn) return fail(OR_ERR_ARG, "connectivity node id out of range");
  if (m->have_hex && m->have_sky) {
    // the assembly plan holds this connectivity: a later upload may change the state only
    std::vector<int32_t> cur((size_t)(8 * ne));
    OR_HIP(hipMemcpy(cur.data(), m->hconn, sizeof(int32_t) * 8 * ne, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < ne; ++s)
      for (int k = 0; k < 8; ++k)
        if (cur[(size_t)(k * ne + s)] != conn[8 * (int64_t)m->hperm[(size_t)s] + k])
          return fail(OR_ERR_ARG, "hex connectivity differs from the one the skyline was uploaded for");
  }
  int rc = choose_order(m, 8, ne, conn, m->hperm);
  if (rc) return rc;
  const std::vector<int32_t> &perm = m->hperm;
  std::vector<int32_t> soa((size_t)(8 * ne));
  for (int64_t s = 0; s < ne; ++s)
    for (int k = 0; k < 8; ++k) soa[(size_t)(k * ne + s)] = conn[8 * (int64_t)perm[(size_t)s] + k];
  OR_HIP(hipMemcpyAsync(m->hconn, soa.data(), sizeof(int32_t) * 8 * ne, hipMemcpyHostToDevice, m->stream));
  OR_HIP(quiesce(m));
  if ((rc = put_state(m, m->hsig, sig, 6 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->heint, eint, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->hrho, rho, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->hvol0, vol0, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->hpla, pla, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->hepsd, epsd, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->hoffg, offg, ne, 1.0, perm))) return rc;
  if ((rc = put_state(m, m->hqvis, qvis, ne, 0.0, perm))) return rc;
  m->have_hex = true;
  return OR_OK;
}

int or_model_upload_q4(or_model *m, const int32_t *conn, const double *thke, const double *offg, const double *thk,
                       const double *forc, const double *mom, const double *gstr, const double *hourg,
                       const double *eint, const double *epsd, const double *sig, const double *pla,
                       const double *epsd_ip) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  const int64_t ne = m->neq;
  if (ne == 0) { m->have_q4 = true; return OR_OK; }
  if (!conn || !thke) return fail(OR_ERR_ARG, "conn and thke are required");
  OR_HIP(hipSetDevice(m->device));
  for (int64_t e = 0; e < ne; ++e) {
    if (!(thke[e] > 0.0)) return fail(OR_ERR_ARG, "shell thickness must be positive");
    for (int k = 0; k < 4; ++k)
      if (conn[4 * e + k] < 0 || conn[4 * e + k] >= m->nn) return fail(OR_ERR_ARG, "connectivity node id out of range");
  }
  if (m->have_q4 && m->have_sky) {
    std::vector<int32_t> cur((size_t)(4 * ne));
    OR_HIP(hipMemcpy(cur.data(), m->qconn, sizeof(int32_t) * 4 * ne, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < ne; ++s)
      for (int k = 0; k < 4; ++k)
        if (cur[(size_t)(k * ne + s)] != conn[4 * (int64_t)m->qperm[(size_t)s] + k])
          return fail(OR_ERR_ARG, "shell connectivity differs from the one the skyline was uploaded for");
  }
  int rc = choose_order(m, 4, ne, conn, m->qperm);
  if (rc) return rc;
  const std::vector<int32_t> &perm = m->qperm;
  std::vector<int32_t> soa((size_t)(4 * ne));
  for (int64_t s = 0; s < ne; ++s)
    for (int k = 0; k < 4; ++k) soa[(size_t)(k * ne + s)] = conn[4 * (int64_t)perm[(size_t)s] + k];
  OR_HIP(hipMemcpyAsync(m->qconn, soa.data(), sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice, m->stream));
  OR_HIP(quiesce(m));
  const int64_t np = m->npt;
  if ((rc = put_state(m, m->qthke, thke, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qoffg, offg, ne, 1.0, perm))) return rc;
  if ((rc = put_state(m, m->qthk, thk ? thk : thke, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qforc, forc, 5 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qmom, mom, 3 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qgstr, gstr, 8 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qhourg, hourg, 5 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qeint, eint, 2 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qepsd, epsd, ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qsig, sig, np * 5 * ne, 0.0, perm))) return rc;
  if ((rc = put_state(m, m->qpla, pla, np * ne, 0.0, perm))) return rc;
  if (m->qepsdip && (rc = put_state(m, m->qepsdip, epsd_ip, np * ne, 0.0, perm))) return rc;
  m->have_q4 = true;
  return OR_OK;
}

int or_model_upload_skyline(or_model *m, const int64_t *adsky, const int64_t *iads, const int64_t *iadc) {
  if (!m || !adsky) return fail(OR_ERR_ARG, "null argument");
  if ((m->neh > 0 && !iads) || (m->neq > 0 && !iadc)) return fail(OR_ERR_ARG, "missing iads / iadc");
  OR_HIP(hipSetDevice(m->device));
  const int64_t nn = m->nn, neh = m->neh, neq = m->neq;
  const int64_t lsky = m->lsky_h + m->lsky_q;
  if (adsky[0] != 0 || adsky[nn] != lsky) return fail(OR_ERR_ARG, "adsky must start at 0 and end at 8*ne_hex+4*ne_q4");
  for (int64_t n = 0; n < nn; ++n)
    if (adsky[n + 1] < adsky[n]) return fail(OR_ERR_ARG, "adsky must be non-decreasing");
  // owner node of every global row, and the family of every row
  std::vector<int32_t> owner((size_t)lsky);
  for (int64_t n = 0; n < nn; ++n)
    for (int64_t k = adsky[n]; k < adsky[n + 1]; ++k) owner[(size_t)k] = (int32_t)n;
  std::vector<int8_t> fam((size_t)lsky, -1);
  std::vector<int32_t> nh((size_t)nn, 0), nq((size_t)nn, 0);
  auto claim = [&](int64_t K, int f) -> int {
    if (K < 0 || K >= lsky) return fail(OR_ERR_ARG, "skyline row out of range");
    if (fam[(size_t)K] != -1) return fail(OR_ERR_ARG, "skyline rows must be distinct");
    fam[(size_t)K] = (int8_t)f;
    return OR_OK;
  };
  int rc;
  for (int64_t e = 0; e < neh; ++e)
    for (int k = 0; k < 8; ++k) {
      const int64_t K = iads[8 * e + k];
      if ((rc = claim(K, 0))) return rc;
      nh[(size_t)owner[(size_t)K]] += 1;
    }
  for (int64_t e = 0; e < neq; ++e)
    for (int k = 0; k < 4; ++k) {
      const int64_t K = iadc[4 * e + k];
      if ((rc = claim(K, 1))) return rc;
      nq[(size_t)owner[(size_t)K]] += 1;
    }
  // FILLCNE order: the solid rows of a node come before its shell rows
  for (int64_t n = 0; n < nn; ++n)
    for (int64_t k = adsky[n]; k < adsky[n + 1]; ++k)
      if ((k - adsky[n] < nh[(size_t)n]) != (fam[(size_t)k] == 0))
        return fail(OR_ERR_ARG, "skyline rows of a node must list solid rows before shell rows (FILLCNE order)");
  // element and corner of every row
  std::vector<int32_t> relem((size_t)lsky);
  std::vector<int8_t> rcorner((size_t)lsky);
  for (int64_t e = 0; e < neh; ++e)
    for (int k = 0; k < 8; ++k) { relem[(size_t)iads[8 * e + k]] = (int32_t)e; rcorner[(size_t)iads[8 * e + k]] = (int8_t)k; }
  for (int64_t e = 0; e < neq; ++e)
    for (int k = 0; k < 4; ++k) { relem[(size_t)iadc[4 * e + k]] = (int32_t)e; rcorner[(size_t)iadc[4 * e + k]] = (int8_t)k; }
  // the storage orders (chosen here if no element upload has chosen them yet)
  if (neh > 0 && m->hperm.empty()) { m->hperm.resize((size_t)neh); std::iota(m->hperm.begin(), m->hperm.end(), 0); }
  if (neq > 0 && m->qperm.empty()) { m->qperm.resize((size_t)neq); std::iota(m->qperm.begin(), m->qperm.end(), 0); }
  // release the previous plan's arrays
  for (void **q : {&m->rowh, &m->rowq, (void **)&m->hfr_node, (void **)&m->hfr_src, (void **)&m->qfr_node,
                   (void **)&m->qfr_src})
    if (*q) { (void)hipFree(*q); *q = nullptr; }
  m->bytes -= m->sky_bytes;
  m->sky_bytes = 0;
  const size_t ss = ssize(m);
  for (int f = 0; f < 2; ++f) {
    const int64_t ne = f == 0 ? neh : neq;
    if (ne == 0) continue;
    const int nk = f == 0 ? 8 : 4, W = f == 0 ? kHexRowWords : kQ4RowWords;
    // the family's rows of every node, in the node's row order
    std::vector<int64_t> rptr((size_t)nn + 1, 0);
    std::vector<int32_t> fe;
    std::vector<int8_t> fk;
    std::vector<char> ok((size_t)nn);
    fe.reserve((size_t)(nk * ne));
    fk.reserve((size_t)(nk * ne));
    for (int64_t n = 0; n < nn; ++n) {
      const int64_t k0 = f == 0 ? adsky[n] : adsky[n] + nh[(size_t)n];
      const int64_t k1 = f == 0 ? adsky[n] + nh[(size_t)n] : adsky[n + 1];
      for (int64_t K = k0; K < k1; ++K) { fe.push_back(relem[(size_t)K]); fk.push_back(rcorner[(size_t)K]); }
      rptr[(size_t)n + 1] = (int64_t)fe.size();
      ok[(size_t)n] = (f == 0 || nh[(size_t)n] == 0) ? 1 : 0;
    }
    orasm::FamilyPlan plan;
    orasm::plan_family(plan, nk, ne, nn, f == 0 ? m->hperm : m->qperm, rptr, fe, fk, ok);
    const int64_t nfr = (int64_t)plan.fr_node.size();
    m->asm_stats[2 * f] = nfr;
    m->asm_stats[2 * f + 1] = plan.nrest;
    void **rows = f == 0 ? &m->rowh : &m->rowq;
    int32_t **frn = f == 0 ? &m->hfr_node : &m->qfr_node;
    uint16_t **frs = f == 0 ? &m->hfr_src : &m->qfr_src;
    const size_t nrow = ss * (size_t)(plan.nrest * W);
    OR_HIP(hipMalloc(rows, nrow ? nrow : ss));
    OR_HIP(hipMalloc((void **)frn, sizeof(int32_t) * (size_t)std::max<int64_t>(nfr, 1)));
    OR_HIP(hipMalloc((void **)frs, sizeof(uint16_t) * orasm::kRunCap * (size_t)std::max<int64_t>(nfr, 1)));
    m->sky_bytes += (int64_t)(nrow + (4 + 2 * orasm::kRunCap) * (size_t)nfr);
    const std::vector<int32_t> &deg = f == 0 ? nh : nq;
    hipStream_t st = m->stream;
    OR_HIP(hipMemcpyAsync(f == 0 ? m->deg_h : m->deg_q, deg.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, st));
    OR_HIP(hipMemcpyAsync(f == 0 ? m->rs_h : m->rs_q, plan.rs.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice, st));
    OR_HIP(hipMemcpyAsync(f == 0 ? m->hfr_start : m->qfr_start, plan.fr_start.data(),
                          sizeof(int32_t) * plan.fr_start.size(), hipMemcpyHostToDevice, st));
    if (nfr > 0) {
      OR_HIP(hipMemcpyAsync(*frn, plan.fr_node.data(), sizeof(int32_t) * nfr, hipMemcpyHostToDevice, st));
      OR_HIP(hipMemcpyAsync(*frs, plan.fr_src.data(), sizeof(uint16_t) * orasm::kRunCap * nfr, hipMemcpyHostToDevice, st));
    }
    OR_HIP(hipMemcpyAsync(f == 0 ? m->hrest : m->qrest, plan.rest.data(), sizeof(int32_t) * nk * ne,
                          hipMemcpyHostToDevice, st));
    OR_HIP(quiesce(m));
  }
  m->bytes += m->sky_bytes;
  m->have_sky = true;
  mark_changed(m);
  return OR_OK;
}

int or_model_set_time(or_model *m, const or_time_state *t) {
  if (!m || !t) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  DevTime dt{};
  dt.tt = t->tt; dt.dt1 = t->dt1; dt.dt2 = t->dt2; dt.dt12 = t->dt12; dt.dt2old = t->dt2old;
  dt.ncycle = t->ncycle; dt.mstop = t->mstop; dt.done = t->mstop ? 1 : 0;
  DevTime two[2] = {dt, dt};
  OR_HIP(hipMemcpy(m->ts, two, sizeof(two), hipMemcpyHostToDevice));
  unsigned long long dm[2];
  const double big = kEP06;
  std::memcpy(&dm[0], &big, 8);
  dm[1] = dm[0];
  OR_HIP(hipMemcpy(m->dtmin, dm, sizeof(dm), hipMemcpyHostToDevice));
  int zero = 0;
  OR_HIP(hipMemcpy(m->flags, &zero, sizeof(int), hipMemcpyHostToDevice));
  m->parity = 0;
  return OR_OK;
}

int or_model_get_time(or_model *m, or_time_state *t) {
  if (!m || !t) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
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
  OR_HIP(hipEventSynchronize(m->ev[4 * (m->ring_used - 1) + 3]));
  for (int j = 0; j < m->ring_used; ++j) {
    float a = 0.f, b = 0.f, c = 0.f;
    OR_HIP(hipEventElapsedTime(&a, m->ev[4 * j], m->ev[4 * j + 1]));
    OR_HIP(hipEventElapsedTime(&b, m->ev[4 * j + 1], m->ev[4 * j + 2]));
    OR_HIP(hipEventElapsedTime(&c, m->ev[4 * j + 2], m->ev[4 * j + 3]));
    m->t_hex += a;
    m->t_q4 += b;
    m->t_node += c;
    m->t_cycles += 1;
  }
  m->ring_used = 0;
  return OR_OK;
}

// Refresh the device copy of a model's kernel view after uploads or parameter changes.
static void fill_dev(or_model *m, ModelDev &d, double tend) {
  d = ModelDev{};
  d.hconn = m->hconn; d.hrest = m->hrest; d.hfr_start = m->hfr_start; d.hfr_node = m->hfr_node;
  d.hfr_src = m->hfr_src; d.hsig = m->hsig; d.heint = m->heint; d.hrho = m->hrho;
  d.hpla = m->hpla; d.hepsd = m->hepsd; d.hoffg = m->hoffg; d.hqvis = m->hqvis; d.hvol0 = m->hvol0;
  d.neh = m->neh;
  d.qconn = m->qconn; d.qrest = m->qrest; d.qfr_start = m->qfr_start; d.qfr_node = m->qfr_node;
  d.qfr_src = m->qfr_src; d.qoffg = m->qoffg; d.qthk = m->qthk; d.qforc = m->qforc;
  d.qmom = m->qmom; d.qgstr = m->qgstr; d.qhourg = m->qhourg; d.qeint = m->qeint; d.qepsd = m->qepsd;
  d.qsig = m->qsig; d.qpla = m->qpla; d.qepsdip = m->qepsdip; d.qthke = m->qthke; d.neq = m->neq;
  d.nd = NodePtrs{m->xv, m->vr, m->d, m->ms, m->in, m->code, m->vimp, m->deg_h, m->deg_q, m->rs_h, m->rs_q,
                  m->preh, m->preq, m->wext, m->nn};
  d.rowh = m->rowh; d.rowq = m->rowq;
  d.preh = m->preh; d.preq = m->preq;
  d.ts = m->ts; d.dtmin = m->dtmin; d.flags = m->flags; d.ecrit = m->ecrit;
  d.dtel_h = m->dtel_h; d.dtel_q = m->dtel_q;
  d.th_slot = m->th_slot; d.th_node = m->th_buf; d.ehou_h = m->ehou_h; d.ehou_q = m->ehou_q;
  d.dtmx = m->neh > 0 ? m->hc.dtmx : 1.0e20;
  d.tend = tend;
  d.hc = m->hc;
  d.qc = m->qc;
}

static int refresh_dev(or_model *m) {
  if (!m->dev_dirty) return OR_OK;
  fill_dev(m, m->hdev, kNoEnd);
  OR_HIP(quiesce(m));
  OR_HIP(hipMemcpy(m->dev, &m->hdev, sizeof(ModelDev), hipMemcpyHostToDevice));
  m->dev_dirty = false;
  return OR_OK;
}

#endif // OR_ILP_TU

// The three launches of one cycle for a set of models of one precision: element
// forces (hexes, then shells) and the node update, each over the set's tile list.
struct LaunchSet {
  const ModelDev *md = nullptr;
  const Tile *th = nullptr, *tq = nullptr, *tn = nullptr;
  int32_t nth = 0, ntq = 0, ntn = 0;
  int vh = 0, vq = 0; // element kernel variants (their block sizes match th / tq)
  bool eh = false, eq = false; // the set keeps EHOUR (hexes, shells): one model with a time history
};

// v & 1 selects the register budget (2 or 3 waves per SIMD); the caller routes v >= 2 to
// the unit built with the ILP-first scheduler.
template <typename R, typename S, bool EH>
static void launch_hex_eh(int v, hipStream_t st, const LaunchSet &L, int parity) {
  constexpr int W0 = variant_waves<R>(0), W1 = variant_waves<R>(1);
  const dim3 g((unsigned)L.nth);
  if ((v & 1) == 0) hipLaunchKernelGGL((k_hex_forces<R, S, kTile, W0, EH>), g, dim3(kTile), 0, st, L.md, L.th, parity);
  else hipLaunchKernelGGL((k_hex_forces<R, S, kTile, W1, EH>), g, dim3(kTile), 0, st, L.md, L.th, parity);
}
template <typename R, typename S>
static void launch_hex(int v, hipStream_t st, const LaunchSet &L, int parity) {
  if (L.eh) launch_hex_eh<R, S, true>(v, st, L, parity);
  else launch_hex_eh<R, S, false>(v, st, L, parity);
}

template <typename R, typename S, bool EH>
static void launch_q4_eh(int v, hipStream_t st, const LaunchSet &L, int parity) {
  constexpr int W0 = variant_waves<R>(0), W1 = variant_waves<R>(1);
  const dim3 g((unsigned)L.ntq);
  if ((v & 1) == 0) hipLaunchKernelGGL((k_q4_forces<R, S, kTile, W0, EH>), g, dim3(kTile), 0, st, L.md, L.tq, parity);
  else hipLaunchKernelGGL((k_q4_forces<R, S, kTile, W1, EH>), g, dim3(kTile), 0, st, L.md, L.tq, parity);
}
template <typename R, typename S>
static void launch_q4(int v, hipStream_t st, const LaunchSet &L, int parity) {
  if (L.eq) launch_q4_eh<R, S, true>(v, st, L, parity);
  else launch_q4_eh<R, S, false>(v, st, L, parity);
}

// Element launches of variants 2..3 (variant - 2 in the unit built with the ILP-first
// scheduler): family 0 hexes, 1 shells; fp64 selects <double,double> or <float,float>.
void or_ilp_elements(int family, int fp64, int variant, hipStream_t st, const void *md, const void *tiles,
                     int32_t ntiles, int parity, bool ehour)
#ifdef OR_ILP_TU
{
  LaunchSet L;
  L.md = (const ModelDev *)md;
  L.eh = L.eq = ehour;
  if (family == 0) {
    L.th = (const Tile *)tiles; L.nth = ntiles;
    if (fp64) launch_hex<double, double>(variant, st, L, parity);
    else launch_hex<float, float>(variant, st, L, parity);
  } else {
    L.tq = (const Tile *)tiles; L.ntq = ntiles;
    if (fp64) launch_q4<double, double>(variant, st, L, parity);
    else launch_q4<float, float>(variant, st, L, parity);
  }
}
#else
;

template <typename R, typename S>
static int launch_cycle(hipStream_t st, const LaunchSet &L, int parity, hipEvent_t after_hex, hipEvent_t after_q4,
                        int output = 0) {
  constexpr int fp64 = sizeof(R) == 8 ? 1 : 0;
  const int po = parity | (output ? 2 : 0); // element kernels: parity and the output-cycle bit
  if (L.nth > 0) {
    if (L.vh >= 2) or_ilp_elements(0, fp64, L.vh - 2, st, L.md, L.th, L.nth, po, L.eh);
    else launch_hex<R, S>(L.vh, st, L, po);
    OR_HIP(hipGetLastError());
  }
  if (after_hex) OR_HIP(hipEventRecord(after_hex, st));
  if (L.ntq > 0) {
    if (L.vq >= 2) or_ilp_elements(1, fp64, L.vq - 2, st, L.md, L.tq, L.ntq, po, L.eq);
    else launch_q4<R, S>(L.vq, st, L, po);
    OR_HIP(hipGetLastError());
  }
  if (after_q4) OR_HIP(hipEventRecord(after_q4, st));
  if (L.ntn > 0) {
    hipLaunchKernelGGL((k_node_update<S>), dim3((unsigned)L.ntn), dim3(kBlock), 0, st, L.md, L.tn, parity,
                       (double)1.1f, output);
    OR_HIP(hipGetLastError());
  }
  return OR_OK;
}

int or_model_run(or_model *m, int64_t ncycles) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  if (ncycles < 0) return fail(OR_ERR_ARG, "ncycles < 0");
  if (!m->have_nodes || !m->have_sky || (m->neh > 0 && (!m->have_hparams || !m->have_hex)) ||
      (m->neq > 0 && (!m->have_qparams || !m->have_q4)))
    return fail(OR_ERR_STATE, "params, nodes, element state and skyline must be uploaded before run");
  if (m->batch) return fail(OR_ERR_STATE, "the model belongs to a batch: advance it with or_batch_run");
  OR_HIP(hipSetDevice(m->device));
  int rc0 = refresh_dev(m);
  if (rc0) return rc0;
  LaunchSet L;
  L.md = m->dev;
  L.th = m->tiles_h; L.nth = m->nt_h; L.vh = m->var_h;
  L.eh = m->ehou_h != nullptr; L.eq = m->ehou_q != nullptr;
  L.tq = m->tiles_q; L.ntq = m->nt_q; L.vq = m->var_q;
  L.tn = m->tiles_n; L.ntn = m->nt_n;
  for (int64_t c = 0; c < ncycles; ++c) {
    if (m->timing) {
      if (m->ring_used * 4 + 4 > (int)m->ev.size()) {
        int rc = flush_timing(m);
        if (rc) return rc;
      }
      OR_HIP(hipEventRecord(m->ev[4 * m->ring_used], m->stream));
    }
    hipEvent_t e1 = m->timing ? m->ev[4 * m->ring_used + 1] : nullptr;
    hipEvent_t e2 = m->timing ? m->ev[4 * m->ring_used + 2] : nullptr;
    const int p = (int)(m->parity & 1);
    int rc = (m->prec == OR_PREC_FP64) ? launch_cycle<double, double>(m->stream, L, p, e1, e2)
                                       : launch_cycle<float, float>(m->stream, L, p, e1, e2);
    if (rc) return rc;
    m->parity ^= 1;
    if (m->timing) {
      OR_HIP(hipEventRecord(m->ev[4 * m->ring_used + 3], m->stream));
      m->ring_used += 1;
    }
  }
  return OR_OK;
}

static int energy_sums(or_model *m, const double *ecrit, double out[kEnergySums]);

// Per-element candidates of the controlling element and each storage slot's element index,
// allocated (and the slot indices uploaded) at the first output cycle.
static int ensure_ctrl(or_model *m) {
  if (m->ctrl) return OR_OK;
  OR_HIP(quiesce(m));
  auto put = [&](double **dt, int32_t **lidx, const std::vector<int32_t> &perm, int64_t ne) -> int {
    if (ne <= 0) return OR_OK;
    std::vector<int32_t> li((size_t)ne);
    for (int64_t s = 0; s < ne; ++s) li[(size_t)s] = perm.empty() ? (int32_t)s : perm[(size_t)s];
    OR_HIP(hipMalloc((void **)dt, sizeof(double) * (size_t)ne));
    OR_HIP(hipMalloc((void **)lidx, sizeof(int32_t) * (size_t)ne));
    OR_HIP(hipMemcpy(*lidx, li.data(), sizeof(int32_t) * (size_t)ne, hipMemcpyHostToDevice));
    m->bytes += (int64_t)(12 * ne);
    return OR_OK;
  };
  int rc = put(&m->dtel_h, &m->lidx_h, m->hperm, m->neh);
  if (rc == OR_OK) rc = put(&m->dtel_q, &m->lidx_q, m->qperm, m->neq);
  if (rc) return rc;
  OR_HIP(hipMalloc((void **)&m->ctrl, 4 * sizeof(unsigned long long)));
  m->bytes += 32;
  mark_changed(m);
  return OR_OK;
}

int or_model_run_ecrit(or_model *m, or_ecrit *out) {
  if (!m || !out) return fail(OR_ERR_ARG, "null argument");
  if (!m->have_nodes || !m->have_sky || (m->neh > 0 && (!m->have_hparams || !m->have_hex)) ||
      (m->neq > 0 && (!m->have_qparams || !m->have_q4)))
    return fail(OR_ERR_STATE, "params, nodes, element state and skyline must be uploaded before run");
  if (m->batch) return fail(OR_ERR_STATE, "the model belongs to a batch: advance it with or_batch_run");
  OR_HIP(hipSetDevice(m->device));
  if (!m->ecrit) {
    OR_HIP(quiesce(m));
    OR_HIP(hipMalloc((void **)&m->ecrit, sizeof(double) * 5 * (size_t)m->nn));
    m->bytes += (int64_t)(sizeof(double) * 5 * m->nn);
    mark_changed(m);
  }
  int rc = ensure_ctrl(m);
  if (rc) return rc;
  // candidates of elements that cannot control the step (and of a stopped model) stay EP06
  auto fill_ep06 = [&](double *p, int64_t n) -> int {
    if (n > 0)
      hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, m->stream, p, n,
                         kEP06);
    OR_HIP(hipGetLastError());
    return OR_OK;
  };
  if ((rc = fill_ep06(m->dtel_h, m->neh)) || (rc = fill_ep06(m->dtel_q, m->neq))) return rc;
  rc = refresh_dev(m);
  if (rc) return rc;
  LaunchSet L;
  L.md = m->dev;
  L.th = m->tiles_h; L.nth = m->nt_h; L.vh = m->var_h;
  L.eh = m->ehou_h != nullptr; L.eq = m->ehou_q != nullptr;
  L.tq = m->tiles_q; L.ntq = m->nt_q; L.vq = m->var_q;
  L.tn = m->tiles_n; L.ntn = m->nt_n;
  const int p = (int)(m->parity & 1);
  rc = (m->prec == OR_PREC_FP64) ? launch_cycle<double, double>(m->stream, L, p, nullptr, nullptr, 1)
                                 : launch_cycle<float, float>(m->stream, L, p, nullptr, nullptr, 1);
  if (rc) return rc;
  m->parity ^= 1;
  double e[kEnergySums];
  if ((rc = energy_sums(m, m->ecrit, e))) return rc;
  out->enint = e[0]; out->encin = e[1]; out->enrot = e[2]; out->wfext = e[3];
  m->momentum[0] = e[4]; m->momentum[1] = e[5]; m->momentum[2] = e[6];
  // the controlling element (k_ctrl_min / k_ctrl_pick)
  unsigned long long w0[4];
  const double big = kEP06;
  std::memcpy(&w0[0], &big, 8);
  w0[1] = w0[0];
  w0[2] = 0;
  w0[3] = ~0ull;
  OR_HIP(hipMemcpyAsync(m->ctrl, w0, sizeof(w0), hipMemcpyHostToDevice, m->stream));
  const unsigned nb = 1024;
  if (m->neh > 0) {
    hipLaunchKernelGGL(k_ctrl_min, dim3(nb), dim3(kBlock), 0, m->stream, m->dtel_h, m->neh, m->ctrl);
    hipLaunchKernelGGL(k_ctrl_pick, dim3(nb), dim3(kBlock), 0, m->stream, m->dtel_h, m->lidx_h, m->neh, m->ctrl, 1,
                       m->ctrl + 2);
  }
  if (m->neq > 0) {
    hipLaunchKernelGGL(k_ctrl_min, dim3(nb), dim3(kBlock), 0, m->stream, m->dtel_q, m->neq, m->ctrl + 1);
    hipLaunchKernelGGL(k_ctrl_pick, dim3(nb), dim3(kBlock), 0, m->stream, m->dtel_q, m->lidx_q, m->neq, m->ctrl + 1,
                       0, m->ctrl + 3);
  }
  OR_HIP(hipGetLastError());
  unsigned long long w[4];
  OR_HIP(hipMemcpyAsync(w, m->ctrl, sizeof(w), hipMemcpyDeviceToHost, m->stream));
  OR_HIP(quiesce(m));
  double mh, mq;
  std::memcpy(&mh, &w[0], 8);
  std::memcpy(&mq, &w[1], 8);
  out->itypts = 0;
  out->pad = 0;
  out->nelts = -1;
  if (mh < kEP06) { out->itypts = 1; out->nelts = (int64_t)w[2]; }
  if (mq < mh) { out->itypts = 3; out->nelts = (int64_t)w[3]; }
  return OR_OK;
}

int or_model_run_output(or_model *m, double out[3]) {
  or_ecrit e;
  const int rc = or_model_run_ecrit(m, &e);
  if (rc) return rc;
  out[0] = e.enint; out[1] = e.encin; out[2] = e.enrot;
  return OR_OK;
}

// ---------------------------------------------------------------- time history (HIST2)
int or_model_th_setup(or_model *m, int64_t n_node, const int32_t *node, int64_t n_hex, const int32_t *hex,
                      int64_t n_q4, const int32_t *q4) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  if (n_node < 0 || n_hex < 0 || n_q4 < 0 || (n_node && !node) || (n_hex && !hex) || (n_q4 && !q4))
    return fail(OR_ERR_ARG, "entity counts must be >= 0 with their lists given");
  if (m->batch) return fail(OR_ERR_STATE, "the model belongs to a batch");
  if (!m->have_nodes || !m->have_sky || (m->neh > 0 && !m->have_hex) || (m->neq > 0 && !m->have_q4))
    return fail(OR_ERR_STATE, "nodes, element state and skyline must be uploaded before the time history");
  std::vector<int32_t> slot((size_t)m->nn, -1);
  for (int64_t j = 0; j < n_node; ++j) {
    if (node[j] < 0 || node[j] >= m->nn) return fail(OR_ERR_ARG, "time-history node out of range");
    if (slot[(size_t)node[j]] >= 0) return fail(OR_ERR_ARG, "time-history node listed twice");
    slot[(size_t)node[j]] = (int32_t)j;
  }
  auto slots = [&](const int32_t *ids, int64_t n, int64_t ne, const std::vector<int32_t> &perm,
                   std::vector<int32_t> &out) -> int {
    std::vector<int32_t> inv((size_t)ne);
    for (int64_t s = 0; s < ne; ++s) inv[(size_t)(perm.empty() ? s : perm[(size_t)s])] = (int32_t)s;
    out.resize((size_t)n);
    for (int64_t j = 0; j < n; ++j) {
      if (ids[j] < 0 || ids[j] >= ne) return fail(OR_ERR_ARG, "time-history element out of range");
      out[(size_t)j] = inv[(size_t)ids[j]];
    }
    return OR_OK;
  };
  std::vector<int32_t> hs, qs;
  int rc = slots(hex, n_hex, m->neh, m->hperm, hs);
  if (rc == OR_OK) rc = slots(q4, n_q4, m->neq, m->qperm, qs);
  if (rc) return rc;
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  for (void **p : {(void **)&m->th_slot, (void **)&m->th_hex, (void **)&m->th_q4, (void **)&m->th_buf})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  const int64_t words = 18 * n_node + 13 * n_hex + (12 + m->npt) * n_q4 + 1;
  OR_HIP(hipMalloc((void **)&m->th_slot, sizeof(int32_t) * (size_t)m->nn));
  OR_HIP(hipMemcpy(m->th_slot, slot.data(), sizeof(int32_t) * (size_t)m->nn, hipMemcpyHostToDevice));
  OR_HIP(hipMalloc((void **)&m->th_buf, sizeof(double) * (size_t)words));
  OR_HIP(hipMemset(m->th_buf, 0, sizeof(double) * (size_t)words));
  if (n_hex) {
    OR_HIP(hipMalloc((void **)&m->th_hex, sizeof(int32_t) * (size_t)n_hex));
    OR_HIP(hipMemcpy(m->th_hex, hs.data(), sizeof(int32_t) * (size_t)n_hex, hipMemcpyHostToDevice));
  }
  if (n_q4) {
    OR_HIP(hipMalloc((void **)&m->th_q4, sizeof(int32_t) * (size_t)n_q4));
    OR_HIP(hipMemcpy(m->th_q4, qs.data(), sizeof(int32_t) * (size_t)n_q4, hipMemcpyHostToDevice));
  }
  // EHOUR accumulates from the first setup on (a later setup keeps it)
  if (m->nt_h > 0 && !m->ehou_h) {
    OR_HIP(hipMalloc((void **)&m->ehou_h, sizeof(double) * (size_t)m->nt_h));
    OR_HIP(hipMemset(m->ehou_h, 0, sizeof(double) * (size_t)m->nt_h));
    m->bytes += 8 * (int64_t)m->nt_h;
  }
  if (m->nt_q > 0 && !m->ehou_q) {
    OR_HIP(hipMalloc((void **)&m->ehou_q, sizeof(double) * (size_t)m->nt_q));
    OR_HIP(hipMemset(m->ehou_q, 0, sizeof(double) * (size_t)m->nt_q));
    m->bytes += 8 * (int64_t)m->nt_q;
  }
  m->bytes += 4 * m->nn + 8 * words + 4 * (n_hex + n_q4);
  m->th_nn = n_node; m->th_nh = n_hex; m->th_nq = n_q4;
  mark_changed(m);
  return OR_OK;
}

static int th_ehour(or_model *m, double *ehour) {
  double *w = m->th_buf + 18 * m->th_nn + 13 * m->th_nh + (12 + m->npt) * m->th_nq;
  hipLaunchKernelGGL(k_th_ehour, dim3(1), dim3(kBlock), 0, m->stream, m->ehou_h, (int64_t)(m->ehou_h ? m->nt_h : 0),
                     m->ehou_q, (int64_t)(m->ehou_q ? m->nt_q : 0), w);
  OR_HIP(hipGetLastError());
  OR_HIP(hipMemcpyAsync(ehour, w, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  OR_HIP(quiesce(m));
  return OR_OK;
}

int or_model_run_th(or_model *m, or_ecrit *ec, double glob[4], double *node_out, double *hex_out, double *q4_out) {
  if (!m || !ec || !glob) return fail(OR_ERR_ARG, "null argument");
  if (!m->th_buf) return fail(OR_ERR_STATE, "or_model_th_setup first");
  if ((m->th_nn && !node_out) || (m->th_nh && !hex_out) || (m->th_nq && !q4_out))
    return fail(OR_ERR_ARG, "an output array of the listed entities is null");
  int rc = or_model_run_ecrit(m, ec);
  if (rc) return rc;
  OR_HIP(hipSetDevice(m->device));
  double *hb = m->th_buf + 18 * m->th_nn, *qb = hb + 13 * m->th_nh;
  const int64_t ne = m->th_nh + m->th_nq;
  if (ne > 0) {
    const dim3 g((unsigned)((ne + kBlock - 1) / kBlock));
    if (m->prec == OR_PREC_FP64)
      hipLaunchKernelGGL(k_th_elements<double>, g, dim3(kBlock), 0, m->stream, m->dev, m->th_hex, m->th_nh, m->th_q4,
                         m->th_nq, m->npt, hb, qb);
    else
      hipLaunchKernelGGL(k_th_elements<float>, g, dim3(kBlock), 0, m->stream, m->dev, m->th_hex, m->th_nh, m->th_q4,
                         m->th_nq, m->npt, hb, qb);
    OR_HIP(hipGetLastError());
  }
  if (m->th_nn) OR_HIP(hipMemcpyAsync(node_out, m->th_buf, sizeof(double) * 18 * m->th_nn, hipMemcpyDeviceToHost,
                                      m->stream));
  if (m->th_nh) OR_HIP(hipMemcpyAsync(hex_out, hb, sizeof(double) * 13 * m->th_nh, hipMemcpyDeviceToHost, m->stream));
  if (m->th_nq)
    OR_HIP(hipMemcpyAsync(q4_out, qb, sizeof(double) * (12 + m->npt) * m->th_nq, hipMemcpyDeviceToHost, m->stream));
  glob[0] = m->momentum[0]; glob[1] = m->momentum[1]; glob[2] = m->momentum[2];
  return th_ehour(m, &glob[3]);
}

int or_model_ehour(or_model *m, double *ehour) {
  if (!m || !ehour) return fail(OR_ERR_ARG, "null argument");
  if (!m->th_buf) return fail(OR_ERR_STATE, "or_model_th_setup first");
  OR_HIP(hipSetDevice(m->device));
  return th_ehour(m, ehour);
}

int or_model_set_ehour(or_model *m, double ehour) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  if (!m->th_buf) return fail(OR_ERR_STATE, "or_model_th_setup first");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  if (m->ehou_h) OR_HIP(hipMemset(m->ehou_h, 0, sizeof(double) * (size_t)m->nt_h));
  if (m->ehou_q) OR_HIP(hipMemset(m->ehou_q, 0, sizeof(double) * (size_t)m->nt_q));
  double *first = m->ehou_h ? m->ehou_h : m->ehou_q;
  if (first) OR_HIP(hipMemcpy(first, &ehour, sizeof(double), hipMemcpyHostToDevice));
  return OR_OK;
}

int or_model_sync(or_model *m) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  int flags = 0;
  OR_HIP(hipMemcpy(&flags, m->flags, sizeof(int), hipMemcpyDeviceToHost));
  if (flags & 1) return fail(OR_ERR_MSTOP, "non-positive Jacobian: element deleted and MSTOP raised");
  return OR_OK;
}

int or_model_set_timing(or_model *m, int enable) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  if (enable && m->ev.empty()) {
    m->ev.resize(4 * 256);
    for (auto &e : m->ev) OR_HIP(hipEventCreate(&e));
  }
  m->timing = enable != 0;
  return OR_OK;
}

int or_model_kernel_times(or_model *m, int reset, double *hex_ms, double *q4_ms, double *node_ms, int64_t *cycles) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  int rc = flush_timing(m);
  if (rc) return rc;
  const double nc = m->t_cycles > 0 ? (double)m->t_cycles : 1.0;
  if (hex_ms) *hex_ms = m->t_hex / nc;
  if (q4_ms) *q4_ms = m->t_q4 / nc;
  if (node_ms) *node_ms = m->t_node / nc;
  if (cycles) *cycles = m->t_cycles;
  if (reset) { m->t_hex = 0.0; m->t_q4 = 0.0; m->t_node = 0.0; m->t_cycles = 0; }
  return OR_OK;
}

int or_model_download_nodes(or_model *m, double *x, double *v, double *vr, double *d) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  const int64_t nn = m->nn;
  std::vector<double> xv((size_t)(6 * nn));
  OR_HIP(hipMemcpy(xv.data(), m->xv, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost));
  for (int64_t n = 0; n < nn; ++n) {
    if (x) { x[3 * n] = xv[6 * n]; x[3 * n + 1] = xv[6 * n + 1]; x[3 * n + 2] = xv[6 * n + 2]; }
    if (v) { v[3 * n] = xv[6 * n + 3]; v[3 * n + 1] = xv[6 * n + 4]; v[3 * n + 2] = xv[6 * n + 5]; }
  }
  if (vr) {
    if (m->vr) {
      std::vector<double> w((size_t)(4 * nn));
      OR_HIP(hipMemcpy(w.data(), m->vr, sizeof(double) * 4 * nn, hipMemcpyDeviceToHost));
      for (int64_t n = 0; n < nn; ++n) { vr[3 * n] = w[4 * n]; vr[3 * n + 1] = w[4 * n + 1]; vr[3 * n + 2] = w[4 * n + 2]; }
    } else {
      std::memset(vr, 0, sizeof(double) * 3 * nn);
    }
  }
  if (d) OR_HIP(hipMemcpy(d, m->d, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost));
  return OR_OK;
}

int or_model_download_hex(or_model *m, double *sig, double *eint, double *rho, double *pla, double *epsd,
                          double *offg, double *qvis) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  const int64_t ne = m->neh;
  int rc;
  if ((rc = get_state(m, sig, m->hsig, 6 * ne, m->hperm))) return rc;
  if ((rc = get_state(m, eint, m->heint, ne, m->hperm))) return rc;
  if ((rc = get_state(m, rho, m->hrho, ne, m->hperm))) return rc;
  if ((rc = get_state(m, pla, m->hpla, ne, m->hperm))) return rc;
  if ((rc = get_state(m, epsd, m->hepsd, ne, m->hperm))) return rc;
  if ((rc = get_state(m, offg, m->hoffg, ne, m->hperm))) return rc;
  if ((rc = get_state(m, qvis, m->hqvis, ne, m->hperm))) return rc;
  return OR_OK;
}

int or_model_download_q4(or_model *m, double *offg, double *thk, double *forc, double *mom, double *gstr,
                         double *hourg, double *eint, double *epsd, double *sig, double *pla, double *epsd_ip) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  const int64_t ne = m->neq, np = m->npt;
  int rc;
  if ((rc = get_state(m, offg, m->qoffg, ne, m->qperm))) return rc;
  if ((rc = get_state(m, thk, m->qthk, ne, m->qperm))) return rc;
  if ((rc = get_state(m, forc, m->qforc, 5 * ne, m->qperm))) return rc;
  if ((rc = get_state(m, mom, m->qmom, 3 * ne, m->qperm))) return rc;
  if ((rc = get_state(m, gstr, m->qgstr, 8 * ne, m->qperm))) return rc;
  if ((rc = get_state(m, hourg, m->qhourg, 5 * ne, m->qperm))) return rc;
  if ((rc = get_state(m, eint, m->qeint, 2 * ne, m->qperm))) return rc;
  if ((rc = get_state(m, epsd, m->qepsd, ne, m->qperm))) return rc;
  if ((rc = get_state(m, sig, m->qsig, np * 5 * ne, m->qperm))) return rc;
  if ((rc = get_state(m, pla, m->qpla, np * ne, m->qperm))) return rc;
  if (epsd_ip) {
    if (m->qepsdip) {
      if ((rc = get_state(m, epsd_ip, m->qepsdip, np * ne, m->qperm))) return rc;
    } else {
      // ASRATE = 1: LBUF%EPSD = 1*EPSD_PG + 0*LBUF%EPSD equals GBUF%EPSD at every point
      std::vector<double> g((size_t)ne);
      if ((rc = get_state(m, g.data(), m->qepsd, ne, m->qperm))) return rc;
      for (int64_t ip = 0; ip < np; ++ip) std::memcpy(epsd_ip + ip * ne, g.data(), sizeof(double) * ne);
    }
  }
  return OR_OK;
}

// Fixed-order reduction of the internal energy and of the kinetic terms: 0.5*MS*|V|^2 from the
// resident state, or the per-node ECRIT terms of the last output cycle (ecrit != null, which
// also gives the momentum sums out[4..6]).
static int energy_sums(or_model *m, const double *ecrit, double out[kEnergySums]) {
  int64_t n = m->nn;
  if (m->neh > n) n = m->neh;
  if (m->neq > n) n = m->neq;
  const int nb = 1024;
  const int64_t per = (n + nb - 1) / nb;
  const int64_t per_block = ((per + kBlock - 1) / kBlock) * kBlock;
  double *partial = m->energy + kEnergySums;
  OR_HIP(quiesce(m));
  if (m->prec == OR_PREC_FP64)
    hipLaunchKernelGGL((k_energy_partial<double>), dim3(nb), dim3(kBlock), 0, m->stream, (const double *)m->heint,
                       (const double *)m->hvol0, m->neh, (const double *)m->qeint, m->neq, m->xv, m->vr, m->ms, m->in,
                       ecrit, m->wext, m->nn, per_block, partial);
  else
    hipLaunchKernelGGL((k_energy_partial<float>), dim3(nb), dim3(kBlock), 0, m->stream, (const float *)m->heint,
                       (const float *)m->hvol0, m->neh, (const float *)m->qeint, m->neq, m->xv, m->vr, m->ms, m->in,
                       ecrit, m->wext, m->nn, per_block, partial);
  OR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, m->stream, (const double *)partial, nb,
                     ecrit ? 1.0 : 0.5, m->energy);
  OR_HIP(hipGetLastError());
  OR_HIP(quiesce(m));
  OR_HIP(hipMemcpy(out, m->energy, kEnergySums * sizeof(double), hipMemcpyDeviceToHost));
  return OR_OK;
}

int or_model_energy(or_model *m, double out[3]) {
  if (!m || !out) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  double e[kEnergySums];
  const int rc = energy_sums(m, nullptr, e);
  if (rc) return rc;
  out[0] = e[0]; out[1] = e[1]; out[2] = e[2];
  return OR_OK;
}

int or_model_wfext(or_model *m, double *wfext) {
  if (!m || !wfext) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(m->device));
  double e[kEnergySums];
  const int rc = energy_sums(m, nullptr, e);
  if (rc) return rc;
  *wfext = e[3];
  return OR_OK;
}

int or_model_footprint_estimate(int64_t nn, int64_t ne_hex, int64_t ne_q4, int32_t npt, int precision,
                                int strain_rate_filter, const int32_t *rows_hex, const int32_t *rows_q4,
                                int64_t *bytes) {
  if (!bytes) return fail(OR_ERR_ARG, "null argument");
  if (nn <= 0 || ne_hex < 0 || ne_q4 < 0) return fail(OR_ERR_ARG, "nn must be > 0, ne_hex and ne_q4 >= 0");
  if (ne_q4 > 0 && (npt < 2 || npt > 10)) return fail(OR_ERR_ARG, "shell NPT must be in 2..10");
  if (precision != OR_PREC_FP64 && precision != OR_PREC_SPLIT) return fail(OR_ERR_ARG, "bad precision");
  if ((ne_hex > 0 && !rows_hex) || (ne_q4 > 0 && !rows_q4)) return fail(OR_ERR_ARG, "missing row counts");
  const int64_t ss = precision == OR_PREC_FP64 ? 8 : 4;
  const int64_t nsl = (nn + kBlock - 1) / kBlock;
  const int64_t th = (ne_hex + kTile - 1) / kTile, tq = (ne_q4 + kTile - 1) / kTile;
  // the allocations of or_model_create, in the same order
  int64_t b = 8 * (6 * nn + 3 * nn + nn + 3 * nn) + 4 * nn;             // xv, d, ms, vimp, code
  b += 16 * nn;                                                          // wext (models with imposed DOFs)
  if (ne_q4 > 0) b += 8 * (4 * nn + nn);                                 // vr, in
  if (ne_hex > 0)                                                        // deg, rs, prefix, tiles, conn, rest, state
    b += 4 * (nn + nn + 1) + 8 * 3 * nn + 4 * (th + 1) + 4 * 16 * ne_hex + ss * 13 * ne_hex;
  if (ne_q4 > 0) {
    b += 4 * (nn + nn + 1) + 8 * 6 * nn + 4 * (tq + 1) + 4 * 8 * ne_q4 + ss * (4 + 5 + 3 + 8 + 5 + 2) * ne_q4 +
         ss * 6 * (int64_t)npt * ne_q4;
    if (strain_rate_filter) b += ss * (int64_t)npt * ne_q4;              // LBUF EPSD (or_model_set_q4_params)
  }
  b += (int64_t)sizeof(DevTime) * 2 + 8 * 2 + 4 + 8 * (kEnergySums * (1 + 1024)) + (int64_t)sizeof(ModelDev);
  b += (int64_t)sizeof(Tile) * (th + tq + nsl);
  // or_model_upload_skyline: rest rows (at most every row of the family) and the prefix lists
  // (at most one prefix per node and family)
  for (int f = 0; f < 2; ++f) {
    const int32_t *rows = f == 0 ? rows_hex : rows_q4;
    if ((f == 0 ? ne_hex : ne_q4) == 0) continue;
    int64_t nrow = 0, nodes = 0;
    for (int64_t n = 0; n < nn; ++n) { nrow += rows[n]; nodes += rows[n] > 0; }
    b += ss * nrow * (f == 0 ? kHexRowWords : kQ4RowWords) + (4 + 2 * orasm::kRunCap) * nodes;
  }
  *bytes = b;
  return OR_OK;
}

int or_model_assembly_stats(const or_model *m, int64_t out[4]) {
  if (!m || !out) return fail(OR_ERR_ARG, "null argument");
  if (!m->have_sky) return fail(OR_ERR_STATE, "upload the skyline first");
  for (int k = 0; k < 4; ++k) out[k] = m->asm_stats[k];
  return OR_OK;
}

int or_model_footprint(const or_model *m, int64_t *bytes) {
  if (!m || !bytes) return fail(OR_ERR_ARG, "null argument");
  *bytes = m->bytes;
  return OR_OK;
}

// ============================================================== group drop-ins
namespace {
// Device staging of one group call: every buffer allocated, copied, freed here.
struct Stage {
  std::vector<void *> ptrs;
  hipError_t e = hipSuccess;
  ~Stage() {
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  }
  template <typename T> T *up(const T *h, size_t n) {
    T *d = nullptr;
    if (e == hipSuccess) e = hipMalloc((void **)&d, sizeof(T) * (n ? n : 1));
    ptrs.push_back(d);
    if (e == hipSuccess && h && n) e = hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice);
    return d;
  }
  template <typename T> void down(T *h, const T *d, size_t n) {
    if (e == hipSuccess && h && n) e = hipMemcpy(h, d, sizeof(T) * n, hipMemcpyDeviceToHost);
  }
};
} // namespace

int or_sforc3_group(int device, const or_hex_params *p, int64_t numnod, int32_t nel, int64_t nft, const double *x,
                    const double *v, const int32_t *ixs, double *sig, double *eint, double *rho, const double *vol,
                    double *pla, double *epsd, double *off, double *qvis, int64_t lsky, double *fsky,
                    const int32_t *iads, double dt1, double *dt2t, int32_t *mstop) {
  if (!p || !x || !v || !ixs || !sig || !eint || !rho || !vol || !pla || !epsd || !off || !qvis || !fsky || !iads ||
      !dt2t || !mstop)
    return fail(OR_ERR_ARG, "null argument");
  if (nel <= 0 || nel > 128) return fail(OR_ERR_ARG, "nel must be in 1..128 (MVSIZ-1)");
  if (numnod <= 0 || lsky <= 0 || nft < 0) return fail(OR_ERR_ARG, "bad sizes");
  int rc = check_hex_params(*p);
  if (rc) return rc;
  for (int i = 0; i < nel; ++i)
    for (int k = 0; k < 8; ++k) {
      const int32_t nid = ixs[11 * (nft + i) + 1 + k];
      if (nid < 1 || nid > numnod) return fail(OR_ERR_ARG, "IXS node id out of range");
      const int32_t K = iads[8 * (nft + i) + k];
      if (K < 1 || K > lsky) return fail(OR_ERR_ARG, "IADS row out of range");
    }
  OR_HIP(hipSetDevice(device));
  HexConsts c = make_hex_consts(*p);
  const int64_t ncol = nft + nel;
  Stage s;
  double *dx = s.up(x, 3 * numnod), *dv = s.up(v, 3 * numnod);
  int32_t *dixs = s.up(ixs, 11 * ncol), *diads = s.up(iads, 8 * ncol);
  double *dsig = s.up(sig, 6 * nel), *deint = s.up(eint, nel), *drho = s.up(rho, nel), *dvol = s.up(vol, nel),
         *dpla = s.up(pla, nel), *depsd = s.up(epsd, nel), *doff = s.up(off, nel), *dq = s.up(qvis, nel),
         *dfsky = s.up(fsky, 8 * lsky);
  double *ddtx = s.up<double>(nullptr, nel);
  int *dms = s.up<int>(nullptr, nel);
  if (s.e == hipSuccess) {
    hipLaunchKernelGGL(k_sforc3_group, dim3((nel + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, nel, nft, dx, dv, dixs,
                       dsig, deint, drho, dvol, dpla, depsd, doff, dq, dfsky, diads, c, dt1, ddtx, dms);
    s.e = hipGetLastError();
  }
  if (s.e == hipSuccess) s.e = hipDeviceSynchronize();
  std::vector<double> dtx((size_t)nel);
  std::vector<int> msv((size_t)nel);
  s.down(sig, dsig, 6 * nel); s.down(eint, deint, nel); s.down(rho, drho, nel); s.down(pla, dpla, nel);
  s.down(epsd, depsd, nel); s.down(off, doff, nel); s.down(qvis, dq, nel); s.down(fsky, dfsky, 8 * lsky);
  s.down(dtx.data(), ddtx, nel); s.down(msv.data(), dms, nel);
  if (s.e != hipSuccess) return fail(OR_ERR_HIP, std::string("or_sforc3_group: ") + hipGetErrorString(s.e));
  for (int i = 0; i < nel; ++i) {
    if (dtx[i] < *dt2t) *dt2t = dtx[i];
    if (msv[i]) *mstop = 1;
  }
  return OR_OK;
}

// ---------------------------------------------------------------------------
// SFORC3 group drop-in without allocation.  The group is staged through one file-scope
// device array instead of per-call device buffers: its own nodes (at most 8 per element,
// renumbered in order of first appearance), its GBUF arrays and its FSKY rows (8 per
// element, gathered from the caller's FSKY and scattered back, so that words the kernel
// does not write keep their values).  k_sforc3_group runs on the staged copy with
// nft = 0.  The array holds kStaticGroupBytes per device, for the whole process.
namespace {
constexpr int kStaticNel = 128;                                   // MVSIZ - 1
constexpr size_t kStaticGroupDoubles = 6 * 8 * kStaticNel         // x, v of the group's nodes
                                       + 13 * kStaticNel           // GBUF (SIG 6, 7 scalars)
                                       + 64 * kStaticNel           // FSKY rows (8 words x 8)
                                       + kStaticNel;               // element time steps
constexpr size_t kStaticGroupInts = 11 * kStaticNel + 8 * kStaticNel + kStaticNel; // IXS, IADS, MSTOP
constexpr size_t kStaticGroupBytes = 8 * kStaticGroupDoubles + 4 * kStaticGroupInts;

static __device__ __align__(256) unsigned char scratch[kStaticGroupBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (host code never uses the array's
// name as an address): looked up once per device, with that device current.
int scratch_of(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices) return fail(OR_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(OR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (unsigned char *)ptr;
  }
  *out = g_scratch[device];
  return OR_OK;
}
} // namespace

int or_sforc3_group_static(int device, const or_hex_params *p, int64_t numnod, int32_t nel, int64_t nft,
                           const double *x, const double *v, const int32_t *ixs, double *sig, double *eint,
                           double *rho, const double *vol, double *pla, double *epsd, double *off, double *qvis,
                           int64_t lsky, double *fsky, const int32_t *iads, double dt1, double *dt2t,
                           int32_t *mstop) {
  if (!p || !x || !v || !ixs || !sig || !eint || !rho || !vol || !pla || !epsd || !off || !qvis || !fsky || !iads ||
      !dt2t || !mstop)
    return fail(OR_ERR_ARG, "null argument");
  if (nel <= 0 || nel > kStaticNel)
    return fail(OR_ERR_ARG, "nel must be in 1..128: the staging array holds one group of at most 128 hexes");
  if (numnod <= 0 || lsky <= 0 || nft < 0) return fail(OR_ERR_ARG, "bad sizes");
  int rc = check_hex_params(*p);
  if (rc) return rc;
  for (int i = 0; i < nel; ++i)
    for (int k = 0; k < 8; ++k) {
      const int32_t nid = ixs[11 * (nft + i) + 1 + k];
      if (nid < 1 || nid > numnod) return fail(OR_ERR_ARG, "IXS node id out of range");
      const int32_t K = iads[8 * (nft + i) + k];
      if (K < 1 || K > lsky) return fail(OR_ERR_ARG, "IADS row out of range");
    }
  OR_HIP(hipSetDevice(device));
  unsigned char *base = nullptr;
  if ((rc = scratch_of(device, &base))) return rc;
  const HexConsts c = make_hex_consts(*p);
  // the staged group on the host, laid out as in the device array
  double hd[kStaticGroupDoubles];
  int32_t hi[kStaticGroupInts];
  double *lx = hd, *lv = hd + 24 * kStaticNel, *lsig = hd + 48 * kStaticNel;
  double *leint = lsig + 6 * nel, *lrho = leint + nel, *lvol = lrho + nel, *lpla = lvol + nel, *lepsd = lpla + nel,
         *loff = lepsd + nel, *lqvis = loff + nel;
  double *lfsky = hd + 61 * kStaticNel, *ldtx = hd + 125 * kStaticNel;
  int32_t *lixs = hi, *liads = hi + 11 * kStaticNel, *lms = hi + 19 * kStaticNel;
  int32_t nloc = 0;
  int32_t gid[8 * kStaticNel]; // global (1-based) id of each local node
  for (int i = 0; i < nel; ++i) {
    for (int k = 0; k < 8; ++k) {
      const int32_t g = ixs[11 * (nft + i) + 1 + k];
      int32_t l = 0;
      while (l < nloc && gid[l] != g) ++l;
      if (l == nloc) {
        gid[nloc++] = g;
        for (int d = 0; d < 3; ++d) {
          lx[3 * l + d] = x[3 * (int64_t)(g - 1) + d];
          lv[3 * l + d] = v[3 * (int64_t)(g - 1) + d];
        }
      }
      lixs[11 * i + 1 + k] = l + 1;
      liads[8 * i + k] = 8 * i + k + 1;
      const int64_t K = iads[8 * (nft + i) + k] - 1;
      for (int w = 0; w < 8; ++w) lfsky[8 * (8 * i + k) + w] = fsky[8 * K + w];
    }
    lixs[11 * i] = ixs[11 * (nft + i)];
    for (int w = 9; w < 11; ++w) lixs[11 * i + w] = ixs[11 * (nft + i) + w];
  }
  for (int k = 0; k < 6; ++k)
    for (int i = 0; i < nel; ++i) lsig[k * nel + i] = sig[(int64_t)k * nel + i];
  for (int i = 0; i < nel; ++i) {
    leint[i] = eint[i]; lrho[i] = rho[i]; lvol[i] = vol[i]; lpla[i] = pla[i];
    lepsd[i] = epsd[i]; loff[i] = off[i]; lqvis[i] = qvis[i];
  }
  double *dd = (double *)base;
  int32_t *di = (int32_t *)(base + 8 * kStaticGroupDoubles);
  OR_HIP(hipMemcpy(dd, hd, sizeof(hd), hipMemcpyHostToDevice));
  OR_HIP(hipMemcpy(di, hi, sizeof(hi), hipMemcpyHostToDevice));
  double *dsig = dd + 48 * kStaticNel;
  double *deint = dsig + 6 * nel, *drho = deint + nel, *dvol = drho + nel, *dpla = dvol + nel, *depsd = dpla + nel,
         *doff = depsd + nel, *dq = doff + nel;
  hipLaunchKernelGGL(k_sforc3_group, dim3(1), dim3(kBlock), 0, 0, nel, (int64_t)0, (const double *)dd,
                     (const double *)(dd + 24 * kStaticNel), (const int32_t *)di, dsig, deint, drho,
                     (const double *)dvol, dpla, depsd, doff, dq, dd + 61 * kStaticNel,
                     (const int32_t *)(di + 11 * kStaticNel), c, dt1, dd + 125 * kStaticNel, (int *)(di + 19 * kStaticNel));
  OR_HIP(hipGetLastError());
  OR_HIP(hipMemcpy(hd, dd, sizeof(hd), hipMemcpyDeviceToHost)); // synchronous after the launch
  OR_HIP(hipMemcpy(hi, di, sizeof(hi), hipMemcpyDeviceToHost));
  for (int k = 0; k < 6; ++k)
    for (int i = 0; i < nel; ++i) sig[(int64_t)k * nel + i] = lsig[k * nel + i];
  for (int i = 0; i < nel; ++i) {
    eint[i] = leint[i]; rho[i] = lrho[i]; pla[i] = lpla[i]; epsd[i] = lepsd[i]; off[i] = loff[i];
    qvis[i] = lqvis[i];
    for (int k = 0; k < 8; ++k) {
      const int64_t K = iads[8 * (nft + i) + k] - 1;
      for (int w = 0; w < 8; ++w) fsky[8 * K + w] = lfsky[8 * (8 * i + k) + w];
    }
    if (ldtx[i] < *dt2t) *dt2t = ldtx[i];
    if (lms[i]) *mstop = 1;
  }
  return OR_OK;
}

int or_cforc3_group(int device, const or_q4_params *p, int64_t numnod, int32_t nel, int64_t nft, const double *x,
                    const double *v, const double *vr, const int32_t *ixc, double *offg, double *thk,
                    const double *thke, double *forc, double *mom, double *gstr, double *hourg, double *eint,
                    double *epsd, double *sig, double *pla, double *epsd_ip, int64_t lsky, double *fsky,
                    const int32_t *iadc, double dt1, double *dt2t) {
  if (!p || !x || !v || !vr || !ixc || !offg || !thk || !thke || !forc || !mom || !gstr || !hourg || !eint ||
      !epsd || !sig || !pla || !epsd_ip || !fsky || !iadc || !dt2t)
    return fail(OR_ERR_ARG, "null argument");
  if (nel <= 0 || nel > 128) return fail(OR_ERR_ARG, "nel must be in 1..128 (MVSIZ-1)");
  if (numnod <= 0 || lsky <= 0 || nft < 0) return fail(OR_ERR_ARG, "bad sizes");
  Q4Consts c;
  int rc = make_q4_consts(*p, c);
  if (rc) return rc;
  for (int i = 0; i < nel; ++i)
    for (int k = 0; k < 4; ++k) {
      const int32_t nid = ixc[7 * (nft + i) + 1 + k];
      if (nid < 1 || nid > numnod) return fail(OR_ERR_ARG, "IXC node id out of range");
      const int32_t K = iadc[4 * (nft + i) + k];
      if (K < 1 || K > lsky) return fail(OR_ERR_ARG, "IADC row out of range");
    }
  OR_HIP(hipSetDevice(device));
  const int64_t ncol = nft + nel, np = p->npt;
  Stage s;
  double *dx = s.up(x, 3 * numnod), *dv = s.up(v, 3 * numnod), *dvr = s.up(vr, 3 * numnod);
  int32_t *dixc = s.up(ixc, 7 * ncol), *diadc = s.up(iadc, 4 * ncol);
  double *doffg = s.up(offg, nel), *dthk = s.up(thk, nel), *dthke = s.up(thke, nel), *dforc = s.up(forc, 5 * nel),
         *dmom = s.up(mom, 3 * nel), *dgstr = s.up(gstr, 8 * nel), *dhourg = s.up(hourg, 5 * nel),
         *deint = s.up(eint, 2 * nel), *depsd = s.up(epsd, nel), *dsig = s.up(sig, np * 5 * nel),
         *dpla = s.up(pla, np * nel), *depi = s.up(epsd_ip, np * nel), *dfsky = s.up(fsky, 8 * lsky);
  double *ddt = s.up<double>(nullptr, nel);
  if (s.e == hipSuccess) {
    hipLaunchKernelGGL(k_cforc3_group, dim3((nel + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, nel, nft, dx, dv, dvr,
                       dixc, doffg, dthk, dthke, dforc, dmom, dgstr, dhourg, deint, depsd, dsig, dpla, depi, dfsky,
                       diadc, c, dt1, ddt);
    s.e = hipGetLastError();
  }
  if (s.e == hipSuccess) s.e = hipDeviceSynchronize();
  std::vector<double> dtv((size_t)nel);
  s.down(offg, doffg, nel); s.down(thk, dthk, nel); s.down(forc, dforc, 5 * nel); s.down(mom, dmom, 3 * nel);
  s.down(gstr, dgstr, 8 * nel); s.down(hourg, dhourg, 5 * nel); s.down(eint, deint, 2 * nel);
  s.down(epsd, depsd, nel); s.down(sig, dsig, np * 5 * nel); s.down(pla, dpla, np * nel);
  s.down(epsd_ip, depi, np * nel); s.down(fsky, dfsky, 8 * lsky); s.down(dtv.data(), ddt, nel);
  if (s.e != hipSuccess) return fail(OR_ERR_HIP, std::string("or_cforc3_group: ") + hipGetErrorString(s.e));
  for (int i = 0; i < nel; ++i)
    if (dtv[i] < *dt2t) *dt2t = dtv[i];
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Device evaluation of the FP64 elementary functions (diagnostic).
namespace {
__global__ void k_math_eval(int fn, int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                            double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i];
  out[i] = fn == 0 ? ormath::or_log(a) : fn == 1 ? ormath::or_exp(a) : ormath::or_pow(a, y[i]);
}
} // namespace

int or_math_eval(int device, int fn, int64_t n, const double *x, const double *y, double *out) {
  if (fn < 0 || fn > 2) return fail(OR_ERR_ARG, "fn must be 0 (log), 1 (exp) or 2 (pow)");
  if (n < 0 || n > (int64_t(1) << 30)) return fail(OR_ERR_ARG, "n must be in 0..2^30");
  if (n == 0) return OR_OK;
  if (!x || !out || (fn == 2 && !y)) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(device));
  Stage s;
  double *dx = s.up(x, n), *dy = fn == 2 ? s.up(y, n) : nullptr, *dout = s.up<double>(nullptr, n);
  if (s.e == hipSuccess) {
    hipLaunchKernelGGL(k_math_eval, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0, fn, n, dx, dy,
                       dout);
    s.e = hipGetLastError();
  }
  if (s.e == hipSuccess) s.e = hipDeviceSynchronize();
  s.down(out, dout, n);
  if (s.e != hipSuccess) return fail(OR_ERR_HIP, std::string("or_math_eval: ") + hipGetErrorString(s.e));
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Batches: many resident models in one launch per kernel and precision class.
namespace {

int read_time(or_model *m, int slot, DevTime &t) {
  OR_HIP(hipMemcpy(&t, m->ts + slot, sizeof(DevTime), hipMemcpyDeviceToHost));
  return OR_OK;
}

// Move a model's cycle scalars to the slot of the given parity, and set its stop flag
// for the end time tend.
int align_model(or_model *m, int64_t parity, double tend) {
  const int from = (int)(m->parity & 1), to = (int)(parity & 1);
  DevTime t;
  int rc = read_time(m, from, t);
  if (rc) return rc;
  t.done = (t.mstop != 0 || t.tt >= tend) ? 1 : 0;
  OR_HIP(hipMemcpy(m->ts + to, &t, sizeof(DevTime), hipMemcpyHostToDevice));
  unsigned long long dm;
  const double big = kEP06;
  std::memcpy(&dm, &big, 8);
  OR_HIP(hipMemcpy(m->dtmin + to, &dm, sizeof(dm), hipMemcpyHostToDevice));
  m->parity = parity & 1;
  return OR_OK;
}

int batch_rebuild(or_batch *b) {
  if (!b->dirty) return OR_OK;
  OR_HIP(hipStreamSynchronize(b->stream));
  const int64_t ns = (int64_t)b->members.size();
  if (ns > b->md_cap) {
    if (b->md) OR_HIP(hipFree(b->md));
    b->md = nullptr;
    b->md_cap = 0;
    OR_HIP(hipMalloc((void **)&b->md, sizeof(ModelDev) * (size_t)ns));
    b->md_cap = ns;
  }
  std::vector<ModelDev> hd((size_t)ns);
  std::vector<Tile> tl[2][3];
  for (int64_t i = 0; i < ns; ++i) {
    or_model *m = b->members[(size_t)i];
    fill_dev(m, hd[(size_t)i], b->tend[(size_t)i]);
    if (!b->live[(size_t)i]) continue;
    const int c = m->prec == OR_PREC_FP64 ? 0 : 1;
    const int64_t cnt[3] = {m->neh, m->neq, m->nn};
    const int step[3] = {kTile, kTile, kBlock};
    for (int f = 0; f < 3; ++f)
      for (int64_t first = 0; first < cnt[f]; first += step[f]) tl[c][f].push_back(Tile{(int32_t)i, (int32_t)first});
  }
  if (ns > 0) OR_HIP(hipMemcpy(b->md, hd.data(), sizeof(ModelDev) * (size_t)ns, hipMemcpyHostToDevice));
  for (int c = 0; c < 2; ++c)
    for (int f = 0; f < 3; ++f) {
      const int64_t n = (int64_t)tl[c][f].size();
      if (n > b->tcap[c][f]) {
        if (b->tiles[c][f]) OR_HIP(hipFree(b->tiles[c][f]));
        b->tiles[c][f] = nullptr;
        b->tcap[c][f] = 0;
        OR_HIP(hipMalloc((void **)&b->tiles[c][f], sizeof(Tile) * (size_t)n));
        b->tcap[c][f] = n;
      }
      if (n > 0) OR_HIP(hipMemcpy(b->tiles[c][f], tl[c][f].data(), sizeof(Tile) * (size_t)n, hipMemcpyHostToDevice));
      b->nt[c][f] = (int32_t)n;
    }
  b->dirty = false;
  return OR_OK;
}

} // namespace

int or_batch_create(or_batch **out, int device) {
  if (!out) return fail(OR_ERR_ARG, "null out");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(OR_ERR_NODEV, "no HIP device");
  if (device < 0 || device >= ndev) return fail(OR_ERR_ARG, "device index out of range");
  OR_HIP(hipSetDevice(device));
  or_batch *b = new or_batch();
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(OR_ERR_HIP, "hipStreamCreate");
  }
  *out = b;
  return OR_OK;
}

int or_batch_destroy(or_batch *b) {
  if (!b) return OR_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  while (!b->members.empty()) (void)or_batch_remove(b, b->members.back());
  if (b->md) (void)hipFree(b->md);
  for (auto &row : b->tiles)
    for (Tile *t : row)
      if (t) (void)hipFree(t);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return OR_OK;
}

int or_batch_add(or_batch *b, or_model *m, double t_end) {
  if (!b || !m) return fail(OR_ERR_ARG, "null argument");
  if (m->batch) return fail(OR_ERR_STATE, "the model already belongs to a batch");
  if (m->device != b->device) return fail(OR_ERR_ARG, "the model lives on another device");
  if (!(t_end > 0.0)) return fail(OR_ERR_ARG, "t_end must be > 0");
  if (!m->have_nodes || !m->have_sky || (m->neh > 0 && (!m->have_hparams || !m->have_hex)) ||
      (m->neq > 0 && (!m->have_qparams || !m->have_q4)))
    return fail(OR_ERR_STATE, "params, nodes, element state and skyline must be uploaded before add");
  OR_HIP(hipSetDevice(b->device));
  OR_HIP(hipStreamSynchronize(m->stream));
  OR_HIP(hipStreamSynchronize(b->stream));
  int rc = align_model(m, b->parity, t_end);
  if (rc) return rc;
  m->batch = b;
  b->members.push_back(m);
  b->tend.push_back(t_end);
  b->live.push_back(1);
  b->dirty = true;
  return OR_OK;
}

int or_batch_remove(or_batch *b, or_model *m) {
  if (!b || !m) return fail(OR_ERR_ARG, "null argument");
  size_t i = 0;
  while (i < b->members.size() && b->members[i] != m) ++i;
  if (i == b->members.size()) return fail(OR_ERR_ARG, "the model is not in this batch");
  OR_HIP(hipSetDevice(b->device));
  OR_HIP(hipStreamSynchronize(b->stream));
  b->members.erase(b->members.begin() + (ptrdiff_t)i);
  b->tend.erase(b->tend.begin() + (ptrdiff_t)i);
  b->live.erase(b->live.begin() + (ptrdiff_t)i);
  m->batch = nullptr;
  m->dev_dirty = true;
  b->dirty = true;
  // alone again: the model runs without an end time unless MSTOP was raised
  return align_model(m, m->parity, kNoEnd);
}

int or_batch_run(or_batch *b, int64_t ncycles) {
  if (!b) return fail(OR_ERR_ARG, "null batch");
  if (ncycles < 0) return fail(OR_ERR_ARG, "ncycles < 0");
  OR_HIP(hipSetDevice(b->device));
  int rc = batch_rebuild(b);
  if (rc) return rc;
  LaunchSet L[2];
  for (int c = 0; c < 2; ++c) {
    L[c].md = b->md;
    L[c].th = b->tiles[c][0]; L[c].nth = b->nt[c][0]; L[c].vh = b->var[c][0];
    L[c].tq = b->tiles[c][1]; L[c].ntq = b->nt[c][1]; L[c].vq = b->var[c][1];
    L[c].tn = b->tiles[c][2]; L[c].ntn = b->nt[c][2];
  }
  for (int64_t k = 0; k < ncycles; ++k) {
    const int p = (int)(b->parity & 1);
    if ((rc = launch_cycle<double, double>(b->stream, L[0], p, nullptr, nullptr))) return rc;
    if ((rc = launch_cycle<float, float>(b->stream, L[1], p, nullptr, nullptr))) return rc;
    b->parity ^= 1;
    b->cycles += 1;
  }
  // members still in the tile lists follow the batch parity; stopped ones keep their slot
  for (size_t i = 0; i < b->members.size(); ++i)
    if (b->live[i]) b->members[i]->parity = b->parity & 1;
  return OR_OK;
}

int or_batch_sync(or_batch *b) {
  if (!b) return fail(OR_ERR_ARG, "null batch");
  OR_HIP(hipSetDevice(b->device));
  OR_HIP(hipStreamSynchronize(b->stream));
  return OR_OK;
}

int or_batch_active(or_batch *b, int32_t *nactive) {
  if (!b || !nactive) return fail(OR_ERR_ARG, "null argument");
  OR_HIP(hipSetDevice(b->device));
  OR_HIP(hipStreamSynchronize(b->stream));
  int32_t n = 0;
  for (size_t i = 0; i < b->members.size(); ++i) {
    DevTime t;
    int rc = read_time(b->members[i], (int)(b->members[i]->parity & 1), t);
    if (rc) return rc;
    const char alive = t.done ? 0 : 1;
    if (alive != b->live[i]) {
      b->live[i] = alive;
      b->dirty = true; // stopped models leave the tile lists
    }
    n += alive;
  }
  *nactive = n;
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Kernel variants and their autotuning
namespace {

// Device time of `ncycles` cycles launched by `launch` on `st` (events around the loop).
template <typename F> int time_cycles(hipStream_t st, int ncycles, F launch, float &ms) {
  hipEvent_t e0, e1;
  OR_HIP(hipEventCreate(&e0));
  OR_HIP(hipEventCreate(&e1));
  int rc = OR_OK;
  if (hipEventRecord(e0, st) != hipSuccess) rc = fail(OR_ERR_HIP, "hipEventRecord");
  for (int k = 0; k < ncycles && rc == OR_OK; ++k) rc = launch();
  if (rc == OR_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(OR_ERR_HIP, "event timing");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

} // namespace

int or_model_set_variant(or_model *m, int family, int variant) {
  if (!m) return fail(OR_ERR_ARG, "null model");
  if (family < 0 || family > 1 || variant < 0 || variant >= kVariants) return fail(OR_ERR_ARG, "bad family / variant");
  OR_HIP(hipSetDevice(m->device));
  OR_HIP(quiesce(m));
  (family == 0 ? m->var_h : m->var_q) = variant;
  return put_elem_tiles(m);
}

int or_model_autotune(or_model *m, int32_t cycles_per_variant, int32_t chosen[2]) {
  if (!m || !chosen) return fail(OR_ERR_ARG, "null argument");
  if (cycles_per_variant < 1) return fail(OR_ERR_ARG, "cycles_per_variant must be >= 1");
  if (m->batch) return fail(OR_ERR_STATE, "the model belongs to a batch: use or_batch_autotune");
  if (!m->have_nodes || !m->have_sky || (m->neh > 0 && (!m->have_hparams || !m->have_hex)) ||
      (m->neq > 0 && (!m->have_qparams || !m->have_q4)))
    return fail(OR_ERR_STATE, "params, nodes, element state and skyline must be uploaded before autotune");
  OR_HIP(hipSetDevice(m->device));
  int rc = refresh_dev(m);
  if (rc) return rc;
  for (int f = 0; f < 2; ++f) {
    if ((f == 0 ? m->neh : m->neq) == 0) continue;
    int best = 0;
    float best_ms = 0.f;
    for (int v = 0; v < kVariants; ++v) {
      if ((rc = or_model_set_variant(m, f, v))) return rc;
      LaunchSet L;
      L.md = m->dev;
      L.th = m->tiles_h; L.nth = m->nt_h; L.vh = m->var_h;
      L.eh = m->ehou_h != nullptr; L.eq = m->ehou_q != nullptr;
      L.tq = m->tiles_q; L.ntq = m->nt_q; L.vq = m->var_q;
      L.tn = m->tiles_n; L.ntn = m->nt_n;
      float ms = 0.f;
      auto one = [&]() -> int {
        const int p = (int)(m->parity & 1);
        int r = (m->prec == OR_PREC_FP64) ? launch_cycle<double, double>(m->stream, L, p, nullptr, nullptr)
                                          : launch_cycle<float, float>(m->stream, L, p, nullptr, nullptr);
        m->parity ^= 1;
        return r;
      };
      if ((rc = time_cycles(m->stream, cycles_per_variant, one, ms))) return rc;
      if (v == 0 || ms < best_ms) { best = v; best_ms = ms; }
    }
    if ((rc = or_model_set_variant(m, f, best))) return rc;
  }
  chosen[0] = m->var_h;
  chosen[1] = m->var_q;
  return OR_OK;
}

int or_batch_autotune(or_batch *b, int32_t cycles_per_variant, int32_t chosen[4]) {
  if (!b || !chosen) return fail(OR_ERR_ARG, "null argument");
  if (cycles_per_variant < 1) return fail(OR_ERR_ARG, "cycles_per_variant must be >= 1");
  OR_HIP(hipSetDevice(b->device));
  int rc;
  for (int c = 0; c < 2; ++c)
    for (int f = 0; f < 2; ++f) {
      if ((rc = batch_rebuild(b))) return rc;
      if (b->nt[c][f] == 0) continue;
      int best = 0;
      float best_ms = 0.f;
      for (int v = 0; v < kVariants; ++v) {
        b->var[c][f] = v;
        b->dirty = true;
        if ((rc = batch_rebuild(b))) return rc;
        float ms = 0.f;
        auto one = [&]() { return or_batch_run(b, 1); };
        if ((rc = time_cycles(b->stream, cycles_per_variant, one, ms))) return rc;
        if (v == 0 || ms < best_ms) { best = v; best_ms = ms; }
      }
      b->var[c][f] = best;
      b->dirty = true;
    }
  if ((rc = batch_rebuild(b))) return rc;
  for (int c = 0; c < 2; ++c)
    for (int f = 0; f < 2; ++f) chosen[2 * c + f] = b->var[c][f];
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Nodal half of a cycle with host arrays (ASSPAR4 -> DEPLA drop-in)
int or_nodal_update(int device, int64_t numnod, int64_t lsky, const int64_t *adsky, const double *fsky,
                    const double *ms, const double *in, const int32_t *code, const double *vimp, double *x,
                    double *v, double *vr, double *d, double *stifn, double *stifr, double dt2t, double dtmx,
                    or_time_state *ts) {
  if (!adsky || !fsky || !ms || !x || !v || !d || !ts) return fail(OR_ERR_ARG, "null argument");
  if (vr && !in) return fail(OR_ERR_ARG, "rotational velocities need the nodal inertias IN");
  if (numnod <= 0 || lsky < 0) return fail(OR_ERR_ARG, "bad sizes");
  if (adsky[0] != 1 || adsky[numnod] != lsky + 1) return fail(OR_ERR_ARG, "ADSKY must run from 1 to LSKY+1");
  for (int64_t n = 0; n < numnod; ++n) {
    if (adsky[n + 1] < adsky[n]) return fail(OR_ERR_ARG, "ADSKY must be non-decreasing");
    if (code && (code[n] & ~511)) return fail(OR_ERR_ARG, "node code has bits outside 0..511");
    if (code && (code[n] & (kIX | kIY | kIZ)) && !vimp) return fail(OR_ERR_ARG, "imposed-velocity bits without vimp");
  }
  OR_HIP(hipSetDevice(device));
  // DT2 pipeline on the host (resol.F:2721-2722, 6125-6129, 6352, 6494-6497)
  const double DT1 = ts->dt2;
  const double DT2 = dt2_pipeline(dt2t, ts->dt2old, (double)1.1f, dtmx);
  const double DT12 = (1.0 / 2.0) * (DT1 + DT2);
  Stage s;
  const int64_t *dads = s.up(adsky, numnod + 1);
  const double *dfsky = s.up(fsky, 8 * lsky), *dms = s.up(ms, numnod);
  const double *din = in ? s.up(in, numnod) : nullptr;
  const int32_t *dcode = code ? s.up(code, numnod) : nullptr;
  const double *dvimp = vimp ? s.up(vimp, 3 * numnod) : nullptr;
  double *dx = s.up(x, 3 * numnod), *dv = s.up(v, 3 * numnod), *dd = s.up(d, 3 * numnod);
  double *dvr = vr ? s.up(vr, 3 * numnod) : nullptr;
  double *dsn = stifn ? s.up<double>(nullptr, numnod) : nullptr;
  double *dsr = stifr ? s.up<double>(nullptr, numnod) : nullptr;
  if (s.e == hipSuccess) {
    hipLaunchKernelGGL(k_nodal_dropin, dim3((unsigned)((numnod + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0, numnod,
                       dads, dfsky, dms, din, dcode, dvimp, dx, dv, dvr, dd, dsn, dsr, DT12, DT2);
    s.e = hipGetLastError();
  }
  if (s.e == hipSuccess) s.e = hipDeviceSynchronize();
  s.down(x, dx, 3 * numnod); s.down(v, dv, 3 * numnod); s.down(d, dd, 3 * numnod);
  if (vr) s.down(vr, dvr, 3 * numnod);
  if (stifn) s.down(stifn, dsn, numnod);
  if (stifr) s.down(stifr, dsr, numnod);
  if (s.e != hipSuccess) return fail(OR_ERR_HIP, std::string("or_nodal_update: ") + hipGetErrorString(s.e));
  ts->dt1 = DT1;
  ts->dt2 = DT2;
  ts->dt12 = DT12;
  ts->dt2old = DT2;
  ts->tt = ts->tt + DT2;
  ts->ncycle += 1;
  if (DT2 <= 0.0) ts->mstop |= kMstopTimeStep; // resol.F:8575, 9599
  return OR_OK;
}
#endif // OR_ILP_TU
