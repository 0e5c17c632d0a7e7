// Host simulation of the tiled assembly (openradioss_amd/csrc/or_assembly.h), for the CPU
// tests: the element kernels' LDS corners, prefixes and rest rows, and the node kernel's
// sum of them, in the same arithmetic as the kernels; compared by the tests with the plain
// ASSPAR4 sum of the rows in skyline order.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../openradioss_amd/csrc/or_assembly.h"

extern "C" {

// conn [ne][nk] (model element order), x [nn][3] or null, rows of the family per node in
// skyline order: rptr [nn+1], relem, rcorner; prefix_ok [nn]; vals [ne][nk][W] corner words
// (model element order, FP64); init [nn][W] or null: the running sum a node without a prefix
// starts from (the sum of the rows of earlier families).  Outputs: sum [nn][W] and stats [3] =
// (prefixes, rest rows, rows in all).  The element order is tile_order's, as in the library.
int asm_simulate(int nk, int W, int64_t ne, int64_t nn, const int32_t *conn, const double *x, const int64_t *rptr,
                 const int32_t *relem, const int8_t *rcorner, const char *prefix_ok, const double *vals,
                 const double *init, double *sum, int64_t *stats) {
  std::vector<int32_t> perm = orasm::tile_order(nk, ne, conn, x);
  std::vector<int64_t> rp(rptr, rptr + nn + 1);
  std::vector<int32_t> re(relem, relem + rp[nn]);
  std::vector<int8_t> rc(rcorner, rcorner + rp[nn]);
  std::vector<char> ok(prefix_ok, prefix_ok + nn);
  orasm::FamilyPlan p;
  orasm::plan_family(p, nk, ne, nn, perm, rp, re, rc, ok);
  const int64_t ntiles = (ne + orasm::kTile - 1) / orasm::kTile;
  std::vector<double> pre((size_t)(nn * W), 0.0), rows((size_t)(p.nrest * W), 0.0);
  std::vector<double> lds((size_t)(nk * W * orasm::kTile));
  for (int64_t t = 0; t < ntiles; ++t) {
    // element phase: every corner of the tile's elements into LDS; rest rows from LDS
    for (int e = 0; e < orasm::kTile; ++e) {
      const int64_t s = t * orasm::kTile + e;
      if (s >= ne) break;
      const int64_t el = perm[(size_t)s];
      for (int k = 0; k < nk; ++k)
        for (int c = 0; c < W; ++c) lds[(size_t)((k * W + c) * orasm::kTile + e)] = vals[(el * nk + k) * W + c];
      for (int k = 0; k < nk; ++k) {
        const int32_t r = p.rest[(size_t)(k * ne + s)];
        if (r >= 0)
          for (int c = 0; c < W; ++c) rows[(size_t)(r * W + c)] = lds[(size_t)((k * W + c) * orasm::kTile + e)];
      }
    }
    // prefixes of the tile, in row order from zero
    for (int32_t j = p.fr_start[(size_t)t]; j < p.fr_start[(size_t)t + 1]; ++j) {
      double a[6] = {0, 0, 0, 0, 0, 0};
      for (int r = 0; r < orasm::kRunCap; ++r) {
        const uint16_t u = p.fr_src[(size_t)(j * orasm::kRunCap + r)];
        if (u == orasm::kNoSrc) break;
        for (int c = 0; c < W; ++c) a[c] = a[c] + lds[(size_t)(((u >> 8) * W + c) * orasm::kTile + (u & 0xFF))];
      }
      for (int c = 0; c < W; ++c) pre[(size_t)(p.fr_node[(size_t)j] * W + c)] = a[c];
    }
  }
  // node phase: the prefix (when the family starts the node's sum), then the rest rows
  for (int64_t n = 0; n < nn; ++n) {
    double a[6] = {0, 0, 0, 0, 0, 0};
    if (rptr[n + 1] > rptr[n] && prefix_ok[n])
      for (int c = 0; c < W; ++c) a[c] = pre[(size_t)(n * W + c)];
    else if (init)
      for (int c = 0; c < W; ++c) a[c] = init[n * W + c];
    for (int32_t r = p.rs[(size_t)n]; r < p.rs[(size_t)n + 1]; ++r)
      for (int c = 0; c < W; ++c) a[c] = a[c] + rows[(size_t)(r * W + c)];
    for (int c = 0; c < W; ++c) sum[n * W + c] = a[c];
  }
  stats[0] = (int64_t)p.fr_node.size();
  stats[1] = p.nrest;
  stats[2] = rptr[nn];
  return 0;
}
}
