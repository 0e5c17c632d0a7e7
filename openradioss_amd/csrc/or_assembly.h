// or_assembly.h -- host-side plan of the tiled Parith/ON assembly (no device code).
//
// ASSPAR4 (asspar4.F:164-180) sums the skyline rows of a node in FILLCNE order
// (starter domdec2.F:2046-2240: solids before shells, ascending user element id
// inside a family), from zero, one row after the other.  Floating-point addition
// is not associative, so any partial sum that is reused must be a prefix of that
// order started from zero.
//
// The resident model stores the elements of a family in tiles of kTile elements
// (one workgroup of the element kernel).  For every node, the leading run of its
// rows whose elements all lie in the tile of its first element (at most kRunCap of
// them) is summed in that tile's LDS, in row order, from zero: that is exactly the
// running sum ASSPAR4 holds after those rows.  The tile writes it once per node
// (the node's prefix); the remaining rows of the node are written as rows and added
// to the prefix by the node kernel, in row order.  Same operations, same order,
// same bits, with far fewer row words through HBM.
//
// Tiles are formed from a space-filling order: elements sorted by the Morton code
// of their centroid on a grid of the median element size, cut into kTile runs,
// ascending element index inside each tile.  On a structured block this gives
// 8 x 8 x 4 bricks; any order is correct, only the share of rows summed in LDS
// depends on it.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace orasm {

constexpr int kTile = 256;          // elements per tile (the element kernel's workgroup)
constexpr int kRunCap = 8;          // rows of a node summed in LDS, at most
constexpr uint16_t kNoSrc = 0xFFFF; // unused entry of a prefix's source list

// Storage order of a family's elements: slot -> element.  conn is [ne][nk] in element
// order; x is [nn][3] or null (then the element order is kept).
inline std::vector<int32_t> tile_order(int nk, int64_t ne, const int32_t *conn, const double *x) {
  std::vector<int32_t> perm((size_t)ne);
  std::iota(perm.begin(), perm.end(), 0);
  if (!x || ne <= kTile) return perm;
  std::vector<double> c((size_t)(3 * ne));
  std::vector<double> h((size_t)ne);
  for (int64_t e = 0; e < ne; ++e) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < nk; ++k)
      for (int d = 0; d < 3; ++d) s[d] += x[3 * (int64_t)conn[nk * e + k] + d];
    for (int d = 0; d < 3; ++d) c[(size_t)(3 * e + d)] = s[d] / nk;
    // element size: the diagonal from local node 1 to the opposite node, over sqrt(dim)
    const int64_t a = conn[nk * e], b = conn[nk * e + (nk == 8 ? 6 : 2)];
    double dd = 0.0;
    for (int d = 0; d < 3; ++d) dd += (x[3 * b + d] - x[3 * a + d]) * (x[3 * b + d] - x[3 * a + d]);
    h[(size_t)e] = std::sqrt(dd) / std::sqrt(nk == 8 ? 3.0 : 2.0);
  }
  std::vector<double> hs(h);
  std::nth_element(hs.begin(), hs.begin() + ne / 2, hs.end());
  double q = hs[(size_t)(ne / 2)];
  double lo[3] = {c[0], c[1], c[2]};
  for (int64_t e = 0; e < ne; ++e)
    for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], c[(size_t)(3 * e + d)]);
  if (!(q > 0.0) || !std::isfinite(q)) return perm;
  auto spread = [](uint64_t v) { // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  std::vector<uint64_t> key((size_t)ne);
  for (int64_t e = 0; e < ne; ++e) {
    uint64_t g[3];
    for (int d = 0; d < 3; ++d) {
      double t = std::floor((c[(size_t)(3 * e + d)] - lo[d]) / q);
      g[d] = (uint64_t)std::min(std::max(t, 0.0), 2097151.0);
    }
    key[(size_t)e] = spread(g[0]) | spread(g[1]) << 1 | spread(g[2]) << 2;
  }
  std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return key[(size_t)a] < key[(size_t)b]; });
  for (int64_t t = 0; t < ne; t += kTile) std::sort(perm.begin() + t, perm.begin() + std::min(ne, t + kTile));
  return perm;
}

// The assembly plan of one family of a model.
struct FamilyPlan {
  std::vector<int32_t> fr_start; // [ntiles + 1] first prefix entry of each tile
  std::vector<int32_t> fr_node;  // [nfr] node of each prefix
  std::vector<uint16_t> fr_src;  // [nfr][kRunCap] LDS sources in row order: corner << 8 | element in tile
  std::vector<int32_t> rest;     // [nk][ne] per corner of a storage slot: its rest row, or -1 (summed in LDS)
  std::vector<int32_t> rs;       // [nn + 1] first rest row of each node
  int64_t nrest = 0;             // rest rows in all
};

// rows: the node's rows of this family in ASSPAR4 order, as CSR over nodes (rptr [nn + 1])
// of (element, corner) pairs in model element numbering.  prefix_ok[n] == 0 when rows
// of an earlier family precede them (the node's sum does not start with this family).
inline void plan_family(FamilyPlan &p, int nk, int64_t ne, int64_t nn, const std::vector<int32_t> &perm,
                        const std::vector<int64_t> &rptr, const std::vector<int32_t> &relem,
                        const std::vector<int8_t> &rcorner, const std::vector<char> &prefix_ok) {
  std::vector<int32_t> slot((size_t)ne);
  for (int64_t s = 0; s < ne; ++s) slot[(size_t)perm[(size_t)s]] = (int32_t)s;
  const int64_t ntiles = (ne + kTile - 1) / kTile;
  p.rest.assign((size_t)(nk * ne), -1);
  p.rs.assign((size_t)(nn + 1), 0);
  std::vector<int32_t> run((size_t)nn, 0), cnt((size_t)ntiles + 1, 0);
  for (int64_t n = 0; n < nn; ++n) {
    const int64_t a = rptr[(size_t)n], b = rptr[(size_t)n + 1];
    int32_t L = 0;
    if (b > a && prefix_ok[(size_t)n]) {
      const int32_t t0 = slot[(size_t)relem[(size_t)a]] / kTile;
      while (a + L < b && L < kRunCap && slot[(size_t)relem[(size_t)(a + L)]] / kTile == t0) ++L;
      cnt[(size_t)t0 + 1] += 1;
    }
    run[(size_t)n] = L;
    p.rs[(size_t)n + 1] = p.rs[(size_t)n] + (int32_t)(b - a - L);
  }
  p.nrest = p.rs[(size_t)nn];
  p.fr_start.assign((size_t)ntiles + 1, 0);
  for (int64_t t = 0; t < ntiles; ++t) p.fr_start[(size_t)t + 1] = p.fr_start[(size_t)t] + cnt[(size_t)t + 1];
  const int64_t nfr = p.fr_start[(size_t)ntiles];
  p.fr_node.assign((size_t)nfr, 0);
  p.fr_src.assign((size_t)(nfr * kRunCap), kNoSrc);
  std::vector<int32_t> fill(p.fr_start.begin(), p.fr_start.end() - 1);
  for (int64_t n = 0; n < nn; ++n) {
    const int64_t a = rptr[(size_t)n], b = rptr[(size_t)n + 1];
    const int32_t L = run[(size_t)n];
    if (L > 0) {
      const int32_t t0 = slot[(size_t)relem[(size_t)a]] / kTile;
      const int32_t j = fill[(size_t)t0]++;
      p.fr_node[(size_t)j] = (int32_t)n;
      for (int32_t r = 0; r < L; ++r) {
        const int32_t s = slot[(size_t)relem[(size_t)(a + r)]];
        p.fr_src[(size_t)(j * kRunCap + r)] = (uint16_t)((int)rcorner[(size_t)(a + r)] << 8 | (s % kTile));
      }
    }
    for (int64_t r = a + L; r < b; ++r) {
      const int32_t s = slot[(size_t)relem[(size_t)r]];
      p.rest[(size_t)((int64_t)rcorner[(size_t)r] * ne + s)] = p.rs[(size_t)n] + (int32_t)(r - a - L);
    }
  }
}

} // namespace orasm
