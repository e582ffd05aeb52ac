// partition.cpp -- vertex partitions into blocks: recursive coordinate bisection (in place of
// the reference's METIS k-way call) and the reference's structured tiling.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "eik_internal.h"

namespace eik {

// Recursive coordinate bisection into exactly `nblocks` blocks.  Each range is cut across its
// longest bounding-box axis.  Near the cut that splits the vertex count in proportion to the
// blocks on each side, the cut goes through the widest gap between consecutive coordinates, as
// long as every block on both sides keeps within kSlack of the mesh-wide block size: on a
// lattice (jittered or not) the cuts then fall between vertex planes instead of through one,
// and the blocks are lattice-aligned boxes with the least halo (a 65^3 cube at 64 vertices per
// block: 1.92 -> 1.75x element copies, halo vertices per block 149 -> 119); on unstructured
// meshes the cut prefers sparse regions.
// Replaces METIS_PartGraphKway at meshFIM3dEikonal.cu:202-203 / meshFIM2d_eikonal.cu:161-162:
// the converged times do not depend on the partition, only the halo size and the number of
// block visits do.
int partition_rcb(Mesh& m, int32_t nblocks) {
  if (nblocks < 1 || nblocks > m.nv) return fail(EIK_ERR_ARG, "partition: bad block count");
  constexpr double kSlack = 0.15;
  const double S = (double)m.nv / nblocks;  // mesh-wide block size
  std::vector<int32_t> idx(m.nv);
  std::iota(idx.begin(), idx.end(), 0);
  m.label.assign(m.nv, 0);
  struct Range {
    int64_t lo, hi;
    int32_t b0, nb;
  };
  std::vector<Range> stack = {{0, m.nv, 0, nblocks}};
  const float* X = m.xyz.data();
  while (!stack.empty()) {
    Range r = stack.back();
    stack.pop_back();
    if (r.nb == 1) {
      for (int64_t i = r.lo; i < r.hi; ++i) m.label[idx[i]] = r.b0;
      continue;
    }
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = r.lo; i < r.hi; ++i)
      for (int c = 0; c < 3; ++c) {
        float x = X[3 * (int64_t)idx[i] + c];
        lo[c] = std::min(lo[c], x);
        hi[c] = std::max(hi[c], x);
      }
    int axis = 0;
    for (int c = 1; c < 3; ++c)
      if (hi[c] - lo[c] > hi[axis] - lo[axis]) axis = c;
    auto less = [&](int32_t a, int32_t b) {
      float xa = X[3 * (int64_t)a + axis], xb = X[3 * (int64_t)b + axis];
      return xa < xb || (xa == xb && a < b);
    };
    const int64_t n = r.hi - r.lo;
    // candidate splits (nb1 blocks on the low side, k vertices): nb1 within two of half, every
    // block on both sides within kSlack of S on average
    const int32_t half = r.nb / 2;
    int64_t wlo = n, whi = 0;
    struct Cand {
      int32_t nb1;
      int64_t k0, k1;
    };
    std::vector<Cand> cands;
    for (int32_t nb1 = std::max(1, half - 2); nb1 <= std::min(r.nb - 1, half + 2); ++nb1) {
      const int32_t nb2 = r.nb - nb1;
      int64_t k0 = std::max<int64_t>((int64_t)std::ceil(nb1 * S * (1 - kSlack)),
                                     n - (int64_t)std::floor(nb2 * S * (1 + kSlack)));
      int64_t k1 = std::min<int64_t>((int64_t)std::floor(nb1 * S * (1 + kSlack)),
                                     n - (int64_t)std::ceil(nb2 * S * (1 - kSlack)));
      k0 = std::max<int64_t>(k0, nb1);
      k1 = std::min<int64_t>(k1, n - nb2);
      if (k0 > k1) continue;
      cands.push_back({nb1, k0, k1});
      wlo = std::min(wlo, k0);
      whi = std::max(whi, k1);
    }
    int32_t nb1 = half;
    int64_t k = (n * half + r.nb / 2) / r.nb;  // proportional cut (always allowed)
    k = std::max<int64_t>(half, std::min<int64_t>(k, n - (r.nb - half)));
    if (!cands.empty() && wlo >= 1 && whi < n) {
      // order the window [wlo - 1, whi] along the axis: everything below it first
      auto b = idx.begin() + r.lo;
      std::nth_element(b, b + (wlo - 1), b + n, less);
      std::nth_element(b + wlo, b + whi, b + n, less);
      std::sort(b + wlo, b + whi, less);
      auto coord = [&](int64_t i) { return X[3 * (int64_t)idx[r.lo + i] + axis]; };
      // the gaps that separate vertex planes are the ones at least half as wide as the widest;
      // of those, the cut nearest the proportional split
      double widest = 0.0;
      for (int64_t kk = wlo; kk <= whi; ++kk)
        widest = std::max(widest, (double)coord(kk) - (double)coord(kk - 1));
      double bestd = INFINITY;
      for (const Cand& c : cands) {
        const double kp = (double)n * c.nb1 / r.nb;
        for (int64_t kk = c.k0; kk <= c.k1; ++kk) {
          const double gap = (double)coord(kk) - (double)coord(kk - 1);
          const double d = std::fabs(kk - kp) + S * std::abs(c.nb1 - half);
          if (gap >= 0.5 * widest && d < bestd) {
            bestd = d;
            k = kk;
            nb1 = c.nb1;
          }
        }
      }
    } else {
      auto mid = idx.begin() + r.lo + k;
      std::nth_element(idx.begin() + r.lo, mid, idx.begin() + r.hi, less);
    }
    stack.push_back({r.lo + k, r.hi, r.b0 + nb1, r.nb - nb1});
    stack.push_back({r.lo, r.lo + k, r.b0, nb1});
  }
  m.nblocks = nblocks;
  m.layout.valid = false;
  m.stats_valid = false;
  m.device_dirty = true;
  return EIK_OK;
}

// The reference's structured tiling (meshFIM3dEikonal.cu:239-283, meshFIM2d_eikonal.cu:206-252).
int partition_structured(Mesh& m, int L, int W, int D, int bl, int bw, int bd) {
  if (L <= 0 || W <= 0 || D <= 0 || bl <= 0 || bw <= 0 || bd <= 0)
    return fail(EIK_ERR_ARG, "structured partition: dimensions must be positive");
  if ((int64_t)L * W * D != m.nv)
    return fail(EIK_ERR_ARG, "structured partition: L*W*D != number of vertices");
  int nbl = L / bl, nbw = W / bw, nbd = D / bd;
  if (nbl <= 0 || nbw <= 0 || nbd <= 0)
    return fail(EIK_ERR_ARG, "structured partition: block larger than the grid");
  // the reference labels with integer division; a grid that is not a multiple of the block
  // size gives the trailing vertices labels past nbl*nbw*nbd, which it cannot represent
  if (L % bl || W % bw || D % bd)
    return fail(EIK_ERR_ARG, "structured partition: grid must be a multiple of the block size");
  m.label.assign(m.nv, 0);
  for (int k = 0; k < D; ++k)
    for (int i = 0; i < W; ++i)
      for (int j = 0; j < L; ++j) {
        int64_t v = (int64_t)k * L * W + (int64_t)i * L + j;
        m.label[v] = (k / bd) * nbl * nbw + (i / bw) * nbl + (j / bl);
      }
  m.nblocks = nbl * nbw * nbd;
  m.layout.valid = false;
  m.stats_valid = false;
  m.device_dirty = true;
  return EIK_OK;
}

int partition_by_size(Mesh& m, int32_t verts_per_block) {
  if (verts_per_block <= 0) return fail(EIK_ERR_ARG, "verts_per_block must be > 0");
  int64_t nb;
  if (m.dim == 3)  // ceil (meshFIM3dEikonal.cu:160)
    nb = (m.nv + verts_per_block - 1) / verts_per_block;
  else  // floor (meshFIM2d_eikonal.cu:119)
    nb = m.nv / verts_per_block;
  if (nb < 2) nb = 2;  // both drivers force at least 2 blocks
  if (nb > m.nv) nb = m.nv;
  return partition_rcb(m, (int32_t)nb);
}

}  // namespace eik
