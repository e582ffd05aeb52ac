// layout.cpp -- block-local layout of a partitioned mesh.
//
// What the reference builds in PartitionTets / GetTetMem / GetVertMem
// (meshFIM3dEikonal.cu:568-923) and in PartitionFaces / GenerateData (meshFIM2d_eikonal.cu:
// 254-614): every element is copied into each block that owns one of its corners (owned +
// halo-shared elements), and each block knows where its vertices' values live.  Here the
// replicated per-corner value slots (tetT / triMem, reconciled by slot lists and CopyOutBack)
// are replaced by one value per vertex:
//   * vertices are renumbered block after block (owned interior vertices, then owned boundary
//     vertices, i.e. those another block reads), so a block's owned values are one contiguous
//     range of the global time array;
//   * a block's elements address vertices by 16-bit local index: [0, nown) owned,
//     [nown, nown + nhalo) halo, the halo being the block's list of foreign vertex ids;
//   * an element carries the squared metric lengths of its edges (6 per tet, 3 per triangle)
//     instead of the reference's 6 Gram products (gettetmem, meshFIM3dEikonal.cu:650-678) or
//     3 edge lengths + speedInv (meshFIM2d_eikonal.cu:412-420);
//   * block dependencies (who must be revisited when a block's boundary values change) are
//     CSR lists, the reference's m_BlockNeighbor sets (meshFIM3dEikonal.cu:633-645).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

#include "eik_internal.h"

namespace eik {

// Runs fn(i, worker) for i in [0, n) on up to 16 host threads (dynamic chunks of `grain`).
template <class Fn>
static void parallel_for(int64_t n, int64_t grain, Fn&& fn) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nt = (int)std::min<int64_t>(hw, std::max<int64_t>(1, (n + grain - 1) / grain));
  if (nt <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(i, 0);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&](int w) {
    for (;;) {
      const int64_t i0 = next.fetch_add(grain);
      if (i0 >= n) return;
      const int64_t i1 = std::min(n, i0 + grain);
      for (int64_t i = i0; i < i1; ++i) fn(i, w);
    }
  };
  std::vector<std::thread> th;
  for (int w = 1; w < nt; ++w) th.emplace_back(work, w);
  work(0);
  for (auto& t : th) t.join();
}

// Corner heights from the (fp32) squared edge lengths the device uses, in fp64, rounded down
// with a margin so that min(other times) + H never exceeds the exact local-solve minimum.
static float round_down(double h) {
  if (!(h > 0.0)) return 0.f;
  float f = (float)(h * (1.0 - 1e-5));
  return f > 0.f ? f : 0.f;
}

// tet: h_k^2 = det(Gram of the edges from k) / det(Gram of two edges of the opposite face)
void tet_heights(const float* L6, float* H) {
  // squared length of edge (a, b), a != b
  auto len = [&](int a, int b) -> double {
    if (a > b) std::swap(a, b);
    static const int idx[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
    return (double)L6[idx[a][b]];
  };
  for (int k = 0; k < 4; ++k) {
    int o[3], n = 0;
    for (int j = 0; j < 4; ++j)
      if (j != k) o[n++] = j;
    // Gram of e_i = X_o[i] - X_k
    double g[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        g[i][j] = i == j ? len(k, o[i]) : 0.5 * (len(k, o[i]) + len(k, o[j]) - len(o[i], o[j]));
    const double det3 = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]) -
                        g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]) +
                        g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
    // Gram of the face edges f1 = X_o1 - X_o0, f2 = X_o2 - X_o0
    const double a = len(o[0], o[1]), b = len(o[0], o[2]);
    const double ab = 0.5 * (a + b - len(o[1], o[2]));
    const double det2 = a * b - ab * ab;
    H[k] = det2 > 0.0 && det3 > 0.0 ? round_down(std::sqrt(det3 / det2)) : 0.f;
  }
}

// triangle: h_k^2 = det(Gram of the two edges from k) / |opposite edge|^2
void tri_heights(const float* L3, float* H) {
  // L3 = {01, 12, 20}
  const double l01 = L3[0], l12 = L3[1], l20 = L3[2];
  const double len[3][3] = {{0, l01, l20}, {l01, 0, l12}, {l20, l12, 0}};
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const double ga = len[k][a], gb = len[k][b];
    const double gab = 0.5 * (ga + gb - len[a][b]);
    const double det = ga * gb - gab * gab;
    H[k] = det > 0.0 && len[a][b] > 0.0 ? round_down(std::sqrt(det / len[a][b])) : 0.f;
  }
}

int build_layout(Mesh& m) {
  if (m.nblocks <= 0 || (int64_t)m.label.size() != m.nv)
    return fail(EIK_ERR_STATE, "mesh is not partitioned");
  const int k = m.dim == 3 ? 4 : 3;
  const int32_t nb = m.nblocks;
  const int64_t nv = m.nv;
  for (int64_t v = 0; v < nv; ++v)
    if (m.label[v] < 0 || m.label[v] >= nb) return fail(EIK_ERR_ARG, "partition label out of range");
  if (m.dim == 2) tri_virtual_faces(m);
  const int64_t nvf = m.dim == 2 ? (int64_t)m.vfaces.size() / 3 : 0;
  Layout L;
  L.nblocks = nb;

  // 1. boundary vertices: corners of an element that spans several blocks, and the non-obtuse
  //    corners of a virtual face that sits in another block
  std::vector<uint8_t> bnd(nv, 0);
  for (int64_t e = 0; e < m.ne; ++e) {
    const int32_t* c = &m.elem[k * e];
    bool multi = false;
    for (int j = 1; j < k; ++j) multi |= m.label[c[j]] != m.label[c[0]];
    if (multi)
      for (int j = 0; j < k; ++j) bnd[c[j]] = 1;
  }
  for (int64_t e = 0; e < nvf; ++e) {
    const int32_t* c = &m.vfaces[3 * e];
    for (int j = 1; j < 3; ++j)
      if (m.label[c[j]] != m.label[c[0]]) bnd[c[j]] = 1;
  }

  // 2. renumbering: block by block, interior before boundary, caller order within each class
  std::vector<int32_t> cnt(nb + 1, 0), cntb(nb, 0);
  for (int64_t v = 0; v < nv; ++v) {
    cnt[m.label[v] + 1]++;
    if (bnd[v]) cntb[m.label[v]]++;
  }
  for (int32_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
  L.perm.assign(nv, 0);
  L.iperm.assign(nv, 0);
  L.owner.assign(nv, 0);
  {
    std::vector<int32_t> fi(nb), fb(nb);
    for (int32_t b = 0; b < nb; ++b) {
      fi[b] = cnt[b];
      fb[b] = cnt[b + 1] - cntb[b];
    }
    for (int64_t v = 0; v < nv; ++v) {
      int32_t b = m.label[v];
      int32_t id = bnd[v] ? fb[b]++ : fi[b]++;
      L.perm[id] = (int32_t)v;
      L.iperm[v] = id;
      L.owner[id] = b;
    }
  }

  // 3. element lists per block: element e goes to every block owning one of its corners;
  //    a virtual face goes to its obtuse vertex's block only (meshFIM2d_eikonal.cu:281-289)
  const int64_t ntot = m.ne + nvf;
  std::vector<int64_t> eoff(nb + 1, 0);
  auto elem_corners = [&](int64_t e) -> const int32_t* {
    return e < m.ne ? &m.elem[k * e] : &m.vfaces[3 * (e - m.ne)];
  };
  auto for_blocks = [&](int64_t e, auto&& fn) {
    const int32_t* c = elem_corners(e);
    if (e >= m.ne) {
      fn(m.label[c[0]]);
      return;
    }
    int32_t seen[4];
    int ns = 0;
    for (int j = 0; j < k; ++j) {
      int32_t b = m.label[c[j]];
      bool dup = false;
      for (int q = 0; q < ns; ++q) dup |= seen[q] == b;
      if (!dup) {
        seen[ns++] = b;
        fn(b);
      }
    }
  };
  for (int64_t e = 0; e < ntot; ++e) for_blocks(e, [&](int32_t b) { eoff[b + 1]++; });
  for (int32_t b = 0; b < nb; ++b) eoff[b + 1] += eoff[b];
  const int64_t total = eoff[nb];
  if (total >= (int64_t)INT32_MAX) return fail(EIK_ERR_CAPACITY, "too many block elements");
  std::vector<int64_t> elist(total);
  {
    std::vector<int64_t> fill(eoff.begin(), eoff.end() - 1);
    for (int64_t e = 0; e < ntot; ++e) for_blocks(e, [&](int32_t b) { elist[fill[b]++] = e; });
  }
  L.total_elems = total;

  // 4. per element (once, in parallel): squared metric edge lengths and corner heights; then
  //    per block (in parallel, each worker with its own vertex map): halo list, local corner
  //    indices and the records; the halo lists are concatenated in block order
  const int nl = m.dim == 3 ? 6 : 3;
  std::vector<float> eL((size_t)ntot * nl), eH((size_t)ntot * k);
  parallel_for(ntot, 4096, [&](int64_t e, int) {
    const int32_t* c = elem_corners(e);
    float* Le = &eL[(size_t)e * nl];
    float* He = &eH[(size_t)e * k];
    if (m.dim == 3) {
      static const int pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
      for (int p = 0; p < 6; ++p) Le[p] = (float)edge_len2(m, e, pairs[p][0], pairs[p][1]);
      tet_heights(Le, He);
    } else {
      const int64_t src = e < m.ne ? e : m.vorigin[e - m.ne];
      for (int p = 0; p < 3; ++p) Le[p] = (float)tri_edge_len2(m, src, c, p, (p + 1) % 3);
      tri_heights(Le, He);
    }
  });
  L.blocks.resize(nb);
  if (m.dim == 3) L.tets.resize(total); else L.tris.resize(total);
  L.rec_elem.resize(total);
  for (int64_t q = 0; q < total; ++q)
    L.rec_elem[q] = (int32_t)(elist[q] < m.ne ? elist[q] : m.vorigin[elist[q] - m.ne]);
  const unsigned nworkers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<int32_t>> local(nworkers);  // new id -> local index in the block
  std::vector<std::vector<int32_t>> bhalo(nb);
  std::vector<int32_t> bad(nb, 0);  // 1: too many local vertices, 2: too many elements
  parallel_for(nb, 16, [&](int64_t bi, int w) {
    const int32_t b = (int32_t)bi;
    std::vector<int32_t>& loc = local[w];
    if (loc.empty()) loc.assign(nv, -1);
    BlockDesc& d = L.blocks[b];
    d.vstart = cnt[b];
    d.nown = cnt[b + 1] - cnt[b];
    d.nbnd0 = d.nown - cntb[b];
    d.estart = (int32_t)eoff[b];
    d.nelem = (int32_t)(eoff[b + 1] - eoff[b]);
    d.pad = 0;
    std::vector<int32_t>& halo = bhalo[b];
    for (int32_t i = 0; i < d.nown; ++i) loc[d.vstart + i] = i;
    int32_t nloc = d.nown;
    for (int64_t q = eoff[b]; q < eoff[b + 1]; ++q) {
      const int64_t e = elist[q];
      const int32_t* c = elem_corners(e);
      int32_t lv[4] = {0, 0, 0, 0};
      for (int j = 0; j < k; ++j) {
        const int32_t id = L.iperm[c[j]];
        if (loc[id] < 0) {
          loc[id] = nloc++;
          halo.push_back(id);
        }
        lv[j] = loc[id];
      }
      if (nloc > 65535) {
        bad[b] = 1;
        break;
      }
      const float* Le = &eL[(size_t)e * nl];
      const float* He = &eH[(size_t)e * k];
      if (m.dim == 3) {
        TetRec& r = L.tets[q];
        for (int j = 0; j < 4; ++j) r.v[j] = (uint16_t)lv[j];
        for (int p = 0; p < 6; ++p) r.L[p] = Le[p];
        for (int j = 0; j < 4; ++j) r.H[j] = He[j];
      } else {
        TriRec& r = L.tris[q];
        for (int j = 0; j < 3; ++j) r.v[j] = (uint16_t)lv[j];
        r.mask = e < m.ne ? 0x7 : 0x1;
        for (int p = 0; p < 3; ++p) r.L[p] = Le[p];
        for (int j = 0; j < 3; ++j) r.H[j] = He[j];
      }
    }
    d.nhalo = nloc - d.nown;
    if (!bad[b] && d.nelem > 16383) bad[b] = 2;
    // reset the map for the worker's next block
    for (int32_t i = 0; i < d.nown; ++i) loc[d.vstart + i] = -1;
    for (int32_t id : halo) loc[id] = -1;
  });
  for (int32_t b = 0; b < nb; ++b) {
    if (bad[b] == 1)
      return fail(EIK_ERR_CAPACITY, "block " + std::to_string(b) +
                                        " has more than 65535 local vertices; use smaller blocks");
    if (bad[b] == 2)
      return fail(EIK_ERR_CAPACITY, "block " + std::to_string(b) +
                                        " has more than 16383 elements; use smaller blocks");
  }
  {
    int64_t nh = 0;
    for (int32_t b = 0; b < nb; ++b) nh += (int64_t)bhalo[b].size();
    L.halo.resize(nh);
    int64_t h0 = 0;
    for (int32_t b = 0; b < nb; ++b) {
      BlockDesc& d = L.blocks[b];
      d.hstart = (int32_t)h0;
      std::copy(bhalo[b].begin(), bhalo[b].end(), L.halo.begin() + h0);
      h0 += (int64_t)bhalo[b].size();
      L.max_local = std::max(L.max_local, d.nown + d.nhalo);
      L.max_own = std::max(L.max_own, d.nown);
      L.max_elem = std::max(L.max_elem, d.nelem);
    }
  }

  // 5. suppliers (owners of a block's halo) and dependents (the transpose)
  {
    std::vector<int32_t> stamp(nb, -1);
    L.sup_start.assign(nb + 1, 0);
    for (int32_t c = 0; c < nb; ++c) {
      const BlockDesc& d = L.blocks[c];
      for (int32_t h = d.hstart; h < d.hstart + d.nhalo; ++h) {
        int32_t b = L.owner[L.halo[h]];
        if (stamp[b] != c) {
          stamp[b] = c;
          L.sup.push_back(b);
        }
      }
      L.sup_start[c + 1] = (int32_t)L.sup.size();
    }
    L.dep_start.assign(nb + 1, 0);
    for (int32_t b : L.sup) L.dep_start[b + 1]++;
    for (int32_t b = 0; b < nb; ++b) L.dep_start[b + 1] += L.dep_start[b];
    L.dep.assign(L.sup.size(), 0);
    std::vector<int32_t> fill(L.dep_start.begin(), L.dep_start.end() - 1);
    for (int32_t c = 0; c < nb; ++c)
      for (int32_t s = L.sup_start[c]; s < L.sup_start[c + 1]; ++s) L.dep[fill[L.sup[s]]++] = c;
  }
  // 6. per boundary vertex: which of its owner's dependents hold it in their halo
  {
    for (int32_t b = 0; b < nb; ++b)
      L.max_dep = std::max(L.max_dep, L.dep_start[b + 1] - L.dep_start[b]);
    if (L.max_dep > 65535) return fail(EIK_ERR_CAPACITY, "a block has more than 65535 dependents");
    std::vector<int32_t> cntv(nv + 1, 0);
    for (int32_t c = 0; c < nb; ++c) {
      const BlockDesc& d = L.blocks[c];
      for (int32_t h = d.hstart; h < d.hstart + d.nhalo; ++h) cntv[L.halo[h] + 1]++;
    }
    for (int64_t v = 0; v < nv; ++v) cntv[v + 1] += cntv[v];
    L.vh_start = cntv;
    L.vh_pos.assign(cntv[nv], 0);
    std::vector<int32_t> fill(cntv.begin(), cntv.end() - 1);
    // position of block c in the dependent list of block b
    std::vector<int32_t> pos(nb, -1);
    for (int32_t b = 0; b < nb; ++b) {
      for (int32_t q = L.dep_start[b]; q < L.dep_start[b + 1]; ++q) pos[L.dep[q]] = q - L.dep_start[b];
      // halo entries of the dependents that refer to vertices of b
      for (int32_t q = L.dep_start[b]; q < L.dep_start[b + 1]; ++q) {
        const BlockDesc& d = L.blocks[L.dep[q]];
        for (int32_t h = d.hstart; h < d.hstart + d.nhalo; ++h) {
          int32_t v = L.halo[h];
          if (L.owner[v] == b) L.vh_pos[fill[v]++] = (uint16_t)pos[L.dep[q]];
        }
      }
      for (int32_t q = L.dep_start[b]; q < L.dep_start[b + 1]; ++q) pos[L.dep[q]] = -1;
    }
    // the same as one 64-bit mask per vertex when every block has at most 64 dependents
    L.vmask.clear();
    if (L.max_dep <= 64) {
      L.vmask.assign(nv, 0);
      for (int64_t v = 0; v < nv; ++v)
        for (int32_t p = L.vh_start[v]; p < L.vh_start[v + 1]; ++p)
          L.vmask[v] |= 1ull << L.vh_pos[p];
    }
  }
  L.valid = true;
  m.layout = std::move(L);
  m.stats_valid = true;
  m.device_dirty = true;
  return EIK_OK;
}

}  // namespace eik

namespace eik {

float half_value(uint16_t h) {
  const int e = (h >> 10) & 0x1f, m = h & 0x3ff;
  if (e == 0) return std::ldexp((float)m, -24);
  return std::ldexp(1.0f + (float)m / 1024.0f, e - 15);
}

uint16_t half_below(float x) {
  if (!(x > 0.f)) return 0;
  if (x >= 65504.f) return 0x7bff;
  int k = 0;
  const double f = std::frexp((double)x, &k);  // x = f 2^k, f in [0.5, 1)
  const int e = k - 1;                          // x = (2f) 2^e, 2f in [1, 2)
  uint16_t h;
  if (e >= -14) {
    const int m = (int)std::floor((2.0 * f - 1.0) * 1024.0);
    h = (uint16_t)(((e + 15) << 10) | m);
  } else {
    h = (uint16_t)std::floor(std::ldexp((double)x, 24));  // subnormal: multiples of 2^-24
  }
  while (h > 0 && half_value(h) > x) --h;  // never above x
  return h;
}

void split_records(const Layout& L, int dim, std::vector<uint32_t>& sel, std::vector<float>& len) {
  const int64_t n = L.total_elems;
  sel.assign(4 * n, 0u);
  len.assign((dim == 3 ? 6 : 4) * n, 0.f);
  for (int64_t e = 0; e < n; ++e) {
    uint32_t* s = &sel[4 * e];
    if (dim == 3) {
      const TetRec& r = L.tets[e];
      s[0] = (uint32_t)r.v[0] | ((uint32_t)r.v[1] << 16);
      s[1] = (uint32_t)r.v[2] | ((uint32_t)r.v[3] << 16);
      s[2] = (uint32_t)half_below(r.H[0]) | ((uint32_t)half_below(r.H[1]) << 16);
      s[3] = (uint32_t)half_below(r.H[2]) | ((uint32_t)half_below(r.H[3]) << 16);
      for (int q = 0; q < 6; ++q) len[6 * e + q] = r.L[q];
    } else {
      const TriRec& r = L.tris[e];
      s[0] = (uint32_t)r.v[0] | ((uint32_t)r.v[1] << 16);
      s[1] = (uint32_t)r.v[2] | ((uint32_t)r.mask << 16);
      s[2] = (uint32_t)half_below(r.H[0]) | ((uint32_t)half_below(r.H[1]) << 16);
      s[3] = (uint32_t)half_below(r.H[2]);
      for (int q = 0; q < 3; ++q) len[4 * e + q] = r.L[q];
    }
  }
}

}  // namespace eik
