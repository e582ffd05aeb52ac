// mesh.cpp -- mesh model: validation, file readers, synthetic generators, edge metrics and the
// 2D virtual faces for obtuse corners.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "eik_internal.h"

namespace eik {

int mesh_validate(const Mesh& m) {
  const int k = m.dim == 3 ? 4 : 3;
  if (m.nv <= 0) return fail(EIK_ERR_ARG, "mesh has no vertices");
  if (m.nv >= (int64_t)INT32_MAX) return fail(EIK_ERR_ARG, "too many vertices");
  if ((int64_t)m.elem.size() != m.ne * k) return fail(EIK_ERR_ARG, "element array size mismatch");
  for (int64_t i = 0; i < m.ne * k; ++i)
    if (m.elem[i] < 0 || m.elem[i] >= m.nv)
      return fail(EIK_ERR_ARG, "element references vertex " + std::to_string(m.elem[i]) +
                                   " outside [0, " + std::to_string(m.nv) + ")");
  for (size_t i = 0; i < m.xyz.size(); ++i)
    if (!std::isfinite(m.xyz[i])) return fail(EIK_ERR_ARG, "non-finite vertex coordinate");
  return EIK_OK;
}

// -------------------------------------------------------------------------------------------
// Edge metrics
// -------------------------------------------------------------------------------------------

// Squared length of edge (a, b) of tet e under its metric: (x_b - x_a)^T M (x_b - x_a), M stored
// upper triangular [0 1 2; 1 3 4; 2 4 5] (the reference's vMv, Vec.h:479-483).  Computed in
// double; the layout stores it as float.
double edge_len2(const Mesh& m, int64_t e, int32_t a, int32_t b) {
  const int32_t* t = &m.elem[4 * e];
  const float* p = &m.xyz[3 * (int64_t)t[a]];
  const float* q = &m.xyz[3 * (int64_t)t[b]];
  double d[3] = {(double)q[0] - p[0], (double)q[1] - p[1], (double)q[2] - p[2]};
  double M[6] = {1, 0, 0, 1, 0, 1};
  if (m.metric_kind == 1) {
    M[0] = M[3] = M[5] = m.metric[e];
  } else if (m.metric_kind == 2) {
    for (int i = 0; i < 6; ++i) M[i] = m.metric[6 * e + i];
  }
  return (d[0] * M[0] + d[1] * M[1] + d[2] * M[2]) * d[0] +
         (d[0] * M[1] + d[1] * M[3] + d[2] * M[4]) * d[1] +
         (d[0] * M[2] + d[1] * M[4] + d[2] * M[5]) * d[2];
}

// Squared metric length of edge (corners[a], corners[b]) of a triangle whose speed / metric is
// that of face `face`.  Isotropic: speedInv^2 |e|^2 (TriMesh_stats.cu:64-68 gives speedInv =
// 1/speed).  Anisotropic: e^T M e with M in the frame of `face`'s own corners.
double tri_edge_len2(const Mesh& m, int64_t face, const int32_t* corners, int32_t a, int32_t b) {
  const float* p = &m.xyz[3 * (int64_t)corners[a]];
  const float* q = &m.xyz[3 * (int64_t)corners[b]];
  double d[3] = {(double)q[0] - p[0], (double)q[1] - p[1], (double)q[2] - p[2]};
  if (m.metric_kind == 2) {
    const int32_t* f = &m.elem[3 * face];
    const float* v0 = &m.xyz[3 * (int64_t)f[0]];
    const float* v1 = &m.xyz[3 * (int64_t)f[1]];
    const float* v2 = &m.xyz[3 * (int64_t)f[2]];
    double e1[3], e2[3], n[3], t1[3], t2[3];
    for (int i = 0; i < 3; ++i) {
      e1[i] = (double)v1[i] - v0[i];
      e2[i] = (double)v2[i] - v0[i];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double n1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int i = 0; i < 3; ++i) {
      n[i] /= nn;
      t1[i] = e1[i] / n1;
    }
    t2[0] = n[1] * t1[2] - n[2] * t1[1];
    t2[1] = n[2] * t1[0] - n[0] * t1[2];
    t2[2] = n[0] * t1[1] - n[1] * t1[0];
    double x = d[0] * t1[0] + d[1] * t1[1] + d[2] * t1[2];
    double y = d[0] * t2[0] + d[1] * t2[1] + d[2] * t2[2];
    const float* M = &m.metric[3 * face];
    return x * x * M[0] + 2.0 * x * y * M[1] + y * y * M[2];
  }
  double F = m.metric_kind == 1 ? (double)m.metric[face] : 1.0;
  return F * F * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// -------------------------------------------------------------------------------------------
// Virtual faces (TriMesh::IsNonObtuse / SplitFace, src/core2d/include/TriMesh.h:191-313;
// need_face_virtual_faces, src/core2d/cuda/TriMesh_connectivity.cu:206-235)
// -------------------------------------------------------------------------------------------
static inline int tri_index(const int32_t* f, int32_t v) {
  return f[0] == v ? 0 : f[1] == v ? 1 : f[2] == v ? 2 : -1;
}

static inline float fdist(const float* a, const float* b) {
  float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// The angle at v is computed in float and compared with pi/2 in double, as the reference does;
// float rounding decides right angles the same way.
bool tri_is_non_obtuse(const float* xyz, const int32_t* f, int32_t v) {
  int iv = tri_index(f, v);
  const float* A = xyz + 3 * (int64_t)v;
  const float* B = xyz + 3 * (int64_t)f[(iv + 1) % 3];
  const float* C = xyz + 3 * (int64_t)f[(iv + 2) % 3];
  float a = fdist(B, C), b = fdist(A, C), c = fdist(A, B);
  float ang = 0.0f;
  if (a > 0 && b > 0 && c > 0) ang = std::acos((b * b + c * c - a * a) / (2 * b * c));
  return (double)ang < 3.14159265358979323846 / 2.0f;
}

namespace {
struct Splitter {
  const Mesh& m;
  const std::vector<int32_t>& across;
  std::vector<int32_t>& rows;
  std::vector<int32_t>& origin;
  int32_t face;
  int guard = 0;

  void split(int32_t v, const int32_t* cf, int32_t adj) {
    if (++guard > 4096) return;  // degenerate fans: stop instead of recursing forever
    int iv = tri_index(cf, v);
    int32_t v1 = v, v2 = cf[(iv + 1) % 3], v4 = cf[(iv + 2) % 3];
    const int32_t* fa = &m.elem[3 * (int64_t)adj];
    int32_t v3 = fa[(tri_index(fa, v2) + 1) % 3];
    const int32_t f1[3] = {v1, v3, v4}, f2[3] = {v1, v2, v3};
    for (int side = 0; side < 2; ++side) {
      const int32_t* nf = side == 0 ? f1 : f2;
      if (tri_is_non_obtuse(m.xyz.data(), nf, v)) {
        rows.insert(rows.end(), nf, nf + 3);
        origin.push_back(face);
      } else {
        int32_t opp = side == 0 ? v2 : v4;
        int32_t next = across[3 * (int64_t)adj + tri_index(fa, opp)];
        if (next > -1) split(v, nf, next);
      }
    }
  }
};
}  // namespace

void tri_virtual_faces(Mesh& m) {
  if (m.vfaces_ready) return;
  m.vfaces.clear();
  m.vorigin.clear();
  const int64_t nf = m.ne;
  // vertex -> incident faces
  std::vector<int32_t> off(m.nv + 1, 0), inc(3 * nf);
  for (int64_t i = 0; i < 3 * nf; ++i) off[m.elem[i] + 1]++;
  for (int64_t v = 0; v < m.nv; ++v) off[v + 1] += off[v];
  {
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t i = 0; i < 3 * nf; ++i) inc[fill[m.elem[i]]++] = (int32_t)(i / 3);
  }
  // across-edge map (TriMesh_connectivity.cu:240-278)
  std::vector<int32_t> across(3 * nf, -1);
  for (int64_t i = 0; i < nf; ++i) {
    for (int j = 0; j < 3; ++j) {
      if (across[3 * i + j] != -1) continue;
      int32_t v1 = m.elem[3 * i + (j + 1) % 3], v2 = m.elem[3 * i + (j + 2) % 3];
      for (int32_t a = off[v1]; a < off[v1 + 1]; ++a) {
        int32_t other = inc[a];
        if (other == i) continue;
        const int32_t* of = &m.elem[3 * (int64_t)other];
        if (tri_index(of, v2) < 0) continue;
        int ind = (tri_index(of, v1) + 1) % 3;
        if (of[(ind + 1) % 3] != v2) continue;
        across[3 * i + j] = other;
        across[3 * (int64_t)other + ind] = (int32_t)i;
        break;
      }
    }
  }
  for (int64_t i = 0; i < nf; ++i) {
    const int32_t* f = &m.elem[3 * i];
    for (int j = 0; j < 3; ++j) {
      if (tri_is_non_obtuse(m.xyz.data(), f, f[j])) continue;
      int32_t adj = across[3 * i + j];
      if (adj < 0) continue;  // obtuse corner on a boundary edge: no virtual face
      Splitter s{m, across, m.vfaces, m.vorigin, (int32_t)i};
      s.split(f[j], f, adj);
    }
  }
  m.vfaces_ready = true;
}

// -------------------------------------------------------------------------------------------
// Readers
// -------------------------------------------------------------------------------------------
namespace {
// next non-empty, non-comment line of a TetGen file, tokenised
bool next_numbers(std::istream& in, std::vector<std::string>& tok) {
  std::string line;
  while (std::getline(in, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream ss(line);
    tok.clear();
    std::string t;
    while (ss >> t) tok.push_back(t);
    if (!tok.empty()) return true;
  }
  return false;
}
}  // namespace

// TetGen .node/.ele (the subset load_tetmesh reads for the solve: src/core3d/cuda/tetgen.cu:
// 1993-2150): node header "n dim attrs markers", node rows "id x y z ...", element header
// "n corners attrs" (corners 4 or 10, the first 4 are the tet), element rows "id c0 c1 c2 c3 ...".
int load_tetgen(const std::string& base, Mesh& m) {
  std::ifstream nf(base + ".node"), ef(base + ".ele");
  if (!nf) return fail(EIK_ERR_IO, "cannot access file " + base + ".node");
  if (!ef) return fail(EIK_ERR_IO, "cannot access file " + base + ".ele");
  std::vector<std::string> tok;
  if (!next_numbers(nf, tok)) return fail(EIK_ERR_IO, "empty .node file");
  int64_t nv = std::stoll(tok[0]);
  if (nv <= 0) return fail(EIK_ERR_IO, "bad vertex count in .node");
  m.dim = 3;
  m.nv = nv;
  m.xyz.assign(3 * nv, 0.f);
  for (int64_t i = 0; i < nv; ++i) {
    if (!next_numbers(nf, tok) || tok.size() < 4) return fail(EIK_ERR_IO, "truncated .node file");
    for (int c = 0; c < 3; ++c) m.xyz[3 * i + c] = std::stof(tok[1 + c]);
  }
  if (!next_numbers(ef, tok)) return fail(EIK_ERR_IO, "empty .ele file");
  int64_t nt = std::stoll(tok[0]);
  int corners = tok.size() > 1 ? std::stoi(tok[1]) : 4;
  if (corners != 4 && corners != 10) return fail(EIK_ERR_IO, "tets must have 4 or 10 corners");
  m.ne = nt;
  m.elem.assign(4 * nt, 0);
  int32_t minidx = INT32_MAX;
  for (int64_t i = 0; i < nt; ++i) {
    if (!next_numbers(ef, tok) || (int)tok.size() < 1 + corners)
      return fail(EIK_ERR_IO, "truncated .ele file");
    for (int c = 0; c < 4; ++c) {
      m.elem[4 * i + c] = (int32_t)std::stol(tok[1 + c]);
      minidx = std::min(minidx, m.elem[4 * i + c]);
    }
  }
  // 0- or 1-based ids: TetMesh::init re-bases on the minimum index (tetmesh.cu:27-52)
  if (minidx == 1)
    for (auto& x : m.elem) x -= 1;
  return mesh_validate(m);
}

// PLY (the vertex x/y/z and face vertex_indices subset of TriMesh::read_ply,
// src/core2d/cuda/TriMesh_io.cu:268-510), ascii and binary.  Polygons are fan-triangulated.
int load_ply(const std::string& path, Mesh& m) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(EIK_ERR_IO, "cannot access file " + path);
  std::string line;
  std::getline(in, line);
  if (line.compare(0, 3, "ply") != 0) return fail(EIK_ERR_IO, path + " is not a PLY file");
  int fmt = 0;  // 0 ascii, 1 binary little endian, 2 binary big endian
  struct Prop {
    std::string name, type, count_type;
    bool list = false;
  };
  struct Elem {
    std::string name;
    int64_t n = 0;
    std::vector<Prop> props;
  };
  std::vector<Elem> elems;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ss(line);
    std::string w;
    ss >> w;
    if (w == "format") {
      std::string f;
      ss >> f;
      fmt = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : 2;
    } else if (w == "element") {
      Elem e;
      ss >> e.name >> e.n;
      elems.push_back(e);
    } else if (w == "property" && !elems.empty()) {
      Prop p;
      std::string t;
      ss >> t;
      if (t == "list") {
        p.list = true;
        ss >> p.count_type >> p.type >> p.name;
      } else {
        p.type = t;
        ss >> p.name;
      }
      elems.back().props.push_back(p);
    } else if (w == "end_header") {
      break;
    }
    // comments, obj_info and malformed lines are ignored
  }
  auto tsize = [](const std::string& t) -> int {
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" ||
        t == "float32")
      return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
  };
  auto read_bin = [&](const std::string& t) -> double {
    unsigned char buf[8];
    int n = tsize(t);
    if (n == 0) return 0;
    in.read((char*)buf, n);
    if (fmt == 2) std::reverse(buf, buf + n);
    if (t == "char" || t == "int8") return (double)*(int8_t*)buf;
    if (t == "uchar" || t == "uint8") return (double)*(uint8_t*)buf;
    if (t == "short" || t == "int16") { int16_t v; memcpy(&v, buf, 2); return v; }
    if (t == "ushort" || t == "uint16") { uint16_t v; memcpy(&v, buf, 2); return v; }
    if (t == "int" || t == "int32") { int32_t v; memcpy(&v, buf, 4); return v; }
    if (t == "uint" || t == "uint32") { uint32_t v; memcpy(&v, buf, 4); return v; }
    if (t == "float" || t == "float32") { float v; memcpy(&v, buf, 4); return v; }
    double v;
    memcpy(&v, buf, 8);
    return v;
  };
  m.dim = 2;
  m.xyz.clear();
  m.elem.clear();
  for (const Elem& e : elems) {
    for (int64_t r = 0; r < e.n; ++r) {
      std::vector<double> scal;
      std::vector<int64_t> list;
      std::istringstream ss;
      if (fmt == 0) {
        if (!std::getline(in, line)) return fail(EIK_ERR_IO, "truncated PLY body in " + path);
        ss.str(line);
      }
      double xyz[3] = {0, 0, 0};
      for (const Prop& p : e.props) {
        if (p.list) {
          double cnt;
          if (fmt == 0) {
            if (!(ss >> cnt)) return fail(EIK_ERR_IO, "bad PLY list in " + path);
          } else {
            cnt = read_bin(p.count_type);
          }
          if (!(cnt >= 0 && cnt <= 65536)) return fail(EIK_ERR_IO, "bad PLY list length in " + path);
          for (int64_t k = 0; k < (int64_t)cnt; ++k) {
            double x;
            if (fmt == 0) {
              if (!(ss >> x)) return fail(EIK_ERR_IO, "bad PLY list in " + path);
            } else {
              x = read_bin(p.type);
            }
            if (p.name == "vertex_indices" || p.name == "vertex_index") list.push_back((int64_t)x);
          }
        } else {
          double x = 0;
          if (fmt == 0) {
            if (!(ss >> x)) return fail(EIK_ERR_IO, "bad PLY value in " + path);
          } else {
            x = read_bin(p.type);
          }
          if (p.name == "x") xyz[0] = x;
          else if (p.name == "y") xyz[1] = x;
          else if (p.name == "z") xyz[2] = x;
        }
      }
      if (e.name == "vertex") {
        m.xyz.insert(m.xyz.end(), {(float)xyz[0], (float)xyz[1], (float)xyz[2]});
      } else if (e.name == "face") {
        for (size_t k = 1; k + 1 < list.size(); ++k) {
          m.elem.push_back((int32_t)list[0]);
          m.elem.push_back((int32_t)list[k]);
          m.elem.push_back((int32_t)list[k + 1]);
        }
      }
      if (!in) return fail(EIK_ERR_IO, "truncated PLY body in " + path);
    }
  }
  m.nv = (int64_t)m.xyz.size() / 3;
  m.ne = (int64_t)m.elem.size() / 3;
  if (m.ne == 0) return fail(EIK_ERR_IO, path + " has no faces");
  // files whose indices run from 1 (or any offset) through nv (+offset) are re-based, as the
  // reference's reader does (check_ind_range, src/core2d/cuda/TriMesh_io.cu:1187-1219)
  {
    const auto mm = std::minmax_element(m.elem.begin(), m.elem.end());
    const int32_t lo = *mm.first, hi = *mm.second;
    if (!(lo == 0 && hi == m.nv - 1) && (int64_t)hi - lo == m.nv - 1)
      for (int32_t& v : m.elem) v -= lo;
  }
  return mesh_validate(m);
}

// -------------------------------------------------------------------------------------------
// Generators
// -------------------------------------------------------------------------------------------
int gen_square(int n, Mesh& m) {
  if (n < 1 || n > 20000) return fail(EIK_ERR_ARG, "square: n out of range");
  const int64_t s = n + 1;
  m.dim = 2;
  m.nv = s * s;
  m.ne = 2LL * n * n;
  m.xyz.resize(3 * m.nv);
  for (int64_t j = 0; j < s; ++j)
    for (int64_t i = 0; i < s; ++i) {
      int64_t v = j * s + i;
      m.xyz[3 * v + 0] = (float)((double)i / n);
      m.xyz[3 * v + 1] = (float)((double)j / n);
      m.xyz[3 * v + 2] = 0.f;
    }
  m.elem.resize(3 * m.ne);
  int64_t f = 0;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      int32_t a = (int32_t)(j * s + i), b = a + 1, c = (int32_t)(a + s), d = c + 1;
      // diagonal (i,j)-(i+1,j+1) = a-d; counter-clockwise faces
      int32_t t0[3] = {a, b, d}, t1[3] = {a, d, c};
      std::copy(t0, t0 + 3, &m.elem[3 * f++]);
      std::copy(t1, t1 + 3, &m.elem[3 * f++]);
    }
  return EIK_OK;
}

namespace {
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }
}  // namespace

int gen_kuhn_cube(int n, float jitter, uint64_t seed, Mesh& m) {
  if (n < 1 || n > 1200) return fail(EIK_ERR_ARG, "kuhn cube: n out of range");
  if (!(jitter >= 0.f && jitter < 0.35f)) return fail(EIK_ERR_ARG, "kuhn cube: jitter must be in [0, 0.35)");
  const int64_t s = n + 1;
  const double h = 1.0 / n;
  m.dim = 3;
  m.nv = s * s * s;
  m.ne = 6LL * n * n * n;
  if (m.ne * 4 >= (int64_t)INT32_MAX) return fail(EIK_ERR_ARG, "kuhn cube too large");
  m.xyz.resize(3 * m.nv);
  uint64_t st = seed;
  for (int64_t k = 0; k < s; ++k)
    for (int64_t j = 0; j < s; ++j)
      for (int64_t i = 0; i < s; ++i) {
        int64_t v = (k * s + j) * s + i;
        int64_t ijk[3] = {i, j, k};
        bool interior = i > 0 && j > 0 && k > 0 && i < n && j < n && k < n;
        for (int c = 0; c < 3; ++c) {
          double x = ijk[c] * h;
          if (interior && jitter > 0.f) x += (2.0 * uniform01(st) - 1.0) * jitter * h;
          m.xyz[3 * v + c] = (float)x;
        }
      }
  // 6 tets per cell along the (0,0,0)-(1,1,1) diagonal, one per axis permutation: the path
  // 0 -> e_p0 -> e_p0 + e_p1 -> (1,1,1).
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  m.elem.resize(4 * m.ne);
  int64_t t = 0;
  for (int64_t k = 0; k < n; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i)
        for (int p = 0; p < 6; ++p) {
          int64_t c[3] = {i, j, k};
          int32_t* out = &m.elem[4 * t++];
          out[0] = (int32_t)((c[2] * s + c[1]) * s + c[0]);
          for (int st2 = 0; st2 < 3; ++st2) {
            c[perms[p][st2]] += 1;
            out[st2 + 1] = (int32_t)((c[2] * s + c[1]) * s + c[0]);
          }
        }
  // every tet must keep a positive |volume| after the jitter
  for (int64_t e = 0; e < m.ne; ++e) {
    const int32_t* v = &m.elem[4 * e];
    double a[3], b[3], c[3];
    for (int q = 0; q < 3; ++q) {
      a[q] = (double)m.xyz[3 * (int64_t)v[1] + q] - m.xyz[3 * (int64_t)v[0] + q];
      b[q] = (double)m.xyz[3 * (int64_t)v[2] + q] - m.xyz[3 * (int64_t)v[0] + q];
      c[q] = (double)m.xyz[3 * (int64_t)v[3] + q] - m.xyz[3 * (int64_t)v[0] + q];
    }
    double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                 a[2] * (b[0] * c[1] - b[1] * c[0]);
    if (!(std::fabs(det) > 1e-3 * h * h * h))
      return fail(EIK_ERR_ARG, "kuhn cube: jitter produced a degenerate tet");
  }
  return EIK_OK;
}

int gen_icosphere(int level, Mesh& m) {
  if (level < 0 || level > 10) return fail(EIK_ERR_ARG, "icosphere: level out of range");
  const double p = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> V = {-1, p, 0, 1, p, 0, -1, -p, 0, 1, -p, 0, 0, -1, p, 0, 1, p,
                           0, -1, -p, 0, 1, -p, p, 0, -1, p, 0, 1, -p, 0, -1, -p, 0, 1};
  std::vector<int32_t> F = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                            4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9,  4, 3,  4,  2, 3, 2, 6, 3,
                            6, 8, 3, 8,  9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  auto normalize = [](double* x) {
    double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    x[0] /= r; x[1] /= r; x[2] /= r;
  };
  for (size_t i = 0; i < V.size(); i += 3) normalize(&V[i]);
  for (int l = 0; l < level; ++l) {
    std::unordered_map<uint64_t, int32_t> mid;
    std::vector<int32_t> G;
    G.reserve(F.size() * 4);
    auto midpoint = [&](int32_t a, int32_t b) -> int32_t {
      uint64_t key = a < b ? ((uint64_t)a << 32) | (uint32_t)b : ((uint64_t)b << 32) | (uint32_t)a;
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double x[3] = {(V[3 * a] + V[3 * b]) / 2, (V[3 * a + 1] + V[3 * b + 1]) / 2,
                     (V[3 * a + 2] + V[3 * b + 2]) / 2};
      normalize(x);
      int32_t id = (int32_t)(V.size() / 3);
      V.insert(V.end(), x, x + 3);
      mid.emplace(key, id);
      return id;
    };
    for (size_t f = 0; f < F.size(); f += 3) {
      int32_t a = F[f], b = F[f + 1], c = F[f + 2];
      int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      int32_t nf[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      G.insert(G.end(), nf, nf + 12);
    }
    F.swap(G);
  }
  m.dim = 2;
  m.nv = (int64_t)V.size() / 3;
  m.ne = (int64_t)F.size() / 3;
  m.xyz.resize(V.size());
  for (size_t i = 0; i < V.size(); ++i) m.xyz[i] = (float)V[i];
  m.elem = F;
  return EIK_OK;
}

}  // namespace eik
