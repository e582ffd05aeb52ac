// capi.cpp -- host-only entry points of the C ABI: errors, meshes, speeds, generators and
// partitions.  The solve entry points live in solver.hip.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>

#include "eik_internal.h"

namespace eik {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace eik

using namespace eik;

namespace {
Mesh* M(eik_mesh_t h) { return reinterpret_cast<Mesh*>(h); }

int finish_create(Mesh* m, eik_mesh_t* out) {
  int rc = mesh_validate(*m);
  if (rc) {
    delete m;
    return rc;
  }
  *out = reinterpret_cast<eik_mesh_t>(m);
  return EIK_OK;
}
// Runs a file reader; a malformed token (std::stof / std::stoll out of range or not a number) or
// a header count too large to allocate throws inside the reader, which is reported as EIK_ERR_IO
// instead of escaping through the C ABI.
template <class F>
int load_mesh(F&& reader, eik_mesh_t* out, const char* what) {
  Mesh* m = new Mesh();
  int rc;
  try {
    rc = reader(*m);
  } catch (const std::exception& e) {
    rc = fail(EIK_ERR_IO, std::string("malformed mesh file ") + what + ": " + e.what());
  }
  if (rc) {
    delete m;
    return rc;
  }
  *out = reinterpret_cast<eik_mesh_t>(m);
  return EIK_OK;
}
// The block layout (element copies, halos) is built with the partition, as the reference's
// PartitionTets / PartitionFaces follow its partition step, so the partition's statistics are
// known at once.  A block too large for the kernels' 16-bit local indices is still a valid
// partition: the solve reports it (EIK_ERR_CAPACITY).
int layout_with_partition(Mesh& m) {
  int rc = build_layout(m);
  return rc == EIK_ERR_CAPACITY ? EIK_OK : rc;
}
}  // namespace

extern "C" {

const char* eik_last_error(void) { return g_last_error.c_str(); }

const char* eik_version(void) { return "eikonal_amd 0.1 (gfx950)"; }

int eik_mesh_create_tet(const float* xyz, int64_t nv, const int32_t* tets, int64_t nt,
                        eik_mesh_t* out) {
  if (!xyz || !tets || !out || nv <= 0 || nt <= 0) return fail(EIK_ERR_ARG, "bad tet mesh arguments");
  Mesh* m = new Mesh();
  m->dim = 3;
  m->nv = nv;
  m->ne = nt;
  m->xyz.assign(xyz, xyz + 3 * nv);
  m->elem.assign(tets, tets + 4 * nt);
  int32_t mn = INT32_MAX;
  for (int32_t x : m->elem) mn = std::min(mn, x);
  if (mn == 1)  // 1-based ids (TetMesh::init, tetmesh.cu:42-49)
    for (auto& x : m->elem) x -= 1;
  return finish_create(m, out);
}

int eik_mesh_create_tri(const float* xyz, int64_t nv, const int32_t* faces, int64_t nf,
                        eik_mesh_t* out) {
  if (!xyz || !faces || !out || nv <= 0 || nf <= 0) return fail(EIK_ERR_ARG, "bad triangle mesh arguments");
  Mesh* m = new Mesh();
  m->dim = 2;
  m->nv = nv;
  m->ne = nf;
  m->xyz.assign(xyz, xyz + 3 * nv);
  m->elem.assign(faces, faces + 3 * nf);
  return finish_create(m, out);
}

int eik_mesh_load_tetgen(const char* basename, eik_mesh_t* out) {
  if (!basename || !out) return fail(EIK_ERR_ARG, "null argument");
  return load_mesh([&](Mesh& m) { return load_tetgen(basename, m); }, out, basename);
}

int eik_mesh_load_ply(const char* path, eik_mesh_t* out) {
  if (!path || !out) return fail(EIK_ERR_ARG, "null argument");
  return load_mesh([&](Mesh& m) { return load_ply(path, m); }, out, path);
}

int eik_mesh_destroy(eik_mesh_t h) {
  Mesh* m = M(h);
  if (!m) return EIK_OK;
  device_release(*m);
  delete m;
  return EIK_OK;
}

int eik_mesh_info(eik_mesh_t h, eik_mesh_info_t* out) {
  Mesh* m = M(h);
  if (!m || !out) return fail(EIK_ERR_ARG, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->dim = m->dim;
  out->nv = m->nv;
  out->nelem = m->ne;
  out->nblocks = m->nblocks;
  if (m->dim == 2 && m->vfaces_ready) out->nvirtual = (int64_t)m->vfaces.size() / 3;
  if (m->stats_valid) {
    out->block_elems = m->layout.total_elems;
    out->block_halo = (int64_t)m->layout.halo.size();
    out->max_block_verts = m->layout.max_own;
    out->max_block_local = m->layout.max_local;
    out->max_block_elems = m->layout.max_elem;
  }
  return EIK_OK;
}

int eik_mesh_get_vertices(eik_mesh_t h, float* xyz_out) {
  Mesh* m = M(h);
  if (!m || !xyz_out) return fail(EIK_ERR_ARG, "null argument");
  std::memcpy(xyz_out, m->xyz.data(), sizeof(float) * m->xyz.size());
  return EIK_OK;
}

int eik_mesh_get_elements(eik_mesh_t h, int32_t* out) {
  Mesh* m = M(h);
  if (!m || !out) return fail(EIK_ERR_ARG, "null argument");
  std::memcpy(out, m->elem.data(), sizeof(int32_t) * m->elem.size());
  return EIK_OK;
}

int eik_mesh_set_speed(eik_mesh_t h, const float* values, int64_t count) {
  Mesh* m = M(h);
  if (!m) return fail(EIK_ERR_ARG, "null mesh");
  if (count > 0 && !values) return fail(EIK_ERR_ARG, "null values");
  for (int64_t i = 0; i < count; ++i)
    if (!std::isfinite(values[i])) return fail(EIK_ERR_ARG, "non-finite speed value");
  m->metric.clear();
  m->metric_kind = 0;
  if (m->dim == 3) {
    // TetMesh::init (tetmesh.cu:54-79)
    if (count == 6 * m->ne) {
      for (int64_t t = 0; t < m->ne; ++t) {
        const float* M6 = values + 6 * t;
        // the metric must be positive definite for the travel times to be lengths
        double a = M6[0], b = M6[1], c = M6[2], d = M6[3], e = M6[4], f = M6[5];
        double m2 = a * d - b * b;
        double m3 = a * (d * f - e * e) - b * (b * f - e * c) + c * (b * e - d * c);
        if (!(a > 0 && m2 > 0 && m3 > 0))
          return fail(EIK_ERR_ARG, "metric tensor of tet " + std::to_string(t) + " is not positive definite");
      }
      m->metric.assign(values, values + count);
      m->metric_kind = 2;
    } else if (count == m->ne) {
      for (int64_t t = 0; t < m->ne; ++t)
        if (!(values[t] > 0)) return fail(EIK_ERR_ARG, "speed multipliers must be > 0");
      m->metric.assign(values, values + count);
      m->metric_kind = 1;
    }
  } else {
    // TriMesh::InitializeAttributes (TriMesh_stats.cu:64-68): speedInv = 1 / speed
    if (count == m->ne) {
      m->metric.resize(count);
      for (int64_t f = 0; f < count; ++f) {
        if (!(values[f] > 0)) return fail(EIK_ERR_ARG, "face speeds must be > 0");
        m->metric[f] = 1.0f / values[f];
      }
      m->metric_kind = 1;
    }
  }
  m->layout.valid = false;
  m->device_dirty = true;
  return EIK_OK;
}

int eik_mesh_set_metric_tri(eik_mesh_t h, const float* m3, int64_t nf) {
  Mesh* m = M(h);
  if (!m || !m3) return fail(EIK_ERR_ARG, "null argument");
  if (m->dim != 2) return fail(EIK_ERR_ARG, "triangle metric on a tet mesh");
  if (nf != m->ne) return fail(EIK_ERR_ARG, "need 3 metric values per face");
  for (int64_t f = 0; f < nf; ++f) {
    double a = m3[3 * f], b = m3[3 * f + 1], c = m3[3 * f + 2];
    if (!(a > 0 && a * c - b * b > 0))
      return fail(EIK_ERR_ARG, "metric of face " + std::to_string(f) + " is not positive definite");
  }
  m->metric.assign(m3, m3 + 3 * nf);
  m->metric_kind = 2;
  m->layout.valid = false;
  m->device_dirty = true;
  return EIK_OK;
}

int eik_gen_square(int32_t n, eik_mesh_t* out) {
  if (!out) return fail(EIK_ERR_ARG, "null argument");
  Mesh* m = new Mesh();
  int rc = gen_square(n, *m);
  if (rc) {
    delete m;
    return rc;
  }
  return finish_create(m, out);
}

int eik_gen_kuhn_cube(int32_t n, float jitter, uint64_t seed, eik_mesh_t* out) {
  if (!out) return fail(EIK_ERR_ARG, "null argument");
  Mesh* m = new Mesh();
  int rc = gen_kuhn_cube(n, jitter, seed, *m);
  if (rc) {
    delete m;
    return rc;
  }
  return finish_create(m, out);
}

int eik_gen_icosphere(int32_t level, eik_mesh_t* out) {
  if (!out) return fail(EIK_ERR_ARG, "null argument");
  Mesh* m = new Mesh();
  int rc = gen_icosphere(level, *m);
  if (rc) {
    delete m;
    return rc;
  }
  return finish_create(m, out);
}

int eik_partition(eik_mesh_t h, int32_t verts_per_block, int32_t* nblocks_out) {
  Mesh* m = M(h);
  if (!m) return fail(EIK_ERR_ARG, "null mesh");
  int rc = partition_by_size(*m, verts_per_block);
  if (rc || (rc = layout_with_partition(*m))) return rc;
  if (nblocks_out) *nblocks_out = m->nblocks;
  return EIK_OK;
}

int eik_partition_structured(eik_mesh_t h, int32_t L, int32_t W, int32_t D, int32_t bl, int32_t bw,
                             int32_t bd, int32_t* nblocks_out) {
  Mesh* m = M(h);
  if (!m) return fail(EIK_ERR_ARG, "null mesh");
  int rc = partition_structured(*m, L, W, D, bl, bw, bd);
  if (rc || (rc = layout_with_partition(*m))) return rc;
  if (nblocks_out) *nblocks_out = m->nblocks;
  return EIK_OK;
}

int eik_partition_set_labels(eik_mesh_t h, const int32_t* labels, int32_t nblocks) {
  Mesh* m = M(h);
  if (!m || !labels) return fail(EIK_ERR_ARG, "null argument");
  if (nblocks <= 0) return fail(EIK_ERR_ARG, "nblocks must be > 0");
  for (int64_t v = 0; v < m->nv; ++v)
    if (labels[v] < 0 || labels[v] >= nblocks) return fail(EIK_ERR_ARG, "label out of range");
  m->label.assign(labels, labels + m->nv);
  m->nblocks = nblocks;
  m->layout.valid = false;
  m->stats_valid = false;
  m->device_dirty = true;
  return layout_with_partition(*m);
}

int eik_partition_labels(eik_mesh_t h, int32_t* labels_out) {
  Mesh* m = M(h);
  if (!m || !labels_out) return fail(EIK_ERR_ARG, "null argument");
  if (m->nblocks <= 0) return fail(EIK_ERR_STATE, "mesh is not partitioned");
  std::memcpy(labels_out, m->label.data(), sizeof(int32_t) * m->nv);
  return EIK_OK;
}

int eik_mesh_virtual_faces(eik_mesh_t h, int32_t* rows_out, int32_t* origin_out, int64_t cap,
                           int64_t* n_out) {
  Mesh* m = M(h);
  if (!m || !n_out) return fail(EIK_ERR_ARG, "null argument");
  if (m->dim != 2) {
    *n_out = 0;
    return EIK_OK;
  }
  tri_virtual_faces(*m);
  int64_t n = (int64_t)m->vfaces.size() / 3;
  *n_out = n;
  if (rows_out && origin_out) {
    int64_t k = std::min(n, cap);
    std::memcpy(rows_out, m->vfaces.data(), sizeof(int32_t) * 3 * k);
    std::memcpy(origin_out, m->vorigin.data(), sizeof(int32_t) * k);
  }
  return EIK_OK;
}

int eik_mesh_layout(eik_mesh_t h, eik_layout_view_t* out) {
  Mesh* m = M(h);
  if (!m || !out) return fail(EIK_ERR_ARG, "null argument");
  if (m->nblocks <= 0) return fail(EIK_ERR_STATE, "mesh is not partitioned");
  if (!m->layout.valid) {
    int rc = build_layout(*m);
    if (rc) return rc;
  }
  const Layout& L = m->layout;
  std::memset(out, 0, sizeof(*out));
  out->dim = m->dim;
  out->nblocks = L.nblocks;
  out->nv = m->nv;
  out->nelems = L.total_elems;
  out->nhalo = (int64_t)L.halo.size();
  out->rec_bytes = m->dim == 3 ? (int32_t)sizeof(TetRec) : (int32_t)sizeof(TriRec);
  out->max_local = L.max_local;
  out->perm = L.perm.data();
  out->iperm = L.iperm.data();
  out->blocks = reinterpret_cast<const int32_t*>(L.blocks.data());
  out->halo = L.halo.data();
  out->elems = m->dim == 3 ? (const void*)L.tets.data() : (const void*)L.tris.data();
  out->sup_start = L.sup_start.data();
  out->sup = L.sup.data();
  out->dep_start = L.dep_start.data();
  out->dep = L.dep.data();
  return EIK_OK;
}

}  // extern "C"
