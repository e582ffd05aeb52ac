// eik_internal.h -- host-side data model shared by the mesh, partition, layout and solver
// translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "eikonal_amd.h"

namespace eik {

// Element record of the block layout, tet meshes.  Corners are local vertex indices of the
// owning block (< nown: owned by the block, else halo slot nown + k).  L holds the squared
// metric lengths of the six edges {01, 02, 03, 12, 13, 23}: every Gram product a local solve
// needs follows from them by polarisation, e_i.e_j = (|e_i|^2 + |e_j|^2 - |e_i - e_j|^2) / 2.
// H[k] is the metric distance from corner k to the plane of the other three, rounded down: a
// solve of corner k can give nothing below min(other times) + H[k], which prunes solves that
// cannot lower the corner.
struct alignas(16) TetRec {
  uint16_t v[4];
  float L[6];
  float H[4];
};
static_assert(sizeof(TetRec) == 48, "TetRec must be 48 bytes");

// Triangle record: corners, a corner mask (bit k set = corner k may be updated; virtual faces
// only update their obtuse corner), squared metric lengths of edges {01, 12, 20}, and H[k] the
// metric distance from corner k to the line of the opposite edge, rounded down.
struct alignas(16) TriRec {
  uint16_t v[3];
  uint16_t mask;
  float L[3];
  float H[3];
};
static_assert(sizeof(TriRec) == 32, "TriRec must be 32 bytes");

// Per-block descriptor, one 32-byte record so a visit reads it with two loads.
struct alignas(16) BlockDesc {
  int32_t vstart;   // first owned vertex (block-contiguous numbering)
  int32_t nown;     // owned vertices
  int32_t nbnd0;    // local index of the first boundary vertex (owned, seen by another block)
  int32_t hstart;   // first halo entry
  int32_t nhalo;    // halo vertices
  int32_t estart;   // first element record
  int32_t nelem;    // element records
  int32_t pad;
};
static_assert(sizeof(BlockDesc) == 32, "BlockDesc must be 32 bytes");

// Host copy of the block layout.
struct Layout {
  int32_t nblocks = 0;
  std::vector<int32_t> perm;     // new id -> caller id
  std::vector<int32_t> iperm;    // caller id -> new id
  std::vector<BlockDesc> blocks;
  std::vector<int32_t> halo;     // new ids of halo vertices, block after block
  std::vector<TetRec> tets;
  std::vector<TriRec> tris;
  std::vector<int32_t> rec_elem;  // per record: its element (a virtual face: its original face)
  std::vector<int32_t> owner;    // new id -> block
  // dependents of block b: blocks whose halo holds a vertex owned by b
  std::vector<int32_t> dep_start, dep;
  // suppliers of block c: blocks owning a vertex of c's halo
  std::vector<int32_t> sup_start, sup;
  // per vertex (new id): positions, within its owner's dependent list, of the blocks that hold
  // the vertex in their halo (empty for interior vertices)
  std::vector<int32_t> vh_start;
  std::vector<uint16_t> vh_pos;
  // the same as a bit mask over the owner's dependent list (bit k = k-th dependent), built when
  // no block has more than 64 dependents
  std::vector<uint64_t> vmask;
  int32_t max_dep = 0;
  int32_t max_local = 0, max_own = 0, max_elem = 0;
  int64_t total_elems = 0;
  bool valid = false;
};

struct DeviceState;  // defined in the solver (HIP) translation unit

struct Mesh {
  int dim = 3;  // 3 = tets, 2 = triangles
  int64_t nv = 0, ne = 0;
  std::vector<float> xyz;        // 3 nv
  std::vector<int32_t> elem;     // 4 ne or 3 ne, 0-based
  // speeds: tets use 6 metric entries per tet (upper triangular); triangles use speedInv
  // (iso) or 3 metric entries per face (aniso)
  std::vector<float> metric;
  int metric_kind = 0;           // 0 = identity / unit speed, 1 = scalar per element, 2 = tensor
  // 2D virtual faces (obtuse vertex first) and their originating faces
  std::vector<int32_t> vfaces, vorigin;
  bool vfaces_ready = false;
  // partition
  std::vector<int32_t> label;
  int32_t nblocks = 0;
  Layout layout;
  // the layout's block statistics describe the current partition (they do not depend on the
  // speeds, so a speed change keeps them while it invalidates the element records)
  bool stats_valid = false;
  DeviceState* dev = nullptr;
  bool device_dirty = true;
};

// Inner sweeps per block visit when the caller leaves the choice to the library: measured on
// the 64^3 cube (1 and 2 equal, 3 slower) and on the level-7 anisotropic icosphere (4 best:
// 2.11 ms vs 2.55 ms at 2; the front needs more sweeps to cross an 8 x 8 block).
inline int32_t default_inner_sweeps(int dim, int32_t requested) {
  return requested > 0 ? requested : (dim == 3 ? 2 : 4);
}

// errors ----------------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// mesh.cpp ----------------------------------------------------------------------------------
int mesh_validate(const Mesh& m);
void tri_virtual_faces(Mesh& m);
bool tri_is_non_obtuse(const float* xyz, const int32_t* f, int32_t v);
int load_tetgen(const std::string& base, Mesh& m);
int load_ply(const std::string& path, Mesh& m);
int gen_square(int n, Mesh& m);
int gen_kuhn_cube(int n, float jitter, uint64_t seed, Mesh& m);
int gen_icosphere(int level, Mesh& m);
// squared metric length of edge (a, b) of element e
double edge_len2(const Mesh& m, int64_t e, int32_t a, int32_t b);
double tri_edge_len2(const Mesh& m, int64_t face, const int32_t* corners, int32_t a, int32_t b);

// partition.cpp -----------------------------------------------------------------------------
int partition_rcb(Mesh& m, int32_t nblocks);
int partition_structured(Mesh& m, int L, int W, int D, int bl, int bw, int bd);
// the reference's block count for a target block size, then partition_rcb
int partition_by_size(Mesh& m, int32_t verts_per_block);

// layout.cpp --------------------------------------------------------------------------------
int build_layout(Mesh& m);
void tet_heights(const float* L6, float* H);
// Split records of the layout for visits that read element data from memory at every use: per
// element a 16-byte selection record (sel, 4 words: corners as u16, then the corner heights as
// fp16 rounded down; triangles: 3 corners and the corner mask, 3 heights) and the squared edge
// lengths (len: 6 floats per tet, 4 per triangle with the last 0).
void split_records(const Layout& L, int dim, std::vector<uint32_t>& sel, std::vector<float>& len);
// fp16 bits of the largest half-precision value <= x (x >= 0; 0 for x <= 0, 65504 above it)
uint16_t half_below(float x);
float half_value(uint16_t h);
void tri_heights(const float* L3, float* H);

// solver (HIP) ------------------------------------------------------------------------------
void device_release(Mesh& m);

}  // namespace eik
