// Eikonal.cpp -- the reference's solver class (src/Eikonal.cu) and its lower-level drivers
// (meshFIM3dEikonal, meshFIM2dEikonal) on top of the C ABI.
//
// Results.  solveEikonal runs the persistent device scheduler (one launch per solve) and keeps
// the converged field: iteration_values_ then holds that one field.  The per-outer-iteration
// history of the reference (GenerateData's return value) is produced on demand: the first call
// of numIterations(), getResultAtIteration() or writeVTK(true) solves again with one launch per
// outer iteration (the reference's schedule, meshFIM3dEikonal.cu:411-542), the snapshots kept in
// device memory and copied once when the solve ends, and fills iteration_values_ with them.
// Callers that only read the final field pay neither the per-iteration launches nor any copy of
// the history.
#include "Eikonal.h"

#include <algorithm>

void eik_facade_check(int rc) {
  if (rc == EIK_OK) return;
  std::fprintf(stderr, "eikonal_amd: %s\n", eik_last_error());
  std::exit(-1);
}

namespace {

void copy_mesh_arrays(eik_mesh_t h, std::vector<point>& verts, int k, int64_t* ne_out,
                      std::vector<int32_t>& el) {
  eik_mesh_info_t info;
  eik_facade_check(eik_mesh_info(h, &info));
  std::vector<float> xyz(3 * info.nv);
  el.resize((size_t)k * info.nelem);
  eik_facade_check(eik_mesh_get_vertices(h, xyz.data()));
  eik_facade_check(eik_mesh_get_elements(h, el.data()));
  verts.resize(info.nv);
  for (int64_t i = 0; i < info.nv; ++i)
    verts[i] = point(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  *ne_out = info.nelem;
}

// One VTK file per time field: result<j>.vtk, legacy ASCII UNSTRUCTURED_GRID with the point
// scalars "traveltime" (meshFIM3dEikonal.cu:116-151, meshFIM2d_eikonal.cu:74-109).
template <size_t K>
void write_vtk_files(const std::vector<point>& V, const std::vector<std::array<int, K>>& cells,
                     const std::vector<std::vector<float>>& values) {
  const int vtk_type = K == 3 ? 5 : 10;  // VTK_TRIANGLE / VTK_TETRA
  for (size_t j = 0; j < values.size(); ++j) {
    const std::string name = "result" + std::to_string(j) + ".vtk";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", name.c_str());
      continue;
    }
    std::fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(f, "POINTS %zu float\n", V.size());
    for (const point& p : V) std::fprintf(f, "%.12f %.12f %.12f\n", p[0], p[1], p[2]);
    std::fprintf(f, "CELLS %zu %zu\n", cells.size(), cells.size() * (K + 1));
    for (const auto& c : cells) {
      std::fprintf(f, "%d", (int)K);
      for (size_t i = 0; i < K; ++i) std::fprintf(f, " %d", c[i]);
      std::fprintf(f, "\n");
    }
    std::fprintf(f, "CELL_TYPES %zu\n", cells.size());
    for (size_t e = 0; e < cells.size(); ++e) std::fprintf(f, "%d\n", vtk_type);
    std::fprintf(f, "POINT_DATA %zu\nSCALARS traveltime float 1\nLOOKUP_TABLE default\n", V.size());
    for (float v : values[j]) std::fprintf(f, "%.12f\n", v);
    std::fclose(f);
  }
}

// Seeds, options and the solve shared by both drivers.  history: one launch per outer iteration
// with the snapshots recorded on the device (returned all, oldest first); otherwise the
// persistent scheduler and the converged field only.
std::vector<std::vector<float>> run_solve(eik_mesh_t h, int64_t nv, const std::vector<int>& seeds,
                                          size_t maxIter, bool history, bool verbose,
                                          eik_solve_stats_t* st) {
  std::vector<std::vector<float>> out;
  if (seeds.empty()) {
    // nothing to propagate from: every value stays unknown
    eik_mesh_info_t info;
    eik_facade_check(eik_mesh_info(h, &info));
    out.emplace_back(nv, info.dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D);
    return out;
  }
  eik_solve_opts_t o;
  eik_solve_opts_default(&o);
  o.scheduler = history ? EIK_SCHED_SYNC : EIK_SCHED_ASYNC;
  o.max_iterations = (int32_t)std::min<size_t>(maxIter, 0x7fffffff);
  o.record_history = history ? 1 : 0;
  std::vector<int32_t> sd(seeds.begin(), seeds.end());
  std::vector<float> T(nv);
  eik_facade_check(eik_solve(h, sd.data(), (int64_t)sd.size(), nullptr, &o, T.data(), st));
  if (history) {
    out.resize((size_t)st->history_len);
    for (int64_t i = 0; i < st->history_len; ++i) {
      out[i].resize(nv);
      eik_facade_check(eik_get_history(h, i, out[i].data()));
    }
  }
  if (out.empty()) out.push_back(std::move(T));
  if (verbose)
    std::printf("solve: %lld outer iterations, %lld block visits, %.3f ms on the device\n",
                (long long)st->outer_iterations, (long long)st->block_visits, st->kernel_ms);
  return out;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Meshes
// ------------------------------------------------------------------------------------------
TriMesh::~TriMesh() {
  if (h_) eik_mesh_destroy(h_);
}

TriMesh* TriMesh::read(const char* filename, bool verbose) {
  eik_mesh_t h = nullptr;
  if (eik_mesh_load_ply(filename, &h) != EIK_OK) {
    if (verbose) std::fprintf(stderr, "%s\n", eik_last_error());
    return nullptr;
  }
  TriMesh* m = new TriMesh();
  m->h_ = h;
  std::vector<int32_t> el;
  int64_t nf = 0;
  copy_mesh_arrays(h, m->vertices, 3, &nf, el);
  m->faces.resize(nf);
  for (int64_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) m->faces[f][c] = el[3 * f + c];
  if (verbose)
    std::printf("mesh %s: %zu vertices, %lld faces\n", filename, m->vertices.size(), (long long)nf);
  return m;
}

TetMesh::~TetMesh() {
  if (h_) eik_mesh_destroy(h_);
}

TetMesh* TetMesh::read(const char* basename, const std::vector<float>& speedMtx, bool verbose) {
  eik_mesh_t h = nullptr;
  if (eik_mesh_load_tetgen(basename, &h) != EIK_OK) {
    if (verbose) std::fprintf(stderr, "%s\n", eik_last_error());
    return nullptr;
  }
  TetMesh* m = new TetMesh();
  m->h_ = h;
  std::vector<int32_t> el;
  int64_t nt = 0;
  copy_mesh_arrays(h, m->vertices, 4, &nt, el);
  m->tets.resize(nt);
  for (int64_t t = 0; t < nt; ++t)
    for (int c = 0; c < 4; ++c) m->tets[t][c] = el[4 * t + c];
  // speeds belong to the tet mesh (TetMesh::init, tetmesh.cu:54-79)
  if (eik_mesh_set_speed(h, speedMtx.data(), (int64_t)speedMtx.size()) != EIK_OK)
    std::fprintf(stderr,
                 "eikonal_amd: speed values not used (%s); the mesh keeps the unit metric\n",
                 eik_last_error());
  if (verbose)
    std::printf("mesh %s: %zu vertices, %lld tets\n", basename, m->vertices.size(), (long long)nt);
  return m;
}

// ------------------------------------------------------------------------------------------
// 3D driver
// ------------------------------------------------------------------------------------------
void meshFIM3dEikonal::SetMesh(TetMesh* mesh) {
  m_meshPtr = mesh;
  m_meshPtr->vertT.resize(m_meshPtr->vertices.size());
}

bool meshFIM3dEikonal::FindSeedPoint(double x, double y, double z, double epsilon) {
  m_SeedPoints.clear();
  for (size_t i = 0; i < m_meshPtr->vertices.size(); ++i) {
    const point& p = m_meshPtr->vertices[i];
    const double d = std::sqrt((x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) +
                               (z - p[2]) * (z - p[2]));
    if (d < epsilon) {
      m_SeedPoints.push_back((index)i);
      return true;
    }
  }
  return false;
}

void meshFIM3dEikonal::InitSpeedMat() { eik_facade_check(eik_mesh_set_speed(m_meshPtr->handle(), nullptr, 0)); }

void meshFIM3dEikonal::GraphPartition_METIS2(int& numBlock, int maxNumBlockVerts, bool verbose) {
  int32_t nb = 0;
  eik_facade_check(eik_partition(m_meshPtr->handle(), maxNumBlockVerts, &nb));
  numBlock = nb;
  m_numBlock = nb;
  if (verbose) std::printf("partition: %d blocks of up to %d vertices\n", nb, maxNumBlockVerts);
}

void meshFIM3dEikonal::GraphPartition_Square(int squareLength, int squareWidth, int squareHeight,
                                             int blockLength, int blockWidth, int blockHeight,
                                             bool verbose) {
  int32_t nb = 0;
  eik_facade_check(eik_partition_structured(m_meshPtr->handle(), squareLength, squareWidth,
                                            squareHeight, blockLength, blockWidth, blockHeight,
                                            &nb));
  m_numBlock = nb;
  if (verbose) std::printf("structured partition: %d blocks\n", nb);
}

void meshFIM3dEikonal::PartitionTets(int numBlock, bool verbose) {
  // the block-local layout (element copies per block, halo lists): built on the host now
  eik_layout_view_t v;
  eik_facade_check(eik_mesh_layout(m_meshPtr->handle(), &v));
  if (verbose) {
    eik_mesh_info_t info;
    eik_facade_check(eik_mesh_info(m_meshPtr->handle(), &info));
    std::printf("layout: %d blocks, %lld element copies (%.2f per tet), at most %d tets per block\n",
                info.nblocks, (long long)info.block_elems,
                (double)info.block_elems / (double)std::max<int64_t>(info.nelem, 1),
                info.max_block_elems);
  }
}

std::vector<std::vector<float>> meshFIM3dEikonal::GenerateData(size_t maxIter, bool verbose) {
  return run_solve(m_meshPtr->handle(), (int64_t)m_meshPtr->vertices.size(), m_SeedPoints, maxIter,
                   true, verbose, &lastStats_);
}

void meshFIM3dEikonal::writeVTK(std::vector<std::vector<float>> values) {
  write_vtk_files(m_meshPtr->vertices, m_meshPtr->tets, values);
}

// ------------------------------------------------------------------------------------------
// 2D driver
// ------------------------------------------------------------------------------------------
// Per-face speeds apply when there is one per face; otherwise every face has speed 1
// (InitializeAttributes, TriMesh_stats.cu:64-68).  Speeds the library rejects (not finite and
// positive) are reported and leave speed 1, as tensors that are not positive definite do in 3D.
void meshFIM2dEikonal::SetMesh(TriMesh* mesh, const std::vector<float>& face_speeds, int speed_type) {
  m_meshPtr = mesh;
  m_meshPtr->speed_type_ = speed_type;
  const bool per_face = face_speeds.size() == m_meshPtr->faces.size();
  if (speed_type != ONE && !per_face)
    std::fprintf(stderr, "speed type %d sets no speeds; using speed 1\n", speed_type);
  if (eik_mesh_set_speed(mesh->handle(), per_face ? face_speeds.data() : nullptr,
                         per_face ? (int64_t)face_speeds.size() : 0) != EIK_OK) {
    std::fprintf(stderr, "eikonal_amd: speed values not used (%s); the mesh keeps speed 1\n",
                 eik_last_error());
    eik_facade_check(eik_mesh_set_speed(mesh->handle(), nullptr, 0));
  }
}

void meshFIM2dEikonal::GraphPartition_METIS2(int& numBlock, int maxNumBlockVerts, bool verbose) {
  int32_t nb = 0;
  eik_facade_check(eik_partition(m_meshPtr->handle(), maxNumBlockVerts, &nb));
  numBlock = nb;
  if (verbose) std::printf("partition: %d blocks of about %d vertices\n", nb, maxNumBlockVerts);
}

void meshFIM2dEikonal::GraphPartition_Square(int squareLength, int squareWidth, int blockLength,
                                             int blockWidth, bool verbose) {
  int32_t nb = 0;
  eik_facade_check(eik_partition_structured(m_meshPtr->handle(), squareLength, squareWidth, 1,
                                            blockLength, blockWidth, 1, &nb));
  if (verbose) std::printf("structured partition: %d blocks\n", nb);
}

void meshFIM2dEikonal::PartitionFaces(int numBlock) {
  eik_layout_view_t v;
  eik_facade_check(eik_mesh_layout(m_meshPtr->handle(), &v));
}

std::vector<std::vector<float>> meshFIM2dEikonal::GenerateData(int numBlock, size_t maxIters,
                                                               bool verbose) {
  return run_solve(m_meshPtr->handle(), (int64_t)m_meshPtr->vertices.size(), m_SeedPoints, maxIters,
                   true, verbose, &lastStats_);
}

void meshFIM2dEikonal::writeVTK(std::vector<std::vector<float>> time_values) {
  write_vtk_files(m_meshPtr->vertices, m_meshPtr->faces, time_values);
}

// ------------------------------------------------------------------------------------------
// Eikonal
// ------------------------------------------------------------------------------------------
Eikonal::Eikonal(bool isTriMesh, std::string fname, bool verbose)
    : verbose_(verbose),
      filename_(fname),
      maxBlocks_(16),
      maxVertsPerBlock_(64),
      stopDistance_(50000.f),
      isStructured_(false),
      userSetInitial_(false),
      speedType_(ONE),
      squareLength_(16),
      squareWidth_(16),
      squareDepth_(16),
      squareBlockLength_(1),
      squareBlockWidth_(1),
      squareBlockDepth_(1),
      maxIterations_(100),
      triMesh_(nullptr),
      tetMesh_(nullptr),
      FIMPtr2d_(nullptr),
      FIMPtr3d_(nullptr),
      isTriMesh_(isTriMesh),
      historyPending_(false) {}

Eikonal::~Eikonal() {
  delete FIMPtr2d_;
  delete FIMPtr3d_;
  delete tetMesh_;
  delete triMesh_;
}

// Builds the per-iteration history of the last solve if it has not been built: solves again with
// one launch per outer iteration and replaces iteration_values_ with its snapshots.
void Eikonal::ensureHistory() {
  if (!historyPending_) return;
  historyPending_ = false;
  iteration_values_ = isTriMesh_ ? FIMPtr2d_->GenerateData(maxBlocks_, maxIterations_, verbose_)
                                 : FIMPtr3d_->GenerateData(maxIterations_, verbose_);
}

std::vector<float>& Eikonal::getFinalResult() { return iteration_values_.at(iteration_values_.size() - 1); }

std::vector<float>& Eikonal::getResultAtIteration(size_t i) {
  ensureHistory();
  return iteration_values_.at(i);
}

size_t Eikonal::numIterations() {
  ensureHistory();
  return iteration_values_.size();
}

void Eikonal::initSpeedMtxMultipliers(std::vector<float> values) { speedMtxMultipliers_ = values; }

void Eikonal::initializeVertices(std::vector<float> values) {
  if (!triMesh_ && !tetMesh_) {
    std::cerr << "You must initialize the mesh first!" << std::endl;
    std::exit(0);
  }
  std::vector<float>& vt = triMesh_ ? triMesh_->vertT : tetMesh_->vertT;
  const size_t nv = triMesh_ ? triMesh_->vertices.size() : tetMesh_->vertices.size();
  if (values.size() != nv) {
    std::cerr << "Initialize values size does not match number of vertices!" << std::endl;
    std::exit(0);
  }
  vt = values;
  userSetInitial_ = true;
}

void Eikonal::initializeMesh() {
  if (isTriMesh_) {
    if (!triMesh_) {
      triMesh_ = TriMesh::read(filename_.c_str(), verbose_);
      if (!triMesh_) {
        std::printf("File open failed!!\n");
        std::exit(1);
      }
    }
  } else if (!tetMesh_) {
    tetMesh_ = TetMesh::read(filename_.c_str(), speedMtxMultipliers_, verbose_);
    if (!tetMesh_) {
      std::printf("File open failed!!\n");
      std::exit(1);
    }
  }
}

void Eikonal::solveEikonal() {
  const clock_t start = clock();
  initializeMesh();
  const size_t nv = isTriMesh_ ? triMesh_->vertices.size() : tetMesh_->vertices.size();
  std::vector<float>& vt = isTriMesh_ ? triMesh_->vertT : tetMesh_->vertT;
  std::vector<int> seeds;
  if (!userSetInitial_) {
    vt.assign(nv, isTriMesh_ ? EIK_LARGENUM_2D : EIK_LARGENUM_3D);
    vt[0] = 0.f;
    seeds.push_back(0);
  } else {
    for (size_t i = 0; i < nv; ++i)
      if (vt[i] == 0.f) seeds.push_back((int)i);
  }
  // a new driver object per solve, the previous one released (the reference leaks it,
  // Eikonal.cu:142, 184)
  delete FIMPtr2d_;
  delete FIMPtr3d_;
  FIMPtr2d_ = nullptr;
  FIMPtr3d_ = nullptr;
  eik_solve_stats_t st;
  if (isTriMesh_) {
    FIMPtr2d_ = new meshFIM2dEikonal();
    FIMPtr2d_->SetSeedPoint(seeds);
    FIMPtr2d_->SetMesh(triMesh_, speedMtxMultipliers_, speedType_);
    FIMPtr2d_->SetStopDistance(stopDistance_);
    if (isStructured_) {
      maxBlocks_ = (squareLength_ / squareBlockLength_) * (squareWidth_ / squareBlockWidth_);
      FIMPtr2d_->GraphPartition_Square(squareLength_, squareWidth_, squareBlockLength_,
                                       squareBlockWidth_, verbose_);
    } else {
      FIMPtr2d_->GraphPartition_METIS2(maxBlocks_, maxVertsPerBlock_, verbose_);
    }
    FIMPtr2d_->PartitionFaces(maxBlocks_);
    FIMPtr2d_->InitializeLabels(maxBlocks_);
    iteration_values_ = run_solve(triMesh_->handle(), (int64_t)nv, seeds, maxIterations_, false,
                                  verbose_, &st);
    FIMPtr2d_->lastStats_ = st;
  } else {
    FIMPtr3d_ = new meshFIM3dEikonal();
    FIMPtr3d_->SetMesh(tetMesh_);
    FIMPtr3d_->SetSeedPoint(seeds);
    if (isStructured_) {
      maxBlocks_ = (squareLength_ / squareBlockLength_) * (squareWidth_ / squareBlockWidth_) *
                   (squareDepth_ / squareBlockDepth_);
      FIMPtr3d_->GraphPartition_Square(squareLength_, squareWidth_, squareDepth_,
                                       squareBlockLength_, squareBlockWidth_, squareBlockDepth_,
                                       verbose_);
    } else {
      FIMPtr3d_->GraphPartition_METIS2(maxBlocks_, maxVertsPerBlock_, verbose_);
    }
    FIMPtr3d_->m_numBlock = maxBlocks_;
    FIMPtr3d_->PartitionTets(maxBlocks_, verbose_);
    iteration_values_ = run_solve(tetMesh_->handle(), (int64_t)nv, seeds, maxIterations_, false,
                                  verbose_, &st);
    FIMPtr3d_->lastStats_ = st;
  }
  historyPending_ = !seeds.empty();
  if (verbose_)
    std::printf("Computing time : %.10lf ms\n", (double)(clock() - start) * 1000.0 / CLOCKS_PER_SEC);
}

void Eikonal::writeVTK(bool all) {
  if (iteration_values_.empty()) return;
  if (all) ensureHistory();
  std::vector<std::vector<float>> vals;
  for (size_t i = all ? 0 : iteration_values_.size() - 1; i < iteration_values_.size(); ++i)
    vals.push_back(iteration_values_[i]);
  if (FIMPtr2d_)
    FIMPtr2d_->writeVTK(vals);
  else if (FIMPtr3d_)
    FIMPtr3d_->writeVTK(vals);
}

void Eikonal::printErrorGraph(std::vector<float> solution) {
  const size_t n = numIterations();
  if (n == 0 || solution.empty()) return;
  std::vector<double> rms(n);
  for (size_t i = 0; i < n; ++i) {
    const std::vector<float>& r = iteration_values_[i];
    double s = 0.0;
    for (size_t j = 0; j < solution.size() && j < r.size(); ++j) {
      const double e = (double)solution[j] - r[j];
      s += e * e;
    }
    rms[i] = std::sqrt(s / (double)solution.size());
  }
  // one row per decade between the largest and smallest error, one column per iteration
  const double hi = *std::max_element(rms.begin(), rms.end());
  const double lo = std::max(*std::min_element(rms.begin(), rms.end()), 1e-12);
  const int top = (int)std::ceil(std::log10(std::max(hi, 1e-12)));
  const int bot = (int)std::floor(std::log10(lo));
  std::printf("\n\nlog(Err)|\n");
  for (int d = top; d >= bot; --d) {
    std::printf((top - d) % 2 == 0 ? "   10^%2d|" : "        |", d);
    for (size_t i = 0; i < n; ++i) {
      const int dec = rms[i] > 0 ? (int)std::floor(std::log10(rms[i])) : bot;
      std::printf("%c", dec == d ? '*' : ' ');
    }
    std::printf("\n");
  }
  std::printf("--------|------------------------------------------  Converged to: %.4f\n",
              rms[n - 1]);
  std::printf("        |1   5    10   15   20   25   30   35\n");
  std::printf("                   Iteration\n");
}
