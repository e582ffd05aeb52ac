// Eikonal.cpp -- the reference's solver class (src/Eikonal.cu) on top of the C ABI.
#include "Eikonal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iostream>

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
      isTriMesh_(isTriMesh),
      scheduler_(EIK_SCHED_SYNC),
      mesh_(nullptr),
      loaded_(nullptr) {}

Eikonal::~Eikonal() {
  delete tetMesh_;
  delete triMesh_;
  if (mesh_) eik_mesh_destroy(mesh_);
}

void Eikonal::check(int rc) {
  if (rc == EIK_OK) return;
  std::fprintf(stderr, "eikonal_amd: %s\n", eik_last_error());
  std::exit(-1);
}

std::vector<float>& Eikonal::getFinalResult() { return iteration_values_.at(iteration_values_.size() - 1); }

std::vector<float>& Eikonal::getResultAtIteration(size_t i) { return iteration_values_.at(i); }

size_t Eikonal::numIterations() { return iteration_values_.size(); }

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
  const bool have = isTriMesh_ ? triMesh_ != nullptr : tetMesh_ != nullptr;
  if (have && mesh_ && loaded_ == (isTriMesh_ ? (const void*)triMesh_ : (const void*)tetMesh_))
    return;
  if (mesh_) {
    eik_mesh_destroy(mesh_);
    mesh_ = nullptr;
  }
  int rc = isTriMesh_ ? eik_mesh_load_ply(filename_.c_str(), &mesh_)
                      : eik_mesh_load_tetgen(filename_.c_str(), &mesh_);
  if (rc != EIK_OK) {
    if (verbose_) std::fprintf(stderr, "%s\n", eik_last_error());
    std::printf("File open failed!!\n");
    std::exit(1);
  }
  eik_mesh_info_t info;
  check(eik_mesh_info(mesh_, &info));
  std::vector<float> xyz(3 * info.nv);
  std::vector<int32_t> el((isTriMesh_ ? 3 : 4) * info.nelem);
  check(eik_mesh_get_vertices(mesh_, xyz.data()));
  check(eik_mesh_get_elements(mesh_, el.data()));
  std::vector<point> verts(info.nv);
  for (int64_t i = 0; i < info.nv; ++i)
    for (int c = 0; c < 3; ++c) verts[i][c] = xyz[3 * i + c];
  if (isTriMesh_) {
    if (!triMesh_) triMesh_ = new TriMesh();
    triMesh_->vertices = std::move(verts);
    triMesh_->faces.resize(info.nelem);
    for (int64_t f = 0; f < info.nelem; ++f)
      for (int c = 0; c < 3; ++c) triMesh_->faces[f][c] = el[3 * f + c];
    loaded_ = triMesh_;
  } else {
    if (!tetMesh_) tetMesh_ = new TetMesh();
    tetMesh_->vertices = std::move(verts);
    tetMesh_->tets.resize(info.nelem);
    for (int64_t t = 0; t < info.nelem; ++t)
      for (int c = 0; c < 4; ++c) tetMesh_->tets[t][c] = el[4 * t + c];
    // tet speeds belong to the mesh (TetMesh::init): applied now, as the reference does
    check(eik_mesh_set_speed(mesh_, speedMtxMultipliers_.data(),
                             (int64_t)speedMtxMultipliers_.size()));
    loaded_ = tetMesh_;
  }
  if (verbose_)
    std::printf("mesh %s: %lld vertices, %lld %s\n", filename_.c_str(), (long long)info.nv,
                (long long)info.nelem, isTriMesh_ ? "faces" : "tets");
}

void Eikonal::solveEikonal() {
  initializeMesh();
  const size_t nv = isTriMesh_ ? triMesh_->vertices.size() : tetMesh_->vertices.size();
  std::vector<float>& vt = isTriMesh_ ? triMesh_->vertT : tetMesh_->vertT;
  std::vector<int32_t> seeds;
  if (!userSetInitial_) {
    vt.assign(nv, isTriMesh_ ? EIK_LARGENUM_2D : EIK_LARGENUM_3D);
    vt[0] = 0.f;
    seeds.push_back(0);
  } else {
    for (size_t i = 0; i < nv; ++i)
      if (vt[i] == 0.f) seeds.push_back((int32_t)i);
  }
  if (isTriMesh_) {
    if (speedType_ != ONE && speedMtxMultipliers_.size() != triMesh_->faces.size())
      std::fprintf(stderr, "speed type %d is not supported; using speed 1\n", speedType_);
    check(eik_mesh_set_speed(mesh_, speedMtxMultipliers_.data(),
                             (int64_t)speedMtxMultipliers_.size()));
  }
  int32_t nb = 0;
  if (isStructured_) {
    if (isTriMesh_)
      check(eik_partition_structured(mesh_, squareLength_, squareWidth_, 1, squareBlockLength_,
                                     squareBlockWidth_, 1, &nb));
    else
      check(eik_partition_structured(mesh_, squareLength_, squareWidth_, squareDepth_,
                                     squareBlockLength_, squareBlockWidth_, squareBlockDepth_,
                                     &nb));
  } else {
    check(eik_partition(mesh_, maxVertsPerBlock_, &nb));
  }
  maxBlocks_ = nb;
  iteration_values_.clear();
  if (seeds.empty()) {
    // nothing to propagate from: every value stays unknown
    iteration_values_.push_back(std::vector<float>(nv, isTriMesh_ ? EIK_LARGENUM_2D : EIK_LARGENUM_3D));
    return;
  }
  eik_solve_opts_t o;
  eik_solve_opts_default(&o);
  o.scheduler = scheduler_;
  o.max_iterations = maxIterations_;
  o.record_history = scheduler_ == EIK_SCHED_SYNC ? 1 : 0;
  std::vector<float> T(nv);
  eik_solve_stats_t st;
  check(eik_solve(mesh_, seeds.data(), (int64_t)seeds.size(), nullptr, &o, T.data(), &st));
  for (int64_t i = 0; i < st.history_len; ++i) {
    std::vector<float> h(nv);
    check(eik_get_history(mesh_, i, h.data()));
    iteration_values_.push_back(std::move(h));
  }
  if (iteration_values_.empty()) iteration_values_.push_back(std::move(T));
  if (verbose_)
    std::printf("solve: %lld outer iterations, %lld block visits, %.3f ms on the device\n",
                (long long)st.outer_iterations, (long long)st.block_visits, st.kernel_ms);
}

void Eikonal::writeVTK(bool all) {
  const bool tri = isTriMesh_;
  if ((tri && !triMesh_) || (!tri && !tetMesh_) || iteration_values_.empty()) return;
  const std::vector<point>& V = tri ? triMesh_->vertices : tetMesh_->vertices;
  const size_t ne = tri ? triMesh_->faces.size() : tetMesh_->tets.size();
  const int k = tri ? 3 : 4;
  for (size_t j = all ? 0 : iteration_values_.size() - 1; j < iteration_values_.size(); ++j) {
    const std::string name = "result" + std::to_string(all ? j : 0) + ".vtk";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", name.c_str());
      continue;
    }
    std::fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(f, "POINTS %zu float\n", V.size());
    for (const point& p : V) std::fprintf(f, "%.12f %.12f %.12f\n", p[0], p[1], p[2]);
    std::fprintf(f, "CELLS %zu %zu\n", ne, ne * (k + 1));
    for (size_t e = 0; e < ne; ++e) {
      std::fprintf(f, "%d", k);
      for (int c = 0; c < k; ++c) std::fprintf(f, " %d", tri ? triMesh_->faces[e][c] : tetMesh_->tets[e][c]);
      std::fprintf(f, "\n");
    }
    std::fprintf(f, "CELL_TYPES %zu\n", ne);
    for (size_t e = 0; e < ne; ++e) std::fprintf(f, "%d\n", tri ? 5 : 10);  // VTK_TRIANGLE / VTK_TETRA
    std::fprintf(f, "POINT_DATA %zu\nSCALARS traveltime float 1\nLOOKUP_TABLE default\n", V.size());
    for (float v : iteration_values_[j]) std::fprintf(f, "%.12f\n", v);
    std::fclose(f);
  }
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
