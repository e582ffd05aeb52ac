// Eikonal.h -- C++ facade with the reference's solver class interface (src/Eikonal.h:14-61), on
// top of the C ABI in eikonal_amd.h.  Code written against the reference's class (its examples
// and tests) compiles against this header unchanged: the same constructor, methods, public option
// members and defaults, and public meshes whose `vertices` can be read and which callers may
// delete and set to NULL between solves (src/test/3d/3D-improve.cc:13-34).
//
// Behaviour kept from the reference (src/Eikonal.cu):
//   * an unreadable mesh file prints a message and exits with status 1 (:102-115);
//   * initializeVertices exits with status 0 on a size mismatch or before initializeMesh
//     (:69-98); seeds are then the vertices whose value is 0 (:145-158), else vertex 0;
//   * tet speeds are taken when the mesh is initialised (TetMesh::init, :118), triangle speeds
//     when the solve starts (SetMesh, :159);
//   * the block partition overwrites maxBlocks_ with the block count (:165-175, :218-221);
//   * getResultAtIteration(i) is the time field after outer iteration i; maxIterations_ caps
//     the outer iterations (:178, :223).
// Differences: the device work runs on the calling process's current HIP device; library errors
// print the library's message and exit with status -1 (the reference's CUDA error path,
// src/core3d/include/cutil.h:36-58); vertices that no front reaches read LARGENUM in 3D as in 2D
// (the reference leaves them at 0 in 3D, meshFIM3dEikonal.cu:389).
#pragma once

#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "eikonal_amd.h"

#ifndef ONE
#define ONE 1  // speed types of the reference (core2d/include/TriMesh.h:21-26)
#endif
#ifndef CURVATURE
#define CURVATURE 2
#endif
#ifndef NOISE
#define NOISE 3
#endif

// 3-vector with the reference's element access (point = Vec<3,float>, core2d/include/Vec.h).
struct point {
  float v[3] = {0.f, 0.f, 0.f};
  float& operator[](int i) { return v[i]; }
  const float& operator[](int i) const { return v[i]; }
};
inline float len(const point& p) { return std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// Host copies of the meshes, with the members callers of the reference read.
class TriMesh {
 public:
  std::vector<point> vertices;
  std::vector<std::array<int, 3>> faces;
  std::vector<float> vertT;
};

class TetMesh {
 public:
  std::vector<point> vertices;
  std::vector<std::array<int, 4>> tets;
  std::vector<float> vertT;
};

class Eikonal {
 public:
  Eikonal(bool isTriMesh, std::string fname = "../src/test/test_data/sphere_74verts.ply",
          bool verbose = false);
  virtual ~Eikonal();
  void initializeMesh();
  void initializeVertices(std::vector<float> values);
  void initSpeedMtxMultipliers(std::vector<float> values);
  std::vector<float>& getFinalResult();
  std::vector<float>& getResultAtIteration(size_t i);
  size_t numIterations();
  // result<j>.vtk (legacy ASCII UNSTRUCTURED_GRID, point scalars "traveltime") for the last
  // iteration, or for every iteration when `all`
  void writeVTK(bool all);
  void solveEikonal();
  // RMS error against `solution` after each iteration, as a log-scale text plot
  void printErrorGraph(std::vector<float> solution);

  // options: names and defaults of the reference (src/Eikonal.cu:4-24)
  bool verbose_;
  std::string filename_;
  int maxBlocks_;
  int maxVertsPerBlock_;
  float stopDistance_;  // kept for source compatibility; the reference's kernels ignore it too
  bool isStructured_;
  bool userSetInitial_;
  int speedType_;
  int squareLength_, squareWidth_, squareDepth_;
  int squareBlockLength_, squareBlockWidth_, squareBlockDepth_;
  int maxIterations_;
  TriMesh* triMesh_;
  TetMesh* tetMesh_;
  std::vector<std::vector<float>> iteration_values_;
  bool isTriMesh_;
  std::vector<float> speedMtxMultipliers_;
  // addition: EIK_SCHED_SYNC (default) keeps the per-iteration history; EIK_SCHED_ASYNC runs the
  // persistent device scheduler and keeps the converged field only (numIterations() == 1)
  int scheduler_;

 private:
  eik_mesh_t mesh_;
  const void* loaded_;  // the TriMesh / TetMesh object mesh_ was built for
  void check(int rc);
};
