// Eikonal.h -- C++ facade with the reference's solver class interface (src/Eikonal.h:14-61), on
// top of the C ABI in eikonal_amd.h.  Code written against the reference's class (its examples
// and tests) compiles against this header unchanged: the same constructor, methods, public option
// members and defaults; public meshes whose `vertices` can be read and which callers may delete
// and set to NULL between solves (src/test/3d/3D-improve.cc:13-34); the lower-level solver
// objects behind FIMPtr2d_ / FIMPtr3d_ (src/Eikonal.h:57-58) with the drivers' public methods
// (meshFIM2d_eikonal.h, meshFIM3dEikonal.h); the point type with the arithmetic the examples use
// (src/core2d/include/Vec.h); and the headers the reference's Eikonal.h brings in (<cstring>,
// the C and C++ I/O headers, <vector>, <time.h>), so examples that rely on them compile.
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
// (the reference leaves them at 0 in 3D, meshFIM3dEikonal.cu:389); a tet metric tensor that is
// not positive definite is reported on stderr and the mesh keeps the unit metric instead (the
// reference takes it as given and its local solver takes square roots of negative lengths).
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "eikonal_amd.h"

// speed types of the reference (core2d/include/TriMesh.h:15-26; NOISE and MANUAL_SPEED share
// the value 3 there)
#ifndef ONE
#define ONE 1
#endif
#ifndef CURVATURE
#define CURVATURE 2
#endif
#ifndef NOISE
#define NOISE 3
#endif
#ifndef MANUAL_SPEED
#define MANUAL_SPEED 3
#endif

// 3-vector of floats (the reference's point = Vec<3,float>): element access, component-wise and
// scalar arithmetic, ^ the dot product, % the cross product, len / dist and normalize.
struct point {
  float v[3];
  point() : v{0.f, 0.f, 0.f} {}
  point(float x, float y, float z) : v{x, y, z} {}
  float& operator[](int i) { return v[i]; }
  const float& operator[](int i) const { return v[i]; }
  point& operator+=(const point& o) { for (int i = 0; i < 3; ++i) v[i] += o.v[i]; return *this; }
  point& operator-=(const point& o) { for (int i = 0; i < 3; ++i) v[i] -= o.v[i]; return *this; }
  point& operator*=(const point& o) { for (int i = 0; i < 3; ++i) v[i] *= o.v[i]; return *this; }
  point& operator/=(const point& o) { for (int i = 0; i < 3; ++i) v[i] /= o.v[i]; return *this; }
  point& operator*=(float s) { for (float& x : v) x *= s; return *this; }
  point& operator/=(float s) { for (float& x : v) x /= s; return *this; }
};
inline point operator+(point a, const point& b) { return a += b; }
inline point operator-(point a, const point& b) { return a -= b; }
inline point operator*(point a, const point& b) { return a *= b; }
inline point operator/(point a, const point& b) { return a /= b; }
inline point operator*(point a, float s) { return a *= s; }
inline point operator*(float s, point a) { return a *= s; }
inline point operator/(point a, float s) { return a /= s; }
inline point operator-(const point& a) { return point(-a[0], -a[1], -a[2]); }
inline const point& operator+(const point& a) { return a; }
inline float operator^(const point& a, const point& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline point operator%(const point& a, const point& b) {
  return point(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline bool operator==(const point& a, const point& b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }
inline bool operator!=(const point& a, const point& b) { return !(a == b); }
inline float len2(const point& p) { return p ^ p; }
inline float len(const point& p) { return std::sqrt(len2(p)); }
inline float dist2(const point& a, const point& b) { return len2(a - b); }
inline float dist(const point& a, const point& b) { return len(a - b); }
inline point normalize(point& p) {
  const float l = len(p);
  if (l > 0.f) p /= l;
  return p;
}
inline std::ostream& operator<<(std::ostream& os, const point& p) {
  return os << "(" << p[0] << ", " << p[1] << ", " << p[2] << ")";
}
inline std::istream& operator>>(std::istream& is, point& p) { return is >> p[0] >> p[1] >> p[2]; }

// Host copies of the meshes, with the members callers of the reference read.  Each owns the
// library's mesh handle (destroyed with the object), so deleting a mesh and setting the pointer
// to NULL makes the next solve load the file again, as with the reference.
class TriMesh {
 public:
  std::vector<point> vertices;
  std::vector<std::array<int, 3>> faces;
  std::vector<float> vertT;
  int speed_type_ = ONE;
  TriMesh() = default;
  TriMesh(const TriMesh&) = delete;
  TriMesh& operator=(const TriMesh&) = delete;
  ~TriMesh();
  // TriMesh::read (src/core2d/cuda/TriMesh_io.cu:154): the PLY file, or NULL if it cannot be read
  static TriMesh* read(const char* filename, bool verbose = false);
  eik_mesh_t handle() const { return h_; }

 private:
  eik_mesh_t h_ = nullptr;
  friend class Eikonal;
};

class TetMesh {
 public:
  std::vector<point> vertices;
  std::vector<std::array<int, 4>> tets;
  std::vector<float> vertT;
  TetMesh() = default;
  TetMesh(const TetMesh&) = delete;
  TetMesh& operator=(const TetMesh&) = delete;
  ~TetMesh();
  // tetgenio::load_tetmesh + TetMesh::init (src/Eikonal.cu:109-127): the TetGen .node/.ele pair
  // `basename`, with the speed multipliers or tensors of TetMesh::init; NULL if it cannot be read
  static TetMesh* read(const char* basename, const std::vector<float>& speedMtx,
                       bool verbose = false);
  eik_mesh_t handle() const { return h_; }

 private:
  eik_mesh_t h_ = nullptr;
  friend class Eikonal;
};

// The 3D driver (meshFIM3dEikonal.h): seeds, partition, solve, VTK output.  GenerateData runs
// the solve and returns the time field after every outer iteration.
class meshFIM3dEikonal {
 public:
  typedef int index;
  void SetSeedPoint(std::vector<index> seeds) { m_SeedPoints = seeds; }
  void SetMesh(TetMesh* mesh);
  // first vertex within epsilon of (x, y, z) becomes the only seed (meshFIM3dEikonal.h:40-55)
  bool FindSeedPoint(double x, double y, double z, double epsilon);
  void InitSpeedMat();  // unit metric on every tet
  void GraphPartition_METIS2(int& numBlock, int maxNumBlockVerts, bool verbose = false);
  void GraphPartition_Square(int squareLength, int squareWidth, int squareHeight, int blockLength,
                             int blockWidth, int blockHeight, bool verbose = false);
  void PartitionTets(int numBlock, bool verbose = false);
  std::vector<std::vector<float>> GenerateData(size_t maxIter = 1000, bool verbose = false);
  void writeVTK(std::vector<std::vector<float>> values);
  int m_numBlock = 0;
  // addition: statistics of the last GenerateData
  eik_solve_stats_t lastStats_ = {};

 private:
  TetMesh* m_meshPtr = nullptr;
  std::vector<index> m_SeedPoints;
};

// The 2D driver (meshFIM2d_eikonal.h).
class meshFIM2dEikonal {
 public:
  typedef int index;
  void SetSeedPoint(std::vector<index> seeds) { m_SeedPoints = seeds; }
  // face speeds (speedInv = 1 / speed per face when there is one per face, else speed 1) as
  // TriMesh::InitializeAttributes (TriMesh_stats.cu:56-69); speed types other than ONE print a
  // note and use speed 1 (in the reference they only reach the virtual faces)
  void SetMesh(TriMesh* mesh, const std::vector<float>& face_speeds, int speed_type = ONE);
  void SetStopDistance(float d) { m_StopDistance = d; }  // kept; the reference's kernel ignores it
  void GraphPartition_METIS2(int& numBlock, int maxNumBlockVerts, bool verbose = false);
  void GraphPartition_Square(int squareLength, int squareWidth, int blockLength, int blockWidth,
                             bool verbose = false);
  void PartitionFaces(int numBlock);
  void InitializeLabels(int numBlock) {}  // labels live on the device
  std::vector<std::vector<float>> GenerateData(int numBlock, size_t maxIters = 1000,
                                               bool verbose = false);
  void writeVTK(std::vector<std::vector<float>> time_values);
  eik_solve_stats_t lastStats_ = {};

 private:
  TriMesh* m_meshPtr = nullptr;
  std::vector<index> m_SeedPoints;
  float m_StopDistance = 50000.f;
};

// The solver class.  solveEikonal keeps the converged field in iteration_values_ (one entry); the
// per-outer-iteration history of the reference is built by the first numIterations(),
// getResultAtIteration() or writeVTK(true) after a solve, which solves again with one launch per
// outer iteration and replaces iteration_values_ with its snapshots (references to the previous
// entries are then invalid, as after a new solve).
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
  meshFIM2dEikonal* FIMPtr2d_;
  meshFIM3dEikonal* FIMPtr3d_;
  bool isTriMesh_;
  std::vector<float> speedMtxMultipliers_;

 private:
  // the last solve kept only its converged field; its per-iteration history is built on first use
  bool historyPending_;
  void ensureHistory();
};

// The facade's error path: the library's message on stderr, then exit(-1) (cutil.h:36-58).
void eik_facade_check(int rc);
