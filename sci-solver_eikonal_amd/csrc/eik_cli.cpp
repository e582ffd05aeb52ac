// eik_cli.cpp -- command-line solver with the flags of the reference's two example programs:
//
//   eik_cli 3d [-h] [-v] [-i INPUT] [-b MAX_BLOCKS] [-m MAX_ITER] [-x MATRIX_FILE ...]
//       as src/examples/example1.cu:39-79: a TetGen .node/.ele pair (INPUT without extension,
//       default ../src/test/test_data/sphere334), tensors (6 per tet) or scalars (1 per tet)
//       read from the -x files; -b sets maxBlocks_ (which the partition overwrites, as in the
//       reference); the error graph against the distance from the origin with -v;
//   eik_cli 2d [-h] [-v] [-i INPUT] [-b MAX_BLK_VERT] [-n MAX_ITER] [-s ONE|CURVATURE|NOISE|MANUAL]
//              [-x SCALAR_FILE ...]
//       as src/examples/example2.cu:34-84: a PLY triangle mesh, per-face speeds from the -x
//       files, the example's square_scalars.txt written next to the results (20 on faces
//       within 6 of (16, 16, 0), else 1, example2.cu:86-97), and the error graph against the
//       geodesic distance on a sphere of radius 100 from vertex 0 with -v.
//
// Both write result0.vtk (Eikonal::writeVTK(false)).  --print-options prints the parsed options
// and exits before any mesh is read (for testing the flag parsing).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "Eikonal.h"

static void usage3d() {
  std::printf("Usage: eik_cli 3d [OPTIONS]\n");
  std::printf("  -h              Show this help.\n");
  std::printf("  -v              Verbose output.\n");
  std::printf("  -i INPUT        Use this tet mesh \n");
  std::printf("  -b MAX_BLOCKS   Max # of blocks to use\n");
  std::printf("  -m MAX_ITER     Max # of iterations before quit\n");
  std::printf("  -x MATRIX_FILE  File of tensor matrices per tet (N*6 floats).\n");
  std::printf("                  N floats if considered scalar speeds per tet.\n");
}

static void usage2d() {
  std::printf("Usage: eik_cli 2d [OPTIONS]\n");
  std::printf("  -h              Show this help.\n");
  std::printf("  -v              Verbose output.\n");
  std::printf("  -i INPUT        Use this triangle mesh \n");
  std::printf("  -b MAX_BLK_VERT Max # of verts/block to use\n");
  std::printf("  -n MAX_ITER     Max # of iterations to run\n");
  std::printf("  -s SPEEDTYPE    Speed type is [ONE], CURVATURE, or NOISE.\n");
  std::printf("  -x SCALAR_FILE  File with speed scalars per face.\n");
}

// every value of the files named after -x (up to the next option), appended
static int read_values(int argc, char** argv, int i, std::vector<float>& out) {
  while (i + 1 < argc && argv[i + 1][0] != '-') {
    std::ifstream in(argv[++i]);
    if (!in) std::fprintf(stderr, "cannot read %s\n", argv[i]);
    float v;
    while (in >> v) out.push_back(v);
  }
  return i;
}

static const char* speed_name(int t) {
  return t == CURVATURE ? "CURVATURE" : t == NOISE ? "NOISE" : "ONE";
}

int main(int argc, char** argv) {
  if (argc < 2 || (std::strcmp(argv[1], "3d") != 0 && std::strcmp(argv[1], "2d") != 0)) {
    std::fprintf(stderr, "usage: eik_cli 3d|2d [OPTIONS]  (-h for the options)\n");
    return 64;
  }
  const bool tri = std::strcmp(argv[1], "2d") == 0;
  Eikonal data(tri);
  if (!tri) data.filename_ = "../src/test/test_data/sphere334";
  bool print = false;
  for (int i = 2; i < argc; ++i) {
    const char* a = argv[i];
    if (std::strcmp(a, "-v") == 0) {
      data.verbose_ = true;
    } else if (std::strcmp(a, "-i") == 0) {
      if (i + 1 >= argc) break;
      data.filename_ = argv[++i];
    } else if (std::strcmp(a, "-b") == 0) {
      if (i + 1 >= argc) break;
      if (tri)
        data.maxVertsPerBlock_ = std::atoi(argv[++i]);
      else
        data.maxBlocks_ = std::atoi(argv[++i]);
    } else if (std::strcmp(a, tri ? "-n" : "-m") == 0) {
      if (i + 1 >= argc) break;
      data.maxIterations_ = std::atoi(argv[++i]);
    } else if (tri && std::strcmp(a, "-s") == 0) {
      if (i + 1 >= argc) break;
      const std::string t = argv[++i];
      if (t == "CURVATURE")
        data.speedType_ = CURVATURE;
      else if (t == "NOISE")
        data.speedType_ = NOISE;
      else if (t == "MANUAL")
        data.speedType_ = MANUAL_SPEED;
    } else if (std::strcmp(a, "-x") == 0) {
      i = read_values(argc, argv, i, data.speedMtxMultipliers_);
    } else if (std::strcmp(a, "-h") == 0) {
      tri ? usage2d() : usage3d();
      return 0;
    } else if (std::strcmp(a, "--print-options") == 0) {
      print = true;
    }
  }
  if (print) {
    std::printf("mesh=%s\ninput=%s\nverbose=%d\nmax_blocks=%d\nmax_verts_per_block=%d\n"
                "max_iterations=%d\nspeed_type=%s\nspeed_values=%zu\n",
                tri ? "2d" : "3d", data.filename_.c_str(), data.verbose_ ? 1 : 0, data.maxBlocks_,
                data.maxVertsPerBlock_, data.maxIterations_, speed_name(data.speedType_),
                data.speedMtxMultipliers_.size());
    return 0;
  }
  std::vector<float> solution;
  if (tri) {
    data.initializeMesh();
    std::ofstream out("square_scalars.txt");
    for (const auto& f : data.triMesh_->faces) {
      const point p = (data.triMesh_->vertices[f[0]] + data.triMesh_->vertices[f[1]] +
                       data.triMesh_->vertices[f[2]]) / 3.f;
      out << (len(p - point(16.f, 16.f, 0.f)) < 6.f ? 20. : 1.) << "\n";
    }
    out.close();
    data.solveEikonal();
    data.writeVTK(false);
    const point first = data.triMesh_->vertices[0];
    for (const point& v : data.triMesh_->vertices)
      solution.push_back(100.f * std::acos(std::fmax(-1.f, std::fmin(1.f, (first ^ v) / len(v) / len(first)))));
  } else {
    data.solveEikonal();
    data.writeVTK(false);
    for (const point& v : data.tetMesh_->vertices) solution.push_back(len(v));
  }
  if (data.verbose_) data.printErrorGraph(solution);
  return 0;
}
