/*
 * fim_oracle.c -- FP64 CPU oracle for the Fast Iterative Method (FIM) Eikonal hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this file: it is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as
 * the reference to compare with (or the timed CPU baseline), never as the thing measured or
 * shipped.
 *
 * It restates, in double precision, the arithmetic of SCIInstitute/SCI-Solver_Eikonal on the
 * solve path:
 *   - orc_tet_local        : localSolverTet1, src/core3d/cuda/CUDA_kernels.cu:69-225
 *   - orc_tet_products     : gettetmem + vMv, src/core3d/cuda/meshFIM3dEikonal.cu:650-678,
 *                            src/core3d/include/Vec.h:479-483; derived products
 *                            src/core3d/cuda/CUDA_kernels.cu:419-437
 *   - orc_tet_vertex_solve : the four rotations of FIMCuda, CUDA_kernels.cu:453-456
 *   - orc_tri_iso_corner   : the 2D corner solve, src/core2d/cuda/CUDA_kernels.cu:134-206
 *   - orc_tri_aniso_corner : the lambda1 = 0 face branch of localSolverTet1 (CUDA_kernels.cu:84-111)
 *                            applied to a triangle (the reference has no anisotropic 2D kernel)
 *   - orc_fim3d / orc_fim2d: a serial vertex-level FIM (active list, Jeong & Whitaker) run to exact
 *                            convergence ("no value decreases"), which is the fixed point the
 *                            block FIM of meshFIM3dEikonal.cu:411-542 / meshFIM2d_eikonal.cu:683-816
 *                            converges to (the reference stops at |dT| < EPS = 1e-5,
 *                            src/core3d/include/CUDADefines.h:20).  Sources start at their
 *                            values (0 in the reference) and, like every vertex, only ever
 *                            decrease: with start times, a source reached earlier from another
 *                            source takes the earlier time
 *   - orc_*_sweep          : a brute-force Bellman-Ford style relaxation over all elements
 *                            (no active list), a second algorithm for the same fixed point
 *   - orc_tri_virtual_faces: IsNonObtuse / SplitFace / need_face_virtual_faces,
 *                            src/core2d/include/TriMesh.h:191-313,
 *                            src/core2d/cuda/TriMesh_connectivity.cu:206-278
 *
 * Pinning (see tests/test_oracle_*.py): known answers of the reference's tests (sphere8 and
 * sphere13 = 19.58, src/test/3d/3D-small.cc:3-22), brute-force minimisation of the Hopf-Lax
 * functional on random elements, plane-wave exactness, the Euclidean lower bound / Dijkstra
 * upper bound on homogeneous meshes, iso == aniso at M = F^2 I, and agreement of the two
 * independent FIM drivers.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_LARGENUM_3D 1e9   /* src/core3d/include/tetmesh.h:11-13 */
#define ORC_LARGENUM_2D 1e7   /* src/core2d/include/TriMesh.h:11-13 */
#define ORC_SMALLNUM 1e-8     /* src/core3d/include/tetmesh.h:14-16 */

static inline double dmin(double a, double b) { return fmin(a, b); }

/* Regularisation added to h in the interior weight lambda2 = (s*lambda1 + t)/(h + reg),
 * CUDA_kernels.cu:200 (reg = SMALLNUM = 1e-8).  With reg = 0 the interior is the exact
 * stationary point of the reference's own derivation, completed where it is ill-conditioned
 * (interior_extra).
 * The absolute 1e-8 misses the interior minimum whenever |h| is not much larger than 1e-8, which
 * happens on unit-scale meshes; orc_set_interior_reg(0) gives the unregularised formula. */
static double g_interior_reg = ORC_SMALLNUM;
void orc_set_interior_reg(double reg) { g_interior_reg = reg; }
double orc_get_interior_reg(void) { return g_interior_reg; }

/* Block layout records of the product (read by the blocked FIM below). */
typedef struct {
  uint16_t v[4];
  float L[6];
  float H[4]; /* corner heights (a pruning bound of the product's kernels; unused here) */
} orc_tetrec_t;
typedef struct {
  uint16_t v[3];
  uint16_t mask;
  float L[3];
  float H[3];
} orc_trirec_t;

/* The local solvers and the blocked FIM, in two precisions from one text (fim_local.inc): FP64
 * (the names without suffix: the oracle) and FP32 (suffix _f32: the multi-core CPU baseline, the
 * reference's arithmetic width).  In the FP32 variant the double literals promote the
 * expressions they appear in, as the reference's own double literals do (the divides by
 * 2.0 * p and by h + SMALLNUM, CUDA_kernels.cu:93-200). */
#define REAL double
#define SQRT sqrt
#define FABS fabs
#define RMIN fmin
#define NAME(f) f
#include "fim_local.inc"
#undef REAL
#undef SQRT
#undef FABS
#undef RMIN
#undef NAME
#define REAL float
#define SQRT sqrtf
#define FABS fabsf
#define RMIN fminf
#define NAME(f) f##_f32
#include "fim_local.inc"
#undef REAL
#undef SQRT
#undef FABS
#undef RMIN
#undef NAME


/* ------------------------------------------------------------------------------------------ */
/* 3D local solver: localSolverTet1 (src/core3d/cuda/CUDA_kernels.cu:69-225), in FP64.         */
/* Inputs: the three known corner times and the six metric products of AC, BC, CD.            */
/* Returns the minimum over the candidate set {face BCD, face ABD, face ACD, their edges,      */
/* interior stationary point}.                                                                 */
/* ------------------------------------------------------------------------------------------ */






/* ------------------------------------------------------------------------------------------ */
/* Tet metric products.                                                                        */
/* vMv(v1, M, v2) with M stored upper-triangular [0 1 2; 1 3 4; 2 4 5] (Vec.h:479-483).        */
/* gettetmem stores ACBC, BCCD, ACCD, ADBD, ACAD, BCBD (meshFIM3dEikonal.cu:669-674); FIMCuda  */
/* derives the rest (CUDA_kernels.cu:429-437).  out10 = {ACAC, ACBC, ACCD, BCBC, BCCD, CDCD,   */
/* ADAD, ADBD, ADCD, BDBD, CDBD} (11 values).                                                  */
/* ------------------------------------------------------------------------------------------ */
static double vMv(const double* a, const double* M, const double* b) {
  return (a[0] * M[0] + a[1] * M[1] + a[2] * M[2]) * b[0] +
         (a[0] * M[1] + a[1] * M[3] + a[2] * M[4]) * b[1] +
         (a[0] * M[2] + a[1] * M[4] + a[2] * M[5]) * b[2];
}

void orc_tet_products(const double* A, const double* B, const double* C, const double* D,
                      const double* M, double* out) {
  double AC[3], AD[3], BC[3], CD[3], BD[3];
  for (int i = 0; i < 3; ++i) {
    AC[i] = C[i] - A[i];
    AD[i] = D[i] - A[i];
    BC[i] = C[i] - B[i];
    CD[i] = D[i] - C[i];
    BD[i] = D[i] - B[i];
  }
  double ACBC = vMv(AC, M, BC), BCCD = vMv(BC, M, CD), ACCD = vMv(AC, M, CD);
  double ADBD = vMv(AD, M, BD), ACAD = vMv(AC, M, AD), BCBD = vMv(BC, M, BD);
  double ADBC = ACBC + BCCD;
  double ACBD = ACBC + ACCD;
  double CDBD = ADBD - ACBD;
  double ADCD = ADBD - ADBC;
  double CDCD = ADCD - ACCD;
  double ADAD = ADCD + ACAD;
  double ACAC = ACAD - ACCD;
  double BDBD = BCBD + CDBD;
  double BCBC = BCBD - BCCD;
  out[0] = ACAC; out[1] = ACBC; out[2] = ACCD; out[3] = BCBC; out[4] = BCCD; out[5] = CDCD;
  out[6] = ADAD; out[7] = ADBD; out[8] = ADCD; out[9] = BDBD; out[10] = CDBD;
}


/* ------------------------------------------------------------------------------------------ */
/* 2D local solvers                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* Isotropic corner solve of the 2D kernel (src/core2d/cuda/CUDA_kernels.cu:134-206):
 * corner C with neighbours A, B; F = speedInv of the face; returns min(TC, candidates).
 * The caller applies the 'all three LARGENUM' skip of CUDA_kernels.cu:129. */
double orc_tri_iso_corner(double TA, double TB, double TC, double LenAB, double LenBC,
                          double LenAC, double F) {
  const double L = ORC_LARGENUM_2D;
  double TC1 = L, TC2 = L;
  double TAB = TB - TA;
  double a = (F * F * LenAB * LenAB - TAB * TAB) * LenAB * LenAB;
  double EdgeTA = TA + LenAC * F;
  double EdgeTB = TB + LenBC * F;
  if (a > 0.0) {
    double cosA = (LenAC * LenAC + LenAB * LenAB - LenBC * LenBC) / (2.0 * LenAC * LenAB);
    double b = 2.0 * LenAB * LenAC * cosA * (TAB * TAB - F * F * LenAB * LenAB);
    double delta = 4.0 * LenAC * LenAC * a * TAB * TAB * (1.0 - cosA * cosA);
    double sd = sqrt(delta);
    double l1 = (-b + sd) / (2.0 * a);
    double l2 = (-b - sd) / (2.0 * a);
    if (l1 >= 0.0 && l1 <= 1.0) {
      double LenAD = l1 * LenAB;
      double LenCD = sqrt(LenAC * LenAC + LenAD * LenAD - 2.0 * LenAC * LenAD * cosA);
      TC1 = l1 * TAB + TA + LenCD * F;
    }
    if (l2 >= 0.0 && l2 <= 1.0) {
      double LenAD = l2 * LenAB;
      double LenCD = sqrt(LenAC * LenAC + LenAD * LenAD - 2.0 * LenAC * LenAD * cosA);
      TC2 = l2 * TAB + TA + LenCD * F;
    }
    return dmin(TC, dmin(TC2, dmin(TC1, dmin(EdgeTA, EdgeTB))));
  }
  return dmin(TC, dmin(EdgeTA, EdgeTB));
}



/* ------------------------------------------------------------------------------------------ */
/* Incidence helpers                                                                           */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int n;
  int* off; /* n+1 */
  int* idx; /* off[n] */
} csr_t;

static void csr_free(csr_t* c) {
  free(c->off);
  free(c->idx);
  c->off = c->idx = NULL;
}

/* element -> vertex incidence (elements of arity k) turned into vertex -> (element*k + corner) */
static csr_t vert_to_corner(int nv, int ne, int k, const int* elems) {
  csr_t c;
  c.n = nv;
  c.off = (int*)calloc((size_t)nv + 1, sizeof(int));
  for (int64_t i = 0; i < (int64_t)ne * k; ++i) c.off[elems[i] + 1]++;
  for (int v = 0; v < nv; ++v) c.off[v + 1] += c.off[v];
  c.idx = (int*)malloc(sizeof(int) * (size_t)c.off[nv] + 1);
  int* fill = (int*)malloc(sizeof(int) * ((size_t)nv + 1));
  memcpy(fill, c.off, sizeof(int) * ((size_t)nv + 1));
  for (int64_t i = 0; i < (int64_t)ne * k; ++i) c.idx[fill[elems[i]]++] = (int)i;
  free(fill);
  return c;
}

/* vertex neighbours (vertices sharing an element) from the corner incidence */
static csr_t vert_neighbors(int nv, int k, const int* elems, const csr_t* v2c) {
  csr_t c;
  c.n = nv;
  c.off = (int*)calloc((size_t)nv + 1, sizeof(int));
  int* mark = (int*)malloc(sizeof(int) * (size_t)nv);
  for (int v = 0; v < nv; ++v) mark[v] = -1;
  /* two passes: count then fill */
  for (int pass = 0; pass < 2; ++pass) {
    int64_t pos = 0;
    for (int v = 0; v < nv; ++v) {
      if (pass == 1) pos = c.off[v];
      int cnt = 0;
      for (int j = v2c->off[v]; j < v2c->off[v + 1]; ++j) {
        int e = v2c->idx[j] / k;
        for (int m = 0; m < k; ++m) {
          int u = elems[(int64_t)e * k + m];
          if (u == v || mark[u] == v + pass * nv) continue;
          mark[u] = v + pass * nv;
          if (pass == 1) c.idx[pos++] = u;
          cnt++;
        }
      }
      if (pass == 0) c.off[v + 1] = cnt;
    }
    if (pass == 0) {
      for (int v = 0; v < nv; ++v) c.off[v + 1] += c.off[v];
      c.idx = (int*)malloc(sizeof(int) * (size_t)c.off[nv] + 1);
      for (int v = 0; v < nv; ++v) mark[v] = -1;
    }
  }
  free(mark);
  return c;
}

/* ------------------------------------------------------------------------------------------ */
/* 3D FIM drivers                                                                              */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int nv, nt;
  const int* tets;
  double* prod; /* 11 per tet */
  csr_t v2c, nbr;
} tetctx_t;

static void tetctx_init(tetctx_t* c, int nv, const double* xyz, int nt, const int* tets,
                        const double* M6) {
  static const double I6[6] = {1, 0, 0, 1, 0, 1};
  c->nv = nv;
  c->nt = nt;
  c->tets = tets;
  c->prod = (double*)malloc(sizeof(double) * 11 * (size_t)nt + 8);
  for (int t = 0; t < nt; ++t) {
    const int* v = tets + 4 * (int64_t)t;
    const double* M = M6 ? M6 + 6 * (int64_t)t : I6;
    orc_tet_products(xyz + 3 * (int64_t)v[0], xyz + 3 * (int64_t)v[1], xyz + 3 * (int64_t)v[2],
                     xyz + 3 * (int64_t)v[3], M, c->prod + 11 * (int64_t)t);
  }
  c->v2c = vert_to_corner(nv, nt, 4, tets);
  c->nbr = vert_neighbors(nv, 4, tets, &c->v2c);
}

static void tetctx_free(tetctx_t* c) {
  free(c->prod);
  csr_free(&c->v2c);
  csr_free(&c->nbr);
}

/* min over the incident tets of the rotated local solve (FIMCuda, CUDA_kernels.cu:453-456 and
 * the slot minimum :460-474) */
static double tet_vertex_solve(const tetctx_t* c, const double* T, int v, int64_t* nsolves) {
  double best = ORC_LARGENUM_3D;
  for (int j = c->v2c.off[v]; j < c->v2c.off[v + 1]; ++j) {
    int code = c->v2c.idx[j];
    int t = code >> 2, k = code & 3;
    const int* tv = c->tets + 4 * (int64_t)t;
    double Tc[4] = {T[tv[0]], T[tv[1]], T[tv[2]], T[tv[3]]};
    best = dmin(best, orc_tet_corner(Tc, c->prod + 11 * (int64_t)t, k));
    (*nsolves)++;
  }
  return best;
}

static void init_times(int nv, double L, int nseeds, const int* seeds, const double* seed_vals,
                       double* T) {
  for (int v = 0; v < nv; ++v) T[v] = L;
  for (int s = 0; s < nseeds; ++s) T[seeds[s]] = dmin(T[seeds[s]], seed_vals ? seed_vals[s] : 0.0);
}

/* Generic active-list FIM over a vertex solve callback. */
typedef double (*vsolve_fn)(const void* ctx, const double* T, int v, int64_t* nsolves);

static int64_t fim_active_list(int nv, const csr_t* nbr, const void* ctx, vsolve_fn solve,
                               int nseeds, const int* seeds, double* T, int64_t* nsolves_out) {
  int* queue = (int*)malloc(sizeof(int) * ((size_t)nv + 1));
  char* inlist = (char*)calloc((size_t)nv, 1);
  int64_t head = 0, count = 0, nupd = 0, nsolves = 0;
  int cap = nv + 1;
  /* initial active list: neighbours of the seeds.  Sources are not frozen: a source whose start
   * time exceeds another source's start time plus the travel time between them takes the
   * earlier time (with zero start times, as in the reference, no source ever changes). */
  for (int s = 0; s < nseeds; ++s) {
    int v = seeds[s];
    for (int j = nbr->off[v]; j < nbr->off[v + 1]; ++j) {
      int u = nbr->idx[j];
      if (!inlist[u]) {
        inlist[u] = 1;
        queue[(head + count) % cap] = u;
        count++;
      }
    }
  }
  while (count > 0) {
    int v = queue[head];
    head = (head + 1) % cap;
    count--;
    double p = T[v];
    double q = solve(ctx, T, v, &nsolves);
    nupd++;
    if (q < p) {
      T[v] = q; /* still changing: stays active */
      queue[(head + count) % cap] = v;
      count++;
      continue;
    }
    /* converged: remove and probe neighbours */
    inlist[v] = 0;
    for (int j = nbr->off[v]; j < nbr->off[v + 1]; ++j) {
      int u = nbr->idx[j];
      if (inlist[u]) continue;
      double pu = T[u];
      double qu = solve(ctx, T, u, &nsolves);
      nupd++;
      if (qu < pu) {
        T[u] = qu;
        inlist[u] = 1;
        queue[(head + count) % cap] = u;
        count++;
      }
    }
  }
  free(queue);
  free(inlist);
  if (nsolves_out) *nsolves_out = nsolves;
  return nupd;
}

static double tet_solve_cb(const void* ctx, const double* T, int v, int64_t* ns) {
  return tet_vertex_solve((const tetctx_t*)ctx, T, v, ns);
}

/* Serial vertex-level FIM to exact convergence.  M6: 6 metric entries per tet (NULL = I).
 * Returns the number of vertex updates (>= 0) or -1 on bad input. */
int64_t orc_fim3d(int nv, const double* xyz, int nt, const int* tets, const double* M6,
                  int nseeds, const int* seeds, const double* seed_vals, double* T,
                  int64_t* nsolves) {
  if (nv <= 0 || nt < 0 || nseeds <= 0) return -1;
  for (int64_t i = 0; i < (int64_t)nt * 4; ++i)
    if (tets[i] < 0 || tets[i] >= nv) return -1;
  for (int s = 0; s < nseeds; ++s)
    if (seeds[s] < 0 || seeds[s] >= nv) return -1;
  tetctx_t c;
  tetctx_init(&c, nv, xyz, nt, tets, M6);
  init_times(nv, ORC_LARGENUM_3D, nseeds, seeds, seed_vals, T);
  int64_t n = fim_active_list(nv, &c.nbr, &c, tet_solve_cb, nseeds, seeds, T, nsolves);
  tetctx_free(&c);
  return n;
}

/* Brute force: sweep every tet corner until no value decreases. */
int64_t orc_fim3d_sweep(int nv, const double* xyz, int nt, const int* tets, const double* M6,
                        int nseeds, const int* seeds, const double* seed_vals, double* T,
                        int max_sweeps) {
  if (nv <= 0 || nt < 0 || nseeds <= 0) return -1;
  tetctx_t c;
  tetctx_init(&c, nv, xyz, nt, tets, M6);
  init_times(nv, ORC_LARGENUM_3D, nseeds, seeds, seed_vals, T);
  int64_t sweeps = 0;
  int changed = 1;
  while (changed && sweeps < max_sweeps) {
    changed = 0;
    sweeps++;
    for (int t = 0; t < nt; ++t) {
      const int* tv = tets + 4 * (int64_t)t;
      for (int k = 0; k < 4; ++k) {
        double Tc[4] = {T[tv[0]], T[tv[1]], T[tv[2]], T[tv[3]]};
        double cand = orc_tet_corner(Tc, c.prod + 11 * (int64_t)t, k);
        if (cand < T[tv[k]]) {
          T[tv[k]] = cand;
          changed = 1;
        }
      }
    }
  }
  tetctx_free(&c);
  return changed ? -2 : sweeps;
}

/* ------------------------------------------------------------------------------------------ */
/* 2D FIM drivers                                                                              */
/* mode 0: isotropic reference corner solve, param = speedInv per face (1 per face).           */
/* mode 1: anisotropic face-branch solve, param = Gram products {e1.e1, e1.e2, e2.e2} per face */
/*         (e1 = v1 - v0, e2 = v2 - v0) under the face metric.                                 */
/* Virtual faces (vf, rows of 3 with the obtuse vertex first) only update their first corner.  */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int nv, nf, nvf, mode;
  const int* faces;
  const int* vfaces;
  const double* xyz;
  const double* param;
  const double* vparam;
  double* edge; /* 3 per face (iso): |01|, |12|, |20| */
  double* vedge;
  csr_t v2c, nbr, v2vf;
} trictx_t;

static void face_edges(const double* xyz, const int* f, double* e) {
  for (int i = 0; i < 3; ++i) {
    const double* a = xyz + 3 * (int64_t)f[i];
    const double* b = xyz + 3 * (int64_t)f[(i + 1) % 3];
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    e[i] = sqrt(dx * dx + dy * dy + dz * dz);
  }
}

/* Corner k of a face: the reference's unrolled order (CUDA_kernels.cu:134-354):
 * C=0: A=1, B=2, LenAB=e1, LenBC=e2, LenAC=e0;  C=1: A=2, B=0, LenAB=e2, LenBC=e0, LenAC=e1;
 * C=2: A=0, B=1, LenAB=e0, LenBC=e1, LenAC=e2. */
static double tri_corner(const trictx_t* c, const double* Tf, const double* e, const double* prm,
                         int k) {
  const double L = c->mode == 0 ? ORC_LARGENUM_2D : ORC_LARGENUM_3D;
  if (Tf[0] >= L && Tf[1] >= L && Tf[2] >= L) return Tf[k]; /* CUDA_kernels.cu:129 */
  if (c->mode == 0) {
    int a = (k + 1) % 3, b = (k + 2) % 3;
    double LenAB = e[a], LenBC = e[b], LenAC = e[k];
    return orc_tri_iso_corner(Tf[a], Tf[b], Tf[k], LenAB, LenBC, LenAC, prm[0]);
  }
  return dmin(Tf[k], orc_tri_aniso_vertex(Tf, prm, k));
}

static double tri_vertex_solve(const void* vctx, const double* T, int v, int64_t* ns) {
  const trictx_t* c = (const trictx_t*)vctx;
  double best = c->mode == 0 ? ORC_LARGENUM_2D : ORC_LARGENUM_3D;
  int np = c->mode == 0 ? 1 : 3;
  for (int j = c->v2c.off[v]; j < c->v2c.off[v + 1]; ++j) {
    int code = c->v2c.idx[j];
    int f = code / 3, k = code % 3;
    const int* fv = c->faces + 3 * (int64_t)f;
    double Tf[3] = {T[fv[0]], T[fv[1]], T[fv[2]]};
    best = dmin(best, tri_corner(c, Tf, c->edge + 3 * (int64_t)f, c->param + np * (int64_t)f, k));
    (*ns)++;
  }
  for (int j = c->v2vf.off[v]; j < c->v2vf.off[v + 1]; ++j) {
    int code = c->v2vf.idx[j];
    int f = code / 3, k = code % 3;
    if (k != 0) continue; /* virtual faces only update the obtuse vertex */
    const int* fv = c->vfaces + 3 * (int64_t)f;
    double Tf[3] = {T[fv[0]], T[fv[1]], T[fv[2]]};
    best = dmin(best, tri_corner(c, Tf, c->vedge + 3 * (int64_t)f, c->vparam + np * (int64_t)f, 0));
    (*ns)++;
  }
  return best;
}

int64_t orc_fim2d(int nv, const double* xyz, int nf, const int* faces, const double* param,
                  int mode, int nvf, const int* vfaces, const double* vparam, int nseeds,
                  const int* seeds, const double* seed_vals, double* T, int64_t* nsolves) {
  if (nv <= 0 || nf < 0 || nseeds <= 0 || (mode != 0 && mode != 1)) return -1;
  for (int64_t i = 0; i < (int64_t)nf * 3; ++i)
    if (faces[i] < 0 || faces[i] >= nv) return -1;
  for (int64_t i = 0; i < (int64_t)nvf * 3; ++i)
    if (vfaces[i] < 0 || vfaces[i] >= nv) return -1;
  trictx_t c;
  memset(&c, 0, sizeof c);
  c.nv = nv; c.nf = nf; c.nvf = nvf; c.mode = mode;
  c.faces = faces; c.vfaces = vfaces; c.xyz = xyz; c.param = param; c.vparam = vparam;
  c.edge = (double*)malloc(sizeof(double) * 3 * (size_t)nf + 8);
  c.vedge = (double*)malloc(sizeof(double) * 3 * (size_t)nvf + 8);
  for (int f = 0; f < nf; ++f) face_edges(xyz, faces + 3 * (int64_t)f, c.edge + 3 * (int64_t)f);
  for (int f = 0; f < nvf; ++f) face_edges(xyz, vfaces + 3 * (int64_t)f, c.vedge + 3 * (int64_t)f);
  c.v2c = vert_to_corner(nv, nf, 3, faces);
  c.v2vf = vert_to_corner(nv, nvf, 3, vfaces);
  {
    /* neighbour graph over real and virtual faces: a virtual face makes its obtuse vertex
     * depend on two vertices that need not share a real face with it */
    int* all = (int*)malloc(sizeof(int) * 3 * ((size_t)nf + (size_t)nvf) + 4);
    memcpy(all, faces, sizeof(int) * 3 * (size_t)nf);
    memcpy(all + 3 * (int64_t)nf, vfaces, sizeof(int) * 3 * (size_t)nvf);
    csr_t v2all = vert_to_corner(nv, nf + nvf, 3, all);
    c.nbr = vert_neighbors(nv, 3, all, &v2all);
    csr_free(&v2all);
    free(all);
  }
  init_times(nv, mode == 0 ? ORC_LARGENUM_2D : ORC_LARGENUM_3D, nseeds, seeds, seed_vals, T);
  int64_t n = fim_active_list(nv, &c.nbr, &c, tri_vertex_solve, nseeds, seeds, T, nsolves);
  free(c.edge);
  free(c.vedge);
  csr_free(&c.v2c);
  csr_free(&c.v2vf);
  csr_free(&c.nbr);
  return n;
}

int64_t orc_fim2d_sweep(int nv, const double* xyz, int nf, const int* faces, const double* param,
                        int mode, int nvf, const int* vfaces, const double* vparam, int nseeds,
                        const int* seeds, const double* seed_vals, double* T, int max_sweeps) {
  if (nv <= 0 || nf < 0 || nseeds <= 0 || (mode != 0 && mode != 1)) return -1;
  trictx_t c;
  memset(&c, 0, sizeof c);
  c.nv = nv; c.nf = nf; c.nvf = nvf; c.mode = mode;
  c.faces = faces; c.vfaces = vfaces; c.xyz = xyz; c.param = param; c.vparam = vparam;
  c.edge = (double*)malloc(sizeof(double) * 3 * (size_t)nf + 8);
  c.vedge = (double*)malloc(sizeof(double) * 3 * (size_t)nvf + 8);
  for (int f = 0; f < nf; ++f) face_edges(xyz, faces + 3 * (int64_t)f, c.edge + 3 * (int64_t)f);
  for (int f = 0; f < nvf; ++f) face_edges(xyz, vfaces + 3 * (int64_t)f, c.vedge + 3 * (int64_t)f);
  init_times(nv, mode == 0 ? ORC_LARGENUM_2D : ORC_LARGENUM_3D, nseeds, seeds, seed_vals, T);
  int np = mode == 0 ? 1 : 3;
  int64_t sweeps = 0;
  int changed = 1;
  while (changed && sweeps < max_sweeps) {
    changed = 0;
    sweeps++;
    for (int pass = 0; pass < 2; ++pass) {
      int n = pass == 0 ? nf : nvf;
      const int* el = pass == 0 ? faces : vfaces;
      const double* ed = pass == 0 ? c.edge : c.vedge;
      const double* pr = pass == 0 ? param : vparam;
      for (int f = 0; f < n; ++f) {
        const int* fv = el + 3 * (int64_t)f;
        for (int k = 0; k < 3; ++k) {
          if (pass == 1 && k != 0) continue;
          double Tf[3] = {T[fv[0]], T[fv[1]], T[fv[2]]};
          double cand = tri_corner(&c, Tf, ed + 3 * (int64_t)f, pr + np * (int64_t)f, k);
          if (cand < T[fv[k]]) {
            T[fv[k]] = cand;
            changed = 1;
          }
        }
      }
    }
  }
  free(c.edge);
  free(c.vedge);
  return changed ? -2 : sweeps;
}

/* ------------------------------------------------------------------------------------------ */
/* Virtual faces for obtuse corners (src/core2d/include/TriMesh.h:191-313,                     */
/* src/core2d/cuda/TriMesh_connectivity.cu:206-278).  Coordinates are taken as float and the   */
/* obtuse test is done in float like the reference (acos in float, compared against            */
/* M_PI/2.0f in double), so right angles are classified exactly as the reference's rounding    */
/* does.  Output rows (v1, v3, v4) / (v1, v2, v3): the obtuse vertex first, plus the index of  */
/* the originating face.  Returns the number of virtual faces, or -1 if cap is exceeded.       */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  const float* xyz;
  const int* faces;
  const int* across; /* 3 per face, -1 on boundary */
  int* out;
  int* origin;
  int count, cap, origin_face, depth_guard;
} vfctx_t;

static float fdist(const float* a, const float* b) {
  float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

static int tri_indexof(const int* f, int v) { return f[0] == v ? 0 : f[1] == v ? 1 : f[2] == v ? 2 : -1; }

/* TriMesh::IsNonObtuse, TriMesh.h:191-212 */
int orc_is_non_obtuse(const float* xyz, const int* f, int v) {
  int iv = tri_indexof(f, v);
  const float* A = xyz + 3 * (int64_t)v;
  const float* B = xyz + 3 * (int64_t)f[(iv + 1) % 3];
  const float* C = xyz + 3 * (int64_t)f[(iv + 2) % 3];
  float a = fdist(B, C), b = fdist(A, C), c = fdist(A, B);
  float ang = 0.0f;
  if (a > 0 && b > 0 && c > 0) ang = acosf((b * b + c * c - a * a) / (2 * b * c));
  return (double)ang < 3.14159265358979323846 / 2.0f;
}

static void split_face(vfctx_t* c, int v, const int* cf, int nfadj) {
  if (++c->depth_guard > 4096) return;
  int iv = tri_indexof(cf, v);
  int v1 = v, v2 = cf[(iv + 1) % 3], v4 = cf[(iv + 2) % 3];
  const int* adj = c->faces + 3 * (int64_t)nfadj;
  int v3 = adj[(tri_indexof(adj, v2) + 1) % 3];
  int f1[3] = {v1, v3, v4}, f2[3] = {v1, v2, v3};
  if (orc_is_non_obtuse(c->xyz, f1, v)) {
    if (c->count < c->cap) {
      memcpy(c->out + 3 * (int64_t)c->count, f1, sizeof f1);
      c->origin[c->count] = c->origin_face;
    }
    c->count++;
  } else {
    int nn = c->across[3 * (int64_t)nfadj + tri_indexof(adj, v2)];
    if (nn > -1) split_face(c, v, f1, nn);
  }
  if (orc_is_non_obtuse(c->xyz, f2, v)) {
    if (c->count < c->cap) {
      memcpy(c->out + 3 * (int64_t)c->count, f2, sizeof f2);
      c->origin[c->count] = c->origin_face;
    }
    c->count++;
  } else {
    int nn = c->across[3 * (int64_t)nfadj + tri_indexof(adj, v4)];
    if (nn > -1) split_face(c, v, f2, nn);
  }
}

/* across-edge map (TriMesh_connectivity.cu:240-278): across[f][j] = face sharing the edge
 * opposite corner j with consistent orientation, -1 if none. */
void orc_tri_across_edge(int nv, int nf, const int* faces, int* across) {
  csr_t v2c = vert_to_corner(nv, nf, 3, faces);
  for (int64_t i = 0; i < (int64_t)nf * 3; ++i) across[i] = -1;
  for (int i = 0; i < nf; ++i) {
    for (int j = 0; j < 3; ++j) {
      if (across[3 * (int64_t)i + j] != -1) continue;
      int v1 = faces[3 * (int64_t)i + (j + 1) % 3];
      int v2 = faces[3 * (int64_t)i + (j + 2) % 3];
      for (int a = v2c.off[v1]; a < v2c.off[v1 + 1]; ++a) {
        int other = v2c.idx[a] / 3;
        if (other == i) continue;
        const int* of = faces + 3 * (int64_t)other;
        if (tri_indexof(of, v2) < 0) continue;
        int ind = (tri_indexof(of, v1) + 1) % 3;
        if (of[(ind + 1) % 3] != v2) continue;
        across[3 * (int64_t)i + j] = other;
        across[3 * (int64_t)other + ind] = i;
        break;
      }
    }
  }
  csr_free(&v2c);
}

int orc_tri_virtual_faces(int nv, const float* xyz, int nf, const int* faces, int cap, int* out,
                          int* origin) {
  int* across = (int*)malloc(sizeof(int) * 3 * (size_t)nf + 4);
  orc_tri_across_edge(nv, nf, faces, across);
  vfctx_t c;
  c.xyz = xyz; c.faces = faces; c.across = across; c.out = out; c.origin = origin;
  c.count = 0; c.cap = cap;
  for (int i = 0; i < nf; ++i) {
    const int* f = faces + 3 * (int64_t)i;
    for (int j = 0; j < 3; ++j) {
      if (!orc_is_non_obtuse(xyz, f, f[j])) {
        int nfae = across[3 * (int64_t)i + j];
        c.origin_face = i;
        c.depth_guard = 0;
        if (nfae > -1) split_face(&c, f[j], f, nfae);
      }
    }
  }
  free(across);
  return c.count > cap ? -1 : c.count;
}

/* ------------------------------------------------------------------------------------------ */
/* Blocked FIM over the product's block layout (CPU, FP64).                                    */
/* Runs the block schedule of meshFIM3dEikonal::GenerateData (meshFIM3dEikonal.cu:411-542):    */
/* active blocks are swept until no owned value decreases, then blocks whose suppliers changed  */
/* boundary values (or that did not converge) are visited in the next outer iteration.  Element */
/* records carry squared edge lengths; the local solves use the reference formulas above with   */
/* the Gram products recovered by polarisation.  Layout, halo and dependency errors show up as  */
/* differences from orc_fim3d / orc_fim2d.                                                      */
/* ------------------------------------------------------------------------------------------ */







