// local_solver.h -- fp32 upwind local solvers for the FIM update, device code.
//
// The update of a target vertex D from an element whose other corners X_i carry times t_i is the
// Hopf-Lax minimum over the opposite simplex:
//     T_D = min_{lambda in simplex} sum_i lambda_i t_i + || D - sum_i lambda_i X_i ||_M .
// The reference evaluates it as the minimum over the interior stationary point, the stationary
// point of each face and the edges (localSolverTet1, src/core3d/cuda/CUDA_kernels.cu:69-225;
// the 2D corner solve, src/core2d/cuda/CUDA_kernels.cu:134-206).  The same minimum is computed
// here in the gradient form of each piece: with edge vectors e_i = X_i - D, Gram matrix
// G = [e_i^T M e_j] and K = G^{-1}, the linear interpolant with |grad T|_{M^-1} = 1 gives
//     (tau - d 1)^T K (tau - d 1) = 1,   tau_i = t_i - t0,  T_D = t0 + d,
// a quadratic in d whose larger root is the candidate; it is the stationary point of the
// Hopf-Lax functional when the foot point lies inside the simplex, i.e. when every component
// of K (d 1 - tau) is >= 0 (the barycentric weights of the foot point up to a positive factor).
// The function is convex over the simplex, so a causal interior point is the minimum and the
// faces need not be evaluated; otherwise the minimum is over the faces and edges.
//
// Every Gram entry comes from squared metric edge lengths by polarisation,
//     e_i . e_j = (L_Di + L_Dj - L_ij) / 2,
// which treats all four rotations of a tet (CUDA_kernels.cu:453-456) alike.
//
// Precision: times are shifted by the smallest known corner time before the solve, the
// discriminant b^2 - a c is evaluated with an fma-compensated product (Kahan), and when it is
// small relative to b^2 (a near-double root) the root is recomputed in fp64.
#pragma once

#include <hip/hip_runtime.h>

namespace eik {

#define EIK_INF __builtin_huge_valf()

// b*b - a*c with the rounding error of a*c compensated (exact product via fma).
__device__ __forceinline__ float disc_comp(float a, float b, float c) {
  float p = a * c;
  float e = __builtin_fmaf(a, c, -p);
  return __builtin_fmaf(b, b, -p) - e;
}

// Two known corners (face of a tet, or a triangle).  g11 = |e1|^2, g22 = |e2|^2, g12 = e1.e2.
// Returns the causal candidate or +inf.
__device__ __forceinline__ float solve_face(float t1, float t2, float g11, float g12, float g22) {
  float t0 = __builtin_fminf(t1, t2);
  float a1 = t1 - t0, a2 = t2 - t0;
  float det = __builtin_fmaf(g11, g22, -g12 * g12);
  if (!(det > 0.f)) return EIK_INF;
  float inv = 1.0f / det;
  float k11 = g22 * inv, k22 = g11 * inv, k12 = -g12 * inv;
  float r1 = k11 + k12, r2 = k12 + k22;           // K 1
  float w1 = k11 * a1 + k12 * a2;                 // K tau
  float w2 = k12 * a1 + k22 * a2;
  float A = r1 + r2;
  float B = w1 + w2;
  float C = a1 * w1 + a2 * w2 - 1.0f;
  float disc = disc_comp(A, B, C);
  if (!(disc >= 0.f) || !(A > 0.f)) return EIK_INF;
  float d;
  if (disc < 1e-4f * B * B) {
    double Ad = (double)A, Bd = (double)B, Cd = (double)C;
    d = (float)((Bd + sqrt(fmax(Bd * Bd - Ad * Cd, 0.0))) / Ad);
  } else {
    d = (B + __builtin_sqrtf(disc)) / A;
  }
  if (d * r1 - w1 < 0.f || d * r2 - w2 < 0.f) return EIK_INF;
  return t0 + d;
}

// Three known corners (interior of a tet's opposite face).  G symmetric 3x3.
__device__ __forceinline__ float solve_tet_interior(float t1, float t2, float t3, float g11, float g12,
                                                    float g13, float g22, float g23, float g33) {
  float t0 = __builtin_fminf(t1, __builtin_fminf(t2, t3));
  float a1 = t1 - t0, a2 = t2 - t0, a3 = t3 - t0;
  float c11 = g22 * g33 - g23 * g23;
  float c22 = g11 * g33 - g13 * g13;
  float c33 = g11 * g22 - g12 * g12;
  float c12 = g13 * g23 - g12 * g33;
  float c13 = g12 * g23 - g13 * g22;
  float c23 = g12 * g13 - g11 * g23;
  float det = g11 * c11 + g12 * c12 + g13 * c13;
  if (!(det > 0.f)) return EIK_INF;
  float inv = 1.0f / det;
  c11 *= inv; c22 *= inv; c33 *= inv; c12 *= inv; c13 *= inv; c23 *= inv;
  float r1 = c11 + c12 + c13, r2 = c12 + c22 + c23, r3 = c13 + c23 + c33;
  float w1 = c11 * a1 + c12 * a2 + c13 * a3;
  float w2 = c12 * a1 + c22 * a2 + c23 * a3;
  float w3 = c13 * a1 + c23 * a2 + c33 * a3;
  float A = r1 + r2 + r3;
  float B = w1 + w2 + w3;
  float C = a1 * w1 + a2 * w2 + a3 * w3 - 1.0f;
  float disc = disc_comp(A, B, C);
  if (!(disc >= 0.f) || !(A > 0.f)) return EIK_INF;
  float d;
  if (disc < 1e-4f * B * B) {
    double Ad = (double)A, Bd = (double)B, Cd = (double)C;
    d = (float)((Bd + sqrt(fmax(Bd * Bd - Ad * Cd, 0.0))) / Ad);
  } else {
    d = (B + __builtin_sqrtf(disc)) / A;
  }
  if (d * r1 - w1 < 0.f || d * r2 - w2 < 0.f || d * r3 - w3 < 0.f) return EIK_INF;
  return t0 + d;
}

// Target D of a tet from corners 1..3.  L1..L3 = |X_i - D|^2, L12, L13, L23 = |X_i - X_j|^2.
// Unknown corners (t >= large) are left out, as the reference's candidates through them are
// never the minimum.  Returns +inf when no corner is known.
__device__ __forceinline__ float tet_corner_solve(float t1, float t2, float t3, float L1, float L2,
                                                  float L3, float L12, float L13, float L23,
                                                  float large) {
  bool k1 = t1 < large, k2 = t2 < large, k3 = t3 < large;
  float g12 = 0.5f * (L1 + L2 - L12);
  float g13 = 0.5f * (L1 + L3 - L13);
  float g23 = 0.5f * (L2 + L3 - L23);
  if (k1 && k2 && k3) {
    float v = solve_tet_interior(t1, t2, t3, L1, g12, g13, L2, g23, L3);
    if (v < EIK_INF) return v;
  }
  float best = EIK_INF;
  if (k1 && k2) best = __builtin_fminf(best, solve_face(t1, t2, L1, g12, L2));
  if (k1 && k3) best = __builtin_fminf(best, solve_face(t1, t3, L1, g13, L3));
  if (k2 && k3) best = __builtin_fminf(best, solve_face(t2, t3, L2, g23, L3));
  if (k1) best = __builtin_fminf(best, t1 + __builtin_sqrtf(L1));
  if (k2) best = __builtin_fminf(best, t2 + __builtin_sqrtf(L2));
  if (k3) best = __builtin_fminf(best, t3 + __builtin_sqrtf(L3));
  return best;
}

// Target Z of a triangle from X1, X2.  L1 = |X1 - Z|^2, L2 = |X2 - Z|^2, L12 = |X1 - X2|^2.
__device__ __forceinline__ float tri_corner_solve(float t1, float t2, float L1, float L2, float L12,
                                                  float large) {
  bool k1 = t1 < large, k2 = t2 < large;
  float best = EIK_INF;
  if (k1 && k2) {
    float g12 = 0.5f * (L1 + L2 - L12);
    best = solve_face(t1, t2, L1, g12, L2);
    if (best < EIK_INF) return best;
  }
  if (k1) best = __builtin_fminf(best, t1 + __builtin_sqrtf(L1));
  if (k2) best = __builtin_fminf(best, t2 + __builtin_sqrtf(L2));
  return best;
}

// Corner c (0..3) of a tet with times t0..t3 and squared edge lengths L0..L5 =
// {01,02,03,12,13,23}.  The other corners are taken in increasing order; the inputs are
// permuted with selects on scalar values so that every lane runs the same solver code whatever
// its corner (no divergent switch, no runtime-indexed arrays, which would live in scratch).
__device__ __forceinline__ float tet_rotation(int c, float t0, float t1, float t2, float t3,
                                              float L0, float L1, float L2, float L3, float L4,
                                              float L5, float large) {
  const bool c0 = c == 0, c1 = c == 1, c2 = c == 2, le1 = c <= 1, le2 = c <= 2;
  float a1 = c0 ? t1 : t0;
  float a2 = le1 ? t2 : t1;
  float a3 = le2 ? t3 : t2;
  float D1 = le1 ? L0 : (c2 ? L1 : L2);
  float D2 = c0 ? L1 : (le2 ? L3 : L4);
  float D3 = c0 ? L2 : (c1 ? L4 : L5);
  float E12 = c0 ? L3 : (c1 ? L1 : L0);
  float E13 = c0 ? L4 : (le2 ? L2 : L1);
  float E23 = le1 ? L5 : (c2 ? L4 : L3);
  return tet_corner_solve(a1, a2, a3, D1, D2, D3, E12, E13, E23, large);
}

// Corner c (0..2) of a triangle with times t0..t2 and squared edge lengths L0..L2 = {01,12,20}.
__device__ __forceinline__ float tri_rotation(int c, float t0, float t1, float t2, float L0,
                                              float L1, float L2, float large) {
  const bool c0 = c == 0, c1 = c == 1;
  float a1 = c0 ? t1 : t0;
  float a2 = (c0 || c1) ? t2 : t1;
  float D1 = (c0 || c1) ? L0 : L2;
  float D2 = c0 ? L2 : L1;
  float E12 = c0 ? L1 : (c1 ? L2 : L0);
  return tri_corner_solve(a1, a2, D1, D2, E12, large);
}

}  // namespace eik
