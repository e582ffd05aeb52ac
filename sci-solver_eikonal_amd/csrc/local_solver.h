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
// discriminant is evaluated with an fma-compensated product (Kahan), and ill-conditioned
// elements (det(G) below 1e-4 of the product of its diagonal: flat tets, needle triangles) are
// solved in fp64.  A candidate below the exact lower bound t0 + dist(D, opposite face) is
// rejected as a rounding artefact.
#pragma once

#include <hip/hip_runtime.h>

namespace eik {

#define EIK_INF __builtin_huge_valf()

// Arithmetic helpers for the fp32 and fp64 variants of the solvers.
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float sqrt_(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double sqrt_(double a) { return __builtin_sqrt(a); }

// b*b - a*c with the rounding error of a*c compensated (exact product via fma, Kahan).
template <typename R>
__device__ __forceinline__ R disc_comp(R a, R b, R c) {
  R p = a * c;
  R e = fma_(a, c, -p);
  return fma_(b, b, -p) - e;
}

// The quadratic (tau - d 1)^T K (tau - d 1) = 1 in the form A d^2 - 2 B d + C = 0 with
// A = 1^T K 1, B = 1^T K tau, C = tau^T K tau - 1; r = K 1, w = K tau.  Its larger root gives
// the foot point of the characteristic: barycentric weights lambda_i = (d r_i - w_i) / S,
// S = d A - B = sqrt(disc), causal when all are >= 0.  The candidate is then the Hopf-Lax cost
// evaluated at that foot point, lambda.tau + |D - E|, with
//     |D - E|^2 = sum_i lambda_i L_Di - sum_{i<j} lambda_i lambda_j L_ij
// from the squared edge lengths: the cost of an actual path through the element, so rounding
// in the root can only raise the candidate, never lower it below the exact minimum (the FIM
// keeps the smallest candidate it ever sees, so a candidate rounded low would bias the result).
// Returns the candidate minus t0, or +inf when not causal.
template <typename R, int N>
__device__ __forceinline__ R causal_cost(R A, R B, R C, const R* r, const R* w, const R* a,
                                         const R* LD, const R* LX) {
  R disc = disc_comp(A, B, C);
  if (!(disc >= R(0)) || !(A > R(0))) return R(EIK_INF);
  R sd = sqrt_(disc);
  R d = (B + sd) / A;
  R lam[N];
  R S = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lam[i] = d * r[i] - w[i];
    if (lam[i] < R(0)) return R(EIK_INF);
    S += lam[i];
  }
  if (!(S > R(0))) return R(EIK_INF);
  const R inv = R(1) / S;
  R lt = R(0), x = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lam[i] *= inv;
    lt += lam[i] * a[i];
    x += lam[i] * LD[i];
  }
  // pairwise terms: (0,1) for N = 2; (0,1), (0,2), (1,2) for N = 3 (LX in that order)
  x -= lam[0] * lam[1] * LX[0];
  if constexpr (N == 3) x -= lam[0] * lam[2] * LX[1] + lam[1] * lam[2] * LX[2];
  return lt + sqrt_(x > R(0) ? x : R(0));
}

// Two known corners (face of a tet, or a triangle).  L1 = |X1 - D|^2, L2 = |X2 - D|^2,
// L12 = |X1 - X2|^2; a1, a2 = t_i - t0 >= 0.  Returns the candidate minus t0, or +inf.
template <typename R>
__device__ __forceinline__ R face_cost(R a1, R a2, R L1, R L2, R L12) {
  const R g12 = R(0.5) * (L1 + L2 - L12);
  R det = fma_(L1, L2, -g12 * g12);
  if (!(det > R(0))) return R(EIK_INF);
  R inv = R(1) / det;
  R k11 = L2 * inv, k22 = L1 * inv, k12 = -g12 * inv;
  R r[2] = {k11 + k12, k12 + k22};                      // K 1
  R w[2] = {k11 * a1 + k12 * a2, k12 * a1 + k22 * a2};  // K tau
  R A = r[0] + r[1];
  R B = w[0] + w[1];
  R C = a1 * w[0] + a2 * w[1] - R(1);
  const R a[2] = {a1, a2}, LD[2] = {L1, L2}, LX[1] = {L12};
  return causal_cost<R, 2>(A, B, C, r, w, a, LD, LX);
}

// Three known corners; L1..L3 = |X_i - D|^2, L12, L13, L23 = |X_i - X_j|^2.
template <typename R>
__device__ __forceinline__ R interior_cost(R a1, R a2, R a3, R L1, R L2, R L3, R L12, R L13,
                                           R L23) {
  const R g11 = L1, g22 = L2, g33 = L3;
  const R g12 = R(0.5) * (L1 + L2 - L12);
  const R g13 = R(0.5) * (L1 + L3 - L13);
  const R g23 = R(0.5) * (L2 + L3 - L23);
  R c11 = g22 * g33 - g23 * g23;
  R c22 = g11 * g33 - g13 * g13;
  R c33 = g11 * g22 - g12 * g12;
  R c12 = g13 * g23 - g12 * g33;
  R c13 = g12 * g23 - g13 * g22;
  R c23 = g12 * g13 - g11 * g23;
  R det = g11 * c11 + g12 * c12 + g13 * c13;
  if (!(det > R(0))) return R(EIK_INF);
  R inv = R(1) / det;
  c11 *= inv; c22 *= inv; c33 *= inv; c12 *= inv; c13 *= inv; c23 *= inv;
  R r[3] = {c11 + c12 + c13, c12 + c22 + c23, c13 + c23 + c33};
  R w[3] = {c11 * a1 + c12 * a2 + c13 * a3, c12 * a1 + c22 * a2 + c23 * a3,
            c13 * a1 + c23 * a2 + c33 * a3};
  R A = r[0] + r[1] + r[2];
  R B = w[0] + w[1] + w[2];
  R C = a1 * w[0] + a2 * w[1] + a3 * w[2] - R(1);
  const R a[3] = {a1, a2, a3}, LD[3] = {L1, L2, L3}, LX[3] = {L12, L13, L23};
  return causal_cost<R, 3>(A, B, C, r, w, a, LD, LX);
}

// Precision of the local solves.  The stationary-point algebra goes through the inverse Gram
// matrix, which loses about log10(1/cond) digits on flat elements (the reference fixtures
// include tets with det(G) / (G11 G22 G33) ~ 1e-10); the solves therefore run in fp64 (on
// MI355X the fp64 vector FMA rate equals the unpacked fp32 rate) from the fp32 times and edge
// metrics, and the candidate is rounded to fp32.
using SolveReal = double;

// Target D of a tet from corners 1..3.  L1..L3 = |X_i - D|^2, L12, L13, L23 = |X_i - X_j|^2.
// Unknown corners (t >= large) are left out, as the reference's candidates through them are
// never the minimum.  Returns +inf when no corner is known.
__device__ __forceinline__ float tet_corner_solve(float t1, float t2, float t3, float L1, float L2,
                                                  float L3, float L12, float L13, float L23,
                                                  float large) {
  using R = SolveReal;
  const bool k1 = t1 < large, k2 = t2 < large, k3 = t3 < large;
  const R t0 = (R)__builtin_fminf(t1, __builtin_fminf(t2, t3));
  const R a1 = (R)t1 - t0, a2 = (R)t2 - t0, a3 = (R)t3 - t0;
  if (k1 && k2 && k3) {
    R v = interior_cost<R>(a1, a2, a3, L1, L2, L3, L12, L13, L23);
    if (v < R(EIK_INF)) return (float)(t0 + v);  // a causal interior point is the minimum
  }
  R best = R(EIK_INF);
  if (k1 && k2) best = fmin(best, face_cost<R>(a1, a2, L1, L2, L12));
  if (k1 && k3) best = fmin(best, face_cost<R>(a1, a3, L1, L3, L13));
  if (k2 && k3) best = fmin(best, face_cost<R>(a2, a3, L2, L3, L23));
  if (k1) best = fmin(best, a1 + sqrt_((R)L1));
  if (k2) best = fmin(best, a2 + sqrt_((R)L2));
  if (k3) best = fmin(best, a3 + sqrt_((R)L3));
  return best < R(EIK_INF) ? (float)(t0 + best) : EIK_INF;
}

// Target Z of a triangle from X1, X2.  L1 = |X1 - Z|^2, L2 = |X2 - Z|^2, L12 = |X1 - X2|^2.
__device__ __forceinline__ float tri_corner_solve(float t1, float t2, float L1, float L2, float L12,
                                                  float large) {
  using R = SolveReal;
  const bool k1 = t1 < large, k2 = t2 < large;
  const R t0 = (R)(k1 && k2 ? __builtin_fminf(t1, t2) : (k1 ? t1 : t2));
  R best = R(EIK_INF);
  if (k1 && k2) {
    best = face_cost<R>((R)t1 - t0, (R)t2 - t0, L1, L2, L12);
    if (best < R(EIK_INF)) return (float)(t0 + best);
  }
  if (k1) best = fmin(best, (R)t1 - t0 + sqrt_((R)L1));
  if (k2) best = fmin(best, (R)t2 - t0 + sqrt_((R)L2));
  return best < R(EIK_INF) ? (float)(t0 + best) : EIK_INF;
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
