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
// Precision: times are shifted by the smallest known corner time and the inputs scaled to unit
// trace before the solve, the discriminant is evaluated with an fma-compensated product
// (Kahan), ill-conditioned elements are solved in fp64, and the candidate is always the path
// cost at the computed foot point (see causal_cost).
#pragma once

#include <hip/hip_runtime.h>

namespace eik {

#define EIK_INF __builtin_huge_valf()

// Arithmetic helpers for the fp32 and fp64 variants of the solvers.
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
// fp32 square root and reciprocal are the hardware instructions (v_sqrt_f32, v_rcp_f32, about
// 1 ulp) rather than the correctly rounded expansions: every candidate is re-evaluated as the
// cost of an actual path (causal_cost), so an ulp in an intermediate only moves the foot point,
// and the fp64 variants keep IEEE operations.
__device__ __forceinline__ float sqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double sqrt_(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float rcp_(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double rcp_(double a) { return 1.0 / a; }

// b*b - a*c with the rounding error of a*c compensated (exact product via fma, Kahan).
template <typename R>
__device__ __forceinline__ R disc_comp(R a, R b, R c) {
  R p = a * c;
  R e = fma_(a, c, -p);
  return fma_(b, b, -p) - e;
}

// Work in adjugate form: with adj = det(G) K the quadratic (tau - d 1)^T K (tau - d 1) = 1
// reads A d^2 - 2 B d + C = 0 with A = 1^T adj 1, B = 1^T adj tau, C = tau^T adj tau - det
// (r = adj 1, w = adj tau).  Its larger root gives the foot point of the characteristic:
// unnormalised barycentric weights lam_i = d r_i - w_i, causal when all are >= 0, summing to
// S = d A - B = sqrt(disc).  The candidate is the Hopf-Lax cost at that foot point,
//     (sum_i lam_i tau_i + sqrt(S sum_i lam_i L_Di - sum_{i<j} lam_i lam_j L_ij)) / S,
// i.e. lambda.tau + |D - E| with |D - E|^2 from the squared edge lengths: the cost of an actual
// path through the element, so rounding in the root can only raise the candidate, never lower
// it below the exact minimum (the FIM keeps the smallest candidate it ever sees, so candidates
// rounded low would bias the result).  Returns the candidate minus t0, or +inf if not causal.
template <typename R, int N>
__device__ __forceinline__ R causal_cost(R A, R B, R C, const R* r, const R* w, const R* a,
                                         const R* LD, const R* LX) {
  // branch-free: lanes of a wave take different exits otherwise
  const R disc = disc_comp(A, B, C);
  bool ok = disc >= R(0) && A > R(0);
  const R d = (B + sqrt_(disc > R(0) ? disc : R(0))) * rcp_(A);
  R lam[N];
  R S = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lam[i] = d * r[i] - w[i];
    ok = ok && lam[i] >= R(0);
    S += lam[i];
  }
  ok = ok && S > R(0);
  R lt = R(0), x = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lt += lam[i] * a[i];
    x += lam[i] * LD[i];
  }
  x *= S;
  // pairwise terms: (0,1) for N = 2; (0,1), (0,2), (1,2) for N = 3 (LX in that order)
  x -= lam[0] * lam[1] * LX[0];
  if constexpr (N == 3) x -= lam[0] * lam[2] * LX[1] + lam[1] * lam[2] * LX[2];
  const R v = (lt + sqrt_(x > R(0) ? x : R(0))) * rcp_(S);
  return ok ? v : R(EIK_INF);
}

// Two known corners (face of a tet, or a triangle).  L1 = |X1 - D|^2, L2 = |X2 - D|^2,
// L12 = |X1 - X2|^2; a1, a2 = t_i - t0 >= 0.  Returns the candidate minus t0, or +inf.
template <typename R>
__device__ __forceinline__ R face_cost(R a1, R a2, R L1, R L2, R L12) {
  const R g12 = R(0.5) * (L1 + L2 - L12);
  const R det = fma_(L1, L2, -g12 * g12);
  // adj = [[L2, -g12], [-g12, L1]]
  R r[2] = {L2 - g12, L1 - g12};
  R w[2] = {L2 * a1 - g12 * a2, L1 * a2 - g12 * a1};
  R A = r[0] + r[1];
  R B = w[0] + w[1];
  R C = a1 * w[0] + a2 * w[1] - det;
  const R a[2] = {a1, a2}, LD[2] = {L1, L2}, LX[1] = {L12};
  const R v = causal_cost<R, 2>(A, B, C, r, w, a, LD, LX);
  return det > R(0) ? v : R(EIK_INF);
}

// Three known corners; L1..L3 = |X_i - D|^2, L12, L13, L23 = |X_i - X_j|^2.  ill receives
// det(G) < kCond G11 G22 G33 (the ratio is 1 for orthogonal edges, 0 for a flat element).
template <typename R>
__device__ __forceinline__ R interior_cost(R a1, R a2, R a3, R L1, R L2, R L3, R L12, R L13,
                                           R L23, R kCond, bool* ill) {
  const R g11 = L1, g22 = L2, g33 = L3;
  const R g12 = R(0.5) * (L1 + L2 - L12);
  const R g13 = R(0.5) * (L1 + L3 - L13);
  const R g23 = R(0.5) * (L2 + L3 - L23);
  const R c11 = fma_(g22, g33, -g23 * g23);
  const R c22 = fma_(g11, g33, -g13 * g13);
  const R c33 = fma_(g11, g22, -g12 * g12);
  const R c12 = fma_(g13, g23, -g12 * g33);
  const R c13 = fma_(g12, g23, -g13 * g22);
  const R c23 = fma_(g12, g13, -g11 * g23);
  const R det = g11 * c11 + g12 * c12 + g13 * c13;
  *ill = det < kCond * (g11 * g22 * g33);
  R r[3] = {c11 + c12 + c13, c12 + c22 + c23, c13 + c23 + c33};
  R w[3] = {c11 * a1 + c12 * a2 + c13 * a3, c12 * a1 + c22 * a2 + c23 * a3,
            c13 * a1 + c23 * a2 + c33 * a3};
  R A = r[0] + r[1] + r[2];
  R B = w[0] + w[1] + w[2];
  R C = a1 * w[0] + a2 * w[1] + a3 * w[2] - det;
  const R a[3] = {a1, a2, a3}, LD[3] = {L1, L2, L3}, LX[3] = {L12, L13, L23};
  const R v = causal_cost<R, 3>(A, B, C, r, w, a, LD, LX);
  return det > R(0) ? v : R(EIK_INF);
}

// Precision of the local solves.  The fp32 path runs on inputs scaled to unit trace (the
// quadratic is homogeneous: L -> L / s, tau -> tau / sqrt(s), d -> d / sqrt(s)), so the
// adjugate entries stay O(1) for any mesh scale.  Elements whose Gram matrix is
// ill-conditioned (det(G) / (G11 G22 G33) below kCondFp64: flat tets, needle triangles; the
// reference fixtures include tets at 1e-10) are solved again in fp64 from the same inputs, and
// the candidate rounded to fp32.
constexpr float kCondFp64 = 1e-3f;

__device__ __noinline__ float interior_cost_fp64(float a1, float a2, float a3, float L1, float L2,
                                                 float L3, float L12, float L13, float L23) {
  bool ill;
  double v = interior_cost<double>(a1, a2, a3, L1, L2, L3, L12, L13, L23, 0.0, &ill);
  return v < double(EIK_INF) ? (float)v : EIK_INF;
}

__device__ __noinline__ float face_cost_fp64(float a1, float a2, float L1, float L2, float L12) {
  double v = face_cost<double>(a1, a2, L1, L2, L12);
  return v < double(EIK_INF) ? (float)v : EIK_INF;
}

// fp32 face with the fp64 fallback; inputs already scaled (trace of the 2x2 Gram = 1 is not
// needed for the face: the 2x2 adjugate entries are products of two lengths).
__device__ __forceinline__ float face_cost_mixed(float a1, float a2, float L1, float L2, float L12) {
  const float g12 = 0.5f * (L1 + L2 - L12);
  const float det = fma_(L1, L2, -g12 * g12);
  if (det < kCondFp64 * L1 * L2) return face_cost_fp64(a1, a2, L1, L2, L12);
  return face_cost<float>(a1, a2, L1, L2, L12);
}

// Target D of a tet from corners 1..3.  L1..L3 = |X_i - D|^2, L12, L13, L23 = |X_i - X_j|^2.
// Unknown corners (t >= large) are left out, as the reference's candidates through them are
// never the minimum.  Returns +inf when no corner is known.
__device__ __forceinline__ float tet_corner_solve(float t1, float t2, float t3, float L1, float L2,
                                                  float L3, float L12, float L13, float L23,
                                                  float large) {
  const bool k1 = t1 < large, k2 = t2 < large, k3 = t3 < large;
  const float t0 = __builtin_fminf(t1, __builtin_fminf(t2, t3));
  // unit-trace scaling
  const float tr = L1 + L2 + L3;
  const float is = __builtin_amdgcn_rcpf(tr);
  const float sq = sqrt_(tr), isq = __builtin_amdgcn_rcpf(sq);
  const float n1 = L1 * is, n2 = L2 * is, n3 = L3 * is;
  const float n12 = L12 * is, n13 = L13 * is, n23 = L23 * is;
  const float b1 = (t1 - t0) * isq, b2 = (t2 - t0) * isq, b3 = (t3 - t0) * isq;
  if (k1 && k2 && k3) {
    bool ill;
    float v = interior_cost<float>(b1, b2, b3, n1, n2, n3, n12, n13, n23, kCondFp64, &ill);
    if (ill) v = interior_cost_fp64(b1, b2, b3, n1, n2, n3, n12, n13, n23);
    if (v < EIK_INF) return t0 + v * sq;  // a causal interior point is the minimum
  }
  float best = EIK_INF;
  if (k1 && k2) best = __builtin_fminf(best, face_cost_mixed(b1, b2, n1, n2, n12));
  if (k1 && k3) best = __builtin_fminf(best, face_cost_mixed(b1, b3, n1, n3, n13));
  if (k2 && k3) best = __builtin_fminf(best, face_cost_mixed(b2, b3, n2, n3, n23));
  if (k1) best = __builtin_fminf(best, b1 + sqrt_(n1));
  if (k2) best = __builtin_fminf(best, b2 + sqrt_(n2));
  if (k3) best = __builtin_fminf(best, b3 + sqrt_(n3));
  return best < EIK_INF ? t0 + best * sq : EIK_INF;
}

// Target Z of a triangle from X1, X2.  L1 = |X1 - Z|^2, L2 = |X2 - Z|^2, L12 = |X1 - X2|^2.
__device__ __forceinline__ float tri_corner_solve(float t1, float t2, float L1, float L2, float L12,
                                                  float large) {
  const bool k1 = t1 < large, k2 = t2 < large;
  const float t0 = k1 && k2 ? __builtin_fminf(t1, t2) : (k1 ? t1 : t2);
  const float tr = L1 + L2;
  const float is = __builtin_amdgcn_rcpf(tr);
  const float sq = sqrt_(tr), isq = __builtin_amdgcn_rcpf(sq);
  const float n1 = L1 * is, n2 = L2 * is, n12 = L12 * is;
  const float b1 = (t1 - t0) * isq, b2 = (t2 - t0) * isq;
  float best = EIK_INF;
  if (k1 && k2) {
    best = face_cost_mixed(b1, b2, n1, n2, n12);
    if (best < EIK_INF) return t0 + best * sq;
  }
  if (k1) best = __builtin_fminf(best, b1 + sqrt_(n1));
  if (k2) best = __builtin_fminf(best, b2 + sqrt_(n2));
  return best < EIK_INF ? t0 + best * sq : EIK_INF;
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
