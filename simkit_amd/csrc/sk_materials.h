// Per-element hyperelastic densities on a row-major 3x3 deformation gradient
// F[3*i + c] (i spatial, c material): energy psi, first Piola-Kirchhoff
// stress P (same layout) and the 9x9 Hessian d2psi/dvecF^2 as a packed
// symmetric matrix over row-major vec(F).
#pragma once

#include "sk_linalg.h"

namespace sk {

enum Material : int { kSNH = 0, kARAP = 1 };

// Levi-Civita sign of (i, k, n) for distinct indices in {0,1,2}.
constexpr __host__ __device__ int levi(int i, int k) { return ((k - i + 3) % 3 == 1) ? 1 : -1; }

template <typename R>
__device__ __forceinline__ void cofactor3(const R (&F)[9], R (&c)[9]) {
  c[0] = F[4] * F[8] - F[5] * F[7];
  c[1] = F[5] * F[6] - F[3] * F[8];
  c[2] = F[3] * F[7] - F[4] * F[6];
  c[3] = F[2] * F[7] - F[1] * F[8];
  c[4] = F[0] * F[8] - F[2] * F[6];
  c[5] = F[1] * F[6] - F[0] * F[7];
  c[6] = F[1] * F[5] - F[2] * F[4];
  c[7] = F[2] * F[3] - F[0] * F[5];
  c[8] = F[0] * F[4] - F[1] * F[3];
}

// Stable neo-Hookean (Smith, de Goes, Kim 2018), d = 3:
//   psi = mu/2 (I_C - 3) - mu/2 log(I_C + 1) + lam/2 (J - alpha)^2,
//   alpha = 1 + 3 mu / (4 lam)                 (stable_neo_hookean.py:84-126)
//   P = mu I_C/(I_C+1) F + lam (J - alpha) cof F       (:192-215)
//   H = mu I_C/(I_C+1) I + 2 mu/(I_C+1)^2 vecF vecF^T + lam vec(cof) vec(cof)^T
//       + lam (J - alpha) d vec(cof)/d vecF            (:315-440)
// with d cof_ij / d F_kl = eps_ikn eps_jlq F_nq.
template <typename R>
struct SNH {
  struct Scal {
    R IC, J, alpha, cof[9];
  };
  static __device__ __forceinline__ Scal scalars(const R (&F)[9], R mu, R lam) {
    Scal s;
    cofactor3(F, s.cof);
    R ic = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) ic += F[k] * F[k];
    s.IC = ic;
    s.J = F[0] * s.cof[0] + F[1] * s.cof[1] + F[2] * s.cof[2];
    s.alpha = R(1) + (R(3) * mu) / (R(4) * lam);
    return s;
  }
  static __device__ __forceinline__ R energy(const R (&F)[9], R mu, R lam) {
    const Scal s = scalars(F, mu, lam);
    const R dj = s.J - s.alpha;
    return R(0.5) * mu * (s.IC - R(3)) - R(0.5) * mu * log(s.IC + R(1)) + R(0.5) * lam * dj * dj;
  }
  static __device__ __forceinline__ void gradient(const R (&F)[9], R mu, R lam, R (&P)[9]) {
    const Scal s = scalars(F, mu, lam);
    const R A = mu * s.IC / (s.IC + R(1));
    const R D = lam * (s.J - s.alpha);
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = A * F[k] + D * s.cof[k];
  }
  static __device__ __forceinline__ void hessian(const R (&F)[9], R mu, R lam, R (&H)[45]) {
    const Scal s = scalars(F, mu, lam);
    const R den = s.IC + R(1);
    const R A = mu * s.IC / den;
    const R B = R(2) * mu / (den * den);
    const R D = lam * (s.J - s.alpha);
#pragma unroll
    for (int a = 0; a < 9; ++a) {
#pragma unroll
      for (int b = a; b < 9; ++b) {
        R h = B * F[a] * F[b] + lam * s.cof[a] * s.cof[b];
        if (a == b) h += A;
        const int i = a / 3, j = a % 3, k = b / 3, l = b % 3;
        if (i != k && j != l) {
          const int n = 3 - i - k, q = 3 - j - l;
          const int sg = levi(i, k) * levi(j, l);
          h += (sg > 0 ? D : -D) * F[3 * n + q];
        }
        H[Sym<9>::idx(a, b)] = h;
      }
    }
  }
};

// ARAP: psi = mu/2 ||F - R||^2, P = mu (F - R), H = mu (I - dR/dF) with R the
// rotation of the signed SVD and dR/dF = sum_k 2/max(s_i + s_j, 1e-8) t_k t_k^T,
// t_k = vec(U T_k V^T)/sqrt(2) for the twists coupling (0,1), (1,2), (0,2)
// (arap.py:86-145, polar_svd.py:209-218, rotation_gradient.py:42-72).
template <typename R>
struct ARAP {
  static __device__ __forceinline__ void rotation(const R (&U)[9], const R (&V)[9], R (&Rm)[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Rm[i * 3 + j] = U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] + U[i * 3 + 2] * V[j * 3 + 2];
  }
  static __device__ __forceinline__ R energy(const R (&F)[9], R mu, R /*lam*/) {
    R U[9], s[3], V[9], Rm[9];
    svd3_signed(F, U, s, V);
    rotation(U, V, Rm);
    R acc = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc += (F[k] - Rm[k]) * (F[k] - Rm[k]);
    return R(0.5) * mu * acc;
  }
  static __device__ __forceinline__ void gradient(const R (&F)[9], R mu, R /*lam*/, R (&P)[9]) {
    R U[9], s[3], V[9], Rm[9];
    svd3_signed(F, U, s, V);
    rotation(U, V, Rm);
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = mu * (F[k] - Rm[k]);
  }
  // Twist directions t_k (unit 9-vectors) and their dR/dF coefficients.
  static __device__ __forceinline__ void twists(const R (&F)[9], R (&t)[3][9], R (&coef)[3]) {
    R U[9], s[3], V[9];
    svd3_signed(F, U, s, V);
    const int pa[3] = {0, 1, 0}, pb[3] = {1, 2, 2};
    const R inv_sqrt2 = R(0.70710678118654752440);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int a = pa[k], b = pb[k];
      // U T_k V^T = u_b v_a^T - u_a v_b^T (sign irrelevant in t t^T).
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          t[k][i * 3 + j] = (U[i * 3 + b] * V[j * 3 + a] - U[i * 3 + a] * V[j * 3 + b]) * inv_sqrt2;
      const R ss = s[a] + s[b];
      coef[k] = R(2) / (ss > R(1e-8) ? ss : R(1e-8));
    }
  }
  static __device__ __forceinline__ void hessian(const R (&F)[9], R mu, R /*lam*/, R (&H)[45]) {
    R t[3][9], coef[3];
    twists(F, t, coef);
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
      for (int b = a; b < 9; ++b) {
        R k = coef[0] * t[0][a] * t[0][b] + coef[1] * t[1][a] * t[1][b] + coef[2] * t[2][a] * t[2][b];
        H[Sym<9>::idx(a, b)] = mu * ((a == b ? R(1) : R(0)) - k);
      }
  }
};

}  // namespace sk
