// Per-element hyperelastic densities on a row-major 3x3 deformation gradient
// F[3*i + c] (i spatial, c material): energy psi, first Piola-Kirchhoff
// stress P (same layout) and the 9x9 Hessian d2psi/dvecF^2 as a packed
// symmetric matrix over row-major vec(F).
#pragma once

#include "sk_linalg.h"

namespace sk {


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

// H = a I + b vecF vecF^T + c vec(cof) vec(cof)^T + d dvec(cof)/dvecF, the
// shape shared by every density built from I_C = ||F||^2 and J = det F, with
// d cof_ij / d F_kl = eps_ikn eps_jlq F_nq.
template <typename R>
__device__ __forceinline__ void cofactor_form_hessian(const R (&F)[9], const R (&cof)[9], R a, R b, R c, R d,
                                                      R (&H)[45]) {
#pragma unroll
  for (int p = 0; p < 9; ++p) {
#pragma unroll
    for (int q = p; q < 9; ++q) {
      R h = b * F[p] * F[q] + c * cof[p] * cof[q];
      if (p == q) h += a;
      const int i = p / 3, j = p % 3, k = q / 3, l = q % 3;
      if (i != k && j != l) {
        const int n = 3 - i - k, m = 3 - j - l;
        const int sg = levi(i, k) * levi(j, l);
        h += (sg > 0 ? d : -d) * F[3 * n + m];
      }
      H[Sym<9>::idx(p, q)] = h;
    }
  }
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
  static constexpr bool kPolar = false;  // no polar decomposition in the density
  static constexpr bool kAnalyticEig = true;  // eigenvalue modification by hessian_psd_analytic
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
    cofactor_form_hessian(F, s.cof, mu * s.IC / den, R(2) * mu / (den * den), lam, lam * (s.J - s.alpha), H);
  }
  // vol H with its eigenvalues e modified as psd_project does (method 1: max(e, floor); 2: |e|), from
  // the Hessian's closed-form eigensystem in the signed SVD F = U diag(s) V^T (U, V rotations):
  // with H = a I + b vecF vecF^T + c vec(cof) vec(cof)^T + d dvec(cof)/dvecF,
  //   twist  vec(U (e_i e_j^T - e_j e_i^T) V^T)/sqrt2: a + d s_k,
  //   flip   vec(U (e_i e_j^T + e_j e_i^T) V^T)/sqrt2: a - d s_k   ({i, j, k} = {0, 1, 2}),
  //   scaling vec(U diag(w) V^T), w an eigenvector of the 3x3 block
  //   a I + b s s^T + c t t^T + d S, t = (s1 s2, s0 s2, s0 s1), S_ij = s_k (i != j), S_ii = 0.
  // H itself is formed as in hessian() and only the modified modes are added, (m(e) - e) v v^T.
  // Returns false (H not written) when the smallest singular value is below 1e-2 of the largest:
  // the SVD (Jacobi on F^T F) then resolves U and V too coarsely for FP64 parity, and the caller
  // projects by the 9x9 Jacobi instead.
  static __device__ __forceinline__ bool hessian_psd_analytic(const R (&F)[9], R mu, R lam, R vol, int method,
                                                              R floor_, R (&H)[45]) {
    R U[9], sv[3], V[9];
    svd3_signed(F, U, sv, V);
    const R amax = fmax(fabs(sv[0]), fmax(fabs(sv[1]), fabs(sv[2])));
    const R amin = fmin(fabs(sv[0]), fmin(fabs(sv[1]), fabs(sv[2])));
    if (!(amin >= R(1e-2) * amax)) return false;  // also for F = 0 and NaN
    const Scal s = scalars(F, mu, lam);
    const R den = s.IC + R(1);
    const R a = mu * s.IC / den, b = R(2) * mu / (den * den), c = lam, d = lam * (s.J - s.alpha);
    cofactor_form_hessian(F, s.cof, a, b, c, d, H);
#pragma unroll
    for (int k = 0; k < 45; ++k) H[k] *= vol;
    auto modified = [&](R e) { return method == 2 ? fabs(e) : (e < floor_ ? floor_ : e); };
    auto add_mode = [&](R cm, const R (&v)[9]) {
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int q = p; q < 9; ++q) H[Sym<9>::idx(p, q)] += cm * v[p] * v[q];
    };
    const R inv_sqrt2 = R(0.70710678118654752440);
    const int pi[3] = {0, 1, 0}, pj[3] = {1, 2, 2}, pk[3] = {2, 0, 1};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int i = pi[m], j = pj[m], k = pk[m];
#pragma unroll
      for (int sgn = 0; sgn < 2; ++sgn) {  // 0: twist, 1: flip
        const R e = vol * (sgn == 0 ? a + d * sv[k] : a - d * sv[k]);
        const R cm = modified(e) - e;
        if (cm != R(0)) {
          R v[9];
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
              const R t1 = U[r * 3 + i] * V[cc * 3 + j], t2 = U[r * 3 + j] * V[cc * 3 + i];
              v[r * 3 + cc] = (sgn == 0 ? t1 - t2 : t1 + t2) * inv_sqrt2;
            }
          add_mode(cm, v);
        }
      }
    }
    const R t3[3] = {sv[1] * sv[2], sv[0] * sv[2], sv[0] * sv[1]};
    R A3[6], W[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j)
        A3[Sym<3>::idx(i, j)] =
            vol * ((i == j ? a : d * sv[3 - i - j]) + b * sv[i] * sv[j] + c * t3[i] * t3[j]);
    jacobi_eigh<3, R>(A3, W, 20);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const R e = A3[Sym<3>::idx(m, m)];
      const R cm = modified(e) - e;
      if (cm != R(0)) {
        R v[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc)
            v[r * 3 + cc] = W[0 * 3 + m] * U[r * 3 + 0] * V[cc * 3 + 0] + W[1 * 3 + m] * U[r * 3 + 1] * V[cc * 3 + 1] +
                            W[2 * 3 + m] * U[r * 3 + 2] * V[cc * 3 + 2];
        add_mode(cm, v);
      }
    }
    return true;
  }
};

// ARAP: psi = mu/2 ||F - R||^2, P = mu (F - R), H = mu (I - dR/dF) with R the
// rotation of the signed SVD and dR/dF = sum_k 2/max(s_i + s_j, 1e-8) t_k t_k^T,
// t_k = vec(U T_k V^T)/sqrt(2) for the twists coupling (0,1), (1,2), (0,2)
// (arap.py:86-145, polar_svd.py:209-218, rotation_gradient.py:42-72).
template <typename R>
struct ARAP {
  static constexpr bool kAnalyticEig = false;
  static constexpr bool kPolar = true;
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

// Linear (small-strain) elasticity: psi = mu ||eps||^2 + lam/2 tr(eps)^2,
// eps = (F + F^T)/2 - I; P = mu (F + F^T - 2 I) + lam tr(F - I) I;
// H = mu (I + T) + lam vec(I) vec(I)^T with T the vec-transpose permutation
// (linear_elasticity.py: energy, gradient and hessian element tiers).
template <typename R>
struct Linear {
  static constexpr bool kAnalyticEig = false;
  static constexpr bool kPolar = false;
  static __device__ __forceinline__ R energy(const R (&F)[9], R mu, R lam) {
    R acc = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const R e = R(0.5) * (F[3 * i + j] + F[3 * j + i]) - (i == j ? R(1) : R(0));
        acc += e * e;
      }
    const R tr = F[0] + F[4] + F[8] - R(3);
    return mu * acc + R(0.5) * lam * tr * tr;
  }
  static __device__ __forceinline__ void gradient(const R (&F)[9], R mu, R lam, R (&P)[9]) {
    const R tr = F[0] + F[4] + F[8] - R(3);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        P[3 * i + j] = mu * (F[3 * i + j] + F[3 * j + i] - (i == j ? R(2) : R(0))) + (i == j ? lam * tr : R(0));
  }
  static __device__ __forceinline__ void hessian(const R (&/*F*/)[9], R mu, R lam, R (&H)[45]) {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int q = p; q < 9; ++q) {
        const int i = p / 3, j = p % 3, k = q / 3, l = q % 3;
        R h = R(0);
        if (i == k && j == l) h += mu;
        if (i == l && j == k) h += mu;
        if (i == j && k == l) h += lam;
        H[Sym<9>::idx(p, q)] = h;
      }
  }
};

// Macklin-Mueller neo-Hookean (macklin_mueller_neo_hookean.py, 3D element tier):
//   psi = mu (1 - J) + lam/2 (1 - J)^2 + mu/2 (I_C - 3)
//   P = mu F + (lam (J - 1) - mu) cof F
//   H = mu I + lam vec(cof) vec(cof)^T + (lam (J - 1) - mu) dvec(cof)/dvecF
template <typename R>
struct MMNH {
  static constexpr bool kAnalyticEig = false;
  static constexpr bool kPolar = false;
  static __device__ __forceinline__ R energy(const R (&F)[9], R mu, R lam) {
    R cof[9];
    cofactor3(F, cof);
    R ic = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) ic += F[k] * F[k];
    const R c = R(1) - (F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2]);
    return mu * c + R(0.5) * lam * c * c + R(0.5) * mu * (ic - R(3));
  }
  static __device__ __forceinline__ void gradient(const R (&F)[9], R mu, R lam, R (&P)[9]) {
    R cof[9];
    cofactor3(F, cof);
    const R J = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2];
    const R d = lam * (J - R(1)) - mu;
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = mu * F[k] + d * cof[k];
  }
  static __device__ __forceinline__ void hessian(const R (&F)[9], R mu, R lam, R (&H)[45]) {
    R cof[9];
    cofactor3(F, cof);
    const R J = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2];
    cofactor_form_hessian(F, cof, mu, R(0), lam, lam * (J - R(1)) - mu, H);
  }
};

// Fixed corotational (fcr.py, 3D element tier): twice the ARAP density plus
// a volume term,
//   psi = mu ||F - R||^2 + lam/2 (J - 1)^2
//   P = 2 mu (F - R) + lam (J - 1) cof F
//   H = 2 mu (I - dR/dF) + lam vec(cof) vec(cof)^T + lam (J - 1) dvec(cof)/dvecF
template <typename R>
struct FCR {
  static constexpr bool kAnalyticEig = false;
  static constexpr bool kPolar = true;
  static __device__ __forceinline__ R energy(const R (&F)[9], R mu, R lam) {
    R cof[9];
    cofactor3(F, cof);
    const R dj = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2] - R(1);
    return ARAP<R>::energy(F, R(2) * mu, lam) + R(0.5) * lam * dj * dj;
  }
  static __device__ __forceinline__ void gradient(const R (&F)[9], R mu, R lam, R (&P)[9]) {
    R cof[9];
    cofactor3(F, cof);
    const R d = lam * (F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2] - R(1));
    ARAP<R>::gradient(F, R(2) * mu, lam, P);
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] += d * cof[k];
  }
  static __device__ __forceinline__ void hessian(const R (&F)[9], R mu, R lam, R (&H)[45]) {
    R cof[9], Hv[45];
    cofactor3(F, cof);
    const R J = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2];
    cofactor_form_hessian(F, cof, R(0), R(0), lam, lam * (J - R(1)), Hv);
    ARAP<R>::hessian(F, R(2) * mu, lam, H);
#pragma unroll
    for (int k = 0; k < 45; ++k) H[k] += Hv[k];
  }
};

// Calls fn(tag) with tag::type<R> naming the material's density template.
template <template <typename> class M>
struct MaterialTag {
  template <typename R>
  using type = M<R>;
};

template <typename Fn>
inline bool dispatch_material(int material, Fn&& fn) {
  switch (material) {
    case 0: fn(MaterialTag<SNH>{}); return true;
    case 1: fn(MaterialTag<ARAP>{}); return true;
    case 2: fn(MaterialTag<Linear>{}); return true;
    case 3: fn(MaterialTag<MMNH>{}); return true;
    case 4: fn(MaterialTag<FCR>{}); return true;
    default: return false;
  }
}

}  // namespace sk
