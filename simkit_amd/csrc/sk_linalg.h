// Small dense linear algebra that lives entirely in registers (one matrix per
// lane): cyclic Jacobi eigen-decomposition of packed symmetric NxN matrices,
// eigenvalue-floor PSD projection, and a signed 3x3 SVD.
#pragma once

#include <hip/hip_runtime.h>

namespace sk {

// Packed upper-triangle storage of a symmetric NxN matrix, row by row.
template <int N>
struct Sym {
  static constexpr int size = N * (N + 1) / 2;
  static constexpr __host__ __device__ int idx(int i, int j) {
    return i <= j ? i * N - i * (i - 1) / 2 + (j - i) : j * N - j * (j - 1) / 2 + (i - j);
  }
};

// Cyclic-by-row Jacobi on a packed symmetric matrix. On return a[] holds the
// eigenvalues on its diagonal (off-diagonals ~0) and v (row-major, v[r*N+k])
// the eigenvectors as columns. Each rotation zeroes a_pq with
// t = sgn(d) 2 a_pq / (|d| + sqrt(d^2 + 4 a_pq^2)), d = a_qq - a_pp (the
// smaller-angle root), c = 1/sqrt(1+t^2), s = t c. Sweeps stop once the
// off-diagonal mass is below 1e-32 of the diagonal mass (or max_sweeps).
template <int N, typename R>
__device__ __forceinline__ int jacobi_eigh(R (&a)[Sym<N>::size], R (&v)[N * N], int max_sweeps) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) v[i * N + j] = (i == j) ? R(1) : R(0);

  // Off-diagonal mass relative to the diagonal mass at which sweeps stop:
  // (1e-16)^2 in FP64; in FP32 roundoff keeps off-diagonals near 1e-7 relative,
  // so the FP32 threshold is (3e-7)^2.
  const R tol = sizeof(R) == 8 ? R(1e-32) : R(1e-13);
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    R off = 0, dia = 0;
#pragma unroll
    for (int p = 0; p < N; ++p) {
      dia += a[Sym<N>::idx(p, p)] * a[Sym<N>::idx(p, p)];
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += a[Sym<N>::idx(p, q)] * a[Sym<N>::idx(p, q)];
    }
    if (!(off > tol * dia)) break;  // also exits on NaN
#pragma unroll
    for (int p = 0; p < N - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const R apq = a[Sym<N>::idx(p, q)];
        const R app = a[Sym<N>::idx(p, p)];
        const R aqq = a[Sym<N>::idx(q, q)];
        const R d = aqq - app;
        const R rt = sqrt(d * d + R(4) * apq * apq);
        R t = R(2) * apq / (fabs(d) + rt);
        t = (d < R(0)) ? -t : t;
        // No rotation when a_pq is zero or so small that d^2 + 4 a_pq^2
        // underflows (rt == 0): a_pq is then negligible, and rotating would
        // divide by zero.
        t = (apq == R(0) || !(rt > R(0))) ? R(0) : t;
        const R c = R(1) / sqrt(R(1) + t * t);
        const R s = t * c;
        const R h = t * apq;
        a[Sym<N>::idx(p, p)] = app - h;
        a[Sym<N>::idx(q, q)] = aqq + h;
        a[Sym<N>::idx(p, q)] = R(0);
#pragma unroll
        for (int r = 0; r < N; ++r) {
          if (r == p || r == q) continue;
          const R g = a[Sym<N>::idx(r, p)];
          const R hh = a[Sym<N>::idx(r, q)];
          a[Sym<N>::idx(r, p)] = c * g - s * hh;
          a[Sym<N>::idx(r, q)] = s * g + c * hh;
        }
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const R g = v[r * N + p];
          const R hh = v[r * N + q];
          v[r * N + p] = c * g - s * hh;
          v[r * N + q] = s * g + c * hh;
        }
      }
    }
  }
  return sweep;
}

// Eigenvalue modification of a packed symmetric matrix, in place:
// method 1 ('proj') floors eigenvalues below `floor_` at `floor_`, method 2
// ('abs') takes absolute values; the result is rebuilt as V diag(s) V^T.
// psd_project.py:30-45. Returns the number of modified eigenvalues.
template <int N, typename R>
__device__ __forceinline__ int psd_project_sym(R (&a)[Sym<N>::size], int method, R floor_, int max_sweeps = 20) {
  R v[N * N];
  jacobi_eigh<N, R>(a, v, max_sweeps);
  R s[N];
  int changed = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const R e = a[Sym<N>::idx(k, k)];
    R m = e;
    if (method == 2) m = fabs(e);
    else if (method == 1) m = (e < floor_) ? floor_ : e;
    changed += (m != e);
    s[k] = m;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i; j < N; ++j) {
      R acc = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += v[i * N + k] * s[k] * v[j * N + k];
      a[Sym<N>::idx(i, j)] = acc;
    }
  return changed;
}

// Signed SVD of a 3x3 matrix F (row-major): F = U diag(s) V^T with U and V
// proper rotations, s0 >= s1 >= |s2| and sign(s2) = sign(det F). This is the
// factorisation the reference's svd_rv produces (svd_rv.py:42-51: the
// smallest singular value takes the sign of det(U V^T)). Method of McAdams et
// al. 2011, "Computing the SVD of 3x3 matrices with minimal branching":
// Jacobi eigenvectors V of F^T F, sorted by decreasing eigenvalue and made a
// rotation, then a Givens QR of F V whose diagonal is the signed singular
// values and whose orthogonal factor is U.
template <typename R>
__device__ __forceinline__ void svd3_signed(const R (&F)[9], R (&U)[9], R (&s)[3], R (&V)[9]) {
  R A[6];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = j; k < 3; ++k)
      A[Sym<3>::idx(j, k)] = F[0 * 3 + j] * F[0 * 3 + k] + F[1 * 3 + j] * F[1 * 3 + k] + F[2 * 3 + j] * F[2 * 3 + k];
  jacobi_eigh<3, R>(A, V, 20);
  R e0 = A[Sym<3>::idx(0, 0)], e1 = A[Sym<3>::idx(1, 1)], e2 = A[Sym<3>::idx(2, 2)];
  auto swapcol = [&](int a, int b) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R t = V[r * 3 + a];
      V[r * 3 + a] = V[r * 3 + b];
      V[r * 3 + b] = t;
    }
  };
  if (e0 < e1) { R t = e0; e0 = e1; e1 = t; swapcol(0, 1); }
  if (e0 < e2) { R t = e0; e0 = e2; e2 = t; swapcol(0, 2); }
  if (e1 < e2) { R t = e1; e1 = e2; e2 = t; swapcol(1, 2); }
  const R detV = V[0] * (V[4] * V[8] - V[5] * V[7]) - V[1] * (V[3] * V[8] - V[5] * V[6]) +
                 V[2] * (V[3] * V[7] - V[4] * V[6]);
  if (detV < R(0)) {
    V[2] = -V[2];
    V[5] = -V[5];
    V[8] = -V[8];
  }
  R B[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) B[i * 3 + k] = F[i * 3 + 0] * V[0 * 3 + k] + F[i * 3 + 1] * V[1 * 3 + k] + F[i * 3 + 2] * V[2 * 3 + k];
#pragma unroll
  for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? R(1) : R(0);
  // Givens rotation on rows (r0, r1) of B that zeroes B[r1][col]; U <- U G^T.
  auto givens = [&](int r0, int r1, int col) {
    const R x = B[r0 * 3 + col], y = B[r1 * 3 + col];
    const R rr = sqrt(x * x + y * y);
    const R c = rr > R(0) ? x / rr : R(1);
    const R sn = rr > R(0) ? y / rr : R(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const R b0 = B[r0 * 3 + k], b1 = B[r1 * 3 + k];
      B[r0 * 3 + k] = c * b0 + sn * b1;
      B[r1 * 3 + k] = -sn * b0 + c * b1;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const R u0 = U[r * 3 + r0], u1 = U[r * 3 + r1];
      U[r * 3 + r0] = c * u0 + sn * u1;
      U[r * 3 + r1] = -sn * u0 + c * u1;
    }
  };
  givens(0, 1, 0);
  givens(0, 2, 0);
  givens(1, 2, 1);
  s[0] = B[0];
  s[1] = B[4];
  s[2] = B[8];
}

}  // namespace sk
