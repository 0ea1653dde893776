/* Complex128 in-place statevector restatement for full-size golden runs.
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/libsv128.so by oracle/Makefile
 * (gcc, OpenMP) and called by tests/golden/make_fullsize_goldens.py and the
 * CPU tests that pin it (tests/test_oracle_c128.py).  The product library never
 * links or loads it.
 *
 * It restates, gate by gate and in complex128, the per-gate contractions of
 * the reference (src/tyxonq/libs/quantum_library/kernels/statevector.py):
 *   sv_apply_1q  <- apply_1q_statevector  statevector.py:28-42
 *                   (out[..a..] = sum_b U[a, b] psi[..b..] on qubit q's axis,
 *                    qubit q = index bit n-1-q, big-endian)
 *   sv_apply_2q  <- apply_2q_statevector  statevector.py:45-59
 *                   (U[(a0 a1), (b0 b1)] with q0 the most significant index;
 *                    q0 == q1 leaves the state unchanged, :46-47)
 *   sv_norm2, sv_pauli_dot   <- the FP64 sums of expect_z_statevector
 *                   (statevector.py:62-68) and of the Pauli-sum expectation
 *                   Re <psi| P |psi> (pauli.py:74-87, dynamics.py:117-126)
 *   sv_pauli_rot <- exp(-i a/2 P) of the Pauli-rotation gate sequences
 *                   (libs/circuits_library/utils.py:29-59)
 * The reference contracts with einsum and a fresh array per gate, which does
 * not fit a 30-qubit complex128 state twice in host memory; this file applies
 * the same matrices in place, in the same order.
 *
 * Pauli strings are (x, z) masks over INDEX bits: P = i^{|x&z|} X^x Z^z, so
 * (P psi)[i] = i^{|x&z|} (-1)^{|(i^x) & z|} psi[i ^ x]  (Y = i X Z).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <omp.h>

typedef double complex cplx;

static inline int parity64(uint64_t v) { return __builtin_popcountll(v) & 1; }

static cplx ipow(int k) {
  switch (k & 3) {
    case 0: return 1.0;
    case 1: return I;
    case 2: return -1.0;
    default: return -I;
  }
}

void sv_set_threads(int t) { if (t > 0) omp_set_num_threads(t); }

void sv_basis(cplx* psi, int n, uint64_t idx) {
  const int64_t dim = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) psi[i] = 0.0;
  psi[idx] = 1.0;
}

/* U: 8 doubles, row-major 2x2, (re, im) pairs. */
void sv_apply_1q(cplx* psi, int n, int q, const double* U) {
  const int bit = n - 1 - q;
  const int64_t stride = (int64_t)1 << bit;
  const int64_t half = (int64_t)1 << (n - 1);
  const cplx u00 = U[0] + I * U[1], u01 = U[2] + I * U[3];
  const cplx u10 = U[4] + I * U[5], u11 = U[6] + I * U[7];
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < half; ++k) {
    const int64_t lo = k & (stride - 1);
    const int64_t i0 = ((k - lo) << 1) | lo;
    const int64_t i1 = i0 | stride;
    const cplx a = psi[i0], b = psi[i1];
    psi[i0] = u00 * a + u01 * b;
    psi[i1] = u10 * a + u11 * b;
  }
}

/* U: 32 doubles, row-major 4x4 in the basis |q0 q1>, (re, im) pairs. */
void sv_apply_2q(cplx* psi, int n, int q0, int q1, const double* U) {
  if (q0 == q1) return;
  const int b0 = n - 1 - q0, b1 = n - 1 - q1;
  const int blo = b0 < b1 ? b0 : b1, bhi = b0 < b1 ? b1 : b0;
  const int64_t m0 = (int64_t)1 << b0, m1 = (int64_t)1 << b1;
  const int64_t quarter = (int64_t)1 << (n - 2);
  cplx u[16];
  for (int r = 0; r < 16; ++r) u[r] = U[2 * r] + I * U[2 * r + 1];
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < quarter; ++k) {
    /* insert zeros at blo, then at bhi */
    int64_t v = k;
    int64_t lo = v & (((int64_t)1 << blo) - 1);
    v = ((v - lo) << 1) | lo;
    lo = v & (((int64_t)1 << bhi) - 1);
    v = ((v - lo) << 1) | lo;
    int64_t idx[4];
    idx[0] = v;            /* q0 = 0, q1 = 0 */
    idx[1] = v | m1;       /* q0 = 0, q1 = 1 */
    idx[2] = v | m0;       /* q0 = 1, q1 = 0 */
    idx[3] = v | m0 | m1;  /* q0 = 1, q1 = 1 */
    cplx a[4];
    for (int c = 0; c < 4; ++c) a[c] = psi[idx[c]];
    for (int r = 0; r < 4; ++r) {
      cplx s = 0.0;
      for (int c = 0; c < 4; ++c) s += u[4 * r + c] * a[c];
      psi[idx[r]] = s;
    }
  }
}

/* psi <- exp(-i a/2 P) psi = cos(a/2) psi - i sin(a/2) P psi. */
void sv_pauli_rot(cplx* psi, int n, uint64_t x, uint64_t z, double a) {
  const int64_t dim = (int64_t)1 << n;
  const double c = cos(0.5 * a), s = sin(0.5 * a);
  const cplx ph = ipow(__builtin_popcountll(x & z));
  if (x == 0) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; ++i) {
      const double sg = parity64((uint64_t)i & z) ? -1.0 : 1.0;
      psi[i] = (c - I * s * sg) * psi[i];
    }
    return;
  }
  const int top = 63 - __builtin_clzll(x); /* i and i^x are paired; visit i with bit top = 0 */
  const int64_t stride = (int64_t)1 << top;
  const int64_t half = dim >> 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < half; ++k) {
    const int64_t lo = k & (stride - 1);
    const int64_t i = ((k - lo) << 1) | lo;
    const int64_t j = i ^ (int64_t)x;
    const cplx pi = psi[i], pj = psi[j];
    /* (P psi)[i] = ph (-1)^{|j&z|} psi[j], (P psi)[j] = ph (-1)^{|i&z|} psi[i] */
    const cplx Ppi = ph * (parity64((uint64_t)j & z) ? -1.0 : 1.0) * pj;
    const cplx Ppj = ph * (parity64((uint64_t)i & z) ? -1.0 : 1.0) * pi;
    psi[i] = c * pi - I * s * Ppi;
    psi[j] = c * pj - I * s * Ppj;
  }
}

/* out += coeff * P psi (out must not alias psi). */
void sv_pauli_apply_acc(const cplx* psi, cplx* out, int n, uint64_t x, uint64_t z, double coeff) {
  const int64_t dim = (int64_t)1 << n;
  const cplx ph = coeff * ipow(__builtin_popcountll(x & z));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) {
    const int64_t j = i ^ (int64_t)x;
    const double sg = parity64((uint64_t)j & z) ? -1.0 : 1.0;
    out[i] += ph * sg * psi[j];
  }
}

/* <u| P |v> into (re, im). */
void sv_pauli_dot(const cplx* u, const cplx* v, int n, uint64_t x, uint64_t z, double* re,
                  double* im) {
  const int64_t dim = (int64_t)1 << n;
  double sr = 0.0, si = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sr, si)
  for (int64_t i = 0; i < dim; ++i) {
    const int64_t j = i ^ (int64_t)x;
    const double sg = parity64((uint64_t)j & z) ? -1.0 : 1.0;
    const cplx t = conj(u[i]) * v[j] * sg;
    sr += creal(t);
    si += cimag(t);
  }
  const cplx r = (sr + I * si) * ipow(__builtin_popcountll(x & z));
  *re = creal(r);
  *im = cimag(r);
}

double sv_norm2(const cplx* psi, int n) {
  const int64_t dim = (int64_t)1 << n;
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < dim; ++i) s += creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
  return s;
}

void sv_copy(const cplx* src, cplx* dst, int n) {
  const int64_t dim = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) dst[i] = src[i];
}

void sv_zero(cplx* psi, int n) {
  const int64_t dim = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dim; ++i) psi[i] = 0.0;
}
