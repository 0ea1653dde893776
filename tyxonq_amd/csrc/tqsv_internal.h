// tqsv_internal.h - declarations shared by the libtqsv translation units.
#ifndef TQSV_INTERNAL_H
#define TQSV_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace tqsv {

// error state (tqsv_abi.cpp)
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define TQSV_HIP(call)                                  \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return ::tqsv::hip_fail(e_, #call); \
  } while (0)

// ---- tile pass kernel (tqsv_pass.hip) ----
struct PassLaunch {
  float2* psi;
  float2* lam;
  const uint32_t* prog;   // device pointer to this pass's record
  double* partials;       // grid * nred doubles (reduction passes)
  uint32_t ntiles;
  uint32_t grid;
  size_t lds_bytes;
};
// mode: MODE_GATE / MODE_RED / MODE_DUAL (tqsv_prog.h)
// lean: the pass has no OP_U2 / OP_PROT steps (smaller kernel variant)
hipError_t launch_pass(int tb, int mode, bool idx64, bool lean, const PassLaunch& pl,
                       hipStream_t st);
size_t pass_lds_bytes(int tb, int mode, uint32_t nred);
bool pass_variant_exists(int tb, int mode, bool idx64);

// ---- single-op kernels (tqsv_kernels.hip) ----
struct Mat2f { float v[8]; };
struct Mat4f { float v[32]; };

hipError_t k_init_basis(float2* psi, int n, uint64_t index, hipStream_t st);
hipError_t k_apply_1q(float2* psi, int n, int bit, uint64_t cmask, uint64_t cval, const Mat2f& u,
                      hipStream_t st);
hipError_t k_apply_2q(float2* psi, int n, int bh, int bl, uint64_t cmask, uint64_t cval,
                      const Mat4f& u, hipStream_t st);
hipError_t k_apply_diag(float2* psi, int n, int nbits, const int* bits, const float* phases,
                        hipStream_t st);
hipError_t k_pauli_rot(float2* psi, int n, uint64_t x, uint64_t z, float c, float s, int k,
                       hipStream_t st);
hipError_t k_project_zero(float2* psi, int n, int bit, int keep, hipStream_t st);
hipError_t k_scale(float2* psi, int n, float re, float im, hipStream_t st);
// dst[i] = i < nsrc ? src[i] : 0 for i < ndst (src may be NULL when nsrc = 0)
hipError_t k_pad_copy(const float2* src, uint64_t nsrc, float2* dst, uint64_t ndst, hipStream_t st);
hipError_t k_apply_kq(const float2* psi, float2* out, int n, int k, const uint32_t* bitpos,
                      const uint64_t* off, const float2* U, hipStream_t st);
hipError_t k_probs(const float2* psi, int n, double* out, hipStream_t st);
// blockwise FP64 partial sums: mode 0 = |psi|^2, 1 = |psi|^2 (-1)^{bit}, 2 = |psi|^2 [bit==keep]
uint32_t reduce_grid(int n);
hipError_t k_reduce_prob(const float2* psi, int n, int mode, int bit, double* partials,
                         hipStream_t st);
// out[r] = sum_g partials[g * nred + r] (fixed order)
hipError_t k_final_sum(const double* partials, uint32_t ngroups, uint32_t nred, double* out,
                       hipStream_t st);
// one final sum per slot over a sweep-wide partials buffer (desc: 4 words per
// slot: offset, blocks, stride, index)
hipError_t k_final_sum_slots(const double* partials, const uint32_t* desc, uint32_t nslots,
                             double* out, hipStream_t st);
// partials[g*2 + {0,1}] = block sums of (-1)^{|i&z|} conj(u[i]) v[i^x]  (grid = reduce_grid(n))
hipError_t k_pauli_dot(const float2* u, const float2* v, int n, uint64_t x, uint64_t z,
                       double* partials, hipStream_t st);
// out2[2t + {0,1}] = full sums of the above for terms t (one block per term; small n)
hipError_t k_pauli_dot_terms(const float2* u, const float2* v, int n, uint32_t nterms,
                             const uint64_t* dx, const uint64_t* dz, double* out2,
                             hipStream_t st);
// out = sum_k coef_k P_k psi, terms given as device arrays, in two parts.
// Table part (n <= 32): subgroups of terms sharing the flip mask x and the Z
// mask outside x, zc; 4 words each {x, zc, table offset, pattern shifts:
// 5 bits each: position of the k-th flip bit (ascending) minus k, at most 4
// flip bits, 31 for unused positions}.  The table holds, per pattern q of the partner
// index j = i ^ x on the bits of x, the summed coefficient
// sum_t c_t i^{k_t} (-1)^{|dep(q) & z_t|}, so that a subgroup adds
// (-1)^{|j & zc|} tab[off + q] psi[j].  Term part: flip groups x with their
// terms' z masks and coefficients c_t i^{k_t}.
struct HamApply {
  uint32_t nsub = 0;
  const uint32_t* sub = nullptr;
  const float2* tab = nullptr;
  size_t ntab = 0;  // table entries
  // blocks of subgroups that share a flip mask: header {x, shifts, first
  // member, end member}, members {zc, table offset}
  uint32_t nblk = 0;
  const uint32_t* blk = nullptr;
  const uint32_t* mem = nullptr;
  // blocks in chunk form: header {x, shifts, z0, first chunk, end chunk},
  // chunks {lowest bit, table offset} (tables of 2^|x| x 16 entries)
  uint32_t ncblk = 0;
  const uint32_t* cblk = nullptr;
  const uint32_t* chk = nullptr;
  uint32_t ngroups = 0;
  const uint64_t* gx = nullptr;
  const uint32_t* gstart = nullptr;
  const uint64_t* tz = nullptr;
  const float2* tc = nullptr;
  // the diagonal (x = 0) terms summed per index (2^n floats), when the term
  // part leaves them out
  const float* diag = nullptr;
};
hipError_t k_diag_build(float* diag, int n, uint32_t nt, const uint64_t* z, const double* c,
                        hipStream_t st);
// epart (optional): per block {Re, Im} of sum_i conj(psi_i) (H psi)_i, FP64
// from the unrounded (H psi)_i; ham_apply_grid(n) blocks.
hipError_t k_ham_apply(const float2* psi, float2* out, int n, const HamApply& H, hipStream_t st,
                       double* epart = nullptr);
uint32_t ham_apply_grid(int n);

// ---- batched energy+gradient of small requests (tqsv_batch.hip) ----
// One work-group per request; every plan is built for BATCH_QUBITS qubits
// (whole state = one tile, identity tile map) and contains tile passes only.
constexpr int BATCH_QUBITS = 13;
struct BatchItem {
  const float2* psi;       // caller's initial state, 2^nu amplitudes (read only)
  uint32_t nu;             // caller's qubits (<= BATCH_QUBITS)
  uint32_t nfwd, nrev;     // pass counts
  const uint32_t* fprog;   // forward program, pass records at fprog + foff[k]
  const uint32_t* rprog;   // reverse program
  const uint32_t* foff;    // device arrays: pass offsets (words) and, for the
  const uint32_t* roff;    // reverse passes, their first reduction slot
  const uint32_t* rred;
  uint32_t ngroups;        // H as flip groups (as k_ham_apply)
  const uint64_t* gx;
  const uint32_t* gstart;
  const uint64_t* tz;
  const float2* tc;
  double* out;             // reduction slots [0, nred), energy (re, im) at [nred, nred + 1]
  uint32_t nred;
};
hipError_t k_batch_adjoint(const BatchItem* items, uint32_t nreq, bool lean, uint32_t max_pass_nred,
                           hipStream_t st);

}  // namespace tqsv

#endif
