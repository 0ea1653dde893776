/*
 * tqsv.h - C ABI of libtqsv, the MI355X (gfx950) statevector engine behind
 * TyxonQ's simulator path.
 *
 * The reference has no native code and no FFI: its "kernels" are NumPy/PyTorch
 * einsums issued from Python.  Each entry point below replaces one reference
 * function (cited as file:line, relative to the reference's src/tyxonq/).  The
 * Python engine (tyxonq_amd/_native.py) binds these with ctypes.
 *
 * Conventions
 *   - A state is 2^n complex64 amplitudes, interleaved (re, im) float pairs,
 *     owned by the caller (a torch tensor on the GPU).  libtqsv never
 *     allocates or frees a state.
 *   - Amplitude index bit b belongs to qubit q = n-1-b (big-endian qubit
 *     order of the reference, engine.py:1049-1057).  All entry points take
 *     BIT positions; the qubit->bit mapping happens in the host engine.
 *   - Matrices are row-major, (re, im) interleaved doubles; they are rounded
 *     to float once on the host.  For a 4x4 matrix on bits (b0, b1) the basis
 *     index is 2*bit(b0) + bit(b1), i.e. b0 is the most significant, matching
 *     the reference's 4x4 convention (gates.py:82-90).
 *   - A Pauli string is a pair of bit masks (x, z): X where only x is set,
 *     Z where only z is set, Y where both are set.  P|i> = i^{|x&z|}
 *     (-1)^{|i&z|} |i^x>.
 *   - `stream` is a hipStream_t (NULL = default stream).  Work is enqueued
 *     asynchronously unless a result is returned in host memory, in which
 *     case the call synchronizes that stream before returning.
 *   - Return value 0 = success, negative = error; tqsv_last_error() then
 *     holds a thread-local message.  No C++ exception crosses the ABI.
 */
#ifndef TQSV_H
#define TQSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQSV_OK 0
#define TQSV_EINVAL -1    /* bad argument (shape, bit index, size) */
#define TQSV_EDEVICE -2   /* HIP runtime error */
#define TQSV_ENOMEM -3    /* host or device allocation failed */
#define TQSV_EUNSUPPORTED -4

/* ---- library / errors ---------------------------------------------------- */

/* Version number, major*10000 + minor*100 + patch. */
int tqsv_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* tqsv_last_error(void);

/* ---- single-operation kernels ------------------------------------------- */

/* |psi> = |index>.  Replaces init_statevector, statevector.py:19-25
 * (which builds a 2^n-element Python list). */
int tqsv_init_basis(void* psi, int n, uint64_t index, void* stream);

/* Apply a 2x2 matrix U (8 doubles) to bit `bit`, in place.
 * Replaces apply_1q_statevector, statevector.py:28-42. */
int tqsv_apply_1q(void* psi, int n, int bit, const double* U, void* stream);

/* Apply a 4x4 matrix U (32 doubles) to bits (bit_hi, bit_lo), in place;
 * bit_hi is the most significant index of U.  bit_hi == bit_lo is a no-op,
 * as in the reference.  Replaces apply_2q_statevector, statevector.py:45-59,
 * and the k<=2 case of apply_kqubit_unitary, statevector.py:71-129. */
int tqsv_apply_2q(void* psi, int n, int bit_hi, int bit_lo, const double* U, void* stream);

/* out = U psi for a k-qubit matrix U (2^k x 2^k complex, row-major (re, im)
 * doubles) on bits[0..k-1], bits[0] the most significant index of U; 1 <= k
 * <= 10, distinct bits; out must be a separate state (psi is not modified).
 * Replaces apply_kqubit_unitary, statevector.py:71-129, for k >= 3 (used by
 * the chemistry state preparation, statevector_ops.py:119-137). */
int tqsv_apply_kq(const void* psi, void* out, int n, int k, const int* bits, const double* U,
                  void* stream);

/* Apply U (2x2) to `bit` on the amplitudes whose index satisfies
 * (index & ctrl_mask) == ctrl_val.  Controlled gates (cx, cry, cz) of the
 * engine dispatch, engine.py:66-80, without building a 4x4 matrix. */
int tqsv_apply_ctrl(void* psi, int n, uint64_t ctrl_mask, uint64_t ctrl_val, int bit,
                    const double* U, void* stream);

/* Multiply amplitude i by phases[k] (re, im) where k is formed from the bits
 * of i listed in bits[0..nbits) (bits[0] most significant).  nbits <= 4.
 * Diagonal gates (rz, s, sdg, cz, rzz) without a dense matrix. */
int tqsv_apply_diag(void* psi, int n, int nbits, const int* bits, const double* phases,
                    void* stream);

/* |psi> <- exp(-i theta/2 P)|psi> for the Pauli string (xmask, zmask).
 * The rotation that evolve_pauli_ops, libs/circuits_library/utils.py:29-59,
 * builds from basis changes, a CX ladder and one rz. */
int tqsv_pauli_rot(void* psi, int n, uint64_t xmask, uint64_t zmask, double theta,
                   void* stream);

/* Zero the half of the state whose bit `bit` differs from `keep` and divide
 * by the norm of the kept half (left as is if that norm is 0).  The norm of
 * the kept half is returned in *kept_norm2 (host) if not NULL.
 * Replaces StatevectorEngine._project_z, engine.py:1075-1087. */
int tqsv_project_z(void* psi, int n, int bit, int keep, double* kept_norm2, void* stream);

/* psi *= (re + i im). */
int tqsv_scale(void* psi, int n, double re, double im, void* stream);

/* ---- reductions (FP64 accumulation) -------------------------------------- */

/* <psi|psi> into *out (host). */
int tqsv_norm2(const void* psi, int n, double* out, void* stream);

/* <Z_bit> = sum |psi_i|^2 (-1)^{bit of i} into *out (host).
 * Replaces expect_z_statevector, statevector.py:62-68. */
int tqsv_expect_z(const void* psi, int n, int bit, double* out, void* stream);

/* out[0] + i out[1] = sum_i (-1)^{|i & zmask|} conj(u[i]) v[i ^ xmask] (host).
 * With x = z = 0 this is <u|v>; <u|P|v> = (-i)^{|x&z|} times it.  The inner
 * product the reference takes with tensordot(conj(psi'), psi),
 * core/ir/circuit.py:1793-1796. */
int tqsv_dot(const void* u, const void* v, int n, uint64_t xmask, uint64_t zmask, double* out,
             void* stream);

/* tqsv_dot for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  No counterpart in
 * the reference.  Same sum as tqsv_dot, written as out_dev[0] (re) and
 * out_dev[1] (im) into a DEVICE buffer of the caller, without synchronising.
 * The block partial sums go to a fixed 32 KiB device array of the library
 * (one per GPU, held until the process ends); nothing is allocated.  Calls
 * that overlap on one GPU (no stream order between them) race on that array. */
int tqsv_dot_noalloc(const void* u, const void* v, int n, uint64_t xmask, uint64_t zmask,
                     double* out_dev, void* stream);

/* probs[i] = |psi_i|^2 as float64 into a DEVICE buffer of 2^n doubles.
 * Replaces StatevectorEngine.probability, engine.py:1041-1047. */
int tqsv_probs(const void* psi, int n, double* probs_dev, void* stream);

/* ---- fused circuits ------------------------------------------------------ */

#define TQSV_OP_MAT1 1 /* 2x2 matrix m[0..8) on t[0], controlled by (cmask, cval) */
#define TQSV_OP_MAT2 2 /* 4x4 matrix m[0..32) on (t[0], t[1]), controlled by (cmask, cval) */
#define TQSV_OP_PROT 3 /* exp(-i theta/2 P), P = (xmask, zmask) */

typedef struct tqsv_op {
  int32_t kind;    /* TQSV_OP_* */
  int32_t t[2];    /* target bits */
  int32_t param;   /* trainable parameter index for tqsv_adjoint, -1 if none */
  uint64_t cmask;  /* control bits (MAT1/MAT2) */
  uint64_t cval;   /* required values of the control bits, subset of cmask */
  uint64_t xmask;  /* PROT: flip mask */
  uint64_t zmask;  /* PROT: phase mask */
  double theta;    /* PROT: angle */
  /* Generator of a parametrised op, used by tqsv_adjoint:
   * op = exp(-i phi/2 G) with G the Pauli string (gx, gz), phi = gcoef*param + const.
   * For PROT ops the generator is (xmask, zmask) and gx/gz are ignored. */
  uint64_t gx;
  uint64_t gz;
  double gcoef;
  double m[32];    /* matrix, (re, im) interleaved, row-major */
} tqsv_op;

typedef struct tqsv_plan_opts {
  int32_t tile_bits;     /* bits per tile T (11..14); 0 = automatic */
  int32_t min_low_bits;  /* contiguous low bits every tile keeps; 0 = automatic */
  int32_t max_cost;      /* VALU budget per pass, ops per amplitude; 0 = automatic
                            (260 for states of 28+ qubits, 400 below) */
  int32_t fuse;          /* 1 = fused tile passes; 0 = one kernel launch per op */
  int32_t jit;           /* 1 = passes run as kernels generated and compiled for this
                            plan (hiprtc, cached per structure); -1 = the precompiled
                            interpreting kernel; 0 = automatic (generated for n >= 22 and
                            at most 64 passes).  TQSV_JIT=0/1 overrides. */
  int32_t min_qubits;    /* states of fewer qubits run as states of this many whose extra
                            high index bits stay zero (copied through a buffer of the
                            plan); 0 = 11; at most 13 (13 for tqsv_batch_adjoint) */
} tqsv_plan_opts;

typedef struct tqsv_plan_info {
  int32_t n;
  int32_t ops_in;          /* ops given */
  int32_t gates_after_merge;
  int32_t passes;          /* state sweeps (kernel launches over the state) */
  int32_t phases;          /* register phases summed over passes */
  int32_t tile_bits;
  int64_t bytes_per_run;   /* algorithmic HBM bytes: passes * 2 * 8 * 2^n */
  int64_t program_bytes;   /* device program size */
  int32_t reverse_passes;  /* dual passes of the adjoint / VJP sweep (0 until first built) */
  int32_t reverse_single;  /* reverse items run as single-op launches */
  int32_t generated;       /* forward passes run on generated (run-time compiled) kernels */
  int32_t reverse_generated; /* the same for the reverse sweep */
} tqsv_plan_info;

typedef struct tqsv_plan tqsv_plan;

/* Plan a circuit: merge 1-qubit runs, pack gates into tile passes and
 * register phases, upload the pass programs.  Replaces the per-op dispatch
 * loop of StatevectorEngine.state, engine.py:914-1038.
 * opts may be NULL.  The plan owns a small device buffer (its program). */
int tqsv_plan_create(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                     tqsv_plan** out);

/* Replace the parameters (matrices, angles) of a plan whose op structure is
 * unchanged (same kinds, targets, controls, masks) and re-upload. */
int tqsv_plan_update(tqsv_plan* plan, const tqsv_op* ops, int nops, void* stream);

int tqsv_plan_get_info(const tqsv_plan* plan, tqsv_plan_info* info);

/* Plan on the host only (no device memory is touched) and report what
 * tqsv_plan_create would build.  For planning inspection and tests. */
int tqsv_plan_inspect(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                      tqsv_plan_info* info);

/* Plan on the host and compile the generated kernel of every tile pass
 * (with adjoint != 0 also those of the reverse sweep) without touching a
 * device: the compiler keeps the code objects in its disk cache, so that a
 * later tqsv_plan_create with generated kernels (opts->jit = 1) of the same
 * circuit structure loads them instead of compiling.  Needs no GPU.
 * *kernels (may be NULL) = distinct kernels compiled.  The reference has no
 * counterpart (it compiles nothing); for serving start-up and tests. */
int tqsv_plan_precompile(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                         int adjoint, int* kernels);

/* Plan on the host and execute the pass programs on a HOST complex64 state
 * (2^n interleaved float pairs, in place, n <= 26) with the device kernel's
 * index arithmetic and step semantics.  For testing the planner and the
 * program format without a GPU. */
int tqsv_plan_emulate(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                      float* state);

/* Run the planned circuit on psi in place. */
int tqsv_plan_run(tqsv_plan* plan, void* psi, void* stream);

void tqsv_plan_destroy(tqsv_plan* plan);

/* One-shot: plan, run, destroy. */
int tqsv_apply_fused(void* psi, int n, const tqsv_op* ops, int nops,
                     const tqsv_plan_opts* opts, void* stream);

/* ---- Pauli sums ---------------------------------------------------------- */

typedef struct tqsv_ham tqsv_ham;

/* H = sum_k coeff[k] P_k with P_k = (xmask[k], zmask[k]).  Plans the fused
 * expectation passes for an n-qubit state.  Replaces the dense
 * pauli_string_sum_dense + psi^H H psi route, pauli.py:74-87 and
 * dynamics.py:117-126, and the per-term re-simulation of
 * Circuit._expectation_statevector, core/ir/circuit.py:1751-1796. */
int tqsv_ham_create(int n, int nterms, const uint64_t* xmask, const uint64_t* zmask,
                    const double* coeff, tqsv_ham** out);

/* Plan a Pauli sum's expectation passes on the host and compile their
 * generated kernels (the ones tqsv_ham_create would use: states of 2^22+
 * amplitudes) into the compiler's disk cache, without a GPU; *kernels (may be
 * NULL) = distinct kernels compiled (0 when the plan uses the precompiled
 * kernel).  No reference counterpart; for start-up and tests, as
 * tqsv_plan_precompile. */
int tqsv_ham_precompile(int n, int nterms, const uint64_t* xmask, const uint64_t* zmask,
                        const double* coeff, int* kernels);

/* Per-term <psi|P_k|psi> into terms[nterms] (host, may be NULL), the energy
 * sum_k coeff_k <P_k> into *energy (host, may be NULL) and <psi|psi> into
 * *norm2 (host, may be NULL).  All accumulated in FP64; the energy is NOT
 * divided by the norm. */
/* Replace the coefficients of a Hamiltonian (same terms, same order), e.g. a
 * field strength that changes between requests; the term structure and the
 * device plans are kept.  No counterpart in the reference (whose Pauli sums
 * are plain lists, vqe.py:9-20). */
int tqsv_ham_set_coeffs(tqsv_ham* ham, int nterms, const double* coeff, void* stream);

int tqsv_ham_expval(tqsv_ham* ham, const void* psi, double* terms, double* energy,
                    double* norm2, void* stream);

typedef struct tqsv_ham_info {
  int32_t n;
  int32_t nterms;
  int32_t passes;       /* state sweeps of one expectation (read-only) */
  int32_t phases;       /* register phases summed over the fused passes */
  int32_t flip_groups;  /* distinct flip masks x */
  int32_t unfused_terms;/* terms evaluated by the one-term-per-launch kernels */
} tqsv_ham_info;

int tqsv_ham_get_info(const tqsv_ham* ham, tqsv_ham_info* info);

/* out = H psi (out must not alias psi). */
int tqsv_ham_apply(tqsv_ham* ham, const void* psi, void* out, void* stream);

void tqsv_ham_destroy(tqsv_ham* ham);

/* One-shot expectation of a Pauli sum. */
int tqsv_expval(const void* psi, int n, int nterms, const uint64_t* xmask,
                const uint64_t* zmask, const double* coeff, double* terms, double* energy,
                double* norm2, void* stream);

/* ---- batched energy and gradient of small requests ----------------------- */

/* Energy and gradient of `nreq` requests in ONE kernel launch: one work-group
 * per request, whose state and lambda stay in LDS for the whole forward sweep,
 * H|psi>, energy and reverse sweep.  Request i: plans[i] built with
 * min_qubits = 13 for a state of at most 13 qubits, hams[i] of the caller's
 * size, psi[i] its initial state (device, read only), grads[i] (host,
 * nparams[i] doubles) and energies[i] receive the results.  Requests of
 * different sizes and depths share the launch; each result depends only on
 * its own request.  No counterpart in the reference, whose serving of
 * energy+gradient requests is one engine call per request
 * (numerics/backends/pytorch_backend.py:446-507 per call). */
typedef struct tqsv_batch tqsv_batch;
int tqsv_batch_create(tqsv_batch** out);
void tqsv_batch_destroy(tqsv_batch* batch);
int tqsv_batch_adjoint(tqsv_batch* batch, int nreq, tqsv_plan* const* plans,
                       tqsv_ham* const* hams, const void* const* psi, const int* nparams,
                       double* const* grads, double* energies, void* stream);

/* ---- density matrices ---------------------------------------------------- */

/* A density matrix of n qubits is a 2n-qubit state vec(rho), vec[(r << n) | c]
 * = rho[r, c]; gates and channels run through the plans on 2n qubits.
 * Tr(rho P_t) for Pauli terms over the n qubits (masks as tqsv_ham_create)
 * into terms[t] (host, may be NULL) and sum_t coeff[t] Tr(rho P_t) into *value
 * (coeff may be NULL: all ones); FP64 reductions.  n <= 16.  Replaces
 * exp_z_density (libs/quantum_library/kernels/density_matrix.py:66-73) and
 * DensityMatrixEngine.expval (devices/simulators/density_matrix/engine.py:
 * 153-183, dense trace of rho H). */
int tqsv_dm_expval(const void* rho, int n, int nterms, const uint64_t* xmask,
                   const uint64_t* zmask, const double* coeff, double* terms, double* value,
                   void* stream);

/* out[r] = sqrt(max(Re rho[r, r], 0)) (an n-qubit complex64 state whose |.|^2
 * is the diagonal of rho), the input of tqsv_sampler_run for counts from a
 * density matrix (engine.py:103-151). */
int tqsv_dm_diag_amplitudes(const void* rho, int n, void* out, void* stream);

/* ---- shot sampling ------------------------------------------------------- */

/* A sampler for n-qubit states: owns the device scratch for up to max_shots
 * shots per call (about 50 B per shot plus 40 B per 4096 amplitudes). */
typedef struct tqsv_sampler tqsv_sampler;
int tqsv_sampler_create(int n, int64_t max_shots, tqsv_sampler** out);
void tqsv_sampler_destroy(tqsv_sampler* sampler);

/* Draw `shots` basis states from |psi|^2 / <psi|psi> (inverse CDF, FP64
 * cumulative sums, Philox4x64-10 stream keyed by `seed`) and return the
 * distinct states with their counts in increasing index order:
 * index[0..*nunique), count[0..*nunique) (host arrays of at least
 * min(shots, 2^n) entries).  Synchronises the stream.  Replaces the counts
 * path of StatevectorEngine.run (devices/simulators/statevector/engine.py:
 * 377-465: |psi|^2 on the host, numpy multinomial, bincount). */
int tqsv_sampler_run(tqsv_sampler* sampler, const void* psi, int64_t shots, uint64_t seed,
                     uint64_t* index, int64_t* count, int64_t* nunique, void* stream);

/* ---- adjoint differentiation -------------------------------------------- */

/* Energy and gradient of E(theta) = <psi(theta)|H|psi(theta)> by one forward
 * sweep and one reverse (uncompute) sweep.
 *   psi  : in = initial state, out = final state psi(theta)
 *   lam  : caller scratch state of the same size (overwritten)
 *   grad : host, nparams doubles, overwritten with dE/dparam
 *   energy: host, may be NULL
 * Replaces PyTorchBackend.value_and_grad through autograd,
 * numerics/backends/pytorch_backend.py:446-507 (which stores every
 * intermediate state), and the 2P-run parameter shift of
 * libs/quantum_library/kernels/common.py:11-23. */
int tqsv_adjoint(tqsv_plan* plan, tqsv_ham* ham, void* psi, void* lam, int nparams,
                 double* grad, double* energy, void* stream);

/* Vector-Jacobian product of psi(theta) = U(theta)|psi0>:
 *   psi : in = final state U|psi0>, out = U^dagger... recovered initial state
 *   lam : in = cotangent dL/dpsi (PyTorch convention), overwritten
 *   grad: host, nparams doubles, overwritten with dL/dparam = Re<lam|dpsi/dparam>
 * Backward of Circuit.state() under torch autograd, core/ir/circuit.py:427-503. */
int tqsv_vjp(tqsv_plan* plan, void* psi, void* lam, int nparams, double* grad, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TQSV_H */
