// tqsv_abi.cpp - error state, single-op entry points and plan management of
// the libtqsv C ABI (declared in include/tqsv.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "tqsv_planner.h"

namespace tqsv {

namespace {
thread_local std::string g_err;

// per-thread reduction scratch for the one-shot reductions
struct Scratch {
  double* d = nullptr;
  size_t cap = 0;
  ~Scratch() {
    if (d) (void)hipFree(d);
  }
};
thread_local Scratch g_scratch;

int scratch(size_t ndoubles, double** out) {
  if (g_scratch.cap < ndoubles) {
    if (g_scratch.d) TQSV_HIP(hipFree(g_scratch.d));
    g_scratch.d = nullptr;
    g_scratch.cap = 0;
    TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&g_scratch.d), ndoubles * sizeof(double)));
    g_scratch.cap = ndoubles;
  }
  *out = g_scratch.d;
  return 0;
}

int check_state(const void* psi, int n, const char* fn) {
  if (!psi) return fail(TQSV_EINVAL, std::string(fn) + ": state pointer is NULL");
  if (n < 1 || n > 34) return fail(TQSV_EINVAL, std::string(fn) + ": n must be in [1, 34]");
  return 0;
}

int check_bit(int b, int n, const char* fn) {
  if (b < 0 || b >= n)
    return fail(TQSV_EINVAL, std::string(fn) + ": bit " + std::to_string(b) + " out of range for n=" +
                                 std::to_string(n));
  return 0;
}

// sum of one reduce_prob pass, to host
int reduce_prob_to_host(const float2* psi, int n, int mode, int bit, double* out, hipStream_t st) {
  const uint32_t g = reduce_grid(n);
  double* d = nullptr;
  int rc = scratch((size_t)g + 1, &d);
  if (rc) return rc;
  TQSV_HIP(k_reduce_prob(psi, n, mode, bit, d, st));
  TQSV_HIP(k_final_sum(d, g, 1, d + g, st));
  TQSV_HIP(hipMemcpyAsync(out, d + g, sizeof(double), hipMemcpyDeviceToHost, st));
  TQSV_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return TQSV_EDEVICE;
}

int plan_upload(uint32_t** dptr, size_t* cap, const std::vector<uint32_t>& prog, hipStream_t st);
int plan_build_forward(tqsv_plan* p, const tqsv_op* ops, int nops, hipStream_t st);

}  // namespace tqsv

using namespace tqsv;

static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
static float2* F(void* p) { return reinterpret_cast<float2*>(p); }
static const float2* CF(const void* p) { return reinterpret_cast<const float2*>(p); }

extern "C" {

int tqsv_version(void) { return 100; }

const char* tqsv_last_error(void) { return g_err.c_str(); }

int tqsv_init_basis(void* psi, int n, uint64_t index, void* stream) {
  int rc = check_state(psi, n, "tqsv_init_basis");
  if (rc) return rc;
  if (index >> n) return fail(TQSV_EINVAL, "tqsv_init_basis: index out of range");
  TQSV_HIP(k_init_basis(F(psi), n, index, S(stream)));
  return 0;
}

int tqsv_apply_1q(void* psi, int n, int bit, const double* U, void* stream) {
  return tqsv_apply_ctrl(psi, n, 0, 0, bit, U, stream);
}

int tqsv_apply_ctrl(void* psi, int n, uint64_t ctrl_mask, uint64_t ctrl_val, int bit,
                    const double* U, void* stream) {
  int rc = check_state(psi, n, "tqsv_apply_1q");
  if (rc) return rc;
  if ((rc = check_bit(bit, n, "tqsv_apply_1q"))) return rc;
  if (!U) return fail(TQSV_EINVAL, "tqsv_apply_1q: matrix is NULL");
  if ((ctrl_mask >> n) || (ctrl_val & ~ctrl_mask) || ((ctrl_mask >> bit) & 1u))
    return fail(TQSV_EINVAL, "tqsv_apply_ctrl: bad control mask");
  Mat2f u;
  for (int j = 0; j < 8; ++j) u.v[j] = (float)U[j];
  TQSV_HIP(k_apply_1q(F(psi), n, bit, ctrl_mask, ctrl_val, u, S(stream)));
  return 0;
}

int tqsv_apply_2q(void* psi, int n, int bit_hi, int bit_lo, const double* U, void* stream) {
  int rc = check_state(psi, n, "tqsv_apply_2q");
  if (rc) return rc;
  if ((rc = check_bit(bit_hi, n, "tqsv_apply_2q")) || (rc = check_bit(bit_lo, n, "tqsv_apply_2q")))
    return rc;
  if (!U) return fail(TQSV_EINVAL, "tqsv_apply_2q: matrix is NULL");
  if (bit_hi == bit_lo) return 0;  // the reference returns the state unchanged
  Mat4f u;
  for (int j = 0; j < 32; ++j) u.v[j] = (float)U[j];
  TQSV_HIP(k_apply_2q(F(psi), n, bit_hi, bit_lo, 0, 0, u, S(stream)));
  return 0;
}

int tqsv_apply_kq(const void* psi, void* out, int n, int k, const int* bits, const double* U,
                  void* stream) {
  int rc = check_state(psi, n, "tqsv_apply_kq");
  if (rc) return rc;
  if (!out || out == psi) return fail(TQSV_EINVAL, "tqsv_apply_kq: out must be a separate state");
  if (k < 1 || k > 10 || k > n || !bits || !U)
    return fail(TQSV_EINVAL, "tqsv_apply_kq: k must be in [1, min(n, 10)]");
  uint64_t seen = 0;
  std::vector<uint32_t> pos(k);
  for (int b = 0; b < k; ++b) {
    if ((rc = check_bit(bits[b], n, "tqsv_apply_kq"))) return rc;
    if ((seen >> bits[b]) & 1ull) return fail(TQSV_EINVAL, "tqsv_apply_kq: repeated target bit");
    seen |= 1ull << bits[b];
    pos[b] = (uint32_t)bits[b];
  }
  const size_t K = (size_t)1 << k;
  std::vector<uint64_t> off(K, 0);
  for (size_t j = 0; j < K; ++j)
    for (int b = 0; b < k; ++b)
      if ((j >> (k - 1 - b)) & 1u) off[j] |= 1ull << pos[b];
  std::vector<float2> u(K * K);
  for (size_t e = 0; e < K * K; ++e) u[e] = make_float2((float)U[2 * e], (float)U[2 * e + 1]);
  // per-call device copies of the matrix and tables (freed after the launch completes)
  const size_t bytes = K * K * sizeof(float2) + K * sizeof(uint64_t) + (size_t)k * sizeof(uint32_t);
  unsigned char* d = nullptr;
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&d), bytes));
  float2* dU = reinterpret_cast<float2*>(d);
  uint64_t* dOff = reinterpret_cast<uint64_t*>(d + K * K * sizeof(float2));
  uint32_t* dPos = reinterpret_cast<uint32_t*>(d + K * K * sizeof(float2) + K * sizeof(uint64_t));
  hipError_t e = hipMemcpyAsync(dU, u.data(), K * K * sizeof(float2), hipMemcpyHostToDevice, S(stream));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dOff, off.data(), K * sizeof(uint64_t), hipMemcpyHostToDevice, S(stream));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dPos, pos.data(), (size_t)k * sizeof(uint32_t), hipMemcpyHostToDevice, S(stream));
  if (e == hipSuccess) e = k_apply_kq(CF(psi), F(out), n, k, dPos, dOff, dU, S(stream));
  if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
  (void)hipFree(d);
  TQSV_HIP(e);
  return 0;
}

int tqsv_apply_diag(void* psi, int n, int nbits, const int* bits, const double* phases,
                    void* stream) {
  int rc = check_state(psi, n, "tqsv_apply_diag");
  if (rc) return rc;
  if (nbits < 1 || nbits > 4 || !bits || !phases)
    return fail(TQSV_EINVAL, "tqsv_apply_diag: nbits must be in [1, 4]");
  for (int j = 0; j < nbits; ++j)
    if ((rc = check_bit(bits[j], n, "tqsv_apply_diag"))) return rc;
  float ph[32];
  for (int k = 0; k < (2 << nbits); ++k) ph[k] = (float)phases[k];
  TQSV_HIP(k_apply_diag(F(psi), n, nbits, bits, ph, S(stream)));
  return 0;
}

int tqsv_pauli_rot(void* psi, int n, uint64_t xmask, uint64_t zmask, double theta, void* stream) {
  int rc = check_state(psi, n, "tqsv_pauli_rot");
  if (rc) return rc;
  if ((xmask | zmask) >> n) return fail(TQSV_EINVAL, "tqsv_pauli_rot: mask out of range");
  const int k = __builtin_popcountll(xmask & zmask) & 3;
  TQSV_HIP(k_pauli_rot(F(psi), n, xmask, zmask, (float)std::cos(0.5 * theta),
                       (float)std::sin(0.5 * theta), k, S(stream)));
  return 0;
}

int tqsv_project_z(void* psi, int n, int bit, int keep, double* kept_norm2, void* stream) {
  int rc = check_state(psi, n, "tqsv_project_z");
  if (rc) return rc;
  if ((rc = check_bit(bit, n, "tqsv_project_z"))) return rc;
  if (keep != 0 && keep != 1) return fail(TQSV_EINVAL, "tqsv_project_z: keep must be 0 or 1");
  double kept = 0.0;
  rc = reduce_prob_to_host(CF(psi), n, 2, bit, &kept, S(stream));
  if (rc) return rc;
  if (keep == 0) {
    double tot = 0.0;
    rc = reduce_prob_to_host(CF(psi), n, 0, bit, &tot, S(stream));
    if (rc) return rc;
    kept = tot - kept;
  }
  TQSV_HIP(k_project_zero(F(psi), n, bit, keep, S(stream)));
  if (kept > 0.0) {
    const double inv = 1.0 / std::sqrt(kept);
    TQSV_HIP(k_scale(F(psi), n, (float)inv, 0.f, S(stream)));
  }
  if (kept_norm2) *kept_norm2 = kept;
  return 0;
}

int tqsv_scale(void* psi, int n, double re, double im, void* stream) {
  int rc = check_state(psi, n, "tqsv_scale");
  if (rc) return rc;
  TQSV_HIP(k_scale(F(psi), n, (float)re, (float)im, S(stream)));
  return 0;
}

int tqsv_norm2(const void* psi, int n, double* out, void* stream) {
  int rc = check_state(psi, n, "tqsv_norm2");
  if (rc) return rc;
  if (!out) return fail(TQSV_EINVAL, "tqsv_norm2: out is NULL");
  return reduce_prob_to_host(CF(psi), n, 0, 0, out, S(stream));
}

int tqsv_expect_z(const void* psi, int n, int bit, double* out, void* stream) {
  int rc = check_state(psi, n, "tqsv_expect_z");
  if (rc) return rc;
  if ((rc = check_bit(bit, n, "tqsv_expect_z"))) return rc;
  if (!out) return fail(TQSV_EINVAL, "tqsv_expect_z: out is NULL");
  return reduce_prob_to_host(CF(psi), n, 1, bit, out, S(stream));
}

int tqsv_dot(const void* u, const void* v, int n, uint64_t xmask, uint64_t zmask, double* out,
             void* stream) {
  int rc = check_state(u, n, "tqsv_dot");
  if (rc) return rc;
  if ((rc = check_state(v, n, "tqsv_dot"))) return rc;
  if (!out) return fail(TQSV_EINVAL, "tqsv_dot: out is NULL");
  if ((xmask | zmask) >> n) return fail(TQSV_EINVAL, "tqsv_dot: mask out of range");
  const uint32_t g = reduce_grid(n);
  double* d = nullptr;
  if ((rc = scratch((size_t)2 * g + 2, &d))) return rc;
  TQSV_HIP(k_pauli_dot(CF(u), CF(v), n, xmask, zmask, d, S(stream)));
  TQSV_HIP(k_final_sum(d, g, 2, d + 2 * g, S(stream)));
  TQSV_HIP(hipMemcpyAsync(out, d + 2 * g, 2 * sizeof(double), hipMemcpyDeviceToHost, S(stream)));
  TQSV_HIP(hipStreamSynchronize(S(stream)));
  return 0;
}

int tqsv_probs(const void* psi, int n, double* probs_dev, void* stream) {
  int rc = check_state(psi, n, "tqsv_probs");
  if (rc) return rc;
  if (!probs_dev) return fail(TQSV_EINVAL, "tqsv_probs: output is NULL");
  TQSV_HIP(k_probs(CF(psi), n, probs_dev, S(stream)));
  return 0;
}

// ---- plans ----

int tqsv_plan_create(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                     tqsv_plan** out) {
  try {
    if (!out) return fail(TQSV_EINVAL, "tqsv_plan_create: out is NULL");
    *out = nullptr;
    if (n < 1 || n > 34) return fail(TQSV_EINVAL, "tqsv_plan_create: n must be in [1, 34]");
    if (nops < 0 || (nops > 0 && !ops)) return fail(TQSV_EINVAL, "tqsv_plan_create: bad op array");
    if (opts && (opts->min_qubits < 0 || opts->min_qubits > 13))
      return fail(TQSV_EINVAL, "tqsv_plan_create: min_qubits must be in [0, 13]");
    auto p = std::unique_ptr<tqsv_plan, void (*)(tqsv_plan*)>(new tqsv_plan(), tqsv_plan_destroy);
    p->nu = n;
    p->nops = nops;
    if (opts) {
      p->so.tile_bits = opts->tile_bits > 0 ? opts->tile_bits : 13;
      p->so.min_low = opts->min_low_bits > 0 ? opts->min_low_bits : 5;
      p->so.max_cost = opts->max_cost > 0 ? opts->max_cost : 0;
      p->so.fuse = opts->fuse != 0;
    }
    p->n = plan_qubits(n, p->so.fuse, opts ? opts->min_qubits : 0);
    p->so.n = p->n;
    if (p->n > n) {
      const size_t bytes = sizeof(float2) << p->n;
      TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&p->pad_a), bytes));
      TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&p->pad_b), bytes));
    }
    p->jit_req = opts ? opts->jit : 0;
    const int rc = plan_build_forward(p.get(), ops, nops, nullptr);
    if (rc) return rc;
    *out = p.release();
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_plan_create: out of host memory");
  } catch (const std::exception& e) {
    return fail(TQSV_EINVAL, std::string("tqsv_plan_create: ") + e.what());
  }
}

int tqsv_plan_update(tqsv_plan* p, const tqsv_op* ops, int nops, void* stream) {
  try {
    if (!p) return fail(TQSV_EINVAL, "tqsv_plan_update: plan is NULL");
    if (nops != p->nops) return fail(TQSV_EINVAL, "tqsv_plan_update: op count changed");
    if (nops > 0 && !ops) return fail(TQSV_EINVAL, "tqsv_plan_update: bad op array");
    return plan_build_forward(p, ops, nops, S(stream));
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_plan_update: out of host memory");
  }
}

int tqsv_plan_get_info(const tqsv_plan* p, tqsv_plan_info* info) {
  if (!p || !info) return fail(TQSV_EINVAL, "tqsv_plan_get_info: NULL argument");
  std::memset(info, 0, sizeof(*info));
  info->n = p->nu;
  info->ops_in = p->nops;
  info->gates_after_merge = (int32_t)p->gates.size();
  int passes = 0, phases = 0;
  for (const PassPlan& pp : p->passes) {
    ++passes;
    phases += (int)pp.phases.size();
  }
  info->passes = passes;
  info->phases = phases;
  info->tile_bits = effective_tile_bits(p->n, p->so.tile_bits, false);
  info->bytes_per_run = (int64_t)passes * 16ll * (1ll << p->n);
  info->program_bytes = (int64_t)(p->prog.size() * sizeof(uint32_t));
  int rt = 0, rs = 0;
  for (const PassPlan& pp : p->rpasses) (pp.tile ? rt : rs) += 1;
  info->reverse_passes = rt;
  info->reverse_single = rs;
  info->generated = p->fgen ? 1 : 0;
  info->reverse_generated = p->rgen ? 1 : 0;
  return 0;
}

int tqsv_plan_inspect(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                      tqsv_plan_info* info) {
  try {
    if (!info) return fail(TQSV_EINVAL, "tqsv_plan_inspect: info is NULL");
    if (n < 1 || n > 34) return fail(TQSV_EINVAL, "tqsv_plan_inspect: n must be in [1, 34]");
    if (nops < 0 || (nops > 0 && !ops)) return fail(TQSV_EINVAL, "tqsv_plan_inspect: bad op array");
    tqsv_plan p;
    p.nu = n;
    p.nops = nops;
    if (opts) {
      p.so.tile_bits = opts->tile_bits > 0 ? opts->tile_bits : 13;
      p.so.min_low = opts->min_low_bits > 0 ? opts->min_low_bits : 5;
      p.so.max_cost = opts->max_cost > 0 ? opts->max_cost : 0;
      p.so.fuse = opts->fuse != 0;
    }
    p.n = plan_qubits(n, p.so.fuse, opts ? opts->min_qubits : 0);
    p.so.n = p.n;
    std::vector<Gate> g;
    int rc = gates_from_ops(n, ops, nops, g);
    if (rc) return rc;
    merge_and_classify(g);
    uint32_t red_total = 0;
    rc = schedule(g, p.so, p.passes, p.prog, red_total);
    if (rc) return rc;
    p.gates.swap(g);
    return tqsv_plan_get_info(&p, info);
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_plan_inspect: out of host memory");
  }
}

int tqsv_plan_precompile(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                         int adjoint, int* kernels) {
  try {
    if (n < 1 || n > 34) return fail(TQSV_EINVAL, "tqsv_plan_precompile: n must be in [1, 34]");
    if (nops < 0 || (nops > 0 && !ops)) return fail(TQSV_EINVAL, "tqsv_plan_precompile: bad op array");
    if (opts && (opts->min_qubits < 0 || opts->min_qubits > 13))
      return fail(TQSV_EINVAL, "tqsv_plan_precompile: min_qubits must be in [0, 13]");
    SchedOpts so;
    if (opts) {
      so.tile_bits = opts->tile_bits > 0 ? opts->tile_bits : 13;
      so.min_low = opts->min_low_bits > 0 ? opts->min_low_bits : 5;
      so.max_cost = opts->max_cost > 0 ? opts->max_cost : 0;
      so.fuse = opts->fuse != 0;
    }
    const int pn = plan_qubits(n, so.fuse, opts ? opts->min_qubits : 0);
    so.n = pn;
    std::vector<Gate> g;
    int rc = gates_from_ops(n, ops, nops, g);
    if (rc) return rc;
    merge_and_classify(g);
    std::vector<PassPlan> passes;
    std::vector<uint32_t> prog;
    uint32_t red_total = 0;
    rc = schedule(g, so, passes, prog, red_total);
    if (rc) return rc;
    int total = 0, k = 0;
    rc = jit_compile_only(passes, prog, pn, &k);
    if (rc) return rc;
    total += k;
    if (adjoint) {
      std::vector<Gate> r;
      reverse_gates(g, r);
      SchedOpts ro = so;
      ro.dual = true;
      std::vector<PassPlan> rpasses;
      std::vector<uint32_t> rprog;
      red_total = 0;
      rc = schedule(r, ro, rpasses, rprog, red_total);
      if (rc) return rc;
      rc = jit_compile_only(rpasses, rprog, pn, &k);
      if (rc) return rc;
      total += k;
    }
    if (kernels) *kernels = total;
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_plan_precompile: out of host memory");
  }
}

int tqsv_plan_run(tqsv_plan* p, void* psi, void* stream) {
  if (!p || !psi) return fail(TQSV_EINVAL, "tqsv_plan_run: NULL argument");
  hipStream_t st = S(stream);
  if (p->n == p->nu)
    return run_passes(p->passes, p->gates, p->n, p->dprog, F(psi), nullptr, nullptr, nullptr, st);
  TQSV_HIP(k_pad_copy(F(psi), 1ull << p->nu, p->pad_a, 1ull << p->n, st));
  const int rc =
      run_passes(p->passes, p->gates, p->n, p->dprog, p->pad_a, nullptr, nullptr, nullptr, st);
  if (rc) return rc;
  TQSV_HIP(hipMemcpyAsync(psi, p->pad_a, sizeof(float2) << p->nu, hipMemcpyDeviceToDevice, st));
  return 0;
}

void tqsv_plan_destroy(tqsv_plan* p) {
  if (!p) return;
  (void)hipFree(p->dprog);
  (void)hipFree(p->drprog);
  (void)hipFree(p->dpartials);
  (void)hipFree(p->dred);
  (void)hipFree(p->dslots);
  (void)hipFree(p->pad_a);
  (void)hipFree(p->pad_b);
  delete p;
}

int tqsv_apply_fused(void* psi, int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                     void* stream) {
  tqsv_plan* p = nullptr;
  int rc = tqsv_plan_create(n, ops, nops, opts, &p);
  if (rc) return rc;
  rc = tqsv_plan_run(p, psi, stream);
  if (!rc && hipStreamSynchronize(S(stream)) != hipSuccess)
    rc = fail(TQSV_EDEVICE, "tqsv_apply_fused: stream synchronize failed");
  tqsv_plan_destroy(p);
  return rc;
}

}  // extern "C"
