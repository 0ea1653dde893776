// tqsv_jit.cpp - tile-pass kernels generated per pass program and compiled at
// run time with hiprtc.
//
// The precompiled kernel (tqsv_pass_impl.h) reads each step's kind and slot
// bits from the program and branches to one of many code arms; with 64 live
// slot registers every such branch costs register copies at the join, scalar
// branches and waits on scalar loads.  For plans that are run many times the
// step sequence of each pass is instead written out as straight-line code: a
// generated kernel fixes the tile geometry, each phase's register positions
// and each step's kind and slot bits, and calls the same step functions
// (tqsv_device.h).  The numbers (angles, phases, matrices) are still read from
// the program buffer, so the compiled code stays valid when the parameters of
// a plan change without changing its structure; the generated source is the
// cache key.
//
// The compiler is the hiprtc of the ROCm installation ($ROCM_PATH/lib, else
// /opt/rocm/lib), loaded into a linker namespace of its own.  A process that
// imported a framework with a bundled HIP runtime (PyTorch ships its own
// libhiprtc and libamd_comgr) would otherwise bind hiprtc to that bundle, and
// with it an older device compiler: the expectation kernels of a 500-term
// 30-qubit Pauli sum came out of ROCm 7.0's compiler with 206-327 registers
// and took 60.6 ms, out of ROCm 7.2's with 178-266 and took 43.5 ms.  Only the
// source-to-code-object step runs in that namespace; code objects are loaded
// by the process's HIP runtime.  TQSV_HIPRTC=<path> selects another hiprtc,
// TQSV_HIPRTC=process the one the process resolved.
#include <dlfcn.h>
#include <unistd.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tqsv_planner.h"

namespace tqsv {

// header texts, embedded at build time (tqsv_jit_headers.cpp, written by _build.py)
extern const char* const kJitHdrDevice;
extern const char* const kJitHdrProg;

namespace {

struct JitKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_mu;
std::unordered_map<std::string, JitKernel> g_cache;  // device id + source -> kernel

// kernel names by mode, so that profiles tell the pass kinds apart
const char* kernel_name(int mode) {
  return mode == MODE_DUAL ? "tqsv_adjoint_pass" : mode == MODE_RED ? "tqsv_expval_pass" : "tqsv_gate_pass";
}

int pass_mode(uint32_t flags) {
  if (flags & PF_DUAL) return MODE_DUAL;
  return (flags & PF_WRITE) ? MODE_GATE : MODE_RED;
}

// slots of the pairs (s, s ^ xr) whose class (the bits of s on xr above its
// lowest bit) is in `act`
uint32_t class_slots(uint32_t xr, uint32_t act) {
  const uint32_t lb = xr & (~xr + 1u), hi = xr & ~lb;
  uint32_t m = 0;
  for (uint32_t s = 0; s < (uint32_t)NSLOT; ++s) {
    if (s & lb) continue;
    uint32_t cls = 0, k = 0;
    for (int b = 0; b < RB; ++b)
      if ((hi >> b) & 1u) {
        if ((s >> b) & 1u) cls |= 1u << k;
        ++k;
      }
    if ((act >> cls) & 1u) m |= (1u << s) | (1u << (s ^ xr));
  }
  return m;
}

// The slots a phase's steps read or write: rotation groups and group
// reductions touch the pairs of their active classes, every other step all
// slots.  ws = offset of the phase's first step in pw.
uint32_t phase_slots(const uint32_t* pw, size_t ws, uint32_t nsteps) {
  uint32_t m = 0;
  for (uint32_t is = 0; is < nsteps; ++is) {
    const uint32_t* w = pw + ws;
    const uint32_t op = w[0] & 0xffffu;
    if (op == OP_PGRP)
      m |= class_slots(w[1], w[5]);
    else if (op == OP_RED && w[2] == RED_GROUP)
      m |= class_slots(w[1], w[RED_HDR + RED_ENT + 1]);
    else
      return 0xffffffffu;
    ws += w[0] >> 16;
  }
  return m;
}

// Source of the kernel for one pass record (pw = the pass's program words).
std::string gen_source(int tb, const uint32_t* pw, int n) {
  const uint32_t flags = pw[PH_FLAGS], L = pw[PH_L], m = pw[PH_M];
  const uint32_t nphase = pw[PH_NPHASE], nred = pw[PH_NRED];
  const int mode = pass_mode(flags);
  const bool write = (flags & PF_WRITE) != 0;
  const int NT = 1 << tb, TSIZE = 1 << (tb + RB), NWAVE = std::max(NT / 64, 1);
  // Streaming (non-temporal) hints on the tile loads and stores of states
  // larger than the last-level cache, which are read and written once per pass
  // (30-qubit brickwork + 500-term expectation: 134.9 -> 130.0 ms per step).
  // TQSV_NT overrides: bit 0 loads, bit 1 stores.
  const char* nte = std::getenv("TQSV_NT");
  const int ntm = nte && nte[0] ? std::atoi(nte) : (n >= 25 ? 3 : 0);
  const bool nt_ld = (ntm & 1) != 0, nt_st = (ntm & 2) != 0;
  // Touch the next tile's lines ahead of its loads in read-only passes over
  // states larger than the last-level cache (2^25 amplitudes = 256 MiB): there
  // each tile load waits on HBM latency with little else to do (500-term
  // 30-qubit expectation 43.4 -> 33.0 ms).  Gate and adjoint passes are bound
  // by VALU issue or cache-resident, and the extra loads cost them 3-7%.
  // TQSV_PREFETCH=1 / 0 forces the touches on / off for every pass.
  const char* pfe = std::getenv("TQSV_PREFETCH");
  // (Two tile buffers with the next tile copied asynchronously into the
  // other one were measured slower: 34.3 against 29.3 ms for the 500-term
  // 30-qubit expectation.  They allow one work-group per CU, and 64 KB in
  // flight per CU does not cover the HBM latency under load.)
  // Pipelined passes: the next tile is loaded into registers while the
  // current one is worked on in LDS, and a finished tile leaves through
  // registers (jit_fetch_tile / jit_put_tile / jit_get_tile / jit_put_global).
  // Gate passes without reductions only (30-qubit brickwork gate part 96.3
  // -> 90.6 ms).  Dual passes staged the same way (psi and lambda, about 110
  // accumulation registers per lane) were slower: with two tiles per
  // work-group at 22 qubits only one load is hidden and the copies cost more,
  // 54.4 -> 58.9 us per dual pass.  Read-only passes keep their FP64
  // accumulators in registers: with 64 more for the next tile, 22
  // accumulators already spill 271 registers at 2 waves per SIMD, so they
  // keep the next-tile touches.  TQSV_PIPE=0 turns it off.
  const char* ppe = std::getenv("TQSV_PIPE");
  const bool pipe = mode == MODE_GATE && write && nred == 0 && !(ppe && ppe[0] == '0');
  // Dual passes with one tile buffer: psi and lambda go through the same
  // 64 KB of LDS in turn (loads, phase changes, stores), so that two
  // work-groups fit per CU instead of one.  Only when every phase loads and
  // stores all slots: phases that skip slots rely on both tiles staying in
  // LDS.  TQSV_DUAL1=0 / 1 forces it off / on where it applies.
  bool all_slots = true;
  for (uint32_t ip = 0, at = PH_SIZE; ip < nphase; ++ip) {
    const uint32_t ns = pw[at + PS_NSTEPS];
    if (phase_slots(pw, at + PS_SIZE, ns) != 0xffffffffu) all_slots = false;
    at += PS_SIZE;
    for (uint32_t k = 0; k < ns; ++k) at += pw[at] >> 16;
  }
  const char* d1e = std::getenv("TQSV_DUAL1");
  const bool one_buf = mode == MODE_DUAL && write && all_slots &&
                       !(d1e && d1e[0] == '0');
  const bool prefetch = !pipe && (pfe && (pfe[0] == '0' || pfe[0] == '1')
                                      ? pfe[0] == '1'
                                      : (mode == MODE_RED && n >= 25));
  std::ostringstream o;
  // read-only passes: at least 2 waves per SIMD (at most 256 registers), as
  // the LDS of two work-groups per CU allows
  o << "#include \"tqsv_device.h\"\nusing namespace tqsv;\n"
    << "extern \"C\" __global__ void __launch_bounds__(" << NT << ") "
    << (mode == MODE_RED || pipe || one_buf ? "__attribute__((amdgpu_waves_per_eu(2))) " : "")
    << kernel_name(mode)
    << "(float2* __restrict__ psi, float2* __restrict__ lam, const uint32_t* __restrict__ prog_g,"
       " double* __restrict__ partials, uint32_t ntiles) {\n"
    << "  typedef " << (n > 32 ? "uint64_t" : "uint32_t") << " IDX;\n"
    << "  typedef TileMap<IDX, " << L;
  for (uint32_t k = 0; k < m; ++k) o << ", " << pw[PH_H + k];
  o << "> TM;\n"
    << "  constexpr int MODE = " << mode << ";\n"
    << "  __shared__ __align__(16) float2 tA[" << TSIZE << "];\n";
  if (mode == MODE_DUAL && !one_buf)
    o << "  __shared__ __align__(16) float2 tB[" << TSIZE << "];\n";
  else
    o << "  float2* tB = nullptr;\n";
  // read-only passes keep per-thread accumulators in registers; dual passes
  // accumulate per wave in LDS
  const bool racc = mode == MODE_RED && nred > 0;
  if (racc)
    o << "  double racc[" << nred << "];\n#pragma unroll\n  for (int k = 0; k < " << nred
      << "; ++k) racc[k] = 0.0;\n";
  else if (nred)
    o << "  __shared__ double acc[" << NWAVE * nred << "];\n";
  o << "  const cw_t prog = (cw_t)prog_g;\n"
    << "  const uint32_t tid = threadIdx.x;\n";
  if (nred && !racc) o << "  jit_zero_acc<" << NWAVE << ", " << nred << ">(acc, tid, " << NT << ");\n";
  const std::string fetch = "jit_fetch_tile<" + std::to_string(tb) + ", IDX, TM, " +
                            (nt_ld ? "true" : "false") + ">";
  const std::string put = "jit_put_tile<" + std::to_string(tb) + ">(tA, tid, nx);";
  if (pipe) {
    // the block's first tile goes to LDS, its second to registers
    o << "  float4 nx[kTileChunks];\n"
      << "  if (blockIdx.x < ntiles) {\n"
      << "    " << fetch << "(psi, TM::base(blockIdx.x), tid, nx);\n"
      << "    " << put << "\n"
      << "    if (blockIdx.x + gridDim.x < ntiles) " << fetch
      << "(psi, TM::base(blockIdx.x + gridDim.x), tid, nx);\n"
      << "  }\n"
      << "  __syncthreads();\n";
  }
  o << "  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {\n"
    << "    const IDX base = TM::base(tile);\n";
  const std::string ld1 = "jit_load_tile<" + std::to_string(tb) + ", MODE_GATE, IDX, TM, " +
                          (nt_ld ? "true" : "false") + ">";
  const std::string st1 = "jit_store_tile<" + std::to_string(tb) + ", MODE_GATE, IDX, TM, " +
                          (nt_st ? "true" : "false") + ">";
  if (one_buf)
    o << "    PhaseGeo<IDX> G, Gp;\n"
      << "    float2 a[NSLOT], b[NSLOT];\n"
      << "    __syncthreads();\n"
      << "    " << ld1 << "(psi, lam, tA, tB, base, tid);\n";
  else if (!pipe)
    o << "    __syncthreads();\n"
      << "    jit_load_tile<" << tb << ", MODE, IDX, TM, " << (nt_ld ? "true" : "false")
      << ">(psi, lam, tA, tB, base, tid);\n";
  if (prefetch)
    o << "    float pf[4] = {0.f, 0.f, 0.f, 0.f};\n"
      << "    if (tile + gridDim.x < ntiles)\n"
      << "      jit_touch_tile<" << tb << ", MODE, IDX, TM>(psi, lam, TM::base(tile + gridDim.x), tid, pf);\n";
  if (!pipe) o << "    __syncthreads();\n";
  // read-only passes write nothing to LDS in their phases: no barriers
  // between them
  size_t at = PH_SIZE;
  for (uint32_t ip = 0; ip < nphase; ++ip) {
    const uint32_t* ph = pw + at;
    const uint32_t nsteps = ph[PS_NSTEPS];
    const uint32_t used = phase_slots(pw, at + PS_SIZE, nsteps);
    const std::string um = used == 0xffffffffu ? "" : ", " + std::to_string(used) + "u";
    if (one_buf) {
      // psi's slots move to this phase's positions through the buffer, then
      // lambda's (at the first phase: psi's slots, then lambda's tile)
      o << "    {\n" << (ip ? "      Gp = G;\n" : "") << "      jit_geo<TM, IDX";
      for (int j = 0; j < RB; ++j) o << ", " << ph[PS_R + j];
      o << ">(G, base, tid);\n";
      if (ip)
        o << "      __syncthreads();\n"
          << "      jit_store_slots<MODE_GATE>(a, b, tA, tB, Gp);\n"
          << "      __syncthreads();\n"
          << "      jit_load_slots<MODE_GATE>(a, b, tA, tB, G);\n"
          << "      __syncthreads();\n"
          << "      jit_store_slots<MODE_GATE>(b, a, tA, tB, Gp);\n"
          << "      __syncthreads();\n"
          << "      jit_load_slots<MODE_GATE>(b, a, tA, tB, G);\n";
      else
        o << "      jit_load_slots<MODE_GATE>(a, b, tA, tB, G);\n"
          << "      __syncthreads();\n"
          << "      " << ld1 << "(lam, psi, tA, tB, base, tid);\n"
          << "      __syncthreads();\n"
          << "      jit_load_slots<MODE_GATE>(b, a, tA, tB, G);\n";
    } else {
      o << "    {\n      PhaseGeo<IDX> G;\n      jit_geo<TM, IDX";
      for (int j = 0; j < RB; ++j) o << ", " << ph[PS_R + j];
      o << ">(G, base, tid);\n"
        << "      float2 a[NSLOT], b[NSLOT];\n"
        << "      jit_load_slots<MODE" << um << ">(a, b, tA, tB, G);\n";
    }
    size_t ws = at + PS_SIZE;
    for (uint32_t is = 0; is < nsteps; ++is) {
      const uint32_t* w = pw + ws;
      const uint32_t op = w[0] & 0xffffu, nw = w[0] >> 16;
      const std::string P = "prog + " + std::to_string(ws);
      switch (op) {
        case OP_ROT:
          o << "      jrot<" << w[1] << ", " << w[2] << ", " << ((w[4] | w[5]) ? "true" : "false")
            << ", MODE>(a, b, " << P << ", G);\n";
          break;
        case OP_SCL: o << "      scl_step<MODE>(a, b, " << P << ");\n"; break;
        case OP_CXP:
          o << "      jcxp<" << w[1] << ", " << w[2] << ", " << w[3] << ", MODE>(a, b, " << P
            << ", G);\n";
          break;
        case OP_PH:
          o << "      jph<" << w[1] << ", " << w[2] << ", MODE>(a, b, " << P << ", G);\n";
          break;
        case OP_SGN: o << "      jsgn<" << w[1] << "u, MODE>(a, b);\n"; break;
        case OP_TAB: o << "      tab_step<MODE>(a, b, " << P << ");\n"; break;
        case OP_THR: o << "      jthr<" << w[1] << ", MODE>(a, b, " << P << ", G);\n"; break;
        case OP_U1:
          o << "      ju1<" << w[1] << ", " << w[2] << "u, " << (w[3] ? "true" : "false")
            << ", MODE>(a, b, " << P << ", G);\n";
          break;
        case OP_U2:
          o << "      ju2<" << w[1] << ", " << w[2] << ", " << (w[3] ? "true" : "false")
            << ", MODE>(a, b, " << P << ", G);\n";
          break;
        case OP_D2: o << "      d2_step<MODE, IDX>(a, b, " << P << ", G);\n"; break;
        case OP_PROT: {
          const uint32_t xr = w[1], k = w[6] & 3u;
          const uint32_t q = xr ? (k + 3u) & 3u : 3u;
          o << "      jprot<" << xr << ", " << q << ", MODE>(a, b, " << P << ", G);\n";
          break;
        }
        case OP_PGRP: {
          const uint64_t z0 = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
          o << "      jpgrp<" << w[1] << ", " << w[5] << "u, " << w[6] << "u, " << w[4] << "u, MODE>(a, b, "
            << P << ", G, (IDX)" << z0 << "ull);\n";
          break;
        }
        case OP_RED: {
          const uint32_t xr = w[1];
          if (racc) {
            const uint32_t kind = xr ? RED_PAIR : RED_DIAG;
            const uint32_t lb = xr & (~xr + 1u);
            o << "      {\n        float2 p[NSLOT];\n        jred_products<" << xr << ", " << kind
              << ">(p, a);\n";
            const uint32_t* e = w + RED_HDR;
            for (uint32_t t = 0; t < w[3]; ++t, e += RED_ENT) {
              const uint64_t zt = (uint64_t)e[1] | ((uint64_t)e[2] << 32);
              o << "        racc[" << e[0] << "] += jred_term<" << e[3] << "u, " << e[4] << ", "
                << lb << ", IDX>(p, G.gb, (IDX)" << zt << "ull);\n";
            }
            o << "      }\n";
            break;
          }
          if (mode == MODE_DUAL && w[2] == RED_GROUP) {
            const uint32_t* e = w + RED_HDR;
            const uint64_t z0 = (uint64_t)e[1] | ((uint64_t)e[2] << 32);
            o << "      jred_group<" << xr << ", " << e[RED_ENT + 1] << "u, " << e[3] << "u, "
              << e[RED_ENT] << "u, " << e[0] << ", " << nred << ", IDX>(a, b, " << P << ", G, (IDX)"
              << z0 << "ull, acc);\n";
          }
          else if (mode == MODE_DUAL) {
            const uint32_t* e = w + RED_HDR;
            const uint64_t zt = (uint64_t)e[1] | ((uint64_t)e[2] << 32);
            o << "      jred_cross<" << xr << ", " << e[4] << ", " << e[3] << "u, " << e[0] << ", "
              << nred << ", IDX>(a, b, G, (IDX)" << zt << "ull, acc);\n";
          }
          else
            o << "      red_x<" << xr << ", " << (xr ? RED_PAIR : RED_DIAG) << ", MODE, IDX>(a, b, "
              << P << ", G, acc, " << nred << ");\n";
          break;
        }
        default: break;
      }
      ws += nw;
    }
    if (one_buf)
      o << "    }\n";
    else if (write)
      o << "      jit_store_slots<MODE" << um << ">(a, b, tA, tB, G);\n      __syncthreads();\n    }\n";
    else
      o << "    }\n";
    at = ws;
  }
  if (one_buf)
    o << "    __syncthreads();\n"
      << "    jit_store_slots<MODE_GATE>(a, b, tA, tB, G);\n"
      << "    __syncthreads();\n"
      << "    " << st1 << "(psi, lam, tA, tB, base, tid);\n"
      << "    __syncthreads();\n"
      << "    jit_store_slots<MODE_GATE>(b, a, tA, tB, G);\n"
      << "    __syncthreads();\n"
      << "    " << st1 << "(lam, psi, tA, tB, base, tid);\n";
  if (pipe) {
    // finished tile out of LDS into registers; then (all reads done) the
    // next tile into LDS, the finished one to memory, the one after into
    // registers.  `more` is the same for the whole block.
    const std::string store = "jit_put_global<" + std::to_string(tb) + ", IDX, TM, " +
                              (nt_st ? "true" : "false") + ">";
    o << "    float4 out[kTileChunks];\n"
      << "    jit_get_tile<" << tb << ">(tA, tid, out);\n"
      << "    const bool more = tile + gridDim.x < ntiles;\n"
      << "    if (more) {\n      __syncthreads();\n      " << put << "\n    }\n"
      << "    " << store << "(psi, base, tid, out);\n"
      << "    if (tile + 2u * gridDim.x < ntiles) " << fetch
      << "(psi, TM::base(tile + 2u * gridDim.x), tid, nx);\n"
      << "    if (more) __syncthreads();\n";
  } else if (write && !one_buf) {
    o << "    jit_store_tile<" << tb << ", MODE, IDX, TM, " << (nt_st ? "true" : "false")
      << ">(psi, lam, tA, tB, base, tid);\n";
  }
  if (prefetch) o << "    __asm__ volatile(\"\" :: \"v\"(pf[0]), \"v\"(pf[1]), \"v\"(pf[2]), \"v\"(pf[3]));\n";
  o << "  }\n";
  if (racc)
    o << "  jred_finish<" << nred << ", " << NWAVE << ">(racc, reinterpret_cast<double*>(tA), partials, tid, "
      << NT << ");\n";
  else if (nred)
    o << "  jit_write_partials<" << NWAVE << ", " << nred << ">(acc, partials, tid, " << NT << ");\n";
  o << "}\n";
  return o.str();
}

// the hiprtc entry points the generator uses
struct Rtc {
  decltype(&hiprtcCreateProgram) create = nullptr;
  decltype(&hiprtcCompileProgram) compile = nullptr;
  decltype(&hiprtcGetProgramLogSize) log_size = nullptr;
  decltype(&hiprtcGetProgramLog) get_log = nullptr;
  decltype(&hiprtcGetCodeSize) code_size = nullptr;
  decltype(&hiprtcGetCode) get_code = nullptr;
  decltype(&hiprtcDestroyProgram) destroy = nullptr;
  std::string origin;
  // the environment pointer of the compiler namespace's own C library
  char*** ns_environ = nullptr;
};

template <typename F>
bool bind(void* h, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(h, name));
  return f != nullptr;
}

const Rtc& rtc() {
  static const Rtc r = [] {
    Rtc x;
    std::string path;
    if (const char* e = std::getenv("TQSV_HIPRTC")) {
      path = e;
    } else {
      const char* root = std::getenv("ROCM_PATH");
      path = std::string(root && root[0] ? root : "/opt/rocm") + "/lib/libhiprtc.so.7";
    }
    if (!path.empty() && path != "process") {
      if (void* h = dlmopen(LM_ID_NEWLM, path.c_str(), RTLD_NOW | RTLD_LOCAL)) {
        if (bind(h, "hiprtcCreateProgram", x.create) && bind(h, "hiprtcCompileProgram", x.compile) &&
            bind(h, "hiprtcGetProgramLogSize", x.log_size) && bind(h, "hiprtcGetProgramLog", x.get_log) &&
            bind(h, "hiprtcGetCodeSize", x.code_size) && bind(h, "hiprtcGetCode", x.get_code) &&
            bind(h, "hiprtcDestroyProgram", x.destroy)) {
          x.origin = path;
          x.ns_environ = reinterpret_cast<char***>(dlsym(h, "__environ"));
          return x;
        }
      }
    }
    x.create = &hiprtcCreateProgram;
    x.compile = &hiprtcCompileProgram;
    x.log_size = &hiprtcGetProgramLogSize;
    x.get_log = &hiprtcGetProgramLog;
    x.code_size = &hiprtcGetCodeSize;
    x.get_code = &hiprtcGetCode;
    x.destroy = &hiprtcDestroyProgram;
    x.origin = "process";
    return x;
  }();
  return r;
}

// The compiler namespace (dlmopen) has a C library of its own, whose
// environment pointer was copied from the process's when the namespace was
// loaded.  setenv in the process may since have moved that array and freed
// the old one, which the compiler's getenv calls would then read (a
// segmentation fault inside the next compile).  Before each batch of
// compiles the namespace is pointed at a private copy of the process's
// current environment.  Copies are kept for the life of the process (one per
// distinct environment, a few KB each): a compile on another thread may
// still read the previous one.
void sync_compiler_environment(const Rtc& R) {
  if (!R.ns_environ) return;
  struct EnvCopy {
    std::vector<std::string> strs;
    std::vector<char*> ptrs;
  };
  static std::mutex mu;
  static std::vector<std::unique_ptr<EnvCopy>> copies;
  std::lock_guard<std::mutex> lk(mu);
  std::vector<std::string> cur;
  for (char** e = environ; e && *e; ++e) cur.emplace_back(*e);
  if (copies.empty() || copies.back()->strs != cur) {
    auto c = std::make_unique<EnvCopy>();
    c->strs = std::move(cur);
    for (std::string& v : c->strs) c->ptrs.push_back(&v[0]);
    c->ptrs.push_back(nullptr);
    copies.push_back(std::move(c));
  }
  *R.ns_environ = copies.back()->ptrs.data();
}

bool compile_source(const std::string& src, std::vector<char>& code, std::string& log) {
  const Rtc& R = rtc();
  hiprtcProgram p;
  const char* hdrs[2] = {kJitHdrDevice, kJitHdrProg};
  const char* names[2] = {"tqsv_device.h", "tqsv_prog.h"};
  if (R.create(&p, src.c_str(), "tqsv_jit_pass.hip", 2, hdrs, names) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return false;
  }
  // no packed FP32 instructions in device code (as for the precompiled kernels)
  const char* opts[] = {"--offload-arch=gfx950", "-O3",          "-std=c++17",       "-Xclang",
                        "-target-feature",       "-Xclang",      "-packed-fp32-ops", "-w"};
  const hiprtcResult r = R.compile(p, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  R.log_size(p, &ls);
  log.assign(ls, '\0');
  if (ls) R.get_log(p, &log[0]);
  bool ok = r == HIPRTC_SUCCESS;
  if (ok) {
    size_t cs = 0;
    ok = R.code_size(p, &cs) == HIPRTC_SUCCESS && cs > 0;
    if (ok) {
      code.resize(cs);
      ok = R.get_code(p, code.data()) == HIPRTC_SUCCESS;
    }
  }
  R.destroy(&p);
  return ok;
}

}  // namespace

int jit_policy(int requested, int n, size_t tile_passes, size_t max_auto_passes) {
  if (const char* e = std::getenv("TQSV_JIT")) {
    if (e[0] == '0') return 0;
    if (e[0] == '1') return 1;
  }
  if (requested > 0) return 1;
  if (requested < 0) return 0;
  // automatic: states of 2^22+ amplitudes, and plans whose compile time
  // (about a second per distinct pass) stays moderate
  return n >= 22 && tile_passes <= max_auto_passes ? 1 : 0;
}

std::string jit_source(int tb, const uint32_t* pass_words, int n) { return gen_source(tb, pass_words, n); }

int jit_compile_only(const std::vector<PassPlan>& passes, const std::vector<uint32_t>& prog, int n,
                     int* compiled) {
  std::vector<std::string> srcs;
  {
    std::unordered_map<std::string, int> seen;
    for (const PassPlan& pp : passes) {
      if (!pp.tile) continue;
      std::string src = gen_source(pp.tb, prog.data() + pp.prog_off, n);
      if (seen.emplace(src, 1).second) srcs.push_back(std::move(src));
    }
  }
  // TQSV_JIT_DUMP=<directory>: also write each source there (for inspection)
  if (const char* dir = std::getenv("TQSV_JIT_DUMP"); dir && dir[0]) {
    for (size_t j = 0; j < srcs.size(); ++j) {
      const std::string path = std::string(dir) + "/tqsv_pass_" + std::to_string(j) + ".hip";
      if (FILE* f = std::fopen(path.c_str(), "w")) {
        std::fputs(srcs[j].c_str(), f);
        std::fclose(f);
      }
    }
  }
  std::vector<std::string> logs(srcs.size());
  std::vector<char> ok(srcs.size(), 0);
  sync_compiler_environment(rtc());
  std::atomic<size_t> next{0};
  auto work = [&]() {
    std::vector<char> code;
    for (size_t j; (j = next.fetch_add(1)) < srcs.size();) ok[j] = compile_source(srcs[j], code, logs[j]) ? 1 : 0;
  };
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nthr = std::min<size_t>(hw, srcs.size());
  std::vector<std::thread> th;
  for (size_t t = 1; t < nthr; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  for (size_t j = 0; j < srcs.size(); ++j)
    if (!ok[j]) return fail(TQSV_EDEVICE, "run-time compilation of a pass kernel failed: " + logs[j]);
  if (compiled) *compiled = (int)srcs.size();
  return 0;
}

int jit_prepare(std::vector<PassPlan>& passes, const std::vector<uint32_t>& prog, int n) {
  int dev = 0;
  TQSV_HIP(hipGetDevice(&dev));
  const std::string pre = std::to_string(dev) + "\n";
  std::vector<std::string> keys(passes.size());
  std::vector<size_t> todo;  // passes whose kernel is neither cached nor already queued
  {
    std::lock_guard<std::mutex> lk(g_mu);
    std::unordered_map<std::string, size_t> queued;
    for (size_t i = 0; i < passes.size(); ++i) {
      PassPlan& pp = passes[i];
      pp.jit_fn = nullptr;
      if (!pp.tile) continue;
      keys[i] = pre + gen_source(pp.tb, prog.data() + pp.prog_off, n);
      if (g_cache.count(keys[i]) || queued.count(keys[i])) continue;
      queued[keys[i]] = i;
      todo.push_back(i);
    }
  }
  if (!todo.empty()) {
    std::vector<std::vector<char>> codes(todo.size());
    std::vector<std::string> logs(todo.size());
    std::vector<char> ok(todo.size(), 0);
    sync_compiler_environment(rtc());
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t j; (j = next.fetch_add(1)) < todo.size();)
        ok[j] = compile_source(keys[todo[j]].substr(pre.size()), codes[j], logs[j]) ? 1 : 0;
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nthr = std::min<size_t>(hw, todo.size());
    std::vector<std::thread> th;
    for (size_t t = 1; t < nthr; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    for (size_t j = 0; j < todo.size(); ++j)
      if (!ok[j]) return fail(TQSV_EDEVICE, "run-time compilation of a pass kernel failed: " + logs[j]);
    std::lock_guard<std::mutex> lk(g_mu);
    for (size_t j = 0; j < todo.size(); ++j) {
      JitKernel k;
      TQSV_HIP(hipModuleLoadData(&k.mod, codes[j].data()));
      const uint32_t* pw = prog.data() + passes[todo[j]].prog_off;
      TQSV_HIP(hipModuleGetFunction(&k.fn, k.mod, kernel_name(pass_mode(pw[PH_FLAGS]))));
      if (const char* v = std::getenv("TQSV_JIT_VERBOSE"); v && v[0] == '1') {
        // pass shape and the compiler's resource use (scratch = register spills)
        int regs = 0, scratch = 0;
        (void)hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, k.fn);
        (void)hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k.fn);
        std::fprintf(stderr, "tqsv jit: pass %zu mode %u L %u m %u phases %u nred %u regs %d scratch %d (hiprtc %s)\n",
                     todo[j], pass_mode(pw[PH_FLAGS]), pw[PH_L], pw[PH_M], pw[PH_NPHASE], pw[PH_NRED],
                     regs, scratch, rtc().origin.c_str());
      }
      g_cache[keys[todo[j]]] = k;
    }
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (size_t i = 0; i < passes.size(); ++i)
    if (passes[i].tile) passes[i].jit_fn = g_cache[keys[i]].fn;
  return 0;
}

hipError_t jit_launch(const PassPlan& pp, const PassLaunch& pl, hipStream_t st) {
  float2* psi = pl.psi;
  float2* lam = pl.lam;
  const uint32_t* prog = pl.prog;
  double* partials = pl.partials;
  uint32_t ntiles = pl.ntiles;
  void* args[] = {&psi, &lam, &prog, &partials, &ntiles};
  return hipModuleLaunchKernel(static_cast<hipFunction_t>(pp.jit_fn), pl.grid, 1, 1, 1u << pp.tb,
                               1, 1, 0, st, args, nullptr);
}

}  // namespace tqsv
