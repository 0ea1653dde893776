// tqsv_batch.hip - energy and gradient of many small requests in one launch.
//
// A request of at most 13 qubits fits one work-group's LDS twice over (psi and
// lambda, 64 KB each), so its whole adjoint runs without leaving the
// work-group: the forward passes act on the LDS-resident state one after the
// other (every pass of a 13-qubit plan has the whole state as its tile, in
// the identity tile map, so the LDS layout is the same for all passes), then
// lambda = H|psi> is formed from LDS, the FP64 energy reduced, and the
// reverse passes run on (psi, lambda) with their <lambda|G|psi> reductions
// summed into the request's output slots.  One launch serves the whole batch,
// one work-group per request; requests of different sizes and depths share it.
#include <hip/hip_runtime.h>

#include "tqsv_internal.h"
#include "tqsv_pass_impl.h"

namespace tqsv {
namespace {

constexpr int BTB = BATCH_QUBITS - RB;  // 256 threads
constexpr int BNT = 1 << BTB;
constexpr uint32_t BTSIZE = 1u << BATCH_QUBITS;
constexpr uint32_t BNWAVE = BNT / 64;

template <bool LEAN>
__global__ __launch_bounds__(BNT) void batch_adjoint_kernel(const BatchItem* __restrict__ items) {
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tA = reinterpret_cast<float2*>(smem);
  float2* tB = tA + BTSIZE;
  // reduction accumulators of the reverse passes; before them, the energy's
  // per-wave sums
  double* acc = reinterpret_cast<double*>(smem + 2 * BTSIZE * sizeof(float2));
  double* esum = acc;
  const BatchItem it = items[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  const uint32_t dimu = 1u << it.nu;

  // psi into tile A (positions beyond the caller's state are zero)
  for (uint32_t p = 2u * tid; p < BTSIZE; p += 2u * BNT) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < dimu) v = reinterpret_cast<const float4*>(it.psi)[p >> 1];
    *reinterpret_cast<float4*>(tA + swz(p)) = v;
  }
  __syncthreads();

  for (uint32_t k = 0; k < it.nfwd; ++k) {
    const cw_t prog = (cw_t)(it.fprog + it.foff[k]);
    run_phases<BTB, MODE_GATE, uint32_t, LEAN>(prog, tA, tB, acc, 0u, 0u, prog[PH_L], prog[PH_M],
                                               prog + PH_H);
  }

  // lambda = H psi into tile B (the arithmetic of k_ham_apply)
  for (uint32_t i = tid; i < BTSIZE; i += BNT) {
    double ar = 0.0, ai = 0.0;
    for (uint32_t g = 0; g < it.ngroups; ++g) {
      const uint32_t j = i ^ (uint32_t)it.gx[g];
      const float2 v = tA[swz(j)];
      float fr = 0.f, fi = 0.f;
      for (uint32_t t = it.gstart[g]; t < it.gstart[g + 1]; ++t) {
        const float sg = (__popcll((uint64_t)j & it.tz[t]) & 1) ? -1.f : 1.f;
        fr += sg * it.tc[t].x;
        fi += sg * it.tc[t].y;
      }
      ar += (double)(fr * v.x) - (double)(fi * v.y);
      ai += (double)(fr * v.y) + (double)(fi * v.x);
    }
    tB[swz(i)] = make_float2((float)ar, (float)ai);
  }
  __syncthreads();

  // E = <psi|lambda> in FP64: per thread, per wave, then fixed order
  {
    double er = 0.0, ei = 0.0;
    for (uint32_t i = tid; i < BTSIZE; i += BNT) {
      const float2 a = tA[swz(i)], b = tB[swz(i)];
      er += (double)a.x * b.x + (double)a.y * b.y;
      ei += (double)a.x * b.y - (double)a.y * b.x;
    }
    er = wave_sum_d(er);
    ei = wave_sum_d(ei);
    if ((tid & 63u) == 0) {
      esum[2 * (tid >> 6)] = er;
      esum[2 * (tid >> 6) + 1] = ei;
    }
    __syncthreads();
    if (tid == 0) {
      double sr = 0.0, si = 0.0;
      for (uint32_t w = 0; w < BNWAVE; ++w) {
        sr += esum[2 * w];
        si += esum[2 * w + 1];
      }
      it.out[it.nred] = sr;
      it.out[it.nred + 1] = si;
    }
    __syncthreads();
  }

  for (uint32_t k = 0; k < it.nrev; ++k) {
    const cw_t prog = (cw_t)(it.rprog + it.roff[k]);
    const uint32_t nred = prog[PH_NRED];
    for (uint32_t i = tid; i < BNWAVE * nred; i += BNT) acc[i] = 0.0;
    __syncthreads();
    run_phases<BTB, MODE_DUAL, uint32_t, LEAN>(prog, tA, tB, acc, nred, 0u, prog[PH_L],
                                               prog[PH_M], prog + PH_H);
    for (uint32_t r = tid; r < nred; r += BNT) {
      double s = 0.0;
      for (uint32_t w = 0; w < BNWAVE; ++w) s += acc[w * nred + r];
      it.out[it.rred[k] + r] = s;
    }
    __syncthreads();
  }
}

template <bool LEAN>
hipError_t launch_batch(const BatchItem* items, uint32_t nreq, size_t lds, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_adjoint_kernel<LEAN>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL(batch_adjoint_kernel<LEAN>, dim3(nreq), dim3(BNT), lds, st, items);
  return hipGetLastError();
}

}  // namespace

hipError_t k_batch_adjoint(const BatchItem* items, uint32_t nreq, bool lean, uint32_t max_pass_nred,
                           hipStream_t st) {
  const size_t nacc = (size_t)BNWAVE * (max_pass_nred > 2u ? max_pass_nred : 2u);
  const size_t lds = 2 * BTSIZE * sizeof(float2) + nacc * sizeof(double);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  return lean ? launch_batch<true>(items, nreq, lds, st) : launch_batch<false>(items, nreq, lds, st);
}

}  // namespace tqsv
