// tqsv_density.hip - reductions of a vectorised density matrix.
//
// A density matrix of n qubits is held as a 2n-qubit state vec(rho) with
// vec[(r << n) | c] = rho[r, c] (row index in the high bits), so gates and
// channels run through the statevector plans (U on the row bits, conj(U) on
// the column bits; a channel as its 4x4 superoperator on a row/column bit
// pair).  What is specific to density matrices is only reading them out:
//   Tr(rho P) = i^{|x&z|} sum_r (-1)^{|r&z|} rho[r, r ^ x]
// for P with flip mask x and phase mask z (P|r> = i^{|x&z|} (-1)^{|r&z|}
// |r ^ x>), and the diagonal as the input of the shot sampler.
// Replaces exp_z_density and the trace / diag readouts of
// libs/quantum_library/kernels/density_matrix.py:66-73 and
// devices/simulators/density_matrix/engine.py:103-151.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "../../include/tqsv.h"
#include "tqsv_device.h"
#include "tqsv_internal.h"

namespace tqsv {
namespace {

constexpr int DTHR = 256;

__global__ __launch_bounds__(DTHR) void dm_pauli_kernel(const float2* __restrict__ rho, uint32_t n,
                                                        uint64_t x, uint64_t z,
                                                        double* __restrict__ partials) {
  __shared__ double ws[2 * DTHR / 64];
  const uint64_t dim = 1ull << n;
  double sr = 0.0, si = 0.0;
  for (uint64_t r = (uint64_t)blockIdx.x * DTHR + threadIdx.x; r < dim;
       r += (uint64_t)gridDim.x * DTHR) {
    const float2 v = rho[(r << n) | (r ^ x)];
    const double s = (__popcll(r & z) & 1) ? -1.0 : 1.0;
    sr += s * (double)v.x;
    si += s * (double)v.y;
  }
  sr = wave_sum_d(sr);
  si = wave_sum_d(si);
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) {
    ws[2 * w] = sr;
    ws[2 * w + 1] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < DTHR / 64; ++k) {
      a += ws[2 * k];
      b += ws[2 * k + 1];
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
}

__global__ void dm_diag_amp_kernel(const float2* __restrict__ rho, uint32_t n,
                                   float2* __restrict__ out) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (1ull << n)) return;
  const float p = rho[(r << n) | r].x;
  out[r] = make_float2(sqrtf(fmaxf(p, 0.f)), 0.f);
}

uint32_t dm_grid(int n) {
  const uint64_t dim = 1ull << n;
  const uint64_t g = (dim + DTHR - 1) / DTHR;
  return (uint32_t)(g < 1024 ? g : 1024);
}

}  // namespace
}  // namespace tqsv

extern "C" {

int tqsv_dm_expval(const void* rho, int n, int nterms, const uint64_t* xmask,
                   const uint64_t* zmask, const double* coeff, double* terms, double* value,
                   void* stream) {
  using namespace tqsv;
  if (!rho || nterms < 0 || (nterms > 0 && (!xmask || !zmask)))
    return fail(TQSV_EINVAL, "tqsv_dm_expval: NULL argument");
  if (n < 1 || n > 16) return fail(TQSV_EINVAL, "tqsv_dm_expval: n must be in [1, 16]");
  const uint64_t full = (1ull << n) - 1ull;
  for (int t = 0; t < nterms; ++t)
    if ((xmask[t] | zmask[t]) & ~full) return fail(TQSV_EINVAL, "tqsv_dm_expval: mask out of range");
  if (nterms == 0) {
    if (value) *value = 0.0;
    return 0;
  }
  try {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const uint32_t grid = dm_grid(n);
    double* dpart = nullptr;
    double* dout = nullptr;
    TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&dpart), sizeof(double) * 2 * grid));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dout), sizeof(double) * 2 * (size_t)nterms);
    if (e != hipSuccess) {
      (void)hipFree(dpart);
      return hip_fail(e, "tqsv_dm_expval: hipMalloc");
    }
    for (int t = 0; t < nterms && e == hipSuccess; ++t) {
      hipLaunchKernelGGL(dm_pauli_kernel, dim3(grid), dim3(DTHR), 0, st,
                         reinterpret_cast<const float2*>(rho), (uint32_t)n, xmask[t], zmask[t], dpart);
      e = hipGetLastError();
      if (e == hipSuccess) e = k_final_sum(dpart, grid, 2, dout + 2 * t, st);
    }
    std::vector<double> h(2 * (size_t)nterms);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dpart);
    (void)hipFree(dout);
    if (e != hipSuccess) return hip_fail(e, "tqsv_dm_expval");
    double v = 0.0;
    for (int t = 0; t < nterms; ++t) {
      const double re = h[2 * t], im = h[2 * t + 1];
      double tv;
      switch (__builtin_popcountll(xmask[t] & zmask[t]) & 3) {  // Re(i^k (re + i im))
        case 0: tv = re; break;
        case 1: tv = -im; break;
        case 2: tv = -re; break;
        default: tv = im; break;
      }
      if (terms) terms[t] = tv;
      v += (coeff ? coeff[t] : 1.0) * tv;
    }
    if (value) *value = v;
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_dm_expval: out of host memory");
  }
}

int tqsv_dm_diag_amplitudes(const void* rho, int n, void* out, void* stream) {
  using namespace tqsv;
  if (!rho || !out) return fail(TQSV_EINVAL, "tqsv_dm_diag_amplitudes: NULL argument");
  if (n < 1 || n > 16) return fail(TQSV_EINVAL, "tqsv_dm_diag_amplitudes: n must be in [1, 16]");
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(dm_diag_amp_kernel, dim3((uint32_t)((dim + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const float2*>(rho),
                     (uint32_t)n, reinterpret_cast<float2*>(out));
  TQSV_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
