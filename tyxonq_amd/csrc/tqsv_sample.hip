// tqsv_sample.hip - shot sampling on the GPU.
//
// Replaces the host path of StatevectorEngine.run with shots > 0
// (devices/simulators/statevector/engine.py:377-465: |psi|^2 copied out,
// numpy multinomial, bincount, big-endian bitstrings).  Here |psi|^2 never
// leaves the device; only the sampled basis states and their counts do.
//
// Inverse-CDF sampling in two levels.  The state is cut into blocks of
// SBLK amplitudes:
//   1. FP64 block sums of |psi_i|^2 and their exclusive prefix P[b] (P[nb] is
//      the total mass).
//   2. Shot j draws u_j in [0, 1) from a counter-based Philox4x64-10 stream
//      (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2,
//      3", SC 2011): block g = j / 4 of the stream is Philox(counter = (g, 0,
//      0, 0), key = (seed, 0)), word j % 4, u_j = (word >> 11) * 2^-53.
//      x_j = u_j * P[nb]; the shot's block is the largest b with P[b] <= x_j
//      (zero-mass blocks skipped downwards).
//   3. Shots are bucketed by block (counting sort).  One work-group per block
//      with shots forms the block's inclusive FP64 prefix c[k] in LDS; the
//      shot's amplitude is the smallest k with c[k] > x_j - P[b] (clamped to
//      the block, zero-width amplitudes skipped downwards), counted in an LDS
//      histogram and written out as (index, count) pairs in index order.
// The result depends only on psi, the shot count and the seed, not on thread
// scheduling: the atomics only count, and every sum has a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/tqsv.h"
#include "tqsv_device.h"
#include "tqsv_internal.h"

namespace tqsv {
namespace {

constexpr int SB = 12;
constexpr uint32_t SBLK = 1u << SB;  // amplitudes per sampling block
constexpr int STHR = 256;            // threads per work-group

struct U4 {
  uint64_t v[4];
};

__device__ __forceinline__ U4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                            uint64_t k0, uint64_t k1) {
  constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t lo0 = M0 * c0, hi0 = __umul64hi(M0, c0);
    const uint64_t lo1 = M1 * c2, hi1 = __umul64hi(M1, c2);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  return {{c0, c1, c2, c3}};
}

__device__ __forceinline__ double prob(float2 a) {
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

__global__ __launch_bounds__(STHR) void block_sums_kernel(const float2* __restrict__ psi,
                                                          uint32_t blk, double* __restrict__ bsum) {
  __shared__ double wsum[STHR / 64];
  const uint64_t base = (uint64_t)blockIdx.x * blk;
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < blk; i += STHR) s += prob(psi[base + i]);
  s = wave_sum_d(s);
  if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < STHR / 64; ++w) t += wsum[w];
    bsum[blockIdx.x] = t;
  }
}

// exclusive prefix of in[0..n) into out[0..n], out[n] = total (one work-group;
// each thread scans a contiguous segment sequentially)
template <typename T>
__global__ __launch_bounds__(1024) void scan_kernel(const T* __restrict__ in, uint32_t n,
                                                    T* __restrict__ out) {
  __shared__ T sh[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (n + 1023u) / 1024u;
  const uint32_t lo = min(t * per, n), hi = min(lo + per, n);
  T s = 0;
  for (uint32_t i = lo; i < hi; ++i) s += in[i];
  sh[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < 1024u; d <<= 1) {
    const T v = t >= d ? sh[t - d] : (T)0;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  T run = t ? sh[t - 1] : (T)0;
  for (uint32_t i = lo; i < hi; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (lo < hi && hi == n) out[n] = run;
  if (n == 0 && t == 0) out[0] = 0;
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ __launch_bounds__(STHR) void shot_blocks_kernel(const double* __restrict__ prefix,
                                                           uint32_t nb, uint64_t shots,
                                                           uint64_t k0, uint64_t k1,
                                                           uint32_t* __restrict__ sblock,
                                                           double* __restrict__ sx,
                                                           uint32_t* __restrict__ bcount) {
  const uint64_t g = (uint64_t)blockIdx.x * STHR + threadIdx.x;
  if (g * 4u >= shots) return;
  const U4 r = philox4x64_10(g, 0, 0, 0, k0, k1);
  const double total = prefix[nb];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint64_t j = g * 4u + (uint64_t)w;
    if (j >= shots) break;
    const double u = (double)(r.v[w] >> 11) * 0x1.0p-53;
    const double x = u * total;
    uint32_t lo = 0, hi = nb;  // prefix[lo] <= x; answer in [lo, hi)
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (prefix[mid] <= x)
        lo = mid;
      else
        hi = mid;
    }
    while (lo > 0u && prefix[lo + 1] == prefix[lo]) --lo;
    sblock[j] = lo;
    sx[j] = x - prefix[lo];
    atomicAdd(&bcount[lo], 1u);
  }
}

__global__ void scatter_kernel(const uint32_t* __restrict__ sblock, const double* __restrict__ sx,
                               uint64_t shots, const uint32_t* __restrict__ boff,
                               uint32_t* __restrict__ bfill, double* __restrict__ bx) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= shots) return;
  const uint32_t b = sblock[j];
  const uint32_t slot = boff[b] + atomicAdd(&bfill[b], 1u);
  bx[slot] = sx[j];
}

// work-group exclusive scan of one value per thread (STHR threads)
template <typename T>
__device__ __forceinline__ T wg_exclusive(T v, T* sh) {
  const uint32_t t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (uint32_t d = 1; d < (uint32_t)STHR; d <<= 1) {
    const T w = t >= d ? sh[t - d] : (T)0;
    __syncthreads();
    sh[t] += w;
    __syncthreads();
  }
  const T r = t ? sh[t - 1] : (T)0;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(STHR) void resolve_kernel(const float2* __restrict__ psi, uint32_t blk,
                                                       const uint32_t* __restrict__ boff,
                                                       const double* __restrict__ bx,
                                                       uint64_t* __restrict__ oidx,
                                                       uint32_t* __restrict__ ocnt,
                                                       uint32_t* __restrict__ nuniq) {
  __shared__ double c[SBLK];
  __shared__ uint32_t cnt[SBLK];
  __shared__ double dsh[STHR];
  __shared__ uint32_t ush[STHR];
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  const uint32_t s0 = boff[b], s1 = boff[b + 1];
  if (s0 == s1) {
    if (t == 0) nuniq[b] = 0;
    return;
  }
  const uint64_t base = (uint64_t)b * blk;
  const uint32_t per = (blk + STHR - 1) / STHR;
  const uint32_t lo = min(t * per, blk), hi = min(lo + per, blk);
  double run = 0.0;
  for (uint32_t i = lo; i < hi; ++i) {
    run += prob(psi[base + i]);
    c[i] = run;
    cnt[i] = 0u;
  }
  const double off = wg_exclusive<double>(run, dsh);
  for (uint32_t i = lo; i < hi; ++i) c[i] += off;
  __syncthreads();
  for (uint32_t j = s0 + t; j < s1; j += STHR) {
    const double x = bx[j];
    uint32_t l = 0, h = blk - 1;  // smallest k with c[k] > x, or blk - 1
    while (l < h) {
      const uint32_t mid = (l + h) >> 1;
      if (c[mid] > x)
        h = mid;
      else
        l = mid + 1;
    }
    while (l > 0u && c[l] == c[l - 1]) --l;
    atomicAdd(&cnt[l], 1u);
  }
  __syncthreads();
  uint32_t nz = 0;
  for (uint32_t i = lo; i < hi; ++i) nz += cnt[i] ? 1u : 0u;
  uint32_t pos = s0 + wg_exclusive<uint32_t>(nz, ush);
  for (uint32_t i = lo; i < hi; ++i)
    if (cnt[i]) {
      oidx[pos] = base + i;
      ocnt[pos] = cnt[i];
      ++pos;
    }
  // the thread that ends the block knows the block's number of states
  if (lo < hi && hi == blk) nuniq[b] = pos - s0;
}

// gather the per-block (index, count) runs contiguously
__global__ __launch_bounds__(STHR) void compact_kernel(const uint32_t* __restrict__ boff,
                                                       const uint32_t* __restrict__ uoff,
                                                       const uint64_t* __restrict__ oidx,
                                                       const uint32_t* __restrict__ ocnt,
                                                       uint64_t* __restrict__ fidx,
                                                       uint32_t* __restrict__ fcnt) {
  const uint32_t b = blockIdx.x;
  const uint32_t s0 = boff[b], u0 = uoff[b], nu = uoff[b + 1] - uoff[b];
  for (uint32_t k = threadIdx.x; k < nu; k += STHR) {
    fidx[u0 + k] = oidx[s0 + k];
    fcnt[u0 + k] = ocnt[s0 + k];
  }
}

}  // namespace
}  // namespace tqsv

struct tqsv_sampler {
  int n = 0;
  uint64_t max_shots = 0;
  uint32_t blk = 0, nb = 0;
  double* bsum = nullptr;      // nb
  double* prefix = nullptr;    // nb + 1
  uint32_t* bcount = nullptr;  // nb
  uint32_t* boff = nullptr;    // nb + 1
  uint32_t* bfill = nullptr;   // nb
  uint32_t* nuniq = nullptr;   // nb
  uint32_t* uoff = nullptr;    // nb + 1
  uint32_t* sblock = nullptr;  // max_shots
  double* sx = nullptr;        // max_shots
  double* bx = nullptr;        // max_shots
  uint64_t* oidx = nullptr;    // max_shots
  uint32_t* ocnt = nullptr;    // max_shots
  uint64_t* fidx = nullptr;    // max_shots
  uint32_t* fcnt = nullptr;    // max_shots
  std::vector<uint32_t> hcnt;
};

namespace {

void free_sampler(tqsv_sampler* s) {
  if (!s) return;
  void* ptrs[] = {s->bsum, s->prefix, s->bcount, s->boff, s->bfill, s->nuniq, s->uoff,
                  s->sblock, s->sx, s->bx, s->oidx, s->ocnt, s->fidx, s->fcnt};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

extern "C" {

int tqsv_sampler_create(int n, int64_t max_shots, tqsv_sampler** out) {
  using namespace tqsv;
  if (!out) return fail(TQSV_EINVAL, "tqsv_sampler_create: NULL output");
  *out = nullptr;
  if (n < 1 || n > 40) return fail(TQSV_EINVAL, "tqsv_sampler_create: n must be in [1, 40]");
  if (max_shots < 1 || max_shots > 0x7fffffffLL)
    return fail(TQSV_EINVAL, "tqsv_sampler_create: max_shots must be in [1, 2^31 - 1]");
  tqsv_sampler* s = new (std::nothrow) tqsv_sampler();
  if (!s) return fail(TQSV_ENOMEM, "tqsv_sampler_create: out of host memory");
  s->n = n;
  s->max_shots = (uint64_t)max_shots;
  s->blk = n >= SB ? SBLK : (1u << n);
  const uint64_t nb = (1ull << n) / s->blk;
  if (nb > (1ull << 28)) {
    delete s;
    return fail(TQSV_EINVAL, "tqsv_sampler_create: state too large");
  }
  s->nb = (uint32_t)nb;
  const size_t m = (size_t)max_shots;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = dalloc(&s->bsum, nb);
  if (e == hipSuccess) e = dalloc(&s->prefix, nb + 1);
  if (e == hipSuccess) e = dalloc(&s->bcount, nb);
  if (e == hipSuccess) e = dalloc(&s->boff, nb + 1);
  if (e == hipSuccess) e = dalloc(&s->bfill, nb);
  if (e == hipSuccess) e = dalloc(&s->nuniq, nb);
  if (e == hipSuccess) e = dalloc(&s->uoff, nb + 1);
  if (e == hipSuccess) e = dalloc(&s->sblock, m);
  if (e == hipSuccess) e = dalloc(&s->sx, m);
  if (e == hipSuccess) e = dalloc(&s->bx, m);
  if (e == hipSuccess) e = dalloc(&s->oidx, m);
  if (e == hipSuccess) e = dalloc(&s->ocnt, m);
  if (e == hipSuccess) e = dalloc(&s->fidx, m);
  if (e == hipSuccess) e = dalloc(&s->fcnt, m);
  if (e != hipSuccess) {
    free_sampler(s);
    return hip_fail(e, "tqsv_sampler_create: hipMalloc");
  }
  *out = s;
  return 0;
}

int tqsv_sampler_run(tqsv_sampler* s, const void* psi, int64_t shots, uint64_t seed,
                     uint64_t* index, int64_t* count, int64_t* nunique, void* stream) {
  using namespace tqsv;
  if (!s || !psi || !index || !count || !nunique)
    return fail(TQSV_EINVAL, "tqsv_sampler_run: NULL argument");
  *nunique = 0;
  if (shots < 0 || (uint64_t)shots > s->max_shots)
    return fail(TQSV_EINVAL, "tqsv_sampler_run: shots must be in [0, max_shots]");
  if (shots == 0) return 0;
  try {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const float2* a = reinterpret_cast<const float2*>(psi);
    const uint32_t nb = s->nb;
    const uint64_t ns = (uint64_t)shots;
    const uint32_t zg = (nb + 255u) / 256u;
    hipLaunchKernelGGL(block_sums_kernel, dim3(nb), dim3(STHR), 0, st, a, s->blk, s->bsum);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_kernel<double>, dim3(1), dim3(1024), 0, st, s->bsum, nb, s->prefix);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(zero_u32_kernel, dim3(zg), dim3(256), 0, st, s->bcount, nb);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(zg), dim3(256), 0, st, s->bfill, nb);
    TQSV_HIP(hipGetLastError());
    const uint64_t ng = (ns + 3u) / 4u;
    hipLaunchKernelGGL(shot_blocks_kernel, dim3((uint32_t)((ng + STHR - 1) / STHR)), dim3(STHR), 0,
                       st, s->prefix, nb, ns, seed, (uint64_t)0, s->sblock, s->sx, s->bcount);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_kernel<uint32_t>, dim3(1), dim3(1024), 0, st, s->bcount, nb, s->boff);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(scatter_kernel, dim3((uint32_t)((ns + 255u) / 256u)), dim3(256), 0, st,
                       s->sblock, s->sx, ns, s->boff, s->bfill, s->bx);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(resolve_kernel, dim3(nb), dim3(STHR), 0, st, a, s->blk, s->boff, s->bx,
                       s->oidx, s->ocnt, s->nuniq);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_kernel<uint32_t>, dim3(1), dim3(1024), 0, st, s->nuniq, nb, s->uoff);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(compact_kernel, dim3(nb), dim3(STHR), 0, st, s->boff, s->uoff, s->oidx,
                       s->ocnt, s->fidx, s->fcnt);
    TQSV_HIP(hipGetLastError());
    double total = 0.0;
    uint32_t nu = 0;
    TQSV_HIP(hipMemcpyAsync(&total, s->prefix + nb, sizeof(double), hipMemcpyDeviceToHost, st));
    TQSV_HIP(hipMemcpyAsync(&nu, s->uoff + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TQSV_HIP(hipStreamSynchronize(st));
    if (!(total > 0.0) || !std::isfinite(total))
      return fail(TQSV_EINVAL, "tqsv_sampler_run: the state has zero or non-finite norm");
    if (nu > ns) return fail(TQSV_EDEVICE, "tqsv_sampler_run: inconsistent histogram");
    s->hcnt.resize(nu);
    TQSV_HIP(hipMemcpyAsync(index, s->fidx, nu * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    TQSV_HIP(hipMemcpyAsync(s->hcnt.data(), s->fcnt, nu * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            st));
    TQSV_HIP(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < nu; ++i) count[i] = (int64_t)s->hcnt[i];
    *nunique = (int64_t)nu;
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_sampler_run: out of host memory");
  }
}

void tqsv_sampler_destroy(tqsv_sampler* s) { free_sampler(s); }

}  // extern "C"
