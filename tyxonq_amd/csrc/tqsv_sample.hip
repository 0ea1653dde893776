// tqsv_sample.hip - shot sampling on the GPU.
//
// Replaces the host path of StatevectorEngine.run with shots > 0
// (devices/simulators/statevector/engine.py:377-465: |psi|^2 copied out,
// numpy multinomial, bincount, big-endian bitstrings).  Here |psi|^2 never
// leaves the device; only the sampled basis states and their counts do.
//
// Inverse-CDF sampling in two levels.  The state is cut into blocks of
// SBLK amplitudes:
//   1. FP64 sums of |psi_i|^2 over segments of 64 amplitudes (64 segments per
//      block) and over blocks, in one coalesced read; the exclusive prefix
//      P[b] of the block sums (P[nb] is the total mass).
//   2. Shot j draws u_j in [0, 1) from a counter-based Philox4x64-10 stream
//      (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2,
//      3", SC 2011): block g = j / 4 of the stream is Philox(counter = (g, 0,
//      0, 0), key = (seed, 0)), word j % 4, u_j = (word >> 11) * 2^-53.
//      x_j = u_j * P[nb]; the shot's block is the largest b with P[b] <= x_j
//      (zero-mass blocks skipped downwards).
//   3. Shots are bucketed by block (counting sort).  One wave per block with
//      shots scans the block's 64 segment sums; a shot's segment is chosen by
//      a ballot over the segment prefixes, its amplitude by a ballot over the
//      in-wave prefix of that segment's 64 probabilities (read from psi; zero
//      probabilities skipped downwards), counted in an LDS histogram and
//      written out as (index, count) pairs in index order.
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

// Segment sums (64 amplitudes each, 64 per block) and block sums (the 64
// segment sums reduced by wave 0), in one coalesced read: every lane loads two
// neighbouring amplitudes (16 bytes), so one wave load covers two segments,
// each reduced over its 32 lanes.
__device__ __forceinline__ double half_wave_sum_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  v += shfl_xor_d(v, 16);
  return v;
}

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(STHR) void seg_sums_kernel(const float2* __restrict__ psi,
                                                        uint32_t blk, double* __restrict__ segsum,
                                                        double* __restrict__ bsum) {
  constexpr uint32_t NW = STHR / 64;
  constexpr uint32_t ROUNDS = SBLK / 128 / NW;  // wave loads of 128 amplitudes
  __shared__ double segS[SBLK / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * blk;
#pragma unroll 4
  for (uint32_t r = 0; r < ROUNDS; ++r) {
    const uint32_t i = w * (SBLK / NW) + r * 128u + 2u * lane;
    double p = 0.0;
    if (i < blk) {
      // streaming load: the state is read once here; the later per-shot reads
      // of single segments are scattered
      const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(psi + base + i));
      p = prob(make_float2(v.x, v.y)) + prob(make_float2(v.z, v.w));
    }
    const double sg = half_wave_sum_d(p);
    if ((lane & 31u) == 0) segS[(w * (SBLK / NW) + r * 128u) / 64u + (lane >> 5)] = sg;
  }
  __syncthreads();
  if (w == 0) {
    const double sg = segS[lane];
    segsum[(uint64_t)blockIdx.x * 64u + lane] = sg;
    const double tot = wave_sum_d(sg);
    if (lane == 0) bsum[blockIdx.x] = tot;
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

// Scans of long arrays in three launches: sums of 1024-entry chunks, a
// one-work-group scan of the chunk sums, then each chunk's scan plus its
// offset.  Every thread handles 4 consecutive entries in both chunk kernels,
// so the order of every sum is fixed.
constexpr uint32_t SCH = 1024;

template <typename T>
__device__ __forceinline__ T wg_exclusive(T v, T* sh);

template <typename T>
__global__ __launch_bounds__(STHR) void chunk_sums_kernel(const T* __restrict__ in, uint32_t n,
                                                          T* __restrict__ csum) {
  __shared__ T sh[STHR];
  const uint32_t lo = min(blockIdx.x * SCH + threadIdx.x * 4u, n), hi = min(lo + 4u, n);
  T s = 0;
  for (uint32_t i = lo; i < hi; ++i) s += in[i];
  const T ex = wg_exclusive<T>(s, sh);
  if (threadIdx.x == STHR - 1) csum[blockIdx.x] = ex + s;
}

template <typename T>
__global__ __launch_bounds__(STHR) void chunk_scan_kernel(const T* __restrict__ in, uint32_t n,
                                                          const T* __restrict__ cpre,
                                                          T* __restrict__ out) {
  __shared__ T sh[STHR];
  const uint32_t lo = min(blockIdx.x * SCH + threadIdx.x * 4u, n), hi = min(lo + 4u, n);
  T s = 0;
  for (uint32_t i = lo; i < hi; ++i) s += in[i];
  T run = cpre[blockIdx.x] + wg_exclusive<T>(s, sh);
  for (uint32_t i = lo; i < hi; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (lo < hi && hi == n) out[n] = run;
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

__device__ __forceinline__ double shfl_up_d(double v, int d) {
  const int lo = __shfl_up(__double2loint(v), d, 64);
  const int hi = __shfl_up(__double2hiint(v), d, 64);
  return __hiloint2double(hi, lo);
}

// inclusive prefix over the 64 lanes (Hillis-Steele, fixed order)
__device__ __forceinline__ double wave_incl_scan_d(double v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double t = shfl_up_d(v, d);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

constexpr uint32_t DENSE_MIN = 32;  // blocks with more shots go to resolve_dense_kernel

// One wave per block with shots.  The block is 64 segments of 64 amplitudes
// (amplitudes beyond a short block count as zero).  Shot x: the segment is
// the first s with E[s] + S[s] > x (segment sums S, exclusive prefix E; the
// last segment if none); the amplitude is the first lane l with
// E[s] + (p_0 + .. + p_l) > x in that segment (lane 63 if none), moved down
// over zero probabilities.  Only the segments that shots land in are read.
// The LDS histogram is zeroed per 64-entry word on first use, and only the
// touched words are compacted.
__device__ __forceinline__ double seg_prob(const float2* __restrict__ psi, uint64_t base,
                                           uint32_t blk, uint32_t i) {
  return i < blk ? prob(psi[base + i]) : 0.0;
}

__global__ __launch_bounds__(64) void resolve_kernel(const float2* __restrict__ psi, uint32_t blk,
                                                     const double* __restrict__ segsum,
                                                     const uint32_t* __restrict__ boff,
                                                     const double* __restrict__ bx,
                                                     uint64_t* __restrict__ oidx,
                                                     uint32_t* __restrict__ ocnt,
                                                     uint32_t* __restrict__ nuniq,
                                                     uint32_t* __restrict__ dense_list,
                                                     uint32_t* __restrict__ dense_n) {
  constexpr uint32_t NWORD = SBLK / 64;
  __shared__ uint32_t cnt[SBLK];
  __shared__ uint64_t touched[NWORD];
  const uint32_t b = blockIdx.x, lane = threadIdx.x;
  const uint32_t s0 = boff[b], s1 = boff[b + 1];
  if (s1 - s0 > DENSE_MIN) {  // left to resolve_dense_kernel
    if (lane == 0) dense_list[atomicAdd(dense_n, 1u)] = b;
    return;
  }
  if (s0 == s1) {
    if (lane == 0) nuniq[b] = 0;
    return;
  }
  const uint64_t base = (uint64_t)b * blk;
  const double S = segsum[(uint64_t)b * 64u + lane];
  const double inc = wave_incl_scan_d(S, lane);
  const double ex = shfl_up_d(inc, 1);
  const double E = lane ? ex : 0.0;
  const double I = E + S;
  touched[lane] = 0ull;
  for (uint32_t j = s0; j < s1; ++j) {
    const double x = bx[j];
    const uint64_t m1 = __ballot(I > x);
    const uint32_t seg = m1 ? (uint32_t)__builtin_ctzll(m1) : NWORD - 1u;
    const double Es = __shfl(E, (int)seg, 64);
    const double p = seg_prob(psi, base, blk, seg * 64u + lane);
    const double c = Es + wave_incl_scan_d(p, lane);
    const uint64_t m2 = __ballot(c > x);
    const uint32_t l = m2 ? (uint32_t)__builtin_ctzll(m2) : 63u;
    const uint64_t upto = l == 63u ? ~0ull : ((2ull << l) - 1ull);
    const uint64_t nz = __ballot(p != 0.0) & upto;
    uint32_t k;
    if (nz) {
      k = seg * 64u + (63u - (uint32_t)__builtin_clzll(nz));
    } else {
      k = seg * 64u;
      while (k > 0u) {
        --k;
        if (seg_prob(psi, base, blk, k) != 0.0) break;
      }
    }
    const uint32_t word = k >> 6;
    if (touched[word] == 0ull) cnt[word * 64u + lane] = 0u;
    if (lane == 0) {
      cnt[k] += 1u;
      touched[word] |= 1ull << (k & 63u);
    }
  }
  uint32_t running = 0;
  for (uint32_t r = 0; r < NWORD; ++r) {
    const uint64_t tw = touched[r];
    if (!tw) continue;
    if ((tw >> lane) & 1ull) {
      const uint32_t pos = s0 + running + (uint32_t)__popcll(tw & ((1ull << lane) - 1ull));
      oidx[pos] = base + r * 64u + lane;
      ocnt[pos] = cnt[r * 64u + lane];
    }
    running += (uint32_t)__popcll(tw);
  }
  if (lane == 0) nuniq[b] = running;
}

// Blocks with more than DENSE_MIN shots: the same rule, with the block's
// probabilities and in-segment prefixes formed once in LDS (by the same wave
// scans, so the values equal those of resolve_kernel) and the shots spread
// over four waves.
__global__ __launch_bounds__(STHR) void resolve_dense_kernel(
    const float2* __restrict__ psi, uint32_t blk, const double* __restrict__ segsum,
    const uint32_t* __restrict__ boff, const double* __restrict__ bx, uint64_t* __restrict__ oidx,
    uint32_t* __restrict__ ocnt, uint32_t* __restrict__ nuniq,
    const uint32_t* __restrict__ dense_list, const uint32_t* __restrict__ dense_n) {
  constexpr uint32_t NSEG = SBLK / 64;
  constexpr uint32_t NW = STHR / 64;
  __shared__ double pl[SBLK];
  __shared__ double cl[SBLK];
  __shared__ uint32_t cnt[SBLK];
  __shared__ double segE[NSEG], segI[NSEG];
  __shared__ uint32_t wtot[NW];
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint32_t ndense = *dense_n;
  for (uint32_t q = blockIdx.x; q < ndense; q += gridDim.x) {
    __syncthreads();
    const uint32_t b = dense_list[q];
    const uint32_t s0 = boff[b], s1 = boff[b + 1];
    const uint64_t base = (uint64_t)b * blk;
    if (w == 0) {
      const double S = segsum[(uint64_t)b * 64u + lane];
      const double inc = wave_incl_scan_d(S, lane);
      const double ex = shfl_up_d(inc, 1);
      const double E = lane ? ex : 0.0;
      segE[lane] = E;
      segI[lane] = E + S;
    }
    for (uint32_t i = t; i < SBLK; i += STHR) cnt[i] = 0u;
    __syncthreads();
    for (uint32_t sg = w; sg < NSEG; sg += NW) {
      const double p = seg_prob(psi, base, blk, sg * 64u + lane);
      pl[sg * 64u + lane] = p;
      cl[sg * 64u + lane] = segE[sg] + wave_incl_scan_d(p, lane);
    }
    __syncthreads();
    const double I = segI[lane];
    for (uint32_t j = s0 + w; j < s1; j += NW) {
      const double x = bx[j];
      const uint64_t m1 = __ballot(I > x);
      const uint32_t seg = m1 ? (uint32_t)__builtin_ctzll(m1) : NSEG - 1u;
      const double p = pl[seg * 64u + lane];
      const uint64_t m2 = __ballot(cl[seg * 64u + lane] > x);
      const uint32_t l = m2 ? (uint32_t)__builtin_ctzll(m2) : 63u;
      const uint64_t upto = l == 63u ? ~0ull : ((2ull << l) - 1ull);
      const uint64_t nz = __ballot(p != 0.0) & upto;
      uint32_t k;
      if (nz) {
        k = seg * 64u + (63u - (uint32_t)__builtin_clzll(nz));
      } else {
        k = seg * 64u;
        while (k > 0u) {
          --k;
          if (pl[k] != 0.0) break;
        }
      }
      if (lane == 0) atomicAdd(&cnt[k], 1u);
    }
    __syncthreads();
    uint32_t running = 0;
    for (uint32_t r = 0; r < SBLK / STHR; ++r) {
      const uint32_t i = r * STHR + t;
      const uint32_t cv = cnt[i];
      const uint64_t m = __ballot(cv != 0u);
      if (lane == 0) wtot[w] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t off = running, tot = 0;
      for (uint32_t q = 0; q < NW; ++q) {
        if (q < w) off += wtot[q];
        tot += wtot[q];
      }
      if (cv) {
        const uint32_t pos = s0 + off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        oidx[pos] = base + i;
        ocnt[pos] = cv;
      }
      running += tot;
      __syncthreads();
    }
    if (t == 0) nuniq[b] = running;
  }
}

// gather the per-block (index, count) runs contiguously
__global__ __launch_bounds__(STHR) void compact_kernel(const uint32_t* __restrict__ boff,
                                                       const uint32_t* __restrict__ uoff,
                                                       const uint64_t* __restrict__ oidx,
                                                       const uint32_t* __restrict__ ocnt,
                                                       uint64_t* __restrict__ fidx,
                                                       int64_t* __restrict__ fcnt) {
  const uint32_t b = blockIdx.x;
  const uint32_t s0 = boff[b], u0 = uoff[b], nu = uoff[b + 1] - uoff[b];
  for (uint32_t k = threadIdx.x; k < nu; k += STHR) {
    fidx[u0 + k] = oidx[s0 + k];
    fcnt[u0 + k] = (int64_t)ocnt[s0 + k];
  }
}

}  // namespace
}  // namespace tqsv

struct tqsv_sampler {
  int n = 0;
  uint64_t max_shots = 0;
  uint32_t blk = 0, nb = 0;
  double* bsum = nullptr;      // nb
  double* segsum = nullptr;    // nb * 64
  double* prefix = nullptr;    // nb + 1
  uint32_t* bcount = nullptr;  // nb
  uint32_t* boff = nullptr;    // nb + 1
  uint32_t* bfill = nullptr;   // nb
  uint32_t* nuniq = nullptr;   // nb
  uint32_t* uoff = nullptr;    // nb + 1
  uint32_t* dlist = nullptr;   // nb: blocks for the dense resolve
  uint32_t* dn = nullptr;      // 1
  uint32_t* sblock = nullptr;  // max_shots
  double* sx = nullptr;        // max_shots
  double* bx = nullptr;        // max_shots
  uint64_t* oidx = nullptr;    // max_shots
  uint32_t* ocnt = nullptr;    // max_shots
  uint64_t* fidx = nullptr;    // max_shots
  int64_t* fcnt = nullptr;     // max_shots
  double* csum_d = nullptr;    // chunk sums / prefixes of the scans
  double* cpre_d = nullptr;
  uint32_t* csum_u = nullptr;
  uint32_t* cpre_u = nullptr;
};

namespace {

void free_sampler(tqsv_sampler* s) {
  if (!s) return;
  void* ptrs[] = {s->bsum, s->segsum, s->prefix, s->bcount, s->boff, s->bfill, s->nuniq, s->uoff,
                  s->sblock, s->sx, s->bx, s->oidx, s->ocnt, s->fidx, s->fcnt,
                  s->csum_d, s->cpre_d, s->csum_u, s->cpre_u, s->dlist, s->dn};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

uint32_t nchunks(uint32_t n) { return (n + tqsv::SCH - 1) / tqsv::SCH; }

// exclusive scan of in[0..n) into out[0..n] (out[n] = total)
template <typename T>
hipError_t scan(const T* in, uint32_t n, T* out, T* csum, T* cpre, hipStream_t st) {
  using namespace tqsv;
  const uint32_t nc = nchunks(n);
  hipLaunchKernelGGL(chunk_sums_kernel<T>, dim3(nc), dim3(STHR), 0, st, in, n, csum);
  hipLaunchKernelGGL(scan_kernel<T>, dim3(1), dim3(1024), 0, st, csum, nc, cpre);
  hipLaunchKernelGGL(chunk_scan_kernel<T>, dim3(nc), dim3(STHR), 0, st, in, n, cpre, out);
  return hipGetLastError();
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
  if (e == hipSuccess) e = dalloc(&s->segsum, nb * 64);
  if (e == hipSuccess) e = dalloc(&s->prefix, nb + 1);
  if (e == hipSuccess) e = dalloc(&s->bcount, nb);
  if (e == hipSuccess) e = dalloc(&s->boff, nb + 1);
  if (e == hipSuccess) e = dalloc(&s->bfill, nb);
  if (e == hipSuccess) e = dalloc(&s->nuniq, nb);
  if (e == hipSuccess) e = dalloc(&s->uoff, nb + 1);
  if (e == hipSuccess) e = dalloc(&s->dlist, nb);
  if (e == hipSuccess) e = dalloc(&s->dn, 1);
  if (e == hipSuccess) e = dalloc(&s->sblock, m);
  if (e == hipSuccess) e = dalloc(&s->sx, m);
  if (e == hipSuccess) e = dalloc(&s->bx, m);
  if (e == hipSuccess) e = dalloc(&s->oidx, m);
  if (e == hipSuccess) e = dalloc(&s->ocnt, m);
  if (e == hipSuccess) e = dalloc(&s->fidx, m);
  if (e == hipSuccess) e = dalloc(&s->fcnt, m);
  if (e == hipSuccess) e = dalloc(&s->csum_d, nchunks(s->nb));
  if (e == hipSuccess) e = dalloc(&s->cpre_d, nchunks(s->nb) + 1);
  if (e == hipSuccess) e = dalloc(&s->csum_u, nchunks(s->nb));
  if (e == hipSuccess) e = dalloc(&s->cpre_u, nchunks(s->nb) + 1);
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
    hipLaunchKernelGGL(seg_sums_kernel, dim3(nb), dim3(STHR), 0, st, a, s->blk, s->segsum,
                       s->bsum);
    TQSV_HIP(hipGetLastError());
    TQSV_HIP(scan<double>(s->bsum, nb, s->prefix, s->csum_d, s->cpre_d, st));
    hipLaunchKernelGGL(zero_u32_kernel, dim3(zg), dim3(256), 0, st, s->bcount, nb);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(zg), dim3(256), 0, st, s->bfill, nb);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(256), 0, st, s->dn, 1u);
    TQSV_HIP(hipGetLastError());
    const uint64_t ng = (ns + 3u) / 4u;
    hipLaunchKernelGGL(shot_blocks_kernel, dim3((uint32_t)((ng + STHR - 1) / STHR)), dim3(STHR), 0,
                       st, s->prefix, nb, ns, seed, (uint64_t)0, s->sblock, s->sx, s->bcount);
    TQSV_HIP(hipGetLastError());
    TQSV_HIP(scan<uint32_t>(s->bcount, nb, s->boff, s->csum_u, s->cpre_u, st));
    hipLaunchKernelGGL(scatter_kernel, dim3((uint32_t)((ns + 255u) / 256u)), dim3(256), 0, st,
                       s->sblock, s->sx, ns, s->boff, s->bfill, s->bx);
    TQSV_HIP(hipGetLastError());
    hipLaunchKernelGGL(resolve_kernel, dim3(nb), dim3(64), 0, st, a, s->blk, s->segsum, s->boff,
                       s->bx, s->oidx, s->ocnt, s->nuniq, s->dlist, s->dn);
    hipLaunchKernelGGL(resolve_dense_kernel, dim3(std::min<uint32_t>(nb, 512u)), dim3(STHR), 0, st,
                       a, s->blk, s->segsum, s->boff, s->bx, s->oidx, s->ocnt, s->nuniq, s->dlist,
                       s->dn);
    TQSV_HIP(hipGetLastError());
    TQSV_HIP(scan<uint32_t>(s->nuniq, nb, s->uoff, s->csum_u, s->cpre_u, st));
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
    TQSV_HIP(hipMemcpyAsync(index, s->fidx, nu * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    TQSV_HIP(hipMemcpyAsync(count, s->fcnt, nu * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    TQSV_HIP(hipStreamSynchronize(st));
    *nunique = (int64_t)nu;
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_sampler_run: out of host memory");
  }
}

void tqsv_sampler_destroy(tqsv_sampler* s) { free_sampler(s); }

}  // extern "C"
