// tqsv_kernels.hip - single-operation kernels and FP64 reductions.
//
// These serve the one-gate entry points of the C ABI (the drop-ins for the
// reference's apply_1q/apply_2q/expect_z kernels) and the unfused launch
// mode of a plan.  The fused tile passes live in tqsv_pass.hip.
#include <hip/hip_runtime.h>

#include "tqsv_internal.h"

namespace tqsv {

namespace {

constexpr int BLK = 256;

__device__ __forceinline__ uint64_t ins0(uint64_t x, uint32_t b) {
  return ((x >> b) << (b + 1)) | (x & ((1ull << b) - 1ull));
}

__device__ __forceinline__ uint32_t par(uint32_t v) { return __popc(v) & 1u; }
__device__ __forceinline__ uint32_t par(uint64_t v) { return __popcll(v) & 1u; }

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 mv2(const Mat2f& u, int r, float2 x0, float2 x1) {
  const float* m = u.v + 4 * r;
  return make_float2(m[0] * x0.x - m[1] * x0.y + m[2] * x1.x - m[3] * x1.y,
                     m[0] * x0.y + m[1] * x0.x + m[2] * x1.y + m[3] * x1.x);
}

uint32_t grid_for(uint64_t work) {
  uint64_t g = (work + BLK - 1) / BLK;
  if (g > 65535ull * 16ull) g = 65535ull * 16ull;
  if (g == 0) g = 1;
  return (uint32_t)g;
}

__global__ void init_kernel(float2* psi, uint64_t dim, uint64_t index) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
    psi[i] = make_float2(i == index ? 1.f : 0.f, 0.f);
}

// one thread per amplitude pair, controlled by (cm, cv)
__global__ void apply_1q_kernel(float2* psi, uint64_t npairs, uint32_t bit, uint64_t cm,
                                uint64_t cv, Mat2f u) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npairs; t += stride) {
    const uint64_t i0 = ins0(t, bit);
    if ((i0 & cm) != cv) continue;
    const uint64_t i1 = i0 | (1ull << bit);
    const float2 x0 = psi[i0], x1 = psi[i1];
    psi[i0] = mv2(u, 0, x0, x1);
    psi[i1] = mv2(u, 1, x0, x1);
  }
}

__global__ void apply_2q_kernel(float2* psi, uint64_t nquads, uint32_t bh, uint32_t bl, uint32_t blo,
                                uint32_t bhi, uint64_t cm, uint64_t cv, Mat4f u) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nquads; t += stride) {
    const uint64_t i00 = ins0(ins0(t, blo), bhi);
    if ((i00 & cm) != cv) continue;
    const uint64_t idx[4] = {i00, i00 | (1ull << bl), i00 | (1ull << bh),
                             i00 | (1ull << bh) | (1ull << bl)};
    float2 x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = psi[idx[c]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float yr = 0.f, yi = 0.f;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float mr = u.v[8 * r + 2 * c], mi = u.v[8 * r + 2 * c + 1];
        yr += mr * x[c].x - mi * x[c].y;
        yi += mr * x[c].y + mi * x[c].x;
      }
      psi[idx[r]] = make_float2(yr, yi);
    }
  }
}

struct DiagArgs {
  int nbits;
  int bits[4];
  float ph[32];
};

__global__ void apply_diag_kernel(float2* psi, uint64_t dim, DiagArgs d) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    uint32_t k = 0;
    for (int j = 0; j < d.nbits; ++j) k = (k << 1) | (uint32_t)((i >> d.bits[j]) & 1ull);
    psi[i] = cmulf(psi[i], make_float2(d.ph[2 * k], d.ph[2 * k + 1]));
  }
}

// exp(-i theta/2 P): new[i] = c psi[i] - i s omega (-1)^{|(i^x)&z|} psi[i^x]
__global__ void pauli_rot_kernel(float2* psi, uint64_t dim, uint64_t x, uint64_t z, float c,
                                 float s, int k) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // kappa = -i * i^k
  const int q = (k + 3) & 3;
  const float kr = (q == 0) ? 1.f : (q == 2 ? -1.f : 0.f);
  const float ki = (q == 1) ? 1.f : (q == 3 ? -1.f : 0.f);
  if (x == 0) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
      const float sg = (__popcll(i & z) & 1) ? -s : s;
      const float2 v = psi[i];
      // (c + sg * kappa) v with kappa = -i for k = 0
      const float2 f = make_float2(c + sg * kr, sg * ki);
      psi[i] = cmulf(f, v);
    }
    return;
  }
  const uint64_t lb = x & (~x + 1ull);
  const uint32_t lbit = (uint32_t)__ffsll((unsigned long long)lb) - 1u;
  const uint64_t npairs = dim >> 1;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npairs; t += stride) {
    const uint64_t i0 = ins0(t, lbit), i1 = i0 ^ x;
    const float s0 = (__popcll(i0 & z) & 1) ? -s : s;
    const float s1 = (__popcll(i1 & z) & 1) ? -s : s;
    const float2 a0 = psi[i0], a1 = psi[i1];
    const float2 r1 = cmulf(make_float2(kr, ki), a1), r0 = cmulf(make_float2(kr, ki), a0);
    psi[i0] = make_float2(c * a0.x + s1 * r1.x, c * a0.y + s1 * r1.y);
    psi[i1] = make_float2(c * a1.x + s0 * r0.x, c * a1.y + s0 * r0.y);
  }
}

__global__ void project_kernel(float2* psi, uint64_t dim, uint32_t bit, uint32_t keep) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
    if (((i >> bit) & 1ull) != keep) psi[i] = make_float2(0.f, 0.f);
}

__global__ void scale_kernel(float2* psi, uint64_t dim, float re, float im) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
    psi[i] = cmulf(psi[i], make_float2(re, im));
}

__global__ void probs_kernel(const float2* psi, uint64_t dim, double* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    const double re = psi[i].x, im = psi[i].y;
    out[i] = re * re + im * im;
  }
}

__device__ double block_sum(double v) {
  __shared__ double red[BLK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < BLK / 64; ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(BLK) void reduce_prob_kernel(const float2* psi, uint64_t dim, int mode,
                                                          uint32_t bit, double* partials) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  float facc = 0.f;
  int cnt = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    const float2 v = psi[i];
    float p = v.x * v.x + v.y * v.y;
    const uint32_t b = (uint32_t)((i >> bit) & 1ull);
    if (mode == 1) p = b ? -p : p;
    if (mode == 2) p = b ? p : 0.f;
    facc += p;
    if (++cnt == 64) {  // bounded FP32 runs, accumulated in FP64
      acc += (double)facc;
      facc = 0.f;
      cnt = 0;
    }
  }
  acc += (double)facc;
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(BLK) void final_sum_kernel(const double* partials, uint32_t ngroups,
                                                        uint32_t nred, double* out) {
  const uint32_t r = blockIdx.x;
  double acc = 0.0;
  for (uint32_t g = threadIdx.x; g < ngroups; g += BLK) acc += partials[(size_t)g * nred + r];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[r] = s;
}

// one block per output slot r: the sum over the slot's blocks of partials
// (entry r of `desc`: {offset, blocks, stride, index}: partials[offset + b *
// stride + index] for b < blocks), in a fixed order
__global__ __launch_bounds__(BLK) void final_sum_slots_kernel(const double* partials,
                                                              const uint32_t* desc, double* out) {
  const uint32_t r = blockIdx.x;
  const uint32_t off = desc[4 * r], nb = desc[4 * r + 1], stride = desc[4 * r + 2],
                 k = desc[4 * r + 3];
  double acc = 0.0;
  for (uint32_t g = threadIdx.x; g < nb; g += BLK) acc += partials[(size_t)off + (size_t)g * stride + k];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[r] = s;
}

// block partial sums of (-1)^{|i&z|} conj(u[i]) v[i^x]
__global__ __launch_bounds__(BLK) void pauli_dot_kernel(const float2* u, const float2* v, uint64_t dim,
                                                        uint64_t x, uint64_t z, double* partials) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double ar = 0.0, ai = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
    const float2 a = u[i], b = v[i ^ x];
    double pr = (double)a.x * b.x + (double)a.y * b.y;
    double pi = (double)a.x * b.y - (double)a.y * b.x;
    if (__popcll(i & z) & 1) {
      pr = -pr;
      pi = -pi;
    }
    ar += pr;
    ai += pi;
  }
  const double sr = block_sum(ar);
  __syncthreads();
  const double si = block_sum(ai);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sr;
    partials[2 * blockIdx.x + 1] = si;
  }
}

// one block per term: full sums for small states
__global__ __launch_bounds__(BLK) void pauli_dot_terms_kernel(const float2* u, const float2* v,
                                                              uint64_t dim, const uint64_t* dx,
                                                              const uint64_t* dz, double* out2) {
  const uint32_t t = blockIdx.x;
  const uint64_t x = dx[t], z = dz[t];
  double ar = 0.0, ai = 0.0;
  for (uint64_t i = threadIdx.x; i < dim; i += BLK) {
    const float2 a = u[i], b = v[i ^ x];
    double pr = (double)a.x * b.x + (double)a.y * b.y;
    double pi = (double)a.x * b.y - (double)a.y * b.x;
    if (__popcll(i & z) & 1) {
      pr = -pr;
      pi = -pi;
    }
    ar += pr;
    ai += pi;
  }
  const double sr = block_sum(ar);
  __syncthreads();
  const double si = block_sum(ai);
  if (threadIdx.x == 0) {
    out2[2 * t] = sr;
    out2[2 * t + 1] = si;
  }
}

// out[i] = sum over flip groups x of c_x(i) psi[i ^ x], c_x(j) = sum_t c_t i^k_t
// (-1)^{|j & z_t|} over the group's terms; one thread per output amplitude.
// Terms that share x and the Z mask outside x (zc) come as one table
// subgroup: c(j) = (-1)^{|j & zc|} tab[pattern of j on the bits of x], one
// coefficient load instead of one parity and sign per term.  Subgroups that
// share a flip mask form a block with one partner load, pattern and complex
// product (the 14251 terms of a 20-spin-orbital Jordan-Wigner Hamiltonian
// are 2445 subgroups alone on their 4-bit flip masks, 90 blocks of 19
// subgroups on 2-bit flip masks, and 211 diagonal terms summed into a vector
// beforehand).  Headers and term arrays are read-only and indexed
// uniformly across the wave, so they come through scalar loads; signs are an
// XOR of the sign bit with a parity (32-bit indices up to 32 qubits).
// OFF32 (states of at most 2^29 amplitudes): the table part gathers through
// 32-bit byte offsets from the uniform base addresses, without 64-bit address
// arithmetic per load.
template <typename IDX, bool OFF32 = false>
__global__ void ham_apply_kernel(const float2* __restrict__ psi, float2* __restrict__ out,
                                 uint64_t dim, uint32_t nsub, const uint32_t* __restrict__ sub,
                                 const float2* __restrict__ tab, uint32_t nblk,
                                 const uint32_t* __restrict__ blk,
                                 const uint32_t* __restrict__ mem, uint32_t ncblk,
                                 const uint32_t* __restrict__ cblk,
                                 const uint32_t* __restrict__ chk, uint32_t ngroups,
                                 const uint64_t* __restrict__ gx,
                                 const uint32_t* __restrict__ gstart,
                                 const uint64_t* __restrict__ tz, const float2* __restrict__ tc,
                                 const float* __restrict__ diag, double* __restrict__ epart) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double er = 0.0, ei = 0.0;  // <psi|H psi> of this thread's amplitudes (with epart)
  // Work-groups are dispatched round-robin over the 8 XCDs (block id mod 8).
  // Block b takes the b/8-th block of the (b mod 8)-th eighth of the index
  // range, so each XCD works on one eighth.  The subgroups and groups come in
  // ascending x, so all of an XCD's work-groups read partners from one eighth
  // of the state at a time (j = i ^ x, top bits of x varying slowest), which
  // its L2 can hold for states up to 2^22 amplitudes.
  uint64_t b = blockIdx.x;
  if ((gridDim.x & 7u) == 0) b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
  for (uint64_t i0 = b * blockDim.x + threadIdx.x; i0 < dim; i0 += stride) {
    const IDX i = (IDX)i0;
    double ar = 0.0, ai = 0.0;  // FP64 across subgroups and groups, one rounding at the end
    if constexpr (sizeof(IDX) == 4) {
      // table part: per subgroup one partner load, the pattern of the partner
      // on the (at most 4) flip bits: bit k = bit k of j >> sh_k, sh_k =
      // position of the k-th flip bit - k, packed 5 bits each; unused
      // positions hold 31, whose bit k >= 1 is always 0 for 32-bit indices),
      // one coefficient load and the sign of the Z chain.  Four subgroups per
      // round: all eight loads are issued before the arithmetic of the first,
      // so their latencies overlap.
      auto pattern = [](uint32_t j, uint32_t, uint32_t pk) {
        return ((j >> (pk & 31u)) & 1u) | ((j >> ((pk >> 5) & 31u)) & 2u) |
               ((j >> ((pk >> 10) & 31u)) & 4u) | ((j >> ((pk >> 15) & 31u)) & 8u);
      };
      auto gather = [](const float2* base, uint32_t k) -> float2 {
        if constexpr (OFF32)
          return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) + (k << 3));
        else
          return base[k];
      };
      uint32_t g = 0;
      for (; g + 4 <= nsub; g += 4) {
        uint32_t jj[4], zc[4];
        float2 v[4], c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t* d = sub + 4 * (g + u);
          jj[u] = (uint32_t)i ^ d[0];
          zc[u] = d[1];
          v[u] = gather(psi, jj[u]);
          c[u] = gather(tab, d[2] + pattern(jj[u], d[0], d[3]));
        }
        // the round's four products are summed in FP32 (as exact as the
        // products themselves), the rounds in FP64
        float sr = 0.f, si = 0.f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t neg = par(jj[u] & zc[u]) << 31;
          const float re = c[u].x * v[u].x - c[u].y * v[u].y;
          const float im = c[u].x * v[u].y + c[u].y * v[u].x;
          sr += __uint_as_float(__float_as_uint(re) ^ neg);
          si += __uint_as_float(__float_as_uint(im) ^ neg);
        }
        ar += (double)sr;
        ai += (double)si;
      }
      for (; g < nsub; ++g) {
        const uint32_t* d = sub + 4 * g;
        const uint32_t j = (uint32_t)i ^ d[0];
        const float2 v = gather(psi, j);
        const float2 c = gather(tab, d[2] + pattern(j, d[0], d[3]));
        const uint32_t neg = par(j & d[1]) << 31;
        const float re = c.x * v.x - c.y * v.y, im = c.x * v.y + c.y * v.x;
        ar += (double)__uint_as_float(__float_as_uint(re) ^ neg);
        ai += (double)__uint_as_float(__float_as_uint(im) ^ neg);
      }
      // subgroups sharing a flip mask: one partner load and pattern, the
      // members' signed coefficients summed (FP32), one complex product
      for (uint32_t bb = 0; bb < nblk; ++bb) {
        const uint32_t* hd = blk + 4 * bb;
        const uint32_t j = (uint32_t)i ^ hd[0];
        const float2 v = gather(psi, j);
        const uint32_t q = pattern(j, hd[0], hd[1]);
        float cr = 0.f, ci = 0.f;
        const uint32_t m1 = hd[3];
#pragma unroll 4
        for (uint32_t m = hd[2]; m < m1; ++m) {
          const uint32_t* md = mem + 2 * m;
          const float2 c = gather(tab, md[1] + q);
          const uint32_t neg = par(j & md[0]) << 31;
          cr += __uint_as_float(__float_as_uint(c.x) ^ neg);
          ci += __uint_as_float(__float_as_uint(c.y) ^ neg);
        }
        ar += (double)(cr * v.x - ci * v.y);
        ai += (double)(cr * v.y + ci * v.x);
      }
      // blocks in chunk form: the members differ from the base Z mask z0 in
      // one bit each, so their signed sum is (-1)^{|j & z0|} times a sum of
      // chunk tables, each indexed by the pattern and 4 consecutive bits of j
      for (uint32_t bb = 0; bb < ncblk; ++bb) {
        const uint32_t* hd = cblk + 5 * bb;
        const uint32_t j = (uint32_t)i ^ hd[0];
        const float2 v = gather(psi, j);
        const uint32_t qs = pattern(j, hd[0], hd[1]) << 4;
        float cr = 0.f, ci = 0.f;
        const uint32_t c1 = hd[4];
#pragma unroll 4
        for (uint32_t cc = hd[3]; cc < c1; ++cc) {
          const uint32_t* cd = chk + 2 * cc;
          const float2 c = gather(tab, cd[1] + qs + ((j >> cd[0]) & 15u));
          cr += c.x;
          ci += c.y;
        }
        const uint32_t neg = par(j & hd[2]) << 31;
        const float re = cr * v.x - ci * v.y, im = cr * v.y + ci * v.x;
        ar += (double)__uint_as_float(__float_as_uint(re) ^ neg);
        ai += (double)__uint_as_float(__float_as_uint(im) ^ neg);
      }
    }
    if (diag) {
      // the diagonal terms (x = 0), summed per index beforehand
      const float d = diag[i0];
      const float2 v = psi[i];
      ar += (double)(d * v.x);
      ai += (double)(d * v.y);
    }
    for (uint32_t g = 0; g < ngroups; ++g) {
      const IDX j = i ^ (IDX)gx[g];
      const float2 v = psi[j];
      float fr = 0.f, fi = 0.f;
      const uint32_t t1 = gstart[g + 1];
#pragma unroll 4
      for (uint32_t t = gstart[g]; t < t1; ++t) {
        const uint32_t neg = par((IDX)(j & (IDX)tz[t])) << 31;
        const float2 c = tc[t];
        fr += __uint_as_float(__float_as_uint(c.x) ^ neg);
        fi += __uint_as_float(__float_as_uint(c.y) ^ neg);
      }
      ar += (double)(fr * v.x) - (double)(fi * v.y);
      ai += (double)(fr * v.y) + (double)(fi * v.x);
    }
    out[i0] = make_float2((float)ar, (float)ai);
    if (epart) {
      // the energy from the unrounded (H psi)_i
      const float2 p = psi[i];
      er += (double)p.x * ar + (double)p.y * ai;
      ei += (double)p.x * ai - (double)p.y * ar;
    }
  }
  if (epart) {
    const double sr = block_sum(er);
    __syncthreads();
    const double si = block_sum(ei);
    if (threadIdx.x == 0) {
      epart[2 * blockIdx.x] = sr;
      epart[2 * blockIdx.x + 1] = si;
    }
  }
}

// out[i] = sum_j U[row(i)][j] psi[base(i) | off[j]] for a k-qubit matrix: row
// is i's target bits (first target most significant), base is i with those
// bits cleared, off[j] the target bits of column j.  One thread per output
// amplitude, FP32 sums (out of place).
template <typename IDX>
__global__ void apply_kq_kernel(const float2* __restrict__ psi, float2* __restrict__ out,
                                uint64_t dim, int k, const uint32_t* __restrict__ bitpos,
                                const uint64_t* __restrict__ off, const float2* __restrict__ U) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint32_t K = 1u << k;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < dim; i0 += stride) {
    const IDX i = (IDX)i0;
    uint32_t row = 0;
    IDX base = i;
    for (int b = 0; b < k; ++b) {
      row = (row << 1) | (uint32_t)((i >> bitpos[b]) & 1u);
      base &= ~(((IDX)1) << bitpos[b]);
    }
    const float2* u = U + (size_t)row * K;
    float re = 0.f, im = 0.f;
    for (uint32_t j = 0; j < K; ++j) {
      const float2 v = psi[base | (IDX)off[j]];
      re += u[j].x * v.x - u[j].y * v.y;
      im += u[j].x * v.y + u[j].y * v.x;
    }
    out[i0] = make_float2(re, im);
  }
}

__global__ void pad_copy_kernel(const float2* __restrict__ src, uint64_t nsrc,
                                float2* __restrict__ dst, uint64_t ndst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ndst) dst[i] = i < nsrc ? src[i] : make_float2(0.f, 0.f);
}

}  // namespace

hipError_t k_pad_copy(const float2* src, uint64_t nsrc, float2* dst, uint64_t ndst,
                      hipStream_t st) {
  hipLaunchKernelGGL(pad_copy_kernel, dim3((uint32_t)((ndst + BLK - 1) / BLK)), dim3(BLK), 0, st,
                     src, nsrc, dst, ndst);
  return hipGetLastError();
}

hipError_t k_apply_kq(const float2* psi, float2* out, int n, int k, const uint32_t* bitpos,
                      const uint64_t* off, const float2* U, hipStream_t st) {
  const uint64_t dim = 1ull << n;
  if (n <= 32)
    hipLaunchKernelGGL(apply_kq_kernel<uint32_t>, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, out,
                       dim, k, bitpos, off, U);
  else
    hipLaunchKernelGGL(apply_kq_kernel<uint64_t>, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, out,
                       dim, k, bitpos, off, U);
  return hipGetLastError();
}

hipError_t k_init_basis(float2* psi, int n, uint64_t index, hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(init_kernel, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, dim, index);
  return hipGetLastError();
}

hipError_t k_apply_1q(float2* psi, int n, int bit, uint64_t cmask, uint64_t cval, const Mat2f& u,
                      hipStream_t st) {
  const uint64_t np = 1ull << (n - 1);
  hipLaunchKernelGGL(apply_1q_kernel, dim3(grid_for(np)), dim3(BLK), 0, st, psi, np, (uint32_t)bit,
                     cmask, cval, u);
  return hipGetLastError();
}

hipError_t k_apply_2q(float2* psi, int n, int bh, int bl, uint64_t cmask, uint64_t cval,
                      const Mat4f& u, hipStream_t st) {
  const uint64_t nq = 1ull << (n - 2);
  const uint32_t blo = (uint32_t)(bh < bl ? bh : bl), bhi = (uint32_t)(bh < bl ? bl : bh);
  hipLaunchKernelGGL(apply_2q_kernel, dim3(grid_for(nq)), dim3(BLK), 0, st, psi, nq, (uint32_t)bh,
                     (uint32_t)bl, blo, bhi, cmask, cval, u);
  return hipGetLastError();
}

hipError_t k_apply_diag(float2* psi, int n, int nbits, const int* bits, const float* phases,
                        hipStream_t st) {
  DiagArgs d{};
  d.nbits = nbits;
  for (int j = 0; j < nbits; ++j) d.bits[j] = bits[j];
  for (int k = 0; k < (2 << nbits); ++k) d.ph[k] = phases[k];
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(apply_diag_kernel, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, dim, d);
  return hipGetLastError();
}

hipError_t k_pauli_rot(float2* psi, int n, uint64_t x, uint64_t z, float c, float s, int k,
                       hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(pauli_rot_kernel, dim3(grid_for(x ? dim / 2 : dim)), dim3(BLK), 0, st, psi,
                     dim, x, z, c, s, k);
  return hipGetLastError();
}

hipError_t k_project_zero(float2* psi, int n, int bit, int keep, hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(project_kernel, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, dim,
                     (uint32_t)bit, (uint32_t)keep);
  return hipGetLastError();
}

hipError_t k_scale(float2* psi, int n, float re, float im, hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, dim, re, im);
  return hipGetLastError();
}

hipError_t k_probs(const float2* psi, int n, double* out, hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(probs_kernel, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, dim, out);
  return hipGetLastError();
}

uint32_t reduce_grid(int n) {
  const uint64_t dim = 1ull << n;
  uint64_t g = (dim + BLK * 16 - 1) / (BLK * 16);
  if (g > 2048) g = 2048;
  if (g == 0) g = 1;
  return (uint32_t)g;
}

hipError_t k_reduce_prob(const float2* psi, int n, int mode, int bit, double* partials,
                         hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(reduce_prob_kernel, dim3(reduce_grid(n)), dim3(BLK), 0, st, psi, dim, mode,
                     (uint32_t)bit, partials);
  return hipGetLastError();
}

hipError_t k_final_sum(const double* partials, uint32_t ngroups, uint32_t nred, double* out,
                       hipStream_t st) {
  if (nred == 0) return hipSuccess;
  hipLaunchKernelGGL(final_sum_kernel, dim3(nred), dim3(BLK), 0, st, partials, ngroups, nred, out);
  return hipGetLastError();
}

hipError_t k_final_sum_slots(const double* partials, const uint32_t* desc, uint32_t nslots,
                             double* out, hipStream_t st) {
  if (nslots == 0) return hipSuccess;
  hipLaunchKernelGGL(final_sum_slots_kernel, dim3(nslots), dim3(BLK), 0, st, partials, desc, out);
  return hipGetLastError();
}

hipError_t k_pauli_dot(const float2* u, const float2* v, int n, uint64_t x, uint64_t z,
                       double* partials, hipStream_t st) {
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(pauli_dot_kernel, dim3(reduce_grid(n)), dim3(BLK), 0, st, u, v, dim, x, z,
                     partials);
  return hipGetLastError();
}

hipError_t k_pauli_dot_terms(const float2* u, const float2* v, int n, uint32_t nterms,
                             const uint64_t* dx, const uint64_t* dz, double* out2,
                             hipStream_t st) {
  if (nterms == 0) return hipSuccess;
  const uint64_t dim = 1ull << n;
  hipLaunchKernelGGL(pauli_dot_terms_kernel, dim3(nterms), dim3(BLK), 0, st, u, v, dim, dx, dz,
                     out2);
  return hipGetLastError();
}

// diag[i] = sum_t c_t (-1)^{|i & z_t|}: the diagonal (x = 0) terms of a Pauli
// sum per index, FP64 sums rounded once
template <typename IDX>
__global__ void diag_build_kernel(float* __restrict__ diag, uint64_t dim, uint32_t nt,
                                  const uint64_t* __restrict__ z, const double* __restrict__ c) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < dim; i0 += stride) {
    const IDX i = (IDX)i0;
    double s = 0.0;
    for (uint32_t t = 0; t < nt; ++t) s += par((IDX)(i & (IDX)z[t])) ? -c[t] : c[t];
    diag[i0] = (float)s;
  }
}

hipError_t k_diag_build(float* diag, int n, uint32_t nt, const uint64_t* z, const double* c,
                        hipStream_t st) {
  const uint64_t dim = 1ull << n;
  if (n <= 32)
    hipLaunchKernelGGL(diag_build_kernel<uint32_t>, dim3(grid_for(dim)), dim3(BLK), 0, st, diag,
                       dim, nt, z, c);
  else
    hipLaunchKernelGGL(diag_build_kernel<uint64_t>, dim3(grid_for(dim)), dim3(BLK), 0, st, diag,
                       dim, nt, z, c);
  return hipGetLastError();
}

uint32_t ham_apply_grid(int n) { return grid_for(1ull << n); }

hipError_t k_ham_apply(const float2* psi, float2* out, int n, const HamApply& H, hipStream_t st,
                       double* epart) {
  const uint64_t dim = 1ull << n;
  if (n <= 29 && H.ntab < (1ull << 29))
    hipLaunchKernelGGL((ham_apply_kernel<uint32_t, true>), dim3(grid_for(dim)), dim3(BLK), 0, st, psi,
                       out, dim, H.nsub, H.sub, H.tab, H.nblk, H.blk, H.mem, H.ncblk, H.cblk,
                       H.chk, H.ngroups, H.gx, H.gstart, H.tz, H.tc, H.diag, epart);
  else if (n <= 32)
    hipLaunchKernelGGL(ham_apply_kernel<uint32_t>, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, out,
                       dim, H.nsub, H.sub, H.tab, H.nblk, H.blk, H.mem, H.ncblk, H.cblk, H.chk,
                       H.ngroups, H.gx, H.gstart, H.tz, H.tc, H.diag, epart);
  else
    hipLaunchKernelGGL(ham_apply_kernel<uint64_t>, dim3(grid_for(dim)), dim3(BLK), 0, st, psi, out,
                       dim, 0u, nullptr, nullptr, 0u, nullptr, nullptr, 0u, nullptr, nullptr,
                       H.ngroups, H.gx, H.gstart, H.tz, H.tc, H.diag, epart);
  return hipGetLastError();
}

}  // namespace tqsv
