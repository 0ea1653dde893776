// tqsv_pass_impl.h - the fused tile-pass kernel (templates).  Each
// tqsv_pass_*.hip translation unit instantiates some of its variants.
//
// One launch sweeps the state once (one HBM read, and one write for gate
// passes).  Each work-group copies a tile of 2^T amplitudes into LDS, then
// runs the pass's phases: a phase loads 2^RB amplitudes per thread from LDS
// into registers (the amplitudes that differ only in the phase's register
// positions), applies the phase's steps entirely in registers, and writes the
// amplitudes back.  Gates whose targets are register positions cost no
// communication; diagonal factors, controls and Pauli signs are evaluated from
// the global index of each amplitude, so their bits may lie anywhere.
//
// The program format is defined in tqsv_prog.h; tqsv_plan.cpp writes it.
#ifndef TQSV_PASS_IMPL_H
#define TQSV_PASS_IMPL_H

#include <hip/hip_runtime.h>

#include "tqsv_internal.h"
#include "tqsv_prog.h"

namespace tqsv {

// LDS swizzle: XOR bits 1..4 with bits 5..8.  Linear over GF(2), so
// swz(a ^ b) = swz(a) ^ swz(b); bit 0 is preserved, so the amplitude pairs
// (2c, 2c+1) stay adjacent and 16-byte aligned.
__device__ __host__ __forceinline__ uint32_t swz(uint32_t p) { return p ^ ((p >> 4) & 30u); }

template <typename IDX>
__device__ __forceinline__ IDX ld_idx(const uint32_t* p);
template <>
__device__ __forceinline__ uint32_t ld_idx<uint32_t>(const uint32_t* p) {
  return p[0];
}
template <>
__device__ __forceinline__ uint64_t ld_idx<uint64_t>(const uint32_t* p) {
  return (uint64_t)p[0] | ((uint64_t)p[1] << 32);
}

__device__ __forceinline__ uint32_t parity(uint32_t v) { return __popc(v) & 1u; }
__device__ __forceinline__ uint32_t parity(uint64_t v) { return __popcll(v) & 1u; }

__device__ __forceinline__ float fl(const uint32_t* p) { return __uint_as_float(p[0]); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cflip(float2 a, uint32_t sb) {
  return make_float2(__uint_as_float(__float_as_uint(a.x) ^ sb),
                     __uint_as_float(__float_as_uint(a.y) ^ sb));
}

// per-phase geometry of one thread
template <typename IDX>
struct PhaseGeo {
  uint32_t bt;           // swizzled LDS index of the thread's slot 0
  uint32_t sph[RB];      // swizzled LDS offset of each register position
  IDX gb;                // global index of the thread's slot 0
  IDX gR[RB];            // global index bit of each register position
};

template <typename IDX>
__device__ __forceinline__ IDX gs_of(const PhaseGeo<IDX>& G, int s) {
  IDX v = 0;
#pragma unroll
  for (int j = 0; j < RB; ++j)
    if (s & (1 << j)) v ^= G.gR[j];
  return v;
}
template <typename IDX>
__device__ __forceinline__ uint32_t ls_of(const PhaseGeo<IDX>& G, int s) {
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < RB; ++j)
    if (s & (1 << j)) v ^= G.sph[j];
  return v;
}

struct M2f {
  float r00, i00, r01, i01, r10, i10, r11, i11;
};

__device__ __forceinline__ M2f load_m2(const uint32_t* w) {
  M2f u;
  u.r00 = fl(w + 0); u.i00 = fl(w + 1); u.r01 = fl(w + 2); u.i01 = fl(w + 3);
  u.r10 = fl(w + 4); u.i10 = fl(w + 5); u.r11 = fl(w + 6); u.i11 = fl(w + 7);
  return u;
}

template <uint32_t TYPE>
__device__ __forceinline__ void mul2(const M2f& u, float2 x0, float2 x1, float2& y0, float2& y1) {
  if constexpr (TYPE == U1_REAL) {
    y0.x = u.r00 * x0.x + u.r01 * x1.x;
    y0.y = u.r00 * x0.y + u.r01 * x1.y;
    y1.x = u.r10 * x0.x + u.r11 * x1.x;
    y1.y = u.r10 * x0.y + u.r11 * x1.y;
  } else if constexpr (TYPE == U1_RXLIKE) {
    y0.x = u.r00 * x0.x - u.i01 * x1.y;
    y0.y = u.r00 * x0.y + u.i01 * x1.x;
    y1.x = u.r11 * x1.x - u.i10 * x0.y;
    y1.y = u.r11 * x1.y + u.i10 * x0.x;
  } else {
    y0.x = u.r00 * x0.x - u.i00 * x0.y + u.r01 * x1.x - u.i01 * x1.y;
    y0.y = u.r00 * x0.y + u.i00 * x0.x + u.r01 * x1.y + u.i01 * x1.x;
    y1.x = u.r10 * x0.x - u.i10 * x0.y + u.r11 * x1.x - u.i11 * x1.y;
    y1.y = u.r10 * x0.y + u.i10 * x0.x + u.r11 * x1.y + u.i11 * x1.x;
  }
}

// ---- OP_U1 --------------------------------------------------------------
template <int K, uint32_t TYPE, bool CTRL, typename IDX>
__device__ __forceinline__ void u1_apply(float2 (&a)[NSLOT], const M2f& u, const PhaseGeo<IDX>& G,
                                         IDX cm, IDX cv) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & (1 << K)) continue;
    const int s1 = s | (1 << K);
    float2 y0, y1;
    mul2<TYPE>(u, a[s], a[s1], y0, y1);
    if constexpr (CTRL) {
      const bool on = (((G.gb ^ gs_of(G, s)) & cm) == cv);
      a[s] = on ? y0 : a[s];
      a[s1] = on ? y1 : a[s1];
    } else {
      a[s] = y0;
      a[s1] = y1;
    }
  }
}

template <int K, bool DUAL, typename IDX>
__device__ __forceinline__ void u1_dispatch(float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                            const uint32_t* w, const PhaseGeo<IDX>& G) {
  const uint32_t type = w[2];
  const bool ctrl = w[3] != 0;
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  const M2f u = load_m2(w + 8);
#define TQSV_U1(TY, CT)                                  \
  {                                                     \
    u1_apply<K, TY, CT, IDX>(a, u, G, cm, cv);          \
    if constexpr (DUAL) u1_apply<K, TY, CT, IDX>(b, u, G, cm, cv); \
  }
  if (ctrl) {
    if (type == U1_REAL) TQSV_U1(U1_REAL, true)
    else TQSV_U1(U1_GENERAL, true)
  } else {
    if (type == U1_REAL) TQSV_U1(U1_REAL, false)
    else if (type == U1_RXLIKE) TQSV_U1(U1_RXLIKE, false)
    else TQSV_U1(U1_GENERAL, false)
  }
#undef TQSV_U1
}

// ---- OP_U2 --------------------------------------------------------------
template <int KH, int KL, bool CTRL, typename IDX>
__device__ __forceinline__ void u2_apply(float2 (&a)[NSLOT], const uint32_t* m,
                                         const PhaseGeo<IDX>& G, IDX cm, IDX cv) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & ((1 << KH) | (1 << KL))) continue;
    const int idx[4] = {s, s | (1 << KL), s | (1 << KH), s | (1 << KH) | (1 << KL)};
    float2 x[4], y[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = a[idx[c]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float yr = 0.f, yi = 0.f;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float mr = fl(m + 8 * r + 2 * c), mi = fl(m + 8 * r + 2 * c + 1);
        yr += mr * x[c].x - mi * x[c].y;
        yi += mr * x[c].y + mi * x[c].x;
      }
      y[r] = make_float2(yr, yi);
    }
    bool on = true;
    if constexpr (CTRL) on = (((G.gb ^ gs_of(G, s)) & cm) == cv);
#pragma unroll
    for (int c = 0; c < 4; ++c) a[idx[c]] = on ? y[c] : x[c];
  }
}

template <int KH, int KL, bool DUAL, typename IDX>
__device__ __forceinline__ void u2_dispatch(float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                            const uint32_t* w, const PhaseGeo<IDX>& G) {
  const bool ctrl = w[3] != 0;
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  if (ctrl) {
    u2_apply<KH, KL, true, IDX>(a, w + 8, G, cm, cv);
    if constexpr (DUAL) u2_apply<KH, KL, true, IDX>(b, w + 8, G, cm, cv);
  } else {
    u2_apply<KH, KL, false, IDX>(a, w + 8, G, cm, cv);
    if constexpr (DUAL) u2_apply<KH, KL, false, IDX>(b, w + 8, G, cm, cv);
  }
}

// ---- OP_DIAG ------------------------------------------------------------
template <bool DUAL, typename IDX>
__device__ __forceinline__ void diag_apply(float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                           const uint32_t* w, const PhaseGeo<IDX>& G) {
  const uint32_t ntab = w[1], nthr = w[2];
  float2 f = make_float2(1.f, 0.f);
  const uint32_t* e = w + DIAG_HDR + (ntab ? DIAG_TAB : 0);
  for (uint32_t k = 0; k < nthr; ++k, e += DIAG_ENT) {
    const uint32_t ba = (uint32_t)((G.gb >> e[0]) & 1u);
    const uint32_t bb = (uint32_t)((G.gb >> e[1]) & 1u);
    const uint32_t sel = 2u * ba + bb;
    // select one of 4 complex entries without dynamic indexing of a register array
    float dr = fl(e + 2), di = fl(e + 3);
    dr = sel == 1u ? fl(e + 4) : dr; di = sel == 1u ? fl(e + 5) : di;
    dr = sel == 2u ? fl(e + 6) : dr; di = sel == 2u ? fl(e + 7) : di;
    dr = sel == 3u ? fl(e + 8) : dr; di = sel == 3u ? fl(e + 9) : di;
    f = cmul(f, make_float2(dr, di));
  }
  if (ntab) {
    const uint32_t* t = w + DIAG_HDR;
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
      const float2 ts = make_float2(fl(t + 2 * s), fl(t + 2 * s + 1));
      const float2 m = nthr ? cmul(ts, f) : ts;
      a[s] = cmul(a[s], m);
      if constexpr (DUAL) b[s] = cmul(b[s], m);
    }
  } else {
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
      a[s] = cmul(a[s], f);
      if constexpr (DUAL) b[s] = cmul(b[s], f);
    }
  }
}

// ---- OP_CZ --------------------------------------------------------------
template <bool DUAL, typename IDX>
__device__ __forceinline__ void cz_apply(float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                         const uint32_t* w, const PhaseGeo<IDX>& G) {
  const uint32_t cnt = w[1];
  uint32_t flip = 0;  // bit s: slot s changes sign
  const uint32_t* e = w + 2;
  for (uint32_t k = 0; k < cnt; ++k, e += 3) {
    const IDX mask = ld_idx<IDX>(e);
    const uint32_t sh = e[2];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
      const IDX g = G.gb ^ gs_of(G, s);
      flip ^= parity((IDX)(g & (g >> sh) & mask)) << s;
    }
  }
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    const uint32_t sb = ((flip >> s) & 1u) << 31;
    a[s] = cflip(a[s], sb);
    if constexpr (DUAL) b[s] = cflip(b[s], sb);
  }
}

// ---- OP_D2 --------------------------------------------------------------
template <bool DUAL, typename IDX>
__device__ __forceinline__ void d2_apply(float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                         const uint32_t* w, const PhaseGeo<IDX>& G) {
  const uint32_t ga = w[1], gbb = w[2];
  const float2 d0 = make_float2(fl(w + 4), fl(w + 5)), d1 = make_float2(fl(w + 6), fl(w + 7));
  const float2 d2 = make_float2(fl(w + 8), fl(w + 9)), d3 = make_float2(fl(w + 10), fl(w + 11));
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    const IDX g = G.gb ^ gs_of(G, s);
    const uint32_t sel = 2u * (uint32_t)((g >> ga) & 1u) + (uint32_t)((g >> gbb) & 1u);
    float2 d = d0;
    d = sel == 1u ? d1 : d;
    d = sel == 2u ? d2 : d;
    d = sel == 3u ? d3 : d;
    a[s] = cmul(a[s], d);
    if constexpr (DUAL) b[s] = cmul(b[s], d);
  }
}

// ---- OP_PROT ------------------------------------------------------------
// new0 = c a0 + s (-1)^{sig1} i^q a1, new1 = c a1 + s (-1)^{sig0} i^q a0
template <int Q>
__device__ __forceinline__ float2 iq(float2 v) {
  if constexpr (Q == 0) return v;
  else if constexpr (Q == 1) return make_float2(-v.y, v.x);
  else if constexpr (Q == 2) return make_float2(-v.x, -v.y);
  else return make_float2(v.y, -v.x);
}

template <int XR, int Q, typename IDX>
__device__ __forceinline__ void prot_apply1(float2 (&a)[NSLOT], const PhaseGeo<IDX>& G, IDX z,
                                            float c, float sn, uint32_t xzpar) {
  if constexpr (XR == 0) {
    // diagonal Pauli string: amp *= c - i s sigma (q = 3 for k = 0)
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
      const uint32_t sg = parity((IDX)((G.gb ^ gs_of(G, s)) & z));
      const float ss = sg ? -sn : sn;
      const float2 v = a[s];
      a[s] = make_float2(c * v.x + ss * v.y, c * v.y - ss * v.x);
    }
  } else {
    constexpr int LB = XR & (-XR);
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
      if (s & LB) continue;
      const int s1 = s ^ XR;
      const uint32_t sg0 = parity((IDX)((G.gb ^ gs_of(G, s)) & z));
      const uint32_t sg1 = sg0 ^ xzpar;
      const float s0f = sg0 ? -sn : sn, s1f = sg1 ? -sn : sn;
      const float2 x0 = a[s], x1 = a[s1];
      const float2 r1 = iq<Q>(x1), r0 = iq<Q>(x0);
      a[s] = make_float2(c * x0.x + s1f * r1.x, c * x0.y + s1f * r1.y);
      a[s1] = make_float2(c * x1.x + s0f * r0.x, c * x1.y + s0f * r0.y);
    }
  }
}

template <int XR, bool DUAL, typename IDX>
__device__ __forceinline__ void prot_x(float2 (&a)[NSLOT], float2 (&b)[NSLOT], const uint32_t* w,
                                       const PhaseGeo<IDX>& G) {
  const IDX z = ld_idx<IDX>(w + 2);
  const float c = fl(w + 4), sn = fl(w + 5);
  const uint32_t k = w[6] & 3u;
  const uint32_t q = (k + 3u) & 3u;  // -i * i^k
  const uint32_t xzpar = k & 1u;
#define TQSV_PR(QQ)                                          \
  {                                                         \
    prot_apply1<XR, QQ, IDX>(a, G, z, c, sn, xzpar);         \
    if constexpr (DUAL) prot_apply1<XR, QQ, IDX>(b, G, z, c, sn, xzpar); \
  }
  if constexpr (XR == 0) {
    TQSV_PR(3)
  } else {
    switch (q) {
      case 0: TQSV_PR(0) break;
      case 1: TQSV_PR(1) break;
      case 2: TQSV_PR(2) break;
      default: TQSV_PR(3) break;
    }
  }
#undef TQSV_PR
}

// ---- OP_RED -------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int XR, int KIND, bool DUAL, typename IDX>
__device__ __forceinline__ void red_x(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                      const uint32_t* w, const PhaseGeo<IDX>& G, double* acc,
                                      uint32_t nred) {
  constexpr int LB = XR & (-XR);
  float2 p[NSLOT];
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if constexpr (KIND == RED_PAIR) {
      if (s & LB) continue;
      const float2 u = a[s ^ XR], v = a[s];  // conj(u) * v
      p[s] = make_float2(u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x);
    } else if constexpr (KIND == RED_DIAG) {
      p[s] = make_float2(a[s].x * a[s].x + a[s].y * a[s].y, 0.f);
    } else {
      if constexpr (DUAL) {
        const float2 u = b[s], v = a[s ^ XR];
        p[s] = make_float2(u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x);
      }
    }
  }
  const uint32_t nterms = w[3];
  const uint32_t* e = w + RED_HDR;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t t = 0; t < nterms; ++t, e += RED_ENT) {
    const uint32_t slot = e[0];
    const IDX zt = ld_idx<IDX>(e + 1);
    const uint32_t smask = e[3];
    const uint32_t part = e[4];
    float accv = 0.f;
    if (part == 0u) {
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        if (KIND == RED_PAIR && (s & LB)) continue;
        const float sg = ((smask >> s) & 1u) ? -1.f : 1.f;
        accv = fmaf(sg, p[s].x, accv);
      }
    } else {
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        if (KIND == RED_PAIR && (s & LB)) continue;
        const float sg = ((smask >> s) & 1u) ? -1.f : 1.f;
        accv = fmaf(sg, p[s].y, accv);
      }
    }
    double d = (double)accv;
    if (parity((IDX)(G.gb & zt))) d = -d;
    d = wave_sum(d);
    if (lane == 0) acc[wave * nred + slot] += d;
  }
}

template <int XR, bool DUAL, typename IDX>
__device__ __forceinline__ void red_dispatch_kind(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                                  const uint32_t* w, const PhaseGeo<IDX>& G,
                                                  double* acc, uint32_t nred) {
  const uint32_t kind = w[2];
  if constexpr (XR == 0) {
    if (kind == RED_CROSS) {
      if constexpr (DUAL) red_x<0, RED_CROSS, DUAL, IDX>(a, b, w, G, acc, nred);
    } else {
      red_x<0, RED_DIAG, DUAL, IDX>(a, b, w, G, acc, nred);
    }
  } else {
    if (kind == RED_CROSS) {
      if constexpr (DUAL) red_x<XR, RED_CROSS, DUAL, IDX>(a, b, w, G, acc, nred);
    } else {
      red_x<XR, RED_PAIR, DUAL, IDX>(a, b, w, G, acc, nred);
    }
  }
}

// ---- dispatch helpers over the runtime register masks --------------------
template <bool DUAL, typename IDX, int X = 0>
__device__ __forceinline__ void prot_switch(uint32_t xr, float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                            const uint32_t* w, const PhaseGeo<IDX>& G) {
  if constexpr (X < NSLOT) {
    if (xr == (uint32_t)X) {
      prot_x<X, DUAL, IDX>(a, b, w, G);
      return;
    }
    prot_switch<DUAL, IDX, X + 1>(xr, a, b, w, G);
  }
}

template <bool DUAL, typename IDX, int X = 0>
__device__ __forceinline__ void red_switch(uint32_t xr, const float2 (&a)[NSLOT],
                                           const float2 (&b)[NSLOT], const uint32_t* w,
                                           const PhaseGeo<IDX>& G, double* acc, uint32_t nred) {
  if constexpr (X < NSLOT) {
    if (xr == (uint32_t)X) {
      red_dispatch_kind<X, DUAL, IDX>(a, b, w, G, acc, nred);
      return;
    }
    red_switch<DUAL, IDX, X + 1>(xr, a, b, w, G, acc, nred);
  }
}

template <bool DUAL, typename IDX>
__device__ __forceinline__ void u1_switch(uint32_t k, float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                          const uint32_t* w, const PhaseGeo<IDX>& G) {
  switch (k) {
    case 0: u1_dispatch<0, DUAL, IDX>(a, b, w, G); break;
    case 1: u1_dispatch<1, DUAL, IDX>(a, b, w, G); break;
    case 2: u1_dispatch<2, DUAL, IDX>(a, b, w, G); break;
    case 3: u1_dispatch<3, DUAL, IDX>(a, b, w, G); break;
    default: u1_dispatch<4, DUAL, IDX>(a, b, w, G); break;
  }
}

template <bool DUAL, typename IDX>
__device__ __forceinline__ void u2_switch(uint32_t kh, uint32_t kl, float2 (&a)[NSLOT],
                                          float2 (&b)[NSLOT], const uint32_t* w,
                                          const PhaseGeo<IDX>& G) {
  switch (kh * 8u + kl) {
    case 8 * 1 + 0: u2_dispatch<1, 0, DUAL, IDX>(a, b, w, G); break;
    case 8 * 2 + 0: u2_dispatch<2, 0, DUAL, IDX>(a, b, w, G); break;
    case 8 * 2 + 1: u2_dispatch<2, 1, DUAL, IDX>(a, b, w, G); break;
    case 8 * 3 + 0: u2_dispatch<3, 0, DUAL, IDX>(a, b, w, G); break;
    case 8 * 3 + 1: u2_dispatch<3, 1, DUAL, IDX>(a, b, w, G); break;
    case 8 * 3 + 2: u2_dispatch<3, 2, DUAL, IDX>(a, b, w, G); break;
    case 8 * 4 + 0: u2_dispatch<4, 0, DUAL, IDX>(a, b, w, G); break;
    case 8 * 4 + 1: u2_dispatch<4, 1, DUAL, IDX>(a, b, w, G); break;
    case 8 * 4 + 2: u2_dispatch<4, 2, DUAL, IDX>(a, b, w, G); break;
    default: u2_dispatch<4, 3, DUAL, IDX>(a, b, w, G); break;
  }
}

// ---- the kernel -----------------------------------------------------------
// tile index -> global index of the tile's position 0: spread the tile number
// over the index bits outside the tile (L and above, skipping H).
template <typename IDX>
__device__ __forceinline__ IDX tile_base(uint32_t tile, uint32_t L, uint32_t m, const uint32_t* H) {
  IDX v = (IDX)tile << L;
  for (uint32_t k = 0; k < m; ++k) {
    const uint32_t h = H[k];
    const IDX lowm = (((IDX)1) << h) - 1;
    v = ((v >> h) << (h + 1)) | (v & lowm);
  }
  return v;
}

// tile position -> global index bits (no tile base)
template <typename IDX>
__device__ __forceinline__ IDX pos_map(uint32_t p, uint32_t L, uint32_t m, const uint32_t* H) {
  IDX v = (IDX)(p & ((1u << L) - 1u));
  for (uint32_t k = 0; k < m; ++k)
    if ((p >> (L + k)) & 1u) v |= ((IDX)1) << H[k];
  return v;
}

template <int TB, bool DUAL, typename IDX>
__global__ __launch_bounds__(1 << TB) void k_pass(float2* __restrict__ psi, float2* __restrict__ lam,
                                                  const uint32_t* __restrict__ prog,
                                                  double* __restrict__ partials, uint32_t ntiles) {
  constexpr int NT = 1 << TB;
  constexpr int T = TB + RB;
  constexpr uint32_t TSIZE = 1u << T;
  constexpr int NCHUNK = (TSIZE / 2) / NT;  // float4 chunks per thread
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tA = reinterpret_cast<float2*>(smem);
  float2* tB = tA + TSIZE;  // only used if DUAL
  double* acc = reinterpret_cast<double*>(smem + (DUAL ? 2 : 1) * TSIZE * sizeof(float2));

  const uint32_t tid = threadIdx.x;
  const uint32_t flags = prog[PH_FLAGS];
  const uint32_t L = prog[PH_L], m = prog[PH_M];
  const uint32_t nphase = prog[PH_NPHASE], nred = prog[PH_NRED];
  const uint32_t* H = prog + PH_H;
  constexpr uint32_t NWAVE = NT / 64 > 0 ? NT / 64 : 1;

  for (uint32_t i = tid; i < NWAVE * nred; i += NT) acc[i] = 0.0;

  // global index bits of the thread's first load chunk (tile position 2*tid)
  const IDX gthr = pos_map<IDX>(2u * tid, L, m, H);
  const uint32_t lthr = swz(2u * tid);

  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const IDX base = tile_base<IDX>(tile, L, m, H);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NCHUNK; ++k) {
      const uint32_t pk = (uint32_t)k << (TB + 1);
      const IDX g = base | gthr | pos_map<IDX>(pk, L, m, H);
      const uint32_t l = lthr ^ swz(pk);
      const float4 v = reinterpret_cast<const float4*>(psi)[g >> 1];
      *reinterpret_cast<float4*>(tA + l) = v;
      if constexpr (DUAL) {
        const float4 u = reinterpret_cast<const float4*>(lam)[g >> 1];
        *reinterpret_cast<float4*>(tB + l) = u;
      }
    }
    __syncthreads();

    const uint32_t* ph = prog + PH_SIZE;
    for (uint32_t iph = 0; iph < nphase; ++iph) {
      const uint32_t nsteps = ph[PS_NSTEPS];
      PhaseGeo<IDX> G;
      uint32_t tpos = tid;
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const uint32_t r = ph[PS_R + j];
        tpos = ((tpos >> r) << (r + 1)) | (tpos & ((1u << r) - 1u));
        G.sph[j] = swz(1u << r);
        G.gR[j] = pos_map<IDX>(1u << r, L, m, H);
      }
      G.bt = swz(tpos);
      G.gb = base | pos_map<IDX>(tpos, L, m, H);

      float2 a[NSLOT], b[NSLOT];
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        const uint32_t l = G.bt ^ ls_of(G, s);
        a[s] = tA[l];
        if constexpr (DUAL) b[s] = tB[l];
      }

      const uint32_t* w = ph + PS_SIZE;
      for (uint32_t is = 0; is < nsteps; ++is) {
        const uint32_t op = w[0] & 0xffffu;
        const uint32_t nw = w[0] >> 16;
        switch (op) {
          case OP_U1: u1_switch<DUAL, IDX>(w[1], a, b, w, G); break;
          case OP_U2: u2_switch<DUAL, IDX>(w[1], w[2], a, b, w, G); break;
          case OP_DIAG: diag_apply<DUAL, IDX>(a, b, w, G); break;
          case OP_CZ: cz_apply<DUAL, IDX>(a, b, w, G); break;
          case OP_D2: d2_apply<DUAL, IDX>(a, b, w, G); break;
          case OP_PROT: prot_switch<DUAL, IDX>(w[1], a, b, w, G); break;
          case OP_RED: red_switch<DUAL, IDX>(w[1], a, b, w, G, acc, nred); break;
          default: break;
        }
        w += nw;
      }

      if (flags & PF_WRITE) {
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
          const uint32_t l = G.bt ^ ls_of(G, s);
          tA[l] = a[s];
          if constexpr (DUAL) tB[l] = b[s];
        }
      }
      __syncthreads();
      ph = w;
    }

    if (flags & PF_WRITE) {
#pragma unroll
      for (int k = 0; k < NCHUNK; ++k) {
        const uint32_t pk = (uint32_t)k << (TB + 1);
        const IDX g = base | gthr | pos_map<IDX>(pk, L, m, H);
        const uint32_t l = lthr ^ swz(pk);
        reinterpret_cast<float4*>(psi)[g >> 1] = *reinterpret_cast<const float4*>(tA + l);
        if constexpr (DUAL)
          reinterpret_cast<float4*>(lam)[g >> 1] = *reinterpret_cast<const float4*>(tB + l);
      }
    }
  }

  if (nred) {
    __syncthreads();
    for (uint32_t r = tid; r < nred; r += NT) {
      double sum = 0.0;
      for (uint32_t wv = 0; wv < NWAVE; ++wv) sum += acc[wv * nred + r];
      partials[(size_t)blockIdx.x * nred + r] = sum;
    }
  }
}

template <int TB, bool DUAL, typename IDX>
hipError_t launch_one(const PassLaunch& pl, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<TB, DUAL, IDX>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_pass<TB, DUAL, IDX>), dim3(pl.grid), dim3(1 << TB), pl.lds_bytes, st,
                     pl.psi, pl.lam, pl.prog, pl.partials, pl.ntiles);
  return hipGetLastError();
}

}  // namespace tqsv

#endif
