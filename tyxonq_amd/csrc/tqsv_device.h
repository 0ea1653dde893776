// tqsv_device.h - device-side building blocks of the tile-pass kernels.
//
// Shared by the precompiled interpreting kernel (tqsv_pass_impl.h), which
// reads the step sequence of a pass from its program at run time, and by the
// kernels that tqsv_jit.cpp generates per pass program, which call the same
// step functions with the step sequence fixed at compile time.  This header
// is also compiled at run time (hiprtc), so it includes no host headers there.
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
// Programs are read through the constant address space: every program word
// is uniform across the work-group, so the compiler fetches it with scalar
// loads and keeps step parameters in scalar registers.
#ifndef TQSV_DEVICE_H
#define TQSV_DEVICE_H

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "tqsv_prog.h"

namespace tqsv {


typedef const __attribute__((address_space(4))) uint32_t* cw_t;

// LDS swizzle: XOR bits 1..4 with bits 5..8.  Linear over GF(2), so
// swz(a ^ b) = swz(a) ^ swz(b); bit 0 is preserved, so the amplitude pairs
// (2c, 2c+1) stay adjacent and 16-byte aligned.
__device__ __host__ __forceinline__ uint32_t swz(uint32_t p) { return p ^ ((p >> 4) & 30u); }

template <typename IDX>
__device__ __forceinline__ IDX ld_idx(cw_t p);
template <>
__device__ __forceinline__ uint32_t ld_idx<uint32_t>(cw_t p) {
  return p[0];
}
template <>
__device__ __forceinline__ uint64_t ld_idx<uint64_t>(cw_t p) {
  return (uint64_t)p[0] | ((uint64_t)p[1] << 32);
}

__device__ __forceinline__ uint32_t parity(uint32_t v) { return __popc(v) & 1u; }
__device__ __forceinline__ uint32_t parity(uint64_t v) { return __popcll(v) & 1u; }

__device__ __forceinline__ float fl(cw_t p) { return __uint_as_float(p[0]); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cneg(float2 a) { return make_float2(-a.x, -a.y); }
// per-component select: a select of whole float2 values can become a select
// of their addresses, which moves the operands to scratch memory
__device__ __forceinline__ float2 sel2(bool c, float2 p, float2 q) {
  return make_float2(c ? p.x : q.x, c ? p.y : q.y);
}

// per-phase geometry of one thread
template <typename IDX>
struct PhaseGeo {
  uint32_t bt;           // swizzled LDS index of the thread's slot 0
  uint32_t sph[RB];      // swizzled LDS offset of each register position
  IDX gb;                // global index of the thread's slot 0 (register bits zero)
  IDX gR[RB];            // global index bit of each register position
  IDX rmask;             // OR of gR
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

// ---- wave reduction -------------------------------------------------------
// Butterfly over the 64 lanes: quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror (lanes i and 7-i of each 8), row_mirror (i and 15-i of each
// 16), then exchanges with lanes i^16 and i^32.
// FP64 butterfly: the same lane pattern, each double moved as two 32-bit halves
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double shfl_xor_d(double v, int o) {
  const int lo = __shfl_xor(__double2loint(v), o, 64);
  const int hi = __shfl_xor(__double2hiint(v), o, 64);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  v += shfl_xor_d(v, 16);
  v += shfl_xor_d(v, 32);
  return v;
}
// DPP on the rows selected by ROWMASK (other lanes get 0)
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_rows_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, false);
  return __hiloint2double(hi, lo);
}
// The wave total in lane 63 only, without the LDS round trips of the two
// cross-row shuffles: after the in-row butterfly, row_bcast:15 adds row 0 to
// row 1 and row 2 to row 3, row_bcast:31 adds rows 0+1 to rows 2 and 3.
__device__ __forceinline__ double wave_sum_last_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  v += dpp_rows_d<0x142, 0xa>(v);
  v += dpp_rows_d<0x143, 0xc>(v);
  return v;
}

// ---- OP_U1 --------------------------------------------------------------
struct M2f {
  float r00, i00, r01, i01, r10, i10, r11, i11;
};

__device__ __forceinline__ M2f load_m2(cw_t w) {
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

// Controls split into the thread part (one test per thread) and the register
// part (uniform per slot: slots whose register bits fail it are skipped).
template <int K, uint32_t TYPE, bool CTRL, typename IDX>
__device__ __forceinline__ void u1_apply(float2 (&a)[NSLOT], const M2f& u, const PhaseGeo<IDX>& G,
                                         IDX cm, IDX cv, bool ton) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & (1 << K)) continue;
    const int s1 = s | (1 << K);
    if constexpr (CTRL) {
      const IDX rs = gs_of(G, s);
      if ((rs & cm & G.rmask) != (cv & G.rmask)) continue;
    }
    float2 y0, y1;
    mul2<TYPE>(u, a[s], a[s1], y0, y1);
    if constexpr (CTRL) {
      a[s] = sel2(ton, y0, a[s]);
      a[s1] = sel2(ton, y1, a[s1]);
    } else {
      a[s] = y0;
      a[s1] = y1;
    }
  }
}

template <int K, int MODE, typename IDX>
__device__ __forceinline__ void u1_dispatch(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                            const PhaseGeo<IDX>& G) {
  const uint32_t type = w[2];
  const bool ctrl = w[3] != 0;
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  const M2f u = load_m2(w + 8);
  const bool ton = (G.gb & cm) == (cv & ~G.rmask);
#define TQSV_U1(TY, CT)                                              \
  {                                                                 \
    u1_apply<K, TY, CT, IDX>(a, u, G, cm, cv, ton);                 \
    if constexpr (MODE == MODE_DUAL) u1_apply<K, TY, CT, IDX>(b, u, G, cm, cv, ton); \
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
__device__ __forceinline__ void u2_apply(float2 (&a)[NSLOT], cw_t m, const PhaseGeo<IDX>& G,
                                         IDX cm, IDX cv, bool ton) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & ((1 << KH) | (1 << KL))) continue;
    if constexpr (CTRL) {
      const IDX rs = gs_of(G, s);
      if ((rs & cm & G.rmask) != (cv & G.rmask)) continue;
    }
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
#pragma unroll
    for (int c = 0; c < 4; ++c) a[idx[c]] = CTRL ? sel2(ton, y[c], x[c]) : y[c];
  }
}

template <int KH, int KL, int MODE, typename IDX>
__device__ __forceinline__ void u2_dispatch(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                            const PhaseGeo<IDX>& G) {
  const bool ctrl = w[3] != 0;
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  const bool ton = (G.gb & cm) == (cv & ~G.rmask);
  if (ctrl) {
    u2_apply<KH, KL, true, IDX>(a, w + 8, G, cm, cv, ton);
    if constexpr (MODE == MODE_DUAL) u2_apply<KH, KL, true, IDX>(b, w + 8, G, cm, cv, ton);
  } else {
    u2_apply<KH, KL, false, IDX>(a, w + 8, G, cm, cv, ton);
    if constexpr (MODE == MODE_DUAL) u2_apply<KH, KL, false, IDX>(b, w + 8, G, cm, cv, ton);
  }
}

// ---- OP_ROT -------------------------------------------------------------
template <int K, int MD>
__device__ __forceinline__ void rot_apply(float2 (&a)[NSLOT], float t) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & (1 << K)) continue;
    const int s1 = s | (1 << K);
    const float2 x0 = a[s], x1 = a[s1];
    if constexpr (MD == 0) {
      a[s] = make_float2(fmaf(-t, x1.x, x0.x), fmaf(-t, x1.y, x0.y));
      a[s1] = make_float2(fmaf(t, x0.x, x1.x), fmaf(t, x0.y, x1.y));
    } else if constexpr (MD == 1) {
      a[s] = make_float2(fmaf(t, x0.x, -x1.x), fmaf(t, x0.y, -x1.y));
      a[s1] = make_float2(fmaf(t, x1.x, x0.x), fmaf(t, x1.y, x0.y));
    } else if constexpr (MD == 2) {
      // x0 - i t x1 = (x0.x + t x1.y, x0.y - t x1.x)
      a[s] = make_float2(fmaf(t, x1.y, x0.x), fmaf(-t, x1.x, x0.y));
      a[s1] = make_float2(fmaf(t, x0.y, x1.x), fmaf(-t, x0.x, x1.y));
    } else {
      // t x0 - i x1 = (t x0.x + x1.y, t x0.y - x1.x)
      a[s] = make_float2(fmaf(t, x0.x, x1.y), fmaf(t, x0.y, -x1.x));
      a[s1] = make_float2(fmaf(t, x1.x, x0.y), fmaf(t, x1.y, -x0.x));
    }
  }
}

template <int MODE, typename IDX>
__device__ __forceinline__ void rot_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                         const PhaseGeo<IDX>& G) {
  float t = fl(w + 3);
  const IDX zs = ld_idx<IDX>(w + 4);
  if (zs != 0) t = parity((IDX)(G.gb & zs)) ? -t : t;
#define TQSV_ROT(K, MD)                                  \
  case 4 * K + MD:                                      \
    rot_apply<K, MD>(a, t);                             \
    if constexpr (MODE == MODE_DUAL) rot_apply<K, MD>(b, t); \
    break;
#define TQSV_ROTK(K) TQSV_ROT(K, 0) TQSV_ROT(K, 1) TQSV_ROT(K, 2) TQSV_ROT(K, 3)
  switch (4u * w[1] + w[2]) {
    TQSV_ROTK(0)
    TQSV_ROTK(1)
    TQSV_ROTK(2)
    TQSV_ROTK(3)
    TQSV_ROTK(4)
    default: break;
  }
#undef TQSV_ROTK
#undef TQSV_ROT
}

// ---- OP_PH --------------------------------------------------------------
template <int K>
__device__ __forceinline__ void ph_apply(float2 (&a)[NSLOT], float2 u) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s)
    if (s & (1 << K)) a[s] = cmul(a[s], u);
}

// The arms of switches over a slot bit end in an empty asm statement that
// differs per arm: otherwise the compiler merges the near-identical arms into
// one loop with a variable slot index, which moves the slot array to scratch.
#define TQSV_ARM_END(TAG, K) __asm__ volatile("; " TAG " arm %0" ::"n"(K))

// Marks every slot register as rewritten in place.  Used at the end of switch
// arms that leave some slots untouched, so that all arms hand the slots over
// in the same registers instead of each needing copies at the join.
__device__ __forceinline__ void touch_all(float2 (&a)[NSLOT]) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) __asm__ volatile("" : "+v"(a[s].x), "+v"(a[s].y));
}

template <int MODE, typename IDX>
__device__ __forceinline__ void ph_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                        const PhaseGeo<IDX>& G) {
  float2 u = make_float2(fl(w + 4), fl(w + 5));
  const uint32_t nent = w[2];
  cw_t e = w + OPH_HDR;
  for (uint32_t i = 0; i < nent; ++i, e += OPH_ENT) {
    const float2 v = cmul(u, make_float2(fl(e + 1), fl(e + 2)));
    const bool on = ((G.gb >> e[0]) & 1u) != 0;
    u = sel2(on, v, u);
  }
#define TQSV_PH(K)                                       \
  case K:                                               \
    ph_apply<K>(a, u);                                  \
    touch_all(a);                                       \
    if constexpr (MODE == MODE_DUAL) {                  \
      ph_apply<K>(b, u);                                \
      touch_all(b);                                     \
    }                                                   \
    TQSV_ARM_END("ph", K);                              \
    break;
  switch (w[1]) {
    TQSV_PH(0)
    TQSV_PH(1)
    TQSV_PH(2)
    TQSV_PH(3)
    TQSV_PH(4)
    default: break;
  }
#undef TQSV_PH
}

// ---- OP_SGN / OP_TAB / OP_THR ---------------------------------------------
template <int MODE>
__device__ __forceinline__ void sgn_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w) {
  const uint32_t mask = w[1];
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (mask & (1u << s)) {
      a[s] = cneg(a[s]);
      if constexpr (MODE == MODE_DUAL) b[s] = cneg(b[s]);
    }
  }
}

template <int MODE>
__device__ __forceinline__ void tab_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w) {
  cw_t t = w + TAB_OFF;
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    const float2 f = make_float2(fl(t + 2 * s), fl(t + 2 * s + 1));
    a[s] = cmul(a[s], f);
    if constexpr (MODE == MODE_DUAL) b[s] = cmul(b[s], f);
  }
}

template <int MODE, typename IDX>
__device__ __forceinline__ void thr_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                         const PhaseGeo<IDX>& G) {
  const uint32_t nent = w[1];
  float2 F = make_float2(fl(w + 2), fl(w + 3));
  cw_t e = w + THR_HDR;
  for (uint32_t i = 0; i < nent; ++i, e += THR_ENT) {
    const float2 v = cmul(F, make_float2(fl(e + 2), fl(e + 3)));
    const bool on = (((G.gb >> e[0]) & (G.gb >> e[1])) & 1u) != 0;
    F = sel2(on, v, F);
  }
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    a[s] = cmul(a[s], F);
    if constexpr (MODE == MODE_DUAL) b[s] = cmul(b[s], F);
  }
}

template <int MODE>
__device__ __forceinline__ void scl_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w) {
  const float f = fl(w + 1);
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    a[s] = make_float2(a[s].x * f, a[s].y * f);
    if constexpr (MODE == MODE_DUAL) b[s] = make_float2(b[s].x * f, b[s].y * f);
  }
}

// ---- OP_CXP -------------------------------------------------------------
// Controlled swap on slot bit K.  With the control a slot bit (CR < RB) or
// absent (CR == CX_NONE) the swap is a renaming of registers; with the
// control a thread bit (CR == CX_THREAD) each lane selects by its own test.
template <int K, int CR, int CV>
__device__ __forceinline__ void cxp_apply(float2 (&a)[NSLOT], bool ton) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & (1 << K)) continue;
    const int s1 = s | (1 << K);
    if constexpr (CR < RB) {
      if (((s >> CR) & 1) != CV) continue;
    }
    const float2 x0 = a[s], x1 = a[s1];
    if constexpr (CR == (int)CX_THREAD) {
      a[s] = sel2(ton, x1, x0);
      a[s1] = sel2(ton, x0, x1);
    } else {
      a[s] = x1;
      a[s1] = x0;
    }
  }
}

template <typename IDX>
__device__ __forceinline__ bool cxp_thread_on(cw_t w, const PhaseGeo<IDX>& G) {
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  return (G.gb & cm) == cv;
}

// interpreter form: smask (bit s: the register part of the control holds
// for slot s) is uniform, ton is the thread part.  Dual passes swap both
// states per pair, so that each pair's selector dies at once.  (Arms per
// control slot would not be cheaper here: every arm must hand the slots over
// in the same registers, so a renaming turns into moves at the join.)
template <int K, bool DUAL>
__device__ __forceinline__ void cxp_apply_rt(float2 (&a)[NSLOT], float2 (&b)[NSLOT], bool ton,
                                             uint32_t smask) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & (1 << K)) continue;
    const int s1 = s | (1 << K);
    const bool on = ((smask >> s) & 1u) != 0 && ton;
    const float2 x0 = a[s], x1 = a[s1];
    a[s] = sel2(on, x1, x0);
    a[s1] = sel2(on, x0, x1);
    if constexpr (DUAL) {
      const float2 y0 = b[s], y1 = b[s1];
      b[s] = sel2(on, y1, y0);
      b[s1] = sel2(on, y0, y1);
    }
  }
}

template <int K, int MODE, typename IDX>
__device__ __forceinline__ void cxp_dispatch(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                             const PhaseGeo<IDX>& G) {
  const uint32_t smask = w[8];
  const bool ton = w[2] == CX_THREAD ? cxp_thread_on<IDX>(w, G) : true;
  cxp_apply_rt<K, MODE == MODE_DUAL>(a, b, ton, smask);
}

template <int MODE, typename IDX>
__device__ __forceinline__ void cxp_switch(uint32_t k, float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                           cw_t w, const PhaseGeo<IDX>& G) {
#define TQSV_CXP(K)                                \
  case K:                                          \
    cxp_dispatch<K, MODE, IDX>(a, b, w, G);        \
    touch_all(a);                                  \
    if constexpr (MODE == MODE_DUAL) touch_all(b); \
    TQSV_ARM_END("cxp", K);                        \
    break;
  switch (k) {
    TQSV_CXP(0)
    TQSV_CXP(1)
    TQSV_CXP(2)
    TQSV_CXP(3)
    TQSV_CXP(4)
    default: break;
  }
#undef TQSV_CXP
}

// generated-kernel form: slot, control slot and value fixed at compile time
template <int K, int CR, int CV, int MODE, typename IDX>
__device__ __forceinline__ void jcxp(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                     const PhaseGeo<IDX>& G) {
  bool ton = true;
  if constexpr (CR == (int)CX_THREAD) ton = cxp_thread_on<IDX>(w, G);
  cxp_apply<K, CR, CV>(a, ton);
  if constexpr (MODE == MODE_DUAL) cxp_apply<K, CR, CV>(b, ton);
}

// ---- OP_D2 --------------------------------------------------------------
template <int MODE, typename IDX>
__device__ __forceinline__ void d2_step(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                        const PhaseGeo<IDX>& G) {
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
    if constexpr (MODE == MODE_DUAL) b[s] = cmul(b[s], d);
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
  const uint32_t tp = parity((IDX)(G.gb & z));
  if constexpr (XR == 0) {
    // diagonal Pauli string: amp *= c - i s sigma
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
      const uint32_t sg = tp ^ parity((IDX)(gs_of(G, s) & z));
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
      const uint32_t sg0 = tp ^ parity((IDX)(gs_of(G, s) & z));
      const uint32_t sg1 = sg0 ^ xzpar;
      const float s0f = sg0 ? -sn : sn, s1f = sg1 ? -sn : sn;
      const float2 x0 = a[s], x1 = a[s1];
      const float2 r1 = iq<Q>(x1), r0 = iq<Q>(x0);
      a[s] = make_float2(c * x0.x + s1f * r1.x, c * x0.y + s1f * r1.y);
      a[s1] = make_float2(c * x1.x + s0f * r0.x, c * x1.y + s0f * r0.y);
    }
  }
}

template <int XR, int MODE, typename IDX>
__device__ __forceinline__ void prot_x(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                       const PhaseGeo<IDX>& G) {
  const IDX z = ld_idx<IDX>(w + 2);
  const float c = fl(w + 4), sn = fl(w + 5);
  const uint32_t k = w[6] & 3u;
  const uint32_t q = (k + 3u) & 3u;  // -i * i^k
  const uint32_t xzpar = k & 1u;
#define TQSV_PR(QQ)                                          \
  {                                                         \
    prot_apply1<XR, QQ, IDX>(a, G, z, c, sn, xzpar);         \
    if constexpr (MODE == MODE_DUAL) prot_apply1<XR, QQ, IDX>(b, G, z, c, sn, xzpar); \
  }
  if constexpr (XR == 0) {
    TQSV_PR(3)
  } else {
    // i^(q+2) = -i^q: the sign goes into s, so two arms per flip mask suffice
    const float snq = q >= 2u ? -sn : sn;
#define TQSV_PRS(QQ)                                           \
  {                                                           \
    prot_apply1<XR, QQ, IDX>(a, G, z, c, snq, xzpar);          \
    if constexpr (MODE == MODE_DUAL) prot_apply1<XR, QQ, IDX>(b, G, z, c, snq, xzpar); \
  }
    if (q & 1u)
      TQSV_PRS(1)
    else
      TQSV_PRS(0)
#undef TQSV_PRS
  }
#undef TQSV_PR
}

// ---- OP_PGRP ------------------------------------------------------------
// compressed index of slot s over the slot bits of mask M (ascending)
template <uint32_t M>
__device__ __forceinline__ constexpr uint32_t pext_slot(uint32_t s) {
  uint32_t r = 0, k = 0;
  for (int b = 0; b < RB; ++b)
    if ((M >> b) & 1u) {
      if ((s >> b) & 1u) r |= 1u << k;
      ++k;
    }
  return r;
}

// A group of commuting Pauli rotations on the pairs (s, s ^ XR): per pair
// class the two-level rotation new_i = a_i + (cm1 a_i + sg(j) u_j a_j),
// new_j = a_j + (cm1 a_j + sg(i) u_i a_i), the small bracket first so that
// each amplitude is rounded once at its own size.  sg(i) = (-1)^{tp + bit s of
// the slot sign mask}, tp the parity of the thread's bits & z0.
// CT (generated kernels): the active class mask ACT, the slot sign mask SM
// and |x & z0| mod 2 KP are compile-time constants, so the slot signs become
// FMA source modifiers and only tp is applied at run time (once per class).
// Otherwise they are the uniform runtime values act, sm, kpar.
template <int XR, bool CT, uint32_t ACT, uint32_t SM, uint32_t KP, typename IDX>
__device__ __forceinline__ void pgrp_apply1(float2 (&a)[NSLOT], const PhaseGeo<IDX>& G, IDX z0,
                                            uint32_t kpar, uint32_t act, uint32_t sm, cw_t tab) {
  constexpr int LB = XR & (-XR);
  constexpr uint32_t HI = (uint32_t)XR & ~(uint32_t)LB;
  const uint32_t tp = parity((IDX)(G.gb & z0));
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & LB) continue;
    const uint32_t cls = pext_slot<HI>((uint32_t)s);
    if (CT ? !((ACT >> cls) & 1u) : !((act >> cls) & 1u)) continue;
    const int s1 = s ^ XR;
    const cw_t e = tab + PG_ENT * cls;
    const float cm1 = fl(e);
    float uir, uii, ujr, uji;
    if constexpr (CT) {
      // the thread sign once per class (common subexpression across the
      // class's pairs), the slot signs at compile time
      const float tir = tp ? -fl(e + 1) : fl(e + 1), tii = tp ? -fl(e + 2) : fl(e + 2);
      const float tjr = tp ? -fl(e + 3) : fl(e + 3), tji = tp ? -fl(e + 4) : fl(e + 4);
      const bool bi = (SM >> s) & 1u, bj = bi != ((KP & 1u) != 0u);
      uir = bi ? -tir : tir;
      uii = bi ? -tii : tii;
      ujr = bj ? -tjr : tjr;
      uji = bj ? -tji : tji;
    } else {
      const uint32_t sgi = tp ^ ((sm >> s) & 1u);
      const uint32_t sgj = sgi ^ kpar;
      uir = sgi ? -fl(e + 1) : fl(e + 1);
      uii = sgi ? -fl(e + 2) : fl(e + 2);
      ujr = sgj ? -fl(e + 3) : fl(e + 3);
      uji = sgj ? -fl(e + 4) : fl(e + 4);
    }
    const float2 x0 = a[s], x1 = a[s1];
    const float t0r = fmaf(-uji, x1.y, fmaf(cm1, x0.x, ujr * x1.x));
    const float t0i = fmaf(uji, x1.x, fmaf(cm1, x0.y, ujr * x1.y));
    const float t1r = fmaf(-uii, x0.y, fmaf(cm1, x1.x, uir * x0.x));
    const float t1i = fmaf(uii, x0.x, fmaf(cm1, x1.y, uir * x0.y));
    a[s] = make_float2(x0.x + t0r, x0.y + t0i);
    a[s1] = make_float2(x1.x + t1r, x1.y + t1i);
  }
}

template <int XR, int MODE, typename IDX>
__device__ __forceinline__ void pgrp_x(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                       const PhaseGeo<IDX>& G) {
  if constexpr (XR != 0) {
    const IDX z0 = ld_idx<IDX>(w + 2);
    const uint32_t kpar = w[4], act = w[5], sm = w[6];
    pgrp_apply1<XR, false, 0u, 0u, 0u, IDX>(a, G, z0, kpar, act, sm, w + PG_HDR);
    if constexpr (MODE == MODE_DUAL)
      pgrp_apply1<XR, false, 0u, 0u, 0u, IDX>(b, G, z0, kpar, act, sm, w + PG_HDR);
  }
}

// ---- OP_RED -------------------------------------------------------------
// +-1.0 from bit s of a uniform mask, built with scalar bit operations
__device__ __forceinline__ float sgn1(uint32_t mask, int s) {
  return __uint_as_float(0x3f800000u | (((mask >> s) & 1u) << 31));
}

// Per term: FP32 sum over the thread's 32 slots, then FP64 over the wave's
// lanes, the LDS accumulator of the wave and beyond.
// Cross reduction of a dual pass (one entry): one part (PART 0 real, 1
// imaginary) of sum_s sg(s) conj(b[s]) a[s ^ XR], the part the gradient of
// the rotation needs.  Two FP32 accumulators (even and odd slots) keep the
// FMA chains short.  CT: the slot sign mask SM is a compile-time constant
// (generated kernels; the signs become FMA source modifiers), else the
// uniform runtime mask `sm`.
template <int XR, int PART, bool CT, uint32_t SM>
__device__ __forceinline__ float cross_part(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                            uint32_t sm) {
  float v[2] = {0.f, 0.f};
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    const float2 u = b[s], w = a[s ^ XR];  // conj(u) * w
    float& acc = v[s & 1];
    if constexpr (CT) {
      const bool neg = (SM >> s) & 1u;
      if constexpr (PART == 0) {
        acc = fmaf(neg ? -u.x : u.x, w.x, acc);
        acc = fmaf(neg ? -u.y : u.y, w.y, acc);
      } else {
        acc = fmaf(neg ? -u.x : u.x, w.y, acc);
        acc = fmaf(neg ? u.y : -u.y, w.x, acc);
      }
    } else {
      const float q = PART == 0 ? u.x * w.x + u.y * w.y : u.x * w.y - u.y * w.x;
      acc = fmaf(sgn1(sm, s), q, acc);
    }
  }
  return v[0] + v[1];
}

template <int XR, typename IDX>
__device__ __forceinline__ void red_cross(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                          cw_t w, const PhaseGeo<IDX>& G, double* acc,
                                          uint32_t nred) {
  cw_t e = w + RED_HDR;
  const IDX zt = ld_idx<IDX>(e + 1);
  float v = e[4] ? cross_part<XR, 1, false, 0u>(a, b, e[3]) : cross_part<XR, 0, false, 0u>(a, b, e[3]);
  if (parity((IDX)(G.gb & zt))) v = -v;
  const double d = wave_sum_last_d((double)v);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 63) acc[wave * nred + e[0]] += d;
}

// the cross reduction of generated kernels: slot, z mask, sign mask and part
// fixed at compile time
template <int XR, int PART, uint32_t SM, uint32_t SLOT, uint32_t NRED, typename IDX>
__device__ __forceinline__ void jred_cross(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                           const PhaseGeo<IDX>& G, IDX zt, double* acc) {
  float v = cross_part<XR, PART, true, SM>(a, b, 0u);
  if (parity((IDX)(G.gb & zt))) v = -v;
  const double d = wave_sum_last_d((double)v);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 63) acc[wave * NRED + SLOT] += d;
}

template <int XR, int KIND, int MODE, typename IDX>
__device__ __forceinline__ void red_x(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT], cw_t w,
                                      const PhaseGeo<IDX>& G, double* acc, uint32_t nred) {
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
      if constexpr (MODE == MODE_DUAL) {
        const float2 u = b[s], v = a[s ^ XR];
        p[s] = make_float2(u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x);
      }
    }
  }
  const uint32_t nterms = w[3];
  cw_t e = w + RED_HDR;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t t = 0; t < nterms; ++t, e += RED_ENT) {
    const uint32_t slot = e[0];
    const IDX zt = ld_idx<IDX>(e + 1);
    const uint32_t smask = e[3];
    const uint32_t part = e[4];
    float v = 0.f;
    if (part == 0u) {
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        if (KIND == RED_PAIR && (s & LB)) continue;
        v = fmaf(sgn1(smask, s), p[s].x, v);
      }
    } else {
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        if (KIND == RED_PAIR && (s & LB)) continue;
        v = fmaf(sgn1(smask, s), p[s].y, v);
      }
    }
    v = parity((IDX)(G.gb & zt)) ? -v : v;
    const double d = wave_sum_last_d((double)v);
    if (lane == 63) acc[wave * nred + slot] += d;
  }
}

// Group reduction of a dual pass (kind RED_GROUP): Im <lambda| G |psi> for
// the generator G = sum_r g_r P_r of a rotation group with one shared
// parameter, in one sweep over the thread's pairs and one wave sum, instead
// of one reduction (two wave sums) per Pauli string.  CT: active class mask,
// slot sign mask and kpar fixed at compile time (as for pgrp_apply1).
template <int XR, bool CT, uint32_t ACT, uint32_t SM, uint32_t KP, typename IDX>
__device__ __forceinline__ float group_part(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                            const PhaseGeo<IDX>& G, IDX z0, uint32_t kpar,
                                            uint32_t act, uint32_t sm, cw_t tab) {
  constexpr int LB = XR & (-XR);
  constexpr uint32_t HI = (uint32_t)XR & ~(uint32_t)LB;
  const uint32_t tp = parity((IDX)(G.gb & z0));
  float vi = 0.f;
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (s & LB) continue;
    const uint32_t cls = pext_slot<HI>((uint32_t)s);
    if (CT ? !((ACT >> cls) & 1u) : !((act >> cls) & 1u)) continue;
    const int s1 = s ^ XR;
    cw_t c = tab + RGRP_ENT * cls;
    float wir, wii, wjr, wji;
    if constexpr (CT) {
      const float tir = tp ? -fl(c) : fl(c), tii = tp ? -fl(c + 1) : fl(c + 1);
      const float tjr = tp ? -fl(c + 2) : fl(c + 2), tji = tp ? -fl(c + 3) : fl(c + 3);
      const bool bi = (SM >> s) & 1u, bj = bi != ((KP & 1u) != 0u);
      wir = bi ? -tir : tir;
      wii = bi ? -tii : tii;
      wjr = bj ? -tjr : tjr;
      wji = bj ? -tji : tji;
    } else {
      const uint32_t sgi = tp ^ ((sm >> s) & 1u);
      const uint32_t sgj = sgi ^ kpar;
      wjr = sgj ? -fl(c + 2) : fl(c + 2);
      wji = sgj ? -fl(c + 3) : fl(c + 3);
      wir = sgi ? -fl(c) : fl(c);
      wii = sgi ? -fl(c + 1) : fl(c + 1);
    }
    // Im(conj(l) w a) = l.x (w.r a.y + w.i a.x) - l.y (w.r a.x - w.i a.y)
    const float2 li = b[s], lj = b[s1], ai = a[s], aj = a[s1];
    const float qjr = wjr * aj.x - wji * aj.y, qji = wjr * aj.y + wji * aj.x;
    const float qir = wir * ai.x - wii * ai.y, qii = wir * ai.y + wii * ai.x;
    vi = fmaf(li.x, qji, vi);
    vi = fmaf(-li.y, qjr, vi);
    vi = fmaf(lj.x, qii, vi);
    vi = fmaf(-lj.y, qir, vi);
  }
  return vi;
}

template <int XR, typename IDX>
__device__ __forceinline__ void red_group(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                          cw_t w, const PhaseGeo<IDX>& G, double* acc,
                                          uint32_t nred) {
  cw_t e = w + RED_HDR;
  cw_t t = e + RED_ENT;
  const float vi = group_part<XR, false, 0u, 0u, 0u, IDX>(a, b, G, ld_idx<IDX>(e + 1), t[0], t[1],
                                                          e[3], t + 2);
  const double d = wave_sum_last_d((double)vi);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 63) acc[wave * nred + e[0]] += d;
}

// the group reduction of generated kernels
template <int XR, uint32_t ACT, uint32_t SM, uint32_t KP, uint32_t SLOT, uint32_t NRED,
          typename IDX>
__device__ __forceinline__ void jred_group(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                           cw_t w, const PhaseGeo<IDX>& G, IDX z0, double* acc) {
  const float vi =
      group_part<XR, true, ACT, SM, KP, IDX>(a, b, G, z0, 0u, 0u, 0u, w + RED_HDR + RED_ENT + 2);
  const double d = wave_sum_last_d((double)vi);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 63) acc[wave * NRED + SLOT] += d;
}

template <int XR, int MODE, typename IDX>
__device__ __forceinline__ void red_dispatch_kind(const float2 (&a)[NSLOT],
                                                  const float2 (&b)[NSLOT], cw_t w,
                                                  const PhaseGeo<IDX>& G, double* acc,
                                                  uint32_t nred) {
  const uint32_t kind = w[2];
  if constexpr (MODE == MODE_DUAL) {
    if (kind == RED_CROSS) red_cross<XR, IDX>(a, b, w, G, acc, nred);
    if constexpr (XR != 0)
      if (kind == RED_GROUP) red_group<XR, IDX>(a, b, w, G, acc, nred);
  } else {
    if constexpr (XR == 0) {
      red_x<0, RED_DIAG, MODE, IDX>(a, b, w, G, acc, nred);
    } else {
      red_x<XR, RED_PAIR, MODE, IDX>(a, b, w, G, acc, nred);
    }
  }
}

// ---- dispatch helpers over the runtime register masks --------------------
template <int MODE, typename IDX, int X = 0>
__device__ __forceinline__ void prot_switch(uint32_t xr, float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                            cw_t w, const PhaseGeo<IDX>& G) {
  if constexpr (X < NSLOT) {
    if (xr == (uint32_t)X) {
      prot_x<X, MODE, IDX>(a, b, w, G);
      return;
    }
    prot_switch<MODE, IDX, X + 1>(xr, a, b, w, G);
  }
}

template <int MODE, typename IDX, int X = 0>
__device__ __forceinline__ void pgrp_switch(uint32_t xr, float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                            cw_t w, const PhaseGeo<IDX>& G) {
  if constexpr (X < NSLOT) {
    if (xr == (uint32_t)X) {
      pgrp_x<X, MODE, IDX>(a, b, w, G);
      __asm__ volatile("; pgrp arm %0" ::"n"(X));
      return;
    }
    pgrp_switch<MODE, IDX, X + 1>(xr, a, b, w, G);
  }
}

template <int MODE, typename IDX, int X = 0>
__device__ __forceinline__ void red_switch(uint32_t xr, const float2 (&a)[NSLOT],
                                           const float2 (&b)[NSLOT], cw_t w,
                                           const PhaseGeo<IDX>& G, double* acc, uint32_t nred) {
  if constexpr (X < NSLOT) {
    if (xr == (uint32_t)X) {
      red_dispatch_kind<X, MODE, IDX>(a, b, w, G, acc, nred);
      // keeps the arms distinct: merged arms index the slot array with a
      // variable, which moves it to scratch memory
      __asm__ volatile("; red arm %0" ::"n"(X));
      return;
    }
    red_switch<MODE, IDX, X + 1>(xr, a, b, w, G, acc, nred);
  }
}

template <int MODE, typename IDX>
__device__ __forceinline__ void u1_switch(uint32_t k, float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                          cw_t w, const PhaseGeo<IDX>& G) {
  switch (k) {
    case 0: u1_dispatch<0, MODE, IDX>(a, b, w, G); break;
    case 1: u1_dispatch<1, MODE, IDX>(a, b, w, G); break;
    case 2: u1_dispatch<2, MODE, IDX>(a, b, w, G); break;
    case 3: u1_dispatch<3, MODE, IDX>(a, b, w, G); break;
    default: u1_dispatch<4, MODE, IDX>(a, b, w, G); break;
  }
}

template <int MODE, typename IDX>
__device__ __forceinline__ void u2_switch(uint32_t kh, uint32_t kl, float2 (&a)[NSLOT],
                                          float2 (&b)[NSLOT], cw_t w, const PhaseGeo<IDX>& G) {
  switch (kh * 8u + kl) {
    case 8 * 1 + 0: u2_dispatch<1, 0, MODE, IDX>(a, b, w, G); break;
    case 8 * 2 + 0: u2_dispatch<2, 0, MODE, IDX>(a, b, w, G); break;
    case 8 * 2 + 1: u2_dispatch<2, 1, MODE, IDX>(a, b, w, G); break;
    case 8 * 3 + 0: u2_dispatch<3, 0, MODE, IDX>(a, b, w, G); break;
    case 8 * 3 + 1: u2_dispatch<3, 1, MODE, IDX>(a, b, w, G); break;
    case 8 * 3 + 2: u2_dispatch<3, 2, MODE, IDX>(a, b, w, G); break;
    case 8 * 4 + 0: u2_dispatch<4, 0, MODE, IDX>(a, b, w, G); break;
    case 8 * 4 + 1: u2_dispatch<4, 1, MODE, IDX>(a, b, w, G); break;
    case 8 * 4 + 2: u2_dispatch<4, 2, MODE, IDX>(a, b, w, G); break;
    default: u2_dispatch<4, 3, MODE, IDX>(a, b, w, G); break;
  }
}

// ---- the kernel -----------------------------------------------------------
// tile index -> global index of the tile's position 0: spread the tile number
// over the index bits outside the tile (L and above, skipping H).
template <typename IDX>
__device__ __forceinline__ IDX tile_base(uint32_t tile, uint32_t L, uint32_t m, cw_t H) {
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
__device__ __forceinline__ IDX pos_map(uint32_t p, uint32_t L, uint32_t m, cw_t H) {
  IDX v = (IDX)(p & ((1u << L) - 1u));
  for (uint32_t k = 0; k < m; ++k)
    if ((p >> (L + k)) & 1u) v |= ((IDX)1) << H[k];
  return v;
}


// ---- building blocks of generated pass kernels ------------------------------
// A generated kernel fixes the tile geometry, the register positions of each
// phase and the sequence and kind of its steps at compile time; the numbers
// (angles, phases, matrices, masks of Pauli strings) are still read from the
// pass program, so a plan whose parameters change keeps its compiled kernels.

// Tile geometry: L low bits and the high bits H (ascending).
template <typename IDX, int L, int... H>
struct TileMap {
  __device__ static __forceinline__ IDX pos(uint32_t p) {
    IDX v = (IDX)(p & ((1u << L) - 1u));
    int k = L;
    ((v |= ((IDX)((p >> k) & 1u)) << H, ++k), ...);
    return v;
  }
  __device__ static __forceinline__ IDX base(uint32_t tile) {
    IDX v = (IDX)tile << L;
    ((v = ((v >> H) << (H + 1)) | (v & ((((IDX)1) << H) - 1))), ...);
    return v;
  }
};

template <typename TM, typename IDX, int R0, int R1, int R2, int R3, int R4>
__device__ __forceinline__ void jit_geo(PhaseGeo<IDX>& G, IDX base, uint32_t tid) {
  constexpr int R[RB] = {R0, R1, R2, R3, R4};
  uint32_t tpos = tid;
  G.rmask = 0;
#pragma unroll
  for (int j = 0; j < RB; ++j) {
    tpos = ((tpos >> R[j]) << (R[j] + 1)) | (tpos & ((1u << R[j]) - 1u));
    G.sph[j] = swz(1u << R[j]);
    G.gR[j] = TM::pos(1u << R[j]);
    G.rmask |= G.gR[j];
  }
  G.bt = swz(tpos);
  G.gb = base | TM::pos(tpos);
}

// USED: the slots the phase's steps read or write (phases made of rotation
// groups touch only the pairs of their active classes); the others are
// neither loaded nor stored.
template <int MODE, uint32_t USED = 0xffffffffu, typename IDX>
__device__ __forceinline__ void jit_load_slots(float2 (&a)[NSLOT], float2 (&b)[NSLOT],
                                               const float2* tA, const float2* tB,
                                               const PhaseGeo<IDX>& G) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (!((USED >> s) & 1u)) continue;
    const uint32_t l = G.bt ^ ls_of(G, s);
    a[s] = tA[l];
    if constexpr (MODE == MODE_DUAL) b[s] = tB[l];
  }
}

template <int MODE, uint32_t USED = 0xffffffffu, typename IDX>
__device__ __forceinline__ void jit_store_slots(const float2 (&a)[NSLOT], const float2 (&b)[NSLOT],
                                                float2* tA, float2* tB, const PhaseGeo<IDX>& G) {
  uint32_t btw = G.bt;
  __asm__ volatile("" : "+v"(btw));
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (!((USED >> s) & 1u)) continue;
    const uint32_t l = btw ^ ls_of(G, s);
    tA[l] = a[s];
    if constexpr (MODE == MODE_DUAL) tB[l] = b[s];
  }
}

// 16-byte global accesses of the tile loops; NT = streaming (non-temporal) hint
typedef float nt_f4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ float4 gload4(const float4* p) {
  if constexpr (NT) {
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void gstore4(float4* p, float4 v) {
  if constexpr (NT) {
    const nt_f4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_f4*>(p));
  } else {
    *p = v;
  }
}

template <int TB, int MODE, typename IDX, typename TM, bool NT = false>
__device__ __forceinline__ void jit_load_tile(const float2* __restrict__ psi,
                                              const float2* __restrict__ lam, float2* tA,
                                              float2* tB, IDX base, uint32_t tid) {
  constexpr int NCHUNK = (1 << (TB + RB)) / 2 / (1 << TB);
  const IDX gthr = base | TM::pos(2u * tid);
  const uint32_t lthr = swz(2u * tid);
#pragma unroll
  for (int k = 0; k < NCHUNK; ++k) {
    const uint32_t pk = (uint32_t)k << (TB + 1);
    const IDX g = gthr | TM::pos(pk);
    const uint32_t l = lthr ^ swz(pk);
    *reinterpret_cast<float4*>(tA + l) = gload4<NT>(reinterpret_cast<const float4*>(psi) + (g >> 1));
    if constexpr (MODE == MODE_DUAL)
      *reinterpret_cast<float4*>(tB + l) = gload4<NT>(reinterpret_cast<const float4*>(lam) + (g >> 1));
  }
}

// Register-staged tile copies of the pipelined gate passes.  A thread moves
// the same 16-byte chunks as jit_load_tile / jit_store_tile (NSLOT / 2 of
// them), but through registers: the next tile's chunks are loaded into
// registers while the current tile is worked on in LDS, and the finished tile
// is read out of LDS into registers before the next one is written there, so
// that neither the loads nor the stores are waited for next to their issue.
constexpr int kTileChunks = NSLOT / 2;

template <int TB, typename IDX, typename TM, bool NT>
__device__ __forceinline__ void jit_fetch_tile(const float2* __restrict__ psi, IDX base,
                                               uint32_t tid, float4 (&r)[kTileChunks]) {
  const IDX gthr = base | TM::pos(2u * tid);
#pragma unroll
  for (int k = 0; k < kTileChunks; ++k) {
    const IDX g = gthr | TM::pos((uint32_t)k << (TB + 1));
    r[k] = gload4<NT>(reinterpret_cast<const float4*>(psi) + (g >> 1));
  }
}

template <int TB, typename IDX, typename TM, bool NT>
__device__ __forceinline__ void jit_put_global(float2* __restrict__ psi, IDX base, uint32_t tid,
                                               const float4 (&r)[kTileChunks]) {
  const IDX gthr = base | TM::pos(2u * tid);
#pragma unroll
  for (int k = 0; k < kTileChunks; ++k) {
    const IDX g = gthr | TM::pos((uint32_t)k << (TB + 1));
    gstore4<NT>(reinterpret_cast<float4*>(psi) + (g >> 1), r[k]);
  }
}

template <int TB>
__device__ __forceinline__ void jit_put_tile(float2* tA, uint32_t tid,
                                             const float4 (&r)[kTileChunks]) {
  const uint32_t lthr = swz(2u * tid);
#pragma unroll
  for (int k = 0; k < kTileChunks; ++k)
    *reinterpret_cast<float4*>(tA + (lthr ^ swz((uint32_t)k << (TB + 1)))) = r[k];
}

template <int TB>
__device__ __forceinline__ void jit_get_tile(const float2* tA, uint32_t tid,
                                             float4 (&r)[kTileChunks]) {
  const uint32_t lthr = swz(2u * tid);
#pragma unroll
  for (int k = 0; k < kTileChunks; ++k)
    r[k] = *reinterpret_cast<const float4*>(tA + (lthr ^ swz((uint32_t)k << (TB + 1))));
}

// One 4-byte load per 128-byte line (16 amplitudes) of a tile, two lines per
// thread: issued for a work-group's next tile while it works on the current
// one, so that the next tile's loads find their lines in the caches.  The
// caller keeps the loaded words live to the end of the iteration (an empty asm
// use), which is where the loads are waited for; barriers only wait for LDS.
template <int TB, int MODE, typename IDX, typename TM>
__device__ __forceinline__ void jit_touch_tile(const float2* __restrict__ psi,
                                               const float2* __restrict__ lam, IDX base,
                                               uint32_t tid, float (&t)[4]) {
  const IDX g0 = base | TM::pos((2u * tid) << 4);
  const IDX g1 = base | TM::pos((2u * tid + 1u) << 4);
  t[0] = reinterpret_cast<const float*>(psi)[2 * g0];
  t[1] = reinterpret_cast<const float*>(psi)[2 * g1];
  if constexpr (MODE == MODE_DUAL) {
    t[2] = reinterpret_cast<const float*>(lam)[2 * g0];
    t[3] = reinterpret_cast<const float*>(lam)[2 * g1];
  }
}

template <int TB, int MODE, typename IDX, typename TM, bool NT = false>
__device__ __forceinline__ void jit_store_tile(float2* __restrict__ psi, float2* __restrict__ lam,
                                               const float2* tA, const float2* tB, IDX base,
                                               uint32_t tid) {
  constexpr int NCHUNK = (1 << (TB + RB)) / 2 / (1 << TB);
  IDX gthr = base | TM::pos(2u * tid);
  uint32_t lthr = swz(2u * tid);
  __asm__ volatile("" : "+v"(gthr), "+v"(lthr));
#pragma unroll
  for (int k = 0; k < NCHUNK; ++k) {
    const uint32_t pk = (uint32_t)k << (TB + 1);
    const IDX g = gthr | TM::pos(pk);
    const uint32_t l = lthr ^ swz(pk);
    gstore4<NT>(reinterpret_cast<float4*>(psi) + (g >> 1), *reinterpret_cast<const float4*>(tA + l));
    if constexpr (MODE == MODE_DUAL)
      gstore4<NT>(reinterpret_cast<float4*>(lam) + (g >> 1), *reinterpret_cast<const float4*>(tB + l));
  }
}

template <int K, int MD, bool ZS, int MODE, typename IDX>
__device__ __forceinline__ void jrot(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                     const PhaseGeo<IDX>& G) {
  float t = fl(w + 3);
  if constexpr (ZS) t = parity((IDX)(G.gb & ld_idx<IDX>(w + 4))) ? -t : t;
  rot_apply<K, MD>(a, t);
  if constexpr (MODE == MODE_DUAL) rot_apply<K, MD>(b, t);
}

template <int K, int NENT, int MODE, typename IDX>
__device__ __forceinline__ void jph(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                    const PhaseGeo<IDX>& G) {
  float2 u = make_float2(fl(w + 4), fl(w + 5));
#pragma unroll
  for (int i = 0; i < NENT; ++i) {
    cw_t e = w + OPH_HDR + OPH_ENT * i;
    const float2 v = cmul(u, make_float2(fl(e + 1), fl(e + 2)));
    u = ((G.gb >> e[0]) & 1u) ? v : u;
  }
  ph_apply<K>(a, u);
  if constexpr (MODE == MODE_DUAL) ph_apply<K>(b, u);
}

template <uint32_t MASK, int MODE>
__device__ __forceinline__ void jsgn(float2 (&a)[NSLOT], float2 (&b)[NSLOT]) {
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if ((MASK >> s) & 1u) {
      a[s] = cneg(a[s]);
      if constexpr (MODE == MODE_DUAL) b[s] = cneg(b[s]);
    }
  }
}

template <int NENT, int MODE, typename IDX>
__device__ __forceinline__ void jthr(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                     const PhaseGeo<IDX>& G) {
  float2 F = make_float2(fl(w + 2), fl(w + 3));
#pragma unroll
  for (int i = 0; i < NENT; ++i) {
    cw_t e = w + THR_HDR + THR_ENT * i;
    const float2 v = cmul(F, make_float2(fl(e + 2), fl(e + 3)));
    F = (((G.gb >> e[0]) & (G.gb >> e[1])) & 1u) ? v : F;
  }
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    a[s] = cmul(a[s], F);
    if constexpr (MODE == MODE_DUAL) b[s] = cmul(b[s], F);
  }
}

template <int K, uint32_t TYPE, bool CTRL, int MODE, typename IDX>
__device__ __forceinline__ void ju1(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                    const PhaseGeo<IDX>& G) {
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  const M2f u = load_m2(w + 8);
  const bool ton = (G.gb & cm) == (cv & ~G.rmask);
  u1_apply<K, TYPE, CTRL, IDX>(a, u, G, cm, cv, ton);
  if constexpr (MODE == MODE_DUAL) u1_apply<K, TYPE, CTRL, IDX>(b, u, G, cm, cv, ton);
}

template <int KH, int KL, bool CTRL, int MODE, typename IDX>
__device__ __forceinline__ void ju2(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                    const PhaseGeo<IDX>& G) {
  const IDX cm = ld_idx<IDX>(w + 4);
  const IDX cv = ld_idx<IDX>(w + 6);
  const bool ton = (G.gb & cm) == (cv & ~G.rmask);
  u2_apply<KH, KL, CTRL, IDX>(a, w + 8, G, cm, cv, ton);
  if constexpr (MODE == MODE_DUAL) u2_apply<KH, KL, CTRL, IDX>(b, w + 8, G, cm, cv, ton);
}

template <int XR, int Q, int MODE, typename IDX>
__device__ __forceinline__ void jprot(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                      const PhaseGeo<IDX>& G) {
  const IDX z = ld_idx<IDX>(w + 2);
  const float c = fl(w + 4), sn = fl(w + 5);
  const uint32_t xzpar = w[6] & 1u;
  prot_apply1<XR, Q, IDX>(a, G, z, c, sn, xzpar);
  if constexpr (MODE == MODE_DUAL) prot_apply1<XR, Q, IDX>(b, G, z, c, sn, xzpar);
}

// rotation group of generated kernels: x, z0 and the class/sign structure
// fixed at compile time, the class constants read from the program
template <int XR, uint32_t ACT, uint32_t SM, uint32_t KP, int MODE, typename IDX>
__device__ __forceinline__ void jpgrp(float2 (&a)[NSLOT], float2 (&b)[NSLOT], cw_t w,
                                      const PhaseGeo<IDX>& G, IDX z0) {
  pgrp_apply1<XR, true, ACT, SM, KP, IDX>(a, G, z0, 0u, 0u, 0u, w + PG_HDR);
  if constexpr (MODE == MODE_DUAL)
    pgrp_apply1<XR, true, ACT, SM, KP, IDX>(b, G, z0, 0u, 0u, 0u, w + PG_HDR);
}

template <int NWAVE, int NRED>
__device__ __forceinline__ void jit_zero_acc(double* acc, uint32_t tid, uint32_t nt) {
  for (uint32_t i = tid; i < (uint32_t)(NWAVE * NRED); i += nt) acc[i] = 0.0;
}

template <int NWAVE, int NRED>
__device__ __forceinline__ void jit_write_partials(const double* acc, double* partials,
                                                   uint32_t tid, uint32_t nt) {
  __syncthreads();
  for (uint32_t r = tid; r < (uint32_t)NRED; r += nt) {
    double sum = 0.0;
    for (int wv = 0; wv < NWAVE; ++wv) sum += acc[wv * NRED + r];
    partials[(size_t)blockIdx.x * NRED + r] = sum;
  }
}


// ---- read-only expectation passes of generated kernels ----------------------
// Every term of the pass has a per-thread FP64 accumulator that lives across
// all tiles of the work-group: per tile the term's +-1 weighted sum over the
// thread's 32 slots (FP32, at most 32 products) is added in FP64; the lanes and
// waves are reduced once, at the end of the kernel.
template <int XR, uint32_t KIND>
__device__ __forceinline__ void jred_products(float2 (&p)[NSLOT], const float2 (&a)[NSLOT]) {
  constexpr int LB = XR & (-XR);
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if constexpr (KIND == RED_PAIR) {
      if (s & LB) continue;
      const float2 u = a[s ^ XR], v = a[s];  // conj(u) * v
      p[s] = make_float2(u.x * v.x + u.y * v.y, u.x * v.y - u.y * v.x);
    } else {
      p[s] = make_float2(a[s].x * a[s].x + a[s].y * a[s].y, 0.f);
    }
  }
}

// sum_s (-1)^{MASK bit s} part(p[s]) over the slots of the group (LB = lowest
// flip bit: pairs are counted once), times (-1)^{|gb & zt|}
template <uint32_t MASK, int PART, int LB, typename IDX>
__device__ __forceinline__ double jred_term(const float2 (&p)[NSLOT], IDX gb, IDX zt) {
  float v = 0.f;
#pragma unroll
  for (int s = 0; s < NSLOT; ++s) {
    if (LB && (s & LB)) continue;
    const float x = PART ? p[s].y : p[s].x;
    v = ((MASK >> s) & 1u) ? v - x : v + x;
  }
  return (double)(parity((IDX)(gb & zt)) ? -v : v);
}

// lanes, then waves (through LDS), then one partial per term and work-group
template <int K, int NWAVE>
__device__ __forceinline__ void jred_finish(double (&acc)[K], double* lds, double* partials,
                                            uint32_t tid, uint32_t nt) {
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += shfl_xor_d(v, o);
    if (lane == 0) lds[wave * K + k] = v;
  }
  __syncthreads();
  for (uint32_t r = tid; r < (uint32_t)K; r += nt) {
    double sum = 0.0;
    for (int wv = 0; wv < NWAVE; ++wv) sum += lds[wv * K + r];
    partials[(size_t)blockIdx.x * K + r] = sum;
  }
}

}  // namespace tqsv

#endif
