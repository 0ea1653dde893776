// tqsv_pass_impl.h - the precompiled tile-pass kernel, which interprets the
// step sequence of a pass program at run time (templates; each
// tqsv_pass_*.hip translation unit instantiates one variant).  The step
// functions are in tqsv_device.h; the program format is in tqsv_prog.h.
//
// Kernel modes: MODE_GATE (psi, gate steps), MODE_RED (psi read-only,
// reduction steps), MODE_DUAL (psi and lambda, gate steps applied to both and
// cross reductions <lambda|P|psi>).
#ifndef TQSV_PASS_IMPL_H
#define TQSV_PASS_IMPL_H

#include <hip/hip_runtime.h>

#include "tqsv_device.h"
#include "tqsv_internal.h"

namespace tqsv {

// The phases of one pass over the tile held in LDS (tA, and tB for lambda in
// dual passes): per phase, every thread gathers its 32 slots, runs the
// phase's steps in registers and writes the slots back.  `base` is the global
// index of the tile's position 0.
template <int TB, int MODE, typename IDX, bool LEAN>
__device__ __forceinline__ void run_phases(cw_t prog, float2* tA, float2* tB, double* acc,
                                           uint32_t nred, IDX base, uint32_t L, uint32_t m,
                                           cw_t H) {
  constexpr bool DUAL = MODE == MODE_DUAL;
  constexpr bool WRITE = MODE != MODE_RED;
  const uint32_t tid = threadIdx.x;
  const uint32_t nphase = prog[PH_NPHASE];
  cw_t ph = prog + PH_SIZE;
  for (uint32_t iph = 0; iph < nphase; ++iph) {
    const uint32_t nsteps = ph[PS_NSTEPS];
    PhaseGeo<IDX> G;
    uint32_t tpos = tid;
    G.rmask = 0;
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const uint32_t r = ph[PS_R + j];
      tpos = ((tpos >> r) << (r + 1)) | (tpos & ((1u << r) - 1u));
      G.sph[j] = swz(1u << r);
      G.gR[j] = pos_map<IDX>(1u << r, L, m, H);
      G.rmask |= G.gR[j];
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

    cw_t w = ph + PS_SIZE;
    for (uint32_t is = 0; is < nsteps; ++is) {
      const uint32_t op = w[0] & 0xffffu;
      const uint32_t nw = w[0] >> 16;
      if constexpr (MODE == MODE_RED) {
        // the slot registers do not change in a read-only pass; marking
        // them as rewritten here keeps the compiler from hoisting the pair
        // products of all 32 reduction arms out of the step loop
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) __asm__ volatile("" : "+v"(a[s].x), "+v"(a[s].y));
        if (op == OP_RED) red_switch<MODE, IDX>(w[1], a, b, w, G, acc, nred);
      } else {
        switch (op) {
          case OP_ROT: rot_step<MODE, IDX>(a, b, w, G); break;
          case OP_PH: ph_step<MODE, IDX>(a, b, w, G); break;
          case OP_SGN: sgn_step<MODE>(a, b, w); break;
          case OP_TAB: tab_step<MODE>(a, b, w); break;
          case OP_THR: thr_step<MODE, IDX>(a, b, w, G); break;
          case OP_SCL: scl_step<MODE>(a, b, w); break;
          case OP_CXP: cxp_switch<MODE, IDX>(w[1], a, b, w, G); break;
          case OP_U1: u1_switch<MODE, IDX>(w[1], a, b, w, G); break;
          case OP_U2:
            if constexpr (!LEAN) u2_switch<MODE, IDX>(w[1], w[2], a, b, w, G);
            break;
          case OP_D2: d2_step<MODE, IDX>(a, b, w, G); break;
          case OP_PROT:
            if constexpr (!LEAN) prot_switch<MODE, IDX>(w[1], a, b, w, G);
            break;
          case OP_PGRP:
            if constexpr (!LEAN) pgrp_switch<MODE, IDX>(w[1], a, b, w, G);
            break;
          case OP_RED:
            if constexpr (DUAL) red_switch<MODE, IDX>(w[1], a, b, w, G, acc, nred);
            break;
          default: break;
        }
      }
      w += nw;
    }

    if constexpr (WRITE) {
      // recompute the slot addresses instead of keeping 32 of them live
      // through the phase
      uint32_t btw = G.bt;
      __asm__ volatile("" : "+v"(btw));
#pragma unroll
      for (int s = 0; s < NSLOT; ++s) {
        const uint32_t l = btw ^ ls_of(G, s);
        tA[l] = a[s];
        if constexpr (DUAL) tB[l] = b[s];
      }
    }
    __syncthreads();
    ph = w;
  }
}

// LEAN variants leave out the dense 4x4 (OP_U2) and Pauli-rotation (OP_PROT, OP_PGRP)
// arms, which are most of the kernel's code: without them the kernel is about
// a fifth of the size and every branch stays within the short-branch range.
// The planner marks passes that contain neither step as lean.
template <int TB, int MODE, typename IDX, bool LEAN>
__global__ __launch_bounds__(1 << TB) void k_pass(float2* __restrict__ psi, float2* __restrict__ lam,
                                                  const uint32_t* __restrict__ prog_g,
                                                  double* __restrict__ partials, uint32_t ntiles) {
  constexpr bool DUAL = MODE == MODE_DUAL;
  constexpr bool WRITE = MODE != MODE_RED;
  constexpr bool REDS = MODE != MODE_GATE;
  constexpr int NT = 1 << TB;
  constexpr int T = TB + RB;
  constexpr uint32_t TSIZE = 1u << T;
  constexpr int NCHUNK = (TSIZE / 2) / NT;  // float4 chunks per thread
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tA = reinterpret_cast<float2*>(smem);
  float2* tB = tA + TSIZE;  // only used in dual passes
  double* acc = reinterpret_cast<double*>(smem + (DUAL ? 2 : 1) * TSIZE * sizeof(float2));

  const cw_t prog = (cw_t)prog_g;
  const uint32_t tid = threadIdx.x;
  const uint32_t L = prog[PH_L], m = prog[PH_M];
  const uint32_t nred = REDS ? prog[PH_NRED] : 0u;
  const cw_t H = prog + PH_H;
  constexpr uint32_t NWAVE = NT / 64 > 0 ? NT / 64 : 1;

  if constexpr (REDS)
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

    run_phases<TB, MODE, IDX, LEAN>(prog, tA, tB, acc, nred, base, L, m, H);

    if constexpr (WRITE) {
      // likewise for the chunk addresses of the tile
      IDX gw = base | gthr;
      uint32_t lw = lthr;
      __asm__ volatile("" : "+v"(gw), "+v"(lw));
#pragma unroll
      for (int k = 0; k < NCHUNK; ++k) {
        const uint32_t pk = (uint32_t)k << (TB + 1);
        const IDX g = gw | pos_map<IDX>(pk, L, m, H);
        const uint32_t l = lw ^ swz(pk);
        reinterpret_cast<float4*>(psi)[g >> 1] = *reinterpret_cast<const float4*>(tA + l);
        if constexpr (DUAL)
          reinterpret_cast<float4*>(lam)[g >> 1] = *reinterpret_cast<const float4*>(tB + l);
      }
    }
  }

  if constexpr (REDS) {
    if (nred) {
      __syncthreads();
      for (uint32_t r = tid; r < nred; r += NT) {
        double sum = 0.0;
        for (uint32_t wv = 0; wv < NWAVE; ++wv) sum += acc[wv * nred + r];
        partials[(size_t)blockIdx.x * nred + r] = sum;
      }
    }
  }
}

template <int TB, int MODE, typename IDX, bool LEAN>
hipError_t launch_one(const PassLaunch& pl, hipStream_t st) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<TB, MODE, IDX, LEAN>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_pass<TB, MODE, IDX, LEAN>), dim3(pl.grid), dim3(1 << TB), pl.lds_bytes,
                     st, pl.psi, pl.lam, pl.prog, pl.partials, pl.ntiles);
  return hipGetLastError();
}

}  // namespace tqsv

#endif
