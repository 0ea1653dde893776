// tqsv_prog.h - word layout of the tile-pass programs that the host planner
// (tqsv_plan.cpp, tqsv_ham.cpp) writes and the pass kernel (tqsv_pass_impl.h)
// interprets.
//
// A state of 2^n amplitudes is swept in tiles of 2^T amplitudes.  A tile is
// the set of amplitudes that agree on every index bit outside the tile's bit
// set { 0 .. L-1 } U { H[0] .. H[m-1] } (T = L + m).  Tile position p maps to
// the global index bits as: p bits [0, L) -> index bits [0, L); p bit L+k ->
// index bit H[k].  A work-group holds one tile in LDS and runs the pass's
// phases over it; in a phase every thread holds 2^RB amplitudes in registers
// ("slots") whose tile positions differ only in the RB "register positions"
// R[0..RB).  Slot s holds the amplitude whose register-position bits equal the
// bits of s; the other index bits ("thread bits") are fixed per thread and
// are read from the thread's global index gb (the index of its slot 0).
#ifndef TQSV_PROG_H
#define TQSV_PROG_H

#ifdef __HIPCC_RTC__
typedef unsigned int uint32_t;
typedef unsigned long long uint64_t;
#else
#include <stdint.h>
#endif

namespace tqsv {

constexpr int RB = 5;              // register bits per phase
constexpr int NSLOT = 1 << RB;     // amplitudes per thread per state
constexpr int MAXH = 9;            // maximum number of non-contiguous tile bits

// pass flags
constexpr uint32_t PF_WRITE = 1u;  // write the tile back (gate pass); else read-only
constexpr uint32_t PF_DUAL = 2u;   // a second state (lambda) travels with psi

// kernel modes (one kernel variant per mode)
constexpr int MODE_GATE = 0;  // psi only, gate steps, written back
constexpr int MODE_RED = 1;   // psi only, reduction steps, read-only
constexpr int MODE_DUAL = 2;  // psi and lambda, gate steps and cross reductions

// pass header (word offsets)
constexpr int PH_FLAGS = 0;
constexpr int PH_L = 1;
constexpr int PH_M = 2;
constexpr int PH_NPHASE = 3;
constexpr int PH_NRED = 4;        // reduction slots written by this pass
constexpr int PH_WORDS = 5;       // total words of the pass record
constexpr int PH_H = 8;           // H[0..MAXH)
constexpr int PH_SIZE = PH_H + MAXH + 3;  // 20 words

// phase header
constexpr int PS_NSTEPS = 0;
constexpr int PS_R = 1;           // R[0..RB), ascending
constexpr int PS_SIZE = 8;

// Step word 0: opcode | (nwords << 16); all offsets below are from the step
// start.  "k" is a slot bit (0..RB-1); "g" is a global index bit.
enum : uint32_t {
  // general 2x2 matrix on slot bit k, optionally controlled.
  // [1]=k [2]=type [3]=ctrl(0/1) [4..6)=cmask(lo,hi) [6..8)=cval(lo,hi) [8..16)=m (floats)
  OP_U1 = 1,
  // 4x4 matrix on slot bits (kh, kl), kh > kl, kh the MSB of the basis.
  // [1]=kh [2]=kl [3]=ctrl [4..6)=cmask [6..8)=cval [8..40)=m (floats)
  OP_U2 = 2,
  // rotation with deferred scale on slot bit k (x0 = a[s], x1 = a[s | 1<<k]):
  //   mode 0: y0 = x0 - t x1,   y1 = x1 + t x0
  //   mode 1: y0 = t x0 - x1,   y1 = t x1 + x0
  //   mode 2: y0 = x0 - i t x1, y1 = x1 - i t x0
  //   mode 3: y0 = t x0 - i x1, y1 = t x1 - i x0
  // t is negated in threads where parity(gb & zsig) = 1.
  // [1]=k [2]=mode [3]=t (float) [4..6)=zsig (lo,hi)
  OP_ROT = 3,
  // one-sided phase on slot bit k: a[s] *= u for slots with bit k set, where
  // u = U0 * prod_{entries e with bit g_e of gb set} E_e.
  // [1]=k [2]=nent [3]=pad [4..6)=U0 (re,im), then nent entries {g, Er, Ei}
  OP_PH = 4,
  // sign flips: a[s] = -a[s] for every slot s whose bit is set in the mask.  [1]=mask
  OP_SGN = 5,
  // uniform slot table: a[s] *= tab[s].  [1..4)=pad [4..68)=tab (32 complex)
  OP_TAB = 6,
  // per-thread complex factor: a[s] *= F for all s, F = F0 * prod_{entries e with bits
  // ga_e and gb_e of gb both set} E_e.
  // [1]=nent [2..4)=F0 (re,im), then nent entries {ga, gb, Er, Ei}
  OP_THR = 7,
  // 2-bit diagonal on arbitrary index bits: amp *= d[2*bit(ga)+bit(gb)]
  // [1]=ga [2]=gb [3]=pad [4..12)=d (4 complex)
  OP_D2 = 8,
  // Pauli rotation exp(-i theta/2 P) up to a real factor, flip bits inside the
  // register slots: a -> C a - i S P a.  [1]=xr (slot mask) [2..4)=z (lo,hi)
  // [4]=C [5]=S (one of them 1, the other tan or cot of theta/2; the factor
  // joins the deferred scale) [6]=k with omega = i^k, k=|x&z| mod 4
  OP_PROT = 9,
  // reduction group.  [1]=xr [2]=kind [3]=nterms, then nterms entries
  // {slot, zt lo, zt hi, smask, part} = 5 words.
  //  kind 0: pairs s (lowest set bit of xr clear): v = conj(a[s^xr]) a[s]
  //  kind 1: all s: v = |a[s]|^2 (xr = 0)
  //  kind 2: all s: v = conj(b[s]) a[s^xr]  (b = lambda, dual passes; one
  //    entry: the one part of <lambda|P|psi> the gradient needs)
  // term value += sum_s (-1)^{smask bit s} (-1)^{|gb & zt|} part(v)
  //  kind 3 (dual passes, one entry; zt = z0 of a rotation group): the
  //    imaginary part of <lambda| G |psi> for the group generator G =
  //    sum_r g_r P_r, summed over the pairs (i, j = i ^ x) of the active
  //    classes: conj(b_i) sg(j) V_j a_j + conj(b_j) sg(i) V_i a_i, with
  //    sg(k) = (-1)^{|k & z0|} (the entry's smask: its register-bit part).
  //    After the entry: [kpar] [active class mask] then per class
  //    {V_i re, V_i im, V_j re, V_j im}.
  OP_RED = 10,
  // real scale: a[s] *= f for all s.  [1]=f (float)
  OP_SCL = 11,
  // controlled swap on slot bit k: a[s] <-> a[s | 1<<k] for the slots s (bit
  // k clear) whose control holds.  [1]=k [2]=creg: the control's slot bit
  // (0..RB-1), CX_THREAD when the control is a thread bit (tested on gb),
  // CX_NONE without control [3]=control value (0/1) [4..6)=cmask (lo,hi)
  // [6..8)=cval (lo,hi), both on global index bits [8]=smask: bit s set for
  // the slots whose register bits pass the control (all slots unless creg is
  // a slot bit)
  OP_CXP = 12,
  // commuting Pauli rotations sharing the flip mask x (register slots),
  // prod_r exp(-i a_r/2 P_r) with P_r = (x, z0 ^ d_r), d_r inside x.  On the
  // pair (i, j = i ^ x), i the member with the lowest bit of x clear, it is
  //   new_i = a_i + cm1 a_i + sg(j) u_j a_j,  new_j = a_j + cm1 a_j + sg(i) u_i a_i
  // with sg(k) = (-1)^{|k & z0|} and the class constants of the pair's
  // pattern on x (cm1 = cos - 1, u = -i sin/w times the generator entry).
  // [1]=xr [2..4)=z0 (lo,hi) [4]=|x & z0| mod 2 [5]=active class mask
  // [6]=slot sign mask (bit s = parity of slot s's register bits & z0; the
  // thread's bits give the rest of sg), then per class (pattern of the x bits
  // above the lowest one, compressed) {cm1, u_i re, u_i im, u_j re, u_j im};
  // inactive classes are the identity.
  OP_PGRP = 13,
};

constexpr uint32_t CX_THREAD = RB;
constexpr uint32_t CX_NONE = RB + 1;

// OP_U1 types
constexpr uint32_t U1_GENERAL = 0;
constexpr uint32_t U1_REAL = 1;    // all imaginary parts zero
constexpr uint32_t U1_RXLIKE = 2;  // real diagonal, imaginary off-diagonal

// reduction kinds
constexpr uint32_t RED_PAIR = 0;
constexpr uint32_t RED_DIAG = 1;
constexpr uint32_t RED_CROSS = 2;
constexpr uint32_t RED_GROUP = 3;
constexpr int RGRP_ENT = 4;

constexpr int OPH_HDR = 6;
constexpr int OPH_ENT = 3;
constexpr int THR_HDR = 4;
constexpr int THR_ENT = 4;
constexpr int TAB_OFF = 4;
constexpr int RED_HDR = 4;
constexpr int RED_ENT = 5;
constexpr int PG_HDR = 7;
constexpr int PG_ENT = 5;

}  // namespace tqsv

#endif
