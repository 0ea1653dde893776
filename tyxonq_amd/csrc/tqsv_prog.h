// tqsv_prog.h - word layout of the tile-pass programs that the host planner
// (tqsv_plan.cpp) writes and the pass kernel (tqsv_pass.hip) interprets.
//
// A state of 2^n amplitudes is swept in tiles of 2^T amplitudes.  A tile is
// the set of amplitudes that agree on every index bit outside the tile's bit
// set { 0 .. L-1 } U { H[0] .. H[m-1] } (T = L + m).  Tile position p maps to
// the global index bits as: p bits [0, L) -> index bits [0, L); p bit L+k ->
// index bit H[k].  A work-group holds one tile in LDS and runs the pass's
// phases over it; in a phase every thread holds 2^RB amplitudes in registers
// whose tile positions differ only in the RB "register positions" R[0..RB).
#ifndef TQSV_PROG_H
#define TQSV_PROG_H

#include <stdint.h>

namespace tqsv {

constexpr int RB = 5;              // register bits per phase
constexpr int NSLOT = 1 << RB;     // amplitudes per thread per state
constexpr int MAXH = 9;            // maximum number of non-contiguous tile bits

// pass flags
constexpr uint32_t PF_WRITE = 1u;  // write the tile back (gate pass); else read-only
constexpr uint32_t PF_DUAL = 2u;   // a second state (lambda) travels with psi

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

// step word 0: opcode | (nwords << 16); all offsets below are from the step start.
enum : uint32_t {
  // 2x2 matrix on register slot bit k.
  // [1]=k [2]=type [3]=ctrl(0/1) [4..6)=cmask(lo,hi) [6..8)=cval(lo,hi) [8..16)=m (floats)
  OP_U1 = 1,
  // 4x4 matrix on register slot bits (kh, kl), kh > kl, kh the MSB of the basis.
  // [1]=kh [2]=kl [3]=ctrl [4..6)=cmask [6..8)=cval [8..40)=m (floats)
  OP_U2 = 2,
  // diagonal: amp[s] *= tab[s] * prod_e d_e[2*bit(ga_e)+bit(gb_e)] where the
  // product runs over entries whose bits are outside the register positions.
  // [1]=ntab(0/1) [2]=nthr [3]=pad [4..68)=tab (32 complex, when ntab)
  // then nthr entries of {ga, gb, d (8 floats)} = 10 words
  OP_DIAG = 3,
  // sign flip by parity of popcount(i & (i >> shift) & mask), for batches of CZ.
  // [1]=cnt then cnt entries {mask lo, mask hi, shift} = 3 words
  OP_CZ = 4,
  // 2-bit diagonal on arbitrary index bits: amp *= d[2*bit(ga)+bit(gb)]
  // [1]=ga [2]=gb [3]=pad [4..12)=d (4 complex)
  OP_D2 = 5,
  // Pauli rotation exp(-i theta/2 P), flip bits inside the register slots.
  // [1]=xr (slot mask) [2..4)=z (lo,hi) [4]=cos(theta/2) [5]=sin(theta/2)
  // [6]=k with omega = i^k, k=|x&z| mod 4
  OP_PROT = 6,
  // reduction group.  [1]=xr [2]=kind [3]=nterms, then nterms entries
  // {slot, zt lo, zt hi, smask, part} = 5 words.
  //  kind 0: pairs s (lowest set bit of xr clear): v = conj(a[s^xr]) a[s]
  //  kind 1: all s: v = |a[s]|^2 (xr = 0)
  //  kind 2: all s: v = conj(b[s]) a[s^xr]  (b = lambda, DUAL passes)
  // term value += sum_s (-1)^{smask bit s} (-1)^{|gbase & zt|} part(v)
  OP_RED = 7,
};

// OP_U1 types
constexpr uint32_t U1_GENERAL = 0;
constexpr uint32_t U1_REAL = 1;    // all imaginary parts zero
constexpr uint32_t U1_RXLIKE = 2;  // real diagonal, imaginary off-diagonal

// reduction kinds
constexpr uint32_t RED_PAIR = 0;
constexpr uint32_t RED_DIAG = 1;
constexpr uint32_t RED_CROSS = 2;

constexpr int DIAG_HDR = 4;
constexpr int DIAG_TAB = 64;
constexpr int DIAG_ENT = 10;
constexpr int RED_HDR = 4;
constexpr int RED_ENT = 5;

}  // namespace tqsv

#endif
