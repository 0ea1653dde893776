// tqsv_plan.cpp - gate model, 1-qubit merging, and the pass/phase scheduler
// that packs a circuit into fused tile passes.
//
// Scheduling model.  Every gate uses some index bits non-diagonally (its
// targets: they must sit in the register positions of the phase that applies
// it) and some only diagonally (controls, diagonal targets, Z parts of Pauli
// strings: evaluated from the global index, they may be anywhere).  Two gates
// commute when every bit they share is used diagonally by both.  A pass is
// built by scanning the remaining gates in order and taking every gate that
// commutes with all gates skipped so far and whose targets fit the tile; the
// phases of a pass are carved out of its gates the same way with the register
// positions as capacity.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>

#include "tqsv_planner.h"

namespace tqsv {

namespace {

inline int popc(uint64_t v) { return __builtin_popcountll(v); }
inline int ctz(uint64_t v) { return __builtin_ctzll(v); }
inline uint64_t bitm(int b) { return 1ull << b; }
inline uint64_t lowmask(int L) { return L >= 64 ? ~0ull : ((1ull << L) - 1ull); }
inline bool is0(cd v) { return v.real() == 0.0 && v.imag() == 0.0; }
inline bool is1(cd v) { return v.real() == 1.0 && v.imag() == 0.0; }

uint32_t mat1_type(const Gate& g) {
  bool real = true;
  for (int j = 0; j < 4; ++j)
    if (g.m[j].imag() != 0.0) real = false;
  if (real) return U1_REAL;
  if (g.m[0].imag() == 0.0 && g.m[3].imag() == 0.0 && g.m[1].real() == 0.0 &&
      g.m[2].real() == 0.0)
    return U1_RXLIKE;
  return U1_GENERAL;
}

// columns orthogonal and of equal norm: a nonzero multiple of a unitary
bool unitary_like(const cd m[4]) {
  const double r2 = std::norm(m[0]) + std::norm(m[2]);
  if (!(r2 > 1e-300) || !std::isfinite(r2)) return false;
  if (std::abs(std::norm(m[1]) + std::norm(m[3]) - r2) > 1e-9 * r2) return false;
  return std::abs(std::conj(m[0]) * m[1] + std::conj(m[2]) * m[3]) <= 1e-9 * r2;
}

// A fixed (unparameterised) gate with at most one control and an
// antidiagonal matrix of unit-modulus entries: a controlled swap times a
// diagonal.  The swap is a register renaming and the diagonal joins the
// pending phase polynomial.
bool perm_like(const Gate& g) {
  if (g.kind != GK_MAT1 || g.param >= 0 || popc(g.cmask) > 1) return false;
  if (!is0(g.m[0]) || !is0(g.m[3])) return false;
  return std::abs(std::abs(g.m[1]) - 1.0) <= 1e-12 && std::abs(std::abs(g.m[2]) - 1.0) <= 1e-12;
}

// approximate VALU work per amplitude, for the per-pass budget
int gate_cost(const Gate& g) {
  switch (g.kind) {
    case GK_MAT1:
      if (perm_like(g)) return 1;
      if (g.cmask == 0 && unitary_like(g.m)) return 4;  // rotation + one-sided phase
      return (mat1_type(g) == U1_GENERAL ? 8 : 4) + (g.cmask ? 4 : 0);
    case GK_MAT2: return g.cmask ? 20 : 16;
    case GK_DIAG2: return 1;
    case GK_PROT: return 8;
    case GK_RED: return g.nmem > 0 ? 6 : 4;
    case GK_PGRP: {
      PgrpClass cls[kMaxGroup];
      const int nc = pgrp_classes(g, cls, false);
      int act = 0;
      for (int c = 0; c < nc; ++c) act += cls[c].active ? 1 : 0;
      return std::max(1, (9 * act + nc - 1) / nc);
    }
  }
  return 8;
}

// usage bits of an unclassified gate (for merging)
uint64_t raw_used(const Gate& g) {
  uint64_t u = g.cmask | g.x | g.z;
  if (g.t0 >= 0) u |= bitm(g.t0);
  if (g.t1 >= 0) u |= bitm(g.t1);
  return u;
}

void to_diag2(Gate& g, int a, int b, const cd d[4]) {
  g.kind = GK_DIAG2;
  g.t0 = a;
  g.t1 = b;
  for (int j = 0; j < 4; ++j) g.m[j] = d[j];
  for (int j = 4; j < 16; ++j) g.m[j] = cd(0.0, 0.0);
  g.cmask = g.cval = 0;
  g.nb = 0;
  g.db = bitm(a) | bitm(b);
}

void classify(Gate& g) {
  g.nb = g.db = 0;
  switch (g.kind) {
    case GK_MAT1: {
      const bool diag = is0(g.m[1]) && is0(g.m[2]);
      if (diag && g.cmask == 0) {
        const cd d[4] = {g.m[0], cd(1, 0), cd(1, 0), g.m[3]};
        to_diag2(g, g.t0, g.t0, d);
        break;
      }
      if (diag && popc(g.cmask) == 1) {
        const int cb = ctz(g.cmask);
        const int cv = (int)((g.cval >> cb) & 1u);
        cd d[4];
        for (int c = 0; c < 2; ++c)
          for (int t = 0; t < 2; ++t) d[2 * c + t] = (c == cv) ? (t ? g.m[3] : g.m[0]) : cd(1, 0);
        to_diag2(g, cb, g.t0, d);
        break;
      }
      g.nb = bitm(g.t0);
      g.db = g.cmask;
      break;
    }
    case GK_MAT2: {
      bool diag = true;
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
          if (r != c && !is0(g.m[4 * r + c])) diag = false;
      if (diag && g.cmask == 0) {
        const cd d[4] = {g.m[0], g.m[5], g.m[10], g.m[15]};
        to_diag2(g, g.t0, g.t1, d);
        break;
      }
      if (!diag) {
        // controlled forms: block diagonal in one bit with an identity block
        for (int which = 0; which < 2; ++which) {
          // which = 0: blocks by t0 (msb), target t1; which = 1: blocks by t1 (lsb), target t0
          auto rc = [&](int v, int tb) { return which == 0 ? (2 * v + tb) : (2 * tb + v); };
          bool blk = true;
          for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
              const int vr = which == 0 ? (r >> 1) : (r & 1);
              const int vc = which == 0 ? (c >> 1) : (c & 1);
              if (vr != vc && !is0(g.m[4 * r + c])) blk = false;
            }
          if (!blk) continue;
          for (int v = 0; v < 2; ++v) {
            // is block v the identity?
            const cd b00 = g.m[4 * rc(v, 0) + rc(v, 0)], b01 = g.m[4 * rc(v, 0) + rc(v, 1)];
            const cd b10 = g.m[4 * rc(v, 1) + rc(v, 0)], b11 = g.m[4 * rc(v, 1) + rc(v, 1)];
            if (!(is1(b00) && is0(b01) && is0(b10) && is1(b11))) continue;
            const int o = 1 - v;
            cd u[4] = {g.m[4 * rc(o, 0) + rc(o, 0)], g.m[4 * rc(o, 0) + rc(o, 1)],
                       g.m[4 * rc(o, 1) + rc(o, 0)], g.m[4 * rc(o, 1) + rc(o, 1)]};
            const int ctrl = which == 0 ? g.t0 : g.t1;
            const int tgt = which == 0 ? g.t1 : g.t0;
            g.kind = GK_MAT1;
            g.t0 = tgt;
            g.t1 = -1;
            g.cmask |= bitm(ctrl);
            if (o) g.cval |= bitm(ctrl);
            for (int j = 0; j < 4; ++j) g.m[j] = u[j];
            for (int j = 4; j < 16; ++j) g.m[j] = cd(0, 0);
            classify(g);
            return;
          }
        }
      }
      g.nb = bitm(g.t0) | bitm(g.t1);
      g.db = g.cmask;
      break;
    }
    case GK_DIAG2: g.db = bitm(g.t0) | bitm(g.t1); break;
    case GK_PROT:
    case GK_RED:
      if (g.kind == GK_RED && g.nmem > 0) {
        uint64_t zu = g.z;
        for (int r = 0; r < g.nmem; ++r) zu |= g.mz[r];
        g.nb = g.x;
        g.db = zu & ~g.x;
        break;
      }
      g.nb = g.x;
      g.db = g.z & ~g.x;
      break;
    case GK_PGRP: {
      uint64_t zu = g.z;
      for (int r = 0; r < g.nmem; ++r) zu |= g.mz[r];
      g.nb = g.x;
      g.db = zu & ~g.x;
      break;
    }
  }
  g.cost = gate_cost(g);
}

bool fusable(const Gate& g) {
  if (g.kind == GK_PROT || g.kind == GK_RED || g.kind == GK_PGRP) return popc(g.x) <= RB;
  return true;
}

struct Emit {
  std::vector<uint32_t>& w;
  std::vector<uint32_t>* fpos;  // positions of value words, or null
  explicit Emit(std::vector<uint32_t>& v, std::vector<uint32_t>* fp = nullptr) : w(v), fpos(fp) {}
  void u(uint32_t v) { w.push_back(v); }
  void f(float x) {
    uint32_t v;
    std::memcpy(&v, &x, 4);
    if (fpos) fpos->push_back((uint32_t)w.size());
    w.push_back(v);
  }
  void c(cd v) {
    f((float)v.real());
    f((float)v.imag());
  }
  void u64(uint64_t v) {
    u((uint32_t)v);
    u((uint32_t)(v >> 32));
  }
  size_t begin(uint32_t op) {
    const size_t at = w.size();
    w.push_back(op);
    return at;
  }
  void end(size_t at) { w[at] |= (uint32_t)((w.size() - at) << 16); }
};

struct TileGeo {
  int L = 0;
  std::vector<int> H;
  int T = 0;
  int pos(int b) const {
    if (b < L) return b;
    for (size_t k = 0; k < H.size(); ++k)
      if (H[k] == b) return L + (int)k;
    return -1;
  }
  uint32_t posmask(uint64_t bits) const {
    uint32_t m = 0;
    for (uint64_t v = bits; v; v &= v - 1) m |= 1u << pos(ctz(v));
    return m;
  }
};

bool conflicts(const Gate& g, uint64_t blkN, uint64_t blkD) {
  return ((g.nb | g.db) & blkN) || (g.nb & blkD);
}
void block(const Gate& g, uint64_t& blkN, uint64_t& blkD) {
  blkN |= g.nb;
  blkD |= g.db;
}

// ---- pending diagonal ------------------------------------------------------
// Diagonal factors are not applied where they occur.  They are kept as the
// exponent of a quadratic pseudo-Boolean polynomial over index bits,
//   D(i) = exp(c0 + sum_q lin_q b_q(i) + sum_{p<q} quad_pq b_p(i) b_q(i)),
// so that the true state is D * (stored state).  A term is applied only when
// a non-diagonal step is about to act on one of its bits (it commutes with
// everything applied in between), and whatever remains is applied once at
// the end of the plan.  Only unit-modulus factors enter lin/quad; c0 also
// carries the real scale deferred by the rotations.
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPi = 3.1415926535897932384626433832795;

inline cd wrap(cd v) { return cd(v.real(), std::remainder(v.imag(), kTwoPi)); }
inline bool trivial(cd v) { return std::abs(v.real()) < 1e-14 && std::abs(v.imag()) < 1e-13; }
// exp(v) == -1 (a sign)
inline bool is_sign(cd v) {
  return std::abs(v.real()) < 1e-14 && std::abs(std::abs(v.imag()) - kPi) < 1e-12;
}
inline cd cexp_snap(cd v) {
  if (trivial(v)) return cd(1, 0);
  if (is_sign(v)) return cd(-1, 0);
  return std::exp(v);
}

struct Pend {
  cd c0{0, 0};
  cd lin[64];
  std::vector<cd> quad;  // [p * 64 + q], p < q
  uint64_t lmask = 0;
  uint64_t qm[64] = {};  // qm[p]: bits q with a quad(p, q) term
  int nrot = 0;          // rotations since the stored state was last rescaled
  Pend() : quad(64 * 64) {}

  bool empty() const {
    if (lmask || !trivial(wrap(c0))) return false;
    for (int p = 0; p < 64; ++p)
      if (qm[p]) return false;
    return true;
  }
  void add_lin(int q, cd v) {
    lin[q] = wrap(lin[q] + v);
    if (trivial(lin[q])) {
      lin[q] = cd(0, 0);
      lmask &= ~bitm(q);
    } else {
      lmask |= bitm(q);
    }
  }
  cd& qref(int p, int q) { return p < q ? quad[p * 64 + q] : quad[q * 64 + p]; }
  void add_quad(int p, int q, cd v) {
    if (p == q) {
      add_lin(p, v);
      return;
    }
    cd& r = qref(p, q);
    r = wrap(r + v);
    if (trivial(r)) {
      r = cd(0, 0);
      qm[p] &= ~bitm(q);
      qm[q] &= ~bitm(p);
    } else {
      qm[p] |= bitm(q);
      qm[q] |= bitm(p);
    }
  }
  void clear_quad(int p, int q) {
    qref(p, q) = cd(0, 0);
    qm[p] &= ~bitm(q);
    qm[q] &= ~bitm(p);
  }
  void clear_lin(int q) {
    lin[q] = cd(0, 0);
    lmask &= ~bitm(q);
  }
  // Moving a permutation pi of the index bits through D: pi D = D' pi with
  // D'(z) = D(pi(z)) (pi is its own inverse here).
  // flip: z_t -> 1 - z_t (an X on bit t).
  void flip(int t) {
    if (lmask & bitm(t)) {
      c0 = wrap(c0 + lin[t]);
      lin[t] = wrap(-lin[t]);
    }
    for (uint64_t u = qm[t]; u; u &= u - 1) {
      const int q = ctz(u);
      cd& r = qref(t, q);
      const cd v = r;
      r = wrap(-v);
      add_lin(q, v);
    }
  }
  // cx: z_t -> z_t xor z_c = z_t + z_c - 2 z_t z_c (a CX with control c).
  // lin_t z_t stays and adds lin_t z_c - 2 lin_t z_t z_c; quad_tc z_t z_c
  // becomes quad_tc z_c - quad_tc z_t z_c; quad_tq z_t z_q (q != c) adds
  // quad_tq z_c z_q and a cubic term -2 quad_tq z_t z_c z_q, which vanishes
  // (exp(-2 i pi) = 1) only when quad_tq is a sign: cx_ok tests that.
  bool cx_ok(int t, int c) const {
    for (uint64_t u = qm[t] & ~bitm(c); u; u &= u - 1) {
      const int q = ctz(u);
      if (!is_sign(t < q ? quad[t * 64 + q] : quad[q * 64 + t])) return false;
    }
    return true;
  }
  void cx(int t, int c) {
    const cd lt = (lmask & bitm(t)) ? lin[t] : cd(0, 0);
    const cd qtc = (qm[t] & bitm(c)) ? qref(t, c) : cd(0, 0);
    for (uint64_t u = qm[t] & ~bitm(c); u; u &= u - 1) {
      const int q = ctz(u);
      add_quad(c, q, qref(t, q));
    }
    add_lin(c, lt + qtc);
    clear_quad(t, c);
    add_quad(t, c, -2.0 * lt - qtc);
  }
  // d[2*bit(a) + bit(b)]; a == b is a one-bit diagonal (d[0], d[3]).
  // Returns false when a factor is zero or not of unit modulus.
  bool add_diag2(int a, int b, const cd d[4]) {
    const int used[4] = {0, a == b ? 3 : 1, 2, 3};
    cd l[4];
    for (int j : used) {
      if (std::abs(std::abs(d[j]) - 1.0) > 1e-12) return false;
      l[j] = cd(0.0, std::arg(d[j]));
    }
    c0 = wrap(c0 + l[0]);
    if (a == b) {
      add_lin(a, l[3] - l[0]);
      return true;
    }
    add_lin(a, l[2] - l[0]);
    add_lin(b, l[1] - l[0]);
    add_quad(a, b, l[3] - l[2] - l[1] + l[0]);
    return true;
  }
};

// Factorisation U = g k D_L K D_R of a 2x2 matrix proportional to a unitary:
// D_L = diag(1, eL), D_R = diag(1, eR), K one of the four rotation forms of
// OP_ROT with parameter t and deferred scale k (cos or sin of the angle).
struct RotDec {
  int mode = 0;      // OP_ROT mode
  double t = 0.0;
  cd eL{1, 0}, eR{1, 0};
  cd logk{0, 0};     // log(g k)
};

inline cd unit(cd v) {
  const double a = std::abs(v);
  return a > 0 ? v / a : cd(1, 0);
}
inline cd snap_unit(cd v) {
  if (std::abs(v - cd(1, 0)) < 1e-12) return cd(1, 0);
  if (std::abs(v + cd(1, 0)) < 1e-12) return cd(-1, 0);
  if (std::abs(v - cd(0, 1)) < 1e-12) return cd(0, 1);
  if (std::abs(v + cd(0, 1)) < 1e-12) return cd(0, -1);
  return v;
}

bool decompose_u1(const cd m[4], RotDec& d, int hint) {
  if (!unitary_like(m)) return false;
  const double r = std::sqrt(std::norm(m[0]) + std::norm(m[2]));
  const double c = std::abs(m[0]) / r, s = std::abs(m[2]) / r;
  const cd g = std::abs(m[0]) > 0 ? r * unit(m[0]) : r * unit(m[2]);
  cd eL, eR;
  if (s >= c) {
    eL = unit(m[2] / g);
    eR = unit(-m[1] / g);
  } else {
    eL = std::abs(m[2]) > 0 ? unit(m[2] / g) : cd(1, 0);
    eR = unit(m[3] / (g * eL));
  }
  eL = snap_unit(eL);
  eR = snap_unit(eR);
  bool cot = s > c;
  if (hint == 0 && c > 0 && s <= 4.0 * c) cot = false;
  if (hint == 1 && s > 0 && c <= 4.0 * s) cot = true;
  // The device receives t in float.  The deferred scale is taken from that
  // rounded t, k = 1/sqrt(1 + t_f^2) (= cos for the tan forms, sin for the cot
  // forms), so that every applied rotation is exactly unitary and rounding t
  // only moves the angle (by ~1e-7 rad) instead of the norm.
  d.t = (double)(float)(cot ? c / s : s / c);
  d.logk = std::log(g) - 0.5 * std::log1p(d.t * d.t);
  const bool rsign = std::abs(eR.imag()) == 0.0;         // eR = +-1: real form
  const bool isign = !rsign && std::abs(eR.real()) == 0.0;  // eR = +-i: imaginary form
  if (isign) {
    // K_i = [[c, -is], [-is, c]]: eL' = i eL, eR' = -i eR
    d.mode = cot ? 3 : 2;
    d.eL = snap_unit(cd(0, 1) * eL);
    d.eR = snap_unit(cd(0, -1) * eR);
  } else {
    d.mode = cot ? 1 : 0;
    d.eL = eL;
    d.eR = eR;
  }
  return true;
}

// Slot map of one phase.
struct Slots {
  int of_bit[64];
  uint64_t reg = 0;
  int bit_of[RB];
  Slots(const std::array<uint32_t, RB>& Rp, const TileGeo& geo) {
    for (int b = 0; b < 64; ++b) of_bit[b] = -1;
    for (int j = 0; j < RB; ++j) bit_of[j] = -1;
    for (int b = 0; b < 64; ++b) {
      const int p = geo.pos(b);
      if (p < 0) continue;
      for (int j = 0; j < RB; ++j)
        if ((int)Rp[j] == p) {
          of_bit[b] = j;
          bit_of[j] = b;
          reg |= bitm(b);
        }
    }
  }
};

struct PhaseEmitter {
  Pend& P;
  const Slots& S;
  Emit& e;

  // apply every pending term whose two bits are both register bits
  void regreg() {
    cd tab[NSLOT];
    for (int s = 0; s < NSLOT; ++s) tab[s] = cd(1, 0);
    bool any = false;
    for (uint64_t v = S.reg; v; v &= v - 1) {
      const int p = ctz(v);
      for (uint64_t u = P.qm[p] & S.reg & ~lowmask(p + 1); u; u &= u - 1) {
        const int q = ctz(u);
        const cd f = cexp_snap(P.qref(p, q));
        const int kp = S.of_bit[p], kq = S.of_bit[q];
        for (int s = 0; s < NSLOT; ++s)
          if (((s >> kp) & 1) && ((s >> kq) & 1)) tab[s] *= f;
        P.clear_quad(p, q);
        any = true;
      }
    }
    if (!any) return;
    bool signs = true;
    uint32_t mask = 0;
    for (int s = 0; s < NSLOT; ++s) {
      if (tab[s] == cd(1, 0)) continue;
      if (tab[s] == cd(-1, 0)) {
        mask |= 1u << s;
        continue;
      }
      signs = false;
    }
    if (signs) {
      if (!mask) return;
      const size_t at = e.begin(OP_SGN);
      e.u(mask);
      e.end(at);
      return;
    }
    const size_t at = e.begin(OP_TAB);
    for (int j = 1; j < TAB_OFF; ++j) e.u(0);
    for (int s = 0; s < NSLOT; ++s) e.c(tab[s]);
    e.end(at);
  }

  // apply every pending term on register bit q
  void extract(int q) {
    bool rr = false;
    for (uint64_t u = P.qm[q]; u; u &= u - 1)
      if (S.reg & bitm(ctz(u))) rr = true;
    if (rr) regreg();
    const cd U0 = cexp_snap(P.lin[q]);
    std::vector<std::pair<int, cd>> ents;
    for (uint64_t u = P.qm[q]; u; u &= u - 1) {
      const int p = ctz(u);
      ents.push_back({p, cexp_snap(P.qref(q, p))});
    }
    for (const auto& x : ents) P.clear_quad(q, x.first);
    P.clear_lin(q);
    if (U0 == cd(1, 0) && ents.empty()) return;
    const size_t at = e.begin(OP_PH);
    e.u((uint32_t)S.of_bit[q]);
    e.u((uint32_t)ents.size());
    e.u(0);
    e.c(U0);
    for (const auto& x : ents) {
      e.u((uint32_t)x.first);
      e.c(x.second);
    }
    e.end(at);
  }

  // Before a rotation on q: when every pending factor on q is a sign, it
  // stays pending and flips the rotation instead (Z K Z = K with t -> -t,
  // times -1 for the cot forms); returns the flip mask and uniform flip.
  bool frame(int q, uint64_t& zsig, bool& uflip) {
    bool rr = false;
    for (uint64_t u = P.qm[q]; u; u &= u - 1)
      if (S.reg & bitm(ctz(u))) rr = true;
    if (rr) regreg();
    zsig = 0;
    uflip = false;
    if (P.lmask & bitm(q)) {
      if (!is_sign(P.lin[q])) return false;
      uflip = true;
    }
    for (uint64_t u = P.qm[q]; u; u &= u - 1) {
      const int p = ctz(u);
      if (!is_sign(P.qref(q, p))) return false;
      zsig |= bitm(p);
    }
    return true;
  }

  // Each rotation grows the stored amplitudes by at most sqrt(1 + t^2) <= 4.2
  // (|t| <= 4); every 16 rotations the accumulated real scale is applied.
  // The step is emitted at fixed places, not when the scale gets small, so
  // that the program structure does not depend on the angles.
  // The factor is rounded to float; what rounding leaves out stays in c0 and
  // is applied with the next factor (error feedback), so that the norm drift
  // stays at one rounding however many rescales a plan has.
  void rescale() {
    if (++P.nrot < 16) return;
    const size_t at = e.begin(OP_SCL);
    const float f = (float)std::exp(P.c0.real());
    e.f(f);
    e.end(at);
    P.c0 = cd(P.c0.real() - std::log((double)f), P.c0.imag());
    P.nrot = 0;
  }

  void rot(const Gate& g, const RotDec& d) {
    const int q = g.t0;
    P.add_lin(q, cd(0, std::arg(d.eR)));
    uint64_t zsig = 0;
    bool uflip = false;
    const bool fr = frame(q, zsig, uflip);
    if (!fr) {
      extract(q);
      zsig = 0;
      uflip = false;
    }
    const size_t at = e.begin(OP_ROT);
    e.u((uint32_t)S.of_bit[q]);
    e.u((uint32_t)d.mode);
    e.f((float)(uflip ? -d.t : d.t));
    e.u64(zsig);
    e.end(at);
    if (d.mode & 1) {
      // cot forms: Z K Z = -K(-t)
      if (uflip) P.c0 = wrap(P.c0 + cd(0, kPi));
      for (uint64_t v = zsig; v; v &= v - 1) P.add_lin(ctz(v), cd(0, kPi));
    }
    P.add_lin(q, cd(0, std::arg(d.eL)));
    P.c0 = wrap(P.c0 + d.logk);
    rescale();
  }

  // a perm_like gate: its diagonal joins the pending factor, the factor moves
  // through the swap, and the swap is emitted
  void cx(const Gate& g) {
    const int t = g.t0;
    const int c = g.cmask ? ctz(g.cmask) : -1;
    const int cv = c >= 0 ? (int)((g.cval >> c) & 1u) : 1;
    // [[0, m1], [m2, 0]] = X diag(m2, m1)
    if (!(is1(g.m[1]) && is1(g.m[2]))) {
      cd d[4];
      if (c < 0) {
        d[0] = g.m[2];
        d[1] = d[2] = cd(1, 0);
        d[3] = g.m[1];
        P.add_diag2(t, t, d);
      } else {
        for (int zc = 0; zc < 2; ++zc)
          for (int zt = 0; zt < 2; ++zt)
            d[2 * zc + zt] = zc == cv ? (zt ? g.m[1] : g.m[2]) : cd(1, 0);
        P.add_diag2(c, t, d);
      }
    }
    if (c >= 0 && !P.cx_ok(t, c)) extract(t);
    if (cv == 0 || c < 0) P.flip(t);
    if (c >= 0) P.cx(t, c);
    const uint32_t creg =
        c < 0 ? CX_NONE : (S.of_bit[c] >= 0 ? (uint32_t)S.of_bit[c] : CX_THREAD);
    uint32_t smask = ~0u;
    if (creg < (uint32_t)RB) {
      smask = 0;
      for (int s = 0; s < NSLOT; ++s)
        if ((int)((s >> creg) & 1) == cv) smask |= 1u << s;
    }
    const size_t at = e.begin(OP_CXP);
    e.u((uint32_t)S.of_bit[t]);
    e.u(creg);
    e.u((uint32_t)cv);
    e.u64(g.cmask);
    e.u64(g.cval);
    e.u(smask);
    e.end(at);
  }

  // apply everything that is pending
  void flush() {
    regreg();
    for (int j = 0; j < RB; ++j)
      if (S.bit_of[j] >= 0) extract(S.bit_of[j]);
    std::vector<std::array<int, 2>> ab;
    std::vector<cd> ev;
    for (uint64_t v = P.lmask; v; v &= v - 1) {
      const int p = ctz(v);
      ab.push_back({p, p});
      ev.push_back(cexp_snap(P.lin[p]));
      P.clear_lin(p);
    }
    for (int p = 0; p < 64; ++p)
      for (uint64_t u = P.qm[p] & ~lowmask(p + 1); u; u &= u - 1) {
        const int q = ctz(u);
        ab.push_back({p, q});
        ev.push_back(cexp_snap(P.qref(p, q)));
        P.clear_quad(p, q);
      }
    const cd F0 = std::exp(P.c0);
    P.c0 = cd(0, 0);
    P.nrot = 0;
    if (ab.empty() && std::abs(F0 - cd(1, 0)) < 1e-15) return;
    const size_t at = e.begin(OP_THR);
    e.u((uint32_t)ab.size());
    e.c(F0);
    for (size_t i = 0; i < ab.size(); ++i) {
      e.u((uint32_t)ab[i][0]);
      e.u((uint32_t)ab[i][1]);
      e.c(ev[i]);
    }
    e.end(at);
  }
};

// emit one phase's steps
void emit_phase(std::vector<Gate>& gates, const std::vector<int>& sel,
                const std::array<uint32_t, RB>& Rp, const TileGeo& geo, Emit& e, uint32_t& nred,
                uint32_t red_base, Pend& P, bool flush_at_end) {
  const Slots S(Rp, geo);
  PhaseEmitter pe{P, S, e};
  auto slotmask = [&](uint64_t bits) -> uint32_t {
    uint32_t m = 0;
    for (uint64_t v = bits; v; v &= v - 1) {
      const int s = S.of_bit[ctz(v)];
      if (s >= 0) m |= 1u << s;
    }
    return m;
  };
  // smask: bit s set iff parity(s & zs) == 1
  auto smask_of = [&](uint32_t zs) -> uint32_t {
    uint32_t m = 0;
    for (int s = 0; s < NSLOT; ++s)
      if (popc((uint64_t)(s & zs)) & 1) m |= 1u << s;
    return m;
  };
  auto extract_bits = [&](uint64_t bits) {
    for (uint64_t v = bits; v; v &= v - 1) pe.extract(ctz(v));
  };

  for (int gi : sel) {
    Gate& g = gates[gi];
    switch (g.kind) {
      case GK_MAT1: {
        if (perm_like(g)) {
          pe.cx(g);
          break;
        }
        RotDec d;
        if (g.cmask == 0 && decompose_u1(g.m, d, g.rot_form)) {
          g.rot_form = d.mode & 1;
          pe.rot(g, d);
          break;
        }
        extract_bits(bitm(g.t0));
        const size_t at = e.begin(OP_U1);
        e.u((uint32_t)S.of_bit[g.t0]);
        e.u(mat1_type(g));
        e.u(g.cmask ? 1u : 0u);
        e.u64(g.cmask);
        e.u64(g.cval);
        for (int j = 0; j < 4; ++j) e.c(g.m[j]);
        e.end(at);
        break;
      }
      case GK_MAT2: {
        extract_bits(bitm(g.t0) | bitm(g.t1));
        int k0 = S.of_bit[g.t0], k1 = S.of_bit[g.t1];
        cd mm[16];
        if (k0 > k1) {
          for (int j = 0; j < 16; ++j) mm[j] = g.m[j];
        } else {
          auto sw = [](int i) { return ((i & 1) << 1) | (i >> 1); };
          for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) mm[4 * r + c] = g.m[4 * sw(r) + sw(c)];
          std::swap(k0, k1);
        }
        const size_t at = e.begin(OP_U2);
        e.u((uint32_t)k0);
        e.u((uint32_t)k1);
        e.u(g.cmask ? 1u : 0u);
        e.u64(g.cmask);
        e.u64(g.cval);
        for (int j = 0; j < 16; ++j) e.c(mm[j]);
        e.end(at);
        break;
      }
      case GK_DIAG2: {
        if (P.add_diag2(g.t0, g.t1, g.m)) break;
        const size_t at = e.begin(OP_D2);
        e.u((uint32_t)g.t0);
        e.u((uint32_t)g.t1);
        e.u(0);
        for (int j = 0; j < 4; ++j) e.c(g.m[j]);
        e.end(at);
        break;
      }
      case GK_PROT: {
        // exp(-i th/2 P) = k (C - i S P) with one of C, S equal to 1 and the
        // other the tangent (or cotangent) in float; k = +-1/sqrt(1 + t_f^2)
        // joins the deferred scale, as for OP_ROT, so that each applied
        // rotation is exactly unitary: c = cos, s = sin in float would change
        // the norm by up to ~1e-7 per rotation, coherently over the state.
        extract_bits(g.x);
        const bool tan_form = std::abs(g.s) <= std::abs(g.c);
        const double tf = (double)(float)(tan_form ? g.s / g.c : g.c / g.s);
        const double lead = tan_form ? g.c : g.s;
        const size_t at = e.begin(OP_PROT);
        e.u(slotmask(g.x));
        e.u64(g.z);
        e.f(tan_form ? 1.0f : (float)tf);
        e.f(tan_form ? (float)tf : 1.0f);
        e.u((uint32_t)g.k);
        e.end(at);
        P.c0 = wrap(P.c0 + cd(-0.5 * std::log1p(tf * tf), lead < 0 ? kPi : 0.0));
        pe.rescale();
        break;
      }
      case GK_PGRP: {
        extract_bits(g.x);
        PgrpClass cls[kMaxGroup];
        const int nc = pgrp_classes(g, cls);
        uint32_t act = 0;
        for (int c = 0; c < nc; ++c)
          if (cls[c].active) act |= 1u << c;
        const size_t at = e.begin(OP_PGRP);
        e.u(slotmask(g.x));
        e.u64(g.z);
        e.u((uint32_t)(popc(g.x & g.z) & 1));
        e.u(act);
        e.u(smask_of(slotmask(g.z)));
        for (int c = 0; c < nc; ++c) {
          e.f((float)cls[c].cm1);
          e.c(cls[c].ui);
          e.c(cls[c].uj);
        }
        e.end(at);
        break;
      }
      case GK_RED: {
        // <lambda|P|psi> with both states carrying the same pending factor D:
        // the terms on the flip bits are applied first; the rest cancels
        // except |exp(c0)|^2, which the host applies to the result.
        extract_bits(g.x);
        if (g.nmem > 0) {
          // group generator G = sum_r mg[r] P_r: one slot, Im <lambda|G|psi>
          Gate gen = g;
          for (int r = 0; r < g.nmem; ++r) gen.ma[r] = g.mg[r];
          PgrpClass cls[kMaxGroup];
          const int nc = pgrp_classes(gen, cls, false);  // active flags only
          int xb[64];
          int m = 0;
          for (uint64_t v = g.x; v; v &= v - 1) xb[m++] = ctz(v);
          static const cd ipow[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
          auto Vg = [&](uint64_t pat) {
            cd acc(0, 0);
            for (int r = 0; r < g.nmem; ++r) {
              const uint64_t d = (g.mz[r] ^ g.z) & g.x;
              acc += g.mg[r] * ((popc(pat & d) & 1) ? -1.0 : 1.0) * ipow[popc(g.x & g.mz[r]) & 3];
            }
            return acc;
          };
          uint32_t act = 0;
          for (int c = 0; c < nc; ++c)
            if (cls[c].active) act |= 1u << c;
          const size_t at = e.begin(OP_RED);
          e.u(slotmask(g.x));
          e.u(RED_GROUP);
          e.u(1);
          e.u(nred);
          e.u64(g.z);
          e.u(smask_of(slotmask(g.z)));
          e.u(1);
          e.u((uint32_t)(popc(g.x & g.z) & 1));
          e.u(act);
          for (int c = 0; c < nc; ++c) {
            uint64_t pi = 0;
            for (int b = 1; b < m; ++b)
              if ((c >> (b - 1)) & 1) pi |= bitm(xb[b]);
            e.c(Vg(pi));
            e.c(Vg(pi ^ g.x));
          }
          e.end(at);
          g.red_slot = (int)(red_base + nred);
          g.red_part = 1;
          g.red_scale = std::exp(2.0 * P.c0.real());
          nred += 1;
          break;
        }
        // the gradient needs one part of <lambda|P|psi> = i^k <lambda|X^x Z^z|psi>:
        // the imaginary part for even k, the real part for odd k
        const uint32_t xr = slotmask(g.x);
        const uint32_t zs = slotmask(g.z);
        const uint32_t part = (g.k & 1) ? 0u : 1u;
        const size_t at = e.begin(OP_RED);
        e.u(xr);
        e.u(RED_CROSS);
        e.u(1);
        e.u(nred);
        e.u64(g.z);
        e.u(smask_of(zs));
        e.u(part);
        e.end(at);
        g.red_slot = (int)(red_base + nred);
        g.red_part = (int)part;
        g.red_scale = std::exp(2.0 * P.c0.real());
        nred += 1;
        break;
      }
    }
  }
  if (flush_at_end) pe.flush();
}

}  // namespace

Gate dagger(const Gate& g) {
  Gate h = g;
  switch (g.kind) {
    case GK_MAT1:
      h.m[0] = std::conj(g.m[0]);
      h.m[1] = std::conj(g.m[2]);
      h.m[2] = std::conj(g.m[1]);
      h.m[3] = std::conj(g.m[3]);
      break;
    case GK_MAT2:
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) h.m[4 * r + c] = std::conj(g.m[4 * c + r]);
      break;
    case GK_DIAG2:
      for (int j = 0; j < 4; ++j) h.m[j] = std::conj(g.m[j]);
      break;
    case GK_PROT: h.s = -g.s; break;
    case GK_PGRP:
      for (int r = 0; r < g.nmem; ++r) h.ma[r] = -g.ma[r];
      break;
    default: break;
  }
  h.cost = gate_cost(h);
  return h;
}

int gates_from_ops(int n, const tqsv_op* ops, int nops, std::vector<Gate>& out) {
  out.clear();
  out.reserve((size_t)nops);
  const uint64_t full = lowmask(n);
  for (int i = 0; i < nops; ++i) {
    const tqsv_op& o = ops[i];
    // the message prefix is built only for a failing op
    auto bad = [i](int code, const std::string& msg) {
      return fail(code, "op " + std::to_string(i) + ": " + msg);
    };
    Gate& g = out.emplace_back();
    g.src = i;
    g.param = o.param;
    g.gx = o.gx;
    g.gz = o.gz;
    g.gcoef = o.gcoef;
    if ((o.cmask & ~full) || (o.cval & ~o.cmask))
      return bad(TQSV_EINVAL, "control mask out of range");
    switch (o.kind) {
      case TQSV_OP_MAT1:
        if (o.t[0] < 0 || o.t[0] >= n) return bad(TQSV_EINVAL, "target bit out of range");
        if (o.cmask & bitm(o.t[0])) return bad(TQSV_EINVAL, "target is also a control");
        g.kind = GK_MAT1;
        g.t0 = o.t[0];
        g.cmask = o.cmask;
        g.cval = o.cval;
        for (int j = 0; j < 4; ++j) g.m[j] = cd(o.m[2 * j], o.m[2 * j + 1]);
        break;
      case TQSV_OP_MAT2:
        if (o.t[0] < 0 || o.t[0] >= n || o.t[1] < 0 || o.t[1] >= n)
          return bad(TQSV_EINVAL, "target bit out of range");
        if (o.t[0] == o.t[1]) return bad(TQSV_EINVAL, "targets must differ");
        if (o.cmask & (bitm(o.t[0]) | bitm(o.t[1])))
          return bad(TQSV_EINVAL, "target is also a control");
        g.kind = GK_MAT2;
        g.t0 = o.t[0];
        g.t1 = o.t[1];
        g.cmask = o.cmask;
        g.cval = o.cval;
        for (int j = 0; j < 16; ++j) g.m[j] = cd(o.m[2 * j], o.m[2 * j + 1]);
        break;
      case TQSV_OP_PROT:
        if ((o.xmask | o.zmask) & ~full) return bad(TQSV_EINVAL, "Pauli mask out of range");
        if (o.cmask) return bad(TQSV_EINVAL, "Pauli rotations take no controls");
        g.kind = GK_PROT;
        g.x = o.xmask;
        g.z = o.zmask;
        g.k = popc(g.x & g.z) & 3;
        g.c = std::cos(0.5 * o.theta);
        g.s = std::sin(0.5 * o.theta);
        if (g.param >= 0) {
          g.gx = g.x;
          g.gz = g.z;
        }
        break;
      default: return bad(TQSV_EINVAL, "unknown op kind " + std::to_string(o.kind));
    }
    if (g.param >= 0) {
      if (g.cmask) return bad(TQSV_EUNSUPPORTED, "controlled ops cannot carry a parameter");
      if ((g.gx | g.gz) & ~full) return bad(TQSV_EINVAL, "generator mask out of range");
      if ((g.gx | g.gz) == 0) return bad(TQSV_EINVAL, "parametrised op without generator");
    }
  }
  return 0;
}

int pgrp_classes(const Gate& g, PgrpClass* out, bool constants) {
  int xb[64];
  int m = 0;
  for (uint64_t v = g.x; v; v &= v - 1) xb[m++] = ctz(v);
  const int nc = m > 0 ? 1 << (m - 1) : 0;
  static const cd ipow[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
  double amax = 0.0, gmax = 0.0;
  for (int r = 0; r < g.nmem; ++r) {
    amax += std::abs(g.ma[r]);
    gmax += std::abs(g.mg[r]);
  }
  // V(p) = sum_r w_r i^{k_r} (-1)^{|p & d_r|} for a pattern p of the x bits,
  // d_r = (z_r ^ z) & x.  Per member once: d_r, the phase-weighted angle and
  // generator coefficient, and |d_r| mod 2, which relates the two members of
  // a pair: V(p ^ x) = sum_r w_r i^{k_r} (-1)^{|p & d_r| + |d_r|}.
  uint64_t dr[kMaxGroup];
  cd wa[kMaxGroup], wg[kMaxGroup];
  bool odd[kMaxGroup];
  for (int r = 0; r < g.nmem; ++r) {
    dr[r] = (g.mz[r] ^ g.z) & g.x;
    const cd ph = ipow[popc(g.x & g.mz[r]) & 3];
    wa[r] = g.ma[r] * ph;
    wg[r] = g.mg[r] * ph;
    odd[r] = (popc(dr[r]) & 1) != 0;
  }
  // one shared parameter: the classes that stay the identity for every
  // value of it are known from the coefficients, so the set of active
  // classes (part of the generated code) does not change with the angle
  bool shared = g.nmem > 0 && g.mp[0] >= 0;
  for (int r = 1; r < g.nmem; ++r) shared = shared && g.mp[r] == g.mp[0];
  for (int c = 0; c < nc; ++c) {
    uint64_t pi = 0;
    for (int b = 1; b < m; ++b)
      if ((c >> (b - 1)) & 1) pi |= bitm(xb[b]);
    cd vi(0, 0), vj(0, 0), vg(0, 0);
    for (int r = 0; r < g.nmem; ++r) {
      const bool ni = (popc(pi & dr[r]) & 1) != 0;
      vi += ni ? -wa[r] : wa[r];
      vj += (ni != odd[r]) ? -wa[r] : wa[r];
      vg += ni ? -wg[r] : wg[r];
    }
    const double w = 0.5 * (std::abs(vi) + std::abs(vj));
    PgrpClass& k = out[c];
    bool act = w > 1e-15 * std::max(amax, 1e-300);
    if (shared && gmax > 0) act = act || std::abs(vg) > 1e-12 * gmax;
    k.active = act;
    if (!constants) continue;
    const double sq = std::sin(0.25 * w);
    k.cm1 = -2.0 * sq * sq;
    const double sow = w > 0 ? std::sin(0.5 * w) / w : 0.5;
    k.ui = cd(0, -sow) * vi;
    k.uj = cd(0, -sow) * vj;
  }
  return nc;
}

// Consecutive Pauli rotations with one flip mask and z masks that differ
// only inside the flip mask and commute (an even number of differing
// positions) become one GK_PGRP gate: on every pair (i, i ^ x)
// their product is one two-level rotation.  The rotations of one UCC
// excitation (the Jordan-Wigner strings of T - T^dagger) are such a run: the
// group touches only the pairs whose occupation pattern the excitation
// changes (1 in 8 for doubles) with one rounding, instead of every amplitude
// once per string.
void group_rotations(std::vector<Gate>& gs) {
  // compacted in place: the write position w never passes the read position i
  size_t w = 0;
  auto put = [&](size_t i) {
    if (w != i) gs[w] = gs[i];
    ++w;
  };
  for (size_t i = 0; i < gs.size(); ++i) {
    const Gate& g = gs[i];
    if (g.kind != GK_PROT || g.x == 0 || popc(g.x) > RB) {
      put(i);
      continue;
    }
    size_t j = i + 1;
    while (j < gs.size() && (int)(j - i) < kMaxGroup) {
      const Gate& h = gs[j];
      const uint64_t d = h.z ^ g.z;
      if (h.kind != GK_PROT || h.x != g.x || (d & ~g.x) || (popc(d) & 1)) break;
      ++j;
    }
    if (j - i < 2) {
      put(i);
      continue;
    }
    Gate grp = g;
    grp.kind = GK_PGRP;
    grp.k = popc(g.x & g.z) & 3;
    grp.nmem = (int)(j - i);
    for (size_t r = i; r < j; ++r) {
      grp.mz[r - i] = gs[r].z;
      grp.ma[r - i] = 2.0 * std::atan2(gs[r].s, gs[r].c);
      grp.mg[r - i] = gs[r].gcoef;
      grp.mp[r - i] = gs[r].param;
      grp.param = std::max(grp.param, gs[r].param);
    }
    classify(grp);
    gs[w++] = grp;
    i = j - 1;
  }
  gs.resize(w);
}

void merge_and_classify(std::vector<Gate>& gs) {
  std::vector<int> open(64, -1), last(64, -1);
  std::vector<char> dead(gs.size(), 0);
  for (size_t i = 0; i < gs.size(); ++i) {
    Gate& g = gs[i];
    if (g.kind == GK_MAT1 && g.cmask == 0 && g.param < 0) {
      const int b = g.t0;
      const int j = open[b];
      if (j >= 0 && last[b] == j) {
        Gate& h = gs[j];
        cd r[4];
        r[0] = g.m[0] * h.m[0] + g.m[1] * h.m[2];
        r[1] = g.m[0] * h.m[1] + g.m[1] * h.m[3];
        r[2] = g.m[2] * h.m[0] + g.m[3] * h.m[2];
        r[3] = g.m[2] * h.m[1] + g.m[3] * h.m[3];
        for (int q = 0; q < 4; ++q) h.m[q] = r[q];
        dead[i] = 1;
        continue;
      }
      open[b] = (int)i;
    }
    for (uint64_t v = raw_used(g); v; v &= v - 1) last[ctz(v)] = (int)i;
  }
  size_t w = 0;  // compacted in place
  for (size_t i = 0; i < gs.size(); ++i)
    if (!dead[i]) {
      classify(gs[i]);
      if (w != i) gs[w] = gs[i];
      ++w;
    }
  gs.resize(w);
  group_rotations(gs);
}

void classify_gate(Gate& g) { classify(g); }

int effective_tile_bits(int n, int requested, bool dual) {
  int T = requested > 0 ? requested : 13;
  if (T < 11) T = 11;
  if (T > 14) T = 14;
  if (dual && T > 13) T = 13;
  if (n > 32) T = 13;  // 64-bit index kernels exist for 2^13-amplitude tiles only
  if (n < T) T = n;
  return T;
}

int select_passes(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan>& plan) {
  const int n = so.n;
  const int T = effective_tile_bits(n, so.tile_bits, so.dual);
  const bool tiles_ok = so.fuse && T >= 11 && !(so.dual && n > 32);
  const int TBv = T - RB;
  const int mode = so.dual ? MODE_DUAL : MODE_GATE;
  int Lmin = so.min_low > 0 ? so.min_low : 5;
  if (Lmin > T - RB) Lmin = T - RB;
  if (Lmin < T - MAXH) Lmin = T - MAXH;  // at most MAXH high bits fit the pass header
  if (Lmin < 1) Lmin = 1;
  const uint64_t full = lowmask(n);
  const int budget0 = so.max_cost > 0 ? so.max_cost : default_max_cost(n);
  // reduction slots per pass: the LDS beyond the tile(s) holds one double per
  // slot and wave
  const size_t lds_tiles = pass_lds_bytes(TBv, mode, 0);
  const size_t nwave = ((size_t)1 << TBv) >= 64 ? ((size_t)1 << TBv) / 64 : 1;
  const uint32_t red_cap =
      lds_tiles < 160 * 1024 ? (uint32_t)((160 * 1024 - lds_tiles) / (nwave * sizeof(double))) : 0;

  std::vector<int> rem(gates.size());
  for (size_t i = 0; i < gates.size(); ++i) rem[i] = (int)i;
  std::vector<char> done(gates.size(), 0);
  // Circuits made mostly of Pauli-string rotations (chemistry ansaetze, their
  // reverse sweeps) reach further per pass when the tile bits are searched
  // by admitted work: the 20-spin-orbital UCCSD forward plan has 79 passes
  // instead of 128.  Layered gate circuits keep the gate-order choice (the
  // 30-qubit brickwork: 21 passes instead of 23).
  size_t nrot = 0;
  for (const Gate& g : gates)
    nrot += (g.kind == GK_PROT || g.kind == GK_PGRP || g.kind == GK_RED) ? 1 : 0;
  const bool search = so.tile_search || 2 * nrot >= gates.size();

  // ---- pass and phase selection
  while (!rem.empty()) {
    PassPlan pp;
    pp.dual = so.dual;
    pp.mode = mode;
    if (!tiles_ok || !fusable(gates[rem[0]])) {
      pp.tile = false;
      pp.single = rem[0];
      plan.push_back(pp);
      rem.erase(rem.begin());
      continue;
    }
    // 1) choose the tile's bit set.  Search: grow the set from the low bits by
    // the candidate (one bit, or the non-tile bits of an upcoming gate) that
    // admits the most work per added bit under the rule of step 2, ties broken
    // by the work admissible without the budget (light-cone reach); then fill
    // the remaining bits in gate order.
    uint64_t S = lowmask(Lmin);
    if (search) {
      auto work = [&](uint64_t bits, bool use_budget) -> long {
        uint64_t bN = 0, bD = 0;
        int bud = budget0;
        long w = 0;
        bool first = true;
        for (int gi : rem) {
          const Gate& g = gates[gi];
          if (bN == full) break;
          if (conflicts(g, bN, bD) || !fusable(g) || (g.nb & ~bits) ||
              (use_budget && !first && g.cost > bud)) {
            block(g, bN, bD);
            continue;
          }
          w += g.cost;
          bud -= g.cost;
          first = false;
        }
        return w;
      };
      long wcur = work(S, true);
      for (;;) {
        const int room = T - popc(S);
        if (room <= 0) break;
        std::vector<uint64_t> cands;
        for (int b = 0; b < n; ++b)
          if (!((S >> b) & 1u)) cands.push_back(bitm(b));
        uint64_t bN = 0, bD = 0;
        size_t seen = 0;
        for (int gi : rem) {
          const Gate& g = gates[gi];
          if (bN == full || seen >= 64) break;
          const uint64_t add = g.nb & ~S;
          if (!conflicts(g, bN, bD) && fusable(g) && popc(add) >= 2 && popc(add) <= room &&
              std::find(cands.begin(), cands.end(), add) == cands.end()) {
            cands.push_back(add);
            ++seen;
          }
          block(g, bN, bD);
        }
        uint64_t best = 0;
        double bscore = -1.0;
        long bpot = -1, bw = 0;
        for (uint64_t c : cands) {
          const long w = work(S | c, true);
          const long pot = work(S | c, false);
          const double score = (double)(w - wcur) / popc(c);
          if (score > bscore || (score == bscore && pot > bpot)) {
            best = c;
            bscore = score;
            bpot = pot;
            bw = w;
          }
        }
        if (!best || (bscore <= 0.0 && bpot <= wcur)) break;
        S |= best;
        wcur = bw;
      }
    }
    uint64_t blkN = 0, blkD = 0;
    int budget = budget0;
    bool any = false;  // the first admissible gate is taken whatever its cost
    for (int gi : rem) {
      const Gate& g = gates[gi];
      if (blkN == full) break;
      if (conflicts(g, blkN, blkD) || !fusable(g)) {
        block(g, blkN, blkD);
        continue;
      }
      const uint64_t need = g.nb & ~S;
      if (popc(S | need) > T || (any && g.cost > budget)) {
        block(g, blkN, blkD);
        continue;
      }
      S |= need;
      budget -= g.cost;
      any = true;
    }
    TileGeo geo;
    geo.T = T;
    int L = Lmin;
    for (;;) {
      const int nL = T - popc(S & ~lowmask(L));
      if (nL <= L) break;
      L = nL;
    }
    // L is the largest low count such that low(L) U (S above L) has <= T bits; pad H if short
    geo.L = L;
    for (uint64_t v = S & ~lowmask(L); v; v &= v - 1) geo.H.push_back(ctz(v));
    for (int b = L; (int)geo.H.size() < T - L && b < n; ++b)
      if (!((S >> b) & 1u)) {
        geo.H.push_back(b);
        S |= bitm(b);
      }
    std::sort(geo.H.begin(), geo.H.end());
    if ((int)geo.H.size() > MAXH || L < 1 || L + (int)geo.H.size() != T)
      return fail(TQSV_EINVAL, "planner built an invalid tile (internal error)");
    const uint64_t tilebits = lowmask(L) | (S & ~lowmask(L));
    // 2) take every admissible gate with the tile fixed; reductions also
    // count against the LDS left for their per-wave accumulators
    std::vector<int> chosen;
    blkN = blkD = 0;
    budget = budget0;
    uint32_t nred_pass = 0;
    for (int gi : rem) {
      const Gate& g = gates[gi];
      if (conflicts(g, blkN, blkD) || !fusable(g) || (g.nb & ~tilebits) ||
          (!chosen.empty() && g.cost > budget) ||
          (g.kind == GK_RED && nred_pass + 1 > red_cap)) {
        block(g, blkN, blkD);
        continue;
      }
      chosen.push_back(gi);
      budget -= g.cost;
      if (g.kind == GK_RED) nred_pass += 1;
    }
    if (chosen.empty()) return fail(TQSV_EINVAL, "planner made no progress (internal error)");
    // 3) phases
    std::vector<int> prem = chosen;
    while (!prem.empty()) {
      uint32_t R = 0;
      blkN = blkD = 0;
      for (int gi : prem) {
        const Gate& g = gates[gi];
        if (conflicts(g, blkN, blkD)) {
          block(g, blkN, blkD);
          continue;
        }
        const uint32_t need = geo.posmask(g.nb);
        if (popc((uint64_t)(R | need)) > RB) {
          block(g, blkN, blkD);
          continue;
        }
        R |= need;
      }
      for (int p = T - 1; popc((uint64_t)R) < RB && p >= 0; --p) R |= 1u << p;
      std::vector<int> sel, keep;
      blkN = blkD = 0;
      for (int gi : prem) {
        const Gate& g = gates[gi];
        if (conflicts(g, blkN, blkD) || (geo.posmask(g.nb) & ~R)) {
          block(g, blkN, blkD);
          keep.push_back(gi);
          continue;
        }
        sel.push_back(gi);
      }
      std::array<uint32_t, RB> Rp{};
      int j = 0;
      for (int p = 0; p < T; ++p)
        if ((R >> p) & 1u) Rp[j++] = (uint32_t)p;
      pp.phases.push_back(sel);
      pp.Rpos.push_back(Rp);
      prem.swap(keep);
    }
    for (int gi : chosen) done[gi] = 1;
    std::vector<int> nrem;
    nrem.reserve(rem.size());
    for (int gi : rem)
      if (!done[gi]) nrem.push_back(gi);
    rem.swap(nrem);

    pp.tile = true;
    pp.tb = TBv;
    pp.L = (uint32_t)geo.L;
    pp.m = (uint32_t)geo.H.size();
    for (size_t k = 0; k < geo.H.size(); ++k) pp.H[k] = (uint32_t)geo.H[k];
    plan.push_back(pp);
  }
  return 0;
}


int emit_passes(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan> plan,
                std::vector<PassPlan>& passes, std::vector<uint32_t>& prog, uint32_t& red_total,
                std::vector<uint32_t>* fpos) {
  const int n = so.n;
  const int T = effective_tile_bits(n, so.tile_bits, so.dual);
  const int TBv = T - RB;
  const int mode = so.dual ? MODE_DUAL : MODE_GATE;
  // ---- emission.  The pending diagonal travels from pass to pass; it is
  // flushed in the last phase of the last tile pass before a single-op pass
  // (which does not know about it) and at the end of the plan.
  Pend P;
  for (size_t i = 0; i < plan.size(); ++i) {
    PassPlan& pp = plan[i];
    if (!pp.tile) {
      Gate& g = gates[pp.single];
      if (g.kind == GK_RED) {
        pp.nred = 2;
        pp.red_base = red_total;
        g.red_slot = (int)red_total;
        g.red_part = -1;
        g.red_scale = 1.0;
        red_total += 2;
      }
      passes.push_back(pp);
      continue;
    }
    const bool flush = i + 1 == plan.size() || !plan[i + 1].tile;
    TileGeo geo;
    geo.T = T;
    geo.L = (int)pp.L;
    for (uint32_t k = 0; k < pp.m; ++k) geo.H.push_back((int)pp.H[k]);
    pp.prog_off = prog.size();
    pp.red_base = red_total;
    Emit e(prog, fpos);
    for (int k = 0; k < PH_SIZE; ++k) e.u(0);
    uint32_t nred = 0;
    for (size_t ip = 0; ip < pp.phases.size(); ++ip) {
      const size_t hdr = prog.size();
      for (int k = 0; k < PS_SIZE; ++k) e.u(0);
      const size_t steps0 = prog.size();
      emit_phase(gates, pp.phases[ip], pp.Rpos[ip], geo, e, nred, red_total, P,
                 flush && ip + 1 == pp.phases.size());
      uint32_t nsteps = 0;
      for (size_t at = steps0; at < prog.size(); at += prog[at] >> 16) ++nsteps;
      prog[hdr + PS_NSTEPS] = nsteps;
      for (int r = 0; r < RB; ++r) prog[hdr + PS_R + r] = pp.Rpos[ip][r];
    }
    uint32_t* h = prog.data() + pp.prog_off;
    h[PH_FLAGS] = PF_WRITE | (so.dual ? PF_DUAL : 0u);
    h[PH_L] = pp.L;
    h[PH_M] = pp.m;
    h[PH_NPHASE] = (uint32_t)pp.phases.size();
    h[PH_NRED] = nred;
    h[PH_WORDS] = (uint32_t)(prog.size() - pp.prog_off);
    for (uint32_t k = 0; k < pp.m; ++k) h[PH_H + k] = pp.H[k];
    pp.prog_words = prog.size() - pp.prog_off;
    pp.lean = true;
    for (size_t at = pp.prog_off + PH_SIZE, ip = 0; ip < pp.phases.size(); ++ip) {
      const uint32_t ns = prog[at + PS_NSTEPS];
      at += PS_SIZE;
      for (uint32_t k = 0; k < ns; ++k) {
        const uint32_t op = prog[at] & 0xffffu;
        if (op == OP_U2 || op == OP_PROT || op == OP_PGRP) pp.lean = false;
        at += prog[at] >> 16;
      }
    }
    pp.nred = nred;
    red_total += nred;
    pp.ntiles = (uint32_t)(1ull << (n - T));
    const uint32_t cap = nred ? 2048u : 65536u;
    pp.grid = pp.ntiles < cap ? pp.ntiles : cap;
    pp.lds = pass_lds_bytes(TBv, mode, nred);
    if (pp.lds > 160 * 1024) return fail(TQSV_EUNSUPPORTED, "pass needs more LDS than available");
    if (!pass_variant_exists(TBv, mode, n > 32))
      return fail(TQSV_EUNSUPPORTED, "no tile-pass kernel variant for this tile size");
    passes.push_back(pp);
  }
  if (!P.empty()) return fail(TQSV_EINVAL, "planner left diagonal factors unapplied (internal error)");
  return 0;
}

int schedule(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan>& passes,
             std::vector<uint32_t>& prog, uint32_t& red_total) {
  std::vector<PassPlan> plan;
  const int rc = select_passes(gates, so, plan);
  if (rc) return rc;
  return emit_passes(gates, so, std::move(plan), passes, prog, red_total);
}

bool same_structure(const std::vector<Gate>& a, const std::vector<Gate>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i) {
    const Gate& x = a[i];
    const Gate& y = b[i];
    if (x.kind != y.kind || x.t0 != y.t0 || x.t1 != y.t1 || x.cmask != y.cmask ||
        x.cval != y.cval || x.x != y.x || x.z != y.z || x.nb != y.nb || x.db != y.db ||
        x.cost != y.cost || (x.param >= 0) != (y.param >= 0) || x.nmem != y.nmem)
      return false;
    for (int r = 0; r < x.nmem; ++r)
      if (x.mz[r] != y.mz[r] || (x.mp[r] >= 0) != (y.mp[r] >= 0)) return false;
    if (x.kind == GK_PGRP) {
      // only the active flags are compared: no class constants
      PgrpClass cx[kMaxGroup], cy[kMaxGroup];
      const int nc = pgrp_classes(x, cx, false);
      pgrp_classes(y, cy, false);
      for (int c = 0; c < nc; ++c)
        if (cx[c].active != cy[c].active) return false;
    }
  }
  return true;
}

static hipError_t apply_single(const Gate& g, int n, float2* psi, hipStream_t st) {
  switch (g.kind) {
    case GK_MAT1: {
      Mat2f u;
      for (int j = 0; j < 4; ++j) {
        u.v[2 * j] = (float)g.m[j].real();
        u.v[2 * j + 1] = (float)g.m[j].imag();
      }
      return k_apply_1q(psi, n, g.t0, g.cmask, g.cval, u, st);
    }
    case GK_MAT2: {
      Mat4f u;
      for (int j = 0; j < 16; ++j) {
        u.v[2 * j] = (float)g.m[j].real();
        u.v[2 * j + 1] = (float)g.m[j].imag();
      }
      return k_apply_2q(psi, n, g.t0, g.t1, g.cmask, g.cval, u, st);
    }
    case GK_DIAG2: {
      float ph[8];
      if (g.t0 == g.t1) {
        const int bits[1] = {g.t0};
        ph[0] = (float)g.m[0].real(); ph[1] = (float)g.m[0].imag();
        ph[2] = (float)g.m[3].real(); ph[3] = (float)g.m[3].imag();
        return k_apply_diag(psi, n, 1, bits, ph, st);
      }
      const int bits[2] = {g.t0, g.t1};
      for (int j = 0; j < 4; ++j) {
        ph[2 * j] = (float)g.m[j].real();
        ph[2 * j + 1] = (float)g.m[j].imag();
      }
      return k_apply_diag(psi, n, 2, bits, ph, st);
    }
    case GK_PROT: return k_pauli_rot(psi, n, g.x, g.z, (float)g.c, (float)g.s, g.k, st);
    case GK_PGRP:
      for (int r = 0; r < g.nmem; ++r) {
        const hipError_t e = k_pauli_rot(psi, n, g.x, g.mz[r], (float)std::cos(0.5 * g.ma[r]),
                                         (float)std::sin(0.5 * g.ma[r]),
                                         popc(g.x & g.mz[r]) & 3, st);
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    default: return hipErrorInvalidValue;
  }
}

int run_passes(const std::vector<PassPlan>& passes, const std::vector<Gate>& gates, int n,
               const uint32_t* dprog, float2* psi, float2* lam, double* dpartials,
               double* dred_out, hipStream_t st, const uint32_t* dslots, uint32_t nslots) {
  const bool one_sum = dslots != nullptr;
  for (const PassPlan& pp : passes) {
    double* part = one_sum ? dpartials + pp.part_off : dpartials;
    if (pp.tile) {
      PassLaunch pl;
      pl.psi = psi;
      pl.lam = lam;
      pl.prog = dprog + pp.prog_off;
      pl.partials = part;
      pl.ntiles = pp.ntiles;
      pl.grid = pp.grid;
      pl.lds_bytes = pp.lds;
      if (pp.jit_fn)
        TQSV_HIP(jit_launch(pp, pl, st));
      else
        TQSV_HIP(launch_pass(pp.tb, pp.mode, n > 32, pp.lean, pl, st));
      if (pp.nred && !one_sum)
        TQSV_HIP(k_final_sum(dpartials, pp.grid, pp.nred, dred_out + pp.red_base, st));
    } else {
      const Gate& g = gates[pp.single];
      if (g.kind == GK_RED) {
        const uint32_t grid = reduce_grid(n);
        TQSV_HIP(k_pauli_dot(lam, psi, n, g.x, g.z, part, st));
        if (!one_sum) TQSV_HIP(k_final_sum(dpartials, grid, 2, dred_out + pp.red_base, st));
      } else {
        TQSV_HIP(apply_single(g, n, psi, st));
        if (pp.dual) TQSV_HIP(apply_single(g, n, lam, st));
      }
    }
  }
  if (one_sum) TQSV_HIP(k_final_sum_slots(dpartials, dslots, nslots, dred_out, st));
  return 0;
}

}  // namespace tqsv
