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

int gate_cost(const Gate& g) {
  switch (g.kind) {
    case GK_MAT1: return (mat1_type(g) == U1_GENERAL ? 8 : 4) + (g.cmask ? 4 : 0);
    case GK_MAT2: return g.cmask ? 20 : 16;
    case GK_DIAG2: return 2;
    case GK_PROT: return 8;
    case GK_RED: return 10;
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
      g.nb = g.x;
      g.db = g.z & ~g.x;
      break;
  }
  g.cost = gate_cost(g);
}

bool fusable(const Gate& g) {
  if (g.kind == GK_PROT || g.kind == GK_RED) return popc(g.x) <= RB;
  return true;
}

struct Emit {
  std::vector<uint32_t>& w;
  explicit Emit(std::vector<uint32_t>& v) : w(v) {}
  void u(uint32_t v) { w.push_back(v); }
  void f(float x) {
    uint32_t v;
    std::memcpy(&v, &x, 4);
    w.push_back(v);
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

// emit one phase's steps
void emit_phase(std::vector<Gate>& gates, const std::vector<int>& sel,
                const std::array<uint32_t, RB>& Rp, const TileGeo& geo, Emit& e, uint32_t& nred,
                uint32_t red_base) {
  auto slot_of_pos = [&](int p) -> int {
    for (int j = 0; j < RB; ++j)
      if ((int)Rp[j] == p) return j;
    return -1;
  };
  auto slot_of_bit = [&](int b) -> int {
    const int p = geo.pos(b);
    return p < 0 ? -1 : slot_of_pos(p);
  };
  auto slotmask = [&](uint64_t bits) -> uint32_t {
    uint32_t m = 0;
    for (uint64_t v = bits; v; v &= v - 1) {
      const int s = slot_of_bit(ctz(v));
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

  cd tab[NSLOT];
  bool tab_used = false;
  uint32_t tab_deps = 0;
  struct Ent {
    int a, b;
    cd d[4];
  };
  std::vector<Ent> ents;
  size_t cz_at = (size_t)-1;  // open OP_CZ step (entries may be appended)
  auto reset_tab = [&]() {
    for (int s = 0; s < NSLOT; ++s) tab[s] = cd(1, 0);
    tab_used = false;
    tab_deps = 0;
    ents.clear();
  };
  reset_tab();
  auto flush = [&]() {
    if (!tab_used && ents.empty()) return;
    cz_at = (size_t)-1;
    const size_t at = e.begin(OP_DIAG);
    e.u(tab_used ? 1u : 0u);
    e.u((uint32_t)ents.size());
    e.u(0);
    if (tab_used)
      for (int s = 0; s < NSLOT; ++s) {
        e.f((float)tab[s].real());
        e.f((float)tab[s].imag());
      }
    for (const Ent& x : ents) {
      e.u((uint32_t)x.a);
      e.u((uint32_t)x.b);
      for (int j = 0; j < 4; ++j) {
        e.f((float)x.d[j].real());
        e.f((float)x.d[j].imag());
      }
    }
    e.end(at);
    reset_tab();
  };

  for (int gi : sel) {
    Gate& g = gates[gi];
    switch (g.kind) {
      case GK_MAT1: {
        const int k = slot_of_bit(g.t0);
        if (tab_deps & (1u << k)) flush();
        cz_at = (size_t)-1;
        const size_t at = e.begin(OP_U1);
        e.u((uint32_t)k);
        e.u(mat1_type(g));
        e.u(g.cmask ? 1u : 0u);
        e.u64(g.cmask);
        e.u64(g.cval);
        for (int j = 0; j < 4; ++j) {
          e.f((float)g.m[j].real());
          e.f((float)g.m[j].imag());
        }
        e.end(at);
        break;
      }
      case GK_MAT2: {
        int k0 = slot_of_bit(g.t0), k1 = slot_of_bit(g.t1);
        if (tab_deps & ((1u << k0) | (1u << k1))) flush();
        cz_at = (size_t)-1;
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
        for (int j = 0; j < 16; ++j) {
          e.f((float)mm[j].real());
          e.f((float)mm[j].imag());
        }
        e.end(at);
        break;
      }
      case GK_DIAG2: {
        const int ka = slot_of_bit(g.t0), kb = slot_of_bit(g.t1);
        if (ka >= 0 && kb >= 0) {
          for (int s = 0; s < NSLOT; ++s) {
            const int ba = (s >> ka) & 1, bb = (s >> kb) & 1;
            tab[s] *= g.m[2 * ba + bb];
          }
          tab_used = true;
          tab_deps |= (1u << ka) | (1u << kb);
        } else if (ka < 0 && kb < 0) {
          Ent x;
          x.a = g.t0;
          x.b = g.t1;
          for (int j = 0; j < 4; ++j) x.d[j] = g.m[j];
          ents.push_back(x);
        } else if (is1(g.m[0]) && is1(g.m[1]) && is1(g.m[2]) && g.m[3] == cd(-1, 0)) {
          const int lo = std::min(g.t0, g.t1), hi = std::max(g.t0, g.t1);
          if (cz_at == (size_t)-1) {
            cz_at = e.begin(OP_CZ);
            e.u(0);
            e.end(cz_at);
          }
          // append an entry to the open CZ step (it is the last step emitted)
          e.w[cz_at + 1] += 1;
          e.u64(bitm(lo));
          e.u((uint32_t)(hi - lo));
          e.w[cz_at] = (e.w[cz_at] & 0xffffu) | (uint32_t)((e.w.size() - cz_at) << 16);
        } else {
          cz_at = (size_t)-1;
          const size_t at = e.begin(OP_D2);
          e.u((uint32_t)g.t0);
          e.u((uint32_t)g.t1);
          e.u(0);
          for (int j = 0; j < 4; ++j) {
            e.f((float)g.m[j].real());
            e.f((float)g.m[j].imag());
          }
          e.end(at);
        }
        break;
      }
      case GK_PROT: {
        const uint32_t xr = slotmask(g.x);
        if (tab_deps & xr) flush();
        cz_at = (size_t)-1;
        const size_t at = e.begin(OP_PROT);
        e.u(xr);
        e.u64(g.z);
        e.f((float)g.c);
        e.f((float)g.s);
        e.u((uint32_t)g.k);
        e.end(at);
        break;
      }
      case GK_RED: {
        flush();
        cz_at = (size_t)-1;
        const uint32_t xr = slotmask(g.x);
        const uint32_t zs = slotmask(g.z);
        const uint32_t sm = smask_of(zs);
        const size_t at = e.begin(OP_RED);
        e.u(xr);
        e.u(RED_CROSS);
        e.u(2);
        for (uint32_t part = 0; part < 2; ++part) {
          e.u(nred + part);
          e.u64(g.z);
          e.u(sm);
          e.u(part);
        }
        e.end(at);
        g.red_slot = (int)(red_base + nred);
        nred += 2;
        break;
      }
    }
  }
  flush();
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
    const std::string where = "op " + std::to_string(i) + ": ";
    Gate g;
    g.src = i;
    g.param = o.param;
    g.gx = o.gx;
    g.gz = o.gz;
    g.gcoef = o.gcoef;
    if ((o.cmask & ~full) || (o.cval & ~o.cmask))
      return fail(TQSV_EINVAL, where + "control mask out of range");
    switch (o.kind) {
      case TQSV_OP_MAT1:
        if (o.t[0] < 0 || o.t[0] >= n) return fail(TQSV_EINVAL, where + "target bit out of range");
        if (o.cmask & bitm(o.t[0])) return fail(TQSV_EINVAL, where + "target is also a control");
        g.kind = GK_MAT1;
        g.t0 = o.t[0];
        g.cmask = o.cmask;
        g.cval = o.cval;
        for (int j = 0; j < 4; ++j) g.m[j] = cd(o.m[2 * j], o.m[2 * j + 1]);
        break;
      case TQSV_OP_MAT2:
        if (o.t[0] < 0 || o.t[0] >= n || o.t[1] < 0 || o.t[1] >= n)
          return fail(TQSV_EINVAL, where + "target bit out of range");
        if (o.t[0] == o.t[1]) return fail(TQSV_EINVAL, where + "targets must differ");
        if (o.cmask & (bitm(o.t[0]) | bitm(o.t[1])))
          return fail(TQSV_EINVAL, where + "target is also a control");
        g.kind = GK_MAT2;
        g.t0 = o.t[0];
        g.t1 = o.t[1];
        g.cmask = o.cmask;
        g.cval = o.cval;
        for (int j = 0; j < 16; ++j) g.m[j] = cd(o.m[2 * j], o.m[2 * j + 1]);
        break;
      case TQSV_OP_PROT:
        if ((o.xmask | o.zmask) & ~full) return fail(TQSV_EINVAL, where + "Pauli mask out of range");
        if (o.cmask) return fail(TQSV_EINVAL, where + "Pauli rotations take no controls");
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
      default: return fail(TQSV_EINVAL, where + "unknown op kind " + std::to_string(o.kind));
    }
    if (g.param >= 0) {
      if (g.cmask) return fail(TQSV_EUNSUPPORTED, where + "controlled ops cannot carry a parameter");
      if ((g.gx | g.gz) & ~full) return fail(TQSV_EINVAL, where + "generator mask out of range");
      if ((g.gx | g.gz) == 0) return fail(TQSV_EINVAL, where + "parametrised op without generator");
    }
    out.push_back(g);
  }
  return 0;
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
  std::vector<Gate> out;
  out.reserve(gs.size());
  for (size_t i = 0; i < gs.size(); ++i)
    if (!dead[i]) {
      classify(gs[i]);
      out.push_back(gs[i]);
    }
  gs.swap(out);
}

void classify_gate(Gate& g) { classify(g); }

int effective_tile_bits(int n, int requested, bool dual) {
  int T = requested > 0 ? requested : 13;
  if (T < 11) T = 11;
  if (T > 14) T = 14;
  if (dual && T > 13) T = 13;
  if (n > 32 && T < 13) T = 13;
  if (n < T) T = n;
  return T;
}

int schedule(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan>& passes,
             std::vector<uint32_t>& prog, uint32_t& red_total) {
  const int n = so.n;
  const int T = effective_tile_bits(n, so.tile_bits, so.dual);
  const bool tiles_ok = so.fuse && T >= 11 && !(so.dual && n > 32);
  const int TBv = T - RB;
  int Lmin = so.min_low > 0 ? so.min_low : 5;
  if (Lmin > T - RB) Lmin = T - RB;
  if (Lmin < T - MAXH) Lmin = T - MAXH;  // at most MAXH high bits fit the pass header
  if (Lmin < 1) Lmin = 1;
  const uint64_t full = lowmask(n);
  const int budget0 = so.max_cost > 0 ? so.max_cost : 96;

  std::vector<int> rem(gates.size());
  for (size_t i = 0; i < gates.size(); ++i) rem[i] = (int)i;
  std::vector<char> done(gates.size(), 0);

  while (!rem.empty()) {
    PassPlan pp;
    pp.dual = so.dual;
    if (!tiles_ok || !fusable(gates[rem[0]])) {
      pp.tile = false;
      pp.single = rem[0];
      if (gates[rem[0]].kind == GK_RED) {
        pp.nred = 2;
        pp.red_base = red_total;
        gates[rem[0]].red_slot = (int)red_total;
        red_total += 2;
      }
      passes.push_back(pp);
      rem.erase(rem.begin());
      continue;
    }
    // 1) choose the tile's bit set
    uint64_t S = lowmask(Lmin);
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
    // 2) take every admissible gate with the tile fixed
    std::vector<int> chosen;
    blkN = blkD = 0;
    budget = budget0;
    for (int gi : rem) {
      const Gate& g = gates[gi];
      if (conflicts(g, blkN, blkD) || !fusable(g) || (g.nb & ~tilebits) ||
          (!chosen.empty() && g.cost > budget)) {
        block(g, blkN, blkD);
        continue;
      }
      chosen.push_back(gi);
      budget -= g.cost;
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

    // 4) emit the pass record
    pp.tile = true;
    pp.tb = TBv;
    pp.L = (uint32_t)geo.L;
    pp.m = (uint32_t)geo.H.size();
    for (size_t k = 0; k < geo.H.size(); ++k) pp.H[k] = (uint32_t)geo.H[k];
    pp.prog_off = prog.size();
    pp.red_base = red_total;
    Emit e(prog);
    for (int k = 0; k < PH_SIZE; ++k) e.u(0);
    uint32_t nred = 0;
    for (size_t ip = 0; ip < pp.phases.size(); ++ip) {
      const size_t hdr = prog.size();
      for (int k = 0; k < PS_SIZE; ++k) e.u(0);
      const size_t steps0 = prog.size();
      emit_phase(gates, pp.phases[ip], pp.Rpos[ip], geo, e, nred, red_total);
      // count steps
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
    pp.nred = nred;
    red_total += nred;
    pp.ntiles = (uint32_t)(1ull << (n - T));
    const uint32_t cap = nred ? 2048u : 65536u;
    pp.grid = pp.ntiles < cap ? pp.ntiles : cap;
    pp.lds = pass_lds_bytes(TBv, so.dual, nred);
    if (pp.lds > 160 * 1024) return fail(TQSV_EUNSUPPORTED, "pass needs more LDS than available");
    passes.push_back(pp);
  }
  return 0;
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
    default: return hipErrorInvalidValue;
  }
}

int run_passes(const std::vector<PassPlan>& passes, const std::vector<Gate>& gates, int n,
               const uint32_t* dprog, float2* psi, float2* lam, double* dpartials,
               double* dred_out, hipStream_t st) {
  for (const PassPlan& pp : passes) {
    if (pp.tile) {
      PassLaunch pl;
      pl.psi = psi;
      pl.lam = lam;
      pl.prog = dprog + pp.prog_off;
      pl.partials = dpartials;
      pl.ntiles = pp.ntiles;
      pl.grid = pp.grid;
      pl.lds_bytes = pp.lds;
      TQSV_HIP(launch_pass(pp.tb, pp.dual, n > 32, pl, st));
      if (pp.nred) TQSV_HIP(k_final_sum(dpartials, pp.grid, pp.nred, dred_out + pp.red_base, st));
    } else {
      const Gate& g = gates[pp.single];
      if (g.kind == GK_RED) {
        const uint32_t grid = reduce_grid(n);
        TQSV_HIP(k_pauli_dot(lam, psi, n, g.x, g.z, dpartials, st));
        TQSV_HIP(k_final_sum(dpartials, grid, 2, dred_out + pp.red_base, st));
      } else {
        TQSV_HIP(apply_single(g, n, psi, st));
        if (pp.dual) TQSV_HIP(apply_single(g, n, lam, st));
      }
    }
  }
  return 0;
}

}  // namespace tqsv
