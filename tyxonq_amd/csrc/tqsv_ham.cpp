// tqsv_ham.cpp - Pauli-sum expectation (fused tile reductions), H|psi>, and
// the adjoint / vector-Jacobian sweeps.
//
// Expectation.  Terms are grouped by flip mask x.  A group is evaluated in a
// register phase whose register positions contain x: every thread then holds
// both members of each pair (i, i^x) and forms conj(psi[i^x]) psi[i] once per
// pair; each term of the group only adds a signed copy of it (the sign of Z
// parts splits into a per-slot constant and a per-thread parity).  Groups are
// packed into read-only tile passes, so one pass serves every group whose
// flip bits lie in its tile.  Sums go through per-wave FP64 accumulators in
// LDS, per-work-group partials and a fixed-order final sum: the result does
// not depend on launch timing.
#include <algorithm>
#include <new>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "tqsv_planner.h"

using namespace tqsv;

struct tqsv_ham {
  int n = 0;
  int nterms = 0;
  int jit = 0;  // expectation passes run as generated kernels
  std::vector<uint64_t> x, z;
  std::vector<double> coeff;
  // expectation plan
  std::vector<PassPlan> passes;
  std::vector<uint32_t> prog;
  uint32_t* dprog = nullptr;
  // raw output slot per term (index nterms is the norm); mode 0: one tile sum,
  // mode 1: (re, im) pair from the dot kernels
  std::vector<int> tslot, tmode;
  uint32_t nout = 0;
  std::vector<int> small_terms;   // evaluated by the one-block-per-term kernel
  uint32_t small_base = 0;
  std::vector<int> big_terms;     // evaluated by a grid-wide dot kernel each
  uint64_t* d_sx = nullptr;
  uint64_t* d_sz = nullptr;
  double* dpartials = nullptr;
  size_t partials_cap = 0;
  double* dout = nullptr;
  std::vector<double> hout;
  // H|psi>
  uint32_t ngroups = 0;
  uint64_t* d_gx = nullptr;
  uint32_t* d_gstart = nullptr;
  uint64_t* d_tz = nullptr;
  float2* d_tc = nullptr;
  std::vector<int> apply_order;  // term of each H|psi> coefficient slot
  // the same H in the two-part form of k_ham_apply (HamApply): table
  // subgroups (n <= 32) and the remaining terms as flip groups
  std::vector<uint32_t> sub;               // 4 words per subgroup
  std::vector<std::vector<int>> sub_terms;  // terms of each subgroup
  // the same subgroups as H|psi> reads them: those alone on their flip mask
  // (4 words each, as in sub), and blocks of the subgroups that share one
  // (header {x, shifts, first member, end member}, members {zc, table offset})
  std::vector<uint32_t> sub1, blk, mem;
  uint32_t* d_sub1 = nullptr;
  uint32_t* d_blk = nullptr;
  uint32_t* d_mem = nullptr;
  // blocks whose members' Z masks differ from one base mask z0 in at most one
  // bit r each: sum_m T_m[q] (-1)^{|j & zc_m|} = (-1)^{|j & z0|} sum over
  // chunks of tables indexed by q and 4 consecutive bits of j (header {x,
  // shifts, z0, first chunk, end chunk}, chunks {lowest bit, table offset});
  // cdefs: per chunk the members (subgroup, bit - lowest bit, -1 for the base)
  std::vector<uint32_t> cblk, chk;
  struct ChunkDef {
    int xbits = 0;
    std::vector<std::pair<int, int>> members;
  };
  std::vector<ChunkDef> cdefs;
  uint32_t* d_cblk = nullptr;
  uint32_t* d_chk = nullptr;
  float2* d_stab = nullptr;
  size_t stab_size = 0;
  uint32_t nrest = 0;
  uint64_t* d_rgx = nullptr;
  uint32_t* d_rgstart = nullptr;
  uint64_t* d_rtz = nullptr;
  float2* d_rtc = nullptr;
  std::vector<int> rest_order;
  // diagonal terms summed per index (2^n floats) for H|psi>, when there are
  // enough of them and the state is small enough (see plan_apply)
  std::vector<int> diag_terms;
  float* d_diag = nullptr;
  uint64_t* d_dz = nullptr;
  double* d_dc = nullptr;
  bool jit_ready = false;        // generated expectation kernels compiled

  ~tqsv_ham() {
    (void)hipFree(dprog);
    (void)hipFree(d_sx);
    (void)hipFree(d_sz);
    (void)hipFree(dpartials);
    (void)hipFree(dout);
    (void)hipFree(d_gx);
    (void)hipFree(d_gstart);
    (void)hipFree(d_tz);
    (void)hipFree(d_tc);
    (void)hipFree(d_sub1);
    (void)hipFree(d_blk);
    (void)hipFree(d_mem);
    (void)hipFree(d_cblk);
    (void)hipFree(d_chk);
    (void)hipFree(d_stab);
    (void)hipFree(d_rgx);
    (void)hipFree(d_rgstart);
    (void)hipFree(d_rtz);
    (void)hipFree(d_rtc);
    (void)hipFree(d_diag);
    (void)hipFree(d_dz);
    (void)hipFree(d_dc);
  }
  tqsv::HamApply apply_args() const {
    tqsv::HamApply a;
    a.nsub = (uint32_t)(sub1.size() / 4);
    a.sub = d_sub1;
    a.nblk = (uint32_t)(blk.size() / 4);
    a.blk = d_blk;
    a.mem = d_mem;
    a.ncblk = (uint32_t)(cblk.size() / 5);
    a.cblk = d_cblk;
    a.chk = d_chk;
    a.tab = d_stab;
    a.ntab = stab_size;
    a.ngroups = nrest;
    a.gx = d_rgx;
    a.gstart = d_rgstart;
    a.tz = d_rtz;
    a.tc = d_rtc;
    a.diag = d_diag;
    return a;
  }
};

namespace {

inline int popc(uint64_t v) { return __builtin_popcountll(v); }
inline int ctz(uint64_t v) { return __builtin_ctzll(v); }
inline uint64_t lowmask(int L) { return L >= 64 ? ~0ull : ((1ull << L) - 1ull); }

template <typename T>
int upload(T** dptr, const std::vector<T>& v) {
  if (v.empty()) return 0;
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(dptr), v.size() * sizeof(T)));
  TQSV_HIP(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

constexpr uint32_t NRED_MAX = 2048;
// generated kernels keep one FP64 accumulator per term and thread in registers
constexpr uint32_t NRED_MAX_GEN = 40;

int plan_expval(tqsv_ham* h) {
  const int n = h->n;
  const uint32_t cap = h->jit ? NRED_MAX_GEN : NRED_MAX;
  const int T = effective_tile_bits(n, 13, false);
  const bool tiles = T >= 11;
  const int nt = h->nterms + 1;  // + norm
  auto tx = [&](int t) { return t < h->nterms ? h->x[t] : 0ull; };
  auto tz = [&](int t) { return t < h->nterms ? h->z[t] : 0ull; };
  h->tslot.assign(nt, -1);
  h->tmode.assign(nt, 0);

  std::map<uint64_t, std::vector<int>> groups;
  for (int t = 0; t < nt; ++t) {
    if (!tiles) {
      h->small_terms.push_back(t);
    } else if (popc(tx(t)) > RB) {
      h->big_terms.push_back(t);
    } else {
      groups[tx(t)].push_back(t);
    }
  }
  uint32_t nout = 0;
  if (tiles) {
    const int TB = T - RB;
    const int Lmin = std::min(5, T - RB);
    // groups split into chunks of at most cap/2 terms, most constrained first
    struct Chunk {
      uint64_t x;
      std::vector<int> terms;
    };
    std::vector<Chunk> chunks;
    for (auto& kv : groups)
      for (size_t i = 0; i < kv.second.size(); i += std::max<uint32_t>(cap / 2, 1)) {
        Chunk ch;
        ch.x = kv.first;
        ch.terms.assign(kv.second.begin() + i,
                        kv.second.begin() + std::min<size_t>(kv.second.size(), i + std::max<uint32_t>(cap / 2, 1)));
        chunks.push_back(ch);
      }
    std::stable_sort(chunks.begin(), chunks.end(), [&](const Chunk& a, const Chunk& b) {
      return popc(a.x & ~lowmask(Lmin)) > popc(b.x & ~lowmask(Lmin));
    });
    std::vector<uint64_t> order;
    for (const Chunk& ch : chunks) order.push_back(ch.x);
    std::vector<char> done(order.size(), 0);
    size_t left = order.size();
    while (left) {
      // tile bit set
      uint64_t S = lowmask(Lmin);
      uint32_t nred = 0;
      for (size_t i = 0; i < order.size(); ++i) {
        if (done[i]) continue;
        const uint64_t need = order[i] & ~S;
        const uint32_t c = (uint32_t)chunks[i].terms.size();
        if (popc(S | need) > T || nred + c > cap) continue;
        S |= need;
        nred += c;
      }
      int L = Lmin;
      for (;;) {
        const int nL = T - popc(S & ~lowmask(L));
        if (nL <= L) break;
        L = nL;
      }
      std::vector<int> H;
      for (uint64_t v = S & ~lowmask(L); v; v &= v - 1) H.push_back(ctz(v));
      if ((int)H.size() > MAXH || L + (int)H.size() != T)
        return fail(TQSV_EINVAL, "expectation planner built an invalid tile (internal error)");
      const uint64_t tileb = lowmask(L) | (S & ~lowmask(L));
      auto pos = [&](int b) {
        if (b < L) return b;
        for (size_t k = 0; k < H.size(); ++k)
          if (H[k] == b) return L + (int)k;
        return -1;
      };
      auto posmask = [&](uint64_t bits) {
        uint32_t m = 0;
        for (uint64_t v = bits; v; v &= v - 1) m |= 1u << pos(ctz(v));
        return m;
      };
      // groups of this pass
      std::vector<size_t> mine;
      nred = 0;
      for (size_t i = 0; i < order.size(); ++i) {
        if (done[i] || (order[i] & ~tileb)) continue;
        const uint32_t c = (uint32_t)chunks[i].terms.size();
        if (nred + c > cap) continue;
        mine.push_back(i);
        nred += c;
        done[i] = 1;
        --left;
      }
      // phases: pack flip masks into register sets
      PassPlan pp;
      pp.tile = true;
      pp.tb = TB;
      pp.L = (uint32_t)L;
      pp.m = (uint32_t)H.size();
      for (size_t k = 0; k < H.size(); ++k) pp.H[k] = (uint32_t)H[k];
      pp.prog_off = h->prog.size();
      pp.red_base = nout;
      std::vector<uint32_t>& w = h->prog;
      for (int k = 0; k < PH_SIZE; ++k) w.push_back(0);
      std::vector<size_t> prem = mine;
      uint32_t local = 0;
      uint32_t nphase = 0;
      while (!prem.empty()) {
        uint32_t R = 0;
        std::vector<size_t> sel, keep;
        for (size_t i : prem) {
          const uint32_t need = posmask(order[i]);
          if (popc((uint64_t)(R | need)) <= RB) R |= need;
        }
        for (int p = T - 1; popc((uint64_t)R) < RB && p >= 0; --p) R |= 1u << p;
        for (size_t i : prem) {
          if (posmask(order[i]) & ~R)
            keep.push_back(i);
          else
            sel.push_back(i);
        }
        std::array<uint32_t, RB> Rp{};
        int j = 0;
        for (int p = 0; p < T; ++p)
          if ((R >> p) & 1u) Rp[j++] = (uint32_t)p;
        auto slot_of_bit = [&](int b) -> int {
          const int p = pos(b);
          for (int q = 0; q < RB; ++q)
            if ((int)Rp[q] == p) return q;
          return -1;
        };
        auto slotmask = [&](uint64_t bits) {
          uint32_t m = 0;
          for (uint64_t v = bits; v; v &= v - 1) {
            const int s = slot_of_bit(ctz(v));
            if (s >= 0) m |= 1u << s;
          }
          return m;
        };
        const size_t hdr = w.size();
        for (int k = 0; k < PS_SIZE; ++k) w.push_back(0);
        w[hdr + PS_NSTEPS] = (uint32_t)sel.size();
        for (int r = 0; r < RB; ++r) w[hdr + PS_R + r] = Rp[r];
        for (size_t i : sel) {
          const uint64_t xm = order[i];
          const std::vector<int>& terms = chunks[i].terms;
          const size_t at = w.size();
          w.push_back(OP_RED);
          w.push_back(slotmask(xm));
          w.push_back(xm ? RED_PAIR : RED_DIAG);
          w.push_back((uint32_t)terms.size());
          for (int t : terms) {
            const uint64_t zz = tz(t);
            const uint32_t zs = slotmask(zz);
            uint32_t sm = 0;
            for (int s = 0; s < NSLOT; ++s)
              if (popc((uint64_t)(s & zs)) & 1) sm |= 1u << s;
            const int k = popc(xm & zz) & 3;
            w.push_back(local);
            w.push_back((uint32_t)zz);
            w.push_back((uint32_t)(zz >> 32));
            w.push_back(sm);
            w.push_back(xm ? (uint32_t)(k & 1) : 0u);
            h->tslot[t] = (int)(nout + local);
            h->tmode[t] = 0;
            ++local;
          }
          w[at] |= (uint32_t)((w.size() - at) << 16);
        }
        ++nphase;
        prem.swap(keep);
      }
      uint32_t* hd = w.data() + pp.prog_off;
      hd[PH_FLAGS] = 0;
      hd[PH_L] = pp.L;
      hd[PH_M] = pp.m;
      hd[PH_NPHASE] = nphase;
      hd[PH_NRED] = local;
      hd[PH_WORDS] = (uint32_t)(w.size() - pp.prog_off);
      for (uint32_t k = 0; k < pp.m; ++k) hd[PH_H + k] = pp.H[k];
      pp.prog_words = w.size() - pp.prog_off;
      pp.nred = local;
      pp.ntiles = (uint32_t)(1ull << (n - T));
      pp.grid = std::min<uint32_t>(pp.ntiles, 2048u);
      pp.lds = pass_lds_bytes(TB, MODE_RED, local);
      h->partials_cap = std::max(h->partials_cap, (size_t)pp.grid * local);
      nout += local;
      h->passes.push_back(pp);
    }
  }
  // dot-kernel terms: (re, im) pairs
  h->small_base = nout;
  for (int t : h->small_terms) {
    h->tslot[t] = (int)nout;
    h->tmode[t] = 1;
    nout += 2;
  }
  for (int t : h->big_terms) {
    h->tslot[t] = (int)nout;
    h->tmode[t] = 1;
    nout += 2;
  }
  h->partials_cap = std::max(h->partials_cap, (size_t)reduce_grid(n) * 2);
  h->nout = nout;
  h->hout.assign(nout, 0.0);
  return 0;
}

// device copies of the expectation plan (after planning is final)
int upload_expval(tqsv_ham* h) {
  if (upload(&h->dprog, h->prog)) return TQSV_EDEVICE;
  std::vector<uint64_t> sx, sz;
  for (int t : h->small_terms) {
    sx.push_back(t < h->nterms ? h->x[t] : 0ull);
    sz.push_back(t < h->nterms ? h->z[t] : 0ull);
  }
  if (upload(&h->d_sx, sx)) return TQSV_EDEVICE;
  if (upload(&h->d_sz, sz)) return TQSV_EDEVICE;
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&h->dpartials), h->partials_cap * sizeof(double)));
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&h->dout), std::max<uint32_t>(h->nout, 1) * sizeof(double)));
  return 0;
}

// coefficients of H|psi> in group order (of `order`): coeff * i^k
std::vector<float2> apply_coeffs(const tqsv_ham* h, const std::vector<int>& order) {
  std::vector<float2> tc;
  tc.reserve(order.size());
  for (int t : order) {
    const int k = popc(h->x[t] & h->z[t]) & 3;
    const double c = h->coeff[t];
    float2 v;
    switch (k) {
      case 0: v = make_float2((float)c, 0.f); break;
      case 1: v = make_float2(0.f, (float)c); break;
      case 2: v = make_float2((float)-c, 0.f); break;
      default: v = make_float2(0.f, (float)-c); break;
    }
    tc.push_back(v);
  }
  return tc;
}

// the coefficient tables of the table subgroups, formed in FP64: entry q of a
// subgroup with flip mask x = sum_t c_t i^{k_t} (-1)^{|dep_x(q) & z_t|}
std::vector<float2> sub_tables(const tqsv_ham* h) {
  static const std::complex<double> ipow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
  std::vector<std::complex<double>> t64;  // every subgroup's table, FP64
  std::vector<size_t> at(h->sub_terms.size());
  for (size_t g = 0; g < h->sub_terms.size(); ++g) {
    at[g] = t64.size();
    const uint32_t x = h->sub[4 * g];
    int pos[32];
    int m = 0;
    for (uint32_t v = x; v; v &= v - 1) pos[m++] = ctz(v);
    for (uint32_t q = 0; q < (1u << m); ++q) {
      uint64_t dep = 0;
      for (int k = 0; k < m; ++k)
        if ((q >> k) & 1u) dep |= 1ull << pos[k];
      std::complex<double> c(0, 0);
      for (int t : h->sub_terms[g]) {
        const double sg = (popc(dep & h->z[t]) & 1) ? -1.0 : 1.0;
        c += sg * h->coeff[t] * ipow[popc(h->x[t] & h->z[t]) & 3];
      }
      t64.push_back(c);
    }
  }
  std::vector<float2> tab;
  tab.reserve(t64.size());
  for (const auto& c : t64) tab.push_back(make_float2((float)c.real(), (float)c.imag()));
  // chunk tables: entry (q, v) = sum over the chunk's members of T_m[q], times
  // (-1)^(bit of v at the member's Z bit) (the base member: no sign)
  for (const auto& cd : h->cdefs) {
    const uint32_t nq = 1u << cd.xbits;
    for (uint32_t q = 0; q < nq; ++q)
      for (uint32_t v = 0; v < 16; ++v) {
        std::complex<double> c(0, 0);
        for (const auto& mb : cd.members) {
          const double sg = mb.second >= 0 && ((v >> mb.second) & 1u) ? -1.0 : 1.0;
          c += sg * t64[at[(size_t)mb.first] + q];
        }
        tab.push_back(make_float2((float)c.real(), (float)c.imag()));
      }
  }
  return tab;
}

// the diagonal vector from the current coefficients (on stream st, waited for)
int build_diag(tqsv_ham* h, hipStream_t st) {
  std::vector<double> dc;
  for (int t : h->diag_terms) dc.push_back(h->coeff[t]);
  TQSV_HIP(hipMemcpyAsync(h->d_dc, dc.data(), dc.size() * sizeof(double), hipMemcpyHostToDevice, st));
  TQSV_HIP(k_diag_build(h->d_diag, h->n, (uint32_t)dc.size(), h->d_dz, h->d_dc, st));
  TQSV_HIP(hipStreamSynchronize(st));
  return 0;
}

int plan_apply(tqsv_ham* h) {
  std::map<uint64_t, std::vector<int>> groups;
  for (int t = 0; t < h->nterms; ++t) groups[h->x[t]].push_back(t);
  std::vector<uint64_t> gx, tzv;
  std::vector<uint32_t> gstart;
  h->apply_order.clear();
  for (auto& kv : groups) {
    gx.push_back(kv.first);
    gstart.push_back((uint32_t)tzv.size());
    for (int t : kv.second) {
      tzv.push_back(h->z[t]);
      h->apply_order.push_back(t);
    }
  }
  gstart.push_back((uint32_t)tzv.size());
  h->ngroups = (uint32_t)gx.size();
  if (upload(&h->d_gx, gx) || upload(&h->d_gstart, gstart) || upload(&h->d_tz, tzv) ||
      upload(&h->d_tc, apply_coeffs(h, h->apply_order)))
    return TQSV_EDEVICE;
  // two-part form: terms sharing x (1..4 bits) and the Z mask outside x,
  // at least 2 of them, become a table subgroup; the rest stay per term
  constexpr int kTabBits = 4;
  std::map<std::pair<uint64_t, uint64_t>, std::vector<int>> subs;
  if (h->n <= 32)
    for (int t = 0; t < h->nterms; ++t)
      if (h->x[t] && popc(h->x[t]) <= kTabBits) subs[{h->x[t], h->z[t] & ~h->x[t]}].push_back(t);
  std::vector<char> in_table(h->nterms, 0);
  h->sub.clear();
  h->sub_terms.clear();
  uint32_t off = 0;
  for (auto& kv : subs) {
    if (kv.second.size() < 2) continue;
    const uint64_t x = kv.first.first;
    // shifts: bit k of the pattern is bit k of j >> (position of the k-th
    // flip bit - k); 5 bits each, 31 for unused positions (bit k >= 1 of
    // j >> 31 is 0)
    uint32_t pk = 0;
    int k = 0;
    for (uint64_t v = x; v; v &= v - 1, ++k) pk |= (uint32_t)(ctz(v) - k) << (5 * k);
    for (; k < kTabBits; ++k) pk |= 31u << (5 * k);
    h->sub.insert(h->sub.end(), {(uint32_t)x, (uint32_t)kv.first.second, off, pk});
    h->sub_terms.push_back(kv.second);
    for (int t : kv.second) in_table[t] = 1;
    off += 1u << popc(x);
  }
  // diagonal terms (x = 0): summed per index into a vector once (and after
  // new coefficients) when there are at least 8 of them and the vector (4 B
  // per amplitude) stays at most 256 MB; H|psi> then reads one float per
  // amplitude instead of forming a parity per term
  constexpr int kDiagMinTerms = 8;
  constexpr int kDiagMaxQubits = 26;
  h->diag_terms.clear();
  if (h->n <= kDiagMaxQubits) {
    for (int t = 0; t < h->nterms; ++t)
      if (!h->x[t]) h->diag_terms.push_back(t);
    if ((int)h->diag_terms.size() < kDiagMinTerms) h->diag_terms.clear();
  }
  for (int t : h->diag_terms) in_table[t] = 1;
  std::map<uint64_t, std::vector<int>> rest;
  for (int t = 0; t < h->nterms; ++t)
    if (!in_table[t]) rest[h->x[t]].push_back(t);
  std::vector<uint64_t> rgx, rtz;
  std::vector<uint32_t> rgstart;
  h->rest_order.clear();
  for (auto& kv : rest) {
    rgx.push_back(kv.first);
    rgstart.push_back((uint32_t)rtz.size());
    for (int t : kv.second) {
      rtz.push_back(h->z[t]);
      h->rest_order.push_back(t);
    }
  }
  rgstart.push_back((uint32_t)rtz.size());
  h->nrest = (uint32_t)rgx.size();
  // subgroups sharing a flip mask (consecutive in sub, which is ordered by x)
  // form a block: one partner load, pattern and complex product for all of
  // them, one table load and Z sign per member
  h->sub1.clear();
  h->blk.clear();
  h->mem.clear();
  h->cblk.clear();
  h->chk.clear();
  h->cdefs.clear();
  const size_t nsubg = h->sub.size() / 4;
  uint32_t coff = off;  // chunk tables follow the subgroup tables
  for (size_t a = 0; a < nsubg;) {
    size_t b = a + 1;
    while (b < nsubg && h->sub[4 * b] == h->sub[4 * a]) ++b;
    // chunk form: a base Z mask z0 among the members from which every
    // member differs in at most one bit
    int base = -1;
    for (size_t k0 = a; b - a >= 3 && k0 < b && base < 0; ++k0) {
      bool ok = true;
      for (size_t k = a; k < b && ok; ++k) ok = popc(h->sub[4 * k + 1] ^ h->sub[4 * k0 + 1]) <= 1;
      if (ok) base = (int)k0;
    }
    std::vector<int> rbits;
    if (base >= 0) {
      const uint32_t z0 = h->sub[4 * (size_t)base + 1];
      for (size_t k = a; k < b; ++k)
        if (const uint32_t d = h->sub[4 * k + 1] ^ z0) rbits.push_back(ctz(d));
      std::sort(rbits.begin(), rbits.end());
      size_t nch = 0;
      for (size_t r = 0; r < rbits.size(); ++nch) {
        const int lo = rbits[r];
        while (r < rbits.size() && rbits[r] < lo + 4) ++r;
      }
      if (nch == 0 || nch >= b - a) base = -1;  // no gain over the member loop
    }
    if (base >= 0) {
      const uint32_t x = h->sub[4 * a], z0 = h->sub[4 * (size_t)base + 1];
      const int xbits = popc(x);
      const uint32_t c0 = (uint32_t)(h->chk.size() / 2);
      size_t r = 0;
      while (r < rbits.size()) {
        const int lo = rbits[r];
        tqsv_ham::ChunkDef cd;
        cd.xbits = xbits;
        for (size_t k = a; k < b; ++k) {
          const uint32_t d = h->sub[4 * k + 1] ^ z0;
          if (d == 0) {
            if (c0 == h->chk.size() / 2) cd.members.push_back({(int)k, -1});  // base: first chunk
          } else if (ctz(d) >= lo && ctz(d) < lo + 4) {
            cd.members.push_back({(int)k, ctz(d) - lo});
          }
        }
        while (r < rbits.size() && rbits[r] < lo + 4) ++r;
        h->chk.insert(h->chk.end(), {(uint32_t)lo, coff});
        coff += (1u << xbits) * 16u;
        h->cdefs.push_back(std::move(cd));
      }
      h->cblk.insert(h->cblk.end(), {x, h->sub[4 * a + 3], z0, c0, (uint32_t)(h->chk.size() / 2)});
    } else if (b - a == 1) {
      h->sub1.insert(h->sub1.end(), h->sub.begin() + 4 * a, h->sub.begin() + 4 * a + 4);
    } else {
      const uint32_t m0 = (uint32_t)(h->mem.size() / 2);
      for (size_t k = a; k < b; ++k) h->mem.insert(h->mem.end(), {h->sub[4 * k + 1], h->sub[4 * k + 2]});
      h->blk.insert(h->blk.end(), {h->sub[4 * a], h->sub[4 * a + 3], m0, m0 + (uint32_t)(b - a)});
    }
    a = b;
  }
  const std::vector<float2> tab = sub_tables(h);
  h->stab_size = tab.size();
  if (upload(&h->d_sub1, h->sub1) || upload(&h->d_blk, h->blk) ||
      upload(&h->d_mem, h->mem) || upload(&h->d_cblk, h->cblk) || upload(&h->d_chk, h->chk) ||
      upload(&h->d_stab, tab) || upload(&h->d_rgx, rgx) ||
      upload(&h->d_rgstart, rgstart) || upload(&h->d_rtz, rtz) ||
      upload(&h->d_rtc, apply_coeffs(h, h->rest_order)))
    return TQSV_EDEVICE;
  if (!h->diag_terms.empty()) {
    std::vector<uint64_t> dz;
    for (int t : h->diag_terms) dz.push_back(h->z[t]);
    if (upload(&h->d_dz, dz)) return TQSV_EDEVICE;
    TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&h->d_dc), h->diag_terms.size() * sizeof(double)));
    TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&h->d_diag), sizeof(float) << h->n));
    const int rc = build_diag(h, nullptr);
    if (rc) return rc;
  }
  return 0;
}

// <P_t> from the raw sums
double term_value(const tqsv_ham* h, int t) {
  const uint64_t xm = t < h->nterms ? h->x[t] : 0ull;
  const uint64_t zm = t < h->nterms ? h->z[t] : 0ull;
  const int k = popc(xm & zm) & 3;
  const int s = h->tslot[t];
  if (h->tmode[t] == 0) {
    const double S = h->hout[s];
    if (!xm) return S;
    static const double f[4] = {2.0, -2.0, -2.0, 2.0};
    return f[k] * S;
  }
  // Re((-i)^k S)
  const double re = h->hout[s], im = h->hout[s + 1];
  switch (k) {
    case 0: return re;
    case 1: return im;
    case 2: return -re;
    default: return -im;
  }
}

int run_expval(tqsv_ham* h, const float2* psi, hipStream_t st) {
  const int n = h->n;
  if (h->jit && !h->jit_ready) {  // compiled on first use: H|psi> alone needs no kernels
    const int rc = jit_prepare(h->passes, h->prog, n);
    if (rc) return rc;
    h->jit_ready = true;
  }
  for (const PassPlan& pp : h->passes) {
    PassLaunch pl;
    pl.psi = const_cast<float2*>(psi);
    pl.lam = nullptr;
    pl.prog = h->dprog + pp.prog_off;
    pl.partials = h->dpartials;
    pl.ntiles = pp.ntiles;
    pl.grid = pp.grid;
    pl.lds_bytes = pp.lds;
    if (pp.jit_fn)
      TQSV_HIP(jit_launch(pp, pl, st));
    else
      TQSV_HIP(launch_pass(pp.tb, MODE_RED, n > 32, false, pl, st));
    TQSV_HIP(k_final_sum(h->dpartials, pp.grid, pp.nred, h->dout + pp.red_base, st));
  }
  if (!h->small_terms.empty())
    TQSV_HIP(k_pauli_dot_terms(psi, psi, n, (uint32_t)h->small_terms.size(), h->d_sx, h->d_sz,
                               h->dout + h->small_base, st));
  for (int t : h->big_terms) {
    TQSV_HIP(k_pauli_dot(psi, psi, n, h->x[t], h->z[t], h->dpartials, st));
    TQSV_HIP(k_final_sum(h->dpartials, reduce_grid(n), 2, h->dout + h->tslot[t], st));
  }
  TQSV_HIP(hipMemcpyAsync(h->hout.data(), h->dout, h->nout * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  TQSV_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

// ---- plan build helpers shared with tqsv_abi.cpp ---------------------------
namespace tqsv {

int plan_upload(uint32_t** dptr, size_t* cap, const std::vector<uint32_t>& prog, hipStream_t st) {
  const size_t need = std::max<size_t>(prog.size(), 1);
  if (*cap < need) {
    if (*dptr) TQSV_HIP(hipFree(*dptr));
    *dptr = nullptr;
    TQSV_HIP(hipMalloc(reinterpret_cast<void**>(dptr), need * sizeof(uint32_t)));
    *cap = need;
  }
  if (!prog.empty()) {
    TQSV_HIP(hipMemcpyAsync(*dptr, prog.data(), prog.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, st));
    TQSV_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

static size_t count_tile(const std::vector<PassPlan>& passes) {
  size_t k = 0;
  for (const PassPlan& pp : passes) k += pp.tile ? 1 : 0;
  return k;
}

// Generated kernels for a freshly emitted program.  The source of a pass
// depends only on its non-value words, so when the program with its value
// words zeroed (the skeleton) equals the one last compiled for this plan, the
// previous build's kernels are kept and no source is generated; otherwise
// they are compiled or taken from the process-wide cache (jit_prepare).
static int attach_kernels(std::vector<PassPlan>& passes, const std::vector<uint32_t>& prog,
                          const std::vector<uint32_t>& fpos, const std::vector<PassPlan>& old,
                          std::vector<uint32_t>& skel, std::vector<uint32_t>& work, int n) {
  work.assign(prog.begin(), prog.end());
  for (uint32_t i : fpos) work[i] = 0;
  bool same = !skel.empty() && work == skel && old.size() == passes.size();
  for (size_t i = 0; same && i < passes.size(); ++i)
    same = passes[i].tile == old[i].tile && passes[i].tb == old[i].tb &&
           passes[i].prog_off == old[i].prog_off && (!passes[i].tile || old[i].jit_fn);
  if (same) {
    for (size_t i = 0; i < passes.size(); ++i) passes[i].jit_fn = old[i].jit_fn;
    return 0;
  }
  skel.clear();
  const int rc = jit_prepare(passes, prog, n);
  if (rc) return rc;
  skel.swap(work);
  return 0;
}

int plan_build_forward(tqsv_plan* p, const tqsv_op* ops, int nops, hipStream_t st) {
  std::vector<Gate>& g = p->gwork;
  int rc = gates_from_ops(p->nu, ops, nops, g);
  if (rc) return rc;
  merge_and_classify(g);
  const bool same = same_structure(g, p->gates) && !p->fsel.empty();
  if (p->fgen && g.size() == p->gates.size())
    for (size_t i = 0; i < g.size(); ++i) g[i].rot_form = p->gates[i].rot_form;
  std::vector<PassPlan> sel;
  if (same) {
    sel = p->fsel;
  } else {
    rc = select_passes(g, p->so, sel);
    if (rc) return rc;
  }
  std::vector<PassPlan>& passes = p->pwork;
  std::vector<uint32_t>& prog = p->progwork;
  passes.clear();
  prog.clear();
  uint32_t red_total = 0;
  std::vector<uint32_t>& fpos = p->fposwork;
  fpos.clear();
  rc = emit_passes(g, p->so, sel, passes, prog, red_total, &fpos);
  if (rc) return rc;
  p->fsel.swap(sel);
  p->max_param = -1;
  for (const Gate& x : g) p->max_param = std::max(p->max_param, x.param);
  p->fgen = jit_policy(p->jit_req, p->n, count_tile(passes), 64) != 0;
  if (p->fgen) {
    rc = attach_kernels(passes, prog, fpos, p->passes, p->fskel, p->skelwork, p->n);
    if (rc) return rc;
  } else {
    p->fskel.clear();
  }
  p->gates.swap(g);
  p->passes.swap(passes);
  p->prog.swap(prog);
  p->adj_built = false;
  return plan_upload(&p->dprog, &p->dprog_cap, p->prog, st);
}

// grow *d to `need` doubles keeping its first `keep`: work queued on st may
// still write that region, so the stream is drained first
static int grow_keep(double** d, size_t* cap, size_t need, size_t keep, hipStream_t st) {
  need = std::max<size_t>(need, 1);
  if (*cap >= need) return 0;
  TQSV_HIP(hipStreamSynchronize(st));
  double* nd = nullptr;
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(&nd), need * sizeof(double)));
  if (*d && keep) {
    const hipError_t e = hipMemcpy(nd, *d, std::min(keep, *cap) * sizeof(double), hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
      (void)hipFree(nd);
      return hip_fail(e, "grow_keep");
    }
  }
  if (*d) TQSV_HIP(hipFree(*d));
  *d = nd;
  *cap = need;
  return 0;
}

static int ensure(double** d, size_t* cap, size_t need) {
  need = std::max<size_t>(need, 1);
  if (*cap >= need) return 0;
  if (*d) TQSV_HIP(hipFree(*d));
  *d = nullptr;
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(d), need * sizeof(double)));
  *cap = need;
  return 0;
}

void reverse_gates(const std::vector<Gate>& gates, std::vector<Gate>& r) {
  r.clear();
  for (int i = (int)gates.size() - 1; i >= 0; --i) {
    const Gate& g = gates[i];
    bool shared = g.kind == GK_PGRP && g.nmem > 0 && g.mp[0] >= 0;
    for (int mi = 1; shared && mi < g.nmem; ++mi) shared = g.mp[mi] == g.mp[0];
    if (shared) {
      // one parameter for the whole group: dE/dtheta = Im <lambda| G |psi>
      // with G = sum_r mg[r] P_r, one reduction over the group's pairs
      Gate red = g;
      red.kind = GK_RED;
      red.k = 0;
      red.param = g.mp[0];
      red.gcoef = 1.0;
      classify_gate(red);
      r.push_back(red);
    } else if (g.kind == GK_PGRP) {
      // every member commutes with the whole group, so all their reductions
      // are taken at one point, before the group is undone
      for (int mi = 0; mi < g.nmem; ++mi) {
        if (g.mp[mi] < 0) continue;
        Gate red;
        red.kind = GK_RED;
        red.x = g.x;
        red.z = g.mz[mi];
        red.k = popc(g.x & g.mz[mi]) & 3;
        red.param = g.mp[mi];
        red.gcoef = g.mg[mi];
        classify_gate(red);
        r.push_back(red);
      }
    } else if (g.param >= 0) {
      Gate red;
      red.kind = GK_RED;
      red.x = g.gx;
      red.z = g.gz;
      red.k = popc(g.gx & g.gz) & 3;
      red.param = g.param;
      red.gcoef = g.gcoef;
      classify_gate(red);
      r.push_back(red);
    }
    r.push_back(dagger(g));
  }
}

int plan_build_adjoint(tqsv_plan* p, hipStream_t st) {
  if (p->adj_built) return 0;
  std::vector<Gate>& r = p->rwork;
  reverse_gates(p->gates, r);
  const bool same = same_structure(r, p->rgates) && !p->rsel.empty();
  if (p->rgen && r.size() == p->rgates.size())
    for (size_t i = 0; i < r.size(); ++i) r[i].rot_form = p->rgates[i].rot_form;
  SchedOpts so = p->so;
  so.dual = true;
  std::vector<PassPlan> sel;
  int rc = 0;
  if (same) {
    sel = p->rsel;
  } else {
    rc = select_passes(r, so, sel);
    if (rc) return rc;
  }
  std::vector<PassPlan>& passes = p->rpwork;
  std::vector<uint32_t>& prog = p->rprogwork;
  passes.clear();
  prog.clear();
  uint32_t red_total = 0;
  std::vector<uint32_t>& fpos = p->fposwork;
  fpos.clear();
  rc = emit_passes(r, so, sel, passes, prog, red_total, &fpos);
  if (rc) return rc;
  p->rsel.swap(sel);
  p->rgen = jit_policy(p->jit_req, p->n, count_tile(passes), 64) != 0;
  if (p->rgen) {
    rc = attach_kernels(passes, prog, fpos, p->rpasses, p->rskel, p->skelwork, p->n);
    if (rc) return rc;
  } else {
    p->rskel.clear();
  }
  p->rgates.swap(r);
  p->rpasses.swap(passes);
  p->rprog.swap(prog);
  p->nred_adj = (int)red_total;
  // every pass's block partials in one buffer, summed by one kernel after the
  // sweep: one entry {offset, blocks, stride, index} per reduction slot; the
  // energy's partials (from H|psi>, tqsv_adjoint) first, its two entries last
  const uint32_t eg = ham_apply_grid(p->nu);  // H|psi> runs on the caller's qubits
  size_t pcap = 2 * (size_t)eg;
  p->hslots.assign(4 * ((size_t)red_total + 2), 0);
  auto slot = [&](uint32_t r, size_t off, uint32_t nb, uint32_t stride, uint32_t k) {
    uint32_t* d = p->hslots.data() + 4 * (size_t)r;
    d[0] = (uint32_t)off;
    d[1] = nb;
    d[2] = stride;
    d[3] = k;
  };
  for (PassPlan& pp : p->rpasses) {
    if (pp.tile && pp.nred) {
      pp.part_off = pcap;
      for (uint32_t k = 0; k < pp.nred; ++k) slot(pp.red_base + k, pcap, pp.grid, pp.nred, k);
      pcap += (size_t)pp.grid * pp.nred;
    } else if (!pp.tile && p->rgates[pp.single].kind == GK_RED) {
      const uint32_t g = reduce_grid(p->n);
      pp.part_off = pcap;
      for (uint32_t k = 0; k < 2; ++k) slot(pp.red_base + k, pcap, g, 2, k);
      pcap += 2 * (size_t)g;
    }
  }
  slot(red_total, 0, eg, 2, 0);
  slot(red_total + 1, 0, eg, 2, 1);
  if (pcap >= (1ull << 32)) return fail(TQSV_EINVAL, "reverse sweep: too many reduction partials");
  rc = plan_upload(&p->drprog, &p->drprog_cap, p->rprog, st);
  if (rc) return rc;
  rc = plan_upload(&p->dslots, &p->dslots_cap, p->hslots, st);
  if (rc) return rc;
  // the energy partials of a queued H|psi> may already be in the buffer
  rc = grow_keep(&p->dpartials, &p->dpartials_cap, pcap, 2 * (size_t)eg, st);
  if (rc) return rc;
  rc = ensure(&p->dred, &p->dred_cap, (size_t)red_total + 2);
  if (rc) return rc;
  p->adj_built = true;
  return 0;
}

// dE/dparam from the reverse sweep's reduction slots:
// grad += scale * gcoef * red_scale * Im((-i)^k S) per parametrised gate
static void collect_grad(const tqsv_plan* p, const double* hred, int nparams, double* grad,
                         double scale) {
  for (int i = 0; i < nparams; ++i) grad[i] = 0.0;
  for (const Gate& g : p->rgates) {
    if (g.kind != GK_RED) continue;
    if (g.nmem > 0) {  // group reduction: one slot holding Im <lambda|G|psi>
      grad[g.param] += scale * g.gcoef * g.red_scale * hred[g.red_slot];
      continue;
    }
    // <lambda|P|psi> = i^k <lambda|X^x Z^z|psi>; the gradient takes its
    // imaginary part: Im(c), -Re(c), -Im(c), Re(c) for k = 0..3.  Tile passes
    // store only that part of c, single-op passes store re and im.
    const double* h = hred + g.red_slot;
    const double c_re = g.red_part == 1 ? 0.0 : h[0];
    const double c_im = g.red_part == 0 ? 0.0 : g.red_part == 1 ? h[0] : h[1];
    double v;
    switch (g.k) {
      case 0: v = c_im; break;
      case 1: v = -c_re; break;
      case 2: v = -c_im; break;
      default: v = c_re; break;
    }
    grad[g.param] += scale * g.gcoef * g.red_scale * v;
  }
}

// reverse sweep + gradient assembly; scale = 1 for energies, 1/2 for VJPs
static int reverse_and_collect(tqsv_plan* p, float2* psi, float2* lam, int nparams, double* grad,
                               double scale, double* hred, hipStream_t st, bool energy) {
  // with `energy` the final sum also forms the energy slots (after the
  // reverse sweep's) from the partials of the queued H|psi>
  int rc = run_passes(p->rpasses, p->rgates, p->n, p->drprog, psi, lam, p->dpartials, p->dred, st,
                      p->dslots, (uint32_t)p->nred_adj + (energy ? 2u : 0u));
  if (rc) return rc;
  const size_t nr = (size_t)p->nred_adj + 2;
  TQSV_HIP(hipMemcpyAsync(hred, p->dred, nr * sizeof(double), hipMemcpyDeviceToHost, st));
  TQSV_HIP(hipStreamSynchronize(st));
  collect_grad(p, hred, nparams, grad, scale);
  return 0;
}

}  // namespace tqsv

struct tqsv_batch {
  uint32_t* doff = nullptr;
  size_t doff_cap = 0;
  tqsv::BatchItem* ditems = nullptr;
  size_t items_cap = 0;
  double* dout = nullptr;
  size_t out_cap = 0;
  ~tqsv_batch() {
    (void)hipFree(doff);
    (void)hipFree(ditems);
    (void)hipFree(dout);
  }
};

namespace tqsv {
template <typename T>
static int grow(T** d, size_t* cap, size_t need) {
  need = std::max<size_t>(need, 1);
  if (*cap >= need) return 0;
  if (*d) TQSV_HIP(hipFree(*d));
  *d = nullptr;
  *cap = 0;
  TQSV_HIP(hipMalloc(reinterpret_cast<void**>(d), need * sizeof(T)));
  *cap = need;
  return 0;
}
}  // namespace tqsv

// ---- C ABI -----------------------------------------------------------------
extern "C" {

int tqsv_ham_create(int n, int nterms, const uint64_t* xmask, const uint64_t* zmask,
                    const double* coeff, tqsv_ham** out) {
  try {
    if (!out) return fail(TQSV_EINVAL, "tqsv_ham_create: out is NULL");
    *out = nullptr;
    if (n < 1 || n > 34) return fail(TQSV_EINVAL, "tqsv_ham_create: n must be in [1, 34]");
    if (nterms < 0 || (nterms > 0 && (!xmask || !zmask || !coeff)))
      return fail(TQSV_EINVAL, "tqsv_ham_create: bad term arrays");
    const uint64_t full = lowmask(n);
    auto h = std::make_unique<tqsv_ham>();
    h->n = n;
    h->nterms = nterms;
    for (int t = 0; t < nterms; ++t) {
      if ((xmask[t] | zmask[t]) & ~full)
        return fail(TQSV_EINVAL, "tqsv_ham_create: term " + std::to_string(t) + " acts outside n qubits");
      h->x.push_back(xmask[t]);
      h->z.push_back(zmask[t]);
      h->coeff.push_back(coeff[t]);
    }
    // generated kernels hold at most NRED_MAX_GEN terms per pass; when that
    // gives too many passes (compile time), plan for the precompiled kernel
    h->jit = jit_policy(0, n, 0, 32);
    int rc = plan_expval(h.get());
    if (rc) return rc;
    if (h->jit && !jit_policy(0, n, h->passes.size(), 32)) {
      h->jit = 0;
      h->passes.clear();
      h->prog.clear();
      h->small_terms.clear();
      h->big_terms.clear();
      h->nout = 0;
      h->partials_cap = 0;
      rc = plan_expval(h.get());
      if (rc) return rc;
    }
    rc = upload_expval(h.get());
    if (rc) return rc;
    rc = plan_apply(h.get());
    if (rc) return rc;
    *out = h.release();
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_ham_create: out of host memory");
  } catch (const std::exception& e) {
    return fail(TQSV_EINVAL, std::string("tqsv_ham_create: ") + e.what());
  }
}

int tqsv_ham_precompile(int n, int nterms, const uint64_t* xmask, const uint64_t* zmask,
                        const double* coeff, int* kernels) {
  try {
    if (n < 1 || n > 34) return fail(TQSV_EINVAL, "tqsv_ham_precompile: n must be in [1, 34]");
    if (nterms < 0 || (nterms > 0 && (!xmask || !zmask || !coeff)))
      return fail(TQSV_EINVAL, "tqsv_ham_precompile: bad term arrays");
    const uint64_t full = lowmask(n);
    tqsv_ham h;  // host planning only: no device buffer is allocated
    h.n = n;
    h.nterms = nterms;
    for (int t = 0; t < nterms; ++t) {
      if ((xmask[t] | zmask[t]) & ~full)
        return fail(TQSV_EINVAL, "tqsv_ham_precompile: term " + std::to_string(t) + " acts outside n qubits");
      h.x.push_back(xmask[t]);
      h.z.push_back(zmask[t]);
      h.coeff.push_back(coeff[t]);
    }
    // the plan tqsv_ham_create makes (same policy and fallback)
    h.jit = jit_policy(0, n, 0, 32);
    int rc = plan_expval(&h);
    if (rc) return rc;
    if (h.jit && !jit_policy(0, n, h.passes.size(), 32)) {
      h.jit = 0;
      h.passes.clear();
      h.prog.clear();
      h.small_terms.clear();
      h.big_terms.clear();
      h.nout = 0;
      h.partials_cap = 0;
      rc = plan_expval(&h);
      if (rc) return rc;
    }
    int k = 0;
    if (h.jit) {
      rc = jit_compile_only(h.passes, h.prog, n, &k);
      if (rc) return rc;
    }
    if (kernels) *kernels = k;
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_ham_precompile: out of host memory");
  }
}

int tqsv_ham_set_coeffs(tqsv_ham* h, int nterms, const double* coeff, void* stream) {
  try {
    if (!h || (nterms > 0 && !coeff)) return fail(TQSV_EINVAL, "tqsv_ham_set_coeffs: NULL argument");
    if (nterms != h->nterms) return fail(TQSV_EINVAL, "tqsv_ham_set_coeffs: term count differs");
    for (int t = 0; t < nterms; ++t) h->coeff[t] = coeff[t];
    if (h->apply_order.empty()) return 0;
    const std::vector<float2> tc = apply_coeffs(h, h->apply_order);
    const std::vector<float2> rtc = apply_coeffs(h, h->rest_order);
    const std::vector<float2> tab = sub_tables(h);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    TQSV_HIP(hipMemcpyAsync(h->d_tc, tc.data(), tc.size() * sizeof(float2), hipMemcpyHostToDevice, st));
    if (!rtc.empty())
      TQSV_HIP(hipMemcpyAsync(h->d_rtc, rtc.data(), rtc.size() * sizeof(float2), hipMemcpyHostToDevice, st));
    if (!tab.empty())
      TQSV_HIP(hipMemcpyAsync(h->d_stab, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice, st));
    TQSV_HIP(hipStreamSynchronize(st));
    if (!h->diag_terms.empty()) return build_diag(h, st);
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_ham_set_coeffs: out of host memory");
  }
}

int tqsv_ham_expval(tqsv_ham* h, const void* psi, double* terms, double* energy, double* norm2,
                    void* stream) {
  if (!h || !psi) return fail(TQSV_EINVAL, "tqsv_ham_expval: NULL argument");
  int rc = run_expval(h, reinterpret_cast<const float2*>(psi), reinterpret_cast<hipStream_t>(stream));
  if (rc) return rc;
  double e = 0.0;
  for (int t = 0; t < h->nterms; ++t) {
    const double v = term_value(h, t);
    if (terms) terms[t] = v;
    e += h->coeff[t] * v;
  }
  if (energy) *energy = e;
  if (norm2) *norm2 = term_value(h, h->nterms);
  return 0;
}

int tqsv_ham_apply(tqsv_ham* h, const void* psi, void* out, void* stream) {
  if (!h || !psi || !out) return fail(TQSV_EINVAL, "tqsv_ham_apply: NULL argument");
  if (psi == out) return fail(TQSV_EINVAL, "tqsv_ham_apply: out must not alias psi");
  TQSV_HIP(k_ham_apply(reinterpret_cast<const float2*>(psi), reinterpret_cast<float2*>(out), h->n,
                       h->apply_args(), reinterpret_cast<hipStream_t>(stream)));
  return 0;
}

int tqsv_ham_get_info(const tqsv_ham* h, tqsv_ham_info* info) {
  if (!h || !info) return fail(TQSV_EINVAL, "tqsv_ham_get_info: NULL argument");
  std::memset(info, 0, sizeof(*info));
  info->n = h->n;
  info->nterms = h->nterms;
  int phases = 0;
  for (const PassPlan& pp : h->passes) phases += (int)h->prog[pp.prog_off + PH_NPHASE];
  info->passes = (int32_t)(h->passes.size() + h->big_terms.size() + (h->small_terms.empty() ? 0 : 1));
  info->phases = phases;
  std::map<uint64_t, int> g;
  for (int t = 0; t < h->nterms; ++t) g[h->x[t]] = 1;
  info->flip_groups = (int32_t)g.size();
  info->unfused_terms = (int32_t)(h->big_terms.size() + h->small_terms.size());
  return 0;
}

void tqsv_ham_destroy(tqsv_ham* h) { delete h; }

int tqsv_expval(const void* psi, int n, int nterms, const uint64_t* xmask, const uint64_t* zmask,
                const double* coeff, double* terms, double* energy, double* norm2, void* stream) {
  tqsv_ham* h = nullptr;
  int rc = tqsv_ham_create(n, nterms, xmask, zmask, coeff, &h);
  if (rc) return rc;
  rc = tqsv_ham_expval(h, psi, terms, energy, norm2, stream);
  tqsv_ham_destroy(h);
  return rc;
}

int tqsv_adjoint(tqsv_plan* p, tqsv_ham* h, void* psi, void* lam, int nparams, double* grad,
                 double* energy, void* stream) {
  try {
    if (!p || !h || !psi || !lam || (nparams > 0 && !grad))
      return fail(TQSV_EINVAL, "tqsv_adjoint: NULL argument");
    if (p->nu != h->n) return fail(TQSV_EINVAL, "tqsv_adjoint: plan and Hamiltonian sizes differ");
    if (nparams <= p->max_param)
      return fail(TQSV_EINVAL, "tqsv_adjoint: nparams smaller than the largest parameter index + 1");
    if (psi == lam) return fail(TQSV_EINVAL, "tqsv_adjoint: lam must not alias psi");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // the energy partials lead the plan's partials buffer (nothing is queued
    // on the stream that still uses it: every sweep ends synchronised)
    const uint32_t eg = ham_apply_grid(h->n);
    int rc = grow_keep(&p->dpartials, &p->dpartials_cap, 2 * (size_t)eg, 0, st);
    if (rc) return rc;
    const bool pad = p->n > p->nu;
    float2* a = pad ? p->pad_a : reinterpret_cast<float2*>(psi);
    float2* b = pad ? p->pad_b : reinterpret_cast<float2*>(lam);
    if (pad) {
      // the caller's state in the low 2^nu amplitudes, zeros above; lambda
      // zero above as well (H|psi> below writes only the low part)
      TQSV_HIP(k_pad_copy(reinterpret_cast<const float2*>(psi), 1ull << p->nu, a, 1ull << p->n, st));
      TQSV_HIP(k_pad_copy(nullptr, 0, b, 1ull << p->n, st));
    }
    rc = run_passes(p->passes, p->gates, p->n, p->dprog, a, nullptr, nullptr, nullptr, st);
    if (rc) return rc;
    // lambda = H psi and the block partials of E = Re <psi|H psi>
    TQSV_HIP(k_ham_apply(a, b, h->n, h->apply_args(), st, p->dpartials));
    // the reverse programs are built on the host while the device runs the
    // forward sweep and H|psi> (after a parameter update they are re-emitted)
    rc = plan_build_adjoint(p, st);
    if (rc) return rc;
    std::vector<double> hred((size_t)p->nred_adj + 2);
    rc = reverse_and_collect(p, a, b, nparams, grad, 1.0, hred.data(), st, true);
    if (rc) return rc;
    if (pad) {
      const size_t bytes = sizeof(float2) << p->nu;
      TQSV_HIP(hipMemcpyAsync(psi, a, bytes, hipMemcpyDeviceToDevice, st));
      TQSV_HIP(hipMemcpyAsync(lam, b, bytes, hipMemcpyDeviceToDevice, st));
    }
    if (energy) *energy = hred[(size_t)p->nred_adj];
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_adjoint: out of host memory");
  }
}

int tqsv_batch_create(tqsv_batch** out) {
  if (!out) return fail(TQSV_EINVAL, "tqsv_batch_create: NULL output");
  *out = new (std::nothrow) tqsv_batch();
  return *out ? 0 : fail(TQSV_ENOMEM, "tqsv_batch_create: out of host memory");
}

void tqsv_batch_destroy(tqsv_batch* b) { delete b; }

int tqsv_batch_adjoint(tqsv_batch* b, int nreq, tqsv_plan* const* plans, tqsv_ham* const* hams,
                       const void* const* psi, const int* nparams, double* const* grads,
                       double* energies, void* stream) {
  try {
    if (!b || nreq < 0 || (nreq > 0 && (!plans || !hams || !psi || !nparams || !grads || !energies)))
      return fail(TQSV_EINVAL, "tqsv_batch_adjoint: NULL argument");
    if (nreq == 0) return 0;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    bool lean = true;
    uint32_t max_nred = 0;
    size_t noff = 0, nout = 0;
    for (int i = 0; i < nreq; ++i) {
      tqsv_plan* p = plans[i];
      const tqsv_ham* h = hams[i];
      if (!p || !h || !psi[i] || (nparams[i] > 0 && !grads[i]))
        return fail(TQSV_EINVAL, "tqsv_batch_adjoint: NULL request field");
      if (p->n != BATCH_QUBITS || p->nu != h->n)
        return fail(TQSV_EINVAL, "tqsv_batch_adjoint: plans must be built for 13 qubits "
                                 "(min_qubits = 13) with a Hamiltonian of the caller's size");
      if (nparams[i] <= p->max_param)
        return fail(TQSV_EINVAL, "tqsv_batch_adjoint: nparams smaller than the largest parameter "
                                 "index + 1");
      int rc = plan_build_adjoint(p, st);
      if (rc) return rc;
      for (const auto* v : {&p->passes, &p->rpasses})
        for (const PassPlan& pp : *v) {
          if (!pp.tile || pp.tb != BATCH_QUBITS - RB)
            return fail(TQSV_EUNSUPPORTED, "tqsv_batch_adjoint: a plan has a pass that is not "
                                           "a whole-state tile pass");
          lean = lean && pp.lean;
          max_nred = std::max(max_nred, pp.nred);
        }
      noff += p->passes.size() + 2 * p->rpasses.size();
      nout += (size_t)p->nred_adj + 2;
    }
    int rc = grow(&b->doff, &b->doff_cap, noff);
    if (!rc) rc = grow(&b->ditems, &b->items_cap, (size_t)nreq);
    if (!rc) rc = grow(&b->dout, &b->out_cap, nout);
    if (rc) return rc;
    std::vector<uint32_t> off;
    off.reserve(noff);
    std::vector<BatchItem> items((size_t)nreq);
    std::vector<size_t> outpos((size_t)nreq);
    size_t opos = 0;
    for (int i = 0; i < nreq; ++i) {
      const tqsv_plan* p = plans[i];
      const tqsv_ham* h = hams[i];
      BatchItem& it = items[(size_t)i];
      it.psi = reinterpret_cast<const float2*>(psi[i]);
      it.nu = (uint32_t)p->nu;
      it.nfwd = (uint32_t)p->passes.size();
      it.nrev = (uint32_t)p->rpasses.size();
      it.fprog = p->dprog;
      it.rprog = p->drprog;
      it.foff = b->doff + off.size();
      for (const PassPlan& pp : p->passes) off.push_back((uint32_t)pp.prog_off);
      it.roff = b->doff + off.size();
      for (const PassPlan& pp : p->rpasses) off.push_back((uint32_t)pp.prog_off);
      it.rred = b->doff + off.size();
      for (const PassPlan& pp : p->rpasses) off.push_back(pp.red_base);
      it.ngroups = h->ngroups;
      it.gx = h->d_gx;
      it.gstart = h->d_gstart;
      it.tz = h->d_tz;
      it.tc = h->d_tc;
      it.out = b->dout + opos;
      it.nred = (uint32_t)p->nred_adj;
      outpos[(size_t)i] = opos;
      opos += (size_t)p->nred_adj + 2;
    }
    TQSV_HIP(hipMemcpyAsync(b->doff, off.data(), off.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, st));
    TQSV_HIP(hipMemcpyAsync(b->ditems, items.data(), items.size() * sizeof(BatchItem),
                            hipMemcpyHostToDevice, st));
    TQSV_HIP(k_batch_adjoint(b->ditems, (uint32_t)nreq, lean, max_nred, st));
    std::vector<double> hout(nout);
    TQSV_HIP(hipMemcpyAsync(hout.data(), b->dout, nout * sizeof(double), hipMemcpyDeviceToHost, st));
    TQSV_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nreq; ++i) {
      const double* hred = hout.data() + outpos[(size_t)i];
      collect_grad(plans[i], hred, nparams[i], grads[i], 1.0);
      energies[i] = hred[plans[i]->nred_adj];
    }
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_batch_adjoint: out of host memory");
  }
}

int tqsv_vjp(tqsv_plan* p, void* psi, void* lam, int nparams, double* grad, void* stream) {
  try {
    if (!p || !psi || !lam || (nparams > 0 && !grad)) return fail(TQSV_EINVAL, "tqsv_vjp: NULL argument");
    if (nparams <= p->max_param)
      return fail(TQSV_EINVAL, "tqsv_vjp: nparams smaller than the largest parameter index + 1");
    if (psi == lam) return fail(TQSV_EINVAL, "tqsv_vjp: lam must not alias psi");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int rc = plan_build_adjoint(p, st);
    if (rc) return rc;
    std::vector<double> hred((size_t)p->nred_adj + 2);
    if (p->n == p->nu)
      return reverse_and_collect(p, reinterpret_cast<float2*>(psi), reinterpret_cast<float2*>(lam),
                                 nparams, grad, 0.5, hred.data(), st, false);
    const uint64_t du = 1ull << p->nu, dn = 1ull << p->n;
    TQSV_HIP(k_pad_copy(reinterpret_cast<const float2*>(psi), du, p->pad_a, dn, st));
    TQSV_HIP(k_pad_copy(reinterpret_cast<const float2*>(lam), du, p->pad_b, dn, st));
    rc = reverse_and_collect(p, p->pad_a, p->pad_b, nparams, grad, 0.5, hred.data(), st, false);
    if (rc) return rc;
    TQSV_HIP(hipMemcpyAsync(psi, p->pad_a, du * sizeof(float2), hipMemcpyDeviceToDevice, st));
    TQSV_HIP(hipMemcpyAsync(lam, p->pad_b, du * sizeof(float2), hipMemcpyDeviceToDevice, st));
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_vjp: out of host memory");
  }
}

}  // extern "C"
