// tqsv_emulate.cpp - host emulation of planned circuits.
//
// Runs a plan's pass programs on a host complex64 state with the same index
// arithmetic as the device kernel (tile base deposit, position map, LDS
// swizzle, register-phase geometry) and the same per-step semantics, one
// emulated thread after another.  Threads of one phase touch disjoint
// amplitudes, so the sequential emulation computes exactly what the device
// does up to float rounding order.  This lets the planner and the program
// format be tested on a machine without a GPU.
#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

#include "tqsv_planner.h"

namespace tqsv {

namespace {

using cf = std::complex<float>;

inline uint32_t swz_h(uint32_t p) { return p ^ ((p >> 4) & 30u); }
inline uint32_t par64(uint64_t v) { return (uint32_t)__builtin_popcountll(v) & 1u; }
inline float fw(const uint32_t* w) {
  float f;
  std::memcpy(&f, w, 4);
  return f;
}
inline uint64_t w64(const uint32_t* w) { return (uint64_t)w[0] | ((uint64_t)w[1] << 32); }

struct Geo {
  uint32_t L, m;
  const uint32_t* H;
  uint64_t pos_map(uint32_t p) const {
    uint64_t v = p & ((1u << L) - 1u);
    for (uint32_t k = 0; k < m; ++k)
      if ((p >> (L + k)) & 1u) v |= 1ull << H[k];
    return v;
  }
  uint64_t tile_base(uint64_t tile) const {
    uint64_t v = tile << L;
    for (uint32_t k = 0; k < m; ++k) {
      const uint32_t h = H[k];
      v = ((v >> h) << (h + 1)) | (v & ((1ull << h) - 1ull));
    }
    return v;
  }
};

void mul2(const uint32_t* w, cf x0, cf x1, cf& y0, cf& y1) {
  const cf u00(fw(w + 0), fw(w + 1)), u01(fw(w + 2), fw(w + 3));
  const cf u10(fw(w + 4), fw(w + 5)), u11(fw(w + 6), fw(w + 7));
  y0 = u00 * x0 + u01 * x1;
  y1 = u10 * x0 + u11 * x1;
}

cf iq(cf v, uint32_t q) {
  switch (q & 3u) {
    case 0: return v;
    case 1: return cf(-v.imag(), v.real());
    case 2: return -v;
    default: return cf(v.imag(), -v.real());
  }
}

// one emulated thread of one phase: applies the phase's steps to a[] (and b[])
void run_steps(const uint32_t* w, uint32_t nsteps, cf* a, cf* b, bool dual, uint64_t gb,
               const uint64_t* gR, std::vector<double>* red) {
  auto gs = [&](int s) {
    uint64_t v = 0;
    for (int j = 0; j < RB; ++j)
      if (s & (1 << j)) v ^= gR[j];
    return v;
  };
  for (uint32_t is = 0; is < nsteps; ++is) {
    const uint32_t op = w[0] & 0xffffu, nw = w[0] >> 16;
    cf* arrs[2] = {a, b};
    const int narr = dual ? 2 : 1;
    switch (op) {
      case OP_U1: {
        const int K = (int)w[1];
        const bool ctrl = w[3] != 0;
        const uint64_t cm = w64(w + 4), cv = w64(w + 6);
        for (int z = 0; z < narr; ++z) {
          cf* x = arrs[z];
          for (int s = 0; s < NSLOT; ++s) {
            if (s & (1 << K)) continue;
            if (ctrl && (((gb ^ gs(s)) & cm) != cv)) continue;
            cf y0, y1;
            mul2(w + 8, x[s], x[s | (1 << K)], y0, y1);
            x[s] = y0;
            x[s | (1 << K)] = y1;
          }
        }
        break;
      }
      case OP_U2: {
        const int KH = (int)w[1], KL = (int)w[2];
        const bool ctrl = w[3] != 0;
        const uint64_t cm = w64(w + 4), cv = w64(w + 6);
        for (int z = 0; z < narr; ++z) {
          cf* x = arrs[z];
          for (int s = 0; s < NSLOT; ++s) {
            if (s & ((1 << KH) | (1 << KL))) continue;
            if (ctrl && (((gb ^ gs(s)) & cm) != cv)) continue;
            const int idx[4] = {s, s | (1 << KL), s | (1 << KH), s | (1 << KH) | (1 << KL)};
            cf xin[4], y[4];
            for (int c = 0; c < 4; ++c) xin[c] = x[idx[c]];
            for (int r = 0; r < 4; ++r) {
              y[r] = 0;
              for (int c = 0; c < 4; ++c)
                y[r] += cf(fw(w + 8 + 8 * r + 2 * c), fw(w + 9 + 8 * r + 2 * c)) * xin[c];
            }
            for (int c = 0; c < 4; ++c) x[idx[c]] = y[c];
          }
        }
        break;
      }
      case OP_ROT: {
        const int K = (int)w[1];
        const uint32_t md = w[2];
        float t = fw(w + 3);
        const uint64_t zs = w64(w + 4);
        if (par64(gb & zs)) t = -t;
        const cf I(0.f, 1.f);
        for (int z = 0; z < narr; ++z) {
          cf* x = arrs[z];
          for (int s = 0; s < NSLOT; ++s) {
            if (s & (1 << K)) continue;
            const int s1 = s | (1 << K);
            const cf x0 = x[s], x1 = x[s1];
            switch (md) {
              case 0: x[s] = x0 - t * x1; x[s1] = x1 + t * x0; break;
              case 1: x[s] = t * x0 - x1; x[s1] = t * x1 + x0; break;
              case 2: x[s] = x0 - I * t * x1; x[s1] = x1 - I * t * x0; break;
              default: x[s] = t * x0 - I * x1; x[s1] = t * x1 - I * x0; break;
            }
          }
        }
        break;
      }
      case OP_PH: {
        const int K = (int)w[1];
        cf u(fw(w + 4), fw(w + 5));
        const uint32_t* e = w + OPH_HDR;
        for (uint32_t i = 0; i < w[2]; ++i, e += OPH_ENT)
          if ((gb >> e[0]) & 1u) u *= cf(fw(e + 1), fw(e + 2));
        for (int z = 0; z < narr; ++z)
          for (int s = 0; s < NSLOT; ++s)
            if (s & (1 << K)) arrs[z][s] *= u;
        break;
      }
      case OP_SGN: {
        for (int z = 0; z < narr; ++z)
          for (int s = 0; s < NSLOT; ++s)
            if ((w[1] >> s) & 1u) arrs[z][s] = -arrs[z][s];
        break;
      }
      case OP_TAB: {
        for (int z = 0; z < narr; ++z)
          for (int s = 0; s < NSLOT; ++s)
            arrs[z][s] *= cf(fw(w + TAB_OFF + 2 * s), fw(w + TAB_OFF + 2 * s + 1));
        break;
      }
      case OP_CXP: {
        // the encoding the interpreting kernel reads: the register part of
        // the control as a slot mask, a thread-bit control by the global masks
        const int K = (int)w[1];
        const uint64_t cm = w64(w + 4), cv = w64(w + 6);
        const bool ton = w[2] == CX_THREAD ? (gb & cm) == cv : true;
        for (int z = 0; z < narr; ++z) {
          cf* x = arrs[z];
          for (int s = 0; s < NSLOT; ++s) {
            if (s & (1 << K)) continue;
            if (((w[8] >> s) & 1u) && ton) std::swap(x[s], x[s | (1 << K)]);
          }
        }
        break;
      }
      case OP_SCL: {
        const float f = fw(w + 1);
        for (int z = 0; z < narr; ++z)
          for (int s = 0; s < NSLOT; ++s) arrs[z][s] *= f;
        break;
      }
      case OP_THR: {
        cf F(fw(w + 2), fw(w + 3));
        const uint32_t* e = w + THR_HDR;
        for (uint32_t i = 0; i < w[1]; ++i, e += THR_ENT)
          if (((gb >> e[0]) & (gb >> e[1])) & 1u) F *= cf(fw(e + 2), fw(e + 3));
        for (int z = 0; z < narr; ++z)
          for (int s = 0; s < NSLOT; ++s) arrs[z][s] *= F;
        break;
      }
      case OP_D2: {
        for (int s = 0; s < NSLOT; ++s) {
          const uint64_t g = gb ^ gs(s);
          const uint32_t sel = 2u * (uint32_t)((g >> w[1]) & 1u) + (uint32_t)((g >> w[2]) & 1u);
          const cf d(fw(w + 4 + 2 * sel), fw(w + 5 + 2 * sel));
          for (int z = 0; z < narr; ++z) arrs[z][s] *= d;
        }
        break;
      }
      case OP_PROT: {
        const uint32_t xr = w[1];
        const uint64_t zm = w64(w + 2);
        const float c = fw(w + 4), sn = fw(w + 5);
        const uint32_t k = w[6] & 3u, q = (k + 3u) & 3u, xzpar = k & 1u;
        for (int z = 0; z < narr; ++z) {
          cf* x = arrs[z];
          if (xr == 0) {
            for (int s = 0; s < NSLOT; ++s) {
              const float ss = par64((gb ^ gs(s)) & zm) ? -sn : sn;
              x[s] *= cf(c, -ss);
            }
          } else {
            const uint32_t lb = xr & (~xr + 1u);
            for (int s = 0; s < NSLOT; ++s) {
              if (s & lb) continue;
              const int s1 = s ^ (int)xr;
              const uint32_t sg0 = par64((gb ^ gs(s)) & zm), sg1 = sg0 ^ xzpar;
              const float s0f = sg0 ? -sn : sn, s1f = sg1 ? -sn : sn;
              const cf x0 = x[s], x1 = x[s1];
              x[s] = c * x0 + s1f * iq(x1, q);
              x[s1] = c * x1 + s0f * iq(x0, q);
            }
          }
        }
        break;
      }
      case OP_PGRP: {
        const uint32_t xr = w[1], kpar = w[4], act = w[5], sm = w[6];
        const uint64_t z0 = w64(w + 2);
        const uint32_t lb = xr & (~xr + 1u), hi = xr & ~lb;
        for (int z = 0; z < narr; ++z) {
          cf* x = arrs[z];
          for (int s = 0; s < NSLOT; ++s) {
            if (s & lb) continue;
            uint32_t cls = 0, k = 0;
            for (int bb = 0; bb < RB; ++bb)
              if ((hi >> bb) & 1u) {
                if ((s >> bb) & 1) cls |= 1u << k;
                ++k;
              }
            if (!((act >> cls) & 1u)) continue;
            const uint32_t* e = w + PG_HDR + PG_ENT * cls;
            const int s1 = s ^ (int)xr;
            const uint32_t sgi = par64(gb & z0) ^ ((sm >> s) & 1u), sgj = sgi ^ kpar;
            const float cm1 = fw(e);
            const cf ui = cf(fw(e + 1), fw(e + 2)) * (sgi ? -1.f : 1.f);
            const cf uj = cf(fw(e + 3), fw(e + 4)) * (sgj ? -1.f : 1.f);
            const cf x0 = x[s], x1 = x[s1];
            x[s] = x0 + (cm1 * x0 + uj * x1);
            x[s1] = x1 + (cm1 * x1 + ui * x0);
          }
        }
        break;
      }
      case OP_RED: {
        const uint32_t xr = w[1], kind = w[2], nterms = w[3];
        const uint32_t lb = xr & (~xr + 1u);
        if (kind == RED_GROUP) {
          const uint32_t* e = w + RED_HDR;
          const uint64_t z0 = w64(e + 1);
          const uint32_t kpar = e[RED_ENT], act = e[RED_ENT + 1];
          const uint32_t* tab = e + RED_ENT + 2;
          const uint32_t hi = xr & ~lb;
          double acc = 0.0;
          for (int s = 0; s < NSLOT; ++s) {
            if (s & lb) continue;
            uint32_t cls = 0, k = 0;
            for (int bb = 0; bb < RB; ++bb)
              if ((hi >> bb) & 1u) {
                if ((s >> bb) & 1) cls |= 1u << k;
                ++k;
              }
            if (!((act >> cls) & 1u)) continue;
            const uint32_t* c = tab + RGRP_ENT * cls;
            const int s1 = s ^ (int)xr;
            const uint32_t sgi = par64(gb & z0) ^ ((e[3] >> s) & 1u), sgj = sgi ^ kpar;
            const cf wi = cf(fw(c), fw(c + 1)) * (sgi ? -1.f : 1.f);
            const cf wj = cf(fw(c + 2), fw(c + 3)) * (sgj ? -1.f : 1.f);
            acc += (std::conj(b[s]) * wj * a[s1] + std::conj(b[s1]) * wi * a[s]).imag();
          }
          if (red) (*red)[e[0]] += acc;
          break;
        }
        const uint32_t* e = w + RED_HDR;
        for (uint32_t t = 0; t < nterms; ++t, e += RED_ENT) {
          double acc = 0.0;
          for (int s = 0; s < NSLOT; ++s) {
            cf p;
            if (kind == RED_PAIR) {
              if (s & lb) continue;
              p = std::conj(a[s ^ xr]) * a[s];
            } else if (kind == RED_DIAG) {
              p = cf(std::norm(a[s]), 0.f);
            } else {
              p = std::conj(b[s]) * a[s ^ xr];
            }
            const float v = e[4] ? p.imag() : p.real();
            acc += ((e[3] >> s) & 1u) ? -v : v;
          }
          if (par64(gb & w64(e + 1))) acc = -acc;
          if (red) (*red)[e[0]] += acc;
        }
        break;
      }
      default: break;
    }
    w += nw;
  }
}

// emulate one tile pass; red (pass-local slots) may be null
void emulate_pass(const PassPlan& pp, const uint32_t* prog, int n, cf* psi, cf* lam,
                  std::vector<double>* red) {
  const int TB = pp.tb, T = TB + RB;
  const uint32_t TSIZE = 1u << T, NT = 1u << TB;
  Geo geo{prog[PH_L], prog[PH_M], prog + PH_H};
  const uint32_t flags = prog[PH_FLAGS], nphase = prog[PH_NPHASE];
  const bool dual = (flags & PF_DUAL) != 0;
  std::vector<cf> tA(TSIZE), tB(dual ? TSIZE : 0);
  const uint64_t ntiles = 1ull << (n - T);
  for (uint64_t tile = 0; tile < ntiles; ++tile) {
    const uint64_t base = geo.tile_base(tile);
    for (uint32_t p = 0; p < TSIZE; ++p) {
      const uint64_t g = base | geo.pos_map(p);
      tA[swz_h(p)] = psi[g];
      if (dual) tB[swz_h(p)] = lam[g];
    }
    const uint32_t* ph = prog + PH_SIZE;
    for (uint32_t iph = 0; iph < nphase; ++iph) {
      const uint32_t nsteps = ph[PS_NSTEPS];
      uint64_t gR[RB];
      uint32_t sph[RB];
      for (int j = 0; j < RB; ++j) {
        sph[j] = swz_h(1u << ph[PS_R + j]);
        gR[j] = geo.pos_map(1u << ph[PS_R + j]);
      }
      const uint32_t* wend = nullptr;
      for (uint32_t tid = 0; tid < NT; ++tid) {
        uint32_t tpos = tid;
        for (int j = 0; j < RB; ++j) {
          const uint32_t r = ph[PS_R + j];
          tpos = ((tpos >> r) << (r + 1)) | (tpos & ((1u << r) - 1u));
        }
        const uint32_t bt = swz_h(tpos);
        const uint64_t gb = base | geo.pos_map(tpos);
        cf a[NSLOT], b[NSLOT];
        uint32_t l[NSLOT];
        for (int s = 0; s < NSLOT; ++s) {
          uint32_t v = bt;
          for (int j = 0; j < RB; ++j)
            if (s & (1 << j)) v ^= sph[j];
          l[s] = v;
          a[s] = tA[v];
          if (dual) b[s] = tB[v];
        }
        run_steps(ph + PS_SIZE, nsteps, a, b, dual, gb, gR, red);
        if (flags & PF_WRITE)
          for (int s = 0; s < NSLOT; ++s) {
            tA[l[s]] = a[s];
            if (dual) tB[l[s]] = b[s];
          }
        if (!wend) {
          const uint32_t* w = ph + PS_SIZE;
          for (uint32_t is = 0; is < nsteps; ++is) w += w[0] >> 16;
          wend = w;
        }
      }
      ph = wend ? wend : ph + PS_SIZE;
    }
    if (flags & PF_WRITE)
      for (uint32_t p = 0; p < TSIZE; ++p) {
        const uint64_t g = base | geo.pos_map(p);
        psi[g] = tA[swz_h(p)];
        if (dual) lam[g] = tB[swz_h(p)];
      }
  }
}

void apply_single_host(const Gate& g, int n, cf* psi) {
  const uint64_t dim = 1ull << n;
  switch (g.kind) {
    case GK_MAT1:
      for (uint64_t i = 0; i < dim; ++i) {
        if ((i >> g.t0) & 1u) continue;
        if ((i & g.cmask) != g.cval) continue;
        const uint64_t j = i | (1ull << g.t0);
        const cf x0 = psi[i], x1 = psi[j];
        psi[i] = cf(g.m[0]) * x0 + cf(g.m[1]) * x1;
        psi[j] = cf(g.m[2]) * x0 + cf(g.m[3]) * x1;
      }
      break;
    case GK_MAT2:
      for (uint64_t i = 0; i < dim; ++i) {
        if (((i >> g.t0) & 1u) || ((i >> g.t1) & 1u) || (i & g.cmask) != g.cval) continue;
        const uint64_t idx[4] = {i, i | (1ull << g.t1), i | (1ull << g.t0),
                                 i | (1ull << g.t0) | (1ull << g.t1)};
        cf x[4];
        for (int c = 0; c < 4; ++c) x[c] = psi[idx[c]];
        for (int r = 0; r < 4; ++r) {
          cf y = 0;
          for (int c = 0; c < 4; ++c) y += cf(g.m[4 * r + c]) * x[c];
          psi[idx[r]] = y;
        }
      }
      break;
    case GK_DIAG2:
      for (uint64_t i = 0; i < dim; ++i) {
        const int s = 2 * (int)((i >> g.t0) & 1u) + (int)((i >> g.t1) & 1u);
        psi[i] *= cf(g.m[s]);
      }
      break;
    case GK_PGRP:
      for (int r = 0; r < g.nmem; ++r) {
        Gate m = g;
        m.kind = GK_PROT;
        m.z = g.mz[r];
        m.k = __builtin_popcountll(g.x & g.mz[r]) & 3;
        m.c = std::cos(0.5 * g.ma[r]);
        m.s = std::sin(0.5 * g.ma[r]);
        apply_single_host(m, n, psi);
      }
      break;
    case GK_PROT: {
      std::vector<cf> out(dim);
      const cf om = std::pow(cf(0, 1), g.k);
      for (uint64_t i = 0; i < dim; ++i) {
        const uint64_t j = i ^ g.x;  // (P psi)[i] = om (-1)^{|j&z|} psi[j]
        const float sg = par64(j & g.z) ? -1.f : 1.f;
        out[i] = (float)g.c * psi[i] - cf(0, (float)g.s) * om * sg * psi[j];
      }
      std::memcpy(psi, out.data(), dim * sizeof(cf));
      break;
    }
    default: break;
  }
}

}  // namespace
}  // namespace tqsv

using namespace tqsv;

extern "C" int tqsv_plan_emulate(int n, const tqsv_op* ops, int nops, const tqsv_plan_opts* opts,
                                 float* state) {
  try {
    if (!state) return fail(TQSV_EINVAL, "tqsv_plan_emulate: state is NULL");
    if (n < 1 || n > 26) return fail(TQSV_EINVAL, "tqsv_plan_emulate: n must be in [1, 26]");
    tqsv_plan p;
    p.nu = n;
    p.nops = nops;
    if (opts) {
      p.so.tile_bits = opts->tile_bits > 0 ? opts->tile_bits : 13;
      p.so.min_low = opts->min_low_bits > 0 ? opts->min_low_bits : 5;
      p.so.max_cost = opts->max_cost > 0 ? opts->max_cost : 0;
      p.so.fuse = opts->fuse != 0;
    }
    p.n = plan_qubits(n, p.so.fuse, opts ? opts->min_qubits : 0);
    p.so.n = p.n;
    std::vector<Gate> g;
    int rc = gates_from_ops(n, ops, nops, g);
    if (rc) return rc;
    merge_and_classify(g);
    uint32_t red_total = 0;
    rc = schedule(g, p.so, p.passes, p.prog, red_total);
    if (rc) return rc;
    // padded like tqsv_plan_run: the caller's amplitudes low, zeros above
    std::vector<cf> padded((size_t)1 << p.n, cf(0.f, 0.f));
    cf* user = reinterpret_cast<cf*>(state);
    std::copy(user, user + ((size_t)1 << n), padded.begin());
    cf* psi = padded.data();
    for (const PassPlan& pp : p.passes) {
      if (pp.tile)
        emulate_pass(pp, p.prog.data() + pp.prog_off, p.n, psi, nullptr, nullptr);
      else
        apply_single_host(g[pp.single], p.n, psi);
    }
    for (size_t i = ((size_t)1 << n); i < padded.size(); ++i)
      if (padded[i] != cf(0.f, 0.f))
        return fail(TQSV_EINVAL, "tqsv_plan_emulate: padding amplitudes changed (internal error)");
    std::copy(padded.begin(), padded.begin() + ((size_t)1 << n), user);
    return 0;
  } catch (const std::bad_alloc&) {
    return fail(TQSV_ENOMEM, "tqsv_plan_emulate: out of host memory");
  }
}
