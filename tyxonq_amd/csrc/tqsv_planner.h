// tqsv_planner.h - host-side gate model and pass/phase scheduler.
#ifndef TQSV_PLANNER_H
#define TQSV_PLANNER_H

#include <array>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tqsv.h"
#include "tqsv_internal.h"
#include "tqsv_prog.h"

namespace tqsv {

using cd = std::complex<double>;

enum GateKind : int {
  GK_MAT1 = 1,   // 2x2 on t0, controls (cmask, cval)
  GK_MAT2 = 2,   // 4x4 on (t0 msb, t1), controls
  GK_DIAG2 = 3,  // diagonal d[2*bit(t0)+bit(t1)] (t0 == t1 for one bit)
  GK_PROT = 4,   // exp(-i theta/2 P)
  GK_RED = 5,    // cross reduction <lambda| P |psi> (adjoint sweeps)
  GK_PGRP = 6,   // commuting Pauli rotations sharing one flip mask (members below)
};

constexpr int kMaxGroup = 16;  // members of a GK_PGRP gate

struct Gate {
  int kind = 0;
  int t0 = -1, t1 = -1;
  uint64_t cmask = 0, cval = 0;
  uint64_t x = 0, z = 0;
  int k = 0;          // |x & z| mod 4
  double c = 1.0, s = 0.0;
  cd m[16];
  // bit usage for commutation: nb = used non-diagonally, db = used only diagonally
  uint64_t nb = 0, db = 0;
  int cost = 0;
  // adjoint bookkeeping
  int param = -1;
  uint64_t gx = 0, gz = 0;
  double gcoef = 0.0;
  int red_slot = -1;
  // which part of <lambda|P|psi> the slot holds: -1 two slots (re, im), 0 one
  // slot with the real part, 1 one slot with the imaginary part
  int red_part = -1;
  double red_scale = 1.0;  // true <lambda|P|psi> = red_scale * (value read from the sweep)
  // rotation form (tan 0 / cot 1) chosen when the gate was last emitted; for
  // plans on generated kernels it is kept across parameter updates while
  // |t| <= 4, so that the compiled code does not change when an angle moves
  // across |t| = 1 (other plans choose by |t| <= 1 alone, so that results do
  // not depend on earlier parameters)
  int rot_form = -1;
  int src = -1;
  // GK_PGRP: prod_r exp(-i ma[r]/2 P_r), P_r = (x, mz[r]); the members commute
  // and their z masks differ only inside x.  mp[r], mg[r]: the member's
  // parameter (-1: none) and generator coefficient (angle = mg[r] * param +
  // const) for the adjoint.  `param` is the largest member parameter.
  int nmem = 0;
  uint64_t mz[kMaxGroup] = {};
  double ma[kMaxGroup] = {};
  double mg[kMaxGroup] = {};
  int mp[kMaxGroup] = {};
};

// Class constants of a GK_PGRP gate for its pairs (i, i ^ x), one class per
// pattern of the x bits above the lowest one (compressed, ascending bits).
struct PgrpClass {
  bool active = false;  // false: the identity on these pairs, whatever the angles
  double cm1 = 0.0;     // cos(w/2) - 1
  cd ui{0, 0}, uj{0, 0};
};
// Returns the number of classes (2^(|x|-1)).  constants = false: only the
// `active` flags (no trigonometry; for cost estimates).
int pgrp_classes(const Gate& g, PgrpClass* out, bool constants = true);

// Convert ABI ops into gates (validates).  Returns 0 or an error code.
int gates_from_ops(int n, const tqsv_op* ops, int nops, std::vector<Gate>& out);
// Merge runs of uncontrolled single-bit gates (param < 0) and classify.
void merge_and_classify(std::vector<Gate>& g);
// Inverse gate (same usage).
Gate dagger(const Gate& g);

struct PassPlan {
  bool tile = false;     // tile pass; else a single-op launch of `single`
  int tb = 0;
  bool dual = false;
  int mode = 0;          // kernel mode (MODE_GATE / MODE_DUAL)
  uint32_t L = 0, m = 0;
  uint32_t H[MAXH] = {};
  std::vector<std::vector<int>> phases;   // gate indices per phase
  std::vector<std::array<uint32_t, RB>> Rpos;  // register positions per phase, ascending
  int single = -1;
  uint32_t nred = 0;      // reduction slots of this pass
  uint32_t red_base = 0;  // first slot in the plan's reduction output
  size_t part_off = 0;    // this pass's block partials in a sweep-wide partials buffer
  // emitted program
  size_t prog_off = 0;   // words
  size_t prog_words = 0;
  uint32_t ntiles = 0, grid = 0;
  size_t lds = 0;
  void* jit_fn = nullptr;  // generated kernel (hipFunction_t), when the plan uses them
  bool lean = false;       // no OP_U2 / OP_PROT steps: the lean precompiled variant runs it
};

// Default per-pass budget (approximate VALU ops per amplitude).  Passes are
// bound by VALU issue as much as by HBM; measured on 18-26 qubit RY/RZ/CX
// adjoint plans, budgets from 96 up to about 400 cut the time (fewer passes,
// less per-pass overhead) and larger ones change nothing.
constexpr int kDefaultMaxCost = 400;
// States of 2^28+ amplitudes stream from HBM in generated kernels bound by VALU
// issue; on the 30-qubit brickwork a budget of 260 (21 passes) ran the gate part
// in 96.4 ms against 98.8 ms at 400 (also 21 passes) and 100.5 ms at 200.
constexpr int kLargeStateMaxCost = 260;
inline int default_max_cost(int n) { return n >= 28 ? kLargeStateMaxCost : kDefaultMaxCost; }

// Smallest state that runs as tile passes (2^11 amplitudes: 64 threads x 32).
constexpr int kPadQubits = 11;
inline int plan_qubits(int n, bool fuse, int min_qubits) {
  const int target = min_qubits > 0 ? min_qubits : kPadQubits;
  return fuse && n < target ? target : n;
}

struct SchedOpts {
  int n = 0;
  int tile_bits = 13;
  int min_low = 5;
  int max_cost = 0;  // 0: default_max_cost(n)
  bool fuse = true;
  bool dual = false;
  bool tile_search = false;  // choose tile bits by admitted work (else in gate order)
};

// Classify one gate (usage bits, diagonal forms, cost).
void classify_gate(Gate& g);

// Plan passes over `gates` (in order).  Appends the programs to `prog`.
// Reduction gates get their output slots (gates[i].red_slot: one double, the
// part of <lambda|P|psi> the gradient needs, in tile passes; two doubles, re
// and im, in single-op passes) allocated from `red_total`.
int schedule(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan>& passes,
             std::vector<uint32_t>& prog, uint32_t& red_total);

// The two halves of schedule(): pass/phase selection (which gates go to which
// pass and phase, tile and register bits) and emission of the programs.  A plan
// whose gates keep their structure (same_structure) when only parameters
// change reuses its selection and re-emits.
int select_passes(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan>& plan);
// fpos (optional) receives the positions in prog of the value words (angles,
// phases, scales: everything written as a float); the other words fix the
// program's structure and the generated source of each pass.
int emit_passes(std::vector<Gate>& gates, const SchedOpts& so, std::vector<PassPlan> plan,
                std::vector<PassPlan>& passes, std::vector<uint32_t>& prog, uint32_t& red_total,
                std::vector<uint32_t>* fpos = nullptr);
bool same_structure(const std::vector<Gate>& a, const std::vector<Gate>& b);

// Run planned passes (device program at dprog; partial-sum scratch for reductions).
// Reductions: without `dslots` each pass's block partials go to the start of
// `dpartials` and are summed into its slots right after the pass; with
// `dslots` (nslots entries of 4 words: partials offset, blocks, stride, index)
// each pass writes at its part_off and one kernel after the sweep sums every
// slot.
int run_passes(const std::vector<PassPlan>& passes, const std::vector<Gate>& gates, int n,
               const uint32_t* dprog, float2* psi, float2* lam, double* dpartials,
               double* dred_out, hipStream_t st, const uint32_t* dslots = nullptr,
               uint32_t nslots = 0);

int effective_tile_bits(int n, int requested, bool dual);

// ---- generated pass kernels (tqsv_jit.cpp) ----
// 1 if a plan of n qubits with the given number of tile passes uses
// generated kernels (requested: 1 on, -1 off, 0 automatic: n >= 22 and at most
// max_auto_passes passes; the environment variable TQSV_JIT=0/1 overrides).
int jit_policy(int requested, int n, size_t tile_passes, size_t max_auto_passes);
// Compile (or take from the process-wide cache) the kernel of every tile pass
// and set PassPlan::jit_fn.
int jit_prepare(std::vector<PassPlan>& passes, const std::vector<uint32_t>& prog, int n);
hipError_t jit_launch(const PassPlan& pp, const PassLaunch& pl, hipStream_t st);
// Compile the kernel of every tile pass without loading it (no device is
// used; the compiler's own disk cache keeps the code objects for later
// processes).  *compiled = distinct kernels compiled.
int jit_compile_only(const std::vector<PassPlan>& passes, const std::vector<uint32_t>& prog, int n,
                     int* compiled);
// generated source of one pass record (for inspection and tests)
std::string jit_source(int tb, const uint32_t* pass_words, int n);

// The reverse sweep of a forward gate list (tqsv_ham.cpp): per gate in
// reverse order, the gradient reduction(s) of its parameter(s), then its
// inverse.
void reverse_gates(const std::vector<Gate>& gates, std::vector<Gate>& r);

}  // namespace tqsv

struct tqsv_plan {
  int n = 0;   // qubits the plan is built for (at least kPadQubits when fusing)
  int nu = 0;  // qubits of the caller's states (n > nu: states are padded, see pad_a)
  int nops = 0;
  int jit_req = 0;  // requested: 1 generated kernels, -1 precompiled, 0 automatic
  bool fgen = false, rgen = false;  // the last build ran generated kernels (forward, reverse)
  tqsv::SchedOpts so;
  std::vector<tqsv::Gate> gates;          // forward gates after merge
  std::vector<tqsv::PassPlan> fsel;       // pass/phase selection of the forward gates
  std::vector<tqsv::PassPlan> passes;
  std::vector<uint32_t> prog;
  uint32_t* dprog = nullptr;
  size_t dprog_cap = 0;
  // adjoint (built lazily)
  bool adj_built = false;
  std::vector<tqsv::Gate> rgates;         // reverse sweep items
  std::vector<tqsv::PassPlan> rsel;
  std::vector<tqsv::PassPlan> rpasses;
  std::vector<uint32_t> rprog;
  uint32_t* drprog = nullptr;
  size_t drprog_cap = 0;
  int nred_adj = 0;
  double* dpartials = nullptr;
  size_t dpartials_cap = 0;
  double* dred = nullptr;
  size_t dred_cap = 0;
  // per reduction slot of the reverse sweep: {partials offset, blocks, stride,
  // index} (one final sum for the whole sweep), then two entries for the
  // energy of tqsv_adjoint, whose block partials (from H|psi>) lead the
  // partials buffer
  std::vector<uint32_t> hslots;
  uint32_t* dslots = nullptr;
  size_t dslots_cap = 0;
  int max_param = -1;
  // Scratch of the rebuilds after a parameter update, swapped with the live
  // vectors above: their capacity is reused, so an update of a large circuit
  // (thousands of gates) touches no fresh memory pages.
  std::vector<tqsv::Gate> gwork, rwork;
  std::vector<tqsv::PassPlan> pwork, rpwork;
  std::vector<uint32_t> progwork, rprogwork;
  // programs with their value words zeroed, as last compiled (generated
  // kernels): an update whose skeleton is unchanged keeps the kernels
  std::vector<uint32_t> fskel, rskel, skelwork, fposwork;
  // States of fewer than kPadQubits qubits run as kPadQubits-qubit states whose
  // extra high index bits stay zero (no gate touches them), so that they use
  // the tile passes instead of one launch per gate: the caller's 2^nu
  // amplitudes are copied into pad_a (pad_b for lambda), the rest is zero.
  float2* pad_a = nullptr;
  float2* pad_b = nullptr;
};

#endif
