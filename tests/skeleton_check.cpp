// Checker for tests/test_skeleton_host.py: plans op files of one circuit
// structure (different parameters), emits their programs with the positions
// of the value words, and reports whether programs with equal skeletons
// (value words zeroed) always have equal generated pass sources -- the
// condition under which a plan update keeps its compiled kernels.
//   skeleton_check <n> <dual 0|1> <max_cost> <ops file>...
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>
#include "tqsv_planner.h"
using namespace tqsv;
static std::vector<char> rd(const char* f) { std::ifstream s(f, std::ios::binary); return std::vector<char>((std::istreambuf_iterator<char>(s)), {}); }
int main(int argc, char** argv) {
  int n = atoi(argv[1]); bool dual = atoi(argv[2]) != 0;
  SchedOpts so; so.n = n; so.tile_bits = 13; so.max_cost = atoi(argv[3]); so.dual = dual;
  std::vector<std::vector<uint32_t>> skels; std::vector<std::vector<std::string>> srcs;
  std::vector<PassPlan> sel0;
  std::vector<Gate> g0;
  for (int a = 4; a < argc; ++a) {
    auto buf = rd(argv[a]); int nops = buf.size() / sizeof(tqsv_op);
    std::vector<Gate> g; gates_from_ops(n, reinterpret_cast<const tqsv_op*>(buf.data()), nops, g); merge_and_classify(g);
    if (dual) { std::vector<Gate> r; reverse_gates(g, r); g.swap(r); }
    if (!g0.empty() && g.size() == g0.size()) for (size_t i = 0; i < g.size(); ++i) g[i].rot_form = g0[i].rot_form;  // kept across updates
    std::vector<PassPlan> sel; if (sel0.empty()) { select_passes(g, so, sel); sel0 = sel; } else sel = sel0;
    std::vector<PassPlan> passes; std::vector<uint32_t> prog, fpos; uint32_t red = 0;
    if (emit_passes(g, so, sel, passes, prog, red, &fpos)) { printf("emit failed\n"); return 1; }
    g0 = g;
    std::vector<uint32_t> sk = prog; for (uint32_t i : fpos) sk[i] = 0;
    std::vector<std::string> ss; for (auto& p : passes) if (p.tile) ss.push_back(jit_source(p.tb, prog.data() + p.prog_off, n));
    skels.push_back(sk); srcs.push_back(ss);
  }
  int bad = 0, same = 0;
  for (size_t a = 1; a < skels.size(); ++a) {
    const bool eqk = skels[a] == skels[0], eqs = srcs[a] == srcs[0];
    same += eqk;
    if (eqk && !eqs) { ++bad; printf("file %zu: equal skeleton, different source\n", a); }
    printf("file %zu: skeleton %s, sources %s\n", a, eqk ? "equal" : "differ", eqs ? "equal" : "differ");
  }
  printf("%s\n", bad ? "FAIL" : "ok");
  return bad;
}
