// pr: pull PageRank app with the reference's command line (src/pr.cc:106-144).
// PageRankPull keeps the GAPBS signature (gapbs-verify/pr.cc:34-35); its body
// is one call into the GPU engine.  DMM_PR_MODE=serial|kahan|fp64 picks the
// arithmetic (default: serial, the reference's FP32 statement order).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

#include "harness.h"

typedef float ScoreT;       // src/pr.cc:38
const float kDamp = 0.85f;  // src/pr.cc:39

static int pr_mode() {
  const char* m = std::getenv("DMM_PR_MODE");
  if (m && std::strcmp(m, "kahan") == 0) return DMM_PR_FP32_KAHAN;
  if (m && std::strcmp(m, "fp64") == 0) return DMM_PR_FP64;
  return DMM_PR_FP32_SERIAL;
}

std::vector<ScoreT> PageRankPull(const Graph& g, int max_iters, double epsilon = 0) {
  const int mode = pr_mode();
  std::vector<ScoreT> scores(g.num_nodes());
  dmm_pr_stats st;
  int rc;
  if (mode == DMM_PR_FP64) {
    std::vector<double> s64(g.num_nodes());
    rc = dmm_pr_pull(g.ctx(), max_iters, epsilon, mode, s64.data(), &st);
    for (size_t i = 0; i < s64.size(); i++) scores[i] = static_cast<ScoreT>(s64[i]);
  } else {
    rc = dmm_pr_pull(g.ctx(), max_iters, epsilon, mode, scores.data(), &st);
  }
  if (rc != DMM_OK) {
    std::cerr << "dmm_pr_pull: " << dmm_last_error(g.ctx()) << std::endl;
    std::exit(1);
  }
  for (int i = 0; i < st.iters && i < DMM_MAX_PR_ITERS; i++) printf(" %2d    %lf\n", i, st.err[i]);
  return scores;
}

// src/pr.cc:73-83 (top 5 scores).
void PrintTopScores(const Graph& g, const std::vector<ScoreT>& scores) {
  std::vector<std::pair<ScoreT, NodeID>> top;
  for (NodeID n = 0; n < g.num_nodes(); n++) top.push_back({scores[n], n});
  const size_t k = std::min<size_t>(5, top.size());
  std::partial_sort(top.begin(), top.begin() + k, top.end(), std::greater<std::pair<ScoreT, NodeID>>());
  for (size_t i = 0; i < k; i++) std::cout << top[i].second << ":" << top[i].first << std::endl;
}

// Verification: the push-direction residual check of gapbs-verify/pr.cc:
// 77-93 (one serial push iteration from the scores must change them by
// less than the tolerance).  Like the reference's DMM build (src/pr.cc:
// 87-103) the scores are also appended to pr_output.txt.
bool PRVerifier(const Graph& g, const std::vector<ScoreT>& scores, double target_error) {
  {
    std::ofstream out("pr_output.txt", std::ios::app);
    for (ScoreT s : scores) out << std::setprecision(5) << s << "\n";
  }
  const ScoreT base_score = (1.0f - kDamp) / g.num_nodes();
  std::vector<ScoreT> incoming(g.num_nodes(), 0);
  for (NodeID u = 0; u < g.num_nodes(); u++) {
    ScoreT outgoing_contrib = scores[u] / g.out_degree(u);
    for (int64_t e = g.offsets()[u]; e < g.offsets()[u + 1]; e++) incoming[g.neighs()[e]] += outgoing_contrib;
  }
  double error = 0;
  for (NodeID n = 0; n < g.num_nodes(); n++) error += std::fabs(base_score + kDamp * incoming[n] - scores[n]);
  PrintTime("Total Error", error);
  return error < target_error;
}

int main(int argc, char* argv[]) {
  CLPageRank cli(argc, argv, "pagerank", 1e-4, 20);
  if (!cli.ParseArgs()) return -1;
  Graph g;
  Builder b(cli);
  if (!b.MakeGraph(&g)) return 1;
  auto PRBound = [&cli](const Graph& gr) { return PageRankPull(gr, cli.max_iters(), cli.tolerance()); };
  auto VerifierBound = [&cli](const Graph& gr, const std::vector<ScoreT>& scores) {
    return PRVerifier(gr, scores, cli.tolerance());
  };
  BenchmarkKernel(cli, g, PRBound, PrintTopScores, VerifierBound);
  return 0;
}
