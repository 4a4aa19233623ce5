// Host harness of the bfs / pr apps: the reference's command line, builder,
// source picker and trial loop (src/command_line.h, src/benchmark.h,
// src/util.h, src/timer.h) over the C ABI in include/dmm_gapbs.h.
//
// The same flags, defaults and output lines as the reference, minus the
// OpenSHMEM runtime: the process is PE 0 of 1 and one GPU holds the graph.
#ifndef DMM_APPS_HARNESS_H_
#define DMM_APPS_HARNESS_H_

#include <getopt.h>

#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "dmm_gapbs.h"

typedef int32_t NodeID;  // src/benchmark.h:320

// ---- src/util.h:28-59 ------------------------------------------------------
inline void PrintLabel(const std::string& label, const std::string& val) {
  printf("PE: %d | %-21s%7s\n", 0, (label + ":").c_str(), val.c_str());
}
inline void PrintTime(const std::string& s, double seconds) {
  printf("PE: %d | %-21s%3.5lf\n", 0, (s + ":").c_str(), seconds);
}
inline void PrintStep(const std::string& s, int64_t count) {
  printf("PE: %d | %-14s%14" PRId64 "\n", 0, (s + ":").c_str(), count);
}
inline void PrintStep(const std::string& s, double seconds, int64_t count = -1) {
  if (count != -1)
    printf("PE: %d | %5s%11" PRId64 "  %10.5lf\n", 0, s.c_str(), count, seconds);
  else
    printf("PE: %d | %5s%23.5lf\n", 0, s.c_str(), seconds);
}

// ---- src/timer.h:26-56 -----------------------------------------------------
class Timer {
 public:
  void Start() { start_ = elapsed_ = std::chrono::steady_clock::now(); }
  void Stop() { elapsed_ = std::chrono::steady_clock::now(); }
  double Seconds() const { return std::chrono::duration<double>(elapsed_ - start_).count(); }

 private:
  std::chrono::steady_clock::time_point start_, elapsed_;
};

// ---- src/command_line.h:28-189 ---------------------------------------------
class CLBase {
 protected:
  int argc_;
  char** argv_;
  std::string name_;
  std::string get_args_ = "f:g:hk:su:";
  std::vector<std::string> help_strings_;
  int scale_ = -1;
  int degree_ = 16;
  std::string filename_ = "";
  bool symmetrize_ = false;
  bool uniform_ = false;

  void AddHelpLine(char opt, std::string opt_arg, std::string text, std::string def = "") {
    char buf[100];
    if (opt_arg != "") opt_arg = "<" + opt_arg + ">";
    if (def != "") def = "[" + def + "]";
    snprintf(buf, sizeof(buf), " -%c %-9s: %-54s%10s", opt, opt_arg.c_str(), text.c_str(), def.c_str());
    help_strings_.push_back(buf);
  }

 public:
  CLBase(int argc, char** argv, std::string name = "") : argc_(argc), argv_(argv), name_(name) {
    AddHelpLine('h', "", "print this help message");
    AddHelpLine('f', "file", "load graph from file");
    AddHelpLine('s', "", "symmetrize input edge list", "false");
    AddHelpLine('g', "scale", "generate 2^scale kronecker graph");
    AddHelpLine('u', "scale", "generate 2^scale uniform-random graph");
    AddHelpLine('k', "degree", "average degree for synthetic graph", std::to_string(degree_));
  }
  virtual ~CLBase() {}

  bool ParseArgs() {
    signed char c_opt;
    while ((c_opt = getopt(argc_, argv_, get_args_.c_str())) != -1) HandleArg(c_opt, optarg);
    if ((filename_ == "") && (scale_ == -1)) {
      std::cout << "No graph input specified. (Use -h for help)" << std::endl;
      return false;
    }
    // Files: the serialized graph cache (.sg, src/reader.h:291-345); the
    // text edge-list readers are not part of this build.
    if (filename_ != "" && (filename_.size() < 3 || filename_.compare(filename_.size() - 3, 3, ".sg") != 0)) {
      std::cout << "Only .sg graph files are supported by this build; use -g or -u otherwise" << std::endl;
      return false;
    }
    if (scale_ != -1) symmetrize_ = true;
    return true;
  }

  virtual void HandleArg(signed char opt, char* opt_arg) {
    switch (opt) {
      case 'f': filename_ = std::string(opt_arg); break;
      case 'g': scale_ = atoi(opt_arg); break;
      case 'h': PrintUsage(); break;
      case 'k': degree_ = atoi(opt_arg); break;
      case 's': symmetrize_ = true; break;
      case 'u': uniform_ = true; scale_ = atoi(opt_arg); break;
    }
  }

  void PrintUsage() {
    std::cout << name_ << std::endl;
    for (const std::string& h : help_strings_) std::cout << h << std::endl;
    std::exit(0);
  }

  int scale() const { return scale_; }
  int degree() const { return degree_; }
  std::string filename() const { return filename_; }
  bool symmetrize() const { return symmetrize_; }
  bool uniform() const { return uniform_; }
};

class CLApp : public CLBase {
  bool do_analysis_ = false;
  int num_trials_ = 16;
  int64_t start_vertex_ = -1;
  bool do_verify_ = false;

 public:
  CLApp(int argc, char** argv, std::string name) : CLBase(argc, argv, name) {
    get_args_ += "an:r:v";
    AddHelpLine('a', "", "output analysis of last run", "false");
    AddHelpLine('n', "n", "perform n trials", std::to_string(num_trials_));
    AddHelpLine('r', "node", "start from node r", "rand");
    AddHelpLine('v', "", "verify the output of each run", "false");
  }
  void HandleArg(signed char opt, char* opt_arg) override {
    switch (opt) {
      case 'a': do_analysis_ = true; break;
      case 'n': num_trials_ = atoi(opt_arg); break;
      case 'r': start_vertex_ = atol(opt_arg); break;
      case 'v': do_verify_ = true; break;
      default: CLBase::HandleArg(opt, opt_arg);
    }
  }
  bool do_analysis() const { return do_analysis_; }
  int num_trials() const { return num_trials_; }
  int64_t start_vertex() const { return start_vertex_; }
  bool do_verify() const { return do_verify_; }
};

class CLPageRank : public CLApp {
  int max_iters_;
  double tolerance_;

 public:
  CLPageRank(int argc, char** argv, std::string name, double tolerance, int max_iters)
      : CLApp(argc, argv, name), max_iters_(max_iters), tolerance_(tolerance) {
    get_args_ += "i:t:";
    AddHelpLine('i', "i", "perform at most i iterations", std::to_string(max_iters_));
    AddHelpLine('t', "t", "use tolerance t", std::to_string(tolerance_));
  }
  void HandleArg(signed char opt, char* opt_arg) override {
    switch (opt) {
      case 'i': max_iters_ = atoi(opt_arg); break;
      case 't': tolerance_ = std::stod(opt_arg); break;
      default: CLApp::HandleArg(opt, opt_arg);
    }
  }
  int max_iters() const { return max_iters_; }
  double tolerance() const { return tolerance_; }
};

// ---- Graph: host CSR + the device context holding its copy ----------------
class Graph {
 public:
  Graph() {}
  ~Graph() {
    if (ctx_) dmm_ctx_destroy(ctx_);
    if (hg_) dmm_host_graph_free(hg_);
  }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  int64_t num_nodes() const { return n_; }
  int64_t num_edges() const { return num_edges_directed() / 2; }
  int64_t num_edges_directed() const { return dmm_host_graph_num_edges_directed(hg_); }
  int64_t out_degree(NodeID v) const { return off_[v + 1] - off_[v]; }
  const int64_t* offsets() const { return off_; }
  const NodeID* neighs() const { return nbr_; }
  dmm_ctx* ctx() const { return ctx_; }

  // src/graph.h:PrintStats
  void PrintStats() const {
    std::cout << "Graph has " << n_ << " nodes and " << num_edges() << " undirected edges for degree: "
              << (n_ ? num_edges() / n_ : 0) << std::endl;
  }

  bool Init(dmm_host_graph* hg, int device) {
    hg_ = hg;
    n_ = dmm_host_graph_num_nodes(hg);
    off_ = dmm_host_graph_out_offsets(hg);
    nbr_ = dmm_host_graph_out_neighs(hg);
    if (dmm_ctx_create(device, &ctx_) != DMM_OK) {
      std::cerr << "dmm_ctx_create: " << dmm_last_error(nullptr) << std::endl;
      return false;
    }
    if (dmm_graph_upload(ctx_, n_, 0, n_, off_, nbr_, nullptr, nullptr, num_edges_directed()) != DMM_OK) {
      std::cerr << "dmm_graph_upload: " << dmm_last_error(ctx_) << std::endl;
      return false;
    }
    return true;
  }

 private:
  dmm_host_graph* hg_ = nullptr;
  dmm_ctx* ctx_ = nullptr;
  int64_t n_ = 0;
  const int64_t* off_ = nullptr;
  const NodeID* nbr_ = nullptr;
};

// ---- BuilderBase::MakeGraph (src/builder.h:314-338) -------------------------
class Builder {
 public:
  explicit Builder(const CLBase& cli) : cli_(cli) {}
  bool MakeGraph(Graph* g, int device = 0) {
    Timer t;
    t.Start();
    dmm_host_graph* hg = nullptr;
    if (cli_.filename() != "") {  // src/builder.h:320-322
      if (dmm_read_serialized_graph(cli_.filename().c_str(), &hg) != DMM_OK) {
        std::cerr << "graph read failed: " << dmm_last_error(nullptr) << std::endl;
        return false;
      }
      t.Stop();
      PrintTime("Read Time", t.Seconds());
      return g->Init(hg, device);
    }
    if (dmm_build_graph(cli_.scale(), cli_.degree(), cli_.uniform() ? 1 : 0, 0, &hg) != DMM_OK) {
      std::cerr << "graph build failed: " << dmm_last_error(nullptr) << std::endl;
      return false;
    }
    t.Stop();
    PrintTime("Build Time", t.Seconds());
    return g->Init(hg, device);
  }

 private:
  const CLBase& cli_;
};

// ---- SourcePicker (src/benchmark.h:49-87) -----------------------------------
class SourcePicker {
 public:
  explicit SourcePicker(const Graph& g, NodeID given_source = -1) : g_(g), given_(given_source) {}
  NodeID PickNext() {
    // Draw i+1 sources from the fresh generator and keep the last: the
    // sequence equals the reference's stateful picker.
    next_++;
    if (cache_.size() < static_cast<size_t>(next_)) {
      cache_.resize(std::max<size_t>(64, 2 * next_));
      dmm_pick_sources(g_.num_nodes(), g_.offsets(), given_, static_cast<int32_t>(cache_.size()), cache_.data());
    }
    return cache_[next_ - 1];
  }

 private:
  const Graph& g_;
  NodeID given_;
  int next_ = 0;
  std::vector<NodeID> cache_;
};

// ---- BenchmarkKernel (src/benchmark.h:118-145) ------------------------------
template <typename GraphFunc, typename AnalysisFunc, typename VerifierFunc>
void BenchmarkKernel(const CLApp& cli, const Graph& g, GraphFunc kernel, AnalysisFunc stats,
                     VerifierFunc verify) {
  g.PrintStats();
  double total_seconds = 0;
  Timer trial_timer;
  for (int iter = 0; iter < cli.num_trials(); iter++) {
    trial_timer.Start();
    auto result = kernel(g);
    trial_timer.Stop();
    PrintTime("Trial Time", trial_timer.Seconds());
    total_seconds += trial_timer.Seconds();
    if (cli.do_analysis() && (iter == (cli.num_trials() - 1))) stats(g, result);
    if (cli.do_verify()) PrintLabel("Verification", verify(g, result) ? "PASS" : "FAIL");
  }
  PrintTime("Average Time", total_seconds / cli.num_trials());
}

#endif  // DMM_APPS_HARNESS_H_
