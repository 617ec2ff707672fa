// Planned, graph-captured batched commit (see batch.hip).
#pragma once
#include <string>
#include <vector>

#include "internal.hpp"

namespace p3g {

struct CommitPlan {
  struct Launch {
    enum Kind { NTT, NTT_SINGLE, LEAF } kind = NTT;
    uint32_t mode = 0;
    bool reg = false;  // register-blocked NTT kernel
    uint32_t K = 0;     // its stage count
    uint32_t nd = 0;
    uint64_t blocks = 0;
    const void* descs = nullptr;
    const uint64_t* starts = nullptr;
    std::vector<uint8_t> args;  // NTT_SINGLE: the pass descriptor, passed by value
    // leaf hashing
    uint64_t h = 0;
    int prev = -1, out_state = -1;
    uint32_t prev_shift = 0;
    bool final = false;
  };

  Ctx* ctx = nullptr;
  uint32_t hash = HASH_P2;
  unsigned log_blowup = 1;
  bool use_graph = true;
  std::vector<DevMat> traces;  // device inputs, filled by the caller before run()
  std::vector<DevMat> ldes;    // bit-reversed LDEs, valid until the next run()
  std::vector<const uint64_t*> tw_inv, tw_fwd, scale_dev;
  std::vector<uint64_t*> scale_owned;
  std::vector<uint64_t> scale_stride;
  uint64_t height = 0;
  unsigned depth = 0;
  uint64_t* digests = nullptr;
  std::vector<uint64_t> layer_off;
  uint64_t* states[2] = {nullptr, nullptr};
  uint8_t* blob = nullptr;
  std::vector<Launch> launches;
  uint32_t kmax = 8;
  int leaf_threads = 128;
  bool grouped = true;  // one launch per pass index over all traces, or one per trace
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t root[4] = {0, 0, 0, 0};
  std::string tune_log;

  ~CommitPlan();
  std::string shape_key() const;
  void build_launches(uint32_t kmax, int leaf_threads, bool grouped);
  void enqueue(hipStream_t s);
  void capture(hipStream_t s);
  double time_once(hipStream_t s);
  void autotune(hipStream_t s);
  void run(hipStream_t s);
  Tree* tree_view();  // non-owning view of the last run's LDEs and digest layers
};

CommitPlan* make_commit_plan(Ctx& ctx, uint32_t hash, const std::vector<uint64_t>& heights,
                             const std::vector<uint64_t>& widths, unsigned log_blowup, bool use_graph, bool autotune,
                             bool grouped = true);

}  // namespace p3g
