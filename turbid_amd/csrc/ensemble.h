// An ensemble of independent DNS contexts (uneven grids, Reynolds numbers,
// settling velocities) advanced together on one stream with one launch per
// kernel and launch shape per position of the step: each member's step is
// recorded once per (alternation, wall-row slot, dt, previous dt, variant) of
// every member, the members' launch sequences are merged into grouped launches
// (group.h), and the merged sequence is captured into one HIP graph.
#pragma once

#include <vector>

#include "context.h"
#include "group.h"

namespace tbd {

class Ensemble {
 public:
  // The members are not owned and must outlive the ensemble.  At most
  // kMaxMembers, every one of them on this device.
  explicit Ensemble(const std::vector<Context*>& members);
  ~Ensemble();
  Ensemble(const Ensemble&) = delete;
  Ensemble& operator=(const Ensemble&) = delete;

  // One RK3 step of every member (member m at time[m] with alternation[m]).
  // Ordered after the work already issued on the members' streams; work
  // issued on a member's stream afterwards is ordered after it.
  void step(const double* time, const int* alternation);
  // of the last recorded step: grouped launches, and the launches of the
  // members together (what one launch per member and kernel would issue)
  int grouped_launches() const { return last_grouped_; }
  int member_launches() const { return last_member_; }
  void sync();
  // the 50 features of every member after the last step (member m at
  // out + 50 m): one gather and one copy, then a wait for the ensemble
  void features(double* out);

 private:
  struct Launch {
    const void* fn = nullptr;
    dim3 block;
    size_t lds = 0;
    int total = 0;                 // workgroups over all members in the launch
    std::vector<char> own;         // argument bytes of the launch's first member (unused by the kernel)
    size_t group_off = 0;          // offset of the launch's Group in the arena
    std::vector<int> members;      // member indices, in the Group's order
  };
  struct Program {
    std::vector<int> alt, slot, bits;
    std::vector<double> dt, olddt;
    char* arena = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int grouped = 0, member = 0;
  };
  void build(Program& p, const double* time, const int* alternation);
  void drop();

  std::vector<Context*> m_;
  std::vector<hipEvent_t> mev_;
  hipStream_t st_ = nullptr;
  hipEvent_t ev_ = nullptr;
  std::vector<Program> progs_;
  static constexpr size_t kMaxPrograms = 8;
  int last_grouped_ = 0, last_member_ = 0;
  double* feat_dev_ = nullptr;   // n x 50
  double* feat_host_ = nullptr;  // pinned copy
};

}  // namespace tbd
