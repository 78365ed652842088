// Ensemble stepping through grouped launches (ensemble.h, group.h).
#include "ensemble.h"

#include <algorithm>
#include <climits>
#include <cstring>

#include "pointwise.h"

namespace tbd {

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

Ensemble::Ensemble(const std::vector<Context*>& members) : m_(members) {
  TBD_CHECK(!m_.empty() && m_.size() <= (size_t)kMaxMembers, "Ensemble > between 1 and 32 members");
  for (Context* c : m_) TBD_CHECK(c != nullptr, "Ensemble > null member");
  TBD_HIP(hipStreamCreate(&st_));
  TBD_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  mev_.resize(m_.size(), nullptr);
  for (auto& e : mev_) TBD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  TBD_HIP(hipMalloc((void**)&feat_dev_, sizeof(double) * 50 * m_.size()));
  TBD_HIP(hipHostMalloc((void**)&feat_host_, sizeof(double) * 50 * m_.size(), hipHostMallocDefault));
}

Ensemble::~Ensemble() {
  drop();
  if (st_) (void)hipStreamSynchronize(st_);
  if (feat_dev_) (void)hipFree(feat_dev_);
  if (feat_host_) (void)hipHostFree(feat_host_);
  for (auto e : mev_) if (e) (void)hipEventDestroy(e);
  if (ev_) (void)hipEventDestroy(ev_);
  if (st_) (void)hipStreamDestroy(st_);
}

void Ensemble::drop() {
  if (progs_.empty()) return;
  (void)hipStreamSynchronize(st_);
  for (auto& p : progs_) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    if (p.arena) (void)hipFree(p.arena);
  }
  progs_.clear();
}

void Ensemble::sync() { TBD_HIP(hipStreamSynchronize(st_)); }

void Ensemble::features(double* out) {
  const int n = (int)m_.size();
  std::vector<const double*> src(n);
  for (int m = 0; m < n; ++m) src[m] = m_[m]->features_device();
  gather_rows(src.data(), n, 50, feat_dev_, st_);
  TBD_HIP(hipMemcpyAsync(feat_host_, feat_dev_, sizeof(double) * 50 * n, hipMemcpyDeviceToHost, st_));
  TBD_HIP(hipStreamSynchronize(st_));
  std::memcpy(out, feat_host_, sizeof(double) * 50 * n);
}

// Records every member's step, merges the launch sequences and captures the
// merged sequence.  Merge by position: the i-th launches of all members that
// have one are grouped by kernel instance and workgroup shape, one grouped
// launch per group, position after position.  Members with alike steps (the
// same variant and features) differ at a position only where the instance
// depends on the grid (transform lengths, contraction depth); members whose
// steps differ (another precision) merge wherever they coincide.  Each
// member's launches keep their order and members share no writable memory,
// so any grouping is correct.
void Ensemble::build(Program& p, const double* time, const int* alternation) {
  const int n = (int)m_.size();
  std::vector<Recorder> rec(n);
  for (int m = 0; m < n; ++m) m_[m]->record_step(time[m], alternation[m], rec[m]);
  std::vector<Launch> launches;
  std::vector<size_t> args_off;
  std::vector<char> image;
  int member_total = 0;
  size_t maxlen = 0;
  for (int m = 0; m < n; ++m) {
    member_total += (int)rec[m].ops.size();
    maxlen = std::max(maxlen, rec[m].ops.size());
  }
  for (size_t pos = 0; pos < maxlen; ++pos) {
    std::vector<bool> taken(n, false);
    for (int m0 = 0; m0 < n; ++m0) {
      if (taken[m0] || pos >= rec[m0].ops.size()) continue;
      const Rec& r0 = rec[m0].ops[pos];
      std::vector<const Rec*> mem;
      std::vector<int> idx;
      for (int m = m0; m < n; ++m) {
        if (taken[m] || pos >= rec[m].ops.size()) continue;
        const Rec& r = rec[m].ops[pos];
        if (r.fn != r0.fn || r.block.x != r0.block.x || r.block.y != r0.block.y || r.block.z != r0.block.z ||
            r.args.size() != r0.args.size())
          continue;
        mem.push_back(&r);
        idx.push_back(m);
        taken[m] = true;
      }
      Launch L;
      L.members = idx;
      L.fn = r0.fn;
      L.block = r0.block;
      L.own = r0.args;
      const size_t stride = align_up(r0.args.size(), 16);
      L.group_off = align_up(image.size(), 256);
      const size_t aoff = align_up(L.group_off + sizeof(Group), 256);
      image.resize(aoff + mem.size() * stride, 0);
      Group g;
      std::memset(&g, 0, sizeof(g));
      g.n = (int)mem.size();
      g.stride = (int)stride;
      long long total = 0;
      for (size_t i = 0; i < mem.size(); ++i) {
        const Rec& r = *mem[i];
        g.start[i] = (int)total;
        g.gx[i] = (int)r.grid.x;
        g.gy[i] = (int)r.grid.y;
        g.gz[i] = (int)r.grid.z;
        total += (long long)r.grid.x * r.grid.y * r.grid.z;
        TBD_CHECK(total < INT_MAX, "Ensemble > too many workgroups in one grouped launch");
        if (r.lds > L.lds) L.lds = r.lds;
        std::memcpy(image.data() + aoff + i * stride, r.args.data(), r.args.size());
      }
      g.start[mem.size()] = (int)total;
      std::memcpy(image.data() + L.group_off, &g, sizeof(g));
      L.total = (int)total;
      launches.push_back(std::move(L));
      args_off.push_back(aoff);
    }
  }
  TBD_HIP(hipMalloc((void**)&p.arena, image.empty() ? 256 : image.size()));
  for (size_t i = 0; i < launches.size(); ++i) {
    Group* g = reinterpret_cast<Group*>(image.data() + launches[i].group_off);
    g->args = (unsigned long long)(p.arena + args_off[i]);
  }
  TBD_HIP(hipMemcpy(p.arena, image.data(), image.size(), hipMemcpyHostToDevice));
  // the graph's edges follow the members: a grouped launch waits for the
  // previous launch of each of its members only, so launches of disjoint
  // member sets (different transform lengths at one position) are free to
  // run side by side
  TBD_HIP(hipGraphCreate(&p.graph, 0));
  std::vector<hipGraphNode_t> last(n, nullptr);
  for (const Launch& L : launches) {
    if (L.total == 0) continue;
    std::vector<hipGraphNode_t> deps;
    for (int m : L.members)
      if (last[m] && std::find(deps.begin(), deps.end(), last[m]) == deps.end()) deps.push_back(last[m]);
    const Group* gp = reinterpret_cast<const Group*>(p.arena + L.group_off);
    void* kargs[2] = {(void*)L.own.data(), (void*)&gp};
    hipKernelNodeParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.func = const_cast<void*>(L.fn);
    kp.gridDim = dim3(L.total);
    kp.blockDim = L.block;
    kp.sharedMemBytes = (unsigned int)L.lds;
    kp.kernelParams = kargs;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    TBD_HIP(hipGraphAddKernelNode(&node, p.graph, deps.empty() ? nullptr : deps.data(), deps.size(), &kp));
    for (int m : L.members) last[m] = node;
  }
  TBD_HIP(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
  p.grouped = (int)launches.size();
  p.member = member_total;
}

void Ensemble::step(const double* time, const int* alternation) {
  const int n = (int)m_.size();
  for (Context* c : m_)
    TBD_CHECK(c->groupable(), "Ensemble::step > a member runs with the overlap or fused-transform variant bit, "
                              "the LDS Stockham transforms or the profiler: clear them first");
  Program* prog = nullptr;
  for (auto& p : progs_) {
    bool same = true;
    for (int m = 0; m < n && same; ++m)
      same = p.alt[m] == alternation[m] && p.slot[m] == m_[m]->wall_slot() && p.dt[m] == m_[m]->dt() &&
             p.olddt[m] == m_[m]->olddt() && p.bits[m] == m_[m]->variant().bits();
    if (same) {
      prog = &p;
      break;
    }
  }
  // the members' work so far, then the forcing pairs the step reads
  for (int m = 0; m < n; ++m) {
    TBD_HIP(hipEventRecord(mev_[m], m_[m]->stream()));
    TBD_HIP(hipStreamWaitEvent(st_, mev_[m], 0));
  }
  std::vector<double*> dst(n);
  std::vector<double> v(2 * n);
  for (int m = 0; m < n; ++m) {
    dst[m] = m_[m]->force_pair();
    m_[m]->forcing(time[m], v[2 * m], v[2 * m + 1]);
  }
  set_pairs(dst.data(), v.data(), n, st_);
  if (prog) {
    for (Context* c : m_) c->replayed_step();
  } else {
    if (progs_.size() >= kMaxPrograms) drop();
    Program p;
    for (int m = 0; m < n; ++m) {
      p.alt.push_back(alternation[m]);
      p.slot.push_back(m_[m]->wall_slot());
      p.dt.push_back(m_[m]->dt());
      p.olddt.push_back(m_[m]->olddt());
      p.bits.push_back(m_[m]->variant().bits());
    }
    try {
      build(p, time, alternation);
    } catch (...) {
      if (p.exec) (void)hipGraphExecDestroy(p.exec);
      if (p.graph) (void)hipGraphDestroy(p.graph);
      if (p.arena) (void)hipFree(p.arena);
      throw;
    }
    progs_.push_back(p);
    prog = &progs_.back();
  }
  last_grouped_ = prog->grouped;
  last_member_ = prog->member;
  TBD_HIP(hipGraphLaunch(prog->exec, st_));
  TBD_HIP(hipEventRecord(ev_, st_));
  for (Context* c : m_) TBD_HIP(hipStreamWaitEvent(c->stream(), ev_, 0));
}

}  // namespace tbd
