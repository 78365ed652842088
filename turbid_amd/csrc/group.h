// Grouped launches: one kernel over the workgroups of several independent
// solver contexts (ensemble members), each with its own grid and its own
// argument struct.
//
// Every kernel of the substage is written once, as `template <..., bool G>
// __global__ void k(A a, GroupP g)`.  With G = false it is the plain
// launch (grid and arguments its own, g unused).  With G = true one launch
// covers the members of a Group: member m owns the linear workgroups
// [start[m], start[m + 1]) and sees, through `member`, the workgroup index
// and grid its own launch would have had and its own argument struct (in
// device memory, written before the launch).  The kernel body is the same in
// both; only blockIdx/gridDim are read through the Blk it returns.
//
// Host side: the wrappers launch through `glaunch`.  While a Recorder is
// active on the calling thread, glaunch records the launch (the grouped
// instance, the argument bytes, grid, block, LDS) instead of issuing it; an
// ensemble replays the members' records merged into grouped launches.
#pragma once

#include <vector>

#include "common.h"

namespace tbd {

constexpr int kMaxMembers = 32;

struct Group {
  unsigned long long args;          // device address of member 0's argument struct; member m's at args + m * stride
  int stride;
  int n;                            // members in this launch
  int start[kMaxMembers + 1];       // first linear workgroup of each member, start[n] = total
  int gx[kMaxMembers], gy[kMaxMembers], gz[kMaxMembers];
};

// Kernels see the Group, and through it the argument table, in the constant
// address space, as they see their own arguments: the compiler keeps what it
// reads there in scalar registers and treats the pointers it holds as global.
typedef const __attribute__((address_space(4))) Group* GroupP;

// a workgroup's index and grid as the member's own launch would give them
struct Blk {
  int x, y, z, gx, gy, gz;
};

template <bool G, class A>
__device__ __forceinline__ const A& member(const A& own, GroupP g, Blk& b) {
  if constexpr (G) {
    const int lin = (int)blockIdx.x;
    int m = 0;
    while (m + 1 < g->n && lin >= g->start[m + 1]) ++m;
    int l = lin - g->start[m];
    b.gx = g->gx[m];
    b.gy = g->gy[m];
    b.gz = g->gz[m];
    b.x = l % b.gx;
    l /= b.gx;
    b.y = l % b.gy;
    b.z = l / b.gy;
    typedef const __attribute__((address_space(4))) char* ConstP;
    const ConstP p = (ConstP)g->args + (size_t)m * g->stride;
    return *reinterpret_cast<const A*>((const char*)p);
  } else {
    b.x = (int)blockIdx.x;
    b.y = (int)blockIdx.y;
    b.z = (int)blockIdx.z;
    b.gx = (int)gridDim.x;
    b.gy = (int)gridDim.y;
    b.gz = (int)gridDim.z;
    return own;
  }
}

// One recorded launch.
struct Rec {
  const void* fn = nullptr;         // the grouped (G = true) instance
  std::vector<char> args;           // the argument struct's bytes
  dim3 grid, block;
  size_t lds = 0;
};

// While active on a thread, glaunch appends to `ops` instead of launching.
struct Recorder {
  std::vector<Rec> ops;
  static Recorder*& active();
};

template <class A>
void glaunch(void (*single)(A, GroupP), void (*grouped)(A, GroupP), dim3 grid, dim3 block, size_t lds,
             hipStream_t st, const A& a) {
  if (Recorder* r = Recorder::active()) {
    Rec rec;
    rec.fn = (const void*)grouped;
    rec.args.assign(reinterpret_cast<const char*>(&a), reinterpret_cast<const char*>(&a) + sizeof(A));
    rec.grid = grid;
    rec.block = block;
    rec.lds = lds;
    r->ops.push_back(std::move(rec));
    return;
  }
  hipLaunchKernelGGL(single, grid, block, lds, st, a, (GroupP)nullptr);
  TBD_HIP(hipGetLastError());
}

}  // namespace tbd
