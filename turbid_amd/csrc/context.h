// One DNS on one GPU: operators, device buffers and the substage sequence of
// FlatCarrierPhase / FlatScalarPhase (reference src/phase_carrier_flat.cc,
// src/phase_scalar_flat.cc), launched on one HIP stream.
#pragma once

#include <algorithm>
#include <vector>

#include "../../include/turbid_amd.h"
#include "common.h"
#include "fft.h"
#include "group.h"
#include "setup.h"
#include "zgemm.h"

namespace tbd {

// Kernel classes of the profiler: one class per kernel (and precision), with
// the algorithmic work each launch does (bytes for the HBM-bound passes,
// flops for the contractions).
enum KClass {
  KC_YB64 = 0, KC_XB64, KC_XF64, KC_YF64,     // y/x transform passes, FP64
  KC_YB32, KC_XB32, KC_XF32, KC_YF32,         // the same in FP32
  KC_XFU64, KC_XFU32,                         // fused x transforms + products
  KC_Z64, KC_Z32,                             // wall-normal contractions
  KC_HELM,                                    // Helmholtz / pressure solves
  KC_PROD64, KC_PROD32,                       // pointwise nonlinear products
  KC_OTHER,
  KC_HELMS,                                   // Helmholtz / pressure solves through the even/odd form
  KC_RHELM, KC_RHELMS,                        // rhs + boundary rows + Helmholtz (full / even/odd form)
  KC_PCHAIN,                                  // divergence + pressure solve + projection
  KC_N
};

struct Profiler {
  bool on = false;
  struct Rec { int kclass; double work; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  long long launches[KC_N] = {};
  double ms[KC_N] = {};
  double work[KC_N] = {};   // bytes or flops
  hipEvent_t get();
  void begin(int kclass, hipStream_t st, Rec& r);
  void end(Rec& r, hipStream_t st);
  void collect();
  void reset();
  ~Profiler();
};

class Context {
 public:
  explicit Context(const tbd_config& cfg);
  ~Context();

  void upload(const double* uvw, const double* p, const double* c);
  void download(double* uvw, double* p, double* c);
  void download_physical_velocity(double* out);
  void substage(int stage, double time, int alternation);
  void step(double time, int alternation);
  void project();
  // phase-binned statistics of the current state (on the device)
  void stats_accumulate(int bin);
  void stats_read(int bin, double* out, long long* count);
  void stats_reset();
  static constexpr int kStatBins = 64;
  void set_dt(double dt);
  void features(double* out50);
  void sync();
  void helmholtz_host(int which, int stage, int mode, const double* x, double* y);
  // swirling strength of the current velocity on the physical grid (R
  // values, x fastest): Gradient of u, v, w, DftBackward, per-point
  // eigenvalues (FlatChannelPostprocessor::Execute), all in FP64
  void swirling(double* out);
  // Replace the wall-normal operators D1, D2 (column-major nz*nz; every
  // factorisation is rebuilt from them with the configured boundary rows) or
  // one factorisation (bordered V~, V~^-1, lambda, fix), then re-upload the
  // device copies and drop the captured steps.
  void set_operators(const double* d1, const double* d2);
  void set_factorisation(int which, const double* V, const double* Vinv, const double* lambda, const double* fix);
  void copy_operators_from(const Context& o);

  // Kernel variants of a step, each tested against the oracle:
  //   fused   x transforms fused around the nonlinear products
  //   rhs     right-hand side + boundary rows + solve as one kernel, and the
  //           divergence + pressure solve + projection chain (needs the
  //           even/odd D1, D2; without them the switch has no effect)
  //   overlap scalar stages on a second stream, the even/odd-form solve of w
  //           beside those of u, v on a third
  //   wide[4] two-level transform kernels of the y backward, y forward,
  //           x backward and x forward pass with twice the default workgroup
  //           size (twice the sequences per workgroup)
  //   swap[4] the same four passes with the two factors of the length in
  //           the other order (the radix split N2 x N1; lengths with two
  //           distinct factors only)
  //   zwide   plain contractions (gradients, D1 of the nonlinear terms,
  //           unfused rhs and projection) with twice the column groups per
  //           panel where the LDS allows
  // A switch set in the environment at creation (TBD_FUSED, TBD_RHS_UNFUSED,
  // TBD_NO_OVERLAP; TBD_FFT_WIDE for the four widths) is pinned: set_variant
  // leaves it and autotune does not vary it.
  static constexpr int kVariants = 4096;
  struct Variant {
    bool fused = false, rhs = true, overlap = true;
    bool wide[4] = {false, false, false, false};
    bool swap[4] = {false, false, false, false};
    bool zwide = false;
    int bits() const {
      int b = (fused ? 1 : 0) | (rhs ? 2 : 0) | (overlap ? 4 : 0) | (zwide ? 2048 : 0);
      for (int i = 0; i < 4; ++i) b |= (wide[i] ? 8 << i : 0) | (swap[i] ? 128 << i : 0);
      return b;
    }
    static Variant of(int bits) {
      Variant v;
      v.fused = (bits & 1) != 0;
      v.rhs = (bits & 2) != 0;
      v.overlap = (bits & 4) != 0;
      v.zwide = (bits & 2048) != 0;
      for (int i = 0; i < 4; ++i) {
        v.wide[i] = (bits & (8 << i)) != 0;
        v.swap[i] = (bits & (128 << i)) != 0;
      }
      return v;
    }
  };
  Variant variant() const;
  void set_variant(const Variant& v);
  // Times variants on a scratch context (same configuration, operators and
  // dt; this context's state uploaded to it before each) for `steps`
  // captured steps after three warm-up steps: every combination of fused,
  // rhs and overlap, then each transform width, each radix order and the
  // contraction panel width in turn from the best so far (kept when
  // faster), and keeps the fastest.
  // ms[bits] = ms per step (NaN where not tried, kVariants entries); returns
  // the kept bits.
  int autotune(int steps, double* ms);

  // One step's launches recorded into r instead of issued (the grouped
  // launches of an ensemble, group.h), with the host-side state a step
  // advances (wall-row slot, previous dt) advanced as by step().  The
  // forcing is read from the device pair force_ (written before each replay).
  // Throws if anything of the step was issued on the context's stream
  // rather than recorded (a second stream, the LDS Stockham transforms).
  void record_step(double time, int alternation, Recorder& r);
  // the host-side state advance of a step replayed from a recording
  void replayed_step();
  // whether record_step can record this context's step: no second stream,
  // no fused x transforms (their grouped instances are switched off: a
  // grouped step with them faulted the GPU once, cause not found), the
  // register transforms in x and y, no profiling
  bool groupable() const {
    return !overlap_ && !fused_ && !prof.on && fft_.two_level && fft_has_two_level(d.nx) &&
           fft_has_two_level(d.ny);
  }
  double* force_pair() const { return force_; }
  const double* features_device() const { return fg_; }
  void forcing(double time, double& f1, double& f2) const { forces(time, f1, f2); }
  hipStream_t stream() const { return st_; }
  int wall_slot() const { return wslot_; }
  double dt() const { return dt_; }
  double olddt() const { return olddt_; }

  tbd_config cfg;
  Dims d;
  Chebyshev ch;
  Factorisation fac[5];
  Profiler prof;

 private:
  template <typename T> void set_d1(ZArgs& a) const;
  void set_d2(ZArgs& a) const;
  template <typename T> void carrier_predict(int stage, double time, int alternation);
  template <typename T> void carrier_correct(int stage);
  template <typename T> void scalar_correct(int stage, int alternation);
  void forces(double time, double& f1, double& f2) const;
  template <class F> void launch(int kclass, double work, F&& f);
  double zwork() const;
  // explicit-term update + rhs + boundary rows + Helmholtz of n fields
  // (factorisations k[i], carrier components comp[i]) through rhs_helmholtz: one launch for the
  // fields with a full factorisation, one for those with an even/odd form
  void rhs_solve(RHArgs& base, int n, const int* k, const int* comp, const double2* const* x,
                 double2* const* h, const void* const* adv, double2* const* y, double sigma);
  // Helmholtz solves of n <= 3 fields, field i with factorisation k[i] and
  // stage coefficient sigma[i]: fields whose factorisation has an even/odd
  // form go through helmholtz_split (one launch each), the rest share one
  // helmholtz launch
  void solve(int mode, int n, const int* k, const double* sigma, const double2* const* x, double2* const* y);
  // profiled transform passes (one launch each)
  template <typename T, typename SpecT>
  void ypb(const SpecT* spec, long long sstride, C2<T>* tmp, int nv, int zoff, int nplanes, int scol = -1,
           const int* gate = nullptr);
  template <typename T>
  void xpb(const C2<T>* tmp, T* phys, long long pstride, int nv, int nplanes, const int* gate = nullptr);
  template <typename T>
  void xpf(const T* phys, long long pstride, C2<T>* tmp, int nv, int nplanes, const int* gate = nullptr);
  template <typename T, typename SpecT>
  void ypf(const C2<T>* tmp, SpecT* spec, long long sstride, int nv, int zoff, int nplanes, int scol = -1,
           const int* gate = nullptr);
  template <typename T, typename SpecT>
  void fwd(const T* phys, long long pstride, SpecT* spec, long long sstride, int nv, int zoff, int nplanes,
           C2<T>* tmp, int scol = -1, const int* gate = nullptr) {
    xpf<T>(phys, pstride, tmp, nv, nplanes, gate);
    ypf<T, SpecT>(tmp, spec, sstride, nv, zoff, nplanes, scol, gate);
  }
  template <typename T, typename SpecT>
  void bwd(const SpecT* spec, long long sstride, T* phys, long long pstride, int nv, int zoff, int nplanes,
           C2<T>* tmp, int scol = -1, const int* gate = nullptr) {
    ypb<T, SpecT>(spec, sstride, tmp, nv, zoff, nplanes, scol, gate);
    xpb<T>(tmp, phys, pstride, nv, nplanes, gate);
  }
  void* alloc(size_t bytes);
  void build_factorisations();
  void install_operators();
  void* alloc_op(size_t bytes);        // operator copies (replaced by install_operators)
  std::vector<void*> opallocs_;

  hipStream_t st_ = nullptr;
  FftPlan fft_;
  int lda_ = 0;
  double xpb_ = 0;          // complex elements of one x-pencil component (nz*mx*ny)
  bool opentop_ = false;
  bool fused_ = false;      // x transforms fused around the nonlinear products (TBD_FUSED=1)
  // on the fused conv path, x derivatives from the velocity's x-pencils
  // (TBD_FUSED_DX=0: the nine-gradient form)
  bool fused_dx_ = true;
  double dt_ = 0, olddt_ = 0;
  double coef_[3][3];

  // operators (row-major, padded lda)
  double* d1_ = nullptr;
  double* d2_ = nullptr;
  float* d1f_ = nullptr;
  bool parity_ = false;     // D1, D2 applied through their even/odd split
  int ldp_ = 0;             // leading dimension of the split operators
  double* d1p_ = nullptr;
  double* d2p_ = nullptr;
  float* d1pf_ = nullptr;
  double* vinv_[5] = {nullptr};
  double* vnorm_[5] = {nullptr};
  double* lam_[5] = {nullptr};
  double* fix_[5] = {nullptr};
  // even/odd form of the factorisations that have one (hsplit_)
  bool hsplit_[5] = {false};
  int pld_[5] = {0};
  double* pinv_[5] = {nullptr};
  double* pnorm_[5] = {nullptr};
  double* plam_[5] = {nullptr};
  double* pfix_[5] = {nullptr};
  double2* pgrad_ = nullptr;   // postprocessor work: 9 spectral gradients (FP64), allocated on first use
  double* pphys_ = nullptr;    // 9 physical gradients + 1 output
  double* vinvp_[5] = {nullptr};   // V~^-1 with permuted columns (rhs_helmholtz)
  bool fuse_rhs_ = false;          // rhs + boundary rows + Helmholtz fused (TBD_RHS_UNFUSED=1 disables)
  bool want_rhs_ = true;           // the variant switch (fuse_rhs_ also needs the even/odd operators)
  bool pin_fused_ = false, pin_rhs_ = false, pin_overlap_ = false, pin_wide_ = false;
  bool zwide_ = false;             // plain contractions on panels of twice the column groups
  bool fuse_p_ = false;            // divergence + pressure solve + projection fused (same switch)
  // The scalar's Correct of stage s reads the carrier's physical velocity of
  // stage s and writes only scalar state; without buoyancy the carrier never
  // reads the scalar.  So in step() the scalar of stage s runs on a second
  // stream beside the carrier Predict of stage s + 1, joined before the
  // carrier Correct that overwrites the physical velocity (TBD_NO_OVERLAP=1
  // disables; off while profiling).  The scalar has its own work buffers.
  bool overlap_ = false;
  hipStream_t st2_ = nullptr;
  hipEvent_t evf_ = nullptr, evj_ = nullptr;
  // the carrier's even/odd-form solve (w) beside its full-form solves (u, v)
  hipStream_t st3_ = nullptr;
  hipEvent_t evf3_ = nullptr, evj3_ = nullptr;
  void *sgrad_ = nullptr, *sprod_ = nullptr, *sgphys_ = nullptr, *saphys_ = nullptr, *stmp_ = nullptr;
  void stages(double time, int alternation);
  void pchain(double q, double cf, const double2* s, double2* p, double2* y, int slot, bool count);
  double* kx_ = nullptr;
  double* ky_ = nullptr;
  double* gridz_ = nullptr;
  double* glz_ = nullptr;
  double glx_ = 0, gly_ = 0;

  // spectral FP64 state and work (S complex per component)
  double2 *uvw_ = nullptr, *p_ = nullptr, *c_ = nullptr;
  double2 *huvw_ = nullptr, *hc_ = nullptr, *suvw_ = nullptr;
  double2 *rhs_ = nullptr, *rhsp_ = nullptr, *rhsc_ = nullptr;
  // nonlinear-term spectral work in the transform precision (C2<T>)
  void* grad_ = nullptr;   // 9 components
  void* adv_ = nullptr;    // 3
  void* prod_ = nullptr;   // 6
  void* advc_ = nullptr;   // 1
  // physical work in the transform precision (T)
  void* uvwnl_ = nullptr;  // 3
  void* gphys_ = nullptr;  // 9
  void* aphys_ = nullptr;  // 3
  double* cphys_ = nullptr;
  void* tmp_ = nullptr;    // transform intermediate (bytes for 9 components)
  // x-pencils of u, v, w left by their inverse transform to physical space
  // (after every Correct and every state upload); the fused conv form takes
  // the x derivatives from them
  void* uxp_ = nullptr;
  // wall rows of grad p: m1/m2 (bottom, top) ping-pong, and open-top rows
  double2* wall_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // [slot][bottom/top], 3P each
  double2* wall2_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // 2P each
  int wslot_ = 0;          // slot holding m1
  // erosion
  // bottom planes of du/dz, dv/dz (T) written by the carrier's Predict and
  // read by the scalar's Correct of the same stage; one pair per stage parity,
  // so the scalar of stage s may run beside the carrier Predict of s + 1
  void* dudz_bot_[2] = {nullptr, nullptr};
  void* dvdz_bot_[2] = {nullptr, nullptr};
  double* ephys_ = nullptr;
  double2* ehat_ = nullptr;
  double2* wrows_ = nullptr;   // 4P: bottom/top rows of D1 u, D1 v
  // features
  double* fl_ = nullptr;       // featurelocal_ (64)
  double* fg_ = nullptr;       // featureglobal_ (64)
  double* wallmean_ = nullptr; // 6: D1 u/v/w rows 0 and nz-1 on the mean column
  double* diagpart_ = nullptr; // per-block maxima of the diagnostics

  // captured steps (graph replay; TBD_GRAPH=0 runs the substages eagerly)
  struct StepGraph {
    int alternation = 0, slot = 0;
    double dt = 0, olddt = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<StepGraph> graphs_;
  static constexpr size_t kMaxGraphs = 8;
  // (dt, previous dt) of the last step: a step whose pair differs from the
  // last one's runs eagerly (time mode moves dt almost every step, and a
  // capture that is never replayed costs more than it saves)
  double last_dt_ = -1, last_olddt_ = -1;
  bool use_graphs_ = true;
  bool capturing_ = false;
  double* force_ = nullptr;    // mean-mode force read by the captured H update
  int* filt_ = nullptr;        // set by the positivity filter when it changes c
  int* need_ = nullptr;        // set when physical c may hold a negative value (the filter's gate)
  double* mmpart_ = nullptr;   // per-block min/max of the negativity check
  double* stats_ = nullptr;    // [kStatBins][kPhaseStats][nz] sums
  long long stat_count_[64] = {};
  void drop_graphs();

  std::vector<void*> allocs_;
};

}  // namespace tbd
