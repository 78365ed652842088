// Substage orchestration on one stream (see context.h).  Each step below
// cites the reference statement it implements.
#include "context.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>

#include "nonlinear.h"
#include "pointwise.h"
#include "zgemm.h"

namespace tbd {

static const double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// profiler: event pairs around launches of a kernel class
// ---------------------------------------------------------------------------
hipEvent_t Profiler::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  TBD_HIP(hipEventCreate(&e));
  return e;
}
void Profiler::begin(int kclass, hipStream_t st, Rec& r) {
  r.kclass = kclass;
  r.a = get();
  r.b = get();
  TBD_HIP(hipEventRecord(r.a, st));
}
void Profiler::end(Rec& r, hipStream_t st) {
  TBD_HIP(hipEventRecord(r.b, st));
  recs.push_back(r);
}
void Profiler::collect() {
  for (auto& r : recs) {
    TBD_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    TBD_HIP(hipEventElapsedTime(&t, r.a, r.b));
    launches[r.kclass] += 1;
    ms[r.kclass] += t;
    work[r.kclass] += r.work;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  recs.clear();
}
void Profiler::reset() {
  collect();
  for (int i = 0; i < KC_N; ++i) { launches[i] = 0; ms[i] = 0; work[i] = 0; }
}
Profiler::~Profiler() {
  for (auto& r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : pool) (void)hipEventDestroy(e);
}

template <class F>
void Context::launch(int kclass, double work, F&& f) {
  if (!prof.on) {
    f();
    return;
  }
  Profiler::Rec r;
  r.work = work;
  prof.begin(kclass, st_, r);
  f();
  prof.end(r, st_);
}

template <typename T> static constexpr bool is32() { return std::is_same<T, float>::value; }

// algorithmic work: a contraction of one field is 4*nz^2*P flops (real
// operator times complex columns); each pass moves its input and its output
static double zflops(const Dims& d) { return 4.0 * d.nz * (double)d.nz * d.P; }

// flops one contraction of one field executes: the even/odd split applies an
// (M+1) x (M+1) and an (M+1) x M block instead of the nz x nz operator
double Context::zwork() const {
  if (!parity_) return zflops(d);
  const double M = (d.nz - 1) / 2;
  return 4.0 * ((M + 1) * (M + 1) + (M + 1) * M) * (double)d.P;
}

void Context::rhs_solve(RHArgs& base, int n, const int* k, const int* comp, const double2* const* x,
                        double2* const* h, const void* const* adv, double2* const* y, double sigma) {
  base.nz = d.nz; base.P = (int)d.P; base.nyc = d.nyc; base.kx = kx_; base.ky = ky_;
  base.d2p = d2p_; base.ldp = ldp_;
  // with fields of both kinds, the even/odd-form launch runs on a stream of
  // its own beside the full-form one (joined before returning)
  int kinds = 0;
  for (int i = 0; i < n; ++i) kinds |= hsplit_[k[i]] ? 2 : 1;
  const bool fork = overlap_ && !prof.on && kinds == 3;
  const hipStream_t main = st_;
  if (fork) {
    TBD_HIP(hipEventRecord(evf3_, main));
    TBD_HIP(hipStreamWaitEvent(st3_, evf3_, 0));
  }
  for (int pass = 0; pass < 2; ++pass) {
    RHArgs a = base;
    a.split = pass;
    a.lda = pass ? 0 : lda_;
    int nf = 0;
    for (int i = 0; i < n; ++i) {
      const int f = k[i];
      if ((int)hsplit_[f] != pass) continue;
      RHField& q = a.f[nf++];
      q.x = x[i]; q.h = h[i]; q.adv = adv[i]; q.y = y[i]; q.comp = comp[i]; q.sigma = sigma; q.fix = fix_[f];
      if (pass) {
        a.lda = pld_[f];
        q.vinv = pinv_[f]; q.vnorm = pnorm_[f]; q.lambda = plam_[f]; q.pfix = pfix_[f];
      } else {
        q.vinv = vinvp_[f]; q.vnorm = vnorm_[f]; q.lambda = lam_[f];
      }
    }
    if (nf == 0) continue;
    const hipStream_t sx = (fork && pass == 1) ? st3_ : main;
    launch(pass ? KC_RHELMS : KC_RHELM, nf * rhs_helmholtz_flops(d.nz, d.P, pass == 1),
           [&] { rhs_helmholtz(a, nf, sx); });
  }
  if (fork) {
    TBD_HIP(hipEventRecord(evj3_, st3_));
    TBD_HIP(hipStreamWaitEvent(main, evj3_, 0));
  }
}

void Context::pchain(double q, double cf, const double2* s, double2* p, double2* y, int slot, bool count) {
  PArgs a;
  a.nz = d.nz; a.P = (int)d.P; a.nyc = d.nyc; a.kx = kx_; a.ky = ky_;
  a.d1p = d1p_; a.ldp = ldp_;
  a.pinv = pinv_[3]; a.pnorm = pnorm_[3]; a.plam = plam_[3]; a.pfix = pfix_[3]; a.fix = fix_[3];
  a.lda = pld_[3]; a.nullrow = fac[3].pnull;
  a.q = q; a.cf = cf; a.s = s; a.p = p; a.y = y; a.S = d.S;
  a.wall_bot = wall_[slot][0]; a.wall_top = wall_[slot][1];
  a.wall2_bot = wall2_[slot][0]; a.wall2_top = wall2_[slot][1];
  a.open_top = opentop_ ? 1 : 0;
  if (count) launch(KC_PCHAIN, pressure_chain_flops(d.nz, d.P), [&] { pressure_chain(a, st_); });
  else pressure_chain(a, st_);
}

void Context::solve(int mode, int n, const int* k, const double* sigma, const double2* const* x,
                    double2* const* y) {
  HArgs h;
  h.nz = d.nz; h.lda = lda_; h.P = (int)d.P; h.nyc = d.nyc; h.mode = mode; h.kx = kx_; h.ky = ky_;
  int nplain = 0;
  for (int i = 0; i < n; ++i) {
    const int f = k[i];
    const int nullrow = sigma[i] == 0.0 ? (hsplit_[f] ? fac[f].pnull : fac[f].nullrow) : -1;
    if (hsplit_[f]) {
      HArgs hs = h;
      hs.lda = pld_[f];
      HField& q = hs.f[0];
      q.vinv = pinv_[f]; q.vnorm = pnorm_[f]; q.lambda = plam_[f]; q.fix = fix_[f]; q.pfix = pfix_[f];
      q.sigma = sigma[i]; q.nullrow = nullrow; q.x = x[i]; q.y = y[i];
      launch(KC_HELMS, helmholtz_split_flops(d.nz, d.P), [&] { helmholtz_split(hs, st_); });
    } else {
      HField& q = h.f[nplain++];
      q.vinv = vinv_[f]; q.vnorm = vnorm_[f]; q.lambda = lam_[f]; q.fix = fix_[f];
      q.sigma = sigma[i]; q.nullrow = nullrow; q.x = x[i]; q.y = y[i];
    }
  }
  if (nplain > 0) launch(KC_HELM, 2 * nplain * zflops(d), [&] { helmholtz(h, nplain, st_); });
}

template <typename T, typename SpecT>
void Context::ypb(const SpecT* spec, long long sstride, C2<T>* tmp, int nv, int zoff, int nplanes, int scol,
                  const int* gate) {
  // a gated launch may do nothing: its work is not counted
  const double b = gate ? 0.0 : (double)nv * nplanes * d.mx * (d.nyc * sizeof(SpecT) + (double)d.ny * sizeof(C2<T>));
  launch(is32<T>() ? KC_YB32 : KC_YB64, b,
         [&] { fft_ypass_backward<T, SpecT>(fft_, spec, sstride, tmp, nv, zoff, nplanes, st_, scol, gate); });
}
template <typename T>
void Context::xpb(const C2<T>* tmp, T* phys, long long pstride, int nv, int nplanes, const int* gate) {
  const double b = gate ? 0.0 : (double)nv * nplanes * d.ny * (d.mx * sizeof(C2<T>) + (double)d.nx * sizeof(T));
  launch(is32<T>() ? KC_XB32 : KC_XB64, b,
         [&] { fft_xpass_backward<T>(fft_, tmp, phys, pstride, nv, nplanes, st_, gate); });
}
template <typename T>
void Context::xpf(const T* phys, long long pstride, C2<T>* tmp, int nv, int nplanes, const int* gate) {
  // a gated launch may do nothing: its work is not counted
  const double b = gate ? 0.0 : (double)nv * nplanes * d.ny * (d.mx * sizeof(C2<T>) + (double)d.nx * sizeof(T));
  launch(is32<T>() ? KC_XF32 : KC_XF64, b,
         [&] { fft_xpass_forward<T>(fft_, phys, pstride, tmp, nv, nplanes, st_, gate); });
}
template <typename T, typename SpecT>
void Context::ypf(const C2<T>* tmp, SpecT* spec, long long sstride, int nv, int zoff, int nplanes, int scol,
                  const int* gate) {
  const double b = gate ? 0.0 : (double)nv * nplanes * d.mx * (d.nyc * sizeof(SpecT) + (double)d.ny * sizeof(C2<T>));
  launch(is32<T>() ? KC_YF32 : KC_YF64, b,
         [&] { fft_ypass_forward<T, SpecT>(fft_, tmp, spec, sstride, nv, zoff, nplanes, st_, scol, gate); });
}

void* Context::alloc(size_t bytes) {
  void* p = nullptr;
  TBD_HIP(hipMalloc(&p, std::max<size_t>(bytes, 256)));
  allocs_.push_back(p);
  fill_zero(p, std::max<size_t>(bytes, 256), st_);
  return p;
}

// row-major copy of a column-major n*n matrix with leading dimension lda
static std::vector<double> rowmajor(const std::vector<double>& cm, int n, int lda) {
  std::vector<double> r((size_t)n * lda, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) r[(size_t)i * lda + j] = cm[i + (size_t)n * j];
  return r;
}

Context::Context(const tbd_config& c) : cfg(c) {
  TBD_CHECK(cfg.nx >= 2 && cfg.ny >= 2 && cfg.nz >= 4, "Context > grid too small");
  TBD_CHECK(std::fabs(cfg.lz - 2) <= 10E-12,
            "Domain(mpi_t,dcp_t,cfg_t) > Dimensionless z-length Should be Equal to Two!");
  TBD_CHECK(cfg.nphase == 1 || cfg.nphase == 2, "Context > total_phase_number must be 1 or 2");
  TBD_CHECK(cfg.precision == 0 || cfg.precision == 1, "Context > precision must be 0 (FP64) or 1 (mixed)");
  d = Dims::make(cfg.nx, cfg.ny, cfg.nz);
  xpb_ = (double)d.nz * d.mx * d.ny;
  TBD_CHECK(d.nyc >= 2 && d.mx >= 1, "Context > grid too small for the 3/2 rule");
  TBD_HIP(hipStreamCreate(&st_));
  TBD_HIP(hipStreamCreate(&st2_));
  TBD_HIP(hipEventCreateWithFlags(&evf_, hipEventDisableTiming));
  TBD_HIP(hipEventCreateWithFlags(&evj_, hipEventDisableTiming));
  TBD_HIP(hipStreamCreate(&st3_));
  TBD_HIP(hipEventCreateWithFlags(&evf3_, hipEventDisableTiming));
  TBD_HIP(hipEventCreateWithFlags(&evj3_, hipEventDisableTiming));
  if (const char* e = std::getenv("TBD_NO_OVERLAP")) {
    overlap_ = std::string(e) != "1";
    pin_overlap_ = true;
  } else {
    overlap_ = true;
  }
  fft_.create(d);
  if (const char* e = std::getenv("TBD_FFT_WIDE")) {
    fft_.wide_yb = fft_.wide_yf = fft_.wide_xb = fft_.wide_xf = std::string(e) == "1";
    pin_wide_ = true;
  }
  if (const char* e = std::getenv("TBD_FUSED")) {
    fused_ = std::string(e) == "1";
    pin_fused_ = true;
  }
  if (const char* e = std::getenv("TBD_RHS_UNFUSED")) {
    want_rhs_ = !(e[0] == '1');
    pin_rhs_ = true;
  }
  if (const char* e = std::getenv("TBD_FUSED_DX")) fused_dx_ = std::string(e) != "0";
  dt_ = olddt_ = cfg.dt;
  const double c1[3] = {0.0 / 1.0, -5.0 / 9.0, -153.0 / 128.0};    // stepper_time_rk3.cc:29-31
  const double c2[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
  const double c3[3] = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};
  for (int s = 0; s < 3; ++s) { coef_[0][s] = c1[s]; coef_[1][s] = c2[s]; coef_[2][s] = c3[s]; }

  // operators
  ch = make_chebyshev(cfg.nz, cfg.mapping, cfg.mapping_coefficient);
  build_factorisations();
  opentop_ = std::fabs(cfg.vt_u1[2] - 0) <= 10E-12 && std::fabs(cfg.vt_u1[3] - 1) <= 10E-12;   // boundary.cc:59-71

  install_operators();
  const int nz = d.nz;
  auto up = [&](const std::vector<double>& h) {
    double* p = (double*)alloc(sizeof(double) * h.size());
    TBD_HIP(hipMemcpy(p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return p;
  };
  // Fourier multipliers with the reference's sign at iy = nyc/2 (stepper_space.cc:181-213)
  std::vector<double> kx(d.mx), ky(d.nyc);
  const double kx1 = 2 * kPi / cfg.lx, ky1 = 2 * kPi / cfg.ly;
  for (int ix = 0; ix < d.mx; ++ix) kx[ix] = kx1 * (ix - 0);
  for (int iy = 0; iy < d.nyc; ++iy) ky[iy] = (iy <= d.nyc / 2) ? +ky1 * (iy - 0) : -ky1 * (d.nyc - iy);
  kx_ = up(kx);
  ky_ = up(ky);
  gridz_ = up(ch.gridz);
  // grid lengths (domain.cc:304-332)
  std::vector<double> gx(d.nx), gy(d.ny);
  for (int i = 0; i < d.nx; ++i) gx[i] = cfg.lx / d.nx * i;
  for (int i = 0; i < d.ny; ++i) gy[i] = cfg.ly / d.ny * i;
  glx_ = 1e300;
  gly_ = 1e300;
  for (int i = 0; i + 1 < d.nx; ++i) glx_ = std::min(glx_, gx[i + 1] - gx[i]);
  for (int i = 0; i + 1 < d.ny; ++i) gly_ = std::min(gly_, gy[i + 1] - gy[i]);
  std::vector<double> glz(nz);
  const std::vector<double>& g = ch.gridzmap;
  for (int i = 0; i < nz; ++i) {
    if (i == 0) glz[i] = (g[1] - g[0]) / 2;
    else if (i == nz - 1) glz[i] = (g[nz - 1] - g[nz - 2]) / 2;
    else glz[i] = std::min(g[i + 1] - g[i], g[i] - g[i - 1]);
  }
  glz_ = up(glz);

  // buffers
  const size_t S16 = sizeof(double2) * d.S;
  const size_t ts = cfg.precision == 1 ? sizeof(float) : sizeof(double);
  const size_t Sc = 2 * ts * d.S;     // spectral component in the transform precision
  const size_t Rt = ts * d.R;
  uvw_ = (double2*)alloc(3 * S16);
  p_ = (double2*)alloc(S16);
  huvw_ = (double2*)alloc(3 * S16);
  suvw_ = (double2*)alloc(3 * S16);
  rhs_ = (double2*)alloc(3 * S16);
  rhsp_ = (double2*)alloc(S16);
  grad_ = alloc(9 * Sc);
  adv_ = alloc(3 * Sc);
  prod_ = alloc(6 * Sc);
  uvwnl_ = alloc(3 * Rt);
  gphys_ = alloc(9 * Rt);
  aphys_ = alloc(3 * Rt);
  tmp_ = alloc(std::max((size_t)9 * 2 * ts, sizeof(double2)) * fft_tmp_elems(d, 1));
  uxp_ = alloc((size_t)3 * 2 * ts * fft_tmp_elems(d, 1));
  for (int s = 0; s < 2; ++s)
    for (int b = 0; b < 2; ++b) {
      wall_[s][b] = (double2*)alloc(sizeof(double2) * 3 * d.P);
      wall2_[s][b] = (double2*)alloc(sizeof(double2) * 2 * d.P);
    }
  if (cfg.nphase > 1) {
    c_ = (double2*)alloc(S16);
    hc_ = (double2*)alloc(S16);
    rhsc_ = (double2*)alloc(S16);
    advc_ = alloc(Sc);
    cphys_ = (double*)alloc(sizeof(double) * d.R);
    const size_t plane = (size_t)d.nx * d.ny;
    for (int k = 0; k < 2; ++k) {
      dudz_bot_[k] = alloc(ts * plane);
      dvdz_bot_[k] = alloc(ts * plane);
    }
    sgrad_ = alloc(3 * Sc);
    sprod_ = alloc(3 * Sc);
    sgphys_ = alloc(3 * Rt);
    saphys_ = alloc(Rt);
    stmp_ = alloc(std::max((size_t)3 * 2 * ts, sizeof(double2)) * fft_tmp_elems(d, 1));
    ephys_ = (double*)alloc(sizeof(double) * plane);
    ehat_ = (double2*)alloc(sizeof(double2) * d.P);
    wrows_ = (double2*)alloc(sizeof(double2) * 4 * d.P);
  }
  fl_ = (double*)alloc(sizeof(double) * 64);
  fg_ = (double*)alloc(sizeof(double) * 64);
  wallmean_ = (double*)alloc(sizeof(double) * 8);
  diagpart_ = (double*)alloc(sizeof(double) * 10 * kDiagMaxBlocks);
  force_ = (double*)alloc(sizeof(double) * 2);
  filt_ = (int*)alloc(sizeof(int) * 2);
  need_ = (int*)alloc(sizeof(int) * 2);
  mmpart_ = (double*)alloc(sizeof(double) * 1024);
  stats_ = (double*)alloc(sizeof(double) * kStatBins * kPhaseStats * d.nz);
  if (const char* e = std::getenv("TBD_GRAPH")) use_graphs_ = std::string(e) != "0";
  TBD_HIP(hipStreamSynchronize(st_));
}

Context::~Context() {
  if (st_) (void)hipStreamSynchronize(st_);
  if (st2_) (void)hipStreamSynchronize(st2_);
  if (st3_) (void)hipStreamSynchronize(st3_);
  drop_graphs();
  for (void* p : allocs_) (void)hipFree(p);
  for (void* p : opallocs_) (void)hipFree(p);
  fft_.destroy();
  if (evf_) (void)hipEventDestroy(evf_);
  if (evj_) (void)hipEventDestroy(evj_);
  if (evf3_) (void)hipEventDestroy(evf3_);
  if (evj3_) (void)hipEventDestroy(evj3_);
  if (st2_) (void)hipStreamDestroy(st2_);
  if (st3_) (void)hipStreamDestroy(st3_);
  if (st_) (void)hipStreamDestroy(st_);
}

void Context::sync() { TBD_HIP(hipStreamSynchronize(st_)); }

// Time mode: the next step runs with dt.  Captured steps are keyed by
// (dt, previous dt), so none is dropped here; step() replays, captures or
// runs eagerly as the pair demands.
void Context::set_dt(double dt) {
  TBD_CHECK(dt > 0 && std::isfinite(dt), "Context::set_dt > dt must be positive and finite");
  dt_ = dt;
}

// Boundary-condition sets of the factorisations: u, v, w, p and (two phases)
// the scalar's Robin rows 1e7*(Re Sc ws, 1) on both walls (phase_scalar_flat.cc:194-200)
void Context::build_factorisations() {
  fac[0] = make_factorisation(ch, cfg.vt_u1);
  fac[1] = std::equal(cfg.vt_u1, cfg.vt_u1 + 4, cfg.vt_u2) ? fac[0] : make_factorisation(ch, cfg.vt_u2);
  for (int i = 0; i < 5; ++i) fac[1].vt[i] = cfg.vt_u2[i];
  fac[2] = make_factorisation(ch, cfg.vt_u3);
  fac[3] = make_factorisation(ch, cfg.vt_p);
  if (cfg.nphase > 1) {
    double bc[5] = {cfg.reynolds * cfg.schmidt * cfg.settling_velocity, 1,
                    cfg.reynolds * cfg.schmidt * cfg.settling_velocity, 1, 0};
    for (int i = 0; i < 4; ++i) bc[i] = bc[i] * 10E+6;
    fac[4] = make_factorisation(ch, bc);
  }
}

void* Context::alloc_op(size_t bytes) {
  void* p = nullptr;
  TBD_HIP(hipMalloc(&p, std::max<size_t>(bytes, 256)));
  opallocs_.push_back(p);
  return p;
}

// Device copies of D1, D2 and the factorisations with their derived forms
// (even/odd splits, permuted V~^-1).  A repeated call (the operators were
// replaced) waits for the device and frees the previous copies first.
void Context::install_operators() {
  if (!opallocs_.empty()) {
    TBD_HIP(hipDeviceSynchronize());
    for (void* q : opallocs_) (void)hipFree(q);
    opallocs_.clear();
  }
  const int nz = d.nz;
  lda_ = (nz + 3) / 4 * 4;
  auto up = [&](const std::vector<double>& h) {
    double* p = (double*)alloc_op(sizeof(double) * h.size());
    TBD_HIP(hipMemcpy(p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return p;
  };
  d1_ = up(rowmajor(ch.d1, nz, lda_));
  d2_ = up(rowmajor(ch.d2, nz, lda_));
  {
    std::vector<double> r = rowmajor(ch.d1, nz, lda_);
    std::vector<float> f(r.begin(), r.end());
    d1f_ = (float*)alloc_op(sizeof(float) * f.size());
    TBD_HIP(hipMemcpy(d1f_, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
  }
  // D1 is centro-antisymmetric and D2 centro-symmetric on the unmapped CGL
  // grid: apply them through their even/odd split (half the MFMA work)
  {
    const std::vector<double> r1 = rowmajor(ch.d1, nz, lda_), r2 = rowmajor(ch.d2, nz, lda_);
    auto centro = [&](const std::vector<double>& A, double sgn) {
      double mx = 0, dv = 0;
      for (int i = 0; i < nz; ++i)
        for (int j = 0; j < nz; ++j) {
          mx = std::max(mx, std::fabs(A[(size_t)i * lda_ + j]));
          dv = std::max(dv, std::fabs(A[(size_t)i * lda_ + j] - sgn * A[(size_t)(nz - 1 - i) * lda_ + nz - 1 - j]));
        }
      return dv <= 1e-12 * mx;
    };
    const char* np = std::getenv("TBD_NO_PARITY");
    d1p_ = d2p_ = nullptr;
    d1pf_ = nullptr;
    parity_ = nz % 2 == 1 && !(np && np[0] == '1') && centro(r1, -1.0) && centro(r2, 1.0);
    if (parity_) {
      d1p_ = up(pack_parity<double>(r1.data(), nz, lda_, &ldp_));
      d2p_ = up(pack_parity<double>(r2.data(), nz, lda_, &ldp_));
      const std::vector<float> f = pack_parity<float>(r1.data(), nz, lda_, &ldp_);
      d1pf_ = (float*)alloc_op(sizeof(float) * f.size());
      TBD_HIP(hipMemcpy(d1pf_, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
    }
  }
  const char* nps = std::getenv("TBD_NO_PARITY");
  for (int k = 0; k < (cfg.nphase > 1 ? 5 : 4); ++k) {
    vinv_[k] = up(rowmajor(fac[k].vinv, nz, lda_));
    // the full-form solves read V~ with the backward fixes folded in
    vnorm_[k] = up(rowmajor(fold_backward_fixes(fac[k].vnorm, fac[k].fix, nz), nz, lda_));
    lam_[k] = up(fac[k].lambda);
    fix_[k] = up(fac[k].fix);
    vinvp_[k] = nz % 2 == 1 ? up(permute_columns(rowmajor(fac[k].vinv, nz, lda_), nz, lda_)) : nullptr;
    hsplit_[k] = fac[k].pM > 0 && !(nps && nps[0] == '1');
    pinv_[k] = pnorm_[k] = plam_[k] = pfix_[k] = nullptr;
    if (hsplit_[k]) {
      pld_[k] = fac[k].pld;
      pinv_[k] = up(fac[k].pinv);
      pnorm_[k] = up(fac[k].pnorm);
      plam_[k] = up(fac[k].plam);
      pfix_[k] = up(fac[k].pfix);
    }
  }
  fuse_rhs_ = parity_ && nz % 2 == 1 && nz >= 7 && want_rhs_;
  fuse_p_ = fuse_rhs_ && hsplit_[3];
}

void Context::copy_operators_from(const Context& o) {
  TBD_CHECK(o.d.nz == d.nz, "Context::copy_operators_from > different nz");
  ch = o.ch;
  for (int k = 0; k < 5; ++k) fac[k] = o.fac[k];
  install_operators();
  drop_graphs();
}

Context::Variant Context::variant() const {
  Variant v;
  v.fused = fused_;
  v.rhs = want_rhs_;
  v.overlap = overlap_;
  v.wide[0] = fft_.wide_yb;
  v.wide[1] = fft_.wide_yf;
  v.wide[2] = fft_.wide_xb;
  v.wide[3] = fft_.wide_xf;
  v.swap[0] = fft_.swap_yb;
  v.swap[1] = fft_.swap_yf;
  v.swap[2] = fft_.swap_xb;
  v.swap[3] = fft_.swap_xf;
  v.zwide = zwide_;
  return v;
}

void Context::set_variant(const Variant& v) {
  sync();
  if (!pin_fused_) fused_ = v.fused;
  if (!pin_rhs_) want_rhs_ = v.rhs;
  if (!pin_overlap_) overlap_ = v.overlap;
  if (!pin_wide_) {
    fft_.wide_yb = v.wide[0];
    fft_.wide_yf = v.wide[1];
    fft_.wide_xb = v.wide[2];
    fft_.wide_xf = v.wide[3];
  }
  // a swapped radix order exists only for some lengths; elsewhere the bit
  // has no kernel to select and stays clear
  fft_.swap_yb = v.swap[0] && fft_has_swap(d.ny);
  fft_.swap_yf = v.swap[1] && fft_has_swap(d.ny);
  fft_.swap_xb = v.swap[2] && fft_has_swap(d.nx);
  fft_.swap_xf = v.swap[3] && fft_has_swap(d.nx);
  zwide_ = v.zwide;
  fuse_rhs_ = parity_ && d.nz % 2 == 1 && d.nz >= 7 && want_rhs_;
  fuse_p_ = fuse_rhs_ && hsplit_[3];
  drop_graphs();
}

int Context::autotune(int steps, double* ms8) {
  TBD_CHECK(steps >= 1 && steps <= 10000, "tbd_autotune > steps out of range");
  for (int i = 0; i < kVariants; ++i) ms8[i] = std::nan("");
  // this context's state, uploaded to the scratch context before each variant
  const size_t n2 = 2 * (size_t)d.S;
  std::vector<double> u(3 * n2), p(n2), c(cfg.nphase > 1 ? n2 : 0);
  download(u.data(), p.data(), c.empty() ? nullptr : c.data());
  Context scratch(cfg);
  scratch.pin_fused_ = scratch.pin_rhs_ = scratch.pin_overlap_ = scratch.pin_wide_ = false;   // the loop below decides
  scratch.copy_operators_from(*this);
  scratch.set_dt(dt_);
  const bool rhs_ok = parity_ && d.nz % 2 == 1 && d.nz >= 7;
  const Variant cur = variant();
  int best = cur.bits();
  double best_ms = 0.0;
  std::vector<double> feat(50);
  auto admissible = [&](const Variant& v) {
    if ((pin_fused_ && v.fused != cur.fused) || (pin_rhs_ && v.rhs != cur.rhs) ||
        (pin_overlap_ && v.overlap != cur.overlap))
      return false;
    for (int i = 0; i < 4; ++i) {
      if (pin_wide_ && v.wide[i] != cur.wide[i]) return false;
      // the wide and swapped kernels exist only for the two-level transform
      // family, the swapped ones for lengths with two distinct factors
      if (v.wide[i] && !fft_.two_level) return false;
      if (v.swap[i] && (!fft_.two_level || !fft_has_swap(i < 2 ? d.ny : d.nx))) return false;
    }
    // the fused right-hand side needs the even/odd operators; without them
    // both settings run the same kernels
    return rhs_ok || v.rhs == cur.rhs;
  };
  auto time_variant = [&](const Variant& v) {
    scratch.set_variant(v);
    scratch.upload(u.data(), p.data(), c.empty() ? nullptr : c.data());
    double time = 0.0;
    int alt = 0;
    auto one = [&] {
      time += dt_;
      scratch.step(time, alt);
      alt ^= 1;
      scratch.features(feat.data());
    };
    for (int i = 0; i < 3; ++i) one();
    scratch.sync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; ++i) one();
    scratch.sync();
    const auto t1 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
    ms8[v.bits()] = ms;
    if (best_ms == 0.0 || ms < best_ms) {
      best_ms = ms;
      best = v.bits();
    }
  };
  // every combination of the three step-level switches at the current widths
  for (int low = 0; low < 8; ++low) {
    Variant v = Variant::of(low);
    for (int i = 0; i < 4; ++i) {
      v.wide[i] = cur.wide[i];
      v.swap[i] = cur.swap[i];
    }
    if (admissible(v)) time_variant(v);
  }
  // then each transform width in turn, from the best so far
  for (int i = 0; i < 4; ++i) {
    Variant v = Variant::of(best);
    v.wide[i] = !v.wide[i];
    if (admissible(v)) time_variant(v);
  }
  // then each pass's radix order
  for (int i = 0; i < 4; ++i) {
    Variant v = Variant::of(best);
    v.swap[i] = !v.swap[i];
    if (admissible(v)) time_variant(v);
  }
  // then the panel width of the plain contractions
  {
    Variant v = Variant::of(best);
    v.zwide = !v.zwide;
    time_variant(v);
  }
  set_variant(Variant::of(best));
  return variant().bits();
}

void Context::set_operators(const double* d1, const double* d2) {
  const size_t n = (size_t)d.nz * d.nz;
  for (size_t i = 0; i < n; ++i)
    TBD_CHECK(std::isfinite(d1[i]) && std::isfinite(d2[i]), "tbd_set_operators > operators are not finite");
  ch.d1.assign(d1, d1 + n);
  ch.d2.assign(d2, d2 + n);
  build_factorisations();
  install_operators();
  drop_graphs();
}

void Context::set_factorisation(int which, const double* V, const double* Vinv, const double* lambda,
                                const double* fix) {
  TBD_CHECK(which >= 0 && which < (cfg.nphase > 1 ? 5 : 4), "tbd_set_factorisation > bad index");
  fac[which] = factorisation_from(d.nz, fac[which].vt, V, Vinv, lambda, fix);
  install_operators();
  drop_graphs();
}

// Destroys the captured steps.  A replay may still be running on the
// context's streams (a step returns before its graph has finished), so the
// streams are drained first: an exec is never destroyed while in flight.
// (Errors of the drain are left to the next checked call: this also runs in
// the destructor.)
void Context::drop_graphs() {
  if (graphs_.empty()) return;
  for (hipStream_t s : {st_, st2_, st3_})
    if (s) (void)hipStreamSynchronize(s);
  for (auto& g : graphs_) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  graphs_.clear();
}

// One RK3 step.  Outside profiling the three substages are captured once per
// (alternation, wall-row slot, dt, previous dt) into a HIP graph and replayed;
// the only per-step input, the mean-mode force, is written to device memory
// before the launch.  The host-side state a step advances (wall-row slot,
// previous dt) is advanced the same way as by the eager substages.  A step
// whose (dt, previous dt) differs from the last step's runs eagerly: in time
// mode dt moves almost every step, and such a step's graph would not be
// replayed.  At most kMaxGraphs captured steps are kept.
void Context::step(double time, int alternation) {
  const bool moved = last_dt_ >= 0 && (dt_ != last_dt_ || olddt_ != last_olddt_);
  last_dt_ = dt_;
  last_olddt_ = olddt_;
  if (!use_graphs_ || prof.on || moved) {
    stages(time, alternation);
    return;
  }
  double f1 = 0, f2 = 0;
  forces(time, f1, f2);
  set_pair(force_, f1, f2, st_);
  const int slot0 = wslot_;
  for (auto& g : graphs_) {
    if (g.alternation == alternation && g.slot == slot0 && g.dt == dt_ && g.olddt == olddt_) {
      TBD_HIP(hipGraphLaunch(g.exec, st_));
      wslot_ = 1 - wslot_;   // three toggles
      olddt_ = dt_;
      return;
    }
  }
  if (graphs_.size() >= kMaxGraphs) drop_graphs();
  StepGraph g;
  g.alternation = alternation;
  g.slot = slot0;
  g.dt = dt_;
  g.olddt = olddt_;
  capturing_ = true;
  TBD_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  try {
    stages(time, alternation);
  } catch (...) {
    capturing_ = false;
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(st_, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  capturing_ = false;
  TBD_HIP(hipStreamEndCapture(st_, &g.graph));
  TBD_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  graphs_.push_back(g);
  TBD_HIP(hipGraphLaunch(g.exec, st_));
}

void Context::record_step(double time, int alternation, Recorder& r) {
  TBD_CHECK(!prof.on, "Context::record_step > not while profiling");
  TBD_CHECK(Recorder::active() == nullptr, "Context::record_step > a recording is already active on this thread");
  last_dt_ = dt_;
  last_olddt_ = olddt_;
  // a capture on the context's own stream catches any work not routed
  // through the recorder: it must come out empty
  TBD_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  Recorder::active() = &r;
  capturing_ = true;
  try {
    stages(time, alternation);
  } catch (...) {
    Recorder::active() = nullptr;
    capturing_ = false;
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(st_, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  Recorder::active() = nullptr;
  capturing_ = false;
  hipGraph_t g = nullptr;
  TBD_HIP(hipStreamEndCapture(st_, &g));
  size_t nodes = 0;
  const hipError_t e = hipGraphGetNodes(g, nullptr, &nodes);
  (void)hipGraphDestroy(g);
  TBD_HIP(e);
  TBD_CHECK(nodes == 0, "Context::record_step > the step issued work outside the recorded launches "
                        "(LDS Stockham transforms or a second stream)");
}

void Context::replayed_step() {
  last_dt_ = dt_;
  last_olddt_ = olddt_;
  wslot_ = 1 - wslot_;   // three toggles
  olddt_ = dt_;
}

void Context::upload(const double* uvw, const double* p, const double* c) {
  const size_t S16 = sizeof(double2) * d.S;
  TBD_HIP(hipMemcpyAsync(uvw_, uvw, 3 * S16, hipMemcpyHostToDevice, st_));
  TBD_HIP(hipMemcpyAsync(p_, p, S16, hipMemcpyHostToDevice, st_));
  if (cfg.nphase > 1) {
    TBD_CHECK(c != nullptr, "FlatScalarVariable::Load > concentration missing");
    TBD_HIP(hipMemcpyAsync(c_, c, S16, hipMemcpyHostToDevice, st_));
    fill_zero(hc_, S16, st_);
  }
  fill_zero(huvw_, 3 * S16, st_);
  for (int s = 0; s < 2; ++s)
    for (int b = 0; b < 2; ++b) {
      fill_zero(wall_[s][b], sizeof(double2) * 3 * d.P, st_);
      fill_zero(wall2_[s][b], sizeof(double2) * 2 * d.P, st_);
    }
  olddt_ = dt_;
  // uvwnonlinear = DftBackward(uvw) (phase_carrier_flat.cc:233-234)
  if (cfg.precision == 1)
    fft_backward<float, double2>(fft_, uvw_, d.S, (float*)uvwnl_, d.R, 3, 0, d.nz, (float2*)uxp_, st_);
  else
    fft_backward<double, double2>(fft_, uvw_, d.S, (double*)uvwnl_, d.R, 3, 0, d.nz, (double2*)uxp_, st_);
  sync();
}

void Context::download(double* uvw, double* p, double* c) {
  sync();
  const size_t S16 = sizeof(double2) * d.S;
  if (uvw) TBD_HIP(hipMemcpy(uvw, uvw_, 3 * S16, hipMemcpyDeviceToHost));
  if (p) TBD_HIP(hipMemcpy(p, p_, S16, hipMemcpyDeviceToHost));
  if (c && cfg.nphase > 1) TBD_HIP(hipMemcpy(c, c_, S16, hipMemcpyDeviceToHost));
}

void Context::download_physical_velocity(double* out) {
  sync();
  if (cfg.precision == 1) {
    std::vector<float> h(3 * d.R);
    TBD_HIP(hipMemcpy(h.data(), uvwnl_, sizeof(float) * 3 * d.R, hipMemcpyDeviceToHost));
    for (long long i = 0; i < 3 * d.R; ++i) out[i] = h[i];
  } else {
    TBD_HIP(hipMemcpy(out, uvwnl_, sizeof(double) * 3 * d.R, hipMemcpyDeviceToHost));
  }
}

void Context::features(double* out50) {
  sync();
  TBD_HIP(hipMemcpy(out50, fg_, sizeof(double) * 50, hipMemcpyDeviceToHost));
}

// driven force on the mean mode (phase_carrier_flat.cc:360-502)
void Context::forces(double time, double& f1, double& f2) const {
  f1 = f2 = 0;
  const double t = time;
  const double dt = dt_;
  const double tm = cfg.output_on_step ? (t - dt / 2) : (t + dt / 2);
  const int d0 = cfg.force_direction[0], d1 = cfg.force_direction[1];
  const double* a = cfg.force_amplitude;
  switch (cfg.force_type) {
    case 0: break;
    case 1:
      if (d0 == 1) f1 = a[0]; else if (d0 == 2) f2 = a[0];
      break;
    case 2: {
      const double f = a[1];
      if (d1 == 1) f1 = f * std::cos(f * tm); else if (d1 == 2) f2 = f * std::cos(f * tm);
    } break;
    case 3: {
      if (d0 == 1) f1 = a[0]; else if (d0 == 2) f2 = a[0];
      const double f = a[1];
      if (d1 == 1) f1 = f1 + f * std::cos(f * tm); else if (d1 == 2) f2 = f2 + f * std::cos(f * tm);
    } break;
    case 4: {
      if (d0 == 1) f1 = a[0]; else if (d0 == 2) f2 = a[0];
      const double f = a[1];
      const double T = a[0], t0 = a[1], last = a[2], tar = a[3];
      const double damp = -std::log(tar) / (last + 0.25) / T;
      const double gv = (f * std::cos(f * tm) - damp * std::sin(f * tm)) * std::exp(-damp * (t - t0));
      if (d1 == 1) f1 = f1 + gv; else if (d1 == 2) f2 = f2 + gv;
    } break;
    default:
      throw Error("FlatCarrierPhase::Predict() > Not Supported Force Type!");
  }
}

// wall-normal operators of a contraction, in the arithmetic type T
template <typename T>
void Context::set_d1(ZArgs& a) const {
  const bool f = std::is_same<T, float>::value;
  a.ncg_mult = zwide_ ? 2 : 1;
  if (parity_) {
    a.op = f ? (const void*)d1pf_ : (const void*)d1p_;
    a.lda = ldp_;
    a.parity = -1;
  } else {
    a.op = f ? (const void*)d1f_ : (const void*)d1_;
    a.lda = lda_;
    a.parity = 0;
  }
}
void Context::set_d2(ZArgs& a) const {
  a.ncg_mult = zwide_ ? 2 : 1;
  a.op = parity_ ? d2p_ : d2_;
  a.lda = parity_ ? ldp_ : lda_;
  a.parity = parity_ ? 1 : 0;
}

// ---------------------------------------------------------------------------
// FlatCarrierPhase::Predict (phase_carrier_flat.cc:272-559)
// ---------------------------------------------------------------------------
template <typename T>
void Context::carrier_predict(int stage, double time, int alternation) {
  typedef C2<T> CT;
  const double c1 = coef_[0][stage], c2 = coef_[1][stage], c3 = coef_[2][stage];
  const long long S = d.S, R = d.R;
  CT* grad = (CT*)grad_;
  CT* adv = (CT*)adv_;
  CT* prod = (CT*)prod_;
  T* uvwnl = (T*)uvwnl_;
  T* gphys = (T*)gphys_;
  T* aphys = (T*)aphys_;
  CT* tmp = (CT*)tmp_;

  // wall gradients of the mean mode for the log (dudxyz(2) rows 0, N-1 col 0)
  launch(KC_OTHER, 0.0, [&] { wall_mean_rows(d1_, lda_, d.nz, uvw_, S, wallmean_, st_); });
  if (!alternation) {
    // convective form (phase_carrier_flat.cc:784-806)
    // on the fused path with the register transforms, the x derivatives are
    // taken from the velocity's x-pencils (i kx commutes with the y pass):
    // the contraction writes only i ky u_i and D1 u_i, and six components
    // take the y pass instead of nine
    const bool dx_fused = fused_ && fused_dx_ && xfused_registers(fft_);
    ZArgs a;
    a.mode = dx_fused ? Z_GRADYZ : Z_GRAD; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<T>(a);
    a.x = uvw_; a.xs = S; a.y = grad; a.ys = S; a.kx = kx_; a.ky = ky_;
    launch(is32<T>() ? KC_Z32 : KC_Z64, 3 * zwork(), [&] { zgemm<T, double2, CT>(a, 3, st_); });
    const bool erosion_planes = cfg.nphase > 1 && cfg.scalar_bc == 1;
    if (fused_) {
      // y^-1 of the gradients; x^-1, u.grad u_i, x in one kernel; y
      ypb<T, CT>(grad, S, tmp, dx_fused ? 6 : 9, 0, d.nz);
      // outputs go to a separate x-pencil buffer: the three outputs of a row
      // tile are written by different workgroups while the inputs are read
      CT* xout = (CT*)gphys_;
      launch(is32<T>() ? KC_XFU32 : KC_XFU64, 12 * xpb_ * sizeof(CT) + 3.0 * R * sizeof(T), [&] {
        xfused_conv<T>(fft_, tmp, uvwnl, R, (T)0, (T)0, (T)0, 3, xout,
                       erosion_planes ? (T*)dudz_bot_[stage & 1] : nullptr,
                       erosion_planes ? (T*)dvdz_bot_[stage & 1] : nullptr, st_,
                       dx_fused ? (const CT*)uxp_ : nullptr, dx_fused ? 2 * kPi / cfg.lx : 0.0);
      });
      ypf<T, CT>(xout, adv, S, 3, 0, d.nz);
    } else {
      bwd<T, CT>(grad, S, gphys, R, 9, 0, d.nz, tmp);
      launch(is32<T>() ? KC_PROD32 : KC_PROD64, 15.0 * R * sizeof(T), [&] { products_conv<T>(uvwnl, gphys, aphys, R, st_); });
      fwd<T, CT>(aphys, R, adv, S, 3, 0, d.nz, tmp);
      if (erosion_planes) {
        const size_t pl = (size_t)d.nx * d.ny * sizeof(T);
        copy_words(dudz_bot_[stage & 1], gphys + 2 * R, pl, st_);
        copy_words(dvdz_bot_[stage & 1], gphys + 5 * R, pl, st_);
      }
    }
  } else {
    // divergence form (phase_carrier_flat.cc:725-779): six distinct products
    if (fused_) {
      launch(is32<T>() ? KC_XFU32 : KC_XFU64, 6 * xpb_ * sizeof(CT) + 3.0 * R * sizeof(T), [&] { xfused_div6<T>(fft_, uvwnl, R, tmp, st_); });
      ypf<T, CT>(tmp, prod, S, 6, 0, d.nz);
    } else {
      launch(is32<T>() ? KC_PROD32 : KC_PROD64, 9.0 * R * sizeof(T), [&] { products_div<T>(uvwnl, gphys, R, st_); });
      fwd<T, CT>(gphys, R, prod, S, 6, 0, d.nz, tmp);
    }
    // adv_i = (i kx F_i0 + i ky F_i1) + D1 F_i2 ; rows (uu,uv,uw) (uv,vv,vw) (uw,vw,ww)
    const int rows[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    {
      ZArgs a;
      a.mode = Z_GRADTMP; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<T>(a);
      for (int f = 0; f < 3; ++f) {
        a.xf[f] = prod + rows[f][2] * S; a.e0f[f] = prod + rows[f][0] * S; a.e1f[f] = prod + rows[f][1] * S;
      }
      a.y = adv; a.ys = S; a.kx = kx_; a.ky = ky_;
      launch(is32<T>() ? KC_Z32 : KC_Z64, 3 * zwork(), [&] { zgemm<T, CT, CT>(a, 3, st_); });
    }
    // Gradient(u), Gradient(v) + DftBackward for the erosion map: only the
    // bottom plane of du/dz, dv/dz is read (phase_scalar_flat.cc:349-351)
    if (cfg.nphase > 1 && cfg.scalar_bc == 1) {
      launch(KC_OTHER, 0.0, [&] {
        zrows_wall(d1_, lda_, d.nz, uvw_, d.P, wrows_, wrows_ + d.P, st_);
        zrows_wall(d1_, lda_, d.nz, uvw_ + S, d.P, wrows_ + 2 * d.P, wrows_ + 3 * d.P, st_);
      });
      const long long plane = (long long)d.nx * d.ny;
      bwd<T, double2>(wrows_, 2 * d.P, (T*)dudz_bot_[stage & 1], plane, 1, 0, 1, tmp, 1);
      bwd<T, double2>(wrows_ + 2 * d.P, 2 * d.P, (T*)dvdz_bot_[stage & 1], plane, 1, 0, 1, tmp, 1);
    }
  }
  // H = c1 H - dt adv, buoyancy, mean-mode force (:294-504)
  HUpd u;
  u.c1 = c1; u.dt = dt_; u.nz = d.nz;
  u.buoy = (cfg.nphase > 1 && cfg.affected_by_scalar_phase) ? 1 : 0;
  u.K = dt_ * cfg.averaged_concentration / std::pow(cfg.froude, 2) * (cfg.specific_gravity - 1);
  u.sn = std::sin(cfg.slope); u.cs = std::cos(cfg.slope); u.slope_dir = cfg.slope_direction;
  forces(time, u.f1, u.f2);
  if (capturing_) u.fdev = force_;
  // on the fused path the update happens inside the right-hand-side kernel
  if (!fuse_rhs_) launch(KC_OTHER, 0.0, [&] { h_update<CT>(huvw_, adv, c_, u, 3, S, st_); });
  // lap = D2 u - kx^2 u - ky^2 u ; rhs = sigma (c2 H + u) - lap (:517-536)
  const double redtc3 = -cfg.reynolds / dt_ / c3;
  if (!fuse_rhs_) {
    ZArgs a;
    a.mode = Z_RHS; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d2(a);
    a.x = uvw_; a.xs = S; a.y = rhs_; a.ys = S; a.h = huvw_; a.hs = S; a.kx = kx_; a.ky = ky_;
    a.s1 = redtc3; a.s2 = redtc3 * c2;
    launch(KC_Z64, 3 * zwork(), [&] { zgemm<double, double2, double2>(a, 3, st_); });
  }
  // boundary rows from the extrapolated wall pressure gradient (boundary_carrier_flat.cc:185-246)
  {
    const double c3m0 = coef_[2][stage], c3m1 = coef_[2][(stage + 2) % 3], c3m2 = coef_[2][(stage + 1) % 3];
    double xm0 = 0, xm1 = 0, xm2 = 0;
    switch (stage) {
      case 0: xm0 = 2 * (0 + c3m0) * dt_; xm1 = 0; xm2 = 2 * (0 - c3m1) * olddt_; break;
      case 1: xm0 = 2 * (c3m0 + c3m1 + 0) * dt_; xm1 = 2 * (0 + c3m1 + 0) * dt_; xm2 = 0; break;
      default: xm0 = 2 * (c3m0 + c3m1 + c3m2) * dt_; xm1 = 2 * (0 + c3m1 + c3m2) * dt_; xm2 = 2 * (0 + +0 + c3m2) * dt_;
    }
    const double bs = (xm0 - xm1) / (xm1 - xm2), bk = 2 * dt_ * c3m0;
    const double sg = -cfg.reynolds / dt_ / coef_[2][stage];
    if (fuse_rhs_) {
      // rhs, boundary rows and the three Helmholtz solves in one pass per field
      RHArgs r;
      r.s1 = redtc3; r.s2 = redtc3 * c2;
      r.bc = 0; r.s = bs; r.k = bk;
      r.m1b = wall_[wslot_][0]; r.m1t = wall_[wslot_][1];
      r.m2b = wall_[1 - wslot_][0]; r.m2t = wall_[1 - wslot_][1];
      r.n1t = wall2_[wslot_][1]; r.n2t = wall2_[1 - wslot_][1];
      r.open_top = opentop_ ? 1 : 0;
      r.top[0] = cfg.vt_u1[4]; r.top[1] = cfg.vt_u2[4]; r.top[2] = cfg.vt_u3[4];
      r.hupd = 1; r.adv32 = is32<T>() ? 1 : 0; r.c1 = u.c1; r.dt = u.dt;
      r.buoy = u.buoy; r.K = u.K; r.sn = u.sn; r.cs = u.cs; r.slope_dir = u.slope_dir; r.cb = c_;
      r.force = 1; r.f1 = u.f1; r.f2 = u.f2; r.fdev = u.fdev;
      const int k[3] = {0, 1, 2};
      const double2* x[3] = {uvw_, uvw_ + S, uvw_ + 2 * S};
      double2* h[3] = {huvw_, huvw_ + S, huvw_ + 2 * S};
      const void* av[3] = {adv, adv + S, adv + 2 * S};
      double2* y[3] = {suvw_, suvw_ + S, suvw_ + 2 * S};
      rhs_solve(r, 3, k, k, x, h, av, y, sg);
      return;
    }
    BcArgs b;
    b.nz = d.nz; b.P = d.P; b.S = S;
    b.s = bs;
    b.k = bk;
    b.m1b = wall_[wslot_][0]; b.m1t = wall_[wslot_][1];
    b.m2b = wall_[1 - wslot_][0]; b.m2t = wall_[1 - wslot_][1];
    b.n1t = wall2_[wslot_][1]; b.n2t = wall2_[1 - wslot_][1];
    b.open_top = opentop_ ? 1 : 0;
    b.top0 = cfg.vt_u1[4]; b.top1 = cfg.vt_u2[4]; b.top2 = cfg.vt_u3[4];
    launch(KC_OTHER, 0.0, [&] { bc_carrier(rhs_, b, st_); });
    // intermediate velocity: three Helmholtz solves (:554, :867-931)
    const int k[3] = {0, 1, 2};
    const double sigma[3] = {sg, sg, sg};
    const double2* x[3] = {rhs_, rhs_ + S, rhs_ + 2 * S};
    double2* y[3] = {suvw_, suvw_ + S, suvw_ + 2 * S};
    solve(0, 3, k, sigma, x, y);
  }
}

// ---------------------------------------------------------------------------
// FlatCarrierPhase::Correct (phase_carrier_flat.cc:566-709)
// ---------------------------------------------------------------------------
template <typename T>
void Context::carrier_correct(int stage) {
  const double c3 = coef_[2][stage];
  const long long S = d.S, R = d.R;
  if (fuse_p_) {
    // divergence, pressure solve, grad p, boundary Update and projection in
    // one pass over the spectral columns (:586-620)
    if (stage == 2) olddt_ = dt_;
    wslot_ = 1 - wslot_;
    pchain(0.5 / dt_ / c3, 2 * dt_ * c3, suvw_, p_, uvw_, wslot_, true);
  } else {
    // div s, rhs_p = div / (2 dt c3) (:586-589)
    {
      ZArgs a;
      a.mode = Z_DIV; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<double>(a);
      a.x = suvw_ + 2 * S; a.e0 = suvw_; a.e1 = suvw_ + S; a.y = rhsp_; a.kx = kx_; a.ky = ky_;
      a.q = 0.5 / dt_ / c3;
      launch(KC_Z64, zwork(), [&] { zgemm<double, double2, double2>(a, 1, st_); });
    }
    // pressure solve (:594, :955-987)
    {
      const int k = 3;
      const double sigma = 0.0;
      const double2* x = rhsp_;
      double2* y = p_;
      solve(1, 1, &k, &sigma, &x, &y);
    }
    // grad p, boundary Update (m2 <- m1, m1 <- walls), u = s - 2 dt c3 grad p,
    // mean w := 0 (:597-620, boundary_carrier_flat.cc:131-170)
    if (stage == 2) olddt_ = dt_;
    wslot_ = 1 - wslot_;
    {
      ZArgs a;
      a.mode = Z_CORRECT; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<double>(a);
      a.x = p_; a.e0 = suvw_; a.es = S; a.y = uvw_; a.ys = S; a.kx = kx_; a.ky = ky_;
      a.cf = 2 * dt_ * c3;
      a.wall_bot = wall_[wslot_][0]; a.wall_top = wall_[wslot_][1];
      a.wall2_bot = wall2_[wslot_][0]; a.wall2_top = wall2_[wslot_][1];
      a.open_top = opentop_ ? 1 : 0;
      launch(KC_Z64, zwork(), [&] { zgemm<double, double2, double2>(a, 1, st_); });
    }
  }
  // uvwnonlinear = DftBackward(uvw) (:625-626)
  bwd<T, double2>(uvw_, S, (T*)uvwnl_, R, 3, 0, d.nz, (C2<T>*)uxp_);
  if (stage == 2) {   // diagnostics (:633-703)
    fill_zero(fl_, sizeof(double) * 20, st_);
    launch(KC_OTHER, 0.0, [&] {
      diagnostics<T>((const T*)uvwnl_, R, d.nx, d.ny, d.nz, glx_, gly_, glz_, fl_, diagpart_, st_);
    });
    // the step's features: fl_, then 10..15: dU/dz, dV/dz, dW/dz on the mean
    // column at the bottom, then the top
    const int src[6] = {0, 2, 4, 1, 3, 5};
    collect_features(fg_, fl_, 50, wallmean_, src, st_);
  }
}

// ---------------------------------------------------------------------------
// FlatScalarPhase::Predict + Correct (phase_scalar_flat.cc:245-533)
// ---------------------------------------------------------------------------
template <typename T>
void Context::scalar_correct(int stage, int alternation) {
  typedef C2<T> CT;
  const double c1 = coef_[0][stage], c2 = coef_[1][stage], c3 = coef_[2][stage];
  const long long S = d.S, R = d.R;
  T* uvwnl = (T*)uvwnl_;
  // own work buffers (not the carrier's): see Context::overlap_
  T* gphys = (T*)sgphys_;
  T* aphys = (T*)saphys_;
  CT* grad = (CT*)sgrad_;
  CT* prod = (CT*)sprod_;
  CT* advc = (CT*)advc_;
  CT* tmp = (CT*)stmp_;
  // advecting velocity = carrier velocity + settling shift on the mean mode
  // (:256-279); its physical form is the carrier's plus constants
  const double ws = cfg.settling_velocity;
  const T su = (T)(cfg.slope_direction == 1 ? ws * std::sin(cfg.slope) : 0.0);
  const T sv = (T)(cfg.slope_direction == 2 ? ws * std::sin(cfg.slope) : 0.0);
  const T sw = (T)(-ws * std::cos(cfg.slope));
  if (alternation) {   // divergence form (:563-580)
    if (fused_) {
      ypb<T, double2>(c_, S, tmp, 1, 0, d.nz);
      launch(is32<T>() ? KC_XFU32 : KC_XFU64, 4 * xpb_ * sizeof(CT) + 3.0 * R * sizeof(T), [&] { xfused_cdiv<T>(fft_, tmp, uvwnl, R, su, sv, sw, tmp, st_); });
      ypf<T, CT>(tmp, prod, S, 3, 0, d.nz);
    } else {
      bwd<T, double2>(c_, S, aphys, R, 1, 0, d.nz, tmp);
      launch(is32<T>() ? KC_PROD32 : KC_PROD64, 7.0 * R * sizeof(T), [&] { scalar_products_div<T, T>(uvwnl, su, sv, sw, aphys, gphys, R, st_); });
      fwd<T, CT>(gphys, R, prod, S, 3, 0, d.nz, tmp);
    }
    ZArgs a;
    a.mode = Z_GRADTMP; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<T>(a);
    a.x = prod + 2 * S; a.e0 = prod; a.e1 = prod + S; a.y = advc; a.kx = kx_; a.ky = ky_;
    launch(is32<T>() ? KC_Z32 : KC_Z64, zwork(), [&] { zgemm<T, CT, CT>(a, 1, st_); });
  } else {             // convective form (:584-592)
    ZArgs a;
    a.mode = Z_GRAD; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<T>(a);
    a.x = c_; a.xs = S; a.y = grad; a.ys = S; a.kx = kx_; a.ky = ky_;
    launch(is32<T>() ? KC_Z32 : KC_Z64, zwork(), [&] { zgemm<T, double2, CT>(a, 1, st_); });
    if (fused_) {
      ypb<T, CT>(grad, S, tmp, 3, 0, d.nz);
      launch(is32<T>() ? KC_XFU32 : KC_XFU64, 4 * xpb_ * sizeof(CT) + 3.0 * R * sizeof(T), [&] {
        xfused_conv<T>(fft_, tmp, uvwnl, R, su, sv, sw, 1, tmp, nullptr, nullptr, st_);
      });
      ypf<T, CT>(tmp, advc, S, 1, 0, d.nz);
    } else {
      bwd<T, CT>(grad, S, gphys, R, 3, 0, d.nz, tmp);
      launch(is32<T>() ? KC_PROD32 : KC_PROD64, 7.0 * R * sizeof(T), [&] { scalar_products_conv<T>(uvwnl, su, sv, sw, gphys, aphys, R, st_); });
      fwd<T, CT>(aphys, R, advc, S, 1, 0, d.nz, tmp);
    }
  }
  // H = c1 H - dt adv (:313-315); on the fused path inside the right-hand-side kernel
  if (!fuse_rhs_) {
    HUpd u;
    u.c1 = c1; u.dt = dt_; u.nz = 0;
    launch(KC_OTHER, 0.0, [&] { h_update<CT>(hc_, advc, c_, u, 1, S, st_); });
  }
  // lap, rhs (:319-329)
  if (!fuse_rhs_) {
    const double redtc3 = -cfg.reynolds * cfg.schmidt / dt_ / c3;
    ZArgs a;
    a.mode = Z_RHS; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d2(a);
    a.x = c_; a.xs = S; a.y = rhsc_; a.ys = S; a.h = hc_; a.hs = S; a.kx = kx_; a.ky = ky_;
    a.s1 = redtc3; a.s2 = redtc3 * c2;
    launch(KC_Z64, zwork(), [&] { zgemm<double, double2, double2>(a, 1, st_); });
  }
  // boundary rows + erosion/deposition (boundary_scalar_flat.cc:151-152, :340-384)
  const int erosion = cfg.scalar_bc == 1 ? 1 : 0;
  if (erosion) {
    const long long plane = (long long)d.nx * d.ny;
    launch(KC_OTHER, 0.0, [&] {
      erosion_map<T>((const T*)dudz_bot_[stage & 1], (const T*)dvdz_bot_[stage & 1], ephys_, plane, cfg.reynolds,
                     cfg.critical_shear_stress, cfg.erosion_rate, st_);
    });
    fwd<double, double2>(ephys_, plane, ehat_, d.P, 1, 0, 1, (double2*)stmp_, 1);
  }
  const double wsc = cfg.settling_velocity * std::cos(cfg.slope);
  if (!fuse_rhs_)
    launch(KC_OTHER, 0.0, [&] {
      bc_scalar(rhsc_, d.nz, d.P, ehat_, c_, wsc, cfg.averaged_concentration, cfg.reynolds, cfg.schmidt, erosion,
                st_);
    });
  if (erosion) {   // feature 28: mean erosion flux, 29: mean deposition flux (:377-380)
    copy_words(fl_ + 28, ehat_, sizeof(double), st_);
    launch(KC_OTHER, 0.0, [&] { scaled_real(c_, wsc, cfg.averaged_concentration, fl_ + 29, st_); });
  }
  // Helmholtz (:399, :640-661)
  {
    const int k = 4;
    const double sigma = -cfg.reynolds * cfg.schmidt / dt_ / coef_[2][stage];
    if (fuse_rhs_) {
      // rhs, boundary rows and the solve in one pass; c is read before it is overwritten
      RHArgs r;
      r.s1 = sigma; r.s2 = sigma * coef_[1][stage];
      r.bc = 1; r.ehat = ehat_; r.ws = wsc; r.cbar = cfg.averaged_concentration; r.re = cfg.reynolds;
      r.sc = cfg.schmidt; r.erosion = erosion;
      r.hupd = 1; r.adv32 = is32<T>() ? 1 : 0; r.c1 = coef_[0][stage]; r.dt = dt_;
      const int comp = 0;
      const double2* x = c_;
      double2* h = hc_;
      const void* av = advc;
      double2* y = c_;
      rhs_solve(r, 1, &k, &comp, &x, &h, &av, &y, sigma);
    } else {
      const double2* x = rhsc_;
      double2* y = c_;
      solve(0, 1, &k, &sigma, &x, &y);
    }
  }
  // Bartnicki positivity filter: physical c, 5 plane iterations, back (:419-478).
  // The filter leaves a field without negative values unchanged, so an FP32
  // transform of c and a min/max check first decide whether the exact FP64
  // transform and filter have anything to do (need_); the filter raises
  // filt_ when it changes a plane; otherwise the transform back would
  // reproduce c except for the kx = 0 plane's projection, which is applied
  // directly
  bwd<float, double2>(c_, S, (float*)saphys_, R, 1, 0, d.nz, (float2*)stmp_);
  launch(KC_OTHER, 0.0, [&] { negativity_check<float>((const float*)saphys_, R, mmpart_, need_, st_); });
  bwd<double, double2>(c_, S, cphys_, R, 1, 0, d.nz, (double2*)stmp_, -1, need_);
  fill_zero(filt_, sizeof(int) * 2, st_);
  launch(KC_OTHER, 0.0, [&] { bartnicki_filter(cphys_, d.nx, d.ny, d.nz, filt_, st_, need_); });
  fwd<double, double2>(cphys_, R, c_, S, 1, 0, d.nz, (double2*)stmp_, -1, filt_);
  launch(KC_OTHER, 0.0, [&] { kx0_round_trip(c_, d.nyc, d.nz, filt_, st_); });
  launch(KC_OTHER, 0.0, [&] { trapz_mean(c_, gridz_, d.nz, fl_ + 21, st_); });
}

void Context::stages(double time, int alternation) {
  const bool buoy = cfg.nphase > 1 && cfg.affected_by_scalar_phase;
  if (!overlap_ || cfg.nphase < 2 || buoy || prof.on) {
    for (int s = 0; s < 3; ++s) substage(s, time, alternation);
    return;
  }
  const hipStream_t main = st_;
  for (int s = 0; s < 3; ++s) {
    if (cfg.precision == 1) carrier_predict<float>(s, time, alternation);
    else carrier_predict<double>(s, time, alternation);
    // the scalar of stage s - 1 has read the physical velocity Correct overwrites
    if (s > 0) TBD_HIP(hipStreamWaitEvent(main, evj_, 0));
    if (cfg.precision == 1) carrier_correct<float>(s);
    else carrier_correct<double>(s);
    TBD_HIP(hipEventRecord(evf_, main));
    TBD_HIP(hipStreamWaitEvent(st2_, evf_, 0));
    st_ = st2_;
    try {
      if (cfg.precision == 1) scalar_correct<float>(s, alternation);
      else scalar_correct<double>(s, alternation);
    } catch (...) {
      st_ = main;
      throw;
    }
    st_ = main;
    TBD_HIP(hipEventRecord(evj_, st2_));
  }
  TBD_HIP(hipStreamWaitEvent(main, evj_, 0));
}

void Context::substage(int stage, double time, int alternation) {
  TBD_CHECK(stage >= 0 && stage < 3, "Context::substage > stage must be 0, 1 or 2");
  if (cfg.precision == 1) {
    carrier_predict<float>(stage, time, alternation);
    carrier_correct<float>(stage);
    if (cfg.nphase > 1) scalar_correct<float>(stage, alternation);
  } else {
    carrier_predict<double>(stage, time, alternation);
    carrier_correct<double>(stage);
    if (cfg.nphase > 1) scalar_correct<double>(stage, alternation);
  }
  if (stage == 2 && cfg.nphase > 1) {
    // the log keeps the scalar values of stage 1 (copied at the carrier's stage 2)
  }
}

void Context::stats_accumulate(int bin) {
  TBD_CHECK(bin >= 0 && bin < kStatBins, "Context::stats_accumulate > phase bin out of range");
  phase_stats(uvw_, cfg.nphase > 1 ? c_ : nullptr, d.S, d.nz, d.nyc, d.P,
              stats_ + (size_t)bin * kPhaseStats * d.nz, st_);
  stat_count_[bin] += 1;
}

void Context::stats_read(int bin, double* out, long long* count) {
  TBD_CHECK(bin >= 0 && bin < kStatBins, "Context::stats_read > phase bin out of range");
  sync();
  TBD_HIP(hipMemcpy(out, stats_ + (size_t)bin * kPhaseStats * d.nz, sizeof(double) * kPhaseStats * d.nz,
                    hipMemcpyDeviceToHost));
  if (count) *count = stat_count_[bin];
}

void Context::stats_reset() {
  fill_zero(stats_, sizeof(double) * kStatBins * kPhaseStats * d.nz, st_);
  for (auto& c : stat_count_) c = 0;
  sync();
}

// Projection of an arbitrary velocity onto divergence-free fields: the
// pressure-type solve of Correct (phase_carrier_flat.cc:586-620) with
// 2 dt c3 = 1, phi in place of p.
void Context::project() {
  const long long S = d.S;
  if (fuse_p_) {
    pchain(1.0, 1.0, uvw_, rhsp_, uvw_, 0, false);
  } else {
    {
      ZArgs a;
      a.mode = Z_DIV; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<double>(a);
      a.x = uvw_ + 2 * S; a.e0 = uvw_; a.e1 = uvw_ + S; a.y = rhsp_; a.kx = kx_; a.ky = ky_;
      a.q = 1.0;
      zgemm<double, double2, double2>(a, 1, st_);
    }
    {
      const int k = 3;
      const double sigma = 0.0;
      const double2* x = rhsp_;
      double2* y = rhsp_;
      solve(1, 1, &k, &sigma, &x, &y);
    }
    {
      ZArgs a;
      a.mode = Z_CORRECT; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<double>(a);
      a.x = rhsp_; a.e0 = uvw_; a.es = S; a.y = uvw_; a.ys = S; a.kx = kx_; a.ky = ky_;
      a.cf = 1.0;
      a.wall_bot = wall_[0][0]; a.wall_top = wall_[0][1];
      a.wall2_bot = wall2_[0][0]; a.wall2_top = wall2_[0][1];
      a.open_top = opentop_ ? 1 : 0;
      zgemm<double, double2, double2>(a, 1, st_);
    }
  }
  for (int s = 0; s < 2; ++s)
    for (int b = 0; b < 2; ++b) {
      fill_zero(wall_[s][b], sizeof(double2) * 3 * d.P, st_);
      fill_zero(wall2_[s][b], sizeof(double2) * 2 * d.P, st_);
    }
  olddt_ = dt_;
  if (cfg.precision == 1)
    fft_backward<float, double2>(fft_, uvw_, d.S, (float*)uvwnl_, d.R, 3, 0, d.nz, (float2*)uxp_, st_);
  else
    fft_backward<double, double2>(fft_, uvw_, d.S, (double*)uvwnl_, d.R, 3, 0, d.nz, (double2*)uxp_, st_);
  sync();
}

void Context::swirling(double* out) {
  const long long S = d.S, R = d.R;
  if (!pgrad_) {
    pgrad_ = (double2*)alloc(sizeof(double2) * 9 * S);
    pphys_ = (double*)alloc(sizeof(double) * 10 * R);
  }
  // Gradient(uvw, dudxyz / dvdxyz / dwdxyz) (postprocessor_channel_flat.cc:203-205)
  ZArgs a;
  a.mode = Z_GRAD; a.nz = d.nz; a.P = d.P; a.nyc = d.nyc; set_d1<double>(a);
  a.x = uvw_; a.xs = S; a.y = pgrad_; a.ys = S; a.kx = kx_; a.ky = ky_;
  zgemm<double, double2, double2>(a, 3, st_);
  // DftBackward of the nine components (:206-208), one at a time: the
  // transform scratch holds one FP64 component in either precision
  for (int k = 0; k < 9; ++k)
    fft_backward<double, double2>(fft_, pgrad_ + k * S, S, pphys_ + k * R, R, 1, 0, d.nz, (double2*)tmp_, st_);
  swirling_strength(pphys_, R, pphys_ + 9 * R, st_);
  TBD_HIP(hipMemcpyAsync(out, pphys_ + 9 * R, sizeof(double) * R, hipMemcpyDeviceToHost, st_));
  sync();
}

void Context::helmholtz_host(int which, int stage, int mode, const double* x, double* y) {
  TBD_CHECK(which >= 0 && which < (cfg.nphase > 1 ? 5 : 4), "Context::helmholtz > bad factorisation index");
  const size_t S16 = sizeof(double2) * d.S;
  // buffers of this call only (hipFree waits for the device)
  struct Tmp {
    void* p = nullptr;
    ~Tmp() { if (p) (void)hipFree(p); }
  } bx, by;
  TBD_HIP(hipMalloc(&bx.p, S16));
  TBD_HIP(hipMalloc(&by.p, S16));
  double2* dx = (double2*)bx.p;
  double2* dy = (double2*)by.p;
  TBD_HIP(hipMemcpyAsync(dx, x, S16, hipMemcpyHostToDevice, st_));
  double sig = 0.0;
  if (which < 3) sig = -cfg.reynolds / dt_ / coef_[2][stage];
  else if (which == 4) sig = -cfg.reynolds * cfg.schmidt / dt_ / coef_[2][stage];
  const double2* xx = dx;
  solve(mode, 1, &which, &sig, &xx, &dy);
  TBD_HIP(hipMemcpyAsync(y, dy, S16, hipMemcpyDeviceToHost, st_));
  sync();
}

}  // namespace tbd
