// C ABI (include/turbid_amd.h).  Exceptions never cross the boundary: they
// become negative return codes with the message kept for tbd_last_error().
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/turbid_amd.h"
#include "context.h"
#include "ensemble.h"
#include "fft.h"
#include "pointwise.h"
#include "zgemm.h"
#include "sno.h"

struct tbd_ctx {
  tbd::Context* c;
};
struct tbd_ens {
  tbd::Ensemble* e;
};

namespace tbd {
void dft_forward_noalloc(int nx, int ny, int nz, int nv, int fp32, const double* phys, double* spec);
}

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -2;
  }
}

// RAII device buffer for the stand-alone entry points
struct DBuf {
  void* p = nullptr;
  explicit DBuf(size_t n) { TBD_HIP(hipMalloc(&p, n < 256 ? 256 : n)); }
  ~DBuf() { if (p) (void)hipFree(p); }
  template <class T> T* as() { return (T*)p; }
};
struct FBuf {
  float* p = nullptr;
  explicit FBuf(size_t n) { TBD_HIP(hipMalloc((void**)&p, n * sizeof(float) < 256 ? 256 : n * sizeof(float))); }
  ~FBuf() { if (p) (void)hipFree(p); }
};
}  // namespace

extern "C" {

const char* tbd_last_error(void) { return g_err.c_str(); }

int tbd_create(const tbd_config* cfg, tbd_ctx** out) {
  return guard([&] {
    TBD_CHECK(cfg && out, "tbd_create > null argument");
    *out = nullptr;
    tbd_ctx* c = new tbd_ctx;
    try {
      c->c = new tbd::Context(*cfg);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

void tbd_destroy(tbd_ctx* ctx) {
  if (!ctx) return;
  try { delete ctx->c; } catch (...) {}
  delete ctx;
}

int tbd_upload_state(tbd_ctx* ctx, const double* uvw, const double* p, const double* c) {
  return guard([&] {
    TBD_CHECK(ctx && uvw && p, "tbd_upload_state > null argument");
    ctx->c->upload(uvw, p, c);
  });
}

int tbd_download_state(tbd_ctx* ctx, double* uvw, double* p, double* c) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_download_state > null context");
    ctx->c->download(uvw, p, c);
  });
}

int tbd_download_physical_velocity(tbd_ctx* ctx, double* uvw) {
  return guard([&] {
    TBD_CHECK(ctx && uvw, "tbd_download_physical_velocity > null argument");
    ctx->c->download_physical_velocity(uvw);
  });
}

int tbd_substage(tbd_ctx* ctx, int stage, double time, int alternation) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_substage > null context");
    ctx->c->substage(stage, time, alternation);
  });
}

int tbd_ensemble_create(tbd_ctx* const* members, int n, tbd_ens** out) {
  return guard([&] {
    TBD_CHECK(members && out && n >= 1, "tbd_ensemble_create > null argument or no members");
    *out = nullptr;
    std::vector<tbd::Context*> m;
    for (int i = 0; i < n; ++i) {
      TBD_CHECK(members[i], "tbd_ensemble_create > null member");
      m.push_back(members[i]->c);
    }
    tbd_ens* e = new tbd_ens;
    try {
      e->e = new tbd::Ensemble(m);
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

int tbd_ensemble_step(tbd_ens* e, const double* time, const int* alternation) {
  return guard([&] {
    TBD_CHECK(e && time && alternation, "tbd_ensemble_step > null argument");
    e->e->step(time, alternation);
  });
}

int tbd_ensemble_launches(tbd_ens* e, int* grouped, int* member) {
  return guard([&] {
    TBD_CHECK(e, "tbd_ensemble_launches > null ensemble");
    if (grouped) *grouped = e->e->grouped_launches();
    if (member) *member = e->e->member_launches();
  });
}

int tbd_ensemble_features(tbd_ens* e, double* out) {
  return guard([&] {
    TBD_CHECK(e && out, "tbd_ensemble_features > null argument");
    e->e->features(out);
  });
}

int tbd_ensemble_sync(tbd_ens* e) {
  return guard([&] {
    TBD_CHECK(e, "tbd_ensemble_sync > null ensemble");
    e->e->sync();
  });
}

void tbd_ensemble_destroy(tbd_ens* e) {
  if (!e) return;
  try { delete e->e; } catch (...) {}
  delete e;
}

int tbd_step(tbd_ctx* ctx, double time, int alternation) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_step > null context");
    ctx->c->step(time, alternation);
  });
}

int tbd_set_dt(tbd_ctx* ctx, double dt) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_set_dt > null context");
    ctx->c->set_dt(dt);
  });
}

int tbd_features(tbd_ctx* ctx, double* out50) {
  return guard([&] {
    TBD_CHECK(ctx && out50, "tbd_features > null argument");
    ctx->c->features(out50);
  });
}

int tbd_sync(tbd_ctx* ctx) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_sync > null context");
    ctx->c->sync();
  });
}

int tbd_profile_enable(tbd_ctx* ctx, int enable) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_profile_enable > null context");
    ctx->c->prof.on = enable != 0;
  });
}

int tbd_profile_read(tbd_ctx* ctx, int kclass, long long* launches, double* total_ms, double* work) {
  return guard([&] {
    TBD_CHECK(ctx && kclass >= 0 && kclass < tbd::KC_N, "tbd_profile_read > bad argument");
    ctx->c->sync();
    ctx->c->prof.collect();
    if (launches) *launches = ctx->c->prof.launches[kclass];
    if (total_ms) *total_ms = ctx->c->prof.ms[kclass];
    if (work) *work = ctx->c->prof.work[kclass];
  });
}

int tbd_profile_reset(tbd_ctx* ctx) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_profile_reset > null context");
    ctx->c->sync();
    ctx->c->prof.reset();
  });
}

int tbd_get_operators(tbd_ctx* ctx, double* gridz, double* d1, double* d2) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_get_operators > null context");
    const auto& ch = ctx->c->ch;
    const size_t n = ch.nz;
    if (gridz) std::memcpy(gridz, ch.gridz.data(), sizeof(double) * n);
    if (d1) std::memcpy(d1, ch.d1.data(), sizeof(double) * n * n);
    if (d2) std::memcpy(d2, ch.d2.data(), sizeof(double) * n * n);
  });
}

int tbd_get_factorisation(tbd_ctx* ctx, int which, double* V, double* Vinv, double* lambda, double* fix) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_get_factorisation > null context");
    TBD_CHECK(which >= 0 && which < (ctx->c->cfg.nphase > 1 ? 5 : 4), "tbd_get_factorisation > bad index");
    const auto& f = ctx->c->fac[which];
    const size_t n = f.nz;
    if (V) std::memcpy(V, f.vnorm.data(), sizeof(double) * n * n);
    if (Vinv) std::memcpy(Vinv, f.vinv.data(), sizeof(double) * n * n);
    if (lambda) std::memcpy(lambda, f.lambda.data(), sizeof(double) * n);
    if (fix) std::memcpy(fix, f.fix.data(), sizeof(double) * 4 * n);
  });
}

int tbd_set_operators(tbd_ctx* ctx, const double* d1, const double* d2) {
  return guard([&] {
    TBD_CHECK(ctx && d1 && d2, "tbd_set_operators > null argument");
    ctx->c->set_operators(d1, d2);
  });
}

int tbd_set_factorisation(tbd_ctx* ctx, int which, const double* V, const double* Vinv, const double* lambda,
                          const double* fix) {
  return guard([&] {
    TBD_CHECK(ctx && V && Vinv && lambda && fix, "tbd_set_factorisation > null argument");
    ctx->c->set_factorisation(which, V, Vinv, lambda, fix);
  });
}

int tbd_autotune(tbd_ctx* ctx, int steps, double* ms_per_step, int* bits) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_autotune > null context");
    double ms[tbd::Context::kVariants];
    const int b = ctx->c->autotune(steps, ms);
    if (ms_per_step) std::memcpy(ms_per_step, ms, sizeof(ms));
    if (bits) *bits = b;
  });
}

int tbd_get_variant(tbd_ctx* ctx, int* bits) {
  return guard([&] {
    TBD_CHECK(ctx && bits, "tbd_get_variant > null argument");
    *bits = ctx->c->variant().bits();
  });
}

int tbd_set_variant(tbd_ctx* ctx, int bits) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_set_variant > null context");
    TBD_CHECK(bits >= 0 && bits < tbd::Context::kVariants, "tbd_set_variant > bits out of range");
    ctx->c->set_variant(tbd::Context::Variant::of(bits));
  });
}

int tbd_dft_forward(int nx, int ny, int nz, int nv, int fp32, const double* phys, double* spec) {
  return guard([&] {
    TBD_CHECK(phys && spec && nv >= 1, "tbd_dft_forward > bad argument");
    const tbd::Dims d = tbd::Dims::make(nx, ny, nz);
    tbd::FftPlan p;
    p.create(d);
    DBuf dphys(sizeof(double) * d.R * nv), dspec(sizeof(double2) * d.S * nv);
    DBuf tmp(sizeof(double2) * tbd::fft_tmp_elems(d, nv));
    TBD_HIP(hipMemcpy(dphys.p, phys, sizeof(double) * d.R * nv, hipMemcpyHostToDevice));
    if (fp32) {
      std::vector<float> h(phys, phys + d.R * nv);
      DBuf fphys(sizeof(float) * d.R * nv);
      TBD_HIP(hipMemcpy(fphys.p, h.data(), sizeof(float) * d.R * nv, hipMemcpyHostToDevice));
      DBuf fspec(sizeof(float2) * d.S * nv);
      tbd::fft_forward<float, float2>(p, fphys.as<float>(), d.R, fspec.as<float2>(), d.S, nv, 0, nz,
                                      tmp.as<float2>(), 0);
      std::vector<float> o(2 * d.S * nv);
      TBD_HIP(hipMemcpy(o.data(), fspec.p, sizeof(float2) * d.S * nv, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < o.size(); ++i) spec[i] = o[i];
    } else {
      tbd::fft_forward<double, double2>(p, dphys.as<double>(), d.R, dspec.as<double2>(), d.S, nv, 0, nz,
                                        tmp.as<double2>(), 0);
      TBD_HIP(hipMemcpy(spec, dspec.p, sizeof(double2) * d.S * nv, hipMemcpyDeviceToHost));
    }
    TBD_HIP(hipDeviceSynchronize());
    p.destroy();
  });
}

int tbd_dft_forward_noalloc(int nx, int ny, int nz, int nv, int fp32, const double* phys, double* spec) {
  return guard([&] { tbd::dft_forward_noalloc(nx, ny, nz, nv, fp32, phys, spec); });
}

int tbd_dft_backward(int nx, int ny, int nz, int nv, int fp32, const double* spec, double* phys) {
  return guard([&] {
    TBD_CHECK(phys && spec && nv >= 1, "tbd_dft_backward > bad argument");
    const tbd::Dims d = tbd::Dims::make(nx, ny, nz);
    tbd::FftPlan p;
    p.create(d);
    DBuf dspec(sizeof(double2) * d.S * nv), tmp(sizeof(double2) * tbd::fft_tmp_elems(d, nv));
    TBD_HIP(hipMemcpy(dspec.p, spec, sizeof(double2) * d.S * nv, hipMemcpyHostToDevice));
    if (fp32) {
      DBuf fphys(sizeof(float) * d.R * nv);
      tbd::fft_backward<float, double2>(p, dspec.as<double2>(), d.S, fphys.as<float>(), d.R, nv, 0, nz,
                                        tmp.as<float2>(), 0);
      std::vector<float> o(d.R * nv);
      TBD_HIP(hipMemcpy(o.data(), fphys.p, sizeof(float) * d.R * nv, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < o.size(); ++i) phys[i] = o[i];
    } else {
      DBuf dphys(sizeof(double) * d.R * nv);
      tbd::fft_backward<double, double2>(p, dspec.as<double2>(), d.S, dphys.as<double>(), d.R, nv, 0, nz,
                                         tmp.as<double2>(), 0);
      TBD_HIP(hipMemcpy(phys, dphys.p, sizeof(double) * d.R * nv, hipMemcpyDeviceToHost));
    }
    TBD_HIP(hipDeviceSynchronize());
    p.destroy();
  });
}

int tbd_zcontract(int nz, long long ncol, int fp32, const double* m, const double* x, double* y) {
  return guard([&] {
    TBD_CHECK(nz >= 4 && ncol >= 1 && m && x && y, "tbd_zcontract > bad argument");
    const int lda = (nz + 3) / 4 * 4;
    std::vector<double> rm((size_t)nz * lda, 0.0);
    for (int i = 0; i < nz; ++i)
      for (int j = 0; j < nz; ++j) rm[(size_t)i * lda + j] = m[i + (size_t)nz * j];
    std::vector<double> kx(1, 0.0), ky(ncol, 0.0);
    DBuf dk(sizeof(double) * 1), dky(sizeof(double) * ncol);
    TBD_HIP(hipMemcpy(dk.p, kx.data(), sizeof(double), hipMemcpyHostToDevice));
    TBD_HIP(hipMemcpy(dky.p, ky.data(), sizeof(double) * ncol, hipMemcpyHostToDevice));
    const size_t n2 = (size_t)nz * ncol;
    DBuf dx(sizeof(double2) * n2), dy(sizeof(double2) * n2);
    TBD_HIP(hipMemcpy(dx.p, x, sizeof(double2) * n2, hipMemcpyHostToDevice));
    tbd::ZArgs a;
    a.mode = tbd::Z_PLAIN; a.nz = nz; a.lda = lda; a.P = ncol; a.nyc = (int)ncol;
    a.x = dx.p; a.y = dy.p; a.kx = dk.as<double>(); a.ky = dky.as<double>();
    if (fp32) {
      std::vector<float> rf(rm.begin(), rm.end());
      DBuf dm(sizeof(float) * rf.size()), fy(sizeof(float2) * n2);
      TBD_HIP(hipMemcpy(dm.p, rf.data(), sizeof(float) * rf.size(), hipMemcpyHostToDevice));
      a.op = dm.p;
      a.y = fy.p;
      tbd::zgemm<float, double2, float2>(a, 1, 0);
      std::vector<float> o(2 * n2);
      TBD_HIP(hipMemcpy(o.data(), fy.p, sizeof(float2) * n2, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < o.size(); ++i) y[i] = o[i];
    } else {
      DBuf dm(sizeof(double) * rm.size());
      TBD_HIP(hipMemcpy(dm.p, rm.data(), sizeof(double) * rm.size(), hipMemcpyHostToDevice));
      a.op = dm.p;
      tbd::zgemm<double, double2, double2>(a, 1, 0);
      TBD_HIP(hipMemcpy(y, dy.p, sizeof(double2) * n2, hipMemcpyDeviceToHost));
    }
    TBD_HIP(hipDeviceSynchronize());
  });
}

int tbd_helmholtz(tbd_ctx* ctx, int which, int stage, int mode, const double* x, double* y) {
  return guard([&] {
    TBD_CHECK(ctx && x && y && stage >= 0 && stage < 3 && (mode == 0 || mode == 1), "tbd_helmholtz > bad argument");
    ctx->c->helmholtz_host(which, stage, mode, x, y);
  });
}

int tbd_swirling_strength(tbd_ctx* ctx, double* out) {
  return guard([&] {
    TBD_CHECK(ctx && out, "tbd_swirling_strength > bad argument");
    ctx->c->swirling(out);
  });
}

int tbd_stats_accumulate(tbd_ctx* ctx, int bin) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_stats_accumulate > null context");
    ctx->c->stats_accumulate(bin);
  });
}
int tbd_stats_read(tbd_ctx* ctx, int bin, double* out, long long* count) {
  return guard([&] {
    TBD_CHECK(ctx && out, "tbd_stats_read > null argument");
    ctx->c->stats_read(bin, out, count);
  });
}
int tbd_stats_reset(tbd_ctx* ctx) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_stats_reset > null context");
    ctx->c->stats_reset();
  });
}

int tbd_project(tbd_ctx* ctx) {
  return guard([&] {
    TBD_CHECK(ctx, "tbd_project > null context");
    ctx->c->project();
  });
}

struct tbd_sno {
  tbd::Sno* n;
};

int tbd_sno_create(int nx, int ny, int nz, int channels, int layers, int cin, int cout, int batch, tbd_sno** out) {
  return guard([&] {
    TBD_CHECK(out, "tbd_sno_create > null output");
    tbd::SnoShape s;
    s.nx = nx; s.ny = ny; s.nz = nz; s.C = channels; s.L = layers; s.cin = cin; s.cout = cout; s.batch = batch;
    std::unique_ptr<tbd::Sno> n(new tbd::Sno(s));
    *out = new tbd_sno{n.release()};
  });
}
void tbd_sno_destroy(tbd_sno* net) {
  if (!net) return;
  delete net->n;
  delete net;
}
int tbd_sno_param_count(tbd_sno* net, long long* count) {
  return guard([&] {
    TBD_CHECK(net && count, "tbd_sno_param_count > null argument");
    *count = net->n->lay.total;
  });
}
int tbd_sno_set_params(tbd_sno* net, const float* params) {
  return guard([&] {
    TBD_CHECK(net && params, "tbd_sno_set_params > null argument");
    net->n->set_params(params);
  });
}
int tbd_sno_get_params(tbd_sno* net, float* params) {
  return guard([&] {
    TBD_CHECK(net && params, "tbd_sno_get_params > null argument");
    net->n->get_params(params);
  });
}
int tbd_sno_get_grads(tbd_sno* net, float* grads) {
  return guard([&] {
    TBD_CHECK(net && grads, "tbd_sno_get_grads > null argument");
    net->n->get_grads(grads);
  });
}
int tbd_sno_forward(tbd_sno* net, const float* x, float* y, int nb) {
  return guard([&] {
    TBD_CHECK(net && x && y, "tbd_sno_forward > null argument");
    net->n->forward(x, y, nb);
  });
}
int tbd_sno_loss_grad(tbd_sno* net, const float* x, const float* t, const float* w, int nb, double* loss) {
  return guard([&] {
    TBD_CHECK(net && x && t && w && loss, "tbd_sno_loss_grad > null argument");
    *loss = net->n->loss_and_grad(x, t, w, nb);
  });
}
int tbd_sno_train_step(tbd_sno* net, const float* x, const float* t, const float* w, int nb, float lr,
                       double* loss) {
  return guard([&] {
    TBD_CHECK(net && x && t && w && loss, "tbd_sno_train_step > null argument");
    *loss = net->n->train_step(x, t, w, nb, lr);
  });
}
int tbd_bf16_gemm(int m, int n, int k, const float* a, const float* b, float* c) {
  return guard([&] {
    TBD_CHECK(m > 0 && n > 0 && k > 0 && a && b && c, "tbd_bf16_gemm > bad argument");
    FBuf da((size_t)m * k), db((size_t)k * n), dc((size_t)m * n);
    TBD_HIP(hipMemcpy(da.p, a, sizeof(float) * m * k, hipMemcpyHostToDevice));
    TBD_HIP(hipMemcpy(db.p, b, sizeof(float) * k * n, hipMemcpyHostToDevice));
    tbd::gemm_bf16(m, n, k, 1, 1, 1.f, tbd::GemmOp{da.p, 0, 0, k, 1}, tbd::GemmOp{db.p, 0, 0, n, 1}, 0.f,
                   tbd::GemmOut{dc.p, 0, 0, n, 1}, 0);
    TBD_HIP(hipDeviceSynchronize());
    TBD_HIP(hipMemcpy(c, dc.p, sizeof(float) * m * n, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
