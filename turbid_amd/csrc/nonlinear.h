// Nonlinear terms with the x-direction transforms fused around the products
// (reference FlatCarrierPhase::NonlinearVelocity, phase_carrier_flat.cc:716-813,
// FlatScalarPhase::NonlinearPressure, phase_scalar_flat.cc:552-599).
//
// The y passes of fft.h move between the spectral layout and the x-pencil
// intermediate T[comp][z][kx][y]; these kernels take a tile of rows of one
// z-plane, inverse-transform the needed components along x in LDS, form the
// products with the physical velocity rows in registers, and forward-
// transform the results along x -- the physical gradient and product fields
// never travel through HBM.  Rows are transformed two at a time as one
// complex sequence (row 2s in the real part, row 2s+1 in the imaginary part).
#pragma once

#include "fft.h"

namespace tbd {

// adv_i = sum_j (U_j + shift_j) g_{i,j}; nout = 3 (carrier: g has 9 components
// ordered du/dx, du/dy, du/dz, dv/dx, ..., dw/dz) or 1 (scalar: g has 3).
// bot0/bot1 (optional, nout = 3): physical bottom plane of du/dz and dv/dz.
// With nout = 3, adv must not overlap g (the outputs of a row tile are written
// by different workgroups); with nout = 1 it may be g itself.
// ux (optional, nout = 3, xfused_registers(p)): the x-pencils of u, v, w
// (the intermediate of their inverse y pass); then g holds the six
// components du/dy, dv/dy, dw/dy, du/dz, dv/dz, dw/dz and du_i/dx is i kx
// times ux_i, kx = dkx * (spectral x index), formed in the kernel.
template <typename T>
void xfused_conv(const FftPlan& p, const C2<T>* g, const T* U, long long R, T su, T sv, T sw, int nout,
                 C2<T>* adv, T* bot0, T* bot1, hipStream_t st, const C2<T>* ux = nullptr, double dkx = 0.0);

// whether the fused kernels run as the two-level register transform at this
// plan's x length (the form that takes ux above)
bool xfused_registers(const FftPlan& p);

// the six products uu, uv, uw, vv, vw, ww of U, transformed along x
template <typename T>
void xfused_div6(const FftPlan& p, const T* U, long long R, C2<T>* prod, hipStream_t st);

// c (one component, inverse-transformed along x) times (U_j + shift_j), j = 0..2
template <typename T>
void xfused_cdiv(const FftPlan& p, const C2<T>* c, const T* U, long long R, T su, T sv, T sw, C2<T>* prod,
                 hipStream_t st);

}  // namespace tbd
