// Per-point 2D updates shared by the single-grid and the batched kernels, so
// both compute bit-identical values.
#pragma once

#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {

// Staggered difference between the half points p+s ("+1/2") and p ("-1/2").
template <typename T, int ORDER>
struct Op;

template <typename T>
struct Op<T, 2> {
  __device__ static T d(const T *f, long p, long s) { return f[p + s] - f[p]; }
};

// 2D4:565  (27*f(+1/2) - 27*f(-1/2) - f(+3/2) + f(-3/2))
template <typename T>
struct Op<T, 4> {
  __device__ static T d(const T *f, long p, long s) {
    return T(27) * f[p + s] - T(27) * f[p] - f[p + 2 * s] + f[p - s];
  }
};

// ADE8:734-735 association: c1*(..) + c2*(..) + c3*(..) + c4*(..)
template <typename T>
struct Op<T, 8> {
  __device__ static T d(const T *f, long p, long s) {
    return T(kHolberg1) * (f[p + s] - f[p]) + T(kHolberg2) * (f[p + 2 * s] - f[p - s]) +
           T(kHolberg3) * (f[p + 3 * s] - f[p - 2 * s]) +
           T(kHolberg4) * (f[p + 4 * s] - f[p - 3 * s]);
  }
};

// stress + C-PML at grid point (i, j): 2D2:556-600 (2D4:557-601)
template <typename T, int ORDER, bool NT>
__device__ __forceinline__ void stress2d_point(const Step2D<T> &p, int i, int j) {
  const long ld = p.ld;
  const long c = i + j * ld;
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const long mxo = p.sx.s(i) + (long)p.sx.w * j;
  const long myo = i + p.mld * p.sy.s(j);
  // sigma_xx, sigma_yy at (i+1/2, j), Fortran i=1..NX-1, j=2..NY
  if (i <= p.nx - 2 && j >= 1) {
    const T lhx = T(0.5) * (p.lam[c + 1] + p.lam[c]);
    const T mhx = T(0.5) * (p.mu[c + 1] + p.mu[c]);
    const T lp2m = lhx + T(2) * mhx;
    T dvx_dx = Op<T, ORDER>::d(p.vx, c, 1) / p.den_x;
    T dvy_dy = Op<T, ORDER>::d(p.vy, c - ld, ld) / p.den_y;
    if (inx) dvx_dx = cpml(dvx_dx, p.mx[0] + mxo, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
    if (iny) dvy_dy = cpml(dvy_dy, p.my[0] + myo, p.py.a[j], p.py.b[j], p.py.K[j]);
    // own-point values (ld_once / st_once: streamed past the L2 when NT)
    st_once<NT>(p.sxx + c, ld_once<NT>(p.sxx + c) + (lp2m * dvx_dx + lhx * dvy_dy) * p.dt);
    st_once<NT>(p.syy + c, ld_once<NT>(p.syy + c) + (lhx * dvx_dx + lp2m * dvy_dy) * p.dt);
  }
  // sigma_xy at (i, j+1/2), Fortran i=2..NX, j=1..NY-1
  if (i >= 1 && j <= p.ny - 2) {
    const T mhy = T(0.5) * (p.mu[c + ld] + p.mu[c]);
    T dvy_dx = Op<T, ORDER>::d(p.vy, c - 1, 1) / p.den_x;
    T dvx_dy = Op<T, ORDER>::d(p.vx, c, ld) / p.den_y;
    if (inx) dvy_dx = cpml(dvy_dx, p.mx[1] + mxo, p.px.a[i], p.px.b[i], p.px.K[i]);
    // 2D4:596 divides by K_y(j) where 2D2:595 uses K_y_half(j)
    if (iny)
      dvx_dy = cpml(dvx_dy, p.my[1] + myo, p.py.ah[j], p.py.bh[j],
                    ORDER == 4 ? p.py.K[j] : p.py.Kh[j]);
    st_once<NT>(p.sxy + c, ld_once<NT>(p.sxy + c) + mhy * (dvy_dx + dvx_dy) * p.dt);
  }
}

// velocity + C-PML, source and rigid edges at grid point (i, j):
// 2D2:606-680 (2D4:607-681)
template <typename T, int ORDER, bool NT>
__device__ __forceinline__ void velocity2d_point(const Step2D<T> &p, int i, int j) {
  const long ld = p.ld;
  const long c = i + j * ld;
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const long mxo = p.sx.s(i) + (long)p.sx.w * j;
  const long myo = i + p.mld * p.sy.s(j);
  T vxn = ld_once<NT>(p.vx + c), vyn = ld_once<NT>(p.vy + c);
  // vx at (i, j), Fortran i=2..NX, j=2..NY
  if (i >= 1 && j >= 1) {
    T dsxx_dx = Op<T, ORDER>::d(p.sxx, c - 1, 1) / p.den_x;
    T dsxy_dy = Op<T, ORDER>::d(p.sxy, c - ld, ld) / p.den_y;
    if (inx) dsxx_dx = cpml(dsxx_dx, p.mx[2] + mxo, p.px.a[i], p.px.b[i], p.px.K[i]);
    if (iny) dsxy_dy = cpml(dsxy_dy, p.my[2] + myo, p.py.a[j], p.py.b[j], p.py.K[j]);
    vxn = vxn + (dsxx_dx + dsxy_dy) * p.dt / p.rho[c];
  }
  // vy at (i+1/2, j+1/2), Fortran i=1..NX-1, j=1..NY-1
  if (i <= p.nx - 2 && j <= p.ny - 2) {
    const T rhh = T(0.25) * (p.rho[c] + p.rho[c + 1] + p.rho[c + 1 + ld] + p.rho[c + ld]);
    T dsxy_dx = Op<T, ORDER>::d(p.sxy, c, 1) / p.den_x;
    T dsyy_dy = Op<T, ORDER>::d(p.syy, c, ld) / p.den_y;
    if (inx) dsxy_dx = cpml(dsxy_dx, p.mx[3] + mxo, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
    if (iny) dsyy_dy = cpml(dsyy_dy, p.my[3] + myo, p.py.ah[j], p.py.bh[j], p.py.Kh[j]);
    vyn = vyn + (dsxy_dx + dsyy_dy) * p.dt / rhh;
  }
  // point force, added after the velocity update (2D2:660-667)
  if (i == p.isrc && j == p.jsrc) {
    const int it = *p.step;
    const T rhh = T(0.25) * (p.rho[c] + p.rho[c + 1] + p.rho[c + 1 + ld] + p.rho[c + ld]);
    vxn = vxn + p.fx[it - 1] * p.dt / p.rho[c];
    vyn = vyn + p.fy[it - 1] * p.dt / rhh;
  }
  // rigid edges (2D2:670-680)
  if (i == 0 || i == p.nx - 1 || j == 0 || j == p.ny - 1) {
    vxn = T(0);
    vyn = T(0);
  }
  st_once<NT>(p.vx + c, vxn);
  st_once<NT>(p.vy + c, vyn);
}

}  // namespace cpml
