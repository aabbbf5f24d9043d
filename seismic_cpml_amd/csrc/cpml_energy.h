// Energy of one cell, shared by the single-grid and the batched energy
// kernels so both sum bit-identical terms.
#pragma once

#include "cpml_internal.h"

namespace cpml {

// One 2D cell's terms, 2D2:694-711 (kinetic summed then halved by the
// caller; potential per cell), in the field precision, accumulated in double.
template <typename T>
__device__ inline void point_energy2d(T vx, T vy, T sxx, T syy, T sxy, T l, T m, T r,
                                      double &kin, double &pot) {
  kin += (double)(r * (vx * vx + vy * vy));
  const T exx = ((l + T(2) * m) * sxx - l * syy) / (T(4) * m * (l + m));
  const T eyy = ((l + T(2) * m) * syy - l * sxx) / (T(4) * m * (l + m));
  const T exy = sxy / (T(2) * m);
  pot += (double)(T(0.5) * (exx * sxx + eyy * syy + T(2) * exy * sxy));
}

template <typename T>
__device__ inline void cell_energy(const Energy<T> &p, long q, double &kin, double &pot) {
  const T r = p.rho[q], l = p.lam[q], m = p.mu[q];
  if (p.ndim == 2) {
    point_energy2d(p.vx[q], p.vy[q], p.sxx[q], p.syy[q], p.sxy[q], l, m, r, kin, pot);
  } else {
    // 3D:1149-1176; the reference counts eps_yy*sigma_yy twice and omits
    // eps_zz*sigma_zz (3D:1170-1172); 'corrected3d' uses the textbook sum.
    const T vx = p.vx[q], vy = p.vy[q], vz = p.vz[q];
    kin += (double)(T(0.5) * r * (vx * vx + vy * vy + vz * vz));
    const T sxx = p.sxx[q], syy = p.syy[q], szz = p.szz[q];
    const T sxy = p.sxy[q], sxz = p.sxz[q], syz = p.syz[q];
    const T den = T(2) * m * (T(3) * l + T(2) * m);
    const T exx = (T(2) * (l + m) * sxx - l * syy - l * szz) / den;
    const T eyy = (T(2) * (l + m) * syy - l * sxx - l * szz) / den;
    const T ezz = (T(2) * (l + m) * szz - l * sxx - l * syy) / den;
    const T exy = sxy / (T(2) * m), exz = sxz / (T(2) * m), eyz = syz / (T(2) * m);
    const T third = p.corrected3d ? ezz * szz : eyy * syy;
    pot += (double)(T(0.5) * (exx * sxx + eyy * syy + third + T(2) * exy * sxy +
                              T(2) * exz * sxz + T(2) * eyz * syz));
  }
}

}  // namespace cpml
