// 3D 4th-order staggered C-PML step for gfx950: the elastic part of
// seismic_CPML_3D_viscoelastic_MPI.f90 (the program run with tau_epsilon =
// tau_sigma, where its attenuation memory variables stay zero):
//   stress    seismic_CPML_3D_viscoelastic_MPI.f90:977-1229
//   velocity  ...viscoelastic_MPI.f90:1244-1327, source 1329-1353,
//             Dirichlet 1355-1398 (planes 1 and N of every axis)
// on the whole grid of one device (no slab interfaces: the reference's MPI
// run reads zero ghost planes across each of its slab interfaces, which a
// single domain does not have).
//
// Difference operator, as written there:
//   (27 f(+1) - 27 f(0) - f(+2) + f(-1)) * ONE_OVER_DELTA / 24
// summed left to right, divided (not multiplied) by 24; the file is built
// with -ffp-contract=off, so float64 results equal the restatement bit for
// bit.  Lame coefficients per point with Mu_nu = 1 (3D-visco:981-985):
//   lambda_u = (lambda + (2/3) mu) - (2/3) mu,  lambda_u + 2 mu,  mu,
// and the stress increment is (sum of coefficient * derivative) * DELTAT.
//
// Mapping: a 64 x 4 block covers 64 points along x (one wavefront per row,
// coalesced) and 4 rows; each (i, j) column marches through a chunk of z
// planes.  The radius-2 neighbours are read from global memory; the x and y
// ones come from the same or adjacent rows of the block and the z ones from
// the planes the column just visited, so they hit in L1/L2.  Tiles are
// dealt to the 8 XCDs in contiguous bands (neighbouring tiles share an L2).
#include <hip/hip_runtime.h>

#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

constexpr int TY4 = 4;  // rows per block

struct Col4 {
  int i, j, kb, ke;
  bool ok;
};

__device__ __forceinline__ Col4 col4(int nx, int ny, int nz, int kchunk) {
  const int ntx = (nx + BX - 1) / BX, nty = (ny + TY4 - 1) / TY4;
  const int ntz = (nz + kchunk - 1) / kchunk;
  const int nt = ntx * nty * ntz;
  const int h = blockIdx.x;
  const int xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, r = nt & 7;
  const int t = xcd * q + min(xcd, r) + loc;  // band of XCD `xcd`, position loc
  const int tx = t % ntx, rest = t / ntx;
  const int ty = rest % nty, tz = zchunk_of(rest / nty, ntz);
  Col4 c;
  c.i = tx * BX + threadIdx.x;
  c.j = ty * TY4 + threadIdx.y;
  c.kb = tz * kchunk;
  c.ke = min(c.kb + kchunk, nz);
  c.ok = c.i < nx && c.j < ny;
  return c;
}

// (27 fp - 27 f0 - fpp + fm) * od / 24, in the reference's order
template <typename T>
__device__ __forceinline__ T d4(T fp, T f0, T fpp, T fm, T od) {
  return (((T(27) * fp - T(27) * f0) - fpp) + fm) * od / T(24);
}

template <typename T>
__global__ __launch_bounds__(BX *TY4) void k_stress3d4(Step3D<T> p) {
  const Col4 c = col4(p.nx, p.ny, p.nz, p.kchunk);
  if (!c.ok) return;
  const int i = c.i, j = c.j, nx = p.nx, ny = p.ny, nz = p.nz;
  const long ld = p.ld, ps = p.ps;
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const long mxo = p.sx.s(i) + (long)p.sx.w * j, mxs = (long)p.sx.w * ny;
  const long myo = i + p.mld * p.sy.s(j), mys = p.mld * p.sy.w;
  const long mzo = i + p.mld * j, mzs = p.mld * ny;
  // loop ranges of 3D-visco:977-979, 1080-1082, 1141-1144, 1185-1186 (0-based)
  const bool r_n = i <= nx - 2 && j >= 1;
  const bool r_xy = i >= 1 && j <= ny - 2;
  const bool r_xz = i >= 1;
  const bool r_yz = j <= ny - 2;
  const T c23 = T(2) / T(3);  // 2.d0/DIM
  for (int k = c.kb; k < c.ke; ++k) {
    const long q = i + j * ld + k * ps;
    const bool inz = p.sz.in(k);
    const long mzk = mzo + (inz ? p.sz.s(k) : 0) * mzs;
    const T *vx = p.vx + q, *vy = p.vy + q, *vz = p.vz + q;
    const T mu = p.mu[q];
    if (r_n && k >= 1) {
      const T lam = p.lam[q];
      const T lu = (lam + c23 * mu) - c23 * mu;
      const T l2u = lu + T(2) * mu;
      T dvx_dx = d4(vx[1], vx[0], vx[2], vx[-1], p.odx);
      T dvy_dy = d4(vy[0], vy[-ld], vy[ld], vy[-2 * ld], p.ody);
      T dvz_dz = d4(vz[0], vz[-ps], vz[ps], vz[-2 * ps], p.odz);
      if (inx) dvx_dx = cpml(dvx_dx, p.mx[0] + mxo + k * mxs, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
      if (iny) dvy_dy = cpml(dvy_dy, p.my[0] + myo + k * mys, p.py.a[j], p.py.b[j], p.py.K[j]);
      if (inz) dvz_dz = cpml(dvz_dz, p.mz[0] + mzk, p.pz.a[k], p.pz.b[k], p.pz.K[k]);
      p.sxx[q] = p.sxx[q] + (l2u * dvx_dx + lu * dvy_dy + lu * dvz_dz) * p.dt;
      p.syy[q] = p.syy[q] + (lu * dvx_dx + l2u * dvy_dy + lu * dvz_dz) * p.dt;
      p.szz[q] = p.szz[q] + (lu * dvx_dx + lu * dvy_dy + l2u * dvz_dz) * p.dt;
    }
    if (r_xy) {
      T dvy_dx = d4(vy[0], vy[-1], vy[1], vy[-2], p.odx);
      T dvx_dy = d4(vx[ld], vx[0], vx[2 * ld], vx[-ld], p.ody);
      if (inx) dvy_dx = cpml(dvy_dx, p.mx[1] + mxo + k * mxs, p.px.a[i], p.px.b[i], p.px.K[i]);
      if (iny)
        dvx_dy = cpml(dvx_dy, p.my[1] + myo + k * mys, p.py.ah[j], p.py.bh[j], p.py.Kh[j]);
      p.sxy[q] = p.sxy[q] + mu * (dvx_dy + dvy_dx) * p.dt;
    }
    if (r_xz && k <= nz - 2) {
      T dvz_dx = d4(vz[0], vz[-1], vz[1], vz[-2], p.odx);
      T dvx_dz = d4(vx[ps], vx[0], vx[2 * ps], vx[-ps], p.odz);
      if (inx) dvz_dx = cpml(dvz_dx, p.mx[2] + mxo + k * mxs, p.px.a[i], p.px.b[i], p.px.K[i]);
      if (inz) dvx_dz = cpml(dvx_dz, p.mz[1] + mzk, p.pz.ah[k], p.pz.bh[k], p.pz.Kh[k]);
      p.sxz[q] = p.sxz[q] + mu * (dvx_dz + dvz_dx) * p.dt;
    }
    if (r_yz && k <= nz - 2) {
      T dvz_dy = d4(vz[ld], vz[0], vz[2 * ld], vz[-ld], p.ody);
      T dvy_dz = d4(vy[ps], vy[0], vy[2 * ps], vy[-ps], p.odz);
      if (iny)
        dvz_dy = cpml(dvz_dy, p.my[2] + myo + k * mys, p.py.ah[j], p.py.bh[j], p.py.Kh[j]);
      if (inz) dvy_dz = cpml(dvy_dz, p.mz[2] + mzk, p.pz.ah[k], p.pz.bh[k], p.pz.Kh[k]);
      p.syz[q] = p.syz[q] + mu * (dvy_dz + dvz_dy) * p.dt;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(BX *TY4) void k_velocity3d4(Step3D<T> p) {
  const Col4 c = col4(p.nx, p.ny, p.nz, p.kchunk);
  if (!c.ok) return;
  const int i = c.i, j = c.j, nx = p.nx, ny = p.ny, nz = p.nz;
  const long ld = p.ld, ps = p.ps;
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const long mxo = p.sx.s(i) + (long)p.sx.w * j, mxs = (long)p.sx.w * ny;
  const long myo = i + p.mld * p.sy.s(j), mys = p.mld * p.sy.w;
  const long mzo = i + p.mld * j, mzs = p.mld * ny;
  // loop ranges of 3D-visco:1244-1246, 1266-1267, 1287-1289 (0-based)
  const bool r_vx = i >= 1 && j >= 1;
  const bool r_vy = i <= nx - 2 && j <= ny - 2;
  const bool r_vz = i <= nx - 2 && j >= 1;
  const bool edge_xy = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
  const bool src_col = i == p.isrc && j == p.jsrc;
  for (int k = c.kb; k < c.ke; ++k) {
    const long q = i + j * ld + k * ps;
    const bool inz = p.sz.in(k);
    const long mzk = mzo + (inz ? p.sz.s(k) : 0) * mzs;
    const T *sxx = p.sxx + q, *syy = p.syy + q, *szz = p.szz + q;
    const T *sxy = p.sxy + q, *sxz = p.sxz + q, *syz = p.syz + q;
    const T cb = p.cbuoy[q];  // DELTAT_over_rho
    T vxn = p.vx[q], vyn = p.vy[q], vzn = p.vz[q];
    if (r_vx && k >= 1) {
      T dsxx_dx = d4(sxx[0], sxx[-1], sxx[1], sxx[-2], p.odx);
      T dsxy_dy = d4(sxy[0], sxy[-ld], sxy[ld], sxy[-2 * ld], p.ody);
      T dsxz_dz = d4(sxz[0], sxz[-ps], sxz[ps], sxz[-2 * ps], p.odz);
      if (inx) dsxx_dx = cpml(dsxx_dx, p.mx[3] + mxo + k * mxs, p.px.a[i], p.px.b[i], p.px.K[i]);
      if (iny) dsxy_dy = cpml(dsxy_dy, p.my[3] + myo + k * mys, p.py.a[j], p.py.b[j], p.py.K[j]);
      if (inz) dsxz_dz = cpml(dsxz_dz, p.mz[3] + mzk, p.pz.a[k], p.pz.b[k], p.pz.K[k]);
      vxn = cb * (dsxx_dx + dsxy_dy + dsxz_dz) + vxn;
    }
    if (r_vy && k >= 1) {
      T dsxy_dx = d4(sxy[1], sxy[0], sxy[2], sxy[-1], p.odx);
      T dsyy_dy = d4(syy[ld], syy[0], syy[2 * ld], syy[-ld], p.ody);
      T dsyz_dz = d4(syz[0], syz[-ps], syz[ps], syz[-2 * ps], p.odz);
      if (inx)
        dsxy_dx = cpml(dsxy_dx, p.mx[4] + mxo + k * mxs, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
      if (iny)
        dsyy_dy = cpml(dsyy_dy, p.my[4] + myo + k * mys, p.py.ah[j], p.py.bh[j], p.py.Kh[j]);
      if (inz) dsyz_dz = cpml(dsyz_dz, p.mz[4] + mzk, p.pz.a[k], p.pz.b[k], p.pz.K[k]);
      vyn = cb * (dsxy_dx + dsyy_dy + dsyz_dz) + vyn;
    }
    if (r_vz && k <= nz - 2) {
      T dsxz_dx = d4(sxz[1], sxz[0], sxz[2], sxz[-1], p.odx);
      T dsyz_dy = d4(syz[0], syz[-ld], syz[ld], syz[-2 * ld], p.ody);
      T dszz_dz = d4(szz[ps], szz[0], szz[2 * ps], szz[-ps], p.odz);
      if (inx)
        dsxz_dx = cpml(dsxz_dx, p.mx[5] + mxo + k * mxs, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
      if (iny) dsyz_dy = cpml(dsyz_dy, p.my[5] + myo + k * mys, p.py.a[j], p.py.b[j], p.py.K[j]);
      if (inz) dszz_dz = cpml(dszz_dz, p.mz[5] + mzk, p.pz.ah[k], p.pz.bh[k], p.pz.Kh[k]);
      vzn = cb * (dsxz_dx + dsyz_dy + dszz_dz) + vzn;
    }
    // point force on the cut plane (3D-visco:1350-1352)
    if (src_col && k == p.ksrc) {
      const int it = *p.step;
      const T rs = p.rho[q];
      vxn = vxn + p.fx[it - 1] * p.dt / rs;
      vyn = vyn + p.fy[it - 1] * p.dt / rs;
    }
    // rigid faces: planes 1 and N of every axis (3D-visco:1355-1398)
    if (edge_xy || k == 0 || k == nz - 1) {
      vxn = T(0);
      vyn = T(0);
      vzn = T(0);
    }
    p.vx[q] = vxn;
    p.vy[q] = vyn;
    p.vz[q] = vzn;
  }
}

template <typename T>
unsigned ntiles4(const Step3D<T> &p) {
  return (unsigned)((p.nx + BX - 1) / BX) * ((p.ny + TY4 - 1) / TY4) *
         ((p.nz + p.kchunk - 1) / p.kchunk);
}

}  // namespace

template <typename T>
hipError_t launch_stress3d4(const Step3D<T> &p, hipStream_t s) {
  hipLaunchKernelGGL(k_stress3d4<T>, dim3(ntiles4(p)), dim3(BX, TY4), 0, s, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity3d4(const Step3D<T> &p, hipStream_t s) {
  hipLaunchKernelGGL(k_velocity3d4<T>, dim3(ntiles4(p)), dim3(BX, TY4), 0, s, p);
  return hipGetLastError();
}

template hipError_t launch_stress3d4<float>(const Step3D<float> &, hipStream_t);
template hipError_t launch_stress3d4<double>(const Step3D<double> &, hipStream_t);
template hipError_t launch_velocity3d4<float>(const Step3D<float> &, hipStream_t);
template hipError_t launch_velocity3d4<double>(const Step3D<double> &, hipStream_t);

}  // namespace cpml
