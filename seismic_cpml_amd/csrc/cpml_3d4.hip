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
// Mapping: a 64 x TY block (TY = 4, 8 or 16 rows, the handle's tile) covers
// 64 points along x (one wavefront per row, coalesced); each (i, j) column
// marches through a chunk of z
// planes.  The column's own z neighbours are carried in registers from plane
// to plane (one new plane per field and step); the x and y neighbours are
// read from global memory and come from the same or adjacent rows of the
// block, so they hit in L1/L2.  Profile values of the column are loaded once.  Tiles are
// dealt to the 8 XCDs in contiguous bands (neighbouring tiles share an L2).
#include <hip/hip_runtime.h>

#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

struct Col4 {
  int i, j, kb, ke;
  bool ok;
};

template <int TY4>
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

// Every read of a point is issued before its arithmetic and stores (all
// offsets stay inside the two halo cells of each side, so the loads are
// unconditional), through restrict-qualified pointers: the loads of one
// point are in flight together instead of waiting behind its stores.
template <typename T, int TY4>
__global__ __launch_bounds__(BX *TY4) void k_stress3d4(Step3D<T> p) {
  const Col4 c = col4<TY4>(p.nx, p.ny, p.nz, p.kchunk);
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
  T xa = 0, xb = 0, xK = 1, xah = 0, xbh = 0, xKh = 1;
  T ya = 0, yb = 0, yK = 1, yah = 0, ybh = 0, yKh = 1;
  if (inx) {
    xa = p.px.a[i]; xb = p.px.b[i]; xK = p.px.K[i];
    xah = p.px.ah[i]; xbh = p.px.bh[i]; xKh = p.px.Kh[i];
  }
  if (iny) {
    ya = p.py.a[j]; yb = p.py.b[j]; yK = p.py.K[j];
    yah = p.py.ah[j]; ybh = p.py.bh[j]; yKh = p.py.Kh[j];
  }
  // the column's own z neighbours, carried from plane to plane
  const long q0 = i + j * ld + c.kb * ps;
  T vx_zm = p.vx[q0 - ps], vx0 = p.vx[q0], vx_zp = p.vx[q0 + ps];
  T vy_zm = p.vy[q0 - ps], vy0 = p.vy[q0], vy_zp = p.vy[q0 + ps];
  T vz_zmm = p.vz[q0 - 2 * ps], vz_zm = p.vz[q0 - ps], vz0 = p.vz[q0];
  for (int k = c.kb; k < c.ke; ++k) {
    const long q = i + j * ld + k * ps;
    const T *__restrict__ vx = p.vx + q;
    const T *__restrict__ vy = p.vy + q;
    const T *__restrict__ vz = p.vz + q;
    T *__restrict__ sxx = p.sxx + q;
    T *__restrict__ syy = p.syy + q;
    T *__restrict__ szz = p.szz + q;
    T *__restrict__ sxy = p.sxy + q;
    T *__restrict__ sxz = p.sxz + q;
    T *__restrict__ syz = p.syz + q;
    const T vx_xp = vx[1], vx_xpp = vx[2], vx_xm = vx[-1];
    const T vx_yp = vx[ld], vx_ypp = vx[2 * ld], vx_ym = vx[-ld];
    const T vx_zpp = vx[2 * ps];
    const T vy_xm = vy[-1], vy_xp = vy[1], vy_xmm = vy[-2];
    const T vy_ym = vy[-ld], vy_yp = vy[ld], vy_ymm = vy[-2 * ld];
    const T vy_zpp = vy[2 * ps];
    const T vz_xm = vz[-1], vz_xp = vz[1], vz_xmm = vz[-2];
    const T vz_yp = vz[ld], vz_ypp = vz[2 * ld], vz_ym = vz[-ld];
    const T vz_zp = vz[ps];
    const T mu = p.mu[q], lam = p.lam[q];
    const T oxx = sxx[0], oyy = syy[0], ozz = szz[0], oxy = sxy[0], oxz = sxz[0], oyz = syz[0];
    const bool inz = p.sz.in(k);
    const long mzk = mzo + (inz ? p.sz.s(k) : 0) * mzs;
    T za = 0, zb = 0, zK = 1, zah = 0, zbh = 0, zKh = 1;
    if (inz) {
      za = p.pz.a[k]; zb = p.pz.b[k]; zK = p.pz.K[k];
      zah = p.pz.ah[k]; zbh = p.pz.bh[k]; zKh = p.pz.Kh[k];
    }
    if (r_n && k >= 1) {
      const T lu = (lam + c23 * mu) - c23 * mu;
      const T l2u = lu + T(2) * mu;
      T dvx_dx = d4(vx_xp, vx0, vx_xpp, vx_xm, p.odx);
      T dvy_dy = d4(vy0, vy_ym, vy_yp, vy_ymm, p.ody);
      T dvz_dz = d4(vz0, vz_zm, vz_zp, vz_zmm, p.odz);
      if (inx) dvx_dx = cpml(dvx_dx, p.mx[0] + mxo + k * mxs, xah, xbh, xKh);
      if (iny) dvy_dy = cpml(dvy_dy, p.my[0] + myo + k * mys, ya, yb, yK);
      if (inz) dvz_dz = cpml(dvz_dz, p.mz[0] + mzk, za, zb, zK);
      sxx[0] = oxx + (l2u * dvx_dx + lu * dvy_dy + lu * dvz_dz) * p.dt;
      syy[0] = oyy + (lu * dvx_dx + l2u * dvy_dy + lu * dvz_dz) * p.dt;
      szz[0] = ozz + (lu * dvx_dx + lu * dvy_dy + l2u * dvz_dz) * p.dt;
    }
    if (r_xy) {
      T dvy_dx = d4(vy0, vy_xm, vy_xp, vy_xmm, p.odx);
      T dvx_dy = d4(vx_yp, vx0, vx_ypp, vx_ym, p.ody);
      if (inx) dvy_dx = cpml(dvy_dx, p.mx[1] + mxo + k * mxs, xa, xb, xK);
      if (iny) dvx_dy = cpml(dvx_dy, p.my[1] + myo + k * mys, yah, ybh, yKh);
      sxy[0] = oxy + mu * (dvx_dy + dvy_dx) * p.dt;
    }
    if (r_xz && k <= nz - 2) {
      T dvz_dx = d4(vz0, vz_xm, vz_xp, vz_xmm, p.odx);
      T dvx_dz = d4(vx_zp, vx0, vx_zpp, vx_zm, p.odz);
      if (inx) dvz_dx = cpml(dvz_dx, p.mx[2] + mxo + k * mxs, xa, xb, xK);
      if (inz) dvx_dz = cpml(dvx_dz, p.mz[1] + mzk, zah, zbh, zKh);
      sxz[0] = oxz + mu * (dvx_dz + dvz_dx) * p.dt;
    }
    if (r_yz && k <= nz - 2) {
      T dvz_dy = d4(vz_yp, vz0, vz_ypp, vz_ym, p.ody);
      T dvy_dz = d4(vy_zp, vy0, vy_zpp, vy_zm, p.odz);
      if (iny) dvz_dy = cpml(dvz_dy, p.my[2] + myo + k * mys, yah, ybh, yKh);
      if (inz) dvy_dz = cpml(dvy_dz, p.mz[2] + mzk, zah, zbh, zKh);
      syz[0] = oyz + mu * (dvy_dz + dvz_dy) * p.dt;
    }
    vx_zm = vx0, vx0 = vx_zp, vx_zp = vx_zpp;
    vy_zm = vy0, vy0 = vy_zp, vy_zp = vy_zpp;
    vz_zmm = vz_zm, vz_zm = vz0, vz0 = vz_zp;
  }
}

// Float: at least 6 waves per SIMD.  The register limit that sets (80 VGPRs,
// against 82 unbounded, which leaves 5) costs no spill here, and the step ran
// 1.2% faster at 768^3; the stress kernel spills under the same bound and
// keeps its 90 registers, and so does the double velocity kernel.
template <typename T, int TY4>
__global__ __launch_bounds__(BX *TY4, sizeof(T) == 4 ? 6 : 1) void k_velocity3d4(Step3D<T> p) {
  const Col4 c = col4<TY4>(p.nx, p.ny, p.nz, p.kchunk);
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
  T xa = 0, xb = 0, xK = 1, xah = 0, xbh = 0, xKh = 1;
  T ya = 0, yb = 0, yK = 1, yah = 0, ybh = 0, yKh = 1;
  if (inx) {
    xa = p.px.a[i]; xb = p.px.b[i]; xK = p.px.K[i];
    xah = p.px.ah[i]; xbh = p.px.bh[i]; xKh = p.px.Kh[i];
  }
  if (iny) {
    ya = p.py.a[j]; yb = p.py.b[j]; yK = p.py.K[j];
    yah = p.py.ah[j]; ybh = p.py.bh[j]; yKh = p.py.Kh[j];
  }
  const long q0 = i + j * ld + c.kb * ps;
  T sxz_zmm = p.sxz[q0 - 2 * ps], sxz_zm = p.sxz[q0 - ps], sxz0 = p.sxz[q0];
  T syz_zmm = p.syz[q0 - 2 * ps], syz_zm = p.syz[q0 - ps], syz0 = p.syz[q0];
  T szz_zm = p.szz[q0 - ps], szz0 = p.szz[q0], szz_zp = p.szz[q0 + ps];
  for (int k = c.kb; k < c.ke; ++k) {
    const long q = i + j * ld + k * ps;
    const T *__restrict__ sxx = p.sxx + q;
    const T *__restrict__ syy = p.syy + q;
    const T *__restrict__ szz = p.szz + q;
    const T *__restrict__ sxy = p.sxy + q;
    const T *__restrict__ sxz = p.sxz + q;
    const T *__restrict__ syz = p.syz + q;
    T *__restrict__ vx = p.vx + q;
    T *__restrict__ vy = p.vy + q;
    T *__restrict__ vz = p.vz + q;
    const T sxx0 = sxx[0], sxx_xm = sxx[-1], sxx_xp = sxx[1], sxx_xmm = sxx[-2];
    const T sxy0 = sxy[0], sxy_ym = sxy[-ld], sxy_yp = sxy[ld], sxy_ymm = sxy[-2 * ld];
    const T sxy_xp = sxy[1], sxy_xpp = sxy[2], sxy_xm = sxy[-1];
    const T sxz_zp = sxz[ps];
    const T sxz_xp = sxz[1], sxz_xpp = sxz[2], sxz_xm = sxz[-1];
    const T syy0 = syy[0], syy_yp = syy[ld], syy_ypp = syy[2 * ld], syy_ym = syy[-ld];
    const T syz_zp = syz[ps];
    const T syz_ym = syz[-ld], syz_yp = syz[ld], syz_ymm = syz[-2 * ld];
    const T szz_zpp = szz[2 * ps];
    const T cb = p.cbuoy[q];  // DELTAT_over_rho
    T vxn = vx[0], vyn = vy[0], vzn = vz[0];
    const bool inz = p.sz.in(k);
    const long mzk = mzo + (inz ? p.sz.s(k) : 0) * mzs;
    T za = 0, zb = 0, zK = 1, zah = 0, zbh = 0, zKh = 1;
    if (inz) {
      za = p.pz.a[k]; zb = p.pz.b[k]; zK = p.pz.K[k];
      zah = p.pz.ah[k]; zbh = p.pz.bh[k]; zKh = p.pz.Kh[k];
    }
    if (r_vx && k >= 1) {
      T dsxx_dx = d4(sxx0, sxx_xm, sxx_xp, sxx_xmm, p.odx);
      T dsxy_dy = d4(sxy0, sxy_ym, sxy_yp, sxy_ymm, p.ody);
      T dsxz_dz = d4(sxz0, sxz_zm, sxz_zp, sxz_zmm, p.odz);
      if (inx) dsxx_dx = cpml(dsxx_dx, p.mx[3] + mxo + k * mxs, xa, xb, xK);
      if (iny) dsxy_dy = cpml(dsxy_dy, p.my[3] + myo + k * mys, ya, yb, yK);
      if (inz) dsxz_dz = cpml(dsxz_dz, p.mz[3] + mzk, za, zb, zK);
      vxn = cb * (dsxx_dx + dsxy_dy + dsxz_dz) + vxn;
    }
    if (r_vy && k >= 1) {
      T dsxy_dx = d4(sxy_xp, sxy0, sxy_xpp, sxy_xm, p.odx);
      T dsyy_dy = d4(syy_yp, syy0, syy_ypp, syy_ym, p.ody);
      T dsyz_dz = d4(syz0, syz_zm, syz_zp, syz_zmm, p.odz);
      if (inx) dsxy_dx = cpml(dsxy_dx, p.mx[4] + mxo + k * mxs, xah, xbh, xKh);
      if (iny) dsyy_dy = cpml(dsyy_dy, p.my[4] + myo + k * mys, yah, ybh, yKh);
      if (inz) dsyz_dz = cpml(dsyz_dz, p.mz[4] + mzk, za, zb, zK);
      vyn = cb * (dsxy_dx + dsyy_dy + dsyz_dz) + vyn;
    }
    if (r_vz && k <= nz - 2) {
      T dsxz_dx = d4(sxz_xp, sxz0, sxz_xpp, sxz_xm, p.odx);
      T dsyz_dy = d4(syz0, syz_ym, syz_yp, syz_ymm, p.ody);
      T dszz_dz = d4(szz_zp, szz0, szz_zpp, szz_zm, p.odz);
      if (inx) dsxz_dx = cpml(dsxz_dx, p.mx[5] + mxo + k * mxs, xah, xbh, xKh);
      if (iny) dsyz_dy = cpml(dsyz_dy, p.my[5] + myo + k * mys, ya, yb, yK);
      if (inz) dszz_dz = cpml(dszz_dz, p.mz[5] + mzk, zah, zbh, zKh);
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
    vx[0] = vxn;
    vy[0] = vyn;
    vz[0] = vzn;
    sxz_zmm = sxz_zm, sxz_zm = sxz0, sxz0 = sxz_zp;
    syz_zmm = syz_zm, syz_zm = syz0, syz0 = syz_zp;
    szz_zm = szz0, szz0 = szz_zp, szz_zp = szz_zpp;
  }
}

template <typename T>
unsigned ntiles4(const Step3D<T> &p, int ty) {
  return (unsigned)((p.nx + BX - 1) / BX) * ((p.ny + ty - 1) / ty) *
         ((p.nz + p.kchunk - 1) / p.kchunk);
}

}  // namespace

// p.tile_y selects the rows per block (4, 8 or 16)
template <typename T>
hipError_t launch_stress3d4(const Step3D<T> &p, hipStream_t s) {
  if (p.tile_y == 16)
    hipLaunchKernelGGL((k_stress3d4<T, 16>), dim3(ntiles4(p, 16)), dim3(BX, 16), 0, s, p);
  else if (p.tile_y == 8)
    hipLaunchKernelGGL((k_stress3d4<T, 8>), dim3(ntiles4(p, 8)), dim3(BX, 8), 0, s, p);
  else
    hipLaunchKernelGGL((k_stress3d4<T, 4>), dim3(ntiles4(p, 4)), dim3(BX, 4), 0, s, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity3d4(const Step3D<T> &p, hipStream_t s) {
  if (p.tile_y == 16)
    hipLaunchKernelGGL((k_velocity3d4<T, 16>), dim3(ntiles4(p, 16)), dim3(BX, 16), 0, s, p);
  else if (p.tile_y == 8)
    hipLaunchKernelGGL((k_velocity3d4<T, 8>), dim3(ntiles4(p, 8)), dim3(BX, 8), 0, s, p);
  else
    hipLaunchKernelGGL((k_velocity3d4<T, 4>), dim3(ntiles4(p, 4)), dim3(BX, 4), 0, s, p);
  return hipGetLastError();
}

template hipError_t launch_stress3d4<float>(const Step3D<float> &, hipStream_t);
template hipError_t launch_stress3d4<double>(const Step3D<double> &, hipStream_t);
template hipError_t launch_velocity3d4<float>(const Step3D<float> &, hipStream_t);
template hipError_t launch_velocity3d4<double>(const Step3D<double> &, hipStream_t);

}  // namespace cpml
