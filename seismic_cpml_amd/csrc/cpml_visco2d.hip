// 2D viscoelastic C-PML step for gfx950:
// seismic_CPML_2D_velocity_and_stress_second_order_viscoelastic.f90 with
// VISCOELASTIC_ATTENUATION (2D-visco2 below), and as ORDER 4 the
// ..._fourth_order_viscoelastic.f90 program (2D-visco4), which differs only
// in its difference operator:
//   stress + attenuation  2D-visco2:736-833
//   velocity              2D-visco2:839-870, source 913-918, edges 921-929
//   pressure seismogram   2D-visco2:958-990
// Arithmetic statement by statement as there (derivatives multiply by
// ONE_OVER_DELTA, 2D-visco2:209-210); built with -ffp-contract=off, so the
// float64 results equal the reference bit for bit.
//
// Attenuation: N_SLS standard linear solids per modulus, memory variables
// R_e1 (dilatation), R_e11 (deviatoric normal) and R_e13 (shear), updated by
// the ADE scheme of Robertsson, Blanch and Symes (1994) eq. 14:
//   R_new = (R_old + d * DELTAT_phi - R_old * DELTAT / (2 tau_sigma))
//           / (1 + DELTAT / (2 tau_sigma))
// and entering the stress increment as (R_new + R_old) / 2.  Each point reads
// only its own old value, so R is updated in place (the reference keeps an
// _old copy of every array and copies it each step, 2D-visco2:736-738).
//
// One point per thread, 64 x 4 blocks; R is stored dense (nx * ny per SLS
// plane) next to the padded wavefields.
#include <hip/hip_runtime.h>

#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

constexpr int VTY = 4;

// XCD-aware tile order (one-dimensional grid): workgroup h runs on XCD h % 8;
// each XCD gets one contiguous band of 64 x VTY tiles (x fastest), so the
// halo rows and columns neighbouring tiles share are read through one L2.
template <typename T>
__device__ __forceinline__ bool point2(const Visco2D<T> &v, int &i, int &j) {
  const int ntx = (v.s.nx + BX - 1) / BX, nt = ntx * ((v.s.ny + VTY - 1) / VTY);
  const int h = blockIdx.x, xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, r = nt & 7;
  const int t = xcd * q + min(xcd, r) + loc;
  i = (t % ntx) * BX + threadIdx.x;
  j = (t / ntx) * VTY + threadIdx.y;
  return i < v.s.nx && j < v.s.ny;
}

// staggered difference: (f+ - f0) * ONE_OVER_DELTA (2D-visco2), or
// (f+ - f0) * NINE_OVER_8_DELTA + (f- - f++) * ONE_OVER_24_DELTA (2D-visco4)
template <int ORDER, typename T>
__device__ __forceinline__ T dv(T fp, T f0, T fm, T fpp, T od, T n8, T o24) {
  if (ORDER == 2) return (fp - f0) * od;
  return (fp - f0) * n8 + (fm - fpp) * o24;
}

template <typename T, int ORDER>
__global__ __launch_bounds__(BX *VTY) void k_stress2dve(Visco2D<T> v) {
  int i, j;
  if (!point2(v, i, j)) return;
  const Step2D<T> &p = v.s;
  const long ld = p.ld, c = i + j * ld;
  const long r = i + (long)j * p.nx;  // dense index of the memory variables
  // the kernel reads v and writes sigma and R: restrict lets the reads of a
  // point issue together instead of waiting behind its stores
  const T *__restrict__ vx = p.vx;
  const T *__restrict__ vy = p.vy;
  T *__restrict__ sxx = p.sxx;
  T *__restrict__ syy = p.syy;
  T *__restrict__ sxy = p.sxy;
  T *__restrict__ re1 = v.re1;
  T *__restrict__ re11 = v.re11;
  T *__restrict__ re13 = v.re13;
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const long mxo = p.sx.s(i) + (long)p.sx.w * j;
  const long myo = i + p.mld * p.sy.s(j);
  // sigma_xx, sigma_yy: Fortran i = 1..NX-1, j = 2..NY (2D-visco2:742-800)
  if (i <= p.nx - 2 && j >= 1) {
    const T lh = T(0.5) * (p.lam[c + 1] + p.lam[c]);
    const T mh = T(0.5) * (p.mu[c + 1] + p.mu[c]);
    const T lpm = lh + mh;
    const T l2m = lh + T(2) * mh;
    T dvx_dx = dv<ORDER>(vx[c + 1], vx[c], vx[c - 1], vx[c + 2], v.odx, v.n8x, v.o24x);
    T dvy_dy = dv<ORDER>(vy[c], vy[c - ld], vy[c - 2 * ld], vy[c + ld], v.ody, v.n8y, v.o24y);
    if (inx) dvx_dx = cpml(dvx_dx, p.mx[0] + mxo, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
    if (iny) dvy_dy = cpml(dvy_dy, p.my[0] + myo, p.py.a[j], p.py.b[j], p.py.K[j]);
    T se1 = 0, se11 = 0;
    for (int l = 0; l < v.nsls; ++l) {
      T *r1 = re1 + r + l * v.nn, *r11 = re11 + r + l * v.nn;
      const T o1 = *r1, o11 = *r11;
      const T n1 = (o1 + (dvx_dx + dvy_dy) * v.P1[l] - o1 * v.H1[l]) * v.M1[l];
      const T n11 = (o11 + T(0.5) * (dvx_dx - dvy_dy) * v.P2[l] - o11 * v.H2[l]) * v.M2[l];
      *r1 = n1;
      *r11 = n11;
      se1 = se1 + n1 + o1;
      se11 = se11 + n11 + o11;
    }
    sxx[c] = sxx[c] + (l2m * dvx_dx + lh * dvy_dy + (T(0.5) * lpm * se1 + mh * se11)) * p.dt;
    syy[c] = syy[c] + (lh * dvx_dx + l2m * dvy_dy + (T(0.5) * lpm * se1 - mh * se11)) * p.dt;
  }
  // sigma_xy: Fortran i = 2..NX, j = 1..NY-1 (2D-visco2:802-833)
  if (i >= 1 && j <= p.ny - 2) {
    const T mhy = T(0.5) * (p.mu[c + ld] + p.mu[c]);
    T dvy_dx = dv<ORDER>(vy[c], vy[c - 1], vy[c - 2], vy[c + 1], v.odx, v.n8x, v.o24x);
    T dvx_dy = dv<ORDER>(vx[c + ld], vx[c], vx[c - ld], vx[c + 2 * ld], v.ody, v.n8y, v.o24y);
    if (inx) dvy_dx = cpml(dvy_dx, p.mx[1] + mxo, p.px.a[i], p.px.b[i], p.px.K[i]);
    if (iny) dvx_dy = cpml(dvx_dy, p.my[1] + myo, p.py.ah[j], p.py.bh[j], p.py.Kh[j]);
    T se13 = 0;
    for (int l = 0; l < v.nsls; ++l) {
      T *r13 = re13 + r + l * v.nn;
      const T o13 = *r13;
      const T n13 = (o13 + (dvy_dx + dvx_dy) * v.P2[l] - o13 * v.H2[l]) * v.M2[l];
      *r13 = n13;
      se13 = se13 + n13 + o13;
    }
    sxy[c] = sxy[c] + mhy * (dvy_dx + dvx_dy + T(0.5) * se13) * p.dt;
  }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(BX *VTY) void k_velocity2dve(Visco2D<T> v) {
  int i, j;
  if (!point2(v, i, j)) return;
  const Step2D<T> &p = v.s;
  const long ld = p.ld, c = i + j * ld;
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const long mxo = p.sx.s(i) + (long)p.sx.w * j;
  const long myo = i + p.mld * p.sy.s(j);
  const T *__restrict__ sxx = p.sxx;
  const T *__restrict__ syy = p.syy;
  const T *__restrict__ sxy = p.sxy;
  T *__restrict__ vx = p.vx;
  T *__restrict__ vy = p.vy;
  T vxn = vx[c], vyn = vy[c];
  // vx: Fortran i = 2..NX, j = 2..NY (2D-visco2:839-851)
  if (i >= 1 && j >= 1) {
    T dsxx_dx = dv<ORDER>(sxx[c], sxx[c - 1], sxx[c - 2], sxx[c + 1], v.odx, v.n8x, v.o24x);
    T dsxy_dy = dv<ORDER>(sxy[c], sxy[c - ld], sxy[c - 2 * ld], sxy[c + ld], v.ody, v.n8y,
                          v.o24y);
    if (inx) dsxx_dx = cpml(dsxx_dx, p.mx[2] + mxo, p.px.a[i], p.px.b[i], p.px.K[i]);
    if (iny) dsxy_dy = cpml(dsxy_dy, p.my[2] + myo, p.py.a[j], p.py.b[j], p.py.K[j]);
    vxn = vxn + (dsxx_dx + dsxy_dy) * p.dt / p.rho[c];
  }
  // vy: Fortran i = 1..NX-1, j = 1..NY-1 (2D-visco2:853-870)
  if (i <= p.nx - 2 && j <= p.ny - 2) {
    const T rhh = T(0.25) * (p.rho[c] + p.rho[c + 1] + p.rho[c + 1 + ld] + p.rho[c + ld]);
    T dsxy_dx = dv<ORDER>(sxy[c + 1], sxy[c], sxy[c - 1], sxy[c + 2], v.odx, v.n8x, v.o24x);
    T dsyy_dy = dv<ORDER>(syy[c + ld], syy[c], syy[c - ld], syy[c + 2 * ld], v.ody, v.n8y,
                          v.o24y);
    if (inx) dsxy_dx = cpml(dsxy_dx, p.mx[3] + mxo, p.px.ah[i], p.px.bh[i], p.px.Kh[i]);
    if (iny) dsyy_dy = cpml(dsyy_dy, p.my[3] + myo, p.py.ah[j], p.py.bh[j], p.py.Kh[j]);
    vyn = vyn + (dsxy_dx + dsyy_dy) * p.dt / rhh;
  }
  // point force (2D-visco2:913-918; the table holds force / (DELTAX*DELTAY))
  if (i == p.isrc && j == p.jsrc) {
    const int it = *p.step;
    const T rhh = T(0.25) * (p.rho[c] + p.rho[c + 1] + p.rho[c + 1 + ld] + p.rho[c + ld]);
    vxn = vxn + p.fx[it - 1] * p.dt / p.rho[c];
    vyn = vyn + p.fy[it - 1] * p.dt / rhh;
  }
  // rigid edges (2D-visco2:921-929)
  if (i == 0 || i == p.nx - 1 || j == 0 || j == p.ny - 1) {
    vxn = T(0);
    vyn = T(0);
  }
  vx[c] = vxn;
  vy[c] = vyn;
}

// Pressure at the receivers from sigma_xx, sigma_yy in plane strain
// (2D-visco2:979-990), before the record kernel advances the step.
template <typename T>
__global__ void k_pressure2d(Visco2D<T> v) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= v.nrec) return;
  const Step2D<T> &p = v.s;
  const long c = v.recidx[r];
  const int it = *p.step;
  const T lh = T(0.5) * (p.lam[c + 1] + p.lam[c]);
  const T mh = T(0.5) * (p.mu[c + 1] + p.mu[c]);
  const T den = T(4) * mh * (lh + mh);
  const T exx = ((lh + T(2) * mh) * p.sxx[c] - lh * p.syy[c]) / den;
  const T eyy = ((lh + T(2) * mh) * p.syy[c] - lh * p.sxx[c]) / den;
  const T two_thirds = T(2) / T(3);
  v.sisp[(long)(it - 1) * v.nrec + r] = -((lh + two_thirds * mh) * (exx + eyy));
}

}  // namespace

template <typename T>
hipError_t launch_stress2dve(int order, const Visco2D<T> &v, hipStream_t s) {
  const dim3 g(((v.s.nx + BX - 1) / BX) * ((v.s.ny + VTY - 1) / VTY));
  if (order == 4)
    hipLaunchKernelGGL((k_stress2dve<T, 4>), g, dim3(BX, VTY), 0, s, v);
  else
    hipLaunchKernelGGL((k_stress2dve<T, 2>), g, dim3(BX, VTY), 0, s, v);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity2dve(int order, const Visco2D<T> &v, hipStream_t s) {
  const dim3 g(((v.s.nx + BX - 1) / BX) * ((v.s.ny + VTY - 1) / VTY));
  if (order == 4)
    hipLaunchKernelGGL((k_velocity2dve<T, 4>), g, dim3(BX, VTY), 0, s, v);
  else
    hipLaunchKernelGGL((k_velocity2dve<T, 2>), g, dim3(BX, VTY), 0, s, v);
  if (v.nrec > 0)
    hipLaunchKernelGGL(k_pressure2d<T>, dim3((v.nrec + 63) / 64), dim3(64), 0, s, v);
  return hipGetLastError();
}

template hipError_t launch_stress2dve<float>(int, const Visco2D<float> &, hipStream_t);
template hipError_t launch_stress2dve<double>(int, const Visco2D<double> &, hipStream_t);
template hipError_t launch_velocity2dve<float>(int, const Visco2D<float> &, hipStream_t);
template hipError_t launch_velocity2dve<double>(int, const Visco2D<double> &, hipStream_t);

}  // namespace cpml
