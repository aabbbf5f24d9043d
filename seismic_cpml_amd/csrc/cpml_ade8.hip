// ADE-PML RK4 step of seismic_ADEPML_2D_elastic_RK4_eighth_order.f90 (ADE8)
// for gfx950: the 8th-order (Holberg) staggered operator, four Runge-Kutta
// stages per time step and the auxiliary-differential-equation PML whose
// coefficients depend on the stage (ADE8:692-906).  Per stage one pass
// (k_ade_stage): velocity rates from the stage state's sigma, ADE memories in
// the PML band, force at t = (it-1 + rk41) * DELTAT, rate borders zeroed;
// stress rates from its v; then at once X += rk42 * rate * DELTAT, velocity
// borders 0, and the next stage's state X0 + rk41 * rate * DELTAT.  The
// rates stay in registers: 4 launches per step instead of the program's 13
// array sweeps.
//
// X ping-pongs between two state buffers: the step reads X0 (the program's
// d3, ADE8:697-703) from the buffer the step starts in, which it never
// writes, and accumulates into the other one.  Stage 1's state is
// X0 + rk41(1) * X * DELTAT with rk41(1) = 0, i.e. X0 itself bit for bit
// (x + 0 * x * dt == x for every finite x, signed zeros included), so stage
// 1 reads it from X0 and no copy of the state is made.
// Arithmetic statement by statement as there (8th-order difference divided
// by DELTA; -ffp-contract=off), so float64 results equal the reference.
// One point per thread over the interior, 64 x 4 blocks; the ghost cells of
// every array stay zero, as the program's 4-cell borders do.
#include <hip/hip_runtime.h>

#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

// rows per 64-column block; measured at 4096^2 (float, full step): 1 row
// 8.84, 2 rows 9.00, 4 rows 8.91, 8 rows 8.32, 16 rows 8.13 Gcell/s
constexpr int ATY = 2;

// XCD-aware tile order (one-dimensional grid): the hardware hands workgroup h
// to XCD h % 8; each XCD gets one contiguous band of 64 x ATY tiles (x
// fastest), so the halo columns and rows that neighbouring tiles share are
// read through the same L2.
template <typename T>
__device__ __forceinline__ bool pt(const Ade8<T> &a, int &i, int &j) {
  const int ntx = (a.s.nx + BX - 1) / BX, nt = ntx * ((a.s.ny + ATY - 1) / ATY);
  const int h = blockIdx.x, xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, r = nt & 7;
  const int t = xcd * q + min(xcd, r) + loc;
  i = (t % ntx) * BX + threadIdx.x;
  j = (t / ntx) * ATY + threadIdx.y;
  return i < a.s.nx && j < a.s.ny;
}

// (c1 d1 + c2 d2 + c3 d3 + c4 d4) / delta, ADE8:263-266, 720-723
template <typename T>
__device__ __forceinline__ T h8(T d1, T d2, T d3, T d4, T del) {
  return (((T(kHolberg1) * d1 + T(kHolberg2) * d2) + T(kHolberg3) * d3) + T(kHolberg4) * d4) / del;
}

__device__ __forceinline__ bool band(int I, int J, int nx, int ny, int npml) {  // Fortran indices
  return I <= npml + 2 || I >= nx - npml - 2 || J <= npml + 2 || J >= ny - npml - 2;
}

template <typename T>
__device__ __forceinline__ T rk41(int inc) {
  return inc == 0 ? T(0) : inc == 3 ? T(1) : T(0.5);
}
template <typename T>
__device__ __forceinline__ T rk42(int inc) {
  return (inc == 0 || inc == 3) ? T(1) / T(6) : T(2) / T(6);
}

// One Runge-Kutta stage in one pass (ADE8:715-874): the rates of the stage
// state S (X0 in stage 1, else d[in]) at this point -- velocity rates from
// S's stresses with the ADE memories in the PML band and the stage-time
// force, rate borders zeroed (ADE8:718-810); stress rates from S's
// velocities, whose border values the program has zeroed by then
// (ADE8:803-861) -- and at once the accumulation X += rk42 * rate * DELTAT
// into the step's output buffer (stage 1 starts from X0) with the velocity
// borders zeroed, and the next stage's state d[out] =
// X0 + rk41(inc+1) * rate * DELTAT (ADE8:863-874; X0's velocity borders
// count as zero, as the program's d3 after ADE8:803-810).  The rates never
// go to memory, and S is read from one buffer while the next state goes to
// another, so neighbours always see the stage's own state.
template <typename T>
__global__ __launch_bounds__(BX *ATY) void k_ade_stage(Ade8<T> a, int inc, int in, int out) {
  int i, j;
  if (!pt(a, i, j)) return;
  const Step2D<T> &p = a.s;
  const long ld = p.ld, c = i + j * ld;
  const int nx = p.nx, ny = p.ny, I = i + 1, J = j + 1;
  const bool bd = band(I, J, nx, ny, a.npml);
  const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
  const T *const X0[5] = {p.vx, p.vy, p.sxx, p.syy, p.sxy};
  const T *const *st = inc == 0 ? X0 : a.d[in];
  const T *sxx = st[2], *syy = st[3], *sxy = st[4];
  const T *vxs = st[0], *vys = st[1];
  T rate[5] = {T(0), T(0), T(0), T(0), T(0)};
  // vx rates: Fortran i = 2..NX, j = 2..NY (ADE8:718-744)
  if (i >= 1 && j >= 1) {
    T vx_ = h8(sxx[c] - sxx[c - 1], sxx[c + 1] - sxx[c - 2], sxx[c + 2] - sxx[c - 3],
               sxx[c + 3] - sxx[c - 4], a.dx);
    T vy_ = h8(sxy[c] - sxy[c - ld], sxy[c + ld] - sxy[c - 2 * ld], sxy[c + 2 * ld] - sxy[c - 3 * ld],
               sxy[c + 3 * ld] - sxy[c - 4 * ld], a.dy);
    if (bd) {
      T *m4 = a.mem[4] + c, *m7 = a.mem[7] + c;
      *m4 = a.bx[inc * nx + i] * *m4 + a.ax[inc * nx + i] * vx_;
      *m7 = a.by[inc * ny + j] * *m7 + a.ay[inc * ny + j] * vy_;
      vx_ = vx_ / p.px.K[i] + *m4;
      vy_ = vy_ / p.py.K[j] + *m7;
    }
    rate[0] = (vx_ + vy_) / p.rho[c];
  }
  // vy rates: Fortran i = 1..NX-1, j = 1..NY-1 (ADE8:746-768)
  if (i <= nx - 2 && j <= ny - 2) {
    const T rh = T(0.25) * (p.rho[c] + p.rho[c + 1] + p.rho[c + 1 + ld] + p.rho[c + ld]);
    T sx_ = h8(sxy[c + 1] - sxy[c], sxy[c + 2] - sxy[c - 1], sxy[c + 3] - sxy[c - 2],
               sxy[c + 4] - sxy[c - 3], a.dx);
    T sy_ = h8(syy[c + ld] - syy[c], syy[c + 2 * ld] - syy[c - ld], syy[c + 3 * ld] - syy[c - 2 * ld],
               syy[c + 4 * ld] - syy[c - 3 * ld], a.dy);
    if (bd) {
      T *m6 = a.mem[6] + c, *m5 = a.mem[5] + c;
      *m6 = a.bxh[inc * nx + i] * *m6 + a.axh[inc * nx + i] * sx_;
      *m5 = a.byh[inc * ny + j] * *m5 + a.ayh[inc * ny + j] * sy_;
      sx_ = sx_ / p.px.Kh[i] + *m6;
      sy_ = sy_ / p.py.Kh[j] + *m5;
    }
    rate[1] = (sx_ + sy_) / rh;
  }
  // force at the stage time (ADE8:786-802)
  if (i == p.isrc && j == p.jsrc) {
    const int it = *p.step;
    const T rh = T(0.25) * (p.rho[c] + p.rho[c + 1] + p.rho[c + 1 + ld] + p.rho[c + ld]);
    rate[0] = rate[0] + a.f4x[(long)(it - 1) * 4 + inc] / p.rho[c];
    rate[1] = rate[1] + a.f4y[(long)(it - 1) * 4 + inc] / rh;
  }
  if (edge) rate[0] = rate[1] = T(0);  // ADE8:803-810
  // S's velocities with the border points zero (ADE8:803-810 ran before)
  auto V = [&](const T *f, int di, int dj) -> T {
    const int ii = i + di, jj = j + dj;
    if (ii == 0 || ii == nx - 1 || jj == 0 || jj == ny - 1) return T(0);
    return f[c + di + (long)dj * ld];
  };
  // sigma_xx, sigma_yy rates: Fortran i = 1..NX-1, j = 2..NY (ADE8:811-834)
  if (i <= nx - 2 && j >= 1) {
    const T lh = T(0.5) * (p.lam[c + 1] + p.lam[c]);
    const T mh = T(0.5) * (p.mu[c + 1] + p.mu[c]);
    const T l2m = lh + T(2) * mh;
    T dvx_dx = h8(V(vxs, 1, 0) - V(vxs, 0, 0), V(vxs, 2, 0) - V(vxs, -1, 0),
                  V(vxs, 3, 0) - V(vxs, -2, 0), V(vxs, 4, 0) - V(vxs, -3, 0), a.dx);
    T dvy_dy = h8(V(vys, 0, 0) - V(vys, 0, -1), V(vys, 0, 1) - V(vys, 0, -2),
                  V(vys, 0, 2) - V(vys, 0, -3), V(vys, 0, 3) - V(vys, 0, -4), a.dy);
    if (bd) {
      T *m0 = a.mem[0] + c, *m3 = a.mem[3] + c;
      *m0 = a.bxh[inc * nx + i] * *m0 + a.axh[inc * nx + i] * dvx_dx;
      *m3 = a.by[inc * ny + j] * *m3 + a.ay[inc * ny + j] * dvy_dy;
      dvx_dx = dvx_dx / p.px.Kh[i] + *m0;
      dvy_dy = dvy_dy / p.py.K[j] + *m3;
    }
    rate[2] = l2m * dvx_dx + lh * dvy_dy;
    rate[3] = lh * dvx_dx + l2m * dvy_dy;
  }
  // sigma_xy rate: Fortran i = 2..NX, j = 1..NY-1 (ADE8:836-861)
  if (i >= 1 && j <= ny - 2) {
    const T mhy = T(0.5) * (p.mu[c + ld] + p.mu[c]);
    T dvx_dy = h8(V(vxs, 0, 1) - V(vxs, 0, 0), V(vxs, 0, 2) - V(vxs, 0, -1),
                  V(vxs, 0, 3) - V(vxs, 0, -2), V(vxs, 0, 4) - V(vxs, 0, -3), a.dy);
    T dvy_dx = h8(V(vys, 0, 0) - V(vys, -1, 0), V(vys, 1, 0) - V(vys, -2, 0),
                  V(vys, 2, 0) - V(vys, -3, 0), V(vys, 3, 0) - V(vys, -4, 0), a.dx);
    if (bd) {
      T *m2 = a.mem[2] + c, *m1 = a.mem[1] + c;
      *m2 = a.bx[inc * nx + i] * *m2 + a.ax[inc * nx + i] * dvy_dx;
      *m1 = a.byh[inc * ny + j] * *m1 + a.ayh[inc * ny + j] * dvx_dy;
      dvy_dx = dvy_dx / p.px.K[i] + *m2;
      dvx_dy = dvx_dy / p.py.Kh[j] + *m1;
    }
    rate[4] = mhy * (dvy_dx + dvx_dy);
  }
  // accumulation and the next stage's state (ADE8:863-874)
  const T r = rk42<T>(inc);
  const bool next = inc < 3;
  const T rn = next ? rk41<T>(inc + 1) : T(0);
  for (int f = 0; f < 5; ++f) {
    const T d2 = rate[f];
    // X0 is read only where used: stage 1 (accumulation base and next
    // state) and stages 2-3 (next state)
    const T xo = inc == 0 ? X0[f][c] : a.xn[f][c];
    a.xn[f][c] = xo + r * d2 * p.dt;
    if (next) {
      const T x0 = (f < 2 && edge) ? T(0) : inc == 0 ? xo : X0[f][c];
      a.d[out][f][c] = x0 + rn * d2 * p.dt;
    }
  }
  if (edge) {
    a.xn[0][c] = T(0);
    a.xn[1][c] = T(0);
  }
}

}  // namespace

template <typename T>
hipError_t launch_ade8_step(const Ade8<T> &a, hipStream_t s) {
  const dim3 g(((a.s.nx + BX - 1) / BX) * ((a.s.ny + ATY - 1) / ATY)), b(BX, ATY);
  // stage 1 reads X0; the states of stages 2..4 alternate between d[1] and d[0]
  for (int inc = 0; inc < 4; ++inc)
    hipLaunchKernelGGL(k_ade_stage<T>, g, b, 0, s, a, inc, inc & 1, (inc & 1) ^ 1);
  return hipGetLastError();
}

template hipError_t launch_ade8_step<float>(const Ade8<float> &, hipStream_t);
template hipError_t launch_ade8_step<double>(const Ade8<double> &, hipStream_t);

}  // namespace cpml
