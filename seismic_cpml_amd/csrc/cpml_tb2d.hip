// Time-blocked 2D leapfrog: NT full time steps (stress + C-PML, velocity +
// C-PML, point force, rigid edges) per launch and per pass over HBM, for
// seismic_CPML_2D_isotropic_second_order.f90:550-686 and
// seismic_CPML_2D_isotropic_fourth_order.f90:551-687 (and the Holberg
// operator of seismic_ADEPML_2D_elastic_RK4_eighth_order.f90:263-266 in the
// same leapfrog).
//
// One 64-lane wavefront is one work item: lane l holds column
// i0 - G + l of a strip (G = NT (2r - 1) ring columns on each side, r = ORDER
// / 2) and the wave marches up the rows of one chunk.  At row iteration s the
// wave
//   * loads row s + r of v, row s of sigma and row s + 1 of the medium (two
//     iterations ahead of their use),
//   * for each time level m = 1..NT computes sigma^m at row p = s - 2(m-1)r
//     from the row window of v^{m-1} (2r + 1 rows) and v^m at row p - r from
//     the row window of sigma^m, x neighbours moving across lanes,
//   * stores level NT (sigma at row s - 2(NT-1)r, v at s - (2NT-1)r) for the
//     owned lanes and rows, and records the receivers of levels 1..NT-1.
// Every intermediate level stays in registers; the only HBM traffic is the
// state read once (v, sigma, lambda, mu, rho: 8 words per point) and written
// once (5 words) per NT steps.  Each level is valid on 2r - 1 fewer lanes on
// each side than the one before, so after NT levels lanes [G, 63 - G] hold
// correct values; the strip owns `ow` <= 64 - 2G of them.  Rows are handled
// the same way by starting (2NT - 1) r rows below the chunk and ending as far
// above it.  The state is read from one buffer and written to the other
// (ping-pong), so wavefronts never see each other's new values.
//
// The arithmetic of every point is the single-step kernels' (cpml_points2d.h),
// statement by statement, so results are bit-identical to them and to the
// reference association.
#include "cpml_device.h"
#include "cpml_energy.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

constexpr int WPB = 4;  // wavefronts (work items) per workgroup

__device__ __forceinline__ float lane_of(float x, int a4) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(a4, __builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double lane_of(double x, int a4) {
  typedef int i32x2 __attribute__((ext_vector_type(2)));
  i32x2 v = __builtin_bit_cast(i32x2, x);
  v.x = __builtin_amdgcn_ds_bpermute(a4, v.x);
  v.y = __builtin_amdgcn_ds_bpermute(a4, v.y);
  return __builtin_bit_cast(double, v);
}

// Staggered difference on values: g(k) = f[p + k*s] for k = -(r-1) .. r,
// in the association of Op<T, ORDER>::d (cpml_points2d.h).
template <typename T, int ORDER, typename G>
__device__ __forceinline__ T dop(G g) {
  if constexpr (ORDER == 2) {
    return g(1) - g(0);
  } else if constexpr (ORDER == 4) {
    return T(27) * g(1) - T(27) * g(0) - g(2) + g(-1);
  } else {
    return T(kHolberg1) * (g(1) - g(0)) + T(kHolberg2) * (g(2) - g(-1)) +
           T(kHolberg3) * (g(3) - g(-2)) + T(kHolberg4) * (g(4) - g(-3));
  }
}

// psi <- b*psi + a*d ; d <- d/K + psi, with the old psi given and the new
// one returned through `out` (cpml2's arithmetic); no update where `on` is
// false (the memory variable keeps its value).
template <typename T>
__device__ __forceinline__ T pml(bool on, T d, T old, T &out, T a, T b, T K) {
  if (!on) {
    out = old;
    return d;
  }
  out = b * old + a * d;
  return over_k(d, K) + out;
}

template <typename T, int N>
__device__ __forceinline__ void push(T (&w)[N], T x) {
#pragma unroll
  for (int k = 0; k < N - 1; ++k) w[k] = w[k + 1];
  w[N - 1] = x;
}

// One work item: strip ts (columns from ts * ow) and rows [jb, je) of the
// grid described by P, advanced NT steps.  EN: also the energy terms of every
// level at the item's owned points inside the energy window, summed over the
// wavefront and written to item slot `slot` of P.epart.
template <typename T, int ORDER, int NT, int CLS, bool EN = false>
__device__ __forceinline__ void tb2d_item(const TB2D<T> &P, int ts, int jb, int je,
                                          int slot = 0) {
  constexpr bool SX = (CLS & 1) != 0, SY = (CLS & 2) != 0;  // slabs this class touches
  constexpr int r = ORDER / 2;
  constexpr int W = 2 * r + 1;                  // row window of one level
  constexpr int G = NT * (2 * r - 1);           // ring lanes on each side
  constexpr int WM = (2 * NT - 1) * r + 2;      // medium rows: s - (2NT-1)r .. s + 1
  constexpr int NL = NT > 1 ? NT - 1 : 1;       // memory-variable windows (levels 1..NT-1)
  const Step2D<T> &p = P.a;
  const int nx = p.nx, ny = p.ny;
  const long ld = p.ld;
  const int lane = threadIdx.x & 63;
  const int i0 = ts * P.ow;
  const int i = i0 - G + lane;
  const bool col_in = i >= 0 && i < nx;
  const bool own_lane = lane >= G && lane < G + P.ow && i < nx;
  int a4[9];  // byte addresses of lanes lane-4 .. lane+4 (wrapping; ring lanes only)
#pragma unroll
  for (int k = 0; k < 9; ++k) a4[k] = ((lane + k - 4) & 63) * 4;
  auto xs = [&](T x, int k) { return k == 0 ? x : lane_of(x, a4[k + 4]); };

  const bool inx = SX && p.sx.in(i);
  const int sxi = inx ? p.sx.s(i) : 0;
  T pxa = T(0), pxb = T(0), pxK = T(1), pxah = T(0), pxbh = T(0), pxKh = T(1);
  if (inx) {
    pxa = p.px.a[i], pxb = p.px.b[i], pxK = p.px.K[i];
    pxah = p.px.ah[i], pxbh = p.px.bh[i], pxKh = p.px.Kh[i];
  }

  // ---- row windows (index W - 1 is the newest row)
  T V[NT][2][W];        // v^m, m = 0..NT-1: rows q_m - 2r .. q_m (q_0 = s + r)
  T S[NT][3][W];        // sigma^m, m = 1..NT: rows p_m - 2r .. p_m
  T Mx[3][WM];          // lambda, mu, rho: rows s - (2NT-1)r .. s + 1
  // memory variables of levels 1..NT-1 (slab launch): x [0] dvx_dx [1] dvy_dx
  // [2] dsxx_dx [3] dsxy_dx, y [0] dvy_dy [1] dvx_dy [2] dsxy_dy [3] dsyy_dy
  T PX[NL][4][W], PY[NL][4][W];
#pragma unroll
  for (int m = 0; m < NT; ++m)
#pragma unroll
    for (int k = 0; k < W; ++k) {
      V[m][0][k] = V[m][1][k] = T(0);
      S[m][0][k] = S[m][1][k] = S[m][2][k] = T(0);
    }
#pragma unroll
  for (int k = 0; k < WM; ++k) Mx[0][k] = Mx[1][k] = Mx[2][k] = T(0);
#pragma unroll
  for (int m = 0; m < NL; ++m)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < W; ++k) PX[m][q][k] = PY[m][q][k] = T(0);

  auto row_in = [&](int j) { return j >= 0 && j < ny; };
  auto at = [&](int j) { return (long)i + (long)j * ld; };
  const int s0 = jb - (2 * NT - 1) * r, s1 = je - 1 + (2 * NT - 1) * r;

  // prefill: v rows s0 - r .. s0 + r - 1, medium rows s0 + 2 - WM .. s0
#pragma unroll
  for (int k = 0; k < W - 1; ++k) {
    const int j = s0 - r + k;
    if (col_in && row_in(j)) {
      push(V[0][0], p.vx[at(j)]);
      push(V[0][1], p.vy[at(j)]);
    } else {
      push(V[0][0], T(0));
      push(V[0][1], T(0));
    }
  }
#pragma unroll
  for (int k = 0; k < WM - 1; ++k) {
    const int j = s0 + 2 - WM + k;
    const bool ok = col_in && row_in(j);
    push(Mx[0], ok ? p.lam[at(j)] : T(0));
    push(Mx[1], ok ? p.mu[at(j)] : T(0));
    push(Mx[2], ok ? p.rho[at(j)] : T(0));
  }
  // loads for iteration s: v row s + r, sigma row s, medium row s + 1, and
  // the memory variables of level 1 (sigma phase row s, velocity phase row
  // s - r)
  struct Row {
    T vx, vy, sxx, syy, sxy, lam, mu, rho;
    T px[4], py[4];
  };
  auto load = [&](Row &L, int s) {
    const int jv = s + r, jm = s + 1, jq = s - r;
    const bool okv = col_in && row_in(jv), oks = col_in && row_in(s), okm = col_in && row_in(jm);
    L.vx = okv ? p.vx[at(jv)] : T(0);
    L.vy = okv ? p.vy[at(jv)] : T(0);
    L.sxx = oks ? p.sxx[at(s)] : T(0);
    L.syy = oks ? p.syy[at(s)] : T(0);
    L.sxy = oks ? p.sxy[at(s)] : T(0);
    L.lam = okm ? p.lam[at(jm)] : T(0);
    L.mu = okm ? p.mu[at(jm)] : T(0);
    L.rho = okm ? p.rho[at(jm)] : T(0);
    {
#pragma unroll
      for (int q = 0; q < 4; ++q) L.px[q] = L.py[q] = T(0);
      if (SX && inx && row_in(s)) {
        const long o = sxi + (long)p.sx.w * s;
        L.px[0] = p.mx[0][o];
        L.px[1] = p.mx[1][o];
      }
      if (SX && inx && row_in(jq)) {
        const long o = sxi + (long)p.sx.w * jq;
        L.px[2] = p.mx[2][o];
        L.px[3] = p.mx[3][o];
      }
      if (SY && col_in && p.sy.in(s)) {
        const long o = i + p.mld * p.sy.s(s);
        L.py[0] = p.my[0][o];
        L.py[1] = p.my[1][o];
      }
      if (SY && col_in && p.sy.in(jq)) {
        const long o = i + p.mld * p.sy.s(jq);
        L.py[2] = p.my[2][o];
        L.py[3] = p.my[3][o];
      }
    }
  };
  // loads run two iterations ahead of their use (the march is a chain of
  // dependent rows: the second row in flight hides the memory latency)
  Row C, N, N2;
  load(C, s0);
  if (s0 + 1 <= s1) load(N, s0 + 1);

  // receivers of this strip: a cursor per recorded level (rows only increase)
  int rc0 = 0, rc1 = 0;
  if (NT > 1 && P.nrec > 0) {
    rc0 = P.rec_start[ts];
    rc1 = P.rec_start[ts + 1];
  }
  int cur[NL];
#pragma unroll
  for (int m = 0; m < NL; ++m) cur[m] = rc0;
  const int it0 = *p.step;
  double ek[NT], ep[NT];  // energy terms per level (EN)
#pragma unroll
  for (int m = 0; m < NT; ++m) ek[m] = ep[m] = 0.0;
  const bool ecol = EN && own_lane && i >= P.ew[0] && i <= P.ew[1];

  for (int s = s0; s <= s1; ++s) {
    if (s + 2 <= s1) load(N2, s + 2);
    push(V[0][0], C.vx);
    push(V[0][1], C.vy);
    push(Mx[0], C.lam);
    push(Mx[1], C.mu);
    push(Mx[2], C.rho);
    // (unrolled over levels: every index below is a compile-time constant)
#pragma unroll
    for (int m = 1; m <= NT; ++m) {
      const int pr = s - 2 * (m - 1) * r;  // sigma row
      const int qr = pr - r;               // velocity row
      const int ip = WM - 2 - 2 * (m - 1) * r;  // medium window index of row pr
      const int mb = m >= 2 ? m - 2 : 0;        // previous level's windows
      // ---------------- sigma^m at row pr: 2D2:556-600 (2D4:557-601)
      T sxx, syy, sxy;
      if (m == 1) {
        sxx = C.sxx, syy = C.syy, sxy = C.sxy;
      } else {
        sxx = S[mb][0][0], syy = S[mb][1][0], sxy = S[mb][2][0];
      }
      T npx0 = T(0), npx1 = T(0), npy0 = T(0), npy1 = T(0);
      {
        const T(&wx)[W] = V[m - 1][0];  // vx rows pr - r .. pr + r
        const T(&wy)[W] = V[m - 1][1];
        const int j = pr;
        const bool inside = col_in && row_in(j);
        const bool iny = SY && p.sy.in(j);
        T ox0 = T(0), ox1 = T(0), oy0 = T(0), oy1 = T(0);
        if (SX) {
          ox0 = m == 1 ? C.px[0] : PX[mb][0][0];
          ox1 = m == 1 ? C.px[1] : PX[mb][1][0];
        }
        if (SY) {
          oy0 = m == 1 ? C.py[0] : PY[mb][0][0];
          oy1 = m == 1 ? C.py[1] : PY[mb][1][0];
        }
        npx0 = ox0, npx1 = ox1, npy0 = oy0, npy1 = oy1;
        T pya = T(0), pyb = T(0), pyK = T(1), pyah = T(0), pybh = T(0), pyKh = T(1);
        if (iny) {
          pya = p.py.a[j], pyb = p.py.b[j], pyK = p.py.K[j];
          pyah = p.py.ah[j], pybh = p.py.bh[j], pyKh = p.py.Kh[j];
        }
        const T lam = Mx[0][ip], mu = Mx[1][ip];
        const T lam_e = xs(lam, 1), mu_e = xs(mu, 1), mu_n = Mx[1][ip + 1];
        // x differences of this row's v (every lane computes them: the
        // permutes need all lanes)
        const T vxc = wx[r], vyc = wy[r];
        const T ddx_vx = dop<T, ORDER>([&](int k) { return xs(vxc, k); });
        const T ddx_vy = dop<T, ORDER>([&](int k) { return xs(vyc, k - 1); });
        if (inside && i <= nx - 2 && j >= 1) {
          const T lhx = T(0.5) * (lam_e + lam);
          const T mhx = T(0.5) * (mu_e + mu);
          const T lp2m = lhx + T(2) * mhx;
          T dvx_dx = ddx_vx / p.den_x;
          // Op(vy, c - ld, ld): rows j-1 .. : g(k) = vy(j - 1 + k) = wy[r - 1 + k]
          T dvy_dy = dop<T, ORDER>([&](int k) { return wy[r - 1 + k]; }) / p.den_y;
          if (SX) dvx_dx = pml(inx, dvx_dx, ox0, npx0, pxah, pxbh, pxKh);
          if (SY) dvy_dy = pml(iny, dvy_dy, oy0, npy0, pya, pyb, pyK);
          sxx = sxx + (lp2m * dvx_dx + lhx * dvy_dy) * p.dt;
          syy = syy + (lhx * dvx_dx + lp2m * dvy_dy) * p.dt;
        }
        if (inside && i >= 1 && j <= ny - 2) {
          const T mhy = T(0.5) * (mu_n + mu);
          T dvy_dx = ddx_vy / p.den_x;
          T dvx_dy = dop<T, ORDER>([&](int k) { return wx[r + k]; }) / p.den_y;
          if (SX) dvy_dx = pml(inx, dvy_dx, ox1, npx1, pxa, pxb, pxK);
          // 2D4:596 divides by K_y(j) where 2D2:595 uses K_y_half(j)
          if (SY) dvx_dy = pml(iny, dvx_dy, oy1, npy1, pyah, pybh, ORDER == 4 ? pyK : pyKh);
          sxy = sxy + mhy * (dvy_dx + dvx_dy) * p.dt;
        }
      }
      push(S[m - 1][0], sxx);
      push(S[m - 1][1], syy);
      push(S[m - 1][2], sxy);
      if (SX && m < NT) {
        push(PX[m - 1][0], npx0);
        push(PX[m - 1][1], npx1);
      }
      if (SY && m < NT) {
        push(PY[m - 1][0], npy0);
        push(PY[m - 1][1], npy1);
      }
      if (m == NT && own_lane && pr >= jb && pr < je) {
        const long c = at(pr);
        P.sxx[c] = sxx;
        P.syy[c] = syy;
        P.sxy[c] = sxy;
        {
          if (SX && inx) {
            const long o = sxi + (long)p.sx.w * pr;
            P.mx[0][o] = npx0;
            P.mx[1][o] = npx1;
          }
          if (SY && p.sy.in(pr)) {
            const long o = i + p.mld * p.sy.s(pr);
            P.my[0][o] = npy0;
            P.my[1][o] = npy1;
          }
        }
      }
      // ---------------- v^m at row qr: 2D2:606-680 (2D4:607-681)
      {
        const T(&wxx)[W] = S[m - 1][0];  // rows qr - r .. qr + r
        const T(&wyy)[W] = S[m - 1][1];
        const T(&wxy)[W] = S[m - 1][2];
        const int j = qr;
        const bool inside = col_in && row_in(j);
        const bool iny = SY && p.sy.in(j);
        T vxn = V[m - 1][0][0], vyn = V[m - 1][1][0];
        T ox2 = T(0), ox3 = T(0), oy2 = T(0), oy3 = T(0);
        if (SX) {
          ox2 = m == 1 ? C.px[2] : PX[mb][2][0];
          ox3 = m == 1 ? C.px[3] : PX[mb][3][0];
        }
        if (SY) {
          oy2 = m == 1 ? C.py[2] : PY[mb][2][0];
          oy3 = m == 1 ? C.py[3] : PY[mb][3][0];
        }
        T npx2 = ox2, npx3 = ox3, npy2 = oy2, npy3 = oy3;
        T pya = T(0), pyb = T(0), pyK = T(1), pyah = T(0), pybh = T(0), pyKh = T(1);
        if (iny) {
          pya = p.py.a[j], pyb = p.py.b[j], pyK = p.py.K[j];
          pyah = p.py.ah[j], pybh = p.py.bh[j], pyKh = p.py.Kh[j];
        }
        const T rho = Mx[2][ip - r], rho_e = xs(rho, 1), rho_n = Mx[2][ip - r + 1],
                rho_ne = xs(rho_n, 1);
        const T sxxc = wxx[r], sxyc = wxy[r];
        const T ddx_sxx = dop<T, ORDER>([&](int k) { return xs(sxxc, k - 1); });
        const T ddx_sxy = dop<T, ORDER>([&](int k) { return xs(sxyc, k); });
        if (inside && i >= 1 && j >= 1) {
          T dsxx_dx = ddx_sxx / p.den_x;
          T dsxy_dy = dop<T, ORDER>([&](int k) { return wxy[r - 1 + k]; }) / p.den_y;
          if (SX) dsxx_dx = pml(inx, dsxx_dx, ox2, npx2, pxa, pxb, pxK);
          if (SY) dsxy_dy = pml(iny, dsxy_dy, oy2, npy2, pya, pyb, pyK);
          vxn = vxn + (dsxx_dx + dsxy_dy) * p.dt / rho;
        }
        const T rhh = T(0.25) * (rho + rho_e + rho_ne + rho_n);
        if (inside && i <= nx - 2 && j <= ny - 2) {
          T dsxy_dx = ddx_sxy / p.den_x;
          T dsyy_dy = dop<T, ORDER>([&](int k) { return wyy[r + k]; }) / p.den_y;
          if (SX) dsxy_dx = pml(inx, dsxy_dx, ox3, npx3, pxah, pxbh, pxKh);
          if (SY) dsyy_dy = pml(iny, dsyy_dy, oy3, npy3, pyah, pybh, pyKh);
          vyn = vyn + (dsxy_dx + dsyy_dy) * p.dt / rhh;
        }
        // point force at step it0 + m - 1, after the update (2D2:660-667)
        if (i == p.isrc && j == p.jsrc) {
          const int it = it0 + m - 1;
          vxn = vxn + p.fx[it - 1] * p.dt / rho;
          vyn = vyn + p.fy[it - 1] * p.dt / rhh;
        }
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) {  // 2D2:670-680
          vxn = T(0);
          vyn = T(0);
        }
        if (!inside) {
          vxn = T(0);
          vyn = T(0);
        }
        // energy of step it0 + m - 1 at (i, j): v^m, sigma^m of this row
        // (the middle of the sigma window) and the row's medium
        if (EN && ecol && j >= jb && j < je && j >= P.ew[2] && j <= P.ew[3])
          point_energy2d(vxn, vyn, wxx[r], wyy[r], wxy[r], Mx[0][ip - r], Mx[1][ip - r], rho,
                         ek[m - 1], ep[m - 1]);
        if (m < NT) {
          push(V[m][0], vxn);
          push(V[m][1], vyn);
          if (SX) {
            push(PX[m - 1][2], npx2);
            push(PX[m - 1][3], npx3);
          }
          if (SY) {
            push(PY[m - 1][2], npy2);
            push(PY[m - 1][3], npy3);
          }
          // receivers of step it0 + m - 1 on this row (owned rows only)
          if (P.nrec > 0 && j >= jb && j < je) {
            int c = cur[m - 1];
            while (c < rc1 && P.rec_j[c] < j) ++c;
            for (; c < rc1 && P.rec_j[c] == j; ++c)
              if (P.rec_i[c] == i && own_lane) {
                const long o = (long)(it0 + m - 2) * P.nrec + P.rec_slot[c];
                P.sisx[o] = vxn;
                P.sisy[o] = vyn;
              }
            cur[m - 1] = c;
          }
        } else if (own_lane && j >= jb && j < je) {
          const long c = at(j);
          P.vx[c] = vxn;
          P.vy[c] = vyn;
          {
            if (SX && inx) {
              const long o = sxi + (long)p.sx.w * j;
              P.mx[2][o] = npx2;
              P.mx[3][o] = npx3;
            }
            if (SY && p.sy.in(j)) {
              const long o = i + p.mld * p.sy.s(j);
              P.my[2][o] = npy2;
              P.my[3][o] = npy3;
            }
          }
        }
      }
    }
    C = N;
    N = N2;
  }
  if (EN) {  // wavefront sums in a fixed order, lane 0 writes the item's slot
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      double a = ek[m], b = ep[m];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
      }
      if (lane == 0) {
        P.epart[((long)m * P.eslots + slot) * 2] = a;
        P.epart[((long)m * P.eslots + slot) * 2 + 1] = b;
      }
    }
  }
}

// Items of class CLS of one grid: strips clean along x ([sx0, sx1)) or not
// (bit 0), chunks clean along y ([cy0, cy1)) or not (bit 1), x fastest; bit 2:
// every item outside the clean box, with both slabs' code.
__device__ __forceinline__ int outside(int u, int c0, int c1) { return u < c0 ? u : u - c0 + c1; }
template <int CLS>
__device__ __forceinline__ int tb2d_class_items(int nstrips, int nchunks, int bx, int cy) {
  if (CLS & 4) return nstrips * nchunks - bx * cy;
  const int sx = (CLS & 1) ? nstrips - bx : bx, sy = (CLS & 2) ? nchunks - cy : cy;
  return sx * sy;
}

template <typename T, int ORDER, int NT, int CLS, bool EN>
__global__ __launch_bounds__(64 * WPB) void k_tb2d(TB2D<T> P) {
  const int wv = threadIdx.x >> 6;
  const int bx = P.sx1 - P.sx0, cy = P.cy1 - P.cy0;
  const int nitem = tb2d_class_items<CLS>(P.nstrips, P.nchunks, bx, cy);
  // XCD-aware order: each XCD takes one contiguous band of items (strip fastest)
  const int nblk = (nitem + WPB - 1) / WPB;
  const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
  const int q8 = nblk >> 3, rem = nblk & 7;
  const int blk = xcd * q8 + min(xcd, rem) + loc;
  const int item = __builtin_amdgcn_readfirstlane(blk * WPB + wv);
  if (item >= nitem) return;
  int ts, tc;
  if (CLS & 4) {
    // chunks below the box, the box's chunks outside its strips, chunks above
    const int n1 = P.cy0 * P.nstrips, wdt = P.nstrips - bx, n2 = cy * wdt;
    if (item < n1) {
      tc = item / P.nstrips;
      ts = item % P.nstrips;
    } else if (item < n1 + n2) {
      const int t = item - n1;
      tc = P.cy0 + t / wdt;
      ts = outside(t % wdt, P.sx0, P.sx1);
    } else {
      const int t = item - n1 - n2;
      tc = P.cy1 + t / P.nstrips;
      ts = t % P.nstrips;
    }
  } else {
    const int wx = (CLS & 1) ? P.nstrips - bx : bx;
    const int us = item % wx, uc = item / wx;
    ts = (CLS & 1) ? outside(us, P.sx0, P.sx1) : P.sx0 + us;
    tc = (CLS & 2) ? outside(uc, P.cy0, P.cy1) : P.cy0 + uc;
  }
  const int jb = tc * P.jc;
  // energy slot: the item's place in the grid (chunk-major), whatever its class
  tb2d_item<T, ORDER, NT, CLS & 3, EN>(P, ts, jb, min(jb + P.jc, P.a.ny), tc * P.nstrips + ts);
}

// Energy of the NT levels of one launch: level b = blockIdx.x, step
// *e.step + b; the items' sums added in slot order, then as a tree.
constexpr int EBT = 256;
template <typename T>
__global__ __launch_bounds__(EBT) void k_tb2d_energy(Energy<T> e, const double *epart,
                                                     int eslots) {
  const int it = *e.step + (int)blockIdx.x;
  if (it % e.cadence != 0) return;
  const double *pp = epart + (long)blockIdx.x * eslots * 2;
  double kin = 0.0, pot = 0.0;
  for (int k = threadIdx.x; k < eslots; k += EBT) {
    kin += pp[2 * k];
    pot += pp[2 * k + 1];
  }
  __shared__ double sk[EBT], sp[EBT];
  sk[threadIdx.x] = kin;
  sp[threadIdx.x] = pot;
  __syncthreads();
  for (int h = EBT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sk[threadIdx.x] += sk[threadIdx.x + h];
      sp[threadIdx.x] += sp[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    e.out_kin[it - 1] = 0.5 * sk[0];  // 2D2:711: kinetic summed, then halved
    e.out_pot[it - 1] = sp[0];
  }
}

// Many grids in one launch: items[k] = (grid, strip, first row, end row) of
// one class, split into 8 contiguous bands (one per XCD); the wavefronts of
// the blocks an XCD receives stride through its band (one item each when the
// grid has one wavefront per item, several when it is a persistent grid).
template <typename T, int ORDER, int NT, int CLS>
__global__ __launch_bounds__(64 * WPB) void k_tb2d_multi(const TB2D<T> *__restrict__ grids,
                                                         const int4 *__restrict__ items,
                                                         int nitems) {
  const int G = gridDim.x, h = blockIdx.x;
  const int xcd = h & 7;
  const int nb = (G - xcd + 7) >> 3;  // blocks this XCD receives
  const int q = nitems >> 3, rem = nitems & 7;
  const int start = xcd * q + min(xcd, rem), count = q + (xcd < rem ? 1 : 0);
  const int wv = threadIdx.x >> 6;
  for (int u = __builtin_amdgcn_readfirstlane((h >> 3) * WPB + wv); u < count; u += nb * WPB) {
    const int4 e = items[start + u];
    tb2d_item<T, ORDER, NT, CLS>(grids[e.x], e.y, e.z, e.w);
  }
}

// Strips (ow columns) or chunks (jc rows) whose points and ring (g before
// and after) avoid both slabs of an axis: [t0, t1), empty when t1 <= t0.
void clean_range(int n, int B, int g, int wl, int wr, int nt, int &t0, int &t1) {
  t0 = t1 = 0;
  bool seen = false;
  for (int t = 0; t < nt; ++t) {
    const int lo = t * B - g, hi = t * B + B - 1 + g;
    const bool clean = lo >= wl && hi < n - wr;
    if (clean && !seen) t0 = t, seen = true;
    if (clean) t1 = t + 1;
    else if (seen) break;
  }
}

template <typename T, int ORDER, int NT>
hipError_t tb2d_launch(const TB2D<T> &P0, hipStream_t s, hipStream_t side) {
  constexpr int r = ORDER / 2;
  constexpr int G = NT * (2 * r - 1);
  TB2D<T> P = P0;
  const Step2D<T> &p = P.a;
  if (P.ow < 1 || P.ow > 64 - 2 * G) return hipErrorInvalidValue;
  P.nstrips = (p.nx + P.ow - 1) / P.ow;
  P.nchunks = (p.ny + P.jc - 1) / P.jc;
  clean_range(p.nx, P.ow, G, p.sx.wl, p.sx.wr, P.nstrips, P.sx0, P.sx1);
  clean_range(p.ny, P.jc, (2 * NT - 1) * r + r, p.sy.wl, p.sy.wr, P.nchunks, P.cy0, P.cy1);
  if (P.sx1 <= P.sx0 || P.cy1 <= P.cy0) P.sx0 = P.sx1 = P.cy0 = P.cy1 = 0;
  // One grid has few slab items, each a serial march: one launch of all of
  // them (both slabs' code) beside the slab-free items keeps the slab side
  // to one march's latency.  (The batch splits its many slab items by axis.)
  const int nclean = (P.sx1 - P.sx0) * (P.cy1 - P.cy0);
  const int nslab = P.nstrips * P.nchunks - nclean;
  const dim3 blk(64 * WPB);
  hipStream_t t = side ? side : s;
  if (P.epart) {  // the energy-summing instances
    if (nslab)
      hipLaunchKernelGGL((k_tb2d<T, ORDER, NT, 7, true>), dim3((nslab + WPB - 1) / WPB), blk, 0,
                         t, P);
    if (nclean)
      hipLaunchKernelGGL((k_tb2d<T, ORDER, NT, 0, true>), dim3((nclean + WPB - 1) / WPB), blk,
                         0, s, P);
    return hipGetLastError();
  }
  if (nslab)
    hipLaunchKernelGGL((k_tb2d<T, ORDER, NT, 7, false>), dim3((nslab + WPB - 1) / WPB), blk, 0,
                       t, P);
  if (nclean)
    hipLaunchKernelGGL((k_tb2d<T, ORDER, NT, 0, false>), dim3((nclean + WPB - 1) / WPB), blk,
                       0, s, P);
  return hipGetLastError();
}

}  // namespace

int tb2d_max_nt(int order, int prec) {
  if (prec == 8) return 2;
  return order == 2 ? 4 : order == 4 ? 3 : 2;
}

template <typename T>
hipError_t launch_tb2d(int order, int nt, const TB2D<T> &p, hipStream_t s, hipStream_t side) {
  if (nt < 1 || nt > tb2d_max_nt(order, sizeof(T))) return hipErrorInvalidValue;
  switch (order * 10 + nt) {
    case 21: return tb2d_launch<T, 2, 1>(p, s, side);
    case 22: return tb2d_launch<T, 2, 2>(p, s, side);
    case 41: return tb2d_launch<T, 4, 1>(p, s, side);
    case 42: return tb2d_launch<T, 4, 2>(p, s, side);
    case 81: return tb2d_launch<T, 8, 1>(p, s, side);
    case 82: return tb2d_launch<T, 8, 2>(p, s, side);
    default: break;
  }
  if constexpr (sizeof(T) == 4) {
    switch (order * 10 + nt) {
      case 23: return tb2d_launch<T, 2, 3>(p, s, side);
      case 24: return tb2d_launch<T, 2, 4>(p, s, side);
      case 43: return tb2d_launch<T, 4, 3>(p, s, side);
      default: break;
    }
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_tb2d<float>(int, int, const TB2D<float> &, hipStream_t, hipStream_t);
template hipError_t launch_tb2d<double>(int, int, const TB2D<double> &, hipStream_t,
                                        hipStream_t);

template <typename T>
hipError_t launch_tb2d_energy(const Energy<T> &e, const double *epart, int eslots, int nt,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_tb2d_energy<T>, dim3(nt), dim3(EBT), 0, s, e, epart, eslots);
  return hipGetLastError();
}
template hipError_t launch_tb2d_energy<float>(const Energy<float> &, const double *, int, int,
                                              hipStream_t);
template hipError_t launch_tb2d_energy<double>(const Energy<double> &, const double *, int, int,
                                               hipStream_t);

namespace {
template <typename T, int ORDER, int NT>
void tb2d_multi(int cls, const TB2D<T> *grids, const int4 *items, int nitems, int blocks,
                hipStream_t s) {
  const dim3 g(blocks), b(64 * WPB);
  switch (cls) {
    case 0: hipLaunchKernelGGL((k_tb2d_multi<T, ORDER, NT, 0>), g, b, 0, s, grids, items, nitems); break;
    case 1: hipLaunchKernelGGL((k_tb2d_multi<T, ORDER, NT, 1>), g, b, 0, s, grids, items, nitems); break;
    case 2: hipLaunchKernelGGL((k_tb2d_multi<T, ORDER, NT, 2>), g, b, 0, s, grids, items, nitems); break;
    default: hipLaunchKernelGGL((k_tb2d_multi<T, ORDER, NT, 3>), g, b, 0, s, grids, items, nitems); break;
  }
}
}  // namespace

template <typename T>
hipError_t launch_tb2d_multi(int order, int nt, int cls, const TB2D<T> *grids,
                             const int4 *items, int nitems, int blocks, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (nt < 1 || nt > tb2d_max_nt(order, sizeof(T)) || blocks < 1) return hipErrorInvalidValue;
  switch (order * 10 + nt) {
    case 21: tb2d_multi<T, 2, 1>(cls, grids, items, nitems, blocks, s); break;
    case 22: tb2d_multi<T, 2, 2>(cls, grids, items, nitems, blocks, s); break;
    case 41: tb2d_multi<T, 4, 1>(cls, grids, items, nitems, blocks, s); break;
    case 42: tb2d_multi<T, 4, 2>(cls, grids, items, nitems, blocks, s); break;
    case 81: tb2d_multi<T, 8, 1>(cls, grids, items, nitems, blocks, s); break;
    case 82: tb2d_multi<T, 8, 2>(cls, grids, items, nitems, blocks, s); break;
    default:
      if constexpr (sizeof(T) == 4) {
        if (order * 10 + nt == 23) tb2d_multi<T, 2, 3>(cls, grids, items, nitems, blocks, s);
        else if (order * 10 + nt == 24) tb2d_multi<T, 2, 4>(cls, grids, items, nitems, blocks, s);
        else if (order * 10 + nt == 43) tb2d_multi<T, 4, 3>(cls, grids, items, nitems, blocks, s);
        else return hipErrorInvalidValue;
      } else {
        return hipErrorInvalidValue;
      }
  }
  return hipGetLastError();
}

template hipError_t launch_tb2d_multi<float>(int, int, int, const TB2D<float> *, const int4 *,
                                             int, int, hipStream_t);
template hipError_t launch_tb2d_multi<double>(int, int, int, const TB2D<double> *, const int4 *,
                                              int, int, hipStream_t);

// Strips / chunks of one grid whose ring avoids the slabs (the box of
// launch_tb2d's clean class), for building item lists on the host.
void tb2d_clean_box(int order, int nt, int nx, int ny, int ow, int jc, const Slab &sx,
                    const Slab &sy, int box[4]) {
  const int r = order / 2, g = nt * (2 * r - 1);
  const int ns = (nx + ow - 1) / ow, nc = (ny + jc - 1) / jc;
  clean_range(nx, ow, g, sx.wl, sx.wr, ns, box[0], box[1]);
  clean_range(ny, jc, (2 * nt - 1) * r + r, sy.wl, sy.wr, nc, box[2], box[3]);
  if (box[1] <= box[0] || box[3] <= box[2]) box[0] = box[1] = box[2] = box[3] = 0;
}

}  // namespace cpml
