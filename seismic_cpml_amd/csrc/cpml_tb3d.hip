// Time-blocked 3D leapfrog: NT full time steps (stress + C-PML, velocity +
// C-PML, point force, rigid faces) per launch and per pass over HBM, for
// seismic_CPML_3D_isotropic_MPI_OpenMP.f90:836-1121 (the two OpenMP phases of
// NT consecutive `do it` iterations).
//
// A workgroup is NW wavefronts; wave w holds grid row j0 - NT + w and lane l
// column i0 - NT + l, and the group marches through the z planes of one
// chunk.  At plane iteration s it
//   * issues the loads of the next iteration (v at s + 2, sigma and the
//     coefficients at s + 1, the level-1 memory variables),
//   * for each time level m = 1..NT computes sigma^m at plane p = s - 2(m-1)
//     from the 3-plane column window of v^{m-1}, then v^m at plane p - 1
//     from the 3-plane window of sigma^m; x neighbours move across lanes,
//     y neighbours through LDS (written the iteration before: one barrier
//     per plane, two buffers),
//   * stores level NT (sigma at s - 2(NT-1), v at s - 2NT + 1) for the owned
//     points and planes, and records the receivers of levels 1..NT-1.
// Every intermediate level stays in registers and LDS; HBM sees the state
// read once (v 3, sigma 6, DELTAT*lambda, DELTAT*mu, DELTAT/rho: 12 words)
// and written once (9 words) per NT steps, plus the slab memory variables
// (read and written once per NT steps).  Each level is valid one lane and
// one row fewer on each side (the staggered differences reach one point),
// so a group owns 64 - 2NT columns (or fewer) x NW - 2NT rows; z is handled
// the same way by starting 2NT - 1 planes below the chunk and ending as far
// above it.  State is read from one buffer and written to the other.
//
// Per point the arithmetic is exactly the two-phase kernels' (same
// expressions, no FMA contraction), so results are bit-identical to them.
//
// Tiles are split by the absorbing slabs their points and ring touch: the
// x/y slabs (XY) and the z slabs (Z); each class is one launch of the kernel
// built with just the C-PML code it needs.
#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

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

template <typename T, int N>
__device__ __forceinline__ void push(T (&w)[N], T x) {
#pragma unroll
  for (int k = 0; k < N - 1; ++k) w[k] = w[k + 1];
  w[N - 1] = x;
}

// psi <- b*psi + a*d ; d <- d/K + psi (old psi given, new one in `out`); no
// update where `on` is false.  `slab`: the point lies in this axis' slab.
template <typename T>
__device__ __forceinline__ T pml(bool slab, bool on, T d, T old, T &out, T a, T b, T K) {
  out = old;
  if (!slab || !on) return d;
  out = b * old + a * d;
  return over_k(d, K) + out;
}

struct Tile {
  int i0, j0, kb, ke, tx, ty;
};

// Index of the u-th tile of one axis' class: the clean tiles [c0, c1) when
// S is false, else the tiles outside them (below c0, then from c1).
template <bool S>
__device__ __forceinline__ int axis_tile(int u, int c0, int c1) {
  if (!S) return c0 + u;
  return u < c0 ? u : c1 + (u - c0);
}
template <bool S>
__device__ __forceinline__ int axis_count(int n, int c0, int c1) {
  return S ? n - (c1 - c0) : c1 - c0;
}

// Tile of this workgroup.  Along each axis the tiles (z: chunks) whose
// points and ring avoid that axis' slabs form a range [cl[2a], cl[2a+1]);
// the launch of class (SX, SY, SZ) covers the product of the clean (flag
// false) or slab (flag true) tiles of each axis, x fastest, then y, then z,
// in XCD-banded order (each XCD owns one contiguous band of the launch).
template <bool SX, bool SY, bool SZ, typename P>
__device__ __forceinline__ Tile tb_tile(const P &Q, int nx, int ny, int nz, int kchunk) {
  const int ntx = (nx + Q.ow - 1) / Q.ow, nty = (ny + Q.ty - 1) / Q.ty;
  const int ntz = Q.zn[0] + Q.zn[1] + Q.zn[2];
  const int cx = axis_count<SX>(ntx, Q.box[0], Q.box[1]);
  const int cy = axis_count<SY>(nty, Q.box[2], Q.box[3]);
  const int cz = axis_count<SZ>(ntz, Q.zn[0], Q.zn[0] + Q.zn[1]);
  const int nt = cx * cy * cz;
  const int h = blockIdx.x;
  const int xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, rem = nt & 7;
  const int t = xcd * q + min(xcd, rem) + loc;
  Tile f;
  f.tx = axis_tile<SX>(t % cx, Q.box[0], Q.box[1]);
  f.ty = axis_tile<SY>((t / cx) % cy, Q.box[2], Q.box[3]);
  const int tz = axis_tile<SZ>(t / (cx * cy), Q.zn[0], Q.zn[0] + Q.zn[1]);
  f.i0 = f.tx * Q.ow;
  f.j0 = f.ty * Q.ty;
  if (tz < Q.zn[0]) {
    f.kb = tz * kchunk;
    f.ke = min(f.kb + kchunk, Q.z0);
  } else if (tz < Q.zn[0] + Q.zn[1]) {
    f.kb = Q.z0 + (tz - Q.zn[0]) * kchunk;
    f.ke = min(f.kb + kchunk, Q.z1);
  } else {
    f.kb = Q.z1 + (tz - Q.zn[0] - Q.zn[1]) * kchunk;
    f.ke = min(f.kb + kchunk, nz);
  }
  return f;
}

template <typename T, int NT, int NW, bool SX, bool SY, bool SZ>
__global__ __launch_bounds__(64 * NW) void k_tb3d(TB3D<T> P) {
  constexpr bool SLAB = SX || SY || SZ;
  constexpr int WM = 2 * NT;                   // coefficient planes s - 2NT + 1 .. s
  constexpr int NL = NT > 1 ? NT - 1 : 1;      // memory-variable windows (levels 1..NT-1)
  constexpr int NS = 6 * NT;                   // LDS slots per buffer
  __shared__ T xl[2][NS][NW][64];
  const Step3D<T> &p = P.a;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const Tile ft = tb_tile<SX, SY, SZ>(P, nx, ny, nz, p.kchunk);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = ft.i0 - NT + lane, j = ft.j0 - NT + w;
  const bool col_in = i >= 0 && i < nx, row_in = j >= 0 && j < ny;
  const bool in_xy = col_in && row_in;
  const bool own = lane >= NT && lane < NT + P.ow && i < nx && w >= NT && w < NT + P.ty && j < ny;
  const int wm = w > 0 ? w - 1 : 0, wp = w < NW - 1 ? w + 1 : NW - 1;
  const int up4 = ((lane + 63) & 63) * 4, dn4 = ((lane + 1) & 63) * 4;
  const int kb = ft.kb, ke = ft.ke;

  // buffers rebased at interior plane kl (the lowest plane loaded, >= 0)
  const unsigned es = sizeof(T);
  const int kl = max(kb - 2 * NT, 0), kh = min(ke + 2 * NT, nz);
  const unsigned psb = (unsigned)(p.ps * es);
  const unsigned bytes = (unsigned)((long)(kh - kl) * p.ps * es);
  const long origin = (long)(kl + p.R) * p.ps - p.base;
  const unsigned vo = (unsigned)(((long)(j + p.R) * p.ld + p.xpad + i) * es);
  auto rs = [&](const T *f) { return make_rsrc(f + origin, bytes); };
  using KArgs = const __attribute__((address_space(4))) TB3D<T> *;
  auto args = [&]() -> KArgs { return kernarg<TB3D<T>>(); };
  // plane k of a field: 0 outside the grid, and outside the planes [kl, kh)
  // the buffer covers (which only the pipeline's first planes ask for)
  auto LDP = [&](const T *f, int k) -> T {
    if (!in_xy || k < kl || k >= kh) return T(0);
    return BufOp<T>::ld(rs(f), vo, (unsigned)(k - kl) * psb);
  };
  auto STP = [&](T x, T *f, int k) { BufOp<T>::st(x, rs(f), vo, (unsigned)(k - kl) * psb); };

  // x/y slab membership and profiles (loop invariant)
  const bool inx = SX && p.sx.in(i), iny = SY && p.sy.in(j);
  T X[6] = {T(0), T(0), T(1), T(0), T(0), T(1)}, Y[6] = {T(0), T(0), T(1), T(0), T(0), T(1)};
  if (inx) {
    X[0] = p.px.a[i], X[1] = p.px.b[i], X[2] = p.px.K[i];
    X[3] = p.px.ah[i], X[4] = p.px.bh[i], X[5] = p.px.Kh[i];
  }
  if (iny) {
    Y[0] = p.py.a[j], Y[1] = p.py.b[j], Y[2] = p.py.K[j];
    Y[3] = p.py.ah[j], Y[4] = p.py.bh[j], Y[5] = p.py.Kh[j];
  }
  const unsigned sxw = (unsigned)p.sx.w, syw = (unsigned)p.sy.w;
  const unsigned sxi = inx ? (unsigned)p.sx.s(i) : 0u, syj = iny ? (unsigned)p.sy.s(j) : 0u;
  auto OX = [&](int k) { return sxi + sxw * ((unsigned)j + (unsigned)ny * (unsigned)k); };
  auto OY = [&](int k) { return (unsigned)i + (unsigned)p.mld * (syj + syw * (unsigned)k); };
  auto OZ = [&](int k) {
    return (unsigned)i + (unsigned)p.mld * ((unsigned)j + (unsigned)ny * (unsigned)p.sz.s(k));
  };

  // ---- column windows (index 2 = newest plane)
  T V[NT][3][3];    // v^m (m = 0..NT-1): vx vy vz, planes q_m - 2 .. q_m (q_0 = s + 1)
  T S[NT][6][3];    // sigma^m (m = 1..NT): xx yy zz xy xz yz, planes p_m - 2 .. p_m
  T Cw[3][WM];      // DELTAT*lambda, DELTAT*mu, DELTAT/rho: planes s - 2NT + 1 .. s
  // memory variables of levels 1..NT-1 per slab axis: [0..2] the stress
  // phase's (x: dvx_dx dvy_dx dvz_dx, ...), [3..5] the velocity phase's
  T MX[NL][6][3], MY[NL][6][3], MZ[NL][6][3];
#pragma unroll
  for (int m = 0; m < NT; ++m)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int f = 0; f < 3; ++f) V[m][f][k] = T(0);
#pragma unroll
      for (int f = 0; f < 6; ++f) S[m][f][k] = T(0);
    }
#pragma unroll
  for (int k = 0; k < WM; ++k) Cw[0][k] = Cw[1][k] = Cw[2][k] = T(0);
#pragma unroll
  for (int m = 0; m < NL; ++m)
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int k = 0; k < 3; ++k) MX[m][f][k] = MY[m][f][k] = MZ[m][f][k] = T(0);

  // loads of one iteration s: v (s + 1), sigma (s), coefficients (s), and the
  // level-1 memory variables (stress phase at s, velocity phase at s - 1)
  struct Ld {
    T v[3], so[6], c[3];
    T mx[6], my[6], mz[6];
  };
  auto load = [&](Ld &L, int s) {
    const KArgs Q = args();
    L.v[0] = LDP(Q->a.vx, s + 1);
    L.v[1] = LDP(Q->a.vy, s + 1);
    L.v[2] = LDP(Q->a.vz, s + 1);
    L.so[0] = LDP(Q->a.sxx, s);
    L.so[1] = LDP(Q->a.syy, s);
    L.so[2] = LDP(Q->a.szz, s);
    L.so[3] = LDP(Q->a.sxy, s);
    L.so[4] = LDP(Q->a.sxz, s);
    L.so[5] = LDP(Q->a.syz, s);
    L.c[0] = LDP(Q->a.clam, s);
    L.c[1] = LDP(Q->a.cmu, s);
    L.c[2] = LDP(Q->a.cbuoy, s);
    if (SLAB) {
#pragma unroll
      for (int f = 0; f < 6; ++f) L.mx[f] = L.my[f] = L.mz[f] = T(0);
      const bool ks = s >= 0 && s < nz, kv = s - 1 >= 0 && s - 1 < nz;
      if (SX && inx && row_in) {
        if (ks)
#pragma unroll
          for (int f = 0; f < 3; ++f) L.mx[f] = Q->a.mx[f][OX(s)];
        if (kv)
#pragma unroll
          for (int f = 3; f < 6; ++f) L.mx[f] = Q->a.mx[f][OX(s - 1)];
      }
      if (SY && iny && col_in) {
        if (ks)
#pragma unroll
          for (int f = 0; f < 3; ++f) L.my[f] = Q->a.my[f][OY(s)];
        if (kv)
#pragma unroll
          for (int f = 3; f < 6; ++f) L.my[f] = Q->a.my[f][OY(s - 1)];
      }
      if (SZ && in_xy) {
        if (ks && p.sz.in(s))
#pragma unroll
          for (int f = 0; f < 3; ++f) L.mz[f] = Q->a.mz[f][OZ(s)];
        if (kv && p.sz.in(s - 1))
#pragma unroll
          for (int f = 3; f < 6; ++f) L.mz[f] = Q->a.mz[f][OZ(s - 1)];
      }
    }
  };

  const int s0 = kb - (2 * NT - 1), s1 = ke - 1 + (2 * NT - 1);
  // prefill: v planes s0 - 1, s0; coefficient planes s0 - 2NT + 1 .. s0 - 1
  {
    const KArgs Q = args();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      push(V[0][0], LDP(Q->a.vx, s0 - 1 + k));
      push(V[0][1], LDP(Q->a.vy, s0 - 1 + k));
      push(V[0][2], LDP(Q->a.vz, s0 - 1 + k));
    }
#pragma unroll
    for (int k = 0; k < WM - 1; ++k) {
      const int kk = s0 - WM + 1 + k;
      push(Cw[0], LDP(Q->a.clam, kk));
      push(Cw[1], LDP(Q->a.cmu, kk));
      push(Cw[2], LDP(Q->a.cbuoy, kk));
    }
    // LDS cleared, then the y neighbours of v at plane s0 for the first
    // iteration (which reads buffer (s0 & 1) ^ 1)
    for (int q = threadIdx.x; q < 2 * NS * NW * 64; q += 64 * NW) (&xl[0][0][0][0])[q] = T(0);
    __syncthreads();
    const int b0 = (s0 & 1) ^ 1;  // plane s0 is the newest of the prefilled window
    xl[b0][0][w][lane] = V[0][0][2];
    xl[b0][1][w][lane] = V[0][1][2];
    xl[b0][2][w][lane] = V[0][2][2];
  }
  Ld C, N;
  load(C, s0);
  __syncthreads();

  // receivers of this tile: a cursor per recorded level (planes only increase)
  const int tix = ft.ty * ((nx + P.ow - 1) / P.ow) + ft.tx;
  int rc1 = 0;
  int cur[NL];
#pragma unroll
  for (int m = 0; m < NL; ++m) cur[m] = 0;
  if (NT > 1 && P.nrec > 0) {
    rc1 = P.rec_start[tix + 1];
#pragma unroll
    for (int m = 0; m < NL; ++m) cur[m] = P.rec_start[tix];
  }
  const int it0 = *p.step;

  for (int s = s0; s <= s1; ++s) {
    const KArgs Q = args();
    const int bw = s & 1, br = bw ^ 1;  // LDS buffer written / read this iteration
    if (s < s1) load(N, s + 1);
    push(V[0][0], C.v[0]);
    push(V[0][1], C.v[1]);
    push(V[0][2], C.v[2]);
    push(Cw[0], C.c[0]);
    push(Cw[1], C.c[1]);
    push(Cw[2], C.c[2]);
    // y neighbours of v at plane s + 1 for the next iteration
    xl[bw][0][w][lane] = C.v[0];
    xl[bw][1][w][lane] = C.v[1];
    xl[bw][2][w][lane] = C.v[2];

#pragma unroll
    for (int m = 1; m <= NT; ++m) {
      const int pk = s - 2 * (m - 1);  // stress plane
      const int qk = pk - 1;           // velocity plane
      const int mb = m >= 2 ? m - 2 : 0;
      const int cp = WM - 1 - 2 * (m - 1);  // coefficient window index of plane pk
      // LDS slots of this level's inputs: v^{m-1} at pk, sigma^m at qk
      const int sv = m == 1 ? 0 : 3 + 6 * (m - 2) + 3;
      const int ss = 3 + 6 * (m - 1);
      // ---------------- sigma^m at plane pk (3D:836-943)
      T sn[6];
      if (m == 1) {
#pragma unroll
        for (int f = 0; f < 6; ++f) sn[f] = C.so[f];
      } else {
#pragma unroll
        for (int f = 0; f < 6; ++f) sn[f] = S[mb][f][0];
      }
      T om[9] = {};  // old memory variables of the stress phase: x 0..2, y 3..5, z 6..8
      if (SLAB) {
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          if (SX) om[f] = m == 1 ? C.mx[f] : MX[mb][f][0];
          if (SY) om[3 + f] = m == 1 ? C.my[f] : MY[mb][f][0];
          if (SZ) om[6 + f] = m == 1 ? C.mz[f] : MZ[mb][f][0];
        }
      }
      T nm[9];
#pragma unroll
      for (int f = 0; f < 9; ++f) nm[f] = om[f];
      {
        const T vxc = V[m - 1][0][1], vyc = V[m - 1][1][1], vzc = V[m - 1][2][1];
        const T vx_xp = lane_of(vxc, dn4), vy_xm = lane_of(vyc, up4), vz_xm = lane_of(vzc, up4);
        const T vx_yp = xl[br][sv + 0][wp][lane];
        const T vy_ym = xl[br][sv + 1][wm][lane];
        const T vz_yp = xl[br][sv + 2][wp][lane];
        const bool inside = in_xy && pk >= 0 && pk < nz;
        const bool inz = SZ && inside && p.sz.in(pk);
        T Zp[6] = {T(0), T(0), T(1), T(0), T(0), T(1)};
        if (inz) {
          Zp[0] = p.pz.a[pk], Zp[1] = p.pz.b[pk], Zp[2] = p.pz.K[pk];
          Zp[3] = p.pz.ah[pk], Zp[4] = p.pz.bh[pk], Zp[5] = p.pz.Kh[pk];
        }
        if (inside) {
          const T cl = Cw[0][cp], cm = Cw[1][cp];
          const bool r_n = i <= nx - 2 && j >= 1 && pk >= 1;
          {
            T dvx_dx = (vx_xp - vxc) * p.odx;
            T dvy_dy = (vyc - vy_ym) * p.ody;
            T dvz_dz = (vzc - V[m - 1][2][0]) * p.odz;
            dvx_dx = pml(inx, r_n, dvx_dx, om[0], nm[0], X[3], X[4], X[5]);
            dvy_dy = pml(iny, r_n, dvy_dy, om[3], nm[3], Y[0], Y[1], Y[2]);
            dvz_dz = pml(inz, r_n, dvz_dz, om[6], nm[6], Zp[0], Zp[1], Zp[2]);
            if (r_n) {
              const T lp2m = cl + T(2) * cm;
              sn[0] = lp2m * dvx_dx + cl * (dvy_dy + dvz_dz) + sn[0];
              sn[1] = cl * (dvx_dx + dvz_dz) + lp2m * dvy_dy + sn[1];
              sn[2] = cl * (dvx_dx + dvy_dy) + lp2m * dvz_dz + sn[2];
            }
          }
          const bool r_xy = i >= 1 && j <= ny - 2;
          {
            T dvy_dx = (vyc - vy_xm) * p.odx;
            T dvx_dy = (vx_yp - vxc) * p.ody;
            dvy_dx = pml(inx, r_xy, dvy_dx, om[1], nm[1], X[0], X[1], X[2]);
            dvx_dy = pml(iny, r_xy, dvx_dy, om[4], nm[4], Y[3], Y[4], Y[5]);
            if (r_xy) sn[3] = cm * (dvy_dx + dvx_dy) + sn[3];
          }
          const bool r_xz = i >= 1 && pk <= nz - 2;
          {
            T dvz_dx = (vzc - vz_xm) * p.odx;
            T dvx_dz = (V[m - 1][0][2] - vxc) * p.odz;
            dvz_dx = pml(inx, r_xz, dvz_dx, om[2], nm[2], X[0], X[1], X[2]);
            dvx_dz = pml(inz, r_xz, dvx_dz, om[7], nm[7], Zp[3], Zp[4], Zp[5]);
            if (r_xz) sn[4] = cm * (dvz_dx + dvx_dz) + sn[4];
          }
          const bool r_yz = j <= ny - 2 && pk <= nz - 2;
          {
            T dvz_dy = (vz_yp - vzc) * p.ody;
            T dvy_dz = (V[m - 1][1][2] - vyc) * p.odz;
            dvz_dy = pml(iny, r_yz, dvz_dy, om[5], nm[5], Y[3], Y[4], Y[5]);
            dvy_dz = pml(inz, r_yz, dvy_dz, om[8], nm[8], Zp[3], Zp[4], Zp[5]);
            if (r_yz) sn[5] = cm * (dvz_dy + dvy_dz) + sn[5];
          }
        }
      }
#pragma unroll
      for (int f = 0; f < 6; ++f) push(S[m - 1][f], sn[f]);
      if (SLAB && m < NT) {
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          if (SX) push(MX[m - 1][f], nm[f]);
          if (SY) push(MY[m - 1][f], nm[3 + f]);
          if (SZ) push(MZ[m - 1][f], nm[6 + f]);
        }
      }
      xl[bw][ss + 0][w][lane] = sn[3];  // sigma_xy, sigma_yz, sigma_yy of plane pk
      xl[bw][ss + 1][w][lane] = sn[5];
      xl[bw][ss + 2][w][lane] = sn[1];
      if (m == NT && own && pk >= kb && pk < ke) {
        STP(sn[0], Q->sxx, pk);
        STP(sn[1], Q->syy, pk);
        STP(sn[2], Q->szz, pk);
        STP(sn[3], Q->sxy, pk);
        STP(sn[4], Q->sxz, pk);
        STP(sn[5], Q->syz, pk);
        if (SX && inx)
#pragma unroll
          for (int f = 0; f < 3; ++f) Q->mx[f][OX(pk)] = nm[f];
        if (SY && iny)
#pragma unroll
          for (int f = 0; f < 3; ++f) Q->my[f][OY(pk)] = nm[3 + f];
        if (SZ && p.sz.in(pk))
#pragma unroll
          for (int f = 0; f < 3; ++f) Q->mz[f][OZ(pk)] = nm[6 + f];
      }

      // ---------------- v^m at plane qk (3D:976-1121)
      T vn[3] = {V[m - 1][0][0], V[m - 1][1][0], V[m - 1][2][0]};
      T ov[9] = {};
      if (SLAB) {
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          if (SX) ov[f] = m == 1 ? C.mx[3 + f] : MX[mb][3 + f][0];
          if (SY) ov[3 + f] = m == 1 ? C.my[3 + f] : MY[mb][3 + f][0];
          if (SZ) ov[6 + f] = m == 1 ? C.mz[3 + f] : MZ[mb][3 + f][0];
        }
      }
      T nv[9];
#pragma unroll
      for (int f = 0; f < 9; ++f) nv[f] = ov[f];
      {
        const T(&sw)[6][3] = S[m - 1];  // planes qk - 1, qk, qk + 1
        const T sxx_xm = lane_of(sw[0][1], up4);
        const T sxy_xp = lane_of(sw[3][1], dn4);
        const T sxz_xp = lane_of(sw[4][1], dn4);
        const T sxy_ym = xl[br][ss + 0][wm][lane];
        const T syz_ym = xl[br][ss + 1][wm][lane];
        const T syy_yp = xl[br][ss + 2][wp][lane];
        const bool inside = in_xy && qk >= 0 && qk < nz;
        const bool inz = SZ && inside && p.sz.in(qk);
        T Zq[6] = {T(0), T(0), T(1), T(0), T(0), T(1)};
        if (inz) {
          Zq[0] = p.pz.a[qk], Zq[1] = p.pz.b[qk], Zq[2] = p.pz.K[qk];
          Zq[3] = p.pz.ah[qk], Zq[4] = p.pz.bh[qk], Zq[5] = p.pz.Kh[qk];
        }
        if (inside) {
          const T cb = Cw[2][cp - 1];
          const bool r_vx = i >= 1 && j >= 1 && qk >= 1;
          {
            T dsxx_dx = (sw[0][1] - sxx_xm) * p.odx;
            T dsxy_dy = (sw[3][1] - sxy_ym) * p.ody;
            T dsxz_dz = (sw[4][1] - sw[4][0]) * p.odz;
            dsxx_dx = pml(inx, r_vx, dsxx_dx, ov[0], nv[0], X[0], X[1], X[2]);
            dsxy_dy = pml(iny, r_vx, dsxy_dy, ov[3], nv[3], Y[0], Y[1], Y[2]);
            dsxz_dz = pml(inz, r_vx, dsxz_dz, ov[6], nv[6], Zq[0], Zq[1], Zq[2]);
            if (r_vx) vn[0] = cb * (dsxx_dx + dsxy_dy + dsxz_dz) + vn[0];
          }
          const bool r_vy = i <= nx - 2 && j <= ny - 2 && qk >= 1;
          {
            T dsxy_dx = (sxy_xp - sw[3][1]) * p.odx;
            T dsyy_dy = (syy_yp - sw[1][1]) * p.ody;
            T dsyz_dz = (sw[5][1] - sw[5][0]) * p.odz;
            dsxy_dx = pml(inx, r_vy, dsxy_dx, ov[1], nv[1], X[3], X[4], X[5]);
            dsyy_dy = pml(iny, r_vy, dsyy_dy, ov[4], nv[4], Y[3], Y[4], Y[5]);
            dsyz_dz = pml(inz, r_vy, dsyz_dz, ov[7], nv[7], Zq[0], Zq[1], Zq[2]);
            if (r_vy) vn[1] = cb * (dsxy_dx + dsyy_dy + dsyz_dz) + vn[1];
          }
          const bool r_vz = i <= nx - 2 && j >= 1 && qk <= nz - 2;
          {
            T dsxz_dx = (sxz_xp - sw[4][1]) * p.odx;
            T dsyz_dy = (sw[5][1] - syz_ym) * p.ody;
            T dszz_dz = (sw[2][2] - sw[2][1]) * p.odz;
            dsxz_dx = pml(inx, r_vz, dsxz_dx, ov[2], nv[2], X[3], X[4], X[5]);
            dsyz_dy = pml(iny, r_vz, dsyz_dy, ov[5], nv[5], Y[0], Y[1], Y[2]);
            dszz_dz = pml(inz, r_vz, dszz_dz, ov[8], nv[8], Zq[3], Zq[4], Zq[5]);
            if (r_vz) vn[2] = cb * (dsxz_dx + dsyz_dy + dszz_dz) + vn[2];
          }
          // point force at step it0 + m - 1 (3D:1074-1082), rigid faces (3D:1087-1121)
          if (i == p.isrc && j == p.jsrc && qk == p.ksrc) {
            const int it = it0 + m - 1;
            const T rsrc = Q->a.rho[i + (long)j * p.ld + (long)qk * p.ps];
            vn[0] = vn[0] + Q->a.fx[it - 1] * p.dt / rsrc;
            vn[1] = vn[1] + Q->a.fy[it - 1] * p.dt / rsrc;
          }
          if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || qk == 0 || qk == nz - 1)
            vn[0] = vn[1] = vn[2] = T(0);
        }
      }
      if (m < NT) {
#pragma unroll
        for (int f = 0; f < 3; ++f) push(V[m][f], vn[f]);
        if (SLAB) {
#pragma unroll
          for (int f = 0; f < 3; ++f) {
            if (SX) push(MX[m - 1][3 + f], nv[f]);
            if (SY) push(MY[m - 1][3 + f], nv[3 + f]);
            if (SZ) push(MZ[m - 1][3 + f], nv[6 + f]);
          }
        }
        xl[bw][ss + 3][w][lane] = vn[0];  // y neighbours of v^m at plane qk
        xl[bw][ss + 4][w][lane] = vn[1];
        xl[bw][ss + 5][w][lane] = vn[2];
        // receivers of step it0 + m - 1 on plane qk (owned points only)
        if (P.nrec > 0 && qk >= kb && qk < ke) {
          int c = cur[m - 1];
          while (c < rc1 && P.rec_k[c] < qk) ++c;
          for (; c < rc1 && P.rec_k[c] == qk; ++c)
            if (P.rec_i[c] == i && P.rec_j[c] == j && own) {
              const long o = (long)(it0 + m - 2) * P.nrec + P.rec_slot[c];
              P.sisx[o] = vn[0];
              P.sisy[o] = vn[1];
              if (P.sisz) P.sisz[o] = vn[2];
            }
          cur[m - 1] = c;
        }
      } else if (own && qk >= kb && qk < ke) {
        STP(vn[0], Q->vx, qk);
        STP(vn[1], Q->vy, qk);
        STP(vn[2], Q->vz, qk);
        if (SX && inx)
#pragma unroll
          for (int f = 0; f < 3; ++f) Q->mx[3 + f][OX(qk)] = nv[f];
        if (SY && iny)
#pragma unroll
          for (int f = 0; f < 3; ++f) Q->my[3 + f][OY(qk)] = nv[3 + f];
        if (SZ && p.sz.in(qk))
#pragma unroll
          for (int f = 0; f < 3; ++f) Q->mz[3 + f][OZ(qk)] = nv[6 + f];
      }
    }
    __syncthreads();
    C = N;
  }
}

// Tiles along one axis (size B, extent n, slab widths wl, wr) whose points and
// ring of g avoid both slabs: [t0, t1), empty when t1 <= t0.
void clean_tiles(int n, int B, int g, int wl, int wr, int &t0, int &t1) {
  const int nt = (n + B - 1) / B;
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

template <typename T, int NT, int NW, bool SX, bool SY, bool SZ>
void tb3d_class(const TB3D<T> &P, const int (&n)[3], hipStream_t s) {
  const int cx = SX ? n[0] - (P.box[1] - P.box[0]) : P.box[1] - P.box[0];
  const int cy = SY ? n[1] - (P.box[3] - P.box[2]) : P.box[3] - P.box[2];
  const int cz = SZ ? n[2] - P.zn[1] : P.zn[1];
  const long nt = (long)cx * cy * cz;
  if (nt > 0)
    hipLaunchKernelGGL((k_tb3d<T, NT, NW, SX, SY, SZ>), dim3((unsigned)nt), dim3(64 * NW), 0, s,
                       P);
}

template <typename T, int NT, int NW>
hipError_t tb3d_launch(const TB3D<T> &P0, hipStream_t s, hipStream_t side) {
  TB3D<T> P = P0;
  const Step3D<T> &p = P.a;
  P.ty = NW - 2 * NT;
  if (P.ow < 1 || P.ow > 64 - 2 * NT) return hipErrorInvalidValue;
  // z regions: planes whose march (2NT - 1 planes of ring) avoids the z slabs
  const int g = 2 * NT - 1, nz = p.nz, kc = p.kchunk;
  int z0 = p.sz.wl > 0 ? p.sz.wl + g : 0;
  int z1 = p.sz.wr > 0 ? nz - p.sz.wr - g : nz;
  if (z1 <= z0) z0 = z1 = 0;
  P.z0 = z0;
  P.z1 = z1;
  P.zn[0] = (z0 + kc - 1) / kc;
  P.zn[1] = (z1 - z0 + kc - 1) / kc;
  P.zn[2] = (nz - z1 + kc - 1) / kc;
  clean_tiles(p.nx, P.ow, NT, p.sx.wl, p.sx.wr, P.box[0], P.box[1]);
  clean_tiles(p.ny, P.ty, NT, p.sy.wl, p.sy.wr, P.box[2], P.box[3]);
  if (P.box[1] < P.box[0]) P.box[0] = P.box[1] = 0;
  if (P.box[3] < P.box[2]) P.box[2] = P.box[3] = 0;
  const int n[3] = {(p.nx + P.ow - 1) / P.ow, (p.ny + P.ty - 1) / P.ty,
                    P.zn[0] + P.zn[1] + P.zn[2]};
  // slab classes first on the side stream (the corners, with the most code,
  // start earliest), the slab-free class on the main stream
  hipStream_t t = side ? side : s;
  tb3d_class<T, NT, NW, true, true, true>(P, n, t);
  tb3d_class<T, NT, NW, true, true, false>(P, n, t);
  tb3d_class<T, NT, NW, true, false, true>(P, n, t);
  tb3d_class<T, NT, NW, false, true, true>(P, n, t);
  tb3d_class<T, NT, NW, true, false, false>(P, n, t);
  tb3d_class<T, NT, NW, false, true, false>(P, n, t);
  tb3d_class<T, NT, NW, false, false, true>(P, n, t);
  tb3d_class<T, NT, NW, false, false, false>(P, n, s);
  return hipGetLastError();
}

}  // namespace

int tb3d_max_nt(int) { return 2; }

// rows per workgroup: 16 wavefronts in float, 8 in double (LDS: 6 NT values
// per point and buffer)
template <typename T>
hipError_t launch_tb3d(int nt, const TB3D<T> &p, hipStream_t s, hipStream_t side) {
  constexpr int NW = sizeof(T) == 4 ? 16 : 8;
  if (nt == 1) return tb3d_launch<T, 1, NW>(p, s, side);
  if (nt == 2) return tb3d_launch<T, 2, NW>(p, s, side);
  return hipErrorInvalidValue;
}

template hipError_t launch_tb3d<float>(int, const TB3D<float> &, hipStream_t, hipStream_t);
template hipError_t launch_tb3d<double>(int, const TB3D<double> &, hipStream_t, hipStream_t);

}  // namespace cpml
