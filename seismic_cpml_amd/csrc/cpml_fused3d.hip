// Fused 3D time step: stress + C-PML and velocity + C-PML + source + rigid
// faces in ONE pass over HBM, for seismic_CPML_3D_isotropic_MPI_OpenMP.f90
// :836-1121 (the two OpenMP phases of one `do it` iteration).
//
// The step reads state A (v at n, sigma at n-1/2, memory variables) and writes
// state B; the host swaps A and B after every step, so neighbouring tiles can
// never see each other's new values.  Per point the arithmetic is exactly the
// two-phase kernels' (same expressions, no FMA contraction), so results are
// bit-identical to them and to the reference association.
//
// Layout of the work: a block is TY + 2 wavefronts; wave w holds grid row
// j = j0 - 1 + w and lane l holds column i = i0 - 1 + l, so the block covers
// 64 columns x (TY + 2) rows and owns the inner 62 x TY of them (lanes 1..62
// of waves 1..TY).  It marches through a chunk of z planes [kb, ke).  At
// plane s it
//   1. issues the global loads of plane s + 1 (register double buffering),
//   2. computes sigma(s) at all 64 x (TY + 2) points (the outer ring is what
//      the velocity update of the owned edge points needs from neighbouring
//      tiles: recomputed here, never exchanged between blocks) and stores the
//      owned points,
//   3. exchanges through LDS (one barrier per plane) the y neighbours of
//      sigma(s) and of v(s + 1); x neighbours move across lanes (ds_bpermute),
//   4. computes v(s - 1) at the owned points from sigma(s - 1) and its
//      neighbours, sigma(s - 2) (xz, yz) and sigma(s) (zz), and stores it.
// The chunk's planes kb - 1 and ke are computed for sigma only (z ring).
//
// The tiles are split by the absorbing slabs their points and ring touch:
// none (the kernel is built without any C-PML code: fewer registers, more
// waves in flight), only the z slabs, only the x and y slabs, any; each class
// runs as one launch of the kernel built with just the C-PML code it needs.
//
// Compulsory traffic per point: read v(3), sigma(6), DELTAT*lambda, DELTAT*mu,
// DELTAT/rho, write sigma(6), v(3) = 21 words, plus the slab memory variables.
#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

constexpr int FW = 64;       // lanes per row
constexpr int FO = FW - 2;   // owned columns per row (lanes 1..62)

struct FTile {
  int i0, j0, kb, ke;  // first owned column and row, z chunk
};

// Tile of this block.  The tiles whose x/y footprint (points and ring) avoids
// the x and y slabs form a box [box[0], box[1]) x [box[2], box[3]) in tile
// coordinates, and the z chunks [box[4], box[5]) avoid the z slabs.  Each
// launch covers one class of tiles: XY = false the tiles of the x/y box,
// XY = true the ring of tiles around it; Z = false the chunks of the z box,
// Z = true the chunks below and above it.  Tiles are handed out in XCD-aware
// order: each XCD owns one contiguous band of the launch's tiles (x fastest,
// then y, then z chunk), so tiles that share rows share an L2.
template <int TY, bool XY, bool Z>
__device__ __forceinline__ FTile ftile(const int (&box)[6], const int (&zc)[5], int nx, int ny,
                                       int nz, int kchunk) {
  const int ntx = (nx + FO - 1) / FO, nty = (ny + TY - 1) / TY;
  const int ntz = zc[2] + zc[3] + zc[4];
  const int bx = box[1] - box[0], by = box[3] - box[2];
  const int per = XY ? ntx * nty - bx * by : bx * by;  // tiles per z chunk
  const int nc = Z ? ntz - (box[5] - box[4]) : box[5] - box[4];
  const int nt = per * nc;
  const int h = blockIdx.x;
  const int xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, rem = nt & 7;
  const int t = xcd * q + min(xcd, rem) + loc;
  const int c = t / per;
  int r = t % per;
  const int tz = Z ? (c < box[4] ? c : box[5] + c - box[4]) : box[4] + c;
  int tx, ty;
  if (!XY) {
    tx = box[0] + r % bx;
    ty = box[2] + r / bx;
  } else {
    // rows below the box, the rows of the box outside its columns, rows above
    const int n1 = box[2] * ntx, wdt = ntx - bx, n2 = by * wdt;
    if (r < n1) {
      ty = r / ntx;
      tx = r % ntx;
    } else if (r < n1 + n2) {
      r -= n1;
      ty = box[2] + r / wdt;
      const int cc = r % wdt;
      tx = cc < box[0] ? cc : cc - box[0] + box[1];
    } else {
      r -= n1 + n2;
      ty = box[3] + r / ntx;
      tx = r % ntx;
    }
  }
  FTile f;
  f.i0 = tx * FO;
  f.j0 = ty * TY;
  // z chunk tz: region below z0, between z0 and z1, above z1
  if (tz < zc[2]) {
    f.kb = tz * kchunk;
    f.ke = min(f.kb + kchunk, zc[0]);
  } else if (tz < zc[2] + zc[3]) {
    f.kb = zc[0] + (tz - zc[2]) * kchunk;
    f.ke = min(f.kb + kchunk, zc[1]);
  } else {
    f.kb = zc[1] + (tz - zc[2] - zc[3]) * kchunk;
    f.ke = min(f.kb + kchunk, nz);
  }
  return f;
}

// Value of x held by lane src (byte address 4 * src) of this wavefront.
__device__ __forceinline__ float lane_val(float x, int src4) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src4, __builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double lane_val(double x, int src4) {
  typedef int i32x2 __attribute__((ext_vector_type(2)));
  i32x2 v = __builtin_bit_cast(i32x2, x);
  v.x = __builtin_amdgcn_ds_bpermute(src4, v.x);
  v.y = __builtin_amdgcn_ds_bpermute(src4, v.y);
  return __builtin_bit_cast(double, v);
}

// One memory-variable step, old value already loaded, new value stored to mn
// (when given): psi <- b*psi + a*d ; d <- d/K + psi (cpml2's arithmetic).
// `on`: the derivative is updated at this point; otherwise the old value is
// carried over.
template <typename T>
__device__ __forceinline__ T pml(bool on, bool slab, T d, T old, T *mn, T a, T b, T K) {
  if (!slab) return d;
  if (on) {
    const T m = b * old + a * d;
    if (mn) *mn = m;
    return over_k(d, K) + m;
  }
  if (mn) *mn = old;
  return d;
}

// Profile values of one grid index: a, b, K at the point and at the half point.
template <typename T>
struct PV {
  T a, b, K, ah, bh, Kh;
};
template <typename T, typename Prof>
__device__ __forceinline__ PV<T> pv(const Prof &q, int x) {
  return PV<T>{q.a[x], q.b[x], q.K[x], q.ah[x], q.bh[x], q.Kh[x]};
}

// Values of one point loaded one plane ahead (for sigma plane s).
template <typename T>
struct FP {
  T vxp, vyp, vzp;  // v at s + 1
  T so[6];          // sigma(s) old: xx yy zz xy xz yz
  T cl, cm;         // DELTAT*lambda, DELTAT*mu at s
  T cb;             // DELTAT/rho at s - 1
};

template <typename T, int TY, bool XY, bool Z>
__global__ __launch_bounds__(FW *(TY + 2)) void k_fused3d(Fused3D<T> P) {
  constexpr bool SLAB = XY || Z;
  constexpr int NW = TY + 2;
  __shared__ T xs[3][3][NW][FW];  // sigma_xy, sigma_yz, sigma_yy of three planes
  __shared__ T xv[2][3][NW][FW];  // vx, vy, vz of plane s + 1
  const Step3D<T> &p = P.a;
  const int box[6] = {P.box[0], P.box[1], P.box[2], P.box[3], P.box[4], P.box[5]};
  const int zc[5] = {P.z0, P.z1, P.zn[0], P.zn[1], P.zn[2]};
  const FTile ft = ftile<TY, XY, Z>(box, zc, p.nx, p.ny, p.nz, p.kchunk);
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int lane = threadIdx.x & (FW - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = ft.i0 - 1 + lane, j = ft.j0 - 1 + w;
  const bool own_lane = lane >= 1 && lane <= FO && i < nx;
  const bool own_row = w >= 1 && w <= TY && j < ny;  // uniform per wave
  const bool own = own_lane && own_row;
  const bool in_xy = i >= 0 && i < nx && j >= 0 && j < ny;
  // addressable in the allocation (halo of 2 on every side): i, j <= n + 1
  const bool ld_ok = i <= nx + 1 && j <= ny + 1;
  const int wm = w > 0 ? w - 1 : 0, wp = w < NW - 1 ? w + 1 : NW - 1;  // LDS rows
  const unsigned es = sizeof(T);
  const unsigned psb = (unsigned)(p.ps * es);
  // buffers rebased at interior plane kb-2; planes kb-2 .. ke+1
  const unsigned bytes = (unsigned)((long)(ft.ke - ft.kb + 4) * p.ps * es);
  const long origin = (long)(ft.kb - 2 + p.R) * p.ps - p.base;
  auto rs = [&](const T *f) { return make_rsrc(f + origin, bytes); };
  // The ~60 pointers of the step (fields, coefficients, 36 memory arrays, 18
  // profiles, next state) do not fit in scalar registers across the z loop:
  // they are re-read where they are used (kernarg()).
  using KArgs = const __attribute__((address_space(4))) Fused3D<T> *;
  auto args = [&]() -> KArgs { return kernarg<Fused3D<T>>(); };
  // byte offset of (i, j) in a buffer plane
  const unsigned vo = (unsigned)(((long)(j + p.R) * p.ld + p.xpad + i) * es);
  auto soff = [&](int k) -> unsigned { return (unsigned)(k - ft.kb + 2) * psb; };
  auto LD = [&](__amdgpu_buffer_rsrc_t r, int k) { return BufOp<T>::ld(r, vo, soff(k)); };
  const int up4 = (lane > 0 ? lane - 1 : 0) * 4, dn4 = (lane < FW - 1 ? lane + 1 : lane) * 4;
  // slab membership of this column / row and their profile values (loop
  // invariant); z is per plane
  const bool inx = XY && p.sx.in(i), iny = XY && p.sy.in(j);
  PV<T> X{}, Y{};
  if (inx) X = pv<T>(p.px, i);
  if (iny) Y = pv<T>(p.py, j);

  FP<T> C, N;
  auto load = [&](FP<T> &L, int s) {  // loads used at iteration s (sigma plane s)
    if (!ld_ok) return;
    const KArgs Q = args();
    L.vxp = LD(rs(Q->a.vx), s + 1);
    L.vyp = LD(rs(Q->a.vy), s + 1);
    L.vzp = LD(rs(Q->a.vz), s + 1);
    // every wave issues the same loads (the ring rows use fewer of them): the
    // waits for the previous plane count the loads issued after it
    L.so[0] = LD(rs(Q->a.sxx), s);
    L.so[1] = LD(rs(Q->a.syy), s);
    L.so[2] = LD(rs(Q->a.szz), s);
    L.so[3] = LD(rs(Q->a.sxy), s);
    L.so[4] = LD(rs(Q->a.sxz), s);
    L.so[5] = LD(rs(Q->a.syz), s);
    L.cl = LD(rs(Q->a.clam), s);
    L.cm = LD(rs(Q->a.cmu), s);
    L.cb = LD(rs(Q->a.cbuoy), s - 1);
  };

  // ---- pipeline fill: plane kb - 1 and the y neighbours of its v
  const T zero = T(0);
  T vx_c = zero, vy_c = zero, vz_c = zero, vz_m = zero;  // v at s; vz at s - 1
  T yv_x = zero, yv_y = zero, yv_z = zero;  // vx(j+1), vy(j-1), vz(j+1) at s
  if (ld_ok) {
    const unsigned ldb = (unsigned)(p.ld * es), so0 = soff(ft.kb - 1);
    vx_c = LD(rs(p.vx), ft.kb - 1);
    vy_c = LD(rs(p.vy), ft.kb - 1);
    vz_c = LD(rs(p.vz), ft.kb - 1);
    vz_m = LD(rs(p.vz), ft.kb - 2);
    if (j <= ny) {
      yv_x = BufOp<T>::ld(rs(p.vx), vo + ldb, so0);
      yv_z = BufOp<T>::ld(rs(p.vz), vo + ldb, so0);
    }
    yv_y = BufOp<T>::ld(rs(p.vy), vo - ldb, so0);
  }
  C.vxp = C.vyp = C.vzp = C.cl = C.cm = C.cb = zero;
#pragma unroll
  for (int q = 0; q < 6; ++q) C.so[q] = zero;
  N = C;
  load(C, ft.kb - 1);
  // history for the velocity update of plane s - 1
  T vo_x = zero, vo_y = zero, vo_z = zero;  // v(s-1) old
  T a_xx = zero, a_yy = zero, a_zz = zero, a_xy = zero, a_xz = zero, a_yz = zero;  // sigma(s-1)
  T b_xz = zero, b_yz = zero;  // sigma(s-2) xz, yz

  int t3 = 0, tv = 0;  // LDS buffers of plane s: sigma (iteration mod 3), v (mod 2)
  for (int s = ft.kb - 1; s <= ft.ke; ++s) {
    const KArgs Q = args();
    const int kv = s - 1;
    const bool inside = s >= 0 && s < nz && in_xy;
    const bool st = own && s >= ft.kb && s < ft.ke;
    const bool dov = own && kv >= ft.kb && kv < ft.ke;
    // memory variables: old values of sigma(s) [0..2] and v(s-1) [3..5] per
    // axis, loaded before the next plane's loads so that waiting for them
    // does not wait for those
    const bool inz = Z && inside && p.sz.in(s), inzv = Z && dov && p.sz.in(kv);
    const unsigned ox = p.sx.s(i) + (unsigned)p.sx.w * (j + (unsigned)ny * s);
    const unsigned oy = i + (unsigned)p.mld * (p.sy.s(j) + (unsigned)p.sy.w * s);
    const unsigned oz = i + (unsigned)p.mld * (j + (unsigned)ny * p.sz.s(s));
    const unsigned oxv = ox - (unsigned)p.sx.w * ny, oyv = oy - (unsigned)p.mld * p.sy.w;
    const unsigned ozv = i + (unsigned)p.mld * (j + (unsigned)ny * p.sz.s(kv));
    T mox[6], moy[6], moz[6];
    PV<T> PZ{}, PZV{};
#pragma unroll
    for (int q = 0; q < 6; ++q) mox[q] = moy[q] = moz[q] = T(0);
    if (SLAB) {
      if (inside && inx)
#pragma unroll
        for (int q = 0; q < 3; ++q) mox[q] = Q->a.mx[q][ox];
      if (inside && iny)
#pragma unroll
        for (int q = 0; q < 3; ++q) moy[q] = Q->a.my[q][oy];
      if (inz) {
#pragma unroll
        for (int q = 0; q < 3; ++q) moz[q] = Q->a.mz[q][oz];
        PZ = pv<T>(p.pz, s);
      }
      if (dov && inx)
#pragma unroll
        for (int q = 3; q < 6; ++q) mox[q] = Q->a.mx[q][oxv];
      if (dov && iny)
#pragma unroll
        for (int q = 3; q < 6; ++q) moy[q] = Q->a.my[q][oyv];
      if (inzv) {
#pragma unroll
        for (int q = 3; q < 6; ++q) moz[q] = Q->a.mz[q][ozv];
        PZV = pv<T>(p.pz, kv);
      }
    }
    if (s + 1 <= ft.ke) load(N, s + 1);
    const T cb = C.cb;  // DELTAT/rho at s - 1
    // x neighbours of v(s) across lanes (lanes 0 and 63 read their own value;
    // the ring columns never use those derivatives)
    const T vx_xp = lane_val(vx_c, dn4);
    const T vy_xm = lane_val(vy_c, up4);
    const T vz_xm = lane_val(vz_c, up4);

    // ---- sigma(s) at every point of the block
    T sn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) sn[q] = C.so[q];
    if (inside) {
      const T cl = C.cl, cm = C.cm;
      // sigma_xx, sigma_yy, sigma_zz: 3D:836-863
      const bool r_n = i <= nx - 2 && j >= 1 && s >= 1;
      {
        T dvx_dx = (vx_xp - vx_c) * p.odx;
        T dvy_dy = (vy_c - yv_y) * p.ody;
        T dvz_dz = (vz_c - vz_m) * p.odz;
        dvx_dx = pml(r_n, inx, dvx_dx, mox[0], st ? Q->mx[0] + ox : nullptr, X.ah, X.bh, X.Kh);
        dvy_dy = pml(r_n, iny, dvy_dy, moy[0], st ? Q->my[0] + oy : nullptr, Y.a, Y.b, Y.K);
        dvz_dz = pml(r_n, inz, dvz_dz, moz[0], st ? Q->mz[0] + oz : nullptr, PZ.a, PZ.b, PZ.K);
        if (r_n) {
          const T lp2m = cl + T(2) * cm;
          sn[0] = lp2m * dvx_dx + cl * (dvy_dy + dvz_dz) + C.so[0];
          sn[1] = cl * (dvx_dx + dvz_dz) + lp2m * dvy_dy + C.so[1];
          sn[2] = cl * (dvx_dx + dvy_dy) + lp2m * dvz_dz + C.so[2];
        }
      }
      // sigma_xy: 3D:877-894
      const bool r_xy = i >= 1 && j <= ny - 2;
      {
        T dvy_dx = (vy_c - vy_xm) * p.odx;
        T dvx_dy = (yv_x - vx_c) * p.ody;
        dvy_dx = pml(r_xy, inx, dvy_dx, mox[1], st ? Q->mx[1] + ox : nullptr, X.a, X.b, X.K);
        dvx_dy = pml(r_xy, iny, dvx_dy, moy[1], st ? Q->my[1] + oy : nullptr, Y.ah, Y.bh, Y.Kh);
        if (r_xy) sn[3] = cm * (dvy_dx + dvx_dy) + C.so[3];
      }
      // sigma_xz: 3D:908-925
      const bool r_xz = i >= 1 && s <= nz - 2;
      {
        T dvz_dx = (vz_c - vz_xm) * p.odx;
        T dvx_dz = (C.vxp - vx_c) * p.odz;
        dvz_dx = pml(r_xz, inx, dvz_dx, mox[2], st ? Q->mx[2] + ox : nullptr, X.a, X.b, X.K);
        dvx_dz = pml(r_xz, inz, dvx_dz, moz[1], st ? Q->mz[1] + oz : nullptr, PZ.ah, PZ.bh, PZ.Kh);
        if (r_xz) sn[4] = cm * (dvz_dx + dvx_dz) + C.so[4];
      }
      // sigma_yz: 3D:927-943
      const bool r_yz = j <= ny - 2 && s <= nz - 2;
      {
        T dvz_dy = (yv_z - vz_c) * p.ody;
        T dvy_dz = (C.vyp - vy_c) * p.odz;
        dvz_dy = pml(r_yz, iny, dvz_dy, moy[2], st ? Q->my[2] + oy : nullptr, Y.ah, Y.bh, Y.Kh);
        dvy_dz = pml(r_yz, inz, dvy_dz, moz[2], st ? Q->mz[2] + oz : nullptr, PZ.ah, PZ.bh, PZ.Kh);
        if (r_yz) sn[5] = cm * (dvz_dy + dvy_dz) + C.so[5];
      }
      if (st) {
        const unsigned so_s = soff(s);
        BufOp<T>::st(sn[0], rs(Q->sxx), vo, so_s);
        BufOp<T>::st(sn[1], rs(Q->syy), vo, so_s);
        BufOp<T>::st(sn[2], rs(Q->szz), vo, so_s);
        BufOp<T>::st(sn[3], rs(Q->sxy), vo, so_s);
        BufOp<T>::st(sn[4], rs(Q->sxz), vo, so_s);
        BufOp<T>::st(sn[5], rs(Q->syz), vo, so_s);
      }
    }

    // ---- exchange: y neighbours of sigma(s) and of v(s + 1)
    xs[t3][0][w][lane] = sn[3];
    xs[t3][1][w][lane] = sn[5];
    xs[t3][2][w][lane] = sn[1];
    xv[tv][0][w][lane] = C.vxp;
    xv[tv][1][w][lane] = C.vyp;
    xv[tv][2][w][lane] = C.vzp;
    // x neighbours of sigma(s - 1) across lanes
    const T n_xx_xm = lane_val(a_xx, up4);
    const T n_xy_xp = lane_val(a_xy, dn4);
    const T n_xz_xp = lane_val(a_xz, dn4);
    __syncthreads();
    // rows outside the block (w - 1 < 0, w + 1 > TY + 1) read the block's own
    // edge rows: values that only the ring rows' unused derivatives take
    const T nyv_x = xv[tv][0][wp][lane];
    const T nyv_y = xv[tv][1][wm][lane];
    const T nyv_z = xv[tv][2][wp][lane];

    // ---- v(s - 1) at the owned points
    if (dov) {
      const int tp = t3 == 0 ? 2 : t3 - 1;  // LDS buffer of plane s - 1
      const T n_xy_ym = xs[tp][0][w - 1][lane];
      const T n_yz_ym = xs[tp][1][w - 1][lane];
      const T n_yy_yp = xs[tp][2][w + 1][lane];
      T vxn = vo_x, vyn = vo_y, vzn = vo_z;
      // vx: 3D:976-993
      const bool r_vx = i >= 1 && j >= 1 && kv >= 1;
      {
        T dsxx_dx = (a_xx - n_xx_xm) * p.odx;
        T dsxy_dy = (a_xy - n_xy_ym) * p.ody;
        T dsxz_dz = (a_xz - b_xz) * p.odz;
        dsxx_dx = pml(r_vx, inx, dsxx_dx, mox[3], Q->mx[3] + oxv, X.a, X.b, X.K);
        dsxy_dy = pml(r_vx, iny, dsxy_dy, moy[3], Q->my[3] + oyv, Y.a, Y.b, Y.K);
        dsxz_dz = pml(r_vx, inzv, dsxz_dz, moz[3], Q->mz[3] + ozv, PZV.a, PZV.b, PZV.K);
        if (r_vx) vxn = cb * (dsxx_dx + dsxy_dy + dsxz_dz) + vxn;
      }
      // vy: 3D:995-1017
      const bool r_vy = i <= nx - 2 && j <= ny - 2 && kv >= 1;
      {
        T dsxy_dx = (n_xy_xp - a_xy) * p.odx;
        T dsyy_dy = (n_yy_yp - a_yy) * p.ody;
        T dsyz_dz = (a_yz - b_yz) * p.odz;
        dsxy_dx = pml(r_vy, inx, dsxy_dx, mox[4], Q->mx[4] + oxv, X.ah, X.bh, X.Kh);
        dsyy_dy = pml(r_vy, iny, dsyy_dy, moy[4], Q->my[4] + oyv, Y.ah, Y.bh, Y.Kh);
        dsyz_dz = pml(r_vy, inzv, dsyz_dz, moz[4], Q->mz[4] + ozv, PZV.a, PZV.b, PZV.K);
        if (r_vy) vyn = cb * (dsxy_dx + dsyy_dy + dsyz_dz) + vyn;
      }
      // vz: 3D:1031-1052
      const bool r_vz = i <= nx - 2 && j >= 1 && kv <= nz - 2;
      {
        T dsxz_dx = (n_xz_xp - a_xz) * p.odx;
        T dsyz_dy = (a_yz - n_yz_ym) * p.ody;
        T dszz_dz = (sn[2] - a_zz) * p.odz;
        dsxz_dx = pml(r_vz, inx, dsxz_dx, mox[5], Q->mx[5] + oxv, X.ah, X.bh, X.Kh);
        dsyz_dy = pml(r_vz, iny, dsyz_dy, moy[5], Q->my[5] + oyv, Y.a, Y.b, Y.K);
        dszz_dz = pml(r_vz, inzv, dszz_dz, moz[5], Q->mz[5] + ozv, PZV.ah, PZV.bh, PZV.Kh);
        if (r_vz) vzn = cb * (dsxz_dx + dsyz_dy + dszz_dz) + vzn;
      }
      // point force (3D:1074-1082) and rigid faces (3D:1087-1121)
      if (i == p.isrc && j == p.jsrc && kv == p.ksrc) {
        const int it = *Q->a.step;
        const T rsrc = Q->a.rho[i + (long)j * p.ld + (long)kv * p.ps];
        vxn = vxn + Q->a.fx[it - 1] * p.dt / rsrc;
        vyn = vyn + Q->a.fy[it - 1] * p.dt / rsrc;
      }
      if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || kv == 0 || kv == nz - 1) {
        vxn = T(0);
        vyn = T(0);
        vzn = T(0);
      }
      const unsigned so_v = soff(kv);
      BufOp<T>::st(vxn, rs(Q->vx), vo, so_v);
      BufOp<T>::st(vyn, rs(Q->vy), vo, so_v);
      BufOp<T>::st(vzn, rs(Q->vz), vo, so_v);
    }

    // ---- shift the planes
    b_xz = a_xz;
    b_yz = a_yz;
    a_xx = sn[0];
    a_yy = sn[1];
    a_zz = sn[2];
    a_xy = sn[3];
    a_xz = sn[4];
    a_yz = sn[5];
    vo_x = vx_c;
    vo_y = vy_c;
    vo_z = vz_c;
    vz_m = vz_c;
    vx_c = C.vxp;
    vy_c = C.vyp;
    vz_c = C.vzp;
    yv_x = nyv_x;
    yv_y = nyv_y;
    yv_z = nyv_z;
    C = N;
    t3 = t3 == 2 ? 0 : t3 + 1;
    tv ^= 1;
  }
}

// Tiles along one axis (tile size B, extent n, slab widths wl, wr) whose
// points and 1-point ring avoid both slabs: [t0, t1), empty when t1 <= t0.
void clean_tiles(int n, int B, int wl, int wr, int &t0, int &t1) {
  const int nt = (n + B - 1) / B;
  t0 = t1 = 0;
  bool seen = false;
  for (int t = 0; t < nt; ++t) {
    const int lo = t * B - 1 > 0 ? t * B - 1 : 0, hi = t * B + B < n - 1 ? t * B + B : n - 1;
    const bool clean = lo >= wl && hi < n - wr;
    if (clean && !seen) t0 = t, seen = true;
    if (clean) t1 = t + 1;
    else if (seen) break;  // slabs lie at the ends: the clean tiles are contiguous
  }
}

// z chunks: the planes whose chunk ring (planes kb - 1 and ke) can avoid the
// z slabs, [z0, z1) with z0 = wl + 1 and z1 = nz - wr - 1 (or the grid end
// when a side has no slab), form the middle region; each region is cut into
// chunks of at most kchunk planes, so the chunks that touch a slab are no
// thicker than the slab.
template <typename T>
void z_chunks(const Step3D<T> &p, Fused3D<T> &Q) {
  const int nz = p.nz, kc = p.kchunk;
  int z0 = p.sz.wl > 0 ? p.sz.wl + 1 : 0;
  int z1 = p.sz.wr > 0 ? nz - p.sz.wr - 1 : nz;
  if (z1 <= z0) z0 = z1 = 0;  // no clean planes: one region of uniform chunks
  Q.z0 = z0;
  Q.z1 = z1;
  Q.zn[0] = (z0 + kc - 1) / kc;
  Q.zn[1] = (z1 - z0 + kc - 1) / kc;
  Q.zn[2] = (nz - z1 + kc - 1) / kc;
}

template <typename T, int TY>
hipError_t fused_ty(const Fused3D<T> &P, hipStream_t s) {
  const Step3D<T> &p = P.a;
  Fused3D<T> Q = P;
  z_chunks(p, Q);
  clean_tiles(p.nx, FO, p.sx.wl, p.sx.wr, Q.box[0], Q.box[1]);
  clean_tiles(p.ny, TY, p.sy.wl, p.sy.wr, Q.box[2], Q.box[3]);
  if (Q.box[1] == Q.box[0] || Q.box[3] == Q.box[2]) Q.box[0] = Q.box[1] = Q.box[2] = Q.box[3] = 0;
  // the middle z region is clean by construction
  Q.box[4] = Q.zn[0];
  Q.box[5] = Q.zn[0] + Q.zn[1];
  const unsigned ntxy = (unsigned)((p.nx + FO - 1) / FO) * ((p.ny + TY - 1) / TY);
  const unsigned nbxy = (unsigned)(Q.box[1] - Q.box[0]) * (Q.box[3] - Q.box[2]);
  const unsigned ncz = Q.box[5] - Q.box[4], ncr = Q.zn[0] + Q.zn[2];
  const dim3 blk(FW * (TY + 2));
  if (nbxy * ncz) hipLaunchKernelGGL((k_fused3d<T, TY, false, false>), dim3(nbxy * ncz), blk, 0, s, Q);
  if (nbxy * ncr) hipLaunchKernelGGL((k_fused3d<T, TY, false, true>), dim3(nbxy * ncr), blk, 0, s, Q);
  if ((ntxy - nbxy) * ncz)
    hipLaunchKernelGGL((k_fused3d<T, TY, true, false>), dim3((ntxy - nbxy) * ncz), blk, 0, s, Q);
  if ((ntxy - nbxy) * ncr)
    hipLaunchKernelGGL((k_fused3d<T, TY, true, true>), dim3((ntxy - nbxy) * ncr), blk, 0, s, Q);
  return hipGetLastError();
}

}  // namespace

// tile rows 6 (8 waves per block) for tile_y <= 8, else 14 (16 waves)
template <typename T>
hipError_t launch_fused3d(const Fused3D<T> &P, hipStream_t s) {
  if (P.a.tile_y > 8) return fused_ty<T, 14>(P, s);
  return fused_ty<T, 6>(P, s);
}

template hipError_t launch_fused3d<float>(const Fused3D<float> &, hipStream_t);
template hipError_t launch_fused3d<double>(const Fused3D<double> &, hipStream_t);

}  // namespace cpml
