// HIP kernels of the C-PML time loop for gfx950 (CDNA4).
//
// Arithmetic follows the reference programs statement by statement (the file
// is compiled with -ffp-contract=off so no multiply-add is fused), which makes
// the float64 path bit-identical to the reference and the float32 path equal
// to the reference compiled in single precision:
//   2D stress       seismic_CPML_2D_isotropic_second_order.f90:556-600
//                   seismic_CPML_2D_isotropic_fourth_order.f90:557-601
//   2D velocity     ...second_order.f90:606-641, ...fourth_order.f90:607-642
//   source          ...second_order.f90:644-667
//   Dirichlet       ...second_order.f90:670-680
//   3D stress       seismic_CPML_3D_isotropic_MPI_OpenMP.f90:836-943
//   3D velocity     ...MPI_OpenMP.f90:976-1052, source 1055-1083, faces 1087-1121
//   receivers       ...second_order.f90:683-686, ...MPI_OpenMP.f90:1123-1129
//   energy          ...second_order.f90:694-713, ...MPI_OpenMP.f90:1131-1180
//
// Thread mapping: a 64x4 block covers 64 points along x (one wavefront per
// row, coalesced along the fastest index) and 4 rows along y.  3D kernels
// march each (i,j) column through a chunk of z planes and carry the z
// neighbours in registers, so every plane of every field is fetched once per
// chunk; x/y neighbours come from the same or adjacent rows of the block and
// hit in L1/L2.
#include "cpml_device.h"
#include "cpml_internal.h"
#include "cpml_energy.h"
#include "cpml_points2d.h"

namespace cpml {

namespace {


// ---------------------------------------------------------------- 2D ----

// XCD-aware tile order: the hardware hands workgroup h to XCD h % 8; each XCD
// gets one contiguous band of tiles (x fastest), so the rows and columns that
// neighbouring tiles share are read through the same L2.  Sets this thread's
// grid point (i, j).
template <int TY, int RG>
__device__ __forceinline__ void tile2d(int nx, int ny, int &i, int &j) {
  const int ntx = (nx + BX - 1) / BX, nt = ntx * ((ny + TY * RG - 1) / (TY * RG));
  const int h = blockIdx.x, xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, r = nt & 7;
  const int t = xcd * q + min(xcd, r) + loc;
  i = (t % ntx) * BX + threadIdx.x;
  j = (t / ntx) * (TY * RG) + threadIdx.y;
}

// A tile is 64 points by TY*RG rows; the block (64 x TY threads) walks its RG
// row groups, so each wavefront lives for several rows.
template <typename T, int ORDER, int TY, int RG, bool NT>
__global__ __launch_bounds__(BX *TY) void k_stress2d(Step2D<T> p) {
  int i, j;
  tile2d<TY, RG>(p.nx, p.ny, i, j);
#pragma unroll
  for (int r = 0; r < RG; ++r)
    if (i < p.nx && j + r * TY < p.ny) stress2d_point<T, ORDER, NT>(p, i, j + r * TY);
}

template <typename T, int ORDER, int TY, int RG, bool NT>
__global__ __launch_bounds__(BX *TY) void k_velocity2d(Step2D<T> p) {
  int i, j;
  tile2d<TY, RG>(p.nx, p.ny, i, j);
#pragma unroll
  for (int r = 0; r < RG; ++r)
    if (i < p.nx && j + r * TY < p.ny) velocity2d_point<T, ORDER, NT>(p, i, j + r * TY);
}

// ---------------------------------------------------------------- 3D ----
//
// A 3D tile is 64 points along x (one wavefront per row) by TY3 rows, and
// marches through a chunk of z planes carrying the z neighbours in
// registers.  Field offsets are unsigned 32-bit byte offsets into a buffer
// rebased at the tile's first halo plane (the host keeps a chunk's planes plus
// 8 halo and pipeline planes below 4 GiB: chunk_fits() in cpml_api.hip),
// added to uniform base pointers, and the loads of plane k+1 are issued
// before the arithmetic and stores of plane k (register double buffering).

struct Tile3 {
  int i, j, kb, ke;
  bool ok;
};

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD one
// contiguous band of tiles (x fastest, then y, then z chunk) so the tiles that
// read each other's halo rows run on the same XCD and share its L2; the z
// chunks are interleaved over the bands (zchunk_of).
template <int TY3>
__device__ __forceinline__ Tile3 tile3(int nx, int ny, int nz, int kchunk) {
  const int ntx = (nx + BX - 1) / BX, nty = (ny + TY3 - 1) / TY3;
  const int ntz = (nz + kchunk - 1) / kchunk;
  const int nt = ntx * nty * ntz;
  const int h = blockIdx.x;
  const int xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, r = nt & 7;
  const int t = xcd * q + min(xcd, r) + loc;
  const int tx = t % ntx, rest = t / ntx;
  const int ty = rest % nty, tz = zchunk_of(rest / nty, ntz);
  Tile3 o;
  o.i = tx * BX + threadIdx.x;
  o.j = ty * TY3 + threadIdx.y;
  o.kb = tz * kchunk;
  o.ke = min(o.kb + kchunk, nz);
  o.ok = o.i < nx && o.j < ny;
  return o;
}

// Geometry of the rebased chunk buffers of one tile.
template <typename T>
struct Chunk3 {
  unsigned v0;   // byte offset of (i-1, j-1) in buffer plane 0
  unsigned ldb;  // row stride in bytes
  unsigned psb;  // plane stride in bytes
  unsigned bytes;
  long origin;   // element offset of buffer plane 0 from the allocation start
  __device__ Chunk3(const Step3D<T> &p, const Tile3 &t) {
    ldb = (unsigned)(p.ld * sizeof(T));
    psb = (unsigned)(p.ps * sizeof(T));
    bytes = (unsigned)((t.ke - t.kb + 2) * p.ps * sizeof(T));
    origin = (long)(t.kb - 1 + p.R) * p.ps;  // allocation plane of interior plane kb-1
    v0 = (unsigned)(((t.j - 1 + p.R) * p.ld + p.xpad + t.i - 1) * sizeof(T));
  }
  // SGPR part of the offset of (i+dx, j+dy, k+dz) for buffer plane kk = k-kb+1
  __device__ unsigned so(int kk, int dy, int dz) const {
    return (unsigned)(kk + dz) * psb + (unsigned)(1 + dy) * ldb;
  }
  __device__ __amdgpu_buffer_rsrc_t rsrc(const T *interior_origin, long base) const {
    return make_rsrc(interior_origin - base + origin, bytes);
  }
};

// load / store of (i+DX, j+dy, k+dz): the x shift goes to the immediate
#define CPML_LD(r, DX, dy, dz) BufOp<T>::ld(r, g.v0 + (1 + (DX)) * (unsigned)sizeof(T), g.so(kk, dy, dz))
// loads of values only this point reads, and the stores (non-temporal when
// the state is larger than the caches)
#define CPML_LDN(r) \
  BufOp<T>::template ldp<NT ? kNonTemporal : 0>(r, g.v0 + (unsigned)sizeof(T), g.so(kk, 0, 0))
#define CPML_ST(x, r) \
  BufOp<T>::template stp<NT ? kNonTemporal : 0>(x, r, g.v0 + (unsigned)sizeof(T), g.so(kk, 0, 0))

// Values the stress kernel reads at one plane.
template <typename T>
struct SPlane {
  T vx_p, vy_p, vz_k;                              // vx, vy at k+1; vz at k
  T vx_xp, vx_yp, vy_ym, vy_xm, vz_xm, vz_yp;      // x/y neighbours at k
  T cl, cm, sxx, syy, szz, sxy, sxz, syz;
};

template <typename T, int TY3, bool NT>
__global__ __launch_bounds__(BX *TY3) void k_stress3d(Step3D<T> p) {
  const Tile3 t = tile3<TY3>(p.nx, p.ny, p.nz, p.kchunk);
  if (!t.ok) return;
  const int i = t.i, j = t.j;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const Chunk3<T> g(p, t);
  const auto rvx = g.rsrc(p.vx, p.base), rvy = g.rsrc(p.vy, p.base), rvz = g.rsrc(p.vz, p.base);
  const auto rcl = g.rsrc(p.clam, p.base), rcm = g.rsrc(p.cmu, p.base);
  const auto rxx = g.rsrc(p.sxx, p.base), ryy = g.rsrc(p.syy, p.base);
  const auto rzz = g.rsrc(p.szz, p.base), rxy = g.rsrc(p.sxy, p.base);
  const auto rxz = g.rsrc(p.sxz, p.base), ryz = g.rsrc(p.syz, p.base);
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const unsigned mxo = p.sx.s(i) + (unsigned)p.sx.w * j, mxs = (unsigned)p.sx.w * ny;
  const unsigned myo = i + (unsigned)p.mld * p.sy.s(j), mys = (unsigned)p.mld * p.sy.w;
  const unsigned mzo = i + (unsigned)p.mld * j, mzs = (unsigned)p.mld * ny;
  const bool r_n = i <= nx - 2 && j >= 1;
  const bool r_xy = i >= 1 && j <= ny - 2;
  const bool r_xz = i >= 1;
  const bool r_yz = j <= ny - 2;
  auto load = [&](SPlane<T> &L, int kk) {
    L.vx_p = CPML_LD(rvx, 0, 0, 1);
    L.vy_p = CPML_LD(rvy, 0, 0, 1);
    L.vz_k = CPML_LD(rvz, 0, 0, 0);
    L.vx_xp = CPML_LD(rvx, 1, 0, 0);
    L.vx_yp = CPML_LD(rvx, 0, 1, 0);
    L.vy_ym = CPML_LD(rvy, 0, -1, 0);
    L.vy_xm = CPML_LD(rvy, -1, 0, 0);
    L.vz_xm = CPML_LD(rvz, -1, 0, 0);
    L.vz_yp = CPML_LD(rvz, 0, 1, 0);
    L.cl = CPML_LDN(rcl);
    L.cm = CPML_LDN(rcm);
    L.sxx = CPML_LDN(rxx);
    L.syy = CPML_LDN(ryy);
    L.szz = CPML_LDN(rzz);
    L.sxy = CPML_LDN(rxy);
    L.sxz = CPML_LDN(rxz);
    L.syz = CPML_LDN(ryz);
  };
  // x / y profile values of this column (loop invariant; only read in the
  // absorbing slabs)
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
  int kk = 1;  // buffer plane of k = kb
  T vz_m = CPML_LD(rvz, 0, 0, -1);
  T vx_k = CPML_LD(rvx, 0, 0, 0), vy_k = CPML_LD(rvy, 0, 0, 0);
  SPlane<T> L;
  load(L, kk);
  for (int k = t.kb; k < t.ke; ++k, ++kk) {
    const auto Q = kernarg<Step3D<T>>();
    SPlane<T> N = L;
    if (k + 1 < t.ke) load(N, kk + 1);
    const bool inz = p.sz.in(k);
    const unsigned mzk = mzo + (unsigned)p.sz.s(k) * mzs;
    const T cl = L.cl, cm = L.cm;
    // sigma_xx, sigma_yy, sigma_zz: 3D:836-863 (global k >= 2)
    if (r_n && k >= 1) {
      T dvx_dx = (L.vx_xp - vx_k) * p.odx;
      T dvy_dy = (vy_k - L.vy_ym) * p.ody;
      T dvz_dz = (L.vz_k - vz_m) * p.odz;
      if (inx) dvx_dx = cpml(dvx_dx, Q->mx[0] + mxo + k * mxs, xah, xbh, xKh);
      if (iny) dvy_dy = cpml(dvy_dy, Q->my[0] + myo + k * mys, ya, yb, yK);
      if (inz) dvz_dz = cpml(dvz_dz, Q->mz[0] + mzk, Q->pz.a[k], Q->pz.b[k], Q->pz.K[k]);
      const T lp2m = cl + T(2) * cm;
      CPML_ST(lp2m * dvx_dx + cl * (dvy_dy + dvz_dz) + L.sxx, rxx);
      CPML_ST(cl * (dvx_dx + dvz_dz) + lp2m * dvy_dy + L.syy, ryy);
      CPML_ST(cl * (dvx_dx + dvy_dy) + lp2m * dvz_dz + L.szz, rzz);
    }
    // sigma_xy: 3D:877-894
    if (r_xy) {
      T dvy_dx = (vy_k - L.vy_xm) * p.odx;
      T dvx_dy = (L.vx_yp - vx_k) * p.ody;
      if (inx) dvy_dx = cpml(dvy_dx, Q->mx[1] + mxo + k * mxs, xa, xb, xK);
      if (iny) dvx_dy = cpml(dvx_dy, Q->my[1] + myo + k * mys, yah, ybh, yKh);
      CPML_ST(cm * (dvy_dx + dvx_dy) + L.sxy, rxy);
    }
    // sigma_xz: 3D:908-925 (global k <= NZ-1)
    if (r_xz && k <= nz - 2) {
      T dvz_dx = (L.vz_k - L.vz_xm) * p.odx;
      T dvx_dz = (L.vx_p - vx_k) * p.odz;
      if (inx) dvz_dx = cpml(dvz_dx, Q->mx[2] + mxo + k * mxs, xa, xb, xK);
      if (inz) dvx_dz = cpml(dvx_dz, Q->mz[1] + mzk, Q->pz.ah[k], Q->pz.bh[k], Q->pz.Kh[k]);
      CPML_ST(cm * (dvz_dx + dvx_dz) + L.sxz, rxz);
    }
    // sigma_yz: 3D:927-943
    if (r_yz && k <= nz - 2) {
      T dvz_dy = (L.vz_yp - L.vz_k) * p.ody;
      T dvy_dz = (L.vy_p - vy_k) * p.odz;
      if (iny) dvz_dy = cpml(dvz_dy, Q->my[2] + myo + k * mys, yah, ybh, yKh);
      if (inz) dvy_dz = cpml(dvy_dz, Q->mz[2] + mzk, Q->pz.ah[k], Q->pz.bh[k], Q->pz.Kh[k]);
      CPML_ST(cm * (dvz_dy + dvy_dz) + L.syz, ryz);
    }
    vz_m = L.vz_k;
    vx_k = L.vx_p;
    vy_k = L.vy_p;
    L = N;
  }
}

// Values the velocity kernel reads at one plane.
template <typename T>
struct VPlane {
  T szz_p, sxz_k, syz_k;                           // szz at k+1; sxz, syz at k
  T sxx_k, sxx_xm, sxy_k, sxy_ym, sxy_xp, syy_k, syy_yp, sxz_xp, syz_ym;
  T cb, vx, vy, vz;
};

template <typename T, int TY3, bool NT>
__global__ __launch_bounds__(BX *TY3) void k_velocity3d(Step3D<T> p) {
  const Tile3 t = tile3<TY3>(p.nx, p.ny, p.nz, p.kchunk);
  if (!t.ok) return;
  const int i = t.i, j = t.j;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const Chunk3<T> g(p, t);
  const auto rxx = g.rsrc(p.sxx, p.base), ryy = g.rsrc(p.syy, p.base);
  const auto rzz = g.rsrc(p.szz, p.base), rxy = g.rsrc(p.sxy, p.base);
  const auto rxz = g.rsrc(p.sxz, p.base), ryz = g.rsrc(p.syz, p.base);
  const auto rcb = g.rsrc(p.cbuoy, p.base);
  const auto rvx = g.rsrc(p.vx, p.base), rvy = g.rsrc(p.vy, p.base), rvz = g.rsrc(p.vz, p.base);
  const bool inx = p.sx.in(i), iny = p.sy.in(j);
  const unsigned mxo = p.sx.s(i) + (unsigned)p.sx.w * j, mxs = (unsigned)p.sx.w * ny;
  const unsigned myo = i + (unsigned)p.mld * p.sy.s(j), mys = (unsigned)p.mld * p.sy.w;
  const unsigned mzo = i + (unsigned)p.mld * j, mzs = (unsigned)p.mld * ny;
  const bool r_vx = i >= 1 && j >= 1;
  const bool r_vy = i <= nx - 2 && j <= ny - 2;
  const bool r_vz = i <= nx - 2 && j >= 1;
  const bool edge_xy = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
  const bool src_col = i == p.isrc && j == p.jsrc;
  auto load = [&](VPlane<T> &L, int kk) {
    L.szz_p = CPML_LD(rzz, 0, 0, 1);
    L.sxz_k = CPML_LD(rxz, 0, 0, 0);
    L.syz_k = CPML_LD(ryz, 0, 0, 0);
    L.sxx_k = CPML_LD(rxx, 0, 0, 0);
    L.sxx_xm = CPML_LD(rxx, -1, 0, 0);
    L.sxy_k = CPML_LD(rxy, 0, 0, 0);
    L.sxy_ym = CPML_LD(rxy, 0, -1, 0);
    L.sxy_xp = CPML_LD(rxy, 1, 0, 0);
    L.syy_k = CPML_LD(ryy, 0, 0, 0);
    L.syy_yp = CPML_LD(ryy, 0, 1, 0);
    L.sxz_xp = CPML_LD(rxz, 1, 0, 0);
    L.syz_ym = CPML_LD(ryz, 0, -1, 0);
    L.cb = CPML_LDN(rcb);
    L.vx = CPML_LDN(rvx);
    L.vy = CPML_LDN(rvy);
    L.vz = CPML_LDN(rvz);
  };
  // x / y profile values of this column (loop invariant; only read in the
  // absorbing slabs)
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
  int kk = 1;
  T sxz_m = CPML_LD(rxz, 0, 0, -1), syz_m = CPML_LD(ryz, 0, 0, -1);
  T szz_k = CPML_LD(rzz, 0, 0, 0);
  VPlane<T> L;
  load(L, kk);
  for (int k = t.kb; k < t.ke; ++k, ++kk) {
    const auto Q = kernarg<Step3D<T>>();
    VPlane<T> N = L;
    if (k + 1 < t.ke) load(N, kk + 1);
    const bool inz = p.sz.in(k);
    const unsigned mzk = mzo + (unsigned)p.sz.s(k) * mzs;
    const T cb = L.cb;
    T vxn = L.vx, vyn = L.vy, vzn = L.vz;
    // vx: 3D:976-993 (global k >= 2)
    if (r_vx && k >= 1) {
      T dsxx_dx = (L.sxx_k - L.sxx_xm) * p.odx;
      T dsxy_dy = (L.sxy_k - L.sxy_ym) * p.ody;
      T dsxz_dz = (L.sxz_k - sxz_m) * p.odz;
      if (inx) dsxx_dx = cpml(dsxx_dx, Q->mx[3] + mxo + k * mxs, xa, xb, xK);
      if (iny) dsxy_dy = cpml(dsxy_dy, Q->my[3] + myo + k * mys, ya, yb, yK);
      if (inz) dsxz_dz = cpml(dsxz_dz, Q->mz[3] + mzk, Q->pz.a[k], Q->pz.b[k], Q->pz.K[k]);
      vxn = cb * (dsxx_dx + dsxy_dy + dsxz_dz) + vxn;
    }
    // vy: 3D:995-1017
    if (r_vy && k >= 1) {
      T dsxy_dx = (L.sxy_xp - L.sxy_k) * p.odx;
      T dsyy_dy = (L.syy_yp - L.syy_k) * p.ody;
      T dsyz_dz = (L.syz_k - syz_m) * p.odz;
      if (inx) dsxy_dx = cpml(dsxy_dx, Q->mx[4] + mxo + k * mxs, xah, xbh, xKh);
      if (iny) dsyy_dy = cpml(dsyy_dy, Q->my[4] + myo + k * mys, yah, ybh, yKh);
      if (inz) dsyz_dz = cpml(dsyz_dz, Q->mz[4] + mzk, Q->pz.a[k], Q->pz.b[k], Q->pz.K[k]);
      vyn = cb * (dsxy_dx + dsyy_dy + dsyz_dz) + vyn;
    }
    // vz: 3D:1031-1052 (global k <= NZ-1)
    if (r_vz && k <= nz - 2) {
      T dsxz_dx = (L.sxz_xp - L.sxz_k) * p.odx;
      T dsyz_dy = (L.syz_k - L.syz_ym) * p.ody;
      T dszz_dz = (L.szz_p - szz_k) * p.odz;
      if (inx) dsxz_dx = cpml(dsxz_dx, Q->mx[5] + mxo + k * mxs, xah, xbh, xKh);
      if (iny) dsyz_dy = cpml(dsyz_dy, Q->my[5] + myo + k * mys, ya, yb, yK);
      if (inz) dszz_dz = cpml(dszz_dz, Q->mz[5] + mzk, Q->pz.ah[k], Q->pz.bh[k], Q->pz.Kh[k]);
      vzn = cb * (dsxz_dx + dsyz_dy + dszz_dz) + vzn;
    }
    // point force on the cut plane, no density average (3D:1074-1082)
    if (src_col && k == p.ksrc) {
      const int it = *Q->step;
      const T rs = Q->rho[i + (long)j * p.ld + (long)k * p.ps];
      vxn = vxn + Q->fx[it - 1] * p.dt / rs;
      vyn = vyn + Q->fy[it - 1] * p.dt / rs;
    }
    // six rigid faces (3D:1087-1121)
    if (edge_xy || k == 0 || k == nz - 1) {
      vxn = T(0);
      vyn = T(0);
      vzn = T(0);
    }
    CPML_ST(vxn, rvx);
    CPML_ST(vyn, rvy);
    CPML_ST(vzn, rvz);
    sxz_m = L.sxz_k;
    syz_m = L.syz_k;
    szz_k = L.szz_p;
    L = N;
  }
}

#undef CPML_LD
#undef CPML_LDN
#undef CPML_ST

// ------------------------------------------------------ receivers, step --

template <typename T>
__global__ __launch_bounds__(kMaxRecBlock) void k_record(Record<T> p) {
  const int it = *p.step + p.skip;
  const long row = (long)(it - 1) * p.nrec;
  for (int r = threadIdx.x; r < p.nrec; r += blockDim.x) {
    const long q = p.idx[r];
    p.sx[row + r] = p.avg4 ? (p.vx[q] + p.vx[q + 1] + p.vx[q + p.ld] + p.vx[q + 1 + p.ld]) / T(4)
                           : p.vx[q];
    p.sy[row + r] = p.vy[q];
    if (p.sz) p.sz[row + r] = p.vz[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) *p.step = it + 1;
}

// -------------------------------------------------------------- energy --

constexpr int EB = 256;

__device__ void block_sum2(double &a, double &b) {
  __shared__ double sa[EB], sb[EB];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = EB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  a = sa[0];
  b = sb[0];
}

template <typename T>
__global__ __launch_bounds__(EB) void k_energy_partial(Energy<T> p) {
  const int it = *p.step + p.offset;
  if (it % p.cadence != 0) return;
  const int wx = p.i1 - p.i0 + 1, wy = p.j1 - p.j0 + 1, wz = p.k1 - p.k0 + 1;
  const long n = (long)wx * wy * wz;
  double kin = 0.0, pot = 0.0;
  for (long t = blockIdx.x * (long)EB + threadIdx.x; t < n; t += (long)gridDim.x * EB) {
    const int i = p.i0 + (int)(t % wx);
    const long r = t / wx;
    const int j = p.j0 + (int)(r % wy);
    const int k = p.k0 + (int)(r / wy);
    cell_energy(p, i + j * p.ld + k * p.ps, kin, pot);
  }
  block_sum2(kin, pot);
  if (threadIdx.x == 0) {
    p.partial[2 * blockIdx.x] = kin;
    p.partial[2 * blockIdx.x + 1] = pot;
  }
}

template <typename T>
__global__ __launch_bounds__(EB) void k_energy_final(Energy<T> p, int nblocks) {
  const int it = *p.step + p.offset;
  if (it % p.cadence != 0) return;
  double kin = 0.0, pot = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += EB) {
    kin += p.partial[2 * b];
    pot += p.partial[2 * b + 1];
  }
  block_sum2(kin, pot);
  if (threadIdx.x == 0) {
    p.out_kin[it - 1] = p.ndim == 2 ? 0.5 * kin : kin;
    p.out_pot[it - 1] = pot;
  }
}

// ------------------------------------------------- max |v| (stability) --

template <typename T>
__global__ __launch_bounds__(EB) void k_vmax(const T *vx, const T *vy, const T *vz, int nx, int ny,
                                             int nz, long ld, long ps,
                                             unsigned long long *out) {
  const long n = (long)nx * ny * nz;
  double m = 0.0;
  for (long t = blockIdx.x * (long)EB + threadIdx.x; t < n; t += (long)gridDim.x * EB) {
    const int i = (int)(t % nx);
    const long r = t / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const long q = i + j * ld + k * ps;
    T s = vx[q] * vx[q] + vy[q] * vy[q];
    if (vz) s = s + vz[q] * vz[q];
    const double d = (double)s;
    // NaN compares false: force it to the top so a blow-up is reported
    m = (d > m || d != d) ? d : m;
  }
  __shared__ double sm[EB];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = EB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sm[threadIdx.x + s];
      if (o > sm[threadIdx.x] || o != o) sm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double v = sm[0];
    unsigned long long bits = v != v ? 0x7ff8000000000000ull : __double_as_longlong(v);
    atomicMax(out, bits);
  }
}

}  // namespace

// ------------------------------------------------------------ launchers --

template <typename T, int TY, int RG, bool NT>
static hipError_t stress2d_nt(int order, const Step2D<T> &p, hipStream_t s) {
  dim3 grid(((p.nx + BX - 1) / BX) * ((p.ny + TY * RG - 1) / (TY * RG))), block(BX, TY);
  if (order == 2) hipLaunchKernelGGL((k_stress2d<T, 2, TY, RG, NT>), grid, block, 0, s, p);
  else if (order == 4) hipLaunchKernelGGL((k_stress2d<T, 4, TY, RG, NT>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((k_stress2d<T, 8, TY, RG, NT>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <typename T, int TY, int RG>
static hipError_t stress2d_ty(int order, const Step2D<T> &p, hipStream_t s) {
  return p.streaming ? stress2d_nt<T, TY, RG, true>(order, p, s)
                     : stress2d_nt<T, TY, RG, false>(order, p, s);
}

template <typename T, int TY, int RG, bool NT>
static hipError_t velocity2d_nt(int order, const Step2D<T> &p, hipStream_t s) {
  dim3 grid(((p.nx + BX - 1) / BX) * ((p.ny + TY * RG - 1) / (TY * RG))), block(BX, TY);
  if (order == 2) hipLaunchKernelGGL((k_velocity2d<T, 2, TY, RG, NT>), grid, block, 0, s, p);
  else if (order == 4) hipLaunchKernelGGL((k_velocity2d<T, 4, TY, RG, NT>), grid, block, 0, s, p);
  else hipLaunchKernelGGL((k_velocity2d<T, 8, TY, RG, NT>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <typename T, int TY, int RG>
static hipError_t velocity2d_ty(int order, const Step2D<T> &p, hipStream_t s) {
  return p.streaming ? velocity2d_nt<T, TY, RG, true>(order, p, s)
                     : velocity2d_nt<T, TY, RG, false>(order, p, s);
}

// p.tile_y: rows per tile, 4, 8, 16, or 32 / 64 walked by 16-row blocks
template <typename T>
hipError_t launch_stress2d(int order, const Step2D<T> &p, hipStream_t s) {
  switch (p.tile_y) {
    case 64: return stress2d_ty<T, 16, 4>(order, p, s);
    case 32: return stress2d_ty<T, 16, 2>(order, p, s);
    case 16: return stress2d_ty<T, 16, 1>(order, p, s);
    case 8: return stress2d_ty<T, 8, 1>(order, p, s);
    default: return stress2d_ty<T, 4, 1>(order, p, s);
  }
}

template <typename T>
hipError_t launch_velocity2d(int order, const Step2D<T> &p, hipStream_t s) {
  switch (p.tile_y) {
    case 64: return velocity2d_ty<T, 16, 4>(order, p, s);
    case 32: return velocity2d_ty<T, 16, 2>(order, p, s);
    case 16: return velocity2d_ty<T, 16, 1>(order, p, s);
    case 8: return velocity2d_ty<T, 8, 1>(order, p, s);
    default: return velocity2d_ty<T, 4, 1>(order, p, s);
  }
}

template <typename T>
unsigned ntiles3(const Step3D<T> &p, int ty) {
  return (unsigned)((p.nx + BX - 1) / BX) * ((p.ny + ty - 1) / ty) *
         ((p.nz + p.kchunk - 1) / p.kchunk);
}

// p.tile_y selects the tile height (4, 8 or 16 rows of 64 points)
template <typename T, bool NT>
static hipError_t stress3d_nt(const Step3D<T> &p, hipStream_t s) {
  if (p.tile_y == 16)
    hipLaunchKernelGGL((k_stress3d<T, 16, NT>), dim3(ntiles3(p, 16)), dim3(BX, 16), 0, s, p);
  else if (p.tile_y == 8)
    hipLaunchKernelGGL((k_stress3d<T, 8, NT>), dim3(ntiles3(p, 8)), dim3(BX, 8), 0, s, p);
  else
    hipLaunchKernelGGL((k_stress3d<T, 4, NT>), dim3(ntiles3(p, 4)), dim3(BX, 4), 0, s, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_stress3d(const Step3D<T> &p, hipStream_t s) {
  return p.streaming ? stress3d_nt<T, true>(p, s) : stress3d_nt<T, false>(p, s);
}

template <typename T, bool NT>
static hipError_t velocity3d_nt(const Step3D<T> &p, hipStream_t s) {
  if (p.tile_y == 16)
    hipLaunchKernelGGL((k_velocity3d<T, 16, NT>), dim3(ntiles3(p, 16)), dim3(BX, 16), 0, s, p);
  else if (p.tile_y == 8)
    hipLaunchKernelGGL((k_velocity3d<T, 8, NT>), dim3(ntiles3(p, 8)), dim3(BX, 8), 0, s, p);
  else
    hipLaunchKernelGGL((k_velocity3d<T, 4, NT>), dim3(ntiles3(p, 4)), dim3(BX, 4), 0, s, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity3d(const Step3D<T> &p, hipStream_t s) {
  return p.streaming ? velocity3d_nt<T, true>(p, s) : velocity3d_nt<T, false>(p, s);
}

template <typename T>
hipError_t launch_record(const Record<T> &p, hipStream_t s) {
  hipLaunchKernelGGL(k_record<T>, dim3(1), dim3(kMaxRecBlock), 0, s, p);
  return hipGetLastError();
}

template <typename T>
int energy_blocks(const Energy<T> &p) {
  const long n = (long)(p.i1 - p.i0 + 1) * (p.j1 - p.j0 + 1) * (p.k1 - p.k0 + 1);
  long b = (n + EB - 1) / EB;
  if (b > 2048) b = 2048;
  return b < 1 ? 1 : (int)b;
}

template <typename T>
hipError_t launch_energy(const Energy<T> &p, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_partial<T>, dim3(nblocks), dim3(EB), 0, s, p);
  hipLaunchKernelGGL(k_energy_final<T>, dim3(1), dim3(EB), 0, s, p, nblocks);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_vmax(const T *vx, const T *vy, const T *vz, int nx, int ny, int nz, long ld,
                       long ps, unsigned long long *out, hipStream_t s) {
  const long n = (long)nx * ny * nz;
  long b = (n + EB - 1) / EB;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(k_vmax<T>, dim3((unsigned)b), dim3(EB), 0, s, vx, vy, vz, nx, ny, nz, ld, ps,
                     out);
  return hipGetLastError();
}

// ------------------------------------------ PNM snapshot (2D2:861-999) --

// max |f| over the interior (as non-negative double bits, for atomicMax)
template <typename T>
__global__ __launch_bounds__(EB) void k_absmax(const T *f, int nx, int ny, long ld,
                                               unsigned long long *out) {
  const long n = (long)nx * ny;
  double m = 0.0;
  for (long t = blockIdx.x * (long)EB + threadIdx.x; t < n; t += (long)gridDim.x * EB) {
    const double d = (double)fabs(f[(t % nx) + (t / nx) * ld]);
    m = d > m ? d : m;
  }
  __shared__ double sm[EB];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = EB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && sm[threadIdx.x + s] > sm[threadIdx.x])
      sm[threadIdx.x] = sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(sm[0]));
}

// One pixel per thread, top row first; the first matching rule wins, then a
// receiver square paints over it.  Arithmetic in T (the single-precision
// build of the reference computes the image in real(4)).  With g.ghost = 1
// the image is the ghost-padded (NX+2) x (NY+2) array of 2D4:729-731: pixel
// ix shows grid point ix - 1 (the zero halo at the border), while the source,
// receivers and PML lines keep their interior indices, as there.  Receivers
// are drawn at their (x, y) whatever their plane (g.ps = plane stride).
template <typename T>
__global__ __launch_bounds__(EB) void k_color_image(const T *f, int gnx, int gny, long ld,
                                                    const unsigned long long *amax_bits,
                                                    Image g, const int *recidx, int nrec,
                                                    unsigned char *rgb) {
  const int nx = gnx + 2 * g.ghost, ny = gny + 2 * g.ghost;  // image size
  const long q = blockIdx.x * (long)EB + threadIdx.x;
  if (q >= (long)nx * ny) return;
  const int ix = (int)(q % nx) + 1, iy = ny - (int)(q / nx);  // Fortran indices
  const T v = f[(ix - 1 - g.ghost) + (long)(iy - 1 - g.ghost) * ld];
  const T amax = (T)__longlong_as_double((long long)*amax_bits);
  T nv = v / amax;
  if (nv < T(-1)) nv = T(-1);
  if (nv > T(1)) nv = T(1);
  int R, G, B;
  const int WC = 5, TC = 1, SQ = 3;  // cross half-width, half-thickness; square half-size
  if ((ix >= g.isrc - WC && ix <= g.isrc + WC && iy >= g.jsrc - TC && iy <= g.jsrc + TC) ||
      (ix >= g.isrc - TC && ix <= g.isrc + TC && iy >= g.jsrc - WC && iy <= g.jsrc + WC)) {
    R = 255, G = 157, B = 0;
  } else if (ix <= 2 || ix >= nx - 1 || iy <= 2 || iy >= ny - 1) {
    R = 0, G = 0, B = 0;
  } else if ((g.use[0] && ix == g.npml) || (g.use[1] && ix == nx - g.npml) ||
             (g.use[2] && iy == g.npml) || (g.use[3] && iy == ny - g.npml)) {
    R = 255, G = 150, B = 0;
  } else if (fabs(v) <= amax * T(0.01)) {
    R = 255, G = 255, B = 255;
  } else {
    const T x = T(255) * pow(fabs(nv), T(0.30));
    const T fl = floor(x);
    const int c = (int)fl + (x - fl >= T(0.5) ? 1 : 0);  // nint, x >= 0
    if (nv >= T(0)) R = c, G = 0, B = 0;
    else R = 0, G = 0, B = c;
  }
  for (int r = 0; r < nrec; ++r) {
    const long rq = recidx[r] % g.ps;  // offset within the receiver's plane
    const int xr = (int)(rq % ld) + 1, yr = (int)(rq / ld) + 1;
    if (ix >= xr - SQ && ix <= xr + SQ && iy >= yr - SQ && iy <= yr + SQ) R = 30, G = 180, B = 60;
  }
  rgb[3 * q] = (unsigned char)R;
  rgb[3 * q + 1] = (unsigned char)G;
  rgb[3 * q + 2] = (unsigned char)B;
}

template <typename T>
hipError_t launch_color_image(const T *f, int nx, int ny, long ld, unsigned long long *amax,
                              const Image &g, const int *recidx, int nrec, unsigned char *rgb,
                              hipStream_t s) {
  const long n = (long)nx * ny;
  const long b = (n + EB - 1) / EB;
  const long bi = ((long)(nx + 2 * g.ghost) * (ny + 2 * g.ghost) + EB - 1) / EB;
  hipLaunchKernelGGL(k_absmax<T>, dim3((unsigned)(b > 2048 ? 2048 : b)), dim3(EB), 0, s, f, nx,
                     ny, ld, amax);
  hipLaunchKernelGGL(k_color_image<T>, dim3((unsigned)bi), dim3(EB), 0, s, f, nx, ny, ld, amax,
                     g, recidx, nrec, rgb);
  return hipGetLastError();
}

#define CPML_INSTANTIATE(T)                                                                     \
  template hipError_t launch_stress2d<T>(int, const Step2D<T> &, hipStream_t);                \
  template hipError_t launch_velocity2d<T>(int, const Step2D<T> &, hipStream_t);              \
  template hipError_t launch_stress3d<T>(const Step3D<T> &, hipStream_t);                     \
  template hipError_t launch_velocity3d<T>(const Step3D<T> &, hipStream_t);                   \
  template hipError_t launch_record<T>(const Record<T> &, hipStream_t);                       \
  template hipError_t launch_energy<T>(const Energy<T> &, int, hipStream_t);                  \
  template int energy_blocks<T>(const Energy<T> &);                                           \
  template hipError_t launch_vmax<T>(const T *, const T *, const T *, int, int, int, long, long, \
                                     unsigned long long *, hipStream_t);                      \
  template hipError_t launch_color_image<T>(const T *, int, int, long, unsigned long long *,    \
                                            const Image &, const int *, int, unsigned char *,  \
                                            hipStream_t);

CPML_INSTANTIATE(float)
CPML_INSTANTIATE(double)

}  // namespace cpml
