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
// A block owns a tile of 64 x TY points (x, y) and marches through a chunk of
// z planes.  For plane k it
//   1. stages v(k) of the tile plus a 2-point x/y border in LDS (one thread
//      per tile point; the border points go to the first and last waves),
//   2. computes sigma(k) on the tile plus a 1-point ring (the ring values are
//      what the velocity update of the tile's edge points needs from the
//      neighbouring tiles; they are recomputed here, not exchanged) and keeps
//      5 of the 6 components in an LDS double buffer,
//   3. computes v(k-1) on the tile from sigma(k-1) (LDS) and the z
//      neighbours sigma(k-2), sigma(k) held in registers.
// Every global load of plane k+1 is issued before the arithmetic of plane k.
// Compulsory traffic per point: read v(3), sigma(6), DELTAT*lambda, DELTAT*mu,
// DELTAT/rho, write sigma(6), v(3) = 21 words, plus the slab memory variables.
#include "cpml_device.h"
#include "cpml_internal.h"

namespace cpml {
namespace {

constexpr int FTX = 64;

// One thread per sigma point of the tile and its 1-point ring (tile points
// first, so tile waves load whole rows), rounded up to whole waves; the v
// points of the 2-point shell go to the last threads.
template <int TY>
struct FG {
  static constexpr int RW = FTX + 2, RH = TY + 2;   // sigma ring: x in [-1,64], y in [-1,TY]
  static constexpr int HW = FTX + 4, HH = TY + 4;   // v region:   x in [-2,65], y in [-2,TY+1]
  static constexpr int NTILE = FTX * TY;            // tile points
  static constexpr int NSIG = RW * RH;              // tile + ring points
  static constexpr int NT = (NSIG + 63) / 64 * 64;  // threads
  static constexpr int NSHELL = HW * HH - RW * RH;  // v points outside the ring
  static_assert(NT <= 1024 && NSHELL <= NT, "block too large");
};

// ring point r -> tile-relative (x, y): row -1, row TY, then columns -1 and 64
__device__ __forceinline__ void ring_xy(int r, int ty, int &x, int &y) {
  const int rw = FTX + 2;
  if (r < rw) {
    x = r - 1;
    y = -1;
  } else if (r < 2 * rw) {
    x = r - rw - 1;
    y = ty;
  } else {
    const int q = r - 2 * rw;
    x = q < ty ? -1 : FTX;
    y = q < ty ? q : q - ty;
  }
}

// shell point s -> (x, y): row -2, row TY+1, then columns -2 and 65 (rows -1..TY)
__device__ __forceinline__ void shell_xy(int s, int ty, int &x, int &y) {
  const int hw = FTX + 4, h = ty + 2;
  if (s < hw) {
    x = s - 2;
    y = -2;
  } else if (s < 2 * hw) {
    x = s - hw - 2;
    y = ty + 1;
  } else {
    const int q = s - 2 * hw;
    x = q < h ? -2 : FTX + 1;
    y = (q < h ? q : q - h) - 1;
  }
}

struct FTile {
  int i0, j0, kb, ke;
};

// XCD-aware tile order (as for the two-phase kernels): each XCD owns one
// contiguous band of tiles, x fastest, then y, then z chunk.
template <int TY>
__device__ __forceinline__ FTile ftile(int nx, int ny, int nz, int kchunk) {
  const int ntx = (nx + FTX - 1) / FTX, nty = (ny + TY - 1) / TY;
  const int ntz = (nz + kchunk - 1) / kchunk;
  const int nt = ntx * nty * ntz;
  const int h = blockIdx.x;
  const int xcd = h & 7, loc = h >> 3;
  const int q = nt >> 3, r = nt & 7;
  const int t = xcd * q + min(xcd, r) + loc;
  const int tx = t % ntx, rest = t / ntx;
  const int ty = rest % nty, tz = zchunk_of(rest / nty, ntz);
  FTile f;
  f.i0 = tx * FTX;
  f.j0 = ty * TY;
  f.kb = tz * kchunk;
  f.ke = min(f.kb + kchunk, nz);
  return f;
}

// One memory-variable step with separate old/new arrays.  `on`: the
// derivative is updated at this point; otherwise the old value is carried
// over (when a new array is given).
template <typename T>
__device__ __forceinline__ T pml(bool on, bool slab, T d, const T *mo, T *mn, const T *a,
                                 const T *b, const T *K, int x) {
  if (!slab) return d;
  if (on) return cpml2(d, mo, mn, a[x], b[x], K[x]);
  if (mn) *mn = *mo;
  return d;
}

// Per sigma point: registers carried along z.
template <typename T>
struct SP {
  T xo1, xo2, yo1, yo2, zo0, zo1;  // vx, vy at k and k+1; vz at k-1 and k
  T so[6], cl, cm;                 // sigma(k) old, DELTAT*lambda, DELTAT*mu at k
  T px2, py2, pz1;                 // prefetch: vx, vy at k+2; vz at k+1
  T pso[6], pcl, pcm;              // prefetch: sigma, coefficients at k+1
};

template <typename T, int TY>
__global__ __launch_bounds__(FG<TY>::NT) void k_fused3d(Fused3D<T> P) {
  using G = FG<TY>;
  __shared__ T vH[3][G::HH * G::HW];
  __shared__ T sR[2][5][G::RH * G::RW];
  const Step3D<T> &p = P.a;
  const FTile ft = ftile<TY>(p.nx, p.ny, p.nz, p.kchunk);
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int tid = threadIdx.x;
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
  auto voff = [&](int x, int y) -> unsigned {
    return (unsigned)(((long)(ft.j0 + y + p.R) * p.ld + p.xpad + ft.i0 + x) * es);
  };
  auto soff = [&](int k) -> unsigned { return (unsigned)(k - ft.kb + 2) * psb; };
  auto LD = [&](__amdgpu_buffer_rsrc_t r, unsigned v, int k) { return BufOp<T>::ld(r, v, soff(k)); };

  // ---- points of this thread: one sigma point (tile points for the first
  // NTILE threads, ring points after them), a shell v point for the last
  // NSHELL threads
  constexpr int NP = 1;
  constexpr int NTP = 1;
  const bool is_tile = tid < G::NTILE;
  int px[NP], py[NP];
  bool has[NP], ld_ok[NP];
  unsigned vo[NP];
  has[0] = tid < G::NSIG;
  if (is_tile) {
    px[0] = tid % FTX;
    py[0] = tid / FTX;
  } else {
    ring_xy(has[0] ? tid - G::NTILE : 0, TY, px[0], py[0]);
  }
  int shx, shy;
  const bool has_sh = tid >= G::NT - G::NSHELL;
  shell_xy(has_sh ? tid - (G::NT - G::NSHELL) : 0, TY, shx, shy);
  // points addressable in the allocation: i <= nx+1 and j <= ny+1
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    ld_ok[a] = has[a] && ft.i0 + px[a] <= nx + 1 && ft.j0 + py[a] <= ny + 1;
    vo[a] = voff(px[a], py[a]);
  }
  const bool sh_ok = has_sh && ft.i0 + shx <= nx + 1 && ft.j0 + shy <= ny + 1;
  const unsigned sho = voff(shx, shy);

  SP<T> S[NP];
  // tile-point history for the velocity phase
  T xo0[NTP], yo0[NTP], cb[NTP], pcb[NTP];
  T sxz1[NTP], sxz2[NTP], syz1[NTP], syz2[NTP], szz1[NTP];
  T shv[3], pshv[3];

  auto prefetch = [&](int k) {  // loads used at iteration k (sigma plane k)
    const KArgs Q = args();
    const auto avx = rs(Q->a.vx), avy = rs(Q->a.vy), avz = rs(Q->a.vz);
    const auto axx = rs(Q->a.sxx), ayy = rs(Q->a.syy), azz = rs(Q->a.szz);
    const auto axy = rs(Q->a.sxy), axz = rs(Q->a.sxz), ayz = rs(Q->a.syz);
    const auto acl = rs(Q->a.clam), acm = rs(Q->a.cmu), acb = rs(Q->a.cbuoy);
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      if (!ld_ok[a]) continue;
      S[a].px2 = LD(avx, vo[a], k + 1);
      S[a].py2 = LD(avy, vo[a], k + 1);
      S[a].pz1 = LD(avz, vo[a], k);
      S[a].pso[0] = LD(axx, vo[a], k);
      S[a].pso[1] = LD(ayy, vo[a], k);
      S[a].pso[2] = LD(azz, vo[a], k);
      S[a].pso[3] = LD(axy, vo[a], k);
      S[a].pso[4] = LD(axz, vo[a], k);
      S[a].pso[5] = LD(ayz, vo[a], k);
      S[a].pcl = LD(acl, vo[a], k);
      S[a].pcm = LD(acm, vo[a], k);
    }
#pragma unroll
    for (int a = 0; a < NTP; ++a)
      if (is_tile && ld_ok[a]) pcb[a] = LD(acb, vo[a], k - 1);
    if (sh_ok) {
      pshv[0] = LD(avx, sho, k);
      pshv[1] = LD(avy, sho, k);
      pshv[2] = LD(avz, sho, k);
    }
  };
  auto take = [&]() {  // prefetched values become current
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      S[a].xo1 = S[a].xo2;
      S[a].xo2 = S[a].px2;
      S[a].yo1 = S[a].yo2;
      S[a].yo2 = S[a].py2;
      S[a].zo0 = S[a].zo1;
      S[a].zo1 = S[a].pz1;
#pragma unroll
      for (int q = 0; q < 6; ++q) S[a].so[q] = S[a].pso[q];
      S[a].cl = S[a].pcl;
      S[a].cm = S[a].pcm;
    }
#pragma unroll
    for (int a = 0; a < NTP; ++a) cb[a] = pcb[a];
#pragma unroll
    for (int q = 0; q < 3; ++q) shv[q] = pshv[q];
  };

  // ---- pipeline fill: registers as they stand before iteration kb-1
  const T zero = T(0);
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    S[a].px2 = S[a].py2 = S[a].pz1 = S[a].pcl = S[a].pcm = zero;
#pragma unroll
    for (int q = 0; q < 6; ++q) S[a].pso[q] = zero;
    S[a].xo1 = S[a].yo1 = S[a].zo0 = zero;
    S[a].xo2 = S[a].yo2 = S[a].zo1 = zero;
    if (ld_ok[a]) {
      S[a].xo2 = LD(rs(p.vx), vo[a], ft.kb - 1);  // becomes xo1 at iteration kb-1
      S[a].yo2 = LD(rs(p.vy), vo[a], ft.kb - 1);
      S[a].zo1 = LD(rs(p.vz), vo[a], ft.kb - 2);  // becomes zo0 at iteration kb-1
    }
  }
#pragma unroll
  for (int a = 0; a < NTP; ++a) {
    pcb[a] = zero;
    xo0[a] = yo0[a] = zero;
    sxz1[a] = sxz2[a] = syz1[a] = syz2[a] = szz1[a] = zero;
  }
  pshv[0] = pshv[1] = pshv[2] = zero;
  prefetch(ft.kb - 1);

  for (int k = ft.kb - 1; k <= ft.ke; ++k) {
    const KArgs Q = args();
    // history for the velocity phase: v(k-1) at the tile points (xo1 still
    // holds plane k-1 until take() shifts the planes)
#pragma unroll
    for (int a = 0; a < NTP; ++a) {
      xo0[a] = S[a].xo1;
      yo0[a] = S[a].yo1;
    }
    take();
    // 1. stage v(k) in LDS
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      if (!has[a]) continue;
      const int h = (py[a] + 2) * G::HW + px[a] + 2;
      vH[0][h] = S[a].xo1;
      vH[1][h] = S[a].yo1;
      vH[2][h] = S[a].zo1;
    }
    if (has_sh) {
      const int h = (shy + 2) * G::HW + shx + 2;
      vH[0][h] = shv[0];
      vH[1][h] = shv[1];
      vH[2][h] = shv[2];
    }
    // loads for the next plane, in flight during this plane's arithmetic
    if (k + 1 <= ft.ke) prefetch(k + 1);
    __syncthreads();

    // 2. sigma(k) on the tile and the ring
    const bool kin = k >= 0 && k < nz;
    const bool inz = kin && p.sz.in(k);
    const bool store_k = k >= ft.kb && k < ft.ke;
    T snew[NTP][6];
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      if (!has[a]) continue;
      const int i = ft.i0 + px[a], j = ft.j0 + py[a];
      // ring points lie one point outside the tile, so also below 0 on the
      // low faces: only points of the grid have memory variables
      const bool inside = kin && i >= 0 && i < nx && j >= 0 && j < ny;
      T sn[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) sn[q] = S[a].so[q];
      if (inside) {
        const int h = (py[a] + 2) * G::HW + px[a] + 2;
        const T vx_c = S[a].xo1, vy_c = S[a].yo1, vz_c = S[a].zo1;
        const bool st = is_tile && store_k;
        const bool inx = p.sx.in(i), iny = p.sy.in(j);
        const unsigned ox = p.sx.s(i) + (unsigned)p.sx.w * (j + (unsigned)ny * k);
        const unsigned oy = i + (unsigned)p.mld * (p.sy.s(j) + (unsigned)p.sy.w * k);
        const unsigned oz = i + (unsigned)p.mld * (j + (unsigned)ny * p.sz.s(k));
        const T cl = S[a].cl, cm = S[a].cm;
        // sigma_xx, sigma_yy, sigma_zz: 3D:836-863
        const bool r_n = i <= nx - 2 && j >= 1 && k >= 1;
        {
          T dvx_dx = (vH[0][h + 1] - vx_c) * p.odx;
          T dvy_dy = (vy_c - vH[1][h - G::HW]) * p.ody;
          T dvz_dz = (vz_c - S[a].zo0) * p.odz;
          dvx_dx = pml(r_n, inx, dvx_dx, Q->a.mx[0] + ox, st ? Q->mx[0] + ox : nullptr,
                       Q->a.px.ah, Q->a.px.bh, Q->a.px.Kh, i);
          dvy_dy = pml(r_n, iny, dvy_dy, Q->a.my[0] + oy, st ? Q->my[0] + oy : nullptr, Q->a.py.a, Q->a.py.b, Q->a.py.K, j);
          dvz_dz = pml(r_n, inz, dvz_dz, Q->a.mz[0] + oz, st ? Q->mz[0] + oz : nullptr, Q->a.pz.a, Q->a.pz.b, Q->a.pz.K, k);
          if (r_n) {
            const T lp2m = cl + T(2) * cm;
            sn[0] = lp2m * dvx_dx + cl * (dvy_dy + dvz_dz) + S[a].so[0];
            sn[1] = cl * (dvx_dx + dvz_dz) + lp2m * dvy_dy + S[a].so[1];
            sn[2] = cl * (dvx_dx + dvy_dy) + lp2m * dvz_dz + S[a].so[2];
          }
        }
        // sigma_xy: 3D:877-894
        const bool r_xy = i >= 1 && j <= ny - 2;
        {
          T dvy_dx = (vy_c - vH[1][h - 1]) * p.odx;
          T dvx_dy = (vH[0][h + G::HW] - vx_c) * p.ody;
          dvy_dx = pml(r_xy, inx, dvy_dx, Q->a.mx[1] + ox, st ? Q->mx[1] + ox : nullptr, Q->a.px.a, Q->a.px.b, Q->a.px.K, i);
          dvx_dy = pml(r_xy, iny, dvx_dy, Q->a.my[1] + oy, st ? Q->my[1] + oy : nullptr,
                       Q->a.py.ah, Q->a.py.bh, Q->a.py.Kh, j);
          if (r_xy) sn[3] = cm * (dvy_dx + dvx_dy) + S[a].so[3];
        }
        // sigma_xz: 3D:908-925
        const bool r_xz = i >= 1 && k <= nz - 2;
        {
          T dvz_dx = (vz_c - vH[2][h - 1]) * p.odx;
          T dvx_dz = (S[a].xo2 - vx_c) * p.odz;
          dvz_dx = pml(r_xz, inx, dvz_dx, Q->a.mx[2] + ox, st ? Q->mx[2] + ox : nullptr, Q->a.px.a, Q->a.px.b, Q->a.px.K, i);
          dvx_dz = pml(r_xz, inz, dvx_dz, Q->a.mz[1] + oz, st ? Q->mz[1] + oz : nullptr,
                       Q->a.pz.ah, Q->a.pz.bh, Q->a.pz.Kh, k);
          if (r_xz) sn[4] = cm * (dvz_dx + dvx_dz) + S[a].so[4];
        }
        // sigma_yz: 3D:927-943
        const bool r_yz = j <= ny - 2 && k <= nz - 2;
        {
          T dvz_dy = (vH[2][h + G::HW] - vz_c) * p.ody;
          T dvy_dz = (S[a].yo2 - vy_c) * p.odz;
          dvz_dy = pml(r_yz, iny, dvz_dy, Q->a.my[2] + oy, st ? Q->my[2] + oy : nullptr,
                       Q->a.py.ah, Q->a.py.bh, Q->a.py.Kh, j);
          dvy_dz = pml(r_yz, inz, dvy_dz, Q->a.mz[2] + oz, st ? Q->mz[2] + oz : nullptr,
                       Q->a.pz.ah, Q->a.pz.bh, Q->a.pz.Kh, k);
          if (r_yz) sn[5] = cm * (dvz_dy + dvy_dz) + S[a].so[5];
        }
        if (st) {
          const unsigned so_k = soff(k);
          BufOp<T>::st(sn[0], rs(Q->sxx), vo[a], so_k);
          BufOp<T>::st(sn[1], rs(Q->syy), vo[a], so_k);
          BufOp<T>::st(sn[2], rs(Q->szz), vo[a], so_k);
          BufOp<T>::st(sn[3], rs(Q->sxy), vo[a], so_k);
          BufOp<T>::st(sn[4], rs(Q->sxz), vo[a], so_k);
          BufOp<T>::st(sn[5], rs(Q->syz), vo[a], so_k);
        }
      }
      const int r = (py[a] + 1) * G::RW + px[a] + 1;
      sR[k & 1][0][r] = sn[0];
      sR[k & 1][1][r] = sn[1];
      sR[k & 1][2][r] = sn[3];
      sR[k & 1][3][r] = sn[4];
      sR[k & 1][4][r] = sn[5];
#pragma unroll
      for (int q = 0; q < 6; ++q) snew[a][q] = sn[q];
    }
    __syncthreads();

    // 3. v(k-1) on the tile
    const int kv = k - 1;
    if (is_tile && kv >= ft.kb && kv < ft.ke) {
      const int it = *p.step;
      const bool inzv = p.sz.in(kv);
#pragma unroll
      for (int a = 0; a < NTP; ++a) {
        const int i = ft.i0 + px[a], j = ft.j0 + py[a];
        if (i >= nx || j >= ny) continue;
        const int r = (py[a] + 1) * G::RW + px[a] + 1;
        const T *s0 = sR[kv & 1][0], *s1 = sR[kv & 1][1], *s2 = sR[kv & 1][2];
        const T *s3 = sR[kv & 1][3], *s4 = sR[kv & 1][4];
        const bool inx = p.sx.in(i), iny = p.sy.in(j);
        const unsigned ox = p.sx.s(i) + (unsigned)p.sx.w * (j + (unsigned)ny * kv);
        const unsigned oy = i + (unsigned)p.mld * (p.sy.s(j) + (unsigned)p.sy.w * kv);
        const unsigned oz = i + (unsigned)p.mld * (j + (unsigned)ny * p.sz.s(kv));
        const T c = cb[a];
        T vxn = xo0[a], vyn = yo0[a], vzn = S[a].zo0;  // v(k-1)
        // vx: 3D:976-993
        const bool r_vx = i >= 1 && j >= 1 && kv >= 1;
        {
          T dsxx_dx = (s0[r] - s0[r - 1]) * p.odx;
          T dsxy_dy = (s2[r] - s2[r - G::RW]) * p.ody;
          T dsxz_dz = (sxz1[a] - sxz2[a]) * p.odz;
          dsxx_dx = pml(r_vx, inx, dsxx_dx, Q->a.mx[3] + ox, Q->mx[3] + ox, Q->a.px.a, Q->a.px.b, Q->a.px.K, i);
          dsxy_dy = pml(r_vx, iny, dsxy_dy, Q->a.my[3] + oy, Q->my[3] + oy, Q->a.py.a, Q->a.py.b, Q->a.py.K, j);
          dsxz_dz = pml(r_vx, inzv, dsxz_dz, Q->a.mz[3] + oz, Q->mz[3] + oz, Q->a.pz.a, Q->a.pz.b, Q->a.pz.K, kv);
          if (r_vx) vxn = c * (dsxx_dx + dsxy_dy + dsxz_dz) + vxn;
        }
        // vy: 3D:995-1017
        const bool r_vy = i <= nx - 2 && j <= ny - 2 && kv >= 1;
        {
          T dsxy_dx = (s2[r + 1] - s2[r]) * p.odx;
          T dsyy_dy = (s1[r + G::RW] - s1[r]) * p.ody;
          T dsyz_dz = (syz1[a] - syz2[a]) * p.odz;
          dsxy_dx = pml(r_vy, inx, dsxy_dx, Q->a.mx[4] + ox, Q->mx[4] + ox, Q->a.px.ah, Q->a.px.bh, Q->a.px.Kh, i);
          dsyy_dy = pml(r_vy, iny, dsyy_dy, Q->a.my[4] + oy, Q->my[4] + oy, Q->a.py.ah, Q->a.py.bh, Q->a.py.Kh, j);
          dsyz_dz = pml(r_vy, inzv, dsyz_dz, Q->a.mz[4] + oz, Q->mz[4] + oz, Q->a.pz.a, Q->a.pz.b, Q->a.pz.K, kv);
          if (r_vy) vyn = c * (dsxy_dx + dsyy_dy + dsyz_dz) + vyn;
        }
        // vz: 3D:1031-1052
        const bool r_vz = i <= nx - 2 && j >= 1 && kv <= nz - 2;
        {
          T dsxz_dx = (s3[r + 1] - s3[r]) * p.odx;
          T dsyz_dy = (s4[r] - s4[r - G::RW]) * p.ody;
          T dszz_dz = (snew[a][2] - szz1[a]) * p.odz;
          dsxz_dx = pml(r_vz, inx, dsxz_dx, Q->a.mx[5] + ox, Q->mx[5] + ox, Q->a.px.ah, Q->a.px.bh, Q->a.px.Kh, i);
          dsyz_dy = pml(r_vz, iny, dsyz_dy, Q->a.my[5] + oy, Q->my[5] + oy, Q->a.py.a, Q->a.py.b, Q->a.py.K, j);
          dszz_dz = pml(r_vz, inzv, dszz_dz, Q->a.mz[5] + oz, Q->mz[5] + oz, Q->a.pz.ah, Q->a.pz.bh, Q->a.pz.Kh, kv);
          if (r_vz) vzn = c * (dsxz_dx + dsyz_dy + dszz_dz) + vzn;
        }
        // point force (3D:1074-1082) and rigid faces (3D:1087-1121)
        if (i == p.isrc && j == p.jsrc && kv == p.ksrc) {
          const T rsrc = p.rho[i + (long)j * p.ld + (long)kv * p.ps];
          vxn = vxn + p.fx[it - 1] * p.dt / rsrc;
          vyn = vyn + p.fy[it - 1] * p.dt / rsrc;
        }
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || kv == 0 || kv == nz - 1) {
          vxn = T(0);
          vyn = T(0);
          vzn = T(0);
        }
        const unsigned so_v = soff(kv);
        BufOp<T>::st(vxn, rs(Q->vx), vo[a], so_v);
        BufOp<T>::st(vyn, rs(Q->vy), vo[a], so_v);
        BufOp<T>::st(vzn, rs(Q->vz), vo[a], so_v);
      }
    }
    // sigma history at the tile points
#pragma unroll
    for (int a = 0; a < NTP; ++a) {
      sxz2[a] = sxz1[a];
      sxz1[a] = snew[a][4];
      syz2[a] = syz1[a];
      syz1[a] = snew[a][5];
      szz1[a] = snew[a][2];
    }
  }
}

}  // namespace

template <typename T>
unsigned fused_tiles3d(const Fused3D<T> &P, int ty) {
  const Step3D<T> &p = P.a;
  return (unsigned)((p.nx + FTX - 1) / FTX) * ((p.ny + ty - 1) / ty) *
         ((p.nz + p.kchunk - 1) / p.kchunk);
}

// tile rows 4 (448 threads), 8 (704) or 12 (960)
template <typename T>
hipError_t launch_fused3d(const Fused3D<T> &P, hipStream_t s) {
  const int ty = P.a.tile_y;
  if (ty >= 12)
    hipLaunchKernelGGL((k_fused3d<T, 12>), dim3(fused_tiles3d(P, 12)), dim3(FG<12>::NT), 0, s,
                       P);
  else if (ty >= 8)
    hipLaunchKernelGGL((k_fused3d<T, 8>), dim3(fused_tiles3d(P, 8)), dim3(FG<8>::NT), 0, s, P);
  else
    hipLaunchKernelGGL((k_fused3d<T, 4>), dim3(fused_tiles3d(P, 4)), dim3(FG<4>::NT), 0, s, P);
  return hipGetLastError();
}

template hipError_t launch_fused3d<float>(const Fused3D<float> &, hipStream_t);
template hipError_t launch_fused3d<double>(const Fused3D<double> &, hipStream_t);

}  // namespace cpml
