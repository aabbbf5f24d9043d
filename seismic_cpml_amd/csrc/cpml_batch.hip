// Batched multi-shot 2D C-PML: many independent shots (each its own grid
// size, medium, profiles, source and receivers) advanced together.
//
// Two ways to step all shots:
//  * single steps, one launch per phase for all shots: every shot is
//    described by the same Step2D block the single-grid kernels use, stored in
//    a device table; the tiles of all shots form one flat list, split into 8
//    contiguous bands, one per XCD; a launch has one workgroup per tile or a
//    fixed "persistent" grid whose workgroups stride through their band;
//  * time-blocked: NT leapfrog steps per launch for all shots (the kernel of
//    cpml_tb2d.hip over a flat list of (shot, strip, chunk) items, TB2D
//    descriptors per shot, persistent or one wavefront per item), the state
//    ping-ponging between two pools; the items that touch a C-PML slab run
//    beside the others on a second stream.
// Per point the arithmetic is the single-grid kernels', so a batched shot
// equals the same shot run on its own, bit for bit.  Each shot has its own
// step counter, energy curve (per-tile partial sums, then a per-shot sum in
// a fixed order) and stability check (per-shot max |v|, reported with the
// shot's index); chunks of steps are captured once per state buffer in a HIP
// graph.
//
// The autotuner times the launch shapes (steps per launch, chunk rows,
// persistent grid size, rows per single-step tile) on the real batch for a
// few steps, keeps the fastest, returns the batch to rest and caches the
// choice in a small text file keyed by device, order, precision and batch
// shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpml_gpu.h"
#include "cpml_device.h"
#include "cpml_energy.h"
#include "cpml_internal.h"
#include "cpml_points2d.h"

namespace cpml {
namespace {

// flat tile index -> (shot, local tile) and the XCD band of this workgroup
struct BandIter {
  int start, count, nb, loc;
  __device__ BandIter(int ntiles) {
    const int G = gridDim.x, h = blockIdx.x;
    const int xcd = h & 7;
    loc = h >> 3;
    nb = (G - xcd + 7) >> 3;
    const int q = ntiles >> 3, r = ntiles & 7;
    start = xcd * q + min(xcd, r);
    count = q + (xcd < r ? 1 : 0);
  }
};

// A tile is 64 points by TY*RG rows; the block (64 x TY threads) walks its RG
// row groups one after another, so the shot's descriptor is read once per
// TY*RG rows.
template <typename T, int ORDER, int TY, int RG, bool STRESS, bool NT>
__global__ __launch_bounds__(BX *TY) void k_batch2d(const Step2D<T> *__restrict__ shots,
                                                   const int2 *__restrict__ ttab, int ntiles) {
  const BandIter b(ntiles);
  for (int u = b.loc; u < b.count; u += b.nb) {
    const int t = b.start + u;
    const int2 e = ttab[t];  // shot, tile within the shot
    const Step2D<T> &p = shots[e.x];
    const int lt = e.y;
    const int ntx = (p.nx + BX - 1) / BX;
    const int i = (lt % ntx) * BX + threadIdx.x;
    const int j0 = (lt / ntx) * (TY * RG) + threadIdx.y;
#pragma unroll
    for (int r = 0; r < RG; ++r) {
      const int j = j0 + r * TY;
      if (i < p.nx && j < p.ny) {
        if (STRESS) stress2d_point<T, ORDER, NT>(p, i, j);
        else velocity2d_point<T, ORDER, NT>(p, i, j);
      }
    }
  }
}

// receivers of shot blockIdx.x for the state of step it + skip (the launch
// before advanced skip + 1 steps), then that shot's step counter advances
template <typename T>
__global__ __launch_bounds__(kMaxRecBlock) void k_batch_record(const Record<T> *__restrict__ r,
                                                               int skip) {
  const Record<T> &p = r[blockIdx.x];
  const int it = *p.step + skip;
  const long row = (long)(it - 1) * p.nrec;
  for (int q = threadIdx.x; q < p.nrec; q += blockDim.x) {
    const long o = p.idx[q];
    p.sx[row + q] = p.vx[o];
    p.sy[row + q] = p.vy[o];
  }
  __syncthreads();
  if (threadIdx.x == 0) *p.step = it + 1;
}

constexpr int BE = 256;  // threads of the energy / max-norm blocks

__device__ void sum2_256(double &a, double &b) {
  __shared__ double sa[BE], sb[BE];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = BE / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  a = sa[0];
  b = sb[0];
}

// energy of one 64 x 16 tile of a shot's energy window (2D2:694-713) for the
// state of step it + skip when that step is on the cadence
template <typename T>
__global__ __launch_bounds__(BE) void k_batch_energy_partial(const Energy<T> *__restrict__ en,
                                                             const int2 *__restrict__ etab,
                                                             int skip, double *partial) {
  const int2 e = etab[blockIdx.x];
  const Energy<T> &p = en[e.x];
  const int it = *p.step + skip;
  if (it % p.cadence != 0) return;
  const int wx = p.i1 - p.i0 + 1;
  const int ntx = (wx + 63) / 64;
  const int i = p.i0 + (e.y % ntx) * 64 + (threadIdx.x & 63);
  const int jt = p.j0 + (e.y / ntx) * 16;
  double kin = 0.0, pot = 0.0;
  for (int r = threadIdx.x >> 6; r < 16; r += BE / 64) {
    const int j = jt + r;
    if (i <= p.i1 && j <= p.j1) cell_energy(p, i + (long)j * p.ld, kin, pot);
  }
  sum2_256(kin, pot);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = kin;
    partial[2 * blockIdx.x + 1] = pot;
  }
}

// per shot (block): the sum of its tiles' partials, in a fixed order
template <typename T>
__global__ __launch_bounds__(BE) void k_batch_energy_final(const Energy<T> *__restrict__ en,
                                                           const int *__restrict__ estart,
                                                           int skip, const double *partial) {
  const Energy<T> &p = en[blockIdx.x];
  const int it = *p.step + skip;
  if (it % p.cadence != 0) return;
  double kin = 0.0, pot = 0.0;
  for (int t = estart[blockIdx.x] + threadIdx.x; t < estart[blockIdx.x + 1]; t += BE) {
    kin += partial[2 * t];
    pot += partial[2 * t + 1];
  }
  sum2_256(kin, pot);
  if (threadIdx.x == 0) {
    p.out_kin[it - 1] = 0.5 * kin;  // 2D2:705 (kinetic summed, then halved)
    p.out_pot[it - 1] = pot;
  }
}

// max over a tile (64 x rows points) of vx^2 + vy^2, into its shot's slot
// (as non-negative double bits; NaN wins so that a blow-up is reported)
template <typename T>
__global__ __launch_bounds__(BE) void k_batch_vmax(const Step2D<T> *__restrict__ shots,
                                                   const int2 *__restrict__ ttab, int rows,
                                                   unsigned long long *vmax) {
  const int2 e = ttab[blockIdx.x];
  const Step2D<T> &p = shots[e.x];
  const int ntx = (p.nx + BX - 1) / BX;
  const int i = (e.y % ntx) * BX + (threadIdx.x & 63);
  const int jt = (e.y / ntx) * rows;
  double m = 0.0;
  for (int r = threadIdx.x >> 6; r < rows; r += BE / 64) {
    const int j = jt + r;
    if (i < p.nx && j < p.ny) {
      const long q = i + (long)j * p.ld;
      const T s = p.vx[q] * p.vx[q] + p.vy[q] * p.vy[q];
      const double d = (double)s;
      m = (d > m || d != d) ? d : m;
    }
  }
  __shared__ double sm[BE];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = BE / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sm[threadIdx.x + s];
      if (o > sm[threadIdx.x] || o != o) sm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = sm[0];
    atomicMax(vmax + e.x, v != v ? 0x7ff8000000000000ull : (unsigned long long)__double_as_longlong(v));
  }
}

template <typename T, int TY, int RG>
hipError_t batch_phase_ty(int order, bool stress, bool streaming, const Step2D<T> *shots,
                          const int2 *ttab, int ntiles, int grid, hipStream_t s) {
  const dim3 g(grid), blk(BX, TY);
#define CPML_B2(O, NT)                                                                       \
  do {                                                                                       \
    if (stress)                                                                              \
      hipLaunchKernelGGL((k_batch2d<T, O, TY, RG, true, NT>), g, blk, 0, s, shots, ttab,    \
                         ntiles);                                                            \
    else                                                                                     \
      hipLaunchKernelGGL((k_batch2d<T, O, TY, RG, false, NT>), g, blk, 0, s, shots, ttab,   \
                         ntiles);                                                            \
  } while (0)
#define CPML_B(O)              \
  do {                         \
    if (streaming)             \
      CPML_B2(O, true);        \
    else                       \
      CPML_B2(O, false);       \
  } while (0)
  if (order == 2) CPML_B(2);
  else if (order == 4) CPML_B(4);
  else CPML_B(8);
#undef CPML_B
#undef CPML_B2
  return hipGetLastError();
}

// rows per tile: 4, 8, 16 (one row group) or 32, 64 (2 or 4 groups of 16)
template <typename T>
hipError_t batch_phase(int order, int rows, bool stress, bool streaming, const void *shots,
                       const int2 *ttab, int ntiles, int grid, hipStream_t s) {
  const auto *sh = static_cast<const Step2D<T> *>(shots);
  switch (rows) {
    case 64: return batch_phase_ty<T, 16, 4>(order, stress, streaming, sh, ttab, ntiles, grid, s);
    case 32: return batch_phase_ty<T, 16, 2>(order, stress, streaming, sh, ttab, ntiles, grid, s);
    case 16: return batch_phase_ty<T, 16, 1>(order, stress, streaming, sh, ttab, ntiles, grid, s);
    case 8: return batch_phase_ty<T, 8, 1>(order, stress, streaming, sh, ttab, ntiles, grid, s);
    default: return batch_phase_ty<T, 4, 1>(order, stress, streaming, sh, ttab, ntiles, grid, s);
  }
}

template <typename T>
hipError_t batch_record(const void *rec, int nshots, int skip, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_record<T>, dim3(nshots), dim3(kMaxRecBlock), 0, s,
                     static_cast<const Record<T> *>(rec), skip);
  return hipGetLastError();
}

template <typename T>
hipError_t batch_energy(const void *en, const int2 *etab, int netiles, const int *estart,
                        int nshots, int skip, double *partial, hipStream_t s) {
  const auto *e = static_cast<const Energy<T> *>(en);
  if (netiles > 0)
    hipLaunchKernelGGL(k_batch_energy_partial<T>, dim3(netiles), dim3(BE), 0, s, e, etab, skip,
                       partial);
  hipLaunchKernelGGL(k_batch_energy_final<T>, dim3(nshots), dim3(BE), 0, s, e, estart, skip,
                     partial);
  return hipGetLastError();
}

template <typename T>
hipError_t batch_vmax(const void *shots, const int2 *ttab, int ntiles, int rows,
                      unsigned long long *vmax, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_vmax<T>, dim3(ntiles), dim3(BE), 0, s,
                     static_cast<const Step2D<T> *>(shots), ttab, rows, vmax);
  return hipGetLastError();
}

long round_up(long v, long m) { return (v + m - 1) / m * m; }

}  // namespace
}  // namespace cpml

using cpml::Slab;

namespace {

thread_local std::string g_batch_init_error;

struct Shot {
  int nx = 0, ny = 0, npml = 0;
  double dx = 0, dy = 0, dt = 0;
  long ld = 0, ps = 0, base = 0, total = 0, mld = 0;
  size_t off[8] = {};  // pool offsets: vx vy sxx syy sxy lambda mu rho
  Slab sx, sy;
  void *prof = nullptr;  // [2 axes][6][nmax]
  void *mem[8] = {};     // x: dvx_dx dvy_dx dsxx_dx dsxy_dx ; y: dvy_dy dvx_dy dsxy_dy dsyy_dy
  void *mem2[8] = {};    // second buffer of the time-blocked steps
  void *fx = nullptr, *fy = nullptr;
  int isrc = -1, jsrc = -1;
  int nrec = 0;
  int *recidx = nullptr;
  std::vector<int> rix, riy;  // receiver indices (0-based), host copy
  int *tbrec = nullptr;       // receivers per strip of the time-blocked kernel
  int tbrec_key = -1;         // strip width they were built for
  void *seis[2] = {};
  bool have_medium = false, have_pml = false;
};

}  // namespace

struct cpml_batch {
  int order = 2, prec = 4, nshots = 0, nstep = 0, R = 1;
  size_t esize = 4;
  int device = 0, ncu = 256;
  hipStream_t stream = nullptr;
  std::vector<Shot> shots;
  std::vector<void *> allocs;
  double bytes = 0;
  char *pool = nullptr;
  void *d_desc = nullptr;   // Step2D<T>[nshots]
  void *d_rec = nullptr;    // Record<T>[nshots]
  int2 *d_ttab = nullptr;   // [tiles]: shot, tile within the shot
  int ttab_cap = 0;         // entries allocated (the most tiles any tile height gives)
  int *d_steps = nullptr;   // [nshots]
  int ntiles = 0;
  int tile_y = 4, blocks_per_cu = 0, graph_steps = 16;
  bool dirty = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // per state buffer the replay starts from
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  std::vector<int> host_steps;
  // ping-pong state of the time-blocked steps
  int cur = 0;               // 1: the fields live in pool2 and the shots' mem2
  char *pool2 = nullptr;     // same layout as pool (its medium slots unused)
  void *d_desc2 = nullptr;   // Step2D<T>[nshots] over pool2
  void *d_rec2 = nullptr;    // Record<T>[nshots] over pool2
  void *d_tbd[2] = {};       // TB2D<T>[nshots]: state in pool (0) / pool2 (1)
  int4 *d_items = nullptr;   // time-blocked items: the slab-free ones, then the others
  // time-blocked items by class: slab-free, x slabs only, y slabs only, both
  int items_cap = 0, nitems[4] = {0, 0, 0, 0};
  int tb = 0, tb_ow = 0, tb_jc = 128;
  int tb_jc_slab = 0;  // rows per slab item (0: tb_jc), set with tb_jc
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // energy (per shot, every `cadence` steps) and the stability check
  int cadence = 0;
  double *d_ekin = nullptr, *d_epot = nullptr;  // [nshots][nstep]
  void *d_en[2] = {};        // Energy<T>[nshots] per state buffer
  int2 *d_etab = nullptr;    // energy tiles (shot, tile of 64 x 16 of its window)
  int *d_estart = nullptr;   // [nshots + 1]
  double *d_epart = nullptr; // [tiles][2]
  int netiles = 0;
  int it_display = 0;
  double threshold = 1e25;
  unsigned long long *d_vmax = nullptr;  // [nshots]
  std::string err;
};

namespace {

int bfail(cpml_batch *b, int code, const std::string &m) {
  if (b)
    b->err = m;
  else
    g_batch_init_error = m;
  return code;
}

#define BCHK(b, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return bfail((b), CPML_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int balloc(cpml_batch *b, void **p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return bfail(b, e == hipErrorOutOfMemory ? CPML_ERR_OOM : CPML_ERR_HIP,
                 std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  b->allocs.push_back(*p);
  b->bytes += (double)bytes;
  // cleared and waited for (the batch's stream does not synchronise with the
  // null stream, on which blocking copies into the buffer may follow)
  BCHK(b, hipMemsetAsync(*p, 0, bytes, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return CPML_OK;
}

void bfree(cpml_batch *b, void *&p) {
  if (!p) return;
  auto it = std::find(b->allocs.begin(), b->allocs.end(), p);
  if (it != b->allocs.end()) b->allocs.erase(it);
  (void)hipFree(p);
  p = nullptr;
}

int check_shot(cpml_batch *b, int shot) {
  if (!b) return CPML_ERR_ARG;
  if (shot < 0 || shot >= b->nshots) return bfail(b, CPML_ERR_ARG, "shot index out of range");
  BCHK(b, hipSetDevice(b->device));
  return CPML_OK;
}

template <typename T>
cpml::Step2D<T> shot_step(cpml_batch *b, int q, int parity) {
  const Shot &s = b->shots[q];
  cpml::Step2D<T> p{};
  T *f[8];
  for (int k = 0; k < 8; ++k)
    f[k] = reinterpret_cast<T *>((parity && k < 5 ? b->pool2 : b->pool) + s.off[k]) + s.base;
  p.vx = f[0];
  p.vy = f[1];
  p.sxx = f[2];
  p.syy = f[3];
  p.sxy = f[4];
  p.lam = f[5];
  p.mu = f[6];
  p.rho = f[7];
  void *const *mem = parity ? s.mem2 : s.mem;
  for (int k = 0; k < 4; ++k) {
    p.mx[k] = static_cast<T *>(mem[k]);
    p.my[k] = static_cast<T *>(mem[4 + k]);
  }
  const long nmax = std::max(s.nx, s.ny);
  const T *pr = static_cast<const T *>(s.prof);
  p.px = cpml::Prof<T>{pr, pr + nmax, pr + 2 * nmax, pr + 3 * nmax, pr + 4 * nmax, pr + 5 * nmax};
  pr += 6 * nmax;
  p.py = cpml::Prof<T>{pr, pr + nmax, pr + 2 * nmax, pr + 3 * nmax, pr + 4 * nmax, pr + 5 * nmax};
  p.sx = s.sx;
  p.sy = s.sy;
  p.nx = s.nx;
  p.ny = s.ny;
  p.ld = s.ld;
  p.mld = s.mld;
  p.den_x = b->order == 4 ? T(24) * T(s.dx) : T(s.dx);
  p.den_y = b->order == 4 ? T(24) * T(s.dy) : T(s.dy);
  p.dt = T(s.dt);
  p.isrc = s.isrc;
  p.jsrc = s.jsrc;
  p.fx = static_cast<const T *>(s.fx);
  p.fy = static_cast<const T *>(s.fy);
  p.step = b->d_steps + q;
  p.tile_y = b->tile_y;
  p.streaming = b->bytes > cpml::kStreamingBytes;
  return p;
}

// Per-strip receiver lists of one shot for the time-blocked kernel (entries
// sorted by strip, then row): start[nstrips + 1], i, j, slot.
int build_shot_receivers(cpml_batch *b, Shot &s) {
  const int key = b->tb_ow;
  if (s.tbrec && s.tbrec_key == key) return CPML_OK;
  const int ns = (s.nx + b->tb_ow - 1) / b->tb_ow, n = s.nrec;
  std::vector<int> order(n);
  for (int r = 0; r < n; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const int sa = s.rix[a] / b->tb_ow, sc = s.rix[c] / b->tb_ow;
    return sa != sc ? sa < sc : s.riy[a] < s.riy[c];
  });
  const size_t m = (size_t)std::max(n, 1);
  std::vector<int> buf((size_t)ns + 1 + 3 * m, 0);
  int *start = buf.data(), *ri = start + ns + 1, *rj = ri + m, *rs = rj + m;
  for (int k = 0; k < n; ++k) {
    const int r = order[k];
    ri[k] = s.rix[r];
    rj[k] = s.riy[r];
    rs[k] = r;
    start[s.rix[r] / b->tb_ow + 1]++;
  }
  for (int t = 0; t < ns; ++t) start[t + 1] += start[t];
  bfree(b, *(void **)&s.tbrec);
  int rc = balloc(b, (void **)&s.tbrec, buf.size() * sizeof(int));
  if (rc) return rc;
  BCHK(b, hipMemcpy(s.tbrec, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice));
  s.tbrec_key = key;
  return CPML_OK;
}

template <typename T>
int upload_tables(cpml_batch *b) {
  std::vector<int2> ttab;
  int nt = 0;
  for (int q = 0; q < b->nshots; ++q) {
    const Shot &s = b->shots[q];
    if (!s.have_medium || !s.have_pml)
      return bfail(b, CPML_ERR_STATE, "shot " + std::to_string(q) +
                                          ": set_medium and set_pml must be called before run");
    const int tq = (int)(((s.nx + cpml::BX - 1) / cpml::BX) * ((s.ny + b->tile_y - 1) / b->tile_y));
    for (int t = 0; t < tq; ++t) ttab.push_back(make_int2(q, t));
    nt += tq;
  }
  b->ntiles = nt;
  if (nt > b->ttab_cap) return bfail(b, CPML_ERR_STATE, "tile table too small");
  BCHK(b, hipMemcpy(b->d_ttab, ttab.data(), ttab.size() * sizeof(int2), hipMemcpyHostToDevice));
  const int npar = b->pool2 ? 2 : 1;
  for (int par = 0; par < npar; ++par) {
    std::vector<cpml::Step2D<T>> desc(b->nshots);
    std::vector<cpml::Record<T>> rec(b->nshots);
    for (int q = 0; q < b->nshots; ++q) {
      const Shot &s = b->shots[q];
      desc[q] = shot_step<T>(b, q, par);
      cpml::Record<T> r{};
      r.vx = desc[q].vx;
      r.vy = desc[q].vy;
      r.ld = s.ld;
      r.nrec = s.nrec;
      r.idx = s.recidx;
      r.sx = static_cast<T *>(s.seis[0]);
      r.sy = static_cast<T *>(s.seis[1]);
      r.step = b->d_steps + q;
      rec[q] = r;
    }
    BCHK(b, hipMemcpy(par ? b->d_desc2 : b->d_desc, desc.data(), desc.size() * sizeof(desc[0]),
                      hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(par ? b->d_rec2 : b->d_rec, rec.data(), rec.size() * sizeof(rec[0]),
                      hipMemcpyHostToDevice));
    if (b->cadence > 0) {
      std::vector<cpml::Energy<T>> en(b->nshots);
      for (int q = 0; q < b->nshots; ++q) {
        const Shot &s = b->shots[q];
        cpml::Energy<T> e{};
        e.vx = desc[q].vx;
        e.vy = desc[q].vy;
        e.sxx = desc[q].sxx;
        e.syy = desc[q].syy;
        e.sxy = desc[q].sxy;
        e.lam = desc[q].lam;
        e.mu = desc[q].mu;
        e.rho = desc[q].rho;
        e.ld = s.ld;
        e.ps = s.ps;
        // 2D2:695 (NPML+1 .. N-NPML) / 2D4:696 (NPML .. N-NPML+1), 0-based
        const int lo = b->order == 4 ? s.npml - 1 : s.npml;
        e.i0 = std::max(lo, 0);
        e.i1 = std::min(b->order == 4 ? s.nx - s.npml : s.nx - s.npml - 1, s.nx - 1);
        e.j0 = std::max(lo, 0);
        e.j1 = std::min(b->order == 4 ? s.ny - s.npml : s.ny - s.npml - 1, s.ny - 1);
        e.ndim = 2;
        e.cadence = b->cadence;
        e.out_kin = b->d_ekin + (size_t)q * b->nstep;
        e.out_pot = b->d_epot + (size_t)q * b->nstep;
        e.step = b->d_steps + q;
        en[q] = e;
      }
      BCHK(b, hipMemcpy(b->d_en[par], en.data(), en.size() * sizeof(en[0]),
                        hipMemcpyHostToDevice));
    }
    if (b->tb > 0) {
      std::vector<cpml::TB2D<T>> tbd(b->nshots);
      for (int q = 0; q < b->nshots; ++q) {
        Shot &s = b->shots[q];
        int rc = build_shot_receivers(b, s);
        if (rc) return rc;
        cpml::TB2D<T> P{};
        P.a = desc[q];
        const cpml::Step2D<T> o = shot_step<T>(b, q, par ^ 1);
        P.vx = o.vx;
        P.vy = o.vy;
        P.sxx = o.sxx;
        P.syy = o.syy;
        P.sxy = o.sxy;
        for (int k = 0; k < 4; ++k) {
          P.mx[k] = o.mx[k];
          P.my[k] = o.my[k];
        }
        P.ow = b->tb_ow;
        P.jc = b->tb_jc;
        P.nstrips = (s.nx + b->tb_ow - 1) / b->tb_ow;
        P.nchunks = (s.ny + b->tb_jc - 1) / b->tb_jc;
        const size_t m = (size_t)std::max(s.nrec, 1);
        P.rec_start = s.tbrec;
        P.rec_i = s.tbrec + P.nstrips + 1;
        P.rec_j = P.rec_i + m;
        P.rec_slot = P.rec_j + m;
        P.sisx = static_cast<T *>(s.seis[0]);
        P.sisy = static_cast<T *>(s.seis[1]);
        P.nrec = s.nrec;
        tbd[q] = P;
      }
      BCHK(b, hipMemcpy(b->d_tbd[par], tbd.data(), tbd.size() * sizeof(tbd[0]),
                        hipMemcpyHostToDevice));
    }
  }
  if (b->tb > 0) {  // items (strip, rows) of every shot by class: slab-free, x, y, both
    // A slab item runs the C-PML code and its class has few items, so its
    // launch is bound by the latency of one wavefront's march: slab chunks
    // are cut into sub-chunks of tb_jc_slab rows (more, shorter marches).
    std::vector<int4> it[4];
    const int js = b->tb_jc_slab > 0 ? std::min(b->tb_jc_slab, b->tb_jc) : b->tb_jc;
    for (int q = 0; q < b->nshots; ++q) {
      const Shot &s = b->shots[q];
      int box[4];
      cpml::tb2d_clean_box(b->order, b->tb, s.nx, s.ny, b->tb_ow, b->tb_jc, s.sx, s.sy, box);
      const int ns = (s.nx + b->tb_ow - 1) / b->tb_ow, nc = (s.ny + b->tb_jc - 1) / b->tb_jc;
      for (int c = 0; c < nc; ++c)
        for (int t = 0; t < ns; ++t) {
          const int cls = (t >= box[0] && t < box[1] ? 0 : 1) | (c >= box[2] && c < box[3] ? 0 : 2);
          const int jb = c * b->tb_jc, je = std::min(jb + b->tb_jc, s.ny);
          const int step = cls ? js : b->tb_jc;
          for (int r = jb; r < je; r += step)
            it[cls].push_back(make_int4(q, t, r, std::min(r + step, je)));
        }
    }
    int n = 0;
    for (int k = 0; k < 4; ++k) n += b->nitems[k] = (int)it[k].size();
    if (n > b->items_cap) {
      bfree(b, *(void **)&b->d_items);
      int rc = balloc(b, (void **)&b->d_items, (size_t)n * sizeof(int4));
      if (rc) return rc;
      b->items_cap = n;
    }
    for (int k = 1; k < 4; ++k) it[0].insert(it[0].end(), it[k].begin(), it[k].end());
    BCHK(b, hipMemcpy(b->d_items, it[0].data(), (size_t)n * sizeof(int4), hipMemcpyHostToDevice));
  }
  if (b->cadence > 0) {  // energy tiles of 64 x 16 points over every shot's window
    std::vector<int2> et;
    std::vector<int> start(b->nshots + 1, 0);
    for (int q = 0; q < b->nshots; ++q) {
      const Shot &s = b->shots[q];
      const int lo = b->order == 4 ? s.npml - 1 : s.npml;
      const int i0 = std::max(lo, 0), j0 = std::max(lo, 0);
      const int i1 = std::min(b->order == 4 ? s.nx - s.npml : s.nx - s.npml - 1, s.nx - 1);
      const int j1 = std::min(b->order == 4 ? s.ny - s.npml : s.ny - s.npml - 1, s.ny - 1);
      const int tiles = i1 < i0 || j1 < j0 ? 0 : ((i1 - i0 + 64) / 64) * ((j1 - j0 + 16) / 16);
      for (int t = 0; t < tiles; ++t) et.push_back(make_int2(q, t));
      start[q + 1] = start[q] + tiles;
    }
    b->netiles = (int)et.size();
    bfree(b, *(void **)&b->d_etab);
    bfree(b, *(void **)&b->d_estart);
    bfree(b, *(void **)&b->d_epart);
    int rc = balloc(b, (void **)&b->d_etab, std::max<size_t>(et.size(), 1) * sizeof(int2));
    if (!rc) rc = balloc(b, (void **)&b->d_estart, start.size() * sizeof(int));
    if (!rc) rc = balloc(b, (void **)&b->d_epart, std::max<size_t>(et.size(), 1) * 2 * sizeof(double));
    if (rc) return rc;
    if (!et.empty())
      BCHK(b, hipMemcpy(b->d_etab, et.data(), et.size() * sizeof(int2), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_estart, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return CPML_OK;
}

int grid_of(const cpml_batch *b) {
  if (b->blocks_per_cu <= 0) return std::max(b->ntiles, 1);
  return std::max(8, std::min(b->ntiles, b->blocks_per_cu * b->ncu));
}

// workgroups of a time-blocked launch over n items: one wavefront per item, or
// a persistent grid of blocks_per_cu per CU; never fewer than 8, since the
// kernel hands every XCD (blockIdx.x % 8) its own eighth of the items
int tb_blocks(const cpml_batch *b, int n) {
  const int one = (n + 3) / 4;
  if (b->blocks_per_cu <= 0) return std::max(one, 8);
  return std::max(8, std::min(one, b->blocks_per_cu * b->ncu));
}

void *cur_desc(const cpml_batch *b) { return b->cur ? b->d_desc2 : b->d_desc; }
void *cur_rec(const cpml_batch *b) { return b->cur ? b->d_rec2 : b->d_rec; }

template <typename T>
hipError_t bstep_t(cpml_batch *b) {
  const int grid = grid_of(b);
  const bool streaming = b->bytes > cpml::kStreamingBytes;
  hipError_t e = cpml::batch_phase<T>(b->order, b->tile_y, true, streaming, cur_desc(b), b->d_ttab,
                                      b->ntiles, grid, b->stream);
  if (e == hipSuccess)
    e = cpml::batch_phase<T>(b->order, b->tile_y, false, streaming, cur_desc(b), b->d_ttab,
                             b->ntiles, grid, b->stream);
  if (e == hipSuccess && b->cadence > 0)
    e = cpml::batch_energy<T>(b->d_en[b->cur], b->d_etab, b->netiles, b->d_estart, b->nshots, 0,
                              b->d_epart, b->stream);
  if (e == hipSuccess) e = cpml::batch_record<T>(cur_rec(b), b->nshots, 0, b->stream);
  return e;
}

// NT = tb steps of every shot: the slab-free items on the main stream, the
// slab classes (both axes first, then x, then y) beside them on the side
// stream, then energy and receivers of the last step (the kernel recorded
// the first tb - 1)
template <typename T>
hipError_t bunit_t(cpml_batch *b) {
  const auto *g = static_cast<const cpml::TB2D<T> *>(b->d_tbd[b->cur]);
  int off[4] = {0, 0, 0, 0};  // start of each class in the item list
  for (int k = 1; k < 4; ++k) off[k] = off[k - 1] + b->nitems[k - 1];
  hipError_t e = hipEventRecord(b->ev_fork, b->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->side, b->ev_fork, 0);
  for (int k : {3, 1, 2})
    if (e == hipSuccess && b->nitems[k])
      e = cpml::launch_tb2d_multi<T>(b->order, b->tb, k, g, b->d_items + off[k], b->nitems[k],
                                     tb_blocks(b, b->nitems[k]), b->side);
  if (e == hipSuccess)
    e = cpml::launch_tb2d_multi<T>(b->order, b->tb, 0, g, b->d_items, b->nitems[0],
                                   tb_blocks(b, b->nitems[0]), b->stream);
  if (e == hipSuccess) e = hipEventRecord(b->ev_join, b->side);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->ev_join, 0);
  b->cur ^= 1;
  if (e == hipSuccess && b->cadence > 0)
    e = cpml::batch_energy<T>(b->d_en[b->cur], b->d_etab, b->netiles, b->d_estart, b->nshots,
                              b->tb - 1, b->d_epart, b->stream);
  if (e == hipSuccess) e = cpml::batch_record<T>(cur_rec(b), b->nshots, b->tb - 1, b->stream);
  return e;
}

hipError_t enqueue_bstep(cpml_batch *b) {
  return b->prec == 4 ? bstep_t<float>(b) : bstep_t<double>(b);
}
hipError_t enqueue_bunit(cpml_batch *b) {
  return b->prec == 4 ? bunit_t<float>(b) : bunit_t<double>(b);
}

// time blocking is used when set and compatible with the energy cadence
bool btb_active(const cpml_batch *b) {
  return b->tb > 0 && (b->cadence == 0 || b->cadence % b->tb == 0);
}
int bunit_steps(const cpml_batch *b) { return btb_active(b) ? b->tb : 1; }
// launch units per graph: an even number of ping-pong units
int bgraph_units(const cpml_batch *b) {
  if (b->graph_steps <= 0) return 0;
  if (!btb_active(b)) return b->graph_steps;
  const int n = (b->graph_steps + b->tb - 1) / b->tb;
  return (n + 1) / 2 * 2;
}
hipError_t enqueue_bany(cpml_batch *b) {
  return btb_active(b) ? enqueue_bunit(b) : enqueue_bstep(b);
}

void drop_bgraphs(cpml_batch *b) {
  for (auto &g : b->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

int build_bgraph(cpml_batch *b) {
  const int slot = b->cur;
  hipGraph_t g = nullptr;
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int u = 0; u < bgraph_units(b) && e == hipSuccess; ++u) e = enqueue_bany(b);
  b->cur = slot;  // capture does not execute
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, CPML_ERR_HIP, std::string("batch graph capture: ") +
                                      hipGetErrorString(e != hipSuccess ? e : e2));
  }
  e = hipGraphInstantiate(&b->gexec[slot], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(b, CPML_ERR_HIP, "hipGraphInstantiate failed");
  return CPML_OK;
}

int prepare(cpml_batch *b) {
  if (!b->dirty) return CPML_OK;
  int rc = b->prec == 4 ? upload_tables<float>(b) : upload_tables<double>(b);
  if (rc) return rc;
  drop_bgraphs(b);
  b->dirty = false;
  return CPML_OK;
}

// n steps of every shot from the current step: whole time-blocked units (an
// energy cadence needs units that end on its steps: leading steps go one at
// a time until the next unit would), graph replays of them, single steps left
int enqueue_bsteps(cpml_batch *b, int it, int n) {
  const int u = bunit_steps(b);
  if (u > 1 && b->cadence > 0)
    while (n > 0 && (it - 1) % u != 0) {
      BCHK(b, enqueue_bstep(b));
      ++it;
      --n;
    }
  const int gl = bgraph_units(b) * u;
  if (gl > 0 && n >= gl) {
    if (!b->gexec[b->cur]) {
      int rc = build_bgraph(b);
      if (rc) return rc;
    }
    for (; n >= gl; n -= gl, it += gl) BCHK(b, hipGraphLaunch(b->gexec[b->cur], b->stream));
  }
  for (; n >= u && btb_active(b); n -= u, it += u) BCHK(b, enqueue_bunit(b));
  for (; n > 0; --n, ++it) BCHK(b, enqueue_bstep(b));
  return CPML_OK;
}

// max |v| of every shot; the first above the threshold (or NaN) fails the run
int check_shots(cpml_batch *b, int step) {
  BCHK(b, hipMemsetAsync(b->d_vmax, 0, (size_t)b->nshots * sizeof(unsigned long long),
                         b->stream));
  BCHK(b, b->prec == 4 ? cpml::batch_vmax<float>(cur_desc(b), b->d_ttab, b->ntiles, b->tile_y,
                                                 b->d_vmax, b->stream)
                       : cpml::batch_vmax<double>(cur_desc(b), b->d_ttab, b->ntiles, b->tile_y,
                                                  b->d_vmax, b->stream));
  std::vector<unsigned long long> bits(b->nshots);
  BCHK(b, hipMemcpyAsync(bits.data(), b->d_vmax, bits.size() * sizeof(bits[0]),
                         hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int q = 0; q < b->nshots; ++q) {
    double v;
    std::memcpy(&v, &bits[q], sizeof(v));
    v = std::sqrt(v);
    if (!(v <= b->threshold)) {
      char m[200];
      std::snprintf(m, sizeof m,
                    "shot %d: code became unstable and blew up (step %d, max |v| %g)", q, step, v);
      return bfail(b, CPML_ERR_BLOWUP, m);
    }
  }
  return CPML_OK;
}

int run_steps(cpml_batch *b, int it_begin, int it_end) {
  int rc = prepare(b);
  if (rc) return rc;
  std::fill(b->host_steps.begin(), b->host_steps.end(), it_begin);
  BCHK(b, hipMemcpyAsync(b->d_steps, b->host_steps.data(), b->nshots * sizeof(int),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipEventRecord(b->ev0, b->stream));
  int it = it_begin;
  while (it <= it_end) {  // segments end at the stability checks (2D2:716-726)
    int seg = it_end;
    bool check = false;
    if (b->it_display > 0) {
      int c = ((it + b->it_display - 1) / b->it_display) * b->it_display;
      if (it <= 5) c = std::min(c, 5);
      if (c <= it_end) seg = c, check = true;
    }
    if ((rc = enqueue_bsteps(b, it, seg - it + 1))) return rc;
    if (check && (rc = check_shots(b, seg))) return rc;
    it = seg + 1;
  }
  BCHK(b, hipEventRecord(b->ev1, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  float ms = 0;
  BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  return CPML_OK;
}

// back to rest: fields, memory variables and seismograms to zero, the state
// in the first buffer
int reset_state(cpml_batch *b) {
  b->cur = 0;
  for (Shot &s : b->shots) {
    for (int k = 0; k < 5; ++k)
      BCHK(b, hipMemsetAsync(b->pool + s.off[k], 0, (size_t)s.total * b->esize, b->stream));
    for (int k = 0; k < 8; ++k)
      if (s.mem[k]) {
        const size_t e = k < 4 ? (size_t)s.sx.w * s.ny : (size_t)s.mld * s.sy.w;
        BCHK(b, hipMemsetAsync(s.mem[k], 0, std::max<size_t>(e, 1) * b->esize, b->stream));
      }
    for (int k = 0; k < 2; ++k)
      if (s.seis[k])
        BCHK(b, hipMemsetAsync(s.seis[k], 0, (size_t)b->nstep * std::max(s.nrec, 1) * b->esize,
                               b->stream));
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return CPML_OK;
}

std::string tune_key(const cpml_batch *b) {
  hipDeviceProp_t prop{};
  (void)hipGetDeviceProperties(&prop, b->device);
  double cells = 0;
  for (const Shot &s : b->shots) cells += (double)s.nx * s.ny;
  std::ostringstream k;
  k << prop.gcnArchName << "|order" << b->order << "|prec" << b->prec << "|shots" << b->nshots
    << "|mcells" << (long)(cells / 1e6);
  std::string r = k.str();
  for (char &c : r)
    if (c == ' ' || c == '\n') c = '_';
  return r;
}

// launch shape of a batch: rows per single-step tile, persistent blocks per CU
// (0: one workgroup per tile / one wavefront per item), steps per
// time-blocked launch (0: single steps), rows per time-blocked chunk
struct Shape {
  int ty = 4, bpc = 0, tb = 0, jc = 128;
};

bool cache_lookup(const char *path, const std::string &key, Shape *sh) {
  if (!path || !*path) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string k, tag;
    Shape c;
    if ((ss >> k >> c.ty >> c.bpc >> tag >> c.tb >> c.jc) && k == key && tag == "tb") {
      *sh = c;
      return true;
    }
  }
  return false;
}

void cache_store(const char *path, const std::string &key, const Shape &c, double ms) {
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  f << key << " " << c.ty << " " << c.bpc << " tb " << c.tb << " " << c.jc << " " << ms << "\n";
}

}  // namespace

extern "C" {

int cpml_batch_init(cpml_batch **bp, int order, int prec, int nshots, const int *NX,
                    const int *NY, const double *DELTAX, const double *DELTAY,
                    const double *DELTAT, const int *NPOINTS_PML, int NSTEP) {
  if (!bp) return bfail(nullptr, CPML_ERR_ARG, "null batch pointer");
  *bp = nullptr;
  if (order != 2 && order != 4 && order != 8) return bfail(nullptr, CPML_ERR_ARG, "order 2, 4 or 8");
  if (prec != 4 && prec != 8) return bfail(nullptr, CPML_ERR_ARG, "prec must be 4 or 8");
  if (nshots < 1 || nshots > 65536) return bfail(nullptr, CPML_ERR_ARG, "nshots out of range");
  if (NSTEP < 1) return bfail(nullptr, CPML_ERR_ARG, "NSTEP must be >= 1");
  if (!NX || !NY || !DELTAX || !DELTAY || !DELTAT || !NPOINTS_PML)
    return bfail(nullptr, CPML_ERR_ARG, "null per-shot parameter array");
  double cells = 0;
  for (int q = 0; q < nshots; ++q) {
    if (NX[q] < 2 || NY[q] < 2) return bfail(nullptr, CPML_ERR_STENCIL, "grid too small");
    if (!(DELTAX[q] > 0) || !(DELTAY[q] > 0) || !(DELTAT[q] > 0))
      return bfail(nullptr, CPML_ERR_ARG, "DELTAX, DELTAY, DELTAT must be positive");
    if (NPOINTS_PML[q] < 0 || 2 * NPOINTS_PML[q] > std::min(NX[q], NY[q]))
      return bfail(nullptr, CPML_ERR_STENCIL, "NPOINTS_PML does not fit the grid");
    cells += (double)NX[q] * NY[q];
  }
  if (cells > 2.0e9) return bfail(nullptr, CPML_ERR_OOM, "batch too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return bfail(nullptr, CPML_ERR_NODEVICE, "no HIP device visible");

  cpml_batch *b = new cpml_batch();
  b->order = order;
  b->prec = prec;
  b->nshots = nshots;
  b->nstep = NSTEP;
  b->esize = prec;
  b->R = order / 2;
  b->shots.resize(nshots);
  b->host_steps.assign(nshots, 1);
  auto bail = [&](int code) {
    g_batch_init_error = b->err;
    cpml_batch_finalize(b);
    return code;
  };
  hipDeviceProp_t prop{};
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
    b->err = "device query failed";
    return bail(CPML_ERR_HIP);
  }
  b->ncu = std::max(prop.multiProcessorCount, 1);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    b->err = "stream/event creation failed";
    return bail(CPML_ERR_HIP);
  }
  const long align = 256 / (long)b->esize;
  size_t off = 0;
  for (int q = 0; q < nshots; ++q) {
    Shot &s = b->shots[q];
    s.nx = NX[q];
    s.ny = NY[q];
    s.dx = DELTAX[q];
    s.dy = DELTAY[q];
    s.dt = DELTAT[q];
    s.npml = NPOINTS_PML[q];
    s.ld = cpml::round_up(align + s.nx + b->R, align);
    s.ps = s.ld * (s.ny + 2 * b->R);
    s.base = align + b->R * s.ld;
    s.total = s.ps;
    s.mld = cpml::round_up(s.nx, align);
    s.sx.n = s.nx;
    s.sy.n = s.ny;
    for (int k = 0; k < 8; ++k) {
      s.off[k] = off;
      off += (size_t)cpml::round_up(s.total * (long)b->esize, 256);
    }
  }
  size_t free_b = 0, tot_b = 0;
  if (hipMemGetInfo(&free_b, &tot_b) == hipSuccess && (double)off > 0.9 * (double)free_b) {
    b->err = "device memory budget exceeded";
    return bail(CPML_ERR_OOM);
  }
  int rc = balloc(b, (void **)&b->pool, off);
  if (!rc) rc = balloc(b, &b->d_desc, (size_t)nshots * sizeof(cpml::Step2D<double>));
  if (!rc) rc = balloc(b, &b->d_rec, (size_t)nshots * sizeof(cpml::Record<double>));
  // tile table sized for the smallest tile height (4 rows: the most tiles)
  long cap = 0;
  for (int q = 0; q < nshots; ++q)
    cap += (long)((b->shots[q].nx + cpml::BX - 1) / cpml::BX) * ((b->shots[q].ny + 3) / 4);
  b->ttab_cap = (int)cap;
  if (!rc) rc = balloc(b, (void **)&b->d_ttab, (size_t)std::max(cap, 1L) * sizeof(int2));
  if (!rc) rc = balloc(b, (void **)&b->d_steps, (size_t)nshots * sizeof(int));
  if (!rc) rc = balloc(b, (void **)&b->d_vmax, (size_t)nshots * sizeof(unsigned long long));
  for (int q = 0; q < nshots && !rc; ++q) {
    Shot &s = b->shots[q];
    const long nmax = std::max(s.nx, s.ny);
    rc = balloc(b, &s.prof, (size_t)12 * nmax * b->esize);
    if (!rc) rc = balloc(b, &s.fx, (size_t)NSTEP * b->esize);
    if (!rc) rc = balloc(b, &s.fy, (size_t)NSTEP * b->esize);
    if (!rc) rc = balloc(b, (void **)&s.recidx, sizeof(int));
    for (int k = 0; k < 2 && !rc; ++k) rc = balloc(b, &s.seis[k], b->esize);
  }
  if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) {
    b->err = "initial clear failed";
    rc = CPML_ERR_HIP;
  }
  if (rc) return bail(rc);
  *bp = b;
  return CPML_OK;
}

int cpml_batch_finalize(cpml_batch *b) {
  if (!b) return CPML_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->side) (void)hipStreamSynchronize(b->side);
  drop_bgraphs(b);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CPML_OK;
}

int cpml_batch_set_medium(cpml_batch *b, int shot, const void *lambda, const void *mu,
                          const void *rho) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  if (!lambda || !mu || !rho) return bfail(b, CPML_ERR_ARG, "null medium array");
  Shot &s = b->shots[shot];
  const size_t n = (size_t)s.nx * s.ny;
  double m = 0;
  for (size_t q = 0; q < n; ++q) {
    double l, u, r;
    if (b->prec == 4) {
      l = ((const float *)lambda)[q];
      u = ((const float *)mu)[q];
      r = ((const float *)rho)[q];
    } else {
      l = ((const double *)lambda)[q];
      u = ((const double *)mu)[q];
      r = ((const double *)rho)[q];
    }
    if (!(r > 0) || !(u > 0) || !(l + 2 * u > 0) || !std::isfinite(l))
      return bfail(b, CPML_ERR_ARG, "shot " + std::to_string(shot) + ": bad medium value");
    m = std::max(m, (l + 2 * u) / r);
  }
  double c = 0;
  cpml_courant_number(2, std::sqrt(m), s.dt, s.dx, s.dy, 1.0, &c);
  if (c > 1.0)
    return bfail(b, CPML_ERR_CFL, "shot " + std::to_string(shot) +
                                      ": time step is too large, simulation will be unstable");
  const void *src[3] = {lambda, mu, rho};
  const size_t es = b->esize;
  for (int k = 0; k < 3; ++k)
    BCHK(b, hipMemcpy2DAsync(b->pool + s.off[5 + k] + s.base * es, s.ld * es, src[k], s.nx * es,
                             s.nx * es, s.ny, hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  s.have_medium = true;
  return CPML_OK;
}

int cpml_batch_set_pml(cpml_batch *b, int shot, const void *a_x, const void *b_x,
                       const void *K_x, const void *a_x_half, const void *b_x_half,
                       const void *K_x_half, const void *a_y, const void *b_y, const void *K_y,
                       const void *a_y_half, const void *b_y_half, const void *K_y_half) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  const void *ax[2][6] = {{a_x, b_x, K_x, a_x_half, b_x_half, K_x_half},
                          {a_y, b_y, K_y, a_y_half, b_y_half, K_y_half}};
  for (auto &row : ax)
    for (const void *q : row)
      if (!q) return bfail(b, CPML_ERR_ARG, "null profile array");
  Shot &s = b->shots[shot];
  const long nmax = std::max(s.nx, s.ny);
  const int n[2] = {s.nx, s.ny};
  Slab *sl[2] = {&s.sx, &s.sy};
  std::vector<char> host((size_t)12 * nmax * b->esize, 0);
  for (int d = 0; d < 2; ++d) {
    for (int q = 0; q < 6; ++q)
      std::memcpy(host.data() + (d * 6 + q) * nmax * b->esize, ax[d][q], n[d] * b->esize);
    if (b->prec == 4)
      cpml::slab_from_profile((const float *)ax[d][0], (const float *)ax[d][3], (const float *)ax[d][2],
              (const float *)ax[d][5], n[d], *sl[d]);
    else
      cpml::slab_from_profile((const double *)ax[d][0], (const double *)ax[d][3], (const double *)ax[d][2],
              (const double *)ax[d][5], n[d], *sl[d]);
    if (sl[d]->w > n[d]) return bfail(b, CPML_ERR_STENCIL, "absorbing slabs overlap");
  }
  BCHK(b, hipMemcpy(s.prof, host.data(), host.size(), hipMemcpyHostToDevice));
  for (int k = 0; k < 8; ++k) {
    bfree(b, s.mem[k]);
    bfree(b, s.mem2[k]);
    const size_t e = k < 4 ? (size_t)s.sx.w * s.ny : (size_t)s.mld * s.sy.w;
    if ((rc = balloc(b, &s.mem[k], std::max<size_t>(e, 1) * b->esize))) return rc;
    if (b->pool2 && (rc = balloc(b, &s.mem2[k], std::max<size_t>(e, 1) * b->esize))) return rc;
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  s.have_pml = true;
  b->dirty = true;
  return CPML_OK;
}

int cpml_batch_set_source(cpml_batch *b, int shot, int ISOURCE, int JSOURCE,
                          const void *force_x, const void *force_y) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  Shot &s = b->shots[shot];
  if (ISOURCE <= 0) {
    s.isrc = s.jsrc = -1;
    b->dirty = true;
    return CPML_OK;
  }
  if (ISOURCE > s.nx - 1 || JSOURCE < 1 || JSOURCE > s.ny - 1)
    return bfail(b, CPML_ERR_ARG, "source position outside the grid");
  if (!force_x || !force_y) return bfail(b, CPML_ERR_ARG, "null force table");
  BCHK(b, hipMemcpy(s.fx, force_x, (size_t)b->nstep * b->esize, hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(s.fy, force_y, (size_t)b->nstep * b->esize, hipMemcpyHostToDevice));
  s.isrc = ISOURCE - 1;
  s.jsrc = JSOURCE - 1;
  b->dirty = true;
  return CPML_OK;
}

int cpml_batch_set_receivers(cpml_batch *b, int shot, int NREC, const int *ix_rec,
                             const int *iy_rec) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  if (NREC < 0 || (NREC > 0 && (!ix_rec || !iy_rec)))
    return bfail(b, CPML_ERR_ARG, "bad receiver arguments");
  Shot &s = b->shots[shot];
  std::vector<int> idx(std::max(NREC, 1), 0);
  for (int r = 0; r < NREC; ++r) {
    if (ix_rec[r] < 1 || ix_rec[r] > s.nx || iy_rec[r] < 1 || iy_rec[r] > s.ny)
      return bfail(b, CPML_ERR_ARG, "receiver outside the grid");
    idx[r] = (int)((ix_rec[r] - 1) + (long)(iy_rec[r] - 1) * s.ld);
  }
  bfree(b, *(void **)&s.recidx);
  for (auto &p : s.seis) bfree(b, p);
  if ((rc = balloc(b, (void **)&s.recidx, idx.size() * sizeof(int)))) return rc;
  for (auto &p : s.seis)
    if ((rc = balloc(b, &p, (size_t)b->nstep * std::max(NREC, 1) * b->esize))) return rc;
  BCHK(b, hipMemcpy(s.recidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  s.nrec = NREC;
  s.rix.assign(ix_rec, ix_rec + NREC);
  s.riy.assign(iy_rec, iy_rec + NREC);
  for (int r = 0; r < NREC; ++r) --s.rix[r], --s.riy[r];
  s.tbrec_key = -1;
  b->dirty = true;
  return CPML_OK;
}

int cpml_batch_set_tiling(cpml_batch *b, int tile_y, int blocks_per_cu, int graph_steps) {
  if (!b) return CPML_ERR_ARG;
  if (tile_y != 4 && tile_y != 8 && tile_y != 16 && tile_y != 32 && tile_y != 64)
    return bfail(b, CPML_ERR_ARG, "tile rows must be 4, 8, 16, 32 or 64");
  if (blocks_per_cu < 0 || blocks_per_cu > 64 || graph_steps < 0 || graph_steps > 4096)
    return bfail(b, CPML_ERR_ARG, "bad launch parameters");
  b->tile_y = tile_y;
  b->blocks_per_cu = blocks_per_cu;
  b->graph_steps = graph_steps;
  b->dirty = true;
  return CPML_OK;
}

int cpml_batch_autotune(cpml_batch *b, int trial_steps, const char *cache_path, int *tile_y,
                        int *blocks_per_cu, double *best_ms) {
  if (!b) return CPML_ERR_ARG;
  BCHK(b, hipSetDevice(b->device));
  if (trial_steps < 1 || trial_steps > b->nstep)
    return bfail(b, CPML_ERR_ARG, "trial_steps must be in 1..NSTEP");
  const std::string key = tune_key(b);
  const int graph_steps = b->graph_steps;
  Shape best_shape;
  double best = -1;
  if (!cache_lookup(cache_path, key, &best_shape)) {
    std::vector<Shape> trials;
    for (int ty : {4, 16, 64})  // single steps
      for (int bpc : {0, 2}) trials.push_back(Shape{ty, bpc, 0, 128});
    const int maxnt = cpml::tb2d_max_nt(b->order, b->prec);
    for (int nt = 1; nt <= maxnt; ++nt)  // time-blocked
      for (int jc : {24, 40, 64, 128})
        for (int bpc : {0, 2}) trials.push_back(Shape{4, bpc, nt, jc});
    for (const Shape &c : trials) {
      int rc = cpml_batch_set_time_block(b, c.tb, 0, c.jc);
      if (!rc) rc = cpml_batch_set_tiling(b, c.ty, c.bpc, std::min(graph_steps, trial_steps));
      if (rc) return rc;
      // whole time-blocked units only in the timed trial
      const int n = c.tb > 0 ? std::max(c.tb, trial_steps / c.tb * c.tb) : trial_steps;
      if (n > b->nstep) continue;
      if ((rc = run_steps(b, 1, n))) return rc;  // warm: first launch of a shape
      if ((rc = run_steps(b, 1, n))) return rc;
      const double per_step = b->last_ms / n;
      if (best < 0 || per_step < best) {
        best = per_step;
        best_shape = c;
      }
      if ((rc = reset_state(b))) return rc;
    }
    cache_store(cache_path, key, best_shape, best);
  }
  int rc = cpml_batch_set_time_block(b, best_shape.tb, 0, best_shape.jc);
  if (!rc) rc = cpml_batch_set_tiling(b, best_shape.ty, best_shape.bpc, graph_steps);
  if (rc) return rc;
  if ((rc = reset_state(b))) return rc;
  if (tile_y) *tile_y = best_shape.ty;
  if (blocks_per_cu) *blocks_per_cu = best_shape.bpc;
  if (best_ms) *best_ms = best < 0 ? -1.0 : best * trial_steps;
  return CPML_OK;
}

int cpml_batch_set_time_block(cpml_batch *b, int steps, int strip_width, int chunk_rows) {
  if (!b) return CPML_ERR_ARG;
  BCHK(b, hipSetDevice(b->device));
  const int r = b->order / 2, g = steps * (2 * r - 1);
  if (steps < 0 || steps > cpml::tb2d_max_nt(b->order, b->prec))
    return bfail(b, CPML_ERR_ARG, "steps per launch out of range for this order and precision");
  if (strip_width < 0 || (steps > 0 && strip_width > 64 - 2 * g) || chunk_rows < 0)
    return bfail(b, CPML_ERR_ARG, "bad strip width or chunk rows");
  if (steps > 0 && !b->pool2) {  // second state buffer: same layout as the pool
    size_t bytes = 0;
    for (const Shot &s : b->shots)
      bytes = std::max(bytes, s.off[7] + (size_t)cpml::round_up(s.total * (long)b->esize, 256));
    size_t free_b = 0, tot_b = 0;
    if (hipMemGetInfo(&free_b, &tot_b) == hipSuccess && (double)bytes > 0.9 * (double)free_b)
      return bfail(b, CPML_ERR_OOM, "not enough device memory for the second state buffer");
    int rc = balloc(b, (void **)&b->pool2, bytes);
    if (!rc) rc = balloc(b, &b->d_desc2, (size_t)b->nshots * sizeof(cpml::Step2D<double>));
    if (!rc) rc = balloc(b, &b->d_rec2, (size_t)b->nshots * sizeof(cpml::Record<double>));
    for (int par = 0; par < 2 && !rc; ++par)
      rc = balloc(b, &b->d_tbd[par], (size_t)b->nshots * sizeof(cpml::TB2D<double>));
    for (Shot &s : b->shots)
      for (int k = 0; k < 8 && !rc; ++k)
        if (s.mem[k] && !s.mem2[k]) {
          const size_t e = k < 4 ? (size_t)s.sx.w * s.ny : (size_t)s.mld * s.sy.w;
          rc = balloc(b, &s.mem2[k], std::max<size_t>(e, 1) * b->esize);
        }
    if (rc) return rc;
    BCHK(b, hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    BCHK(b, hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    BCHK(b, hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
    BCHK(b, hipStreamSynchronize(b->stream));
  }
  if (steps == 0 && b->cur) {  // bring the state back to the first buffer
    for (Shot &s : b->shots) {
      for (int k = 0; k < 5; ++k)
        BCHK(b, hipMemcpyAsync(b->pool + s.off[k], b->pool2 + s.off[k],
                               (size_t)s.total * b->esize, hipMemcpyDeviceToDevice, b->stream));
      for (int k = 0; k < 8; ++k)
        if (s.mem[k]) {
          const size_t e = k < 4 ? (size_t)s.sx.w * s.ny : (size_t)s.mld * s.sy.w;
          BCHK(b, hipMemcpyAsync(s.mem[k], s.mem2[k], std::max<size_t>(e, 1) * b->esize,
                                 hipMemcpyDeviceToDevice, b->stream));
        }
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    b->cur = 0;
  }
  b->tb = steps;
  b->tb_ow = strip_width > 0 ? strip_width : 64 - 2 * g;  // every lane the ring leaves
  if (steps > 0 && b->tb_ow > 64 - 2 * g) b->tb_ow = 64 - 2 * g;
  b->tb_jc = chunk_rows > 0 ? chunk_rows : 40;
  // slab items: half chunks from 48 rows up (128 shots of 512..2048, 64-row
  // chunks: 32-row slab items 82.4 Gcell/s, 64: 80.5, 16: 78.0, 8: 70.9)
  b->tb_jc_slab = b->tb_jc >= 48 ? (b->tb_jc + 1) / 2 : 0;
  b->dirty = true;
  return CPML_OK;
}

int cpml_batch_launch_shape(cpml_batch *b, int *tile_y, int *blocks_per_cu, int *time_block,
                            int *chunk_rows) {
  if (!b) return CPML_ERR_ARG;
  if (tile_y) *tile_y = b->tile_y;
  if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
  if (time_block) *time_block = b->tb;
  if (chunk_rows) *chunk_rows = b->tb_jc;
  return CPML_OK;
}

int cpml_batch_set_energy(cpml_batch *b, int cadence) {
  if (!b) return CPML_ERR_ARG;
  if (cadence < 0) return bfail(b, CPML_ERR_ARG, "cadence must be >= 0");
  BCHK(b, hipSetDevice(b->device));
  if (cadence > 0 && !b->d_ekin) {
    const size_t n = (size_t)b->nshots * b->nstep * sizeof(double);
    int rc = balloc(b, (void **)&b->d_ekin, n);
    if (!rc) rc = balloc(b, (void **)&b->d_epot, n);
    for (int par = 0; par < 2 && !rc; ++par)
      rc = balloc(b, &b->d_en[par], (size_t)b->nshots * sizeof(cpml::Energy<double>));
    if (rc) return rc;
    BCHK(b, hipStreamSynchronize(b->stream));
  }
  b->cadence = cadence;
  b->dirty = true;
  return CPML_OK;
}

int cpml_batch_get_energy(cpml_batch *b, int shot, double *kinetic, double *potential) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  if (!b->d_ekin) return bfail(b, CPML_ERR_STATE, "energy is off (cpml_batch_set_energy)");
  const size_t n = (size_t)b->nstep;
  if (kinetic)
    BCHK(b, hipMemcpy(kinetic, b->d_ekin + (size_t)shot * n, n * sizeof(double),
                      hipMemcpyDeviceToHost));
  if (potential)
    BCHK(b, hipMemcpy(potential, b->d_epot + (size_t)shot * n, n * sizeof(double),
                      hipMemcpyDeviceToHost));
  return CPML_OK;
}

int cpml_batch_set_stability(cpml_batch *b, int IT_DISPLAY, double threshold) {
  if (!b) return CPML_ERR_ARG;
  b->it_display = IT_DISPLAY;
  b->threshold = threshold;
  return CPML_OK;
}

int cpml_batch_run(cpml_batch *b, int it_begin, int it_end) {
  if (!b) return CPML_ERR_ARG;
  BCHK(b, hipSetDevice(b->device));
  if (it_begin < 1 || it_end > b->nstep || it_begin > it_end)
    return bfail(b, CPML_ERR_ARG, "steps must satisfy 1 <= it_begin <= it_end <= NSTEP");
  return run_steps(b, it_begin, it_end);
}

int cpml_batch_get_seismograms(cpml_batch *b, int shot, void *sisvx, void *sisvy) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  const Shot &s = b->shots[shot];
  if (s.nrec == 0) return CPML_OK;
  const size_t es = b->esize, n = (size_t)b->nstep * s.nrec;
  std::vector<char> tmp(n * es);
  void *outs[2] = {sisvx, sisvy};
  for (int q = 0; q < 2; ++q) {
    if (!outs[q]) continue;
    BCHK(b, hipMemcpy(tmp.data(), s.seis[q], tmp.size(), hipMemcpyDeviceToHost));
    char *o = static_cast<char *>(outs[q]);
    for (int it = 0; it < b->nstep; ++it)
      for (int r = 0; r < s.nrec; ++r)
        std::memcpy(o + ((size_t)it + (size_t)r * b->nstep) * es,
                    tmp.data() + ((size_t)it * s.nrec + r) * es, es);
  }
  return CPML_OK;
}

int cpml_batch_get_field(cpml_batch *b, int shot, int which, void *out) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  static const int slot[9] = {0, 1, -1, 2, 3, -1, 4, -1, -1};
  if (which < 0 || which > 8 || slot[which] < 0 || !out)
    return bfail(b, CPML_ERR_ARG, "bad field selector");
  const Shot &s = b->shots[shot];
  const size_t es = b->esize;
  const char *pool = b->cur ? b->pool2 : b->pool;
  BCHK(b, hipMemcpy2D(out, s.nx * es, pool + s.off[slot[which]] + s.base * es, s.ld * es,
                      s.nx * es, s.ny, hipMemcpyDeviceToHost));
  return CPML_OK;
}

int cpml_batch_check_kind(cpml_batch *b, int kind) {
  if (!b) return CPML_ERR_ARG;
  if (kind != b->prec)
    return bfail(b, CPML_ERR_ARG, "arrays of real(" + std::to_string(kind) +
                                      ") given to a batch of precision " +
                                      std::to_string(b->prec));
  return CPML_OK;
}

int cpml_batch_grid(cpml_batch *b, int shot, int *NX, int *NY) {
  int rc = check_shot(b, shot);
  if (rc) return rc;
  if (!NX || !NY) return bfail(b, CPML_ERR_ARG, "null output");
  *NX = b->shots[shot].nx;
  *NY = b->shots[shot].ny;
  return CPML_OK;
}

int cpml_batch_elapsed_ms(cpml_batch *b, double *ms) {
  if (!b || !ms) return CPML_ERR_ARG;
  *ms = b->last_ms;
  return CPML_OK;
}

int cpml_batch_last_error(cpml_batch *b, char *buf, int len) {
  if (!buf || len <= 0) return CPML_ERR_ARG;
  const std::string &m = b ? b->err : g_batch_init_error;
  const size_t n = std::min(m.size(), (size_t)len - 1);
  std::memcpy(buf, m.data(), n);
  buf[n] = '\0';
  return CPML_OK;
}

}  // extern "C"
