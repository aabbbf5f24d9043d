// Batched multi-shot 2D C-PML: many independent shots (each its own grid
// size, medium, profiles, source and receivers) advanced together, one
// launch per phase for all shots.
//
// Every shot is described by the same Step2D block the single-grid kernels
// use, stored in a device table; the tiles of all shots form one flat list
// (per-shot prefix sums), split into 8 contiguous bands, one per XCD, so the
// halo rows shared by neighbouring tiles stay in one L2.  A launch either has
// one workgroup per tile or a fixed "persistent" grid whose workgroups stride
// through their XCD's band.  Per point the arithmetic is the single-grid
// kernels' (cpml_points2d.h), so a batched shot equals the same shot run on
// its own, bit for bit.  Each shot has its own step counter; chunks of steps
// are captured once in a HIP graph.
//
// The autotuner times the launch shapes (rows per tile, persistent grid size)
// on the real batch for a few steps, keeps the fastest, returns the batch to
// rest and caches the choice in a small text file keyed by device, order,
// precision and batch shape.
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

// receivers of shot blockIdx.x, then that shot's step counter advances
template <typename T>
__global__ __launch_bounds__(kMaxRecBlock) void k_batch_record(const Record<T> *__restrict__ r) {
  const Record<T> &p = r[blockIdx.x];
  const int it = *p.step;
  const long row = (long)(it - 1) * p.nrec;
  for (int q = threadIdx.x; q < p.nrec; q += blockDim.x) {
    const long o = p.idx[q];
    p.sx[row + q] = p.vx[o];
    p.sy[row + q] = p.vy[o];
  }
  __syncthreads();
  if (threadIdx.x == 0) *p.step = it + 1;
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
hipError_t batch_record(const void *rec, int nshots, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_record<T>, dim3(nshots), dim3(kMaxRecBlock), 0, s,
                     static_cast<const Record<T> *>(rec));
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
  void *fx = nullptr, *fy = nullptr;
  int isrc = -1, jsrc = -1;
  int nrec = 0;
  int *recidx = nullptr;
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
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  std::vector<int> host_steps;
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
  BCHK(b, hipMemsetAsync(*p, 0, bytes, b->stream));
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
cpml::Step2D<T> shot_step(cpml_batch *b, int q) {
  const Shot &s = b->shots[q];
  cpml::Step2D<T> p{};
  T *f[8];
  for (int k = 0; k < 8; ++k) f[k] = reinterpret_cast<T *>(b->pool + s.off[k]) + s.base;
  p.vx = f[0];
  p.vy = f[1];
  p.sxx = f[2];
  p.syy = f[3];
  p.sxy = f[4];
  p.lam = f[5];
  p.mu = f[6];
  p.rho = f[7];
  for (int k = 0; k < 4; ++k) {
    p.mx[k] = static_cast<T *>(s.mem[k]);
    p.my[k] = static_cast<T *>(s.mem[4 + k]);
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

template <typename T>
int upload_tables(cpml_batch *b) {
  std::vector<cpml::Step2D<T>> desc(b->nshots);
  std::vector<cpml::Record<T>> rec(b->nshots);
  std::vector<int2> ttab;
  int nt = 0;
  for (int q = 0; q < b->nshots; ++q) {
    const Shot &s = b->shots[q];
    if (!s.have_medium || !s.have_pml)
      return bfail(b, CPML_ERR_STATE, "shot " + std::to_string(q) +
                                          ": set_medium and set_pml must be called before run");
    desc[q] = shot_step<T>(b, q);
    const int tq = (int)(((s.nx + cpml::BX - 1) / cpml::BX) * ((s.ny + b->tile_y - 1) / b->tile_y));
    for (int t = 0; t < tq; ++t) ttab.push_back(make_int2(q, t));
    nt += tq;
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
  b->ntiles = nt;
  BCHK(b, hipMemcpy(b->d_desc, desc.data(), desc.size() * sizeof(desc[0]), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_rec, rec.data(), rec.size() * sizeof(rec[0]), hipMemcpyHostToDevice));
  if (nt > b->ttab_cap) return bfail(b, CPML_ERR_STATE, "tile table too small");
  BCHK(b, hipMemcpy(b->d_ttab, ttab.data(), ttab.size() * sizeof(int2), hipMemcpyHostToDevice));
  return CPML_OK;
}

int grid_of(const cpml_batch *b) {
  if (b->blocks_per_cu <= 0) return std::max(b->ntiles, 1);
  return std::max(8, std::min(b->ntiles, b->blocks_per_cu * b->ncu));
}

hipError_t enqueue_bstep(cpml_batch *b) {
  const int grid = grid_of(b);
  const bool streaming = b->bytes > cpml::kStreamingBytes;
  hipError_t e;
  if (b->prec == 4) {
    e = cpml::batch_phase<float>(b->order, b->tile_y, true, streaming, b->d_desc, b->d_ttab, b->ntiles,
                                 grid, b->stream);
    if (e == hipSuccess)
      e = cpml::batch_phase<float>(b->order, b->tile_y, false, streaming, b->d_desc, b->d_ttab,
                                   b->ntiles, grid, b->stream);
    if (e == hipSuccess) e = cpml::batch_record<float>(b->d_rec, b->nshots, b->stream);
  } else {
    e = cpml::batch_phase<double>(b->order, b->tile_y, true, streaming, b->d_desc, b->d_ttab, b->ntiles,
                                  grid, b->stream);
    if (e == hipSuccess)
      e = cpml::batch_phase<double>(b->order, b->tile_y, false, streaming, b->d_desc, b->d_ttab,
                                    b->ntiles, grid, b->stream);
    if (e == hipSuccess) e = cpml::batch_record<double>(b->d_rec, b->nshots, b->stream);
  }
  return e;
}

int prepare(cpml_batch *b) {
  if (!b->dirty) return CPML_OK;
  int rc = b->prec == 4 ? upload_tables<float>(b) : upload_tables<double>(b);
  if (rc) return rc;
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if (b->graph_steps > 0) {
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < b->graph_steps && e == hipSuccess; ++s) e = enqueue_bstep(b);
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return bfail(b, CPML_ERR_HIP, std::string("batch graph capture: ") +
                                        hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, CPML_ERR_HIP, "hipGraphInstantiate failed");
  }
  b->dirty = false;
  return CPML_OK;
}

int run_steps(cpml_batch *b, int it_begin, int it_end) {
  int rc = prepare(b);
  if (rc) return rc;
  std::fill(b->host_steps.begin(), b->host_steps.end(), it_begin);
  BCHK(b, hipMemcpyAsync(b->d_steps, b->host_steps.data(), b->nshots * sizeof(int),
                         hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipEventRecord(b->ev0, b->stream));
  int n = it_end - it_begin + 1;
  if (b->gexec)
    for (; n >= b->graph_steps; n -= b->graph_steps) BCHK(b, hipGraphLaunch(b->gexec, b->stream));
  for (; n > 0; --n) BCHK(b, enqueue_bstep(b));
  BCHK(b, hipEventRecord(b->ev1, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  float ms = 0;
  BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  return CPML_OK;
}

// back to rest: fields, memory variables and seismograms to zero
int reset_state(cpml_batch *b) {
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

bool cache_lookup(const char *path, const std::string &key, int *ty, int *bpc) {
  if (!path || !*path) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string k;
    int a, c;
    if ((ss >> k >> a >> c) && k == key) {
      *ty = a;
      *bpc = c;
      return true;
    }
  }
  return false;
}

void cache_store(const char *path, const std::string &key, int ty, int bpc, double ms) {
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  f << key << " " << ty << " " << bpc << " " << ms << "\n";
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
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (void *p : b->allocs) (void)hipFree(p);
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
    const size_t e = k < 4 ? (size_t)s.sx.w * s.ny : (size_t)s.mld * s.sy.w;
    if ((rc = balloc(b, &s.mem[k], std::max<size_t>(e, 1) * b->esize))) return rc;
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
  int ty = 4, bpc = 0;
  double best = -1;
  if (!cache_lookup(cache_path, key, &ty, &bpc)) {
    const int tys[5] = {4, 8, 16, 32, 64};
    const int bpcs[4] = {0, 2, 4, 8};
    for (int a : tys)
      for (int c : bpcs) {
        int rc = cpml_batch_set_tiling(b, a, c, std::min(graph_steps, trial_steps));
        if (rc) return rc;
        if ((rc = run_steps(b, 1, trial_steps))) return rc;  // warm: first launch of a shape
        if ((rc = run_steps(b, 1, trial_steps))) return rc;
        if (best < 0 || b->last_ms < best) {
          best = b->last_ms;
          ty = a;
          bpc = c;
        }
      }
    cache_store(cache_path, key, ty, bpc, best);
  }
  int rc = cpml_batch_set_tiling(b, ty, bpc, graph_steps);
  if (rc) return rc;
  if ((rc = reset_state(b))) return rc;
  if (tile_y) *tile_y = ty;
  if (blocks_per_cu) *blocks_per_cu = bpc;
  if (best_ms) *best_ms = best;
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
  BCHK(b, hipMemcpy2D(out, s.nx * es, b->pool + s.off[slot[which]] + s.base * es, s.ld * es,
                      s.nx * es, s.ny, hipMemcpyDeviceToHost));
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
