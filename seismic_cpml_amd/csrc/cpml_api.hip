// C ABI and host runtime of libcpml_gpu (see include/cpml_gpu.h).
//
// A handle owns one non-blocking HIP stream, the padded wavefield/media
// arrays, the slab-only C-PML memory variables, the source and receiver
// tables and a HIP graph that replays a fixed number of time steps.  The
// current step number lives in device memory and is advanced by the receiver
// kernel at the end of every step, so one captured graph serves every
// position in the time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cpml_gpu.h"
#include "cpml_internal.h"

using cpml::Slab;

namespace {

thread_local std::string g_init_error;

// index of each memory variable in the reference naming (include/cpml_gpu.h)
// and its axis (0 x, 1 y, 2 z)
constexpr int kMemAxis[18] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 0, 2, 1, 2};
// kernel slots: x-slab, y-slab, z-slab memories in the order the kernels use
constexpr int kSlot3Dx[6] = {CPML_MEM_DVX_DX,      CPML_MEM_DVY_DX,      CPML_MEM_DVZ_DX,
                             CPML_MEM_DSIGMAXX_DX, CPML_MEM_DSIGMAXY_DX, CPML_MEM_DSIGMAXZ_DX};
constexpr int kSlot3Dy[6] = {CPML_MEM_DVY_DY,      CPML_MEM_DVX_DY,      CPML_MEM_DVZ_DY,
                             CPML_MEM_DSIGMAXY_DY, CPML_MEM_DSIGMAYY_DY, CPML_MEM_DSIGMAYZ_DY};
constexpr int kSlot3Dz[6] = {CPML_MEM_DVZ_DZ,      CPML_MEM_DVX_DZ,      CPML_MEM_DVY_DZ,
                             CPML_MEM_DSIGMAXZ_DZ, CPML_MEM_DSIGMAYZ_DZ, CPML_MEM_DSIGMAZZ_DZ};
constexpr int kSlot2Dx[4] = {CPML_MEM_DVX_DX, CPML_MEM_DVY_DX, CPML_MEM_DSIGMAXX_DX,
                             CPML_MEM_DSIGMAXY_DX};
constexpr int kSlot2Dy[4] = {CPML_MEM_DVY_DY, CPML_MEM_DVX_DY, CPML_MEM_DSIGMAXY_DY,
                             CPML_MEM_DSIGMAYY_DY};

bool mem_used(int ndim, int which) {
  if (ndim == 3) return which >= 0 && which < 18;
  for (int q = 0; q < 4; ++q)
    if (kSlot2Dx[q] == which || kSlot2Dy[q] == which) return true;
  return false;
}

__global__ void k_coeffs3d(const float *lam, const float *mu, const float *rho, float *cl,
                           float *cm, float *cb, float dt, int nx, int ny, int nz, long ld,
                           long ps) {
  const long n = (long)nx * ny * nz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const long q = (t % nx) + ((t / nx) % ny) * ld + (t / ((long)nx * ny)) * ps;
    cl[q] = dt * lam[q];
    cm[q] = dt * mu[q];
    cb[q] = dt / rho[q];
  }
}

__global__ void k_coeffs3d_d(const double *lam, const double *mu, const double *rho, double *cl,
                             double *cm, double *cb, double dt, int nx, int ny, int nz, long ld,
                             long ps) {
  const long n = (long)nx * ny * nz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const long q = (t % nx) + ((t / nx) % ny) * ld + (t / ((long)nx * ny)) * ps;
    cl[q] = dt * lam[q];
    cm[q] = dt * mu[q];
    cb[q] = dt / rho[q];
  }
}

// Scatter an interior-only array (device, Fortran order) into a padded one.
template <typename T>
__global__ void k_scatter(const T *src, T *dst, int nx, int ny, int nz, long ld, long ps) {
  const long n = (long)nx * ny * nz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const long q = (t % nx) + ((t / nx) % ny) * ld + (t / ((long)nx * ny)) * ps;
    dst[q] = src[t];
  }
}

// Medium check on the device: flags bad points and reduces max (l+2m)/rho.
template <typename T>
__global__ void k_medium_check(const T *lam, const T *mu, const T *rho, int nx, int ny, int nz,
                               long ld, long ps, unsigned long long *vmax, int *bad) {
  const long n = (long)nx * ny * nz;
  double m = 0.0;
  int b = 0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const long q = (t % nx) + ((t / nx) % ny) * ld + (t / ((long)nx * ny)) * ps;
    const double r = rho[q], u = mu[q], l = lam[q];
    if (!(r > 0.0) || !(u > 0.0) || !(l + 2.0 * u > 0.0) || !std::isfinite(l)) b = 1;
    const double v = (l + 2.0 * u) / r;
    if (v > m) m = v;
  }
  if (b) atomicOr(bad, 1);
  atomicMax(vmax, (unsigned long long)__double_as_longlong(m));
}

// Guard cells of a row-padded array: the columns c of every row (length ld)
// with c < lo or c >= hi.  op 1: fill them with quiet NaN; op 0: count those
// that no longer hold a NaN.
template <typename T>
__global__ void k_guard(T *a, long n, long ld, long lo, long hi, int op,
                        unsigned long long *changed) {
  unsigned long long c = 0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const long col = t % ld;
    if (col >= lo && col < hi) continue;
    if (op) a[t] = std::numeric_limits<T>::quiet_NaN();
    else if (a[t] == a[t]) ++c;
  }
  if (c) atomicAdd(changed, c);
}

long round_up(long v, long m) { return (v + m - 1) / m * m; }

}  // namespace

struct cpml_solver {
  int ndim = 2, order = 2, prec = 8, nx = 0, ny = 0, nz = 1, npml = 0, nstep = 0;
  double dx = 0, dy = 0, dz = 0, dt = 0;
  int device = 0;
  size_t esize = 8;
  int R = 1;
  long xpad = 0, ld = 0, ps = 0, base = 0, total = 0, mld = 0;
  // host grid arrays span (1-hlo : N+hhi) along each grid axis
  int hlo = 0, hhi = 0;
  Slab sx, sy, sz;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  double bytes = 0;
  // device arrays (padded arrays point at their allocation start)
  void *field[9] = {};  // vx vy vz sxx syy szz sxy sxz syz
  void *field2[9] = {}; // second state buffer of the fused 3D step
  int cur = 0;          // 1: the state lives in field2 / mem2
  bool fused = false;   // 3D: one fused kernel per step (ping-pong state)
  // 2D time blocking: tb steps per launch (0: off), strips of tb_ow columns
  // (0: every lane the ring leaves; 3D: a multiple of 8), chunks of tb_jc rows
  int tb = 0, tb_ow = 0, tb_jc = 40;
  // side stream of the time-blocked launches (slab classes beside the rest)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int *tb_rec = nullptr;        // per-strip receiver lists (device)
  double *tb_epart = nullptr;   // 2D time-blocked energy: per-item sums of every level
  size_t tb_epart_cap = 0;      // its size in doubles
  int tb_rec_ow = -1;           // strip width they were built for (-1: stale)
  std::vector<int> rec_ix, rec_iy, rec_iz;  // receiver grid indices (0-based), host copy
  void *media[6] = {};  // lambda mu rho, and in 3D: DELTAT*lambda DELTAT*mu DELTAT/rho
  void *prof = nullptr; // [3 axes][6][n]
  unsigned char *img = nullptr;  // PNM snapshot pixels (2D), allocated on first use
  void *mem[18] = {};
  void *mem2[18] = {};
  size_t mem_elems[18] = {};
  void *fx = nullptr, *fy = nullptr;
  int *step = nullptr;
  int *recidx = nullptr;
  void *seis[3] = {};
  int nrec = 0;
  double *ekin = nullptr, *epot = nullptr, *epart = nullptr;
  int eblocks = 0;
  unsigned long long *scratch_max = nullptr;
  int *scratch_flag = nullptr;
  bool have_medium = false, have_pml = false;
  int isrc = -1, jsrc = -1, ksrc = -1;
  int cadence = 0, ewin[6] = {}, eflags = 0;
  int it_display = 0;
  double threshold = 1e25;
  int graph_steps = 16;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // per starting state buffer
  bool graph_dirty = true;
  int kchunk = 8;
  int tile_y = 4;
  bool ktiming = false;
  std::vector<hipEvent_t> evpool;
  std::vector<int> ev_tb;  // event triples of this run that bracket a time-blocked launch
  double ktime[2] = {0, 0};
  int kcount[2] = {0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  double last_vnorm = 0;
  int host_step = 0;
  cpml_abc_net *abc = nullptr;  // learned boundary corrector (not owned)
  void *abc_corr = nullptr;     // its per-step correction buffer
  // 2D viscoelastic attenuation (cpml_gpu_set_attenuation)
  int nsls = 0;                 // 0: elastic
  double tau[4][cpml::kMaxSls] = {};  // tau_epsilon_nu1, tau_sigma_nu1, _nu2, _nu2
  void *re[3] = {};             // R_e1, R_e11, R_e13: nsls dense nx*ny planes each
  // ADE-PML RK4 step of the 8th-order program (cpml_gpu_set_ade_rk4)
  bool ade = false;
  void *ade_d[10] = {};         // stage states [2][field], padded like the fields
  void *ade_mem[8] = {};        // memory arrays, padded like the fields
  void *ade_tab[8] = {};        // a_x b_x a_x_half b_x_half a_y b_y a_y_half b_y_half: [4][n]
  void *ade_f4[2] = {};         // force at each stage time: [NSTEP][4]
  std::string err;
};

namespace {

int fail(cpml_solver *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_init_error = msg;
  return code;
}

// The 3D kernels address a z chunk through one buffer resource with unsigned
// 32-bit byte offsets: the chunk's planes plus up to 8 halo and pipeline planes
// must stay below 4 GiB.
bool chunk_fits(const cpml_solver *h, int kchunk) {
  if (h->ndim != 3) return true;
  const double planes = (double)std::min(kchunk, h->nz) + 8.0;
  return planes * (double)h->ps * (double)h->esize <= 4294967295.0;
}

#define HIPCHK(h, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((h), CPML_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int dev_alloc(cpml_solver *h, void **p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(h, e == hipErrorOutOfMemory ? CPML_ERR_OOM : CPML_ERR_HIP,
                std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  h->allocs.push_back(*p);
  h->bytes += (double)bytes;
  // cleared on the handle's stream and waited for: the handle's stream does
  // not synchronise with the null stream, so a blocking hipMemcpy into the
  // new buffer right after could otherwise be overwritten by the clear
  HIPCHK(h, hipMemsetAsync(*p, 0, bytes, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CPML_OK;
}

// n arrays of `bytes` each carved from one allocation (the arrays of one
// state or the coefficients: one contiguous range of device memory, which
// measured steadier and faster for the 3D step than separate allocations).
int dev_alloc_group(cpml_solver *h, void **p, int n, size_t bytes, const bool *want) {
  void *base = nullptr;
  int used = 0;
  for (int q = 0; q < n; ++q) used += want ? want[q] : 1;
  int rc = dev_alloc(h, &base, bytes * (used > 0 ? used : 1));
  if (rc) return rc;
  for (int q = 0, k = 0; q < n; ++q)
    if (!want || want[q]) p[q] = static_cast<char *>(base) + bytes * k++;
  return CPML_OK;
}

void dev_free(cpml_solver *h, void *&p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  (void)hipFree(p);
  p = nullptr;
}

int set_device(cpml_solver *h) {
  HIPCHK(h, hipSetDevice(h->device));
  return CPML_OK;
}

void *fbuf(cpml_solver *h, int q) { return (h->cur ? h->field2 : h->field)[q]; }
void *mbuf(cpml_solver *h, int w) { return (h->cur ? h->mem2 : h->mem)[w]; }

template <typename T>
T *F(cpml_solver *h, int q) {
  return static_cast<T *>(fbuf(h, q)) + h->base;
}
template <typename T>
T *M(cpml_solver *h, int q) {
  return static_cast<T *>(h->media[q]) + h->base;
}

// Host grid arrays: x fastest, (1-hlo : N+hhi) along x, y and (3D) z.
// host_x / host_y are the padded extents; host_at is the element offset of
// the 0-based interior point (i, j, k).
size_t host_x(const cpml_solver *h) { return (size_t)h->nx + h->hlo + h->hhi; }
size_t host_y(const cpml_solver *h) { return (size_t)h->ny + h->hlo + h->hhi; }
size_t host_at(const cpml_solver *h, long i, long j, long k) {
  const size_t kk = h->ndim == 3 ? (size_t)(k + h->hlo) : 0;
  return (size_t)(i + h->hlo) + host_x(h) * ((size_t)(j + h->hlo) + host_y(h) * kk);
}

// Copy the interior of a host grid array into / out of a padded device array
// (ghost cells of the host array are neither read nor written).
int upload_padded(cpml_solver *h, void *dst_alloc, const void *src) {
  const size_t es = h->esize;
  char *dst = static_cast<char *>(dst_alloc) + h->base * es;
  const char *s = static_cast<const char *>(src);
  const int nzz = h->ndim == 3 ? h->nz : 1;
  for (int k = 0; k < nzz; ++k)
    HIPCHK(h, hipMemcpy2DAsync(dst + k * h->ps * es, h->ld * es, s + host_at(h, 0, 0, k) * es,
                               host_x(h) * es, h->nx * es, h->ny, hipMemcpyHostToDevice,
                               h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CPML_OK;
}

int download_padded(cpml_solver *h, const void *src_alloc, void *dst) {
  const size_t es = h->esize;
  const char *src = static_cast<const char *>(src_alloc) + h->base * es;
  char *d = static_cast<char *>(dst);
  const int nzz = h->ndim == 3 ? h->nz : 1;
  for (int k = 0; k < nzz; ++k)
    HIPCHK(h, hipMemcpy2DAsync(d + host_at(h, 0, 0, k) * es, host_x(h) * es,
                               src + k * h->ps * es, h->ld * es, h->nx * es, h->ny,
                               hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CPML_OK;
}

bool field_valid(const cpml_solver *h, int which) {
  if (which < 0 || which > 8) return false;
  if (h->ndim == 2)
    return which == CPML_VX || which == CPML_VY || which == CPML_SIGMAXX ||
           which == CPML_SIGMAYY || which == CPML_SIGMAXY;
  return true;
}

// ----- kernel parameter blocks -----

template <typename T>
cpml::Prof<T> prof_of(cpml_solver *h, int axis) {
  const long nmax = std::max(h->nx, std::max(h->ny, h->nz));
  const T *b = static_cast<const T *>(h->prof) + axis * 6 * nmax;
  return cpml::Prof<T>{b, b + nmax, b + 2 * nmax, b + 3 * nmax, b + 4 * nmax, b + 5 * nmax};
}

template <typename T>
cpml::Step2D<T> step2d(cpml_solver *h) {
  cpml::Step2D<T> p{};
  p.vx = F<T>(h, CPML_VX);
  p.vy = F<T>(h, CPML_VY);
  p.sxx = F<T>(h, CPML_SIGMAXX);
  p.syy = F<T>(h, CPML_SIGMAYY);
  p.sxy = F<T>(h, CPML_SIGMAXY);
  p.lam = M<T>(h, 0);
  p.mu = M<T>(h, 1);
  p.rho = M<T>(h, 2);
  for (int q = 0; q < 4; ++q) {
    p.mx[q] = static_cast<T *>(mbuf(h, kSlot2Dx[q]));
    p.my[q] = static_cast<T *>(mbuf(h, kSlot2Dy[q]));
  }
  p.px = prof_of<T>(h, 0);
  p.py = prof_of<T>(h, 1);
  p.sx = h->sx;
  p.sy = h->sy;
  p.nx = h->nx;
  p.ny = h->ny;
  p.ld = h->ld;
  p.mld = h->mld;
  // 2D2:564 divides by DELTAX; 2D4:565 by (24*DELTAX); ADE8:734 by DELTAX
  p.den_x = h->order == 4 ? T(24) * T(h->dx) : T(h->dx);
  p.den_y = h->order == 4 ? T(24) * T(h->dy) : T(h->dy);
  p.dt = T(h->dt);
  p.isrc = h->isrc;
  p.jsrc = h->jsrc;
  p.fx = static_cast<const T *>(h->fx);
  p.fy = static_cast<const T *>(h->fy);
  p.step = h->step;
  p.tile_y = h->tile_y;
  p.streaming = h->bytes > cpml::kStreamingBytes;
  return p;
}

// 2D-visco2:374-381: per-SLS coefficients in the handle's precision
template <typename T>
cpml::Visco2D<T> visco2d(cpml_solver *h) {
  cpml::Visco2D<T> v{};
  v.s = step2d<T>(h);
  v.re1 = static_cast<T *>(h->re[0]);
  v.re11 = static_cast<T *>(h->re[1]);
  v.re13 = static_cast<T *>(h->re[2]);
  v.nn = (long)h->nx * h->ny;
  v.nsls = h->nsls;
  const T dt = T(h->dt);
  T s1 = 0, s2 = 0;
  for (int l = 0; l < h->nsls; ++l) {
    s1 = s1 + T(h->tau[0][l]) / T(h->tau[1][l]);
    s2 = s2 + T(h->tau[2][l]) / T(h->tau[3][l]);
  }
  for (int l = 0; l < h->nsls; ++l) {
    const T te1 = T(h->tau[0][l]), ts1 = T(h->tau[1][l]);
    const T te2 = T(h->tau[2][l]), ts2 = T(h->tau[3][l]);
    const T o1 = T(1) / ts1, o2 = T(1) / ts2;
    v.H1[l] = T(0.5) * dt / ts1;
    v.H2[l] = T(0.5) * dt / ts2;
    v.M1[l] = T(1) / (T(1) + T(0.5) * dt * o1);
    v.M2[l] = T(1) / (T(1) + T(0.5) * dt * o2);
    v.P1[l] = dt * (T(1) - te1 / ts1) / ts1 / s1;
    v.P2[l] = dt * (T(1) - te2 / ts2) / ts2 / s2;
  }
  v.odx = T(1) / T(h->dx);  // 2D-visco2:209-210
  v.ody = T(1) / T(h->dy);
  v.n8x = T(9) / (T(8) * T(h->dx));  // 2D-visco4: NINE_OVER_8_DELTAX
  v.n8y = T(9) / (T(8) * T(h->dy));
  v.o24x = T(1) / (T(24) * T(h->dx));  // ONE_OVER_24_DELTAX
  v.o24y = T(1) / (T(24) * T(h->dy));
  v.sisp = static_cast<T *>(h->seis[2]);
  v.recidx = h->recidx;
  v.nrec = h->nrec;
  return v;
}

template <typename T>
cpml::Step3D<T> step3d(cpml_solver *h) {
  cpml::Step3D<T> p{};
  for (int q = 0; q < 9; ++q) (&p.vx)[q] = F<T>(h, q);
  p.clam = M<T>(h, 3);
  p.cmu = M<T>(h, 4);
  p.cbuoy = M<T>(h, 5);
  p.rho = M<T>(h, 2);
  p.lam = M<T>(h, 0);
  p.mu = M<T>(h, 1);
  for (int q = 0; q < 6; ++q) {
    p.mx[q] = static_cast<T *>(mbuf(h, kSlot3Dx[q]));
    p.my[q] = static_cast<T *>(mbuf(h, kSlot3Dy[q]));
    p.mz[q] = static_cast<T *>(mbuf(h, kSlot3Dz[q]));
  }
  p.px = prof_of<T>(h, 0);
  p.py = prof_of<T>(h, 1);
  p.pz = prof_of<T>(h, 2);
  p.sx = h->sx;
  p.sy = h->sy;
  p.sz = h->sz;
  p.nx = h->nx;
  p.ny = h->ny;
  p.nz = h->nz;
  p.ld = h->ld;
  p.ps = h->ps;
  p.mld = h->mld;
  p.mps = h->mld * h->ny;
  p.base = h->base;
  p.xpad = h->xpad;
  p.R = h->R;
  // 3D:134-137 ONE_OVER_DELTA = 1/DELTA, applied by multiplication (3D:841)
  p.odx = T(1) / T(h->dx);
  p.ody = T(1) / T(h->dy);
  p.odz = T(1) / T(h->dz);
  p.dt = T(h->dt);
  p.isrc = h->isrc;
  p.jsrc = h->jsrc;
  p.ksrc = h->ksrc;
  p.fx = static_cast<const T *>(h->fx);
  p.fy = static_cast<const T *>(h->fy);
  p.step = h->step;
  p.kchunk = h->kchunk;
  p.tile_y = h->tile_y;
  p.streaming = h->bytes > cpml::kStreamingBytes;
  return p;
}

template <typename T>
cpml::Record<T> record_of(cpml_solver *h) {
  cpml::Record<T> r{};
  r.vx = F<T>(h, CPML_VX);
  r.vy = F<T>(h, CPML_VY);
  r.vz = h->ndim == 3 ? F<T>(h, CPML_VZ) : nullptr;
  r.ld = h->ld;
  r.ps = h->ps;
  r.nrec = h->nrec;
  r.idx = h->recidx;
  r.sx = static_cast<T *>(h->seis[0]);
  r.sy = static_cast<T *>(h->seis[1]);
  r.sz = h->ndim == 3 ? static_cast<T *>(h->seis[2]) : nullptr;
  r.step = h->step;
  r.avg4 = h->ade ? 1 : 0;
  return r;
}

template <typename T>
cpml::Ade8<T> ade_of(cpml_solver *h) {
  cpml::Ade8<T> a{};
  a.s = step2d<T>(h);
  for (int l = 0; l < 2; ++l)
    for (int f = 0; f < 5; ++f) a.d[l][f] = static_cast<T *>(h->ade_d[l * 5 + f]) + h->base;
  // the step accumulates X into the other state buffer
  void **nf = h->cur ? h->field : h->field2;
  const int slot[5] = {CPML_VX, CPML_VY, CPML_SIGMAXX, CPML_SIGMAYY, CPML_SIGMAXY};
  for (int f = 0; f < 5; ++f) a.xn[f] = static_cast<T *>(nf[slot[f]]) + h->base;
  for (int q = 0; q < 8; ++q) a.mem[q] = static_cast<T *>(h->ade_mem[q]) + h->base;
  const T **tab[8] = {&a.ax, &a.bx, &a.axh, &a.bxh, &a.ay, &a.by, &a.ayh, &a.byh};
  for (int q = 0; q < 8; ++q) *tab[q] = static_cast<const T *>(h->ade_tab[q]);
  a.f4x = static_cast<const T *>(h->ade_f4[0]);
  a.f4y = static_cast<const T *>(h->ade_f4[1]);
  a.dx = T(h->dx);  // ADE8 divides by DELTAX, DELTAY (ADE8:721-723)
  a.dy = T(h->dy);
  a.npml = h->npml;
  return a;
}

template <typename T>
cpml::Energy<T> energy_of(cpml_solver *h) {
  cpml::Energy<T> e{};
  for (int q = 0; q < 9; ++q)
    (&e.vx)[q] = fbuf(h, q) ? F<T>(h, q) : nullptr;
  e.lam = M<T>(h, 0);
  e.mu = M<T>(h, 1);
  e.rho = M<T>(h, 2);
  e.ld = h->ld;
  e.ps = h->ps;
  e.i0 = h->ewin[0] - 1;
  e.i1 = h->ewin[1] - 1;
  e.j0 = h->ewin[2] - 1;
  e.j1 = h->ewin[3] - 1;
  e.k0 = h->ewin[4] - 1;
  e.k1 = h->ewin[5] - 1;
  e.ndim = h->ndim;
  e.corrected3d = h->eflags & 1;
  e.cadence = h->cadence;
  e.partial = h->epart;
  e.out_kin = h->ekin;
  e.out_pot = h->epot;
  e.step = h->step;
  return e;
}

template <typename T>
cpml::Fused3D<T> fused_of(cpml_solver *h) {
  cpml::Fused3D<T> P{};
  P.a = step3d<T>(h);
  void **nf = h->cur ? h->field : h->field2;
  void **nm = h->cur ? h->mem : h->mem2;
  for (int q = 0; q < 9; ++q) (&P.vx)[q] = static_cast<T *>(nf[q]) + h->base;
  for (int q = 0; q < 6; ++q) {
    P.mx[q] = static_cast<T *>(nm[kSlot3Dx[q]]);
    P.my[q] = static_cast<T *>(nm[kSlot3Dy[q]]);
    P.mz[q] = static_cast<T *>(nm[kSlot3Dz[q]]);
  }
  return P;
}

// The fused step writes the other state buffer; the host flips `cur` as the
// launch is enqueued, so later launches (and captures) read the new state.
template <typename T>
hipError_t enqueue_fused(cpml_solver *h) {
  hipError_t e = cpml::launch_fused3d<T>(fused_of<T>(h), h->stream);
  h->cur ^= 1;
  return e;
}

template <typename T>
hipError_t enqueue_stress(cpml_solver *h) {
  if (h->ndim == 2 && h->nsls > 0)
    return cpml::launch_stress2dve<T>(h->order, visco2d<T>(h), h->stream);
  if (h->ndim == 2) return cpml::launch_stress2d<T>(h->order, step2d<T>(h), h->stream);
  return h->order == 4 ? cpml::launch_stress3d4<T>(step3d<T>(h), h->stream)
                       : cpml::launch_stress3d<T>(step3d<T>(h), h->stream);
}
template <typename T>
hipError_t enqueue_velocity(cpml_solver *h) {
  if (h->ndim == 2 && h->nsls > 0)
    return cpml::launch_velocity2dve<T>(h->order, visco2d<T>(h), h->stream);
  if (h->ndim == 2) return cpml::launch_velocity2d<T>(h->order, step2d<T>(h), h->stream);
  return h->order == 4 ? cpml::launch_velocity3d4<T>(step3d<T>(h), h->stream)
                       : cpml::launch_velocity3d<T>(step3d<T>(h), h->stream);
}
template <typename T>
hipError_t enqueue_tail(cpml_solver *h) {
  hipError_t e = hipSuccess;
  if (h->cadence > 0) {
    e = cpml::launch_energy<T>(energy_of<T>(h), h->eblocks, h->stream);
    if (e != hipSuccess) return e;
  }
  return cpml::launch_record<T>(record_of<T>(h), h->stream);
}

cpml::SolverView view_of(cpml_solver *h) {
  cpml::SolverView v{};
  v.ndim = h->ndim;
  v.order = h->order;
  v.prec = h->prec;
  v.nx = h->nx;
  v.ny = h->ny;
  v.nz = h->nz;
  v.ld = h->ld;
  v.ps = h->ps;
  for (int q = 0; q < 9; ++q)
    v.f[q] = fbuf(h, q) ? static_cast<char *>(fbuf(h, q)) + h->base * h->esize : nullptr;
  const cpml::Slab *sl[3] = {&h->sx, &h->sy, &h->sz};
  for (int a = 0; a < 3; ++a) {
    v.slab_w[2 * a] = sl[a]->wl;
    v.slab_w[2 * a + 1] = sl[a]->wr;
  }
  v.learned = h->abc != nullptr;
  v.stream = h->stream;
  v.device = h->device;
  return v;
}

// learned boundary: correct the strip after the velocity phase (source and
// rigid edges included), before receivers and energy see the state
hipError_t enqueue_boundary(cpml_solver *h) {
  if (!h->abc) return hipSuccess;
  return cpml::abc_enqueue(h->abc, static_cast<float *>(h->abc_corr), view_of(h));
}

// Time blocking is used when set and the step has nothing it does not
// cover: attenuation, the RK4 scheme, and in 3D energy at steps inside a
// block (the cadence must be a multiple of the steps per launch there; the
// 2D kernels sum the energy of every level themselves).
bool tb_active(const cpml_solver *h) {
  if (h->tb <= 0) return false;
  if (h->ndim == 3 && h->cadence != 0 && h->cadence % h->tb != 0) return false;
  if (h->ndim == 3) return h->order == 2 && !h->abc;
  return !h->ade && h->nsls == 0;
}
// rows of a 3D time-blocked tile (wavefronts per workgroup - 2 tb)
int tb3d_rows(const cpml_solver *h) { return (h->esize == 4 ? 16 : 8) - 2 * h->tb; }
// steps advanced by one launch unit
int unit_steps(const cpml_solver *h) { return tb_active(h) ? h->tb : 1; }

// Per-strip receiver lists of the time-blocked kernel, sorted by row.
// 2D: per strip of tb_ow columns, sorted by row; 3D: per tile of tb_ow
// columns x tb3d_rows rows, sorted by plane.  Layout: start[ntiles + 1], then
// i, j, slot and (3D) k arrays of nrec entries each.
int tb_tiles(const cpml_solver *h) {
  const int ns = (h->nx + h->tb_ow - 1) / h->tb_ow;
  return h->ndim == 3 ? ns * ((h->ny + tb3d_rows(h) - 1) / tb3d_rows(h)) : ns;
}
int build_tb_receivers(cpml_solver *h) {
  const int key = h->tb_ow * 64 + h->tb;
  if (h->tb_rec_ow == key) return CPML_OK;
  const int ns = (h->nx + h->tb_ow - 1) / h->tb_ow;
  const int nt = tb_tiles(h);
  const int rows = h->ndim == 3 ? tb3d_rows(h) : 1;
  auto tile = [&](int r) {
    return h->ndim == 3 ? (h->rec_iy[r] / rows) * ns + h->rec_ix[r] / h->tb_ow
                        : h->rec_ix[r] / h->tb_ow;
  };
  const int n = h->nrec;
  std::vector<int> order(n);
  for (int r = 0; r < n; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int sa = tile(a), sb = tile(b);
    if (sa != sb) return sa < sb;
    return h->ndim == 3 ? h->rec_iz[a] < h->rec_iz[b] : h->rec_iy[a] < h->rec_iy[b];
  });
  const size_t m = (size_t)std::max(n, 1);
  std::vector<int> buf((size_t)nt + 1 + 4 * m, 0);
  int *start = buf.data(), *ri = start + nt + 1, *rj = ri + m, *rs = rj + m, *rk = rs + m;
  for (int k = 0; k < n; ++k) {
    const int r = order[k];
    ri[k] = h->rec_ix[r];
    rj[k] = h->rec_iy[r];
    rk[k] = h->rec_iz[r];
    rs[k] = r;
    start[tile(r) + 1]++;
  }
  for (int t = 0; t < nt; ++t) start[t + 1] += start[t];
  dev_free(h, *(void **)&h->tb_rec);
  int rc = dev_alloc(h, (void **)&h->tb_rec, buf.size() * sizeof(int));
  if (rc) return rc;
  HIPCHK(h, hipMemcpy(h->tb_rec, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice));
  h->tb_rec_ow = key;
  return CPML_OK;
}

// Items of a 2D time-blocked launch (strips x chunks): the energy slots.
int tb2d_slots(const cpml_solver *h) {
  return ((h->nx + h->tb_ow - 1) / h->tb_ow) * ((h->ny + h->tb_jc - 1) / h->tb_jc);
}

// The per-item energy sums of a 2D time-blocked launch (allocated before any
// graph capture: enqueue_steps calls this first).
int ensure_tb_energy(cpml_solver *h) {
  if (h->ndim != 2 || h->cadence <= 0) return CPML_OK;
  const size_t need = (size_t)h->tb * tb2d_slots(h) * 2;
  if (need <= h->tb_epart_cap) return CPML_OK;
  dev_free(h, *(void **)&h->tb_epart);
  h->tb_epart_cap = 0;
  int rc = dev_alloc(h, (void **)&h->tb_epart, need * sizeof(double));
  if (rc) return rc;
  h->tb_epart_cap = need;
  h->graph_dirty = true;
  return CPML_OK;
}

template <typename T>
cpml::TB2D<T> tb_of(cpml_solver *h) {
  cpml::TB2D<T> P{};
  P.a = step2d<T>(h);
  void **nf = h->cur ? h->field : h->field2;
  void **nm = h->cur ? h->mem : h->mem2;
  const int fq[5] = {CPML_VX, CPML_VY, CPML_SIGMAXX, CPML_SIGMAYY, CPML_SIGMAXY};
  T **out[5] = {&P.vx, &P.vy, &P.sxx, &P.syy, &P.sxy};
  for (int q = 0; q < 5; ++q) *out[q] = static_cast<T *>(nf[fq[q]]) + h->base;
  for (int q = 0; q < 4; ++q) {
    P.mx[q] = static_cast<T *>(nm[kSlot2Dx[q]]);
    P.my[q] = static_cast<T *>(nm[kSlot2Dy[q]]);
  }
  P.ow = h->tb_ow;
  P.jc = h->tb_jc;
  const int nt = tb_tiles(h);
  P.rec_start = h->tb_rec;
  P.rec_i = h->tb_rec + nt + 1;
  P.rec_j = P.rec_i + std::max(h->nrec, 1);
  P.rec_slot = P.rec_j + std::max(h->nrec, 1);
  P.sisx = static_cast<T *>(h->seis[0]);
  P.sisy = static_cast<T *>(h->seis[1]);
  P.nrec = h->nrec;
  if (h->cadence > 0) {  // energy of every level, summed inside the kernel
    P.epart = h->tb_epart;
    for (int q = 0; q < 4; ++q) P.ew[q] = h->ewin[q] - 1;
    P.eslots = tb2d_slots(h);
  }
  return P;
}

template <typename T>
cpml::TB3D<T> tb3_of(cpml_solver *h) {
  cpml::TB3D<T> P{};
  P.a = step3d<T>(h);
  P.a.kchunk = h->tb_jc;  // z planes per chunk of the time-blocked march
  void **nf = h->cur ? h->field : h->field2;
  void **nm = h->cur ? h->mem : h->mem2;
  for (int q = 0; q < 9; ++q) (&P.vx)[q] = static_cast<T *>(nf[q]) + h->base;
  for (int q = 0; q < 6; ++q) {
    P.mx[q] = static_cast<T *>(nm[kSlot3Dx[q]]);
    P.my[q] = static_cast<T *>(nm[kSlot3Dy[q]]);
    P.mz[q] = static_cast<T *>(nm[kSlot3Dz[q]]);
  }
  P.ow = h->tb_ow;
  const int nt = tb_tiles(h);
  const int m = std::max(h->nrec, 1);
  P.rec_start = h->tb_rec;
  P.rec_i = h->tb_rec + nt + 1;
  P.rec_j = P.rec_i + m;
  P.rec_slot = P.rec_j + m;
  P.rec_k = P.rec_slot + m;
  P.sisx = static_cast<T *>(h->seis[0]);
  P.sisy = static_cast<T *>(h->seis[1]);
  P.sisz = static_cast<T *>(h->seis[2]);
  P.nrec = h->nrec;
  return P;
}

// One time-blocked launch (tb steps): the kernel writes the other state
// buffer and records the receivers of its first tb - 1 steps; the record
// kernel takes the last one and advances the step counter by tb.
template <typename T>
hipError_t enqueue_tb(cpml_solver *h) {
  hipError_t e = hipEventRecord(h->ev_fork, h->stream);  // fork: side waits for main
  if (e == hipSuccess) e = hipStreamWaitEvent(h->side, h->ev_fork, 0);
  if (e == hipSuccess)
    e = h->ndim == 3 ? cpml::launch_tb3d<T>(h->tb, tb3_of<T>(h), h->stream, h->side)
                     : cpml::launch_tb2d<T>(h->order, h->tb, tb_of<T>(h), h->stream, h->side);
  if (e == hipSuccess) e = hipEventRecord(h->ev_join, h->side);  // join: main waits for side
  if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->ev_join, 0);
  h->cur ^= 1;
  return e;
}
template <typename T>
hipError_t enqueue_tb_tail(cpml_solver *h) {
  if (h->cadence > 0) {
    hipError_t err;
    if (h->ndim == 2) {  // every level, from the sums of the block's items
      err = cpml::launch_tb2d_energy<T>(energy_of<T>(h), h->tb_epart, tb2d_slots(h), h->tb,
                                        h->stream);
    } else {  // the block's last step, from the state in memory
      cpml::Energy<T> e = energy_of<T>(h);
      e.offset = h->tb - 1;
      err = cpml::launch_energy<T>(e, h->eblocks, h->stream);
    }
    if (err != hipSuccess) return err;
  }
  cpml::Record<T> r = record_of<T>(h);
  r.skip = h->tb - 1;
  return cpml::launch_record<T>(r, h->stream);
}

// The RK4 step writes X to the other state buffer; the host flips `cur` as
// the step is enqueued (as for the fused step).
template <typename T>
hipError_t enqueue_ade(cpml_solver *h) {
  hipError_t e = cpml::launch_ade8_step<T>(ade_of<T>(h), h->stream);
  h->cur ^= 1;
  return e;
}

template <typename T>
hipError_t enqueue_step(cpml_solver *h) {
  hipError_t e;
  if (h->ade) {
    e = enqueue_ade<T>(h);
  } else if (h->fused) {
    e = enqueue_fused<T>(h);
  } else {
    e = enqueue_stress<T>(h);
    if (e == hipSuccess) e = enqueue_velocity<T>(h);
  }
  if (e == hipSuccess) e = enqueue_boundary(h);
  if (e == hipSuccess) e = enqueue_tail<T>(h);
  return e;
}

// launch units per captured graph: an even number of ping-pong launches
// returns the state to the buffer it started in
int graph_units(const cpml_solver *h) {
  const int u = unit_steps(h);
  if (h->graph_steps <= 0) return 0;
  if (!h->fused && !h->ade && !tb_active(h)) return h->graph_steps;
  const int n = (h->graph_steps + u - 1) / u;
  return (n + 1) / 2 * 2;
}

hipError_t enqueue_step_any(cpml_solver *h) {
  return h->prec == 4 ? enqueue_step<float>(h) : enqueue_step<double>(h);
}
// one launch unit: unit_steps(h) steps
hipError_t enqueue_unit_any(cpml_solver *h) {
  if (!tb_active(h)) return enqueue_step_any(h);
  hipError_t e = h->prec == 4 ? enqueue_tb<float>(h) : enqueue_tb<double>(h);
  if (e == hipSuccess) e = h->prec == 4 ? enqueue_tb_tail<float>(h) : enqueue_tb_tail<double>(h);
  return e;
}

void drop_graphs(cpml_solver *h) {
  for (auto &g : h->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Capture graph_len() steps starting from the current state buffer.
int build_graph(cpml_solver *h) {
  const int slot = h->cur;
  hipGraph_t g = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int s = 0; s < graph_units(h) && e == hipSuccess; ++s) e = enqueue_unit_any(h);
  h->cur = slot;  // capture does not execute: the state is still where it was
  hipError_t e2 = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(h, CPML_ERR_HIP,
                std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  e = hipGraphInstantiate(&h->gexec[slot], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(h, CPML_ERR_HIP, std::string("hipGraphInstantiate: ") +
                                                        hipGetErrorString(e));
  return CPML_OK;
}

int ensure_events(cpml_solver *h, size_t n) {
  while (h->evpool.size() < n) {
    hipEvent_t e;
    HIPCHK(h, hipEventCreate(&e));
    h->evpool.push_back(e);
  }
  return CPML_OK;
}

// Enqueue n steps; in timing mode event triples bracket each stress and
// velocity kernel (evpool[3*s .. 3*s+2] for step s of this segment).
int enqueue_steps(cpml_solver *h, int n, int *ev_used) {
  if (tb_active(h)) {
    int rc = build_tb_receivers(h);
    if (!rc) rc = ensure_tb_energy(h);
    if (rc) return rc;
  }
  const int u = unit_steps(h);
  if (h->ktiming) {
    int rc = ensure_events(h, (size_t)(*ev_used + 3 * n));
    if (rc) return rc;
    for (; n >= u && tb_active(h); n -= u) {  // time-blocked launches count as the stress kernel
      hipEvent_t *e = &h->evpool[*ev_used];
      HIPCHK(h, hipEventRecord(e[0], h->stream));
      HIPCHK(h, h->prec == 4 ? enqueue_tb<float>(h) : enqueue_tb<double>(h));
      HIPCHK(h, hipEventRecord(e[1], h->stream));
      HIPCHK(h, hipEventRecord(e[2], h->stream));
      HIPCHK(h, h->prec == 4 ? enqueue_tb_tail<float>(h) : enqueue_tb_tail<double>(h));
      h->ev_tb.push_back(*ev_used);
      *ev_used += 3;
    }
    for (int s = 0; s < n; ++s) {
      hipEvent_t *e = &h->evpool[*ev_used];
      HIPCHK(h, hipEventRecord(e[0], h->stream));
      if (h->ade) {  // the whole RK4 step counts as the stress kernel
        HIPCHK(h, h->prec == 4 ? enqueue_ade<float>(h) : enqueue_ade<double>(h));
        HIPCHK(h, hipEventRecord(e[1], h->stream));
      } else if (h->fused) {
        HIPCHK(h, h->prec == 4 ? enqueue_fused<float>(h) : enqueue_fused<double>(h));
        HIPCHK(h, hipEventRecord(e[1], h->stream));
      } else {
        HIPCHK(h, h->prec == 4 ? enqueue_stress<float>(h) : enqueue_stress<double>(h));
        HIPCHK(h, hipEventRecord(e[1], h->stream));
        HIPCHK(h, h->prec == 4 ? enqueue_velocity<float>(h) : enqueue_velocity<double>(h));
      }
      HIPCHK(h, hipEventRecord(e[2], h->stream));
      HIPCHK(h, enqueue_boundary(h));
      HIPCHK(h, h->prec == 4 ? enqueue_tail<float>(h) : enqueue_tail<double>(h));
      *ev_used += 3;
    }
    return CPML_OK;
  }
  const int gl = graph_units(h) * u;  // steps per graph replay
  if (gl > 0 && n >= gl) {
    if (h->graph_dirty) {
      drop_graphs(h);
      h->graph_dirty = false;
    }
    if (!h->gexec[h->cur]) {
      int rc = build_graph(h);
      if (rc) return rc;
    }
    while (n >= gl) {
      HIPCHK(h, hipGraphLaunch(h->gexec[h->cur], h->stream));
      n -= gl;  // an even number of ping-pong launches: the state is back in its buffer
    }
  }
  for (; n >= u && tb_active(h); n -= u) HIPCHK(h, enqueue_unit_any(h));
  for (int s = 0; s < n; ++s) HIPCHK(h, enqueue_step_any(h));
  return CPML_OK;
}

int velocity_norm(cpml_solver *h, double *out) {
  HIPCHK(h, hipMemsetAsync(h->scratch_max, 0, sizeof(unsigned long long), h->stream));
  if (h->prec == 4)
    HIPCHK(h, cpml::launch_vmax<float>(F<float>(h, 0), F<float>(h, 1),
                                       h->ndim == 3 ? F<float>(h, 2) : nullptr, h->nx, h->ny,
                                       h->ndim == 3 ? h->nz : 1, h->ld, h->ps, h->scratch_max,
                                       h->stream));
  else
    HIPCHK(h, cpml::launch_vmax<double>(F<double>(h, 0), F<double>(h, 1),
                                        h->ndim == 3 ? F<double>(h, 2) : nullptr, h->nx, h->ny,
                                        h->ndim == 3 ? h->nz : 1, h->ld, h->ps,
                                        h->scratch_max, h->stream));
  unsigned long long bits = 0;
  HIPCHK(h, hipMemcpyAsync(&bits, h->scratch_max, sizeof(bits), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double v;
  std::memcpy(&v, &bits, sizeof(v));
  *out = std::sqrt(v);
  return CPML_OK;
}

template <typename T>
int check_medium_host(cpml_solver *h, const T *lam, const T *mu, const T *rho, double *cpmax) {
  const int nzz = h->ndim == 3 ? h->nz : 1;
  double m = 0.0;
  for (int k = 0; k < nzz; ++k)
    for (int j = 0; j < h->ny; ++j) {
      const size_t row = host_at(h, 0, j, k);
      for (int i = 0; i < h->nx; ++i) {
        const size_t q = row + i;
        const double r = rho[q], u = mu[q], l = lam[q];
        if (!(r > 0.0) || !(u > 0.0) || !(l + 2.0 * u > 0.0) || !std::isfinite(l))
          return fail(h, CPML_ERR_ARG,
                      "medium: rho and mu must be positive and lambda+2mu > 0 at every point "
                      "(first bad point i,j,k = " + std::to_string(i + 1) + "," +
                          std::to_string(j + 1) + "," + std::to_string(k + 1) + ")");
        m = std::max(m, (l + 2.0 * u) / r);
      }
    }
  *cpmax = std::sqrt(m);
  return CPML_OK;
}

int cfl_check(cpml_solver *h, double cpmax) {
  double c = 0;
  cpml_courant_number(h->ndim, cpmax, h->dt, h->dx, h->dy, h->dz, &c);
  if (c > 1.0) {
    char b[200];
    std::snprintf(b, sizeof b,
                  "time step is too large, simulation will be unstable (Courant number %.6g "
                  "with cp = %.6g)", c, cpmax);
    return fail(h, CPML_ERR_CFL, b);
  }
  return CPML_OK;
}

int finish_medium(cpml_solver *h) {
  if (h->ndim == 3) {
    const int nb = 4096;
    if (h->prec == 4)
      hipLaunchKernelGGL(k_coeffs3d, dim3(nb), dim3(256), 0, h->stream, M<float>(h, 0),
                         M<float>(h, 1), M<float>(h, 2), M<float>(h, 3), M<float>(h, 4),
                         M<float>(h, 5), (float)h->dt, h->nx, h->ny, h->nz, h->ld, h->ps);
    else
      hipLaunchKernelGGL(k_coeffs3d_d, dim3(nb), dim3(256), 0, h->stream, M<double>(h, 0),
                         M<double>(h, 1), M<double>(h, 2), M<double>(h, 3), M<double>(h, 4),
                         M<double>(h, 5), h->dt, h->nx, h->ny, h->nz, h->ld, h->ps);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  h->have_medium = true;
  return CPML_OK;
}

template <typename T>
int profile_impl(int n, double delta_d, double deltat_d, int npml, double cp_d, double f0_d,
                 int use_min, int use_max, int origin_max_uses_n, int clamp_alpha, T *a, T *b,
                 T *K, T *ah, T *bh, T *Kh) {
  // 2D2:283-394: parameters and the per-point formulas of the damping profile
  const T delta = T(delta_d), deltat = T(deltat_d), cp = T(cp_d), f0 = T(f0_d);
  const T pi = std::is_same<T, float>::value ? T(3.141592653589793238462643f)
                                            : T(3.141592653589793238462643);
  const T npower = T(2);
  const T kmax = T(1);
  const T alpha_max = T(2) * pi * (f0 / T(2));
  const T L = T(npml) * delta;
  const T rcoef = std::is_same<T, float>::value ? T(0.001f) : T(0.001);
  const T d0 = -(((npower + T(1)) * cp * std::log(rcoef)) / (T(2) * L));
  const T omin = L;
  const T omax = T(origin_max_uses_n ? n : n - 1) * delta - L;
  const T eps = std::is_same<T, float>::value ? T(1.0e-6f) : T(1.0e-6);
  for (int i = 0; i < n; ++i) {
    T d[2] = {0, 0}, k[2] = {1, 1}, al[2] = {0, 0};
    const T x[2] = {delta * T(i), delta * T(i) + delta / T(2)};
    for (int h = 0; h < 2; ++h) {
      // min side first, max side overwrites (2D2:333-381)
      T dist[2] = {omin - x[h], x[h] - omax};
      for (int side = 0; side < 2; ++side) {
        if (!(side == 0 ? use_min : use_max)) continue;
        if (dist[side] >= T(0)) {
          const T xn = dist[side] / L;
          d[h] = d0 * (xn * xn);
          k[h] = T(1) + (kmax - T(1)) * (xn * xn);
          al[h] = alpha_max * (T(1) - xn);
        }
      }
      if (clamp_alpha && al[h] < T(0)) al[h] = T(0);
    }
    T bb[2], aa[2];
    for (int h = 0; h < 2; ++h) {
      bb[h] = std::exp(-(d[h] / k[h] + al[h]) * deltat);
      aa[h] = std::fabs(d[h]) > eps ? d[h] * (bb[h] - T(1)) / (k[h] * (d[h] + k[h] * al[h])) : T(0);
    }
    a[i] = aa[0];
    b[i] = bb[0];
    K[i] = k[0];
    ah[i] = aa[1];
    bh[i] = bb[1];
    Kh[i] = k[1];
  }
  return CPML_OK;
}

template <typename T>
int stf_impl(int nstep, double deltat_d, double f0_d, double t0_d, double factor_d,
             double angle_d, int kind, T *fx, T *fy) {
  // 2D2:644-657
  const T deltat = T(deltat_d), f0 = T(f0_d), t0 = T(t0_d), factor = T(factor_d);
  const T pi = std::is_same<T, float>::value ? T(3.141592653589793238462643f)
                                            : T(3.141592653589793238462643);
  const T rad = T(angle_d) * (pi / T(180));
  const T a = pi * pi * f0 * f0;
  const T sa = std::sin(rad), ca = std::cos(rad);
  for (int it = 1; it <= nstep; ++it) {
    const T t = T(it - 1) * deltat;
    const T u = t - t0;
    const T g = std::exp(-(a * (u * u)));
    T s;
    if (kind == 0)
      s = factor * g;
    else if (kind == 1)
      s = -(factor * T(2) * a * u * g);
    else
      s = factor * (T(1) - T(2) * a * (u * u)) * g;
    fx[it - 1] = sa * s;
    fy[it - 1] = ca * s;
  }
  return CPML_OK;
}

template <typename T>
int receivers_impl(int nx, int ny, double dx_d, double dy_d, int nrec, double xdeb_d,
                   double ydeb_d, double xfin_d, double yfin_d, int *ix, int *iy) {
  // 2D2:483-509
  const T dx = T(dx_d), dy = T(dy_d), xdeb = T(xdeb_d), ydeb = T(ydeb_d);
  const T xsp = nrec > 1 ? (T(xfin_d) - xdeb) / T(nrec - 1) : T(0);
  const T ysp = nrec > 1 ? (T(yfin_d) - ydeb) / T(nrec - 1) : T(0);
  for (int r = 0; r < nrec; ++r) {
    const T xr = xdeb + T(r) * xsp, yr = ydeb + T(r) * ysp;
    T best = T(1e30);
    ix[r] = iy[r] = 1;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const T ex = dx * T(i) - xr, ey = dy * T(j) - yr;
        const T dist = std::sqrt(ex * ex + ey * ey);
        if (dist < best) {
          best = dist;
          ix[r] = i + 1;
          iy[r] = j + 1;
        }
      }
  }
  return CPML_OK;
}

}  // namespace

extern "C" {

int cpml_gpu_init(cpml_solver **hp, int ndim, int order, int prec, int NX, int NY, int NZ,
                  double DELTAX, double DELTAY, double DELTAZ, double DELTAT, int NPOINTS_PML,
                  int NSTEP) {
  if (!hp) return fail(nullptr, CPML_ERR_ARG, "null handle pointer");
  *hp = nullptr;
  if (ndim != 2 && ndim != 3) return fail(nullptr, CPML_ERR_ARG, "ndim must be 2 or 3");
  if (prec != 4 && prec != 8) return fail(nullptr, CPML_ERR_ARG, "prec must be 4 or 8");
  if (ndim == 2 && order != 2 && order != 4 && order != 8)
    return fail(nullptr, CPML_ERR_ARG, "2D order must be 2, 4 or 8");
  if (ndim == 3 && order != 2 && order != 4)
    return fail(nullptr, CPML_ERR_ARG, "3D order must be 2 or 4");
  if (NX < 2 || NY < 2 || (ndim == 3 && NZ < 2))
    return fail(nullptr, CPML_ERR_STENCIL, "grid must have at least 2 points per axis");
  if (!(DELTAX > 0) || !(DELTAY > 0) || !(DELTAT > 0) || (ndim == 3 && !(DELTAZ > 0)))
    return fail(nullptr, CPML_ERR_ARG, "DELTAX, DELTAY, DELTAZ and DELTAT must be positive");
  if (NSTEP < 1) return fail(nullptr, CPML_ERR_ARG, "NSTEP must be >= 1");
  if (NPOINTS_PML < 0 || 2 * NPOINTS_PML > std::min(NX, NY) ||
      (ndim == 3 && 2 * NPOINTS_PML > NZ))
    return fail(nullptr, CPML_ERR_STENCIL, "NPOINTS_PML does not fit the grid");
  const double cells = (double)NX * NY * (ndim == 3 ? NZ : 1);
  if (cells > 4.0e10) return fail(nullptr, CPML_ERR_OOM, "grid too large");
  // the 3D kernels index every array with unsigned 32-bit offsets
  if (ndim == 3 && ((double)NX + 80) * (NY + 2) * (NZ + 2) > 4.0e9)
    return fail(nullptr, CPML_ERR_OOM, "3D grid too large for 32-bit offsets (about 1500^3)");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, CPML_ERR_NODEVICE, "no HIP device visible");

  cpml_solver *h = new cpml_solver();
  h->ndim = ndim;
  h->order = order;
  h->prec = prec;
  h->nx = NX;
  h->ny = NY;
  h->nz = ndim == 3 ? NZ : 1;
  h->dx = DELTAX;
  h->dy = DELTAY;
  h->dz = ndim == 3 ? DELTAZ : 1.0;
  h->dt = DELTAT;
  h->npml = NPOINTS_PML;
  h->nstep = NSTEP;
  h->esize = prec;
  h->R = ndim == 3 ? 2 : order / 2;  // 3D: two halo cells (the fused step reads i-2, j-2, k-2)
  const long align = 256 / (long)h->esize;  // elements per 256 bytes
  h->xpad = align;
  h->ld = round_up(h->xpad + NX + h->R, align);
  h->ps = h->ld * (NY + 2 * h->R);
  h->base = h->xpad + h->R * h->ld + (ndim == 3 ? h->R * h->ps : 0);
  h->total = h->ps * (ndim == 3 ? NZ + 2 * h->R : 1);
  h->mld = round_up(NX, align);
  // default chunk depth, reduced where a chunk's buffer would pass 4 GiB
  while (h->kchunk > 1 && !chunk_fits(h, h->kchunk)) --h->kchunk;
  if (!chunk_fits(h, 1)) {
    delete h;
    return fail(nullptr, CPML_ERR_OOM, "3D grid planes too large for 32-bit buffer offsets");
  }
  h->sx.n = NX;
  h->sy.n = NY;
  h->sz.n = h->nz;
  // default energy window: 2D2:695 / 2D4:696 / 3D:1137-1139
  const int p = NPOINTS_PML;
  if (ndim == 2 && order == 4) {
    const int w[6] = {p, NX - p + 1, p, NY - p + 1, 1, 1};
    std::copy(w, w + 6, h->ewin);
  } else {
    const int w[6] = {p + 1, NX - p, p + 1, NY - p, ndim == 3 ? p + 1 : 1, ndim == 3 ? NZ - p : 1};
    std::copy(w, w + 6, h->ewin);
  }
  for (int q = 0; q < 6; ++q) h->ewin[q] = std::max(h->ewin[q], 1);

  int rc = CPML_OK;
  auto bail = [&](int code) {
    g_init_error = h->err;
    cpml_gpu_finalize(h);
    return code;
  };
  if (hipGetDevice(&h->device) != hipSuccess) {
    h->err = "hipGetDevice failed";
    return bail(CPML_ERR_HIP);
  }
  size_t free_b = 0, tot_b = 0;
  if (hipMemGetInfo(&free_b, &tot_b) == hipSuccess) {
    const double need = (double)h->total * h->esize * (ndim == 3 ? 15 : 8);
    if (need > 0.95 * (double)free_b) {
      h->err = "device memory budget exceeded: need " + std::to_string(need / 1e9) +
               " GB, free " + std::to_string(free_b / 1e9) + " GB";
      return bail(CPML_ERR_OOM);
    }
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    h->err = "hipStreamCreate failed";
    return bail(CPML_ERR_HIP);
  }
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
    h->err = "hipEventCreate failed";
    return bail(CPML_ERR_HIP);
  }
  const size_t fbytes = (size_t)h->total * h->esize;
  bool want[9];
  for (int q = 0; q < 9; ++q) want[q] = field_valid(h, q);
  rc = dev_alloc_group(h, h->field, 9, fbytes, want);
  if (rc == CPML_OK) rc = dev_alloc_group(h, h->media, ndim == 3 ? 6 : 3, fbytes, nullptr);
  const long nmax = std::max(NX, std::max(NY, h->nz));
  if (rc == CPML_OK) rc = dev_alloc(h, &h->prof, (size_t)18 * nmax * h->esize);
  if (rc == CPML_OK) rc = dev_alloc(h, &h->fx, (size_t)NSTEP * h->esize);
  if (rc == CPML_OK) rc = dev_alloc(h, &h->fy, (size_t)NSTEP * h->esize);
  if (rc == CPML_OK) rc = dev_alloc(h, (void **)&h->step, sizeof(int) * 64);
  if (rc == CPML_OK) rc = dev_alloc(h, (void **)&h->scratch_max, 256);
  if (rc == CPML_OK) rc = dev_alloc(h, (void **)&h->scratch_flag, 256);
  if (rc == CPML_OK) rc = dev_alloc(h, (void **)&h->recidx, sizeof(int));
  for (int q = 0; q < 3 && rc == CPML_OK; ++q) rc = dev_alloc(h, &h->seis[q], h->esize);
  if (rc == CPML_OK && hipStreamSynchronize(h->stream) != hipSuccess) {
    h->err = "initial clear failed";
    rc = CPML_ERR_HIP;
  }
  if (rc != CPML_OK) return bail(rc);
  *hp = h;
  return CPML_OK;
}

int cpml_gpu_finalize(cpml_solver *h) {
  if (!h) return CPML_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto g : h->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
  if (h->side) (void)hipStreamSynchronize(h->side);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CPML_OK;
}

int cpml_gpu_set_medium(cpml_solver *h, const void *lambda, const void *mu, const void *rho) {
  if (!h) return CPML_ERR_ARG;
  if (!lambda || !mu || !rho) return fail(h, CPML_ERR_ARG, "null medium array");
  int rc = set_device(h);
  if (rc) return rc;
  double cpmax = 0;
  rc = h->prec == 4 ? check_medium_host(h, (const float *)lambda, (const float *)mu,
                                        (const float *)rho, &cpmax)
                    : check_medium_host(h, (const double *)lambda, (const double *)mu,
                                        (const double *)rho, &cpmax);
  if (rc) return rc;
  if ((rc = cfl_check(h, cpmax))) return rc;
  const void *src[3] = {lambda, mu, rho};
  for (int q = 0; q < 3; ++q)
    if ((rc = upload_padded(h, h->media[q], src[q]))) return rc;
  h->graph_dirty = true;
  return finish_medium(h);
}

int cpml_gpu_set_medium_device(cpml_solver *h, const void *d_lambda, const void *d_mu,
                               const void *d_rho) {
  if (!h) return CPML_ERR_ARG;
  if (!d_lambda || !d_mu || !d_rho) return fail(h, CPML_ERR_ARG, "null medium array");
  int rc = set_device(h);
  if (rc) return rc;
  const void *src[3] = {d_lambda, d_mu, d_rho};
  for (int q = 0; q < 3; ++q) {
    if (h->prec == 4)
      hipLaunchKernelGGL(k_scatter<float>, dim3(4096), dim3(256), 0, h->stream,
                         (const float *)src[q], M<float>(h, q), h->nx, h->ny, h->nz, h->ld,
                         h->ps);
    else
      hipLaunchKernelGGL(k_scatter<double>, dim3(4096), dim3(256), 0, h->stream,
                         (const double *)src[q], M<double>(h, q), h->nx, h->ny, h->nz, h->ld,
                         h->ps);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipMemsetAsync(h->scratch_max, 0, 256, h->stream));
  HIPCHK(h, hipMemsetAsync(h->scratch_flag, 0, 256, h->stream));
  if (h->prec == 4)
    hipLaunchKernelGGL(k_medium_check<float>, dim3(2048), dim3(256), 0, h->stream,
                       M<float>(h, 0), M<float>(h, 1), M<float>(h, 2), h->nx, h->ny, h->nz,
                       h->ld, h->ps, h->scratch_max, h->scratch_flag);
  else
    hipLaunchKernelGGL(k_medium_check<double>, dim3(2048), dim3(256), 0, h->stream,
                       M<double>(h, 0), M<double>(h, 1), M<double>(h, 2), h->nx, h->ny, h->nz,
                       h->ld, h->ps, h->scratch_max, h->scratch_flag);
  HIPCHK(h, hipGetLastError());
  unsigned long long bits = 0;
  int bad = 0;
  HIPCHK(h, hipMemcpyAsync(&bits, h->scratch_max, 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(&bad, h->scratch_flag, 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (bad)
    return fail(h, CPML_ERR_ARG,
                "medium: rho and mu must be positive and lambda+2mu > 0 at every point");
  double v;
  std::memcpy(&v, &bits, 8);
  if ((rc = cfl_check(h, std::sqrt(v)))) return rc;
  h->graph_dirty = true;
  return finish_medium(h);
}

int cpml_gpu_set_pml(cpml_solver *h, const void *a_x, const void *b_x, const void *K_x,
                     const void *a_x_half, const void *b_x_half, const void *K_x_half,
                     const void *a_y, const void *b_y, const void *K_y, const void *a_y_half,
                     const void *b_y_half, const void *K_y_half, const void *a_z,
                     const void *b_z, const void *K_z, const void *a_z_half,
                     const void *b_z_half, const void *K_z_half) {
  if (!h) return CPML_ERR_ARG;
  const void *ax[3][6] = {{a_x, b_x, K_x, a_x_half, b_x_half, K_x_half},
                          {a_y, b_y, K_y, a_y_half, b_y_half, K_y_half},
                          {a_z, b_z, K_z, a_z_half, b_z_half, K_z_half}};
  for (int d = 0; d < h->ndim; ++d)
    for (int q = 0; q < 6; ++q)
      if (!ax[d][q]) return fail(h, CPML_ERR_ARG, "null profile array");
  int rc = set_device(h);
  if (rc) return rc;
  const int n[3] = {h->nx, h->ny, h->nz};
  Slab *sl[3] = {&h->sx, &h->sy, &h->sz};
  const long nmax = std::max(h->nx, std::max(h->ny, h->nz));
  std::vector<char> host((size_t)18 * nmax * h->esize, 0);
  for (int d = 0; d < 3; ++d) {
    if (d >= h->ndim) {  // unused axis: K = 1, a = 0
      for (int q = 0; q < 6; ++q)
        for (long i = 0; i < nmax; ++i) {
          char *dst = host.data() + ((d * 6 + q) * nmax + i) * h->esize;
          const double v = (q == 2 || q == 5) ? 1.0 : 0.0;
          if (h->prec == 4) {
            float f = (float)v;
            std::memcpy(dst, &f, 4);
          } else {
            std::memcpy(dst, &v, 8);
          }
        }
      *sl[d] = Slab{0, 0, 0, n[d]};
      continue;
    }
    for (int q = 0; q < 6; ++q)
      std::memcpy(host.data() + (d * 6 + q) * nmax * h->esize, ax[d][q], n[d] * h->esize);
    if (h->prec == 4)
      cpml::slab_from_profile((const float *)ax[d][0], (const float *)ax[d][3],
                        (const float *)ax[d][2], (const float *)ax[d][5], n[d], *sl[d]);
    else
      cpml::slab_from_profile((const double *)ax[d][0], (const double *)ax[d][3],
                        (const double *)ax[d][2], (const double *)ax[d][5], n[d], *sl[d]);
    if (sl[d]->w > n[d]) return fail(h, CPML_ERR_STENCIL, "absorbing slabs overlap");
  }
  HIPCHK(h, hipMemcpy(h->prof, host.data(), host.size(), hipMemcpyHostToDevice));
  // (re)allocate the slab-only memory variables, zeroed
  const size_t ny = h->ny, nzz = h->nz;
  for (int w = 0; w < 18; ++w) {
    if (!mem_used(h->ndim, w)) continue;
    size_t elems = 0;
    switch (kMemAxis[w]) {
      case 0: elems = (size_t)h->sx.w * ny * nzz; break;
      case 1: elems = (size_t)h->mld * h->sy.w * nzz; break;
      default: elems = (size_t)h->mld * ny * h->sz.w; break;
    }
    dev_free(h, h->mem[w]);
    dev_free(h, h->mem2[w]);
    h->mem_elems[w] = elems;
    if ((rc = dev_alloc(h, &h->mem[w], elems * h->esize))) return rc;
    if ((h->fused || h->tb) && (rc = dev_alloc(h, &h->mem2[w], elems * h->esize))) return rc;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_pml = true;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_source(cpml_solver *h, int ISOURCE, int JSOURCE, int KSOURCE,
                        const void *force_x, const void *force_y) {
  if (!h) return CPML_ERR_ARG;
  int rc = set_device(h);
  if (rc) return rc;
  if (ISOURCE <= 0) {
    h->isrc = h->jsrc = h->ksrc = -1;
    h->graph_dirty = true;
    return CPML_OK;
  }
  // 2D: the density average at (i+1/2, j+1/2) needs i < NX and j < NY (2D2:663)
  const int imax = h->ndim == 2 ? h->nx - 1 : h->nx;
  const int jmax = h->ndim == 2 ? h->ny - 1 : h->ny;
  if (ISOURCE > imax || JSOURCE < 1 || JSOURCE > jmax ||
      (h->ndim == 3 && (KSOURCE < 1 || KSOURCE > h->nz)))
    return fail(h, CPML_ERR_ARG, "source position outside the grid");
  if (!force_x || !force_y) return fail(h, CPML_ERR_ARG, "null force table");
  HIPCHK(h, hipMemcpy(h->fx, force_x, (size_t)h->nstep * h->esize, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->fy, force_y, (size_t)h->nstep * h->esize, hipMemcpyHostToDevice));
  h->isrc = ISOURCE - 1;
  h->jsrc = JSOURCE - 1;
  h->ksrc = h->ndim == 3 ? KSOURCE - 1 : 0;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_receivers(cpml_solver *h, int NREC, const int *ix_rec, const int *iy_rec,
                           const int *iz_rec) {
  if (!h) return CPML_ERR_ARG;
  if (NREC < 0) return fail(h, CPML_ERR_ARG, "NREC must be >= 0");
  if (NREC > 0 && (!ix_rec || !iy_rec || (h->ndim == 3 && !iz_rec)))
    return fail(h, CPML_ERR_ARG, "null receiver index array");
  int rc = set_device(h);
  if (rc) return rc;
  std::vector<int> idx(std::max(NREC, 1), 0);
  for (int r = 0; r < NREC; ++r) {
    const int i = ix_rec[r], j = iy_rec[r], k = h->ndim == 3 ? iz_rec[r] : 1;
    if (i < 1 || i > h->nx || j < 1 || j > h->ny || k < 1 || k > h->nz)
      return fail(h, CPML_ERR_ARG, "receiver " + std::to_string(r + 1) + " outside the grid");
    const long q = (i - 1) + (long)(j - 1) * h->ld + (long)(k - 1) * h->ps;
    if (q > 0x7fffffffL) return fail(h, CPML_ERR_ARG, "receiver offset overflow");
    idx[r] = (int)q;
  }
  dev_free(h, *(void **)&h->recidx);
  for (int q = 0; q < 3; ++q) dev_free(h, h->seis[q]);
  if ((rc = dev_alloc(h, (void **)&h->recidx, idx.size() * sizeof(int)))) return rc;
  const size_t sbytes = (size_t)h->nstep * std::max(NREC, 1) * h->esize;
  for (int q = 0; q < 3; ++q)
    if ((rc = dev_alloc(h, &h->seis[q], sbytes)))  // 2D: [2] holds pressure (viscoelastic)
      return rc;
  HIPCHK(h, hipMemcpy(h->recidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  h->nrec = NREC;
  h->rec_ix.assign(ix_rec, ix_rec + NREC);
  h->rec_iy.assign(iy_rec, iy_rec + NREC);
  h->rec_iz.assign((size_t)NREC, 1);
  if (h->ndim == 3) h->rec_iz.assign(iz_rec, iz_rec + NREC);
  for (int r = 0; r < NREC; ++r) --h->rec_ix[r], --h->rec_iy[r], --h->rec_iz[r];
  h->tb_rec_ow = -1;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_energy(cpml_solver *h, int cadence, const int *window, int flags) {
  if (!h) return CPML_ERR_ARG;
  if (cadence < 0) return fail(h, CPML_ERR_ARG, "cadence must be >= 0");
  int rc = set_device(h);
  if (rc) return rc;
  if (window) {
    const int n[3] = {h->nx, h->ny, h->nz};
    for (int d = 0; d < 3; ++d) {
      const int lo = window[2 * d], hi = window[2 * d + 1];
      if (d >= h->ndim) continue;
      if (lo < 1 || hi > n[d] || lo > hi) return fail(h, CPML_ERR_ARG, "bad energy window");
    }
    std::copy(window, window + 2 * h->ndim, h->ewin);
    if (h->ndim == 2) h->ewin[4] = h->ewin[5] = 1;
  }
  h->cadence = cadence;
  h->eflags = flags;
  dev_free(h, *(void **)&h->ekin);
  dev_free(h, *(void **)&h->epot);
  dev_free(h, *(void **)&h->epart);
  if (cadence > 0) {
    h->eblocks = h->prec == 4 ? cpml::energy_blocks<float>(energy_of<float>(h))
                              : cpml::energy_blocks<double>(energy_of<double>(h));
    if ((rc = dev_alloc(h, (void **)&h->ekin, (size_t)h->nstep * sizeof(double)))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->epot, (size_t)h->nstep * sizeof(double)))) return rc;
    if ((rc = dev_alloc(h, (void **)&h->epart, (size_t)h->eblocks * 2 * sizeof(double))))
      return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_stability(cpml_solver *h, int IT_DISPLAY, double threshold) {
  if (!h) return CPML_ERR_ARG;
  h->it_display = IT_DISPLAY;
  h->threshold = threshold;
  return CPML_OK;
}

int cpml_gpu_run(cpml_solver *h, int it_begin, int it_end) {
  if (!h) return CPML_ERR_ARG;
  if (!h->have_medium || !h->have_pml)
    return fail(h, CPML_ERR_STATE, "set_medium and set_pml must be called before run");
  if (it_begin < 1 || it_end > h->nstep || it_begin > it_end)
    return fail(h, CPML_ERR_ARG, "steps must satisfy 1 <= it_begin <= it_end <= NSTEP");
  int rc = set_device(h);
  if (rc) return rc;
  if (h->abc) {
    std::string m;
    if ((rc = cpml::abc_prepare(h->abc, view_of(h), &m))) return fail(h, rc, m);
  }
  h->host_step = it_begin;
  HIPCHK(h, hipMemcpyAsync(h->step, &h->host_step, sizeof(int), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  int ev_used = 0;
  h->ev_tb.clear();
  int it = it_begin;
  while (it <= it_end) {
    int seg_end = it_end;
    bool check = false;
    if (h->it_display > 0) {
      int c = ((it + h->it_display - 1) / h->it_display) * h->it_display;
      if (it <= 5) c = std::min(c, 5);
      if (c <= it_end) {
        seg_end = c;
        check = true;
      }
    }
    if ((rc = enqueue_steps(h, seg_end - it + 1, &ev_used))) return rc;
    it = seg_end + 1;
    if (check) {
      double v = 0;
      if ((rc = velocity_norm(h, &v))) return rc;
      h->last_vnorm = v;
      if (!(v <= h->threshold)) {
        char b[160];
        std::snprintf(b, sizeof b, "code became unstable and blew up (step %d, max |v| %g)",
                      seg_end, v);
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return fail(h, CPML_ERR_BLOWUP, b);
      }
    }
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  size_t tbk = 0;
  for (int s = 0; s + 2 < ev_used; s += 3) {
    float a = 0, b = 0;
    HIPCHK(h, hipEventElapsedTime(&a, h->evpool[s], h->evpool[s + 1]));
    h->ktime[0] += a;
    h->kcount[0] += 1;
    const bool tbl = tbk < h->ev_tb.size() && h->ev_tb[tbk] == s;
    if (tbl) ++tbk;
    if (!h->fused && !tbl) {
      HIPCHK(h, hipEventElapsedTime(&b, h->evpool[s + 1], h->evpool[s + 2]));
      h->ktime[1] += b;
      h->kcount[1] += 1;
    }
  }
  return CPML_OK;
}

int cpml_gpu_get_seismograms(cpml_solver *h, void *sisvx, void *sisvy, void *sisvz) {
  if (!h) return CPML_ERR_ARG;
  int rc = set_device(h);
  if (rc) return rc;
  if (h->nrec == 0) return CPML_OK;
  const size_t n = (size_t)h->nstep * h->nrec;
  std::vector<char> tmp(n * h->esize);
  void *outs[3] = {sisvx, sisvy, sisvz};
  for (int q = 0; q < 3; ++q) {
    if (!outs[q] || (q == 2 && h->ndim == 2)) continue;
    HIPCHK(h, hipMemcpy(tmp.data(), h->seis[q], tmp.size(), hipMemcpyDeviceToHost));
    char *o = static_cast<char *>(outs[q]);
    for (int it = 0; it < h->nstep; ++it)
      for (int r = 0; r < h->nrec; ++r)
        std::memcpy(o + ((size_t)it + (size_t)r * h->nstep) * h->esize,
                    tmp.data() + ((size_t)it * h->nrec + r) * h->esize, h->esize);
  }
  return CPML_OK;
}

int cpml_gpu_get_energy(cpml_solver *h, double *kinetic, double *potential) {
  if (!h) return CPML_ERR_ARG;
  if (h->cadence <= 0) return fail(h, CPML_ERR_STATE, "energy is off (set_energy)");
  int rc = set_device(h);
  if (rc) return rc;
  if (kinetic)
    HIPCHK(h, hipMemcpy(kinetic, h->ekin, h->nstep * sizeof(double), hipMemcpyDeviceToHost));
  if (potential)
    HIPCHK(h, hipMemcpy(potential, h->epot, h->nstep * sizeof(double), hipMemcpyDeviceToHost));
  return CPML_OK;
}

int cpml_gpu_get_field(cpml_solver *h, int which, void *out) {
  if (!h) return CPML_ERR_ARG;
  if (!field_valid(h, which) || !out) return fail(h, CPML_ERR_ARG, "bad field selector");
  int rc = set_device(h);
  if (rc) return rc;
  return download_padded(h, fbuf(h, which), out);
}

int cpml_gpu_set_field(cpml_solver *h, int which, const void *in) {
  if (!h) return CPML_ERR_ARG;
  if (!field_valid(h, which) || !in) return fail(h, CPML_ERR_ARG, "bad field selector");
  int rc = set_device(h);
  if (rc) return rc;
  return upload_padded(h, fbuf(h, which), in);
}

int cpml_gpu_get_memory(cpml_solver *h, int which, void *out) {
  if (!h) return CPML_ERR_ARG;
  if (!mem_used(h->ndim, which) || !out) return fail(h, CPML_ERR_ARG, "bad memory selector");
  if (!h->have_pml) return fail(h, CPML_ERR_STATE, "set_pml first");
  int rc = set_device(h);
  if (rc) return rc;
  const size_t es = h->esize;
  std::vector<char> tmp(std::max<size_t>(h->mem_elems[which], 1) * es);
  if (h->mem_elems[which])
    HIPCHK(h, hipMemcpy(tmp.data(), mbuf(h, which), h->mem_elems[which] * es,
                        hipMemcpyDeviceToHost));
  char *o = static_cast<char *>(out);
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) std::memset(o + host_at(h, 0, j, k) * es, 0, (size_t)nx * es);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long src = -1;
        switch (kMemAxis[which]) {
          case 0:
            if (h->sx.in(i)) src = h->sx.s(i) + (long)h->sx.w * (j + (long)ny * k);
            break;
          case 1:
            if (h->sy.in(j)) src = i + h->mld * (h->sy.s(j) + (long)h->sy.w * k);
            break;
          default:
            if (h->sz.in(k)) src = i + h->mld * (j + (long)ny * h->sz.s(k));
            break;
        }
        if (src >= 0) std::memcpy(o + host_at(h, i, j, k) * es, tmp.data() + src * es, es);
      }
  return CPML_OK;
}

int cpml_gpu_check_kind(cpml_solver *h, int kind) {
  if (!h) return CPML_ERR_ARG;
  if (kind != h->prec)
    return fail(h, CPML_ERR_ARG, "arrays of real(" + std::to_string(kind) +
                                     ") given to a handle of precision " +
                                     std::to_string(h->prec));
  return CPML_OK;
}

int cpml_gpu_guard_cells(cpml_solver *h, int op, long long *count) {
  if (!h || !count || (op != 0 && op != 1)) return CPML_ERR_ARG;
  int rc = set_device(h);
  if (rc) return rc;
  struct Arr {
    void *p;
    long n, ld, lo, hi;
  };
  std::vector<Arr> arrs;
  // fields and media: x halo of xpad columns, of which the R next to the
  // grid are its ghost points
  const long flo = h->xpad - h->R, fhi = h->xpad + h->nx + h->R;
  for (int q = 0; q < 9; ++q)
    for (void *p : {h->field[q], h->field2[q]})
      if (p) arrs.push_back({p, h->total, h->ld, flo, fhi});
  for (void *p : h->media)
    if (p) arrs.push_back({p, h->total, h->ld, flo, fhi});
  // y- and z-slab memory variables: rows of mld >= nx columns
  for (int w = 0; w < 18; ++w)
    if (kMemAxis[w] != 0 && h->mem_elems[w])
      for (void *p : {h->mem[w], h->mem2[w]})
        if (p) arrs.push_back({p, (long)h->mem_elems[w], h->mld, 0, h->nx});
  long long guard = 0;
  for (const Arr &a : arrs) guard += (a.n / a.ld) * (a.ld - (a.hi - a.lo));
  HIPCHK(h, hipMemsetAsync(h->scratch_max, 0, sizeof(unsigned long long), h->stream));
  for (const Arr &a : arrs) {
    const int blocks = (int)std::min<long>(4096, (a.n + 255) / 256);
    if (h->esize == 4)
      hipLaunchKernelGGL(k_guard<float>, dim3(blocks), dim3(256), 0, h->stream, (float *)a.p,
                         a.n, a.ld, a.lo, a.hi, op, h->scratch_max);
    else
      hipLaunchKernelGGL(k_guard<double>, dim3(blocks), dim3(256), 0, h->stream, (double *)a.p,
                         a.n, a.ld, a.lo, a.hi, op, h->scratch_max);
    HIPCHK(h, hipGetLastError());
  }
  unsigned long long changed = 0;
  HIPCHK(h, hipMemcpyAsync(&changed, h->scratch_max, sizeof(changed), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *count = op ? guard : (long long)changed;
  return CPML_OK;
}

int cpml_gpu_set_host_layout(cpml_solver *h, int lo, int hi) {
  if (!h) return CPML_ERR_ARG;
  if (lo < 0 || hi < 0 || lo > 8 || hi > 8)
    return fail(h, CPML_ERR_ARG, "host ghost widths must be in 0..8");
  h->hlo = lo;
  h->hhi = hi;
  return CPML_OK;
}

int cpml_gpu_check_grid_array(cpml_solver *h, int kind, int rank, const int64_t *shape) {
  if (!h) return CPML_ERR_ARG;
  int rc = cpml_gpu_check_kind(h, kind);
  if (rc) return rc;
  if (rank < 1 || !shape) return fail(h, CPML_ERR_ARG, "grid array: rank must be >= 1");
  const int nd = h->ndim;
  const int64_t want[3] = {(int64_t)host_x(h), (int64_t)host_y(h),
                           (int64_t)h->nz + h->hlo + h->hhi};
  int64_t total = 1, expect = 1;
  for (int d = 0; d < rank; ++d) total *= shape[d];
  for (int d = 0; d < nd; ++d) expect *= want[d];
  bool ok = total == expect;
  if (rank == nd)
    for (int d = 0; d < nd; ++d) ok = ok && shape[d] == want[d];
  if (!ok) {
    std::string got, exp;
    for (int d = 0; d < rank; ++d) got += (d ? "," : "") + std::to_string(shape[d]);
    for (int d = 0; d < nd; ++d) exp += (d ? "," : "") + std::to_string(want[d]);
    return fail(h, CPML_ERR_ARG, "grid array of shape (" + got + ") given where (" + exp +
                                     ") is expected (grid plus host ghost cells " +
                                     std::to_string(h->hlo) + " below, " +
                                     std::to_string(h->hhi) + " above)");
  }
  return CPML_OK;
}

int cpml_gpu_grid(cpml_solver *h, int *NX, int *NY, int *NZ) {
  if (!h || !NX || !NY || !NZ) return CPML_ERR_ARG;
  *NX = h->nx;
  *NY = h->ny;
  *NZ = h->nz;
  return CPML_OK;
}

int cpml_gpu_set_memory(cpml_solver *h, int which, const void *in) {
  if (!h) return CPML_ERR_ARG;
  if (!mem_used(h->ndim, which) || !in) return fail(h, CPML_ERR_ARG, "bad memory selector");
  if (!h->have_pml) return fail(h, CPML_ERR_STATE, "set_pml first");
  int rc = set_device(h);
  if (rc) return rc;
  if (!h->mem_elems[which]) return CPML_OK;  // no slab along this axis
  const size_t es = h->esize;
  // start from the device copy, so that the row padding keeps its contents
  std::vector<char> tmp(h->mem_elems[which] * es, 0);
  HIPCHK(h, hipMemcpy(tmp.data(), mbuf(h, which), tmp.size(), hipMemcpyDeviceToHost));
  const char *src = static_cast<const char *>(in);
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  // the slab points of the full-grid array (the rest is the reference's
  // exact zero outside the slab and is not stored)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long dst = -1;
        switch (kMemAxis[which]) {
          case 0:
            if (h->sx.in(i)) dst = h->sx.s(i) + (long)h->sx.w * (j + (long)ny * k);
            break;
          case 1:
            if (h->sy.in(j)) dst = i + h->mld * (h->sy.s(j) + (long)h->sy.w * k);
            break;
          default:
            if (h->sz.in(k)) dst = i + h->mld * (j + (long)ny * h->sz.s(k));
            break;
        }
        if (dst >= 0) std::memcpy(tmp.data() + dst * es, src + host_at(h, i, j, k) * es, es);
      }
  HIPCHK(h, hipMemcpy(mbuf(h, which), tmp.data(), tmp.size(), hipMemcpyHostToDevice));
  return CPML_OK;
}

int cpml_gpu_color_image(cpml_solver *h, int field_number, unsigned char *rgb) {
  if (!h || !rgb) return CPML_ERR_ARG;
  if (field_number != 1 && field_number != 2)
    return fail(h, CPML_ERR_ARG, "field_number must be 1 (Vx) or 2 (Vy)");
  int rc = set_device(h);
  if (rc) return rc;
  // 2D4:729-731 images the ghost-padded arrays (3D-visco:1494 the interior)
  const int gh = h->ndim == 2 && h->order == 4 ? 1 : 0;
  const size_t bytes = (size_t)(h->nx + 2 * gh) * (h->ny + 2 * gh) * 3;
  if (!h->img && (rc = dev_alloc(h, (void **)&h->img, bytes))) return rc;
  cpml::Image g{};
  g.ghost = gh;
  g.isrc = h->isrc + 1;
  g.jsrc = h->jsrc + 1;
  g.npml = h->npml;
  // a layer is drawn where the profiles absorb (USE_PML_XMIN ... 2D2:206-209)
  g.use[0] = h->sx.wl > 0;
  g.use[1] = h->sx.wr > 0;
  g.use[2] = h->sy.wl > 0;
  g.use[3] = h->sy.wr > 0;
  g.ps = h->ps;
  // 3D: plane k = NZ/2, the top plane of the reference's cut-plane rank
  // (3D:346 rank_cut_plane = nb_procs/2 - 1, images of vx(:,:,NZ_LOCAL) 3D:1231-1234)
  const long plane = h->ndim == 3 ? (long)(h->nz / 2 - 1) * h->ps : 0;
  unsigned long long *amax = h->scratch_max + 1;
  HIPCHK(h, hipMemsetAsync(amax, 0, sizeof(unsigned long long), h->stream));
  const int q = field_number == 1 ? CPML_VX : CPML_VY;
  if (h->prec == 4)
    HIPCHK(h, cpml::launch_color_image<float>(F<float>(h, q) + plane, h->nx, h->ny, h->ld, amax,
                                              g, h->recidx, h->nrec, h->img, h->stream));
  else
    HIPCHK(h, cpml::launch_color_image<double>(F<double>(h, q) + plane, h->nx, h->ny, h->ld,
                                               amax, g, h->recidx, h->nrec, h->img, h->stream));
  HIPCHK(h, hipMemcpyAsync(rgb, h->img, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CPML_OK;
}

int cpml_pnm_write(const char *path, int nx, int ny, const unsigned char *rgb) {
  if (!path || !rgb || nx < 1 || ny < 1) return CPML_ERR_ARG;
  FILE *fp = std::fopen(path, "w");
  if (!fp) return CPML_ERR_ARG;
  // the reference's text (2D2:900-903, 989): 'P3', list-directed NX NY and
  // '255' (one leading blank each, as amdflang writes them), then one
  // (i3,' ',i3,' ',i3) line per pixel
  std::fprintf(fp, "P3\n %d %d\n 255\n", nx, ny);
  const size_t n = (size_t)nx * ny;
  for (size_t p = 0; p < n; ++p)
    std::fprintf(fp, "%3d %3d %3d\n", rgb[3 * p], rgb[3 * p + 1], rgb[3 * p + 2]);
  const bool ok = std::ferror(fp) == 0;
  return std::fclose(fp) == 0 && ok ? CPML_OK : CPML_ERR_ARG;
}

int cpml_su_write(const char *path, int ntr, int ns, double dt_value, const void *traces,
                  int prec) {
  if (!path || !traces || ntr < 1 || ns < 1 || ns > 32767 || (prec != 4 && prec != 8))
    return CPML_ERR_ARG;
  const double dtw = std::trunc(dt_value);
  if (!(dtw >= 0 && dtw <= 32767)) return CPML_ERR_ARG;
  short head[120] = {};
  head[57] = (short)ns;   // head(58) = nstep (ADE8:992) or NY (ADE8:938)
  head[58] = (short)dtw;  // head(59) = deltat*1e6 (ADE8:993) or DELTAY*1E3 (ADE8:939)
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return CPML_ERR_ARG;
  std::vector<float> tr(ns);
  bool ok = true;
  for (int j = 0; j < ntr && ok; ++j) {
    if (prec == 4) {
      std::memcpy(tr.data(), static_cast<const float *>(traces) + (size_t)j * ns,
                  (size_t)ns * sizeof(float));
    } else {
      const double *src = static_cast<const double *>(traces) + (size_t)j * ns;
      for (int k = 0; k < ns; ++k) tr[k] = (float)src[k];  // real(): nearest
    }
    ok = std::fwrite(head, sizeof(head), 1, fp) == 1 &&
         std::fwrite(tr.data(), sizeof(float), ns, fp) == (size_t)ns;
  }
  return std::fclose(fp) == 0 && ok ? CPML_OK : CPML_ERR_ARG;
}

int cpml_gpu_set_attenuation(cpml_solver *h, int N_SLS, const double *tau_epsilon_nu1,
                             const double *tau_sigma_nu1, const double *tau_epsilon_nu2,
                             const double *tau_sigma_nu2) {
  if (!h) return CPML_ERR_ARG;
  if (h->ndim != 2 || (h->order != 2 && h->order != 4))
    return fail(h, CPML_ERR_ARG, "attenuation is for the 2D second- and fourth-order steps");
  if (N_SLS < 0 || N_SLS > cpml::kMaxSls)
    return fail(h, CPML_ERR_ARG, "N_SLS must be 0.." + std::to_string(cpml::kMaxSls));
  const double *t[4] = {tau_epsilon_nu1, tau_sigma_nu1, tau_epsilon_nu2, tau_sigma_nu2};
  for (int q = 0; q < 4 && N_SLS > 0; ++q) {
    if (!t[q]) return fail(h, CPML_ERR_ARG, "null tau array");
    for (int l = 0; l < N_SLS; ++l)
      if (!(t[q][l] > 0) || !std::isfinite(t[q][l]))
        return fail(h, CPML_ERR_ARG, "tau values must be finite and > 0");
  }
  int rc = set_device(h);
  if (rc) return rc;
  for (int k = 0; k < 3; ++k) {
    dev_free(h, h->re[k]);
    h->re[k] = nullptr;
  }
  h->nsls = N_SLS;
  for (int q = 0; q < 4; ++q)
    for (int l = 0; l < N_SLS; ++l) h->tau[q][l] = t[q][l];
  const size_t bytes = (size_t)N_SLS * h->nx * h->ny * h->esize;
  for (int k = 0; k < 3 && N_SLS > 0; ++k) {
    if ((rc = dev_alloc(h, &h->re[k], bytes))) return rc;
    HIPCHK(h, hipMemsetAsync(h->re[k], 0, bytes, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->graph_dirty = true;
  return CPML_OK;
}

static int ensure_second_state(cpml_solver *h);

int cpml_gpu_set_ade_rk4(cpml_solver *h, const void *a_x, const void *b_x,
                         const void *a_x_half, const void *b_x_half, const void *a_y,
                         const void *b_y, const void *a_y_half, const void *b_y_half,
                         const void *force_x4, const void *force_y4) {
  if (!h) return CPML_ERR_ARG;
  if (h->ndim != 2 || h->order != 8)
    return fail(h, CPML_ERR_ARG, "the ADE-PML RK4 step is for 2D order-8 handles");
  if (!h->have_pml) return fail(h, CPML_ERR_STATE, "call set_pml first (K profiles)");
  const void *tab[8] = {a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, b_y_half};
  for (auto t : tab)
    if (!t) return fail(h, CPML_ERR_ARG, "null ADE coefficient table");
  if (!force_x4 || !force_y4) return fail(h, CPML_ERR_ARG, "null stage force table");
  int rc = set_device(h);
  if (rc) return rc;
  const size_t fb = (size_t)h->total * h->esize;
  if ((rc = ensure_second_state(h))) return rc;  // X ping-pongs between the two
  for (int q = 0; q < 10; ++q)
    if (!h->ade_d[q]) {
      if ((rc = dev_alloc(h, &h->ade_d[q], fb))) return rc;
    }
  for (int q = 0; q < 8; ++q)
    if (!h->ade_mem[q]) {
      if ((rc = dev_alloc(h, &h->ade_mem[q], fb))) return rc;
    }
  for (int q = 0; q < 10; ++q) HIPCHK(h, hipMemsetAsync(h->ade_d[q], 0, fb, h->stream));
  for (int q = 0; q < 8; ++q) HIPCHK(h, hipMemsetAsync(h->ade_mem[q], 0, fb, h->stream));
  // tables: Fortran a(4, N) (stage fastest) stored as [stage][i]
  for (int q = 0; q < 8; ++q) {
    const int n = q < 4 ? h->nx : h->ny;
    const size_t bytes = (size_t)4 * n * h->esize;
    if (!h->ade_tab[q] && (rc = dev_alloc(h, &h->ade_tab[q], bytes))) return rc;
    std::vector<char> t(bytes);
    const char *src = static_cast<const char *>(tab[q]);
    for (int i = 0; i < n; ++i)
      for (int s = 0; s < 4; ++s)
        std::memcpy(t.data() + ((size_t)s * n + i) * h->esize,
                    src + ((size_t)i * 4 + s) * h->esize, h->esize);
    HIPCHK(h, hipMemcpy(h->ade_tab[q], t.data(), bytes, hipMemcpyHostToDevice));
  }
  const size_t f4 = (size_t)4 * h->nstep * h->esize;
  const void *fsrc[2] = {force_x4, force_y4};
  for (int q = 0; q < 2; ++q) {
    if (!h->ade_f4[q] && (rc = dev_alloc(h, &h->ade_f4[q], f4))) return rc;
    HIPCHK(h, hipMemcpy(h->ade_f4[q], fsrc[q], f4, hipMemcpyHostToDevice));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ade = true;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_get_attenuation_memory(cpml_solver *h, int which, void *out) {
  if (!h || !out) return CPML_ERR_ARG;
  if (h->nsls == 0) return fail(h, CPML_ERR_STATE, "attenuation is off");
  if (which < 0 || which > 2) return fail(h, CPML_ERR_ARG, "which must be 0, 1 or 2");
  int rc = set_device(h);
  if (rc) return rc;
  HIPCHK(h, hipMemcpy(out, h->re[which], (size_t)h->nsls * h->nx * h->ny * h->esize,
                      hipMemcpyDeviceToHost));
  return CPML_OK;
}

int cpml_gpu_get_pressure(cpml_solver *h, void *sispressure) {
  if (!h || !sispressure) return CPML_ERR_ARG;
  if (h->ndim != 2 || h->nsls == 0)
    return fail(h, CPML_ERR_STATE, "pressure is recorded by the 2D viscoelastic step only");
  int rc = set_device(h);
  if (rc) return rc;
  if (h->nrec == 0) return CPML_OK;
  const size_t n = (size_t)h->nstep * h->nrec;
  std::vector<char> tmp(n * h->esize);
  HIPCHK(h, hipMemcpy(tmp.data(), h->seis[2], tmp.size(), hipMemcpyDeviceToHost));
  char *o = static_cast<char *>(sispressure);
  for (int it = 0; it < h->nstep; ++it)
    for (int r = 0; r < h->nrec; ++r)
      std::memcpy(o + ((size_t)it + (size_t)r * h->nstep) * h->esize,
                  tmp.data() + ((size_t)it * h->nrec + r) * h->esize, h->esize);
  return CPML_OK;
}

int cpml_gpu_velocity_norm(cpml_solver *h, double *vmax) {
  if (!h || !vmax) return CPML_ERR_ARG;
  int rc = set_device(h);
  if (rc) return rc;
  return velocity_norm(h, vmax);
}

int cpml_gpu_set_graph_steps(cpml_solver *h, int steps) {
  if (!h) return CPML_ERR_ARG;
  if (steps < 0 || steps > 4096) return fail(h, CPML_ERR_ARG, "graph steps out of range");
  h->graph_steps = steps;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_kchunk(cpml_solver *h, int kchunk) {
  if (!h) return CPML_ERR_ARG;
  if (kchunk < 1) return fail(h, CPML_ERR_ARG, "kchunk must be >= 1");
  if (!chunk_fits(h, kchunk))
    return fail(h, CPML_ERR_ARG,
                "kchunk too large: a chunk's planes (with their halo planes) must stay below "
                "4 GiB, the reach of the kernels' 32-bit buffer offsets");
  h->kchunk = kchunk;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_boundary(cpml_solver *h, int mode, cpml_abc_net *net) {
  if (!h) return CPML_ERR_ARG;
  int rc = set_device(h);
  if (rc) return rc;
  if (mode == CPML_BOUNDARY_CPML) {
    h->abc = nullptr;
    dev_free(h, h->abc_corr);
    h->graph_dirty = true;
    return CPML_OK;
  }
  if (mode != CPML_BOUNDARY_LEARNED || !net)
    return fail(h, CPML_ERR_ARG, "mode must be CPML_BOUNDARY_CPML or CPML_BOUNDARY_LEARNED "
                                 "with a corrector");
  if (!h->have_pml) return fail(h, CPML_ERR_STATE, "call set_pml first");
  if (h->ndim == 3 && h->order != 2)
    return fail(h, CPML_ERR_ARG, "the learned boundary is for the 3D second-order step");
  std::string m;
  if ((rc = cpml::abc_prepare(net, view_of(h), &m))) return fail(h, rc, m);
  dev_free(h, h->abc_corr);
  if ((rc = dev_alloc(h, &h->abc_corr,
                      (size_t)cpml::abc_corr_floats(view_of(h)) * sizeof(float))))
    return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->abc = net;
  h->graph_dirty = true;
  return CPML_OK;
}

// Second state buffer (fields and memory variables) of the ping-pong steps.
static int ensure_second_state(cpml_solver *h) {
  if (h->field2[0]) return CPML_OK;
  const int nf = h->ndim == 3 ? 9 : 5;
  size_t free_b = 0, tot_b = 0;
  if (hipMemGetInfo(&free_b, &tot_b) == hipSuccess &&
      (double)h->total * h->esize * nf > 0.95 * (double)free_b)
    return fail(h, CPML_ERR_OOM, "not enough device memory for the second state buffer");
  bool want[9];
  for (int q = 0; q < 9; ++q) want[q] = h->field[q] != nullptr;
  int rc = dev_alloc_group(h, h->field2, 9, (size_t)h->total * h->esize, want);
  if (rc) return rc;
  for (int w = 0; w < 18 && h->have_pml; ++w)
    if (mem_used(h->ndim, w) && !h->mem2[w] &&
        (rc = dev_alloc(h, &h->mem2[w], h->mem_elems[w] * h->esize)))
      return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CPML_OK;
}

// Bring the state back to the first buffer.
static int restore_first_state(cpml_solver *h) {
  if (!h->cur) return CPML_OK;
  for (int q = 0; q < 9; ++q)
    if (h->field[q])
      HIPCHK(h, hipMemcpyAsync(h->field[q], h->field2[q], (size_t)h->total * h->esize,
                               hipMemcpyDeviceToDevice, h->stream));
  for (int w = 0; w < 18; ++w)
    if (h->mem[w] && h->mem_elems[w])
      HIPCHK(h, hipMemcpyAsync(h->mem[w], h->mem2[w], h->mem_elems[w] * h->esize,
                               hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->cur = 0;
  return CPML_OK;
}

int cpml_gpu_set_fused(cpml_solver *h, int on) {
  if (!h) return CPML_ERR_ARG;
  if (on && (h->ndim != 3 || h->order != 2))
    return fail(h, CPML_ERR_ARG, "the fused step is 3D second order only");
  int rc = set_device(h);
  if (rc) return rc;
  if (on && (rc = ensure_second_state(h))) return rc;
  if (!on && (rc = restore_first_state(h))) return rc;
  h->fused = on != 0;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_time_block(cpml_solver *h, int steps, int strip_width, int chunk_rows) {
  if (!h) return CPML_ERR_ARG;
  if (steps < 0) return fail(h, CPML_ERR_ARG, "steps must be >= 0");
  if (steps > 0 && h->ndim == 3 && h->order != 2)
    return fail(h, CPML_ERR_ARG, "3D time blocking is built for the second-order step");
  const int nmax = h->ndim == 3 ? cpml::tb3d_max_nt((int)h->esize)
                                : cpml::tb2d_max_nt(h->order, (int)h->esize);
  if (steps > nmax)
    return fail(h, CPML_ERR_ARG, "at most " + std::to_string(nmax) +
                                     " steps per launch for this order and precision");
  const int r = h->order / 2;
  const int g = steps * (2 * r - 1);
  if (strip_width < 0 || (steps > 0 && strip_width > 64 - 2 * g))
    return fail(h, CPML_ERR_ARG, "strip width must be 0 (default) or 1.." +
                                     std::to_string(64 - 2 * g) + " for this order and step count");
  if (chunk_rows < 0) return fail(h, CPML_ERR_ARG, "chunk rows must be >= 0");
  const int jc = chunk_rows > 0 ? chunk_rows : (h->ndim == 3 ? 64 : 40);
  if (!chunk_fits(h, jc))
    return fail(h, CPML_ERR_ARG, "chunk too deep for the kernels' 32-bit buffer offsets");
  int rc = set_device(h);
  if (rc) return rc;
  if (steps > 0 && (rc = ensure_second_state(h))) return rc;
  if (steps == 0 && (rc = restore_first_state(h))) return rc;
  if (steps > 0 && !h->side) {
    HIPCHK(h, hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    HIPCHK(h, hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  }
  h->tb = steps;
  // default strip: every lane the ring leaves (2D, measured fastest: 52 columns
  // at order 4 and 2 steps gave 79.4 Gcell/s against 75.7 for 48); 3D rounds
  // it down to a multiple of 8 (56 and 48 measured alike)
  const int widest = 64 - 2 * g;
  h->tb_ow = strip_width > 0 ? strip_width
                             : (h->ndim == 3 ? std::max(8, widest / 8 * 8) : widest);
  if (h->tb_ow > widest) h->tb_ow = widest;
  h->tb_jc = jc;
  h->tb_rec_ow = -1;
  h->graph_dirty = true;
  return CPML_OK;
}

// Steps per launch and chunk depth chosen by timing the handle itself: the
// time-blocked kernels give one wavefront (2D) or workgroup (3D) per strip x
// chunk item, so a small grid can leave too few of them to hide latency,
// and single steps win there.  Every field, memory variable, seismogram
// sample and timer the trial runs touch is saved first and put back after.
int cpml_gpu_autotune(cpml_solver *h, int trial_steps, int *steps_per_launch, int *chunk,
                      double *best_ms) {
  if (!h) return CPML_ERR_ARG;
  if (!h->have_medium || !h->have_pml)
    return fail(h, CPML_ERR_STATE, "set_medium and set_pml must be called before autotune");
  if (trial_steps < 1 || trial_steps > h->nstep)
    return fail(h, CPML_ERR_ARG, "trial_steps must be in 1..NSTEP");
  int rc = set_device(h);
  if (rc) return rc;
  struct Cand {
    int nt, jc;
  };
  std::vector<Cand> cands{{0, 0}};
  const bool blockable = !h->ade && h->nsls == 0 && !h->abc && !h->fused &&
                         (h->ndim == 2 || h->order == 2);
  if (blockable) {
    const int nmax = h->ndim == 3 ? cpml::tb3d_max_nt((int)h->esize)
                                  : cpml::tb2d_max_nt(h->order, (int)h->esize);
    const std::vector<int> depths =
        h->ndim == 3 ? std::vector<int>{32, 64} : std::vector<int>{16, 24, 40, 64};
    for (int nt = 1; nt <= nmax; ++nt)
      if (h->ndim == 2 || h->cadence == 0 || h->cadence % nt == 0)  // 3D energy: block ends
        for (int jc : depths)
          if (chunk_fits(h, jc)) cands.push_back({nt, jc});
  }
  auto report = [&](int nt, int jc, double ms) {
    if (steps_per_launch) *steps_per_launch = nt;
    if (chunk) *chunk = nt > 0 ? jc : 0;
    if (best_ms) *best_ms = ms;
  };
  if (cands.size() == 1) {  // nothing to choose: the step has no time-blocked form
    report(h->tb, h->tb_jc, -1.0);
    return CPML_OK;
  }

  // save the state (where `cur` says it is) and what the trials overwrite
  struct Saved {
    void *dev;
    size_t bytes;
    int kind, idx;  // 0 field, 1 memory variable, 2 seismogram, 3 energy curve
  };
  std::vector<Saved> saved;
  auto release = [&]() {
    for (Saved &s : saved) (void)hipFree(s.dev);
    saved.clear();
  };
  auto where = [&](const Saved &s) -> void * {
    if (s.kind == 3) return s.idx ? (void *)h->epot : (void *)h->ekin;
    return s.kind == 0 ? fbuf(h, s.idx) : s.kind == 1 ? mbuf(h, s.idx) : h->seis[s.idx];
  };
  std::vector<Saved> want;
  for (int q = 0; q < 9; ++q)
    if (fbuf(h, q)) want.push_back({nullptr, (size_t)h->total * h->esize, 0, q});
  for (int w = 0; w < 18; ++w)
    if (mbuf(h, w) && h->mem_elems[w]) want.push_back({nullptr, h->mem_elems[w] * h->esize, 1, w});
  if (h->nrec > 0)
    for (int q = 0; q < 3; ++q)
      want.push_back({nullptr, (size_t)h->nstep * h->nrec * h->esize, 2, q});
  if (h->cadence > 0)
    for (int q = 0; q < 2; ++q) want.push_back({nullptr, (size_t)h->nstep * sizeof(double), 3, q});
  for (Saved s : want) {
    if (hipMalloc(&s.dev, s.bytes) != hipSuccess) {
      release();
      return fail(h, CPML_ERR_OOM, "not enough device memory to save the state for autotune");
    }
    saved.push_back(s);
    if (hipMemcpyAsync(s.dev, where(s), s.bytes, hipMemcpyDeviceToDevice, h->stream) !=
        hipSuccess) {
      release();
      return fail(h, CPML_ERR_HIP, "autotune: saving the state failed");
    }
  }
  const double ktime[2] = {h->ktime[0], h->ktime[1]};
  const int kcount[2] = {h->kcount[0], h->kcount[1]};
  const double last_ms = h->last_ms, last_vnorm = h->last_vnorm;
  const int prev_tb = h->tb, prev_ow = h->tb_ow, prev_jc = h->tb_jc;

  Cand best_c{prev_tb, prev_jc};
  double best = -1;
  for (const Cand &c : cands) {
    const int u = c.nt > 0 ? c.nt : 1;
    const int n = std::max(u, trial_steps / u * u);  // whole launches only
    if (n > h->nstep) continue;
    if ((rc = cpml_gpu_set_time_block(h, c.nt, 0, c.jc))) break;
    if ((rc = cpml_gpu_run(h, 1, n))) break;  // warm: graphs, first launch of a shape
    if ((rc = cpml_gpu_run(h, 1, n))) break;
    const double per_step = h->last_ms / n;
    if (best < 0 || per_step < best) {
      best = per_step;
      best_c = c;
    }
  }
  const std::string err = rc ? h->err : std::string();
  // the choice (or, after a failed trial, the previous setting), then the state
  int rc2 = rc ? cpml_gpu_set_time_block(h, prev_tb, prev_tb > 0 ? prev_ow : 0, prev_jc)
               : cpml_gpu_set_time_block(h, best_c.nt, 0, best_c.nt > 0 ? best_c.jc : 0);
  for (const Saved &s : saved)
    if (hipMemcpyAsync(where(s), s.dev, s.bytes, hipMemcpyDeviceToDevice, h->stream) !=
        hipSuccess)
      rc2 = rc2 ? rc2 : CPML_ERR_HIP;
  if (hipStreamSynchronize(h->stream) != hipSuccess) rc2 = rc2 ? rc2 : CPML_ERR_HIP;
  release();
  h->ktime[0] = ktime[0], h->ktime[1] = ktime[1];
  h->kcount[0] = kcount[0], h->kcount[1] = kcount[1];
  h->last_ms = last_ms, h->last_vnorm = last_vnorm;
  if (rc) return fail(h, rc, "autotune trial: " + err);
  if (rc2) return fail(h, rc2, "autotune: restoring the state failed");
  report(best_c.nt, best_c.jc, best * trial_steps);
  return CPML_OK;
}

int cpml_gpu_set_tile(cpml_solver *h, int tile_y) {
  if (!h) return CPML_ERR_ARG;
  const bool ok3 = tile_y == 4 || tile_y == 8 || tile_y == 16;
  if (!(ok3 || (h->ndim == 2 && (tile_y == 32 || tile_y == 64))))
    return fail(h, CPML_ERR_ARG, "tile rows must be 4, 8 or 16 (2D also 32 or 64)");
  h->tile_y = tile_y;
  h->graph_dirty = true;
  return CPML_OK;
}

int cpml_gpu_set_kernel_timing(cpml_solver *h, int on) {
  if (!h) return CPML_ERR_ARG;
  h->ktiming = on != 0;
  h->ktime[0] = h->ktime[1] = 0;
  h->kcount[0] = h->kcount[1] = 0;
  return CPML_OK;
}

int cpml_gpu_kernel_time(cpml_solver *h, int kernel, int *launches, double *total_ms) {
  if (!h || kernel < 0 || kernel > 1) return CPML_ERR_ARG;
  if (launches) *launches = h->kcount[kernel];
  if (total_ms) *total_ms = h->ktime[kernel];
  return CPML_OK;
}

int cpml_gpu_elapsed_ms(cpml_solver *h, double *ms) {
  if (!h || !ms) return CPML_ERR_ARG;
  *ms = h->last_ms;
  return CPML_OK;
}

int cpml_gpu_device_bytes(cpml_solver *h, double *bytes) {
  if (!h || !bytes) return CPML_ERR_ARG;
  *bytes = h->bytes;
  return CPML_OK;
}

int cpml_gpu_slab_widths(cpml_solver *h, int *w) {
  if (!h || !w) return CPML_ERR_ARG;
  const Slab *s[3] = {&h->sx, &h->sy, &h->sz};
  for (int d = 0; d < 3; ++d) {
    w[2 * d] = s[d]->wl;
    w[2 * d + 1] = s[d]->wr;
  }
  return CPML_OK;
}

int cpml_gpu_last_error(cpml_solver *h, char *buf, int len) {
  if (!buf || len <= 0) return CPML_ERR_ARG;
  const std::string &m = h ? h->err : g_init_error;
  const size_t n = std::min(m.size(), (size_t)len - 1);
  std::memcpy(buf, m.data(), n);
  buf[n] = '\0';
  return CPML_OK;
}

int cpml_pml_profile(int prec, int n, double delta, double deltat, int npoints_pml, double cp,
                     double f0, int use_min, int use_max, int origin_max_uses_n,
                     int clamp_alpha, void *a, void *b, void *K, void *a_half, void *b_half,
                     void *K_half) {
  if ((prec != 4 && prec != 8) || n < 1 || npoints_pml < 1 || !(delta > 0) || !a || !b || !K ||
      !a_half || !b_half || !K_half)
    return fail(nullptr, CPML_ERR_ARG, "cpml_pml_profile: bad argument");
  if (prec == 4)
    return profile_impl<float>(n, delta, deltat, npoints_pml, cp, f0, use_min, use_max,
                               origin_max_uses_n, clamp_alpha, (float *)a, (float *)b,
                               (float *)K, (float *)a_half, (float *)b_half, (float *)K_half);
  return profile_impl<double>(n, delta, deltat, npoints_pml, cp, f0, use_min, use_max,
                              origin_max_uses_n, clamp_alpha, (double *)a, (double *)b,
                              (double *)K, (double *)a_half, (double *)b_half, (double *)K_half);
}

int cpml_source_time_function(int prec, int nstep, double deltat, double f0, double t0,
                              double factor, double angle_force, int kind, void *force_x,
                              void *force_y) {
  if ((prec != 4 && prec != 8) || nstep < 1 || kind < 0 || kind > 2 || !force_x || !force_y)
    return fail(nullptr, CPML_ERR_ARG, "cpml_source_time_function: bad argument");
  if (prec == 4)
    return stf_impl<float>(nstep, deltat, f0, t0, factor, angle_force, kind, (float *)force_x,
                           (float *)force_y);
  return stf_impl<double>(nstep, deltat, f0, t0, factor, angle_force, kind, (double *)force_x,
                          (double *)force_y);
}

int cpml_nearest_receivers(int prec, int NX, int NY, double DELTAX, double DELTAY, int NREC,
                           double xdeb, double ydeb, double xfin, double yfin, int *ix_rec,
                           int *iy_rec) {
  if ((prec != 4 && prec != 8) || NX < 1 || NY < 1 || NREC < 0 || (NREC && (!ix_rec || !iy_rec)))
    return fail(nullptr, CPML_ERR_ARG, "cpml_nearest_receivers: bad argument");
  if (prec == 4)
    return receivers_impl<float>(NX, NY, DELTAX, DELTAY, NREC, xdeb, ydeb, xfin, yfin, ix_rec,
                                 iy_rec);
  return receivers_impl<double>(NX, NY, DELTAX, DELTAY, NREC, xdeb, ydeb, xfin, yfin, ix_rec,
                                iy_rec);
}

int cpml_courant_number(int ndim, double cp, double deltat, double deltax, double deltay,
                        double deltaz, double *courant) {
  if (!courant || !(deltax > 0) || !(deltay > 0) || (ndim == 3 && !(deltaz > 0)))
    return fail(nullptr, CPML_ERR_ARG, "cpml_courant_number: bad argument");
  double s = 1.0 / (deltax * deltax) + 1.0 / (deltay * deltay);
  if (ndim == 3) s += 1.0 / (deltaz * deltaz);
  *courant = cp * deltat * std::sqrt(s);
  return CPML_OK;
}

}  // extern "C"

void cpml::set_thread_error(const std::string &msg) { g_init_error = msg; }

int cpml::solver_view(cpml_solver *h, cpml::SolverView *v) {
  if (!h || !v) return CPML_ERR_ARG;
  *v = view_of(h);
  return CPML_OK;
}
