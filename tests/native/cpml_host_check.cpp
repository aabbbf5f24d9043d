// Host-layer exercise of the C ABI, built against an AddressSanitizer build
// of the library's host code (the device code is the ordinary gfx950 build).
//
//   cpml_host_check cpu <dir>   argument validation, NULL handles, error
//                               messages, the host set-up helpers and the
//                               PNM/SU writers (files under <dir>): needs no GPU
//   cpml_host_check gpu <dir>   every set-up, run and read-back path of the
//                               solver (2D orders 2/4/8, 3D orders 2 and 4,
//                               single-step, fused and time-blocked, host
//                               ghost layouts, checkpoint get/set), of the
//                               batch (time blocking, energy, stability,
//                               autotune cache) and of the viscoelastic and
//                               ADE-RK4 steps, on small grids
//
// Every call must return the status the check names; the program prints one
// line per failed check, "HOST CHECK OK" when all pass, and exits non-zero
// otherwise (ASan aborts on the first memory error it finds).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cpml_gpu.h"

static int g_fail = 0, g_checks = 0;

#define EXPECT(call, want)                                                           \
  do {                                                                               \
    const int rc_ = (call);                                                          \
    ++g_checks;                                                                      \
    if (rc_ != (want)) {                                                             \
      ++g_fail;                                                                      \
      std::printf("FAILED %s:%d: %s returned %d, expected %d\n", __FILE__, __LINE__, \
                  #call, rc_, (int)(want));                                          \
    }                                                                                \
  } while (0)

// ---- needs no GPU -------------------------------------------------------

static void check_messages() {
  char buf[512];
  for (int len : {1, 2, 7, 512}) EXPECT(cpml_gpu_last_error(nullptr, buf, len), CPML_OK);
  EXPECT(cpml_gpu_last_error(nullptr, buf, 0), CPML_ERR_ARG);
  EXPECT(cpml_gpu_last_error(nullptr, nullptr, 8), CPML_ERR_ARG);
}

static void check_init_validation() {
  cpml_solver *h = nullptr;
  EXPECT(cpml_gpu_init(nullptr, 2, 2, 8, 10, 10, 1, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 4, 2, 8, 10, 10, 1, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 2, 6, 8, 10, 10, 1, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 3, 8, 8, 10, 10, 10, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 2, 2, 3, 10, 10, 1, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 2, 2, 8, 1, 10, 1, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_STENCIL);
  EXPECT(cpml_gpu_init(&h, 2, 2, 8, 10, 10, 1, -1, 1, 1, 1e-3, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 2, 2, 8, 10, 10, 1, 1, 1, 1, 0.0, 2, 10), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 2, 2, 8, 10, 10, 1, 1, 1, 1, 1e-3, 2, 0), CPML_ERR_ARG);
  EXPECT(cpml_gpu_init(&h, 2, 2, 8, 10, 10, 1, 1, 1, 1, 1e-3, 6, 10), CPML_ERR_STENCIL);
  EXPECT(cpml_gpu_init(&h, 3, 2, 4, 4000, 4000, 4000, 1, 1, 1, 1e-3, 2, 10), CPML_ERR_OOM);
  if (h) ++g_fail, std::printf("FAILED: handle set by a refused init\n");

  const int nx[2] = {64, 1}, ny[2] = {64, 64}, np[2] = {4, 4};
  const double d[2] = {1, 1}, dt[2] = {1e-3, 1e-3};
  cpml_batch *b = nullptr;
  EXPECT(cpml_batch_init(&b, 2, 4, 2, nx, ny, d, d, dt, np, 10), CPML_ERR_STENCIL);
  EXPECT(cpml_batch_init(&b, 3, 4, 2, nx, ny, d, d, dt, np, 10), CPML_ERR_ARG);
  EXPECT(cpml_batch_init(&b, 2, 4, 0, nx, ny, d, d, dt, np, 10), CPML_ERR_ARG);
  EXPECT(cpml_batch_init(&b, 2, 4, 2, nullptr, ny, d, d, dt, np, 10), CPML_ERR_ARG);
}

static void check_null_handles() {
  double v = 0;
  int i = 0;
  long long c = 0;
  int64_t shape[2] = {1, 1};
  EXPECT(cpml_gpu_finalize(nullptr), CPML_OK);  // as free(NULL)
  EXPECT(cpml_gpu_set_medium(nullptr, &v, &v, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_source(nullptr, 1, 1, 1, &v, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_receivers(nullptr, 1, &i, &i, &i), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_energy(nullptr, 1, nullptr, 0), CPML_ERR_ARG);
  EXPECT(cpml_gpu_run(nullptr, 1, 1), CPML_ERR_ARG);
  EXPECT(cpml_gpu_get_field(nullptr, 0, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_field(nullptr, 0, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_get_memory(nullptr, 0, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_memory(nullptr, 0, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_velocity_norm(nullptr, &v), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_time_block(nullptr, 2, 0, 0), CPML_ERR_ARG);
  EXPECT(cpml_gpu_autotune(nullptr, 10, nullptr, nullptr, nullptr), CPML_ERR_ARG);
  EXPECT(cpml_gpu_check_kind(nullptr, 4), CPML_ERR_ARG);
  EXPECT(cpml_gpu_grid(nullptr, &i, &i, &i), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_host_layout(nullptr, 1, 1), CPML_ERR_ARG);
  EXPECT(cpml_gpu_check_grid_array(nullptr, 4, 2, shape), CPML_ERR_ARG);
  EXPECT(cpml_gpu_guard_cells(nullptr, 1, &c), CPML_ERR_ARG);
  EXPECT(cpml_batch_finalize(nullptr), CPML_OK);
  EXPECT(cpml_batch_run(nullptr, 1, 1), CPML_ERR_ARG);
  EXPECT(cpml_batch_check_kind(nullptr, 4), CPML_ERR_ARG);
  EXPECT(cpml_batch_grid(nullptr, 0, &i, &i), CPML_ERR_ARG);
  double dv = 0;
  EXPECT(cpml_velocity_norm_noalloc(3, 2, 8, 8, 1, &dv, &dv, nullptr, &dv, nullptr),
         CPML_ERR_ARG);
  EXPECT(cpml_velocity_norm_noalloc(4, 2, 0, 8, 1, &dv, &dv, nullptr, &dv, nullptr),
         CPML_ERR_ARG);
  EXPECT(cpml_velocity_norm_noalloc(4, 3, 1 << 13, 1 << 13, 1 << 12, &dv, &dv, &dv, &dv,
                                    nullptr),
         CPML_ERR_ARG);
}

template <typename T>
static void check_helpers(int prec, const std::string &dir) {
  for (int n : {2, 3, 11, 101, 1001}) {
    std::vector<T> a(n), b(n), K(n), ah(n), bh(n), Kh(n);
    EXPECT(cpml_pml_profile(prec, n, 10.0, 1e-3, 0, 3000.0, 7.0, 1, 1, 0, 0, a.data(), b.data(),
                            K.data(), ah.data(), bh.data(), Kh.data()),
           CPML_ERR_ARG);
    for (int npml : {1, std::max(1, n / 2)}) {
      for (int flags = 0; flags < 8; ++flags)
        EXPECT(cpml_pml_profile(prec, n, 10.0, 1e-3, npml, 3000.0, 7.0, flags & 1, (flags >> 1) & 1,
                                (flags >> 2) & 1, flags & 1, a.data(), b.data(), K.data(),
                                ah.data(), bh.data(), Kh.data()),
               CPML_OK);
    }
  }
  for (int nstep : {1, 2, 1000}) {
    std::vector<T> fx(nstep), fy(nstep);
    for (int kind = 0; kind < 3; ++kind)
      EXPECT(cpml_source_time_function(prec, nstep, 1e-3, 7.0, 0.17, 1e7, 135.0, kind, fx.data(),
                                       fy.data()),
             CPML_OK);
  }
  for (int nrec : {1, 2, 50}) {
    std::vector<int> ix(nrec), iy(nrec);
    EXPECT(cpml_nearest_receivers(prec, 101, 641, 10, 10, nrec, 100, 2300, 800, 300, ix.data(),
                                  iy.data()),
           CPML_OK);
  }
  double c = 0;
  EXPECT(cpml_courant_number(2, 3300, 2e-3, 10, 10, 10, &c), CPML_OK);
  EXPECT(cpml_courant_number(3, 3300, 2e-3, 10, 10, 10, &c), CPML_OK);
  // writers
  const int nx = 13, ny = 7;
  std::vector<unsigned char> rgb(nx * ny * 3, 200);
  EXPECT(cpml_pnm_write((dir + "/x.pnm").c_str(), nx, ny, rgb.data()), CPML_OK);
  EXPECT(cpml_pnm_write((dir + "/no/such/dir/x.pnm").c_str(), nx, ny, rgb.data()),
         CPML_ERR_ARG);
  std::vector<T> tr(nx * ny, (T)0.25);
  EXPECT(cpml_su_write((dir + "/x.su").c_str(), ny, nx, 2000.0, tr.data(), prec), CPML_OK);
  EXPECT(cpml_su_write((dir + "/y.su").c_str(), ny, nx, 1e9, tr.data(), prec), CPML_ERR_ARG);
  EXPECT(cpml_su_write((dir + "/z.su").c_str(), ny, 0, 10.0, tr.data(), prec), CPML_ERR_ARG);
}

// ---- needs a GPU ------------------------------------------------------

template <typename T>
struct Profiles {
  std::vector<T> v[18];
  Profiles(int prec, const int n[3], int npml) {
    for (int d = 0; d < 3; ++d) {
      for (int k = 0; k < 6; ++k) v[6 * d + k].assign(std::max(n[d], 1), (T)0);
      cpml_pml_profile(prec, n[d], 10.0, 1e-3, npml, 3000.0, 10.0, 1, 1, 0, 0, v[6 * d].data(),
                       v[6 * d + 1].data(), v[6 * d + 2].data(), v[6 * d + 3].data(),
                       v[6 * d + 4].data(), v[6 * d + 5].data());
    }
  }
  const void *p(int q) const { return v[q].data(); }
};

template <typename T>
static void run_solver(int prec, int ndim, int order, int tb, int fused, int lo, int hi) {
  const int n[3] = {ndim == 3 ? 36 : 90, ndim == 3 ? 40 : 70, ndim == 3 ? 34 : 1};
  const int npml = 6, nstep = 30;
  cpml_solver *h = nullptr;
  EXPECT(cpml_gpu_init(&h, ndim, order, prec, n[0], n[1], n[2], 10, 10, 10, 5e-4, npml, nstep),
         CPML_OK);
  if (!h) return;
  EXPECT(cpml_gpu_set_host_layout(h, lo, hi), CPML_OK);
  const size_t hx = n[0] + lo + hi, hy = n[1] + lo + hi, hz = ndim == 3 ? n[2] + lo + hi : 1;
  const size_t cells = hx * hy * hz;
  std::vector<T> lam(cells, (T)8e9), mu(cells, (T)8e9), rho(cells, (T)2500);
  int64_t shape[3] = {(int64_t)hx, (int64_t)hy, (int64_t)hz};
  EXPECT(cpml_gpu_check_grid_array(h, prec, ndim, shape), CPML_OK);
  shape[0] += 1;
  EXPECT(cpml_gpu_check_grid_array(h, prec, ndim, shape), CPML_ERR_ARG);
  EXPECT(cpml_gpu_check_kind(h, 12 - prec), CPML_ERR_ARG);
  EXPECT(cpml_gpu_set_medium(h, lam.data(), mu.data(), rho.data()), CPML_OK);
  Profiles<T> P(prec, n, npml);
  EXPECT(cpml_gpu_set_pml(h, P.p(0), P.p(1), P.p(2), P.p(3), P.p(4), P.p(5), P.p(6), P.p(7),
                          P.p(8), P.p(9), P.p(10), P.p(11), P.p(12), P.p(13), P.p(14), P.p(15),
                          P.p(16), P.p(17)),
         CPML_OK);
  std::vector<T> fx(nstep), fy(nstep);
  cpml_source_time_function(prec, nstep, 5e-4, 20.0, 0.05, 1e7, 45.0, 1, fx.data(), fy.data());
  EXPECT(cpml_gpu_set_source(h, n[0] / 2, n[1] / 2, n[2] / 2 + 1, fx.data(), fy.data()), CPML_OK);
  const int nrec = 5;
  int ix[nrec], iy[nrec], iz[nrec];
  for (int r = 0; r < nrec; ++r) ix[r] = 3 + 7 * r, iy[r] = 5 + 3 * r, iz[r] = 1 + r;
  EXPECT(cpml_gpu_set_receivers(h, nrec, ix, iy, ndim == 3 ? iz : nullptr), CPML_OK);
  // energy on, except for 3D time blocks (which need a cadence that is a
  // multiple of the steps per launch)
  const bool energy = !(tb && ndim == 3);
  EXPECT(cpml_gpu_set_energy(h, energy ? 1 : 0, nullptr, 0), CPML_OK);
  EXPECT(cpml_gpu_set_stability(h, 10, 1e25), CPML_OK);
  if (fused) EXPECT(cpml_gpu_set_fused(h, 1), CPML_OK);
  if (tb) EXPECT(cpml_gpu_set_time_block(h, tb, 0, 0), CPML_OK);
  long long guards = 0;
  EXPECT(cpml_gpu_guard_cells(h, 1, &guards), CPML_OK);
  EXPECT(cpml_gpu_run(h, 1, 13), CPML_OK);
  // the autotuner saves and restores the state around its trials; its trials
  // may allocate the second state buffer, whose guard cells the first
  // poisoning did not reach, so the guards are counted before and poisoned
  // again after it
  {
    EXPECT(cpml_gpu_guard_cells(h, 0, &guards), CPML_OK);
    if (guards) ++g_fail, std::printf("FAILED: %lld guard cells overwritten\n", guards);
    int snt = -1, sjc = -1;
    double sms = 0;
    EXPECT(cpml_gpu_autotune(h, 0, &snt, &sjc, &sms), CPML_ERR_ARG);
    EXPECT(cpml_gpu_autotune(h, 4, &snt, &sjc, &sms), CPML_OK);
    if (snt < 0 || sjc < 0) ++g_fail, std::printf("FAILED: autotune outputs %d %d\n", snt, sjc);
    if (tb) EXPECT(cpml_gpu_set_time_block(h, tb, 0, 0), CPML_OK);  // back to the tested shape
    else EXPECT(cpml_gpu_set_time_block(h, 0, 0, 0), CPML_OK);
    EXPECT(cpml_gpu_guard_cells(h, 1, &guards), CPML_OK);
  }
  // checkpoint every field and memory variable through the host layout
  std::vector<T> buf(cells), back(cells);
  for (int f = 0; f < 9; ++f) {
    const int rc = cpml_gpu_get_field(h, f, buf.data());
    if (rc == CPML_OK) {
      EXPECT(cpml_gpu_set_field(h, f, buf.data()), CPML_OK);
      EXPECT(cpml_gpu_get_field(h, f, back.data()), CPML_OK);
    }
  }
  for (int w = 0; w < 18; ++w) {
    const int rc = cpml_gpu_get_memory(h, w, buf.data());
    if (rc == CPML_OK) EXPECT(cpml_gpu_set_memory(h, w, buf.data()), CPML_OK);
  }
  EXPECT(cpml_gpu_get_field(h, 99, buf.data()), CPML_ERR_ARG);
  EXPECT(cpml_gpu_get_memory(h, -1, buf.data()), CPML_ERR_ARG);
  EXPECT(cpml_gpu_run(h, 14, nstep), CPML_OK);
  EXPECT(cpml_gpu_guard_cells(h, 0, &guards), CPML_OK);
  if (guards) ++g_fail, std::printf("FAILED: %lld guard cells overwritten\n", guards);
  std::vector<T> sx(nstep * nrec), sy(nstep * nrec), sz(nstep * nrec);
  EXPECT(cpml_gpu_get_seismograms(h, sx.data(), sy.data(), ndim == 3 ? sz.data() : nullptr),
         CPML_OK);
  std::vector<double> ek(nstep), ep(nstep);
  EXPECT(cpml_gpu_get_energy(h, ek.data(), ep.data()), energy ? CPML_OK : CPML_ERR_STATE);
  double v = 0;
  EXPECT(cpml_gpu_velocity_norm(h, &v), CPML_OK);
  if (!(v > 0)) ++g_fail, std::printf("FAILED: velocity norm %g\n", v);
  if (ndim == 2 || order == 2) {
    const int g = (ndim == 2 && order == 4) ? 1 : 0;
    std::vector<unsigned char> rgb((size_t)(n[0] + 2 * g) * (n[1] + 2 * g) * 3);
    EXPECT(cpml_gpu_color_image(h, 1, rgb.data()), CPML_OK);
    EXPECT(cpml_gpu_color_image(h, 3, rgb.data()), CPML_ERR_ARG);
  }
  int w[6];
  EXPECT(cpml_gpu_slab_widths(h, w), CPML_OK);
  char msg[64];
  EXPECT(cpml_gpu_last_error(h, msg, sizeof msg), CPML_OK);
  EXPECT(cpml_gpu_finalize(h), CPML_OK);
}

template <typename T>
static void run_batch(int prec, const std::string &dir) {
  const int ns = 3, nstep = 24;
  const int nx[ns] = {70, 96, 64}, ny[ns] = {80, 64, 130}, np[ns] = {6, 8, 5};
  const double d[ns] = {10, 10, 10}, dt[ns] = {5e-4, 5e-4, 5e-4};
  cpml_batch *b = nullptr;
  EXPECT(cpml_batch_init(&b, 4, prec, ns, nx, ny, d, d, dt, np, nstep), CPML_OK);
  if (!b) return;
  for (int q = 0; q < ns; ++q) {
    const size_t cells = (size_t)nx[q] * ny[q];
    std::vector<T> lam(cells, (T)8e9), mu(cells, (T)8e9), rho(cells, (T)2500);
    EXPECT(cpml_batch_set_medium(b, q, lam.data(), mu.data(), rho.data()), CPML_OK);
    const int n[3] = {nx[q], ny[q], 1};
    Profiles<T> P(prec, n, np[q]);
    EXPECT(cpml_batch_set_pml(b, q, P.p(0), P.p(1), P.p(2), P.p(3), P.p(4), P.p(5), P.p(6),
                              P.p(7), P.p(8), P.p(9), P.p(10), P.p(11)),
           CPML_OK);
    std::vector<T> fx(nstep), fy(nstep);
    cpml_source_time_function(prec, nstep, 5e-4, 20.0, 0.05, 1e7, 45.0, 1, fx.data(), fy.data());
    EXPECT(cpml_batch_set_source(b, q, nx[q] / 2, ny[q] / 3, fx.data(), fy.data()), CPML_OK);
    const int nrec = 2 + q;
    std::vector<int> ix(nrec), iy(nrec);
    for (int r = 0; r < nrec; ++r) ix[r] = 4 + r, iy[r] = 6 + 2 * r;
    EXPECT(cpml_batch_set_receivers(b, q, nrec, ix.data(), iy.data()), CPML_OK);
  }
  EXPECT(cpml_batch_set_source(b, ns, 1, 1, nullptr, nullptr), CPML_ERR_ARG);
  EXPECT(cpml_batch_set_energy(b, 2), CPML_OK);
  EXPECT(cpml_batch_set_stability(b, 10, 1e25), CPML_OK);
  EXPECT(cpml_batch_set_time_block(b, 2, 0, 32), CPML_OK);
  EXPECT(cpml_batch_run(b, 1, 11), CPML_OK);
  EXPECT(cpml_batch_set_time_block(b, 0, 0, 0), CPML_OK);
  EXPECT(cpml_batch_run(b, 12, nstep), CPML_OK);
  int ty = 0, bpc = 0, tb = 0, jc = 0;
  double ms = 0;
  const std::string cache = dir + "/tune.txt";
  EXPECT(cpml_batch_autotune(b, 6, cache.c_str(), &ty, &bpc, &ms), CPML_OK);
  EXPECT(cpml_batch_autotune(b, 6, cache.c_str(), &ty, &bpc, &ms), CPML_OK);
  EXPECT(cpml_batch_launch_shape(b, &ty, &bpc, &tb, &jc), CPML_OK);
  EXPECT(cpml_batch_run(b, 1, nstep), CPML_OK);
  for (int q = 0; q < ns; ++q) {
    std::vector<T> sx(nstep * (2 + q)), sy(nstep * (2 + q)), f((size_t)nx[q] * ny[q]);
    EXPECT(cpml_batch_get_seismograms(b, q, sx.data(), sy.data()), CPML_OK);
    std::vector<double> ek(nstep), ep(nstep);
    EXPECT(cpml_batch_get_energy(b, q, ek.data(), ep.data()), CPML_OK);
    EXPECT(cpml_batch_get_field(b, q, CPML_VX, f.data()), CPML_OK);
    int gx = 0, gy = 0;
    EXPECT(cpml_batch_grid(b, q, &gx, &gy), CPML_OK);
  }
  char msg[16];
  EXPECT(cpml_batch_last_error(b, msg, sizeof msg), CPML_OK);
  EXPECT(cpml_batch_finalize(b), CPML_OK);
}

template <typename T>
static void run_visco_and_ade(int prec) {
  const int n[3] = {80, 76, 1}, npml = 8, nstep = 20;
  for (int order : {2, 4, 8}) {
    cpml_solver *h = nullptr;
    EXPECT(cpml_gpu_init(&h, 2, order, prec, n[0], n[1], 1, 10, 10, 1, 5e-4, npml, nstep),
           CPML_OK);
    if (!h) continue;
    const size_t cells = (size_t)n[0] * n[1];
    std::vector<T> lam(cells, (T)8e9), mu(cells, (T)8e9), rho(cells, (T)2500);
    EXPECT(cpml_gpu_set_medium(h, lam.data(), mu.data(), rho.data()), CPML_OK);
    Profiles<T> P(prec, n, npml);
    EXPECT(cpml_gpu_set_pml(h, P.p(0), P.p(1), P.p(2), P.p(3), P.p(4), P.p(5), P.p(6), P.p(7),
                            P.p(8), P.p(9), P.p(10), P.p(11), nullptr, nullptr, nullptr, nullptr,
                            nullptr, nullptr),
           CPML_OK);
    std::vector<T> fx(4 * nstep, (T)1), fy(4 * nstep, (T)1);
    EXPECT(cpml_gpu_set_source(h, 40, 38, 1, fx.data(), fy.data()), CPML_OK);
    int ix[2] = {10, 20}, iy[2] = {12, 30};
    EXPECT(cpml_gpu_set_receivers(h, 2, ix, iy, nullptr), CPML_OK);
    if (order == 8) {
      std::vector<T> tx(4 * n[0], (T)0.5), ty(4 * n[1], (T)0.5);
      EXPECT(cpml_gpu_set_ade_rk4(h, tx.data(), tx.data(), tx.data(), tx.data(), ty.data(),
                                  ty.data(), ty.data(), ty.data(), fx.data(), fy.data()),
             CPML_OK);
    } else {
      const double te[3] = {1.1e-2, 1.2e-2, 1.3e-2}, ts[3] = {1.0e-2, 1.1e-2, 1.2e-2};
      EXPECT(cpml_gpu_set_attenuation(h, 3, te, ts, te, ts), CPML_OK);
      EXPECT(cpml_gpu_set_attenuation(h, 9, te, ts, te, ts), CPML_ERR_ARG);
    }
    EXPECT(cpml_gpu_run(h, 1, nstep), CPML_OK);
    std::vector<T> s(2 * nstep), att((size_t)3 * cells);
    if (order != 8) {
      EXPECT(cpml_gpu_get_pressure(h, s.data()), CPML_OK);
      for (int k = 0; k < 3; ++k) EXPECT(cpml_gpu_get_attenuation_memory(h, k, att.data()), CPML_OK);
    }
    EXPECT(cpml_gpu_finalize(h), CPML_OK);
  }
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s cpu|gpu <dir>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1], dir = argv[2];
  check_messages();
  check_init_validation();
  check_null_handles();
  check_helpers<float>(4, dir);
  check_helpers<double>(8, dir);
  if (mode == "gpu") {
    std::printf("gpu paths\n");
    std::fflush(stdout);
    for (int order : {2, 4, 8}) {
      run_solver<float>(4, 2, order, 0, 0, 0, 0);
      run_solver<double>(8, 2, order, 2, 0, 1, 1);
      run_solver<float>(4, 2, order, 2, 0, 5, 4);
    }
    run_solver<float>(4, 3, 2, 0, 0, 0, 0);
    run_solver<double>(8, 3, 2, 0, 1, 2, 1);
    run_solver<float>(4, 3, 2, 2, 0, 1, 1);
    run_solver<double>(8, 3, 4, 0, 0, 0, 2);
    run_batch<float>(4, dir);
    run_batch<double>(8, dir);
    run_visco_and_ade<float>(4);
  } else {
    // without a GPU the valid init reports CPML_ERR_NODEVICE (after the host checks)
    cpml_solver *h = nullptr;
    const int rc = cpml_gpu_init(&h, 2, 2, 8, 10, 10, 1, 1, 1, 1, 1e-3, 2, 10);
    ++g_checks;
    if (rc != CPML_ERR_NODEVICE && !(rc == CPML_OK && h))
      ++g_fail, std::printf("FAILED: valid init returned %d\n", rc);
    cpml_gpu_finalize(h);
  }
  std::printf("%d checks, %d failed\n", g_checks, g_fail);
  if (g_fail) return 1;
  std::printf("HOST CHECK OK\n");
  return 0;
}
