/*
 * Body of the CPU oracle, included twice by cpml_oracle.c with
 *   REAL  = double | float
 *   SFX   = d | f
 *   LIT(x) = literal x in REAL precision (x or x##f)
 *   EXP, SQRT, FABS = libm routine of that precision.
 *
 * TEST INFRASTRUCTURE ONLY (see cpml_oracle.h).
 *
 * Expression association follows the reference Fortran statement by statement
 * (Fortran evaluates a*b*c as (a*b)*c and -a*b as -(a*b)); the file is compiled
 * with -ffp-contract=off so no multiply-add is fused, matching the x86-64
 * reference build, which has no FMA.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define ORC(name) CAT(oracle_##name, SFX)

/* --------------------------------------------------------------------------
 * a-1  C-PML damping profile along one axis.
 * 2D2:283-394 (x), 2D2:396-465 (y), 3D:399-667 (x, y, z).
 *   origin_max_uses_n : 2D4:401 sets the max-side origin to N*DELTA - L instead
 *                       of (N-1)*DELTA - L (quirk kept as an option).
 *   clamp_alpha       : the reference clamps alpha >= 0 only along x
 *                       (2D2:385-386, 3D:493-494); y and z are not clamped.
 * ------------------------------------------------------------------------ */
void ORC(profile)(int n, REAL delta, REAL deltat, int npml, REAL cp, REAL f0, int use_min,
                  int use_max, int origin_max_uses_n, int clamp_alpha, REAL *a, REAL *b,
                  REAL *K, REAL *a_half, REAL *b_half, REAL *K_half)
{
  const REAL PI = LIT(3.141592653589793238462643);
  const REAL NPOWER = LIT(2.0);
  const REAL K_MAX_PML = LIT(1.0);
  const REAL ALPHA_MAX_PML = LIT(2.0) * PI * (f0 / LIT(2.0)); /* 2D2:218 */
  const REAL ZERO = LIT(0.0);
  const REAL thickness = (REAL)npml * delta;                  /* 2D2:284 */
  const REAL Rcoef = LIT(0.001);                              /* 2D2:288 */
  /* 2D2:296: -(NPOWER+1)*cp*log(Rcoef)/(2*thickness) */
  const REAL d0 = -(((NPOWER + LIT(1.0)) * cp * LOG(Rcoef)) / (LIT(2.0) * thickness));
  const REAL origin_min = thickness;                                          /* 2D2:324 */
  const REAL origin_max = (REAL)(origin_max_uses_n ? n : n - 1) * delta - thickness; /* 2D2:325 */

  for (int i = 1; i <= n; ++i) {
    REAL d = ZERO, dh = ZERO, k = LIT(1.0), kh = LIT(1.0), al = ZERO, alh = ZERO;
    REAL av = ZERO, avh = ZERO;
    const REAL xval = delta * (REAL)(i - 1); /* 2D2:330 */
    REAL absc, xn;
    if (use_min) { /* 2D2:333-356 */
      absc = origin_min - xval;
      if (absc >= ZERO) {
        xn = absc / thickness;
        d = d0 * (xn * xn);
        k = LIT(1.0) + (K_MAX_PML - LIT(1.0)) * (xn * xn);
        al = ALPHA_MAX_PML * (LIT(1.0) - xn);
      }
      absc = origin_min - (xval + delta / LIT(2.0));
      if (absc >= ZERO) {
        xn = absc / thickness;
        dh = d0 * (xn * xn);
        kh = LIT(1.0) + (K_MAX_PML - LIT(1.0)) * (xn * xn);
        alh = ALPHA_MAX_PML * (LIT(1.0) - xn);
      }
    }
    if (use_max) { /* 2D2:359-381 */
      absc = xval - origin_max;
      if (absc >= ZERO) {
        xn = absc / thickness;
        d = d0 * (xn * xn);
        k = LIT(1.0) + (K_MAX_PML - LIT(1.0)) * (xn * xn);
        al = ALPHA_MAX_PML * (LIT(1.0) - xn);
      }
      absc = xval + delta / LIT(2.0) - origin_max;
      if (absc >= ZERO) {
        xn = absc / thickness;
        dh = d0 * (xn * xn);
        kh = LIT(1.0) + (K_MAX_PML - LIT(1.0)) * (xn * xn);
        alh = ALPHA_MAX_PML * (LIT(1.0) - xn);
      }
    }
    if (clamp_alpha) { /* 2D2:385-386 */
      if (al < ZERO) al = ZERO;
      if (alh < ZERO) alh = ZERO;
    }
    const REAL bv = EXP(-(d / k + al) * deltat); /* 2D2:388-389 */
    const REAL bvh = EXP(-(dh / kh + alh) * deltat);
    /* 2D2:392-394 */
    if (FABS(d) > LIT(1.0e-6)) av = d * (bv - LIT(1.0)) / (k * (d + k * al));
    if (FABS(dh) > LIT(1.0e-6)) avh = dh * (bvh - LIT(1.0)) / (kh * (dh + kh * alh));
    a[i - 1] = av;
    b[i - 1] = bv;
    K[i - 1] = k;
    a_half[i - 1] = avh;
    b_half[i - 1] = bvh;
    K_half[i - 1] = kh;
  }
  (void)NPOWER;
}

/* --------------------------------------------------------------------------
 * a-6  Source time function and force split, 2D2:644-657 (3D:1059-1070).
 *   kind 0: Gaussian  (2D2:649, commented in the reference)
 *   kind 1: first derivative of a Gaussian (2D2:651, what the programs run)
 *   kind 2: Ricker    (2D2:654, commented in the reference)
 * fx[it-1], fy[it-1] = force_x, force_y at step it.
 * ------------------------------------------------------------------------ */
void ORC(stf)(int nstep, REAL deltat, REAL f0, REAL t0, REAL factor, REAL angle_deg, int kind,
              REAL *fx, REAL *fy)
{
  const REAL PI = LIT(3.141592653589793238462643);
  const REAL DEG2RAD = PI / LIT(180.0);
  const REAL a = PI * PI * f0 * f0;
  const REAL sa = SIN(angle_deg * DEG2RAD);
  const REAL ca = COS(angle_deg * DEG2RAD);
  for (int it = 1; it <= nstep; ++it) {
    const REAL t = (REAL)(it - 1) * deltat;
    const REAL tt = t - t0;
    REAL st;
    if (kind == 0)
      st = factor * EXP(-(a * (tt * tt)));
    else if (kind == 1)
      st = -(factor * LIT(2.0) * a * tt * EXP(-(a * (tt * tt))));
    else
      st = factor * (LIT(1.0) - LIT(2.0) * a * (tt * tt)) * EXP(-(a * (tt * tt)));
    fx[it - 1] = sa * st;
    fy[it - 1] = ca * st;
  }
}

/* Same source time function at given times t[0..n-1] (the RK4 stage times
 * of ADE8:786-796); kind as above. */
void ORC(stf_at)(int n, const REAL *t, REAL f0, REAL t0, REAL factor, REAL angle_deg, int kind,
                 REAL *fx, REAL *fy)
{
  const REAL PI = LIT(3.141592653589793238462643);
  const REAL DEG2RAD = PI / LIT(180.0);
  const REAL a = PI * PI * f0 * f0;
  const REAL sa = SIN(angle_deg * DEG2RAD);
  const REAL ca = COS(angle_deg * DEG2RAD);
  for (int q = 0; q < n; ++q) {
    const REAL tt = t[q] - t0;
    REAL st;
    if (kind == 0)
      st = factor * EXP(-(a * (tt * tt)));
    else if (kind == 1)
      st = -(factor * LIT(2.0) * a * tt * EXP(-(a * (tt * tt))));
    else
      st = factor * (LIT(1.0) - LIT(2.0) * a * (tt * tt)) * EXP(-(a * (tt * tt)));
    fx[q] = sa * st;
    fy[q] = ca * st;
  }
}

/* --------------------------------------------------------------------------
 * a-8  Receiver placement: nearest grid node to points evenly spaced on the
 * segment (xdeb,ydeb)-(xfin,yfin); first minimum in j-outer/i-inner order.
 * 2D2:486-509.
 * ------------------------------------------------------------------------ */
void ORC(receivers)(int nx, int ny, REAL dx, REAL dy, int nrec, REAL xdeb, REAL ydeb, REAL xfin,
                    REAL yfin, int *ix, int *iy)
{
  const REAL xs = nrec > 1 ? (xfin - xdeb) / (REAL)(nrec - 1) : LIT(0.0);
  const REAL ys = nrec > 1 ? (yfin - ydeb) / (REAL)(nrec - 1) : LIT(0.0);
  for (int r = 1; r <= nrec; ++r) {
    const REAL xr = xdeb + (REAL)(r - 1) * xs;
    const REAL yr = ydeb + (REAL)(r - 1) * ys;
    REAL dist = LIT(1.0e30);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const REAL ex = dx * (REAL)(i - 1) - xr;
        const REAL ey = dy * (REAL)(j - 1) - yr;
        const REAL dv = SQRT(ex * ex + ey * ey);
        if (dv < dist) {
          dist = dv;
          ix[r - 1] = i;
          iy[r - 1] = j;
        }
      }
  }
}

/* --------------------------------------------------------------------------
 * a-3 .. a-11, 2D.  One call runs steps it_begin..it_end of the time loop of
 * 2D2:550-735 (order 2), 2D4:551-736 (order 4), or the leapfrog loop with the
 * ADE8:734-738 8th-order operator (order 8, no reference trajectory).
 *
 * prof_x = [a_x, b_x, K_x, a_x_half, b_x_half, K_x_half] (6*nx)
 * prof_y = the same along y (6*ny)
 * state  = [vx, vy, sxx, syy, sxy,
 *           m_dvx_dx, m_dvx_dy, m_dvy_dx, m_dvy_dy,
 *           m_dsxx_dx, m_dsyy_dy, m_dsxy_dx, m_dsxy_dy]   each nx*ny, in/out
 * ewin   = energy window i0,i1,j0,j1 (1-based, inclusive) or NULL for none;
 *          2D2:695-711 uses NPML+1..N-NPML, 2D4:696-710 uses NPML..N-NPML+1.
 * ------------------------------------------------------------------------ */
typedef struct {
  REAL *p;
  int g, ld;
} ORC(grid2);

int ORC(run2d)(int nx, int ny, int order, REAL dx, REAL dy, REAL dt, int it_begin, int it_end,
               const REAL *lambda, const REAL *mu, const REAL *rho, const REAL *prof_x,
               const REAL *prof_y, int isrc, int jsrc, const REAL *fx, const REAL *fy, int nrec,
               const int *ixr, const int *iyr, int ldseis, REAL *sisvx, REAL *sisvy,
               const int *ewin, REAL *ekin, REAL *epot, REAL *state, int nthreads)
{
  if (order != 2 && order != 4 && order != 8) return 1;
  if (nx < 2 || ny < 2) return 1;
  const int g = order / 2;
  const int ld = nx + 2 * g;
  const size_t plane = (size_t)ld * (size_t)(ny + 2 * g);
  const size_t nn = (size_t)nx * ny;
  REAL *buf = (REAL *)calloc(plane * 16, sizeof(REAL));
  if (!buf) return 2;
  /* padded copies: 5 fields, 8 memory arrays, lambda, mu, rho */
  REAL *F[16];
  for (int q = 0; q < 16; ++q) F[q] = buf + plane * q;
#define P(arr, i, j) (arr)[(size_t)((i) - 1 + g) + (size_t)((j) - 1 + g) * ld]
  for (int q = 0; q < 16; ++q) {
    const REAL *src = q < 13 ? state + nn * q : (q == 13 ? lambda : (q == 14 ? mu : rho));
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) P(F[q], i, j) = src[(size_t)(i - 1) + (size_t)(j - 1) * nx];
  }
  REAL *vx = F[0], *vy = F[1], *sxx = F[2], *syy = F[3], *sxy = F[4];
  REAL *m_dvx_dx = F[5], *m_dvx_dy = F[6], *m_dvy_dx = F[7], *m_dvy_dy = F[8];
  REAL *m_dsxx_dx = F[9], *m_dsyy_dy = F[10], *m_dsxy_dx = F[11], *m_dsxy_dy = F[12];
  REAL *lam = F[13], *muu = F[14], *rh = F[15];
  const REAL *a_x = prof_x, *b_x = prof_x + nx, *K_x = prof_x + 2 * nx;
  const REAL *a_x_half = prof_x + 3 * nx, *b_x_half = prof_x + 4 * nx, *K_x_half = prof_x + 5 * nx;
  const REAL *a_y = prof_y, *b_y = prof_y + ny, *K_y = prof_y + 2 * ny;
  const REAL *a_y_half = prof_y + 3 * ny, *b_y_half = prof_y + 4 * ny, *K_y_half = prof_y + 5 * ny;
  /* 2D4:596 divides dvx/dy by K_y(j) where 2D2:595 uses K_y_half(j). */
  const REAL *K_y_for_dvx_dy = order == 4 ? K_y : K_y_half;
  const REAL h1 = (REAL)oracle_holberg8[0], h2 = (REAL)oracle_holberg8[1];
  const REAL h3 = (REAL)oracle_holberg8[2], h4 = (REAL)oracle_holberg8[3];
  const REAL den_x = order == 2 ? dx : (order == 4 ? LIT(24.0) * dx : dx);
  const REAL den_y = order == 2 ? dy : (order == 4 ? LIT(24.0) * dy : dy);

  /* Forward staggered difference f(s+1/2) - f(s-1/2) over the 2r points of the
   * operator; lo = f(s-1/2 - m), hi = f(s+1/2 + m) accessed through an
   * index function.  Order 2: 2D2:564; order 4: 2D4:565; order 8: ADE8:734. */
#define DIFF(f, im, ip, im2, ip2, im3, ip3, im4, ip4, den)                                   \
  (order == 2   ? ((ip) - (im)) / (den)                                                   \
   : order == 4 ? (LIT(27.0) * (ip) - LIT(27.0) * (im) - (ip2) + (im2)) / (den)          \
                : (h1 * ((ip) - (im)) + h2 * ((ip2) - (im2)) + h3 * ((ip3) - (im3)) +       \
                   h4 * ((ip4) - (im4))) / (den))

#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

  for (int it = it_begin; it <= it_end; ++it) {
    /* ---- stress: sigma_xx, sigma_yy  (2D2:556-580) ---- */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        const REAL lambda_half_x = LIT(0.5) * (P(lam, i + 1, j) + P(lam, i, j));
        const REAL mu_half_x = LIT(0.5) * (P(muu, i + 1, j) + P(muu, i, j));
        const REAL lp2m = lambda_half_x + LIT(2.0) * mu_half_x;
        REAL dvx_dx = DIFF(vx, P(vx, i, j), P(vx, i + 1, j), P(vx, i - 1, j), P(vx, i + 2, j),
                           P(vx, i - 2, j), P(vx, i + 3, j), P(vx, i - 3, j), P(vx, i + 4, j),
                           den_x);
        REAL dvy_dy = DIFF(vy, P(vy, i, j - 1), P(vy, i, j), P(vy, i, j - 2), P(vy, i, j + 1),
                           P(vy, i, j - 3), P(vy, i, j + 2), P(vy, i, j - 4), P(vy, i, j + 3),
                           den_y);
        P(m_dvx_dx, i, j) = b_x_half[i - 1] * P(m_dvx_dx, i, j) + a_x_half[i - 1] * dvx_dx;
        P(m_dvy_dy, i, j) = b_y[j - 1] * P(m_dvy_dy, i, j) + a_y[j - 1] * dvy_dy;
        dvx_dx = dvx_dx / K_x_half[i - 1] + P(m_dvx_dx, i, j);
        dvy_dy = dvy_dy / K_y[j - 1] + P(m_dvy_dy, i, j);
        P(sxx, i, j) = P(sxx, i, j) + (lp2m * dvx_dx + lambda_half_x * dvy_dy) * dt;
        P(syy, i, j) = P(syy, i, j) + (lambda_half_x * dvx_dx + lp2m * dvy_dy) * dt;
      }
    /* ---- stress: sigma_xy  (2D2:582-600) ---- */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 2; i <= nx; ++i) {
        const REAL mu_half_y = LIT(0.5) * (P(muu, i, j + 1) + P(muu, i, j));
        REAL dvy_dx = DIFF(vy, P(vy, i - 1, j), P(vy, i, j), P(vy, i - 2, j), P(vy, i + 1, j),
                           P(vy, i - 3, j), P(vy, i + 2, j), P(vy, i - 4, j), P(vy, i + 3, j),
                           den_x);
        REAL dvx_dy = DIFF(vx, P(vx, i, j), P(vx, i, j + 1), P(vx, i, j - 1), P(vx, i, j + 2),
                           P(vx, i, j - 2), P(vx, i, j + 3), P(vx, i, j - 3), P(vx, i, j + 4),
                           den_y);
        P(m_dvy_dx, i, j) = b_x[i - 1] * P(m_dvy_dx, i, j) + a_x[i - 1] * dvy_dx;
        P(m_dvx_dy, i, j) = b_y_half[j - 1] * P(m_dvx_dy, i, j) + a_y_half[j - 1] * dvx_dy;
        dvy_dx = dvy_dx / K_x[i - 1] + P(m_dvy_dx, i, j);
        dvx_dy = dvx_dy / K_y_for_dvx_dy[j - 1] + P(m_dvx_dy, i, j);
        P(sxy, i, j) = P(sxy, i, j) + mu_half_y * (dvy_dx + dvx_dy) * dt;
      }
    /* ---- velocity: vx  (2D2:606-621) ---- */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny; ++j)
      for (int i = 2; i <= nx; ++i) {
        REAL dsxx_dx = DIFF(sxx, P(sxx, i - 1, j), P(sxx, i, j), P(sxx, i - 2, j), P(sxx, i + 1, j),
                            P(sxx, i - 3, j), P(sxx, i + 2, j), P(sxx, i - 4, j),
                            P(sxx, i + 3, j), den_x);
        REAL dsxy_dy = DIFF(sxy, P(sxy, i, j - 1), P(sxy, i, j), P(sxy, i, j - 2), P(sxy, i, j + 1),
                            P(sxy, i, j - 3), P(sxy, i, j + 2), P(sxy, i, j - 4),
                            P(sxy, i, j + 3), den_y);
        P(m_dsxx_dx, i, j) = b_x[i - 1] * P(m_dsxx_dx, i, j) + a_x[i - 1] * dsxx_dx;
        P(m_dsxy_dy, i, j) = b_y[j - 1] * P(m_dsxy_dy, i, j) + a_y[j - 1] * dsxy_dy;
        dsxx_dx = dsxx_dx / K_x[i - 1] + P(m_dsxx_dx, i, j);
        dsxy_dy = dsxy_dy / K_y[j - 1] + P(m_dsxy_dy, i, j);
        P(vx, i, j) = P(vx, i, j) + (dsxx_dx + dsxy_dy) * dt / P(rh, i, j);
      }
    /* ---- velocity: vy  (2D2:623-641) ---- */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        const REAL rho_hh =
            LIT(0.25) * (P(rh, i, j) + P(rh, i + 1, j) + P(rh, i + 1, j + 1) + P(rh, i, j + 1));
        REAL dsxy_dx = DIFF(sxy, P(sxy, i, j), P(sxy, i + 1, j), P(sxy, i - 1, j), P(sxy, i + 2, j),
                            P(sxy, i - 2, j), P(sxy, i + 3, j), P(sxy, i - 3, j),
                            P(sxy, i + 4, j), den_x);
        REAL dsyy_dy = DIFF(syy, P(syy, i, j), P(syy, i, j + 1), P(syy, i, j - 1), P(syy, i, j + 2),
                            P(syy, i, j - 2), P(syy, i, j + 3), P(syy, i, j - 3),
                            P(syy, i, j + 4), den_y);
        P(m_dsxy_dx, i, j) = b_x_half[i - 1] * P(m_dsxy_dx, i, j) + a_x_half[i - 1] * dsxy_dx;
        P(m_dsyy_dy, i, j) = b_y_half[j - 1] * P(m_dsyy_dy, i, j) + a_y_half[j - 1] * dsyy_dy;
        dsxy_dx = dsxy_dx / K_x_half[i - 1] + P(m_dsxy_dx, i, j);
        dsyy_dy = dsyy_dy / K_y_half[j - 1] + P(m_dsyy_dy, i, j);
        P(vy, i, j) = P(vy, i, j) + (dsxy_dx + dsyy_dy) * dt / rho_hh;
      }
    /* ---- source (2D2:644-667) ---- */
    if (isrc >= 1 && fx && fy) {
      const int i = isrc, j = jsrc;
      const REAL rho_hh =
          LIT(0.25) * (P(rh, i, j) + P(rh, i + 1, j) + P(rh, i + 1, j + 1) + P(rh, i, j + 1));
      P(vx, i, j) = P(vx, i, j) + fx[it - 1] * dt / P(rh, i, j);
      P(vy, i, j) = P(vy, i, j) + fy[it - 1] * dt / rho_hh;
    }
    /* ---- Dirichlet edges (2D2:670-680) ---- */
    for (int j = 1; j <= ny; ++j) {
      P(vx, 1, j) = 0;
      P(vx, nx, j) = 0;
      P(vy, 1, j) = 0;
      P(vy, nx, j) = 0;
    }
    for (int i = 1; i <= nx; ++i) {
      P(vx, i, 1) = 0;
      P(vx, i, ny) = 0;
      P(vy, i, 1) = 0;
      P(vy, i, ny) = 0;
    }
    /* ---- seismograms (2D2:683-686) ---- */
    for (int r = 0; r < nrec; ++r) {
      sisvx[(size_t)(it - 1) + (size_t)r * ldseis] = P(vx, ixr[r], iyr[r]);
      sisvy[(size_t)(it - 1) + (size_t)r * ldseis] = P(vy, ixr[r], iyr[r]);
    }
    /* ---- energy (2D2:694-713) ---- */
    if (ewin && ekin && epot) {
      const int i0 = ewin[0], i1 = ewin[1], j0 = ewin[2], j1 = ewin[3];
      REAL s = 0;
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i)
          s = s + P(rh, i, j) * (P(vx, i, j) * P(vx, i, j) + P(vy, i, j) * P(vy, i, j));
      ekin[it - 1] = LIT(0.5) * s;
      REAL ep = 0;
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          const REAL l = P(lam, i, j), m = P(muu, i, j);
          const REAL exx = ((l + LIT(2.0) * m) * P(sxx, i, j) - l * P(syy, i, j)) /
                           (LIT(4.0) * m * (l + m));
          const REAL eyy = ((l + LIT(2.0) * m) * P(syy, i, j) - l * P(sxx, i, j)) /
                           (LIT(4.0) * m * (l + m));
          const REAL exy = P(sxy, i, j) / (LIT(2.0) * m);
          ep = ep + LIT(0.5) * (exx * P(sxx, i, j) + eyy * P(syy, i, j) +
                                LIT(2.0) * exy * P(sxy, i, j));
        }
      epot[it - 1] = ep;
    }
  }
#undef DIFF
  for (int q = 0; q < 13; ++q)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        state[nn * q + (size_t)(i - 1) + (size_t)(j - 1) * nx] = P(F[q], i, j);
#undef P
  free(buf);
  return 0;
}

/* --------------------------------------------------------------------------
 * a-5 .. a-11, 3D: the loop of 3D:802-1180 on one global domain.
 *
 * The reference splits z over MPI ranks; with its 1-plane halos exchanged
 * before each phase (3D:811-823, 951-963) the result equals one global sweep
 * with the global loop bounds k2begin (global k >= 2, 3D:792-793) and
 * kminus1end (global k <= NZ-1, 3D:795-796), which is what is done here.
 *
 * The reference medium is homogeneous (3D:139-144, 296-300).  This restatement
 * takes per-point lambda, mu, rho and lp2m (lambda+2mu, passed separately so a
 * homogeneous run can use the reference's rho*cp*cp, 3D:145); pass lp2m=NULL
 * to use DELTAT*lambda + 2*(DELTAT*mu) for the (lambda+2mu) coefficient.
 *
 * state = [vx, vy, vz, sxx, syy, szz, sxy, sxz, syz,
 *          m_dvx_dx, m_dvx_dy, m_dvx_dz, m_dvy_dx, m_dvy_dy, m_dvy_dz,
 *          m_dvz_dx, m_dvz_dy, m_dvz_dz, m_dsxx_dx, m_dsyy_dy, m_dszz_dz,
 *          m_dsxy_dx, m_dsxy_dy, m_dsxz_dx, m_dsxz_dz, m_dsyz_dy, m_dsyz_dz]
 *         each nx*ny*nz, in/out.
 * Energy (3D:1131-1180) keeps the reference formula, in which the
 * eps_yy*sigma_yy term appears twice and eps_zz*sigma_zz not at all
 * (3D:1170-1172); energy_slabs = number of MPI ranks whose partial sums are
 * added (3D:1179), so the summation order matches an energy_slabs-rank run.
 * ------------------------------------------------------------------------ */
int ORC(run3d)(int nx, int ny, int nz, REAL dx, REAL dy, REAL dz, REAL dt, int it_begin,
               int it_end, const REAL *lambda, const REAL *mu, const REAL *rho, const REAL *lp2m,
               const REAL *prof_x, const REAL *prof_y, const REAL *prof_z, int isrc, int jsrc,
               int ksrc, const REAL *fx, const REAL *fy, int nrec, const int *ixr, const int *iyr,
               const int *izr, int ldseis, REAL *sisvx, REAL *sisvy, REAL *sisvz, const int *ewin,
               int energy_slabs, REAL *etotal, REAL *state, int nthreads)
{
  if (nx < 2 || ny < 2 || nz < 2) return 1;
  const size_t nn = (size_t)nx * ny * nz;
  REAL *cl = (REAL *)malloc(nn * 4 * sizeof(REAL));
  if (!cl) return 2;
  REAL *c_lambda = cl, *c_mu = cl + nn, *c_lp2m = cl + 2 * nn, *c_or = cl + 3 * nn;
  /* 3D:296-300 precomputed coefficients, per point */
  for (size_t q = 0; q < nn; ++q) {
    c_lambda[q] = dt * lambda[q];
    c_mu[q] = dt * mu[q];
    c_lp2m[q] = lp2m ? dt * lp2m[q] : c_lambda[q] + LIT(2.0) * c_mu[q];
    c_or[q] = dt / rho[q];
  }
  const REAL odx = LIT(1.0) / dx, ody = LIT(1.0) / dy, odz = LIT(1.0) / dz; /* 3D:135-137 */
  REAL *vx = state, *vy = state + nn, *vz = state + 2 * nn;
  REAL *sxx = state + 3 * nn, *syy = state + 4 * nn, *szz = state + 5 * nn;
  REAL *sxy = state + 6 * nn, *sxz = state + 7 * nn, *syz = state + 8 * nn;
  REAL *m_dvx_dx = state + 9 * nn, *m_dvx_dy = state + 10 * nn, *m_dvx_dz = state + 11 * nn;
  REAL *m_dvy_dx = state + 12 * nn, *m_dvy_dy = state + 13 * nn, *m_dvy_dz = state + 14 * nn;
  REAL *m_dvz_dx = state + 15 * nn, *m_dvz_dy = state + 16 * nn, *m_dvz_dz = state + 17 * nn;
  REAL *m_dsxx_dx = state + 18 * nn, *m_dsyy_dy = state + 19 * nn, *m_dszz_dz = state + 20 * nn;
  REAL *m_dsxy_dx = state + 21 * nn, *m_dsxy_dy = state + 22 * nn, *m_dsxz_dx = state + 23 * nn;
  REAL *m_dsxz_dz = state + 24 * nn, *m_dsyz_dy = state + 25 * nn, *m_dsyz_dz = state + 26 * nn;
  const REAL *a_x = prof_x, *b_x = prof_x + nx, *K_x = prof_x + 2 * nx;
  const REAL *a_x_half = prof_x + 3 * nx, *b_x_half = prof_x + 4 * nx, *K_x_half = prof_x + 5 * nx;
  const REAL *a_y = prof_y, *b_y = prof_y + ny, *K_y = prof_y + 2 * ny;
  const REAL *a_y_half = prof_y + 3 * ny, *b_y_half = prof_y + 4 * ny, *K_y_half = prof_y + 5 * ny;
  const REAL *a_z = prof_z, *b_z = prof_z + nz, *K_z = prof_z + 2 * nz;
  const REAL *a_z_half = prof_z + 3 * nz, *b_z_half = prof_z + 4 * nz, *K_z_half = prof_z + 5 * nz;
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
#define Q(i, j, k) ((size_t)((i) - 1) + (size_t)((j) - 1) * sy + (size_t)((k) - 1) * sz)

#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

  for (int it = it_begin; it <= it_end; ++it) {
    /* sigma_xx, sigma_yy, sigma_zz  (3D:836-863) */
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nz; ++k)
      for (int j = 2; j <= ny; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t q = Q(i, j, k);
          REAL dvx_dx = (vx[q + 1] - vx[q]) * odx;
          REAL dvy_dy = (vy[q] - vy[q - sy]) * ody;
          REAL dvz_dz = (vz[q] - vz[q - sz]) * odz;
          m_dvx_dx[q] = b_x_half[i - 1] * m_dvx_dx[q] + a_x_half[i - 1] * dvx_dx;
          m_dvy_dy[q] = b_y[j - 1] * m_dvy_dy[q] + a_y[j - 1] * dvy_dy;
          m_dvz_dz[q] = b_z[k - 1] * m_dvz_dz[q] + a_z[k - 1] * dvz_dz;
          dvx_dx = dvx_dx / K_x_half[i - 1] + m_dvx_dx[q];
          dvy_dy = dvy_dy / K_y[j - 1] + m_dvy_dy[q];
          dvz_dz = dvz_dz / K_z[k - 1] + m_dvz_dz[q];
          sxx[q] = c_lp2m[q] * dvx_dx + c_lambda[q] * (dvy_dy + dvz_dz) + sxx[q];
          syy[q] = c_lambda[q] * (dvx_dx + dvz_dz) + c_lp2m[q] * dvy_dy + syy[q];
          szz[q] = c_lambda[q] * (dvx_dx + dvy_dy) + c_lp2m[q] * dvz_dz + szz[q];
        }
    /* sigma_xy  (3D:877-894) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t q = Q(i, j, k);
          REAL dvy_dx = (vy[q] - vy[q - 1]) * odx;
          REAL dvx_dy = (vx[q + sy] - vx[q]) * ody;
          m_dvy_dx[q] = b_x[i - 1] * m_dvy_dx[q] + a_x[i - 1] * dvy_dx;
          m_dvx_dy[q] = b_y_half[j - 1] * m_dvx_dy[q] + a_y_half[j - 1] * dvx_dy;
          dvy_dx = dvy_dx / K_x[i - 1] + m_dvy_dx[q];
          dvx_dy = dvx_dy / K_y_half[j - 1] + m_dvx_dy[q];
          sxy[q] = c_mu[q] * (dvy_dx + dvx_dy) + sxy[q];
        }
    /* sigma_xz, sigma_yz  (3D:908-943) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) {
      for (int j = 1; j <= ny; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t q = Q(i, j, k);
          REAL dvz_dx = (vz[q] - vz[q - 1]) * odx;
          REAL dvx_dz = (vx[q + sz] - vx[q]) * odz;
          m_dvz_dx[q] = b_x[i - 1] * m_dvz_dx[q] + a_x[i - 1] * dvz_dx;
          m_dvx_dz[q] = b_z_half[k - 1] * m_dvx_dz[q] + a_z_half[k - 1] * dvx_dz;
          dvz_dx = dvz_dx / K_x[i - 1] + m_dvz_dx[q];
          dvx_dz = dvx_dz / K_z_half[k - 1] + m_dvx_dz[q];
          sxz[q] = c_mu[q] * (dvz_dx + dvx_dz) + sxz[q];
        }
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t q = Q(i, j, k);
          REAL dvz_dy = (vz[q + sy] - vz[q]) * ody;
          REAL dvy_dz = (vy[q + sz] - vy[q]) * odz;
          m_dvz_dy[q] = b_y_half[j - 1] * m_dvz_dy[q] + a_y_half[j - 1] * dvz_dy;
          m_dvy_dz[q] = b_z_half[k - 1] * m_dvy_dz[q] + a_z_half[k - 1] * dvy_dz;
          dvz_dy = dvz_dy / K_y_half[j - 1] + m_dvz_dy[q];
          dvy_dz = dvy_dz / K_z_half[k - 1] + m_dvy_dz[q];
          syz[q] = c_mu[q] * (dvz_dy + dvy_dz) + syz[q];
        }
    }
    /* vx, vy  (3D:976-1017) */
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nz; ++k) {
      for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t q = Q(i, j, k);
          REAL dsxx_dx = (sxx[q] - sxx[q - 1]) * odx;
          REAL dsxy_dy = (sxy[q] - sxy[q - sy]) * ody;
          REAL dsxz_dz = (sxz[q] - sxz[q - sz]) * odz;
          m_dsxx_dx[q] = b_x[i - 1] * m_dsxx_dx[q] + a_x[i - 1] * dsxx_dx;
          m_dsxy_dy[q] = b_y[j - 1] * m_dsxy_dy[q] + a_y[j - 1] * dsxy_dy;
          m_dsxz_dz[q] = b_z[k - 1] * m_dsxz_dz[q] + a_z[k - 1] * dsxz_dz;
          dsxx_dx = dsxx_dx / K_x[i - 1] + m_dsxx_dx[q];
          dsxy_dy = dsxy_dy / K_y[j - 1] + m_dsxy_dy[q];
          dsxz_dz = dsxz_dz / K_z[k - 1] + m_dsxz_dz[q];
          vx[q] = c_or[q] * (dsxx_dx + dsxy_dy + dsxz_dz) + vx[q];
        }
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t q = Q(i, j, k);
          REAL dsxy_dx = (sxy[q + 1] - sxy[q]) * odx;
          REAL dsyy_dy = (syy[q + sy] - syy[q]) * ody;
          REAL dsyz_dz = (syz[q] - syz[q - sz]) * odz;
          m_dsxy_dx[q] = b_x_half[i - 1] * m_dsxy_dx[q] + a_x_half[i - 1] * dsxy_dx;
          m_dsyy_dy[q] = b_y_half[j - 1] * m_dsyy_dy[q] + a_y_half[j - 1] * dsyy_dy;
          m_dsyz_dz[q] = b_z[k - 1] * m_dsyz_dz[q] + a_z[k - 1] * dsyz_dz;
          dsxy_dx = dsxy_dx / K_x_half[i - 1] + m_dsxy_dx[q];
          dsyy_dy = dsyy_dy / K_y_half[j - 1] + m_dsyy_dy[q];
          dsyz_dz = dsyz_dz / K_z[k - 1] + m_dsyz_dz[q];
          vy[q] = c_or[q] * (dsxy_dx + dsyy_dy + dsyz_dz) + vy[q];
        }
    }
    /* vz  (3D:1031-1052) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k)
      for (int j = 2; j <= ny; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t q = Q(i, j, k);
          REAL dsxz_dx = (sxz[q + 1] - sxz[q]) * odx;
          REAL dsyz_dy = (syz[q] - syz[q - sy]) * ody;
          REAL dszz_dz = (szz[q + sz] - szz[q]) * odz;
          m_dsxz_dx[q] = b_x_half[i - 1] * m_dsxz_dx[q] + a_x_half[i - 1] * dsxz_dx;
          m_dsyz_dy[q] = b_y[j - 1] * m_dsyz_dy[q] + a_y[j - 1] * dsyz_dy;
          m_dszz_dz[q] = b_z_half[k - 1] * m_dszz_dz[q] + a_z_half[k - 1] * dszz_dz;
          dsxz_dx = dsxz_dx / K_x_half[i - 1] + m_dsxz_dx[q];
          dsyz_dy = dsyz_dy / K_y[j - 1] + m_dsyz_dy[q];
          dszz_dz = dszz_dz / K_z_half[k - 1] + m_dszz_dz[q];
          vz[q] = c_or[q] * (dsxz_dx + dsyz_dy + dszz_dz) + vz[q];
        }
    /* source (3D:1055-1083): force_x, force_y only, no density average */
    if (isrc >= 1 && fx && fy) {
      const size_t q = Q(isrc, jsrc, ksrc);
      vx[q] = vx[q] + fx[it - 1] * dt / rho[q];
      vy[q] = vy[q] + fy[it - 1] * dt / rho[q];
    }
    /* Dirichlet on the six faces (3D:1087-1121) */
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j) {
        vx[Q(1, j, k)] = vy[Q(1, j, k)] = vz[Q(1, j, k)] = 0;
        vx[Q(nx, j, k)] = vy[Q(nx, j, k)] = vz[Q(nx, j, k)] = 0;
      }
    for (int k = 1; k <= nz; ++k)
      for (int i = 1; i <= nx; ++i) {
        vx[Q(i, 1, k)] = vy[Q(i, 1, k)] = vz[Q(i, 1, k)] = 0;
        vx[Q(i, ny, k)] = vy[Q(i, ny, k)] = vz[Q(i, ny, k)] = 0;
      }
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        vx[Q(i, j, 1)] = vy[Q(i, j, 1)] = vz[Q(i, j, 1)] = 0;
        vx[Q(i, j, nz)] = vy[Q(i, j, nz)] = vz[Q(i, j, nz)] = 0;
      }
    /* seismograms (3D:1123-1129) */
    for (int r = 0; r < nrec; ++r) {
      const size_t q = Q(ixr[r], iyr[r], izr[r]);
      sisvx[(size_t)(it - 1) + (size_t)r * ldseis] = vx[q];
      sisvy[(size_t)(it - 1) + (size_t)r * ldseis] = vy[q];
      if (sisvz) sisvz[(size_t)(it - 1) + (size_t)r * ldseis] = vz[q];
    }
    /* energy (3D:1131-1180) */
    if (ewin && etotal) {
      const int i0 = ewin[0], i1 = ewin[1], j0 = ewin[2], j1 = ewin[3], k0 = ewin[4], k1 = ewin[5];
      const int nsl = energy_slabs > 0 ? energy_slabs : 1;
      const int nzl = nz / nsl;
      REAL tot = 0;
      for (int s = 0; s < nsl; ++s) {
        const int ka = s * nzl + 1 > k0 ? s * nzl + 1 : k0;
        const int kb = (s + 1) * nzl < k1 ? (s + 1) * nzl : k1;
        REAL ek = 0, ep = 0;
        for (int k = ka; k <= kb; ++k)
          for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
              const size_t q = Q(i, j, k);
              const REAL r = rho[q], l = lambda[q], m = mu[q];
              ek = ek + LIT(0.5) * r * (vx[q] * vx[q] + vy[q] * vy[q] + vz[q] * vz[q]);
              const REAL den = LIT(2.0) * m * (LIT(3.0) * l + LIT(2.0) * m);
              const REAL exx =
                  (LIT(2.0) * (l + m) * sxx[q] - l * syy[q] - l * szz[q]) / den;
              const REAL eyy =
                  (LIT(2.0) * (l + m) * syy[q] - l * sxx[q] - l * szz[q]) / den;
              const REAL exy = sxy[q] / (LIT(2.0) * m);
              const REAL exz = sxz[q] / (LIT(2.0) * m);
              const REAL eyz = syz[q] / (LIT(2.0) * m);
              ep = ep + LIT(0.5) * (exx * sxx[q] + eyy * syy[q] + eyy * syy[q] +
                                    LIT(2.0) * exy * sxy[q] + LIT(2.0) * exz * sxz[q] +
                                    LIT(2.0) * eyz * syz[q]);
            }
        tot = s == 0 ? ek + ep : tot + (ek + ep);
      }
      etotal[it - 1] = tot;
    }
  }
#undef Q
  free(cl);
  return 0;
}

/* --------------------------------------------------------------------------
 * 3D, 4th order in space: the elastic part of
 * seismic_CPML_3D_viscoelastic_MPI.f90 (3D-visco), i.e. that program with
 * tau_epsilon = tau_sigma (phi = 0, Mu_nu = 1: 3D-visco:466-477), in which
 * the memory variables e* stay zero and only the unrelaxed update remains.
 *   stress   3D-visco:977-1229, velocity 3D-visco:1244-1327,
 *   source   3D-visco:1329-1353, Dirichlet 3D-visco:1355-1398 (two planes
 *   per face), seismograms 3D-visco:1371-1378, energy 3D-visco:1380-1423.
 * Staggered 4th-order difference: (27 f(+) - 27 f(-) - f(++) + f(--)) *
 * ONE_OVER_DELTA / 24, summed left to right as written there.  Loop ranges
 * as 3D-visco in global indices (NPROC slabs concatenated):
 *   sxx syy szz: i 1..NX-1, j 2..NY, k 2..NZ     sxy: i 2..NX, j 1..NY-1, all k
 *   sxz: i 2..NX, j 1..NY, k 1..NZ-1              syz: i 1..NX, j 1..NY-1, k 1..NZ-1
 *   vx: i 2..NX, j 2..NY, k 2..NZ    vy: i 1..NX-1, j 1..NY-1, k 2..NZ
 *   vz: i 1..NX-1, j 2..NY, k 1..NZ-1
 * Lame coefficients as 3D-visco:981-985 with Mu_nu = 1:
 *   lambda_u = (lambda + (2/3) mu) - (2/3) mu, lambda_u + 2 mu, mu.
 * zcut > 0 reproduces the reference's MPI run with a slab interface after
 * global plane zcut: its halo exchange sends two planes in one direction
 * per field (3D-visco:963-975, 1230-1244), so across the interface it reads
 * a never-filled (zero) ghost plane for vz(k+1) in sxx/syy/szz and for
 * sxz(k+1), syz(k+1) in vx, vy on the lower side (k = zcut), and for
 * vx(k-1), vy(k-1) in sxz, syz and szz(k-1) in vz on the upper side
 * (k = zcut+1).  zcut = 0 is the single-domain stencil the GPU computes.
 * Reads outside 1..N use the zero ghost planes (arrays 0:N+1 in x and y,
 * -1:NZ_LOCAL+2 in z).
 * ------------------------------------------------------------------------ */
int ORC(run3d4)(int nx, int ny, int nz, REAL dx, REAL dy, REAL dz, REAL dt, int it_begin,
                int it_end, const REAL *lambda, const REAL *mu, const REAL *rho,
                const REAL *prof_x, const REAL *prof_y, const REAL *prof_z, int isrc, int jsrc,
                int ksrc, const REAL *fx, const REAL *fy, int nrec, const int *ixr,
                const int *iyr, const int *izr, int ldseis, REAL *sisvx, REAL *sisvy,
                const int *ewin, int energy_slabs, REAL *etotal, int zcut, REAL *state,
                int nthreads)
{
  if (nx < 4 || ny < 4 || nz < 4) return 1;
  const size_t nn = (size_t)nx * ny * nz;
  /* ghost-padded copies of the nine fields: 2 planes each side, every axis */
  const int G = 2;
  const size_t px = (size_t)nx + 2 * G, py = (size_t)ny + 2 * G, pz = (size_t)nz + 2 * G;
  const size_t pn = px * py * pz;
  REAL *buf = (REAL *)calloc(pn * 9, sizeof(REAL));
  if (!buf) return 2;
  REAL *f[9];
  for (int q = 0; q < 9; ++q) f[q] = buf + (size_t)q * pn;
  REAL *vx = f[0], *vy = f[1], *vz = f[2], *sxx = f[3], *syy = f[4], *szz = f[5];
  REAL *sxy = f[6], *sxz = f[7], *syz = f[8];
  const size_t sy = px, sz = px * py;
#define P(i, j, k) ((size_t)((i) - 1 + G) + (size_t)((j) - 1 + G) * sy + (size_t)((k) - 1 + G) * sz)
#define Q(i, j, k) ((size_t)((i) - 1) + (size_t)((j) - 1) * nx + (size_t)((k) - 1) * (size_t)nx * ny)
  for (int q = 0; q < 9; ++q)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) f[q][P(i, j, k)] = state[(size_t)q * nn + Q(i, j, k)];
  REAL *m_dvx_dx = state + 9 * nn, *m_dvx_dy = state + 10 * nn, *m_dvx_dz = state + 11 * nn;
  REAL *m_dvy_dx = state + 12 * nn, *m_dvy_dy = state + 13 * nn, *m_dvy_dz = state + 14 * nn;
  REAL *m_dvz_dx = state + 15 * nn, *m_dvz_dy = state + 16 * nn, *m_dvz_dz = state + 17 * nn;
  REAL *m_dsxx_dx = state + 18 * nn, *m_dsyy_dy = state + 19 * nn, *m_dszz_dz = state + 20 * nn;
  REAL *m_dsxy_dx = state + 21 * nn, *m_dsxy_dy = state + 22 * nn, *m_dsxz_dx = state + 23 * nn;
  REAL *m_dsxz_dz = state + 24 * nn, *m_dsyz_dy = state + 25 * nn, *m_dsyz_dz = state + 26 * nn;
  const REAL *a_x = prof_x, *b_x = prof_x + nx, *K_x = prof_x + 2 * nx;
  const REAL *a_x_half = prof_x + 3 * nx, *b_x_half = prof_x + 4 * nx, *K_x_half = prof_x + 5 * nx;
  const REAL *a_y = prof_y, *b_y = prof_y + ny, *K_y = prof_y + 2 * ny;
  const REAL *a_y_half = prof_y + 3 * ny, *b_y_half = prof_y + 4 * ny, *K_y_half = prof_y + 5 * ny;
  const REAL *a_z = prof_z, *b_z = prof_z + nz, *K_z = prof_z + 2 * nz;
  const REAL *a_z_half = prof_z + 3 * nz, *b_z_half = prof_z + 4 * nz, *K_z_half = prof_z + 5 * nz;
  const REAL odx = LIT(1.0) / dx, ody = LIT(1.0) / dy, odz = LIT(1.0) / dz; /* 3D-visco:162-164 */
  const REAL c23 = LIT(2.0) / LIT(3.0);                                      /* 2.d0/DIM */
  const REAL Z = LIT(0.0);
  /* 4th-order staggered difference, 3D-visco:987 and the like */
#define D4(fp, fm, fpp, fmm, od) ((((LIT(27.0) * (fp) - LIT(27.0) * (fm)) - (fpp)) + (fmm)) * (od) / LIT(24.0))

#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

  for (int it = it_begin; it <= it_end; ++it) {
    /* normal stresses (3D-visco:977-1078) */
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nz; ++k)
      for (int j = 2; j <= ny; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          const REAL m = mu[q], l = lambda[q];
          const REAL lu = (l + c23 * m) - c23 * m;
          const REAL l2u = lu + LIT(2.0) * m;
          const REAL vzp = (zcut > 0 && k == zcut) ? Z : vz[p + sz];
          REAL dvx_dx = D4(vx[p + 1], vx[p], vx[p + 2], vx[p - 1], odx);
          REAL dvy_dy = D4(vy[p], vy[p - sy], vy[p + sy], vy[p - 2 * sy], ody);
          REAL dvz_dz = D4(vz[p], vz[p - sz], vzp, vz[p - 2 * sz], odz);
          m_dvx_dx[q] = b_x_half[i - 1] * m_dvx_dx[q] + a_x_half[i - 1] * dvx_dx;
          m_dvy_dy[q] = b_y[j - 1] * m_dvy_dy[q] + a_y[j - 1] * dvy_dy;
          m_dvz_dz[q] = b_z[k - 1] * m_dvz_dz[q] + a_z[k - 1] * dvz_dz;
          dvx_dx = dvx_dx / K_x_half[i - 1] + m_dvx_dx[q];
          dvy_dy = dvy_dy / K_y[j - 1] + m_dvy_dy[q];
          dvz_dz = dvz_dz / K_z[k - 1] + m_dvz_dz[q];
          sxx[p] = sxx[p] + (l2u * dvx_dx + lu * dvy_dy + lu * dvz_dz) * dt;
          syy[p] = syy[p] + (lu * dvx_dx + l2u * dvy_dy + lu * dvz_dz) * dt;
          szz[p] = szz[p] + (lu * dvx_dx + lu * dvy_dy + l2u * dvz_dz) * dt;
        }
    /* sigma_xy (3D-visco:1080-1139) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          REAL dvy_dx = D4(vy[p], vy[p - 1], vy[p + 1], vy[p - 2], odx);
          REAL dvx_dy = D4(vx[p + sy], vx[p], vx[p + 2 * sy], vx[p - sy], ody);
          m_dvy_dx[q] = b_x[i - 1] * m_dvy_dx[q] + a_x[i - 1] * dvy_dx;
          m_dvx_dy[q] = b_y_half[j - 1] * m_dvx_dy[q] + a_y_half[j - 1] * dvx_dy;
          dvy_dx = dvy_dx / K_x[i - 1] + m_dvy_dx[q];
          dvx_dy = dvx_dy / K_y_half[j - 1] + m_dvx_dy[q];
          sxy[p] = sxy[p] + mu[q] * (dvx_dy + dvy_dx) * dt;
        }
    /* sigma_xz, sigma_yz (3D-visco:1141-1226) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) {
      const int cut = zcut > 0 && k == zcut + 1;
      for (int j = 1; j <= ny; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          REAL dvz_dx = D4(vz[p], vz[p - 1], vz[p + 1], vz[p - 2], odx);
          REAL dvx_dz = D4(vx[p + sz], vx[p], vx[p + 2 * sz], cut ? Z : vx[p - sz], odz);
          m_dvz_dx[q] = b_x[i - 1] * m_dvz_dx[q] + a_x[i - 1] * dvz_dx;
          m_dvx_dz[q] = b_z_half[k - 1] * m_dvx_dz[q] + a_z_half[k - 1] * dvx_dz;
          dvz_dx = dvz_dx / K_x[i - 1] + m_dvz_dx[q];
          dvx_dz = dvx_dz / K_z_half[k - 1] + m_dvx_dz[q];
          sxz[p] = sxz[p] + mu[q] * (dvx_dz + dvz_dx) * dt;
        }
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          REAL dvz_dy = D4(vz[p + sy], vz[p], vz[p + 2 * sy], vz[p - sy], ody);
          REAL dvy_dz = D4(vy[p + sz], vy[p], vy[p + 2 * sz], cut ? Z : vy[p - sz], odz);
          m_dvz_dy[q] = b_y_half[j - 1] * m_dvz_dy[q] + a_y_half[j - 1] * dvz_dy;
          m_dvy_dz[q] = b_z_half[k - 1] * m_dvy_dz[q] + a_z_half[k - 1] * dvy_dz;
          dvz_dy = dvz_dy / K_y_half[j - 1] + m_dvz_dy[q];
          dvy_dz = dvy_dz / K_z_half[k - 1] + m_dvy_dz[q];
          syz[p] = syz[p] + mu[q] * (dvy_dz + dvz_dy) * dt;
        }
    }
    /* vx, vy (3D-visco:1244-1285) */
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nz; ++k) {
      const int cut = zcut > 0 && k == zcut;
      for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          REAL dsxx_dx = D4(sxx[p], sxx[p - 1], sxx[p + 1], sxx[p - 2], odx);
          REAL dsxy_dy = D4(sxy[p], sxy[p - sy], sxy[p + sy], sxy[p - 2 * sy], ody);
          REAL dsxz_dz = D4(sxz[p], sxz[p - sz], cut ? Z : sxz[p + sz], sxz[p - 2 * sz], odz);
          m_dsxx_dx[q] = b_x[i - 1] * m_dsxx_dx[q] + a_x[i - 1] * dsxx_dx;
          m_dsxy_dy[q] = b_y[j - 1] * m_dsxy_dy[q] + a_y[j - 1] * dsxy_dy;
          m_dsxz_dz[q] = b_z[k - 1] * m_dsxz_dz[q] + a_z[k - 1] * dsxz_dz;
          dsxx_dx = dsxx_dx / K_x[i - 1] + m_dsxx_dx[q];
          dsxy_dy = dsxy_dy / K_y[j - 1] + m_dsxy_dy[q];
          dsxz_dz = dsxz_dz / K_z[k - 1] + m_dsxz_dz[q];
          vx[p] = dt / rho[q] * (dsxx_dx + dsxy_dy + dsxz_dz) + vx[p];
        }
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          REAL dsxy_dx = D4(sxy[p + 1], sxy[p], sxy[p + 2], sxy[p - 1], odx);
          REAL dsyy_dy = D4(syy[p + sy], syy[p], syy[p + 2 * sy], syy[p - sy], ody);
          REAL dsyz_dz = D4(syz[p], syz[p - sz], cut ? Z : syz[p + sz], syz[p - 2 * sz], odz);
          m_dsxy_dx[q] = b_x_half[i - 1] * m_dsxy_dx[q] + a_x_half[i - 1] * dsxy_dx;
          m_dsyy_dy[q] = b_y_half[j - 1] * m_dsyy_dy[q] + a_y_half[j - 1] * dsyy_dy;
          m_dsyz_dz[q] = b_z[k - 1] * m_dsyz_dz[q] + a_z[k - 1] * dsyz_dz;
          dsxy_dx = dsxy_dx / K_x_half[i - 1] + m_dsxy_dx[q];
          dsyy_dy = dsyy_dy / K_y_half[j - 1] + m_dsyy_dy[q];
          dsyz_dz = dsyz_dz / K_z[k - 1] + m_dsyz_dz[q];
          vy[p] = dt / rho[q] * (dsxy_dx + dsyy_dy + dsyz_dz) + vy[p];
        }
    }
    /* vz (3D-visco:1287-1305) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) {
      const int cut = zcut > 0 && k == zcut + 1;
      for (int j = 2; j <= ny; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t p = P(i, j, k), q = Q(i, j, k);
          REAL dsxz_dx = D4(sxz[p + 1], sxz[p], sxz[p + 2], sxz[p - 1], odx);
          REAL dsyz_dy = D4(syz[p], syz[p - sy], syz[p + sy], syz[p - 2 * sy], ody);
          REAL dszz_dz = D4(szz[p + sz], szz[p], szz[p + 2 * sz], cut ? Z : szz[p - sz], odz);
          m_dsxz_dx[q] = b_x_half[i - 1] * m_dsxz_dx[q] + a_x_half[i - 1] * dsxz_dx;
          m_dsyz_dy[q] = b_y[j - 1] * m_dsyz_dy[q] + a_y[j - 1] * dsyz_dy;
          m_dszz_dz[q] = b_z_half[k - 1] * m_dszz_dz[q] + a_z_half[k - 1] * dszz_dz;
          dsxz_dx = dsxz_dx / K_x_half[i - 1] + m_dsxz_dx[q];
          dsyz_dy = dsyz_dy / K_y[j - 1] + m_dsyz_dy[q];
          dszz_dz = dszz_dz / K_z_half[k - 1] + m_dszz_dz[q];
          vz[p] = dt / rho[q] * (dsxz_dx + dsyz_dy + dszz_dz) + vz[p];
        }
    }
    /* source (3D-visco:1329-1353) */
    if (isrc >= 1 && fx && fy) {
      const size_t p = P(isrc, jsrc, ksrc), q = Q(isrc, jsrc, ksrc);
      vx[p] = vx[p] + fx[it - 1] * dt / rho[q];
      vy[p] = vy[p] + fy[it - 1] * dt / rho[q];
    }
    /* Dirichlet: planes 1 and N of every axis (the ghost planes stay zero),
     * 3D-visco:1355-1398 */
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
            const size_t p = P(i, j, k);
            vx[p] = vy[p] = vz[p] = Z;
          }
    /* seismograms (3D-visco:1371-1378) */
    for (int r = 0; r < nrec; ++r) {
      const size_t p = P(ixr[r], iyr[r], izr[r]);
      sisvx[(size_t)(it - 1) + (size_t)r * ldseis] = vx[p];
      sisvy[(size_t)(it - 1) + (size_t)r * ldseis] = vy[p];
    }
    /* energy (3D-visco:1380-1429): per-slab partial sums added in slab
     * order; sigma_R = sigma in the elastic limit up to the rounding of the
     * relaxed Lame coefficients */
    if (ewin && etotal) {
      const int i0 = ewin[0], i1 = ewin[1], j0 = ewin[2], j1 = ewin[3], k0 = ewin[4], k1 = ewin[5];
      const int nsl = energy_slabs > 0 ? energy_slabs : 1;
      const int nzl = nz / nsl;
      REAL tot = 0;
      for (int s = 0; s < nsl; ++s) {
        const int ka = s * nzl + 1 > k0 ? s * nzl + 1 : k0;
        const int kb = (s + 1) * nzl < k1 ? (s + 1) * nzl : k1;
        REAL ek = 0, ep = 0;
        for (int k = ka; k <= kb; ++k)
          for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
              const size_t p = P(i, j, k), q = Q(i, j, k);
              const REAL r = rho[q], l = lambda[q], m = mu[q];
              ek = ek + LIT(0.5) * r * (vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p]);
              const REAL den = LIT(2.0) * m * (LIT(3.0) * l + LIT(2.0) * m);
              const REAL exx = (LIT(2.0) * (l + m) * sxx[p] - l * syy[p] - l * szz[p]) / den;
              const REAL eyy = (LIT(2.0) * (l + m) * syy[p] - l * sxx[p] - l * szz[p]) / den;
              const REAL exy = sxy[p] / (LIT(2.0) * m);
              const REAL exz = sxz[p] / (LIT(2.0) * m);
              const REAL eyz = syz[p] / (LIT(2.0) * m);
              ep = ep + LIT(0.5) * (exx * sxx[p] + eyy * syy[p] + eyy * syy[p] +
                                    LIT(2.0) * exy * sxy[p] + LIT(2.0) * exz * sxz[p] +
                                    LIT(2.0) * eyz * syz[p]);
            }
        tot = s == 0 ? ek + ep : tot + (ek + ep);
      }
      etotal[it - 1] = tot;
    }
  }
  for (int q = 0; q < 9; ++q)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) state[(size_t)q * nn + Q(i, j, k)] = f[q][P(i, j, k)];
#undef D4
#undef P
#undef Q
  free(buf);
  return 0;
}

/* --------------------------------------------------------------------------
 * 2D viscoelastic C-PML: seismic_CPML_2D_velocity_and_stress_
 * second_order_viscoelastic.f90 (2D-visco2; order 2) and
 * ..._fourth_order_viscoelastic.f90 (2D-visco4; order 4, which differs only
 * in its difference operator and its 0:NX+1 ghost-padded fields) with
 * VISCOELASTIC_ATTENUATION.
 *   coefficients 2D-visco2:374-381 (from tau_epsilon, tau_sigma per SLS),
 *   stress + ADE memory variables 2D-visco2:736-833 (Robertsson et al.
 *   1994 eq. 13-14: R_new = (R_old + d*DELTAT_phi - R_old*DELTAT/(2 tau_s))
 *   * 1/(1 + DELTAT/(2 tau_s)), stress increment uses (R_new + R_old)/2),
 *   velocity 2D-visco2:839-870, source 2D-visco2:873-918 (force table given),
 *   Dirichlet 2D-visco2:921-929, seismograms and pressure 2D-visco2:932-990.
 * Derivatives multiply by ONE_OVER_DELTA (2D-visco2:209-210).  The memory
 * variables R are updated in place: each point reads only its own old value.
 * state: vx vy sxx syy sxy, 8 C-PML memories (as run2d), then R_e1, R_e11,
 * R_e13, each nsls planes; all (ny, nx).
 * ------------------------------------------------------------------------ */
int ORC(run2dve)(int nx, int ny, int order, REAL dx, REAL dy, REAL dt, int it_begin, int it_end,
                 const REAL *lambda, const REAL *mu, const REAL *rho, const REAL *prof_x,
                 const REAL *prof_y, int nsls, const REAL *tau_eps1, const REAL *tau_sig1,
                 const REAL *tau_eps2, const REAL *tau_sig2, int isrc, int jsrc,
                 const REAL *fx, const REAL *fy, int nrec, const int *ixr, const int *iyr,
                 int ldseis, REAL *sisvx, REAL *sisvy, REAL *sisp, REAL *state, int nthreads)
{
  if (nx < 2 || ny < 2 || nsls < 1 || nsls > 8 || (order != 2 && order != 4)) return 1;
  const size_t nn = (size_t)nx * ny;
  /* ghost-padded copies of the five fields (two cells each side; the
   * 4th-order program's arrays are 0:NX+1, 0:NY+1 and its ghosts stay 0) */
  const int G = 2;
  const size_t px = (size_t)nx + 2 * G, pn = px * ((size_t)ny + 2 * G);
  REAL *buf = (REAL *)calloc(pn * 5, sizeof(REAL));
  if (!buf) return 2;
  REAL *vx = buf, *vy = buf + pn, *sxx = buf + 2 * pn, *syy = buf + 3 * pn, *sxy = buf + 4 * pn;
  for (int f = 0; f < 5; ++f)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        buf[(size_t)f * pn + (size_t)(i + G) + (size_t)(j + G) * px] =
            state[(size_t)f * nn + (size_t)i + (size_t)j * nx];
  REAL *m_dvx_dx = state + 5 * nn, *m_dvx_dy = state + 6 * nn, *m_dvy_dx = state + 7 * nn;
  REAL *m_dvy_dy = state + 8 * nn, *m_dsxx_dx = state + 9 * nn, *m_dsyy_dy = state + 10 * nn;
  REAL *m_dsxy_dx = state + 11 * nn, *m_dsxy_dy = state + 12 * nn;
  REAL *re1 = state + 13 * nn, *re11 = re1 + (size_t)nsls * nn, *re13 = re11 + (size_t)nsls * nn;
  const REAL *a_x = prof_x, *b_x = prof_x + nx, *K_x = prof_x + 2 * nx;
  const REAL *a_x_half = prof_x + 3 * nx, *b_x_half = prof_x + 4 * nx, *K_x_half = prof_x + 5 * nx;
  const REAL *a_y = prof_y, *b_y = prof_y + ny, *K_y = prof_y + 2 * ny;
  const REAL *a_y_half = prof_y + 3 * ny, *b_y_half = prof_y + 4 * ny, *K_y_half = prof_y + 5 * ny;
  const REAL odx = LIT(1.0) / dx, ody = LIT(1.0) / dy; /* 2D-visco2:209-210 */
  /* 2D-visco4: NINE_OVER_8_DELTA, ONE_OVER_24_DELTA */
  const REAL n8x = LIT(9.0) / (LIT(8.0) * dx), n8y = LIT(9.0) / (LIT(8.0) * dy);
  const REAL o24x = LIT(1.0) / (LIT(24.0) * dx), o24y = LIT(1.0) / (LIT(24.0) * dy);
  const REAL two_thirds = LIT(2.0) / LIT(3.0);          /* 2D-visco2:221 */
  /* 2D-visco2:374-381 */
  REAL H1[8], H2[8], M1[8], M2[8], P1[8], P2[8], S1 = 0, S2 = 0;
  for (int l = 0; l < nsls; ++l) {
    S1 = S1 + tau_eps1[l] / tau_sig1[l];
    S2 = S2 + tau_eps2[l] / tau_sig2[l];
  }
  for (int l = 0; l < nsls; ++l) {
    const REAL o1 = LIT(1.0) / tau_sig1[l], o2 = LIT(1.0) / tau_sig2[l];
    H1[l] = LIT(0.5) * dt / tau_sig1[l];
    H2[l] = LIT(0.5) * dt / tau_sig2[l];
    M1[l] = LIT(1.0) / (LIT(1.0) + LIT(0.5) * dt * o1);
    M2[l] = LIT(1.0) / (LIT(1.0) + LIT(0.5) * dt * o2);
    P1[l] = dt * (LIT(1.0) - tau_eps1[l] / tau_sig1[l]) / tau_sig1[l] / S1;
    P2[l] = dt * (LIT(1.0) - tau_eps2[l] / tau_sig2[l]) / tau_sig2[l] / S2;
  }
#define I2(i, j) ((size_t)((i) - 1) + (size_t)((j) - 1) * nx)
#define F2(i, j) ((size_t)((i) - 1 + G) + (size_t)((j) - 1 + G) * px)
/* staggered difference: (f+ - f0) * ONE_OVER_DELTA (order 2), or
 * (f+ - f0) * NINE_OVER_8_DELTA + (f- - f++) * ONE_OVER_24_DELTA (order 4) */
#define DV(fp, f0, fm, fpp, od, n8, o24) \
  (order == 2 ? ((fp) - (f0)) * (od) : ((fp) - (f0)) * (n8) + ((fm) - (fpp)) * (o24))
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int it = it_begin; it <= it_end; ++it) {
    /* sigma_xx, sigma_yy (2D-visco2:742-800) */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        const size_t q = I2(i, j), f = F2(i, j);
        const REAL lh = LIT(0.5) * (lambda[q + 1] + lambda[q]);
        const REAL mh = LIT(0.5) * (mu[q + 1] + mu[q]);
        const REAL lpm = lh + mh;
        const REAL l2m = lh + LIT(2.0) * mh;
        REAL dvx_dx = DV(vx[f + 1], vx[f], vx[f - 1], vx[f + 2], odx, n8x, o24x);
        REAL dvy_dy = DV(vy[f], vy[f - px], vy[f - 2 * px], vy[f + px], ody, n8y, o24y);
        m_dvx_dx[q] = b_x_half[i - 1] * m_dvx_dx[q] + a_x_half[i - 1] * dvx_dx;
        m_dvy_dy[q] = b_y[j - 1] * m_dvy_dy[q] + a_y[j - 1] * dvy_dy;
        dvx_dx = dvx_dx / K_x_half[i - 1] + m_dvx_dx[q];
        dvy_dy = dvy_dy / K_y[j - 1] + m_dvy_dy[q];
        REAL se1 = 0, se11 = 0;
        for (int l = 0; l < nsls; ++l) {
          REAL *r1 = re1 + (size_t)l * nn + q, *r11 = re11 + (size_t)l * nn + q;
          const REAL o1 = *r1, o11 = *r11;
          const REAL n1 = (o1 + (dvx_dx + dvy_dy) * P1[l] - o1 * H1[l]) * M1[l];
          const REAL n11 = (o11 + LIT(0.5) * (dvx_dx - dvy_dy) * P2[l] - o11 * H2[l]) * M2[l];
          *r1 = n1;
          *r11 = n11;
          se1 = se1 + n1 + o1;
          se11 = se11 + n11 + o11;
        }
        sxx[f] = sxx[f] + (l2m * dvx_dx + lh * dvy_dy +
                           (LIT(0.5) * lpm * se1 + mh * se11)) * dt;
        syy[f] = syy[f] + (lh * dvx_dx + l2m * dvy_dy +
                           (LIT(0.5) * lpm * se1 - mh * se11)) * dt;
      }
    /* sigma_xy (2D-visco2:802-833) */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 2; i <= nx; ++i) {
        const size_t q = I2(i, j), f = F2(i, j);
        const REAL mhy = LIT(0.5) * (mu[q + nx] + mu[q]);
        REAL dvy_dx = DV(vy[f], vy[f - 1], vy[f - 2], vy[f + 1], odx, n8x, o24x);
        REAL dvx_dy = DV(vx[f + px], vx[f], vx[f - px], vx[f + 2 * px], ody, n8y, o24y);
        m_dvy_dx[q] = b_x[i - 1] * m_dvy_dx[q] + a_x[i - 1] * dvy_dx;
        m_dvx_dy[q] = b_y_half[j - 1] * m_dvx_dy[q] + a_y_half[j - 1] * dvx_dy;
        dvy_dx = dvy_dx / K_x[i - 1] + m_dvy_dx[q];
        dvx_dy = dvx_dy / K_y_half[j - 1] + m_dvx_dy[q];
        REAL se13 = 0;
        for (int l = 0; l < nsls; ++l) {
          REAL *r13 = re13 + (size_t)l * nn + q;
          const REAL o13 = *r13;
          const REAL n13 = (o13 + (dvy_dx + dvx_dy) * P2[l] - o13 * H2[l]) * M2[l];
          *r13 = n13;
          se13 = se13 + n13 + o13;
        }
        sxy[f] = sxy[f] + mhy * (dvy_dx + dvx_dy + LIT(0.5) * se13) * dt;
      }
    /* vx (2D-visco2:839-851) */
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= ny; ++j)
      for (int i = 2; i <= nx; ++i) {
        const size_t q = I2(i, j), f = F2(i, j);
        REAL dsxx_dx = DV(sxx[f], sxx[f - 1], sxx[f - 2], sxx[f + 1], odx, n8x, o24x);
        REAL dsxy_dy = DV(sxy[f], sxy[f - px], sxy[f - 2 * px], sxy[f + px], ody, n8y, o24y);
        m_dsxx_dx[q] = b_x[i - 1] * m_dsxx_dx[q] + a_x[i - 1] * dsxx_dx;
        m_dsxy_dy[q] = b_y[j - 1] * m_dsxy_dy[q] + a_y[j - 1] * dsxy_dy;
        dsxx_dx = dsxx_dx / K_x[i - 1] + m_dsxx_dx[q];
        dsxy_dy = dsxy_dy / K_y[j - 1] + m_dsxy_dy[q];
        vx[f] = vx[f] + (dsxx_dx + dsxy_dy) * dt / rho[q];
      }
    /* vy (2D-visco2:853-870) */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        const size_t q = I2(i, j), f = F2(i, j);
        const REAL rh = LIT(0.25) * (rho[q] + rho[q + 1] + rho[q + 1 + nx] + rho[q + nx]);
        REAL dsxy_dx = DV(sxy[f + 1], sxy[f], sxy[f - 1], sxy[f + 2], odx, n8x, o24x);
        REAL dsyy_dy = DV(syy[f + px], syy[f], syy[f - px], syy[f + 2 * px], ody, n8y, o24y);
        m_dsxy_dx[q] = b_x_half[i - 1] * m_dsxy_dx[q] + a_x_half[i - 1] * dsxy_dx;
        m_dsyy_dy[q] = b_y_half[j - 1] * m_dsyy_dy[q] + a_y_half[j - 1] * dsyy_dy;
        dsxy_dx = dsxy_dx / K_x_half[i - 1] + m_dsxy_dx[q];
        dsyy_dy = dsyy_dy / K_y_half[j - 1] + m_dsyy_dy[q];
        vy[f] = vy[f] + (dsxy_dx + dsyy_dy) * dt / rh;
      }
    /* source (2D-visco2:913-918): fx, fy are force_x, force_y of the step */
    if (isrc >= 1 && fx && fy) {
      const size_t q = I2(isrc, jsrc), f = F2(isrc, jsrc);
      const REAL rh = LIT(0.25) * (rho[q] + rho[q + 1] + rho[q + 1 + nx] + rho[q + nx]);
      vx[f] = vx[f] + fx[it - 1] * dt / rho[q];
      vy[f] = vy[f] + fy[it - 1] * dt / rh;
    }
    /* Dirichlet (2D-visco2:921-929) */
    for (int j = 1; j <= ny; ++j) {
      vx[F2(1, j)] = vy[F2(1, j)] = 0;
      vx[F2(nx, j)] = vy[F2(nx, j)] = 0;
    }
    for (int i = 1; i <= nx; ++i) {
      vx[F2(i, 1)] = vy[F2(i, 1)] = 0;
      vx[F2(i, ny)] = vy[F2(i, ny)] = 0;
    }
    /* seismograms and pressure (2D-visco2:932-990) */
    for (int r = 0; r < nrec; ++r) {
      const size_t q = I2(ixr[r], iyr[r]), f = F2(ixr[r], iyr[r]);
      sisvx[(size_t)(it - 1) + (size_t)r * ldseis] = vx[f];
      sisvy[(size_t)(it - 1) + (size_t)r * ldseis] = vy[f];
      if (sisp) {
        const REAL lh = LIT(0.5) * (lambda[q + 1] + lambda[q]);
        const REAL mh = LIT(0.5) * (mu[q + 1] + mu[q]);
        const REAL den = LIT(4.0) * mh * (lh + mh);
        const REAL exx = ((lh + LIT(2.0) * mh) * sxx[f] - lh * syy[f]) / den;
        const REAL eyy = ((lh + LIT(2.0) * mh) * syy[f] - lh * sxx[f]) / den;
        sisp[(size_t)(it - 1) + (size_t)r * ldseis] = -((lh + two_thirds * mh) * (exx + eyy));
      }
    }
  }
  for (int f = 0; f < 5; ++f)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        state[(size_t)f * nn + (size_t)i + (size_t)j * nx] =
            buf[(size_t)f * pn + (size_t)(i + G) + (size_t)(j + G) * px];
  free(buf);
#undef DV
#undef F2
#undef I2
  return 0;
}

/* --------------------------------------------------------------------------
 * ADE-PML, RK4 in time, 8th-order (Holberg) staggered operator:
 * seismic_ADEPML_2D_elastic_RK4_eighth_order.f90 (ADE8) time loop 692-906.
 * Per step: dX(2) = dX(3) = X; for each of the 4 stages: dX(1) = dX(3) +
 * rk41 * dX(2) * DELTAT (all arrays), velocity rates from dsigma(1) with the
 * ADE memory update inside the PML band (i <= NPML+2, i >= NX-NPML-2, same
 * in j) using the stage coefficients a(inc), b(inc) (ADE8:718-796), the
 * force at t = (it-1 + rk41) * DELTAT added to the rates (797-802), the rate
 * arrays' borders zeroed at every level (803-810), stress rates from dv(1)
 * (811-861), X += rk42 * dX(2) * DELTAT and the velocity borders zeroed
 * (863-874); then vx is recorded as the mean of 4 nodes, vy at the node
 * (877-884).  Arrays keep the 4-cell ghost border of the program
 * (-4:N+4): index (i+4) + (j+4)*(nx+9) for Fortran (i, j); profiles
 * a, b as [stage][i+4], K as [i+4].  state: vx vy sxx syy sxy and the 8
 * memory arrays, each (ny+9)*(nx+9).
 * ------------------------------------------------------------------------ */
int ORC(runade8)(int nx, int ny, REAL dx, REAL dy, REAL dt, int it_begin, int it_end, int npml,
                 const REAL *lam, const REAL *mu, const REAL *rho, const REAL *ax,
                 const REAL *bx, const REAL *Kx, const REAL *axh, const REAL *bxh,
                 const REAL *Kxh, const REAL *ay, const REAL *by, const REAL *Ky,
                 const REAL *ayh, const REAL *byh, const REAL *Kyh, int isrc, int jsrc, REAL f0,
                 REAL t0, REAL factor, REAL angle_deg, int nrec, const int *ixr, const int *iyr,
                 int ldseis, REAL *sisvx, REAL *sisvy, REAL *state, int nthreads)
{
  if (nx < 10 || ny < 10) return 1;
  const long W = nx + 9, gx = nx + 9, gy = ny + 9;
  const size_t nn = (size_t)gx * gy;
#define A(i, j) ((size_t)((i) + 4) + (size_t)((j) + 4) * W)
  REAL *X[5], *D[3][5];
  for (int f = 0; f < 5; ++f) X[f] = state + (size_t)f * nn;
  REAL *m_dvx_dx = state + 5 * nn, *m_dvx_dy = state + 6 * nn, *m_dvy_dx = state + 7 * nn;
  REAL *m_dvy_dy = state + 8 * nn, *m_dsxx_dx = state + 9 * nn, *m_dsyy_dy = state + 10 * nn;
  REAL *m_dsxy_dx = state + 11 * nn, *m_dsxy_dy = state + 12 * nn;
  REAL *buf = (REAL *)calloc(nn * 15, sizeof(REAL));
  if (!buf) return 2;
  for (int l = 0; l < 3; ++l)
    for (int f = 0; f < 5; ++f) D[l][f] = buf + (size_t)(l * 5 + f) * nn;
  /* ADE8:263-266, 403-410 */
  const REAL c1 = LIT(1.231666), c2 = LIT(-1.041182e-1), c3 = LIT(2.063707e-2),
             c4 = LIT(-3.570998e-3);
  const REAL rk41[4] = {LIT(0.0), LIT(0.5), LIT(0.5), LIT(1.0)};
  const REAL rk42[4] = {LIT(1.0) / LIT(6.0), LIT(2.0) / LIT(6.0), LIT(2.0) / LIT(6.0),
                        LIT(1.0) / LIT(6.0)};
  const REAL PI = LIT(3.141592653589793238462643);
  const REAL DEG2RAD = PI / LIT(180.0);
  /* 8th-order difference, (c1 d1 + c2 d2 + c3 d3 + c4 d4) / delta */
#define H8(d1, d2, d3, d4, del) ((((c1 * (d1) + c2 * (d2)) + c3 * (d3)) + c4 * (d4)) / (del))
#define BAND(i, j) ((i) <= npml + 2 || (i) >= nx - npml - 2 || (j) <= npml + 2 || (j) >= ny - npml - 2)
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int it = it_begin; it <= it_end; ++it) {
    for (int f = 0; f < 5; ++f) {
      memcpy(D[1][f], X[f], nn * sizeof(REAL));
      memcpy(D[2][f], X[f], nn * sizeof(REAL));
    }
    for (int inc = 0; inc < 4; ++inc) {
      for (int f = 0; f < 5; ++f)
        for (size_t q = 0; q < nn; ++q) D[0][f][q] = D[2][f][q] + rk41[inc] * D[1][f][q] * dt;
      REAL *vx1 = D[0][0], *vy1 = D[0][1], *sxx1 = D[0][2], *syy1 = D[0][3], *sxy1 = D[0][4];
      REAL *dvx2 = D[1][0], *dvy2 = D[1][1], *dsxx2 = D[1][2], *dsyy2 = D[1][3], *dsxy2 = D[1][4];
      const REAL *bxs = bx + inc * W, *axs = ax + inc * W, *bxhs = bxh + inc * W,
                 *axhs = axh + inc * W;
      const REAL *bys = by + inc * (ny + 9), *ays = ay + inc * (ny + 9),
                 *byhs = byh + inc * (ny + 9), *ayhs = ayh + inc * (ny + 9);
      /* velocity rates (ADE8:718-796) */
#pragma omp parallel for schedule(static)
      for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t q = A(i, j);
          REAL vx_ = H8(sxx1[q] - sxx1[q - 1], sxx1[q + 1] - sxx1[q - 2], sxx1[q + 2] - sxx1[q - 3],
                        sxx1[q + 3] - sxx1[q - 4], dx);
          REAL vy_ = H8(sxy1[q] - sxy1[q - W], sxy1[q + W] - sxy1[q - 2 * W],
                        sxy1[q + 2 * W] - sxy1[q - 3 * W], sxy1[q + 3 * W] - sxy1[q - 4 * W], dy);
          if (BAND(i, j)) {
            m_dsxx_dx[q] = bxs[i + 4] * m_dsxx_dx[q] + axs[i + 4] * vx_;
            m_dsxy_dy[q] = bys[j + 4] * m_dsxy_dy[q] + ays[j + 4] * vy_;
            vx_ = vx_ / Kx[i + 4] + m_dsxx_dx[q];
            vy_ = vy_ / Ky[j + 4] + m_dsxy_dy[q];
          }
          dvx2[q] = (vx_ + vy_) / rho[q];
        }
#pragma omp parallel for schedule(static)
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t q = A(i, j);
          const REAL rh = LIT(0.25) * (rho[q] + rho[q + 1] + rho[q + 1 + W] + rho[q + W]);
          REAL sx_ = H8(sxy1[q + 1] - sxy1[q], sxy1[q + 2] - sxy1[q - 1], sxy1[q + 3] - sxy1[q - 2],
                        sxy1[q + 4] - sxy1[q - 3], dx);
          REAL sy_ = H8(syy1[q + W] - syy1[q], syy1[q + 2 * W] - syy1[q - W],
                        syy1[q + 3 * W] - syy1[q - 2 * W], syy1[q + 4 * W] - syy1[q - 3 * W], dy);
          if (BAND(i, j)) {
            m_dsxy_dx[q] = bxhs[i + 4] * m_dsxy_dx[q] + axhs[i + 4] * sx_;
            m_dsyy_dy[q] = byhs[j + 4] * m_dsyy_dy[q] + ayhs[j + 4] * sy_;
            sx_ = sx_ / Kxh[i + 4] + m_dsxy_dx[q];
            sy_ = sy_ / Kyh[j + 4] + m_dsyy_dy[q];
          }
          dvy2[q] = (sx_ + sy_) / rh;
        }
      /* force at the stage time (ADE8:786-802) */
      {
        const REAL a = PI * PI * f0 * f0;
        const REAL t = ((REAL)(it - 1) + rk41[inc]) * dt;
        const REAL tt = t - t0;
        const REAL st = -(factor * LIT(2.0) * a * tt * EXP(-(a * (tt * tt))));
        const REAL fx = SIN(angle_deg * DEG2RAD) * st;
        const REAL fy = COS(angle_deg * DEG2RAD) * st;
        const size_t q = A(isrc, jsrc);
        const REAL rh = LIT(0.25) * (rho[q] + rho[q + 1] + rho[q + 1 + W] + rho[q + W]);
        dvx2[q] = dvx2[q] + fx / rho[q];
        dvy2[q] = dvy2[q] + fy / rh;
      }
      /* rate borders zeroed at all three levels (ADE8:803-810) */
      for (int l = 0; l < 3; ++l)
        for (int f = 0; f < 2; ++f)
          for (int j = -4; j <= ny + 4; ++j)
            for (int i = -4; i <= nx + 4; ++i)
              if (i <= 1 || i >= nx || j <= 1 || j >= ny) D[l][f][A(i, j)] = 0;
      /* stress rates from the stage velocities (ADE8:811-861) */
#pragma omp parallel for schedule(static)
      for (int j = 2; j <= ny; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
          const size_t q = A(i, j);
          const REAL lh = LIT(0.5) * (lam[q + 1] + lam[q]);
          const REAL mh = LIT(0.5) * (mu[q + 1] + mu[q]);
          const REAL l2m = lh + LIT(2.0) * mh;
          REAL dvx_dx = H8(vx1[q + 1] - vx1[q], vx1[q + 2] - vx1[q - 1], vx1[q + 3] - vx1[q - 2],
                           vx1[q + 4] - vx1[q - 3], dx);
          REAL dvy_dy = H8(vy1[q] - vy1[q - W], vy1[q + W] - vy1[q - 2 * W],
                           vy1[q + 2 * W] - vy1[q - 3 * W], vy1[q + 3 * W] - vy1[q - 4 * W], dy);
          if (BAND(i, j)) {
            m_dvx_dx[q] = bxhs[i + 4] * m_dvx_dx[q] + axhs[i + 4] * dvx_dx;
            m_dvy_dy[q] = bys[j + 4] * m_dvy_dy[q] + ays[j + 4] * dvy_dy;
            dvx_dx = dvx_dx / Kxh[i + 4] + m_dvx_dx[q];
            dvy_dy = dvy_dy / Ky[j + 4] + m_dvy_dy[q];
          }
          dsxx2[q] = l2m * dvx_dx + lh * dvy_dy;
          dsyy2[q] = lh * dvx_dx + l2m * dvy_dy;
        }
#pragma omp parallel for schedule(static)
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 2; i <= nx; ++i) {
          const size_t q = A(i, j);
          const REAL mhy = LIT(0.5) * (mu[q + W] + mu[q]);
          REAL dvx_dy = H8(vx1[q + W] - vx1[q], vx1[q + 2 * W] - vx1[q - W],
                           vx1[q + 3 * W] - vx1[q - 2 * W], vx1[q + 4 * W] - vx1[q - 3 * W], dy);
          REAL dvy_dx = H8(vy1[q] - vy1[q - 1], vy1[q + 1] - vy1[q - 2], vy1[q + 2] - vy1[q - 3],
                           vy1[q + 3] - vy1[q - 4], dx);
          if (BAND(i, j)) {
            m_dvy_dx[q] = bxs[i + 4] * m_dvy_dx[q] + axs[i + 4] * dvy_dx;
            m_dvx_dy[q] = byhs[j + 4] * m_dvx_dy[q] + ayhs[j + 4] * dvx_dy;
            dvy_dx = dvy_dx / Kx[i + 4] + m_dvy_dx[q];
            dvx_dy = dvx_dy / Kyh[j + 4] + m_dvx_dy[q];
          }
          dsxy2[q] = mhy * (dvy_dx + dvx_dy);
        }
      /* accumulate and zero the velocity borders (ADE8:863-874) */
      for (int f = 0; f < 5; ++f)
        for (size_t q = 0; q < nn; ++q) X[f][q] = X[f][q] + rk42[inc] * D[1][f][q] * dt;
      for (int f = 0; f < 2; ++f)
        for (int j = -4; j <= ny + 4; ++j)
          for (int i = -4; i <= nx + 4; ++i)
            if (i <= 1 || i >= nx || j <= 1 || j >= ny) X[f][A(i, j)] = 0;
    }
    /* seismograms (ADE8:877-884) */
    for (int r = 0; r < nrec; ++r) {
      const size_t q = A(ixr[r], iyr[r]);
      sisvx[(size_t)(it - 1) + (size_t)r * ldseis] =
          (X[0][q] + X[0][q + 1] + X[0][q + W] + X[0][q + 1 + W]) / LIT(4.0);
      sisvy[(size_t)(it - 1) + (size_t)r * ldseis] = X[1][q];
    }
  }
#undef BAND
#undef H8
#undef A
  free(buf);
  (void)gy;
  return 0;
}

#undef ORC
#undef CAT
#undef CAT_
