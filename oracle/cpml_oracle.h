/*
 * CPU oracle for the seismic C-PML time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran programs' algorithm, used as the checker for the HIP path.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product library never links or calls it.
 *
 * Every function follows the reference file:line cited beside it.  Aliases:
 *   2D2 = seismic_CPML_2D_isotropic_second_order.f90
 *   2D4 = seismic_CPML_2D_isotropic_fourth_order.f90
 *   3D  = seismic_CPML_3D_isotropic_MPI_OpenMP.f90
 *   ADE8 = seismic_ADEPML_2D_elastic_RK4_eighth_order.f90
 *
 * Conventions: all grid arrays are Fortran (column-major) order with the x
 * index fastest, interior points only (nx*ny or nx*ny*nz), 1-based indices
 * in the argument lists (as the reference programs use them).
 *
 * Each routine exists in a double (suffix _d) and a float (suffix _f)
 * version.  The float version is the restatement of the reference compiled
 * in single precision (the reference's own instruction at 2D2:131-133):
 * every constant and intermediate is rounded to float.
 *
 * Pinning: profile, STF, receivers and the 2D/3D loops are checked against
 * the compiled reference's outputs (tests/golden) and against closed forms.
 * The 8th-order operator has no reference trajectory; it is pinned by its
 * closed-form dispersion relation on sine waves and by the 2nd/4th-order
 * paths it shares code with.
 */
#ifndef CPML_ORACLE_H
#define CPML_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_DECLS(REAL, SFX)                                                          \
  void oracle_profile_##SFX(int n, REAL delta, REAL deltat, int npml, REAL cp, REAL f0,  \
                            int use_min, int use_max, int origin_max_uses_n,             \
                            int clamp_alpha, REAL *a, REAL *b, REAL *K, REAL *a_half,    \
                            REAL *b_half, REAL *K_half);                                 \
  void oracle_stf_##SFX(int nstep, REAL deltat, REAL f0, REAL t0, REAL factor,           \
                        REAL angle_deg, int kind, REAL *fx, REAL *fy);                   \
  void oracle_receivers_##SFX(int nx, int ny, REAL dx, REAL dy, int nrec, REAL xdeb,     \
                              REAL ydeb, REAL xfin, REAL yfin, int *ix, int *iy);        \
  int oracle_run2d_##SFX(int nx, int ny, int order, REAL dx, REAL dy, REAL dt,           \
                         int it_begin, int it_end, const REAL *lambda, const REAL *mu,   \
                         const REAL *rho, const REAL *prof_x, const REAL *prof_y,        \
                         int isrc, int jsrc, const REAL *fx, const REAL *fy, int nrec,   \
                         const int *ixr, const int *iyr, int ldseis, REAL *sisvx,        \
                         REAL *sisvy, const int *ewin, REAL *ekin, REAL *epot,           \
                         REAL *state, int nthreads);                                     \
  int oracle_run3d_##SFX(int nx, int ny, int nz, REAL dx, REAL dy, REAL dz, REAL dt,     \
                         int it_begin, int it_end, const REAL *lambda, const REAL *mu,   \
                         const REAL *rho, const REAL *lp2m, const REAL *prof_x,          \
                         const REAL *prof_y, const REAL *prof_z, int isrc, int jsrc,     \
                         int ksrc, const REAL *fx, const REAL *fy, int nrec,             \
                         const int *ixr, const int *iyr, const int *izr, int ldseis,     \
                         REAL *sisvx, REAL *sisvy, REAL *sisvz, const int *ewin,         \
                         int energy_slabs, REAL *etotal, REAL *state, int nthreads);    \
  int oracle_run3d4_##SFX(int nx, int ny, int nz, REAL dx, REAL dy, REAL dz, REAL dt,    \
                          int it_begin, int it_end, const REAL *lambda, const REAL *mu,  \
                          const REAL *rho, const REAL *prof_x, const REAL *prof_y,       \
                          const REAL *prof_z, int isrc, int jsrc, int ksrc,              \
                          const REAL *fx, const REAL *fy, int nrec, const int *ixr,      \
                          const int *iyr, const int *izr, int ldseis, REAL *sisvx,       \
                          REAL *sisvy, const int *ewin, int energy_slabs, REAL *etotal,  \
                          int zcut, REAL *state, int nthreads);                          \
  int oracle_run2dve_##SFX(int nx, int ny, int order, REAL dx, REAL dy, REAL dt,          \
                           int it_begin,                                                  \
                           int it_end, const REAL *lambda, const REAL *mu, const REAL *rho, \
                           const REAL *prof_x, const REAL *prof_y, int nsls,              \
                           const REAL *tau_eps1, const REAL *tau_sig1,                    \
                           const REAL *tau_eps2, const REAL *tau_sig2, int isrc, int jsrc, \
                           const REAL *fx, const REAL *fy, int nrec, const int *ixr,      \
                           const int *iyr, int ldseis, REAL *sisvx, REAL *sisvy,          \
                           REAL *sisp, REAL *state, int nthreads);                      \
  void oracle_stf_at_##SFX(int n, const REAL *t, REAL f0, REAL t0, REAL factor,           \
                           REAL angle_deg, int kind, REAL *fx, REAL *fy);                 \
  int oracle_runade8_##SFX(int nx, int ny, REAL dx, REAL dy, REAL dt, int it_begin,       \
                           int it_end, int npml, const REAL *lam, const REAL *mu,         \
                           const REAL *rho, const REAL *ax, const REAL *bx, const REAL *Kx, \
                           const REAL *axh, const REAL *bxh, const REAL *Kxh,             \
                           const REAL *ay, const REAL *by, const REAL *Ky,                \
                           const REAL *ayh, const REAL *byh, const REAL *Kyh, int isrc,   \
                           int jsrc, REAL f0, REAL t0, REAL factor, REAL angle_deg,       \
                           int nrec, const int *ixr, const int *iyr, int ldseis,          \
                           REAL *sisvx, REAL *sisvy, REAL *state, int nthreads);

ORACLE_DECLS(double, d)
ORACLE_DECLS(float, f)

/* Holberg (1987) 8th-order staggered coefficients as printed in ADE8:263-266. */
extern const double oracle_holberg8[4];

#ifdef __cplusplus
}
#endif
#endif
