! cpml_gpu -- Fortran binding (ISO_C_BINDING) of libcpml_gpu, the MI355X
! C-PML time-loop library (C declarations: include/cpml_gpu.h).
!
! The reference programs (seismic_CPML_2D_isotropic_second_order.f90,
! seismic_CPML_2D_isotropic_fourth_order.f90,
! seismic_CPML_3D_isotropic_MPI_OpenMP.f90) keep their parameter block, their
! set-up (profiles, medium, receivers) and their output writers; the time loop
!     do it = 1,NSTEP ... enddo        (2D2:550-735, 2D4:551-736, 3D:802-1245)
! becomes
!     call cpml_gpu_run(h, it_begin, it_end, ierr)
! after the arrays have been handed over with the cpml_gpu_set_* calls below.
!
! Every routine returns ierr (0 = ok, see the CPML_* constants); the message
! of a failure is cpml_gpu_error_message(h).  Array arguments take real(4) or
! real(8) arrays (generic names over kind-specific procedures) that must match
! the precision given to cpml_gpu_init (checked: CPML_ERR_ARG otherwise), in
! the reference's Fortran layout (x fastest).  Grid arrays (medium, fields,
! memory variables) may have any rank and are checked against the grid: by
! default they hold the interior points only, (NX,NY[,NZ]); after
! cpml_gpu_set_host_layout(h, lo, hi) they are the reference's
! explicit-lower-bound arrays (1-lo:NX+hi, 1-lo:NY+hi[, 1-lo:NZ+hi]), e.g.
! (0:NX+1,0:NY+1) of 2D4:205 or (-4:NX+4,-4:NY+4) of ADE8:272, handed over
! in place (a contiguous actual argument is not copied).
module cpml_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: CPML_OK = 0, CPML_ERR_ARG = 1, CPML_ERR_CFL = 2, &
       CPML_ERR_STENCIL = 3, CPML_ERR_OOM = 4, CPML_ERR_HIP = 5, CPML_ERR_BLOWUP = 6, &
       CPML_ERR_STATE = 7, CPML_ERR_NODEVICE = 8
  integer, parameter, public :: CPML_VX = 0, CPML_VY = 1, CPML_VZ = 2, CPML_SIGMAXX = 3, &
       CPML_SIGMAYY = 4, CPML_SIGMAZZ = 5, CPML_SIGMAXY = 6, CPML_SIGMAXZ = 7, CPML_SIGMAYZ = 8
  ! memory variables, named as in the reference (3D:222-240)
  integer, parameter, public :: CPML_MEM_DVX_DX = 0, CPML_MEM_DVX_DY = 1, CPML_MEM_DVX_DZ = 2, &
       CPML_MEM_DVY_DX = 3, CPML_MEM_DVY_DY = 4, CPML_MEM_DVY_DZ = 5, CPML_MEM_DVZ_DX = 6, &
       CPML_MEM_DVZ_DY = 7, CPML_MEM_DVZ_DZ = 8, CPML_MEM_DSIGMAXX_DX = 9, &
       CPML_MEM_DSIGMAYY_DY = 10, CPML_MEM_DSIGMAZZ_DZ = 11, CPML_MEM_DSIGMAXY_DX = 12, &
       CPML_MEM_DSIGMAXY_DY = 13, CPML_MEM_DSIGMAXZ_DX = 14, CPML_MEM_DSIGMAXZ_DZ = 15, &
       CPML_MEM_DSIGMAYZ_DY = 16, CPML_MEM_DSIGMAYZ_DZ = 17

  public :: cpml_gpu_init, cpml_gpu_finalize, cpml_gpu_set_medium, cpml_gpu_set_pml_2d, &
       cpml_gpu_set_pml_3d, cpml_gpu_set_source, cpml_gpu_set_receivers, cpml_gpu_set_energy, &
       cpml_gpu_set_stability, cpml_gpu_run, cpml_gpu_get_seismograms, cpml_gpu_get_energy, &
       cpml_gpu_get_field, cpml_gpu_set_field, cpml_gpu_get_memory, cpml_gpu_set_memory, &
       cpml_gpu_set_time_block, cpml_gpu_autotune, cpml_gpu_velocity_norm, &
       cpml_gpu_elapsed_ms, cpml_gpu_set_host_layout, &
       cpml_gpu_error_message, cpml_gpu_create_color_image, cpml_gpu_write_su, &
       cpml_gpu_set_attenuation, cpml_gpu_get_pressure, cpml_gpu_set_ade_rk4
  ! many 2D shots in one launch per phase (no reference counterpart: a survey
  ! is a loop of reference runs)
  public :: cpml_gpu_batch_init, cpml_gpu_batch_finalize, cpml_gpu_batch_set_medium, &
       cpml_gpu_batch_set_pml, cpml_gpu_batch_set_source, cpml_gpu_batch_set_receivers, &
       cpml_gpu_batch_autotune, cpml_gpu_batch_run, cpml_gpu_batch_get_seismograms
  ! learned thin absorbing boundary (replaces the C-PML of 3D:399-667)
  integer, parameter, public :: CPML_BOUNDARY_CPML = 0, CPML_BOUNDARY_LEARNED = 1
  public :: cpml_gpu_abc_create, cpml_gpu_abc_destroy, cpml_gpu_set_boundary

  interface
    integer(c_int) function c_init(h, ndim, order, prec, nx, ny, nz, dx, dy, dz, dt, npml, &
                                   nstep) bind(C, name="cpml_gpu_init")
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: ndim, order, prec, nx, ny, nz, npml, nstep
      real(c_double), value :: dx, dy, dz, dt
    end function
    integer(c_int) function c_finalize(h) bind(C, name="cpml_gpu_finalize")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function c_set_medium(h, lam, mu, rho) bind(C, name="cpml_gpu_set_medium")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, lam, mu, rho
    end function
    integer(c_int) function c_set_pml(h, ax, bx, kx, axh, bxh, kxh, ay, by, ky, ayh, byh, kyh, &
                                      az, bz, kz, azh, bzh, kzh) bind(C, name="cpml_gpu_set_pml")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      type(*), dimension(*), intent(in) :: ax, bx, kx, axh, bxh, kxh, ay, by, ky, ayh, byh, &
           kyh, az, bz, kz, azh, bzh, kzh
    end function
    integer(c_int) function c_set_source(h, is, js, ks, fx, fy) bind(C, name="cpml_gpu_set_source")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: is, js, ks
      type(*), dimension(*), intent(in) :: fx, fy
    end function
    integer(c_int) function c_set_receivers(h, nrec, ix, iy, iz) &
         bind(C, name="cpml_gpu_set_receivers")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: nrec
      integer(c_int), dimension(*), intent(in) :: ix, iy, iz
    end function
    integer(c_int) function c_set_energy(h, cadence, window, flags) &
         bind(C, name="cpml_gpu_set_energy")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: cadence, flags
      integer(c_int), dimension(6), intent(in), optional :: window
    end function
    integer(c_int) function c_set_stability(h, it_display, threshold) &
         bind(C, name="cpml_gpu_set_stability")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: it_display
      real(c_double), value :: threshold
    end function
    integer(c_int) function c_run(h, it_begin, it_end) bind(C, name="cpml_gpu_run")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: it_begin, it_end
    end function
    integer(c_int) function c_get_seismograms(h, sx, sy, sz) &
         bind(C, name="cpml_gpu_get_seismograms")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, sx, sy, sz
    end function
    integer(c_int) function c_get_energy(h, kin, pot) bind(C, name="cpml_gpu_get_energy")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), dimension(*) :: kin, pot
    end function
    integer(c_int) function c_get_field(h, which, out) bind(C, name="cpml_gpu_get_field")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, out
      integer(c_int), value :: which
    end function
    integer(c_int) function c_set_field(h, which, in) bind(C, name="cpml_gpu_set_field")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, in
      integer(c_int), value :: which
    end function
    integer(c_int) function c_get_memory(h, which, out) bind(C, name="cpml_gpu_get_memory")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, out
      integer(c_int), value :: which
    end function
    integer(c_int) function c_set_memory(h, which, in) bind(C, name="cpml_gpu_set_memory")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, in
      integer(c_int), value :: which
    end function
    integer(c_int) function c_set_time_block(h, steps, width, rows) &
         bind(C, name="cpml_gpu_set_time_block")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: steps, width, rows
    end function
    integer(c_int) function c_autotune(h, trial_steps, nt, jc, ms) &
         bind(C, name="cpml_gpu_autotune")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: trial_steps
      integer(c_int), intent(out) :: nt, jc
      real(c_double), intent(out) :: ms
    end function
    integer(c_int) function c_set_attenuation(h, nsls, te1, ts1, te2, ts2) &
         bind(C, name="cpml_gpu_set_attenuation")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nsls
      real(c_double), dimension(*), intent(in) :: te1, ts1, te2, ts2
    end function
    integer(c_int) function c_set_ade_rk4(h, ax, bx, axh, bxh, ay, by, ayh, byh, fx4, fy4) &
         bind(C, name="cpml_gpu_set_ade_rk4")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      type(*), dimension(*), intent(in) :: ax, bx, axh, bxh, ay, by, ayh, byh, fx4, fy4
    end function
    integer(c_int) function c_get_pressure(h, sp) bind(C, name="cpml_gpu_get_pressure")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, sp
    end function
    integer(c_int) function c_velocity_norm(h, v) bind(C, name="cpml_gpu_velocity_norm")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function c_color_image(h, fnum, rgb) bind(C, name="cpml_gpu_color_image")
      import :: c_ptr, c_int, c_int8_t
      type(c_ptr), value :: h
      integer(c_int), value :: fnum
      integer(c_int8_t), dimension(*), intent(out) :: rgb
    end function
    integer(c_int) function c_pnm_write(path, nx, ny, rgb) bind(C, name="cpml_pnm_write")
      import :: c_int, c_int8_t, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: nx, ny
      integer(c_int8_t), dimension(*), intent(in) :: rgb
    end function
    integer(c_int) function c_su_write(path, ntr, ns, dt_value, traces, prec) &
         bind(C, name="cpml_su_write")
      import :: c_ptr, c_int, c_char, c_double
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: ntr, ns, prec
      real(c_double), value :: dt_value
      type(c_ptr), value :: traces
    end function
    integer(c_int) function c_elapsed_ms(h, v) bind(C, name="cpml_gpu_elapsed_ms")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function c_batch_init(b, order, prec, nshots, nx, ny, dx, dy, dt, npml, &
                                         nstep) bind(C, name="cpml_batch_init")
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: b
      integer(c_int), value :: order, prec, nshots, nstep
      integer(c_int), dimension(*), intent(in) :: nx, ny, npml
      real(c_double), dimension(*), intent(in) :: dx, dy, dt
    end function
    integer(c_int) function c_batch_finalize(b) bind(C, name="cpml_batch_finalize")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
    end function
    integer(c_int) function c_batch_set_medium(b, shot, lam, mu, rho) &
         bind(C, name="cpml_batch_set_medium")
      import :: c_ptr, c_int
      type(c_ptr), value :: b, lam, mu, rho
      integer(c_int), value :: shot
    end function
    integer(c_int) function c_batch_set_pml(b, shot, ax, bx, kx, axh, bxh, kxh, ay, by, ky, &
                                            ayh, byh, kyh) bind(C, name="cpml_batch_set_pml")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: shot
      type(*), dimension(*), intent(in) :: ax, bx, kx, axh, bxh, kxh, ay, by, ky, ayh, byh, kyh
    end function
    integer(c_int) function c_batch_set_source(b, shot, is, js, fx, fy) &
         bind(C, name="cpml_batch_set_source")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: shot, is, js
      type(*), dimension(*), intent(in) :: fx, fy
    end function
    integer(c_int) function c_batch_set_receivers(b, shot, nrec, ix, iy) &
         bind(C, name="cpml_batch_set_receivers")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: shot, nrec
      integer(c_int), dimension(*), intent(in) :: ix, iy
    end function
    integer(c_int) function c_batch_autotune(b, trial, path, ty, bpc, ms) &
         bind(C, name="cpml_batch_autotune")
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: trial
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), intent(out) :: ty, bpc
      real(c_double), intent(out) :: ms
    end function
    integer(c_int) function c_batch_run(b, it_begin, it_end) bind(C, name="cpml_batch_run")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: it_begin, it_end
    end function
    integer(c_int) function c_batch_get_seismograms(b, shot, sx, sy) &
         bind(C, name="cpml_batch_get_seismograms")
      import :: c_ptr, c_int
      type(c_ptr), value :: b, sx, sy
      integer(c_int), value :: shot
    end function
    integer(c_int) function c_abc_create(net, w, scales) bind(C, name="cpml_abc_net_create")
      import :: c_ptr, c_int, c_float
      type(c_ptr), intent(out) :: net
      real(c_float), dimension(*), intent(in) :: w, scales
    end function
    integer(c_int) function c_abc_destroy(net) bind(C, name="cpml_abc_net_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: net
    end function
    integer(c_int) function c_set_boundary(h, mode, net) bind(C, name="cpml_gpu_set_boundary")
      import :: c_ptr, c_int
      type(c_ptr), value :: h, net
      integer(c_int), value :: mode
    end function
    integer(c_int) function c_last_error(h, buf, n) bind(C, name="cpml_gpu_last_error")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*) :: buf
      integer(c_int), value :: n
    end function
    integer(c_int) function c_check_kind(h, k) bind(C, name="cpml_gpu_check_kind")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: k
    end function
    integer(c_int) function c_batch_check_kind(b, k) bind(C, name="cpml_batch_check_kind")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: k
    end function
    integer(c_int) function c_check_grid_array(h, k, r, shp) &
         bind(C, name="cpml_gpu_check_grid_array")
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: k, r
      integer(c_int64_t), dimension(*), intent(in) :: shp
    end function
    integer(c_int) function c_set_host_layout(h, lo, hi) bind(C, name="cpml_gpu_set_host_layout")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: lo, hi
    end function
    integer(c_int) function c_batch_grid(b, shot, nx, ny) bind(C, name="cpml_batch_grid")
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: shot
      integer(c_int), intent(out) :: nx, ny
    end function
  end interface

  ! real(4) / real(8) specific procedures under the generic names: the
  ! compiler picks the kind of the arrays given, and each checks it against
  ! the precision of the handle (cpml_gpu_check_kind)
  interface cpml_gpu_set_medium
    module procedure cpml_gpu_set_medium_r4, cpml_gpu_set_medium_r8
  end interface
  interface cpml_gpu_set_pml_2d
    module procedure cpml_gpu_set_pml_2d_r4, cpml_gpu_set_pml_2d_r8
  end interface
  interface cpml_gpu_set_pml_3d
    module procedure cpml_gpu_set_pml_3d_r4, cpml_gpu_set_pml_3d_r8
  end interface
  interface cpml_gpu_set_source
    module procedure cpml_gpu_set_source_r4, cpml_gpu_set_source_r8
  end interface
  interface cpml_gpu_get_seismograms
    module procedure cpml_gpu_get_seismograms_r4, cpml_gpu_get_seismograms_r8
  end interface
  interface cpml_gpu_get_field
    module procedure cpml_gpu_get_field_r4, cpml_gpu_get_field_r8
  end interface
  interface cpml_gpu_set_field
    module procedure cpml_gpu_set_field_r4, cpml_gpu_set_field_r8
  end interface
  interface cpml_gpu_get_memory
    module procedure cpml_gpu_get_memory_r4, cpml_gpu_get_memory_r8
  end interface
  interface cpml_gpu_set_memory
    module procedure cpml_gpu_set_memory_r4, cpml_gpu_set_memory_r8
  end interface
  interface cpml_gpu_set_ade_rk4
    module procedure cpml_gpu_set_ade_rk4_r4, cpml_gpu_set_ade_rk4_r8
  end interface
  interface cpml_gpu_get_pressure
    module procedure cpml_gpu_get_pressure_r4, cpml_gpu_get_pressure_r8
  end interface
  interface cpml_gpu_write_su
    module procedure cpml_gpu_write_su_r4, cpml_gpu_write_su_r8
  end interface
  interface cpml_gpu_batch_set_medium
    module procedure cpml_gpu_batch_set_medium_r4, cpml_gpu_batch_set_medium_r8
  end interface
  interface cpml_gpu_batch_set_pml
    module procedure cpml_gpu_batch_set_pml_r4, cpml_gpu_batch_set_pml_r8
  end interface
  interface cpml_gpu_batch_set_source
    module procedure cpml_gpu_batch_set_source_r4, cpml_gpu_batch_set_source_r8
  end interface
  interface cpml_gpu_batch_get_seismograms
    module procedure cpml_gpu_batch_get_seismograms_r4, cpml_gpu_batch_get_seismograms_r8
  end interface

contains

  ! Parameter block of the reference (2D2:137-218, 3D:123-218) -> handle.
  ! prec = 4 or 8 (kind of the arrays passed later); order 2, 4 or 8 (2D).
  subroutine cpml_gpu_init(h, ndim, order, prec, NX, NY, NZ, DELTAX, DELTAY, DELTAZ, DELTAT, &
                           NPOINTS_PML, NSTEP, ierr)
    type(c_ptr), intent(out) :: h
    integer, intent(in) :: ndim, order, prec, NX, NY, NZ, NPOINTS_PML, NSTEP
    double precision, intent(in) :: DELTAX, DELTAY, DELTAZ, DELTAT
    integer, intent(out) :: ierr
    ierr = c_init(h, ndim, order, prec, NX, NY, NZ, DELTAX, DELTAY, DELTAZ, DELTAT, &
                  NPOINTS_PML, NSTEP)
  end subroutine

  subroutine cpml_gpu_finalize(h, ierr)
    type(c_ptr), intent(inout) :: h
    integer, intent(out) :: ierr
    ierr = c_finalize(h)
    h = c_null_ptr
  end subroutine

  ! lambda, mu, rho (2D2:204, 468-474); checks the Courant number (2D2:513-516)
  subroutine cpml_gpu_set_medium_r4(h, lambda, mu, rho, ierr)
    type(c_ptr), intent(in) :: h
    real(4), dimension(..), contiguous, intent(in), target :: lambda, mu, rho
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 4, rank(lambda), int(shape(lambda), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_check_grid_array(h, 4, rank(mu), int(shape(mu), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_check_grid_array(h, 4, rank(rho), int(shape(rho), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_set_medium(h, c_loc(lambda), c_loc(mu), c_loc(rho))
  end subroutine

  subroutine cpml_gpu_set_medium_r8(h, lambda, mu, rho, ierr)
    type(c_ptr), intent(in) :: h
    real(8), dimension(..), contiguous, intent(in), target :: lambda, mu, rho
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 8, rank(lambda), int(shape(lambda), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_check_grid_array(h, 8, rank(mu), int(shape(mu), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_check_grid_array(h, 8, rank(rho), int(shape(rho), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_set_medium(h, c_loc(lambda), c_loc(mu), c_loc(rho))
  end subroutine

  ! damping profiles of 2D2:283-465 (x and y)
  subroutine cpml_gpu_set_pml_2d_r4(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                                 a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, ierr)
    type(c_ptr), intent(in) :: h
    real(4), dimension(*), intent(in) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
         a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 4)
    if (ierr /= CPML_OK) return
    ierr = c_set_pml(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, a_y, b_y, K_y, &
                     a_y_half, b_y_half, K_y_half, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half)
  end subroutine

  subroutine cpml_gpu_set_pml_2d_r8(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                                 a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, ierr)
    type(c_ptr), intent(in) :: h
    real(8), dimension(*), intent(in) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
         a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 8)
    if (ierr /= CPML_OK) return
    ierr = c_set_pml(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, a_y, b_y, K_y, &
                     a_y_half, b_y_half, K_y_half, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half)
  end subroutine

  ! damping profiles of 3D:399-667 (x, y and z)
  subroutine cpml_gpu_set_pml_3d_r4(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                                 a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, &
                                 a_z, b_z, K_z, a_z_half, b_z_half, K_z_half, ierr)
    type(c_ptr), intent(in) :: h
    real(4), dimension(*), intent(in) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
         a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, a_z, b_z, K_z, a_z_half, b_z_half, K_z_half
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 4)
    if (ierr /= CPML_OK) return
    ierr = c_set_pml(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, a_y, b_y, K_y, &
                     a_y_half, b_y_half, K_y_half, a_z, b_z, K_z, a_z_half, b_z_half, K_z_half)
  end subroutine

  subroutine cpml_gpu_set_pml_3d_r8(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                                 a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, &
                                 a_z, b_z, K_z, a_z_half, b_z_half, K_z_half, ierr)
    type(c_ptr), intent(in) :: h
    real(8), dimension(*), intent(in) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
         a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, a_z, b_z, K_z, a_z_half, b_z_half, K_z_half
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 8)
    if (ierr /= CPML_OK) return
    ierr = c_set_pml(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, a_y, b_y, K_y, &
                     a_y_half, b_y_half, K_y_half, a_z, b_z, K_z, a_z_half, b_z_half, K_z_half)
  end subroutine

  ! force_x(it), force_y(it) at (ISOURCE, JSOURCE[, KSOURCE]) (2D2:644-667)
  subroutine cpml_gpu_set_source_r4(h, ISOURCE, JSOURCE, KSOURCE, force_x, force_y, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: ISOURCE, JSOURCE, KSOURCE
    real(4), dimension(*), intent(in) :: force_x, force_y
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 4)
    if (ierr /= CPML_OK) return
    ierr = c_set_source(h, ISOURCE, JSOURCE, KSOURCE, force_x, force_y)
  end subroutine

  subroutine cpml_gpu_set_source_r8(h, ISOURCE, JSOURCE, KSOURCE, force_x, force_y, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: ISOURCE, JSOURCE, KSOURCE
    real(8), dimension(*), intent(in) :: force_x, force_y
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 8)
    if (ierr /= CPML_OK) return
    ierr = c_set_source(h, ISOURCE, JSOURCE, KSOURCE, force_x, force_y)
  end subroutine

  ! ix_rec, iy_rec[, iz_rec] (2D2:494-509)
  subroutine cpml_gpu_set_receivers(h, NREC, ix_rec, iy_rec, iz_rec, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: NREC
    integer, dimension(NREC), intent(in) :: ix_rec, iy_rec
    integer, dimension(NREC), intent(in), optional :: iz_rec
    integer, intent(out) :: ierr
    if (present(iz_rec)) then
      ierr = c_set_receivers(h, NREC, ix_rec, iy_rec, iz_rec)
    else
      ierr = c_set_receivers(h, NREC, ix_rec, iy_rec, ix_rec)
    endif
  end subroutine

  ! energy every `cadence` steps over window(1:6) = i0,i1,j0,j1,k0,k1 (absent:
  ! the reference's window, 2D2:695 / 2D4:696 / 3D:1137-1139)
  subroutine cpml_gpu_set_energy(h, cadence, ierr, window, corrected3d)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: cadence
    integer, intent(out) :: ierr
    integer, dimension(6), intent(in), optional :: window
    logical, intent(in), optional :: corrected3d
    integer :: flags
    flags = 0
    if (present(corrected3d)) then
      if (corrected3d) flags = 1
    endif
    ierr = c_set_energy(h, cadence, window, flags)
  end subroutine

  ! max |v| check at it % IT_DISPLAY == 0 and it == 5 (2D2:716-726)
  subroutine cpml_gpu_set_stability(h, IT_DISPLAY, STABILITY_THRESHOLD, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: IT_DISPLAY
    double precision, intent(in) :: STABILITY_THRESHOLD
    integer, intent(out) :: ierr
    ierr = c_set_stability(h, IT_DISPLAY, STABILITY_THRESHOLD)
  end subroutine

  ! steps it_begin..it_end of the time loop
  subroutine cpml_gpu_run(h, it_begin, it_end, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: it_begin, it_end
    integer, intent(out) :: ierr
    ierr = c_run(h, it_begin, it_end)
  end subroutine

  ! sisvx(NSTEP,NREC), sisvy(NSTEP,NREC)[, sisvz]
  subroutine cpml_gpu_get_seismograms_r4(h, sisvx, sisvy, ierr, sisvz)
    type(c_ptr), intent(in) :: h
    real(4), dimension(..), contiguous, intent(inout), target :: sisvx, sisvy
    integer, intent(out) :: ierr
    real(4), dimension(..), contiguous, intent(inout), target, optional :: sisvz
    type(c_ptr) :: pz
    ierr = c_check_kind(h, 4)
    if (ierr /= CPML_OK) return
    pz = c_null_ptr
    if (present(sisvz)) pz = c_loc(sisvz)
    ierr = c_get_seismograms(h, c_loc(sisvx), c_loc(sisvy), pz)
  end subroutine

  subroutine cpml_gpu_get_seismograms_r8(h, sisvx, sisvy, ierr, sisvz)
    type(c_ptr), intent(in) :: h
    real(8), dimension(..), contiguous, intent(inout), target :: sisvx, sisvy
    integer, intent(out) :: ierr
    real(8), dimension(..), contiguous, intent(inout), target, optional :: sisvz
    type(c_ptr) :: pz
    ierr = c_check_kind(h, 8)
    if (ierr /= CPML_OK) return
    pz = c_null_ptr
    if (present(sisvz)) pz = c_loc(sisvz)
    ierr = c_get_seismograms(h, c_loc(sisvx), c_loc(sisvy), pz)
  end subroutine

  ! total_energy_kinetic(NSTEP), total_energy_potential(NSTEP) (double precision)
  subroutine cpml_gpu_get_energy(h, total_energy_kinetic, total_energy_potential, ierr)
    type(c_ptr), intent(in) :: h
    double precision, dimension(*) :: total_energy_kinetic, total_energy_potential
    integer, intent(out) :: ierr
    ierr = c_get_energy(h, total_energy_kinetic, total_energy_potential)
  end subroutine

  subroutine cpml_gpu_get_field_r4(h, which, field, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(4), dimension(..), contiguous, intent(inout), target :: field
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 4, rank(field), int(shape(field), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_get_field(h, which, c_loc(field))
  end subroutine

  subroutine cpml_gpu_get_field_r8(h, which, field, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(8), dimension(..), contiguous, intent(inout), target :: field
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 8, rank(field), int(shape(field), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_get_field(h, which, c_loc(field))
  end subroutine

  subroutine cpml_gpu_set_field_r4(h, which, field, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(4), dimension(..), contiguous, intent(in), target :: field
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 4, rank(field), int(shape(field), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_set_field(h, which, c_loc(field))
  end subroutine

  subroutine cpml_gpu_set_field_r8(h, which, field, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(8), dimension(..), contiguous, intent(in), target :: field
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 8, rank(field), int(shape(field), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_set_field(h, which, c_loc(field))
  end subroutine

  ! one C-PML memory variable (CPML_MEM_...) as the reference's full-grid array
  ! (zero outside its absorbing slab); with get_field / set_field a run can be
  ! checkpointed and resumed exactly
  subroutine cpml_gpu_get_memory_r4(h, which, memory, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(4), dimension(..), contiguous, intent(inout), target :: memory
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 4, rank(memory), int(shape(memory), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_get_memory(h, which, c_loc(memory))
  end subroutine

  subroutine cpml_gpu_get_memory_r8(h, which, memory, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(8), dimension(..), contiguous, intent(inout), target :: memory
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 8, rank(memory), int(shape(memory), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_get_memory(h, which, c_loc(memory))
  end subroutine

  subroutine cpml_gpu_set_memory_r4(h, which, memory, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(4), dimension(..), contiguous, intent(in), target :: memory
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 4, rank(memory), int(shape(memory), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_set_memory(h, which, c_loc(memory))
  end subroutine

  subroutine cpml_gpu_set_memory_r8(h, which, memory, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: which
    real(8), dimension(..), contiguous, intent(in), target :: memory
    integer, intent(out) :: ierr
    ierr = c_check_grid_array(h, 8, rank(memory), int(shape(memory), c_int64_t))
    if (ierr /= CPML_OK) return
    ierr = c_set_memory(h, which, c_loc(memory))
  end subroutine

  ! `steps` leapfrog steps per launch (time blocking; 0 = off); optional strip
  ! width (columns per wavefront) and chunk depth (rows in 2D, planes in 3D)
  ! grid arrays span (1-lo : N+hi) along each axis (lo = hi = 1 for the
  ! (0:NX+1,0:NY+1) arrays of 2D4:205-225; lo = 5, hi = 4 for ADE8:272-301)
  subroutine cpml_gpu_set_host_layout(h, lo, hi, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: lo, hi
    integer, intent(out) :: ierr
    ierr = c_set_host_layout(h, lo, hi)
  end subroutine

  subroutine cpml_gpu_set_time_block(h, steps, ierr, strip_width, chunk)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: steps
    integer, intent(out) :: ierr
    integer, intent(in), optional :: strip_width, chunk
    integer :: w, c
    w = 0
    c = 0
    if (present(strip_width)) w = strip_width
    if (present(chunk)) c = chunk
    ierr = c_set_time_block(h, steps, w, c)
  end subroutine

  ! time single steps against every time-blocked shape on this grid and keep
  ! the fastest; the state is left as it was (no reference counterpart)
  subroutine cpml_gpu_autotune(h, trial_steps, ierr, steps, chunk, ms)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: trial_steps
    integer, intent(out) :: ierr
    integer, intent(out), optional :: steps, chunk
    double precision, intent(out), optional :: ms
    integer(c_int) :: nt, jc
    real(c_double) :: t
    ierr = c_autotune(h, trial_steps, nt, jc, t)
    if (present(steps)) steps = nt
    if (present(chunk)) chunk = jc
    if (present(ms)) ms = t
  end subroutine

  ! maxval(sqrt(vx**2 + vy**2 [+ vz**2])) (2D2:719, 3D:1185)
  subroutine cpml_gpu_velocity_norm(h, velocnorm, ierr)
    type(c_ptr), intent(in) :: h
    double precision, intent(out) :: velocnorm
    integer, intent(out) :: ierr
    ierr = c_velocity_norm(h, velocnorm)
  end subroutine

  ! viscoelastic step of seismic_CPML_2D_velocity_and_stress_second_order_
  ! viscoelastic.f90 (2D, order 2): N_SLS relaxation times per modulus, as
  ! its SolvOpt fit gives them (nu1: Qp, nu2: Qs); N_SLS = 0: elastic
  subroutine cpml_gpu_set_attenuation(h, N_SLS, tau_epsilon_nu1, tau_sigma_nu1, &
                                      tau_epsilon_nu2, tau_sigma_nu2, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: N_SLS
    double precision, dimension(*), intent(in) :: tau_epsilon_nu1, tau_sigma_nu1, &
         tau_epsilon_nu2, tau_sigma_nu2
    integer, intent(out) :: ierr
    ierr = c_set_attenuation(h, N_SLS, tau_epsilon_nu1, tau_sigma_nu1, tau_epsilon_nu2, &
                             tau_sigma_nu2)
  end subroutine

  ! ADE-PML RK4 scheme of seismic_ADEPML_2D_elastic_RK4_eighth_order.f90 (2D,
  ! order 8): stage tables a_x(4, NX) ... b_y_half(4, NY) (the interior of its
  ! (4, -4:N+4) arrays) and the force at each stage time force_x4(4, NSTEP)
  subroutine cpml_gpu_set_ade_rk4_r4(h, a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, &
                                  b_y_half, force_x4, force_y4, ierr)
    type(c_ptr), intent(in) :: h
    real(4), dimension(*), intent(in) :: a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, &
         b_y_half, force_x4, force_y4
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 4)
    if (ierr /= CPML_OK) return
    ierr = c_set_ade_rk4(h, a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, b_y_half, &
                         force_x4, force_y4)
  end subroutine

  subroutine cpml_gpu_set_ade_rk4_r8(h, a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, &
                                  b_y_half, force_x4, force_y4, ierr)
    type(c_ptr), intent(in) :: h
    real(8), dimension(*), intent(in) :: a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, &
         b_y_half, force_x4, force_y4
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 8)
    if (ierr /= CPML_OK) return
    ierr = c_set_ade_rk4(h, a_x, b_x, a_x_half, b_x_half, a_y, b_y, a_y_half, b_y_half, &
                         force_x4, force_y4)
  end subroutine

  ! sispressure(NSTEP, NREC) of the viscoelastic step (its kind is the handle's)
  subroutine cpml_gpu_get_pressure_r4(h, sispressure, ierr)
    type(c_ptr), intent(in) :: h
    real(4), dimension(..), contiguous, intent(inout), target :: sispressure
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 4)
    if (ierr /= CPML_OK) return
    ierr = c_get_pressure(h, c_loc(sispressure))
  end subroutine

  subroutine cpml_gpu_get_pressure_r8(h, sispressure, ierr)
    type(c_ptr), intent(in) :: h
    real(8), dimension(..), contiguous, intent(inout), target :: sispressure
    integer, intent(out) :: ierr
    ierr = c_check_kind(h, 8)
    if (ierr /= CPML_OK) return
    ierr = c_get_pressure(h, c_loc(sispressure))
  end subroutine

  ! create_color_image of 2D2:861-999 (2D4:729-731 for order 4, ghost-padded;
  ! 3D:1231-1234 for a 3D handle, the plane k = NZ/2): the image of Vx
  ! (field_number 1) or Vy (2) is computed on the device and written as the
  ! reference's file image<it>_V[xy].pnm (2D2:890-896)
  subroutine cpml_gpu_create_color_image(h, NX, NY, order, it, field_number, ierr)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: NX, NY, order, it, field_number
    integer, intent(out) :: ierr
    integer(c_int8_t), allocatable :: rgb(:)
    character(len=100) :: file_name
    integer :: g
    g = 0
    if (order == 4) g = 1
    allocate(rgb(3 * (NX + 2*g) * (NY + 2*g)))
    ierr = c_color_image(h, field_number, rgb)
    if (ierr /= 0) return
    if (field_number == 1) then
      write(file_name, "('image',i6.6,'_Vx.pnm')") it
    else
      write(file_name, "('image',i6.6,'_Vy.pnm')") it
    endif
    ierr = c_pnm_write(trim(file_name) // c_null_char, NX + 2*g, NY + 2*g, rgb)
    deallocate(rgb)
  end subroutine

  ! SU file of ADE8:987-1003 (seismograms: dt_value = DELTAT*1e6) and
  ! ADE8:935-953 (snapshots: dt_value = DELTAY*1e3): traces(ns, ntr) of kind
  ! prec (4 or 8), each written as float32 after a 120-word int16 header
  subroutine cpml_gpu_write_su_r4(file_name, traces, ns, ntr, prec, dt_value, ierr)
    character(len=*), intent(in) :: file_name
    real(4), dimension(..), contiguous, intent(in), target :: traces
    integer, intent(in) :: ns, ntr, prec
    double precision, intent(in) :: dt_value
    integer, intent(out) :: ierr
    ierr = CPML_ERR_ARG
    if (prec /= 4 .or. size(traces) < int(ns, 8) * ntr) return
    ierr = c_su_write(trim(file_name) // c_null_char, ntr, ns, dt_value, c_loc(traces), prec)
  end subroutine

  subroutine cpml_gpu_write_su_r8(file_name, traces, ns, ntr, prec, dt_value, ierr)
    character(len=*), intent(in) :: file_name
    real(8), dimension(..), contiguous, intent(in), target :: traces
    integer, intent(in) :: ns, ntr, prec
    double precision, intent(in) :: dt_value
    integer, intent(out) :: ierr
    ierr = CPML_ERR_ARG
    if (prec /= 8 .or. size(traces) < int(ns, 8) * ntr) return
    ierr = c_su_write(trim(file_name) // c_null_char, ntr, ns, dt_value, c_loc(traces), prec)
  end subroutine

  subroutine cpml_gpu_elapsed_ms(h, ms, ierr)
    type(c_ptr), intent(in) :: h
    double precision, intent(out) :: ms
    integer, intent(out) :: ierr
    ierr = c_elapsed_ms(h, ms)
  end subroutine

  ! ---- batched 2D shots: per-shot arrays NX(nshots), ... (shot = 1..nshots)
  subroutine cpml_gpu_batch_init(b, order, prec, nshots, NX, NY, DELTAX, DELTAY, DELTAT, &
                                 NPOINTS_PML, NSTEP, ierr)
    type(c_ptr), intent(out) :: b
    integer, intent(in) :: order, prec, nshots, NSTEP
    integer, dimension(nshots), intent(in) :: NX, NY, NPOINTS_PML
    double precision, dimension(nshots), intent(in) :: DELTAX, DELTAY, DELTAT
    integer, intent(out) :: ierr
    ierr = c_batch_init(b, order, prec, nshots, NX, NY, DELTAX, DELTAY, DELTAT, NPOINTS_PML, &
                        NSTEP)
  end subroutine

  subroutine cpml_gpu_batch_finalize(b, ierr)
    type(c_ptr), intent(inout) :: b
    integer, intent(out) :: ierr
    ierr = c_batch_finalize(b)
    b = c_null_ptr
  end subroutine

  subroutine cpml_gpu_batch_set_medium_r4(b, shot, lambda, mu, rho, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot
    real(4), dimension(..), contiguous, intent(in), target :: lambda, mu, rho
    integer, intent(out) :: ierr
    integer(c_int) :: nx, ny
    ierr = c_batch_check_kind(b, 4)
    if (ierr /= CPML_OK) return
    ierr = c_batch_grid(b, shot - 1, nx, ny)
    if (ierr /= CPML_OK) return
    if (size(lambda) /= int(nx, 8) * ny) ierr = CPML_ERR_ARG
    if (size(mu) /= int(nx, 8) * ny) ierr = CPML_ERR_ARG
    if (size(rho) /= int(nx, 8) * ny) ierr = CPML_ERR_ARG
    if (ierr /= CPML_OK) return
    ierr = c_batch_set_medium(b, shot - 1, c_loc(lambda), c_loc(mu), c_loc(rho))
  end subroutine

  subroutine cpml_gpu_batch_set_medium_r8(b, shot, lambda, mu, rho, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot
    real(8), dimension(..), contiguous, intent(in), target :: lambda, mu, rho
    integer, intent(out) :: ierr
    integer(c_int) :: nx, ny
    ierr = c_batch_check_kind(b, 8)
    if (ierr /= CPML_OK) return
    ierr = c_batch_grid(b, shot - 1, nx, ny)
    if (ierr /= CPML_OK) return
    if (size(lambda) /= int(nx, 8) * ny) ierr = CPML_ERR_ARG
    if (size(mu) /= int(nx, 8) * ny) ierr = CPML_ERR_ARG
    if (size(rho) /= int(nx, 8) * ny) ierr = CPML_ERR_ARG
    if (ierr /= CPML_OK) return
    ierr = c_batch_set_medium(b, shot - 1, c_loc(lambda), c_loc(mu), c_loc(rho))
  end subroutine

  subroutine cpml_gpu_batch_set_pml_r4(b, shot, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                                    a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot
    real(4), dimension(*), intent(in) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
         a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
    integer, intent(out) :: ierr
    ierr = c_batch_check_kind(b, 4)
    if (ierr /= CPML_OK) return
    ierr = c_batch_set_pml(b, shot - 1, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, a_y, b_y, &
                           K_y, a_y_half, b_y_half, K_y_half)
  end subroutine

  subroutine cpml_gpu_batch_set_pml_r8(b, shot, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                                    a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot
    real(8), dimension(*), intent(in) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
         a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
    integer, intent(out) :: ierr
    ierr = c_batch_check_kind(b, 8)
    if (ierr /= CPML_OK) return
    ierr = c_batch_set_pml(b, shot - 1, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, a_y, b_y, &
                           K_y, a_y_half, b_y_half, K_y_half)
  end subroutine

  subroutine cpml_gpu_batch_set_source_r4(b, shot, ISOURCE, JSOURCE, force_x, force_y, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot, ISOURCE, JSOURCE
    real(4), dimension(*), intent(in) :: force_x, force_y
    integer, intent(out) :: ierr
    ierr = c_batch_check_kind(b, 4)
    if (ierr /= CPML_OK) return
    ierr = c_batch_set_source(b, shot - 1, ISOURCE, JSOURCE, force_x, force_y)
  end subroutine

  subroutine cpml_gpu_batch_set_source_r8(b, shot, ISOURCE, JSOURCE, force_x, force_y, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot, ISOURCE, JSOURCE
    real(8), dimension(*), intent(in) :: force_x, force_y
    integer, intent(out) :: ierr
    ierr = c_batch_check_kind(b, 8)
    if (ierr /= CPML_OK) return
    ierr = c_batch_set_source(b, shot - 1, ISOURCE, JSOURCE, force_x, force_y)
  end subroutine

  subroutine cpml_gpu_batch_set_receivers(b, shot, NREC, ix_rec, iy_rec, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot, NREC
    integer, dimension(NREC), intent(in) :: ix_rec, iy_rec
    integer, intent(out) :: ierr
    ierr = c_batch_set_receivers(b, shot - 1, NREC, ix_rec, iy_rec)
  end subroutine

  ! picks rows per tile and the persistent grid size; cache_path may be ''
  subroutine cpml_gpu_batch_autotune(b, trial_steps, cache_path, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: trial_steps
    character(len=*), intent(in) :: cache_path
    integer, intent(out) :: ierr
    integer(c_int) :: ty, bpc
    real(c_double) :: ms
    ierr = c_batch_autotune(b, trial_steps, trim(cache_path) // c_null_char, ty, bpc, ms)
  end subroutine

  subroutine cpml_gpu_batch_run(b, it_begin, it_end, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: it_begin, it_end
    integer, intent(out) :: ierr
    ierr = c_batch_run(b, it_begin, it_end)
  end subroutine

  subroutine cpml_gpu_batch_get_seismograms_r4(b, shot, sisvx, sisvy, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot
    real(4), dimension(..), contiguous, intent(inout), target :: sisvx, sisvy
    integer, intent(out) :: ierr
    ierr = c_batch_check_kind(b, 4)
    if (ierr /= CPML_OK) return
    ierr = c_batch_get_seismograms(b, shot - 1, c_loc(sisvx), c_loc(sisvy))
  end subroutine

  subroutine cpml_gpu_batch_get_seismograms_r8(b, shot, sisvx, sisvy, ierr)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: shot
    real(8), dimension(..), contiguous, intent(inout), target :: sisvx, sisvy
    integer, intent(out) :: ierr
    ierr = c_batch_check_kind(b, 8)
    if (ierr /= CPML_OK) return
    ierr = c_batch_get_seismograms(b, shot - 1, c_loc(sisvx), c_loc(sisvy))
  end subroutine

  ! ---- learned thin boundary: weights(6*3*27*16*16), scales(9), real(4)
  subroutine cpml_gpu_abc_create(net, weights, scales, ierr)
    type(c_ptr), intent(out) :: net
    real(c_float), dimension(*), intent(in) :: weights, scales
    integer, intent(out) :: ierr
    ierr = c_abc_create(net, weights, scales)
  end subroutine

  subroutine cpml_gpu_abc_destroy(net, ierr)
    type(c_ptr), intent(inout) :: net
    integer, intent(out) :: ierr
    ierr = c_abc_destroy(net)
    net = c_null_ptr
  end subroutine

  ! mode CPML_BOUNDARY_LEARNED with a corrector, or CPML_BOUNDARY_CPML
  subroutine cpml_gpu_set_boundary(h, mode, net, ierr)
    type(c_ptr), intent(in) :: h, net
    integer, intent(in) :: mode
    integer, intent(out) :: ierr
    ierr = c_set_boundary(h, mode, net)
  end subroutine

  function cpml_gpu_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=512) :: msg
    character(kind=c_char), dimension(512) :: buf
    integer :: i, rc
    buf = c_null_char
    rc = c_last_error(h, buf, 512)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    enddo
  end function

end module cpml_gpu
