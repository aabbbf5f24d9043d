! 2D P-SV viscoelastic C-PML simulation with the time loop on the GPU.
!
! Configuration and set-up of
! seismic_CPML_2D_velocity_and_stress_second_order_viscoelastic.f90
! (parameters 2D-visco2:140-315: Qp 65, Qs 55, 3 standard linear solids,
! Ricker source divided by the cell area) with the time loop (2D-visco2:699-
! 1060) replaced by cpml_gpu calls.  The relaxation times are those the
! reference's SolvOpt fit gives for this Qp, Qs and f0 (2D-visco2:360-372),
! entered as data.  The grid, step count, source and receivers can be
! overridden at compile time (NX_, NY_, NSTEP_, XSRC_, YSRC_, NREC_, XDEB_,
! YDEB_, XFIN_, YFIN_).
!
! Output: gpu_dump.bin (vx, vy and pressure seismograms, final vx, vy; float64).
#ifndef NX_
#define NX_ 2001
#endif
#ifndef NY_
#define NY_ 2001
#endif
#ifndef NSTEP_
#define NSTEP_ 5200
#endif
#ifndef XSRC_
#define XSRC_ 1500
#endif
#ifndef YSRC_
#define YSRC_ 1500
#endif
#ifndef NREC_
#define NREC_ 1
#endif
#ifndef XDEB_
#define XDEB_ 2301
#endif
#ifndef YDEB_
#define YDEB_ 2301
#endif
#ifndef XFIN_
#define XFIN_ 2301
#endif
#ifndef YFIN_
#define YFIN_ 2301
#endif
program seismic_cpml_2d_visco_gpu
  use cpml_gpu
  use, intrinsic :: iso_c_binding, only: c_ptr
  implicit none

  integer, parameter :: wp = 8
  integer, parameter :: NX = NX_
  integer, parameter :: NY = NY_
  real(wp), parameter :: DELTAX = 1.5_wp
  real(wp), parameter :: DELTAY = DELTAX
  integer, parameter :: NPOINTS_PML = 10
  real(wp), parameter :: cp_unrelaxed = 2000._wp
  real(wp), parameter :: cs_unrelaxed = cp_unrelaxed / 1.732_wp
  real(wp), parameter :: density = 2000._wp
  real(wp), parameter :: DELTAT = 2.2e-4_wp
  integer, parameter :: NSTEP = NSTEP_
  real(wp), parameter :: f0 = 35._wp
  real(wp), parameter :: t0 = 1.20_wp / f0
  real(wp), parameter :: factor = 1._wp
  real(wp), parameter :: xsource = real(XSRC_, wp)
  real(wp), parameter :: ysource = real(YSRC_, wp)
  integer, parameter :: ISOURCE = xsource / DELTAX + 1
  integer, parameter :: JSOURCE = ysource / DELTAY + 1
  real(wp), parameter :: ANGLE_FORCE = 0._wp
  integer, parameter :: NREC = NREC_
  real(wp), parameter :: xdeb = real(XDEB_, wp), ydeb = real(YDEB_, wp)
  real(wp), parameter :: xfin = real(XFIN_, wp), yfin = real(YFIN_, wp)
  integer, parameter :: N_SLS = 3
  ! SolvOpt fit of the reference for Qp = 65 (nu1) and Qs = 55 (nu2) around f0
  double precision, parameter :: tau_epsilon_nu1(N_SLS) = &
       (/ 0.024081581857537462d0, 0.0046996089908613505d0, 0.0009567997872435829d0 /)
  double precision, parameter :: tau_sigma_nu1(N_SLS) = &
       (/ 0.022560146386368645d0, 0.004508471279712249d0, 0.000893787640376875d0 /)
  double precision, parameter :: tau_epsilon_nu2(N_SLS) = &
       (/ 0.024305444805272164d0, 0.0047281078292263955d0, 0.0009667252695863502d0 /)
  double precision, parameter :: tau_sigma_nu2(N_SLS) = &
       (/ 0.022509197794294895d0, 0.0045013880073380965d0, 0.0008917332095369118d0 /)

  real(wp), dimension(NX,NY) :: lambda_unrelaxed, mu_unrelaxed, rho, vx, vy
  real(wp), dimension(NX) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half
  real(wp), dimension(NY) :: a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
  real(wp), dimension(NSTEP) :: force_x, force_y
  real(wp), dimension(NSTEP,NREC) :: sisvx, sisvy, sispressure
  integer, dimension(NREC) :: ix_rec, iy_rec
  type(c_ptr) :: h
  integer :: i, j, ierr

  print *, '2D viscoelastic finite-difference code in velocity and stress formulation'
  print *, 'with C-PML, time loop on the GPU (cpml_gpu)'
  print *, 'NX = ', NX, ' NY = ', NY, ' N_SLS = ', N_SLS

  ! C-PML profiles (2D-visco2:383-470: the 2D2 formulas with cp_unrelaxed)
  call pml_profile(NX, DELTAX, DELTAT, NPOINTS_PML, cp_unrelaxed, f0, .true., .true., &
                   .false., .true., a_x, b_x, K_x, a_x_half, b_x_half, K_x_half)
  call pml_profile(NY, DELTAY, DELTAT, NPOINTS_PML, cp_unrelaxed, f0, .true., .true., &
                   .false., .false., a_y, b_y, K_y, a_y_half, b_y_half, K_y_half)
  ! unrelaxed moduli and density (2D-visco2:590-597)
  do j = 1, NY
    do i = 1, NX
      rho(i,j) = density
      mu_unrelaxed(i,j) = density*cs_unrelaxed*cs_unrelaxed
      lambda_unrelaxed(i,j) = density*cp_unrelaxed*cp_unrelaxed - 2._wp*mu_unrelaxed(i,j)
    enddo
  enddo
  call place_receivers(NX, NY, DELTAX, DELTAY, NREC, xdeb, ydeb, xfin, yfin, ix_rec, iy_rec)
  ! Ricker divided by the cell area (2D-visco2:895-899), along y (angle 0)
  call source_forces(NSTEP, DELTAT, f0, t0, factor, ANGLE_FORCE, 2, force_x, force_y)
  force_x = force_x / (DELTAX * DELTAY)
  force_y = force_y / (DELTAX * DELTAY)

  call cpml_gpu_init(h, 2, 2, wp, NX, NY, 1, dble(DELTAX), dble(DELTAY), 1.d0, &
                     dble(DELTAT), NPOINTS_PML, NSTEP, ierr)
  call check(h, ierr, 'init')
  call cpml_gpu_set_medium(h, lambda_unrelaxed, mu_unrelaxed, rho, ierr)
  call check(h, ierr, 'set_medium')
  call cpml_gpu_set_pml_2d(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                           a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, ierr)
  call check(h, ierr, 'set_pml')
  call cpml_gpu_set_attenuation(h, N_SLS, tau_epsilon_nu1, tau_sigma_nu1, tau_epsilon_nu2, &
                                tau_sigma_nu2, ierr)
  call check(h, ierr, 'set_attenuation')
  call cpml_gpu_set_source(h, ISOURCE, JSOURCE, 1, force_x, force_y, ierr)
  call check(h, ierr, 'set_source')
  call cpml_gpu_set_receivers(h, NREC, ix_rec, iy_rec, ierr=ierr)
  call check(h, ierr, 'set_receivers')

  call cpml_gpu_run(h, 1, NSTEP, ierr)
  call check(h, ierr, 'run')

  call cpml_gpu_get_seismograms(h, sisvx, sisvy, ierr)
  call check(h, ierr, 'get_seismograms')
  call cpml_gpu_get_pressure(h, sispressure, ierr)
  call check(h, ierr, 'get_pressure')
  call cpml_gpu_get_field(h, CPML_VX, vx, ierr)
  call cpml_gpu_get_field(h, CPML_VY, vy, ierr)
  call check(h, ierr, 'get_field')
  call cpml_gpu_finalize(h, ierr)

  open(unit=21, file='gpu_dump.bin', access='stream', form='unformatted', status='replace')
  write(21) sisvx, sisvy, sispressure, vx, vy
  close(21)
  print *, 'End of the simulation'

contains

  include 'cpml_setup.inc'

end program seismic_cpml_2d_visco_gpu
