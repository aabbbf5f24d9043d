! 3D elastic C-PML simulation, 4th order in space, with the time loop on one GPU.
!
! Configuration and set-up of seismic_CPML_3D_viscoelastic_MPI.f90
! (3D-visco:152-245; K_MAX_PML 7, Rcoef 1e-4) in its elastic form, i.e. the
! update of that program with the attenuation switched off (tau_epsilon =
! tau_sigma, where its d0 factor sqrt(taumax) is 1), on the whole grid of
! one device: the source at (ISOURCE, JSOURCE, NZ/2) and the receivers on
! that plane, as its cut-plane rank has them (3D-visco:1350, 1374-1377).  The
! time loop (3D-visco:940-1500) is replaced by cpml_gpu calls.  The grid,
! step count, PML width and receiver offsets can be overridden at compile
! time (NX_, NY_, NZ_, NSTEP_, NPML_, RX1_, RY1_, RY2_ in metres from the
! source).
!
! Output: gpu_dump.bin (profiles, receiver indices, vx and vy seismograms,
! final vx on the cut plane; float64).
#ifndef NX_
#define NX_ 210
#endif
#ifndef NY_
#define NY_ 800
#endif
#ifndef NZ_
#define NZ_ 220
#endif
#ifndef NSTEP_
#define NSTEP_ 100000
#endif
#ifndef NPML_
#define NPML_ 10
#endif
#ifndef RX1_
#define RX1_ 500
#endif
#ifndef RY1_
#define RY1_ 500
#endif
#ifndef RY2_
#define RY2_ 2260
#endif
program seismic_cpml_3d4_gpu
  use cpml_gpu
  use, intrinsic :: iso_c_binding, only: c_ptr
  implicit none

  integer, parameter :: wp = 8
  integer, parameter :: NX = NX_, NY = NY_, NZ = NZ_
  real(wp), parameter :: DELTAX = 4._wp, DELTAY = DELTAX, DELTAZ = DELTAX
  real(wp), parameter :: cp = 3000._wp, cs = 2000._wp, rho = 2000._wp
  real(wp), parameter :: mu = rho*cs*cs
  real(wp), parameter :: lambda = rho*(cp*cp - 2._wp*cs*cs)
  integer, parameter :: NSTEP = NSTEP_
  real(wp), parameter :: DELTAT = 4.e-4_wp
  real(wp), parameter :: f0 = 18._wp
  real(wp), parameter :: t0 = 1.20_wp / f0
  real(wp), parameter :: factor = 1.e7_wp
  integer, parameter :: NPOINTS_PML = NPML_
  integer, parameter :: ISOURCE = NPOINTS_PML + 20
  integer, parameter :: JSOURCE = NY / 5 + 1
  integer, parameter :: KSOURCE = NZ / 2
  real(wp), parameter :: xsource = ISOURCE * DELTAX
  real(wp), parameter :: ysource = JSOURCE * DELTAY
  real(wp), parameter :: ANGLE_FORCE = 0._wp
  integer, parameter :: NREC = 3

  real(wp), allocatable, dimension(:,:,:) :: alam, amu, arho, vx
  real(wp), dimension(NX) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half
  real(wp), dimension(NY) :: a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
  real(wp), dimension(NZ) :: a_z, b_z, K_z, a_z_half, b_z_half, K_z_half
  real(wp), dimension(NSTEP) :: force_x, force_y
  real(wp), dimension(NSTEP,NREC) :: sisvx, sisvy
  real(wp), dimension(NREC) :: xrec, yrec
  integer, dimension(NREC) :: ix_rec, iy_rec, iz_rec
  type(c_ptr) :: h
  integer :: i, j, irec, ierr
  real(wp) :: dist, distval

  print *, '3D elastic finite-difference code in velocity and stress formulation with C-PML,'
  print *, '4th order in space, time loop on one GPU (cpml_gpu)'
  print *, 'NX = ', NX, ' NY = ', NY, ' NZ = ', NZ

  ! profiles (3D-visco:537-800): K_MAX_PML 7, Rcoef 1e-4, alpha clamped along x
  call pml_profile(NX, DELTAX, DELTAT, NPOINTS_PML, cp, f0, .true., .true., .false., .true., &
                   a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, 7._wp, 0.0001_wp)
  call pml_profile(NY, DELTAY, DELTAT, NPOINTS_PML, cp, f0, .true., .true., .false., .false., &
                   a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, 7._wp, 0.0001_wp)
  call pml_profile(NZ, DELTAZ, DELTAT, NPOINTS_PML, cp, f0, .true., .true., .false., .false., &
                   a_z, b_z, K_z, a_z_half, b_z_half, K_z_half, 7._wp, 0.0001_wp)
  ! receivers (3D-visco:824-845): nearest node, distances from DELTA*i
  xrec = (/ xsource + RX1_, xsource, xsource + RX1_ /)
  yrec = (/ ysource + RY1_, ysource + RY2_, ysource + RY2_ /)
  do irec = 1, NREC
    dist = huge(1._wp)
    do j = 1, NY
      do i = 1, NX
        distval = sqrt((DELTAX*real(i, wp) - xrec(irec))**2 + (DELTAY*real(j, wp) - yrec(irec))**2)
        if (distval < dist) then
          dist = distval
          ix_rec(irec) = i
          iy_rec(irec) = j
        endif
      enddo
    enddo
  enddo
  iz_rec = KSOURCE
  call source_forces(NSTEP, DELTAT, f0, t0, factor, ANGLE_FORCE, 1, force_x, force_y)

  allocate(alam(NX,NY,NZ), amu(NX,NY,NZ), arho(NX,NY,NZ))
  alam = lambda
  amu = mu
  arho = rho
  call cpml_gpu_init(h, 3, 4, wp, NX, NY, NZ, dble(DELTAX), dble(DELTAY), dble(DELTAZ), &
                     dble(DELTAT), NPOINTS_PML, NSTEP, ierr)
  call check(h, ierr, 'init')
  call cpml_gpu_set_medium(h, alam, amu, arho, ierr)
  call check(h, ierr, 'set_medium')
  deallocate(alam, amu, arho)
  call cpml_gpu_set_pml_3d(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                           a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, &
                           a_z, b_z, K_z, a_z_half, b_z_half, K_z_half, ierr)
  call check(h, ierr, 'set_pml')
  call cpml_gpu_set_source(h, ISOURCE, JSOURCE, KSOURCE, force_x, force_y, ierr)
  call check(h, ierr, 'set_source')
  call cpml_gpu_set_receivers(h, NREC, ix_rec, iy_rec, iz_rec, ierr)
  call check(h, ierr, 'set_receivers')
  call cpml_gpu_run(h, 1, NSTEP, ierr)
  call check(h, ierr, 'run')
  call cpml_gpu_get_seismograms(h, sisvx, sisvy, ierr)
  call check(h, ierr, 'get_seismograms')
  allocate(vx(NX,NY,NZ))
  call cpml_gpu_get_field(h, CPML_VX, vx, ierr)
  call check(h, ierr, 'get_field')
  call cpml_gpu_finalize(h, ierr)

  open(unit=21, file='gpu_dump.bin', access='stream', form='unformatted', status='replace')
  write(21) a_x, b_x, K_x, a_x_half, b_x_half, K_x_half
  write(21) a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
  write(21) a_z, b_z, K_z, a_z_half, b_z_half, K_z_half
  write(21) ix_rec, iy_rec
  write(21) sisvx, sisvy, vx(:,:,KSOURCE)
  close(21)
  print *, 'End of the simulation'

contains

  include 'cpml_setup.inc'

end program seismic_cpml_3d4_gpu
