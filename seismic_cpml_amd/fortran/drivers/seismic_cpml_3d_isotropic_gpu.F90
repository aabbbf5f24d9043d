! 3D elastic C-PML simulation with the time loop on one GPU.
!
! Configuration, set-up and outputs of seismic_CPML_3D_isotropic_MPI_OpenMP.f90
! (parameters 3D:123-218) in its single-rank form: the whole grid on one
! device, the source on the cut plane k = NZ/2 and the receivers on that
! plane (3D:1074-1081, 1123-1129).  The time loop (3D:802-1245) is replaced
! by cpml_gpu calls.  Built for WP=8 and WP=4; the grid and step count can be
! overridden at compile time (NX_, NY_, NZ_, NSTEP_, NPML_, YDEB_, YFIN_).
!
! Outputs: at each display step the timestamp file timestampNNNNNN
! (3D:1219-1229) and the PNM snapshots of Vx and Vy on the cut plane
! (3D:1231-1234); at the end Vx_file_NNN.dat, Vy_file_NNN.dat, energy.dat (one
! column of total energy as 3D:1253-1257) and gpu_dump.bin (full precision).
#ifndef WP
#define WP 8
#endif
#ifndef NX_
#define NX_ 101
#endif
#ifndef NY_
#define NY_ 641
#endif
#ifndef NZ_
#define NZ_ 640
#endif
#ifndef NSTEP_
#define NSTEP_ 2500
#endif
#ifndef NPML_
#define NPML_ 10
#endif
#ifndef YDEB_
#define YDEB_ 2300
#endif
#ifndef YFIN_
#define YFIN_ 300
#endif
program seismic_cpml_3d_isotropic_gpu
  use cpml_gpu
  use, intrinsic :: iso_c_binding, only: c_ptr
  implicit none

  integer, parameter :: wp = WP
  integer, parameter :: NX = NX_
  integer, parameter :: NY = NY_
  integer, parameter :: NZ = NZ_
  real(wp), parameter :: DELTAX = 10._wp, DELTAY = DELTAX, DELTAZ = DELTAX
  real(wp), parameter :: cp = 3300._wp
  real(wp), parameter :: cs = cp / 1.732_wp
  real(wp), parameter :: rho = 2800._wp
  real(wp), parameter :: mu = rho*cs*cs
  real(wp), parameter :: lambda = rho*(cp*cp - 2._wp*cs*cs)
  integer, parameter :: NSTEP = NSTEP_
  real(wp), parameter :: DELTAT = 1.6e-3_wp
  real(wp), parameter :: f0 = 7._wp
  real(wp), parameter :: t0 = 1.20_wp / f0
  real(wp), parameter :: factor = 1.e7_wp
  logical, parameter :: USE_PML = .true.
  integer, parameter :: NPOINTS_PML = NPML_
  integer, parameter :: ISOURCE = NX - 2*NPOINTS_PML - 1
  integer, parameter :: JSOURCE = 2 * NY / 3 + 1
  integer, parameter :: KSOURCE = NZ / 2
  real(wp), parameter :: xsource = (ISOURCE - 1) * DELTAX
  real(wp), parameter :: ysource = (JSOURCE - 1) * DELTAY
  real(wp), parameter :: ANGLE_FORCE = 135._wp
  integer, parameter :: NREC = 2
  real(wp), parameter :: xdeb = xsource - 100._wp
  real(wp), parameter :: ydeb = real(YDEB_, wp)
  real(wp), parameter :: xfin = xsource
  real(wp), parameter :: yfin = real(YFIN_, wp)
  integer, parameter :: IT_DISPLAY = 100
  double precision, parameter :: STABILITY_THRESHOLD = 1.d+25

  real(wp), allocatable, dimension(:,:,:) :: alam, amu, arho, vx
  real(wp), dimension(NX) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half
  real(wp), dimension(NY) :: a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
  real(wp), dimension(NZ) :: a_z, b_z, K_z, a_z_half, b_z_half, K_z_half
  real(wp), dimension(NSTEP) :: force_x, force_y
  real(wp), dimension(NSTEP,NREC) :: sisvx, sisvy
  double precision, dimension(NSTEP) :: ekin, epot
  integer, dimension(NREC) :: ix_rec, iy_rec, iz_rec
  type(c_ptr) :: h
  integer :: it, it_end, ierr
  double precision :: velocnorm, courant_number, ms
  integer(8) :: clock_start, clock_now, clock_rate
  double precision :: tCPU
  integer :: int_tCPU, ihours, iminutes, iseconds
  character(len=100) :: outputname
  integer, parameter :: IOUT = 41

  print *, '3D elastic finite-difference code in velocity and stress formulation with C-PML'
  print *, 'time loop on one GPU (cpml_gpu), real kind ', wp
  print *, 'NX = ', NX, ' NY = ', NY, ' NZ = ', NZ

  ! profiles (3D:399-667): alpha clamped along x only
  call pml_profile(NX, DELTAX, DELTAT, NPOINTS_PML, cp, f0, USE_PML, USE_PML, .false., .true., &
                   a_x, b_x, K_x, a_x_half, b_x_half, K_x_half)
  call pml_profile(NY, DELTAY, DELTAT, NPOINTS_PML, cp, f0, USE_PML, USE_PML, .false., .false., &
                   a_y, b_y, K_y, a_y_half, b_y_half, K_y_half)
  call pml_profile(NZ, DELTAZ, DELTAT, NPOINTS_PML, cp, f0, USE_PML, USE_PML, .false., .false., &
                   a_z, b_z, K_z, a_z_half, b_z_half, K_z_half)
  ! receivers on the cut plane (3D:669-708)
  call place_receivers(NX, NY, DELTAX, DELTAY, NREC, xdeb, ydeb, xfin, yfin, ix_rec, iy_rec)
  iz_rec = KSOURCE
  courant_number = cp * DELTAT * sqrt(1._wp/DELTAX**2 + 1._wp/DELTAY**2 + 1._wp/DELTAZ**2)
  print *, 'Courant number is ', courant_number
  call source_forces(NSTEP, DELTAT, f0, t0, factor, ANGLE_FORCE, 1, force_x, force_y)

  ! homogeneous medium of the reference, as per-point arrays
  allocate(alam(NX,NY,NZ), amu(NX,NY,NZ), arho(NX,NY,NZ))
  alam = lambda
  amu = mu
  arho = rho

  call cpml_gpu_init(h, 3, 2, wp, NX, NY, NZ, dble(DELTAX), dble(DELTAY), dble(DELTAZ), &
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
  call cpml_gpu_set_energy(h, 1, ierr)
  call check(h, ierr, 'set_energy')
  call cpml_gpu_set_stability(h, IT_DISPLAY, STABILITY_THRESHOLD, ierr)
  call check(h, ierr, 'set_stability')

  call system_clock(clock_start, clock_rate)
  it = 1
  do while (it <= NSTEP)
    it_end = (it + IT_DISPLAY - 1) / IT_DISPLAY * IT_DISPLAY
    if (it <= 5) it_end = 5
    it_end = min(it_end, NSTEP)
    call cpml_gpu_run(h, it, it_end, ierr)
    call check(h, ierr, 'run')
    if (mod(it_end, IT_DISPLAY) == 0 .or. it_end == 5) then
      call cpml_gpu_velocity_norm(h, velocnorm, ierr)
      call cpml_gpu_get_energy(h, ekin, epot, ierr)
      print *, 'Time step # ', it_end, ' out of ', NSTEP
      print *, 'Time: ', real((it_end-1)*DELTAT, 4), ' seconds'
      print *, 'Max norm velocity vector V (m/s) = ', velocnorm
      print *, 'Total energy = ', ekin(it_end) + epot(it_end)
      ! elapsed wall-clock time and the progress file (3D:1196-1229)
      call system_clock(clock_now)
      tCPU = dble(clock_now - clock_start) / dble(clock_rate)
      int_tCPU = int(tCPU)
      ihours = int_tCPU / 3600
      iminutes = (int_tCPU - 3600*ihours) / 60
      iseconds = int_tCPU - 3600*ihours - 60*iminutes
      write(*,*) 'Elapsed time in seconds = ', real(tCPU, wp)
      write(*,"(' Elapsed time in hh:mm:ss = ',i4,' h ',i2.2,' m ',i2.2,' s')") &
           ihours, iminutes, iseconds
      write(*,*) 'Mean elapsed time per time step in seconds = ', real(tCPU / dble(it_end), wp)
      write(*,*)
      write(outputname, "('timestamp',i6.6)") it_end
      open(unit=IOUT, file=outputname, status='unknown')
      write(IOUT,*) 'Time step # ', it_end
      write(IOUT,*) 'Time: ', real((it_end-1)*DELTAT, 4), ' seconds'
      write(IOUT,*) 'Max norm velocity vector V (m/s) = ', real(velocnorm, wp)
      write(IOUT,*) 'Total energy = ', real(ekin(it_end) + epot(it_end), wp)
      write(IOUT,*) 'Elapsed time in seconds = ', real(tCPU, wp)
      write(IOUT,"(' Elapsed time in hh:mm:ss = ',i4,' h ',i2.2,' m ',i2.2,' s')") &
           ihours, iminutes, iseconds
      write(IOUT,*) 'Mean elapsed time per time step in seconds = ', &
           real(tCPU / dble(it_end), wp)
      close(IOUT)
      ! snapshots of the cut plane (3D:1231-1234)
      call cpml_gpu_create_color_image(h, NX, NY, 2, it_end, 1, ierr)
      call check(h, ierr, 'create_color_image')
      call cpml_gpu_create_color_image(h, NX, NY, 2, it_end, 2, ierr)
      call check(h, ierr, 'create_color_image')
    endif
    it = it_end + 1
  enddo

  call cpml_gpu_get_seismograms(h, sisvx, sisvy, ierr)
  call check(h, ierr, 'get_seismograms')
  call cpml_gpu_get_energy(h, ekin, epot, ierr)
  call check(h, ierr, 'get_energy')
  allocate(vx(NX,NY,NZ))
  call cpml_gpu_get_field(h, CPML_VX, vx, ierr)
  call check(h, ierr, 'get_field')
  call cpml_gpu_elapsed_ms(h, ms, ierr)
  call cpml_gpu_finalize(h, ierr)

  call write_seismograms(sisvx, sisvy, NSTEP, NREC, DELTAT)
  open(unit=20, file='energy.dat', status='unknown')
  do it = 1, NSTEP
    write(20,*) real(real(it-1, wp)*DELTAT, 4), ekin(it) + epot(it)
  enddo
  close(20)
  open(unit=21, file='gpu_dump.bin', access='stream', form='unformatted', status='replace')
  write(21) sisvx, sisvy, ekin, epot, vx(:,:,KSOURCE)
  close(21)
  print *, 'End of the simulation'

contains

  include 'cpml_setup.inc'

end program seismic_cpml_3d_isotropic_gpu
