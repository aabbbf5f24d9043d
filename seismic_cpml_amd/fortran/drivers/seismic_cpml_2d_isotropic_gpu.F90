! 2D P-SV elastic C-PML simulation with the time loop on the GPU.
!
! Same configuration, set-up and outputs as the reference programs
!   ORDER=2: seismic_CPML_2D_isotropic_second_order.f90 (parameters 2D2:137-218)
!   ORDER=4: seismic_CPML_2D_isotropic_fourth_order.f90 (parameters 2D4:131-219)
! with the time loop (2D2:550-735) replaced by cpml_gpu calls.  Built for
! WP=8 (double precision, as shipped) and WP=4 (the single-precision form
! the reference recommends at 2D2:131-133).
!
! The grid arrays keep the reference's bounds: (NX,NY) at order 2 and the
! ghost-padded (0:NX+1,0:NY+1) of 2D4:205 at order 4, handed over in place
! after cpml_gpu_set_host_layout(h, 1, 1).
!
! Outputs: Vx_file_NNN.dat, Vy_file_NNN.dat, energy.dat, imageNNNNNN_V[xy].pnm
! (reference formats) and gpu_dump.bin (seismograms, energy and final vx, vy
! at full precision).
!
! Built with RESUME_AT=n, the driver checkpoints after step n (fields and
! C-PML memory variables, with get_field / get_memory), closes the handle,
! opens a new one and restores the state (set_field / set_memory) before it
! goes on: the outputs equal the uninterrupted run's.
#ifndef WP
#define WP 8
#endif
#ifndef ORDER
#define ORDER 2
#endif
program seismic_cpml_2d_isotropic_gpu
  use cpml_gpu
  use, intrinsic :: iso_c_binding, only: c_ptr
  implicit none

  integer, parameter :: wp = WP
  integer, parameter :: NX = 101
  integer, parameter :: NY = 641
  real(wp), parameter :: DELTAX = 10._wp
  real(wp), parameter :: DELTAY = DELTAX
  logical, parameter :: USE_PML_XMIN = .true., USE_PML_XMAX = .true.
  logical, parameter :: USE_PML_YMIN = .true., USE_PML_YMAX = .true.
  integer, parameter :: NPOINTS_PML = 10
  real(wp), parameter :: cp = 3300._wp
  real(wp), parameter :: cs = cp / 1.732_wp
  real(wp), parameter :: density = 2800._wp
#if ORDER == 4
  integer, parameter :: NSTEP = 2000 * 2
  real(wp), parameter :: DELTAT = 2.e-3_wp / 2
  integer, parameter :: IT_DISPLAY = 100 * 2
  logical, parameter :: YORIGIN_N = .true.
#else
  integer, parameter :: NSTEP = 2000
  real(wp), parameter :: DELTAT = 2.e-3_wp
  integer, parameter :: IT_DISPLAY = 100
  logical, parameter :: YORIGIN_N = .false.
#endif
  real(wp), parameter :: f0 = 7._wp
  real(wp), parameter :: t0 = 1.20_wp / f0
  real(wp), parameter :: factor = 1.e7_wp
  integer, parameter :: ISOURCE = NX - 2*NPOINTS_PML - 1
  integer, parameter :: JSOURCE = 2 * NY / 3 + 1
  real(wp), parameter :: xsource = (ISOURCE - 1) * DELTAX
  real(wp), parameter :: ysource = (JSOURCE - 1) * DELTAY
  real(wp), parameter :: ANGLE_FORCE = 135._wp
  integer, parameter :: NREC = 2
  real(wp), parameter :: xdeb = xsource - 100._wp
  real(wp), parameter :: ydeb = 2300._wp
  real(wp), parameter :: xfin = xsource
  real(wp), parameter :: yfin = 300._wp
  double precision, parameter :: STABILITY_THRESHOLD = 1.d+25

  ! ghost cells per side of the grid arrays (2D4:205 declares 0:NX+1,0:NY+1)
  integer, parameter :: G = ORDER / 4
  real(wp), dimension(1-G:NX+G,1-G:NY+G) :: lambda, mu, rho, vx, vy
  real(wp), dimension(NX) :: a_x, b_x, K_x, a_x_half, b_x_half, K_x_half
  real(wp), dimension(NY) :: a_y, b_y, K_y, a_y_half, b_y_half, K_y_half
  real(wp), dimension(NSTEP) :: force_x, force_y
  real(wp), dimension(NSTEP,NREC) :: sisvx, sisvy
  double precision, dimension(NSTEP) :: total_energy_kinetic, total_energy_potential
  integer, dimension(NREC) :: ix_rec, iy_rec
  type(c_ptr) :: h
  integer :: i, j, it, it_end, ierr
  double precision :: velocnorm, courant_number, ms
#ifdef RESUME_AT
  ! the state of the 2D loop: 5 fields and 8 memory variables (2D2:222-230)
  integer, parameter :: FIELDS(5) = [CPML_VX, CPML_VY, CPML_SIGMAXX, CPML_SIGMAYY, &
       CPML_SIGMAXY]
  integer, parameter :: MEMS(8) = [CPML_MEM_DVX_DX, CPML_MEM_DVX_DY, CPML_MEM_DVY_DX, &
       CPML_MEM_DVY_DY, CPML_MEM_DSIGMAXX_DX, CPML_MEM_DSIGMAYY_DY, CPML_MEM_DSIGMAXY_DX, &
       CPML_MEM_DSIGMAXY_DY]
  real(wp), dimension(1-G:NX+G,1-G:NY+G,13) :: saved
  real(wp), dimension(NSTEP,NREC) :: sisvx0, sisvy0
  double precision, dimension(NSTEP) :: ekin0, epot0
  integer :: q
#endif

  print *, '2D elastic finite-difference code in velocity and stress formulation with C-PML'
  print *, 'time loop on the GPU (cpml_gpu), order ', ORDER, ', real kind ', wp
  print *, 'NX = ', NX, ' NY = ', NY

  ! C-PML profiles (2D2:283-465)
  call pml_profile(NX, DELTAX, DELTAT, NPOINTS_PML, cp, f0, USE_PML_XMIN, USE_PML_XMAX, &
                   .false., .true., a_x, b_x, K_x, a_x_half, b_x_half, K_x_half)
  call pml_profile(NY, DELTAY, DELTAT, NPOINTS_PML, cp, f0, USE_PML_YMIN, USE_PML_YMAX, &
                   YORIGIN_N, .false., a_y, b_y, K_y, a_y_half, b_y_half, K_y_half)
  ! Lame parameters and density (2D2:468-474)
  do j = 1-G, NY+G
    do i = 1-G, NX+G
      rho(i,j) = density
      mu(i,j) = density*cs*cs
      lambda(i,j) = density*(cp*cp - 2._wp*cs*cs)
    enddo
  enddo
  call place_receivers(NX, NY, DELTAX, DELTAY, NREC, xdeb, ydeb, xfin, yfin, ix_rec, iy_rec)
  do i = 1, NREC
    print *, 'receiver ', i, ' at grid point i,j = ', ix_rec(i), iy_rec(i)
  enddo
  courant_number = cp * DELTAT * sqrt(1._wp/DELTAX**2 + 1._wp/DELTAY**2)
  print *, 'Courant number is ', courant_number
  call source_forces(NSTEP, DELTAT, f0, t0, factor, ANGLE_FORCE, 1, force_x, force_y)

  ! hand the arrays to the GPU library
  call open_handle()

  ! time loop, in the display chunks of 2D2:716 (it == 5 and every IT_DISPLAY)
  it = 1
  do while (it <= NSTEP)
    it_end = (it + IT_DISPLAY - 1) / IT_DISPLAY * IT_DISPLAY
    if (it <= 5) it_end = 5
    it_end = min(it_end, NSTEP)
    call cpml_gpu_run(h, it, it_end, ierr)
    call check(h, ierr, 'run')
    if (mod(it_end, IT_DISPLAY) == 0 .or. it_end == 5) then
      call cpml_gpu_velocity_norm(h, velocnorm, ierr)
      call cpml_gpu_get_energy(h, total_energy_kinetic, total_energy_potential, ierr)
      print *, 'Time step # ', it_end, ' out of ', NSTEP
      print *, 'Time: ', real((it_end-1)*DELTAT, 4), ' seconds'
      print *, 'Max norm velocity vector V (m/s) = ', velocnorm
      print *, 'total energy = ', total_energy_kinetic(it_end) + total_energy_potential(it_end)
      ! PNM snapshots of Vx and Vy (2D2:728-731), computed on the device
      call cpml_gpu_create_color_image(h, NX, NY, ORDER, it_end, 1, ierr)
      call check(h, ierr, 'color_image')
      call cpml_gpu_create_color_image(h, NX, NY, ORDER, it_end, 2, ierr)
      call check(h, ierr, 'color_image')
    endif
#ifdef RESUME_AT
    if (it_end == RESUME_AT) then
      ! checkpoint, a new handle, restart from the saved state
      call cpml_gpu_get_seismograms(h, sisvx0, sisvy0, ierr)
      call cpml_gpu_get_energy(h, ekin0, epot0, ierr)
      do q = 1, 5
        call cpml_gpu_get_field(h, FIELDS(q), saved(:,:,q), ierr)
      enddo
      do q = 1, 8
        call cpml_gpu_get_memory(h, MEMS(q), saved(:,:,5+q), ierr)
      enddo
      call check(h, ierr, 'checkpoint')
      call cpml_gpu_finalize(h, ierr)
      call open_handle()
      do q = 1, 5
        call cpml_gpu_set_field(h, FIELDS(q), saved(:,:,q), ierr)
      enddo
      do q = 1, 8
        call cpml_gpu_set_memory(h, MEMS(q), saved(:,:,5+q), ierr)
      enddo
      call check(h, ierr, 'restore')
      print *, 'resumed from the checkpoint of step ', RESUME_AT
    endif
#endif
    it = it_end + 1
  enddo

  call cpml_gpu_get_seismograms(h, sisvx, sisvy, ierr)
  call check(h, ierr, 'get_seismograms')
  call cpml_gpu_get_energy(h, total_energy_kinetic, total_energy_potential, ierr)
  call check(h, ierr, 'get_energy')
#ifdef RESUME_AT
  sisvx(1:RESUME_AT,:) = sisvx0(1:RESUME_AT,:)
  sisvy(1:RESUME_AT,:) = sisvy0(1:RESUME_AT,:)
  total_energy_kinetic(1:RESUME_AT) = ekin0(1:RESUME_AT)
  total_energy_potential(1:RESUME_AT) = epot0(1:RESUME_AT)
#endif
  vx = 0._wp
  vy = 0._wp
  call cpml_gpu_get_field(h, CPML_VX, vx, ierr)
  call cpml_gpu_get_field(h, CPML_VY, vy, ierr)
  call check(h, ierr, 'get_field')
  call cpml_gpu_elapsed_ms(h, ms, ierr)
  call cpml_gpu_finalize(h, ierr)

  call write_seismograms(sisvx, sisvy, NSTEP, NREC, DELTAT)
  open(unit=20, file='energy.dat', status='unknown')
  do it = 1, NSTEP
    write(20,*) real(real(it-1, wp)*DELTAT, 4), real(total_energy_kinetic(it), 4), &
         real(total_energy_potential(it), 4), &
         real(total_energy_kinetic(it) + total_energy_potential(it), 4)
  enddo
  close(20)
  open(unit=21, file='gpu_dump.bin', access='stream', form='unformatted', status='replace')
  write(21) sisvx, sisvy, total_energy_kinetic, total_energy_potential, vx(1:NX,1:NY), &
       vy(1:NX,1:NY)
  close(21)
  print *, 'End of the simulation'

contains

  include 'cpml_setup.inc'

  ! handle with the program's medium, profiles, source, receivers, energy and
  ! stability check
  subroutine open_handle()
    call cpml_gpu_init(h, 2, ORDER, wp, NX, NY, 1, dble(DELTAX), dble(DELTAY), 1.d0, &
                       dble(DELTAT), NPOINTS_PML, NSTEP, ierr)
    call check(h, ierr, 'init')
    call cpml_gpu_set_host_layout(h, G, G, ierr)
    call check(h, ierr, 'set_host_layout')
    call cpml_gpu_set_medium(h, lambda, mu, rho, ierr)
    call check(h, ierr, 'set_medium')
    call cpml_gpu_set_pml_2d(h, a_x, b_x, K_x, a_x_half, b_x_half, K_x_half, &
                             a_y, b_y, K_y, a_y_half, b_y_half, K_y_half, ierr)
    call check(h, ierr, 'set_pml')
    call cpml_gpu_set_source(h, ISOURCE, JSOURCE, 1, force_x, force_y, ierr)
    call check(h, ierr, 'set_source')
    call cpml_gpu_set_receivers(h, NREC, ix_rec, iy_rec, ierr=ierr)
    call check(h, ierr, 'set_receivers')
    call cpml_gpu_set_energy(h, 1, ierr)
    call check(h, ierr, 'set_energy')
    call cpml_gpu_set_stability(h, IT_DISPLAY, STABILITY_THRESHOLD, ierr)
    call check(h, ierr, 'set_stability')
  end subroutine

end program seismic_cpml_2d_isotropic_gpu
