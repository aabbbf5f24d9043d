! Checks of the cpml_gpu Fortran binding itself (no simulation): the kind and
! shape checks of the generic wrappers, and the explicit-lower-bound host
! layout of the reference's ghost-padded arrays (2D4:205 (0:NX+1,0:NY+1),
! ADE8:272 (-4:NX+4,-4:NY+4)) on get/set of fields and memory variables.
! Prints one line per check and "BINDING OK" when all pass; stops with
! status 1 on the first failure.
program cpml_gpu_binding_check
  use cpml_gpu
  use, intrinsic :: iso_c_binding, only: c_ptr
  implicit none

  integer, parameter :: NX = 40, NY = 30, NPML = 6, NSTEP = 10
  real(8), dimension(NX,NY) :: lam, mu, rho
  real(8), dimension(-4:NX+4,-4:NY+4) :: plam, pmu, prho, pin, pout
  real(8), dimension(0:NX+1,0:NY+1) :: p1
  real(8), dimension(NX*NY) :: flat
  real(4), dimension(NX,NY) :: lam4
  real(8), dimension(NX+1,NY) :: wide
  real(8), dimension(NX) :: ax, bx, kx
  real(8), dimension(NY) :: ay, by, ky
  type(c_ptr) :: h
  integer :: ierr, i, j

  lam = 2.e10_8
  mu = 1.e10_8
  rho = 2500._8
  lam4 = 2.e10
  plam = 2.e10_8
  pmu = 1.e10_8
  prho = 2500._8

  call cpml_gpu_init(h, 2, 2, 8, NX, NY, 1, 10.d0, 10.d0, 1.d0, 1.d-4, NPML, NSTEP, ierr)
  call expect(ierr == CPML_OK, 'init')

  ! real(4) arrays on a double-precision handle
  call cpml_gpu_set_medium(h, lam4, lam4, lam4, ierr)
  call expect(ierr == CPML_ERR_ARG, 'real(4) medium refused by a prec=8 handle')
  print *, '  message: ', trim(cpml_gpu_error_message(h))
  ! a grid array one column too wide
  call cpml_gpu_set_medium(h, wide, wide, wide, ierr)
  call expect(ierr == CPML_ERR_ARG, '(NX+1,NY) medium refused')
  print *, '  message: ', trim(cpml_gpu_error_message(h))
  ! padded arrays without the layout that declares them
  call cpml_gpu_set_medium(h, plam, pmu, prho, ierr)
  call expect(ierr == CPML_ERR_ARG, 'padded medium refused under the default layout')

  call cpml_gpu_set_medium(h, lam, mu, rho, ierr)
  call expect(ierr == CPML_OK, 'interior medium accepted')
  ax = 0; bx = 0; kx = 1; ay = 0; by = 0; ky = 1
  do i = 1, NPML
    bx(i) = 0.5_8; ax(i) = -0.25_8; bx(NX+1-i) = 0.5_8; ax(NX+1-i) = -0.25_8
  enddo
  do j = 1, NPML
    by(j) = 0.5_8; ay(j) = -0.25_8; by(NY+1-j) = 0.5_8; ay(NY+1-j) = -0.25_8
  enddo
  call cpml_gpu_set_pml_2d(h, ax, bx, kx, ax, bx, kx, ay, by, ky, ay, by, ky, ierr)
  call expect(ierr == CPML_OK, 'set_pml_2d')

  ! ADE8 layout: (-4:NX+4,-4:NY+4) is lo = 5, hi = 4
  call cpml_gpu_set_host_layout(h, 5, 4, ierr)
  call expect(ierr == CPML_OK, 'set_host_layout(5, 4)')
  call cpml_gpu_set_medium(h, plam, pmu, prho, ierr)
  call expect(ierr == CPML_OK, '(-4:NX+4,-4:NY+4) medium accepted')
  do j = -4, NY+4
    do i = -4, NX+4
      pin(i,j) = -1._8
      if (i >= 1 .and. i <= NX .and. j >= 1 .and. j <= NY) pin(i,j) = i + 1000._8*j
    enddo
  enddo
  call cpml_gpu_set_field(h, CPML_VX, pin, ierr)
  call expect(ierr == CPML_OK, 'set_field from the padded array')
  pout = 7._8
  call cpml_gpu_get_field(h, CPML_VX, pout, ierr)
  call expect(ierr == CPML_OK, 'get_field into the padded array')
  call expect(all(pout(1:NX,1:NY) == pin(1:NX,1:NY)), 'padded field round trip: interior')
  call expect(all(pout(-4:0,:) == 7._8) .and. all(pout(NX+1:,:) == 7._8) .and. &
              all(pout(:,-4:0) == 7._8) .and. all(pout(:,NY+1:) == 7._8), &
              'padded field round trip: ghost cells untouched')
  ! a memory variable of the x slabs: kept inside the slab, zero elsewhere
  call cpml_gpu_set_memory(h, CPML_MEM_DVX_DX, pin, ierr)
  call expect(ierr == CPML_OK, 'set_memory from the padded array')
  pout = 7._8
  call cpml_gpu_get_memory(h, CPML_MEM_DVX_DX, pout, ierr)
  call expect(ierr == CPML_OK, 'get_memory into the padded array')
  call expect(all(pout(1:NPML,1:NY) == pin(1:NPML,1:NY)) .and. &
              all(pout(NX+1-NPML:NX,1:NY) == pin(NX+1-NPML:NX,1:NY)) .and. &
              all(pout(NX/2,1:NY) == 0._8) .and. &
              all(pout(1:NX,1:NY) == pin(1:NX,1:NY) .or. pout(1:NX,1:NY) == 0._8), &
              'padded memory round trip')
  call expect(all(pout(-4:0,:) == 7._8) .and. all(pout(:,NY+1:) == 7._8), &
              'padded memory: ghost cells untouched')

  ! 2D4 layout: (0:NX+1,0:NY+1) is lo = hi = 1
  call cpml_gpu_set_host_layout(h, 1, 1, ierr)
  p1 = 3._8
  call cpml_gpu_get_field(h, CPML_VX, p1, ierr)
  call expect(ierr == CPML_OK .and. all(p1(1:NX,1:NY) == pin(1:NX,1:NY)) .and. &
              all(p1(0,:) == 3._8) .and. all(p1(NX+1,:) == 3._8), '(0:NX+1,0:NY+1) get_field')
  call cpml_gpu_get_field(h, CPML_VX, pout, ierr)
  call expect(ierr == CPML_ERR_ARG, 'ADE8-padded array refused under the 2D4 layout')

  ! back to interior arrays; rank 1 and a strided section
  call cpml_gpu_set_host_layout(h, 0, 0, ierr)
  call cpml_gpu_get_field(h, CPML_VX, flat, ierr)
  call expect(ierr == CPML_OK .and. all(reshape(flat, [NX,NY]) == pin(1:NX,1:NY)), &
              'rank-1 get_field')
  pout = 7._8
  call cpml_gpu_get_field(h, CPML_VX, pout(1:NX,1:NY), ierr)
  call expect(ierr == CPML_OK .and. all(pout(1:NX,1:NY) == pin(1:NX,1:NY)) .and. &
              all(pout(-4:0,:) == 7._8), 'get_field into an interior section (copy-out)')
  call cpml_gpu_set_host_layout(h, 9, 0, ierr)
  call expect(ierr == CPML_ERR_ARG, 'ghost width above 8 refused')

  call cpml_gpu_finalize(h, ierr)
  print *, 'BINDING OK'

contains

  subroutine expect(cond, what)
    logical, intent(in) :: cond
    character(len=*), intent(in) :: what
    if (.not. cond) then
      print *, 'FAILED: ', what
      stop 1
    endif
    print *, 'ok: ', what
  end subroutine

end program cpml_gpu_binding_check
