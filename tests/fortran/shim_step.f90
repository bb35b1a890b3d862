! Test driver for the Fortran binding: a main-loop fragment that keeps its particles as the
! reference does, in arrays of TYPE particle (pic2d_Modules.f90:432-439), hands them to the
! device through GPU_PARTICLES_UPLOAD_AOS and runs one step of the hot path through the
! edgpu_drop_in wrappers, in the reference main loop's order (pic2d_MainProgram.f90:151-363).

module shim_particles
  implicit none
  type particle
     real(8) X
     real(8) Y
     real(8) VX
     real(8) VY
     real(8) VZ
     integer tag
  end type particle
end module shim_particles

! p: edgpu_params; e(ne, 5), ions(ni, 5): X, Y, VX, VY, VZ columns, overwritten with the
! particles after the step (ne, ni updated); grids (nx, ny); hits(nobj, 2) electron and ion
! wall hits per object; stride_out: the byte stride of TYPE particle as this compiler lays it out
subroutine shim_one_step(p, ne, e, ni, ions, nx, ny, rho_e, rho_i, phi, EX, EY, nobj, hits, stride_out) &
     bind(C, name="shim_one_step")
  use, intrinsic :: iso_c_binding
  use edgpu
  use edgpu_drop_in
  use shim_particles
  implicit none
  type(edgpu_params), intent(in) :: p
  integer(c_int32_t), intent(inout) :: ne, ni
  integer(c_int32_t), value :: nx, ny, nobj
  real(c_double), intent(inout) :: e(ne, 5), ions(ni, 5)
  real(c_double), intent(out), target :: rho_e(nx, ny), rho_i(nx, ny), phi(nx, ny), EX(nx, ny), EY(nx, ny)
  integer(c_int32_t), intent(out) :: hits(nobj, 2)
  integer(c_int64_t), intent(out) :: stride_out
  type(particle), allocatable, target :: electron(:), ion(:)
  integer, allocatable :: ehits(:), ihits(:)
  integer(c_int64_t) :: stride
  integer :: k, n_e, n_i, cap_e, cap_i

  call edgpu_check(edgpu_init(p), "edgpu_init")
  cap_e = ne
  cap_i = ni
  allocate(electron(cap_e), ion(cap_i))
  do k = 1, ne
     electron(k) = particle(e(k, 1), e(k, 2), e(k, 3), e(k, 4), e(k, 5), k)
  end do
  do k = 1, ni
     ion(k) = particle(ions(k, 1), ions(k, 2), ions(k, 3), ions(k, 4), ions(k, 5), k)
  end do
  stride = transfer(c_loc(electron(2)), 0_c_intptr_t) - transfer(c_loc(electron(1)), 0_c_intptr_t)
  stride_out = stride
  call GPU_PARTICLES_UPLOAD_AOS(0, ne, c_loc(electron(1)), stride)
  call GPU_PARTICLES_UPLOAD_AOS(1, ni, c_loc(ion(1)), stride)

  allocate(ehits(nobj), ihits(nobj))
  ehits = 0
  ihits = 0
  call GPU_GATHER_ION_CHARGE_DENSITY(rho_i)
  call GPU_GATHER_ELECTRON_CHARGE_DENSITY(rho_e)
  call GPU_SOLVE_POISSON_FFTX_LINSYSY(phi)
  call GPU_CALCULATE_ELECTRIC_FIELD(EX, EY)
  call GPU_ADVANCE_ELECTRONS(ehits)
  call GPU_ADVANCE_IONS(ihits)
  call GPU_PROCESS_ADDED_ELECTRONS()
  call GPU_PROCESS_ADDED_IONS(1)
  call GPU_CLEAR_ACCUMULATED_FIELDS()
  hits(:, 1) = ehits
  hits(:, 2) = ihits

  call GPU_PARTICLES_DOWNLOAD_AOS(0, cap_e, n_e, c_loc(electron(1)), stride)
  call GPU_PARTICLES_DOWNLOAD_AOS(1, cap_i, n_i, c_loc(ion(1)), stride)
  ne = n_e
  ni = n_i
  do k = 1, n_e
     e(k, :) = [electron(k)%X, electron(k)%Y, electron(k)%VX, electron(k)%VY, electron(k)%VZ]
  end do
  do k = 1, n_i
     ions(k, :) = [ion(k)%X, ion(k)%Y, ion(k)%VX, ion(k)%VY, ion(k)%VZ]
  end do
  call edgpu_check(edgpu_finalize(), "edgpu_finalize")
end subroutine shim_one_step
