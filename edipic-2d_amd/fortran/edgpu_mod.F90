! Fortran side of the edgpu C ABI (include/edgpu.h): the iso_c_binding interfaces and
! drop-in subroutines that a Fortran main loop calls in place of the reference's own
! ADVANCE_ELECTRONS, GATHER_*_CHARGE_DENSITY, SOLVE_POISSON_FFTX_LINSYSY, ... (the name of
! each replaced subroutine is given next to its wrapper).
!
! Error convention: every entry point returns 0 on success; on any other status the
! wrappers print edgpu_last_error() and abort the run, as the reference's own error
! branches call MPI_ABORT.  Compile with -DEDGPU_WITH_MPI inside an MPI build to abort
! through MPI_ABORT; without it the wrappers use ERROR STOP.
!
! Index conventions of the C ABI: 0-based object and species numbers, grids node (i,j) at
! i + j*(gmi+1), the memory order of c_rho(c_indx_x_min:c_indx_x_max, c_indx_y_min:...)
! for the single cluster that covers the domain.

module edgpu
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: EDGPU_OK = 0
  integer(c_int), parameter :: EDGPU_FP64 = 0, EDGPU_MIXED = 1
  integer, parameter :: EDGPU_MAX_SEG = 16, EDGPU_MAX_INNER = 8, EDGPU_MAX_OBJ = 24
  integer, parameter :: EDGPU_MAX_ION_SPECIES = 4, EDGPU_MAX_PROC = 8
  integer(c_int32_t), parameter :: EDGPU_OBJ_METAL = 0, EDGPU_OBJ_DIELECTRIC = 1, EDGPU_OBJ_VACUUM_GAP = 2
  integer(c_int), parameter :: EDGPU_SIDE_LEFT = 0, EDGPU_SIDE_ABOVE = 1, EDGPU_SIDE_RIGHT = 2, EDGPU_SIDE_BELOW = 3

  ! mirror of edgpu_params (field order and types as in include/edgpu.h)
  type, bind(C) :: edgpu_params
     integer(c_int32_t) :: gmi, gmj
     integer(c_int32_t) :: periodic_x
     integer(c_int32_t) :: N_of_particles_cell
     integer(c_int32_t) :: N_subcycles
     integer(c_int32_t) :: precision
     real(c_double) :: Bx_ext, By_ext, Bz_ext
     real(c_double) :: Ez_ext
     real(c_double) :: factor1_E_from_F
     real(c_double) :: factor2_E_from_F
     integer(c_int32_t) :: n_boundary_objects
     integer(c_int32_t) :: n_segments
     integer(c_int32_t) :: seg_side(EDGPU_MAX_SEG)
     integer(c_int32_t) :: seg_start(EDGPU_MAX_SEG)
     integer(c_int32_t) :: seg_end(EDGPU_MAX_SEG)
     integer(c_int32_t) :: seg_object(EDGPU_MAX_SEG)
     integer(c_int32_t) :: n_inner_objects
     integer(c_int32_t) :: inner_box(4, EDGPU_MAX_INNER)   ! C: inner_box[8][4]
     real(c_double) :: object_phi(EDGPU_MAX_OBJ)
     integer(c_int32_t) :: n_ion_species
     integer(c_int32_t) :: Qs(EDGPU_MAX_ION_SPECIES)
     real(c_double) :: QM2s(EDGPU_MAX_ION_SPECIES)
     real(c_double) :: Ms(EDGPU_MAX_ION_SPECIES)
     integer(c_int32_t) :: ions_sense_magnetic_field
     integer(c_int32_t) :: ions_sense_Ez
     integer(c_int64_t) :: max_electrons
     integer(c_int64_t) :: max_ions(EDGPU_MAX_ION_SPECIES)
     integer(c_int64_t) :: add_capacity
     integer(c_int64_t) :: seed
     integer(c_int32_t) :: fuse_collisions
     integer(c_int32_t) :: sort_interval
     integer(c_int32_t) :: ion_sort_interval
     integer(c_int32_t) :: object_type(EDGPU_MAX_OBJ)    ! EDGPU_OBJ_METAL / EDGPU_OBJ_DIELECTRIC
     real(c_double) :: object_eps(EDGPU_MAX_OBJ)          ! whole_object%eps_diel
  end type edgpu_params

  type, bind(C) :: edgpu_advance_stats
     integer(c_int64_t) :: n_before, n_after
     integer(c_int64_t) :: hit_count(EDGPU_MAX_OBJ)
     integer(c_int64_t) :: n_removed
     integer(c_int64_t) :: collisions(EDGPU_MAX_PROC)
     integer(c_int64_t) :: emit_count(EDGPU_MAX_OBJ)
  end type edgpu_advance_stats

  type, bind(C) :: edgpu_see_params
     integer(c_int32_t) :: model_elastic, elastic_reflection, model_inelastic, model_true
     real(c_double) :: setD_elastic, minE_elastic_eV, maxE_elastic_eV
     real(c_double) :: E_elast_0_eV, E_elast_max_eV, maxD_elast, dE_elast_eV, frac_elast_highenergy
     real(c_double) :: setD_inelastic, minE_inelastic_eV, maxE_inelastic_eV
     real(c_double) :: frac_inelastic
     real(c_double) :: T_true_eV
     real(c_double) :: setD_true, minE_true_eV, maxE_true_eV
     real(c_double) :: E_see_0_eV, E_see_max_eV, maxD_see_classic, k_smooth
  end type edgpu_see_params

  interface
     integer(c_int) function edgpu_init(p) bind(C, name="edgpu_init")
       import :: c_int, edgpu_params
       type(edgpu_params), intent(in) :: p
     end function
     integer(c_int) function edgpu_finalize() bind(C, name="edgpu_finalize")
       import :: c_int
     end function
     type(c_ptr) function edgpu_last_error() bind(C, name="edgpu_last_error")
       import :: c_ptr
     end function
     integer(c_int) function edgpu_device() bind(C, name="edgpu_device")
       import :: c_int
     end function
     integer(c_int) function edgpu_particles_upload(species, n, X, Y, VX, VY, VZ, tag) &
          bind(C, name="edgpu_particles_upload")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: species
       integer(c_int64_t), value :: n
       type(c_ptr), value :: X, Y, VX, VY, VZ, tag
     end function
     integer(c_int) function edgpu_particles_download(species, capacity, n, X, Y, VX, VY, VZ, tag) &
          bind(C, name="edgpu_particles_download")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: species
       integer(c_int64_t), value :: capacity
       integer(c_int64_t), intent(out) :: n
       type(c_ptr), value :: X, Y, VX, VY, VZ, tag
     end function
     ! the reference's array of TYPE particle itself: base = c_loc(electron(1)), stride in bytes
     integer(c_int) function edgpu_particles_upload_aos(species, n, base, stride) &
          bind(C, name="edgpu_particles_upload_aos")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: species
       integer(c_int64_t), value :: n, stride
       type(c_ptr), value :: base
     end function
     integer(c_int) function edgpu_particles_download_aos(species, capacity, n, base, stride) &
          bind(C, name="edgpu_particles_download_aos")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: species
       integer(c_int64_t), value :: capacity, stride
       integer(c_int64_t), intent(out) :: n
       type(c_ptr), value :: base
     end function
     integer(c_int) function edgpu_solve_poisson_fftx_noalloc(N, gmj, N_of_particles_cell, phi_below, phi_above, &
          rho_e, rho_i, phi, stream) bind(C, name="edgpu_solve_poisson_fftx_noalloc")
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: N, gmj, N_of_particles_cell
       real(c_double), value :: phi_below, phi_above
       type(c_ptr), value :: rho_e, rho_i, phi, stream
     end function
     integer(c_int) function edgpu_particle_count(species, n) bind(C, name="edgpu_particle_count")
       import :: c_int, c_int32_t, c_int64_t
       integer(c_int32_t), value :: species
       integer(c_int64_t), intent(out) :: n
     end function
     integer(c_int) function edgpu_distribute_particles(species, n, x0, x1, y0, y1, factor_convert, copy_pos) &
          bind(C, name="edgpu_distribute_particles")
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int32_t), value :: species
       integer(c_int64_t), value :: n
       real(c_double), value :: x0, x1, y0, y1, factor_convert
       integer(c_int32_t), value :: copy_pos
     end function
     integer(c_int) function edgpu_gather_charge_density(species, c_rho) bind(C, name="edgpu_gather_charge_density")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: species
       type(c_ptr), value :: c_rho
     end function
     integer(c_int) function edgpu_update_wall_potentials(n_obj, phi) bind(C, name="edgpu_update_wall_potentials")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: n_obj
       real(c_double), intent(in) :: phi(*)
     end function
     integer(c_int) function edgpu_get_surface_charge(below, above) bind(C, name="edgpu_get_surface_charge")
       import :: c_int, c_ptr
       type(c_ptr), value :: below, above
     end function
     integer(c_int) function edgpu_set_surface_charge(below, above) bind(C, name="edgpu_set_surface_charge")
       import :: c_int, c_ptr
       type(c_ptr), value :: below, above
     end function
     integer(c_int) function edgpu_solve_poisson_fftx(c_phi) bind(C, name="edgpu_solve_poisson_fftx")
       import :: c_int, c_ptr
       type(c_ptr), value :: c_phi
     end function
     integer(c_int) function edgpu_solve_potential_krylov(rtol, max_iter, warm_start, iterations, relres, c_phi) &
          bind(C, name="edgpu_solve_potential_krylov")
       import :: c_int, c_int32_t, c_double, c_ptr
       real(c_double), value :: rtol
       integer(c_int32_t), value :: max_iter, warm_start
       integer(c_int32_t), intent(out) :: iterations
       real(c_double), intent(out) :: relres
       type(c_ptr), value :: c_phi
     end function
     integer(c_int) function edgpu_set_krylov_preconditioner(kind) bind(C, name="edgpu_set_krylov_preconditioner")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: kind
     end function
     ! learned warm start: the uint64_t seed passes as a 64-bit integer of the same bits
     integer(c_int) function edgpu_cnn_init(seed, max_pairs) bind(C, name="edgpu_cnn_init")
       import :: c_int, c_int32_t, c_int64_t
       integer(c_int64_t), value :: seed
       integer(c_int32_t), value :: max_pairs
     end function
     integer(c_int) function edgpu_cnn_record_pair() bind(C, name="edgpu_cnn_record_pair")
       import :: c_int
     end function
     integer(c_int) function edgpu_cnn_train(steps, lr, mean_loss) bind(C, name="edgpu_cnn_train")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: steps
       real(c_double), value :: lr
       real(c_double), intent(out) :: mean_loss
     end function
     integer(c_int) function edgpu_calculate_electric_field(EX, EY) bind(C, name="edgpu_calculate_electric_field")
       import :: c_int, c_ptr
       type(c_ptr), value :: EX, EY
     end function
     integer(c_int) function edgpu_clear_accumulated_fields() bind(C, name="edgpu_clear_accumulated_fields")
       import :: c_int
     end function
     integer(c_int) function edgpu_get_grid(which, out) bind(C, name="edgpu_get_grid")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: which
       type(c_ptr), value :: out
     end function
     integer(c_int) function edgpu_set_grid(which, in) bind(C, name="edgpu_set_grid")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: which
       type(c_ptr), value :: in
     end function
     integer(c_int) function edgpu_advance_electrons(stats) bind(C, name="edgpu_advance_electrons")
       import :: c_int, edgpu_advance_stats
       type(edgpu_advance_stats), intent(out) :: stats
     end function
     integer(c_int) function edgpu_advance_ions(stats) bind(C, name="edgpu_advance_ions")
       import :: c_int, edgpu_advance_stats
       type(edgpu_advance_stats), intent(out) :: stats
     end function
     integer(c_int) function edgpu_get_removed(species, capacity, n, idx, obj) bind(C, name="edgpu_get_removed")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       integer(c_int32_t), value :: species
       integer(c_int64_t), value :: capacity
       integer(c_int64_t), intent(out) :: n
       type(c_ptr), value :: idx, obj
     end function
     integer(c_int) function edgpu_process_added_particles(species) bind(C, name="edgpu_process_added_particles")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: species
     end function
     ! whole_object%setD_ii_ee_true(s) etc. of a material with ion-wall model 2 (pic2d_Materials.f90:151-180)
     integer(c_int) function edgpu_setup_ion_induced_emission(object, species, setD, minE_eV, maxE_eV, T_eV, T_e_eV, &
          N_max_vel, V_scale_ms) bind(C, name="edgpu_setup_ion_induced_emission")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: object, species
       real(c_double), value :: setD, minE_eV, maxE_eV, T_eV, T_e_eV
       integer(c_int32_t), value :: N_max_vel
       real(c_double), value :: V_scale_ms
     end function
     integer(c_int) function edgpu_setup_ion_reflection(object, reflects) bind(C, name="edgpu_setup_ion_reflection")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: object, reflects
     end function
     integer(c_int) function edgpu_setup_electron_induced_emission(object, p, T_e_eV, N_max_vel, V_scale_ms) &
          bind(C, name="edgpu_setup_electron_induced_emission")
       import :: c_int, c_int32_t, c_double, edgpu_see_params
       integer(c_int32_t), value :: object
       type(edgpu_see_params), intent(in) :: p
       real(c_double), value :: T_e_eV
       integer(c_int32_t), value :: N_max_vel
       real(c_double), value :: V_scale_ms
     end function
     integer(c_int) function edgpu_setup_constant_emission(object, N_per_step, model, Te_normal_eV, Te_parallel_eV, &
          We_beam_eV, T_e_eV, N_max_vel) bind(C, name="edgpu_setup_constant_emission")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: object, model, N_max_vel
       real(c_double), value :: N_per_step, Te_normal_eV, Te_parallel_eV, We_beam_eV, T_e_eV
     end function
     integer(c_int) function edgpu_perform_electron_emission(emitted) bind(C, name="edgpu_perform_electron_emission")
       import :: c_int, c_ptr
       type(c_ptr), value :: emitted
     end function
     ! generator position for SAVE_CHECKPOINT_POSIX / READ_CHECKPOINT_POSIX in place of
     ! get_rng_state / set_rng_state (pic2d_RandomNumberInterface.f90:9-19)
     integer(c_int) function edgpu_get_rng_state(seed, counter, n_sub) bind(C, name="edgpu_get_rng_state")
       import :: c_int, c_int32_t, c_int64_t
       integer(c_int64_t), intent(out) :: seed
       integer(c_int32_t), intent(out) :: counter, n_sub
     end function
     integer(c_int) function edgpu_set_rng_state(seed, counter, n_sub) bind(C, name="edgpu_set_rng_state")
       import :: c_int, c_int32_t, c_int64_t
       integer(c_int64_t), value :: seed
       integer(c_int32_t), value :: counter, n_sub
     end function
     integer(c_int) function edgpu_setup_electron_neutral_collisions(n_proc, ptype, npts, energy_eV, sigma_m2, &
          ion_species, ion_velocity_factor, M_amu, N_m3, delta_t_s, n_seg, seg_value, seg_step, V_scale_ms) &
          bind(C, name="edgpu_setup_electron_neutral_collisions")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: n_proc, n_seg
       integer(c_int32_t), intent(in) :: ptype(*), npts(*), ion_species(*)
       real(c_double), intent(in) :: energy_eV(*), sigma_m2(*), ion_velocity_factor(*), seg_value(*), seg_step(*)
       real(c_double), value :: M_amu, N_m3, delta_t_s, V_scale_ms
     end function
     integer(c_int) function edgpu_get_collision_table(n_values, energy_eV, prob, pmax) &
          bind(C, name="edgpu_get_collision_table")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), intent(out) :: n_values
       type(c_ptr), value :: energy_eV, prob, pmax
     end function
     integer(c_int) function edgpu_setup_charge_exchange(species, sigma_1eV, alpha, N_m3, T_K, delta_t_s, T_e_eV, &
          N_max_vel, V_scale_ms) bind(C, name="edgpu_setup_charge_exchange")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: species, N_max_vel
       real(c_double), value :: sigma_1eV, alpha, N_m3, T_K, delta_t_s, T_e_eV, V_scale_ms
     end function
     integer(c_int) function edgpu_perform_electron_neutral_collisions(counters) &
          bind(C, name="edgpu_perform_electron_neutral_collisions")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(inout) :: counters(*)
     end function
     integer(c_int) function edgpu_perform_resonant_charge_exchange(counter) &
          bind(C, name="edgpu_perform_resonant_charge_exchange")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(inout) :: counter
     end function
     integer(c_int) function edgpu_set_uniform_stream(u, n) bind(C, name="edgpu_set_uniform_stream")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: u
       integer(c_int64_t), value :: n
     end function
     subroutine edgpu_philox4x32_10(ctr, key, out) bind(C, name="edgpu_philox4x32_10")
       import :: c_int32_t
       integer(c_int32_t), intent(in) :: ctr(4), key(2)
       integer(c_int32_t), intent(out) :: out(4)
     end subroutine
     integer(c_int) function edgpu_get_maxwell_tables(v, v_inj) bind(C, name="edgpu_get_maxwell_tables")
       import :: c_int, c_double
       real(c_double), intent(out) :: v(0:9000), v_inj(0:4500)
     end function
     integer(c_int) function edgpu_collect_moments(species, acc) bind(C, name="edgpu_collect_moments")
       import :: c_int, c_int32_t, c_float
       integer(c_int32_t), value :: species
       real(c_float), intent(out) :: acc(*)
     end function
     integer(c_int) function edgpu_advance_electrons_and_moments(stats, acc) &
          bind(C, name="edgpu_advance_electrons_and_moments")
       import :: c_int, c_float, edgpu_advance_stats
       type(edgpu_advance_stats), intent(out) :: stats
       real(c_float), intent(out) :: acc(*)
     end function
     integer(c_int) function edgpu_advance_ions_and_moments(stats, acc) bind(C, name="edgpu_advance_ions_and_moments")
       import :: c_int, c_float, edgpu_advance_stats
       type(edgpu_advance_stats), intent(out) :: stats
       real(c_float), intent(out) :: acc(*)
     end function
     integer(c_int) function edgpu_step(krylov_rtol, krylov_warm_start, e_stats, i_stats) bind(C, name="edgpu_step")
       import :: c_int, c_int32_t, c_double, edgpu_advance_stats
       real(c_double), value :: krylov_rtol
       integer(c_int32_t), value :: krylov_warm_start
       type(edgpu_advance_stats), intent(out) :: e_stats, i_stats
     end function
     integer(c_int) function edgpu_last_step_timing(ms8) bind(C, name="edgpu_last_step_timing")
       import :: c_int, c_double
       real(c_double), intent(out) :: ms8(8)
     end function
     integer(c_int) function edgpu_sort_particles(species) bind(C, name="edgpu_sort_particles")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: species
     end function
     integer(c_int) function edgpu_sort_with_next_push(species) bind(C, name="edgpu_sort_with_next_push")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: species
     end function
     integer(c_int) function edgpu_set_sorting_push(on) bind(C, name="edgpu_set_sorting_push")
       import :: c_int, c_int32_t
       integer(c_int32_t), value :: on
     end function
     integer(c_int) function edgpu_copy_probe(nbytes, repeats, ms) bind(C, name="edgpu_copy_probe")
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int64_t), value :: nbytes
       integer(c_int32_t), value :: repeats
       real(c_double), intent(out) :: ms
     end function
  end interface

contains

  ! print the library's message for a failed call and stop the run
  subroutine edgpu_check(status, where)
    integer(c_int), intent(in) :: status
    character(*), intent(in) :: where
#ifdef EDGPU_WITH_MPI
    include 'mpif.h'
    integer :: ierr
#endif
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == EDGPU_OK) return
    call c_f_pointer(edgpu_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    print '("Error in ",A," (status ",i3,"): ",1024A1)', where, status, msg(1:n)
#ifdef EDGPU_WITH_MPI
    call MPI_ABORT(MPI_COMM_WORLD, int(status), ierr)
#else
    error stop 1
#endif
  end subroutine edgpu_check

end module edgpu


! Drop-in replacements of the reference's parameterless hot-path subroutines.  The device
! context owns particles and grids; the optional arguments copy a grid back to a Fortran
! array for diagnostics (snapshots, probes) and are otherwise left out.
module edgpu_drop_in
  use, intrinsic :: iso_c_binding
  use edgpu
  implicit none

contains

  ! ADVANCE_ELECTRONS (pic2d_ElectronDynamics.f90:3-513); electron_hit_count(1:n_obj)
  ! receives the wall hits per object of this call.
  subroutine GPU_ADVANCE_ELECTRONS(electron_hit_count)
    integer, intent(inout), optional :: electron_hit_count(:)
    type(edgpu_advance_stats) :: st
    call edgpu_check(edgpu_advance_electrons(st), "ADVANCE_ELECTRONS")
    if (present(electron_hit_count)) electron_hit_count = electron_hit_count + &
         int(st%hit_count(1:size(electron_hit_count)))
  end subroutine

  ! ADVANCE_IONS (pic2d_IonDynamics.f90:3-526)
  subroutine GPU_ADVANCE_IONS(ion_hit_count)
    integer, intent(inout), optional :: ion_hit_count(:)
    type(edgpu_advance_stats) :: st
    call edgpu_check(edgpu_advance_ions(st), "ADVANCE_IONS")
    if (present(ion_hit_count)) ion_hit_count = ion_hit_count + int(st%hit_count(1:size(ion_hit_count)))
  end subroutine

  ! GATHER_ELECTRON_CHARGE_DENSITY (pic2d_ElectronDynamics.f90:1535-1975)
  subroutine GPU_GATHER_ELECTRON_CHARGE_DENSITY(c_rho)
    real(c_double), intent(out), target, optional :: c_rho(:,:)
    type(c_ptr) :: p
    p = c_null_ptr
    if (present(c_rho)) p = c_loc(c_rho)
    call edgpu_check(edgpu_gather_charge_density(0_c_int32_t, p), "GATHER_ELECTRON_CHARGE_DENSITY")
  end subroutine

  ! GATHER_ION_CHARGE_DENSITY (pic2d_IonDynamics.f90:1540-2003), all ion species summed
  subroutine GPU_GATHER_ION_CHARGE_DENSITY(c_rho_i)
    real(c_double), intent(out), target, optional :: c_rho_i(:,:)
    type(c_ptr) :: p
    p = c_null_ptr
    if (present(c_rho_i)) p = c_loc(c_rho_i)
    call edgpu_check(edgpu_gather_charge_density(1_c_int32_t, p), "GATHER_ION_CHARGE_DENSITY")
  end subroutine

  ! SOLVE_POISSON_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:3-1320)
  subroutine GPU_SOLVE_POISSON_FFTX_LINSYSY(c_phi)
    real(c_double), intent(out), target, optional :: c_phi(:,:)
    type(c_ptr) :: p
    p = c_null_ptr
    if (present(c_phi)) p = c_loc(c_phi)
    call edgpu_check(edgpu_solve_poisson_fftx(p), "SOLVE_POISSON_FFTX_LINSYSY")
  end subroutine

  ! SOLVE_POTENTIAL_WITH_PETSC (pic2d_ElectricFieldCalc_PETSc.F90:7-355); rtol as -ksp_rtol
  subroutine GPU_SOLVE_POTENTIAL_WITH_PETSC(rtol, warm_start, iterations, c_phi)
    real(c_double), intent(in) :: rtol
    integer, intent(in) :: warm_start
    integer, intent(out), optional :: iterations
    real(c_double), intent(out), target, optional :: c_phi(:,:)
    integer(c_int32_t) :: it
    real(c_double) :: rr
    type(c_ptr) :: p
    p = c_null_ptr
    if (present(c_phi)) p = c_loc(c_phi)
    call edgpu_check(edgpu_solve_potential_krylov(rtol, 100000_c_int32_t, int(warm_start, c_int32_t), it, rr, p), &
         "SOLVE_POTENTIAL_WITH_PETSC")
    if (present(iterations)) iterations = it
  end subroutine

  ! CALCULATE_ELECTRIC_FIELD_FFTX_LINSYSY / CALCULATE_ELECTRIC_FIELD
  subroutine GPU_CALCULATE_ELECTRIC_FIELD(EX, EY)
    real(c_double), intent(out), target, optional :: EX(:,:), EY(:,:)
    type(c_ptr) :: px, py
    px = c_null_ptr
    py = c_null_ptr
    if (present(EX)) px = c_loc(EX)
    if (present(EY)) py = c_loc(EY)
    call edgpu_check(edgpu_calculate_electric_field(px, py), "CALCULATE_ELECTRIC_FIELD")
  end subroutine

  ! CLEAR_ACCUMULATED_FIELDS (pic2d_ElectricFieldCalc_PETSc.F90:809-829)
  subroutine GPU_CLEAR_ACCUMULATED_FIELDS()
    call edgpu_check(edgpu_clear_accumulated_fields(), "CLEAR_ACCUMULATED_FIELDS")
  end subroutine

  ! Hand the reference's particle array to the device (after DISTRIBUTE_PARTICLES or a
  ! checkpoint read) and back (before a snapshot or checkpoint): base = c_loc(part(1)) of a
  ! TYPE(particle), TARGET array, stride = the byte distance between part(1) and part(2),
  ! species 0 for electron(:), s for ion(s)%part(:).  The device arrays grow as needed.
  subroutine GPU_PARTICLES_UPLOAD_AOS(species, n, base, stride)
    integer, intent(in) :: species, n
    type(c_ptr), intent(in) :: base
    integer(c_int64_t), intent(in) :: stride
    call edgpu_check(edgpu_particles_upload_aos(int(species, c_int32_t), int(n, c_int64_t), base, stride), &
         "PARTICLES_UPLOAD")
  end subroutine

  ! n receives the particle count (N_electrons / N_ions(s)); capacity is the size of the array
  subroutine GPU_PARTICLES_DOWNLOAD_AOS(species, capacity, n, base, stride)
    integer, intent(in) :: species, capacity
    integer, intent(out) :: n
    type(c_ptr), intent(in) :: base
    integer(c_int64_t), intent(in) :: stride
    integer(c_int64_t) :: m
    call edgpu_check(edgpu_particles_download_aos(int(species, c_int32_t), int(capacity, c_int64_t), m, base, &
         stride), "PARTICLES_DOWNLOAD")
    n = int(m)
  end subroutine

  ! UPDATE_WALL_POTENTIALS (pic2d_WallPotentials.f90:1-54): whole_object(1:n)%phi, normalised
  subroutine GPU_UPDATE_WALL_POTENTIALS(phi)
    real(c_double), intent(in) :: phi(:)
    call edgpu_check(edgpu_update_wall_potentials(int(size(phi), c_int32_t), phi), "UPDATE_WALL_POTENTIALS")
  end subroutine

  ! c_local_object_part%surface_charge of the dielectric walls below and above (0:gmi each),
  ! read for a checkpoint and restored from one (pic2d_Checkpoints.f90)
  subroutine GPU_GET_SURFACE_CHARGE(below, above)
    real(c_double), intent(out), target :: below(0:), above(0:)
    call edgpu_check(edgpu_get_surface_charge(c_loc(below), c_loc(above)), "GET_SURFACE_CHARGE")
  end subroutine

  subroutine GPU_SET_SURFACE_CHARGE(below, above)
    real(c_double), intent(in), target :: below(0:), above(0:)
    call edgpu_check(edgpu_set_surface_charge(c_loc(below), c_loc(above)), "SET_SURFACE_CHARGE")
  end subroutine

  ! PROCESS_ADDED_ELECTRONS (pic2d_ElectronDynamics.f90:1429-1531)
  subroutine GPU_PROCESS_ADDED_ELECTRONS()
    call edgpu_check(edgpu_process_added_particles(0_c_int32_t), "PROCESS_ADDED_ELECTRONS")
  end subroutine

  ! PROCESS_ADDED_IONS (pic2d_IonDynamics.f90:1468-1537), ion species s (1-based)
  subroutine GPU_PROCESS_ADDED_IONS(s)
    integer, intent(in) :: s
    call edgpu_check(edgpu_process_added_particles(int(s, c_int32_t)), "PROCESS_ADDED_IONS")
  end subroutine

  ! PERFORM_ELECTRON_NEUTRAL_COLLISIONS (pic2d_enCollisionsGeneralProc.f90:332-654)
  subroutine GPU_PERFORM_ELECTRON_NEUTRAL_COLLISIONS(counters)
    integer(c_int64_t), intent(inout) :: counters(:)
    call edgpu_check(edgpu_perform_electron_neutral_collisions(counters), "PERFORM_ELECTRON_NEUTRAL_COLLISIONS")
  end subroutine

  ! PERFORM_RESONANT_CHARGE_EXCHANGE (pic2d_inCollisionsSpecificProc.f90:48-126)
  subroutine GPU_PERFORM_RESONANT_CHARGE_EXCHANGE(counter)
    integer(c_int64_t), intent(inout) :: counter
    call edgpu_check(edgpu_perform_resonant_charge_exchange(counter), "PERFORM_RESONANT_CHARGE_EXCHANGE")
  end subroutine

  ! COLLECT_ELECTRON_MOMENTS (pic2d_ElectronMoments.f90:3-424) / COLLECT_ION_MOMENTS
  ! (pic2d_IonMoments.f90:3-347): raw FP32 accumulators, 13 node arrays back to back
  subroutine GPU_COLLECT_MOMENTS(species, acc)
    integer, intent(in) :: species
    real(c_float), intent(out) :: acc(:)
    call edgpu_check(edgpu_collect_moments(int(species, c_int32_t), acc), "COLLECT_MOMENTS")
  end subroutine

  ! ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS_2D (pic2d_ElectronDynamicsWithVDFMoments.f90:156-1010)
  ! and ADVANCE_IONS_AND_CALCULATE_MOMENTS_2D (pic2d_IonDynamicsWithVDFMoments.f90:180-1026): the
  ! pushes of snapshot steps, which also sum the moments (half-step velocity, positions before
  ! the step) in the layout of GPU_COLLECT_MOMENTS, one block of 13 node arrays per species
  subroutine GPU_ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS(acc, electron_hit_count)
    real(c_float), intent(out) :: acc(:)
    integer, intent(inout), optional :: electron_hit_count(:)
    type(edgpu_advance_stats) :: st
    call edgpu_check(edgpu_advance_electrons_and_moments(st, acc), "ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS_2D")
    if (present(electron_hit_count)) electron_hit_count = electron_hit_count + &
         int(st%hit_count(1:size(electron_hit_count)))
  end subroutine

  subroutine GPU_ADVANCE_IONS_AND_CALCULATE_MOMENTS(acc, ion_hit_count)
    real(c_float), intent(out) :: acc(:)
    integer, intent(inout), optional :: ion_hit_count(:)
    type(edgpu_advance_stats) :: st
    call edgpu_check(edgpu_advance_ions_and_moments(st, acc), "ADVANCE_IONS_AND_CALCULATE_MOMENTS_2D")
    if (present(ion_hit_count)) ion_hit_count = ion_hit_count + int(st%hit_count(1:size(ion_hit_count)))
  end subroutine

  ! preconditioner of the Krylov solve (the -pc_type of petsc.rc): 0 Jacobi, 1 multigrid
  ! (default), 2 multigrid with every sweep launched separately (same results)
  subroutine GPU_SET_KRYLOV_PRECONDITIONER(kind)
    integer, intent(in) :: kind
    call edgpu_check(edgpu_set_krylov_preconditioner(int(kind, c_int32_t)), "SET_KRYLOV_PRECONDITIONER")
  end subroutine

  ! the electrons' sort done by the next GPU_ADVANCE_ELECTRONS, which writes them in the tile
  ! order counted here (no counterpart in the reference, whose particles stay unordered)
  subroutine GPU_SORT_ELECTRONS_WITH_NEXT_PUSH()
    call edgpu_check(edgpu_sort_with_next_push(0_c_int32_t), "SORT_ELECTRONS_WITH_NEXT_PUSH")
  end subroutine

  ! learned warm start (no counterpart in the reference): allocate the network, record the
  ! (features, target) pair of each warm-up solve, then train it on the device; solves with
  ! warm_start = 2 start from its prediction
  subroutine GPU_CNN_INIT(seed, max_pairs)
    integer(c_int64_t), intent(in) :: seed
    integer, intent(in) :: max_pairs
    call edgpu_check(edgpu_cnn_init(seed, int(max_pairs, c_int32_t)), "CNN_INIT")
  end subroutine

  subroutine GPU_CNN_RECORD_PAIR()
    call edgpu_check(edgpu_cnn_record_pair(), "CNN_RECORD_PAIR")
  end subroutine

  subroutine GPU_CNN_TRAIN(steps, lr, mean_loss)
    integer, intent(in) :: steps
    real(c_double), intent(in) :: lr
    real(c_double), intent(out), optional :: mean_loss
    real(c_double) :: loss
    call edgpu_check(edgpu_cnn_train(int(steps, c_int32_t), lr, loss), "CNN_TRAIN")
    if (present(mean_loss)) mean_loss = loss
  end subroutine

end module edgpu_drop_in
