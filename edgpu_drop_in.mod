﻿!mod$ v1 sum:b2207d31d65bd6e4
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ ac8a25b2ed536ae6 n edgpu
module edgpu_drop_in
use edgpu,only:edgpu_ok
use edgpu,only:edgpu_fp64
use edgpu,only:edgpu_mixed
use edgpu,only:edgpu_max_seg
use edgpu,only:edgpu_max_inner
use edgpu,only:edgpu_max_obj
use edgpu,only:edgpu_max_ion_species
use edgpu,only:edgpu_max_proc
use edgpu,only:edgpu_side_left
use edgpu,only:edgpu_side_above
use edgpu,only:edgpu_side_right
use edgpu,only:edgpu_side_below
use edgpu,only:edgpu_params
use edgpu,only:edgpu_advance_stats
use edgpu,only:edgpu_see_params
use edgpu,only:edgpu_init
use edgpu,only:edgpu_finalize
use edgpu,only:edgpu_last_error
use edgpu,only:edgpu_device
use edgpu,only:edgpu_particles_upload
use edgpu,only:edgpu_particles_download
use edgpu,only:edgpu_particle_count
use edgpu,only:edgpu_distribute_particles
use edgpu,only:edgpu_gather_charge_density
use edgpu,only:edgpu_update_wall_potentials
use edgpu,only:edgpu_solve_poisson_fftx
use edgpu,only:edgpu_solve_potential_krylov
use edgpu,only:edgpu_calculate_electric_field
use edgpu,only:edgpu_clear_accumulated_fields
use edgpu,only:edgpu_get_grid
use edgpu,only:edgpu_set_grid
use edgpu,only:edgpu_advance_electrons
use edgpu,only:edgpu_advance_ions
use edgpu,only:edgpu_get_removed
use edgpu,only:edgpu_process_added_particles
use edgpu,only:edgpu_setup_electron_induced_emission
use edgpu,only:edgpu_setup_constant_emission
use edgpu,only:edgpu_perform_electron_emission
use edgpu,only:edgpu_get_rng_state
use edgpu,only:edgpu_set_rng_state
use edgpu,only:edgpu_setup_electron_neutral_collisions
use edgpu,only:edgpu_get_collision_table
use edgpu,only:edgpu_setup_charge_exchange
use edgpu,only:edgpu_perform_electron_neutral_collisions
use edgpu,only:edgpu_perform_resonant_charge_exchange
use edgpu,only:edgpu_set_uniform_stream
use edgpu,only:edgpu_philox4x32_10
use edgpu,only:edgpu_get_maxwell_tables
use edgpu,only:edgpu_collect_moments
use edgpu,only:edgpu_step
use edgpu,only:edgpu_last_step_timing
use edgpu,only:edgpu_sort_particles
use edgpu,only:edgpu_copy_probe
use edgpu,only:edgpu_check
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
contains
subroutine gpu_advance_electrons(electron_hit_count)
integer(4),intent(inout),optional::electron_hit_count(:)
end
subroutine gpu_advance_ions(ion_hit_count)
integer(4),intent(inout),optional::ion_hit_count(:)
end
subroutine gpu_gather_electron_charge_density(c_rho)
real(8),intent(out),optional,target::c_rho(:,:)
end
subroutine gpu_gather_ion_charge_density(c_rho_i)
real(8),intent(out),optional,target::c_rho_i(:,:)
end
subroutine gpu_solve_poisson_fftx_linsysy(c_phi)
real(8),intent(out),optional,target::c_phi(:,:)
end
subroutine gpu_solve_potential_with_petsc(rtol,warm_start,iterations,c_phi)
real(8),intent(in)::rtol
integer(4),intent(in)::warm_start
integer(4),intent(out),optional::iterations
real(8),intent(out),optional,target::c_phi(:,:)
end
subroutine gpu_calculate_electric_field(ex,ey)
real(8),intent(out),optional,target::ex(:,:)
real(8),intent(out),optional,target::ey(:,:)
end
subroutine gpu_clear_accumulated_fields()
end
subroutine gpu_process_added_electrons()
end
subroutine gpu_process_added_ions(s)
integer(4),intent(in)::s
end
subroutine gpu_perform_electron_neutral_collisions(counters)
integer(8),intent(inout)::counters(:)
end
subroutine gpu_perform_resonant_charge_exchange(counter)
integer(8),intent(inout)::counter
end
subroutine gpu_collect_moments(species,acc)
integer(4),intent(in)::species
real(4),intent(out)::acc(:)
end
end
