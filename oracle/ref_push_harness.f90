! Test harness (checker only) around the reference's own particle pushes, ADVANCE_ELECTRONS
! (pic2d_ElectronDynamics.f90:1-515) and ADVANCE_IONS (pic2d_IonDynamics.f90:3-524), with the
! wall and inner-object collision routines they call (pic2d_ElectronWallCollisions.f90,
! pic2d_IonWallCollisions.f90), compiled from the reference's sources where they lie.
!
! The reference sets its module state in INITIATE_PARAMETERS and ANALYZE_BOUNDARY_OBJECTS
! (pic2d_CurProblemValues.f90), which cannot be built here (the file uses the PETSc solver
! module).  refh_setup sets the same variables for one cluster that holds the whole domain:
! a walled box, or an X-periodic pipeline connected to itself between walls below and above,
! with optional metal inner boxes that touch neither each other nor the walls (so no
! cell of their faces is covered).  refh_advance_electrons copies caller arrays into
! electron(:), runs ADVANCE_ELECTRONS and copies out what is left (in the reference's own
! REMOVE_ELECTRON order), the hit counts per object and the list of hits that
! ADD_ELECTRON_TO_BO_COLLS_LIST keeps (token = tag + 100 * segment, the REAL coordinate).
!
! Objects are numbered as the reference numbers them (1-based); walls are given as segments
! (side 0 left, 1 above, 2 right, 3 below; node range along the wall; 0-based object).

subroutine refh_setup(gmi, gmj, periodic, n_seg, seg_side, seg_start, seg_end, seg_obj, n_boundary_obj, &
                      n_inner, boxes, bx_in, by_in, bz_in, ez_in) bind(C, name="refh_setup")
  use, intrinsic :: iso_c_binding
  USE ParallelOperationValues
  USE CurrentProblemValues
  USE ElectronParticles
  USE ClusterAndItsBoundaries
  USE ExternalFields
  USE Snapshots
  implicit none
  integer(c_int32_t), value :: gmi, gmj, periodic, n_seg, n_boundary_obj, n_inner
  integer(c_int32_t), intent(in) :: seg_side(*), seg_start(*), seg_end(*), seg_obj(*)
  integer(c_int32_t), intent(in) :: boxes(4, *)
  real(c_double), value :: bx_in, by_in, bz_in, ez_in
  integer :: n, m, nobj, b(4)

  Rank_of_process = 0
  Rank_of_master_left = -1
  Rank_of_master_right = -1
  Rank_of_master_above = -1
  Rank_of_master_below = -1

  c_X_area_min = 0.0_8
  c_X_area_max = DBLE(gmi)
  c_Y_area_min = 0.0_8
  c_Y_area_max = DBLE(gmj)
  c_indx_x_min = 0
  c_indx_x_max = gmi
  c_indx_y_min = 0
  c_indx_y_max = gmj
  periodic_boundary_X_left = periodic /= 0
  periodic_boundary_X_right = periodic /= 0
  periodic_boundary_Y_below = .FALSE.
  periodic_boundary_Y_above = .FALSE.
  L_period_X = DBLE(gmi - 1)
  symmetry_plane_X_left = .FALSE.
  c_left_bottom_corner_type = SURROUNDED_BY_WALL
  c_left_top_corner_type = SURROUNDED_BY_WALL
  c_right_bottom_corner_type = SURROUNDED_BY_WALL
  c_right_top_corner_type = SURROUNDED_BY_WALL

  Bx_ext = bx_in
  By_ext = by_in
  Bz_ext = bz_in
  Ez_ext = ez_in
  N_JZ_wires = 0

  N_of_boundary_objects = n_boundary_obj
  N_of_inner_objects = n_inner
  N_of_boundary_and_inner_objects = n_boundary_obj + n_inner
  nobj = N_of_boundary_and_inner_objects
  if (allocated(whole_object)) deallocate(whole_object)
  allocate(whole_object(1:nobj))
  do n = 1, nobj
     whole_object(n)%object_type = METAL_WALL
     whole_object(n)%SEE_enabled = .FALSE.
     whole_object(n)%electron_hit_count = 0
     whole_object(n)%electron_emit_count = 0
     whole_object(n)%number_of_segments = 0
  end do
  if (periodic /= 0) then
     whole_object(1)%object_type = PERIODIC_PIPELINE_X
     if (n_boundary_obj >= 3) whole_object(3)%object_type = PERIODIC_PIPELINE_X
  end if
  ! inner boxes: segments 1 left, 2 top, 3 right, 4 bottom, none of their cells covered
  do n = 1, n_inner
     b = boxes(:, n)
     associate (o => whole_object(n_boundary_obj + n))
       o%ileft = b(1)
       o%iright = b(2)
       o%jbottom = b(3)
       o%jtop = b(4)
       o%Xmin = DBLE(b(1))
       o%Xmax = DBLE(b(2))
       o%Ymin = DBLE(b(3))
       o%Ymax = DBLE(b(4))
       o%number_of_segments = 4
       allocate(o%segment(1:4))
       o%segment(1)%istart = b(1); o%segment(1)%jstart = b(3); o%segment(1)%iend = b(1); o%segment(1)%jend = b(4)
       o%segment(2)%istart = b(1); o%segment(2)%jstart = b(4); o%segment(2)%iend = b(2); o%segment(2)%jend = b(4)
       o%segment(3)%istart = b(2); o%segment(3)%jstart = b(4); o%segment(3)%iend = b(2); o%segment(3)%jend = b(3)
       o%segment(4)%istart = b(2); o%segment(4)%jstart = b(3); o%segment(4)%iend = b(1); o%segment(4)%jend = b(3)
       allocate(o%segment(1)%cell_is_covered(b(3):b(4) - 1))
       allocate(o%segment(2)%cell_is_covered(b(1):b(2) - 1))
       allocate(o%segment(3)%cell_is_covered(b(3):b(4) - 1))
       allocate(o%segment(4)%cell_is_covered(b(1):b(2) - 1))
       do m = 1, 4
          o%segment(m)%cell_is_covered = .FALSE.
       end do
     end associate
  end do

  ! the cluster's parts of the walls
  c_N_of_local_object_parts = n_seg
  c_N_of_local_object_parts_left = 0
  c_N_of_local_object_parts_right = 0
  c_N_of_local_object_parts_above = 0
  c_N_of_local_object_parts_below = 0
  do m = 1, n_seg
     c_local_object_part(m)%object_number = seg_obj(m) + 1
     c_local_object_part(m)%segment_number = 1
     select case (seg_side(m))
     case (0, 2)
        c_local_object_part(m)%istart = merge(0, gmi, seg_side(m) == 0)
        c_local_object_part(m)%iend = c_local_object_part(m)%istart
        c_local_object_part(m)%jstart = seg_start(m)
        c_local_object_part(m)%jend = seg_end(m)
     case default
        c_local_object_part(m)%jstart = merge(gmj, 0, seg_side(m) == 1)
        c_local_object_part(m)%jend = c_local_object_part(m)%jstart
        c_local_object_part(m)%istart = seg_start(m)
        c_local_object_part(m)%iend = seg_end(m)
     end select
     select case (seg_side(m))
     case (0)
        c_N_of_local_object_parts_left = c_N_of_local_object_parts_left + 1
        c_index_of_local_object_part_left(c_N_of_local_object_parts_left) = m
     case (1)
        c_N_of_local_object_parts_above = c_N_of_local_object_parts_above + 1
        c_index_of_local_object_part_above(c_N_of_local_object_parts_above) = m
     case (2)
        c_N_of_local_object_parts_right = c_N_of_local_object_parts_right + 1
        c_index_of_local_object_part_right(c_N_of_local_object_parts_right) = m
     case default
        c_N_of_local_object_parts_below = c_N_of_local_object_parts_below + 1
        c_index_of_local_object_part_below(c_N_of_local_object_parts_below) = m
     end select
  end do

  ! every hit is kept in e_colls_with_bo (ADD_ELECTRON_TO_BO_COLLS_LIST)
  if (allocated(e_colls_with_bo)) deallocate(e_colls_with_bo)
  allocate(e_colls_with_bo(1:nobj))
  do n = 1, nobj
     e_colls_with_bo(n)%must_be_saved = .TRUE.
     e_colls_with_bo(n)%max_N_of_saved_parts = 0
     e_colls_with_bo(n)%N_of_saved_parts = 0
  end do
  N_of_all_snaps = 1
  current_snap = 1
  if (allocated(save_e_collided_with_bo)) deallocate(save_e_collided_with_bo)
  allocate(save_e_collided_with_bo(1:1))
  save_e_collided_with_bo(1) = .TRUE.

  if (allocated(EX)) deallocate(EX)
  if (allocated(EY)) deallocate(EY)
  allocate(EX(0:gmi, 0:gmj), EY(0:gmi, 0:gmj))
  EX = 0.0_8
  EY = 0.0_8
  N_e_to_send_left = 0
  N_e_to_send_right = 0
  N_e_to_send_above = 0
  N_e_to_send_below = 0
  max_N_e_to_send_left = 0
  max_N_e_to_send_right = 0
  max_N_e_to_send_above = 0
  max_N_e_to_send_below = 0
end subroutine refh_setup

! One ADVANCE_ELECTRONS over n electrons (tags 1..n) in the field ex, ey (node arrays with x
! fastest).  Returns the number left; out_* receive the electrons left in the reference's
! order, hits(1:nobj) the hit count per object, and the first max_rec wall hits in the order
! the reference saves them, object by object: rec_obj (1-based), rec_token, rec_coord.
integer(c_int64_t) function refh_advance_electrons(n, x, y, vx, vy, vz, ex_in, ey_in, out_x, out_y, out_vx, &
                                                   out_vy, out_vz, out_tag, hits, max_rec, n_rec, rec_obj, &
                                                   rec_token, rec_coord) bind(C, name="refh_advance_electrons")
  use, intrinsic :: iso_c_binding
  USE CurrentProblemValues
  USE ElectronParticles
  USE ClusterAndItsBoundaries
  implicit none
  integer(c_int64_t), value :: n, max_rec
  real(c_double), intent(in) :: x(*), y(*), vx(*), vy(*), vz(*)
  real(c_double), intent(in) :: ex_in(*), ey_in(*)
  real(c_double), intent(out) :: out_x(*), out_y(*), out_vx(*), out_vy(*), out_vz(*)
  integer(c_int32_t), intent(out) :: out_tag(*), hits(*), rec_obj(*), rec_token(*)
  real(c_float), intent(out) :: rec_coord(*)
  integer(c_int64_t), intent(out) :: n_rec
  integer :: k, nn, i, j, p
  integer(c_int64_t) :: r

  nn = int(n)
  if (allocated(electron)) deallocate(electron)
  allocate(electron(1:max(nn, 1)))
  max_N_electrons = max(nn, 1)
  N_electrons = nn
  do k = 1, nn
     electron(k)%X = x(k)
     electron(k)%Y = y(k)
     electron(k)%VX = vx(k)
     electron(k)%VY = vy(k)
     electron(k)%VZ = vz(k)
     electron(k)%tag = k
  end do
  p = 0
  do j = c_indx_y_min, c_indx_y_max
     do i = c_indx_x_min, c_indx_x_max
        p = p + 1
        EX(i, j) = ex_in(p)
        EY(i, j) = ey_in(p)
     end do
  end do

  CALL ADVANCE_ELECTRONS

  do k = 1, N_electrons
     out_x(k) = electron(k)%X
     out_y(k) = electron(k)%Y
     out_vx(k) = electron(k)%VX
     out_vy(k) = electron(k)%VY
     out_vz(k) = electron(k)%VZ
     out_tag(k) = electron(k)%tag
  end do
  r = 0
  do i = 1, N_of_boundary_and_inner_objects
     hits(i) = whole_object(i)%electron_hit_count
     do k = 1, e_colls_with_bo(i)%N_of_saved_parts
        r = r + 1
        if (r <= max_rec) then
           rec_obj(r) = i
           rec_token(r) = e_colls_with_bo(i)%part(k)%token
           rec_coord(r) = e_colls_with_bo(i)%part(k)%coll_coord
        end if
     end do
  end do
  n_rec = r
  refh_advance_electrons = N_electrons
end function refh_advance_electrons

! The ion side of the same cluster: n_spec species with charges qs(:) and QM2s(:) as the
! reference defines them (SET_ION_SPECIES), N_subcycles, and the two switches of the ion push.
! Call after refh_setup.
subroutine refh_setup_ions(n_spec_in, qs_in, qm2s_in, n_subcycles_in, sense_b, sense_ez) &
     bind(C, name="refh_setup_ions")
  use, intrinsic :: iso_c_binding
  USE CurrentProblemValues
  USE IonParticles
  USE ClusterAndItsBoundaries
  USE Snapshots
  implicit none
  integer(c_int32_t), value :: n_spec_in, n_subcycles_in, sense_b, sense_ez
  integer(c_int32_t), intent(in) :: qs_in(*)
  real(c_double), intent(in) :: qm2s_in(*)
  integer :: n, s

  N_spec = n_spec_in
  N_subcycles = n_subcycles_in
  ions_sense_magnetic_field = sense_b /= 0
  ions_sense_EZ = sense_ez /= 0
  if (allocated(Qs)) deallocate(Qs, QM2s, QM2sNsub, N_ions, max_N_ions)
  if (allocated(N_ions_to_send_left)) deallocate(N_ions_to_send_left, N_ions_to_send_right, &
       N_ions_to_send_above, N_ions_to_send_below, max_N_ions_to_send_left, max_N_ions_to_send_right, &
       max_N_ions_to_send_above, max_N_ions_to_send_below)
  allocate(Qs(N_spec), QM2s(N_spec), QM2sNsub(N_spec), N_ions(N_spec), max_N_ions(N_spec))
  allocate(N_ions_to_send_left(N_spec), N_ions_to_send_right(N_spec), N_ions_to_send_above(N_spec), &
           N_ions_to_send_below(N_spec), max_N_ions_to_send_left(N_spec), max_N_ions_to_send_right(N_spec), &
           max_N_ions_to_send_above(N_spec), max_N_ions_to_send_below(N_spec))
  do s = 1, N_spec
     Qs(s) = qs_in(s)
     QM2s(s) = qm2s_in(s)
     QM2sNsub(s) = QM2s(s) * DBLE(N_subcycles)
  end do
  N_ions = 0
  max_N_ions = 0
  max_N_ions_to_send_left = 0
  max_N_ions_to_send_right = 0
  max_N_ions_to_send_above = 0
  max_N_ions_to_send_below = 0
  if (allocated(ion)) deallocate(ion)
  allocate(ion(1:N_spec))
  do n = 1, N_of_boundary_and_inner_objects
     if (allocated(whole_object(n)%ion_hit_count)) deallocate(whole_object(n)%ion_hit_count)
     allocate(whole_object(n)%ion_hit_count(1:N_spec))
     whole_object(n)%ion_hit_count = 0
     whole_object(n)%reflects_all_ions = .FALSE.
     whole_object(n)%ion_induced_EE_enabled = .FALSE.
  end do
  if (allocated(ion_colls_with_bo)) deallocate(ion_colls_with_bo)
  allocate(ion_colls_with_bo(1:N_of_boundary_and_inner_objects))
  do n = 1, N_of_boundary_and_inner_objects
     ion_colls_with_bo(n)%must_be_saved = .TRUE.
     ion_colls_with_bo(n)%max_N_of_saved_parts = 0
     ion_colls_with_bo(n)%N_of_saved_parts = 0
  end do
  if (allocated(save_ions_collided_with_bo)) deallocate(save_ions_collided_with_bo)
  allocate(save_ions_collided_with_bo(1:1))
  save_ions_collided_with_bo(1) = .TRUE.
  if (allocated(acc_EX)) deallocate(acc_EX)
  if (allocated(acc_EY)) deallocate(acc_EY)
  allocate(acc_EX(c_indx_x_min:c_indx_x_max, c_indx_y_min:c_indx_y_max))
  allocate(acc_EY(c_indx_x_min:c_indx_x_max, c_indx_y_min:c_indx_y_max))
end subroutine refh_setup_ions

! One ADVANCE_IONS (pic2d_IonDynamics.f90:3-524) with n ions of species s (tags 1..n, the other
! species empty) in the accumulated field acc_ex, acc_ey.  Outputs as refh_advance_electrons;
! ion records carry token = tag + 100 * segment + 10000 * s.
integer(c_int64_t) function refh_advance_ions(s, n, x, y, vx, vy, vz, ex_in, ey_in, out_x, out_y, out_vx, &
                                              out_vy, out_vz, out_tag, hits, max_rec, n_rec, rec_obj, &
                                              rec_token, rec_coord) bind(C, name="refh_advance_ions")
  use, intrinsic :: iso_c_binding
  USE CurrentProblemValues
  USE IonParticles
  USE ClusterAndItsBoundaries
  implicit none
  integer(c_int32_t), value :: s
  integer(c_int64_t), value :: n, max_rec
  real(c_double), intent(in) :: x(*), y(*), vx(*), vy(*), vz(*)
  real(c_double), intent(in) :: ex_in(*), ey_in(*)
  real(c_double), intent(out) :: out_x(*), out_y(*), out_vx(*), out_vy(*), out_vz(*)
  integer(c_int32_t), intent(out) :: out_tag(*), hits(*), rec_obj(*), rec_token(*)
  real(c_float), intent(out) :: rec_coord(*)
  integer(c_int64_t), intent(out) :: n_rec
  integer :: k, nn, i, j, p, q
  integer(c_int64_t) :: r

  nn = int(n)
  do q = 1, N_spec
     if (allocated(ion(q)%part)) deallocate(ion(q)%part)
     allocate(ion(q)%part(1:max(nn, 1)))
     N_ions(q) = 0
     max_N_ions(q) = max(nn, 1)
  end do
  N_ions(s) = nn
  do k = 1, nn
     ion(s)%part(k)%X = x(k)
     ion(s)%part(k)%Y = y(k)
     ion(s)%part(k)%VX = vx(k)
     ion(s)%part(k)%VY = vy(k)
     ion(s)%part(k)%VZ = vz(k)
     ion(s)%part(k)%tag = k
  end do
  p = 0
  do j = c_indx_y_min, c_indx_y_max
     do i = c_indx_x_min, c_indx_x_max
        p = p + 1
        acc_EX(i, j) = ex_in(p)
        acc_EY(i, j) = ey_in(p)
     end do
  end do

  CALL ADVANCE_IONS

  do k = 1, N_ions(s)
     out_x(k) = ion(s)%part(k)%X
     out_y(k) = ion(s)%part(k)%Y
     out_vx(k) = ion(s)%part(k)%VX
     out_vy(k) = ion(s)%part(k)%VY
     out_vz(k) = ion(s)%part(k)%VZ
     out_tag(k) = ion(s)%part(k)%tag
  end do
  r = 0
  do i = 1, N_of_boundary_and_inner_objects
     hits(i) = whole_object(i)%ion_hit_count(s)
     do k = 1, ion_colls_with_bo(i)%N_of_saved_parts
        r = r + 1
        if (r <= max_rec) then
           rec_obj(r) = i
           rec_token(r) = ion_colls_with_bo(i)%part(k)%token
           rec_coord(r) = ion_colls_with_bo(i)%part(k)%coll_coord
        end if
     end do
  end do
  n_rec = r
  refh_advance_ions = N_ions(s)
end function refh_advance_ions
