/*
 * edgpu -- MI355X (gfx950) implementation of the EDIPIC-2D per-time-step
 * particle hot path behind a plain C ABI.
 *
 * Every entry point replaces one parameterless Fortran subroutine of the
 * reference (cited per declaration, paths relative to the reference root).
 * The reference's subroutines work on module globals; here the module state
 * lives in one device-resident context created by edgpu_init and owned by the
 * library.  Host arrays cross the boundary only in edgpu_particles_upload /
 * _download and in the optional host copies of grids.
 *
 * Conventions
 *  - Return value: 0 (EDGPU_OK) on success, one of EDGPU_ERR_* otherwise.
 *    A non-zero status from a per-step entry point corresponds to a branch
 *    of the reference that calls MPI_ABORT; the context state is then
 *    undefined and the caller is expected to stop (the Fortran shim calls
 *    MPI_ABORT).  edgpu_last_error() returns a human-readable message.
 *  - Indices are 0-based in this ABI (Fortran object n is object n-1 here,
 *    ion species s is s-1).
 *  - Grids are FP64, node (i,j), 0<=i<=gmi, 0<=j<=gmj, stored x fastest:
 *    index i + j*(gmi+1) -- the memory order of the reference's
 *    c_rho(c_indx_x_min:c_indx_x_max, c_indx_y_min:c_indx_y_max) for the
 *    single cluster that covers the whole domain.
 *  - All calls are synchronous with respect to the host and must be made
 *    from one host thread (the reference is one thread per rank).
 *  - Quantities are in the reference's normalised units (positions in
 *    cells, velocities in V_scale_ms per step, fields as in EX/EY).
 */
#ifndef EDGPU_H
#define EDGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDGPU_OK 0
#define EDGPU_ERR_ARG 1            /* invalid argument or configuration            */
#define EDGPU_ERR_HIP 2            /* HIP runtime failure / no device              */
#define EDGPU_ERR_OUT_OF_BOUNDS 3  /* particle outside the area (ADVANCE_* Error-1) */
#define EDGPU_ERR_INDEX 4          /* cell index outside the grid (Error-2)         */
#define EDGPU_ERR_NO_WALL 5        /* wall hit with no boundary object part there   */
#define EDGPU_ERR_CAPACITY 6       /* particle array or add list full              */
#define EDGPU_ERR_TABLE 7          /* MCC energy table inconsistency                */
#define EDGPU_ERR_STATE 8          /* no context / call out of order               */
#define EDGPU_ERR_NOT_CONVERGED 9  /* Krylov solve did not reach the tolerance      */
#define EDGPU_ERR_SEE 10           /* elastic + inelastic emission yield above 1    */

#define EDGPU_FP64 0   /* particles as FP64 SoA (x, y, vx, vy, vz)                      */
#define EDGPU_MIXED 1  /* int32 cell + FP32 in-cell offsets + FP32 velocities           */

#define EDGPU_MAX_SEG 16
#define EDGPU_MAX_INNER 8
#define EDGPU_MAX_OBJ (EDGPU_MAX_SEG + EDGPU_MAX_INNER)
#define EDGPU_MAX_ION_SPECIES 4
#define EDGPU_MAX_PROC 8

#define EDGPU_SIDE_LEFT 0
#define EDGPU_SIDE_ABOVE 1
#define EDGPU_SIDE_RIGHT 2
#define EDGPU_SIDE_BELOW 3

/* boundary object types (edgpu_params.object_type); the reference's whole_object%object_type
 * METAL_WALL = 1, DIELECTRIC = 4 and VACUUM_GAP = 0 (pic2d_Modules.f90:143-148) */
#define EDGPU_OBJ_METAL 0
#define EDGPU_OBJ_DIELECTRIC 1
#define EDGPU_OBJ_VACUUM_GAP 2

/* Problem description: the subset of the reference's module globals the hot
 * path reads (pic2d_Modules.f90 CurrentProblemValues / ClusterAndItsBoundaries /
 * IonParticles / ExternalFields). */
typedef struct edgpu_params {
    int32_t gmi, gmj;               /* global_maximal_i / global_maximal_j                    */
    int32_t periodic_x;             /* 1: self-connected X-periodic cluster, L_period_X=gmi-1 */
    int32_t N_of_particles_cell;    /* N_of_particles_cell (rho normalisation)                 */
    int32_t N_subcycles;            /* ion sub-cycling (odd)                                  */
    int32_t precision;              /* EDGPU_FP64 or EDGPU_MIXED                              */
    double Bx_ext, By_ext, Bz_ext;  /* normalised uniform B (pic2d_ExternalFields.f90:196-256) */
    double Ez_ext;                  /* normalised uniform Ez (pic2d_ExternalFields.f90:258)    */
    double factor1_E_from_F;        /* F_scale_V/(E_scale_Vm*delta_x_m)                        */
    double factor2_E_from_F;        /* F_scale_V/(E_scale_Vm*2*delta_x_m)                      */
    /* boundary: objects 0..n_boundary_objects-1 are walls described by segments,
     * objects n_boundary_objects.. are inner metal boxes (same order as the deck) */
    int32_t n_boundary_objects;
    int32_t n_segments;
    int32_t seg_side[EDGPU_MAX_SEG];   /* EDGPU_SIDE_*                                       */
    int32_t seg_start[EDGPU_MAX_SEG];  /* node range along the wall (inclusive)              */
    int32_t seg_end[EDGPU_MAX_SEG];
    int32_t seg_object[EDGPU_MAX_SEG]; /* object index (0-based)                             */
    int32_t n_inner_objects;
    int32_t inner_box[EDGPU_MAX_INNER][4]; /* ileft, iright, jbottom, jtop                   */
    double object_phi[EDGPU_MAX_OBJ];      /* normalised potential (phi/F_scale_V) per object */
    /* ion species (pic2d_CurProblemValues.f90:1467-1510) */
    int32_t n_ion_species;
    int32_t Qs[EDGPU_MAX_ION_SPECIES];
    double QM2s[EDGPU_MAX_ION_SPECIES];
    double Ms[EDGPU_MAX_ION_SPECIES];
    int32_t ions_sense_magnetic_field;
    int32_t ions_sense_Ez;
    /* capacities (particles) of the device arrays; the add lists get add_capacity */
    int64_t max_electrons;
    int64_t max_ions[EDGPU_MAX_ION_SPECIES];
    int64_t add_capacity;
    uint64_t seed;                  /* key of the counter-based RNG                         */
    int32_t fuse_collisions;        /* 1: e-n MCC and RCX run inside the push kernels       */
    int32_t sort_interval;          /* electron steps between per-tile counting sorts (0=off)*/
    int32_t ion_sort_interval;      /* steps between sorts of the ions (0: sort_interval)   */
    /* per boundary object: EDGPU_OBJ_METAL (0, the default of a zeroed struct),
     * EDGPU_OBJ_DIELECTRIC with its relative permittivity whole_object%eps_diel
     * (pic2d_Materials.f90:45,82), or EDGPU_OBJ_VACUUM_GAP.  Dielectric objects lie on the y
     * walls of an x-periodic system solved by FFT_X (pic2d_Prepare_FFTX_SYSY.f90:776-897).  A
     * vacuum gap has one segment, whose ends touch other boundary objects; its nodes hold
     * the potential interpolated linearly between those objects' potentials
     * (pic2d_CurProblemValues.f90:334-390, pic2d_WallPotentials.f90:29-52). */
    int32_t object_type[EDGPU_MAX_OBJ];
    double object_eps[EDGPU_MAX_OBJ];
} edgpu_params;

/* Per-call result of ADVANCE_ELECTRONS / ADVANCE_IONS. */
typedef struct edgpu_advance_stats {
    int64_t n_before, n_after;                 /* particles before/after the call          */
    int64_t hit_count[EDGPU_MAX_OBJ];          /* electron_hit_count / ion_hit_count        */
    int64_t n_removed;
    int64_t collisions[EDGPU_MAX_PROC];        /* fused MCC events per process (e-) or RCX (ions, [0]) */
    int64_t emit_count[EDGPU_MAX_OBJ];         /* electron_emit_count: electrons emitted by SEE     */
} edgpu_advance_stats;

/* Electron-induced emission of one boundary object: the material parameters that
 * pic2d_Materials.f90:70-245 reads from init_material_*.dat, energies in eV as in the file
 * (the library makes them dimensionless as the reference does). */
typedef struct edgpu_see_params {
    int32_t model_elastic;       /* Emitted_model(1): 0 off, 1 constant coefficient in (E_min, E_max], 2 model 2 */
    int32_t elastic_reflection;  /* Elast_refl_type: 0 specular, otherwise random angle                       */
    int32_t model_inelastic;     /* Emitted_model(2)                                                           */
    int32_t model_true;          /* Emitted_model(3)                                                           */
    double setD_elastic, minE_elastic_eV, maxE_elastic_eV;
    double E_elast_0_eV, E_elast_max_eV, maxD_elast, dE_elast_eV, frac_elast_highenergy;
    double setD_inelastic, minE_inelastic_eV, maxE_inelastic_eV;
    double frac_inelastic;
    double T_true_eV;            /* T_see_true_eV: temperature of the injected true secondaries                 */
    double setD_true, minE_true_eV, maxE_true_eV;
    double E_see_0_eV, E_see_max_eV, maxD_see_classic, k_smooth;
} edgpu_see_params;

/* ---- lifetime --------------------------------------------------------------------------- */

/* INITIATE_PARAMETERS device part (pic2d_CurProblemValues.f90:59-1253): allocates the
 * context, the grids, the particle arrays, builds the Maxwell tables
 * (PrepareMaxwellDistribIntegral, pic2d_ElectronWallCollisions.f90:835-929) and the
 * Thomas coefficients (PREPARE_SYS_Y, pic2d_Prepare_FFTX_SYSY.f90:805-897). */
int edgpu_init(const edgpu_params *p);
int edgpu_finalize(void);
const char *edgpu_last_error(void);
/* HIP device ordinal the context runs on, or -1 */
int edgpu_device(void);

/* ---- particles at the boundary ------------------------------------------------------------ */

/* species: 0 = electrons (electron(:)), 1..n_ion_species = ion(s)%part(:).
 * Host arrays are the columns X, Y, VX, VY, VZ, tag of the reference's TYPE particle
 * (pic2d_Modules.f90:432-439); tag may be NULL (zero tags).  An upload larger than the
 * capacity, like an append that does not fit (PROCESS_ADDED_*), grows the device arrays as
 * the reference grows electron(:) (pic2d_ElectronDynamics.f90:1458-1482): by what is missing
 * or a tenth of the capacity, whichever is more. */
int edgpu_particles_upload(int32_t species, int64_t n, const double *X, const double *Y, const double *VX,
                           const double *VY, const double *VZ, const int32_t *tag);
int edgpu_particles_download(int32_t species, int64_t capacity, int64_t *n, double *X, double *Y, double *VX,
                             double *VY, double *VZ, int32_t *tag);
/* The same for the reference's array of TYPE particle itself (pic2d_Modules.f90:432-439: REAL(8)
 * X, Y, VX, VY, VZ at byte offsets 0..32, INTEGER tag at 40): base is the first element
 * (c_loc(electron(1))) and stride the distance in bytes between consecutive elements (48 with
 * the usual padding), so the caller hands its array over without splitting it into columns. */
int edgpu_particles_upload_aos(int32_t species, int64_t n, const void *base, int64_t stride);
int edgpu_particles_download_aos(int32_t species, int64_t capacity, int64_t *n, void *base, int64_t stride);
int edgpu_particle_count(int32_t species, int64_t *n);
/* DISTRIBUTE_PARTICLES (pic2d_CurProblemValues.f90:3127-3324) on the device: uniform
 * positions in [x0,x1]x[y0,y1], Maxwellian velocities times factor_convert, drawn from
 * the counter-based RNG (statistically equivalent, not the WELL sequence).  Of n electrons
 * (species 0) drawn, those strictly inside an inner object are removed as the reference
 * removes them (pic2d_CurProblemValues.f90:3245-3258): edgpu_particle_count then reports
 * fewer than n, and that count is the n to pass for the ions.  For ions (species>0) with
 * copy_electron_positions=1 the first n electron positions are reused and VZ=0 as in the
 * reference. */
int edgpu_distribute_particles(int32_t species, int64_t n, double x0, double x1, double y0, double y1,
                               double factor_convert, int32_t copy_electron_positions);

/* ---- charge density and field solve -------------------------------------------------------- */

/* GATHER_ELECTRON_CHARGE_DENSITY (pic2d_ElectronDynamics.f90:1535-1975) for species 0,
 * GATHER_ION_CHARGE_DENSITY (pic2d_IonDynamics.f90:1540-2003) for species>0: bilinear count
 * deposit, periodic self-fold and wall doubling.  c_rho may be NULL (device only). */
int edgpu_gather_charge_density(int32_t species, double *c_rho);
/* UPDATE_WALL_POTENTIALS (pic2d_WallPotentials.f90:1-54): new normalised object potentials */
int edgpu_update_wall_potentials(int32_t n_obj, const double *phi);
/* The surface charge of the dielectric parts of the walls below and above, gmi+1 nodes each
 * in macroparticle charges (c_local_object_part%surface_charge, pic2d_CurProblemValues.f90:
 * 1640-1650; accumulated by the wall hits of both species, pic2d_ElectronWallCollisions.f90:
 * 184-195, pic2d_IonWallCollisions.f90:130-140, and by electron-induced emission,
 * pic2d_SEEmission.f90:259-280; read by SOLVE_POISSON_FFTX_LINSYSY, pic2d_ElectricFieldCalc_
 * FFT_X.f90:100-145).  Read, or set as a checkpoint restores it (pic2d_Checkpoints.f90).
 * Either pointer may be NULL. */
int edgpu_get_surface_charge(double *below, double *above);
int edgpu_set_surface_charge(const double *below, const double *above);
/* SOLVE_POISSON_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:3-1320); requires periodic_x,
 * no inner objects (they switch the reference to PETSc, pic2d_CurProblemValues.f90:460-465;
 * edgpu_step then uses the Krylov solve) and gmi-1 a power of two <= 4096.  c_phi may be NULL. */
int edgpu_solve_poisson_fftx(double *c_phi);
/* SOLVE_POTENTIAL_WITH_PETSC (pic2d_ElectricFieldCalc_PETSc.F90:7-355) as a GPU
 * preconditioned CG on the same 5-point operator with Dirichlet walls and metal inner
 * objects, to relative residual rtol (the role of -ksp_rtol).  warm_start: 0 zero guess (the
 * reference's KSP default), 1 previous potential, 2 the learned guess (edgpu_cnn_*).  c_phi
 * may be NULL. */
int edgpu_solve_potential_krylov(double rtol, int32_t max_iter, int32_t warm_start, int32_t *iterations,
                                 double *relres, double *c_phi);
/* preconditioner of the Krylov solve: 0 Jacobi, 1 geometric multigrid V-cycle (default; the
 * reference's petsc.rc preconditions its GMRES with hypre's algebraic multigrid; its coarse
 * levels run in one LDS-resident workgroup), 2 the same V-cycle with every level's sweeps
 * launched one by one (same results bit for bit, slower) */
int edgpu_set_krylov_preconditioner(int32_t kind);
/* CALCULATE_ELECTRIC_FIELD_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:1324-1547) and
 * CALCULATE_ELECTRIC_FIELD (pic2d_ElectricFieldCalc_PETSc.F90:360-807): EX, EY from the
 * current potential and acc_EX/acc_EY += EX/EY.  EX/EY may be NULL. */
int edgpu_calculate_electric_field(double *EX, double *EY);
/* CLEAR_ACCUMULATED_FIELDS (pic2d_ElectricFieldCalc_PETSc.F90:809-829) */
int edgpu_clear_accumulated_fields(void);
/* test/diagnostic access to the device grids: which = 0 c_rho(e) 1 c_rho_i 2 c_phi 3 EX 4 EY
 * 5 acc_EX 6 acc_EY.  set_grid overwrites the device copy. */
int edgpu_get_grid(int32_t which, double *out);
int edgpu_set_grid(int32_t which, const double *in);

/* ---- particle push ---------------------------------------------------------------------- */

/* ADVANCE_ELECTRONS (pic2d_ElectronDynamics.f90:3-513): bilinear gather of EX/EY, Boris
 * push, move, inner-object and wall hits (counted per object), periodic wrap, removal.
 * The positions after the push are deposited on the fly for the next
 * GATHER_ELECTRON_CHARGE_DENSITY.  With fuse_collisions the e-n collisions of this step
 * are performed here for the pushed electrons. */
int edgpu_advance_electrons(edgpu_advance_stats *stats);
/* ADVANCE_IONS (pic2d_IonDynamics.f90:3-526) for every ion species (acc_EX/acc_EY gather,
 * N_subcycles move).  With fuse_collisions RCX runs here. */
int edgpu_advance_ions(edgpu_advance_stats *stats);
/* indices (before compaction) and object of the particles removed by the last advance of
 * this species -- the wall-crossing set */
int edgpu_get_removed(int32_t species, int64_t capacity, int64_t *n, int64_t *index, int32_t *object);
/* the rest of each record of that list, in the same order: the position (for a boundary wall
 * the position after the push and a periodic wrap; for an inner object the crossing of the
 * step's trajectory with its face, as TRY_ELECTRON_COLL_WITH_INNER_OBJECT finds it,
 * pic2d_ElectronWallCollisions.f90:470-557) and the side of the wall crossed, EDGPU_SIDE_*
 * (for an inner object the side of a wall facing the same way as the face hit: left face
 * EDGPU_SIDE_RIGHT, top face EDGPU_SIDE_BELOW, right face EDGPU_SIDE_LEFT, bottom face
 * EDGPU_SIDE_ABOVE) -- the boundary-collision list of ADD_ELECTRON_TO_BO_COLLS_LIST */
int edgpu_get_removed_records(int32_t species, int64_t capacity, int64_t *n, double *x, double *y, int32_t *side);
/* PROCESS_ADDED_ELECTRONS (pic2d_ElectronDynamics.f90:1429-1531) / PROCESS_ADDED_IONS
 * (pic2d_IonDynamics.f90:1468-1537): append the add list to the particle array */
int edgpu_process_added_particles(int32_t species);

/* ---- wall emission ------------------------------------------------------------------------ */

/* Material of object `object` (0-based; the reference's object_id_number - 1; boundary walls
 * and metal inner boxes): pic2d_Materials.f90:49-245.  Electrons absorbed by an object with
 * any model on then go through PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION
 * (pic2d_SEEmission.f90:4-151) inside edgpu_advance_electrons: elastic reflection
 * (:155-557), inelastic backscattering (:561-790) and true secondaries (:794-1050), yields
 * Coeff_SEE_* (:1054-1208); at an inner box from the crossing of the electron's step with the
 * face it hit.  The emitted electrons join the electron add list with tag object+1 and count
 * in emit_count. */
int edgpu_setup_electron_induced_emission(int32_t object, const edgpu_see_params *p, double T_e_eV, int32_t N_max_vel,
                                          double V_scale_ms);
/* Ion-induced electron emission of object `object` (boundary wall or inner box) for ion species
 * `species` (0-based), the material's ion-wall interaction model 2 (pic2d_Materials.f90:151-
 * 180): emission coefficient setD for incident energies in (minE, maxE) eV, electrons emitted
 * at temperature T_eV.  Ions that the object absorbs then go through PROCESS_ION_INDUCED_
 * ELECTRON_EMISSION (pic2d_IonEmissions.f90:4-87) inside edgpu_advance_ions: floor(setD)
 * electrons, one more with probability setD - floor(setD), each leaving the wall
 * (INJECT_ION_IMPACT_EXTRACTED_ELECTRON, :91-348) while the ion's energy above minE lasts; they
 * join the electron add list with tag object+1, count in the electron emit_count of the
 * object (reported in the ion push's emit_count) and charge a dielectric wall +1 each.
 * setD = 0 turns the object's emission for the species off. */
int edgpu_setup_ion_induced_emission(int32_t object, int32_t species, double setD, double minE_eV, double maxE_eV,
                                     double T_eV, double T_e_eV, int32_t N_max_vel, double V_scale_ms);
/* Ion-wall interaction model 1 of object `object` (pic2d_Materials.f90:156-160,
 * reflects_all_ions): every ion the object absorbs is counted as a hit and comes back
 * specularly 1e-6 off the wall or face it hit (INJECT_REFLECTED_ION, pic2d_IonEmissions.f90:
 * 374-492) through the ion add list, tag object+1, instead of charging a dielectric or emitting
 * electrons (pic2d_IonWallCollisions.f90:43-47).  reflects = 0 restores absorption. */
int edgpu_setup_ion_reflection(int32_t object, int32_t reflects);
/* Constant emission of boundary object `object`: the init_bo_NN.dat values of
 * PREPARE_SETUP_VALUES (pic2d_Setup.f90:45-126): macroparticles per step, model (0 thermal,
 * 1 beam), normal temperature / beam half-energy spread, parallel temperature, beam energy. */
int edgpu_setup_constant_emission(int32_t object, double N_per_step, int32_t model, double Te_normal_eV,
                                  double Te_parallel_eV, double We_beam_eV, double T_e_eV, int32_t N_max_vel);
/* PERFORM_ELECTRON_EMISSION_SETUP (pic2d_Setup.f90:599-1001): the constant emission of every
 * boundary object into the electron add list; emitted[object] += electrons emitted (may be
 * NULL).  edgpu_step calls it after the ion push, as the reference main loop does
 * (pic2d_MainProgram.f90:353). */
int edgpu_perform_electron_emission(int64_t *emitted);

/* ---- collisions --------------------------------------------------------------------------- */

/* INITIATE_ELECTRON_NEUTRAL_COLLISIONS table part (pic2d_enCollisionsGeneralProc.f90:200-325):
 * builds prob_colproc_energy from the activated processes' cross sections (energy_eV and
 * sigma concatenated per process, npts[p] points each), frequencies scaled by
 * N_m3*N_subcycles*delta_t_s. */
int edgpu_setup_electron_neutral_collisions(int32_t n_proc, const int32_t *type, const int32_t *npts,
                                            const double *energy_eV, const double *sigma_m2, const int32_t *ion_species,
                                            const double *ion_velocity_factor, double M_amu, double N_m3,
                                            double delta_t_s, int32_t n_seg, const double *seg_value,
                                            const double *seg_step, double V_scale_ms);
/* read back the table (tests): sizes first with NULL pointers */
int edgpu_get_collision_table(int32_t *n_values, double *energy_eV, double *prob, double *max_colliding_fraction);
/* upper bounds of the total selection probability over the energy bins the fused push uses
 * to drop certain null collisions (bin b covers [(0.5 + (b%4)/8) 2^(b/4-20), (0.5 +
 * (b%4+1)/8) 2^(b/4-20)) eV); sizes first with bound = NULL */
int edgpu_get_collision_bounds(int32_t *nbins, double *bound);
/* INITIATE_ION_NEUTRAL_COLLISIONS + calculate_thermal_cx_probab
 * (pic2d_inCollisionsGeneralProc.f90:3-90, pic2d_inCollisionsSpecificProc.f90:4-44) for ion species */
int edgpu_setup_charge_exchange(int32_t species, double sigma_1eV, double alpha, double N_m3, double T_K,
                                double delta_t_s, double T_e_eV, int32_t N_max_vel, double V_scale_ms);
/* PERFORM_ELECTRON_NEUTRAL_COLLISIONS (pic2d_enCollisionsGeneralProc.f90:332-654) as
 * per-electron null-collision selection with probability max_colliding_fraction; process
 * selection and kinematics as the reference (pic2d_enCollisionsSpecificProc.f90:61-639).
 * New electrons/ions go to the add lists.  counters[p] += events of process p. */
int edgpu_perform_electron_neutral_collisions(int64_t *counters);
/* PERFORM_RESONANT_CHARGE_EXCHANGE (pic2d_inCollisionsSpecificProc.f90:48-126) */
int edgpu_perform_resonant_charge_exchange(int64_t *counter);
/* test-only: replace the counter-based RNG by an explicit uniform stream.  For e-n MCC the
 * electron in slot k uses u[k*11 .. k*11+10] (candidacy, random_r, random_n, 8 kinematic
 * draws); for RCX the ion in slot k uses u[k*4 .. k*4+3]; the emission of an electron
 * absorbed from slot k uses u[k*32 .. k*32+31] in the order the reference draws them; the
 * constant emission uses u[4t .. 4t+3] for its electron t (position, normal and two
 * parallel velocities) and u[4 INT(N)] for the fractional one's test, that electron then
 * using u[4 INT(N) + 4 ..].  n = number of doubles; NULL restores the RNG. */
int edgpu_set_uniform_stream(const double *u, int64_t n);
/* Bound of vz^2 over the particles of an ion species that the ion push keeps when the push
 * cannot change vz: ions whose exchange test fails with any vz^2 up to the bound skip reading
 * vz.  No counterpart in the reference; reported for checks of the bound. */
int edgpu_get_vz_bound(int32_t species, double *vz2max);
/* Counter-based RNG position for checkpoints (the reference saves its WELL state instead,
 * SAVE_CHECKPOINT_POSIX pic2d_Checkpoints.f90:735-740): the seed and the number of
 * streams drawn so far, and the electron sub-cycle counter n_sub. */
int edgpu_get_rng_state(uint64_t *seed, uint32_t *counter, int32_t *n_sub);
int edgpu_set_rng_state(uint64_t seed, uint32_t counter, int32_t n_sub);
/* Philox4x32-10 as used on the device (for known-answer tests) */
void edgpu_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Philox2x32-10, the generator of the two per-particle collision-test uniforms */
void edgpu_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]);
/* Maxwell tables built at init (tests compare them with PrepareMaxwellDistribIntegral) */
int edgpu_get_maxwell_tables(double *v /*9001*/, double *v_inj /*4501*/);

/* ---- moments ------------------------------------------------------------------------------ */

/* COLLECT_ELECTRON_MOMENTS (pic2d_ElectronMoments.f90:3-424, half-step velocity) and
 * COLLECT_ION_MOMENTS (pic2d_IonMoments.f90:3-347): 13 FP32 raw accumulators per node
 * (N, VX, VY, VZ, WX, WY, WZ, VXVY, VXVZ, VYVZ, QX, QY, QZ) before normalisation,
 * 13*(gmi+1)*(gmj+1) floats. */
int edgpu_collect_moments(int32_t species, float *acc);
/* ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS_2D (pic2d_ElectronDynamicsWithVDFMoments.f90:156-1010)
 * and ADVANCE_IONS_AND_CALCULATE_MOMENTS_2D (pic2d_IonDynamicsWithVDFMoments.f90:180-1026), the
 * pushes of snapshot steps: edgpu_advance_electrons / edgpu_advance_ions that also sum, in the
 * same pass over the particles, the 13 moments of every particle at its position before the
 * step with the half-step velocity 0.5 (v + v') (for ions too, unlike COLLECT_ION_MOMENTS).
 * acc receives 13*(gmi+1)*(gmj+1) floats per species (ions: one block per ion species, in
 * species order) laid out as edgpu_collect_moments'; NULL keeps them on the device. */
int edgpu_advance_electrons_and_moments(edgpu_advance_stats *stats, float *acc);
int edgpu_advance_ions_and_moments(edgpu_advance_stats *stats, float *acc);

/* ---- one whole step ----------------------------------------------------------------------- */

/* The hot-path part of one electron step of the reference main loop
 * (pic2d_MainProgram.f90:151-363), in its order: gather rho (ions, electrons), solve (FFT_X if
 * periodic_x else Krylov with rtol/warm_start), electric field, advance electrons; on the last
 * electron step of each ion cycle of N_subcycles steps: advance ions, process added electrons
 * (those emitted by the electrons' wall hits), e-n MCC over the electrons not yet tested this
 * step (with fuse_collisions the pushed electrons were tested inside the push and their
 * ionisation electrons are not collided in their birth step), RCX, process added ions, clear
 * the accumulated fields; on every step constant emission and a final process added
 * electrons.  i_stats is zero on the other electron steps. */
int edgpu_step(double krylov_rtol, int32_t krylov_warm_start, edgpu_advance_stats *e_stats,
               edgpu_advance_stats *i_stats);
/* edgpu_step of a snapshot step: its pushes are the fused ones above (the reference switches to
 * them on snapshot steps); acc_e and acc_i as for those (acc_i is left alone on a step that
 * does not push the ions, N_subcycles > 1). */
int edgpu_step_with_moments(double krylov_rtol, int32_t krylov_warm_start, edgpu_advance_stats *e_stats,
                            edgpu_advance_stats *i_stats, float *acc_e, float *acc_i);
/* Timing of the last edgpu_step (ms, HIP events on the context stream): [0] electron push
 * kernel alone, [1] ion push kernels alone, [2] gather+fold, [3] field solve, [4] E field,
 * [5] MCC+RCX outside the push, [6] appends, constant emission and clearing of the
 * accumulated fields, [7] sort. */
int edgpu_last_step_timing(double *ms8);

/* ---- learned warm start for the Krylov solve --------------------------------------------------
 * No counterpart in the reference, which cold-starts every KSPSolve (no
 * -ksp_initial_guess_nonzero in any petsc.rc).  A 6-layer, 32-channel, 3x3 BF16 convolutional
 * network (MFMA, FP32 accumulation) maps (residual of phi^{n-1} for the new charge,
 * phi^{n-1} - phi^{n-2}, the right-hand side, the free-node mask) to phi^n - phi^{n-1}; the
 * Krylov solve with warm_start=2 starts from phi^{n-1} + that prediction and always finishes
 * to its tolerance, so the plain solve remains the judge and the fallback. */

/* allocate the network (He-normal weights from seed) and room for max_pairs training pairs */
int edgpu_cnn_init(uint64_t seed, int32_t max_pairs);
/* store the (features, target) pair of the last Krylov solve, unnormalised FP32 (ring of
 * max_pairs pairs); the solve must have had a previous potential (not the first one) */
int edgpu_cnn_record_pair(void);
/* `steps` Adam steps (learning rate lr) on randomly drawn recorded pairs, one full grid per
 * step, forward and backward on the device; mean_loss = mean MSE of the last 16 steps */
int edgpu_cnn_train(int32_t steps, double lr, double *mean_loss);
/* input and target scales (root mean squares over the free nodes of the pairs recorded
 * before the first training call): s[0] residual, s[1] phi^{n-1} - phi^{n-2}, s[2]
 * right-hand side, s[3] target phi^n - phi^{n-1}; nfree = free nodes per pair */
int edgpu_cnn_scales(double *s4, int64_t *nfree);
/* FP32 master parameters: per layer the weights [tap][cout][cin] then the bias */
int edgpu_cnn_get_params(int64_t capacity, int64_t *n, float *out);
int edgpu_cnn_set_params(int64_t n, const float *in);
/* test hook: forward, MSE loss and backward on an H x W grid of host features (4 channels
 * per node: residual, increment, right-hand side, free mask) against a host target;
 * returns the prediction (H*W), the parameter gradient (n params) and the loss */
int edgpu_cnn_eval(int32_t H, int32_t W, const float *features, const float *target, float *pred, float *grad,
                   double *loss);

/* ---- diagnostics without a reference counterpart ------------------------------------------- */

/* D (16x16) = A (16x32) B (32x16), row-major FP32 matrices rounded to BF16, computed by one
 * v_mfma_f32_16x16x32_bf16 (checks the lane layout the network kernels assume) */
int edgpu_mfma_probe(const float *A, const float *B, float *D);

/* Per-tile counting sort of one species (order only: by tiles of 16 cells of one row, any
 * order inside a tile; edgpu_step sorts every sort_interval steps).  The reference keeps
 * particles unordered; it has no counterpart. */
int edgpu_sort_particles(int32_t species);
/* Whether edgpu_step sorts the electrons of a sort step with their push (on by default): the
 * step counts the electrons per tile at its start and the electron push writes them to the
 * spare arrays in that order, so the sort makes no pass over the particles of its own.  The
 * order is that of the positions before the push.  Steps that sum moments, the first sort and
 * the ions use the separate sort.  No counterpart in the reference. */
int edgpu_set_sorting_push(int32_t on);
/* Count the electrons per tile now and let the next electron push (edgpu_advance_electrons)
 * write them in that order: the sorting push for callers that run the phases of a step
 * themselves.  Any call that changes the electron array before that push (upload,
 * distribution, appends, a sort) drops the count; electrons never sorted are sorted now
 * instead.  species must be 0.  No counterpart in the reference. */
int edgpu_sort_with_next_push(int32_t species);

/* The FFT_X solve of edgpu_solve_poisson_fftx for callers that cannot allocate device memory
 * (code under HIP graph capture, tight loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Needs no context: the caller passes device arrays rho_e,
 * rho_i (the gathered charge densities, as GATHER_*_CHARGE_DENSITY leaves them) and phi, of
 * (N+2) x (gmj+1) nodes stored x fastest, the X-periodic domain of N cells (a power of two,
 * 4 <= N <= 2048) between metal walls at rows 0 and gmj (2 <= gmj <= 1024) held at phi_below
 * and phi_above; larger grids are refused with EDGPU_ERR_ARG.  Every work array lives in one
 * static device buffer (edgpu_noalloc_scratch_bytes: about 25 MB, held on every device the
 * library runs on until the process ends); nothing is allocated or freed.  The work is queued
 * on `stream` (a hipStream_t; NULL is the null stream) and the call returns without waiting.
 * Calls that overlap on one device share that buffer and race: serialise them (one stream, or
 * events).  The first call on a device looks up the buffer's address (hipGetSymbolAddress);
 * make it outside graph capture. */
int edgpu_solve_poisson_fftx_noalloc(int32_t N, int32_t gmj, int32_t N_of_particles_cell, double phi_below,
                                     double phi_above, const double *rho_e, const double *rho_i, double *phi,
                                     void *stream);
/* size in bytes of that static buffer */
int edgpu_noalloc_scratch_bytes(int64_t *bytes);

/* Bandwidth probe: average time (ms) of a plain device-to-device copy kernel moving nbytes
 * (read nbytes + write nbytes), over `repeats` launches after one warm-up launch.  Used to
 * calibrate HBM byte counters; needs no context. */
int edgpu_copy_probe(int64_t nbytes, int32_t repeats, double *ms);

#ifdef __cplusplus
}
#endif
#endif
