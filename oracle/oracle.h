/*
 * CPU oracle for the EDIPIC-2D particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C, FP64 restatement of
 * the reference Fortran routines that the GPU library replaces.  It is used
 * by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py as the checker and as the timed CPU baseline.  The product path
 * (libedgpu.so) never links, loads or calls it.
 *
 * Every function cites the reference routine it follows (paths relative to
 * the reference repository root).  Pinning status per function is listed in
 * oracle.c next to each definition; functions whose results no closed form,
 * invariant or brute-force check constrains are marked "parity unpinned".
 *
 * Grid arrays are FP64, x fastest: node (i, j) with ixmin<=i<=ixmax,
 * iymin<=j<=iymax lives at [(i-ixmin) + (j-iymin)*nx], nx = ixmax-ixmin+1,
 * the column-major layout of the reference's EX(c_indx_x_min:..., ...).
 */
#ifndef EDIPIC_ORACLE_H
#define EDIPIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- WELL19937a (src/WELL19937a_new.c) ------------------ */
typedef struct orc_well {
    uint32_t s[624];
    int32_t i;
} orc_well;

void orc_well_seed(orc_well *w, int32_t seed);          /* pic2d_RandomNumberInterface.f90:39-57 */
double orc_well_next(orc_well *w);                      /* WELL19937a_new.c:117-204 */
void orc_well_fill(orc_well *w, int64_t n, double *out);

/* ---------------- Maxwell sampler (pic2d_ElectronWallCollisions.f90:835-985) */
#define ORC_R_MAX 9000
#define ORC_R_MAX_INJ 4500
int orc_maxwell_tables(double *v /*[9001]*/, double *v_inj /*[4501]*/);
double orc_maxwell_sample(const double *v, double u);          /* GetMaxwellVelocity */
double orc_maxwell_inj_sample(const double *v_inj, double u);  /* GetInjMaxwellVelocity */

/* ---------------- geometry of the single (whole-domain) cluster ----------- */
#define ORC_MAX_SEG 16
#define ORC_MAX_INNER 8
#define ORC_SIDE_LEFT 0
#define ORC_SIDE_ABOVE 1
#define ORC_SIDE_RIGHT 2
#define ORC_SIDE_BELOW 3

typedef struct orc_geom {
    int32_t ixmin, ixmax, iymin, iymax; /* c_indx_x_min.. (node indices)   */
    int32_t periodic_x;                 /* self-connected X-periodic cluster */
    double  L_period_x;                 /* c_indx_x_max - c_indx_x_min - 1  */
    /* wall segments: side, [start, end] node range along the wall, object id */
    int32_t n_seg;
    int32_t seg_side[ORC_MAX_SEG];
    int32_t seg_start[ORC_MAX_SEG];
    int32_t seg_end[ORC_MAX_SEG];
    int32_t seg_obj[ORC_MAX_SEG];
    /* objects 0..n_obj-1: boundary objects first, then inner objects */
    int32_t n_obj;
    int32_t n_boundary_obj;
    double  obj_phi[ORC_MAX_SEG + ORC_MAX_INNER]; /* normalised potential */
    /* inner metal boxes [ileft, iright] x [jbottom, jtop] (objects n_boundary_obj..) */
    int32_t n_inner;
    int32_t inner_box[ORC_MAX_INNER][4];
    /* vacuum gaps: segment s of a gap object holds the potential interpolated linearly from
     * object seg_gap_from[s] at its start to seg_gap_to[s] at its end (seg_gap[s] = 1) */
    int32_t seg_gap[ORC_MAX_SEG];
    int32_t seg_gap_from[ORC_MAX_SEG];
    int32_t seg_gap_to[ORC_MAX_SEG];
} orc_geom;

/* per-species push parameters */
typedef struct orc_species {
    int32_t is_electron;      /* 1: ADVANCE_ELECTRONS, 0: ADVANCE_IONS        */
    int32_t Qs;               /* ion charge number (deposit weight)            */
    double  QM2s;             /* 0.5*Qs/Ms                                     */
    double  QM2sNsub;         /* QM2s*N_subcycles                              */
    int32_t N_subcycles;
    int32_t sense_B;          /* ions_sense_magnetic_field (always 1 for e-)   */
    int32_t sense_Ez;         /* ions_sense_Ez (always 1 for e-)               */
} orc_species;

/* push status codes (non-zero mirror the reference MPI_ABORT branches) */
#define ORC_OK 0
#define ORC_ERR_OUT_OF_BOUNDS 1
#define ORC_ERR_INDEX 2
#define ORC_ERR_NO_WALL 3
#define ORC_ERR_TABLE 4

/* ADVANCE_ELECTRONS (pic2d_ElectronDynamics.f90:3-513) / ADVANCE_IONS
 * (pic2d_IonDynamics.f90:3-526) without the swap-with-last compaction:
 * particles are updated in place, removed particles get removed[k]=1 and
 * hit_obj[k]=object index; hit_count[obj] counts hits. */
int orc_advance(const orc_geom *g, const orc_species *sp,
                const double *EX, const double *EY,
                double Bx, double By, double Bz, double Ez,
                int64_t n, double *X, double *Y, double *VX, double *VY, double *VZ,
                uint8_t *removed, int32_t *hit_obj, int64_t *hit_count);

/* Raw bilinear deposit (GATHER_*_CHARGE_DENSITY hot loop,
 * pic2d_ElectronDynamics.f90:1597-1638, pic2d_IonDynamics.f90:1653-1686). */
/* orc_advance, and for every removed particle the position the reference records and the
 * wall crossed as its dirflag (1 left, 2 above, 3 right, 4 below): for a boundary wall the
 * position after the push (after a periodic wrap); for an inner object the crossing of the
 * step's trajectory with the object's surface (orc_inner_crossing).  Arrays may be NULL. */
int orc_advance_ex(const orc_geom *g, const orc_species *sp, const double *EX, const double *EY,
                   double Bx, double By, double Bz, double Ez, int64_t n,
                   double *X, double *Y, double *VX, double *VY, double *VZ,
                   uint8_t *removed, int32_t *hit_obj, int64_t *hit_count, double *rem_x, double *rem_y,
                   int32_t *rem_dir);
/* TRY_ELECTRON_COLL_WITH_INNER_OBJECT's search (pic2d_ElectronWallCollisions.f90:470-557, 648-921):
 * the crossing of (xorg, yorg)-(x, y) with the segments of the inner boxes closest to the
 * origin, skipping covered cells; the object, the dirflag of the hit face as the reference
 * sets it (left face 3, top face 4, right face 1, bottom face 2) and the crossing point.
 * ORC_ERR_NO_WALL when there is none (the reference aborts). */
int orc_inner_crossing(const orc_geom *g, double xorg, double yorg, double x, double y, int32_t *obj,
                       int32_t *dirflag, double *xc, double *yc);
int orc_deposit(const orc_geom *g, int32_t Qs, int64_t n, const double *X, const double *Y,
                const uint8_t *skip, double *rho);
/* periodic self-fold + wall doubling (pic2d_ElectronDynamics.f90:1660-1668,1804-1880) */
void orc_fold_walls(const orc_geom *g, double *rho);

/* Numerical-Recipes real FFT as used by the reference (pic2d_ElectricFieldCalc_FFT_X.f90:1568-1710) */
void orc_real8ft(double *data, int32_t n, int32_t isign);
/* PREPARE_SYS_Y metal-wall branch (pic2d_Prepare_FFTX_SYSY.f90:805-897): a_eq[(j)*(N/2+1)+n] */
void orc_prepare_sys_y(int32_t N, int32_t gmj, double *a_eq);
/* PREPARE_SYS_Y with dielectric walls (pic2d_Prepare_FFTX_SYSY.f90:776-897) */
void orc_prepare_sys_y_walls(int32_t N, int32_t gmj, int32_t diel_below, int32_t diel_above, double eps_below,
                             double eps_above, double *a_eq);
/* SOLVE_POISSON_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:3-1320): c_phi over the grid */
int orc_solve_poisson_fftx(const orc_geom *g, int32_t N_ppc, const double *rho_e, const double *rho_i,
                           const double *a_eq, double *phi);
/* the same with the objects of diel_mask dielectric (surface charge rows over the x nodes;
 * two_dielectric: both walls dielectric, the mode-0 gauge row) */
int orc_solve_poisson_fftx_walls(const orc_geom *g, int32_t N_ppc, const double *rho_e, const double *rho_i,
                                 const double *a_eq, uint32_t diel_mask, const double *sigma_below,
                                 const double *sigma_above, int32_t two_dielectric, double *phi);
/* surface charge of dielectric y walls from absorbed-particle records (object, dirflag, x),
 * q = -1 electrons, Qs ions, +1 emitted electrons (count per record, NULL = 1 each) */
void orc_surface_charge(const orc_geom *g, uint32_t diel_mask, int64_t n, const int32_t *obj, const int32_t *dirflag,
                        const double *x, const int32_t *count, int32_t q, double *sigma_below, double *sigma_above);
/* PROCESS_ION_INDUCED_ELECTRON_EMISSION of one ion absorbed by a boundary wall (energies
 * dimensionless, uniforms u[0..nu) in the order the routine draws them) */
int orc_iiee_emit(double setD, double minE, double maxE, double vinj, double Ms, const double *maxw,
                  const double *maxw_inj, int32_t dirflag, double xmax, double ymax, double x, double y, double vx,
                  double vy, double vz, const double *u, int32_t nu, int32_t max_out, double *out, int32_t *n_out);
/* INJECT_REFLECTED_ION at a boundary wall: out = x, y, vx, vy, vz of the reflected ion */
void orc_ion_reflect(int32_t dirflag, double xmax, double ymax, double x, double y, double vx, double vy, double vz,
                     double *out);
/* the periodic exchange of the end nodes of an x-periodic wall's surface charge */
void orc_surface_charge_fold(const orc_geom *g, double *sigma);
/* CALCULATE_ELECTRIC_FIELD_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:1324-1547) */
void orc_efield(const orc_geom *g, const double *phi, double f1, double f2,
                double *EX, double *EY, double *accEX, double *accEY);

/* Krylov path: operator of SolverInitialization (pic2d_PETSc_Solver.F90:42-1346), eps=1,
 * Dirichlet identity rows on walls and metal inner boxes; RHS of SOLVE_POTENTIAL_WITH_PETSC
 * (pic2d_ElectricFieldCalc_PETSc.F90:7-205).  Returns iterations or <0 on failure. */
void orc_krylov_rhs(const orc_geom *g, int32_t N_ppc, const double *rho_e, const double *rho_i, double *b);
void orc_krylov_apply(const orc_geom *g, const double *x, double *y);
int orc_krylov_cg(const orc_geom *g, const double *b, double *x, double rtol, int32_t max_iter, double *relres);

/* ---------------- electron-neutral MCC ------------------------------------ */
#define ORC_MAX_PROC 8
typedef struct orc_mcc {
    int32_t n_proc;                       /* N_of_activated_colproc */
    int32_t type[ORC_MAX_PROC];           /* 10, 11, 20, 30 */
    double  threshold_eV[ORC_MAX_PROC];
    double  ion_velocity_factor[ORC_MAX_PROC];
    int32_t ion_species[ORC_MAX_PROC];
    double  M_amu;                        /* neutral mass */
    int32_t n_seg;                        /* energy segments */
    double  seg_value[33];                /* 0..n_seg */
    double  seg_step[33];                 /* 1..n_seg */
    int32_t seg_index[33];                /* 0..n_seg */
    int32_t n_values;                     /* N_of_energy_values */
    double *energy_eV;                    /* [n_values+1] */
    double *prob;                         /* [(n_values+1)*n_proc], prob[e*n_proc+p] */
    double  max_colliding_fraction;
    double  energy_factor;                /* V_scale^2*m_e*0.5/e_Cl */
} orc_mcc;

/* frequency_of_en_collision (pic2d_enCollisionsSpecificProc.f90:4-57) */
double orc_en_frequency(double energy_eV, int32_t type, int32_t npts, const double *E_eV, const double *sigma,
                        double N_m3, int32_t N_subcycles, double dt_s);
/* INITIATE_ELECTRON_NEUTRAL_COLLISIONS table part (pic2d_enCollisionsGeneralProc.f90:200-325);
 * sigma tables concatenated: npts[p], offsets implied. energy_eV/prob must hold the sizes
 * returned by orc_mcc_table_size. */
int32_t orc_mcc_table_size(int32_t n_seg, const double *seg_value, const double *seg_step);
int orc_mcc_build(orc_mcc *t, int32_t n_proc, const int32_t *type, const int32_t *npts,
                  const double *E_eV, const double *sigma, double N_m3, int32_t N_subcycles, double dt_s);
/* energy -> (indx_energy, a0, a1), process index (1-based; n_proc+1 = null collision)
 * (pic2d_enCollisionsGeneralProc.f90:449-485) */
int orc_mcc_select(const orc_mcc *t, double energy_eV, double random_r, int32_t *indx_energy,
                   double *a0, double *a1, int32_t *indx_coll);
/* kinematics (pic2d_enCollisionsSpecificProc.f90:61-639).  u[] supplies the uniforms in
 * the order the reference draws them.  Returns the number of uniforms consumed. */
int orc_en_scatter(int32_t type, double M_amu, double threshold_eV, double ion_velocity_factor,
                   const double *maxw, double energy_eV, const double *u,
                   double *vx, double *vy, double *vz, int32_t *produced,
                   double *ej /*[3] ejected electron v*/, double *iv /*[3] new ion v*/);

/* per-particle null-collision form fed an explicit uniform stream, the form the GPU runs:
 * uniforms for slot k are u[k*11 .. k*11+10]: [0] candidacy (< max_colliding_fraction),
 * [1] random_r, [2] random_n, [3..10] kinematics.  kind[k] = 0 none, else 1-based process.
 * New particles are appended to the *_new arrays (X,Y,VX,VY,VZ interleaved by 5). */
int orc_mcc_en_stream(const orc_mcc *t, const double *maxw, int64_t n,
                      const double *X, const double *Y, double *VX, double *VY, double *VZ,
                      const double *u, int32_t *kind, int64_t *counters,
                      int64_t cap_new, int64_t *n_new, double *e_new, double *i_new);
/* reference-faithful sampled form: P_max*N picks without replacement from one WELL stream
 * (pic2d_enCollisionsGeneralProc.f90:332-654). */
int orc_mcc_en_reference(const orc_mcc *t, const double *maxw, orc_well *w, int64_t n,
                         const double *X, const double *Y, double *VX, double *VY, double *VZ,
                         int32_t *kind, int64_t *counters,
                         int64_t cap_new, int64_t *n_new, double *e_new, double *i_new);

/* ---------------- resonant charge exchange ------------------------------- */
typedef struct orc_rcx {
    double sigma_1eV, alpha, N_m3, dt_s;
    int32_t N_subcycles;
    double Ms;             /* ion mass / m_e */
    double energy_factor;  /* 0.5*m_e*V_scale^2/e */
    double V_scale;
    double vfactor;        /* Maxwell sample -> normalised velocity */
    double probab_thermal;
} orc_rcx;
/* INITIATE_ION_NEUTRAL_COLLISIONS + calculate_thermal_cx_probab
 * (pic2d_inCollisionsGeneralProc.f90:3-90, pic2d_inCollisionsSpecificProc.f90:4-44) */
void orc_rcx_setup(orc_rcx *r, double sigma_1eV, double alpha, double N_m3, double T_K,
                   double dt_s, int32_t N_subcycles, double Ms, double energy_factor,
                   double V_scale, double T_e_eV, int32_t N_max_vel);
/* PERFORM_RESONANT_CHARGE_EXCHANGE (pic2d_inCollisionsSpecificProc.f90:48-126), per-ion uniforms
 * u[k*4..k*4+3]: [0] test, [1..3] Maxwell draws */
int64_t orc_rcx_stream(const orc_rcx *r, const double *maxw, int64_t n, double *VX, double *VY, double *VZ,
                       const double *u, uint8_t *flag);
int64_t orc_rcx_reference(const orc_rcx *r, const double *maxw, orc_well *w, int64_t n,
                          double *VX, double *VY, double *VZ);

/* ---------------- moments (COLLECT_ELECTRON_MOMENTS, pic2d_ElectronMoments.f90:3-424) --------- */
/* 13 FP32 accumulators per node, raw sums before normalisation: N, VX, VY, VZ, WX, WY, WZ,
 * VXVY, VXVZ, VYVZ, QX, QY, QZ (node arrays back to back). */
void orc_moments_raw(const orc_geom *g, int64_t n, const double *X, const double *Y,
                     const double *VX, const double *VY, const double *VZ, float *acc);

/* ---------------- initial loading (DISTRIBUTE_PARTICLES) ------------------ */
/* pic2d_CurProblemValues.f90:3127-3324, uniform in [xl, xr] x [yb, yt] and Maxwellian
 * velocities scaled by factor_convert; uses the WELL stream in the reference's draw order. */
void orc_distribute(orc_well *w, const double *maxw, int64_t n, double xl, double xr, double yb, double yt,
                    double xmin, double xmax, double ymin, double ymax, double factor_convert,
                    double *X, double *Y, double *VX, double *VY, double *VZ);
int64_t orc_remove_inside_inner(const orc_geom *g, int64_t n, double *X, double *Y, double *VX, double *VY,
                                double *VZ);
void orc_distribute_ions(orc_well *w, const double *maxw, int64_t n, const double *eX, const double *eY,
                         double factor_convert, double *X, double *Y, double *VX, double *VY, double *VZ);

/* ---------------- wall emission ------------------------------------------------------------ */
/* Electron-induced emission parameters of one boundary object (whole_object fields after the
 * conversions of pic2d_Materials.f90:180-245; energies dimensionless).  Index 0/1/2 of the
 * model-1 arrays: elastic, inelastic, true secondaries. */
typedef struct orc_see {
    int32_t model[3];
    int32_t refl_type;
    double setD[3], minE[3], maxE[3];
    double E_elast_0, E_elast_max, maxD_elast, dE_elast, frac_elast_high, frac_inelastic;
    double E_see_0, E_see_max, maxD_classic, k_smooth;
    double vinj_true;
    double lowest;
} orc_see;
#define ORC_ERR_SEE 5
/* Coeff_SEE_Elastic / _Inelastic / _True / _Classic (pic2d_SEEmission.f90:1054-1208):
 * which = 0, 1, 2, 3; model 1 outside its window gives 0 */
double orc_see_coeff(const orc_see *o, int32_t which, double energy, double theta);
/* PROCESS_ELECTRON_COLL_WITH_BOUNDARY_{LEFT,ABOVE,RIGHT,BELOW} clamp + PROCESS_ELECTRON_INDUCED_
 * ELECTRON_EMISSION and its INJECT_* routines for a boundary wall (pic2d_SEEmission.f90:4-1050).
 * dirflag 1 left, 2 above, 3 right, 4 below; (x, y) the absorbed electron's position, the
 * uniforms u[0..nu-1] taken in the reference's order (beyond nu: 0.5).  Emitted electrons go
 * to out[5*m..5*m+4] (x, y, vx, vy, vz), *n_out of them. */
int orc_see_emit(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag, double xmax,
                 double ymax, double x, double y, double vx, double vy, double vz, const double *u, int32_t nu,
                 int32_t max_out, double *out, int32_t *n_out);
/* the same for an electron that hit inner box box[4] = {ileft, iright, jbottom, jtop} at the
 * crossing (x, y), dirflag as orc_inner_crossing gives it */
int orc_see_emit_inner(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag,
                       const int32_t box[4], double x, double y, double vx, double vy, double vz, const double *u,
                       int32_t nu, int32_t max_out, double *out, int32_t *n_out);
/* PERFORM_ELECTRON_EMISSION_SETUP (pic2d_Setup.f90:599-1001) for one wall segment [istart,
 * iend] of an object of length L (pic2d_CurProblemValues.f90:918-941) emitting N_emit (a
 * default REAL) macroparticles per step; dirflag as above; periodic: the X-periodic single
 * cluster (FLAT_WALL corners), else SURROUNDED_BY_WALL corners.  Uniform layout: electron t
 * uses u[4t..4t+3] (position, normal velocity, the two parallel velocities in the order the
 * reference draws them); u[4 INT(n)] tests the fractional electron, which then uses
 * u[4 INT(n) + 4 ..]. */
int orc_constant_emission(int32_t dirflag, int32_t periodic, int32_t gmi, int32_t gmj, int32_t istart, int32_t iend,
                          int32_t L, float N_emit, int32_t model, double f_normal, double f_parallel, double v_beam,
                          const double *maxw, const double *maxw_inj, const double *u, int64_t nu, int64_t max_out,
                          double *out, int64_t *n_out);

#ifdef __cplusplus
}
#endif
#endif
