// Kernel argument structs and host launchers of the edgpu HIP sources.
#pragma once

#include "edgpu_internal.h"

namespace edgpu {

struct SortWork;

// 3x3 Boris rotation matrix passed by value
struct Kmat {
    double k[9];
};

// FP64 SoA particle arrays of one species
struct ParticlesF64 {
    double *x, *y, *vx, *vy, *vz;
    int32_t *tag;
    void *base;  // the one allocation holding the arrays
};

// mixed-precision particle arrays: cell = i + j*gmi (cells, not nodes), in-cell offsets
// and velocities in FP32
struct ParticlesMixed {
    int32_t *cell;
    float *dx, *dy, *vx, *vy, *vz;
    int32_t *tag;
    void *base;  // the one allocation holding the arrays
};

// tiles of the last counting sort: particle ranges [start[t], start[t+1]) of kTileW-cell
// row tiles (see tile.h); particles [covered, n) are untiled
struct Tiles {
    const int32_t *start;
    int32_t ntx, ntiles;
    int64_t covered;
};

// moment sums per node (N, VX, VY, VZ, VX^2, VY^2, VZ^2, VXVY, VXVZ, VYVZ, v^2 VX, v^2 VY, v^2 VZ)
constexpr int kMoments = 13;

struct AdvanceArgs {
    Tiles tiles;
    int tail_blocks;
    ParticlesF64 p;
    ParticlesMixed m;
    int64_t n;
    const double *EX, *EY;  // EX/EY for electrons, acc_EX/acc_EY for ions
    const float *EXf, *EYf; // FP32 copies of the gathered field (mixed path)
    double K[9];            // Boris rotation matrix
    double EZ;
    int sense_B;
    double QM2s, Nsub;
    Geo g;
    double *rho;            // deposit target (positions after the push), may be null
    float *mom;             // 13 moment accumulators (nx*ny each) on snapshot steps, else null
    int Qs;
    int64_t *removed_idx;
    int32_t *removed_obj;   // object | (side + 1) << 8 (side -1: inner object)
    double *removed_x, *removed_y;  // position at removal (after a periodic wrap)
    int64_t removed_cap;
    Counters *cnt;
    const EnTable *en;      // device copy of the collision table
    double en_pmax;
    double en_V2;           // V_scale_ms^2 of the collision table
    RcxPar rcx;
    unsigned long long *vz2max;  // bound of vz^2 over the species (ions; see vz_bound)
    const double *maxw;
    Rng rng;
    AddList add_e;
    IonAddLists add_i;  // per ion species
    int64_t *cand;          // slots of e-n collision candidates found by the push
    int64_t cand_cap;
    int64_t *sel;           // candidates whose selection picked a process (capacity cand_cap)
};

// ---- FP64 particles (particles_fp64.hip)
// workgroups of a tiled push launch: one per tile plus tail workgroups for untiled particles
int advance_grid(const AdvanceArgs &a);
// How the push treats vz (chosen on the host by push_vz_mode):
//   kVzReadWrite  vz is read and written;
//   kVzRead       the push cannot change vz (no field along z and a rotation whose z row is
//                 (0, 0, 1)), so it is read for the collision tests but not written;
//   kVzSkip       ions whose vz does not enter vx, vy either and whose exchange test can be
//                 bounded without it: vz is read only by the ions that pass the bounded test.
constexpr int kVzReadWrite = 0, kVzRead = 1, kVzSkip = 2;
void launch_advance_fp64(const AdvanceArgs &a, bool electron, int coll, int vz_mode, hipStream_t s);
// after a push: the object, face and crossing of the particles it left inside inner boxes
void launch_inner_hits_fp64(const AdvanceArgs &a, bool electron, hipStream_t s);
// the sorting push of the electrons, FP64 (see launch_advance_sort_mixed)
void launch_advance_sort_fp64(const AdvanceArgs &a, const ParticlesF64 &out, const SortWork &w, int coll,
                              int vz_mode, hipStream_t s);
void launch_vz_bound_fp64(const ParticlesF64 &p, int64_t begin, int64_t end, unsigned long long *bound,
                          hipStream_t s);
void launch_vz_bound_mixed(const ParticlesMixed &p, int64_t begin, int64_t end, unsigned long long *bound,
                           hipStream_t s);
// collisions of the candidates the fused electron push listed: process selection over the
// candidate list, then the kinematics of the selected slots
void launch_candidates_fp64(const AdvanceArgs &a, hipStream_t s);
void launch_deposit_fp64(const ParticlesF64 &p, int64_t begin, int64_t n, const Geo &g, double *rho, int Qs,
                         Counters *cnt, hipStream_t s);
void launch_mcc_en_fp64(const ParticlesF64 &p, int64_t begin, int64_t n, const EnTable *t, double pmax, const double *maxw,
                        const Rng &rng, const AddList &ae, const IonAddLists &ai, Counters *cnt, hipStream_t s);
void launch_rcx_fp64(const ParticlesF64 &p, int64_t begin, int64_t n, const RcxPar &r, const double *maxw,
                     const Rng &rng, Counters *cnt, unsigned long long *vz2max, hipStream_t s);
void launch_compact_fp64(const ParticlesF64 &p, const int64_t *removed, int64_t nrem, int64_t old_n, int64_t *holes,
                         int64_t *movers, Counters *cnt, hipStream_t s);
void launch_append_fp64(const ParticlesF64 &p, int64_t n0, const AddList &a, int64_t count, hipStream_t s);
void launch_distribute_fp64(const ParticlesF64 &p, int64_t n0, int64_t n, double x0, double x1, double y0, double y1,
                            double xmax, double ymax, double fc, const double *maxw, const Rng &rng, const double *ex,
                            const double *ey, hipStream_t s);
// removal of the loaded particles that lie inside inner objects: mark them and count them in
// cnt->n_removed, then fill the marked slots below new_n = n - count from the slots above it
// (n_holes and n_movers zero on entry)
void launch_mark_inside_fp64(const ParticlesF64 &p, int64_t n, const Geo &g, Counters *cnt, hipStream_t s);
void launch_fill_marked_fp64(const ParticlesF64 &p, int64_t new_n, int64_t old_n, int64_t *holes, int64_t *movers,
                             Counters *cnt, hipStream_t s);

// ---- mixed precision particles (particles_mixed.hip)
void launch_advance_mixed(const AdvanceArgs &a, bool electron, int coll, int vz_mode, hipStream_t s);
void launch_inner_hits_mixed(const AdvanceArgs &a, bool electron, hipStream_t s);
// the sorting push of the electrons (coll 0 or 1; vz_mode kVzRead or kVzReadWrite): a.m is
// read, the pushed particles are written to out in the tile order counted by
// launch_tile_counts_mixed, whose cursors (w.count) it advances and whose window counts
// (w.wcount) it reserves; events and draws are keyed by the slot in out
void launch_advance_sort_mixed(const AdvanceArgs &a, const ParticlesMixed &out, const SortWork &w, int coll,
                               int vz_mode, hipStream_t s);
void launch_candidates_mixed(const AdvanceArgs &a, hipStream_t s);
void launch_deposit_mixed(const ParticlesMixed &p, int64_t begin, int64_t n, const Geo &g, double *rho, int Qs,
                          hipStream_t s);
void launch_mcc_en_mixed(const ParticlesMixed &p, int64_t begin, int64_t n, const Geo &g, const EnTable *t, double pmax,
                         const double *maxw, const Rng &rng, const AddList &ae, const IonAddLists &ai, Counters *cnt,
                         hipStream_t s);
void launch_rcx_mixed(const ParticlesMixed &p, int64_t begin, int64_t n, const RcxPar &r, const double *maxw,
                      const Rng &rng, Counters *cnt, unsigned long long *vz2max, hipStream_t s);
void launch_compact_mixed(const ParticlesMixed &p, const int64_t *removed, int64_t nrem, int64_t old_n,
                          int64_t *holes, int64_t *movers, Counters *cnt, hipStream_t s);
// the same for a mixed-precision chunk converted from the FP64 stage src (tested on src's positions)
void launch_mark_inside_mixed(const ParticlesF64 &src, const ParticlesMixed &dst, int64_t n, const Geo &g,
                              Counters *cnt, hipStream_t s);
void launch_fill_marked_mixed(const ParticlesMixed &p, int64_t new_n, int64_t old_n, int64_t *holes, int64_t *movers,
                              Counters *cnt, hipStream_t s);
void launch_append_mixed(const ParticlesMixed &p, int64_t n0, const Geo &g, const AddList &a, int64_t count,
                         hipStream_t s);
void launch_to_mixed(const ParticlesF64 &src, const ParticlesMixed &dst, int64_t n, const Geo &g, hipStream_t s);
void launch_to_fp64(const ParticlesMixed &src, const ParticlesF64 &dst, int64_t n, const Geo &g, hipStream_t s);
void launch_field_to_f32(const double *a, const double *b, float *fa, float *fb, int64_t n, hipStream_t s);

// ---- per-tile counting sort (sort.hip)
struct SortWork {
    int32_t *key;         // packed tile (row, column) per particle
    int64_t key_cap;      // particles key holds
    int32_t *count;       // per-tile counts (sized ncell+1, ntiles+1 used)
    int32_t *offset;      // exclusive prefix (sized ncell+1, ntiles+1 used)
    void *scan_tmp;
    size_t scan_tmp_bytes;
    int32_t *wcount;      // sorting push: counts per (source tile, destination in its window)
};
size_t sort_scan_bytes(int32_t ncell);
int tiles_in_x(const Geo &g);
// entries of SortWork::wcount for the grid
int64_t sort_window_counts(const Geo &g);
// sort p into tmp using the previous tiles for locality; writes the new tile starts
hipError_t launch_sort_fp64(const ParticlesF64 &p, const ParticlesF64 &tmp, int64_t n, const Geo &g, SortWork &w,
                            const Tiles &old_tiles, int32_t *new_start, hipStream_t s);
hipError_t launch_sort_mixed(const ParticlesMixed &p, const ParticlesMixed &tmp, int64_t n, const Geo &g,
                             SortWork &w, const Tiles &old_tiles, int32_t *new_start, hipStream_t s);
// the counts of a sorting push: per-tile histogram of the particles' current tiles, scanned;
// the cursors (w.count) start at the new tile starts, which are also written to new_start
hipError_t launch_tile_counts_fp64(const ParticlesF64 &p, int64_t n, const Geo &g, SortWork &w,
                                   const Tiles &old_tiles, int32_t *new_start, hipStream_t s);
hipError_t launch_tile_counts_mixed(const ParticlesMixed &p, int64_t n, const Geo &g, SortWork &w,
                                    const Tiles &old_tiles, int32_t *new_start, hipStream_t s);

// ---- grids (field.hip)
void launch_fill(double *a, int64_t n, double v, hipStream_t s);
void launch_copy_probe(const void *a, void *b, int64_t nbytes, hipStream_t s);
void launch_fold_walls(double *rho, const Geo &g, hipStream_t s);
// forward FFT in x of (rho_i - rho_e)/N_ppc with wall rows replaced by the wall potential, or
// on dielectric wall nodes (no potential) by half the row plus the surface charge (sigma rows
// of gmi+1 nodes, NULL without a dielectric on that wall)
void launch_poisson_forward(const double *rho_e, const double *rho_i, double factor_rho, const double *wall_below,
                            const double *wall_above, const double2 *twiddle, double2 *spec, const Geo &g, int N,
                            hipStream_t s, const double *sigma_below = nullptr, const double *sigma_above = nullptr);
void launch_poisson_thomas(double2 *spec, const double *a_eq, int N, int gmj, hipStream_t s, int two_dielectric = 0);
// twiddles (N/2) and Thomas coefficients ((gmj+1) x (N/2+1)) of the FFT_X solve, built on the
// device; diel_below/diel_above: the wall is a dielectric of permittivity eps_below/eps_above
void launch_fftx_tables(double2 *tw, double *a_eq, int N, int gmj, hipStream_t s, int diel_below = 0,
                        int diel_above = 0, double eps_below = 1.0, double eps_above = 1.0);
void launch_poisson_inverse(const double2 *spec, const double2 *twiddle, double *phi, const Geo &g, int N,
                            hipStream_t s);
void launch_efield(const double *phi, double *EX, double *EY, double *accEX, double *accEY, double f1, double f2,
                   const Geo &g, hipStream_t s);

// ---- Krylov (krylov.hip)
// one level of the multigrid preconditioner: columns 0..ncol-1 (the unique columns when
// x-periodic, period ncol), rows 0..ny-1, row stride `stride`; the operator in fine-grid
// units is scale * (4u - sum of the four neighbours) on the free nodes
struct MgLevel {
    int ncol, ny, stride;
    double scale;
    double *u, *u2, *f, *r;  // correction (Jacobi ping-pong), right-hand side, residual
    uint8_t *mask;
};
constexpr int kMgMaxLevels = 12;
// kPrecondMultigridLaunched: the same V-cycle with every sweep of every level launched on
// its own (no k_mg_fused); it gives the same bits and serves as the fused cycle's check
constexpr int kPrecondJacobi = 0, kPrecondMultigrid = 1, kPrecondMultigridLaunched = 2;

// control block of the PCG loop (krylov.hip)
struct KrylovCtl {
    unsigned int ticket;  // workgroups of the running reduction that stored their partials
    int32_t stop;         // set once the residual test passes
    int32_t iters;        // iterations whose update ran
    int32_t pad;
};

struct KrylovWork {
    double *r, *z, *p, *q, *b, *diag;
    uint8_t *mask;       // 1 = free node, 0 = Dirichlet / periodic image
    double *red;         // per-iteration scalars, 4 per iteration
    double *hred;        // page-locked: 8 of them (residual readbacks), then a KrylovCtl
    KrylovCtl *ctl;      // device
    double *part;        // per-workgroup partial sums: kKrylovRedGrid per quantity, 2 quantities,
                         // or one per level-0 tile of the multigrid smoothing (r.z)
    int32_t nred;
    int precond = kPrecondMultigrid;
    // iterations of the previous solve per preconditioner and warm start (the first batch the
    // loop queues before it looks)
    int last_iters[3][3] = {};
    int nlev;            // multigrid levels (level 0 = the grid itself)
    int mg_fused_first;  // levels mg_fused_first..nlev-1 run in k_mg_fused (== nlev: none)
    int mg_fused_lds;    // its dynamic LDS bytes
    MgLevel lev[kMgMaxLevels];
};
// builds the multigrid hierarchy on the mask (needs w.mask); returns levels or -1 on failure
int krylov_setup_multigrid(KrylovWork &w, const Geo &g, hipStream_t s);
// krylov_pcg: iterations to rtol, -1 when max_iter ran out, kKrylovHipError when a HIP call failed
constexpr int kKrylovHipError = -2;
constexpr int kKrylovRedGrid = 4096;  // workgroups of the reducing PCG kernels at most
void launch_krylov_setup_mask(uint8_t *mask, const Geo &g, hipStream_t s);
void launch_krylov_rhs(const double *rho_e, const double *rho_i, double factor_rho, const double *dirichlet_val,
                       const uint8_t *mask, double *b, double *x, const Geo &g, hipStream_t s);
int krylov_pcg(KrylovWork &w, double *x, const Geo &g, double rtol, int max_iter, double *relres, int &batch_hint,
               hipStream_t s);

// the library's last-error message (edgpu_last_error): sets it and returns code
int edgpu_set_error(int code, const char *msg);

// ---- moments (moments.hip)
void launch_moments_fp64(const ParticlesF64 &p, int64_t n, const Geo &g, const double *EX, const double *EY,
                         const double K[9], double EZ, bool electron, float *acc, hipStream_t s);

}  // namespace edgpu
