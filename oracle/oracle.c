/*
 * CPU oracle for the EDIPIC-2D particle hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference routines the GPU library
 * replaces.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it (through oracle/oracle_ffi.py); the GPU library never does.
 *
 * Compiled with gcc -O2 -ffp-contract=off so that every expression rounds in
 * the order written, which is the order of the reference Fortran.
 *
 * Pinning (which independent check constrains each function) -- see
 * tests/test_oracle.py for the checks themselves:
 *   orc_well_*            pinned: bit-exact against the reference's own
 *                         WELL19937a_new.c compiled into oracle/_ref.
 *   orc_maxwell_*         pinned: tables rebuilt from the closed-form CDF
 *                         (erf) to 1e-4, sample moments, KS test.
 *   orc_advance           pinned: E-only closed form v'=v-E, x'=x+v';
 *                         B-only |v| conservation and rotation angle
 *                         2*atan(|alpha|); linear-E exactness of the gather;
 *                         hand-worked wall/periodic routing cases.
 *   orc_deposit/fold      pinned: single-particle closed-form weights,
 *                         total charge = N, hand-computed fold on a 4x4 grid.
 *   orc_real8ft           pinned: against numpy.fft.rfft up to the NR sign
 *                         and packing convention.
 *   orc_prepare_sys_y,
 *   orc_solve_poisson_fftx pinned: discrete eigenfunction test (exact to
 *                         1e-13) and dense-matrix brute force on tiny grids.
 *   orc_efield            pinned: linear potential gives constant field.
 *   orc_krylov_*          pinned: against scipy.sparse.linalg.spsolve.
 *   orc_en_frequency,
 *   orc_mcc_build         pinned: constant sigma closed form sigma*sqrt(2eE/m).
 *   orc_mcc_select        pinned: brute-force enumeration against an
 *                         independent evaluation of the interpolated table.
 *   orc_en_scatter        pinned: kinematic invariants (elastic energy loss
 *                         1-(2m/M)(1-cos chi), inelastic E-E_th, ionisation
 *                         energy balance E_sc+E_ej = E-E_th).
 *   orc_mcc_en_stream,
 *   orc_mcc_en_reference  pinned statistically: event rate = P_max*N*<p>.
 *   orc_rcx_*             pinned: closed-form probability with alpha=0.
 *   orc_moments_raw       pinned: the 13 accumulators of single particles worked by hand
 *                         (node, cell centre, quarter point, upper-edge cells); N moment
 *                         equals the charge deposit.
 *   orc_distribute*       pinned: draw order, clamps and stream advance against a second
 *                         generator of the same seed; shape through the Maxwell sampler.
 *   orc_remove_inside_inner pinned: hand cases (box faces and corners stay, a run of inside
 *                         slots at the end, the last electron moved into the slot and tested
 *                         again) and, on random loads, the kept multiset equals a filter.
 *   orc_see_coeff         pinned: model-1 windows, the classic yield's maximum at E_max
 *                         for normal incidence and its angular factors, model-2 elastic
 *                         continuity and values at hand-computed energies.
 *   orc_see_emit          pinned: specular reflection flips the normal component only,
 *                         random-angle reflection keeps |v| and leaves the wall, the
 *                         inelastic speed v sqrt(u), the process chosen by the first uniform
 *                         at the yield thresholds, the number of true secondaries for yields
 *                         above 1, energy conservation of the true secondaries.
 *   orc_constant_emission pinned: counts INT(n) + Bernoulli at the fractional threshold,
 *                         positions on the wall segment, beam speed and clipping.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* physical constants, pic2d_Modules.f90:161-168 */
static const double e_Cl = 1.602189e-19;
static const double m_e_kg = 9.109534e-31;
static const double kB_JK = 1.38064852e-23;
static const double pi_c = 3.141592653589793;

/* ======================================================================= */
/* WELL19937a, restated from its published definition with the tempering of
 * the maximally-equidistributed variant; the six unrolled cases of
 * src/WELL19937a_new.c:117-204 are the same recurrence on a circular buffer. */
#define W_R 624
#define W_M1 70
#define W_M2 179
#define W_M3 449

void orc_well_seed(orc_well *w, int32_t seed) {
    /* well_random_seed, pic2d_RandomNumberInterface.f90:39-57.  ISHFT(x,30)
     * with a positive shift is a LEFT shift of the 32-bit word. */
    uint32_t p = (uint32_t)seed;
    w->s[0] = p;
    for (int k = 1; k < W_R; k++) {
        p = w->s[k - 1];
        w->s[k] = 1812433253u * (p ^ (p << 30)) + (uint32_t)k;
    }
    w->i = 0;
}

double orc_well_next(orc_well *w) {
    uint32_t *S = w->s;
    int i = w->i;
    uint32_t V0 = S[i];
    uint32_t VM1 = S[(i + W_M1) % W_R];
    uint32_t VM2 = S[(i + W_M2) % W_R];
    uint32_t VM3 = S[(i + W_M3) % W_R];
    uint32_t VRm1 = S[(i + W_R - 1) % W_R];
    uint32_t VRm2 = S[(i + W_R - 2) % W_R];
    uint32_t z0 = (VRm1 & 0x80000000u) | (VRm2 & 0x7fffffffu);
    uint32_t z1 = (V0 ^ (V0 << 25)) ^ (VM1 ^ (VM1 >> 27));
    uint32_t z2 = (VM2 >> 9) ^ (VM3 ^ (VM3 >> 1));
    uint32_t nv1 = z1 ^ z2;
    S[i] = nv1;
    uint32_t nv0 = z0 ^ (z1 ^ (z1 << 9)) ^ (z2 ^ (z2 << 21)) ^ (nv1 ^ (nv1 >> 21));
    int ni = (i + W_R - 1) % W_R;
    S[ni] = nv0;
    w->i = ni;
    uint32_t y = S[ni] ^ (S[(ni + W_M2 + 1) % W_R] & 0x41180000u);
    return (double)y * 0x1.00000001p-32;
}

void orc_well_fill(orc_well *w, int64_t n, double *out) {
    for (int64_t k = 0; k < n; k++) out[k] = orc_well_next(w);
}

/* ======================================================================= */
/* PrepareMaxwellDistribIntegral, pic2d_ElectronWallCollisions.f90:835-929 */
int orc_maxwell_tables(double *v, double *v_inj) {
    const double U_max = 3.0;
    int ok1 = 0, ok2 = 0;
    double *F = (double *)calloc(180004, sizeof(double));
    if (!F) return -1;
    int N = 180000;
    double Vmin = -U_max, Vmax = U_max;
    double dV = (Vmax - Vmin) / N;
    F[0] = 0.0;
    for (int i = 1; i <= N + 3; i++) {
        double t = Vmin + ((double)i - 0.5) * dV;
        F[i] = F[i - 1] + exp(-(t * t));
    }
    double temp = F[N];
    for (int i = 0; i <= N + 3; i++) F[i] = F[i] * (double)ORC_R_MAX / temp;
    v[0] = Vmin;
    int count = 0;
    for (int i = 1; i <= N + 3; i++) {
        if (((int)F[i] - count) == 1) {
            count++;
            v[count] = Vmin + i * dV;
            if (count == ORC_R_MAX) { ok1 = 1; break; }
        }
    }
    N = 90000;
    Vmin = 0.0;
    Vmax = U_max;
    dV = (Vmax - Vmin) / N;
    memset(F, 0, 180004 * sizeof(double));
    for (int i = 1; i <= N + 3; i++) {
        double t = Vmin + ((double)(float)i - 0.5) * dV;
        F[i] = F[i - 1] + exp(-(t * t)) * t;
    }
    temp = F[N];
    for (int i = 1; i <= N + 3; i++) F[i] = F[i] * (double)ORC_R_MAX_INJ / temp;
    v_inj[0] = Vmin;
    count = 0;
    for (int i = 1; i <= N + 3; i++) {
        if (((int)F[i] - count) == 1) {
            count++;
            v_inj[count] = Vmin + i * dV;
            if (count == ORC_R_MAX_INJ) { ok2 = 1; break; }
        }
    }
    free(F);
    return (ok1 && ok2) ? 0 : -2;
}

/* GetMaxwellVelocity, pic2d_ElectronWallCollisions.f90:961-985 */
double orc_maxwell_sample(const double *v, double u) {
    double R = (double)ORC_R_MAX * u;
    int indx = (int)R;
    if (indx < ORC_R_MAX) return v[indx] + (R - indx) * (v[indx + 1] - v[indx]);
    return v[ORC_R_MAX];
}

/* GetInjMaxwellVelocity, pic2d_ElectronWallCollisions.f90:933-957 */
double orc_maxwell_inj_sample(const double *v_inj, double u) {
    double R = (double)ORC_R_MAX_INJ * u;
    int indx = (int)R;
    if (indx < ORC_R_MAX_INJ) return v_inj[indx] + (R - indx) * (v_inj[indx + 1] - v_inj[indx]);
    return v_inj[ORC_R_MAX_INJ];
}

/* ======================================================================= */
/* wall hit: PROCESS_ELECTRON_COLL_WITH_BOUNDARY_{LEFT,RIGHT,BELOW,ABOVE},
 * pic2d_ElectronWallCollisions.f90:3-287 (clamp, then first segment that
 * contains the clamped coordinate). */
static int wall_hit(const orc_geom *g, int side, double x, double y, int32_t *obj) {
    double c;
    if (side == ORC_SIDE_LEFT || side == ORC_SIDE_RIGHT)
        c = fmin(fmax(y, (double)g->iymin), (double)g->iymax);
    else
        c = fmin(fmax(x, (double)g->ixmin), (double)g->ixmax);
    for (int s = 0; s < g->n_seg; s++) {
        if (g->seg_side[s] != side) continue;
        if (c >= (double)g->seg_start[s] && c <= (double)g->seg_end[s]) {
            *obj = g->seg_obj[s];
            return ORC_OK;
        }
    }
    return ORC_ERR_NO_WALL;
}

/* ---- inner metal objects: the crossing of the step's trajectory with the object's surface
 * (TRY_ELECTRON_COLL_WITH_INNER_OBJECT, FIND_CLOSEST_INTERSECTION_WITH_OBJECT,
 * CHECK_INTERSECTION_WITH_VERTICAL/HORIZONTAL_SEGMENT, pic2d_ElectronWallCollisions.f90:470-831) */

/* cell [c, c+1) of segment m (1 left, 2 top, 3 right, 4 bottom) of inner box b is covered when a
 * boundary wall part lies on the same line over it or another inner box overlaps it
 * (pic2d_CurProblemValues.f90:600-680) */
static int inner_cell_covered(const orc_geom *g, int b, int m, int c) {
    const int32_t *B = g->inner_box[b]; /* ileft, iright, jbottom, jtop */
    const int vertical = (m == 1 || m == 3);
    const int line = vertical ? (m == 1 ? B[0] : B[1]) : (m == 2 ? B[3] : B[2]);
    for (int s = 0; s < g->n_seg; s++) {
        int side = g->seg_side[s];
        if (vertical) {
            if (side != ORC_SIDE_LEFT && side != ORC_SIDE_RIGHT) continue;
            if ((side == ORC_SIDE_LEFT ? g->ixmin : g->ixmax) != line) continue;
        } else {
            if (side != ORC_SIDE_BELOW && side != ORC_SIDE_ABOVE) continue;
            if ((side == ORC_SIDE_BELOW ? g->iymin : g->iymax) != line) continue;
        }
        if (c >= g->seg_start[s] && c < g->seg_end[s]) return 1;
    }
    for (int n = 0; n < g->n_inner; n++) {
        if (n == b) continue;
        const int32_t *N = g->inner_box[n];
        if (vertical) {
            if (line < N[0] || line > N[1]) continue;
            if (c >= N[2] && c < N[3]) return 1;
        } else {
            if (line < N[2] || line > N[3]) continue;
            if (c >= N[0] && c < N[1]) return 1;
        }
    }
    return 0;
}

/* CHECK_INTERSECTION_WITH_VERTICAL_SEGMENT (:839-878): 1 and the crossing's y when the
 * trajectory (xorg, yorg)-(x, y) crosses x = xs within [y0, y1] */
static int cross_vertical(double xorg, double yorg, double x, double y, double xs, double y0, double y1, double *yc) {
    if (fmax(xorg, x) < xs) return 0;
    if (fmin(xorg, x) > xs) return 0;
    if (fmax(yorg, y) < y0) return 0;
    if (fmin(yorg, y) > y1) return 0;
    if (xorg == x) return 0;
    double c = yorg + (y - yorg) * (xs - xorg) / (x - xorg);
    if (!isfinite(c)) return 0;
    if (c < y0 || c > y1) return 0;
    if (c < fmin(y, yorg) || c > fmax(y, yorg)) return 0;
    *yc = c;
    return 1;
}

int orc_inner_crossing(const orc_geom *g, double xorg, double yorg, double x, double y, int32_t *obj,
                       int32_t *dirflag, double *xc, double *yc) {
    int best_b = -1, best_m = -1;
    double best_d = 0.0, bx = 0.0, by = 0.0;
    for (int b = 0; b < g->n_inner; b++) {
        const int32_t *B = g->inner_box[b];
        int mcross = -1;
        double dist = 0.0, mx = 0.0, my = 0.0;
        for (int m = 1; m <= 4; m++) {
            double cx, cy;
            if (m == 1 || m == 3) {
                cx = (double)(m == 1 ? B[0] : B[1]);
                if (!cross_vertical(xorg, yorg, x, y, cx, (double)B[2], (double)B[3], &cy)) continue;
                int jc = (int)cy;
                if (jc > B[3] - 1) jc = B[3] - 1;
                if (jc < B[2]) jc = B[2];
                if (inner_cell_covered(g, b, m, jc)) continue;
            } else {
                cy = (double)(m == 2 ? B[3] : B[2]);
                /* CHECK_INTERSECTION_WITH_HORIZONTAL_SEGMENT (:882-921) is the vertical test with x and y swapped */
                if (!cross_vertical(yorg, xorg, y, x, cy, (double)B[0], (double)B[1], &cx)) continue;
                int ic = (int)cx;
                if (ic > B[1] - 1) ic = B[1] - 1;
                if (ic < B[0]) ic = B[0];
                if (inner_cell_covered(g, b, m, ic)) continue;
            }
            double d = (cx - xorg) * (cx - xorg) + (cy - yorg) * (cy - yorg);
            if (mcross == -1 || d < dist) {
                mcross = m;
                dist = d;
                mx = cx;
                my = cy;
            }
        }
        if (mcross < 0) continue;
        if (best_m == -1 || dist < best_d) {
            best_b = b;
            best_m = mcross;
            best_d = dist;
            bx = mx;
            by = my;
        }
    }
    if (best_m < 0) return ORC_ERR_NO_WALL; /* the reference aborts: Error-1 */
    static const int32_t flag_of_segment[5] = {0, 3, 4, 1, 2};
    *obj = g->n_boundary_obj + best_b;
    *dirflag = flag_of_segment[best_m];
    *xc = bx;
    *yc = by;
    return ORC_OK;
}

/* routing of a particle that left [xmin,xmax]x[ymin,ymax] for a single cluster
 * (all neighbour ranks negative after ANALYZE_CORNERS, corners SURROUNDED_BY_WALL
 * when not X-periodic), pic2d_ElectronDynamics.f90:177-506 */
static int route_exit(const orc_geom *g, double *px, double y, int *removed, int32_t *obj, int *side) {
    const double xmin = g->ixmin, xmax = g->ixmax, ymin = g->iymin, ymax = g->iymax;
    double x = *px;
    *removed = 0;
    if (x < xmin) {
        if (g->periodic_x) {
            x = x + g->L_period_x;
            *px = x;
            if (y < ymin) { *removed = 1; return (*side = ORC_SIDE_BELOW, wall_hit(g, ORC_SIDE_BELOW, x, y, obj)); }
            if (y > ymax) { *removed = 1; return (*side = ORC_SIDE_ABOVE, wall_hit(g, ORC_SIDE_ABOVE, x, y, obj)); }
            return ORC_OK;
        }
        *removed = 1;
        if (y >= ymin + 1.0 && y <= ymax - 1.0) return (*side = ORC_SIDE_LEFT, wall_hit(g, ORC_SIDE_LEFT, x, y, obj));
        if (y < ymin + 1.0) {
            if ((xmin - x) < (ymin - y)) return (*side = ORC_SIDE_BELOW, wall_hit(g, ORC_SIDE_BELOW, x, y, obj));
            return (*side = ORC_SIDE_LEFT, wall_hit(g, ORC_SIDE_LEFT, x, y, obj));
        }
        if ((xmin - x) < (y - ymax)) return (*side = ORC_SIDE_ABOVE, wall_hit(g, ORC_SIDE_ABOVE, x, y, obj));
        return (*side = ORC_SIDE_LEFT, wall_hit(g, ORC_SIDE_LEFT, x, y, obj));
    }
    if (x > xmax) {
        if (g->periodic_x) {
            x = x - g->L_period_x;
            *px = x;
            if (y < ymin) { *removed = 1; return (*side = ORC_SIDE_BELOW, wall_hit(g, ORC_SIDE_BELOW, x, y, obj)); }
            if (y > ymax) { *removed = 1; return (*side = ORC_SIDE_ABOVE, wall_hit(g, ORC_SIDE_ABOVE, x, y, obj)); }
            return ORC_OK;
        }
        *removed = 1;
        if (y >= ymin + 1.0 && y <= ymax - 1.0) return (*side = ORC_SIDE_RIGHT, wall_hit(g, ORC_SIDE_RIGHT, x, y, obj));
        if (y < ymin + 1.0) {
            if ((x - xmax) < (ymin - y)) return (*side = ORC_SIDE_BELOW, wall_hit(g, ORC_SIDE_BELOW, x, y, obj));
            return (*side = ORC_SIDE_RIGHT, wall_hit(g, ORC_SIDE_RIGHT, x, y, obj));
        }
        if ((x - xmax) < (y - ymax)) return (*side = ORC_SIDE_ABOVE, wall_hit(g, ORC_SIDE_ABOVE, x, y, obj));
        return (*side = ORC_SIDE_RIGHT, wall_hit(g, ORC_SIDE_RIGHT, x, y, obj));
    }
    if (y > ymax) { *removed = 1; return (*side = ORC_SIDE_ABOVE, wall_hit(g, ORC_SIDE_ABOVE, x, y, obj)); }
    if (y < ymin) { *removed = 1; return (*side = ORC_SIDE_BELOW, wall_hit(g, ORC_SIDE_BELOW, x, y, obj)); }
    return ORC_OK;
}

/* Boris rotation matrix, pic2d_ElectronDynamics.f90:91-112 */
static void boris_K(double ax, double ay, double az, double K[9]) {
    double ax2 = ax * ax, ay2 = ay * ay, az2 = az * az;
    double theta2 = ax2 + ay2 + az2;
    double invtheta = 1.0 / (1.0 + theta2);
    K[0] = (1.0 - theta2 + 2.0 * ax2) * invtheta;
    K[1] = 2.0 * (ax * ay + az) * invtheta;
    K[2] = 2.0 * (ax * az - ay) * invtheta;
    K[3] = 2.0 * (ax * ay - az) * invtheta;
    K[4] = (1.0 - theta2 + 2.0 * ay2) * invtheta;
    K[5] = 2.0 * (ay * az + ax) * invtheta;
    K[6] = 2.0 * (ax * az + ay) * invtheta;
    K[7] = 2.0 * (ay * az - ax) * invtheta;
    K[8] = (1.0 - theta2 + 2.0 * az2) * invtheta;
}

int orc_advance(const orc_geom *g, const orc_species *sp, const double *EX, const double *EY,
                double Bx, double By, double Bz, double Ez, int64_t n,
                double *X, double *Y, double *VX, double *VY, double *VZ,
                uint8_t *removed, int32_t *hit_obj, int64_t *hit_count) {
    return orc_advance_ex(g, sp, EX, EY, Bx, By, Bz, Ez, n, X, Y, VX, VY, VZ, removed, hit_obj, hit_count, NULL, NULL,
                          NULL);
}

/* the wall a removal crossed, as the reference's dirflag (1 left, 2 above, 3 right, 4 below) */
static int32_t dirflag_of_side(int side) {
    return side == ORC_SIDE_LEFT ? 1 : side == ORC_SIDE_ABOVE ? 2 : side == ORC_SIDE_RIGHT ? 3 : 4;
}

int orc_advance_ex(const orc_geom *g, const orc_species *sp, const double *EX, const double *EY,
                   double Bx, double By, double Bz, double Ez, int64_t n,
                   double *X, double *Y, double *VX, double *VY, double *VZ,
                   uint8_t *removed, int32_t *hit_obj, int64_t *hit_count, double *rem_x, double *rem_y,
                   int32_t *rem_dir) {
    const double xmin = g->ixmin, xmax = g->ixmax, ymin = g->iymin, ymax = g->iymax;
    const int nx = g->ixmax - g->ixmin + 1;
    double K[9];
    double EZ;
    if (sp->is_electron) {
        boris_K(-0.5 * Bx, -0.5 * By, -0.5 * Bz, K);
        EZ = Ez;
    } else {
        boris_K(sp->QM2sNsub * Bx, sp->QM2sNsub * By, sp->QM2sNsub * Bz, K);
        EZ = sp->sense_Ez ? Ez * sp->N_subcycles : 0.0;
    }
    for (int64_t k = 0; k < n; k++) {
        removed[k] = 0;
        hit_obj[k] = -1;
        double x = X[k], y = Y[k];
        if (x < xmin || x > xmax || y < ymin || y > ymax) return ORC_ERR_OUT_OF_BOUNDS;
        int i = (int)x, j = (int)y;
        if (x == xmax) i = g->ixmax - 1;
        if (y == ymax) j = g->iymax - 1;
        if (i < g->ixmin || i > g->ixmax - 1 || j < g->iymin || j > g->iymax - 1) return ORC_ERR_INDEX;
        double ax_ip1 = x - (double)i, ax_i = 1.0 - ax_ip1;
        double ay_jp1 = y - (double)j, ay_j = 1.0 - ay_jp1;
        int64_t p = (int64_t)(i - g->ixmin) + (int64_t)(j - g->iymin) * nx;
        double E_X = EX[p] * ax_i * ay_j + EX[p + 1] * ax_ip1 * ay_j + EX[p + nx] * ax_i * ay_jp1 +
                     EX[p + nx + 1] * ax_ip1 * ay_jp1;
        double E_Y = EY[p] * ax_i * ay_j + EY[p + 1] * ax_ip1 * ay_j + EY[p + nx] * ax_i * ay_jp1 +
                     EY[p + nx + 1] * ax_ip1 * ay_jp1;
        double vx = VX[k], vy = VY[k], vz = VZ[k];
        if (sp->is_electron) {
            /* pic2d_ElectronDynamics.f90:116-132 */
            double vxm = vx - 0.5 * E_X, vym = vy - 0.5 * E_Y, vzm = vz - 0.5 * EZ;
            double vxp = K[0] * vxm + K[1] * vym + K[2] * vzm;
            double vyp = K[3] * vxm + K[4] * vym + K[5] * vzm;
            double vzp = K[6] * vxm + K[7] * vym + K[8] * vzm;
            vx = vxp - 0.5 * E_X;
            vy = vyp - 0.5 * E_Y;
            vz = vzp - 0.5 * EZ;
            x = x + vx;
            y = y + vy;
        } else if (sp->sense_B) {
            /* pic2d_IonDynamics.f90:128-142 */
            double q = sp->QM2s;
            double vxm = vx + q * E_X, vym = vy + q * E_Y, vzm = vz + q * EZ;
            double vxp = K[0] * vxm + K[1] * vym + K[2] * vzm;
            double vyp = K[3] * vxm + K[4] * vym + K[5] * vzm;
            double vzp = K[6] * vxm + K[7] * vym + K[8] * vzm;
            vx = vxp + q * E_X;
            vy = vyp + q * E_Y;
            vz = vzp + q * EZ;
            x = x + vx * (double)sp->N_subcycles;
            y = y + vy * (double)sp->N_subcycles;
        } else {
            /* pic2d_IonDynamics.f90:144-150 */
            double q2 = sp->QM2s + sp->QM2s;
            vx = vx + q2 * E_X;
            vy = vy + q2 * E_Y;
            vz = vz + q2 * EZ;
            x = x + vx * (double)sp->N_subcycles;
            y = y + vy * (double)sp->N_subcycles;
        }
        VX[k] = vx;
        VY[k] = vy;
        VZ[k] = vz;
        /* inner metal objects (open box test, pic2d_ElectronDynamics.f90:161-174); the hit is
         * charged to the object whose surface the step's trajectory crosses first, at the
         * crossing (TRY_ELECTRON_COLL_WITH_INNER_OBJECT; ions go back N_subcycles steps,
         * TRY_ION_COLL_WITH_INNER_OBJECT, pic2d_IonWallCollisions.f90:425-426) */
        int hit_inner = -1;
        for (int m = 0; m < g->n_inner; m++) {
            if (x <= (double)g->inner_box[m][0]) continue;
            if (x >= (double)g->inner_box[m][1]) continue;
            if (y <= (double)g->inner_box[m][2]) continue;
            if (y >= (double)g->inner_box[m][3]) continue;
            hit_inner = m;
            break;
        }
        if (hit_inner >= 0) {
            const double back = sp->is_electron ? 1.0 : (double)sp->N_subcycles;
            int32_t o, dir;
            double xc, yc;
            int rc = orc_inner_crossing(g, x - vx * back, y - vy * back, x, y, &o, &dir, &xc, &yc);
            if (rc != ORC_OK) return rc;
            X[k] = x;
            Y[k] = y;
            removed[k] = 1;
            hit_obj[k] = o;
            hit_count[o]++;
            if (rem_x) rem_x[k] = xc;
            if (rem_y) rem_y[k] = yc;
            if (rem_dir) rem_dir[k] = dir;
            continue;
        }
        if (x >= xmin && x <= xmax && y >= ymin && y <= ymax) {
            X[k] = x;
            Y[k] = y;
            continue;
        }
        int rem = 0, side = -1;
        int32_t o = -1;
        int st = route_exit(g, &x, y, &rem, &o, &side);
        if (st != ORC_OK) return st;
        X[k] = x;
        Y[k] = y;
        if (rem) {
            removed[k] = 1;
            hit_obj[k] = o;
            hit_count[o]++;
            if (rem_x) rem_x[k] = x;
            if (rem_y) rem_y[k] = y;
            if (rem_dir) rem_dir[k] = dirflag_of_side(side);
        }
    }
    return ORC_OK;
}

/* ======================================================================= */
int orc_deposit(const orc_geom *g, int32_t Qs, int64_t n, const double *X, const double *Y,
                const uint8_t *skip, double *rho) {
    const double xmax = g->ixmax, ymax = g->iymax;
    const int nx = g->ixmax - g->ixmin + 1;
    for (int64_t k = 0; k < n; k++) {
        if (skip && skip[k]) continue;
        double x = X[k], y = Y[k];
        int i = (int)x, j = (int)y;
        if (x == xmax) i = g->ixmax - 1;
        if (y == ymax) j = g->iymax - 1;
        if (i < g->ixmin || i > g->ixmax - 1 || j < g->iymin || j > g->iymax - 1) return ORC_ERR_INDEX;
        int64_t p = (int64_t)(i - g->ixmin) + (int64_t)(j - g->iymin) * nx;
        double ax_ip1 = x - (double)i, ax_i = 1.0 - ax_ip1;
        double ay_jp1 = y - (double)j, ay_j = 1.0 - ay_jp1;
        double vij = ax_i * ay_j, vip1j = ax_ip1 * ay_j, vijp1 = ax_i * ay_jp1;
        switch (Qs) {
        case 1:
            rho[p] += vij;
            rho[p + 1] += vip1j;
            rho[p + nx] += vijp1;
            rho[p + nx + 1] += 1.0 - vij - vip1j - vijp1;
            break;
        case 2:
            rho[p] += vij + vij;
            rho[p + 1] += vip1j + vip1j;
            rho[p + nx] += vijp1 + vijp1;
            rho[p + nx + 1] += 2.0 - vij - vip1j - vijp1 - vij - vip1j - vijp1;
            break;
        case 3:
            rho[p] += vij + vij + vij;
            rho[p + 1] += vip1j + vip1j + vip1j;
            rho[p + nx] += vijp1 + vijp1 + vijp1;
            rho[p + nx + 1] += 3.0 - vij - vip1j - vijp1 - vij - vip1j - vijp1 - vij - vip1j - vijp1;
            break;
        case -1:
            rho[p] -= vij;
            rho[p + 1] -= vip1j;
            rho[p + nx] -= vijp1;
            rho[p + nx + 1] = rho[p + nx + 1] - 1.0 + vij + vip1j + vijp1;
            break;
        default:
            return ORC_ERR_TABLE;
        }
    }
    return ORC_OK;
}

void orc_fold_walls(const orc_geom *g, double *rho) {
    const int nx = g->ixmax - g->ixmin + 1;
    const int ny = g->iymax - g->iymin + 1;
#define R(i, j) rho[(int64_t)((i) - g->ixmin) + (int64_t)((j) - g->iymin) * nx]
    if (g->periodic_x) {
        for (int j = g->iymin; j <= g->iymax; j++) {
            R(g->ixmin + 1, j) = R(g->ixmin + 1, j) + R(g->ixmax, j);
            R(g->ixmax - 1, j) = R(g->ixmax - 1, j) + R(g->ixmin, j);
            R(g->ixmin, j) = R(g->ixmax - 1, j);
            R(g->ixmax, j) = R(g->ixmin + 1, j);
        }
        for (int i = g->ixmin; i <= g->ixmax; i++) {
            R(i, g->iymax) = 2.0 * R(i, g->iymax);
            R(i, g->iymin) = 2.0 * R(i, g->iymin);
        }
    } else {
        for (int j = g->iymin + 1; j <= g->iymax - 1; j++) {
            R(g->ixmin, j) = 2.0 * R(g->ixmin, j);
            R(g->ixmax, j) = 2.0 * R(g->ixmax, j);
        }
        for (int i = g->ixmin + 1; i <= g->ixmax - 1; i++) {
            R(i, g->iymax) = 2.0 * R(i, g->iymax);
            R(i, g->iymin) = 2.0 * R(i, g->iymin);
        }
        R(g->ixmin, g->iymin) = 4.0 * R(g->ixmin, g->iymin);
        R(g->ixmin, g->iymax) = 4.0 * R(g->ixmin, g->iymax);
        R(g->ixmax, g->iymin) = 4.0 * R(g->ixmax, g->iymin);
        R(g->ixmax, g->iymax) = 4.0 * R(g->ixmax, g->iymax);
    }
    (void)ny;
#undef R
}

/* ======================================================================= */
/* Numerical Recipes four1/realft (published method), as the reference uses them:
 * forward sign +1, trigonometric recurrences, inverse unnormalised. */
static void four1(double *data1, int nn, int isign) {
    double *data = data1 - 1; /* 1-based */
    int n = 2 * nn, j = 1;
    for (int i = 1; i <= n; i += 2) {
        if (j > i) {
            double tr = data[j], ti = data[j + 1];
            data[j] = data[i];
            data[j + 1] = data[i + 1];
            data[i] = tr;
            data[i + 1] = ti;
        }
        int m = nn;
        while (m >= 2 && j > m) {
            j -= m;
            m /= 2;
        }
        j += m;
    }
    int mmax = 2;
    while (n > mmax) {
        int istep = 2 * mmax;
        double theta = 6.28318530717959 / (isign * mmax);
        double s = sin(0.5 * theta);
        double wpr = -2.0 * s * s;
        double wpi = sin(theta);
        double wr = 1.0, wi = 0.0;
        for (int m = 1; m <= mmax; m += 2) {
            for (int i = m; i <= n; i += istep) {
                int jj = i + mmax;
                double tempr = wr * data[jj] - wi * data[jj + 1];
                double tempi = wr * data[jj + 1] + wi * data[jj];
                data[jj] = data[i] - tempr;
                data[jj + 1] = data[i + 1] - tempi;
                data[i] = data[i] + tempr;
                data[i + 1] = data[i + 1] + tempi;
            }
            double wt = wr;
            wr = wr * wpr - wi * wpi + wr;
            wi = wi * wpr + wt * wpi + wi;
        }
        mmax = istep;
    }
}

void orc_real8ft(double *data1, int32_t n, int32_t isign) {
    double *data = data1 - 1;
    double theta = 3.141592653589793 / (double)(n / 2);
    double c1 = 0.5, c2;
    if (isign == 1) {
        c2 = -0.5;
        four1(data1, n / 2, +1);
    } else {
        c2 = 0.5;
        theta = -theta;
    }
    double s = sin(0.5 * theta);
    double wpr = -2.0 * s * s;
    double wpi = sin(theta);
    double wr = 1.0 + wpr, wi = wpi;
    int n2p3 = n + 3;
    for (int i = 2; i <= n / 4; i++) {
        int i1 = 2 * i - 1, i2 = i1 + 1, i3 = n2p3 - i2, i4 = i3 + 1;
        double h1r = c1 * (data[i1] + data[i3]);
        double h1i = c1 * (data[i2] - data[i4]);
        double h2r = -c2 * (data[i2] + data[i4]);
        double h2i = c2 * (data[i1] - data[i3]);
        data[i1] = h1r + wr * h2r - wi * h2i;
        data[i2] = h1i + wr * h2i + wi * h2r;
        data[i3] = h1r - wr * h2r + wi * h2i;
        data[i4] = -h1i + wr * h2i + wi * h2r;
        double wt = wr;
        wr = wr * wpr - wi * wpi + wr;
        wi = wi * wpr + wt * wpi + wi;
    }
    if (isign == 1) {
        double h1r = data[1];
        data[1] = h1r + data[2];
        data[2] = h1r - data[2];
    } else {
        double h1r = data[1];
        data[1] = c1 * (h1r + data[2]);
        data[2] = c1 * (h1r - data[2]);
        four1(data1, n / 2, -1);
    }
}

void orc_prepare_sys_y(int32_t N, int32_t gmj, double *a_eq) {
    orc_prepare_sys_y_walls(N, gmj, 0, 0, 1.0, 1.0, a_eq);
}

/* PREPARE_SYS_Y with either wall metal or dielectric (pic2d_Prepare_FFTX_SYSY.f90:776-897).
 * A dielectric wall row is the half-cell Gauss law eps_d dphi/dy + (phi_1 - phi_0) = ...,
 * whose diagonal term in mode n is -1 - 2 pi n (pi n + eps) / N^2; the elimination starts
 * at row 0 from 1/prim_lower (metal: 0; mode 0 between two dielectrics: 0, the gauge row)
 * and ends at row gmj with 1/(prim_upper - a_{gmj-1}) (metal: 0). */
void orc_prepare_sys_y_walls(int32_t N, int32_t gmj, int32_t diel_below, int32_t diel_above, double eps_below,
                             double eps_above, double *a_eq) {
    const int nm = N / 2 + 1;
    const double N2 = (double)N * (double)N;
    for (int n = 0; n < nm; n++) {
        const double dn = (double)n;
        const double main_diag = -2.0 - 4.0 * pi_c * pi_c * dn * dn / N2;
        const double prim_lower = -1.0 - 2.0 * pi_c * (pi_c * dn + eps_below) * dn / N2;
        const double prim_upper = -1.0 - 2.0 * pi_c * (pi_c * dn + eps_above) * dn / N2;
        double a0 = 0.0;
        if (diel_below && !(diel_above && n == 0)) a0 = 1.0 / prim_lower;
        a_eq[0 * nm + n] = a0;
        for (int j = 1; j <= gmj - 1; j++) a_eq[j * nm + n] = 1.0 / (main_diag - a_eq[(j - 1) * nm + n]);
        a_eq[gmj * nm + n] = diel_above ? 1.0 / (prim_upper - a_eq[(gmj - 1) * nm + n]) : 0.0;
    }
}

/* the potential segment s gives the node at coordinate c along its wall: its object's, or on a
 * vacuum gap phi_start + (phi_end - phi_start) (c - start) / (end - start) with the end values
 * exact (UPDATE_WALL_POTENTIALS, pic2d_WallPotentials.f90:29-52) */
static double seg_value(const orc_geom *g, int s, int c) {
    if (!g->seg_gap[s]) return g->obj_phi[g->seg_obj[s]];
    const double p0 = g->obj_phi[g->seg_gap_from[s]], p1 = g->obj_phi[g->seg_gap_to[s]];
    if (c == g->seg_start[s]) return p0;
    if (c == g->seg_end[s]) return p1;
    return p0 + (p1 - p0) * (double)(c - g->seg_start[s]) / (double)(g->seg_end[s] - g->seg_start[s]);
}

/* the wall row of SOLVE_POISSON_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:100-145): a metal
 * segment's nodes hold its potential, a dielectric segment's nodes half their charge density
 * plus factor_rho times the surface charge */
static void wall_row_values(const orc_geom *g, int side, uint32_t diel_mask, const double *sigma, double factor_rho,
                            double *row) {
    for (int s = 0; s < g->n_seg; s++) {
        if (g->seg_side[s] != side) continue;
        const int o = g->seg_obj[s];
        const int diel = o < 32 && ((diel_mask >> o) & 1u);
        for (int i = g->seg_start[s]; i <= g->seg_end[s]; i++) {
            double *v = &row[i - g->ixmin];
            *v = diel ? 0.5 * *v + factor_rho * sigma[i - g->ixmin] : seg_value(g, s, i);
        }
    }
}

int orc_solve_poisson_fftx(const orc_geom *g, int32_t N_ppc, const double *rho_e, const double *rho_i,
                           const double *a_eq, double *phi) {
    return orc_solve_poisson_fftx_walls(g, N_ppc, rho_e, rho_i, a_eq, 0u, NULL, NULL, 0, phi);
}

int orc_solve_poisson_fftx_walls(const orc_geom *g, int32_t N_ppc, const double *rho_e, const double *rho_i,
                                 const double *a_eq, uint32_t diel_mask, const double *sigma_below,
                                 const double *sigma_above, int32_t two_dielectric, double *phi) {
    const int nx = g->ixmax - g->ixmin + 1;
    const int N = g->ixmax - g->ixmin - 1;
    const int gmj = g->iymax - g->iymin;
    const int nm = N / 2 + 1;
    if (N < 4 || (N & (N - 1))) return ORC_ERR_TABLE;
    double factor_rho = 1.0 / (double)N_ppc;
    double *rho = (double *)malloc(sizeof(double) * (size_t)nx * (gmj + 1));
    double *xb = (double *)malloc(sizeof(double) * (size_t)(N + 2) * (gmj + 1));
    double *yb = (double *)malloc(sizeof(double) * (size_t)(2 * gmj + 2) * nm);
    double *rhs = (double *)malloc(sizeof(double) * (size_t)(2 * gmj + 2));
    if (!rho || !xb || !yb || !rhs) return -1;
    for (int64_t q = 0; q < (int64_t)nx * (gmj + 1); q++) rho[q] = factor_rho * (rho_i[q] - rho_e[q]);
    wall_row_values(g, ORC_SIDE_BELOW, diel_mask, sigma_below, factor_rho, rho);
    wall_row_values(g, ORC_SIDE_ABOVE, diel_mask, sigma_above, factor_rho, rho + (int64_t)gmj * nx);
    for (int j = 0; j <= gmj; j++) {
        double *r = xb + (int64_t)j * (N + 2);
        for (int i = 1; i <= N; i++) r[i] = rho[(int64_t)j * nx + i];
        r[0] = r[N];
        orc_real8ft(r, N, 1);
        r[N] = r[1];
        r[N + 1] = 0.0;
        r[1] = 0.0;
    }
    for (int n = 0; n < nm; n++)
        for (int j = 0; j <= gmj; j++) {
            yb[(int64_t)n * (2 * gmj + 2) + 2 * j] = xb[(int64_t)j * (N + 2) + 2 * n];
            yb[(int64_t)n * (2 * gmj + 2) + 2 * j + 1] = xb[(int64_t)j * (N + 2) + 2 * n + 1];
        }
    /* Thomas sweep per mode, pic2d_ElectricFieldCalc_FFT_X.f90:676-759: a wall row with a zero
     * coefficient (metal, or the gauge row of mode 0 between two dielectrics) keeps its
     * value, a dielectric row takes part in the elimination */
    for (int n = 0; n < nm; n++) {
        double *y = yb + (int64_t)n * (2 * gmj + 2);
        const double a0 = a_eq[n];
        if (two_dielectric && n == 0) {
            rhs[0] = 0.0;
            rhs[1] = 0.0;
        } else if (a0 == 0.0) {
            rhs[0] = y[0];
            rhs[1] = y[1];
        } else {
            rhs[0] = -y[0] * a0;
            rhs[1] = -y[1] * a0;
        }
        for (int j = 1; j <= gmj - 1; j++) {
            double a = a_eq[j * nm + n];
            rhs[2 * j] = (-y[2 * j] - rhs[2 * j - 2]) * a;
            rhs[2 * j + 1] = (-y[2 * j + 1] - rhs[2 * j - 1]) * a;
        }
        const double ag = a_eq[gmj * nm + n];
        if (ag == 0.0) {
            rhs[2 * gmj] = y[2 * gmj];
            rhs[2 * gmj + 1] = y[2 * gmj + 1];
        } else {
            rhs[2 * gmj] = (-y[2 * gmj] - rhs[2 * gmj - 2]) * ag;
            rhs[2 * gmj + 1] = (-y[2 * gmj + 1] - rhs[2 * gmj - 1]) * ag;
        }
        y[2 * gmj] = rhs[2 * gmj];
        y[2 * gmj + 1] = rhs[2 * gmj + 1];
        for (int j = gmj - 1; j >= 0; j--) {
            double a = a_eq[j * nm + n];
            y[2 * j] = rhs[2 * j] - a * y[2 * j + 2];
            y[2 * j + 1] = rhs[2 * j + 1] - a * y[2 * j + 3];
        }
    }
    for (int n = 0; n < nm; n++)
        for (int j = 0; j <= gmj; j++) {
            xb[(int64_t)j * (N + 2) + 2 * n] = yb[(int64_t)n * (2 * gmj + 2) + 2 * j];
            xb[(int64_t)j * (N + 2) + 2 * n + 1] = yb[(int64_t)n * (2 * gmj + 2) + 2 * j + 1];
        }
    double invfactor = 2.0 / (double)N;
    for (int j = 0; j <= gmj; j++) {
        double *r = xb + (int64_t)j * (N + 2);
        r[1] = r[N];
        orc_real8ft(r, N, -1);
        for (int i = 0; i <= N - 1; i++) r[i] = r[i] * invfactor;
        r[N] = r[0];
        r[N + 1] = r[1];
        for (int i = 0; i <= N + 1; i++) phi[(int64_t)j * nx + i] = r[i];
    }
    free(rho);
    free(xb);
    free(yb);
    free(rhs);
    return ORC_OK;
}

void orc_efield(const orc_geom *g, const double *phi, double f1, double f2, double *EX, double *EY,
                double *accEX, double *accEY) {
    const int nx = g->ixmax - g->ixmin + 1;
    const int ny = g->iymax - g->iymin + 1;
#define P(i, j) phi[(int64_t)((i) - g->ixmin) + (int64_t)((j) - g->iymin) * nx]
#define EXr(i, j) EX[(int64_t)((i) - g->ixmin) + (int64_t)((j) - g->iymin) * nx]
#define EYr(i, j) EY[(int64_t)((i) - g->ixmin) + (int64_t)((j) - g->iymin) * nx]
    for (int j = g->iymin + 1; j <= g->iymax - 1; j++)
        for (int i = g->ixmin + 1; i <= g->ixmax - 1; i++) {
            EXr(i, j) = f2 * (P(i - 1, j) - P(i + 1, j));
            EYr(i, j) = f2 * (P(i, j - 1) - P(i, j + 1));
        }
    for (int i = g->ixmin + 1; i <= g->ixmax - 1; i++) {
        int j = g->iymax;
        EXr(i, j) = f2 * (P(i - 1, j) - P(i + 1, j));
        EYr(i, j) = f1 * (P(i, j - 1) - P(i, j));
        j = g->iymin;
        EXr(i, j) = f2 * (P(i - 1, j) - P(i + 1, j));
        EYr(i, j) = f1 * (P(i, j) - P(i, j + 1));
    }
    if (g->periodic_x) {
        /* x-edge nodes receive the periodic copies the inter-cluster halo
         * delivers (pic2d_ElectricFieldCalc_FFT_X.f90:1386-1420) */
        for (int j = g->iymin; j <= g->iymax; j++) {
            EXr(g->ixmin, j) = EXr(g->ixmax - 1, j);
            EYr(g->ixmin, j) = EYr(g->ixmax - 1, j);
            EXr(g->ixmax, j) = EXr(g->ixmin + 1, j);
            EYr(g->ixmax, j) = EYr(g->ixmin + 1, j);
        }
    } else {
        /* walls on the left/right: one-sided differences as at the y walls */
        for (int j = g->iymin + 1; j <= g->iymax - 1; j++) {
            int i = g->ixmin;
            EXr(i, j) = f1 * (P(i, j) - P(i + 1, j));
            EYr(i, j) = f2 * (P(i, j - 1) - P(i, j + 1));
            i = g->ixmax;
            EXr(i, j) = f1 * (P(i - 1, j) - P(i, j));
            EYr(i, j) = f2 * (P(i, j - 1) - P(i, j + 1));
        }
        int cx[2] = {g->ixmin, g->ixmax}, cy[2] = {g->iymin, g->iymax};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                int i = cx[a], j = cy[b];
                EXr(i, j) = (a == 0) ? f1 * (P(i, j) - P(i + 1, j)) : f1 * (P(i - 1, j) - P(i, j));
                EYr(i, j) = (b == 0) ? f1 * (P(i, j) - P(i, j + 1)) : f1 * (P(i, j - 1) - P(i, j));
            }
    }
    if (accEX && accEY)
        for (int64_t q = 0; q < (int64_t)nx * ny; q++) {
            accEX[q] = accEX[q] + EX[q];
            accEY[q] = accEY[q] + EY[q];
        }
#undef P
#undef EXr
#undef EYr
}

/* ======================================================================= */
/* Krylov path: Dirichlet nodes = domain boundary nodes on walls (all four
 * sides when not X-periodic; y walls when X-periodic) and nodes of metal inner
 * boxes (inclusive ranges, FIND_INNER_OBJECT_CONTAINING_POINT position_flag 9,
 * pic2d_PETSc_Solver.F90:1492-1540). */
static int krylov_dirichlet(const orc_geom *g, int i, int j, double *val) {
    int inner = 0;
    for (int m = 0; m < g->n_inner; m++) {
        if (i < g->inner_box[m][0] || i > g->inner_box[m][1] || j < g->inner_box[m][2] || j > g->inner_box[m][3])
            continue;
        /* later objects overwrite earlier ones, pic2d_ElectricFieldCalc_PETSc.F90:180-190 */
        *val = g->obj_phi[g->n_boundary_obj + m];
        inner = 1;
    }
    if (inner) return 1;
    int side = -1;
    if (j == g->iymin) side = ORC_SIDE_BELOW;
    else if (j == g->iymax) side = ORC_SIDE_ABOVE;
    else if (!g->periodic_x && i == g->ixmin) side = ORC_SIDE_LEFT;
    else if (!g->periodic_x && i == g->ixmax) side = ORC_SIDE_RIGHT;
    if (side < 0) return 0;
    *val = 0.0;
    int c = (side == ORC_SIDE_BELOW || side == ORC_SIDE_ABOVE) ? i : j;
    /* first segment of the side that contains the node (EXIT on the first match,
     * pic2d_ElectricFieldCalc_PETSc.F90:66-80) */
    for (int s = 0; s < g->n_seg; s++)
        if (g->seg_side[s] == side && c >= g->seg_start[s] && c <= g->seg_end[s]) {
            *val = seg_value(g, s, c);
            break;
        }
    return 1;
}

/* unknowns: periodic x -> i in [ixmin, ixmax-2] (N columns), else all nodes */
void orc_krylov_rhs(const orc_geom *g, int32_t N_ppc, const double *rho_e, const double *rho_i, double *b) {
    const int nx = g->ixmax - g->ixmin + 1;
    const double factor_rho = -1.0 / (double)N_ppc;
    for (int j = g->iymin; j <= g->iymax; j++)
        for (int i = g->ixmin; i <= g->ixmax; i++) {
            int64_t q = (int64_t)(i - g->ixmin) + (int64_t)(j - g->iymin) * nx;
            double v;
            if (krylov_dirichlet(g, i, j, &v)) b[q] = v;
            else b[q] = factor_rho * (rho_i[q] - rho_e[q]);
        }
}

/* y = A x with Dirichlet rows kept as identity; in periodic x the column
 * ixmax-1 is the image of ixmin and ixmax of ixmin+1 */
void orc_krylov_apply(const orc_geom *g, const double *x, double *y) {
    const int nx = g->ixmax - g->ixmin + 1;
    for (int j = g->iymin; j <= g->iymax; j++)
        for (int i = g->ixmin; i <= g->ixmax; i++) {
            int64_t q = (int64_t)(i - g->ixmin) + (int64_t)(j - g->iymin) * nx;
            double v;
            if (krylov_dirichlet(g, i, j, &v)) { y[q] = x[q]; continue; }
            int il = i - 1, ir = i + 1;
            if (g->periodic_x) {
                if (il < g->ixmin) il = g->ixmax - 2;
                if (ir > g->ixmax - 2) ir = g->ixmin + (ir - (g->ixmax - 1));
            }
            y[q] = x[q - (i - il)] + x[q + (ir - i)] + x[q - nx] + x[q + nx] - 4.0 * x[q];
        }
}

/* CG on the operator restricted to free nodes (Dirichlet values moved to the
 * right-hand side), applied to -A which is symmetric positive definite. */
int orc_krylov_cg(const orc_geom *g, const double *b, double *x, double rtol, int32_t max_iter, double *relres) {
    const int nx = g->ixmax - g->ixmin + 1, ny = g->iymax - g->iymin + 1;
    const int64_t nn = (int64_t)nx * ny;
    double *r = calloc(nn, sizeof(double)), *p = calloc(nn, sizeof(double)), *Ap = calloc(nn, sizeof(double));
    uint8_t *fr = calloc(nn, 1);
    if (!r || !p || !Ap || !fr) return -1;
    for (int j = g->iymin; j <= g->iymax; j++)
        for (int i = g->ixmin; i <= g->ixmax; i++) {
            int64_t q = (int64_t)(i - g->ixmin) + (int64_t)(j - g->iymin) * nx;
            double v;
            int d = krylov_dirichlet(g, i, j, &v);
            int img = g->periodic_x && (i >= g->ixmax - 1);
            fr[q] = (!d && !img);
            if (d) x[q] = b[q];
        }
    /* images of the periodic columns follow their sources */
#define SYNC_IMG(vec)                                                                  \
    if (g->periodic_x)                                                                 \
        for (int j = g->iymin; j <= g->iymax; j++) {                                   \
            int64_t q0 = (int64_t)(j - g->iymin) * nx;                                 \
            vec[q0 + nx - 2] = vec[q0 + 0];                                            \
            vec[q0 + nx - 1] = vec[q0 + 1];                                            \
        }
    SYNC_IMG(x);
    orc_krylov_apply(g, x, Ap);
    double bn = 0.0, rr = 0.0;
    for (int64_t q = 0; q < nn; q++)
        if (fr[q]) {
            r[q] = -(b[q] - Ap[q]);
            p[q] = r[q];
            rr += r[q] * r[q];
            bn += b[q] * b[q];
        } else {
            r[q] = 0.0;
            p[q] = 0.0;
        }
    /* b-norm includes the Dirichlet contribution of neighbours (||A x0 - b|| at x0 = 0) */
    {
        double *z = calloc(nn, sizeof(double)), *Az = calloc(nn, sizeof(double));
        for (int64_t q = 0; q < nn; q++) z[q] = fr[q] ? 0.0 : x[q];
        SYNC_IMG(z);
        orc_krylov_apply(g, z, Az);
        bn = 0.0;
        for (int64_t q = 0; q < nn; q++)
            if (fr[q]) bn += (b[q] - Az[q]) * (b[q] - Az[q]);
        free(z);
        free(Az);
    }
    double bnorm = sqrt(bn);
    if (bnorm == 0.0) bnorm = 1.0;
    int it = 0;
    double res = sqrt(rr) / bnorm;
    while (res > rtol && it < max_iter) {
        SYNC_IMG(p);
        orc_krylov_apply(g, p, Ap);
        double pAp = 0.0;
        for (int64_t q = 0; q < nn; q++)
            if (fr[q]) pAp += p[q] * (-Ap[q] + (0.0));
        /* p is zero on Dirichlet nodes so A p on free nodes is the restricted product */
        double alpha = rr / pAp;
        double rr1 = 0.0;
        for (int64_t q = 0; q < nn; q++)
            if (fr[q]) {
                x[q] += alpha * p[q];
                r[q] -= alpha * (-Ap[q]);
                rr1 += r[q] * r[q];
            }
        double beta = rr1 / rr;
        rr = rr1;
        for (int64_t q = 0; q < nn; q++)
            if (fr[q]) p[q] = r[q] + beta * p[q];
        it++;
        res = sqrt(rr) / bnorm;
    }
    SYNC_IMG(x);
#undef SYNC_IMG
    if (relres) *relres = res;
    free(r);
    free(p);
    free(Ap);
    free(fr);
    return res <= rtol ? it : -it - 1;
}

/* ======================================================================= */
double orc_en_frequency(double energy_eV, int32_t type, int32_t npts, const double *E, const double *S,
                        double N_m3, int32_t N_subcycles, double dt_s) {
    double f;
    if (energy_eV >= E[npts - 1]) {
        f = S[npts - 1] * sqrt(2.0 * energy_eV * e_Cl / m_e_kg);
    } else if (energy_eV < E[0]) {
        f = (type < 20) ? S[0] * sqrt(2.0 * energy_eV * e_Cl / m_e_kg) : 0.0;
    } else {
        int j;
        for (j = 1; j <= npts - 1; j++)
            if (energy_eV >= E[j - 1] && energy_eV < E[j]) break;
        if (j > npts - 1) j = npts - 1;
        double ej = E[j - 1], ej1 = E[j];
        double fj = S[j - 1] * sqrt(2.0 * ej * e_Cl / m_e_kg);
        double fj1 = S[j] * sqrt(2.0 * ej1 * e_Cl / m_e_kg);
        f = fj + (fj1 - fj) * (energy_eV - ej) / (ej1 - ej);
    }
    return f * N_m3 * (double)N_subcycles * dt_s;
}

int32_t orc_mcc_table_size(int32_t n_seg, const double *seg_value, const double *seg_step) {
    int32_t nv = 0;
    for (int i = 1; i <= n_seg; i++) nv += (int32_t)((seg_value[i] - seg_value[i - 1]) / seg_step[i] + 0.001);
    return nv;
}

int orc_mcc_build(orc_mcc *t, int32_t n_proc, const int32_t *type, const int32_t *npts, const double *E_eV,
                  const double *sigma, double N_m3, int32_t N_subcycles, double dt_s) {
    t->n_proc = n_proc;
    int nv = 0;
    t->seg_index[0] = 0;
    for (int i = 1; i <= t->n_seg; i++) {
        int d = (int)((t->seg_value[i] - t->seg_value[i - 1]) / t->seg_step[i] + 0.001);
        nv += d;
        t->seg_index[i] = t->seg_index[i - 1] + d;
    }
    if (nv != t->seg_index[t->n_seg]) return ORC_ERR_TABLE;
    t->n_values = nv;
    t->energy_eV[0] = t->seg_value[0];
    for (int i = 1; i <= t->n_seg; i++) {
        for (int j = t->seg_index[i - 1] + 1; j <= t->seg_index[i] - 1; j++)
            t->energy_eV[j] = t->seg_value[i - 1] + (double)(j - t->seg_index[i - 1]) * t->seg_step[i];
        t->energy_eV[t->seg_index[i]] = t->seg_value[i];
    }
    const double *Ep[ORC_MAX_PROC], *Sp[ORC_MAX_PROC];
    int off = 0;
    for (int p = 0; p < n_proc; p++) {
        Ep[p] = E_eV + off;
        Sp[p] = sigma + off;
        off += npts[p];
        t->type[p] = type[p];
        t->threshold_eV[p] = (type[p] >= 20) ? Ep[p][0] : 0.0;
    }
    double *P = t->prob;
    for (int e = 0; e <= nv; e++) {
        double en = t->energy_eV[e];
        P[e * n_proc + 0] = orc_en_frequency(en, type[0], npts[0], Ep[0], Sp[0], N_m3, N_subcycles, dt_s);
        for (int p = 1; p < n_proc; p++)
            P[e * n_proc + p] =
                P[e * n_proc + p - 1] + orc_en_frequency(en, type[p], npts[p], Ep[p], Sp[p], N_m3, N_subcycles, dt_s);
    }
    int emax = 0;
    for (int e = 1; e <= nv; e++)
        if (P[e * n_proc + n_proc - 1] > P[emax * n_proc + n_proc - 1]) emax = e;
    double temp = P[emax * n_proc + n_proc - 1];
    t->max_colliding_fraction = 1.0 - exp(-temp);
    for (int e = 0; e <= nv; e++)
        for (int p = 0; p < n_proc; p++) P[e * n_proc + p] = fmax(0.0, fmin(1.0, P[e * n_proc + p] / temp));
    P[emax * n_proc + n_proc - 1] = 1.0;
    for (int e = 0; e <= nv; e++)
        for (int p = 0; p < n_proc - 1; p++) P[e * n_proc + p + 1] = fmax(P[e * n_proc + p], P[e * n_proc + p + 1]);
    return ORC_OK;
}

int orc_mcc_select(const orc_mcc *t, double energy_eV, double random_r, int32_t *indx_energy, double *pa0,
                   double *pa1, int32_t *indx_coll) {
    int idx, seg = 1;
    double a0, a1;
    if (energy_eV >= t->seg_value[t->n_seg]) {
        idx = t->n_values - 1;
        a1 = 1.0;
        a0 = 0.0;
    } else if (energy_eV < t->seg_value[0]) {
        idx = 0;
        a1 = 0.0;
        a0 = 1.0;
    } else {
        for (int i = t->n_seg - 1; i >= 0; i--)
            if (energy_eV >= t->seg_value[i]) {
                seg = i + 1;
                break;
            }
        double r = (double)t->seg_index[seg - 1] + (energy_eV - t->seg_value[seg - 1]) / t->seg_step[seg];
        idx = (int)r;
        if (idx > t->n_values - 1) idx = t->n_values - 1;
        if (idx < 0) idx = 0;
        if (energy_eV < t->energy_eV[idx] || energy_eV > t->energy_eV[idx + 1]) return ORC_ERR_TABLE;
        a1 = (energy_eV - t->energy_eV[idx]) / t->seg_step[seg];
        a0 = 1.0 - a1;
    }
    int k;
    for (k = t->n_proc; k >= 1; k--)
        if (random_r > (a0 * t->prob[idx * t->n_proc + k - 1] + a1 * t->prob[(idx + 1) * t->n_proc + k - 1])) break;
    *indx_energy = idx;
    *pa0 = a0;
    *pa1 = a1;
    *indx_coll = k + 1;
    return ORC_OK;
}

/* shared rotation of pic2d_enCollisionsSpecificProc.f90:114-130 */
static void scatter_rotate(double Vx, double Vy, double Vz, double CosKsi, double SinKsi, double CosFi, double SinFi,
                           double *ox, double *oy, double *oz) {
    double V = sqrt(Vx * Vx + Vy * Vy + Vz * Vz);
    double Vxy = sqrt(Vx * Vx + Vy * Vy);
    if (Vxy > 1.0e-20) {
        double a = Vx / Vxy, b = Vy / Vxy;
        *ox = Vx * CosKsi + (SinFi * V * b + CosFi * Vz * a) * SinKsi;
        *oy = Vy * CosKsi - (SinFi * V * a - CosFi * Vz * b) * SinKsi;
        *oz = Vz * CosKsi - Vxy * CosFi * SinKsi;
    } else {
        *ox = fabs(Vz) * SinKsi * CosFi;
        *oy = fabs(Vz) * SinKsi * SinFi;
        *oz = Vz * CosKsi;
    }
}

static double okhrimovskyy_cos(double energy_eV, double R) {
    double c = 1.0 - 2.0 * R / (1.0 + 8.0 * (energy_eV / 27.21) * (1.0 - R));
    return fmax(fmin(0.999999999999, c), -0.999999999999);
}

int orc_en_scatter(int32_t type, double M_amu, double threshold_eV, double ion_velocity_factor, const double *maxw,
                   double energy_eV, const double *u, double *vx, double *vy, double *vz, int32_t *produced,
                   double *ej, double *iv) {
    *produced = 0;
    double Vx = *vx, Vy = *vy, Vz = *vz;
    if (type == 10 || type == 11) {
        double CosKsi = (type == 10) ? okhrimovskyy_cos(energy_eV, u[0])
                                     : fmax(fmin(0.999999999999, 1.0 - 2.0 * u[0]), -0.999999999999);
        double Ksi = acos(CosKsi);
        double SinKsi = sin(Ksi);
        double Fi = u[1] * 6.28318530718;
        double CosFi = cos(Fi), SinFi = sin(Fi);
        double ox, oy, oz;
        scatter_rotate(Vx, Vy, Vz, CosKsi, SinKsi, CosFi, SinFi, &ox, &oy, &oz);
        double delta_energy = (0.001097161 / M_amu) * (1.0 - CosKsi);
        double alpha = sqrt(1.0 - delta_energy);
        *vx = ox * alpha;
        *vy = oy * alpha;
        *vz = oz * alpha;
        return 2;
    }
    if (type == 20) {
        if (energy_eV <= threshold_eV) return 0;
        double esc = fmax(0.0, energy_eV - threshold_eV);
        double CosKsi = okhrimovskyy_cos(esc, u[0]);
        double Ksi = acos(CosKsi);
        double SinKsi = sin(Ksi);
        double Fi = u[1] * 6.28318530718;
        double CosFi = cos(Fi), SinFi = sin(Fi);
        double ox, oy, oz;
        scatter_rotate(Vx, Vy, Vz, CosKsi, SinKsi, CosFi, SinFi, &ox, &oy, &oz);
        double alpha = sqrt(esc / energy_eV);
        *vx = ox * alpha;
        *vy = oy * alpha;
        *vz = oz * alpha;
        return 2;
    }
    if (type == 30) {
        const double B = 10.0;
        if (energy_eV <= threshold_eV) return 0;
        double eej = B * tan(u[0] * atan(0.5 * (energy_eV - threshold_eV) / B));
        eej = fmax(0.0, fmin(eej, 0.5 * (energy_eV - threshold_eV)));
        double esc = fmax(0.0, energy_eV - threshold_eV - eej);
        double CosKsi_sc = okhrimovskyy_cos(esc, u[1]);
        double SinKsi_sc = sin(acos(CosKsi_sc));
        double Fi_sc = u[2] * 6.28318530718;
        double CosFi_sc = cos(Fi_sc), SinFi_sc = sin(Fi_sc);
        double CosKsi_ej = okhrimovskyy_cos(eej, u[3]);
        double SinKsi_ej = sin(acos(CosKsi_ej));
        double Fi_ej = u[4] * 6.28318530718;
        double CosFi_ej = cos(Fi_ej), SinFi_ej = sin(Fi_ej);
        double ox, oy, oz;
        scatter_rotate(Vx, Vy, Vz, CosKsi_sc, SinKsi_sc, CosFi_sc, SinFi_sc, &ox, &oy, &oz);
        double alpha = sqrt(esc / energy_eV);
        *vx = ox * alpha;
        *vy = oy * alpha;
        *vz = oz * alpha;
        double ex, ey, ez;
        scatter_rotate(Vx, Vy, Vz, CosKsi_ej, SinKsi_ej, CosFi_ej, SinFi_ej, &ex, &ey, &ez);
        alpha = sqrt(eej / energy_eV);
        ej[0] = ex * alpha;
        ej[1] = ey * alpha;
        ej[2] = ez * alpha;
        iv[0] = orc_maxwell_sample(maxw, u[5]) * ion_velocity_factor;
        iv[1] = orc_maxwell_sample(maxw, u[6]) * ion_velocity_factor;
        iv[2] = orc_maxwell_sample(maxw, u[7]) * ion_velocity_factor;
        *produced = 1;
        return 8;
    }
    return -1;
}

/* energy of pic2d_enCollisionsGeneralProc.f90:449 with V_scale^2 passed in t->energy_factor */
static double en_energy(const orc_mcc *t, double vx, double vy, double vz) {
    return (vx * vx + vy * vy + vz * vz) * t->energy_factor * m_e_kg * 0.5 / e_Cl;
}

int orc_mcc_en_stream(const orc_mcc *t, const double *maxw, int64_t n, const double *X, const double *Y,
                      double *VX, double *VY, double *VZ, const double *u, int32_t *kind, int64_t *counters,
                      int64_t cap_new, int64_t *n_new, double *e_new, double *i_new) {
    for (int64_t k = 0; k < n; k++) {
        const double *uk = u + k * 11;
        kind[k] = 0;
        if (!(uk[0] < t->max_colliding_fraction)) continue;
        if (uk[2] > 1.0) continue; /* neutral_density_normalized == 1 */
        double en = en_energy(t, VX[k], VY[k], VZ[k]);
        int32_t idx, coll;
        double a0, a1;
        if (orc_mcc_select(t, en, uk[1], &idx, &a0, &a1, &coll) != ORC_OK) return ORC_ERR_TABLE;
        if (coll > t->n_proc) continue;
        int p = coll - 1;
        int32_t produced = 0;
        double ej[3], iv[3];
        int used = orc_en_scatter(t->type[p], t->M_amu, t->threshold_eV[p], t->ion_velocity_factor[p], maxw, en,
                                  uk + 3, &VX[k], &VY[k], &VZ[k], &produced, ej, iv);
        if (used < 0) return ORC_ERR_TABLE;
        kind[k] = coll;
        if (used > 0) counters[p]++;
        if (produced) {
            if (*n_new >= cap_new) return ORC_ERR_TABLE;
            double *e = e_new + 5 * (*n_new), *q = i_new + 5 * (*n_new);
            e[0] = X[k]; e[1] = Y[k]; e[2] = ej[0]; e[3] = ej[1]; e[4] = ej[2];
            q[0] = X[k]; q[1] = Y[k]; q[2] = iv[0]; q[3] = iv[1]; q[4] = iv[2];
            (*n_new)++;
        }
    }
    return ORC_OK;
}

int orc_mcc_en_reference(const orc_mcc *t, const double *maxw, orc_well *w, int64_t n, const double *X,
                         const double *Y, double *VX, double *VY, double *VZ, int32_t *kind, int64_t *counters,
                         int64_t cap_new, int64_t *n_new, double *e_new, double *i_new) {
    if (n <= 0) return ORC_OK;
    double Rc = t->max_colliding_fraction * (double)n;
    int64_t Ic = (int64_t)Rc;
    double Fc = Rc - (double)Ic;
    int first = 1;
    for (int64_t k = 0; k < n; k++) kind[k] = 0;
    for (int64_t jj = 0; jj <= Ic; jj++) {
        if (jj == 0 && orc_well_next(w) > Fc) continue;
        double random_r = orc_well_next(w);
        double random_n = orc_well_next(w);
        int64_t rj;
        for (;;) {
            rj = (int64_t)(orc_well_next(w) * (double)n);
            if (rj < 1) rj = 1;
            if (rj > n) rj = n;
            if (first || kind[rj - 1] == 0) break;
        }
        int64_t k = rj - 1;
        if (random_n > 1.0) continue;
        double en = en_energy(t, VX[k], VY[k], VZ[k]);
        int32_t idx, coll;
        double a0, a1;
        if (orc_mcc_select(t, en, random_r, &idx, &a0, &a1, &coll) != ORC_OK) return ORC_ERR_TABLE;
        if (coll > t->n_proc) continue;
        first = 0;
        kind[k] = coll;
        int p = coll - 1;
        /* draw exactly the uniforms the reference process consumes */
        double uu[8];
        int need = 0;
        if (t->type[p] == 10 || t->type[p] == 11) need = 2;
        else if (t->type[p] == 20) need = (en <= t->threshold_eV[p]) ? 0 : 2;
        else if (t->type[p] == 30) need = (en <= t->threshold_eV[p]) ? 0 : 8;
        for (int q = 0; q < need; q++) uu[q] = orc_well_next(w);
        int32_t produced = 0;
        double ej[3], iv[3];
        int used = orc_en_scatter(t->type[p], t->M_amu, t->threshold_eV[p], t->ion_velocity_factor[p], maxw, en, uu,
                                  &VX[k], &VY[k], &VZ[k], &produced, ej, iv);
        if (used > 0) counters[p]++;
        if (produced) {
            if (*n_new >= cap_new) return ORC_ERR_TABLE;
            double *e = e_new + 5 * (*n_new), *q = i_new + 5 * (*n_new);
            e[0] = X[k]; e[1] = Y[k]; e[2] = ej[0]; e[3] = ej[1]; e[4] = ej[2];
            q[0] = X[k]; q[1] = Y[k]; q[2] = iv[0]; q[3] = iv[1]; q[4] = iv[2];
            (*n_new)++;
        }
    }
    return ORC_OK;
}

/* ======================================================================= */
static double sigma_rcx(double energy_eV, double s1, double alpha) {
    double t = 1.0 - alpha * log(fmax(energy_eV, 0.001));
    return s1 * (t * t);
}

void orc_rcx_setup(orc_rcx *r, double sigma_1eV, double alpha, double N_m3, double T_K, double dt_s,
                   int32_t N_subcycles, double Ms, double energy_factor, double V_scale, double T_e_eV,
                   int32_t N_max_vel) {
    r->sigma_1eV = sigma_1eV;
    r->alpha = alpha;
    r->N_m3 = N_m3;
    r->dt_s = dt_s;
    r->N_subcycles = N_subcycles;
    r->Ms = Ms;
    r->energy_factor = energy_factor;
    r->V_scale = V_scale;
    double Tgas_eV = T_K * kB_JK / e_Cl;
    double sth = sigma_rcx(1.5 * Tgas_eV, sigma_1eV, alpha);
    double Vmean = sqrt(2.54647908947033 * Tgas_eV * e_Cl / (Ms * m_e_kg));
    r->probab_thermal = N_m3 * sth * Vmean * dt_s * (double)N_subcycles;
    r->vfactor = sqrt(T_K * kB_JK / (T_e_eV * e_Cl * Ms)) / (double)N_max_vel;
}

static int rcx_test(const orc_rcx *r, double vx, double vy, double vz, double u) {
    double pth2 = r->probab_thermal * r->probab_thermal;
    double factor_eV = r->Ms * r->energy_factor;
    double prob_factor = r->N_m3 * r->dt_s * (double)r->N_subcycles;
    double vsq = vx * vx + vy * vy + vz * vz;
    double energy = vsq * factor_eV;
    double vabs = sqrt(vsq) * r->V_scale;
    double p = prob_factor * vabs * sigma_rcx(energy, r->sigma_1eV, r->alpha);
    p = 1.0 * sqrt(p * p + pth2);
    return u <= p;
}

int64_t orc_rcx_stream(const orc_rcx *r, const double *maxw, int64_t n, double *VX, double *VY, double *VZ,
                       const double *u, uint8_t *flag) {
    int64_t c = 0;
    for (int64_t k = 0; k < n; k++) {
        const double *uk = u + 4 * k;
        flag[k] = 0;
        if (rcx_test(r, VX[k], VY[k], VZ[k], uk[0])) {
            VX[k] = orc_maxwell_sample(maxw, uk[1]) * r->vfactor;
            VY[k] = orc_maxwell_sample(maxw, uk[2]) * r->vfactor;
            VZ[k] = orc_maxwell_sample(maxw, uk[3]) * r->vfactor;
            flag[k] = 1;
            c++;
        }
    }
    return c;
}

int64_t orc_rcx_reference(const orc_rcx *r, const double *maxw, orc_well *w, int64_t n, double *VX, double *VY,
                          double *VZ) {
    int64_t c = 0;
    for (int64_t k = 0; k < n; k++) {
        if (rcx_test(r, VX[k], VY[k], VZ[k], orc_well_next(w))) {
            VX[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * r->vfactor;
            VY[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * r->vfactor;
            VZ[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * r->vfactor;
            c++;
        }
    }
    return c;
}

/* ======================================================================= */
void orc_moments_raw(const orc_geom *g, int64_t n, const double *X, const double *Y, const double *VX,
                     const double *VY, const double *VZ, float *acc) {
    const int nx = g->ixmax - g->ixmin + 1, ny = g->iymax - g->iymin + 1;
    const int64_t nn = (int64_t)nx * ny;
    for (int64_t k = 0; k < n; k++) {
        double x = X[k], y = Y[k];
        int i = (int)x, j = (int)y;
        if (x == (double)g->ixmax) i = g->ixmax - 1;
        if (y == (double)g->iymax) j = g->iymax - 1;
        int64_t p = (int64_t)(i - g->ixmin) + (int64_t)(j - g->iymin) * nx;
        float ax_ip1 = (float)(x - (double)i), ax_i = 1.0f - ax_ip1;
        float ay_jp1 = (float)(y - (double)j), ay_j = 1.0f - ay_jp1;
        float w[4];
        w[0] = ax_i * ay_j;
        w[1] = ax_ip1 * ay_j;
        w[2] = ax_i * ay_jp1;
        w[3] = 1.0f - w[0] - w[1] - w[2];
        int64_t q[4] = {p, p + 1, p + nx, p + nx + 1};
        float vx = (float)VX[k], vy = (float)VY[k], vz = (float)VZ[k];
        float v2 = vx * vx + vy * vy + vz * vz;
        float m[13] = {1.0f, vx, vy, vz, vx * vx, vy * vy, vz * vz, vx * vy, vx * vz, vy * vz, v2 * vx, v2 * vy, v2 * vz};
        for (int a = 0; a < 13; a++)
            for (int c = 0; c < 4; c++) acc[a * nn + q[c]] += (a == 0) ? w[c] : m[a] * w[c];
    }
}

/* ======================================================================= */
void orc_distribute(orc_well *w, const double *maxw, int64_t n, double xl, double xr, double yb, double yt,
                    double xmin, double xmax, double ymin, double ymax, double fc, double *X, double *Y, double *VX,
                    double *VY, double *VZ) {
    for (int64_t k = 0; k < n; k++) {
        X[k] = fmax(xmin, fmin(xmax, xl + (xr - xl) * orc_well_next(w)));
        Y[k] = fmax(ymin, fmin(ymax, yb + (yt - yb) * orc_well_next(w)));
        VX[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * fc;
        VY[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * fc;
        VZ[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * fc;
    }
}

/* The removal of the electrons placed inside inner objects, pic2d_CurProblemValues.f90:3245-3258:
 * slots in order; a slot strictly inside a box takes the last electron (REMOVE_ELECTRON,
 * pic2d_ElectronDynamics.f90:517-553) and is tested again.  Returns the new count. */
int64_t orc_remove_inside_inner(const orc_geom *g, int64_t n, double *X, double *Y, double *VX, double *VY,
                                double *VZ) {
    int64_t k = 0;
    while (k < n) {
        int inside = 0;
        for (int b = 0; b < g->n_inner && !inside; b++) {
            const int32_t *B = g->inner_box[b];
            if (X[k] <= (double)B[0] || X[k] >= (double)B[1]) continue;
            if (Y[k] <= (double)B[2] || Y[k] >= (double)B[3]) continue;
            inside = 1;
        }
        if (!inside) {
            k++;
            continue;
        }
        n--;
        X[k] = X[n];
        Y[k] = Y[n];
        VX[k] = VX[n];
        VY[k] = VY[n];
        VZ[k] = VZ[n];
    }
    return n;
}

void orc_distribute_ions(orc_well *w, const double *maxw, int64_t n, const double *eX, const double *eY, double fc,
                         double *X, double *Y, double *VX, double *VY, double *VZ) {
    for (int64_t k = 0; k < n; k++) {
        X[k] = eX[k];
        Y[k] = eY[k];
        VX[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * fc;
        VY[k] = orc_maxwell_sample(maxw, orc_well_next(w)) * fc;
        VZ[k] = 0.0;
    }
}

/* ======================================================================= */
/* wall emission */

typedef struct {
    const double *u;
    int32_t nu, q;
} uniform_seq;

static double seq_next(uniform_seq *s) {
    double r = s->q < s->nu ? s->u[s->q] : 0.5;
    s->q++;
    return r;
}

/* Coeff_SEE_Classic, pic2d_SEEmission.f90:1171-1208 (the exponent k is a default REAL) */
static double see_classic(const orc_see *o, double energy, double theta) {
    if (energy > o->E_see_0) {
        double energy_max_theta = o->E_see_max * (1.0 + 0.318309886 * o->k_smooth * (theta * theta));
        double v = (energy - o->E_see_0) / (energy_max_theta - o->E_see_0);
        double coeff_max_theta = o->maxD_classic * (1.0 + 0.159154943 * o->k_smooth * (theta * theta));
        float k = v < 1.0 ? 0.62f : 0.25f;
        return coeff_max_theta * pow(v * exp(1.0 - v), (double)k);
    }
    return 0.0;
}

double orc_see_coeff(const orc_see *o, int32_t which, double energy, double theta) {
    switch (which) {
    case 0: /* Coeff_SEE_Elastic :1054-1099 */
        if (o->model[0] == 1) return (energy > o->minE[0] && energy <= o->maxE[0]) ? o->setD[0] : 0.0;
        if (o->model[0] == 2) {
            if (energy >= o->E_elast_max)
                return o->maxD_elast * exp(-(energy - o->E_elast_max) / o->dE_elast) *
                           (energy - o->E_elast_max + o->dE_elast) / o->dE_elast +
                       o->frac_elast_high * see_classic(o, energy, theta);
            if (energy > o->E_elast_0)
                return o->maxD_elast * exp(-(energy - o->E_elast_max) / (o->E_elast_max - o->E_elast_0)) *
                           (energy - o->E_elast_0) / (o->E_elast_max - o->E_elast_0) +
                       o->frac_elast_high * see_classic(o, energy, theta);
        }
        return 0.0;
    case 1: /* Coeff_SEE_Inelastic :1103-1134 */
        if (o->model[1] == 1) return (energy > o->minE[1] && energy <= o->maxE[1]) ? o->setD[1] : 0.0;
        if (o->model[1] == 2) return o->frac_inelastic * see_classic(o, energy, theta);
        return 0.0;
    case 2: /* Coeff_SEE_True :1139-1167 */
        if (o->model[2] == 1) return (energy > o->minE[2] && energy <= o->maxE[2]) ? o->setD[2] : 0.0;
        if (o->model[2] == 2) {
            double c = see_classic(o, energy, theta) * (1.0 - o->frac_elast_high - o->frac_inelastic);
            return c < 0.0 ? 0.0 : c;
        }
        return 0.0;
    default:
        return see_classic(o, energy, theta);
    }
}

static void see_push(double *out, int32_t *n, int32_t max_out, double x, double y, double vx, double vy, double vz) {
    if (*n < max_out) {
        double *p = out + 5 * (*n);
        p[0] = x;
        p[1] = y;
        p[2] = vx;
        p[3] = vy;
        p[4] = vz;
    }
    (*n)++;
}

/* emitted-electron position at a boundary wall (INJECT_*, m > 0 branches) */
static void see_position(int32_t dirflag, double xmax, double ymax, double x, double y, double *xn, double *yn) {
    if (dirflag == 1) {
        *xn = 0.0 + 1.0e-6;
        *yn = y;
    } else if (dirflag == 2) {
        *xn = x;
        *yn = ymax - 1.0e-6;
    } else if (dirflag == 3) {
        *xn = xmax - 1.0e-6;
        *yn = y;
    } else {
        *xn = x;
        *yn = 0.0 + 1.0e-6;
    }
}

/* velocity at angle theta from the wall normal, azimuth fi (INJECT_ELASTIC_REFLECTED_ELECTRON,
 * random branch :218-300; INJECT_INELASTIC_BACKSCATTERED_ELECTRON :597-680) */
static void see_angle(int32_t dirflag, double v, double theta, double fi, double *vx, double *vy, double *vz) {
    if (dirflag == 1) {
        *vx = v * cos(theta);
        *vy = v * sin(theta) * sin(fi);
        *vz = v * sin(theta) * cos(fi);
    } else if (dirflag == 2) {
        *vx = v * sin(theta) * sin(fi);
        *vy = -v * cos(theta);
        *vz = v * sin(theta) * cos(fi);
    } else if (dirflag == 3) {
        *vx = -v * cos(theta);
        *vy = v * sin(theta) * sin(fi);
        *vz = v * sin(theta) * cos(fi);
    } else {
        *vx = v * sin(theta) * sin(fi);
        *vy = v * cos(theta);
        *vz = v * sin(theta) * cos(fi);
    }
}

/* INJECT_TRUE_SECONDARY_ELECTRON, pic2d_SEEmission.f90:794-1050 (boundary walls) */
static void see_true(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag, double xn,
                     double yn, double *energy_inc, uniform_seq *s, double *out, int32_t *n, int32_t max_out) {
    double vx, vy, vz;
    if (*energy_inc <= 0.0) return;
    if (dirflag == 1 || dirflag == 3) {
        vx = orc_maxwell_inj_sample(maxw_inj, seq_next(s));
        vy = orc_maxwell_sample(maxw, seq_next(s));
        vz = orc_maxwell_sample(maxw, seq_next(s));
        vx = (dirflag == 1) ? vx * o->vinj_true : -vx * o->vinj_true;
        vy = vy * o->vinj_true;
        vz = vz * o->vinj_true;
    } else {
        vx = orc_maxwell_sample(maxw, seq_next(s));
        vy = orc_maxwell_inj_sample(maxw_inj, seq_next(s));
        vz = orc_maxwell_sample(maxw, seq_next(s));
        vx = vx * o->vinj_true;
        vy = (dirflag == 4) ? vy * o->vinj_true : -vy * o->vinj_true;
        vz = vz * o->vinj_true;
    }
    double energy_new = vx * vx + vy * vy + vz * vz;
    if (energy_new > *energy_inc) {
        double alpha = sqrt(*energy_inc / energy_new);
        vx = vx * alpha;
        vy = vy * alpha;
        vz = vz * alpha;
        *energy_inc = 0.0;
    } else {
        *energy_inc = *energy_inc - energy_new;
    }
    see_push(out, n, max_out, xn, yn, vx, vy, vz);
}

static int see_emit_at(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag, double xn,
                       double yn, double vx, double vy, double vz, const double *u, int32_t nu, int32_t max_out,
                       double *out, int32_t *n_out);

int orc_see_emit(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag, double xmax,
                 double ymax, double x, double y, double vx, double vy, double vz, const double *u, int32_t nu,
                 int32_t max_out, double *out, int32_t *n_out) {
    /* clamp along the wall, PROCESS_ELECTRON_COLL_WITH_BOUNDARY_*, pic2d_ElectronWallCollisions.f90:24,98,166,238 */
    if (dirflag == 1 || dirflag == 3) {
        if (y < 0.0) y = 0.0;
        if (y > ymax) y = ymax;
    } else {
        if (x < 0.0) x = 0.0;
        if (x > xmax) x = xmax;
    }
    double xn, yn;
    see_position(dirflag, xmax, ymax, x, y, &xn, &yn);
    return see_emit_at(o, maxw, maxw_inj, dirflag, xn, yn, vx, vy, vz, u, nu, max_out, out, n_out);
}

int orc_see_emit_inner(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag,
                       const int32_t box[4], double x, double y, double vx, double vy, double vz, const double *u,
                       int32_t nu, int32_t max_out, double *out, int32_t *n_out) {
    /* the inner-object branches of INJECT_* (pic2d_SEEmission.f90:372-549, 696-790, 830-1050):
     * the new electron leaves the hit face 1e-6 outside the box */
    double xn = x, yn = y;
    if (dirflag == 1) xn = (double)box[1] + 1.0e-6;
    else if (dirflag == 2) yn = (double)box[2] - 1.0e-6;
    else if (dirflag == 3) xn = (double)box[0] - 1.0e-6;
    else yn = (double)box[3] + 1.0e-6;
    return see_emit_at(o, maxw, maxw_inj, dirflag, xn, yn, vx, vy, vz, u, nu, max_out, out, n_out);
}

static int see_emit_at(const orc_see *o, const double *maxw, const double *maxw_inj, int32_t dirflag, double xn,
                       double yn, double vx, double vy, double vz, const double *u, int32_t nu, int32_t max_out,
                       double *out, int32_t *n_out) {
    uniform_seq s = {u, nu, 0};
    *n_out = 0;
    /* PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION, pic2d_SEEmission.f90:44-149 */
    double energy_inc = vx * vx + vy * vy + vz * vz;
    if (energy_inc <= o->lowest) return ORC_OK;
    double v = sqrt(energy_inc);
    double theta_inc;
    if (v == 0.0) {
        theta_inc = 0.0;
    } else {
        double c = fabs((dirflag == 1 || dirflag == 3) ? vx : vy) / v;
        theta_inc = acos(c < 1.0 ? c : 1.0);
    }
    double ce = orc_see_coeff(o, 0, energy_inc, theta_inc);
    double ci = orc_see_coeff(o, 1, energy_inc, theta_inc);
    if (ce + ci > 1.0) return ORC_ERR_SEE;
    double ct = orc_see_coeff(o, 2, energy_inc, theta_inc);
    double total = ce + ci + ct;
    if (total == 0.0) return ORC_OK;
    double R = seq_next(&s);
    if (R <= ce) {
        /* INJECT_ELASTIC_REFLECTED_ELECTRON :155-557 */
        double nx, ny, nz;
        if (o->refl_type == 0) {
            nx = (dirflag == 1 || dirflag == 3) ? -vx : vx;
            ny = (dirflag == 2 || dirflag == 4) ? -vy : vy;
            nz = vz;
        } else {
            double r1 = seq_next(&s);
            double theta = asin(r1 < 1.0 ? r1 : 1.0);
            double fi = 6.283185307179 * seq_next(&s);
            see_angle(dirflag, v, theta, fi, &nx, &ny, &nz);
        }
        see_push(out, n_out, max_out, xn, yn, nx, ny, nz);
        return ORC_OK;
    }
    if (R > ce && R <= ce + ci) {
        /* INJECT_INELASTIC_BACKSCATTERED_ELECTRON :561-790 */
        double r1 = seq_next(&s);
        double theta = asin(r1 < 1.0 ? r1 : 1.0);
        double fi = 6.283185307179 * seq_next(&s);
        double v_new = v * sqrt(seq_next(&s));
        double nx, ny, nz;
        see_angle(dirflag, v_new, theta, fi, &nx, &ny, &nz);
        see_push(out, n_out, max_out, xn, yn, nx, ny, nz);
        return ORC_OK;
    }
    if (ct <= 0.0) return ORC_OK;
    if (o->model[2] == 1) energy_inc = fmax(0.0, energy_inc - o->minE[2]);
    else if (o->model[2] == 2) energy_inc = fmax(0.0, energy_inc - o->E_see_0);
    if (total <= 1.0) {
        if (R > ce + ci && R <= total) see_true(o, maxw, maxw_inj, dirflag, xn, yn, &energy_inc, &s, out, n_out, max_out);
        return ORC_OK;
    }
    ct = ct / (1.0 - (ce + ci));
    while (ct >= 1.0) {
        see_true(o, maxw, maxw_inj, dirflag, xn, yn, &energy_inc, &s, out, n_out, max_out);
        ct = ct - 1.0;
    }
    R = seq_next(&s);
    if (R < ct) see_true(o, maxw, maxw_inj, dirflag, xn, yn, &energy_inc, &s, out, n_out, max_out);
    return ORC_OK;
}

/* one emitted electron of PERFORM_ELECTRON_EMISSION_SETUP from uniforms u[0..3] */
static void emit_one(int32_t dirflag, int32_t flat, int32_t gmi, int32_t gmj, int32_t istart, int32_t iend,
                     int32_t model, double f_normal, double f_parallel, double v_beam, const double *maxw,
                     const double *maxw_inj, const double *u, double *p) {
    double c, vn, hi;
    if (dirflag == 1 || dirflag == 3) {
        /* left/right walls: SURROUNDED_BY_WALL corners in the single cluster (:627-650, :779-800) */
        c = (double)istart + u[0] * (double)(iend - istart);
        hi = (double)gmj;
    } else if (flat) {
        /* FLAT_WALL_BELOW / FLAT_WALL_ABOVE: the overlapping periodic cell is left out */
        int e = iend < gmi - 1 ? iend : gmi - 1;
        c = (double)istart + u[0] * (double)(e - istart);
        hi = (double)(gmi - 1);
    } else {
        c = (double)istart + u[0] * (double)(iend - istart);
        hi = (double)gmi;
    }
    if (c < 0.0) c = 0.0;
    if (c > hi) c = hi;
    if (model == 0) vn = orc_maxwell_inj_sample(maxw_inj, u[1]) * f_normal;
    else vn = fmax(0.0, v_beam + orc_maxwell_sample(maxw, u[1]) * f_normal);
    double a1 = orc_maxwell_sample(maxw, u[2]) * f_parallel;
    double a2 = orc_maxwell_sample(maxw, u[3]) * f_parallel;
    if (dirflag == 1) {
        p[0] = 0.0 + 1.0e-6; p[1] = c; p[2] = vn; p[3] = a1; p[4] = a2;
    } else if (dirflag == 3) {
        p[0] = (double)gmi - 1.0e-6; p[1] = c; p[2] = -vn; p[3] = a1; p[4] = a2;
    } else if (dirflag == 2) {
        p[0] = c; p[1] = (double)gmj - 1.0e-6; p[2] = a1; p[3] = -vn; p[4] = a2;
    } else {
        p[0] = c; p[1] = 0.0 + 1.0e-6; p[2] = a1; p[3] = vn; p[4] = a2;
    }
}

int orc_constant_emission(int32_t dirflag, int32_t periodic, int32_t gmi, int32_t gmj, int32_t istart, int32_t iend,
                          int32_t L, float N_emit, int32_t model, double f_normal, double f_parallel, double v_beam,
                          const double *maxw, const double *maxw_inj, const double *u, int64_t nu, int64_t max_out,
                          double *out, int64_t *n_out) {
    const int32_t flat = periodic && (dirflag == 2 || dirflag == 4);
    int32_t len;
    if (flat) len = (iend < gmi - 1 ? iend : gmi - 1) - istart;
    else len = iend - istart;
    /* add_N_e_to_emit: default REAL product and quotient, pic2d_Setup.f90:633-638 */
    float nf = N_emit * (float)len / (float)L;
    double add = (double)nf;
    int64_t n_int = (int64_t)add, m = 0;
    double uu[4];
    for (int64_t t = 0; t < n_int; t++) {
        for (int q = 0; q < 4; q++) uu[q] = 4 * t + q < nu ? u[4 * t + q] : 0.5;
        if (m < max_out) emit_one(dirflag, flat, gmi, gmj, istart, iend, model, f_normal, f_parallel, v_beam, maxw,
                                  maxw_inj, uu, out + 5 * m);
        m++;
    }
    add = add - (double)n_int;
    double r = 4 * n_int < nu ? u[4 * n_int] : 0.5;
    if (r < add) {
        for (int q = 0; q < 4; q++) uu[q] = 4 * (n_int + 1) + q < nu ? u[4 * (n_int + 1) + q] : 0.5;
        if (m < max_out) emit_one(dirflag, flat, gmi, gmj, istart, iend, model, f_normal, f_parallel, v_beam, maxw,
                                  maxw_inj, uu, out + 5 * m);
        m++;
    }
    *n_out = m;
    return ORC_OK;
}

/* ---- surface charge of dielectric y walls -------------------------------------------------
 * The absorbed particles of one push, given as (object, dirflag, x) records, charge the
 * dielectric walls below (dirflag 4) and above (dirflag 2): x is clamped into the area,
 * ileft = max(INT(x), istart), iright = min(ileft+1, iend) of the first segment of the object on
 * that wall whose node range holds x, and the hit adds q (1 - (x - ileft)) at ileft and
 * q (x - ileft) at iright (pic2d_ElectronWallCollisions.f90:168-195 for q = -1,
 * pic2d_IonWallCollisions.f90:200-219 as Qs (x - ileft) and Qs minus that, pic2d_SEEmission.f90:
 * 226-237 for +1 per emitted electron, where count gives the number emitted at the record). */
void orc_surface_charge(const orc_geom *g, uint32_t diel_mask, int64_t n, const int32_t *obj, const int32_t *dirflag,
                        const double *x, const int32_t *count, int32_t q, double *sigma_below, double *sigma_above) {
    for (int64_t k = 0; k < n; k++) {
        const int o = obj[k];
        if (o < 0 || o >= g->n_boundary_obj || o >= 32 || !((diel_mask >> o) & 1u)) continue;
        int side;
        double *row;
        if (dirflag[k] == 4) {
            side = ORC_SIDE_BELOW;
            row = sigma_below;
        } else if (dirflag[k] == 2) {
            side = ORC_SIDE_ABOVE;
            row = sigma_above;
        } else {
            continue;
        }
        const double c = count ? (double)count[k] : 1.0;
        if (c == 0.0) continue;
        double xx = x[k];
        if (xx < (double)g->ixmin) xx = (double)g->ixmin;
        if (xx > (double)g->ixmax) xx = (double)g->ixmax;
        for (int s = 0; s < g->n_seg; s++) {
            if (g->seg_side[s] != side || g->seg_obj[s] != o) continue;
            if (!(xx >= (double)g->seg_start[s] && xx <= (double)g->seg_end[s])) continue;
            int ileft = (int)xx;
            if (ileft < g->seg_start[s]) ileft = g->seg_start[s];
            int iright = ileft + 1;
            if (iright > g->seg_end[s]) iright = g->seg_end[s];
            double dqleft, dqright;
            if (q == 1 || q == -1) {
                dqright = xx - (double)ileft;
                dqleft = 1.0 - dqright;
                dqright *= (double)q;
                dqleft *= (double)q;
            } else {
                dqright = (double)q * (xx - (double)ileft);
                dqleft = (double)q - dqright;
            }
            row[ileft - g->ixmin] += c * dqleft;
            row[iright - g->ixmin] += c * dqright;
            break;
        }
    }
}

/* GATHER_SURFACE_CHARGE_DENSITY between the two ends of an x-periodic wall
 * (pic2d_IonWallCollisions.f90:717-766 with the neighbour clusters of
 * pic2d_CurProblemValues.f90:1773-1782): the charge at the right end node goes to node 1 of the
 * left end, the charge at the left end node to node gmi-1, and both end nodes are cleared. */
void orc_surface_charge_fold(const orc_geom *g, double *sigma) {
    if (!g->periodic_x) return;
    const int nx = g->ixmax - g->ixmin + 1;
    const double right = sigma[nx - 1], left = sigma[0];
    sigma[nx - 1] = 0.0;
    sigma[0] = 0.0;
    sigma[1] += right;
    sigma[nx - 2] += left;
}

/* ---- ion-induced electron emission --------------------------------------------------------
 * PROCESS_ION_INDUCED_ELECTRON_EMISSION (pic2d_IonEmissions.f90:4-87) of one ion absorbed by a
 * boundary wall (dirflag 1 left, 2 above, 3 right, 4 below): incident energy Ms v^2; outside
 * (minE, maxE) nothing (:38-39, Coeff_ii_EE_True :352-368 gives setD inside); the energy above
 * minE feeds INJECT_ION_IMPACT_EXTRACTED_ELECTRON (:91-348), which leaves the wall like a true
 * secondary at the ion species' temperature (vinj = sqrt(T / T_e) / N_max_vel); floor(setD)
 * electrons, then one more if the next uniform is below the remainder.  The wall coordinate is
 * clamped into the area first (PROCESS_ION_COLL_WITH_BOUNDARY_*). */
int orc_iiee_emit(double setD, double minE, double maxE, double vinj, double Ms, const double *maxw,
                  const double *maxw_inj, int32_t dirflag, double xmax, double ymax, double x, double y, double vx,
                  double vy, double vz, const double *u, int32_t nu, int32_t max_out, double *out, int32_t *n_out) {
    *n_out = 0;
    if (dirflag == 1 || dirflag == 3) {
        if (y < 0.0) y = 0.0;
        if (y > ymax) y = ymax;
    } else {
        if (x < 0.0) x = 0.0;
        if (x > xmax) x = xmax;
    }
    double xn, yn;
    see_position(dirflag, xmax, ymax, x, y, &xn, &yn);
    const double energy = Ms * (vx * vx + vy * vy + vz * vz);
    if (energy <= minE || energy >= maxE) return ORC_OK;
    double coef = setD;
    if (coef == 0.0) return ORC_OK;
    orc_see o;
    memset(&o, 0, sizeof(o));
    o.vinj_true = vinj;
    double energy_inc = energy - minE > 0.0 ? energy - minE : 0.0;
    uniform_seq sq = {u, nu, 0};
    while (coef >= 1.0) {
        see_true(&o, maxw, maxw_inj, dirflag, xn, yn, &energy_inc, &sq, out, n_out, max_out);
        coef = coef - 1.0;
    }
    if (coef != 0.0 && seq_next(&sq) < coef)
        see_true(&o, maxw, maxw_inj, dirflag, xn, yn, &energy_inc, &sq, out, n_out, max_out);
    return ORC_OK;
}

/* INJECT_REFLECTED_ION (pic2d_IonEmissions.f90:374-492) for a boundary wall: the ion comes back
 * 1e-6 off the wall at its clamped coordinate along it, the velocity component normal to the
 * wall reversed.  out = x, y, vx, vy, vz. */
void orc_ion_reflect(int32_t dirflag, double xmax, double ymax, double x, double y, double vx, double vy, double vz,
                     double *out) {
    if (dirflag == 1 || dirflag == 3) {
        if (y < 0.0) y = 0.0;
        if (y > ymax) y = ymax;
    } else {
        if (x < 0.0) x = 0.0;
        if (x > xmax) x = xmax;
    }
    see_position(dirflag, xmax, ymax, x, y, &out[0], &out[1]);
    out[2] = (dirflag == 1 || dirflag == 3) ? -vx : vx;
    out[3] = (dirflag == 2 || dirflag == 4) ? -vy : vy;
    out[4] = vz;
}
