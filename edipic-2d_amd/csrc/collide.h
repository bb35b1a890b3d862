// Device functions shared by the FP64 and mixed-precision particle kernels: wall routing,
// cell lookup and the per-particle collision physics.
#pragma once

#include "kernels.h"

namespace edgpu {

// ---------------------------------------------------------------------------------
// wall hit (PROCESS_ELECTRON_COLL_WITH_BOUNDARY_*, pic2d_ElectronWallCollisions.f90:3-287):
// clamp the coordinate along the wall, first segment of that side containing it.
__device__ inline int wall_object(const Geo &g, int side, double x, double y) {
    double c = (side == EDGPU_SIDE_LEFT || side == EDGPU_SIDE_RIGHT) ? fmin(fmax(y, 0.0), g.ymax)
                                                                      : fmin(fmax(x, 0.0), g.xmax);
    for (int s = 0; s < g.n_seg; ++s) {
        if (g.seg_side[s] != side) continue;
        if (c >= (double)g.seg_start[s] && c <= (double)g.seg_end[s]) return g.seg_obj[s];
    }
    return -1;
}

// Routing of a particle that left the area, single cluster (pic2d_ElectronDynamics.f90:177-506,
// all neighbour ranks negative; corners SURROUNDED_BY_WALL when not X-periodic).
// Returns the side hit (or -1 when the particle stays after a periodic wrap).
__device__ inline int exit_side(const Geo &g, double &x, double y) {
    const double xmax = g.xmax, ymax = g.ymax;
    if (x < 0.0) {
        if (g.periodic_x) {
            x = x + g.Lx;
            if (y < 0.0) return EDGPU_SIDE_BELOW;
            if (y > ymax) return EDGPU_SIDE_ABOVE;
            return -1;
        }
        if (y >= 0.0 + 1.0 && y <= ymax - 1.0) return EDGPU_SIDE_LEFT;
        if (y < 0.0 + 1.0) return ((0.0 - x) < (0.0 - y)) ? EDGPU_SIDE_BELOW : EDGPU_SIDE_LEFT;
        return ((0.0 - x) < (y - ymax)) ? EDGPU_SIDE_ABOVE : EDGPU_SIDE_LEFT;
    }
    if (x > xmax) {
        if (g.periodic_x) {
            x = x - g.Lx;
            if (y < 0.0) return EDGPU_SIDE_BELOW;
            if (y > ymax) return EDGPU_SIDE_ABOVE;
            return -1;
        }
        if (y >= 0.0 + 1.0 && y <= ymax - 1.0) return EDGPU_SIDE_RIGHT;
        if (y < 0.0 + 1.0) return ((x - xmax) < (0.0 - y)) ? EDGPU_SIDE_BELOW : EDGPU_SIDE_RIGHT;
        return ((x - xmax) < (y - ymax)) ? EDGPU_SIDE_ABOVE : EDGPU_SIDE_RIGHT;
    }
    if (y > ymax) return EDGPU_SIDE_ABOVE;
    return EDGPU_SIDE_BELOW;  // y < 0
}

// ---------------------------------------------------------------------------------
// Inner metal objects: the crossing of a step's trajectory with the boxes' surfaces
// (TRY_ELECTRON_COLL_WITH_INNER_OBJECT, FIND_CLOSEST_INTERSECTION_WITH_OBJECT,
// CHECK_INTERSECTION_WITH_*_SEGMENT, pic2d_ElectronWallCollisions.f90:470-557, 648-921).
// Segments m = 1 left, 2 top, 3 right, 4 bottom of box b = {ileft, iright, jbottom, jtop}.

// strictly inside one of the boxes (the open-box test of the pushes and of the initial loading,
// pic2d_ElectronDynamics.f90:161-174, pic2d_CurProblemValues.f90:3245-3258)
__device__ inline bool inside_inner(const Geo &g, double x, double y) {
    for (int m = 0; m < g.n_inner; ++m) {
        if (x <= (double)g.box[m][0]) continue;
        if (x >= (double)g.box[m][1]) continue;
        if (y <= (double)g.box[m][2]) continue;
        if (y >= (double)g.box[m][3]) continue;
        return true;
    }
    return false;
}

// cell [c, c+1) of segment m of box b is covered by a boundary wall part on the same line or
// by another box (pic2d_CurProblemValues.f90:600-680)
__device__ inline bool inner_cell_covered(const Geo &g, int b, int m, int c) {
    const bool vertical = (m == 1 || m == 3);
    const int line = vertical ? (m == 1 ? g.box[b][0] : g.box[b][1]) : (m == 2 ? g.box[b][3] : g.box[b][2]);
    for (int s = 0; s < g.n_seg; ++s) {
        const int side = g.seg_side[s];
        if (vertical) {
            if (side != EDGPU_SIDE_LEFT && side != EDGPU_SIDE_RIGHT) continue;
            if ((side == EDGPU_SIDE_LEFT ? 0 : g.gmi) != line) continue;
        } else {
            if (side != EDGPU_SIDE_BELOW && side != EDGPU_SIDE_ABOVE) continue;
            if ((side == EDGPU_SIDE_BELOW ? 0 : g.gmj) != line) continue;
        }
        if (c >= g.seg_start[s] && c < g.seg_end[s]) return true;
    }
    for (int n = 0; n < g.n_inner; ++n) {
        if (n == b) continue;
        if (vertical) {
            if (line < g.box[n][0] || line > g.box[n][1]) continue;
            if (c >= g.box[n][2] && c < g.box[n][3]) return true;
        } else {
            if (line < g.box[n][2] || line > g.box[n][3]) continue;
            if (c >= g.box[n][0] && c < g.box[n][1]) return true;
        }
    }
    return false;
}

// trajectory (xorg, yorg)-(x, y) against the segment x = xs, y0 <= y <= y1; the horizontal
// test is the same with x and y exchanged
__device__ inline bool cross_vertical(double xorg, double yorg, double x, double y, double xs, double y0, double y1,
                                      double &yc) {
    if (fmax(xorg, x) < xs) return false;
    if (fmin(xorg, x) > xs) return false;
    if (fmax(yorg, y) < y0) return false;
    if (fmin(yorg, y) > y1) return false;
    if (xorg == x) return false;
    const double c = yorg + (y - yorg) * (xs - xorg) / (x - xorg);
    if (!isfinite(c)) return false;
    if (c < y0 || c > y1) return false;
    if (c < fmin(y, yorg) || c > fmax(y, yorg)) return false;
    yc = c;
    return true;
}

// The crossing nearest the origin over the faces of all boxes (ties keep the first found):
// object index, the side a boundary wall facing the same way would have (the reference's
// coll_direction_flag: left face 3 = EDGPU_SIDE_RIGHT, top face 4 = EDGPU_SIDE_BELOW, right
// face 1 = EDGPU_SIDE_LEFT, bottom face 2 = EDGPU_SIDE_ABOVE) and the crossing point.
__device__ inline bool inner_crossing(const Geo &g, double xorg, double yorg, double x, double y, int &obj,
                                      int &side, double &xc, double &yc) {
    int best_b = -1, best_m = -1;
    double best_d = 0.0;
    for (int b = 0; b < g.n_inner; ++b) {
        for (int m = 1; m <= 4; ++m) {
            double cx, cy;
            if (m == 1 || m == 3) {
                cx = (double)(m == 1 ? g.box[b][0] : g.box[b][1]);
                if (!cross_vertical(xorg, yorg, x, y, cx, (double)g.box[b][2], (double)g.box[b][3], cy)) continue;
                const int jc = max(g.box[b][2], min(g.box[b][3] - 1, (int)cy));
                if (inner_cell_covered(g, b, m, jc)) continue;
            } else {
                cy = (double)(m == 2 ? g.box[b][3] : g.box[b][2]);
                if (!cross_vertical(yorg, xorg, y, x, cy, (double)g.box[b][0], (double)g.box[b][1], cx)) continue;
                const int ic = max(g.box[b][0], min(g.box[b][1] - 1, (int)cx));
                if (inner_cell_covered(g, b, m, ic)) continue;
            }
            const double d = (cx - xorg) * (cx - xorg) + (cy - yorg) * (cy - yorg);
            // within a box the nearest face, across boxes the nearest box: both keep the first of
            // equal distances, so one comparison in the global order gives the same choice
            if (best_m == -1 || d < best_d) {
                best_b = b;
                best_m = m;
                best_d = d;
                xc = cx;
                yc = cy;
            }
        }
    }
    if (best_m < 0) return false;
    obj = g.n_boundary_obj + best_b;
    side = best_m == 1 ? EDGPU_SIDE_RIGHT : best_m == 2 ? EDGPU_SIDE_BELOW : best_m == 3 ? EDGPU_SIDE_LEFT
                                                                                       : EDGPU_SIDE_ABOVE;
    return true;
}

__device__ inline void set_error(Counters *c, int code, int64_t k) {
    if (atomicCAS(&c->err, 0, code) == 0) c->err_index = (long long)k;
}

// cell of a position (pic2d_ElectronDynamics.f90:62-66)
__device__ inline bool cell_of(const Geo &g, double x, double y, int &i, int &j) {
    i = (int)x;
    j = (int)y;
    if (x == g.xmax) i = g.gmi - 1;
    if (y == g.ymax) j = g.gmj - 1;
    return !(i < 0 || i > g.gmi - 1 || j < 0 || j > g.gmj - 1);
}

// Per-workgroup collision counts in LDS, added to the global counters once per workgroup:
// one global atomic per collision on a handful of addresses would serialise in L2.
__device__ inline void coll_count_init(unsigned int *c) {
    for (int q = threadIdx.x; q < EDGPU_MAX_PROC; q += blockDim.x) c[q] = 0u;
    __syncthreads();
}

__device__ inline void coll_count_flush(const unsigned int *c, Counters *cnt) {
    __syncthreads();
    for (int q = threadIdx.x; q < EDGPU_MAX_PROC; q += blockDim.x)
        if (c[q]) atomicAdd(&cnt->coll[q], (unsigned long long)c[q]);
}

// Append of list entries from a workgroup: buffered in LDS, one global reservation per
// workgroup; entries beyond the buffer go to the global list directly.
template <int N>
struct BlockList {
    int n;
    unsigned long long base;
    int64_t v[N];
};

template <int N>
__device__ inline void block_list_init(BlockList<N> *l) {
    if (threadIdx.x == 0) l->n = 0;
    __syncthreads();
}

template <int N>
__device__ inline void block_list_push(BlockList<N> *l, int64_t v, unsigned long long *gcount, int64_t *g,
                                       int64_t cap, Counters *cnt) {
    const int r = atomicAdd(&l->n, 1);
    if (r < N) {
        l->v[r] = v;
        return;
    }
    const unsigned long long at = atomicAdd(gcount, 1ull);
    if ((int64_t)at < cap) g[at] = v;
    else set_error(cnt, EDGPU_ERR_CAPACITY, v);
}

template <int N>
__device__ inline void block_list_flush(BlockList<N> *l, unsigned long long *gcount, int64_t *g, int64_t cap,
                                        Counters *cnt) {
    __syncthreads();
    const int m = min(l->n, N);
    if (threadIdx.x == 0) l->base = m ? atomicAdd(gcount, (unsigned long long)m) : 0ull;
    __syncthreads();
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        const int64_t at = (int64_t)l->base + q;
        if (at < cap) g[at] = l->v[q];
        else set_error(cnt, EDGPU_ERR_CAPACITY, l->v[q]);
    }
}

// ---------------------------------------------------------------------------------
// electron-neutral collision of one electron (per-particle null-collision form).
// Selection: pic2d_enCollisionsGeneralProc.f90:449-485; kinematics:
// pic2d_enCollisionsSpecificProc.f90:61-639.
__device__ inline double okhrimovskyy_cos(double energy_eV, double R) {
    double c = 1.0 - 2.0 * R / (1.0 + 8.0 * (energy_eV / 27.21) * (1.0 - R));
    return fmax(fmin(0.999999999999, c), -0.999999999999);
}

__device__ inline void scatter_rotate(double Vx, double Vy, double Vz, double CosKsi, double SinKsi, double CosFi,
                                      double SinFi, double &ox, double &oy, double &oz) {
    double V = sqrt(Vx * Vx + Vy * Vy + Vz * Vz);
    double Vxy = sqrt(Vx * Vx + Vy * Vy);
    if (Vxy > 1.0e-20) {
        double a = Vx / Vxy, b = Vy / Vxy;
        ox = Vx * CosKsi + (SinFi * V * b + CosFi * Vz * a) * SinKsi;
        oy = Vy * CosKsi - (SinFi * V * a - CosFi * Vz * b) * SinKsi;
        oz = Vz * CosKsi - Vxy * CosFi * SinKsi;
    } else {
        ox = fabs(Vz) * SinKsi * CosFi;
        oy = fabs(Vz) * SinKsi * SinFi;
        oz = Vz * CosKsi;
    }
}

__device__ inline double maxwell_sample(const double *v, double u) {
    double R = 9000.0 * u;
    int indx = (int)R;
    if (indx < 9000) return v[indx] + (R - indx) * (v[indx + 1] - v[indx]);
    return v[9000];
}

// Energy (eV) and process of a candidate electron (u0[0] < max_colliding_fraction already
// tested): returns the process index+1, n_proc+1 for a null collision, 0 for no collision
// test (density draw above the normalised density) or a table error.
__device__ inline int en_select(const EnTable &t, const double *u0, double vx, double vy, double vz, int64_t slot,
                                Counters *cnt, double &en) {
    en = 0.0;
    if (u0[2] > 1.0) return 0;  // neutral_density_normalized == 1
    en = (vx * vx + vy * vy + vz * vz) * t.V2 * kMeKg * 0.5 / kECl;
    int idx, seg = 1;
    double a0, a1;
    if (en >= t.seg_value[t.n_seg]) {
        idx = t.n_values - 1;
        a1 = 1.0;
        a0 = 0.0;
    } else if (en < t.seg_value[0]) {
        idx = 0;
        a1 = 0.0;
        a0 = 1.0;
    } else {
        for (int i = t.n_seg - 1; i >= 0; --i)
            if (en >= t.seg_value[i]) {
                seg = i + 1;
                break;
            }
        double r = (double)t.seg_index[seg - 1] + (en - t.seg_value[seg - 1]) / t.seg_step[seg];
        idx = (int)r;
        idx = max(0, min(idx, t.n_values - 1));
        if (en < t.energy[idx] || en > t.energy[idx + 1]) {
            set_error(cnt, EDGPU_ERR_TABLE, slot);
            return 0;
        }
        a1 = (en - t.energy[idx]) / t.seg_step[seg];
        a0 = 1.0 - a1;
    }
    int k;
    for (k = t.n_proc; k >= 1; --k)
        if (u0[1] > (a0 * t.prob[idx * t.n_proc + k - 1] + a1 * t.prob[(idx + 1) * t.n_proc + k - 1])) break;
    return k + 1;
}

// Collision of a candidate electron (u0[0] < max_colliding_fraction already tested):
// returns process index+1 of a real collision, 0 for a null collision; appends ionisation
// products.  Inline: an out-of-line call makes every wave of the kernel write its whole
// scratch frame, called or not.
__device__ inline int en_collide_candidate(const EnTable *tp, const double *u0, const double *maxw,
                                                               const Rng &rng, int64_t slot, double x, double y,
                                                               double &vx, double &vy, double &vz, AddList ae,
                                                               const IonAddLists &ions, Counters *cnt,
                                                               unsigned int *coll_count) {
    const EnTable &t = *tp;
    double en;
    const int coll = en_select(t, u0, vx, vy, vz, slot, cnt, en);
    if (coll < 1 || coll > t.n_proc) return 0;
    double u[12];
    u[0] = u0[0];
    u[1] = u0[1];
    u[2] = u0[2];
    u[3] = u0[3];
    const int p = coll - 1;
    const int type = t.type[p];
    draw4(rng, slot, 1, u + 4);
    draw4(rng, slot, 2, u + 8);
    const double *w = u + 3;
    const double Vx = vx, Vy = vy, Vz = vz;
    if (type == 10 || type == 11) {
        double CosKsi = (type == 10) ? okhrimovskyy_cos(en, w[0]) : fmax(fmin(0.999999999999, 1.0 - 2.0 * w[0]), -0.999999999999);
        double SinKsi = sin(acos(CosKsi));
        double Fi = w[1] * 6.28318530718;
        double ox, oy, oz;
        scatter_rotate(Vx, Vy, Vz, CosKsi, SinKsi, cos(Fi), sin(Fi), ox, oy, oz);
        double alpha = sqrt(1.0 - (0.001097161 / t.M_amu) * (1.0 - CosKsi));
        vx = ox * alpha;
        vy = oy * alpha;
        vz = oz * alpha;
    } else if (type == 20) {
        if (en <= t.thr[p]) return coll;  // selected, no energy transfer (returns before counting)
        double esc = fmax(0.0, en - t.thr[p]);
        double CosKsi = okhrimovskyy_cos(esc, w[0]);
        double SinKsi = sin(acos(CosKsi));
        double Fi = w[1] * 6.28318530718;
        double ox, oy, oz;
        scatter_rotate(Vx, Vy, Vz, CosKsi, SinKsi, cos(Fi), sin(Fi), ox, oy, oz);
        double alpha = sqrt(esc / en);
        vx = ox * alpha;
        vy = oy * alpha;
        vz = oz * alpha;
    } else if (type == 30) {
        if (en <= t.thr[p]) return coll;
        const double B = 10.0;
        double eej = B * tan(w[0] * atan(0.5 * (en - t.thr[p]) / B));
        eej = fmax(0.0, fmin(eej, 0.5 * (en - t.thr[p])));
        double esc = fmax(0.0, en - t.thr[p] - eej);
        double CosKsi_sc = okhrimovskyy_cos(esc, w[1]);
        double SinKsi_sc = sin(acos(CosKsi_sc));
        double Fi_sc = w[2] * 6.28318530718;
        double CosKsi_ej = okhrimovskyy_cos(eej, w[3]);
        double SinKsi_ej = sin(acos(CosKsi_ej));
        double Fi_ej = w[4] * 6.28318530718;
        double ox, oy, oz;
        scatter_rotate(Vx, Vy, Vz, CosKsi_sc, SinKsi_sc, cos(Fi_sc), sin(Fi_sc), ox, oy, oz);
        double alpha = sqrt(esc / en);
        vx = ox * alpha;
        vy = oy * alpha;
        vz = oz * alpha;
        double ex, ey, ez;
        scatter_rotate(Vx, Vy, Vz, CosKsi_ej, SinKsi_ej, cos(Fi_ej), sin(Fi_ej), ex, ey, ez);
        alpha = sqrt(eej / en);
        const AddList &ai = ions.s[t.ion_sp[p]];
        unsigned long long s = atomicAdd(ae.count, 1ull);
        unsigned long long si = atomicAdd(ai.count, 1ull);
        if ((int64_t)s >= ae.cap || (int64_t)si >= ai.cap) {
            set_error(cnt, EDGPU_ERR_CAPACITY, slot);
        } else {
            ae.x[s] = x;
            ae.y[s] = y;
            ae.tag[s] = 0;
            ai.tag[si] = 0;
            ae.vx[s] = ex * alpha;
            ae.vy[s] = ey * alpha;
            ae.vz[s] = ez * alpha;
            const double f = t.ivf[p];
            ai.x[si] = x;
            ai.y[si] = y;
            ai.vx[si] = maxwell_sample(maxw, w[5]) * f;
            ai.vy[si] = maxwell_sample(maxw, w[6]) * f;
            ai.vz[si] = maxwell_sample(maxw, w[7]) * f;
        }
    }
    atomicAdd(&coll_count[p], 1u);
    return coll;
}

// full per-electron test: candidacy draw, then the collision
__device__ inline int en_collide(const EnTable *t, double pmax, const double *maxw, const Rng &rng, int64_t slot, double x,
                                 double y, double &vx, double &vy, double &vz, AddList ae, const IonAddLists &ai, Counters *cnt,
                                 unsigned int *coll_count) {
    double u0[4];
    draw_en(rng, slot, u0);
    if (!(u0[0] < pmax)) return 0;
    return en_collide_candidate(t, u0, maxw, rng, slot, x, y, vx, vy, vz, ae, ai, cnt, coll_count);
}

// True when the exchange test of rcx_collide certainly rejects an ion whose vz^2 is at most
// vz2b, given vxy2 = vx^2 + vy^2 evaluated as rcx_collide does: the same uniform and, since
// vxy2 + vz2b >= vxy2 + vz^2 in floating point too, a bound at least the one rcx_collide
// compares with.  Only for a constant cross-section (alpha == 0).
__device__ inline bool rcx_rejects_any_vz(const RcxPar &r, const Rng &rng, int64_t slot, double vxy2, double vz2b) {
    if (r.alpha != 0.0) return false;
    double u[2];
    draw2(rng, slot, u);
    return u[0] * u[0] > (r.c2 * (vxy2 + vz2b) + r.pth2) * (1.0 + 1e-9);
}

// resonant charge exchange of one ion (pic2d_inCollisionsSpecificProc.f90:95-122)
__device__ inline bool rcx_collide(const RcxPar &r, const double *maxw, const Rng &rng, int64_t slot, double &vx,
                                   double &vy, double &vz) {
    double u[4];
    draw2(rng, slot, u);  // the test uniform; the velocity draws below only on exchange
    double vsq = vx * vx + vy * vy + vz * vz;
    if (r.alpha == 0.0) {
        // constant cross-section: p^2 = c2 |v|^2 + pth2.  A uniform whose square exceeds that
        // by a relative margin far above the few-ulp rounding of the evaluation below is
        // rejected by it too, so most ions skip both square roots; the decision is unchanged.
        if (u[0] * u[0] > (r.c2 * vsq + r.pth2) * (1.0 + 1e-9)) return false;
    }
    double energy = vsq * r.factor_eV;
    double vabs = sqrt(vsq) * r.V_scale;
    // t = 1 exactly when alpha == 0 (the logarithm is finite)
    double t = r.alpha == 0.0 ? 1.0 : 1.0 - r.alpha * log(fmax(energy, 0.001));
    double p = r.prob_factor * vabs * (r.sigma_1eV * (t * t));
    p = 1.0 * sqrt(p * p + r.pth2);
    if (u[0] <= p) {
        draw4(rng, slot, 0, u);  // words 1-3 (words 1-3 of an injected stream)
        vx = maxwell_sample(maxw, u[1]) * r.vfactor;
        vy = maxwell_sample(maxw, u[2]) * r.vfactor;
        vz = maxwell_sample(maxw, u[3]) * r.vfactor;
        return true;
    }
    return false;
}

}  // namespace edgpu
