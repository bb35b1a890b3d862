// Wall emission.
//
// Electron-induced emission (PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION,
// pic2d_SEEmission.f90:4-151) of the electrons the push absorbed at a boundary wall: one
// thread per entry of the removed list, before the compaction overwrites the absorbed slots.
// The yields (Coeff_SEE_Elastic/Inelastic/True/Classic, :1054-1208) and the injections
// (:155-557 elastic, :561-790 inelastic, :794-1050 true secondaries) follow the reference,
// boundary walls only; the uniforms of one absorbed electron are drawn in the reference's
// order from its own counter-based sequence.
//
// Constant emission (PERFORM_ELECTRON_EMISSION_SETUP, pic2d_Setup.f90:599-1001): one thread
// per emitted electron of a wall segment.
//
// Both append to the electron add list with tag = emitting object + 1 (object_id_number).
// Compiled with -ffp-contract=off: expressions round as written, as in the reference.
#include "collide.h"
#include "emission.h"

namespace edgpu {

// Uniforms of one absorbed electron: Philox blocks (slot, step, purpose, block 0, 1, ...) or
// stream[slot*stride + q] in an injected stream.
struct UniformSeq {
    const Rng &rng;
    int64_t slot;
    int q = 0;
    double buf[4];
    __device__ UniformSeq(const Rng &r, int64_t s) : rng(r), slot(s) {}
    __device__ double next() {
        if (rng.stream) {
            const int64_t at = slot * rng.stride + q;
            const double u = (q < rng.stride && at < rng.stream_n) ? rng.stream[at] : 0.5;
            ++q;
            return u;
        }
        if ((q & 3) == 0) draw4(rng, slot, q >> 2, buf);
        return buf[q++ & 3];
    }
};

// GetMaxwellVelocity / GetInjMaxwellVelocity (pic2d_ElectronWallCollisions.f90:933-985)
__device__ inline double maxwell_inj_sample(const double *v, double u) {
    double R = 4500.0 * u;
    int indx = (int)R;
    if (indx < 4500) return v[indx] + (R - indx) * (v[indx + 1] - v[indx]);
    return v[4500];
}

// Coeff_SEE_Classic (pic2d_SEEmission.f90:1171-1208); its exponent k is a default REAL
__device__ inline double coeff_classic(double energy, double theta, const SeeObj &o) {
    if (!(energy > o.E_see_0)) return 0.0;
    const double energy_max_theta = o.E_see_max * (1.0 + 0.318309886 * o.k_smooth * (theta * theta));
    const double v = (energy - o.E_see_0) / (energy_max_theta - o.E_see_0);
    const double coeff_max_theta = o.maxD_classic * (1.0 + 0.159154943 * o.k_smooth * (theta * theta));
    const double k = v < 1.0 ? (double)0.62f : (double)0.25f;
    return coeff_max_theta * pow(v * exp(1.0 - v), k);
}

// Coeff_SEE_Elastic (:1054-1099).  Model 1 outside its energy window: 0 (the reference leaves
// the function value unset there).
__device__ inline double coeff_elastic(double energy, double theta, const SeeObj &o) {
    if (o.model[0] == 1) return (energy > o.minE[0] && energy <= o.maxE[0]) ? o.setD[0] : 0.0;
    if (o.model[0] == 2) {
        if (energy >= o.E_elast_max) {
            return o.maxD_elast * exp(-(energy - o.E_elast_max) / o.dE_elast) *
                       (energy - o.E_elast_max + o.dE_elast) / o.dE_elast +
                   o.frac_elast_high * coeff_classic(energy, theta, o);
        }
        if (energy > o.E_elast_0) {
            return o.maxD_elast * exp(-(energy - o.E_elast_max) / (o.E_elast_max - o.E_elast_0)) *
                       (energy - o.E_elast_0) / (o.E_elast_max - o.E_elast_0) +
                   o.frac_elast_high * coeff_classic(energy, theta, o);
        }
    }
    return 0.0;
}

// Coeff_SEE_Inelastic (:1103-1134)
__device__ inline double coeff_inelastic(double energy, double theta, const SeeObj &o) {
    if (o.model[1] == 1) return (energy > o.minE[1] && energy <= o.maxE[1]) ? o.setD[1] : 0.0;
    if (o.model[1] == 2) return o.frac_inelastic * coeff_classic(energy, theta, o);
    return 0.0;
}

// Coeff_SEE_True (:1139-1167)
__device__ inline double coeff_true(double energy, double theta, const SeeObj &o) {
    if (o.model[2] == 1) return (energy > o.minE[2] && energy <= o.maxE[2]) ? o.setD[2] : 0.0;
    if (o.model[2] == 2) {
        const double c = coeff_classic(energy, theta, o) * (1.0 - o.frac_elast_high - o.frac_inelastic);
        return c < 0.0 ? 0.0 : c;
    }
    return 0.0;
}

__device__ inline void emit_electron(const AddList &ae, Counters *cnt, int obj, int64_t slot, double x, double y,
                                     double vx, double vy, double vz) {
    const unsigned long long s = atomicAdd(ae.count, 1ull);
    if ((int64_t)s >= ae.cap) {
        set_error(cnt, EDGPU_ERR_CAPACITY, slot);
        return;
    }
    ae.x[s] = x;
    ae.y[s] = y;
    ae.vx[s] = vx;
    ae.vy[s] = vy;
    ae.vz[s] = vz;
    ae.tag[s] = obj + 1;
    atomicAdd(&cnt->emit[obj], 1ull);
}

// wall geometry of an emission: dirflag 1 left, 2 above, 3 right, 4 below (the reference's)
struct WallAt {
    int dir;
    double x, y;  // position of the emitted electron
};

__device__ inline WallAt wall_at(const Geo &g, int side, double x, double y) {
    WallAt w;
    if (side == EDGPU_SIDE_LEFT) w = WallAt{1, 0.0 + 1.0e-6, y};
    else if (side == EDGPU_SIDE_ABOVE) w = WallAt{2, x, g.ymax - 1.0e-6};
    else if (side == EDGPU_SIDE_RIGHT) w = WallAt{3, g.xmax - 1.0e-6, y};
    else w = WallAt{4, x, 0.0 + 1.0e-6};
    return w;
}

// The face of inner box b that an electron hit, as the wall facing the same way (side as in
// inner_crossing): the new electron leaves 1e-6 outside that face at the crossing's other
// coordinate (the inner-object branches of INJECT_*, pic2d_SEEmission.f90:372-549, 696-790)
__device__ inline WallAt inner_wall_at(const Geo &g, int b, int side, double x, double y) {
    WallAt w;
    if (side == EDGPU_SIDE_LEFT) w = WallAt{1, (double)g.box[b][1] + 1.0e-6, y};
    else if (side == EDGPU_SIDE_ABOVE) w = WallAt{2, x, (double)g.box[b][2] - 1.0e-6};
    else if (side == EDGPU_SIDE_RIGHT) w = WallAt{3, (double)g.box[b][0] - 1.0e-6, y};
    else w = WallAt{4, x, (double)g.box[b][3] + 1.0e-6};
    return w;
}

// velocity of speed v leaving the wall at angle theta from its normal, azimuth fi
// (INJECT_ELASTIC_REFLECTED_ELECTRON random branch, INJECT_INELASTIC_BACKSCATTERED_ELECTRON)
__device__ inline void leave_wall(int dir, double v, double theta, double fi, double &vx, double &vy, double &vz) {
    const double c = v * cos(theta), ss = v * sin(theta) * sin(fi), sc = v * sin(theta) * cos(fi);
    if (dir == 1) {
        vx = c;
        vy = ss;
        vz = sc;
    } else if (dir == 2) {
        vx = ss;
        vy = -c;
        vz = sc;
    } else if (dir == 3) {
        vx = -c;
        vy = ss;
        vz = sc;
    } else {
        vx = ss;
        vy = c;
        vz = sc;
    }
}

// INJECT_TRUE_SECONDARY_ELECTRON (:794-1050), boundary walls
// (true: an electron was emitted; none once the primary's energy is used up)
__device__ inline bool inject_true(const SeeObj &o, const WallAt &w, int obj, int64_t slot, UniformSeq &u,
                                   const double *maxw, const double *maxw_inj, double &energy_inc,
                                   const AddList &ae, Counters *cnt) {
    if (energy_inc <= 0.0) return false;
    double vx, vy, vz;
    if (w.dir == 1 || w.dir == 3) {
        vx = maxwell_inj_sample(maxw_inj, u.next());
        vy = maxwell_sample(maxw, u.next());
        vz = maxwell_sample(maxw, u.next());
        vx = (w.dir == 1 ? vx : -vx) * o.vinj_true;
        vy = vy * o.vinj_true;
        vz = vz * o.vinj_true;
    } else {
        vx = maxwell_sample(maxw, u.next());
        vy = maxwell_inj_sample(maxw_inj, u.next());
        vz = maxwell_sample(maxw, u.next());
        vx = vx * o.vinj_true;
        vy = (w.dir == 4 ? vy : -vy) * o.vinj_true;
        vz = vz * o.vinj_true;
    }
    const double energy_new = vx * vx + vy * vy + vz * vz;
    if (energy_new > energy_inc) {
        const double alpha = sqrt(energy_inc / energy_new);
        vx = vx * alpha;
        vy = vy * alpha;
        vz = vz * alpha;
        energy_inc = 0.0;
    } else {
        energy_inc = energy_inc - energy_new;
    }
    emit_electron(ae, cnt, obj, slot, w.x, w.y, vx, vy, vz);
    return true;
}

// ---------------------------------------------------------------------------------
// Surface charge of the dielectric parts of the y walls, in macroparticle charges per node.
// A hit (or an emission) at x on segment [istart, iend] of the wall moves dq to the nodes
// ileft = max(INT(x), istart) and iright = min(ileft + 1, iend) with the weights
// 1 - (x - ileft) and x - ileft (pic2d_ElectronWallCollisions.f90:184-194,
// pic2d_IonWallCollisions.f90:209-216, pic2d_SEEmission.f90:268-275): an absorbed electron
// adds -1, an absorbed ion of charge Qs adds Qs (as Qs (x - ileft) and Qs minus that), an
// electron emitted from the wall adds +1.  The segment is the first of the object on that
// side whose node range holds x, the part the reference's loop finds.  With x periodic, node
// 0 is the image of node gmi-1 and node gmi that of node 1: their shares go to those nodes,
// which leaves the two end nodes at zero as the reference's exchange between the periodic
// neighbour clusters does (pic2d_IonWallCollisions.f90:717-766,
// pic2d_CurProblemValues.f90:1773-1782).
__device__ inline double *sigma_row(int side, double *sb, double *sa) {
    return side == EDGPU_SIDE_BELOW ? sb : (side == EDGPU_SIDE_ABOVE ? sa : nullptr);
}

__device__ inline bool dielectric_wall_hit(const Geo &g, int obj, int side) {
    return obj < g.n_boundary_obj && ((g.diel_mask >> obj) & 1u) && (side == EDGPU_SIDE_BELOW || side == EDGPU_SIDE_ABOVE);
}

__device__ inline int periodic_image(const Geo &g, int i) {
    if (!g.periodic_x) return i;
    return i == 0 ? g.nx - 2 : (i == g.nx - 1 ? 1 : i);
}

// charge q (-1 electron, +1 emitted electron, Qs ion) at the clamped x of a hit on a wall;
// count identical contributions add count times the weights
__device__ inline void surface_charge_at(const Geo &g, int obj, int side, double x, int q, double count, double *sb,
                                         double *sa) {
    double *row = sigma_row(side, sb, sa);
    if (!row) return;
    for (int k = 0; k < g.n_seg; ++k) {
        if (g.seg_obj[k] != obj || g.seg_side[k] != side) continue;
        if (!(x >= (double)g.seg_start[k] && x <= (double)g.seg_end[k])) continue;
        const int ileft = max((int)x, g.seg_start[k]);
        const int iright = min(ileft + 1, g.seg_end[k]);
        double dql, dqr;
        if (q == 1 || q == -1) {
            dqr = x - (double)ileft;
            dql = 1.0 - dqr;
            if (q < 0) {
                dql = -dql;
                dqr = -dqr;
            }
        } else {
            dqr = (double)q * (x - (double)ileft);
            dql = (double)q - dqr;
        }
        unsafeAtomicAdd(&row[periodic_image(g, ileft)], count * dql);
        unsafeAtomicAdd(&row[periodic_image(g, iright)], count * dqr);
        return;
    }
}

// the wall hits of one push, records [0, n) of its removal list (x clamped into the area as
// PROCESS_*_COLL_WITH_BOUNDARY_BELOW/ABOVE do before anything else)
__global__ void k_surface_charge(Geo g, const int32_t *robj, const double *rx, int64_t n, int q, double *sb,
                                 double *sa, uint32_t skip_mask) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int code = robj[t];
    const int obj = code & 0xff, side = (code >> 8) - 1;
    if (!dielectric_wall_hit(g, obj, side) || ((skip_mask >> obj) & 1u)) return;
    const double x = fmin(fmax(rx[t], 0.0), g.xmax);
    surface_charge_at(g, obj, side, x, q, 1.0, sb, sa);
}

void launch_surface_charge(const Geo &g, const int32_t *robj, const double *rx, int64_t n, int q, double *sigma_below,
                           double *sigma_above, hipStream_t s, uint32_t skip_mask) {
    if (n <= 0) return;
    k_surface_charge<<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(g, robj, rx, n, q, sigma_below, sigma_above,
                                                                         skip_mask);
}

// PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION of one absorbed electron; returns the number of
// electrons it emits (at the clamped wall position *xw)
template <bool MIXED>
__device__ inline int see_one(const SeeTable *tab, const Geo &g, const ParticlesF64 &pf, const ParticlesMixed &pm,
                              const int64_t *ridx, const int32_t *robj, const double *rx, const double *ry,
                              int64_t t, const double *maxw, const double *maxw_inj, const Rng &rng,
                              const AddList &ae, Counters *cnt, int &obj_out, int &side_out, double &xw) {
    const int code = robj[t];
    const int obj = code & 0xff, side = (code >> 8) - 1;
    obj_out = obj;
    side_out = side;
    if (side < 0 || obj >= EDGPU_MAX_OBJ || !tab->obj[obj].on) return 0;
    const SeeObj &o = tab->obj[obj];
    const int64_t k = ridx[t];
    double vx, vy, vz;
    if (MIXED) {
        vx = (double)pm.vx[k];
        vy = (double)pm.vy[k];
        vz = (double)pm.vz[k];
    } else {
        vx = pf.vx[k];
        vy = pf.vy[k];
        vz = pf.vz[k];
    }
    double x = rx[t], y = ry[t];
    WallAt w;
    if (obj >= g.n_boundary_obj) {
        // an inner object: the record holds the crossing of the trajectory with its face
        w = inner_wall_at(g, obj - g.n_boundary_obj, side, x, y);
    } else {
        // the coordinate along the wall is clamped into the area before the emission
        // (PROCESS_ELECTRON_COLL_WITH_BOUNDARY_*, pic2d_ElectronWallCollisions.f90:24,98,166,238)
        if (side == EDGPU_SIDE_LEFT || side == EDGPU_SIDE_RIGHT) y = fmin(fmax(y, 0.0), g.ymax);
        else x = fmin(fmax(x, 0.0), g.xmax);
        w = wall_at(g, side, x, y);
    }
    xw = x;
    // PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION (pic2d_SEEmission.f90:44-149)
    double energy_inc = vx * vx + vy * vy + vz * vz;
    if (energy_inc <= o.lowest) return 0;
    const double v = sqrt(energy_inc);
    double theta = 0.0;
    if (v != 0.0) theta = acos(fmin(1.0, fabs((w.dir == 1 || w.dir == 3) ? vx : vy) / v));
    const double ce = coeff_elastic(energy_inc, theta, o);
    const double ci = coeff_inelastic(energy_inc, theta, o);
    if (ce + ci > 1.0) {
        set_error(cnt, EDGPU_ERR_SEE, k);
        return 0;
    }
    double ct = coeff_true(energy_inc, theta, o);
    const double total = ce + ci + ct;
    if (total == 0.0) return 0;
    UniformSeq u(rng, k);
    double R = u.next();
    if (R <= ce) {
        // INJECT_ELASTIC_REFLECTED_ELECTRON
        double nx, ny, nz;
        if (o.refl_type == 0) {
            nx = (w.dir == 1 || w.dir == 3) ? -vx : vx;
            ny = (w.dir == 2 || w.dir == 4) ? -vy : vy;
            nz = vz;
        } else {
            const double th = asin(fmin(u.next(), 1.0));
            const double fi = 6.283185307179 * u.next();
            leave_wall(w.dir, v, th, fi, nx, ny, nz);
        }
        emit_electron(ae, cnt, obj, k, w.x, w.y, nx, ny, nz);
        return 1;
    }
    if (R > ce && R <= ce + ci) {
        // INJECT_INELASTIC_BACKSCATTERED_ELECTRON
        const double th = asin(fmin(u.next(), 1.0));
        const double fi = 6.283185307179 * u.next();
        const double v_new = v * sqrt(u.next());
        double nx, ny, nz;
        leave_wall(w.dir, v_new, th, fi, nx, ny, nz);
        emit_electron(ae, cnt, obj, k, w.x, w.y, nx, ny, nz);
        return 1;
    }
    if (ct <= 0.0) return 0;
    if (o.model[2] == 1) energy_inc = fmax(0.0, energy_inc - o.minE[2]);
    else if (o.model[2] == 2) energy_inc = fmax(0.0, energy_inc - o.E_see_0);
    if (total <= 1.0) {
        if (R > ce + ci && R <= total) return inject_true(o, w, obj, k, u, maxw, maxw_inj, energy_inc, ae, cnt) ? 1 : 0;
        return 0;
    }
    int emitted = 0;
    ct = ct / (1.0 - (ce + ci));
    while (ct >= 1.0) {
        if (inject_true(o, w, obj, k, u, maxw, maxw_inj, energy_inc, ae, cnt)) ++emitted;
        ct = ct - 1.0;
    }
    R = u.next();
    if (R < ct && inject_true(o, w, obj, k, u, maxw, maxw_inj, energy_inc, ae, cnt)) ++emitted;
    return emitted;
}

template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_see(const SeeTable *tab, Geo g, ParticlesF64 pf, ParticlesMixed pm,
                                                const int64_t *ridx, const int32_t *robj, const double *rx,
                                                const double *ry, int64_t nrem, const double *maxw,
                                                const double *maxw_inj, Rng rng, AddList ae, Counters *cnt,
                                                double *sigma_below, double *sigma_above) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrem) return;
    int obj = -1, side = -1;
    double xw = 0.0;
    const int emitted =
        see_one<MIXED>(tab, g, pf, pm, ridx, robj, rx, ry, t, maxw, maxw_inj, rng, ae, cnt, obj, side, xw);
    // each electron emitted from a dielectric leaves +1 at the hit (pic2d_SEEmission.f90:268-275)
    if (emitted > 0 && dielectric_wall_hit(g, obj, side))
        surface_charge_at(g, obj, side, xw, 1, (double)emitted, sigma_below, sigma_above);
}

// PROCESS_ION_INDUCED_ELECTRON_EMISSION (pic2d_IonEmissions.f90:4-87) of absorbed ion t of
// species s: energy Ms v^2; inside (minE, maxE) the coefficient setD gives floor(setD)
// electrons, then one more if a uniform falls below the remainder; each is
// INJECT_ION_IMPACT_EXTRACTED_ELECTRON (:91-348), the true-secondary injection with the
// species' temperature, out of the energy above minE.  Positions as for electron-induced
// emission: the clamped wall coordinate, or the crossing of an inner box's face.
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_iiee(const IieeTable *tab, Geo g, int s, ParticlesF64 pf,
                                                 ParticlesMixed pm, const int64_t *ridx, const int32_t *robj,
                                                 const double *rx, const double *ry, int64_t nrem,
                                                 const double *maxw, const double *maxw_inj, Rng rng, AddList ae,
                                                 AddList ai, Counters *cnt, double *sigma_below, double *sigma_above) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrem) return;
    const int code = robj[t];
    const int obj = code & 0xff, side = (code >> 8) - 1;
    if (side < 0 || obj >= EDGPU_MAX_OBJ) return;
    const IieeObj &io = tab->obj[obj][s];
    if (!io.on && !tab->reflect[obj]) return;
    const int64_t k = ridx[t];
    double vx, vy, vz;
    if (MIXED) {
        vx = (double)pm.vx[k];
        vy = (double)pm.vy[k];
        vz = (double)pm.vz[k];
    } else {
        vx = pf.vx[k];
        vy = pf.vy[k];
        vz = pf.vz[k];
    }
    const double energy = tab->Ms[s] * (vx * vx + vy * vy + vz * vz);
    double x = rx[t], y = ry[t];
    WallAt w;
    if (obj >= g.n_boundary_obj) {
        w = inner_wall_at(g, obj - g.n_boundary_obj, side, x, y);
    } else {
        if (side == EDGPU_SIDE_LEFT || side == EDGPU_SIDE_RIGHT) y = fmin(fmax(y, 0.0), g.ymax);
        else x = fmin(fmax(x, 0.0), g.xmax);
        w = wall_at(g, side, x, y);
    }
    if (tab->reflect[obj]) {
        // INJECT_REFLECTED_ION: the velocity component normal to the wall or face reversed
        const unsigned long long a = atomicAdd(ai.count, 1ull);
        if ((int64_t)a >= ai.cap) {
            set_error(cnt, EDGPU_ERR_CAPACITY, k);
            return;
        }
        const bool normal_x = w.dir == 1 || w.dir == 3;
        ai.x[a] = w.x;
        ai.y[a] = w.y;
        ai.vx[a] = normal_x ? -vx : vx;
        ai.vy[a] = normal_x ? vy : -vy;
        ai.vz[a] = vz;
        ai.tag[a] = obj + 1;
        return;
    }
    if (energy <= io.minE || energy >= io.maxE) return;
    double coef = io.setD;
    if (coef == 0.0) return;
    SeeObj o{};
    o.vinj_true = io.vinj;
    double energy_inc = fmax(0.0, energy - io.minE);
    UniformSeq u(rng, k);
    int emitted = 0;
    while (coef >= 1.0) {
        if (inject_true(o, w, obj, k, u, maxw, maxw_inj, energy_inc, ae, cnt)) ++emitted;
        coef = coef - 1.0;
    }
    if (coef != 0.0) {
        const double R = u.next();
        if (R < coef && inject_true(o, w, obj, k, u, maxw, maxw_inj, energy_inc, ae, cnt)) ++emitted;
    }
    // each emitted electron leaves +1 on a dielectric wall (pic2d_IonEmissions.f90:150-225)
    if (emitted > 0 && dielectric_wall_hit(g, obj, side))
        surface_charge_at(g, obj, side, x, 1, (double)emitted, sigma_below, sigma_above);
}

void launch_iiee(const IieeTable *t, const Geo &g, bool mixed, int s, const ParticlesF64 &pf,
                 const ParticlesMixed &pm, const int64_t *ridx, const int32_t *robj, const double *rx,
                 const double *ry, int64_t nrem, const double *maxw, const double *maxw_inj, const Rng &rng,
                 const AddList &ae, const AddList &ai, Counters *cnt, double *sigma_below, double *sigma_above,
                 hipStream_t st) {
    if (nrem <= 0) return;
    const int blocks = (int)((nrem + kBlock - 1) / kBlock);
    if (mixed)
        k_iiee<true><<<blocks, kBlock, 0, st>>>(t, g, s, pf, pm, ridx, robj, rx, ry, nrem, maxw, maxw_inj, rng, ae, ai, cnt,
                                                 sigma_below, sigma_above);
    else
        k_iiee<false><<<blocks, kBlock, 0, st>>>(t, g, s, pf, pm, ridx, robj, rx, ry, nrem, maxw, maxw_inj, rng, ae,
                                                  ai, cnt, sigma_below, sigma_above);
}

// PERFORM_ELECTRON_EMISSION_SETUP for one wall segment: thread t < n_int emits electron t;
// thread n_int tests the fractional electron and emits it with the uniforms of slot n_int+1
__global__ __launch_bounds__(kBlock) void k_constant_emission(EmitPart p, Geo g, const double *maxw,
                                                              const double *maxw_inj, Rng rng, AddList ae,
                                                              Counters *cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > p.n_int) return;
    int64_t slot = t;
    double u[4];
    if (t == p.n_int) {
        draw4(rng, slot, 0, u);
        if (!(u[0] < p.frac)) return;
        slot = (int64_t)t + 1;
    }
    draw4(rng, slot, 0, u);
    const double c = fmin(fmax(p.start + u[0] * p.len, p.lo), p.hi);
    double vn;
    if (p.model == 0) vn = maxwell_inj_sample(maxw_inj, u[1]) * p.f_normal;
    else vn = fmax(0.0, p.v_beam + maxwell_sample(maxw, u[1]) * p.f_normal);
    const double a1 = maxwell_sample(maxw, u[2]) * p.f_parallel;
    const double a2 = maxwell_sample(maxw, u[3]) * p.f_parallel;
    double x, y, vx, vy, vz;
    if (p.side == EDGPU_SIDE_LEFT) {
        x = 0.0 + 1.0e-6;
        y = c;
        vx = vn;
        vy = a1;
        vz = a2;
    } else if (p.side == EDGPU_SIDE_RIGHT) {
        x = g.xmax - 1.0e-6;
        y = c;
        vx = -vn;
        vy = a1;
        vz = a2;
    } else if (p.side == EDGPU_SIDE_ABOVE) {
        x = c;
        y = g.ymax - 1.0e-6;
        vy = -vn;
        vx = a1;
        vz = a2;
    } else {
        x = c;
        y = 0.0 + 1.0e-6;
        vy = vn;
        vx = a1;
        vz = a2;
    }
    emit_electron(ae, cnt, p.obj, slot, x, y, vx, vy, vz);
}

void launch_see(const SeeTable *t, const Geo &g, bool mixed, const ParticlesF64 &pf, const ParticlesMixed &pm,
                const int64_t *ridx, const int32_t *robj, const double *rx, const double *ry, int64_t nrem,
                const double *maxw, const double *maxw_inj, const Rng &rng, const AddList &ae, Counters *cnt,
                hipStream_t s, double *sigma_below, double *sigma_above) {
    if (nrem <= 0) return;
    const int blocks = (int)((nrem + kBlock - 1) / kBlock);
    if (mixed)
        k_see<true><<<blocks, kBlock, 0, s>>>(t, g, pf, pm, ridx, robj, rx, ry, nrem, maxw, maxw_inj, rng, ae, cnt,
                                              sigma_below, sigma_above);
    else
        k_see<false><<<blocks, kBlock, 0, s>>>(t, g, pf, pm, ridx, robj, rx, ry, nrem, maxw, maxw_inj, rng, ae, cnt,
                                               sigma_below, sigma_above);
}

void launch_constant_emission(const EmitPart &part, const Geo &g, const double *maxw, const double *maxw_inj,
                              const Rng &rng, const AddList &ae, Counters *cnt, hipStream_t s) {
    const int blocks = (part.n_int + 1 + kBlock - 1) / kBlock;
    k_constant_emission<<<blocks, kBlock, 0, s>>>(part, g, maxw, maxw_inj, rng, ae, cnt);
}

}  // namespace edgpu
