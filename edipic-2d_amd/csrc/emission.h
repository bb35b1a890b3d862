// Wall emission: electron-induced emission at the boundary walls and the constant
// (thermal or beam) emission of boundary objects (emission.hip).
#pragma once

#include "kernels.h"

namespace edgpu {

// Electron-induced emission parameters of one boundary object with dimensionless energies
// (the conversions of pic2d_Materials.f90:180-245).  Index 0/1/2 of the model-1 arrays:
// elastic reflection, inelastic backscattering, true secondaries.
struct SeeObj {
    int32_t on;
    int32_t model[3];
    int32_t refl_type;  // Elast_refl_type: 0 specular
    int32_t pad;
    double lowest;      // lowest_energy_for_see
    double setD[3], minE[3], maxE[3];
    double E_elast_0, E_elast_max, maxD_elast, dE_elast, frac_elast_high, frac_inelastic;
    double E_see_0, E_see_max, maxD_classic, k_smooth;
    double vinj_true;   // factor_convert_seetrue_vinj
};

struct SeeTable {
    SeeObj obj[EDGPU_MAX_OBJ];
};

// Ion-induced electron emission of one (object, ion species): constant coefficient setD for
// dimensionless incident energies in (minE, maxE), emitted electrons scaled by vinj
// (pic2d_Materials.f90:160-180)
struct IieeObj {
    int32_t on;
    int32_t pad;
    double setD, minE, maxE, vinj;
};

struct IieeTable {
    IieeObj obj[EDGPU_MAX_OBJ][EDGPU_MAX_ION_SPECIES];
    double Ms[EDGPU_MAX_ION_SPECIES];
    int32_t reflect[EDGPU_MAX_OBJ];  // reflects_all_ions
};

// uniforms reserved per absorbed electron in an injected stream
constexpr int kSeeStride = 32;

// Constant emission of one wall segment (PERFORM_ELECTRON_EMISSION_SETUP): electron t takes
// the along-wall coordinate start + u len clamped to [lo, hi]; n_int whole electrons, then
// one more with probability frac.
struct EmitPart {
    int32_t obj, side, model, n_int;
    double frac;
    double start, len, lo, hi;
    double f_normal, f_parallel, v_beam;
};

// emission of the electrons absorbed by the last electron push (entries [0, nrem) of the
// removed list, before compaction): velocities read from the absorbed slots
void launch_see(const SeeTable *t, const Geo &g, bool mixed, const ParticlesF64 &pf, const ParticlesMixed &pm,
                const int64_t *ridx, const int32_t *robj, const double *rx, const double *ry, int64_t nrem,
                const double *maxw, const double *maxw_inj, const Rng &rng, const AddList &ae, Counters *cnt,
                hipStream_t s, double *sigma_below = nullptr, double *sigma_above = nullptr);
// the surface charge the wall hits of one push (removal records [0, n)) leave on the dielectric
// parts of the y walls: q = -1 for electrons, Qs for ions (sigma rows of gmi+1 nodes)
// (hits on the objects of skip_mask leave no charge: ions those objects reflect)
void launch_surface_charge(const Geo &g, const int32_t *robj, const double *rx, int64_t n, int q, double *sigma_below,
                           double *sigma_above, hipStream_t s, uint32_t skip_mask = 0);
// emission of the ions of species s (0-based) absorbed by the last push of that species
// (entries [0, nrem) of its removed list, before compaction)
void launch_iiee(const IieeTable *t, const Geo &g, bool mixed, int s, const ParticlesF64 &pf,
                 const ParticlesMixed &pm, const int64_t *ridx, const int32_t *robj, const double *rx,
                 const double *ry, int64_t nrem, const double *maxw, const double *maxw_inj, const Rng &rng,
                 const AddList &ae, const AddList &ai, Counters *cnt, double *sigma_below, double *sigma_above,
                 hipStream_t st);
void launch_constant_emission(const EmitPart &part, const Geo &g, const double *maxw, const double *maxw_inj,
                              const Rng &rng, const AddList &ae, Counters *cnt, hipStream_t s);

}  // namespace edgpu
