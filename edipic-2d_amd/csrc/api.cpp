// C-ABI of the edgpu library: one device-resident context that holds the state the
// reference keeps in Fortran module globals, and one entry point per reference subroutine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <random>

#include "cnn.h"
#include "emission.h"
#include "kernels.h"

using namespace edgpu;

namespace {

std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

int edgpu::edgpu_set_error(int code, const char *msg) { return fail(code, msg ? msg : ""); }

namespace {

#define HIPTRY(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return fail(EDGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int64_t kChunk = 1 << 23;  // staging chunk (particles) for conversions in mixed mode

struct SpeciesState {
    int64_t n = 0, cap = 0;
    ParticlesF64 f{};     // FP64 arrays (FP64 mode)
    ParticlesF64 ftmp{};  // sort spare (FP64 mode)
    ParticlesMixed m{};   // mixed arrays (mixed mode)
    ParticlesMixed mtmp{};
    AddList add{};
    unsigned long long *add_count = nullptr;
    int64_t *removed_idx = nullptr, *holes = nullptr, *movers = nullptr, *cand = nullptr, *sel = nullptr;
    int64_t cand_cap = 0;
    int32_t *removed_obj = nullptr;
    double *removed_x = nullptr, *removed_y = nullptr;
    int64_t removed_cap = 0;
    int64_t last_removed = 0;
    // electrons [coll_done, n) have not had their collision test of this step yet; the first
    // add_done entries of the add list are excluded from it (ionisation products of the push)
    int64_t coll_done = 0;
    int64_t add_done = 0;
    int32_t *tile_start[2] = {nullptr, nullptr};  // tile ranges of the last sort (double buffer)
    int tile_cur = 0;
    bool tiled = false;
    int64_t n_sorted = 0;
    int Qs = -1;
    double QM2s = 0.0, Ms = 0.0;
    double K[9];
    RcxPar rcx{};
};

// learned warm start: network, grid-sized activations, training pairs, potential history
struct CnnState {
    bool on = false;
    CnnNet net;
    CnnActs a;
    uint16_t *feat_own = nullptr;  // features of the current solve (a.feat points here or into the ring)
    int H = 0, W = 0;
    CnnScales sc{1.0, 1.0, 1.0, 1.0};
    bool scaled = false;
    int max_pairs = 0, n_pairs = 0, next = 0;
    float *ring_raw = nullptr;  // raw features, 4 per node per pair
    float *ring_tgt = nullptr;  // raw targets, 1 per node per pair
    float *raw_now = nullptr;   // raw features of the current solve
    float *tgt_n = nullptr;     // normalised target of the current training step
    double *acc = nullptr;      // 5 FP64 accumulators for the scales
    double *phim1 = nullptr, *phim2 = nullptr;  // potentials of the two previous solves
    bool hist = false;                          // phim1/phim2 describe the last Krylov solve
    bool prev_valid = false;                    // that solve had a previous potential
    double *dloss = nullptr;
    int64_t nfree = 0;
    std::vector<void *> owned;
};

struct Ctx {
    edgpu_params P;
    Geo g;
    int device = -1;
    hipStream_t s = nullptr;
    int nsp = 1;
    SpeciesState sp[1 + EDGPU_MAX_ION_SPECIES];
    int64_t nn = 0;  // grid nodes
    double *rho_raw_e = nullptr, *rho_raw_i = nullptr, *rho_e = nullptr, *rho_i = nullptr, *phi = nullptr;
    double *EX = nullptr, *EY = nullptr, *accEX = nullptr, *accEY = nullptr;
    float *EXf = nullptr, *EYf = nullptr, *accEXf = nullptr, *accEYf = nullptr;
    bool raw_valid_e = false, raw_valid_i = false;
    int N = 0;
    bool fft_ok = false;
    double2 *spec = nullptr, *tw = nullptr;
    double *a_eq = nullptr, *wall_below = nullptr, *wall_above = nullptr;
    // dielectric y walls: the wall counts as dielectric (flag), and the surface charge of its
    // dielectric nodes (gmi+1 per row, allocated when any object is a dielectric)
    bool diel_below = false, diel_above = false;
    double *sigma_below = nullptr, *sigma_above = nullptr;
    double *dval = nullptr;
    KrylovWork kw{};
    bool have_phi = false;
    double *maxw = nullptr, *maxw_inj = nullptr;
    std::vector<double> h_maxw, h_maxw_inj;
    EnTable en{};
    EnTable *en_dev = nullptr;
    double *en_energy = nullptr, *en_prob = nullptr;
    bool en_on = false;
    std::vector<double> h_en_energy, h_en_prob;
    Counters *cnt = nullptr;
    unsigned long long *vz2max = nullptr;  // per species: bound of vz^2 (bits of a double), ions only
    double *stream = nullptr;
    int64_t stream_n = 0;
    uint32_t rng_step = 0;
    SortWork sw{};
    int steps_since_sort[EDGPU_MAX_ION_SPECIES + 1] = {};
    ParticlesF64 stage{}, stage2{};  // mixed-mode staging
    float *mom = nullptr;
    // page-locked host memory for the step's small readbacks (counters, list lengths): copies
    // into it are DMA transfers, where pageable memory goes through a staging copy kernel
    struct HostReadback {
        Counters cnt;
        unsigned long long u64[4];
    } *pin = nullptr;
    // moments accumulated by the pushes of a snapshot step: 13 node arrays per species
    float *mom_push = nullptr;
    bool mom_request = false;
    hipEvent_t ev[11];
    hipEvent_t kev[2];
    double kernel_ms[2] = {0, 0};  // last push kernel of electrons / sum over ion species
    double timing[8] = {0};
    CnnState cnn;
    // wall emission
    SeeTable see{};
    SeeTable *see_dev = nullptr;
    bool see_on = false;
    IieeTable iiee{};
    IieeTable *iiee_dev = nullptr;
    bool iiee_on = false;
    uint32_t reflect_mask = 0;  // boundary objects that reflect every ion
    // electron sub-cycling of edgpu_step: n_sub counts electron steps within an ion cycle;
    // fuse_now switches the fused e-n collisions to the ion steps
    int n_sub = 0;
    bool fuse_now = true;
    bool in_step = false;  // inside edgpu_step, which synchronises once at its end
    // The electron sort of a step is done by the step's electron push (the sorting push) when
    // sorting_push is on: the step counts the tiles of the sort_in_push_n electrons at its
    // start and the push writes them in that tile order.  A pending count is dropped by every
    // call that changes the electron array before the push (drop_sort_in_push).
    bool sorting_push = true;
    bool sort_in_push = false;
    int64_t sort_in_push_n = 0;
    struct ConstEmit {
        double N = 0.0;  // macroparticles per step (a default REAL in the reference)
        int model = 0;
        double f_normal = 0.0, f_parallel = 0.0, v_beam = 0.0;
    } emit[EDGPU_MAX_OBJ];
};

Ctx *C = nullptr;

// ------------------------------------------------------------------------------------
// host-side table builders (shared algorithms with the reference, run once at init)

// PrepareMaxwellDistribIntegral, pic2d_ElectronWallCollisions.f90:835-929
bool build_maxwell(std::vector<double> &v, std::vector<double> &vinj) {
    const int Rm = 9000, Rmi = 4500;
    v.assign(Rm + 1, 0.0);
    vinj.assign(Rmi + 1, 0.0);
    std::vector<double> F(180004, 0.0);
    int N = 180000;
    double Vmin = -3.0, Vmax = 3.0, dV = (Vmax - Vmin) / N;
    for (int i = 1; i <= N + 3; ++i) {
        double t = Vmin + ((double)i - 0.5) * dV;
        F[i] = F[i - 1] + std::exp(-(t * t));
    }
    double temp = F[N];
    for (int i = 0; i <= N + 3; ++i) F[i] = F[i] * (double)Rm / temp;
    v[0] = Vmin;
    int count = 0;
    bool ok1 = false, ok2 = false;
    for (int i = 1; i <= N + 3; ++i)
        if (((int)F[i] - count) == 1) {
            ++count;
            v[count] = Vmin + i * dV;
            if (count == Rm) { ok1 = true; break; }
        }
    N = 90000;
    Vmin = 0.0;
    Vmax = 3.0;
    dV = (Vmax - Vmin) / N;
    std::fill(F.begin(), F.end(), 0.0);
    for (int i = 1; i <= N + 3; ++i) {
        double t = Vmin + ((double)(float)i - 0.5) * dV;
        F[i] = F[i - 1] + std::exp(-(t * t)) * t;
    }
    temp = F[N];
    for (int i = 1; i <= N + 3; ++i) F[i] = F[i] * (double)Rmi / temp;
    vinj[0] = Vmin;
    count = 0;
    for (int i = 1; i <= N + 3; ++i)
        if (((int)F[i] - count) == 1) {
            ++count;
            vinj[count] = Vmin + i * dV;
            if (count == Rmi) { ok2 = true; break; }
        }
    return ok1 && ok2;
}

// Boris rotation matrix, pic2d_ElectronDynamics.f90:91-112
void boris_K(double ax, double ay, double az, double K[9]) {
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

// frequency_of_en_collision, pic2d_enCollisionsSpecificProc.f90:4-57
double en_frequency(double E, int type, int npts, const double *Et, const double *S, double N_m3, int Nsub,
                    double dt) {
    double f;
    if (E >= Et[npts - 1]) {
        f = S[npts - 1] * std::sqrt(2.0 * E * kECl / kMeKg);
    } else if (E < Et[0]) {
        f = (type < 20) ? S[0] * std::sqrt(2.0 * E * kECl / kMeKg) : 0.0;
    } else {
        int j;
        for (j = 1; j <= npts - 1; ++j)
            if (E >= Et[j - 1] && E < Et[j]) break;
        if (j > npts - 1) j = npts - 1;
        double ej = Et[j - 1], ej1 = Et[j];
        double fj = S[j - 1] * std::sqrt(2.0 * ej * kECl / kMeKg);
        double fj1 = S[j] * std::sqrt(2.0 * ej1 * kECl / kMeKg);
        f = fj + (fj1 - fj) * (E - ej) / (ej1 - ej);
    }
    return f * N_m3 * (double)Nsub * dt;
}

void free_species(SpeciesState &sp) {
    void *ptrs[] = {sp.f.base,     sp.ftmp.base, sp.m.base,      sp.mtmp.base, sp.add.x,    sp.add.y,
                    sp.add.vx,     sp.add.vy,   sp.add.vz,       sp.add_count, sp.removed_idx, sp.holes,  sp.movers,
                    sp.removed_obj, sp.cand,  sp.sel,   sp.tile_start[0], sp.tile_start[1], sp.removed_x,
                    sp.removed_y,   sp.add.tag};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    sp = SpeciesState{};
}

void free_ctx(Ctx *c) {
    if (!c) return;
    for (int s = 0; s < 1 + EDGPU_MAX_ION_SPECIES; ++s) free_species(c->sp[s]);
    void *ptrs[] = {c->rho_raw_e, c->rho_raw_i, c->rho_e,    c->rho_i,     c->phi,       c->EX,       c->EY,
                    c->accEX,     c->accEY,     c->EXf,      c->EYf,       c->accEXf,    c->accEYf,   c->spec,
                    c->tw,        c->a_eq,      c->wall_below, c->wall_above, c->dval,    c->kw.r,     c->kw.z,
                    c->kw.p,      c->kw.q,      c->kw.b,     c->kw.mask,   c->kw.red,    c->kw.part,  c->kw.ctl,   c->maxw,     c->maxw_inj,
                    c->en_energy, c->en_prob,   c->en_dev, c->cnt,      c->stream,    c->sw.key,    c->sw.count, c->sw.offset, c->sw.wcount,
                    c->sw.scan_tmp, c->stage.base, c->stage2.base, c->mom,  c->mom_push,
                    c->see_dev,   c->vz2max,      c->sigma_below, c->sigma_above, c->iiee_dev};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (void *p : c->cnn.owned)
        if (p) (void)hipFree(p);
    for (int l = 0; l < c->kw.nlev; ++l) {
        MgLevel &L = c->kw.lev[l];
        void *mp[] = {L.u, L.u2, L.r, l > 0 ? (void *)L.f : nullptr, l > 0 ? (void *)L.mask : nullptr};
        for (void *p : mp)
            if (p) (void)hipFree(p);
    }
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    for (auto &e : c->kev) (void)hipEventDestroy(e);
    if (c->s) (void)hipStreamDestroy(c->s);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->kw.hred) (void)hipHostFree(c->kw.hred);
    delete c;
}

template <typename T>
int dalloc(T **p, size_t count) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) return fail(EDGPU_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return EDGPU_OK;
}

// the grid node at the start (end = false) or end of segment k
void segment_point(const edgpu_params &P, int k, bool end, int &i, int &j) {
    const int c = end ? P.seg_end[k] : P.seg_start[k];
    switch (P.seg_side[k]) {
    case EDGPU_SIDE_BELOW: i = c, j = 0; break;
    case EDGPU_SIDE_ABOVE: i = c, j = P.gmj; break;
    case EDGPU_SIDE_LEFT: i = 0, j = c; break;
    default: i = P.gmi, j = c; break;
    }
}

// The boundary object a vacuum gap's segment k touches at its start or end: the first other
// boundary object, in object order, with a segment (in its own order) starting or ending at
// that node, or -1 (pic2d_CurProblemValues.f90:350-380)
int gap_neighbour(const edgpu_params &P, int k, bool end) {
    int gi, gj;
    segment_point(P, k, end, gi, gj);
    for (int o = 0; o < P.n_boundary_objects; ++o) {
        if (o == P.seg_object[k]) continue;
        for (int m = 0; m < P.n_segments; ++m) {
            if (P.seg_object[m] != o) continue;
            int si, sj, ei, ej;
            segment_point(P, m, false, si, sj);
            segment_point(P, m, true, ei, ej);
            if ((si == gi && sj == gj) || (ei == gi && ej == gj)) return o;
        }
    }
    return -1;
}

// The Dirichlet value segment k gives the node at coordinate c along its wall: its object's
// potential, or on a vacuum gap the potential interpolated linearly between the objects at
// its two ends, those potentials exactly at the ends (pic2d_WallPotentials.f90:29-52)
double segment_value(const edgpu_params &P, int k, int c) {
    const int o = P.seg_object[k];
    if (P.object_type[o] != EDGPU_OBJ_VACUUM_GAP) return P.object_phi[o];
    const double p0 = P.object_phi[gap_neighbour(P, k, false)], p1 = P.object_phi[gap_neighbour(P, k, true)];
    const int s0 = P.seg_start[k], s1 = P.seg_end[k];
    if (c == s0) return p0;
    if (c == s1) return p1;
    return p0 + (p1 - p0) * (double)(c - s0) / (double)(s1 - s0);
}

// The potential rows of the y walls for FFT_X: the nodes of metal and vacuum-gap segments hold
// their values, later segments over earlier ones (pic2d_ElectricFieldCalc_FFT_X.f90:100-145),
// every other node NaN (a dielectric node takes half its row plus its surface charge in the
// solve)
void wall_rows(const edgpu_params &P, std::vector<double> &wb, std::vector<double> &wa) {
    wb.assign((size_t)P.gmi + 1, NAN);
    wa.assign((size_t)P.gmi + 1, NAN);
    for (int k = 0; k < P.n_segments; ++k) {
        if (P.seg_side[k] != EDGPU_SIDE_BELOW && P.seg_side[k] != EDGPU_SIDE_ABOVE) continue;
        if (P.object_type[P.seg_object[k]] == EDGPU_OBJ_DIELECTRIC) continue;
        std::vector<double> &w = P.seg_side[k] == EDGPU_SIDE_BELOW ? wb : wa;
        for (int i = std::max(0, P.seg_start[k]); i <= std::min(P.gmi, P.seg_end[k]); ++i)
            w[i] = segment_value(P, k, i);
    }
}

// Dirichlet values of the Krylov solve over the grid, NaN on free nodes: a wall node takes the
// first segment of its side holding it (pic2d_ElectricFieldCalc_PETSc.F90:60-110, 0 where none
// does), nodes of inner boxes their box's potential, later boxes over earlier ones
std::vector<double> dirichlet_values(const edgpu_params &P, int nx) {
    std::vector<double> dv((size_t)nx * (P.gmj + 1), NAN);
    for (int j = 0; j <= P.gmj; ++j)
        for (int i = 0; i <= P.gmi; ++i) {
            int side = -1;
            if (j == 0) side = EDGPU_SIDE_BELOW;
            else if (j == P.gmj) side = EDGPU_SIDE_ABOVE;
            else if (!P.periodic_x && i == 0) side = EDGPU_SIDE_LEFT;
            else if (!P.periodic_x && i == P.gmi) side = EDGPU_SIDE_RIGHT;
            if (side < 0) continue;
            const int c = (side == EDGPU_SIDE_BELOW || side == EDGPU_SIDE_ABOVE) ? i : j;
            double v = 0.0;
            for (int k = 0; k < P.n_segments; ++k)
                if (P.seg_side[k] == side && c >= P.seg_start[k] && c <= P.seg_end[k]) {
                    v = segment_value(P, k, c);
                    break;
                }
            dv[(size_t)j * nx + i] = v;
        }
    for (int m = 0; m < P.n_inner_objects; ++m)
        for (int j = P.inner_box[m][2]; j <= P.inner_box[m][3]; ++j)
            for (int i = P.inner_box[m][0]; i <= P.inner_box[m][1]; ++i)
                if (i >= 0 && i <= P.gmi && j >= 0 && j <= P.gmj)
                    dv[(size_t)j * nx + i] = P.object_phi[P.n_boundary_objects + m];
    return dv;
}

#define TRY(x)                     \
    do {                           \
        int rc_ = (x);             \
        if (rc_ != EDGPU_OK) return rc_; \
    } while (0)

// The arrays of one particle set share one allocation.  Array a starts at a 2 MiB boundary
// plus a * kArrayStagger bytes, so that element k of the arrays a thread reads together
// does not sit at the same offset within a 2 MiB page in every array.
constexpr size_t kArrayStagger = 4352;
constexpr size_t kArenaAlign = size_t(2) << 20;

int alloc_arena(void **base, int n, const size_t *bytes, char **out) {
    size_t off[8], total = 0;
    for (int a = 0; a < n; ++a) {
        off[a] = total + (size_t)a * kArrayStagger;
        total = (off[a] + std::max(bytes[a], size_t(256)) + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    }
    hipError_t e = hipMalloc(base, total);
    if (e != hipSuccess) return fail(EDGPU_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    for (int a = 0; a < n; ++a) out[a] = (char *)*base + off[a];
    return EDGPU_OK;
}

int alloc_f64(ParticlesF64 &p, int64_t cap) {
    const size_t d = sizeof(double) * (size_t)cap, t = sizeof(int32_t) * (size_t)cap;
    const size_t bytes[6] = {d, d, d, d, d, t};
    char *q[6];
    TRY(alloc_arena(&p.base, 6, bytes, q));
    p.x = (double *)q[0], p.y = (double *)q[1], p.vx = (double *)q[2], p.vy = (double *)q[3];
    p.vz = (double *)q[4], p.tag = (int32_t *)q[5];
    return EDGPU_OK;
}

int alloc_mixed(ParticlesMixed &p, int64_t cap) {
    const size_t f = sizeof(float) * (size_t)cap;
    const size_t bytes[7] = {f, f, f, f, f, f, f};
    char *q[7];
    TRY(alloc_arena(&p.base, 7, bytes, q));
    p.cell = (int32_t *)q[0], p.dx = (float *)q[1], p.dy = (float *)q[2], p.vx = (float *)q[3];
    p.vy = (float *)q[4], p.vz = (float *)q[5], p.tag = (int32_t *)q[6];
    return EDGPU_OK;
}

ParticlesF64 offset(const ParticlesF64 &p, int64_t o) {
    ParticlesF64 q = p;
    q.x += o;
    q.y += o;
    q.vx += o;
    q.vy += o;
    q.vz += o;
    q.tag += o;
    return q;
}

ParticlesMixed offset(const ParticlesMixed &p, int64_t o) {
    ParticlesMixed q = p;
    q.cell += o;
    q.dx += o;
    q.dy += o;
    q.vx += o;
    q.vy += o;
    q.vz += o;
    q.tag += o;
    return q;
}

bool mixed() { return C->P.precision == EDGPU_MIXED; }

Rng make_rng(int stride, uint32_t purpose) {
    Rng r{};
    r.stream = C->stream;
    r.stream_n = C->stream_n;
    r.stride = stride;
    r.key0 = (uint32_t)C->P.seed;
    r.key1 = (uint32_t)(C->P.seed >> 32);
    r.key2 = r.key0 ^ (r.key1 * 0x9E3779B1u) ^ (purpose * 0x85EBCA77u);
    r.step = C->rng_step++;
    r.purpose = purpose;
    return r;
}

// the final synchronisation of an entry point: always when the caller will read what it
// copied to the host, otherwise only outside edgpu_step
int sync_unless_in_step(bool host_reads) {
    if (host_reads || !C->in_step) HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

// The error fields persist through zero_counters (see there): a kernel's error is reported by
// the next check, which clears it.
int check_counters(Counters &h, const char *where) {
    HIPTRY(hipMemcpyAsync(&C->pin->cnt, C->cnt, sizeof(Counters), hipMemcpyDeviceToHost, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    h = C->pin->cnt;
    if (h.err != 0) {
        HIPTRY(hipMemsetAsync(&C->cnt->err, 0, sizeof(Counters) - offsetof(Counters, err), C->s));
        char buf[256];
        const char *what = h.err == EDGPU_ERR_OUT_OF_BOUNDS ? "particle out of bounds"
                           : h.err == EDGPU_ERR_INDEX        ? "index out of bounds"
                           : h.err == EDGPU_ERR_NO_WALL      ? "no boundary object part at the wall hit"
                           : h.err == EDGPU_ERR_CAPACITY     ? "capacity exceeded"
                           : h.err == EDGPU_ERR_TABLE        ? "collision energy table inconsistency"
                                                             : "error";
        std::snprintf(buf, sizeof(buf), "%s: %s (particle slot %lld)", where, what, (long long)h.err_index);
        return fail(h.err, buf);
    }
    return EDGPU_OK;
}

// a copy ordered after everything queued on the context stream, finished on return (the
// context's kernels run on a non-blocking stream, which a null-stream hipMemcpy does not wait for)
int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    HIPTRY(hipMemcpyAsync(dst, src, bytes, kind, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

// zeroes the event counts and keeps the error fields, so that an error survives until a check
// reads it even when the check comes after later kernels (the step checks appends once, at its
// end)
int zero_counters() {
    HIPTRY(hipMemsetAsync(C->cnt, 0, offsetof(Counters, err), C->s));
    return EDGPU_OK;
}

AddList add_of(int s) { return C->sp[s].add; }

// the add lists of all ion species (ionisation appends each ion to its process' species)
IonAddLists ion_adds() {
    IonAddLists l{};
    for (int s = 1; s < C->nsp; ++s) l.s[s - 1] = add_of(s);
    return l;
}

Tiles tiles_of(const SpeciesState &sp) {
    Tiles t{};
    t.ntx = tiles_in_x(C->g);
    if (!sp.tiled) return t;
    t.start = sp.tile_start[sp.tile_cur];
    t.ntiles = t.ntx * C->g.gmj;
    t.covered = sp.n_sorted;
    return t;
}

// deposit particles [begin, n) of species s into the raw grid of its kind
void deposit_range(int s, int64_t begin, int64_t n) {
    SpeciesState &sp = C->sp[s];
    double *raw = s == 0 ? C->rho_raw_e : C->rho_raw_i;
    int Qs = s == 0 ? 1 : sp.Qs;
    if (mixed()) launch_deposit_mixed(sp.m, begin, n, C->g, raw, Qs, C->s);
    else launch_deposit_fp64(sp.f, begin, n, C->g, raw, Qs, C->cnt, C->s);
}

// True when the push cannot change vz, so the kernels skip storing it: no field along z and
// a rotation matrix whose z row is (0, 0, 1) exactly (pic2d_ElectronDynamics.f90:116-149,
// pic2d_IonDynamics.f90:128-155); with Bx = By = 0 the reference's matrix has that row.
// Then vz' = 0 vx- + 0 vy- + 1 vz = vz for every particle (in the precision of the push).
bool vz_fixed_under_push(const AdvanceArgs &a, bool electron) {
    const bool mx = mixed();
    auto is = [mx](double v, double want) { return mx ? (float)v == (float)want : v == want; };
    if (!is(a.EZ, 0.0)) return false;
    if (!electron && !a.sense_B) return true;
    return is(a.K[6], 0.0) && is(a.K[7], 0.0) && is(a.K[8], 1.0);
}

// kVzSkip needs, besides a fixed vz, that vz does not enter vx and vy (K[2] = K[5] = 0, or
// no magnetic rotation at all) and that only the ions' exchange test reads it, with a
// constant cross-section so the test can be bounded (rcx_rejects_any_vz).  Electrons keep
// reading vz: their collision candidacy needs it for about one electron in a hundred per
// step, and loading it only for those lanes stalls half the waves on a dependent load.
int push_vz_mode(const AdvanceArgs &a, bool electron, int coll) {
    if (!vz_fixed_under_push(a, electron)) return kVzReadWrite;
    if (electron || (coll == 2 && a.rcx.alpha != 0.0)) return kVzRead;
    const bool mx = mixed();
    auto zero = [mx](double v) { return mx ? (float)v == 0.0f : v == 0.0; };
    if (a.sense_B && !(zero(a.K[2]) && zero(a.K[5]))) return kVzRead;
    return kVzSkip;
}

// Resets the vz^2 bound of an ion species to the maximum over its particles (after an upload
// or a distribution replaced them); appends and exchanges only raise it afterwards.
int reset_vz_bound(int s) {
    if (s == 0) return EDGPU_OK;
    SpeciesState &sp = C->sp[s];
    HIPTRY(hipMemsetAsync(C->vz2max + s, 0, sizeof(unsigned long long), C->s));
    if (mixed()) launch_vz_bound_mixed(sp.m, 0, sp.n, C->vz2max + s, C->s);
    else launch_vz_bound_fp64(sp.f, 0, sp.n, C->vz2max + s, C->s);
    HIPTRY(hipGetLastError());
    return EDGPU_OK;
}

int sort_species(int s);

int do_advance(int s, bool electron, edgpu_advance_stats *st) {
    SpeciesState &sp = C->sp[s];
    AdvanceArgs a{};
    a.tiles = tiles_of(sp);
    a.p = sp.f;
    a.m = sp.m;
    a.n = sp.n;
    a.EX = electron ? C->EX : C->accEX;
    a.EY = electron ? C->EY : C->accEY;
    a.EXf = electron ? C->EXf : C->accEXf;
    a.EYf = electron ? C->EYf : C->accEYf;
    std::memcpy(a.K, sp.K, sizeof(a.K));
    a.EZ = electron ? C->P.Ez_ext : (C->P.ions_sense_Ez ? C->P.Ez_ext * C->P.N_subcycles : 0.0);
    a.sense_B = electron ? 1 : C->P.ions_sense_magnetic_field;
    a.QM2s = sp.QM2s;
    a.Nsub = (double)C->P.N_subcycles;
    a.g = C->g;
    a.rho = electron ? C->rho_raw_e : C->rho_raw_i;
    a.mom = C->mom_request ? C->mom_push + (size_t)s * kMoments * (size_t)C->nn : nullptr;
    a.Qs = electron ? 1 : sp.Qs;
    a.removed_idx = sp.removed_idx;
    a.removed_obj = sp.removed_obj;
    a.removed_x = sp.removed_x;
    a.removed_y = sp.removed_y;
    a.removed_cap = sp.removed_cap;
    a.cnt = C->cnt;
    a.maxw = C->maxw;
    int coll = 0;
    if (C->P.fuse_collisions) {
        if (electron && C->en_on && C->fuse_now) {
            coll = 1;
            a.en = C->en_dev;
            a.en_pmax = C->en.pmax;
            a.en_V2 = C->en.V2;
            a.rng = make_rng(11, 1);
            a.add_e = add_of(0);
            a.add_i = ion_adds();
        } else if (!electron && sp.rcx.on) {
            coll = 2;
            a.rcx = sp.rcx;
            a.rng = make_rng(4, 2);
        }
    }
    a.cand = sp.cand;
    a.cand_cap = sp.cand_cap;
    a.sel = sp.sel;
    HIPTRY(hipEventRecord(C->kev[0], C->s));
    a.vz2max = C->vz2max + s;
    int vz_mode = push_vz_mode(a, electron, coll);
    if (a.mom && vz_mode == kVzSkip) vz_mode = kVzRead;  // the moments need every particle's vz
    bool sorting = electron && C->sort_in_push;
    if (sorting && a.mom) {
        // a push that sums moments sorts first (the moments' push has no sorting form)
        TRY(sort_species(0));
        a.tiles = tiles_of(sp);
        a.p = sp.f;
        a.m = sp.m;
        sorting = false;
    }
    if (sorting) {
        C->sort_in_push = false;
        if (sp.n != C->sort_in_push_n || vz_mode == kVzSkip)
            return fail(EDGPU_ERR_STATE, "the electrons changed between the tile count and the sorting push");
        // the pushed particles go to the spare arrays in the counted tile order, which they keep
        if (mixed()) {
            launch_advance_sort_mixed(a, sp.mtmp, C->sw, coll, vz_mode, C->s);
            std::swap(sp.m, sp.mtmp);
            a.m = sp.m;
        } else {
            launch_advance_sort_fp64(a, sp.ftmp, C->sw, coll, vz_mode, C->s);
            std::swap(sp.f, sp.ftmp);
            a.p = sp.f;
        }
        sp.tile_cur = 1 - sp.tile_cur;
        sp.tiled = true;
        sp.n_sorted = sp.n;
    } else if (mixed()) {
        launch_advance_mixed(a, electron, coll, vz_mode, C->s);
    } else {
        launch_advance_fp64(a, electron, coll, vz_mode, C->s);
    }
    HIPTRY(hipEventRecord(C->kev[1], C->s));
    if (mixed()) launch_inner_hits_mixed(a, electron, C->s);
    else launch_inner_hits_fp64(a, electron, C->s);
    unsigned long long n_tested_adds = 0;
    if (coll == 1) {
        if (mixed()) launch_candidates_mixed(a, C->s);
        else launch_candidates_fp64(a, C->s);
        // the add list now holds the ionisation electrons of this push's collisions, which the
        // reference creates after PERFORM_ELECTRON_NEUTRAL_COLLISIONS and does not collide in
        // their birth step; the emitted electrons appended next are collided
        HIPTRY(hipMemcpyAsync(&C->pin->u64[0], sp.add_count, sizeof(n_tested_adds), hipMemcpyDeviceToHost, C->s));
    }
    Counters h;
    TRY(check_counters(h, electron ? "ADVANCE_ELECTRONS" : "ADVANCE_IONS"));
    if (coll == 1) n_tested_adds = C->pin->u64[0];
    if (coll == 1) sp.add_done = (int64_t)n_tested_adds;
    {
        float kms = 0.0f;
        HIPTRY(hipEventElapsedTime(&kms, C->kev[0], C->kev[1]));
        if (electron) C->kernel_ms[0] = kms;
        else C->kernel_ms[1] += kms;
    }
    const int64_t nrem = (int64_t)h.n_removed;
    if (C->g.diel_mask && nrem > 0)
        // the charge the absorbed particles leave on dielectric walls
        // (pic2d_ElectronWallCollisions.f90:184-194, pic2d_IonWallCollisions.f90:209-216)
        launch_surface_charge(C->g, sp.removed_obj, sp.removed_x, std::min(nrem, sp.removed_cap), electron ? -1 : sp.Qs,
                              C->sigma_below, C->sigma_above, C->s, electron ? 0u : C->reflect_mask);
    if (electron && C->see_on && nrem > 0) {
        // PROCESS_ELECTRON_INDUCED_ELECTRON_EMISSION of the absorbed electrons, whose slots the
        // compaction below overwrites
        launch_see(C->see_dev, C->g, mixed(), sp.f, sp.m, sp.removed_idx, sp.removed_obj, sp.removed_x, sp.removed_y,
                   std::min(nrem, sp.removed_cap), C->maxw, C->maxw_inj, make_rng(kSeeStride, 3), add_of(0), C->cnt,
                   C->s, C->sigma_below, C->sigma_above);
        TRY(check_counters(h, "ADVANCE_ELECTRONS (electron-induced emission)"));
    }
    if (!electron && C->iiee_on && nrem > 0) {
        // the absorbed ions of reflecting objects back into the ion add list, then
        // PROCESS_ION_INDUCED_ELECTRON_EMISSION of the others, into the electron add list
        launch_iiee(C->iiee_dev, C->g, mixed(), s - 1, sp.f, sp.m, sp.removed_idx, sp.removed_obj, sp.removed_x,
                    sp.removed_y, std::min(nrem, sp.removed_cap), C->maxw, C->maxw_inj, make_rng(kSeeStride, 5),
                    add_of(0), add_of(s), C->cnt, C->sigma_below, C->sigma_above, C->s);
        TRY(check_counters(h, "ADVANCE_IONS (ion-induced emission)"));
    }
    if (nrem > 0) {
        // compaction uses the hole/mover counters of the same Counters block
        if (mixed()) launch_compact_mixed(sp.m, sp.removed_idx, nrem, sp.n, sp.holes, sp.movers, C->cnt, C->s);
        else launch_compact_fp64(sp.f, sp.removed_idx, nrem, sp.n, sp.holes, sp.movers, C->cnt, C->s);
    }
    if (st) {
        st->n_before += sp.n;
        for (int o = 0; o < EDGPU_MAX_OBJ; ++o) st->hit_count[o] += (int64_t)h.hit[o];
        st->n_removed += nrem;
        for (int p = 0; p < EDGPU_MAX_PROC; ++p) st->collisions[p] += (int64_t)h.coll[p];
        for (int o = 0; o < EDGPU_MAX_OBJ; ++o) st->emit_count[o] += (int64_t)h.emit[o];
    }
    sp.last_removed = nrem;
    sp.n -= nrem;
    sp.coll_done = coll ? sp.n : 0;
    if (st) st->n_after += sp.n;
    return EDGPU_OK;
}

int upload_range(int s, int64_t n0, int64_t n, const double *X, const double *Y, const double *VX, const double *VY,
                 const double *VZ, const int32_t *tag) {
    SpeciesState &sp = C->sp[s];
    if (!mixed()) {
        ParticlesF64 d = offset(sp.f, n0);
        HIPTRY(hipMemcpyAsync(d.x, X, sizeof(double) * n, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(d.y, Y, sizeof(double) * n, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(d.vx, VX, sizeof(double) * n, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(d.vy, VY, sizeof(double) * n, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(d.vz, VZ, sizeof(double) * n, hipMemcpyHostToDevice, C->s));
        if (tag) HIPTRY(hipMemcpyAsync(d.tag, tag, sizeof(int32_t) * n, hipMemcpyHostToDevice, C->s));
        else HIPTRY(hipMemsetAsync(d.tag, 0, sizeof(int32_t) * n, C->s));
        HIPTRY(hipStreamSynchronize(C->s));
        return EDGPU_OK;
    }
    for (int64_t o = 0; o < n; o += kChunk) {
        int64_t c = std::min(kChunk, n - o);
        HIPTRY(hipMemcpyAsync(C->stage.x, X + o, sizeof(double) * c, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(C->stage.y, Y + o, sizeof(double) * c, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(C->stage.vx, VX + o, sizeof(double) * c, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(C->stage.vy, VY + o, sizeof(double) * c, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipMemcpyAsync(C->stage.vz, VZ + o, sizeof(double) * c, hipMemcpyHostToDevice, C->s));
        if (tag) HIPTRY(hipMemcpyAsync(C->stage.tag, tag + o, sizeof(int32_t) * c, hipMemcpyHostToDevice, C->s));
        else HIPTRY(hipMemsetAsync(C->stage.tag, 0, sizeof(int32_t) * c, C->s));
        launch_to_mixed(C->stage, offset(sp.m, n0 + o), c, C->g, C->s);
        HIPTRY(hipStreamSynchronize(C->s));
    }
    return EDGPU_OK;
}

int species_ok(int32_t s) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (s < 0 || s >= C->nsp) return fail(EDGPU_ERR_ARG, "species index out of range");
    return EDGPU_OK;
}

// Room for `need` more particles of species s: when the arrays are full they grow as the
// reference grows electron(:) (pic2d_ElectronDynamics.f90:1458-1482): by what is missing or a
// tenth of the capacity, whichever is more.  The particles keep their slots (and the sort's
// tiles stay valid); the sort spare and the sort keys grow with them.
int ensure_capacity(int s, int64_t need) {
    SpeciesState &sp = C->sp[s];
    const int64_t free_slots = sp.cap - sp.n;
    if (need <= free_slots) return EDGPU_OK;
    const int64_t newcap = sp.cap + std::max(need - free_slots, sp.cap / 10);
    if (C->P.sort_interval > 0 && newcap >= ((int64_t)1 << 31))
        return fail(EDGPU_ERR_CAPACITY, "particle arrays would grow past the 2^31 particles the sort indexes");
    HIPTRY(hipStreamSynchronize(C->s));
    if (mixed()) {
        ParticlesMixed m{};
        if (alloc_mixed(m, newcap) != EDGPU_OK)
            return fail(EDGPU_ERR_CAPACITY, "growing the particle arrays failed: " + g_err);
        const size_t f = sizeof(float) * (size_t)sp.n;
        HIPTRY(hipMemcpyAsync(m.cell, sp.m.cell, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(m.dx, sp.m.dx, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(m.dy, sp.m.dy, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(m.vx, sp.m.vx, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(m.vy, sp.m.vy, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(m.vz, sp.m.vz, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(m.tag, sp.m.tag, f, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipStreamSynchronize(C->s));
        HIPTRY(hipFree(sp.m.base));
        sp.m = m;
        if (sp.mtmp.base) {
            HIPTRY(hipFree(sp.mtmp.base));
            sp.mtmp = ParticlesMixed{};
            if (alloc_mixed(sp.mtmp, newcap) != EDGPU_OK)
                return fail(EDGPU_ERR_CAPACITY, "growing the sort spare failed: " + g_err);
        }
    } else {
        ParticlesF64 f{};
        if (alloc_f64(f, newcap) != EDGPU_OK)
            return fail(EDGPU_ERR_CAPACITY, "growing the particle arrays failed: " + g_err);
        const size_t d = sizeof(double) * (size_t)sp.n;
        HIPTRY(hipMemcpyAsync(f.x, sp.f.x, d, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(f.y, sp.f.y, d, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(f.vx, sp.f.vx, d, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(f.vy, sp.f.vy, d, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(f.vz, sp.f.vz, d, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(f.tag, sp.f.tag, sizeof(int32_t) * (size_t)sp.n, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipStreamSynchronize(C->s));
        HIPTRY(hipFree(sp.f.base));
        sp.f = f;
        if (sp.ftmp.base) {
            HIPTRY(hipFree(sp.ftmp.base));
            sp.ftmp = ParticlesF64{};
            if (alloc_f64(sp.ftmp, newcap) != EDGPU_OK)
                return fail(EDGPU_ERR_CAPACITY, "growing the sort spare failed: " + g_err);
        }
    }
    // the per-push lists are sized from the capacity as at init
    const int64_t rcap = std::max<int64_t>(1 << 16, newcap / 16);
    if (rcap > sp.removed_cap) {
        void *old[] = {sp.removed_idx, sp.removed_obj, sp.removed_x, sp.removed_y, sp.holes, sp.movers};
        for (void *q : old) HIPTRY(hipFree(q));
        sp.removed_idx = sp.holes = sp.movers = nullptr;
        sp.removed_obj = nullptr;
        sp.removed_x = sp.removed_y = nullptr;
        TRY(dalloc(&sp.removed_idx, rcap));
        TRY(dalloc(&sp.removed_obj, rcap));
        TRY(dalloc(&sp.removed_x, rcap));
        TRY(dalloc(&sp.removed_y, rcap));
        TRY(dalloc(&sp.holes, rcap));
        TRY(dalloc(&sp.movers, rcap));
        sp.removed_cap = rcap;
        sp.last_removed = 0;
    }
    if (sp.cand && std::max<int64_t>(1 << 16, newcap / 8) > sp.cand_cap) {
        HIPTRY(hipFree(sp.cand));
        HIPTRY(hipFree(sp.sel));
        sp.cand = sp.sel = nullptr;
        sp.cand_cap = std::max<int64_t>(1 << 16, newcap / 8);
        TRY(dalloc(&sp.cand, sp.cand_cap));
        TRY(dalloc(&sp.sel, sp.cand_cap));
    }
    if (C->sw.key && newcap > C->sw.key_cap) {
        HIPTRY(hipFree(C->sw.key));
        C->sw.key = nullptr;
        TRY(dalloc(&C->sw.key, newcap));
        C->sw.key_cap = newcap;
    }
    sp.cap = newcap;
    return EDGPU_OK;
}

// copies particles [n0, n0 + n) of species s to host SoA arrays (tag may be null)
int download_range(int s, int64_t n0, int64_t n, double *X, double *Y, double *VX, double *VY, double *VZ,
                   int32_t *tag) {
    SpeciesState &sp = C->sp[s];
    if (!mixed()) {
        const ParticlesF64 d = offset(sp.f, n0);
        HIPTRY(hipMemcpyAsync(X, d.x, sizeof(double) * n, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(Y, d.y, sizeof(double) * n, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(VX, d.vx, sizeof(double) * n, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(VY, d.vy, sizeof(double) * n, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(VZ, d.vz, sizeof(double) * n, hipMemcpyDeviceToHost, C->s));
        if (tag) HIPTRY(hipMemcpyAsync(tag, d.tag, sizeof(int32_t) * n, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipStreamSynchronize(C->s));
        return EDGPU_OK;
    }
    for (int64_t o = 0; o < n; o += kChunk) {
        const int64_t c = std::min(kChunk, n - o);
        launch_to_fp64(offset(sp.m, n0 + o), C->stage, c, C->g, C->s);
        HIPTRY(hipMemcpyAsync(X + o, C->stage.x, sizeof(double) * c, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(Y + o, C->stage.y, sizeof(double) * c, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(VX + o, C->stage.vx, sizeof(double) * c, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(VY + o, C->stage.vy, sizeof(double) * c, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipMemcpyAsync(VZ + o, C->stage.vz, sizeof(double) * c, hipMemcpyDeviceToHost, C->s));
        if (tag) HIPTRY(hipMemcpyAsync(tag + o, C->stage.tag, sizeof(int32_t) * c, hipMemcpyDeviceToHost, C->s));
        HIPTRY(hipStreamSynchronize(C->s));
    }
    return EDGPU_OK;
}

// byte offsets of the fields of the reference's TYPE particle (pic2d_Modules.f90:432-439):
// REAL(8) X, Y, VX, VY, VZ, then INTEGER tag
constexpr size_t kAosTagOffset = 40;
constexpr int64_t kAosChunk = 1 << 20;  // particles converted per host chunk

double *grid_ptr(int which) {
    switch (which) {
    case 0: return C->rho_e;
    case 1: return C->rho_i;
    case 2: return C->phi;
    case 3: return C->EX;
    case 4: return C->EY;
    case 5: return C->accEX;
    case 6: return C->accEY;
    default: return nullptr;
    }
}

void refresh_f32_fields() {
    if (!mixed()) return;
    launch_field_to_f32(C->EX, C->EY, C->EXf, C->EYf, C->nn, C->s);
    launch_field_to_f32(C->accEX, C->accEY, C->accEXf, C->accEYf, C->nn, C->s);
}

// A pending count of the sorting push no longer describes the electrons once their array changes
void drop_sort_in_push(int s) {
    if (s == 0) C->sort_in_push = false;
}

// The counts of a sorting push of the electrons (the tiles the particles are in now) and
// the new tile ranges; the next electron push places the particles.
int count_tiles_for_push() {
    SpeciesState &sp = C->sp[0];
    const Tiles old = tiles_of(sp);
    int32_t *dst = sp.tile_start[1 - sp.tile_cur];
    if (mixed()) HIPTRY(launch_tile_counts_mixed(sp.m, sp.n, C->g, C->sw, old, dst, C->s));
    else HIPTRY(launch_tile_counts_fp64(sp.f, sp.n, C->g, C->sw, old, dst, C->s));
    HIPTRY(hipGetLastError());
    C->sort_in_push = true;
    C->sort_in_push_n = sp.n;
    return EDGPU_OK;
}

// The electrons of a step whose sort is due are sorted by their push when they were sorted
// before (the push's chunks follow the previous tiles), no moments are summed in the push and
// there is a spare set of arrays to write to.
bool electrons_sort_in_push() {
    const SpeciesState &sp = C->sp[0];
    return C->sorting_push && sp.tiled && sp.n > 0 && !C->mom_request && (mixed() ? sp.mtmp.base : sp.ftmp.base);
}

int sort_species(int s) {
    drop_sort_in_push(s);
    SpeciesState &sp = C->sp[s];
    if (sp.n <= 0 || !sp.tile_start[0]) return EDGPU_OK;
    const Tiles old = tiles_of(sp);
    int32_t *dst = sp.tile_start[1 - sp.tile_cur];
    if (mixed()) {
        HIPTRY(launch_sort_mixed(sp.m, sp.mtmp, sp.n, C->g, C->sw, old, dst, C->s));
        std::swap(sp.m, sp.mtmp);
    } else {
        HIPTRY(launch_sort_fp64(sp.f, sp.ftmp, sp.n, C->g, C->sw, old, dst, C->s));
        std::swap(sp.f, sp.ftmp);
    }
    HIPTRY(hipGetLastError());
    sp.tile_cur = 1 - sp.tile_cur;
    sp.tiled = true;
    sp.n_sorted = sp.n;
    return EDGPU_OK;
}

// ---- learned warm start helpers ---------------------------------------------------------

template <typename T>
int cnn_alloc(T **p, size_t count) {
    TRY(dalloc(p, count));
    C->cnn.owned.push_back((void *)*p);
    return EDGPU_OK;
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int cnn_alloc_acts(CnnActs &a, int64_t n, bool own_feat) {
    if (own_feat) TRY(cnn_alloc(&a.feat, (size_t)n * 16));
    for (int l = 0; l < kCnnLayers - 1; ++l) TRY(cnn_alloc(&a.act[l], (size_t)n * 32));
    TRY(cnn_alloc(&a.pred, (size_t)n));
    TRY(cnn_alloc(&a.dout16, (size_t)n * 16));
    TRY(cnn_alloc(&a.g[0], (size_t)n * 32));
    TRY(cnn_alloc(&a.g[1], (size_t)n * 32));
    TRY(cnn_alloc(&a.wflip, (size_t)9 * 32 * 32));
    TRY(cnn_alloc(&a.partial, (size_t)kCnnWgradBlocks * (9 * 32 * 32 + 32)));
    TRY(cnn_alloc(&a.gsum, (size_t)(9 * 32 * 32 + 32)));
    return EDGPU_OK;
}

// scales of the inputs and the target: root mean squares over the free nodes of all recorded
// pairs (fixed at the first training call; a zero scale, e.g. of a constant channel, is 1)
int cnn_set_scales() {
    CnnState &c = C->cnn;
    const int64_t n = C->nn;
    HIPTRY(hipMemsetAsync(c.acc, 0, sizeof(double) * 5, C->s));
    for (int k = 0; k < c.n_pairs; ++k)
        launch_cnn_sumsq(c.ring_raw + (size_t)k * n * 4, c.ring_tgt + (size_t)k * n, n, c.acc, C->s);
    double h[5];
    HIPTRY(hipMemcpyAsync(h, c.acc, sizeof(h), hipMemcpyDeviceToHost, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    if (!(h[4] > 0.0)) return fail(EDGPU_ERR_STATE, "no free nodes in the recorded pairs");
    auto rms = [&](double v) { const double r = std::sqrt(v / h[4]); return r > 0.0 ? r : 1.0; };
    c.sc = CnnScales{rms(h[0]), rms(h[1]), rms(h[2]), rms(h[3])};
    c.nfree = (int64_t)std::llround(h[4] / (double)c.n_pairs);
    c.scaled = true;
    return EDGPU_OK;
}

}  // namespace

// ======================================================================================
extern "C" {

const char *edgpu_last_error(void) { return g_err.c_str(); }

int edgpu_device(void) { return C ? C->device : -1; }

void edgpu_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

void edgpu_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) { philox2x32_10(ctr[0], ctr[1], key, out); }

int edgpu_get_vz_bound(int32_t species, double *vz2max) {
    TRY(species_ok(species));
    if (!vz2max) return fail(EDGPU_ERR_ARG, "null output");
    unsigned long long b = 0;
    HIPTRY(hipMemcpyAsync(&b, C->vz2max + species, sizeof(b), hipMemcpyDeviceToHost, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    std::memcpy(vz2max, &b, sizeof(double));
    return EDGPU_OK;
}

int edgpu_get_rng_state(uint64_t *seed, uint32_t *counter, int32_t *n_sub) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (seed) *seed = C->P.seed;
    if (counter) *counter = C->rng_step;
    if (n_sub) *n_sub = C->n_sub;
    return EDGPU_OK;
}

int edgpu_set_rng_state(uint64_t seed, uint32_t counter, int32_t n_sub) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (n_sub < 0 || n_sub >= C->P.N_subcycles) return fail(EDGPU_ERR_ARG, "n_sub outside the ion cycle");
    C->P.seed = seed;
    C->rng_step = counter;
    C->n_sub = n_sub;
    return EDGPU_OK;
}

int edgpu_finalize(void) {
    if (C) {
        (void)hipStreamSynchronize(C->s);
        free_ctx(C);
        C = nullptr;
    }
    return EDGPU_OK;
}

int edgpu_init(const edgpu_params *p) {
    if (!p) return fail(EDGPU_ERR_ARG, "null params");
    if (C) edgpu_finalize();
    const edgpu_params &P = *p;
    if (P.gmi < 3 || P.gmj < 2 || P.gmi > 65535 || P.gmj > 32767) return fail(EDGPU_ERR_ARG, "grid extents out of range");
    if (P.n_ion_species < 1 || P.n_ion_species > EDGPU_MAX_ION_SPECIES) return fail(EDGPU_ERR_ARG, "n_ion_species");
    if (P.n_segments < 0 || P.n_segments > EDGPU_MAX_SEG) return fail(EDGPU_ERR_ARG, "n_segments");
    if (P.n_inner_objects < 0 || P.n_inner_objects > EDGPU_MAX_INNER) return fail(EDGPU_ERR_ARG, "n_inner_objects");
    if (P.n_boundary_objects < 0 || P.n_boundary_objects + P.n_inner_objects > EDGPU_MAX_OBJ)
        return fail(EDGPU_ERR_ARG, "object count");
    if (P.N_subcycles < 1 || P.N_of_particles_cell < 1) return fail(EDGPU_ERR_ARG, "N_subcycles/N_of_particles_cell");
    if (P.precision != EDGPU_FP64 && P.precision != EDGPU_MIXED) return fail(EDGPU_ERR_ARG, "precision");
    if (P.max_electrons < 1) return fail(EDGPU_ERR_ARG, "max_electrons");
    for (int s = 0; s < P.n_ion_species; ++s) {
        if (P.max_ions[s] < 1) return fail(EDGPU_ERR_ARG, "max_ions");
        if (!(P.Qs[s] == 1 || P.Qs[s] == 2 || P.Qs[s] == 3 || P.Qs[s] == -1)) return fail(EDGPU_ERR_ARG, "Qs");
    }
    for (int k = 0; k < P.n_segments; ++k)
        if (P.seg_side[k] < 0 || P.seg_side[k] > 3 || P.seg_object[k] < 0 || P.seg_object[k] >= P.n_boundary_objects)
            return fail(EDGPU_ERR_ARG, "segment description");
    // dielectrics: boundary objects on the y walls of an x-periodic system without inner objects,
    // the geometry whose solve (FFT_X) models them (pic2d_Prepare_FFTX_SYSY.f90:776-897)
    for (int o = 0; o < EDGPU_MAX_OBJ; ++o) {
        if (P.object_type[o] == EDGPU_OBJ_METAL) continue;
        if (P.object_type[o] == EDGPU_OBJ_VACUUM_GAP) {
            // one segment whose two ends touch other boundary objects
            // (pic2d_CurProblemValues.f90:341-390)
            int nseg = 0, k0 = -1;
            for (int k = 0; k < P.n_segments; ++k)
                if (P.seg_object[k] == o) ++nseg, k0 = k;
            if (o >= P.n_boundary_objects || nseg != 1 || P.seg_end[k0] <= P.seg_start[k0])
                return fail(EDGPU_ERR_ARG, "a vacuum gap is a boundary object of one segment");
            if (gap_neighbour(P, k0, false) < 0 || gap_neighbour(P, k0, true) < 0)
                return fail(EDGPU_ERR_ARG, "an end of a vacuum gap is not connected to any other boundary object");
            continue;
        }
        if (P.object_type[o] != EDGPU_OBJ_DIELECTRIC) return fail(EDGPU_ERR_ARG, "object type");
        if (o >= P.n_boundary_objects || !P.periodic_x || P.n_inner_objects > 0)
            return fail(EDGPU_ERR_ARG, "dielectric objects are walls below or above an x-periodic system without inner objects");
        if (!(P.object_eps[o] > 0.0)) return fail(EDGPU_ERR_ARG, "dielectric permittivity");
        for (int k = 0; k < P.n_segments; ++k)
            if (P.seg_object[k] == o && P.seg_side[k] != EDGPU_SIDE_BELOW && P.seg_side[k] != EDGPU_SIDE_ABOVE)
                return fail(EDGPU_ERR_ARG, "dielectric objects are walls below or above an x-periodic system without inner objects");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(EDGPU_ERR_HIP, "no HIP device available");

    C = new Ctx();
    C->P = P;
    HIPTRY(hipGetDevice(&C->device));
    HIPTRY(hipStreamCreateWithFlags(&C->s, hipStreamNonBlocking));
    for (auto &e : C->ev) HIPTRY(hipEventCreate(&e));
    for (auto &e : C->kev) HIPTRY(hipEventCreate(&e));
    C->nsp = 1 + P.n_ion_species;
    Geo &g = C->g;
    g.gmi = P.gmi;
    g.gmj = P.gmj;
    g.nx = P.gmi + 1;
    g.ny = P.gmj + 1;
    g.periodic_x = P.periodic_x ? 1 : 0;
    g.xmax = (double)P.gmi;
    g.ymax = (double)P.gmj;
    g.Lx = (double)(P.gmi - 1);
    g.n_seg = P.n_segments;
    for (int k = 0; k < P.n_segments; ++k) {
        g.seg_side[k] = P.seg_side[k];
        g.seg_start[k] = P.seg_start[k];
        g.seg_end[k] = P.seg_end[k];
        g.seg_obj[k] = P.seg_object[k];
    }
    g.n_boundary_obj = P.n_boundary_objects;
    g.n_inner = P.n_inner_objects;
    for (int m = 0; m < P.n_inner_objects; ++m)
        for (int q = 0; q < 4; ++q) g.box[m][q] = P.inner_box[m][q];
    g.diel_mask = 0;
    for (int o = 0; o < P.n_boundary_objects; ++o)
        if (P.object_type[o] == EDGPU_OBJ_DIELECTRIC) g.diel_mask |= 1u << o;
    C->nn = (int64_t)g.nx * g.ny;
    const int64_t nn = C->nn;

    // species
    const int64_t addcap = P.add_capacity > 0 ? P.add_capacity : 1 << 20;
    for (int s = 0; s < C->nsp; ++s) {
        SpeciesState &sp = C->sp[s];
        sp.cap = s == 0 ? P.max_electrons : P.max_ions[s - 1];
        if (s > 0) {
            sp.Qs = P.Qs[s - 1];
            sp.QM2s = P.QM2s[s - 1];
            sp.Ms = P.Ms[s - 1];
            double QM2sNsub = (double)P.N_subcycles * sp.QM2s;
            boris_K(QM2sNsub * P.Bx_ext, QM2sNsub * P.By_ext, QM2sNsub * P.Bz_ext, sp.K);
        } else {
            sp.Qs = 1;
            boris_K(-0.5 * P.Bx_ext, -0.5 * P.By_ext, -0.5 * P.Bz_ext, sp.K);
        }
        if (mixed()) {
            TRY(alloc_mixed(sp.m, sp.cap));
            if (P.sort_interval > 0) TRY(alloc_mixed(sp.mtmp, sp.cap));
        } else {
            TRY(alloc_f64(sp.f, sp.cap));
            if (P.sort_interval > 0) TRY(alloc_f64(sp.ftmp, sp.cap));
        }
        sp.add.cap = addcap;
        TRY(dalloc(&sp.add.x, addcap));
        TRY(dalloc(&sp.add.y, addcap));
        TRY(dalloc(&sp.add.vx, addcap));
        TRY(dalloc(&sp.add.vy, addcap));
        TRY(dalloc(&sp.add.vz, addcap));
        TRY(dalloc(&sp.add.tag, addcap));
        TRY(dalloc(&sp.add_count, 1));
        HIPTRY(hipMemset(sp.add_count, 0, sizeof(unsigned long long)));
        sp.add.count = sp.add_count;
        sp.removed_cap = std::max<int64_t>(1 << 16, sp.cap / 16);
        TRY(dalloc(&sp.removed_idx, sp.removed_cap));
        TRY(dalloc(&sp.removed_obj, sp.removed_cap));
        TRY(dalloc(&sp.removed_x, sp.removed_cap));
        TRY(dalloc(&sp.removed_y, sp.removed_cap));
        TRY(dalloc(&sp.holes, sp.removed_cap));
        TRY(dalloc(&sp.movers, sp.removed_cap));
        if (s == 0) {
            sp.cand_cap = std::max<int64_t>(1 << 16, sp.cap / 8);
            TRY(dalloc(&sp.cand, sp.cand_cap));
            TRY(dalloc(&sp.sel, sp.cand_cap));
        }
    }
    if (mixed()) {
        TRY(alloc_f64(C->stage, kChunk));
        TRY(alloc_f64(C->stage2, kChunk));
        TRY(dalloc(&C->EXf, nn));
        TRY(dalloc(&C->EYf, nn));
        TRY(dalloc(&C->accEXf, nn));
        TRY(dalloc(&C->accEYf, nn));
    }
    // grids
    TRY(dalloc(&C->rho_raw_e, nn));
    TRY(dalloc(&C->rho_raw_i, nn));
    TRY(dalloc(&C->rho_e, nn));
    TRY(dalloc(&C->rho_i, nn));
    TRY(dalloc(&C->phi, nn));
    TRY(dalloc(&C->EX, nn));
    TRY(dalloc(&C->EY, nn));
    TRY(dalloc(&C->accEX, nn));
    TRY(dalloc(&C->accEY, nn));
    double *zero_these[] = {C->rho_raw_e, C->rho_raw_i, C->rho_e, C->rho_i, C->phi, C->EX, C->EY, C->accEX, C->accEY};
    for (double *z : zero_these) HIPTRY(hipMemset(z, 0, sizeof(double) * nn));
    refresh_f32_fields();
    TRY(dalloc(&C->cnt, 1));
    HIPTRY(hipMemset(C->cnt, 0, sizeof(Counters)));
    HIPTRY(hipHostMalloc((void **)&C->pin, sizeof(*C->pin), hipHostMallocDefault));
    TRY(dalloc(&C->vz2max, 1 + EDGPU_MAX_ION_SPECIES));
    HIPTRY(hipMemset(C->vz2max, 0, sizeof(unsigned long long) * (1 + EDGPU_MAX_ION_SPECIES)));

    // Maxwell tables
    if (!build_maxwell(C->h_maxw, C->h_maxw_inj)) return fail(EDGPU_ERR_TABLE, "PrepareMaxwellDistribIntegral failed");
    TRY(dalloc(&C->maxw, 9001));
    TRY(dalloc(&C->maxw_inj, 4501));
    HIPTRY(hipMemcpy(C->maxw, C->h_maxw.data(), sizeof(double) * 9001, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(C->maxw_inj, C->h_maxw_inj.data(), sizeof(double) * 4501, hipMemcpyHostToDevice));

    // FFT_X solver (periodic x, N a power of two that fits the LDS transform)
    C->N = P.gmi - 1;
    // inner objects switch an x-periodic system to the Krylov solver, as the reference does
    // (PERIODICITY_X -> PERIODICITY_X_PETSC, pic2d_CurProblemValues.f90:460-465)
    C->fft_ok = P.periodic_x && P.n_inner_objects == 0 && C->N >= 4 && (C->N & (C->N - 1)) == 0 && C->N <= 4096;
    std::vector<double> wb, wa;
    wall_rows(P, wb, wa);
    TRY(dalloc(&C->wall_below, g.nx));
    TRY(dalloc(&C->wall_above, g.nx));
    HIPTRY(hipMemcpy(C->wall_below, wb.data(), sizeof(double) * g.nx, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(C->wall_above, wa.data(), sizeof(double) * g.nx, hipMemcpyHostToDevice));
    // a wall is dielectric when a dielectric object's first segment lies on it, and then takes
    // that object's permittivity (pic2d_Prepare_FFTX_SYSY.f90:776-795; the reference reads the
    // permittivities of objects 2 and 4, the y walls of its decks)
    double eps_below = 1.0, eps_above = 1.0;
    for (int o = 0; o < P.n_boundary_objects; ++o) {
        if (P.object_type[o] != EDGPU_OBJ_DIELECTRIC) continue;
        for (int k = 0; k < P.n_segments; ++k) {
            if (P.seg_object[k] != o) continue;
            if (P.seg_side[k] == EDGPU_SIDE_BELOW && !C->diel_below) {
                C->diel_below = true;
                eps_below = P.object_eps[o];
            } else if (P.seg_side[k] == EDGPU_SIDE_ABOVE && !C->diel_above) {
                C->diel_above = true;
                eps_above = P.object_eps[o];
            }
            break;
        }
    }
    if (g.diel_mask) {
        if (!C->fft_ok) return fail(EDGPU_ERR_ARG, "dielectric walls need the FFT_X solve (gmi-1 a power of two <= 4096)");
        TRY(dalloc(&C->sigma_below, g.nx));
        TRY(dalloc(&C->sigma_above, g.nx));
        HIPTRY(hipMemset(C->sigma_below, 0, sizeof(double) * g.nx));
        HIPTRY(hipMemset(C->sigma_above, 0, sizeof(double) * g.nx));
    }
    if (C->fft_ok) {
        const int N = C->N, nm = N / 2 + 1;
        // PREPARE_SYS_Y (pic2d_Prepare_FFTX_SYSY.f90:776-897), and the twiddles
        TRY(dalloc(&C->a_eq, (size_t)(P.gmj + 1) * nm));
        TRY(dalloc(&C->tw, N / 2));
        TRY(dalloc(&C->spec, (size_t)(P.gmj + 1) * nm));
        launch_fftx_tables(C->tw, C->a_eq, N, P.gmj, C->s, C->diel_below, C->diel_above, eps_below, eps_above);
        HIPTRY(hipGetLastError());
    }
    // Krylov: Dirichlet values
    {
        std::vector<double> dv = dirichlet_values(P, g.nx);
        TRY(dalloc(&C->dval, nn));
        HIPTRY(hipMemcpy(C->dval, dv.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    }
    // sort work
    if (P.sort_interval > 0) {
        const int32_t ncell = P.gmi * P.gmj;
        int64_t maxcap = 0;
        for (int s = 0; s < C->nsp; ++s) maxcap = std::max(maxcap, C->sp[s].cap);
        if (maxcap >= (int64_t)1 << 31) return fail(EDGPU_ERR_ARG, "sorting supports < 2^31 particles per species");
        TRY(dalloc(&C->sw.key, maxcap));
        C->sw.key_cap = maxcap;
        const int ntiles = tiles_in_x(g) * P.gmj;
        for (int s = 0; s < C->nsp; ++s) {
            TRY(dalloc(&C->sp[s].tile_start[0], ntiles + 1));
            TRY(dalloc(&C->sp[s].tile_start[1], ntiles + 1));
        }
        TRY(dalloc(&C->sw.count, ncell + 1));
        TRY(dalloc(&C->sw.offset, ncell + 1));
        TRY(dalloc(&C->sw.wcount, sort_window_counts(g)));
        C->sw.scan_tmp_bytes = sort_scan_bytes(ncell);
        if (C->sw.scan_tmp_bytes == 0) return fail(EDGPU_ERR_HIP, "hipcub scan size query failed");
        HIPTRY(hipMalloc(&C->sw.scan_tmp, C->sw.scan_tmp_bytes > 0 ? C->sw.scan_tmp_bytes : 16));
    }
    HIPTRY(hipDeviceSynchronize());
    return EDGPU_OK;
}

int edgpu_particles_upload(int32_t species, int64_t n, const double *X, const double *Y, const double *VX,
                           const double *VY, const double *VZ, const int32_t *tag) {
    TRY(species_ok(species));
    drop_sort_in_push(species);
    SpeciesState &sp = C->sp[species];
    if (n < 0) return fail(EDGPU_ERR_ARG, "negative particle count");
    if (n > 0 && (!X || !Y || !VX || !VY || !VZ)) return fail(EDGPU_ERR_ARG, "null particle array");
    sp.n = 0;
    TRY(ensure_capacity(species, n));
    TRY(upload_range(species, 0, n, X, Y, VX, VY, VZ, tag));
    sp.n = n;
    sp.coll_done = 0;
    sp.add_done = 0;
    sp.tiled = false;
    TRY(reset_vz_bound(species));
    if (species == 0) C->raw_valid_e = false;
    else C->raw_valid_i = false;
    return EDGPU_OK;
}

int edgpu_particles_download(int32_t species, int64_t capacity, int64_t *n, double *X, double *Y, double *VX,
                             double *VY, double *VZ, int32_t *tag) {
    TRY(species_ok(species));
    SpeciesState &sp = C->sp[species];
    if (n) *n = sp.n;
    if (!X) return EDGPU_OK;
    if (capacity < sp.n) return fail(EDGPU_ERR_CAPACITY, "download buffer too small");
    return download_range(species, 0, sp.n, X, Y, VX, VY, VZ, tag);
}

int edgpu_particles_upload_aos(int32_t species, int64_t n, const void *base, int64_t stride) {
    TRY(species_ok(species));
    drop_sort_in_push(species);
    SpeciesState &sp = C->sp[species];
    if (n < 0) return fail(EDGPU_ERR_ARG, "negative particle count");
    if (n > 0 && (!base || stride < (int64_t)(kAosTagOffset + sizeof(int32_t))))
        return fail(EDGPU_ERR_ARG, "null particle array or a stride below the 44 bytes of one particle");
    sp.n = 0;
    TRY(ensure_capacity(species, n));
    std::vector<double> col((size_t)5 * std::min(n, kAosChunk));
    std::vector<int32_t> tag((size_t)std::min(n, kAosChunk));
    const size_t m = col.size() / 5;
    for (int64_t o = 0; o < n; o += kAosChunk) {
        const int64_t c = std::min(kAosChunk, n - o);
        const char *rec = (const char *)base + o * stride;
        for (int64_t k = 0; k < c; ++k, rec += stride) {
            for (int q = 0; q < 5; ++q) std::memcpy(&col[q * m + k], rec + 8 * q, sizeof(double));
            std::memcpy(&tag[k], rec + kAosTagOffset, sizeof(int32_t));
        }
        TRY(upload_range(species, o, c, &col[0], &col[m], &col[2 * m], &col[3 * m], &col[4 * m], tag.data()));
    }
    sp.n = n;
    sp.coll_done = 0;
    sp.add_done = 0;
    sp.tiled = false;
    TRY(reset_vz_bound(species));
    if (species == 0) C->raw_valid_e = false;
    else C->raw_valid_i = false;
    return EDGPU_OK;
}

int edgpu_particles_download_aos(int32_t species, int64_t capacity, int64_t *n, void *base, int64_t stride) {
    TRY(species_ok(species));
    SpeciesState &sp = C->sp[species];
    if (n) *n = sp.n;
    if (!base) return EDGPU_OK;
    if (stride < (int64_t)(kAosTagOffset + sizeof(int32_t)))
        return fail(EDGPU_ERR_ARG, "stride below the 44 bytes of one particle");
    if (capacity < sp.n) return fail(EDGPU_ERR_CAPACITY, "download buffer too small");
    std::vector<double> col((size_t)5 * std::min(sp.n, kAosChunk));
    std::vector<int32_t> tag((size_t)std::min(sp.n, kAosChunk));
    const size_t m = col.size() / 5;
    for (int64_t o = 0; o < sp.n; o += kAosChunk) {
        const int64_t c = std::min(kAosChunk, sp.n - o);
        TRY(download_range(species, o, c, &col[0], &col[m], &col[2 * m], &col[3 * m], &col[4 * m], tag.data()));
        char *rec = (char *)base + o * stride;
        for (int64_t k = 0; k < c; ++k, rec += stride) {
            for (int q = 0; q < 5; ++q) std::memcpy(rec + 8 * q, &col[q * m + k], sizeof(double));
            std::memcpy(rec + kAosTagOffset, &tag[k], sizeof(int32_t));
        }
    }
    return EDGPU_OK;
}

int edgpu_particle_count(int32_t species, int64_t *n) {
    TRY(species_ok(species));
    if (n) *n = C->sp[species].n;
    return EDGPU_OK;
}

int edgpu_distribute_particles(int32_t species, int64_t n, double x0, double x1, double y0, double y1,
                               double factor_convert, int32_t copy_electron_positions) {
    TRY(species_ok(species));
    drop_sort_in_push(species);
    SpeciesState &sp = C->sp[species];
    if (n < 0) return fail(EDGPU_ERR_ARG, "negative particle count");
    if (n > sp.cap) {
        sp.n = 0;
        TRY(ensure_capacity(species, n));
    }
    const bool copy = species > 0 && copy_electron_positions;
    if (copy && n > C->sp[0].n) return fail(EDGPU_ERR_ARG, "not enough electrons to copy positions from");
    // fresh electrons placed inside an inner object are removed (pic2d_CurProblemValues.f90:3245-3258)
    const bool drop_inside = !copy && species == 0 && C->g.n_inner > 0;
    if (drop_inside) TRY(zero_counters());
    if (!mixed()) {
        Rng r = make_rng(8, 3);
        launch_distribute_fp64(sp.f, 0, n, x0, x1, y0, y1, C->g.xmax, C->g.ymax, factor_convert, C->maxw, r,
                               copy ? C->sp[0].f.x : nullptr, copy ? C->sp[0].f.y : nullptr, C->s);
        if (drop_inside) launch_mark_inside_fp64(sp.f, n, C->g, C->cnt, C->s);
    } else {
        for (int64_t o = 0; o < n; o += kChunk) {
            int64_t c = std::min(kChunk, n - o);
            Rng r = make_rng(8, 3);
            if (copy) launch_to_fp64(offset(C->sp[0].m, o), C->stage2, c, C->g, C->s);
            launch_distribute_fp64(C->stage, 0, c, x0, x1, y0, y1, C->g.xmax, C->g.ymax, factor_convert, C->maxw, r,
                                   copy ? C->stage2.x : nullptr, copy ? C->stage2.y : nullptr, C->s);
            launch_to_mixed(C->stage, offset(sp.m, o), c, C->g, C->s);
            if (drop_inside) launch_mark_inside_mixed(C->stage, offset(sp.m, o), c, C->g, C->cnt, C->s);
        }
    }
    HIPTRY(hipStreamSynchronize(C->s));
    HIPTRY(hipGetLastError());
    if (drop_inside) {
        Counters h;
        TRY(check_counters(h, "DISTRIBUTE_PARTICLES"));
        const int64_t m = (int64_t)h.n_removed;
        if (m > 0) {
            // one-off lists of the marked slots below n - m and the kept slots above it
            int64_t *holes = nullptr, *movers = nullptr;
            TRY(dalloc(&holes, (size_t)m));
            if (dalloc(&movers, (size_t)m) != EDGPU_OK) {
                (void)hipFree(holes);
                return EDGPU_ERR_HIP;
            }
            if (mixed()) launch_fill_marked_mixed(sp.m, n - m, n, holes, movers, C->cnt, C->s);
            else launch_fill_marked_fp64(sp.f, n - m, n, holes, movers, C->cnt, C->s);
            const hipError_t e = hipStreamSynchronize(C->s);
            (void)hipFree(holes);
            (void)hipFree(movers);
            HIPTRY(e);
            HIPTRY(hipGetLastError());
            n -= m;
        }
    }
    sp.n = n;
    sp.coll_done = 0;
    sp.add_done = 0;
    sp.tiled = false;
    if (species == 0) C->raw_valid_e = false;
    else C->raw_valid_i = false;
    return reset_vz_bound(species);
}

int edgpu_gather_charge_density(int32_t species, double *c_rho) {
    TRY(species_ok(species));
    const bool el = species == 0;
    double *raw = el ? C->rho_raw_e : C->rho_raw_i;
    double *rho = el ? C->rho_e : C->rho_i;
    bool &valid = el ? C->raw_valid_e : C->raw_valid_i;
    if (!valid) {
        TRY(zero_counters());
        HIPTRY(hipMemsetAsync(raw, 0, sizeof(double) * C->nn, C->s));
        if (el) deposit_range(0, 0, C->sp[0].n);
        else
            for (int s = 1; s < C->nsp; ++s) deposit_range(s, 0, C->sp[s].n);
        Counters h;
        TRY(check_counters(h, el ? "GATHER_ELECTRON_CHARGE_DENSITY" : "GATHER_ION_CHARGE_DENSITY"));
        valid = true;
    }
    HIPTRY(hipMemcpyAsync(rho, raw, sizeof(double) * C->nn, hipMemcpyDeviceToDevice, C->s));
    launch_fold_walls(rho, C->g, C->s);
    if (c_rho) HIPTRY(hipMemcpyAsync(c_rho, rho, sizeof(double) * C->nn, hipMemcpyDeviceToHost, C->s));
    TRY(sync_unless_in_step(c_rho != nullptr));
    return EDGPU_OK;
}

int edgpu_update_wall_potentials(int32_t n_obj, const double *phi) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (n_obj < 0 || n_obj > EDGPU_MAX_OBJ || (n_obj > 0 && !phi)) return fail(EDGPU_ERR_ARG, "object potentials");
    edgpu_params P = C->P;
    for (int k = 0; k < n_obj; ++k) P.object_phi[k] = phi[k];
    // rebuild wall rows and Dirichlet values with the new potentials
    const Geo &g = C->g;
    std::vector<double> wb, wa;
    wall_rows(P, wb, wa);
    // the copies go on the context stream, behind every kernel already queued that reads the
    // wall rows or the Dirichlet values, and the host vectors live until the stream has drained
    HIPTRY(hipMemcpyAsync(C->wall_below, wb.data(), sizeof(double) * g.nx, hipMemcpyHostToDevice, C->s));
    HIPTRY(hipMemcpyAsync(C->wall_above, wa.data(), sizeof(double) * g.nx, hipMemcpyHostToDevice, C->s));
    const std::vector<double> dv = dirichlet_values(P, g.nx);
    HIPTRY(hipMemcpyAsync(C->dval, dv.data(), sizeof(double) * C->nn, hipMemcpyHostToDevice, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    C->P = P;
    return EDGPU_OK;
}

int edgpu_get_surface_charge(double *below, double *above) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    const int nx = C->g.nx;
    double *out[2] = {below, above};
    const double *dev[2] = {C->sigma_below, C->sigma_above};
    for (int w = 0; w < 2; ++w) {
        if (!out[w]) continue;
        if (dev[w]) HIPTRY(hipMemcpyAsync(out[w], dev[w], sizeof(double) * nx, hipMemcpyDeviceToHost, C->s));
        else std::fill(out[w], out[w] + nx, 0.0);
    }
    HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

int edgpu_set_surface_charge(const double *below, const double *above) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (!C->g.diel_mask) return fail(EDGPU_ERR_STATE, "no dielectric object");
    const int nx = C->g.nx;
    const double *in[2] = {below, above};
    double *dev[2] = {C->sigma_below, C->sigma_above};
    std::vector<double> row((size_t)nx);
    for (int w = 0; w < 2; ++w) {
        if (!in[w]) continue;
        row.assign(in[w], in[w] + nx);
        if (C->g.periodic_x) {
            // the end nodes are periodic images of nodes gmi-1 and 1
            row[nx - 2] += row[0];
            row[1] += row[nx - 1];
            row[0] = row[nx - 1] = 0.0;
        }
        HIPTRY(hipMemcpyAsync(dev[w], row.data(), sizeof(double) * nx, hipMemcpyHostToDevice, C->s));
        HIPTRY(hipStreamSynchronize(C->s));
    }
    return EDGPU_OK;
}

int edgpu_solve_poisson_fftx(double *c_phi) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (!C->fft_ok)
        return fail(EDGPU_ERR_ARG, "FFT_X solver needs periodic x, no inner objects and gmi-1 a power of two <= 4096");
    const double factor_rho = 1.0 / (double)C->P.N_of_particles_cell;
    launch_poisson_forward(C->rho_e, C->rho_i, factor_rho, C->wall_below, C->wall_above, C->tw, C->spec, C->g, C->N,
                           C->s, C->diel_below ? C->sigma_below : nullptr, C->diel_above ? C->sigma_above : nullptr);
    launch_poisson_thomas(C->spec, C->a_eq, C->N, C->g.gmj, C->s, C->diel_below && C->diel_above);
    launch_poisson_inverse(C->spec, C->tw, C->phi, C->g, C->N, C->s);
    HIPTRY(hipGetLastError());
    if (c_phi) HIPTRY(hipMemcpyAsync(c_phi, C->phi, sizeof(double) * C->nn, hipMemcpyDeviceToHost, C->s));
    TRY(sync_unless_in_step(c_phi != nullptr));
    C->have_phi = true;
    return EDGPU_OK;
}

int edgpu_solve_potential_krylov(double rtol, int32_t max_iter, int32_t warm_start, int32_t *iterations,
                                 double *relres, double *c_phi) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (!(rtol > 0.0) || max_iter < 1) return fail(EDGPU_ERR_ARG, "rtol/max_iter");
    const int64_t nn = C->nn;
    KrylovWork &w = C->kw;
    if (!w.r) {
        TRY(dalloc(&w.r, nn));
        TRY(dalloc(&w.z, nn));
        TRY(dalloc(&w.p, nn));
        TRY(dalloc(&w.q, nn));
        TRY(dalloc(&w.b, nn));
        TRY(dalloc(&w.mask, nn));
        w.nred = 20003;
        TRY(dalloc(&w.red, 4 * (size_t)w.nred));
        if (!w.hred) HIPTRY(hipHostMalloc((void **)&w.hred, sizeof(double) * 8 + sizeof(KrylovCtl), hipHostMallocDefault));
        // two quantities of the reducing PCG kernels, or one partial per 64x4 tile of level 0
        TRY(dalloc(&w.part, std::max<size_t>(2 * (size_t)kKrylovRedGrid,
                                              (size_t)(C->g.nx / 64 + 1) * (size_t)(C->g.ny / 4 + 1))));
        TRY(dalloc(&w.ctl, 1));
        launch_krylov_setup_mask(w.mask, C->g, C->s);
        w.nlev = 0;
        if (krylov_setup_multigrid(w, C->g, C->s) < 0) return fail(EDGPU_ERR_HIP, "multigrid allocation failed");
        HIPTRY(hipStreamSynchronize(C->s));
    }
    CnnState &cn = C->cnn;
    if (warm_start < 0 || warm_start > 2) return fail(EDGPU_ERR_ARG, "warm_start must be 0, 1 or 2");
    if (warm_start == 2 && !(cn.on && cn.scaled && cn.net.t > 0))
        return fail(EDGPU_ERR_STATE, "learned warm start needs a trained network (edgpu_cnn_train)");
    if (cn.on) {
        // history for the network: phim1 = phi^{n-1} (the potential now held), phim2 = phi^{n-2}
        cn.prev_valid = C->have_phi;
        if (!C->have_phi) HIPTRY(hipMemsetAsync(C->phi, 0, sizeof(double) * nn, C->s));
        HIPTRY(hipMemcpyAsync(cn.phim2, cn.hist ? cn.phim1 : C->phi, sizeof(double) * nn, hipMemcpyDeviceToDevice, C->s));
        HIPTRY(hipMemcpyAsync(cn.phim1, C->phi, sizeof(double) * nn, hipMemcpyDeviceToDevice, C->s));
        cn.hist = true;
    }
    if (warm_start == 0 || (warm_start == 1 && !C->have_phi)) HIPTRY(hipMemsetAsync(C->phi, 0, sizeof(double) * nn, C->s));
    launch_krylov_rhs(C->rho_e, C->rho_i, -1.0 / (double)C->P.N_of_particles_cell, C->dval, w.mask, w.b, C->phi, C->g,
                      C->s);
    if (warm_start == 2) {
        // guess on the free nodes: phi^{n-1} + s_d * net(residual of phi^{n-1}, phi^{n-1} - phi^{n-2}, b, mask)
        cn.a.feat = cn.feat_own;
        launch_cnn_features_raw(w.b, cn.phim1, cn.phim2, w.mask, C->g, cn.raw_now, C->s);
        launch_cnn_pack(cn.raw_now, cn.sc, cn.a.feat, nn, C->s);
        cnn_forward(cn.net, cn.a, cn.H, cn.W, C->s);
        launch_cnn_apply(cn.a.pred, cn.phim1, w.mask, cn.sc.s_d, C->phi, nn, C->s);
        HIPTRY(hipGetLastError());
    }
    double rr = 0.0;
    int it = krylov_pcg(w, C->phi, C->g, rtol, max_iter, &rr, w.last_iters[w.precond][warm_start], C->s);
    HIPTRY(hipGetLastError());
    if (it == kKrylovHipError) return fail(EDGPU_ERR_HIP, "HIP call failed in the Krylov solve");
    if (iterations) *iterations = it < 0 ? max_iter : it;
    if (relres) *relres = rr;
    if (c_phi) TRY(copy_sync(c_phi, C->phi, sizeof(double) * nn, hipMemcpyDeviceToHost));
    C->have_phi = true;
    if (it < 0) return fail(EDGPU_ERR_NOT_CONVERGED, "Krylov solve did not reach rtol");
    return EDGPU_OK;
}

int edgpu_set_sorting_push(int32_t on) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    C->sorting_push = on != 0;
    return EDGPU_OK;
}

int edgpu_sort_with_next_push(int32_t species) {
    TRY(species_ok(species));
    if (species != 0) return fail(EDGPU_ERR_ARG, "only the electrons are sorted by their push");
    if (C->P.sort_interval <= 0) return fail(EDGPU_ERR_STATE, "sorting needs sort_interval > 0 at init");
    SpeciesState &sp = C->sp[0];
    if (sp.n <= 0) return EDGPU_OK;
    if (!sp.tiled) return sort_species(0);  // never sorted: the push has no tiles to follow yet
    TRY(count_tiles_for_push());
    return sync_unless_in_step(false);
}

int edgpu_set_krylov_preconditioner(int32_t kind) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (kind != kPrecondJacobi && kind != kPrecondMultigrid && kind != kPrecondMultigridLaunched)
        return fail(EDGPU_ERR_ARG, "preconditioner kind");
    C->kw.precond = kind;
    return EDGPU_OK;
}

int edgpu_calculate_electric_field(double *EX, double *EY) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    launch_efield(C->phi, C->EX, C->EY, C->accEX, C->accEY, C->P.factor1_E_from_F, C->P.factor2_E_from_F, C->g, C->s);
    refresh_f32_fields();
    HIPTRY(hipGetLastError());
    if (EX) HIPTRY(hipMemcpyAsync(EX, C->EX, sizeof(double) * C->nn, hipMemcpyDeviceToHost, C->s));
    if (EY) HIPTRY(hipMemcpyAsync(EY, C->EY, sizeof(double) * C->nn, hipMemcpyDeviceToHost, C->s));
    TRY(sync_unless_in_step(EX != nullptr || EY != nullptr));
    return EDGPU_OK;
}

int edgpu_clear_accumulated_fields(void) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    launch_fill(C->accEX, C->nn, 0.0, C->s);
    launch_fill(C->accEY, C->nn, 0.0, C->s);
    refresh_f32_fields();
    TRY(sync_unless_in_step(false));
    return EDGPU_OK;
}

int edgpu_get_grid(int32_t which, double *out) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    double *p = grid_ptr(which);
    if (!p || !out) return fail(EDGPU_ERR_ARG, "grid selector");
    HIPTRY(hipMemcpyAsync(out, p, sizeof(double) * C->nn, hipMemcpyDeviceToHost, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

int edgpu_set_grid(int32_t which, const double *in) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    double *p = grid_ptr(which);
    if (!p || !in) return fail(EDGPU_ERR_ARG, "grid selector");
    HIPTRY(hipMemcpyAsync(p, in, sizeof(double) * C->nn, hipMemcpyHostToDevice, C->s));
    refresh_f32_fields();
    HIPTRY(hipStreamSynchronize(C->s));
    if (which == 2) C->have_phi = true;
    return EDGPU_OK;
}

int edgpu_advance_electrons(edgpu_advance_stats *stats) {
    TRY(species_ok(0));
    if (stats) std::memset(stats, 0, sizeof(*stats));
    TRY(zero_counters());
    HIPTRY(hipMemsetAsync(C->rho_raw_e, 0, sizeof(double) * C->nn, C->s));
    TRY(do_advance(0, true, stats));
    C->raw_valid_e = true;
    return EDGPU_OK;
}

int edgpu_advance_ions(edgpu_advance_stats *stats) {
    TRY(species_ok(0));
    if (stats) std::memset(stats, 0, sizeof(*stats));
    HIPTRY(hipMemsetAsync(C->rho_raw_i, 0, sizeof(double) * C->nn, C->s));
    C->kernel_ms[1] = 0.0;
    for (int s = 1; s < C->nsp; ++s) {
        TRY(zero_counters());
        TRY(do_advance(s, false, stats));
    }
    C->raw_valid_i = true;
    return EDGPU_OK;
}

// The pushes of the calls made while a MomentsRequest lives also accumulate the moments of
// their species (the fused routines of snapshot steps); the accumulators start at zero.
struct MomentsRequest {
    int status = EDGPU_OK;
    MomentsRequest() {
        const size_t per = (size_t)kMoments * (size_t)C->nn;
        if (!C->mom_push) status = dalloc(&C->mom_push, per * (size_t)C->nsp);
        if (status == EDGPU_OK) {
            hipError_t e = hipMemsetAsync(C->mom_push, 0, sizeof(float) * per * (size_t)C->nsp, C->s);
            if (e != hipSuccess) status = fail(EDGPU_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
        }
        C->mom_request = status == EDGPU_OK;
    }
    ~MomentsRequest() { C->mom_request = false; }
};

// copies the moments of species [s0, s1) to the host (13 node arrays each, species after species)
int copy_push_moments(int s0, int s1, float *acc) {
    if (!acc || s1 <= s0) return EDGPU_OK;
    const size_t per = (size_t)kMoments * (size_t)C->nn;
    return copy_sync(acc, C->mom_push + per * (size_t)s0, sizeof(float) * per * (size_t)(s1 - s0),
                     hipMemcpyDeviceToHost);
}

int edgpu_advance_electrons_and_moments(edgpu_advance_stats *stats, float *acc) {
    TRY(species_ok(0));
    MomentsRequest req;
    TRY(req.status);
    TRY(edgpu_advance_electrons(stats));
    return copy_push_moments(0, 1, acc);
}

int edgpu_advance_ions_and_moments(edgpu_advance_stats *stats, float *acc) {
    TRY(species_ok(0));
    MomentsRequest req;
    TRY(req.status);
    TRY(edgpu_advance_ions(stats));
    return copy_push_moments(1, C->nsp, acc);
}

int edgpu_step_with_moments(double krylov_rtol, int32_t krylov_warm_start, edgpu_advance_stats *e_stats,
                            edgpu_advance_stats *i_stats, float *acc_e, float *acc_i) {
    TRY(species_ok(0));
    MomentsRequest req;
    TRY(req.status);
    TRY(edgpu_step(krylov_rtol, krylov_warm_start, e_stats, i_stats));
    TRY(copy_push_moments(0, 1, acc_e));
    // the ions were pushed when the step closed an ion cycle (n_sub back at 0)
    return C->n_sub == 0 ? copy_push_moments(1, C->nsp, acc_i) : EDGPU_OK;
}

int edgpu_get_removed(int32_t species, int64_t capacity, int64_t *n, int64_t *index, int32_t *object) {
    TRY(species_ok(species));
    SpeciesState &sp = C->sp[species];
    const int64_t m = std::min(sp.last_removed, sp.removed_cap);
    if (n) *n = m;
    if (!index && !object) return EDGPU_OK;
    if (capacity < m) return fail(EDGPU_ERR_CAPACITY, "removed-list buffer too small");
    if (index) TRY(copy_sync(index, sp.removed_idx, sizeof(int64_t) * m, hipMemcpyDeviceToHost));
    if (object) {
        TRY(copy_sync(object, sp.removed_obj, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < m; ++q) object[q] &= 0xff;  // the side is packed above the object
    }
    return EDGPU_OK;
}

int edgpu_get_removed_records(int32_t species, int64_t capacity, int64_t *n, double *x, double *y, int32_t *side) {
    TRY(species_ok(species));
    SpeciesState &sp = C->sp[species];
    const int64_t m = std::min(sp.last_removed, sp.removed_cap);
    if (n) *n = m;
    if (!x && !y && !side) return EDGPU_OK;
    if (capacity < m) return fail(EDGPU_ERR_CAPACITY, "removed-list buffer too small");
    if (x) TRY(copy_sync(x, sp.removed_x, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (y) TRY(copy_sync(y, sp.removed_y, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (side) {
        TRY(copy_sync(side, sp.removed_obj, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < m; ++q) side[q] = (side[q] >> 8) - 1;  // the object sits in the low byte
    }
    return EDGPU_OK;
}

int edgpu_process_added_particles(int32_t species) {
    TRY(species_ok(species));
    SpeciesState &sp = C->sp[species];
    HIPTRY(hipMemcpyAsync(&C->pin->u64[0], sp.add_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    const int64_t c = (int64_t)C->pin->u64[0];
    if (c == 0) return EDGPU_OK;
    drop_sort_in_push(species);
    if (c > sp.add.cap) return fail(EDGPU_ERR_CAPACITY, "add list overflowed");
    TRY(ensure_capacity(species, c));
    if (mixed()) launch_append_mixed(sp.m, sp.n, C->g, sp.add, c, C->s);
    else launch_append_fp64(sp.f, sp.n, sp.add, c, C->s);
    // the leading add_done entries (ionisation products of a fused push) join the tested part
    // when everything before them has been tested
    if (sp.add_done > 0 && sp.coll_done == sp.n) sp.coll_done = sp.n + std::min(sp.add_done, c);
    sp.add_done = 0;
    if (species > 0) {
        if (mixed()) launch_vz_bound_mixed(sp.m, sp.n, sp.n + c, C->vz2max + species, C->s);
        else launch_vz_bound_fp64(sp.f, sp.n, sp.n + c, C->vz2max + species, C->s);
    }
    const bool valid = species == 0 ? C->raw_valid_e : C->raw_valid_i;
    if (valid) {
        TRY(zero_counters());
        deposit_range(species, sp.n, sp.n + c);
    }
    HIPTRY(hipMemsetAsync(sp.add_count, 0, sizeof(unsigned long long), C->s));
    sp.n += c;
    if (C->in_step) return EDGPU_OK;  // the step checks the counters at its end
    Counters h;
    TRY(check_counters(h, "PROCESS_ADDED_PARTICLES"));
    return EDGPU_OK;
}

int edgpu_setup_electron_neutral_collisions(int32_t n_proc, const int32_t *type, const int32_t *npts,
                                            const double *energy_eV, const double *sigma_m2, const int32_t *ion_species,
                                            const double *ion_velocity_factor, double M_amu, double N_m3,
                                            double delta_t_s, int32_t n_seg, const double *seg_value,
                                            const double *seg_step, double V_scale_ms) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (n_proc < 1 || n_proc > EDGPU_MAX_PROC || n_seg < 1 || n_seg > 32 || !type || !npts || !energy_eV ||
        !sigma_m2 || !seg_value || !seg_step)
        return fail(EDGPU_ERR_ARG, "collision table description");
    EnTable &t = C->en;
    t = EnTable{};
    t.n_proc = n_proc;
    t.M_amu = M_amu;
    t.n_seg = n_seg;
    t.seg_value[0] = seg_value[0];
    for (int i = 1; i <= n_seg; ++i) {
        t.seg_value[i] = seg_value[i];
        t.seg_step[i] = seg_step[i - 1];
    }
    // energy grid, pic2d_enCollisionsGeneralProc.f90:232-281
    int nv = 0;
    t.seg_index[0] = 0;
    for (int i = 1; i <= n_seg; ++i) {
        int d = (int)((t.seg_value[i] - t.seg_value[i - 1]) / t.seg_step[i] + 0.001);
        nv += d;
        t.seg_index[i] = t.seg_index[i - 1] + d;
    }
    if (nv < 2) return fail(EDGPU_ERR_TABLE, "energy grid too small");
    t.n_values = nv;
    std::vector<double> &E = C->h_en_energy;
    E.assign(nv + 1, 0.0);
    E[0] = t.seg_value[0];
    for (int i = 1; i <= n_seg; ++i) {
        for (int j = t.seg_index[i - 1] + 1; j <= t.seg_index[i] - 1; ++j)
            E[j] = t.seg_value[i - 1] + (double)(j - t.seg_index[i - 1]) * t.seg_step[i];
        E[t.seg_index[i]] = t.seg_value[i];
    }
    // cumulative normalised probabilities, :283-321
    std::vector<const double *> Ep(n_proc), Sp(n_proc);
    int off = 0;
    for (int p = 0; p < n_proc; ++p) {
        if (npts[p] < 1) return fail(EDGPU_ERR_ARG, "cross-section table without points");
        Ep[p] = energy_eV + off;
        Sp[p] = sigma_m2 + off;
        off += npts[p];
        t.type[p] = type[p];
        if (!(type[p] == 10 || type[p] == 11 || type[p] == 20 || type[p] == 30))
            return fail(EDGPU_ERR_ARG, "unsupported collision type");
        t.thr[p] = type[p] >= 20 ? Ep[p][0] : 0.0;
        t.ivf[p] = ion_velocity_factor ? ion_velocity_factor[p] : 0.0;
        t.ion_sp[p] = ion_species ? ion_species[p] : 0;
        if (type[p] >= 30 && (t.ion_sp[p] < 0 || t.ion_sp[p] >= C->nsp - 1))
            return fail(EDGPU_ERR_ARG, "ionisation produces an unknown ion species");
    }
    std::vector<double> &Pt = C->h_en_prob;
    Pt.assign((size_t)(nv + 1) * n_proc, 0.0);
    const int Nsub = C->P.N_subcycles;
    for (int e = 0; e <= nv; ++e) {
        Pt[(size_t)e * n_proc] = en_frequency(E[e], type[0], npts[0], Ep[0], Sp[0], N_m3, Nsub, delta_t_s);
        for (int p = 1; p < n_proc; ++p)
            Pt[(size_t)e * n_proc + p] =
                Pt[(size_t)e * n_proc + p - 1] + en_frequency(E[e], type[p], npts[p], Ep[p], Sp[p], N_m3, Nsub, delta_t_s);
    }
    int emax = 0;
    for (int e = 1; e <= nv; ++e)
        if (Pt[(size_t)e * n_proc + n_proc - 1] > Pt[(size_t)emax * n_proc + n_proc - 1]) emax = e;
    const double temp = Pt[(size_t)emax * n_proc + n_proc - 1];
    if (!(temp > 0.0)) return fail(EDGPU_ERR_TABLE, "collision frequencies are all zero");
    t.pmax = 1.0 - std::exp(-temp);
    for (int e = 0; e <= nv; ++e)
        for (int p = 0; p < n_proc; ++p)
            Pt[(size_t)e * n_proc + p] = std::max(0.0, std::min(1.0, Pt[(size_t)e * n_proc + p] / temp));
    Pt[(size_t)emax * n_proc + n_proc - 1] = 1.0;
    for (int e = 0; e <= nv; ++e)
        for (int p = 0; p < n_proc - 1; ++p)
            Pt[(size_t)e * n_proc + p + 1] = std::max(Pt[(size_t)e * n_proc + p], Pt[(size_t)e * n_proc + p + 1]);
    t.V2 = V_scale_ms * V_scale_ms;
    // upper bound of the interpolated total selection probability per energy bin: the
    // interpolation is piecewise linear in E, so its maximum over a bin is at the bin ends or
    // at table energies inside; the margin covers rounding of the device evaluation
    {
        const int np = n_proc;
        auto ptot = [&](double en) {
            int idx, seg = 1;
            double a0, a1;
            if (en >= t.seg_value[t.n_seg]) {
                idx = nv - 1;
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
                idx = (int)((double)t.seg_index[seg - 1] + (en - t.seg_value[seg - 1]) / t.seg_step[seg]);
                idx = std::max(0, std::min(idx, nv - 1));
                a1 = (en - E[idx]) / t.seg_step[seg];
                a0 = 1.0 - a1;
            }
            return a0 * Pt[(size_t)idx * np + np - 1] + a1 * Pt[(size_t)(idx + 1) * np + np - 1];
        };
        for (int b = 0; b < kEnBins; ++b) {
            const int e = b / 4 - 20, sub = b % 4;
            const double lo = std::ldexp(0.5 + sub / 8.0, e), hi = std::ldexp(0.5 + (sub + 1) / 8.0, e);
            double m = std::max(ptot(lo), ptot(std::nextafter(hi, 0.0)));
            for (int q = 0; q <= nv; ++q)
                if (E[q] >= lo && E[q] < hi) m = std::max(m, ptot(E[q]));
            t.bound[b] = std::min(1.0, m * (1.0 + 1e-9) + 1e-12);
        }
    }
    if (C->en_energy) (void)hipFree(C->en_energy);
    if (C->en_prob) (void)hipFree(C->en_prob);
    C->en_energy = nullptr;
    C->en_prob = nullptr;
    TRY(dalloc(&C->en_energy, E.size()));
    TRY(dalloc(&C->en_prob, Pt.size()));
    TRY(copy_sync(C->en_energy, E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice));
    TRY(copy_sync(C->en_prob, Pt.data(), sizeof(double) * Pt.size(), hipMemcpyHostToDevice));
    t.energy = C->en_energy;
    t.prob = C->en_prob;
    if (!C->en_dev) TRY(dalloc(&C->en_dev, 1));
    TRY(copy_sync(C->en_dev, &t, sizeof(EnTable), hipMemcpyHostToDevice));
    C->en_on = true;
    return EDGPU_OK;
}

int edgpu_get_collision_bounds(int32_t *nbins, double *bound) {
    if (!C || !C->en_on) return fail(EDGPU_ERR_STATE, "no electron-neutral collision table");
    if (nbins) *nbins = kEnBins;
    if (bound) std::memcpy(bound, C->en.bound, sizeof(double) * kEnBins);
    return EDGPU_OK;
}

int edgpu_get_collision_table(int32_t *n_values, double *energy_eV, double *prob, double *max_colliding_fraction) {
    if (!C || !C->en_on) return fail(EDGPU_ERR_STATE, "no collision table");
    if (n_values) *n_values = C->en.n_values;
    if (energy_eV) std::memcpy(energy_eV, C->h_en_energy.data(), sizeof(double) * C->h_en_energy.size());
    if (prob) std::memcpy(prob, C->h_en_prob.data(), sizeof(double) * C->h_en_prob.size());
    if (max_colliding_fraction) *max_colliding_fraction = C->en.pmax;
    return EDGPU_OK;
}

int edgpu_setup_charge_exchange(int32_t species, double sigma_1eV, double alpha, double N_m3, double T_K,
                                double delta_t_s, double T_e_eV, int32_t N_max_vel, double V_scale_ms) {
    TRY(species_ok(species));
    if (species < 1) return fail(EDGPU_ERR_ARG, "charge exchange applies to ion species");
    SpeciesState &sp = C->sp[species];
    RcxPar &r = sp.rcx;
    auto sigma = [&](double e) {
        double t = 1.0 - alpha * std::log(std::max(e, 0.001));
        return sigma_1eV * (t * t);
    };
    const int Nsub = C->P.N_subcycles;
    const double Tgas_eV = T_K * kBJK / kECl;
    const double Vmean = std::sqrt(2.54647908947033 * Tgas_eV * kECl / (sp.Ms * kMeKg));
    const double pth = N_m3 * sigma(1.5 * Tgas_eV) * Vmean * delta_t_s * (double)Nsub;
    const double energy_factor_eV = 0.5 * kMeKg * (V_scale_ms * V_scale_ms) / kECl;
    r.on = 1;
    r.sigma_1eV = sigma_1eV;
    r.alpha = alpha;
    r.prob_factor = N_m3 * delta_t_s * (double)Nsub;
    r.pth2 = pth * pth;
    r.factor_eV = sp.Ms * energy_factor_eV;
    r.V_scale = V_scale_ms;
    r.vfactor = std::sqrt(T_K * kBJK / (T_e_eV * kECl * sp.Ms)) / (double)N_max_vel;
    const double c = r.prob_factor * V_scale_ms * sigma_1eV;
    r.c2 = c * c;
    return EDGPU_OK;
}

// ------------------------------------------------------------------------------------
// wall emission

int edgpu_setup_ion_induced_emission(int32_t object, int32_t species, double setD, double minE_eV, double maxE_eV,
                                     double T_eV, double T_e_eV, int32_t N_max_vel, double V_scale_ms) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (object < 0 || object >= C->g.n_boundary_obj + C->g.n_inner || species < 0 || species >= C->nsp - 1 ||
        T_e_eV <= 0.0 || N_max_vel <= 0 || V_scale_ms <= 0.0 || !(setD >= 0.0) || T_eV < 0.0)
        return fail(EDGPU_ERR_ARG, "ion-induced emission: object, species and normalisation");
    // pic2d_Materials.f90:160-180: energies made dimensionless by energy_factor_eV
    const double efac = 0.5 * kMeKg * (V_scale_ms * V_scale_ms) / kECl;
    IieeObj o{};
    o.on = setD > 0.0;
    o.setD = setD;
    o.minE = minE_eV / efac;
    o.maxE = maxE_eV / efac;
    o.vinj = std::sqrt(T_eV / T_e_eV) / (double)N_max_vel;
    C->iiee.obj[object][species] = o;
    for (int q = 0; q < C->nsp - 1; ++q) C->iiee.Ms[q] = C->P.Ms[q];
    C->iiee_on = C->reflect_mask != 0;
    for (int q = 0; q < EDGPU_MAX_OBJ; ++q)
        for (int r = 0; r < EDGPU_MAX_ION_SPECIES; ++r) C->iiee_on = C->iiee_on || C->iiee.obj[q][r].on;
    if (!C->iiee_dev) TRY(dalloc(&C->iiee_dev, 1));
    TRY(copy_sync(C->iiee_dev, &C->iiee, sizeof(IieeTable), hipMemcpyHostToDevice));
    return EDGPU_OK;
}

int edgpu_setup_ion_reflection(int32_t object, int32_t reflects) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (object < 0 || object >= C->g.n_boundary_obj + C->g.n_inner || object >= 32)
        return fail(EDGPU_ERR_ARG, "ion reflection: object");
    C->iiee.reflect[object] = reflects ? 1 : 0;
    if (reflects) C->reflect_mask |= 1u << object;
    else C->reflect_mask &= ~(1u << object);
    bool on = C->reflect_mask != 0;
    for (int q = 0; q < EDGPU_MAX_OBJ; ++q)
        for (int r = 0; r < EDGPU_MAX_ION_SPECIES; ++r) on = on || C->iiee.obj[q][r].on;
    C->iiee_on = on;
    for (int q = 0; q < C->nsp - 1; ++q) C->iiee.Ms[q] = C->P.Ms[q];
    if (!C->iiee_dev) TRY(dalloc(&C->iiee_dev, 1));
    TRY(copy_sync(C->iiee_dev, &C->iiee, sizeof(IieeTable), hipMemcpyHostToDevice));
    return EDGPU_OK;
}

int edgpu_setup_electron_induced_emission(int32_t object, const edgpu_see_params *p, double T_e_eV, int32_t N_max_vel,
                                          double V_scale_ms) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (!p || object < 0 || object >= C->g.n_boundary_obj + C->g.n_inner || T_e_eV <= 0.0 || N_max_vel <= 0 ||
        V_scale_ms <= 0.0)
        return fail(EDGPU_ERR_ARG, "electron-induced emission: object and normalisation");
    const int m[3] = {p->model_elastic, p->model_inelastic, p->model_true};
    for (int q = 0; q < 3; ++q)
        if (m[q] < 0 || m[q] > 2) return fail(EDGPU_ERR_ARG, "emission model must be 0, 1 or 2");
    // pic2d_Materials.f90:180-245: energies made dimensionless by energy_factor_eV
    const double efac = 0.5 * kMeKg * (V_scale_ms * V_scale_ms) / kECl;
    SeeObj o{};
    for (int q = 0; q < 3; ++q) o.model[q] = m[q];
    o.on = (m[0] + m[1] + m[2]) > 0;
    o.refl_type = p->elastic_reflection;
    o.setD[0] = p->setD_elastic;
    o.minE[0] = p->minE_elastic_eV / efac;
    o.maxE[0] = p->maxE_elastic_eV / efac;
    o.E_elast_0 = p->E_elast_0_eV / efac;
    o.E_elast_max = p->E_elast_max_eV / efac;
    o.maxD_elast = p->maxD_elast;
    o.dE_elast = p->dE_elast_eV / efac;
    o.frac_elast_high = m[0] == 0 ? 0.0 : p->frac_elast_highenergy;
    o.setD[1] = p->setD_inelastic;
    o.minE[1] = p->minE_inelastic_eV / efac;
    o.maxE[1] = p->maxE_inelastic_eV / efac;
    o.frac_inelastic = m[1] == 0 ? 0.0 : p->frac_inelastic;
    o.setD[2] = p->setD_true;
    o.minE[2] = p->minE_true_eV / efac;
    o.maxE[2] = p->maxE_true_eV / efac;
    o.E_see_0 = p->E_see_0_eV / efac;
    o.E_see_max = p->E_see_max_eV / efac;
    o.maxD_classic = p->maxD_see_classic;
    o.k_smooth = p->k_smooth;
    o.vinj_true = std::sqrt(p->T_true_eV / T_e_eV) / (double)N_max_vel;
    o.lowest = 1.0e6 / efac;
    if (m[0] == 1) o.lowest = std::min(o.lowest, o.minE[0]);
    if (m[0] == 2) o.lowest = std::min(o.lowest, o.E_elast_0);
    if (m[1] == 1) o.lowest = std::min(o.lowest, o.minE[1]);
    if (m[1] == 2) o.lowest = std::min(o.lowest, o.E_see_0);
    if (m[2] == 1) o.lowest = std::min(o.lowest, o.minE[2]);
    if (m[2] == 2) o.lowest = std::min(o.lowest, o.E_see_0);
    C->see.obj[object] = o;
    C->see_on = false;
    for (int q = 0; q < EDGPU_MAX_OBJ; ++q) C->see_on = C->see_on || C->see.obj[q].on;
    if (!C->see_dev) TRY(dalloc(&C->see_dev, 1));
    TRY(copy_sync(C->see_dev, &C->see, sizeof(SeeTable), hipMemcpyHostToDevice));
    return EDGPU_OK;
}

int edgpu_setup_constant_emission(int32_t object, double N_per_step, int32_t model, double Te_normal_eV,
                                  double Te_parallel_eV, double We_beam_eV, double T_e_eV, int32_t N_max_vel) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (object < 0 || object >= C->g.n_boundary_obj || T_e_eV <= 0.0 || N_max_vel <= 0 || N_per_step < 0.0 ||
        (model != 0 && model != 1) || Te_normal_eV < 0.0 || Te_parallel_eV < 0.0)
        return fail(EDGPU_ERR_ARG, "constant emission: boundary object, model and temperatures");
    // PREPARE_SETUP_VALUES, pic2d_Setup.f90:102-126
    Ctx::ConstEmit e;
    e.N = (double)(float)N_per_step;
    e.model = model;
    if (model == 0) {
        e.f_normal = std::sqrt(Te_normal_eV / T_e_eV) / (double)N_max_vel;
    } else {
        if (We_beam_eV <= 1.5 * Te_normal_eV)
            return fail(EDGPU_ERR_ARG, "beam energy must exceed 1.5 times the beam energy spread");
        e.f_normal = std::sqrt(0.25 * (Te_normal_eV * Te_normal_eV) / (We_beam_eV * T_e_eV)) / (double)N_max_vel;
        e.v_beam = std::sqrt(We_beam_eV / T_e_eV) / (double)N_max_vel;
    }
    e.f_parallel = std::sqrt(Te_parallel_eV / T_e_eV) / (double)N_max_vel;
    C->emit[object] = e;
    return EDGPU_OK;
}

int edgpu_perform_electron_emission(int64_t *emitted) {
    TRY(species_ok(0));
    const Geo &g = C->g;
    TRY(zero_counters());
    bool any = false;
    for (int o = 0; o < g.n_boundary_obj; ++o) {
        const Ctx::ConstEmit &e = C->emit[o];
        if (!(e.N > 0.0)) continue;
        // object length L (pic2d_CurProblemValues.f90:918-941): periodic systems leave out the
        // overlapping cell of horizontal segments
        int L = 0;
        for (int s = 0; s < g.n_seg; ++s) {
            if (g.seg_obj[s] != o) continue;
            const bool horiz = g.seg_side[s] == EDGPU_SIDE_BELOW || g.seg_side[s] == EDGPU_SIDE_ABOVE;
            if (g.periodic_x && horiz)
                L += std::abs(std::min(g.gmi - 1, g.seg_end[s]) - std::min(g.gmi - 1, g.seg_start[s]));
            else
                L += std::abs(g.seg_end[s] - g.seg_start[s]);
        }
        if (L <= 0) continue;
        for (int s = 0; s < g.n_seg; ++s) {
            if (g.seg_obj[s] != o) continue;
            // one wall segment (PERFORM_ELECTRON_EMISSION_SETUP, pic2d_Setup.f90:627-1001): in the
            // single cluster an X-periodic wall is a FLAT_WALL corner (the overlapping cell is
            // left out), a walled box has SURROUNDED_BY_WALL corners
            EmitPart part{};
            part.obj = o;
            part.side = g.seg_side[s];
            part.model = e.model;
            part.f_normal = e.f_normal;
            part.f_parallel = e.f_parallel;
            part.v_beam = e.v_beam;
            const bool horiz = part.side == EDGPU_SIDE_BELOW || part.side == EDGPU_SIDE_ABOVE;
            int len;
            if (horiz && g.periodic_x) {
                len = std::min(g.seg_end[s], g.gmi - 1) - g.seg_start[s];
                part.lo = 0.0;
                part.hi = (double)(g.gmi - 1);
            } else {
                len = g.seg_end[s] - g.seg_start[s];
                part.lo = 0.0;
                part.hi = horiz ? (double)g.gmi : (double)g.gmj;
            }
            part.start = (double)g.seg_start[s];
            part.len = (double)len;
            // the reference forms the count in default REAL arithmetic
            const float nf = (float)e.N * (float)len / (float)L;
            const double add = (double)nf;
            part.n_int = (int)add;
            part.frac = add - (double)part.n_int;
            launch_constant_emission(part, g, C->maxw, C->maxw_inj, make_rng(4, 4), add_of(0), C->cnt, C->s);
            any = true;
        }
    }
    if (!any) return EDGPU_OK;
    Counters h;
    TRY(check_counters(h, "PERFORM_ELECTRON_EMISSION_SETUP"));
    if (emitted)
        for (int o = 0; o < EDGPU_MAX_OBJ; ++o) emitted[o] += (int64_t)h.emit[o];
    return EDGPU_OK;
}

int edgpu_perform_electron_neutral_collisions(int64_t *counters) {
    TRY(species_ok(0));
    if (!C->en_on) return EDGPU_OK;
    SpeciesState &sp = C->sp[0];
    if (sp.coll_done >= sp.n) {  // every electron was tested in the fused push
        sp.coll_done = 0;
        return EDGPU_OK;
    }
    TRY(zero_counters());
    Rng r = make_rng(11, 1);
    const IonAddLists ai = ion_adds();
    if (mixed()) launch_mcc_en_mixed(sp.m, sp.coll_done, sp.n, C->g, C->en_dev, C->en.pmax, C->maxw, r, add_of(0), ai, C->cnt, C->s);
    else launch_mcc_en_fp64(sp.f, sp.coll_done, sp.n, C->en_dev, C->en.pmax, C->maxw, r, add_of(0), ai, C->cnt, C->s);
    Counters h;
    TRY(check_counters(h, "PERFORM_ELECTRON_NEUTRAL_COLLISIONS"));
    if (counters)
        for (int p = 0; p < C->en.n_proc; ++p) counters[p] += (int64_t)h.coll[p];
    sp.coll_done = 0;
    return EDGPU_OK;
}

int edgpu_perform_resonant_charge_exchange(int64_t *counter) {
    TRY(species_ok(0));
    for (int s = 1; s < C->nsp; ++s) {
        SpeciesState &sp = C->sp[s];
        if (!sp.rcx.on) continue;
        if (sp.coll_done >= sp.n) {  // every ion was tested in the fused push
            sp.coll_done = 0;
            continue;
        }
        TRY(zero_counters());
        Rng r = make_rng(4, 2);
        if (mixed()) launch_rcx_mixed(sp.m, sp.coll_done, sp.n, sp.rcx, C->maxw, r, C->cnt, C->vz2max + s, C->s);
        else launch_rcx_fp64(sp.f, sp.coll_done, sp.n, sp.rcx, C->maxw, r, C->cnt, C->vz2max + s, C->s);
        Counters h;
        TRY(check_counters(h, "PERFORM_RESONANT_CHARGE_EXCHANGE"));
        if (counter) *counter += (int64_t)h.coll[0];
        sp.coll_done = 0;
    }
    return EDGPU_OK;
}

int edgpu_set_uniform_stream(const double *u, int64_t n) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (C->stream) (void)hipFree(C->stream);
    C->stream = nullptr;
    C->stream_n = 0;
    if (!u || n <= 0) return EDGPU_OK;
    TRY(dalloc(&C->stream, n));
    TRY(copy_sync(C->stream, u, sizeof(double) * n, hipMemcpyHostToDevice));
    C->stream_n = n;
    return EDGPU_OK;
}

int edgpu_get_maxwell_tables(double *v, double *v_inj) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (v) std::memcpy(v, C->h_maxw.data(), sizeof(double) * 9001);
    if (v_inj) std::memcpy(v_inj, C->h_maxw_inj.data(), sizeof(double) * 4501);
    return EDGPU_OK;
}

int edgpu_collect_moments(int32_t species, float *acc) {
    TRY(species_ok(species));
    if (!acc) return fail(EDGPU_ERR_ARG, "null moments buffer");
    const int64_t nn = C->nn;
    if (!C->mom) TRY(dalloc(&C->mom, 13 * (size_t)nn));
    HIPTRY(hipMemsetAsync(C->mom, 0, sizeof(float) * 13 * (size_t)nn, C->s));
    SpeciesState &sp = C->sp[species];
    const bool el = species == 0;
    if (!mixed()) {
        launch_moments_fp64(sp.f, sp.n, C->g, C->EX, C->EY, sp.K, C->P.Ez_ext, el, C->mom, C->s);
    } else {
        for (int64_t o = 0; o < sp.n; o += kChunk) {
            int64_t c = std::min(kChunk, sp.n - o);
            launch_to_fp64(offset(sp.m, o), C->stage, c, C->g, C->s);
            launch_moments_fp64(C->stage, c, C->g, C->EX, C->EY, sp.K, C->P.Ez_ext, el, C->mom, C->s);
        }
    }
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(acc, C->mom, sizeof(float) * 13 * (size_t)nn, hipMemcpyDeviceToHost, C->s));
    HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

int edgpu_step(double krylov_rtol, int32_t krylov_warm_start, edgpu_advance_stats *e_stats,
               edgpu_advance_stats *i_stats) {
    TRY(species_ok(0));
    // the entry points called below skip their final synchronisation when they hand
    // nothing back to the host; the step synchronises once at its end
    struct InStep {
        InStep() { C->in_step = true; }
        ~InStep() { C->in_step = false; }
    } in_step;
    hipEvent_t *ev = C->ev;
    HIPTRY(hipEventRecord(ev[8], C->s));
    if (C->P.sort_interval > 0) {
        bool count_e = false;
        for (int s = 0; s < C->nsp; ++s) {
            const int interval = (s > 0 && C->P.ion_sort_interval > 0) ? C->P.ion_sort_interval : C->P.sort_interval;
            if (!C->sp[s].tiled || ++C->steps_since_sort[s] >= interval) {
                if (s == 0 && electrons_sort_in_push()) count_e = true;
                else TRY(sort_species(s));
                C->steps_since_sort[s] = 0;
            }
        }
        // after the other sorts, which use the same work arrays as the electrons' counts
        if (count_e) TRY(count_tiles_for_push());
    }
    HIPTRY(hipEventRecord(ev[9], C->s));
    HIPTRY(hipEventRecord(ev[0], C->s));
    TRY(edgpu_gather_charge_density(1, nullptr));
    TRY(edgpu_gather_charge_density(0, nullptr));
    HIPTRY(hipEventRecord(ev[1], C->s));
    if (C->P.periodic_x && C->fft_ok) {
        TRY(edgpu_solve_poisson_fftx(nullptr));
    } else {
        int32_t it = 0;
        double rr = 0.0;
        TRY(edgpu_solve_potential_krylov(krylov_rtol, 20000, krylov_warm_start, &it, &rr, nullptr));
    }
    HIPTRY(hipEventRecord(ev[2], C->s));
    TRY(edgpu_calculate_electric_field(nullptr, nullptr));
    HIPTRY(hipEventRecord(ev[3], C->s));
    // sub-cycling (pic2d_MainProgram.f90:151-306): ions, collisions and the accumulated field
    // reset on the last electron step of each ion cycle
    const bool ion_step = C->n_sub + 1 == C->P.N_subcycles;
    C->fuse_now = ion_step;
    const int adv = edgpu_advance_electrons(e_stats);
    C->fuse_now = true;
    TRY(adv);
    HIPTRY(hipEventRecord(ev[4], C->s));
    // the reference's order on an ion step (pic2d_MainProgram.f90:229-314): ADVANCE_IONS, then
    // PROCESS_ADDED_ELECTRONS (the electrons emitted by the push's wall hits; the ionisation
    // electrons of a fused collision step join as already tested), PERFORM_ELECTRON_NEUTRAL_
    // COLLISIONS over the electrons not yet tested, PERFORM_RESONANT_CHARGE_EXCHANGE,
    // PROCESS_ADDED_IONS and CLEAR_ACCUMULATED_FIELDS; on every step PERFORM_ELECTRON_EMISSION_
    // SETUP and the final PROCESS_ADDED_ELECTRONS (:353-363) close it
    int64_t cc[EDGPU_MAX_PROC] = {0};
    int64_t rc = 0;
    if (ion_step) {
        TRY(edgpu_advance_ions(i_stats));
        HIPTRY(hipEventRecord(ev[5], C->s));
        TRY(edgpu_process_added_particles(0));
        HIPTRY(hipEventRecord(ev[6], C->s));
        TRY(edgpu_perform_electron_neutral_collisions(cc));
        TRY(edgpu_perform_resonant_charge_exchange(&rc));
        HIPTRY(hipEventRecord(ev[7], C->s));
        for (int s = 1; s < C->nsp; ++s) TRY(edgpu_process_added_particles(s));
        TRY(edgpu_clear_accumulated_fields());
    } else {
        if (i_stats) std::memset(i_stats, 0, sizeof(*i_stats));
        for (int q = 5; q <= 7; ++q) HIPTRY(hipEventRecord(ev[q], C->s));
    }
    bool emits = false;
    for (int o = 0; o < C->g.n_boundary_obj; ++o) emits = emits || C->emit[o].N > 0.0;
    if (emits) TRY(edgpu_perform_electron_emission(e_stats ? e_stats->emit_count : nullptr));
    TRY(edgpu_process_added_particles(0));
    C->n_sub = (C->n_sub + 1) % C->P.N_subcycles;
    HIPTRY(hipEventRecord(ev[10], C->s));
    {
        // errors of the appends and of anything else not checked inside the step
        Counters h;
        TRY(check_counters(h, "edgpu_step"));
    }
    HIPTRY(hipEventSynchronize(ev[10]));
    float t;
    // [2] gather, [3] solve, [4] field, [5] MCC + RCX, [6] appends, emission and clearing, [7] sort
    const int map[6][2] = {{0, 1}, {1, 2}, {2, 3}, {6, 7}, {5, 6}, {8, 9}};
    for (int k = 0; k < 6; ++k) {
        HIPTRY(hipEventElapsedTime(&t, ev[map[k][0]], ev[map[k][1]]));
        C->timing[2 + k] = t;
    }
    HIPTRY(hipEventElapsedTime(&t, ev[7], ev[10]));
    C->timing[6] += t;
    C->timing[0] = C->kernel_ms[0];
    C->timing[1] = C->kernel_ms[1];
    if (e_stats)
        for (int p = 0; p < EDGPU_MAX_PROC; ++p) e_stats->collisions[p] += cc[p];
    if (i_stats) i_stats->collisions[0] += rc;
    return EDGPU_OK;
}

int edgpu_sort_particles(int32_t species) {
    TRY(species_ok(species));
    if (C->P.sort_interval <= 0) return fail(EDGPU_ERR_STATE, "sorting needs sort_interval > 0 at init");
    TRY(sort_species(species));
    HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

int edgpu_copy_probe(int64_t nbytes, int32_t repeats, double *ms) {
    if (nbytes < 16 || repeats < 1) return fail(EDGPU_ERR_ARG, "copy probe size");
    int dev = 0;
    HIPTRY(hipGetDevice(&dev));
    hipStream_t s;
    HIPTRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    void *a = nullptr, *b = nullptr;
    HIPTRY(hipMalloc(&a, nbytes));
    HIPTRY(hipMalloc(&b, nbytes));
    HIPTRY(hipMemsetAsync(a, 0, nbytes, s));
    hipEvent_t e0, e1;
    HIPTRY(hipEventCreate(&e0));
    HIPTRY(hipEventCreate(&e1));
    launch_copy_probe(a, b, nbytes, s);
    HIPTRY(hipEventRecord(e0, s));
    for (int r = 0; r < repeats; ++r) launch_copy_probe(a, b, nbytes, s);
    HIPTRY(hipEventRecord(e1, s));
    HIPTRY(hipEventSynchronize(e1));
    float t = 0.0f;
    HIPTRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipStreamDestroy(s);
    return EDGPU_OK;
}

int edgpu_mfma_probe(const float *A, const float *B, float *D) {
    if (!A || !B || !D) return fail(EDGPU_ERR_ARG, "null matrix");
    float *d = nullptr;
    HIPTRY(hipMalloc(&d, sizeof(float) * (512 + 512 + 256)));
    HIPTRY(hipMemcpy(d, A, sizeof(float) * 512, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(d + 512, B, sizeof(float) * 512, hipMemcpyHostToDevice));
    launch_mfma_probe(d, d + 512, d + 1024, nullptr);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpy(D, d + 1024, sizeof(float) * 256, hipMemcpyDeviceToHost));
    HIPTRY(hipFree(d));
    return EDGPU_OK;
}

int edgpu_cnn_init(uint64_t seed, int32_t max_pairs) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (C->cnn.on) return fail(EDGPU_ERR_STATE, "network already initialised");
    if (max_pairs < 1) return fail(EDGPU_ERR_ARG, "max_pairs must be >= 1");
    CnnState &c = C->cnn;
    c.H = C->g.ny;
    c.W = C->g.nx;
    const int64_t n = C->nn;
    CnnNet &net = c.net;
    int64_t off = 0;
    for (int l = 0; l < kCnnLayers; ++l) {
        int cin, cout;
        cnn_layer_shape(l, cin, cout);
        net.off_w[l] = off;
        off += 9 * cin * cout;
        net.off_b[l] = off;
        off += cout;
    }
    net.nparam = off;
    TRY(cnn_alloc(&net.master, off));
    TRY(cnn_alloc(&net.grad, off));
    TRY(cnn_alloc(&net.m1, off));
    TRY(cnn_alloc(&net.m2, off));
    TRY(cnn_alloc(&net.wbf, off));
    TRY(cnn_alloc(&net.frozen, off));
    for (int l = 0; l < kCnnLayers; ++l) {
        net.wb[l] = net.wbf + net.off_w[l];
        net.b[l] = net.master + net.off_b[l];
        net.gw[l] = net.grad + net.off_w[l];
        net.gb[l] = net.grad + net.off_b[l];
    }
    // He-normal weights; the last layer starts near zero (the guess starts as phi^{n-1});
    // padded input channels of layer 0 and output channels 1.. of the last layer are frozen at 0
    std::vector<float> w(off, 0.0f);
    std::vector<uint8_t> fr(off, 0);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int l = 0; l < kCnnLayers; ++l) {
        int cin, cout;
        cnn_layer_shape(l, cin, cout);
        const int cin_real = l == 0 ? 4 : cin;
        const double sd = std::sqrt(2.0 / (9.0 * cin_real)) * (l == kCnnLayers - 1 ? 0.01 : 1.0);
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < cout; ++co)
                for (int ci = 0; ci < cin; ++ci) {
                    const int64_t q = net.off_w[l] + ((int64_t)tap * cout + co) * cin + ci;
                    const bool dead = (l == 0 && ci >= 4) || (l == kCnnLayers - 1 && co >= 1);
                    fr[q] = dead ? 1 : 0;
                    w[q] = dead ? 0.0f : (float)(sd * nd(rng));
                }
        for (int co = 0; co < cout; ++co) fr[net.off_b[l] + co] = (l == kCnnLayers - 1 && co >= 1) ? 1 : 0;
    }
    TRY(copy_sync(net.master, w.data(), sizeof(float) * off, hipMemcpyHostToDevice));
    TRY(copy_sync(net.frozen, fr.data(), off, hipMemcpyHostToDevice));
    HIPTRY(hipMemset(net.m1, 0, sizeof(float) * off));
    HIPTRY(hipMemset(net.m2, 0, sizeof(float) * off));
    HIPTRY(hipMemset(net.grad, 0, sizeof(float) * off));
    cnn_sync_bf16(net, C->s);
    TRY(cnn_alloc_acts(c.a, n, true));
    c.feat_own = c.a.feat;
    c.max_pairs = max_pairs;
    TRY(cnn_alloc(&c.ring_raw, (size_t)max_pairs * n * 4));
    TRY(cnn_alloc(&c.ring_tgt, (size_t)max_pairs * n));
    TRY(cnn_alloc(&c.raw_now, (size_t)n * 4));
    TRY(cnn_alloc(&c.tgt_n, (size_t)n));
    TRY(cnn_alloc(&c.acc, 5));
    TRY(cnn_alloc(&c.phim1, n));
    TRY(cnn_alloc(&c.phim2, n));
    TRY(cnn_alloc(&c.dloss, 1));
    HIPTRY(hipStreamSynchronize(C->s));
    c.on = true;
    return EDGPU_OK;
}

int edgpu_cnn_record_pair(void) {
    if (!C || !C->cnn.on) return fail(EDGPU_ERR_STATE, "edgpu_cnn_init has not been called");
    CnnState &c = C->cnn;
    if (!c.hist || !C->have_phi) return fail(EDGPU_ERR_STATE, "no Krylov solve to record");
    if (!c.prev_valid) return fail(EDGPU_ERR_STATE, "the last Krylov solve had no previous potential to start from");
    const int64_t n = C->nn;
    const int slot = c.next;
    launch_cnn_features_raw(C->kw.b, c.phim1, c.phim2, C->kw.mask, C->g, c.ring_raw + (size_t)slot * n * 4, C->s);
    launch_cnn_target_raw(C->phi, c.phim1, C->kw.mask, c.ring_tgt + (size_t)slot * n, n, C->s);
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(C->s));
    c.next = (c.next + 1) % c.max_pairs;
    c.n_pairs = std::min(c.n_pairs + 1, c.max_pairs);
    return EDGPU_OK;
}

int edgpu_cnn_train(int32_t steps, double lr, double *mean_loss) {
    if (!C || !C->cnn.on) return fail(EDGPU_ERR_STATE, "edgpu_cnn_init has not been called");
    CnnState &c = C->cnn;
    if (c.n_pairs < 1) return fail(EDGPU_ERR_STATE, "no training pairs recorded");
    if (steps < 1 || !(lr > 0.0)) return fail(EDGPU_ERR_ARG, "steps/lr");
    const int64_t n = C->nn;
    if (!c.scaled) TRY(cnn_set_scales());
    std::mt19937 pick((uint32_t)(c.net.t + 12345));
    double acc = 0.0;
    int cnt = 0;
    c.a.feat = c.feat_own;
    for (int it = 0; it < steps; ++it) {
        const int slot = (int)(pick() % (uint32_t)c.n_pairs);
        launch_cnn_pack(c.ring_raw + (size_t)slot * n * 4, c.sc, c.a.feat, n, C->s);
        launch_cnn_scale(c.ring_tgt + (size_t)slot * n, (float)(1.0 / c.sc.s_d), c.tgt_n, n, C->s);
        cnn_forward(c.net, c.a, c.H, c.W, C->s);
        double L = 0.0;
        HIPTRY(cnn_loss(c.a, c.tgt_n, n, c.nfree, c.dloss, &L, C->s));
        HIPTRY(cnn_backward(c.net, c.a, c.H, c.W, C->s));
        cnn_adam(c.net, (float)lr, C->s);
        if (it >= steps - std::min(steps, 16)) {
            acc += L;
            ++cnt;
        }
    }
    c.a.feat = c.feat_own;
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(C->s));
    if (mean_loss) *mean_loss = acc / cnt;
    return EDGPU_OK;
}

int edgpu_cnn_scales(double *s4, int64_t *nfree) {
    if (!C || !C->cnn.on) return fail(EDGPU_ERR_STATE, "edgpu_cnn_init has not been called");
    if (s4) {
        s4[0] = C->cnn.sc.s_r;
        s4[1] = C->cnn.sc.s_i;
        s4[2] = C->cnn.sc.s_b;
        s4[3] = C->cnn.sc.s_d;
    }
    if (nfree) *nfree = C->cnn.nfree;
    return EDGPU_OK;
}

int edgpu_cnn_get_params(int64_t capacity, int64_t *n, float *out) {
    if (!C || !C->cnn.on) return fail(EDGPU_ERR_STATE, "edgpu_cnn_init has not been called");
    const int64_t m = C->cnn.net.nparam;
    if (n) *n = m;
    if (!out) return EDGPU_OK;
    if (capacity < m) return fail(EDGPU_ERR_CAPACITY, "parameter buffer too small");
    TRY(copy_sync(out, C->cnn.net.master, sizeof(float) * m, hipMemcpyDeviceToHost));
    return EDGPU_OK;
}

int edgpu_cnn_set_params(int64_t n, const float *in) {
    if (!C || !C->cnn.on) return fail(EDGPU_ERR_STATE, "edgpu_cnn_init has not been called");
    if (n != C->cnn.net.nparam || !in) return fail(EDGPU_ERR_ARG, "parameter count");
    TRY(copy_sync(C->cnn.net.master, in, sizeof(float) * n, hipMemcpyHostToDevice));
    cnn_sync_bf16(C->cnn.net, C->s);
    HIPTRY(hipStreamSynchronize(C->s));
    return EDGPU_OK;
}

int edgpu_cnn_eval(int32_t H, int32_t W, const float *features, const float *target, float *pred, float *grad,
                   double *loss) {
    if (!C || !C->cnn.on) return fail(EDGPU_ERR_STATE, "edgpu_cnn_init has not been called");
    if (H < 1 || W < 1 || !features || !target) return fail(EDGPU_ERR_ARG, "eval grid");
    const int64_t n = (int64_t)H * W;
    std::vector<uint16_t> f16((size_t)n * 16, 0);
    int64_t nfree = 0;
    for (int64_t q = 0; q < n; ++q) {
        for (int ch = 0; ch < 4; ++ch) f16[q * 16 + ch] = host_bf16(features[q * 4 + ch]);
        if (features[q * 4 + 3] > 0.5f) ++nfree;
    }
    if (nfree == 0) return fail(EDGPU_ERR_ARG, "no free nodes in the eval features");
    CnnActs a{};
    const size_t mark = C->cnn.owned.size();
    TRY(cnn_alloc_acts(a, n, true));
    float *t = nullptr;
    TRY(cnn_alloc(&t, n));
    TRY(copy_sync(a.feat, f16.data(), sizeof(uint16_t) * f16.size(), hipMemcpyHostToDevice));
    TRY(copy_sync(t, target, sizeof(float) * n, hipMemcpyHostToDevice));
    cnn_forward(C->cnn.net, a, H, W, C->s);
    double L = 0.0;
    HIPTRY(cnn_loss(a, t, n, nfree, C->cnn.dloss, &L, C->s));
    HIPTRY(cnn_backward(C->cnn.net, a, H, W, C->s));
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(C->s));
    if (pred) TRY(copy_sync(pred, a.pred, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (grad) TRY(copy_sync(grad, C->cnn.net.grad, sizeof(float) * C->cnn.net.nparam, hipMemcpyDeviceToHost));
    if (loss) *loss = L;
    for (size_t k = mark; k < C->cnn.owned.size(); ++k) (void)hipFree(C->cnn.owned[k]);
    C->cnn.owned.resize(mark);
    return EDGPU_OK;
}

int edgpu_last_step_timing(double *ms8) {
    if (!C) return fail(EDGPU_ERR_STATE, "edgpu_init has not been called");
    if (ms8) std::memcpy(ms8, C->timing, sizeof(C->timing));
    return EDGPU_OK;
}

}  // extern "C"
