"""Python host binding of libedgpu.so, the MI355X implementation of the EDIPIC-2D
particle hot path.

`Runtime` mirrors the reference's per-step subroutines (same names, lower case, same
argument meaning); a non-zero status of the C ABI raises `EdgpuError`, the analogue of
the reference's MPI_ABORT.  The shared library must have been built (edgpu.build); the
binding refuses to run without it -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import config

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EDGPU_LIB") or os.path.join(PKG, "libedgpu.so")

MAX_SEG, MAX_INNER, MAX_OBJ, MAX_ION, MAX_PROC = 16, 8, 24, 4, 8
FP64, MIXED = 0, 1
OBJ_METAL, OBJ_DIELECTRIC, OBJ_VACUUM_GAP = 0, 1, 2

ERRORS = {1: "invalid argument", 2: "HIP failure", 3: "particle out of bounds", 4: "index out of bounds",
          5: "no wall object part", 6: "capacity exceeded", 7: "collision table inconsistency",
          8: "state error", 9: "Krylov solve not converged", 10: "emission yields above one"}


class EdgpuError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__("%s failed with status %d (%s): %s" % (where, code, ERRORS.get(code, "?"), msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("gmi", C.c_int32), ("gmj", C.c_int32), ("periodic_x", C.c_int32), ("N_of_particles_cell", C.c_int32),
        ("N_subcycles", C.c_int32), ("precision", C.c_int32),
        ("Bx_ext", C.c_double), ("By_ext", C.c_double), ("Bz_ext", C.c_double), ("Ez_ext", C.c_double),
        ("factor1_E_from_F", C.c_double), ("factor2_E_from_F", C.c_double),
        ("n_boundary_objects", C.c_int32), ("n_segments", C.c_int32),
        ("seg_side", C.c_int32 * MAX_SEG), ("seg_start", C.c_int32 * MAX_SEG), ("seg_end", C.c_int32 * MAX_SEG),
        ("seg_object", C.c_int32 * MAX_SEG),
        ("n_inner_objects", C.c_int32), ("inner_box", (C.c_int32 * 4) * MAX_INNER),
        ("object_phi", C.c_double * MAX_OBJ),
        ("n_ion_species", C.c_int32), ("Qs", C.c_int32 * MAX_ION), ("QM2s", C.c_double * MAX_ION),
        ("Ms", C.c_double * MAX_ION), ("ions_sense_magnetic_field", C.c_int32), ("ions_sense_Ez", C.c_int32),
        ("max_electrons", C.c_int64), ("max_ions", C.c_int64 * MAX_ION), ("add_capacity", C.c_int64),
        ("seed", C.c_uint64), ("fuse_collisions", C.c_int32), ("sort_interval", C.c_int32),
        ("ion_sort_interval", C.c_int32),
        ("object_type", C.c_int32 * MAX_OBJ), ("object_eps", C.c_double * MAX_OBJ),
    ]


class AdvanceStats(C.Structure):
    _fields_ = [("n_before", C.c_int64), ("n_after", C.c_int64), ("hit_count", C.c_int64 * MAX_OBJ),
                ("n_removed", C.c_int64), ("collisions", C.c_int64 * MAX_PROC), ("emit_count", C.c_int64 * MAX_OBJ)]

    def as_dict(self):
        return dict(n_before=self.n_before, n_after=self.n_after, hit_count=list(self.hit_count),
                    n_removed=self.n_removed, collisions=list(self.collisions), emit_count=list(self.emit_count))


class SeeParams(C.Structure):
    """edgpu_see_params: the init_material_*.dat values of one boundary object (eV)."""
    _fields_ = [("model_elastic", C.c_int32), ("elastic_reflection", C.c_int32), ("model_inelastic", C.c_int32),
                ("model_true", C.c_int32)] + [(n, C.c_double) for n in (
                    "setD_elastic", "minE_elastic_eV", "maxE_elastic_eV", "E_elast_0_eV", "E_elast_max_eV",
                    "maxD_elast", "dE_elast_eV", "frac_elast_highenergy", "setD_inelastic", "minE_inelastic_eV",
                    "maxE_inelastic_eV", "frac_inelastic", "T_true_eV", "setD_true", "minE_true_eV", "maxE_true_eV",
                    "E_see_0_eV", "E_see_max_eV", "maxD_see_classic", "k_smooth")]


_lib = None
D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
I32 = C.POINTER(C.c_int32)
I64 = C.POINTER(C.c_int64)
U32 = C.POINTER(C.c_uint32)


def lib():
    """Load libedgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libedgpu.so is not built (run __graft_entry__.build() or edgpu/build.py)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    sig = {
        "edgpu_init": [C.POINTER(Params)],
        "edgpu_finalize": [],
        "edgpu_device": [],
        "edgpu_particles_upload": [C.c_int32, C.c_int64, D, D, D, D, D, I32],
        "edgpu_particles_download": [C.c_int32, C.c_int64, I64, D, D, D, D, D, I32],
        "edgpu_particle_count": [C.c_int32, I64],
        "edgpu_particles_upload_aos": [C.c_int32, C.c_int64, C.c_void_p, C.c_int64],
        "edgpu_particles_download_aos": [C.c_int32, C.c_int64, I64, C.c_void_p, C.c_int64],
        "edgpu_distribute_particles": [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int32],
        "edgpu_gather_charge_density": [C.c_int32, D],
        "edgpu_update_wall_potentials": [C.c_int32, D],
        "edgpu_solve_poisson_fftx": [D],
        "edgpu_get_surface_charge": [D, D],
        "edgpu_set_surface_charge": [D, D],
        "edgpu_solve_potential_krylov": [C.c_double, C.c_int32, C.c_int32, I32, D, D],
        "edgpu_set_krylov_preconditioner": [C.c_int32],
        "edgpu_calculate_electric_field": [D, D],
        "edgpu_clear_accumulated_fields": [],
        "edgpu_get_grid": [C.c_int32, D],
        "edgpu_set_grid": [C.c_int32, D],
        "edgpu_advance_electrons": [C.POINTER(AdvanceStats)],
        "edgpu_advance_ions": [C.POINTER(AdvanceStats)],
        "edgpu_get_removed": [C.c_int32, C.c_int64, I64, I64, I32],
        "edgpu_process_added_particles": [C.c_int32],
        "edgpu_get_removed_records": [C.c_int32, C.c_int64, I64, D, D, I32],
        "edgpu_setup_electron_neutral_collisions": [C.c_int32, I32, I32, D, D, I32, D, C.c_double, C.c_double,
                                                    C.c_double, C.c_int32, D, D, C.c_double],
        "edgpu_get_collision_table": [I32, D, D, D],
        "edgpu_get_collision_bounds": [I32, D],
        "edgpu_setup_charge_exchange": [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int32, C.c_double],
        "edgpu_perform_electron_neutral_collisions": [I64],
        "edgpu_setup_electron_induced_emission": [C.c_int32, C.POINTER(SeeParams), C.c_double, C.c_int32, C.c_double],
        "edgpu_setup_ion_induced_emission": [C.c_int32, C.c_int32] + [C.c_double] * 5 + [C.c_int32, C.c_double],
        "edgpu_setup_ion_reflection": [C.c_int32, C.c_int32],
        "edgpu_setup_constant_emission": [C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_int32],
        "edgpu_perform_electron_emission": [I64],
        "edgpu_get_rng_state": [C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), I32],
        "edgpu_get_vz_bound": [C.c_int32, D],
        "edgpu_set_rng_state": [C.c_uint64, C.c_uint32, C.c_int32],
        "edgpu_perform_resonant_charge_exchange": [I64],
        "edgpu_set_uniform_stream": [D, C.c_int64],
        "edgpu_get_maxwell_tables": [D, D],
        "edgpu_collect_moments": [C.c_int32, F],
        "edgpu_advance_electrons_and_moments": [C.POINTER(AdvanceStats), F],
        "edgpu_advance_ions_and_moments": [C.POINTER(AdvanceStats), F],
        "edgpu_step_with_moments": [C.c_double, C.c_int32, C.POINTER(AdvanceStats), C.POINTER(AdvanceStats), F, F],
        "edgpu_step": [C.c_double, C.c_int32, C.POINTER(AdvanceStats), C.POINTER(AdvanceStats)],
        "edgpu_last_step_timing": [D],
        "edgpu_copy_probe": [C.c_int64, C.c_int32, D],
        "edgpu_sort_particles": [C.c_int32],
        "edgpu_set_sorting_push": [C.c_int32],
        "edgpu_sort_with_next_push": [C.c_int32],
        "edgpu_cnn_init": [C.c_uint64, C.c_int32],
        "edgpu_cnn_record_pair": [],
        "edgpu_cnn_train": [C.c_int32, C.c_double, D],
        "edgpu_cnn_get_params": [C.c_int64, I64, F],
        "edgpu_cnn_scales": [D, I64],
        "edgpu_cnn_set_params": [C.c_int64, F],
        "edgpu_cnn_eval": [C.c_int32, C.c_int32, F, F, F, F, D],
        "edgpu_mfma_probe": [F, F, F],
        "edgpu_solve_poisson_fftx_noalloc": [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p],
        "edgpu_noalloc_scratch_bytes": [I64],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.edgpu_last_error.argtypes = []
    L.edgpu_last_error.restype = C.c_char_p
    L.edgpu_philox4x32_10.argtypes = [U32, U32, U32]
    L.edgpu_philox4x32_10.restype = None
    L.edgpu_philox2x32_10.argtypes = [U32, C.c_uint32, U32]
    L.edgpu_philox2x32_10.restype = None


def _dp(a):
    return a.ctypes.data_as(D) if a is not None else None


def _check(rc, where):
    if rc != 0:
        raise EdgpuError(rc, where, lib().edgpu_last_error().decode(errors="replace"))


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().edgpu_philox4x32_10(c, k, o)
    return list(o)


def philox2x32_10(ctr, key):
    c = (C.c_uint32 * 2)(*ctr)
    o = (C.c_uint32 * 2)()
    lib().edgpu_philox2x32_10(c, key, o)
    return list(o)


def mfma_probe(A, B):
    """D (16x16) = A (16x32) B (32x16), BF16 inputs, FP32 result, through one MFMA instruction."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    if A.shape != (16, 32) or B.shape != (32, 16):
        raise ValueError("mfma_probe takes A 16x32 and B 32x16")
    D = np.zeros((16, 16), np.float32)
    _check(lib().edgpu_mfma_probe(A.ctypes.data_as(F), B.ctypes.data_as(F), D.ctypes.data_as(F)),
           "edgpu_mfma_probe")
    return D


def copy_probe(nbytes, repeats=10):
    """Average ms of a plain copy kernel moving nbytes (bandwidth / counter calibration)."""
    ms = C.c_double()
    _check(lib().edgpu_copy_probe(int(nbytes), int(repeats), C.byref(ms)), "edgpu_copy_probe")
    return ms.value


def make_params(prob, precision=FP64, capacity_factor=1.25, seed=20261203, fuse_collisions=1, sort_interval=0,
                add_capacity=None, ion_sort_interval=0):
    """edgpu_params for a config.Problem."""
    p = Params()
    p.gmi, p.gmj = prob.gmi, prob.gmj
    p.periodic_x = 1 if prob.periodic_x else 0
    p.N_of_particles_cell = prob.N_of_particles_cell
    p.N_subcycles = prob.N_subcycles
    p.precision = precision
    p.Bx_ext, p.By_ext, p.Bz_ext, p.Ez_ext = prob.Bx, prob.By, prob.Bz, prob.Ez
    p.factor1_E_from_F, p.factor2_E_from_F = prob.factor1_E_from_F, prob.factor2_E_from_F
    p.n_boundary_objects = prob.n_boundary_obj
    p.n_segments = len(prob.segments)
    for k, (side, s, e, o) in enumerate(prob.segments):
        p.seg_side[k], p.seg_start[k], p.seg_end[k], p.seg_object[k] = side, s, e, o
    p.n_inner_objects = len(prob.inner_boxes)
    for k, b in enumerate(prob.inner_boxes):
        for q in range(4):
            p.inner_box[k][q] = b[q]
    for k, ph in enumerate(prob.object_phi):
        p.object_phi[k] = ph
    for k in range(MAX_OBJ):
        p.object_eps[k] = 1.0
    for k, eps in enumerate(getattr(prob, "dielectric_eps", [None] * prob.n_boundary_obj)):
        if eps is not None:
            p.object_type[k] = OBJ_DIELECTRIC
            p.object_eps[k] = eps
    for k in getattr(prob, "vacuum_gaps", ()):
        p.object_type[k] = OBJ_VACUUM_GAP
    species = getattr(prob, "ion_species", [(prob.Qs, prob.QM2s, prob.Ms)])
    if not 1 <= len(species) <= MAX_ION:
        raise ValueError("between 1 and %d ion species" % MAX_ION)
    p.n_ion_species = len(species)
    for s, (q, qm, ms) in enumerate(species):
        p.Qs[s], p.QM2s[s], p.Ms[s] = q, qm, ms
    p.ions_sense_magnetic_field = prob.ions_sense_magnetic_field
    p.ions_sense_Ez = getattr(prob, "ions_sense_Ez", 1)
    cap = max(1024, int(prob.n_particles * capacity_factor))
    p.max_electrons = cap
    for s in range(len(species)):
        p.max_ions[s] = cap
    p.add_capacity = add_capacity if add_capacity else max(1 << 16, cap // 64)
    p.seed = seed
    p.fuse_collisions = fuse_collisions
    p.sort_interval = sort_interval
    p.ion_sort_interval = ion_sort_interval
    return p


class Runtime:
    """One device context; methods are the reference entry points the hot path replaces."""

    def __init__(self, prob, params=None, **kw):
        self.prob = prob
        self.params = params if params is not None else make_params(prob, **kw)
        _check(lib().edgpu_init(C.byref(self.params)), "edgpu_init")
        self.nn = prob.nx * prob.ny

    def close(self):
        lib().edgpu_finalize()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- particles
    def upload(self, species, X, Y, VX, VY, VZ, tag=None):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (X, Y, VX, VY, VZ)]
        t = None if tag is None else np.ascontiguousarray(tag, np.int32)
        _check(lib().edgpu_particles_upload(species, arrs[0].size, *[_dp(a) for a in arrs],
                                            t.ctypes.data_as(I32) if t is not None else None),
               "edgpu_particles_upload")

    # the reference's TYPE particle as a numpy record (pic2d_Modules.f90:432-439), padded to 48 B
    PARTICLE = np.dtype([("X", "f8"), ("Y", "f8"), ("VX", "f8"), ("VY", "f8"), ("VZ", "f8"), ("tag", "i4"),
                         ("pad", "i4")])

    def upload_aos(self, species, rec):
        """Upload an array of particle records (any dtype with the fields X..VZ, tag at the
        reference's byte offsets)."""
        rec = np.ascontiguousarray(rec)
        _check(lib().edgpu_particles_upload_aos(species, rec.size, rec.ctypes.data, rec.dtype.itemsize),
               "edgpu_particles_upload_aos")

    def download_aos(self, species, dtype=None):
        dtype = np.dtype(dtype or self.PARTICLE)
        n = self.count(species)
        rec = np.zeros(n, dtype)
        m = C.c_int64()
        _check(lib().edgpu_particles_download_aos(species, n, C.byref(m), rec.ctypes.data, dtype.itemsize),
               "edgpu_particles_download_aos")
        return rec

    def sort_particles(self, species):
        _check(lib().edgpu_sort_particles(species), "edgpu_sort_particles")

    def set_sorting_push(self, on):
        """Sort the electrons of a step's due sort with the step's electron push (default on)."""
        _check(lib().edgpu_set_sorting_push(int(bool(on))), "edgpu_set_sorting_push")

    def sort_with_next_push(self):
        """Count the electrons per tile now; the next electron push writes them in that order."""
        _check(lib().edgpu_sort_with_next_push(0), "edgpu_sort_with_next_push")

    def count(self, species):
        n = C.c_int64()
        _check(lib().edgpu_particle_count(species, C.byref(n)), "edgpu_particle_count")
        return n.value

    def download(self, species):
        n = self.count(species)
        out = [np.empty(n, np.float64) for _ in range(5)]
        tag = np.empty(n, np.int32)
        m = C.c_int64()
        _check(lib().edgpu_particles_download(species, n, C.byref(m), *[_dp(a) for a in out],
                                              tag.ctypes.data_as(I32)), "edgpu_particles_download")
        return out + [tag]

    def distribute_particles(self, species, n, x_limits, y_limits, factor_convert, copy_electron_positions=False):
        _check(lib().edgpu_distribute_particles(species, n, x_limits[0], x_limits[1], y_limits[0], y_limits[1],
                                                factor_convert, 1 if copy_electron_positions else 0),
               "edgpu_distribute_particles")

    def setup_wall_emission(self):
        """Electron-induced and constant emission as the problem describes them."""
        we = self.prob.wall_emission
        for obj, m in we.get("see", {}).items():
            self.setup_electron_induced_emission(obj, **m)
        for obj, c in we.get("constant", {}).items():
            self.setup_constant_emission(obj, **c)
        for obj, per_species in we.get("iiee", {}).items():
            for s, m in enumerate(per_species):
                self.setup_ion_induced_emission(obj, s, **m)
        for obj in we.get("ion_reflection", []):
            self.setup_ion_reflection(obj)

    def load_initial_particles(self):
        """DISTRIBUTE_PARTICLES on the device: electrons (less those placed inside inner
        objects), then as many ions on the electron positions."""
        pr = self.prob
        self.distribute_particles(0, pr.n_particles, pr.x_limits, pr.y_limits, pr.factor_convert_e)
        self.distribute_particles(1, self.count(0), pr.x_limits, pr.y_limits, pr.factor_convert_i, True)

    # ---- grids
    def _grid_out(self, host):
        return np.empty(self.nn, np.float64) if host else None

    def gather_electron_charge_density(self, host=True):
        out = self._grid_out(host)
        _check(lib().edgpu_gather_charge_density(0, _dp(out)), "GATHER_ELECTRON_CHARGE_DENSITY")
        return out

    def gather_ion_charge_density(self, host=True):
        out = self._grid_out(host)
        _check(lib().edgpu_gather_charge_density(1, _dp(out)), "GATHER_ION_CHARGE_DENSITY")
        return out

    def update_wall_potentials(self, phi):
        a = np.ascontiguousarray(phi, np.float64)
        _check(lib().edgpu_update_wall_potentials(a.size, _dp(a)), "UPDATE_WALL_POTENTIALS")

    def get_surface_charge(self):
        """Surface charge of the dielectric parts of the walls below and above (gmi+1 nodes
        each, macroparticle charges)."""
        nx = self.prob.gmi + 1
        below, above = np.zeros(nx), np.zeros(nx)
        _check(lib().edgpu_get_surface_charge(_dp(below), _dp(above)), "edgpu_get_surface_charge")
        return below, above

    def set_surface_charge(self, below=None, above=None):
        nx = self.prob.gmi + 1
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (below, above)]
        for a in arrs:
            if a is not None and a.size != nx:
                raise ValueError("surface charge rows have gmi+1 nodes")
        _check(lib().edgpu_set_surface_charge(*[None if a is None else _dp(a) for a in arrs]),
               "edgpu_set_surface_charge")

    def solve_poisson_fftx(self, host=True):
        out = self._grid_out(host)
        _check(lib().edgpu_solve_poisson_fftx(_dp(out)), "SOLVE_POISSON_FFTX_LINSYSY")
        return out

    def solve_potential_krylov(self, rtol=1e-8, max_iter=20000, warm_start=0, host=True):
        out = self._grid_out(host)
        it = C.c_int32()
        rr = C.c_double()
        _check(lib().edgpu_solve_potential_krylov(rtol, max_iter, warm_start, C.byref(it), C.byref(rr), _dp(out)),
               "SOLVE_POTENTIAL_WITH_PETSC")
        return out, it.value, rr.value

    def set_krylov_preconditioner(self, kind):
        """0 Jacobi, 1 multigrid (default), 2 multigrid with every level launched separately."""
        _check(lib().edgpu_set_krylov_preconditioner(int(kind)), "edgpu_set_krylov_preconditioner")

    def calculate_electric_field(self, host=True):
        EX, EY = self._grid_out(host), self._grid_out(host)
        _check(lib().edgpu_calculate_electric_field(_dp(EX), _dp(EY)), "CALCULATE_ELECTRIC_FIELD")
        return EX, EY

    def clear_accumulated_fields(self):
        _check(lib().edgpu_clear_accumulated_fields(), "CLEAR_ACCUMULATED_FIELDS")

    GRIDS = dict(rho_e=0, rho_i=1, phi=2, EX=3, EY=4, accEX=5, accEY=6)

    def get_grid(self, name):
        out = np.empty(self.nn, np.float64)
        _check(lib().edgpu_get_grid(self.GRIDS[name], _dp(out)), "edgpu_get_grid")
        return out

    def set_grid(self, name, values):
        a = np.ascontiguousarray(values, np.float64)
        assert a.size == self.nn
        _check(lib().edgpu_set_grid(self.GRIDS[name], _dp(a)), "edgpu_set_grid")

    # ---- push
    def advance_electrons(self):
        st = AdvanceStats()
        _check(lib().edgpu_advance_electrons(C.byref(st)), "ADVANCE_ELECTRONS")
        return st.as_dict()

    def advance_ions(self):
        st = AdvanceStats()
        _check(lib().edgpu_advance_ions(C.byref(st)), "ADVANCE_IONS")
        return st.as_dict()

    def removed(self, species):
        n = C.c_int64()
        _check(lib().edgpu_get_removed(species, 0, C.byref(n), None, None), "edgpu_get_removed")
        idx = np.empty(n.value, np.int64)
        obj = np.empty(n.value, np.int32)
        _check(lib().edgpu_get_removed(species, n.value, C.byref(n), idx.ctypes.data_as(I64),
                                       obj.ctypes.data_as(I32)), "edgpu_get_removed")
        return idx, obj

    def removed_records(self, species):
        """Positions and wall sides of the last push's removals, in the order of removed()."""
        n = C.c_int64()
        _check(lib().edgpu_get_removed_records(species, 0, C.byref(n), None, None, None), "edgpu_get_removed_records")
        x, y = np.empty(n.value), np.empty(n.value)
        side = np.empty(n.value, np.int32)
        _check(lib().edgpu_get_removed_records(species, n.value, C.byref(n), _dp(x), _dp(y),
                                               side.ctypes.data_as(I32)), "edgpu_get_removed_records")
        return x, y, side

    def process_added_electrons(self):
        _check(lib().edgpu_process_added_particles(0), "PROCESS_ADDED_ELECTRONS")

    def process_added_ions(self, species=None):
        """PROCESS_ADDED_IONS of one ion species (1-based as the particle arrays), or of all."""
        for s in ([species] if species else range(1, self.params.n_ion_species + 1)):
            _check(lib().edgpu_process_added_particles(s), "PROCESS_ADDED_IONS")

    # ---- collisions
    def initiate_electron_neutral_collisions(self, processes=None, segments=config.ENERGY_SEGMENTS, M_amu=None):
        pr = self.prob
        procs = processes if processes is not None else pr.cross_sections()
        self._nproc = len(procs)
        types = np.array([p["type"] for p in procs], np.int32)
        npts = np.array([len(p["E"]) for p in procs], np.int32)
        E = np.concatenate([np.asarray(p["E"], np.float64) for p in procs])
        S = np.concatenate([np.asarray(p["sigma"], np.float64) for p in procs])
        isp = np.array([p.get("ion_species", 0) for p in procs], np.int32)
        ivf = np.array([pr.ion_velocity_factor() if p["type"] >= 30 else 0.0 for p in procs], np.float64)
        sv = np.array([segments[0]] + [s[0] for s in segments[1]], np.float64)
        ss = np.array([s[1] for s in segments[1]], np.float64)
        _check(lib().edgpu_setup_electron_neutral_collisions(
            len(procs), types.ctypes.data_as(I32), npts.ctypes.data_as(I32), _dp(E), _dp(S), isp.ctypes.data_as(I32),
            _dp(ivf), M_amu if M_amu is not None else pr.M_i_amu, pr.neutral_N_m3, pr.dt, len(ss), _dp(sv), _dp(ss),
            pr.V_scale), "INITIATE_ELECTRON_NEUTRAL_COLLISIONS")

    def collision_bounds(self):
        n = C.c_int32()
        _check(lib().edgpu_get_collision_bounds(C.byref(n), None), "edgpu_get_collision_bounds")
        out = np.empty(n.value)
        _check(lib().edgpu_get_collision_bounds(C.byref(n), _dp(out)), "edgpu_get_collision_bounds")
        return out

    def initiate_ion_neutral_collisions(self):
        pr = self.prob
        s1, alpha = pr.rcx_sigma()
        _check(lib().edgpu_setup_charge_exchange(1, s1, alpha, pr.neutral_N_m3, pr.neutral_T_K, pr.dt, pr.T_e_eV,
                                                 pr.N_max_vel, pr.V_scale), "INITIATE_ION_NEUTRAL_COLLISIONS")

    def setup_electron_induced_emission(self, obj, **kw):
        """Material of boundary object `obj` (0-based): SeeParams fields by name (eV)."""
        sp = SeeParams()
        for k, v in kw.items():
            if not hasattr(sp, k):
                raise TypeError("unknown emission parameter %r" % k)
            setattr(sp, k, v)
        pr = self.prob
        _check(lib().edgpu_setup_electron_induced_emission(obj, C.byref(sp), pr.T_e_eV, pr.N_max_vel, pr.V_scale),
               "edgpu_setup_electron_induced_emission")

    def setup_ion_induced_emission(self, obj, species=0, setD=0.0, minE_eV=0.0, maxE_eV=0.0, T_eV=0.0):
        """Ion-induced electron emission of object `obj` (0-based) for ion species `species`."""
        pr = self.prob
        _check(lib().edgpu_setup_ion_induced_emission(obj, species, setD, minE_eV, maxE_eV, T_eV, pr.T_e_eV,
                                                      pr.N_max_vel, pr.V_scale), "edgpu_setup_ion_induced_emission")

    def setup_ion_reflection(self, obj, reflects=True):
        """Object `obj` (0-based) reflects every ion it absorbs (ion-wall model 1)."""
        _check(lib().edgpu_setup_ion_reflection(obj, 1 if reflects else 0), "edgpu_setup_ion_reflection")

    def setup_constant_emission(self, obj, n_per_step, model=0, Te_normal_eV=0.0, Te_parallel_eV=0.0, We_beam_eV=0.0):
        pr = self.prob
        _check(lib().edgpu_setup_constant_emission(obj, n_per_step, model, Te_normal_eV, Te_parallel_eV, We_beam_eV,
                                                   pr.T_e_eV, pr.N_max_vel), "edgpu_setup_constant_emission")

    def vz_bound(self, species):
        v = C.c_double()
        _check(lib().edgpu_get_vz_bound(species, C.byref(v)), "edgpu_get_vz_bound")
        return v.value

    def rng_state(self):
        seed, ctr, nsub = C.c_uint64(), C.c_uint32(), C.c_int32()
        _check(lib().edgpu_get_rng_state(C.byref(seed), C.byref(ctr), C.byref(nsub)), "edgpu_get_rng_state")
        return seed.value, ctr.value, nsub.value

    def set_rng_state(self, seed, counter, n_sub=0):
        _check(lib().edgpu_set_rng_state(seed, counter, n_sub), "edgpu_set_rng_state")

    def perform_electron_emission(self):
        """PERFORM_ELECTRON_EMISSION_SETUP: electrons emitted per object."""
        out = np.zeros(MAX_OBJ, np.int64)
        _check(lib().edgpu_perform_electron_emission(out.ctypes.data_as(I64)), "PERFORM_ELECTRON_EMISSION_SETUP")
        return out

    def collision_table(self):
        n = C.c_int32()
        pm = C.c_double()
        _check(lib().edgpu_get_collision_table(C.byref(n), None, None, C.byref(pm)), "edgpu_get_collision_table")
        E = np.empty(n.value + 1)
        P = np.empty((n.value + 1) * MAX_PROC)
        _check(lib().edgpu_get_collision_table(C.byref(n), _dp(E), _dp(P), C.byref(pm)), "edgpu_get_collision_table")
        nproc = self._nproc
        return E, P[:(n.value + 1) * nproc].reshape(n.value + 1, nproc), pm.value

    def perform_electron_neutral_collisions(self):
        c = np.zeros(MAX_PROC, np.int64)
        _check(lib().edgpu_perform_electron_neutral_collisions(c.ctypes.data_as(I64)),
               "PERFORM_ELECTRON_NEUTRAL_COLLISIONS")
        return c

    def perform_resonant_charge_exchange(self):
        c = np.zeros(1, np.int64)
        _check(lib().edgpu_perform_resonant_charge_exchange(c.ctypes.data_as(I64)),
               "PERFORM_RESONANT_CHARGE_EXCHANGE")
        return int(c[0])

    def set_uniform_stream(self, u):
        if u is None:
            _check(lib().edgpu_set_uniform_stream(None, 0), "edgpu_set_uniform_stream")
            return
        self._stream = np.ascontiguousarray(u, np.float64)
        _check(lib().edgpu_set_uniform_stream(_dp(self._stream), self._stream.size), "edgpu_set_uniform_stream")

    def maxwell_tables(self):
        v = np.empty(9001)
        vi = np.empty(4501)
        _check(lib().edgpu_get_maxwell_tables(_dp(v), _dp(vi)), "edgpu_get_maxwell_tables")
        return v, vi

    # ---- diagnostics and the whole step
    def collect_moments(self, species):
        acc = np.empty(13 * self.nn, np.float32)
        _check(lib().edgpu_collect_moments(species, acc.ctypes.data_as(F)), "COLLECT_MOMENTS")
        return acc.reshape(13, self.nn)

    def advance_electrons_and_moments(self):
        """ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS_2D: the push statistics and the 13 raw
        moment arrays summed in the same pass (half-step velocity, positions before the step)."""
        st = AdvanceStats()
        acc = np.empty(13 * self.nn, np.float32)
        _check(lib().edgpu_advance_electrons_and_moments(C.byref(st), acc.ctypes.data_as(F)),
               "ADVANCE_ELECTRONS_AND_CALCULATE_MOMENTS_2D")
        return st.as_dict(), acc.reshape(13, self.nn)

    def advance_ions_and_moments(self):
        """ADVANCE_IONS_AND_CALCULATE_MOMENTS_2D of every ion species: statistics and the raw
        moments, shape (species, 13, nodes)."""
        st = AdvanceStats()
        ns = self.params.n_ion_species
        acc = np.empty(ns * 13 * self.nn, np.float32)
        _check(lib().edgpu_advance_ions_and_moments(C.byref(st), acc.ctypes.data_as(F)),
               "ADVANCE_IONS_AND_CALCULATE_MOMENTS_2D")
        return st.as_dict(), acc.reshape(ns, 13, self.nn)

    def step(self, krylov_rtol=1e-8, krylov_warm_start=1):
        self._update_electrodes()
        se, si = AdvanceStats(), AdvanceStats()
        _check(lib().edgpu_step(krylov_rtol, krylov_warm_start, C.byref(se), C.byref(si)), "edgpu_step")
        return se.as_dict(), si.as_dict()

    def _update_electrodes(self):
        """UPDATE_WALL_POTENTIALS(T_cntr) of a problem with time-dependent electrodes, before the
        step's solve; T_cntr counts the steps this runtime has made (T_cntr_start sets its
        first value)."""
        wp = getattr(self.prob, "wall_potentials", None)
        t = getattr(self, "T_cntr", getattr(self.prob, "T_cntr_start", 0))
        if wp is not None:
            self.update_wall_potentials(wp.phi(t))
        self.T_cntr = t + 1

    def step_with_moments(self, krylov_rtol=1e-8, krylov_warm_start=1):
        """A snapshot step: edgpu_step with the fused pushes.  Returns the statistics and the raw
        moments of the electrons (13, nodes) and of the ion species (species, 13, nodes; None
        on a step that does not push the ions)."""
        self._update_electrodes()
        se, si = AdvanceStats(), AdvanceStats()
        ns = self.params.n_ion_species
        acc_e = np.empty(13 * self.nn, np.float32)
        acc_i = np.full(ns * 13 * self.nn, np.nan, np.float32)
        _check(lib().edgpu_step_with_moments(krylov_rtol, krylov_warm_start, C.byref(se), C.byref(si),
                                             acc_e.ctypes.data_as(F), acc_i.ctypes.data_as(F)),
               "edgpu_step_with_moments")
        mi = None if np.isnan(acc_i[0]) else acc_i.reshape(ns, 13, self.nn)
        return se.as_dict(), si.as_dict(), acc_e.reshape(13, self.nn), mi

    def last_step_timing(self):
        t = np.zeros(8)
        _check(lib().edgpu_last_step_timing(_dp(t)), "edgpu_last_step_timing")
        return dict(zip(["push_e", "push_i", "gather", "solve", "efield", "mcc", "append", "sort"], t.tolist()))

    # ---- learned warm start of the Krylov solve (no reference counterpart)
    def cnn_init(self, seed=1, max_pairs=64):
        _check(lib().edgpu_cnn_init(int(seed), int(max_pairs)), "edgpu_cnn_init")

    def cnn_record_pair(self):
        _check(lib().edgpu_cnn_record_pair(), "edgpu_cnn_record_pair")

    def cnn_train(self, steps, lr=1e-3):
        loss = C.c_double()
        _check(lib().edgpu_cnn_train(int(steps), float(lr), C.byref(loss)), "edgpu_cnn_train")
        return loss.value

    def cnn_scales(self):
        s = np.zeros(4)
        n = C.c_int64()
        _check(lib().edgpu_cnn_scales(_dp(s), C.byref(n)), "edgpu_cnn_scales")
        return dict(residual=s[0], increment=s[1], rhs=s[2], target=s[3], nfree=n.value)

    def cnn_params(self):
        n = C.c_int64()
        _check(lib().edgpu_cnn_get_params(0, C.byref(n), None), "edgpu_cnn_get_params")
        out = np.empty(n.value, np.float32)
        _check(lib().edgpu_cnn_get_params(n.value, C.byref(n), out.ctypes.data_as(F)), "edgpu_cnn_get_params")
        return out

    def cnn_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().edgpu_cnn_set_params(p.size, p.ctypes.data_as(F)), "edgpu_cnn_set_params")

    def cnn_eval(self, features, target):
        """features (H, W, 4) and target (H, W): prediction (H, W), gradient, loss."""
        f = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        H, W = t.shape
        pred = np.empty((H, W), np.float32)
        grad = np.empty(self.cnn_params().size, np.float32)
        loss = C.c_double()
        _check(lib().edgpu_cnn_eval(H, W, f.ctypes.data_as(F), t.ctypes.data_as(F), pred.ctypes.data_as(F),
                                    grad.ctypes.data_as(F), C.byref(loss)), "edgpu_cnn_eval")
        return pred, grad, loss.value
