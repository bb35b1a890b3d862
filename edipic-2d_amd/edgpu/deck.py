"""Reader of the reference's input decks (the init_*.dat family), so that a run of the
device path starts from the reference's own files.

Each file is read in the order and with the fixed formats of the reference's readers:
  init_configuration.dat  INITIATE_PARAMETERS, pic2d_CurProblemValues.f90:99-267
  init_simcontrol.dat     pic2d_CurProblemValues.f90:830-850
  init_particles.dat      pic2d_CurProblemValues.f90:1005-1030
  init_extfields.dat      PREPARE_EXTERNAL_FIELDS, pic2d_ExternalFields.f90:39-126
  init_bo_NN.dat          PREPARE_SETUP_VALUES, pic2d_Setup.f90:57-126
  init_material_*.dat     pic2d_Materials.f90:70-180
  init_neutrals.dat, init_neutral_AAAAAA.dat, init_neutral_AAAAAA_crsect_coll_id_NN_type_TT.dat
                          INITIATE_ELECTRON_NEUTRAL_COLLISIONS, pic2d_enCollisionsGeneralProc.f90:57-198
  init_neutral_AAAAAA_rcx_param.dat
                          INITIATE_ION_NEUTRAL_COLLISIONS, pic2d_inCollisionsGeneralProc.f90:40-70
Every header line is skipped as the reference's READ (9, '(A1)') does.  What the device
path does not model is rejected with a message: object types other than metal walls,
X-periodic boundaries and metal inner boxes, more than one ion or neutral species, and
wall sections of the boundary not covered by objects.
"""
import math
import glob
import os
import re

from . import config
from .wallpotentials import WallPotentials

# object types, pic2d_Modules.f90
VACUUM_GAP, METAL_WALL, PERIODIC_PIPELINE_X, PERIODIC_PIPELINE_Y, DIELECTRIC, SYMMETRY_PLANE = 0, 1, 2, 3, 4, 5


class DeckError(ValueError):
    pass


def fortran_read(line, fmt):
    """Values of one record read with a Fortran edit list of nX, Iw, Fw.d, Ew.d and Aw
    descriptors (columns taken as the format gives them; a blank field reads as zero)."""
    out, pos = [], 0
    for item in re.findall(r"(\d*)([xXiIfFeEaA])(\d*)(?:\.\d+)?", fmt):
        rep, kind, width = item
        kind = kind.lower()
        if kind == "x":
            pos += int(rep or 1)
            continue
        for _ in range(int(rep or 1)):
            w = int(width)
            field = line[pos:pos + w]
            pos += w
            s = field.strip()
            if kind == "i":
                out.append(int(s) if s else 0)
            elif kind in "fe":
                out.append(float(s.replace("d", "e").replace("D", "e")) if s else 0.0)
            else:
                out.append(field)
    return out


class _File:
    def __init__(self, path):
        if not os.path.exists(path):
            raise DeckError("%s not found" % path)
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.k = 0
        self.path = path

    def header(self):
        self._next()

    def read(self, fmt):
        return fortran_read(self._next(), fmt)

    def _next(self):
        if self.k >= len(self.lines):
            raise DeckError("%s: unexpected end of file" % self.path)
        line = self.lines[self.k]
        self.k += 1
        return line


def _read_configuration(d):
    f = _File(os.path.join(d, "init_configuration.dat"))
    c = {}
    f.header(); c["T_e_eV"], = f.read("(3x,f12.7)")
    f.header(); c["N_plasma_m3"], = f.read("(3x,e14.7)")
    f.header(); c["N_of_cells_debye"], = f.read("(3x,i3)")
    f.header(); c["N_max_vel"], = f.read("(3x,i3)")
    f.header(); c["N_blocks_x"], = f.read("(3x,i3)")
    f.header(); c["N_blocks_y"], = f.read("(3x,i3)")
    f.header(); c["N_grid_block_x"], = f.read("(3x,i3)")
    f.header(); c["N_grid_block_y"], = f.read("(3x,i3)")
    f.header(); c["N_of_particles_cell"], = f.read("(2x,i4)")
    f.header(); c["cluster_N_blocks_x"], = f.read("(5x,i1)")
    f.header(); c["cluster_N_blocks_y"], = f.read("(5x,i1)")
    f.header(); nb, ni = f.read("(3x,i3,3x,i3)")
    objs = []
    for _ in range(nb):
        f.header(); otype, nseg = f.read("(3x,i2,3x,i2)")
        f.header()
        segs = []
        for _ in range(nseg):
            i0, j0, i1, j1 = f.read("(3x,i6,3x,i6,3x,i6,3x,i6)")
            # endpoints ordered left-right / bottom-top (pic2d_CurProblemValues.f90:182-200)
            if i0 == i1:
                j0, j1 = min(j0, j1), max(j0, j1)
            elif j0 == j1:
                i0, i1 = min(i0, i1), max(i0, i1)
            else:
                raise DeckError("segment (%d,%d)-(%d,%d) is neither vertical nor horizontal" % (i0, j0, i1, j1))
            segs.append((i0, j0, i1, j1))
        objs.append(dict(type=otype, segments=segs))
    inner = []
    for _ in range(ni):
        f.header(); otype, = f.read("(3x,i2)")
        f.header(); il, jb, ir, jt = f.read("(3x,i6,3x,i6,3x,i6,3x,i6)")
        inner.append(dict(type=otype, box=(il, ir, jb, jt)))
    c["boundary_objects"], c["inner_objects"] = objs, inner
    return c


def _read_simcontrol(d):
    f = _File(os.path.join(d, "init_simcontrol.dat"))
    s = {}
    f.header(); s["t_sim_ns"], = f.read("(3x,f11.3)")
    f.header(); s["N_subcycles"], = f.read("(8x,i2)")
    f.header(); s["dT_cluster_load_balance"], = f.read("(6x,i4)")
    f.header(); s["dT_global_load_balance"], = f.read("(6x,i4)")
    f.header(); s["dT_save_checkpoint"], = f.read("(3x,i7)")
    f.header(); s["use_checkpoint"], = f.read("(9x,i1)")
    f.header(); s["T_cntr_to_continue"], = f.read("(2x,i8)")
    return s


def _read_particles(d):
    f = _File(os.path.join(d, "init_particles.dat"))
    p = {}
    f.header(); p["init_Te_eV"], = f.read("(3x,f8.3)")
    f.header(); p["init_Ne_m3"], = f.read("(3x,e10.3)")
    f.header(); n_spec, = f.read("(6x,i1)")
    f.header()
    p["species"] = [dict(zip(("M_i_amu", "Qs", "init_Ti_eV", "init_NiNe"), f.read("(3x,f5.1,3x,i2,3x,f8.3,3x,f6.2)")))
                    for _ in range(n_spec)]
    f.header(); p["seed"], = f.read("(3x,i9)")
    return p


def _read_extfields(d):
    path = os.path.join(d, "init_extfields.dat")
    if not os.path.exists(path):
        return dict(Bx_G=0.0, By_G=0.0, Bz_G=0.0, Ez_Vcm=0.0, ions_sense_B=0, ions_sense_Ez=0)
    f = _File(path)
    e = {}
    f.header(); e["Bx_G"], = f.read("(3x,f9.3)")
    f.header(); e["By_G"], = f.read("(3x,f9.3)")
    f.header()
    f.header(); f.read("(3x,f9.3)")  # y of the Bz maximum
    f.header(); f.read("(3x,f9.3)")  # Bz at y = 0
    # the non-uniform profile is disabled in the reference: Bz_ext = Bz at y_Bmax
    f.header(); e["Bz_G"], = f.read("(3x,f9.3)")
    f.header(); f.read("(3x,f9.3)")
    f.header(); f.read("(3x,f9.3)")
    f.header(); f.read("(3x,f9.3)")
    f.header(); e["Ez_Vcm"], = f.read("(3x,f9.3)")
    f.header(); e["ions_sense_B"], = f.read("(7x,i1)")
    f.header(); e["ions_sense_Ez"], = f.read("(7x,i1)")
    return e


def _read_bo(d, n):
    """init_bo_NN.dat of object n (1-based); defaults when absent (pic2d_Setup.f90:40-49,64-69)."""
    b = dict(material="      ", phi_V=0.0, N_emit=0.0, model=0, Te_normal_eV=0.0, Te_parallel_eV=0.0,
             We_beam_eV=0.0)
    path = os.path.join(d, "init_bo_%02d.dat" % n)
    if not os.path.exists(path):
        return b
    f = _File(path)
    f.header(); b["material"], = f.read("(6x,A6)")
    f.header(); b["phi_V"], = f.read("(3x,f9.3)")
    f.header(); b["phi_var_V"], = f.read("(3x,f9.3)")
    f.header(); b["omega_MHz"], = f.read("(3x,f9.3)")
    f.header(); b["phase_deg"], = f.read("(3x,f9.3)")
    f.header(); b["N_emit"], = f.read("(3x,f9.3)")
    f.header(); b["model"], = f.read("(7x,i1)")
    f.header(); b["Te_normal_eV"], = f.read("(4x,f8.3)")
    f.header(); b["Te_parallel_eV"], = f.read("(4x,f8.3)")
    f.header(); b["We_beam_eV"], = f.read("(4x,f8.3)")
    return b


def _read_material(d, name, n_spec=1):
    """init_material_AAAAAA.dat, or None (pic2d_Materials.f90:63-180); eV as in the file."""
    path = os.path.join(d, "init_material_%s.dat" % name)
    if not os.path.exists(path):
        return None
    f = _File(path)
    m = {}
    f.header(); m["eps_diel"], = f.read("(6x,f6.3)")  # whole_object%eps_diel, pic2d_Materials.f90:82
    f.header(); m["model_elastic"], = f.read("(7x,i1)")
    f.header(); m["elastic_reflection"], = f.read("(7x,i1)")
    f.header()
    f.header(); m["setD_elastic"], = f.read("(7x,f5.3)")
    f.header(); m["minE_elastic_eV"], = f.read("(2x,f10.3)")
    f.header(); m["maxE_elastic_eV"], = f.read("(2x,f10.3)")
    f.header()
    f.header(); m["E_elast_0_eV"], = f.read("(2x,f10.3)")
    f.header(); m["E_elast_max_eV"], = f.read("(2x,f10.3)")
    f.header(); m["maxD_elast"], = f.read("(7x,f5.3)")
    f.header(); m["dE_elast_eV"], = f.read("(2x,f10.3)")
    f.header(); m["frac_elast_highenergy"], = f.read("(7x,f5.3)")
    f.header(); m["model_inelastic"], = f.read("(7x,i1)")
    f.header()
    f.header(); m["setD_inelastic"], = f.read("(7x,f5.3)")
    f.header(); m["minE_inelastic_eV"], = f.read("(2x,f10.3)")
    f.header(); m["maxE_inelastic_eV"], = f.read("(2x,f10.3)")
    f.header()
    f.header(); m["frac_inelastic"], = f.read("(7x,f5.3)")
    f.header(); m["model_true"], = f.read("(7x,i1)")
    f.header(); m["T_true_eV"], = f.read("(2x,f10.3)")
    f.header()
    f.header(); m["setD_true"], = f.read("(7x,f5.3)")
    f.header(); m["minE_true_eV"], = f.read("(2x,f10.3)")
    f.header(); m["maxE_true_eV"], = f.read("(2x,f10.3)")
    f.header()
    f.header(); m["E_see_0_eV"], = f.read("(2x,f10.3)")
    f.header(); m["E_see_max_eV"], = f.read("(2x,f10.3)")
    f.header(); m["maxD_see_classic"], = f.read("(6x,f6.3)")
    f.header(); m["k_smooth"], = f.read("(7x,f5.3)")
    # ion-material interaction: 0 adsorbs every ion, 1 reflects them, 2 ion-induced emission
    # (pic2d_Materials.f90:151-170)
    try:
        f.header(); m["ion_wall_interaction"], = f.read("(7x,i1)")
    except DeckError:
        m["ion_wall_interaction"] = 0
    m["ion_wall_interaction"] = max(0, min(m["ion_wall_interaction"], 2))
    if m["ion_wall_interaction"] == 2:
        # ion-induced emission parameters per ion species (pic2d_Materials.f90:171-180)
        m["iiee"] = []
        for _ in range(n_spec):
            f.header()
            q = {}
            f.header(); q["setD"], = f.read("(7x,f5.3)")
            f.header(); q["minE_eV"], = f.read("(2x,f10.3)")
            f.header(); q["maxE_eV"], = f.read("(2x,f10.3)")
            f.header(); q["T_eV"], = f.read("(2x,f10.3)")
            m["iiee"].append(q)
    return m


def _read_neutrals(d):
    path = os.path.join(d, "init_neutrals.dat")
    if not os.path.exists(path):
        return []
    f = _File(path)
    f.header(); n_spec, = f.read("(10x,i2)")
    out = []
    for _ in range(n_spec):
        g = {}
        f.header(); g["name"], = f.read("(6x,A6)")
        f.header(); g["N_m3"], = f.read("(2x,e10.3)")
        f.header(); g["T_K"], = f.read("(5x,f7.1)")
        out.append(g)
    for g in out:
        h = _File(os.path.join(d, "init_neutral_%s.dat" % g["name"]))
        h.header(); g["M_amu"], = h.read("(3x,f7.3)")
        h.header(); nproc, = h.read("(8x,i2)")
        procs = []
        for p in range(1, nproc + 1):
            h.header()
            ptype, active, ion_sp, _ = h.read("(3x,i2,2x,i1,2x,i2,2x,i1)")
            procs.append(dict(id=p, type=ptype, active=active, ion_species=ion_sp))
        h.header(); nseg, = h.read("(8x,i2)")
        h.header(); e0, = h.read("(2x,f9.3)")
        segs = []
        for _ in range(nseg):
            h.header()
            segs.append(tuple(h.read("(2x,f9.3,3x,f7.3)")))
        g["segments"] = (e0, segs)
        g["processes"] = []
        for p in procs:
            if not p["active"]:
                continue
            c = _File(os.path.join(d, "init_neutral_%s_crsect_coll_id_%02d_type_%02d.dat" % (g["name"], p["id"],
                                                                                       p["type"])))
            npts, = c.read("(2x,i4)")
            E, S = [], []
            for _ in range(npts):
                e, s = c.read("(4x,f9.3,2x,e10.3)")
                E.append(e)
                S.append(s)
            g["processes"].append(dict(type=p["type"], E=E, sigma=S, ion_species=max(p["ion_species"] - 1, 0)))
        rcx = os.path.join(d, "init_neutral_%s_rcx_param.dat" % g["name"])
        g["rcx"] = None
        if os.path.exists(rcx):
            with open(rcx) as fr:
                lines = [ln for ln in fr.read().splitlines() if ln.strip()]
            s, sigma, alpha = lines[1].split()[:3]
            if int(s) > 0:
                g["rcx"] = dict(ion_species=int(s), sigma_1eV=float(sigma.replace("d", "e")),
                                alpha=float(alpha.replace("d", "e")))
    return out


class Deck:
    """The parsed deck; problem() gives the workload, apply(runtime) the collision and
    emission set-up."""

    def __init__(self, directory):
        self.dir = directory
        self.conf = _read_configuration(directory)
        self.sim = _read_simcontrol(directory)
        self.parts = _read_particles(directory)
        self.ext = _read_extfields(directory)
        c = self.conf
        self.n_boundary = len(c["boundary_objects"])
        self.bo = [_read_bo(directory, n + 1) for n in range(self.n_boundary + len(c["inner_objects"]))]
        self.materials = {}
        for b in self.bo:
            name = b["material"]
            if name.strip() and name not in self.materials:
                self.materials[name] = _read_material(directory, name, len(self.parts["species"]))
        self.neutrals = _read_neutrals(directory)
        if len(self.parts["species"]) != 1:
            raise DeckError("the device path runs one ion species (the deck has %d)" % len(self.parts["species"]))
        if len(self.neutrals) > 1:
            raise DeckError("the device path runs one neutral species (the deck has %d)" % len(self.neutrals))

    def problem(self, name=None):
        c, p, e = self.conf, self.parts, self.ext
        sp = p["species"][0]
        gmi = c["N_blocks_x"] * c["N_grid_block_x"] + 1
        gmj = c["N_blocks_y"] * c["N_grid_block_y"] + 1
        if os.path.exists(os.path.join(self.dir, "init_ext_circuit.dat")):
            raise DeckError("the external circuit (init_ext_circuit.dat) is not modelled on the device")
        types = [o["type"] for o in c["boundary_objects"]]
        for t in types:
            if t not in (METAL_WALL, PERIODIC_PIPELINE_X, DIELECTRIC, VACUUM_GAP):
                raise DeckError("boundary object type %d is not modelled on the device" % t)
        for o in c["inner_objects"]:
            if o["type"] != METAL_WALL:
                raise DeckError("inner object type %d is not modelled on the device" % o["type"])
        periodic = PERIODIC_PIPELINE_X in types
        gas = self.neutrals[0] if self.neutrals else None
        prob = config.Problem(
            T_e_eV=c["T_e_eV"], N_plasma_m3=c["N_plasma_m3"], N_of_cells_debye=c["N_of_cells_debye"],
            N_max_vel=c["N_max_vel"], N_of_particles_cell=c["N_of_particles_cell"], gmi=gmi, gmj=gmj,
            periodic_x=periodic, init_Te_eV=p["init_Te_eV"], init_Ti_eV=sp["init_Ti_eV"], M_i_amu=sp["M_i_amu"],
            Qs=sp["Qs"], N_subcycles=self.sim["N_subcycles"], init_Ne_m3=p["init_Ne_m3"],
            neutral_N_m3=gas["N_m3"] if gas else 0.0, neutral_T_K=gas["T_K"] if gas else 300.0,
            collisions=gas is not None, ions_sense_magnetic_field=e["ions_sense_B"], name=name or "deck")
        # uniform external fields (Gauss, V/cm in the file), pic2d_ExternalFields.f90:71-101
        prob.Bx = e["Bx_G"] * 1.0e-4 / prob.B_scale
        prob.By = e["By_G"] * 1.0e-4 / prob.B_scale
        prob.Bz = e["Bz_G"] * 1.0e-4 / prob.B_scale
        prob.Ez = e["Ez_Vcm"] * 100.0 / prob.E_scale
        prob.ions_sense_Ez = e["ions_sense_Ez"]
        # walls: one device segment per wall segment; periodic objects keep their numbers
        segs = []
        for n, o in enumerate(c["boundary_objects"]):
            if o["type"] == PERIODIC_PIPELINE_X:
                continue
            for (i0, j0, i1, j1) in o["segments"]:
                if j0 == j1:
                    side = config.SIDE_BELOW if j0 == 0 else config.SIDE_ABOVE
                    if j0 not in (0, gmj):
                        raise DeckError("horizontal segment at j=%d is not on the boundary" % j0)
                    segs.append((side, i0, i1, n))
                else:
                    side = config.SIDE_LEFT if i0 == 0 else config.SIDE_RIGHT
                    if i0 not in (0, gmi):
                        raise DeckError("vertical segment at i=%d is not on the boundary" % i0)
                    segs.append((side, j0, j1, n))
        prob.segments = segs
        prob.n_boundary_obj = self.n_boundary
        prob.inner_boxes = [o["box"] for o in c["inner_objects"]]
        prob.object_phi = [b["phi_V"] / prob.F_scale for b in self.bo]
        # electrodes whose potential follows an oscillation, a waveform or an amplitude profile
        metal = {n for n, o in enumerate(c["boundary_objects"]) if o["type"] == METAL_WALL}
        metal |= {self.n_boundary + m for m in range(len(c["inner_objects"]))}
        wp = WallPotentials(self.dir, self.bo, metal, prob.object_phi, prob.dt, prob.F_scale)
        if wp.varies:
            prob.wall_potentials = wp
        # dielectric walls take the permittivity of their material (1 without a material file,
        # pic2d_Materials.f90:45,82)
        eps = []
        for n, o in enumerate(c["boundary_objects"]):
            if o["type"] != DIELECTRIC:
                eps.append(None)
                continue
            m = self.materials.get(self.bo[n]["material"]) if n < len(self.bo) else None
            eps.append(m["eps_diel"] if m else 1.0)
        if any(v is not None for v in eps):
            prob.dielectric_eps = eps
        gaps = [n for n, o in enumerate(c["boundary_objects"]) if o["type"] == VACUUM_GAP]
        if gaps:
            prob.vacuum_gaps = gaps
        if gas is not None:
            prob.deck_processes = gas["processes"]
            prob.deck_segments = gas["segments"]
            prob.deck_neutral_M_amu = gas["M_amu"]
            prob.deck_rcx = gas["rcx"]
        # emission described by init_bo_NN.dat and the materials
        see, const, iiee = {}, {}, {}
        for n, b in enumerate(self.bo):
            m = self.materials.get(b["material"])
            if m and (m["model_elastic"] or m["model_inelastic"] or m["model_true"]):
                see[n] = {k: v for k, v in m.items() if k not in ("eps_diel", "ion_wall_interaction", "iiee")}
            if m and m.get("iiee"):
                iiee[n] = m["iiee"]
            if b["N_emit"] > 0.0:
                const[n] = dict(n_per_step=b["N_emit"], model=b["model"], Te_normal_eV=b["Te_normal_eV"],
                                Te_parallel_eV=b["Te_parallel_eV"], We_beam_eV=b["We_beam_eV"])
        reflect = [n for n, b in enumerate(self.bo)
                   if self.materials.get(b["material"]) and self.materials[b["material"]]["ion_wall_interaction"] == 1]
        prob.wall_emission = {"see": see, "constant": const, "iiee": iiee, "ion_reflection": reflect}
        prob.seed = self.parts["seed"]
        return prob

    def apply(self, rt):
        """Collisions and wall emission of the deck on a runtime built from problem()."""
        prob = rt.prob
        if getattr(prob, "deck_processes", None):
            rt.initiate_electron_neutral_collisions(processes=prob.deck_processes, segments=prob.deck_segments,
                                                    M_amu=prob.deck_neutral_M_amu)
        if getattr(prob, "deck_rcx", None):
            r = prob.deck_rcx
            prob.rcx_sigma = lambda: (r["sigma_1eV"], r["alpha"])
            rt.initiate_ion_neutral_collisions()
        rt.setup_wall_emission()


def read_deck(directory):
    return Deck(directory)
