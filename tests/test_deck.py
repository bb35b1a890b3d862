"""Input-deck reader (CPU): the Fortran fixed-format field reads, the reference's sample
deck for a periodic system with an electron beam (a data fixture copied from the
reference's input_data_sample_periodic_ebeam) read value by value, the workload built
from it, and -- when the reference tree is present -- every shipped deck either read or
rejected with the feature the device path does not model."""
import glob
import os

import pytest

from edgpu import config
from edgpu import deck as D

HERE = os.path.dirname(os.path.abspath(__file__))
EBEAM = os.path.join(HERE, "golden", "deck_periodic_ebeam")
SEE = os.path.join(HERE, "golden", "deck_periodic_see")


def test_fortran_fixed_format_fields():
    assert D.fortran_read("     10.0000000", "(3x,f12.7)") == [10.0]
    assert D.fortran_read("    1.0000000E+17", "(3x,e14.7)") == [1.0e17]
    assert D.fortran_read("        0      481      513      481", "(3x,i6,3x,i6,3x,i6,3x,i6)") == [0, 481, 513, 481]
    assert D.fortran_read("    2    1", "(3x,i2,3x,i2)") == [2, 1]
    assert D.fortran_read("      METAL0", "(6x,A6)") == ["METAL0"]
    assert D.fortran_read("   10  1   0  0", "(3x,i2,2x,i1,2x,i2,2x,i1)") == [10, 1, 0, 0]
    # columns, not tokens: a value beyond the field width is not read, a blank field is 0
    assert D.fortran_read("     2     0", "(3x,i3)") == [2]
    assert D.fortran_read("", "(3x,i3)") == [0]


def test_sample_ebeam_deck_values():
    d = D.read_deck(EBEAM)
    c = d.conf
    assert (c["T_e_eV"], c["N_plasma_m3"], c["N_of_cells_debye"], c["N_max_vel"]) == (10.0, 1.0e17, 2, 10)
    assert (c["N_blocks_x"], c["N_blocks_y"], c["N_grid_block_x"], c["N_grid_block_y"]) == (8, 4, 64, 120)
    assert c["N_of_particles_cell"] == 50
    assert [o["type"] for o in c["boundary_objects"]] == [2, 1, 2, 1, 1, 1]
    assert [o["segments"] for o in c["boundary_objects"]] == [
        [(0, 0, 0, 481)], [(0, 481, 513, 481)], [(513, 0, 513, 481)], [(313, 0, 513, 0)], [(200, 0, 313, 0)],
        [(0, 0, 200, 0)]]
    assert c["inner_objects"] == []
    assert d.sim["N_subcycles"] == 11 and d.sim["t_sim_ns"] == 10.01
    assert d.parts["init_Te_eV"] == 5.0 and d.parts["seed"] == 123456789
    assert d.parts["species"] == [dict(M_i_amu=40.0, Qs=1, init_Ti_eV=0.03, init_NiNe=100.0)]
    g = d.neutrals[0]
    assert (g["name"], g["N_m3"], g["T_K"], g["M_amu"]) == ("Argon0", 1.0e20, 300.0, 40.0)
    assert [p["type"] for p in g["processes"]] == [10, 20, 30]
    assert g["processes"][2]["ion_species"] == 0
    assert g["processes"][0]["E"][:2] == [0.0, 0.01] and g["processes"][0]["sigma"][0] == 6.3e-20
    assert len(g["processes"][0]["E"]) == 69
    e0, segs = g["segments"]
    assert e0 == 0.0 and len(segs) == 10 and segs[0] == (0.04, 0.005) and segs[-1] == (10000.0, 50.0)
    assert d.bo[4]["N_emit"] == 10.0 and d.bo[4]["model"] == 1 and d.bo[1]["N_emit"] == 0.0


def test_sample_ebeam_workload():
    p = D.read_deck(EBEAM).problem()
    assert (p.gmi, p.gmj, p.periodic_x, p.N_subcycles) == (513, 481, True, 11)
    assert p.n_particles == 50 * 512 * 481  # N x gmj cells
    # walls only: the periodic objects 0 and 2 keep their numbers but have no segments
    assert p.segments == [(config.SIDE_ABOVE, 0, 513, 1), (config.SIDE_BELOW, 313, 513, 3),
                          (config.SIDE_BELOW, 200, 313, 4), (config.SIDE_BELOW, 0, 200, 5)]
    assert p.n_boundary_obj == 6 and p.object_phi == [0.0] * 6
    assert list(p.wall_emission["constant"]) == [4]
    # the same normalisation as the workloads built by hand
    ref = config.Problem(T_e_eV=10.0, N_plasma_m3=1.0e17, N_of_cells_debye=2, N_max_vel=10, N_of_particles_cell=50,
                         gmi=513, gmj=481, periodic_x=True)
    assert p.dt == ref.dt and p.dx == ref.dx and p.V_scale == ref.V_scale


def test_sample_periodic_see_deck_dielectric_walls():
    """input_data_sample_periodic_SEE: the walls above (#2) and below (#4) are dielectrics
    (object type 4) of BN ceramic, permittivity 1.000 from init_material_BNCER0.dat, both with
    electron-induced emission; the two periodic objects have no segments."""
    d = D.read_deck(SEE)
    p = d.problem()
    assert (p.gmi, p.gmj, p.periodic_x) == (513, 481, True)
    assert p.segments == [(config.SIDE_ABOVE, 0, 513, 1), (config.SIDE_BELOW, 0, 513, 3)]
    assert p.dielectric_eps == [None, 1.0, None, 1.0]
    assert sorted(p.wall_emission["see"]) == [1, 3] and not p.wall_emission["constant"]
    assert d.materials["BNCER0"]["eps_diel"] == 1.0
    assert p.n_particles == 20 * 512 * 481


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference tree absent")
def test_every_shipped_deck_is_read_or_rejected_with_a_reason():
    read, rejected = [], []
    for d in sorted(glob.glob("/root/reference/input_data_*")):
        try:
            D.read_deck(d).problem()
            read.append(os.path.basename(d))
        except D.DeckError as e:
            assert "not modelled" in str(e) or "one ion species" in str(e), (d, e)
            rejected.append(os.path.basename(d))
    assert {"input_data_sample_periodic_ebeam", "input_data_sample_periodic_SEE", "input_data_sample"} <= set(read)
    # the others need ion-induced emission, time-varying potentials or an external circuit
    assert len(rejected) == 11 - len(read)


def test_sample_deck_vacuum_gaps():
    """input_data_sample (when the reference tree is present): a walled box of four electrodes
    at 0, 80, 0 and -20 V joined by four vacuum gaps, objects 2, 4, 6 and 8."""
    d = "/root/reference/input_data_sample"
    if not os.path.isdir(d):
        pytest.skip("reference tree absent")
    p = D.read_deck(d).problem()
    assert not p.periodic_x and p.vacuum_gaps == [1, 3, 5, 7]
    assert [round(v * p.F_scale, 6) for v in p.object_phi[:8]] == [0.0, 0.0, 80.0, 0.0, 0.0, 0.0, -20.0, 0.0]


def test_ion_induced_emission_decks():
    """The dc decks' METAL2 material (ion-wall model 2): coefficient 0.2 for ions between 100 eV
    and 10 keV, electrons at 2 eV, on the walls of the periodic deck (objects 2 and 4) and on
    the inner box of the inner-object decks (object 5)."""
    base = "/root/reference"
    if not os.path.isdir(base):
        pytest.skip("reference tree absent")
    want = dict(setD=0.2, minE_eV=100.0, maxE_eV=10000.0, T_eV=2.0)
    p = D.read_deck(os.path.join(base, "input_data_dc_periodic_ion_induced_EE")).problem()
    assert p.wall_emission["iiee"] == {1: [want], 3: [want]}
    p = D.read_deck(os.path.join(base, "input_data_dc_inner_object_ion_ind_EE")).problem()
    assert p.wall_emission["iiee"] == {4: [want]}


def test_electrode_schedule(tmp_path):
    """UPDATE_WALL_POTENTIALS on the host: the harmonic term, and a waveform table of
    init_bo_NN_waveform.dat's form read with the reference's single-precision time column,
    forced ends, and the float32 interpolation weight."""
    import math
    import numpy as np
    from edgpu import wallpotentials as W
    el = W.Electrode(0.1, 0.5, 0.2, 0.3)
    for T in (0, 7, 123):
        assert el.phi(T) == 0.1 + 0.5 * math.sin(0.2 * T + 0.3)
    wf = tmp_path / "init_bo_01_waveform.dat"
    wf.write_text("  50.0  potential amplitude\n  1.0e9   frequency Hz\n# c\n# c\n"
                  "  0.0    -1.0\n  0.49   -0.5\n  0.51    1.0\n  1.0     0.5\n")
    el = W.Electrode()
    dt, F = 1.0e-11, 2.0                     # a 100-step period
    W.read_waveform(el, str(wf), dt, F)
    assert el.wf_T == [0, 49, 51, 100]
    ai = float(np.float32((49 - 25) / 49))
    assert el.phi(25) == 25.0 * (-1.0 * ai + (-0.5) * (1.0 - ai))
    assert el.phi(125) == el.phi(25) and el.phi(100) == el.phi(0) == -25.0


def test_waveform_deck_schedules_its_electrode():
    d = "/root/reference/input_data_waveform_amplitude_profiling"
    if not os.path.isdir(d):
        pytest.skip("reference tree absent")
    p = D.read_deck(d).problem()
    wp = p.wall_potentials
    el = wp.el[0]
    assert el.wf_T is not None and el.ap_T is not None
    phis = [wp.phi(T)[0] for T in range(0, 4000, 7)]
    assert max(phis) > 0.0 > min(phis) and phis.count(phis[0]) < len(phis)
