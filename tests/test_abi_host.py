"""Host-side checks that need no GPU: the C-ABI library loads and exports every entry
point include/edgpu.h declares, the ctypes mirror of edgpu_params has the C layout, the
counter-based RNG matches the published Philox4x32-10 known answers, the product
package never touches the oracle, and the workload scales match the reference's
normalisation."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "edgpu.h")
PKG = os.path.join(ROOT, "edipic-2d_amd")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(edgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for must in ("edgpu_init", "edgpu_advance_electrons", "edgpu_advance_ions", "edgpu_gather_charge_density",
                 "edgpu_solve_poisson_fftx", "edgpu_solve_potential_krylov", "edgpu_calculate_electric_field",
                 "edgpu_perform_electron_neutral_collisions", "edgpu_perform_resonant_charge_exchange",
                 "edgpu_collect_moments", "edgpu_set_uniform_stream", "edgpu_step"):
        assert must in names


def test_library_exports_every_declared_symbol(edgpu):
    L = edgpu.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", edgpu.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (edgpu_[a-z0-9_]+)\b", out.stdout))
    assert set(_declared()) <= exported


def test_python_mirror_binds_every_symbol(edgpu):
    src = open(os.path.join(PKG, "edgpu", "__init__.py")).read()
    unbound = [n for n in _declared() if n not in src]
    assert not unbound, unbound


@pytest.mark.parametrize("cname,pyname", [("edgpu_params", "Params"), ("edgpu_advance_stats", "AdvanceStats"),
                                           ("edgpu_see_params", "SeeParams")])
def test_struct_layout_matches_c(edgpu, tmp_path, cname, pyname):
    """sizeof / offsetof of the ABI structs from the C compiler against the ctypes mirror."""
    cls = getattr(edgpu, pyname)
    fields = [f[0] for f in cls._fields_]
    prog = tmp_path / "layout.c"
    body = "\n".join('printf("%s %%zu\\n", offsetof(%s, %s));' % (f, cname, f) for f in fields)
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "edgpu.h"\nint main(void){\n'
                    'printf("sizeof %%zu\\n", sizeof(%s));\n%s\nreturn 0;}\n' % (cname, body))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    assert int(out["sizeof"]) == C.sizeof(cls)
    for f in fields:
        assert int(out[f]) == getattr(cls, f).offset, f


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(edgpu, ctr, key, want):
    """Philox4x32-10 known-answer vectors (Salmon et al., SC'11, Random123 kat_vectors)."""
    assert edgpu.philox4x32_10(ctr, key) == want


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0], 0, [0xff1dae59, 0x6cd10df2]),
    ([0xffffffff] * 2, 0xffffffff, [0x2c3f628b, 0xab4fd7ad]),
    ([0x243f6a88, 0x85a308d3], 0x13198a2e, [0xdd7ce038, 0xf62a4c12]),
])
def test_philox2x32_known_answers(edgpu, ctr, key, want):
    """Philox2x32-10 known-answer vectors (same source)."""
    assert edgpu.philox2x32_10(ctr, key) == want


def test_product_package_does_not_use_the_oracle():
    pat = re.compile(r"oracle_ffi|liboracle|import oracle|from oracle|orc_[a-z]")
    hits = []
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".f90", ".F90")):
                p = os.path.join(dirpath, f)
                for k, line in enumerate(open(p, errors="replace"), 1):
                    if pat.search(line):
                        hits.append("%s:%d: %s" % (os.path.relpath(p, ROOT), k, line.strip()))
    assert not hits, hits


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """With no built library the package raises instead of falling back to anything."""
    code = ("import sys; sys.path.insert(0, %r); import edgpu\n"
            "try:\n    edgpu.lib()\nexcept RuntimeError as e:\n    print('raised', e)\n" % PKG)
    env = dict(os.environ, EDGPU_LIB=str(tmp_path / "absent.so"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env)
    assert "raised" in out.stdout


def test_workload_scales():
    """Scales of the benchmark configurations (reference normalisation,
    pic2d_CurProblemValues.f90:719-723, 893-903)."""
    sys.path.insert(0, PKG)
    from edgpu import config
    p = config.workload(2)
    assert p.dx == pytest.approx(2.351e-4, rel=1e-3)
    assert p.dt == pytest.approx(5.605e-11, rel=1e-3)
    assert p.B_scale == pytest.approx(0.1014, rel=1e-3)
    assert p.Bz == pytest.approx(0.0986, rel=2e-3)
    assert p.F_scale == pytest.approx(1.000, rel=1e-3)
    assert p.n_particles == 419430400 and p.gmi == 2049 and p.gmj == 1024
    assert config.workload(1).n_particles == 3276800
    p4 = config.workload(4)
    assert p4.dx == pytest.approx(1.288e-4, rel=1e-3) and p4.dt == pytest.approx(2.507e-11, rel=1e-3)
    assert p4.F_scale == pytest.approx(1.5, rel=1e-3)
    assert p4.n_particles == 104857600
    assert config.workload(5).n_particles == 838860800
    # factor1 = F/(E dx) and factor2 = factor1/2
    assert p.factor1_E_from_F == pytest.approx(2 * p.factor2_E_from_F)


FORTRAN_FIELDS = ["gmi", "gmj", "periodic_x", "N_of_particles_cell", "N_subcycles", "precision", "Bx_ext",
                  "By_ext", "Bz_ext", "Ez_ext", "factor1_E_from_F", "factor2_E_from_F", "n_boundary_objects",
                  "n_segments", "seg_side", "seg_start", "seg_end", "seg_object", "n_inner_objects", "inner_box",
                  "object_phi", "n_ion_species", "Qs", "QM2s", "Ms", "ions_sense_magnetic_field", "ions_sense_Ez",
                  "max_electrons", "max_ions", "add_capacity", "seed", "fuse_collisions", "sort_interval",
                  "ion_sort_interval", "object_type", "object_eps"]


@pytest.mark.parametrize("ftype,pyname", [("edgpu_params", "Params"), ("edgpu_advance_stats", "AdvanceStats"),
                                           ("edgpu_see_params", "SeeParams")])
def test_fortran_shim_layout(edgpu, tmp_path, ftype, pyname):
    """The Fortran shim compiles with amdflang and its bind(C) types have the C layout
    (field addresses taken with c_loc in Fortran, compared with the ctypes offsets)."""
    import shutil
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        pytest.skip("no Fortran compiler")
    cls = getattr(edgpu, pyname)
    fields = FORTRAN_FIELDS if pyname == "Params" else [f[0] for f in cls._fields_]
    shim = os.path.join(PKG, "fortran", "edgpu_mod.F90")
    n = len(fields)
    probe = tmp_path / "probe.f90"
    probe.write_text("subroutine edgpu_layout_probe(p, addr) bind(C, name='edgpu_layout_probe')\n"
                     "  use, intrinsic :: iso_c_binding\n  use edgpu\n  implicit none\n"
                     "  type(%s), target, intent(in) :: p\n  type(c_ptr), intent(out) :: addr(%d)\n"
                     "  addr(1) = c_loc(p)\n" % (ftype, n + 1)
                     + "".join("  addr(%d) = c_loc(p%%%s)\n" % (k + 2, f) for k, f in enumerate(fields))
                     + "end subroutine\n")
    subprocess.run([fc, "-c", "-fPIC", "-cpp", shim, "-o", str(tmp_path / "shim.o")], check=True, cwd=tmp_path)
    subprocess.run([fc, "-c", "-fPIC", str(probe), "-o", str(tmp_path / "probe.o")], check=True, cwd=tmp_path)
    so = tmp_path / "probe.so"
    subprocess.run(["gcc", "-shared", "-o", str(so), str(tmp_path / "probe.o")], check=True)
    L = C.CDLL(str(so))
    p = cls()
    addr = (C.c_void_p * (n + 1))()
    L.edgpu_layout_probe(C.byref(p), addr)
    got = [a - addr[0] for a in addr[1:]]
    want = [getattr(cls, f).offset for f in fields]
    assert got == want
    assert set(fields) == set(f[0] for f in cls._fields_)


@pytest.mark.parametrize("src", ["particles_fp64.hip", "particles_mixed.hip"])
def test_push_kernels_stay_in_registers(src):
    """The compiler's resource report of the last build: every push kernel keeps its state in
    registers (at most a few bytes of scratch per lane, no struct copied to private memory)
    and the pushes of ordinary steps run at least 4 waves per SIMD (those of snapshot steps,
    which also sum the moments, at least 2; the sorting pushes, which keep two chunks of
    particle states in registers, at least 3).  A rare path that needs a large private copy of
    the push arguments (1.3 KB per lane) once made both pushes 7x slower while every result
    stayed right; such paths run in kernels of their own after the push."""
    from edgpu import build as B
    path = os.path.join(B.BUILD, os.path.splitext(src)[0] + ".resources")
    if not os.path.exists(path):
        pytest.skip("no resource report: the library was not built here")
    res = {k: v for k, v in B.kernel_resources(src).items() if "k_advance_" in k}
    assert len(res) >= 20
    assert sum("k_advance_sort_" in name for name in res) == 4
    for name, r in res.items():
        moments = name.endswith("ELb1EEEvNS_11AdvanceArgsE")  # the last template flag: MOM
        sorting = "k_advance_sort_" in name
        assert r["ScratchSize"] <= (128 if moments else 64), (name, r)
        assert r["Occupancy"] >= (2 if moments else 3 if sorting else 4), (name, r)
