"""Byte layout of the checkpoint files against offsets worked by hand from the reference's
writers (pic2d_Checkpoints.f90:760-925 POSIX, :58-240 MPI-IO), the generator-state words
against the reference's well_random_seed recurrence worked by hand, and the error paths."""
import struct

import numpy as np
import pytest

from edgpu import checkpoint as K


def _species(n, base, tag0):
    k = np.arange(n, dtype=np.float64)
    return [base + k, base + 0.5 + k, -k, 2 * k, 1e-3 * k, (tag0 + np.arange(n)).astype(np.int32)]


def _header(n_spec=1):
    words, ind, func = K.encode_rng_state(0x1122334455667788, 77)
    return K.Header(n_spec, words, ind, func,
                    {"T_cntr": 5000, "T_cntr_global_load_balance": 5001, "T_cntr_cluster_load_balance": 5002,
                     "Save_probes_e_data_T_cntr": 5003, "N_of_saved_records": 6, "current_snap": 7,
                     "particle_master": 0},
                    max_N_electrons=100, max_N_ions=[200] * n_spec, total_cathode_N_e_to_inject=9,
                    reals=[1.5, -2.25])


def test_names():
    assert K.posix_name(2500, 3) == "checkdir_00002500/Tcntr_00002500_proc_003.check"
    assert len(K.posix_name(2500, 3)) == 47
    assert K.mpiio_name(12) == "Tcntr_00000012.check" and len(K.mpiio_name(12)) == 20


def test_header_length():
    # 636 = 624 RNG words + state_ind + func + 7 counters + 2 electron counts + cathode count
    assert K.header_length(1) == 639 and K.header_length(3) == 643


def test_well_seed_recurrence_worked_by_hand():
    # seed 1: w2 = 1812433253 (1 XOR 2^30) + 1 mod 2^32; 1812433253 * 2^30 mod 2^32 = 2^30
    # (1812433253 = 1 mod 4), so w2 = 2^30 + 1812433253 + 1 = 2886175078
    w = K.well_seed_words([1])
    assert w[0] == 1 and w[1] == 2886175078
    # w3 from w2 = 2886175078 = 0xAC07C766: w2 << 30 keeps bits 0-1 of w2 (= 2) as bits 30-31
    x = 2886175078 ^ ((2886175078 & 3) << 30)
    assert w[2] == (1812433253 * x + 2) % 2 ** 32
    assert w.dtype == np.uint32 and w.size == 624


def test_rng_state_round_trip_and_foreign_state():
    words, ind, func = K.encode_rng_state(0xDEADBEEF12345678, 4242)
    assert (ind, func) == (0, 1) and words[3] == K.MARKER
    assert K.decode_rng_state(words, ind, func) == (0xDEADBEEF12345678, 4242, True)
    # a reference-seeded WELL state (well_random_seed(5), resumed at state_ind 17) is foreign:
    # seeded by its hash, counter 0, and the hash depends on every word
    ref = K.well_seed_words([5])
    s1, c1, ours = K.decode_rng_state(ref, 17, 1)
    assert not ours and c1 == 0
    ref2 = ref.copy()
    ref2[600] ^= 1
    assert K.decode_rng_state(ref2, 17, 1)[0] != s1
    assert K.decode_rng_state(ref, 16, 1)[0] != s1
    # one altered word of ours is no longer recognised
    bad = words.copy()
    bad[100] ^= 4
    assert not K.decode_rng_state(bad, 0, 1)[2]


def test_fnv1a_known_answers():
    # FNV-1a 64 test vectors (Fowler, Noll, Vo): "" and "a"
    assert K._fnv1a64(b"") == 0xCBF29CE484222325
    assert K._fnv1a64(b"a") == 0xAF63DC4C8601EC8C


@pytest.mark.parametrize("layout", ["posix", "mpiio"])
def test_byte_layout(tmp_path, layout):
    h = _header(1)
    e, i = _species(3, 10.0, 100), _species(2, 20.0, 0)
    path = str(tmp_path / "c.check")
    K.write_checkpoint_file(path, h, [e, i], layout)
    b = open(path, "rb").read()
    nint = 639
    ints = struct.unpack("<%di" % nint, b[:4 * nint])
    assert ints[0] == 0x55667788 and ints[1] == 0x11223344 and ints[2] == 77
    assert ints[3] == K.MARKER and (ints[624], ints[625]) == (0, 1)
    assert ints[626:633] == (5000, 5001, 5002, 5003, 6, 7, 0)
    assert ints[633:636] == (100, 3, 9)          # max_N_electrons, N_electrons, cathode count
    assert ints[636:638] == (200, 2)             # max_N_ions(1), N_ions(1)
    assert ints[638] == 2                        # len_surf_charge
    off = 4 * nint
    assert struct.unpack("<2d", b[off:off + 16]) == (1.5, -2.25)
    off += 16
    if layout == "posix":
        assert len(b) == off + 44 * 5
        # electron 2 (k = 1): X, Y, VX, VY, VZ, tag at 44-byte stride
        assert struct.unpack("<5di", b[off + 44:off + 88]) == (11.0, 11.5, -1.0, 2.0, 1e-3, 101)
        # ion 1 follows the three electrons
        assert struct.unpack("<5di", b[off + 132:off + 176]) == (20.0, 20.5, 0.0, 0.0, 0.0, 0)
    else:
        assert len(b) == off + 40 * 5 + 4 * 5
        assert struct.unpack("<5d", b[off + 40:off + 80]) == (11.0, 11.5, -1.0, 2.0, 1e-3)
        assert struct.unpack("<3i", b[off + 120:off + 132]) == (100, 101, 102)
        io = off + 132
        assert struct.unpack("<5d", b[io + 40:io + 80]) == (21.0, 21.5, -1.0, 2.0, 1e-3)
        assert struct.unpack("<2i", b[io + 80:io + 88]) == (0, 1)
    h2, sp = K.read_checkpoint_file(path, 1, layout)
    assert h2.counters == h.counters and h2.N_electrons == 3 and h2.N_ions == [2]
    assert h2.len_surf_charge == 2 and list(h2.reals) == [1.5, -2.25]
    for got, want in zip(sp, [e, i]):
        for a, b_ in zip(got, want):
            assert np.array_equal(a, b_)


def test_chunked_io_matches(tmp_path, monkeypatch):
    monkeypatch.setattr(K, "CHUNK", 7)
    e, i, j = _species(20, 1.0, 0), _species(9, 2.0, 3), _species(0, 0.0, 0)
    for layout in ("posix", "mpiio"):
        path = str(tmp_path / ("c_" + layout))
        K.write_checkpoint_file(path, _header(2), [e, i, j], layout)
        _, sp = K.read_checkpoint_file(path, 2, layout)
        for got, want in zip(sp, [e, i, j]):
            for a, b in zip(got, want):
                assert np.array_equal(a, b)


def test_errors(tmp_path):
    path = str(tmp_path / "c.check")
    K.write_checkpoint_file(path, _header(1), [_species(3, 0.0, 0), _species(2, 0.0, 0)])
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-1])
    with pytest.raises(K.CheckpointError, match="inside the particle records"):
        K.read_checkpoint_file(path, 1)
    open(path, "wb").write(raw + b"\0")
    with pytest.raises(K.CheckpointError, match="trailing"):
        K.read_checkpoint_file(path, 1)
    # a header read for the wrong number of species runs into the reals and particles
    with pytest.raises(K.CheckpointError):
        K.read_checkpoint_file(path, 3)
    open(path, "wb").write(raw[:400])
    with pytest.raises(K.CheckpointError, match="short"):
        K.read_checkpoint_file(path, 1)
    h = _header(1)
    h.max_N_electrons = 2 ** 31
    with pytest.raises(K.CheckpointError, match="4-byte INTEGER"):
        K.write_checkpoint_file(path, h, [_species(1, 0.0, 0), _species(1, 0.0, 0)])
