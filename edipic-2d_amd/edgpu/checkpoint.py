"""Checkpoints in the reference's two binary layouts, written from and read back into the
device particle arrays.

POSIX layout (SAVE_CHECKPOINT_POSIX / READ_CHECKPOINT_POSIX, pic2d_Checkpoints.f90:702-925,
928-1124): file ``checkdir_TTTTTTTT/Tcntr_TTTTTTTT_proc_PPP.check``, Fortran stream access,
little-endian, no record markers:

  * ``636 + 2 N_spec + 1`` default (4-byte) integers: 624 words of the RNG state, state_ind,
    func, T_cntr, T_cntr_global_load_balance, T_cntr_cluster_load_balance,
    Save_probes_e_data_T_cntr, N_of_saved_records, current_snap, particle_master,
    max_N_electrons, N_electrons, total_cathode_N_e_to_inject, then max_N_ions(s), N_ions(s)
    per ion species, then len_surf_charge;
  * len_surf_charge REAL(8) values (dielectric surface charges, then the potentials,
    charges and plasma charge changes of the objects whose potentials are solved for);
  * per electron X, Y, VX, VY, VZ (REAL(8)) and tag (INTEGER), 44 bytes with no padding;
  * the same per ion of each species in turn.

MPI-IO layout (SAVE_CHECKPOINT_MPIIO_2, pic2d_Checkpoints.f90:58-398, one process): file
``Tcntr_TTTTTTTT.check``; the same integers and reals, then per species the 5 N doubles
X, Y, VX, VY, VZ particle by particle, followed by the N tags.

The reference keeps its WELL19937a generator state in the first 626 words.  The device
generator is counter based (Philox keyed by a 64-bit seed and counted by the step), so the
writer stores seed_lo, seed_hi, step counter and a marker word in words 1-4 and fills words
5-624 by the recurrence of the reference's own well_random_seed
(pic2d_RandomNumberInterface.f90:39-56), with state_ind = 0 and func = 1: the words form a
valid WELL state that the reference can resume from, and the reader recognises them by the
marker and the recurrence.  A state the reference wrote carries no such signature; the
reader then seeds the device generator with a 64-bit FNV-1a hash of the 626 words and a
step counter of 0 (the continued run is statistically, not bitwise, the reference's).

Saving requires n_sub = 0, the start of an ion cycle, as in the reference
(pic2d_Checkpoints.f90:755-758): the accumulated ion fields are empty there, so particles,
counters and the generator position are the whole state.  Fields are not saved; the step
recomputes them from the particles, as the reference's main loop does after a restart.
"""
import os

import numpy as np

MAX_INT4 = 2 ** 31 - 1
MARKER = 0x45444750  # "EDGP"
KNUTH = 1812433253
_M32 = 0xFFFFFFFF

POSIX_RECORD = np.dtype([("X", "<f8"), ("Y", "<f8"), ("VX", "<f8"), ("VY", "<f8"), ("VZ", "<f8"), ("tag", "<i4")])
assert POSIX_RECORD.itemsize == 44

COUNTER_NAMES = ("T_cntr", "T_cntr_global_load_balance", "T_cntr_cluster_load_balance",
                 "Save_probes_e_data_T_cntr", "N_of_saved_records", "current_snap", "particle_master")

CHUNK = 1 << 22  # particles per write / read block


class CheckpointError(RuntimeError):
    pass


def header_length(n_spec):
    """Number of 4-byte integers before the reals (ibufer_length, pic2d_Checkpoints.f90:760)."""
    return 636 + 2 * n_spec + 1


def posix_name(T_cntr, rank=0):
    """checkdir_TTTTTTTT/Tcntr_TTTTTTTT_proc_PPP.check (pic2d_Checkpoints.f90:875-880)."""
    return "checkdir_%08d/Tcntr_%08d_proc_%03d.check" % (T_cntr, T_cntr, rank)


def mpiio_name(T_cntr):
    """Tcntr_TTTTTTTT.check (pic2d_Checkpoints.f90:197-198)."""
    return "Tcntr_%08d.check" % T_cntr


def well_seed_words(first, start=1):
    """Words of well_random_seed (pic2d_RandomNumberInterface.f90:39-56) continued from
    ``first`` (a list of the leading words): w(j) = 1812433253 (w(j-1) XOR ISHFT(w(j-1), 30))
    + j - 1 modulo 2^32, with ISHFT(., 30) the left shift of a 4-byte integer."""
    w = [int(v) & _M32 for v in first]
    for k in range(len(w), 624):
        p = w[k - 1]
        w.append((KNUTH * (p ^ ((p << 30) & _M32)) + k) & _M32)
    return np.array(w[:624], dtype=np.uint32)


def encode_rng_state(seed, counter):
    """(state_var(624), state_ind, func) carrying the device generator position."""
    words = well_seed_words([seed & _M32, (seed >> 32) & _M32, counter & _M32, MARKER])
    return words, 0, 1


def _fnv1a64(data):
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def decode_rng_state(words, state_ind, func):
    """(seed, counter, written_by_this_library) from the 626 RNG words of a checkpoint."""
    words = np.asarray(words, dtype=np.uint32)
    if (words.size == 624 and int(words[3]) == MARKER and state_ind == 0 and func == 1
            and np.array_equal(words, well_seed_words([int(v) for v in words[:4]]))):
        return int(words[0]) | (int(words[1]) << 32), int(words[2]), True
    raw = words.astype("<u4").tobytes() + np.array([state_ind, func], dtype="<i4").tobytes()
    return _fnv1a64(raw), 0, False


class Header:
    """The integer block and the reals of a checkpoint."""

    def __init__(self, n_spec, rng_words=None, state_ind=0, func=1, counters=None, max_N_electrons=0,
                 N_electrons=0, total_cathode_N_e_to_inject=0, max_N_ions=None, N_ions=None, reals=None):
        self.n_spec = n_spec
        self.rng_words = np.zeros(624, np.uint32) if rng_words is None else np.asarray(rng_words, np.uint32)
        self.state_ind, self.func = int(state_ind), int(func)
        self.counters = {k: 0 for k in COUNTER_NAMES}
        self.counters.update(counters or {})
        self.max_N_electrons, self.N_electrons = int(max_N_electrons), int(N_electrons)
        self.total_cathode_N_e_to_inject = int(total_cathode_N_e_to_inject)
        self.max_N_ions = list(max_N_ions or [0] * n_spec)
        self.N_ions = list(N_ions or [0] * n_spec)
        self.reals = np.zeros(0) if reals is None else np.asarray(reals, np.float64)

    def ints(self):
        """ibufer(1:636+2 N_spec+1) in the order of pic2d_Checkpoints.f90:762-786, 819."""
        v = [int(w) for w in self.rng_words] + [self.state_ind, self.func]
        v += [int(self.counters[k]) for k in COUNTER_NAMES]
        v += [self.max_N_electrons, self.N_electrons, self.total_cathode_N_e_to_inject]
        for s in range(self.n_spec):
            v += [self.max_N_ions[s], self.N_ions[s]]
        v.append(self.reals.size)
        for k, x in enumerate(v[624:]):
            if not -2 ** 31 <= x <= MAX_INT4:
                raise CheckpointError("integer word %d = %d does not fit the 4-byte INTEGER" % (625 + k, x))
        out = np.empty(len(v), "<i4")
        out[:624] = np.asarray(v[:624], np.uint32).view(np.int32)
        out[624:] = v[624:]
        return out

    @classmethod
    def parse(cls, ints, n_spec):
        ints = np.asarray(ints, "<i4")
        if ints.size != header_length(n_spec):
            raise CheckpointError("short checkpoint header")
        h = cls(n_spec, rng_words=ints[:624].view(np.uint32), state_ind=ints[624], func=ints[625])
        h.counters = {k: int(ints[626 + q]) for q, k in enumerate(COUNTER_NAMES)}
        h.max_N_electrons, h.N_electrons, h.total_cathode_N_e_to_inject = (int(x) for x in ints[633:636])
        pos = 636
        h.max_N_ions, h.N_ions = [], []
        for _ in range(n_spec):
            h.max_N_ions.append(int(ints[pos]))
            h.N_ions.append(int(ints[pos + 1]))
            pos += 2
        h.len_surf_charge = int(ints[pos])
        if h.N_electrons < 0 or min(h.N_ions, default=0) < 0 or h.len_surf_charge < 0:
            raise CheckpointError("negative count in the checkpoint header")
        return h


def _write_species(f, arrays, layout):
    X, Y, VX, VY, VZ, tag = arrays
    n = len(X)
    if layout == "posix":
        rec = np.empty(min(n, CHUNK), POSIX_RECORD)
        for o in range(0, n, CHUNK):
            c = min(CHUNK, n - o)
            r = rec[:c]
            for name, a in zip(POSIX_RECORD.names, (X, Y, VX, VY, VZ, tag)):
                r[name] = a[o:o + c]
            r.tofile(f)
    else:
        buf = np.empty((min(n, CHUNK), 5), "<f8")
        for o in range(0, n, CHUNK):
            c = min(CHUNK, n - o)
            for q, a in enumerate((X, Y, VX, VY, VZ)):
                buf[:c, q] = a[o:o + c]
            buf[:c].tofile(f)
        np.asarray(tag, "<i4").tofile(f)


def _read_species(f, n, layout):
    out = [np.empty(n, np.float64) for _ in range(5)] + [np.empty(n, np.int32)]
    if layout == "posix":
        for o in range(0, n, CHUNK):
            c = min(CHUNK, n - o)
            r = np.fromfile(f, POSIX_RECORD, c)
            if r.size != c:
                raise CheckpointError("checkpoint ends inside the particle records")
            for a, name in zip(out, POSIX_RECORD.names):
                a[o:o + c] = r[name]
    else:
        for o in range(0, n, CHUNK):
            c = min(CHUNK, n - o)
            r = np.fromfile(f, "<f8", 5 * c)
            if r.size != 5 * c:
                raise CheckpointError("checkpoint ends inside the particle records")
            r = r.reshape(c, 5)
            for q in range(5):
                out[q][o:o + c] = r[:, q]
        t = np.fromfile(f, "<i4", n)
        if t.size != n:
            raise CheckpointError("checkpoint ends inside the tags")
        out[5][:] = t
    return out


def write_checkpoint_file(path, header, species, layout="posix"):
    """Write a checkpoint file: header, its reals, then species[0] (electrons) and the ion
    species, each a sequence (X, Y, VX, VY, VZ, tag)."""
    if layout not in ("posix", "mpiio"):
        raise CheckpointError("layout is 'posix' or 'mpiio'")
    if len(species) != header.n_spec + 1:
        raise CheckpointError("one particle set per species")
    header.N_electrons = len(species[0][0])
    header.N_ions = [len(s[0]) for s in species[1:]]
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        header.ints().tofile(f)
        np.asarray(header.reals, "<f8").tofile(f)
        for arrays in species:
            _write_species(f, arrays, layout)


def read_checkpoint_file(path, n_spec, layout="posix"):
    """(Header, [electrons, ions of species 1..n_spec]) of a checkpoint file."""
    with open(path, "rb") as f:
        ints = np.fromfile(f, "<i4", header_length(n_spec))
        h = Header.parse(ints, n_spec)
        h.reals = np.fromfile(f, "<f8", h.len_surf_charge)
        if h.reals.size != h.len_surf_charge:
            raise CheckpointError("checkpoint ends inside the reals")
        species = [_read_species(f, h.N_electrons, layout)]
        for s in range(n_spec):
            species.append(_read_species(f, h.N_ions[s], layout))
        if f.read(1):
            raise CheckpointError("trailing bytes after the last particle record")
    return h, species


def _n_spec(rt):
    return int(rt.params.n_ion_species)


def _dielectric_parts(rt):
    """(side row, istart, iend) of each dielectric wall segment, in the order of the object
    parts the reference writes (pic2d_Checkpoints.f90:106-126): the segments of the boundary
    objects in object order, each its nodes istart..iend."""
    from . import config
    eps = getattr(rt.prob, "dielectric_eps", None) or []
    parts = []
    for (side, i0, i1, o) in sorted(rt.prob.segments, key=lambda t: t[3]):
        if o < len(eps) and eps[o] is not None and side in (config.SIDE_BELOW, config.SIDE_ABOVE):
            parts.append((0 if side == config.SIDE_BELOW else 1, i0, i1))
    return parts


def surface_charge_reals(rt):
    """The surface-charge part of a checkpoint's reals: the dielectric walls' charge node by
    node, part after part (an empty array without dielectric walls)."""
    parts = _dielectric_parts(rt)
    if not parts:
        return np.zeros(0)
    rows = rt.get_surface_charge()
    return np.concatenate([rows[w][i0:i1 + 1] for (w, i0, i1) in parts])


def restore_surface_charge(rt, reals):
    """Inverse of surface_charge_reals: sets the device rows from the first reals; returns how
    many it used."""
    parts = _dielectric_parts(rt)
    if not parts:
        return 0
    need = sum(i1 - i0 + 1 for (_, i0, i1) in parts)
    if len(reals) < need:
        raise CheckpointError("the checkpoint holds %d reals, the dielectric walls need %d" % (len(reals), need))
    nx = rt.prob.gmi + 1
    rows = [np.zeros(nx), np.zeros(nx)]
    pos = 0
    for (w, i0, i1) in parts:
        rows[w][i0:i1 + 1] = reals[pos:pos + i1 - i0 + 1]
        pos += i1 - i0 + 1
    rt.set_surface_charge(*rows)
    return need


def save_checkpoint(rt, T_cntr, directory=".", layout="posix", rank=0, counters=None, reals=None):
    """SAVE_CHECKPOINT_POSIX / SAVE_CHECKPOINT_MPIIO_2 of the device state, called where the
    reference calls it: at the top of time step T_cntr, before the step's work.  counters
    holds the diagnostic counters the caller keeps (COUNTER_NAMES; the load-balance counters
    default to T_cntr, so a reference run resumes balancing at once).  reals defaults to the
    dielectric walls' surface charge.  Returns the path."""
    seed, ctr, n_sub = rt.rng_state()
    if n_sub != 0:
        raise CheckpointError("checkpoint saved at n_sub = %d; the reference saves at the start of an ion "
                              "cycle only" % n_sub)
    n_spec = _n_spec(rt)
    words, ind, func = encode_rng_state(seed, ctr)
    c = {"T_cntr": T_cntr, "T_cntr_global_load_balance": T_cntr, "T_cntr_cluster_load_balance": T_cntr,
         "particle_master": rank}
    c.update(counters or {})
    species = [rt.download(s) for s in range(n_spec + 1)]
    for s, arrays in enumerate(species):
        if len(arrays[0]) > MAX_INT4:
            raise CheckpointError("species %d holds more particles than a 4-byte INTEGER counts" % s)
    p = rt.params
    if reals is None:
        reals = surface_charge_reals(rt)
    h = Header(n_spec, words, ind, func, c,
               max_N_electrons=min(max(int(p.max_electrons), len(species[0][0])), MAX_INT4),
               max_N_ions=[min(max(int(p.max_ions[s]), len(species[s + 1][0])), MAX_INT4) for s in range(n_spec)],
               reals=reals)
    name = posix_name(T_cntr, rank) if layout == "posix" else mpiio_name(T_cntr)
    path = os.path.join(directory, name)
    write_checkpoint_file(path, h, species, layout)
    return path


def read_checkpoint(rt, T_cntr_to_continue, directory=".", rank=0):
    """READ_A_CHECKPOINT (pic2d_Checkpoints.f90:1127-1190): the MPI-IO file if it exists, else
    the POSIX file of this rank; uploads every species, restores the generator position and
    n_sub = 0, and returns the Header (its counters give Start_T_cntr = counters["T_cntr"])."""
    n_spec = _n_spec(rt)
    mp = os.path.join(directory, mpiio_name(T_cntr_to_continue))
    px = os.path.join(directory, posix_name(T_cntr_to_continue, rank))
    if os.path.exists(mp):
        path, layout = mp, "mpiio"
    elif os.path.exists(px):
        path, layout = px, "posix"
    else:
        raise CheckpointError("no checkpoint %s or %s" % (mp, px))
    h, species = read_checkpoint_file(path, n_spec, layout)
    for s, arrays in enumerate(species):
        rt.upload(s, *arrays)
    restore_surface_charge(rt, h.reals)
    seed, ctr, ours = decode_rng_state(h.rng_words, h.state_ind, h.func)
    rt.set_rng_state(seed, ctr, 0)
    h.layout, h.path, h.rng_restored = layout, path, ours
    return h
