"""Wall-flux history files in the reference's format.

history_bo_NN.dat (INITIATE_WALL_DIAGNOSTICS / SAVE_BOUNDARY_PARTICLE_HITS_EMISSIONS,
pic2d_ElectronWallCollisions.f90:341-466): one file per object NN = 01, 02, ... (1-based),
one line per electron time step written with the Fortran format '(2x,i8,10(2x,i8))':
the step counter, the electron hits, the hits of each ion species, the electrons emitted.
"""
import os


def _i8(v):
    s = "%d" % int(v)
    return s.rjust(8) if len(s) <= 8 else "*" * 8


def history_line(T_cntr, electron_hits, ion_hits, electron_emits):
    """One record: '(2x,i8,10(2x,i8))' of T_cntr, electron_hit_count, ion_hit_count(1:N_spec),
    electron_emit_count."""
    vals = [T_cntr, electron_hits] + list(ion_hits) + [electron_emits]
    if len(vals) > 11:
        raise ValueError("the record format holds at most 9 ion species")
    return "".join("  " + _i8(v) for v in vals)


def history_file(directory, k):
    """history_bo_NN.dat of object k (0-based)"""
    return os.path.join(directory, "history_bo_%02d.dat" % (k + 1))


def initiate_history(directory, n_objects):
    """Fresh start: empty files for objects 0..n_objects-1."""
    for k in range(n_objects):
        open(history_file(directory, k), "w").close()


def append_history(directory, T_cntr, electron_hits, ion_hits, electron_emits):
    """Append the records of one step.  electron_hits[k], electron_emits[k] per object,
    ion_hits[s][k] per ion species and object."""
    for k in range(len(electron_hits)):
        with open(history_file(directory, k), "a") as f:
            f.write(history_line(T_cntr, electron_hits[k], [h[k] for h in ion_hits], electron_emits[k]) + "\n")


def read_history(directory, k):
    """Records of object k as lists of integers."""
    with open(history_file(directory, k)) as f:
        return [[int(t) for t in line.split()] for line in f if line.strip()]
