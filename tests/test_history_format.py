"""history_bo_NN.dat writer: the Fortran record format '(2x,i8,10(2x,i8))' of
SAVE_BOUNDARY_PARTICLE_HITS_EMISSIONS (pic2d_ElectronWallCollisions.f90:450-462), worked by hand."""
import os

from edgpu import diagnostics as D


def test_record_layout_by_hand():
    # 2 blanks + 8-wide right-justified integers, one field per count
    assert D.history_line(12, 345, [67], 8) == "        12       345        67         8"
    assert D.history_line(1, 0, [2, 3], 4) == "         1         0         2         3         4"
    # a count wider than the field prints as asterisks, as Fortran's I8 does
    assert D.history_line(123456789, 5, [6], 7).startswith("  ********")


def test_files_per_object_and_append(tmp_path):
    d = str(tmp_path)
    D.initiate_history(d, 3)
    assert sorted(os.listdir(d)) == ["history_bo_01.dat", "history_bo_02.dat", "history_bo_03.dat"]
    D.append_history(d, 1, [10, 20, 30], [[1, 2, 3]], [0, 5, 0])
    D.append_history(d, 2, [11, 21, 31], [[4, 5, 6]], [0, 6, 0])
    assert D.read_history(d, 1) == [[1, 20, 2, 5], [2, 21, 5, 6]]
    assert open(os.path.join(d, "history_bo_03.dat")).read().splitlines()[1] == \
        "         2        31         6         0"
