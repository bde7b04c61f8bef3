"""MAT-file (level 5) reader / writer (ps_matio.cpp) against scipy.io, an
independent implementation of the same published format: files written by
scipy are read back exactly, files written here load in scipy, with and
without compression, for doubles, singles, integer storage types, 3-D arrays
(column-major on disk, C order in memory) and cell arrays; plus the joint
files of objectdetect_learnparam.cpp and loadJoints' flip
(objectdetect_aux.cpp:96-136)."""
import numpy as np
import pytest
import scipy.io as sio

from partapp_amd._lib import PsError
from partapp_amd.api import (JOINT_POS_GAUSSIAN, JOINT_ROT_GAUSSIAN, MatFile, load_joint_mat, load_joints,
                             load_part_marginal, load_score_grid, mat_write, save_joint_mat, save_part_marginal)


@pytest.mark.parametrize("compress", [False, True])
def test_reads_scipy_written_files(tmp_path, compress):
    rng = np.random.default_rng(0)
    a3 = rng.normal(size=(3, 4, 5))
    data = {
        "scalar": np.array([[2.5]]),
        "vec": np.arange(7.0).reshape(7, 1),
        "mat": rng.normal(size=(4, 6)),
        "grid3": a3,
        "single": rng.normal(size=(5, 2)).astype(np.float32),
        "bytes": np.arange(12, dtype=np.uint8).reshape(3, 4),
        "ints": (np.arange(6, dtype=np.int32) - 3).reshape(2, 3),
    }
    p = tmp_path / "a.mat"
    sio.savemat(p, data, do_compression=compress)
    with MatFile(p) as m:
        for k, v in data.items():
            got = m.get(k)
            assert got.shape == v.shape, k
            np.testing.assert_array_equal(got, v.astype(np.float64), err_msg=k)
        with pytest.raises(PsError) as e:
            m.get("missing")
        assert e.value.status == -11


def test_reads_cell_arrays(tmp_path):
    rng = np.random.default_rng(1)
    cells = np.empty((2, 3), dtype=object)
    for i in range(2):
        for j in range(3):
            cells[i, j] = rng.normal(size=(4 + i, 5 + j)).astype(np.float32)
    T = rng.normal(size=(2, 3, 3, 3))
    p = tmp_path / "imgidx0-pidx1-o0-scoregrid.mat"
    sio.savemat(p, {"cell_scoregrid": cells, "transform_Ti2": T, "transform_T2g": T + 1}, do_compression=True)
    grids, ti2, t2g = load_score_grid(p)
    for i in range(2):
        for j in range(3):
            np.testing.assert_array_equal(grids[i][j], cells[i, j])
    np.testing.assert_array_equal(ti2, T)
    np.testing.assert_array_equal(t2g, T + 1)


@pytest.mark.parametrize("compress", [False, True])
def test_written_files_load_in_scipy(tmp_path, compress):
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 4, 5))
    b = rng.normal(size=(6, 2)).astype(np.float32)
    p = tmp_path / "w.mat"
    mat_write(p, "a", a, compress=compress)
    mat_write(p, "b", b, append=True, compress=compress)
    mat_write(p, "s", 7.0, append=True, compress=compress)
    mat_write(p, "v", np.arange(4.0), append=True, compress=compress)  # 1-D -> 4 x 1
    d = sio.loadmat(p)
    np.testing.assert_array_equal(d["a"], a)
    assert d["b"].dtype == np.float32
    np.testing.assert_array_equal(d["b"], b)
    assert d["s"].shape == (1, 1) and d["s"][0, 0] == 7.0
    np.testing.assert_array_equal(d["v"], np.arange(4.0).reshape(4, 1))
    with MatFile(p) as m:  # and back through the reader
        np.testing.assert_array_equal(m.get("a"), a)


def test_part_marginal_dump_round_trip(tmp_path):
    g = np.random.default_rng(3).normal(size=(24, 9, 11)).astype(np.float32)
    p = tmp_path / "log_part_posterior_final_imgidx0_scaleidx0_o0_pidx4.mat"
    save_part_marginal(p, g)
    np.testing.assert_array_equal(load_part_marginal(p), g)
    assert sio.loadmat(p)["log_prob_grid"].shape == (24, 9, 11)


def test_joint_files_and_flip(tmp_path):
    C = [[20.0, 6.0], [6.0, 12.0]]
    save_joint_mat(tmp_path / "joint_2_5.mat", 2, 5, [3.5, -4.0], [-1.25, 2.0], C, rot_mean=0.3, rot_sigma=0.2)
    j, t = load_joint_mat(tmp_path / "joint_2_5.mat")
    assert t == JOINT_ROT_GAUSSIAN
    assert (j.child_idx, j.parent_idx) == (1, 4)  # part ids are 1-based on disk
    assert tuple(j.offset_c) == (3.5, -4.0) and tuple(j.offset_p) == (-1.25, 2.0)
    assert tuple(j.C) == (20.0, 6.0, 6.0, 12.0) and (j.rot_mean, j.rot_sigma) == (0.3, 0.2)
    # flip mirrors x: T C T and T v with T = diag(-1, 1), rot_mean negated
    f, _ = load_joint_mat(tmp_path / "joint_2_5.mat", flip=True)
    assert tuple(f.C) == (20.0, -6.0, -6.0, 12.0)
    assert tuple(f.offset_c) == (-3.5, -4.0) and tuple(f.offset_p) == (1.25, 2.0)
    assert f.rot_mean == -0.3 and f.rot_sigma == 0.2
    # a scipy-written position-only joint: rotation fields keep the Joint() defaults
    sio.savemat(tmp_path / "joint_3_2.mat", {"type": [[float(JOINT_POS_GAUSSIAN)]], "child_idx": [[3.0]],
                                           "parent_idx": [[2.0]], "offset_c": [[1.0, 2.0]],
                                           "offset_p": [[-1.0], [-2.0]], "C": [[4.0, 0.0], [0.0, 9.0]],
                                           "rot_mean": [[0.7]]})
    js = load_joints(tmp_path, [(2, 5), (3, 2)], flip=False)
    assert (js[1].child_idx, js[1].parent_idx, js[1].rot_mean, js[1].rot_sigma) == (2, 1, 0.0, 0.0)
    assert tuple(js[1].offset_c) == (1.0, 2.0) and tuple(js[1].offset_p) == (-1.0, -2.0)


def test_rejects_non_mat_and_bad_joints(tmp_path):
    p = tmp_path / "x.mat"
    p.write_bytes(b"not a mat file")
    with pytest.raises(PsError) as e:
        MatFile(p)
    assert e.value.status == -10
    sio.savemat(tmp_path / "joint_1_2.mat", {"type": [[2.0]], "child_idx": [[1.0]], "parent_idx": [[2.0]],
                                           "offset_c": [[0.0, 0.0]], "offset_p": [[0.0, 0.0]],
                                           "C": [[1.0, 2.0], [2.0, 1.0]], "rot_mean": [[0.0]],
                                           "rot_sigma": [[0.1]]})
    with pytest.raises(PsError) as e:
        load_joint_mat(tmp_path / "joint_1_2.mat")  # det C < 0 (objectdetect_aux.cpp:134-135)
    assert e.value.status == -3
