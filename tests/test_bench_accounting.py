"""CPU checks of the bench's byte accounting: the grids each kernel class of
the sweep must move for the 10-part person tree, and the committed column-pass
traffic summary that the bench reads for roofline.traffic."""
import os
import re

import bench
from partapp_amd.api import ExpParam, person_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def model(**kw):
    return person_model(ExpParam(-180.0, 180.0, 24, 1.0, 1.0, 1, 0.0), **kw)


def test_grid_counts_of_the_person_tree():
    # 9 edges, chains of length 2 (4) and 1 (head): per message the row pass
    # reads its source and writes T; the column pass reads T and writes the
    # destination, plus the unary of a detected chain node (up), the unary
    # (down) and, for a chain node, the stored belief read and next source write
    fp32 = bench.grid_traffic(model(), 4)
    assert fp32 == {"rowpass": 36, "colpass": 57, "root": 12, "leaf_max": 5}
    fp16 = bench.grid_traffic(model(), 2)
    assert fp16 == {"rowpass": 27, "colpass": 48, "root": 12, "leaf_max": 5}
    # B_alg = 224 bytes per state and scale: reading unaries and writing
    # marginals once (2P) and one grid in, one out per message (4E)
    assert 4 * (2 * 10 + 4 * 9) == 224


def test_grid_counts_without_detectors():
    # an undetected leaf is a zero source (no read, no leaf max); an undetected
    # chain node adds no unary to its belief
    is_detect = [1] * 10
    is_detect[9] = 0  # leaf
    is_detect[7] = 0  # chain node
    g = bench.grid_traffic(model(is_detect=is_detect), 4)
    assert g["leaf_max"] == 4
    assert g["rowpass"] == 36 - 1
    assert g["colpass"] == 57 - 1


def test_committed_column_pass_traffic_is_consistent():
    path = os.path.join(ROOT, "profiles", "traffic_colpass_bf16mix_32x300x300x24.txt")
    text = open(path).read()
    fetch = [int(v) for v in re.search(r"FETCH_SIZE KiB per launch[^:]*:([\d ]+)", text).group(1).split()]
    write = [int(v) for v in re.search(r"WRITE_SIZE KiB per launch[^:]*:([\d ]+)", text).group(1).split()]
    assert len(fetch) == len(write) > 0
    per_launch = (2 * sum(fetch) + sum(write)) * 1024 / len(fetch)
    got, src = bench.measured_traffic("bf16mix", "colpass")
    assert got == int(per_launch)
    assert src.endswith("traffic_colpass_bf16mix_32x300x300x24.txt")
    # within 15 % of the algorithmic bytes: 48 grids of 32 x 300 x 300 x 24
    # states (4 bytes) per sweep over the sweep's 4 launches (the file lists
    # whole sweeps)
    assert len(fetch) % 4 == 0
    alg = 48 * 32 * 300 * 300 * 24 * 4 / 4
    assert 1.0 <= per_launch / alg < 1.15
