"""The committed lattice fixture (tests/golden/lattice_counts.json, written by
tests/golden/make_lattice_counts.py with shift-overlap counting) against the
workload generators and the oracle's cell-list family build."""
import json
import os

import numpy as np

from peridyn_amd import workloads as W

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lattice_counts.json")))


def test_particle_counts():
    assert W.lattice_box(20, 20, 20).shape[0] == GOLD["tensile_box_20"]["N"]
    assert W.lattice_box(50, 50, 20).shape[0] == GOLD["plate_50x50x20"]["N"]
    assert W.lattice_disk(150.0, 10).shape[0] == GOLD["disk_r150_t10"]["N"]
    assert W.lattice_sphere(15.0).shape[0] == GOLD["sphere_r15"]["N"]
    assert W.lattice_ball_count(GOLD["horizon"]) == GOLD["interior_degree"]


def test_oracle_families_match_fixture(orc):
    for key, x in (("tensile_box_20", W.lattice_box(20, 20, 20)), ("plate_50x50x20", W.lattice_box(50, 50, 20))):
        fam, intact = orc.cal_family(x, GOLD["horizon"], 200)
        assert int(np.count_nonzero(fam)) == GOLD[key]["nnz"], key
