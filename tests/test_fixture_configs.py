"""fixtures/configs/*.json record the parameters the workload descriptions
leave open; the workload generators must produce exactly those."""
import json
import math
import os

from peridyn_amd import workloads as W

D = os.path.join(os.path.dirname(os.path.dirname(__file__)), "fixtures", "configs")


def load(name):
    return json.load(open(os.path.join(D, name + ".json")))


def test_tensile_config():
    c = load("0_tensile_bb")
    s = W.tensile_bb()
    b = s["blocks"][0]
    assert s["dt"] == c["dt"] and b["horizon"] == c["horizon"] and b["critical_stretch"] == c["critical_stretch"]
    assert b["bulk_modulus"] == c["bulk_modulus"] and b["density"] == c["density"]
    assert b["x"].shape[0] == math.prod(c["lattice"])
    assert abs(s["bcs"][1]["rate"][2] * s["dt"] - 0.01) < 1e-15


def test_plates_config():
    c = load("1_plates_ep")
    s = W.plates_ep()
    b = s["blocks"][0]
    assert s["dt"] == c["dt"] and b["yield_stress"] == c["yield_stress"] and b["shear_modulus"] == c["shear_modulus"]
    ct = s["contacts"][0]
    for k in ("exponent", "distanceD", "distanceX", "max_neighs", "stiffness"):
        assert ct[k] == c["contact"][k], k


def test_disk_config():
    c = load("2_disk_impact")
    s = W.disk_impact()
    assert s["dt"] == c["dt"]
    assert s["blocks"][0]["critical_stretch"] == c["disk_critical_stretch"]
    assert s["blocks"][1]["bulk_modulus"] == c["sphere_bulk_modulus"]
    assert abs(-s["blocks"][1]["v"][0, 2] * s["dt"] - c["speed_per_step"]) < 1e-12


def test_batched_and_served_configs():
    c3 = load("3_batched")
    specs = W.batched_specimens(c3["count"], seed=c3["seed"])
    assert len(specs) == 256 and all(s["dt"] == c3["dt"] for s in specs)
    c4 = load("4_served")
    reqs = W.served_requests(count=c4["count"], seed=c4["seed"], rate=c4["arrival_rate_per_s"])
    assert all(r["steps"] % c4["neigh_update_freq"] == 0 and 200 <= r["steps"] <= 2000 for r in reqs)
    assert all(2.5 <= r["horizon"] <= 4.0 and 1000 <= r["n"] <= 200000 for r in reqs)


def test_served_config_matches_bench_defaults():
    """fixtures/configs/4_served.json describes the stream bench.py serves by default."""
    import bench
    c4 = load("4_served")
    d = bench.make_parser().parse_args([])
    assert d.served_rate == c4["arrival_rate_per_s"]
    assert d.served_precision == c4["precision"]
    assert d.served_lanes == c4["lanes_per_particle"]
    assert c4["count"] == 64 and c4["chunk_steps"] == 50
