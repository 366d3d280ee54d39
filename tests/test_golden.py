"""The committed fixtures in tests/golden/golden_small.npz (written by tests/golden/make_golden.py):
the oracle must reproduce them (CPU), and the library must match them to the parity bars (GPU)."""
import os

import numpy as np
import pytest

from tests.conftest import ROOT, rel_err

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_small.npz")
POTS = ["harmonic", "wca", "lj", "adhesive"]
INTERS = ["vertexVertex", "vertexSmooth"]
COSTS = (1.0, 1.0, 0.01, 1.0)


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def state(gold, prefix):
    return {k: gold[f"{prefix}_{k}"] for k in ("nvp", "box", "pos", "rad", "l0", "theta0", "a0")}


def setup(obj, st, pot="harmonic", inter="vertexVertex", dt=5e-3, seed=7, oracle=False):
    obj.load_state(st)
    obj.set_types(pot, inter)
    obj.set_energy_costs(*COSTS)
    obj.set_attraction(0.1, 0.2)
    obj.set_lj_cutoff(1.5)
    obj.set_dt(dt)
    if oracle:
        obj.set_rng(seed, 0)
    else:
        obj.set_seed(seed)
    cut = obj.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    obj.calc_neighbor_list(cut)
    return obj


def run_cases(make, gold, oracle):
    st, lat = state(gold, "state"), state(gold, "lattice")
    out = {}
    for pot in POTS:
        for inter in INTERS:
            o = setup(make(st["nvp"]), st, pot, inter, oracle=oracle)
            o.calc_force_energy()
            key = f"{pot}_{inter}"
            out[f"force_{key}"], out[f"energy_{key}"] = o.get("force"), o.get("energy")
            out[f"penergy_{key}"] = o.get("penergy")
            u = o.potential_energy()
            out[f"U_{key}"] = np.atleast_1d(u)[:1]
    o = setup(make(lat["nvp"]), lat, oracle=oracle)
    o.calc_force_energy()
    o.init_nve(1e-3, zero_momentum=True)
    o.nve(50)
    out["nve50_pos"], out["nve50_vel"] = o.get("pos"), o.get("vel")
    o = setup(make(lat["nvp"]), lat, seed=9, oracle=oracle)
    o.calc_force_energy()
    o.init_langevin(1e-2, 1.0)
    o.langevin(50)
    out["langevin50_pos"], out["langevin50_vel"] = o.get("pos"), o.get("vel")
    o = setup(make(st["nvp"]), st, oracle=oracle)
    o.calc_force_energy()
    o.init_fire([0.2, 0.5, 1.1, 0.99, 5e-3, 5e-2, 0.2], 20, 0)
    o.fire(100)
    fs = o.fire_state()
    out["fire100_pos"] = o.get("pos")
    out["fire100_state"] = np.array([np.atleast_1d(fs["dt"])[0], np.atleast_1d(fs["alpha"])[0],
                                     np.atleast_1d(fs["numStep"])[0]], dtype=np.float64)
    return out


def test_oracle_reproduces_golden(oracle_mod, gold):
    out = run_cases(lambda nvp: oracle_mod.Oracle(nvp), gold, oracle=True)
    for k, v in out.items():
        np.testing.assert_allclose(v, gold[k], rtol=1e-13, atol=1e-15, err_msg=k)


@pytest.mark.gpu
def test_library_matches_golden(gold):
    from cudaring_amd import DPMSet
    out = run_cases(lambda nvp: DPMSet(nvp), gold, oracle=False)
    for pot in POTS:
        for inter in INTERS:
            key = f"{pot}_{inter}"
            assert rel_err(out[f"force_{key}"], gold[f"force_{key}"]) < 1e-12, key
            assert rel_err(out[f"energy_{key}"], gold[f"energy_{key}"]) < 1e-12, key
            assert rel_err(out[f"penergy_{key}"], gold[f"penergy_{key}"]) < 1e-12, key
            assert abs(out[f"U_{key}"][0] - gold[f"U_{key}"][0]) <= 1e-12 * abs(gold[f"U_{key}"][0]), key
    for k in ("nve50_pos", "langevin50_pos", "fire100_pos"):
        assert rel_err(out[k], gold[k], scale=1.0) < 1e-11, k
    for k in ("nve50_vel", "langevin50_vel"):
        assert rel_err(out[k], gold[k]) < 1e-9, k
    np.testing.assert_allclose(out["fire100_state"], gold["fire100_state"], rtol=1e-13)
