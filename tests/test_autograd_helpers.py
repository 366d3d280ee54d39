"""The torch helpers of cudaring_amd.autograd on the CPU: the differentiable shape parameter
against the oracle's getParticleShapes, and the calA0 -> l0 rule against the set-up sizing."""
import numpy as np
import pytest

from cudaring_amd import packing
from tests.helpers import oracle_from_state

torch = pytest.importorskip("torch")


def test_shape_parameter_matches_oracle(oracle_mod):
    from cudaring_amd.autograd import shape_parameter
    st = packing.random_system(10, 0.7, 12)
    rng = np.random.default_rng(0)
    pos = st["pos"].reshape(-1, 2) + 0.05 * (rng.random((len(st["rad"]), 2)) - 0.5)
    # shift part of the system across the periodic boundary: the walk must unwrap it
    pos[:40] += st["box"] * 0.999
    st["pos"] = pos.ravel()
    o = oracle_from_state(oracle_mod, st, interaction="all")
    o.calc_force_energy()
    s3 = o.shapes()
    want = float(np.mean(np.asarray(s3).reshape(-1, 3)[:, 2]))
    got = float(shape_parameter(torch.as_tensor(st["pos"]), st["nvp"], st["box"]))
    assert abs(got - want) < 1e-12 * want


def test_shape_parameter_gradient_fd():
    from cudaring_amd.autograd import shape_parameter
    st = packing.random_system(6, 0.7, 3)
    x = torch.as_tensor(st["pos"]).clone().requires_grad_()
    f = shape_parameter(x, st["nvp"], st["box"])
    f.backward()
    d = torch.as_tensor(np.random.default_rng(1).standard_normal(x.numel()))
    h = 1e-6
    with torch.no_grad():
        fd = (shape_parameter(x + h * d, st["nvp"], st["box"]) - shape_parameter(x - h * d, st["nvp"], st["box"])) / (2 * h)
    assert abs(float(x.grad @ d) - float(fd)) < 1e-6 * abs(float(fd)) + 1e-9  # FD round-off ~1e-10


def test_l0_from_calA0_matches_sizing():
    from cudaring_amd.autograd import l0_from_calA0
    nvp = np.array([16, 22, 16, 22, 19])
    calA0 = np.array([1.0, 1.05, 1.1, 1.15, 1.2])
    a0, l0, _, _ = packing.ring_shape_params(nvp, calA0, n_min=16)
    got = l0_from_calA0(torch.as_tensor(calA0), nvp, a0).numpy()
    np.testing.assert_allclose(got, l0, rtol=1e-14)
