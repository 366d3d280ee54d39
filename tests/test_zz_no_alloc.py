"""The allocation-free entry point dpm_energy_param_grad_noalloc, called through the C ABI: its
results equal dpm_energy_param_grad's bit for bit (the same kernels, a fixed buffer instead of an
allocated one), on two calls in a row whose second set is smaller (the buffer keeps the first
call's partials past the second set's end), and sets larger than the buffer are refused."""
import ctypes as C

import numpy as np
import pytest

from cudaring_amd import DPMSet, packing
from tests.helpers import gpu_from_state

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MAX_VERTICES = 262144  # DPM_NOALLOC_MAX_VERTICES, include/dpmring.h
DPM_ERR_ARG = -1  # include/dpmring.h


def prepared(nrings, seed):
    st = packing.lattice_system(nrings, 0.62, seed, perturb=0.1)
    rng = np.random.default_rng(seed)
    st["a0"] = st["a0"] * (1 + 0.05 * rng.random(len(st["a0"])))
    g = gpu_from_state(st)
    nr = len(st["nvp"])
    g.set_ring_costs(0.5 + rng.random(nr), 0.5 + rng.random(nr), 0.005 + 0.02 * rng.random(nr))
    cut = g.set_displacement_cutoff(0.5, 2 * float(np.mean(st["rad"])))
    g.calc_neighbor_list(cut)
    g.calc_force_energy()
    g.synchronize()
    return g


def grads(g, noalloc):
    fn = g.L.dpm_energy_param_grad_noalloc if noalloc else g.L.dpm_energy_param_grad
    gring = torch.full((3 * g.np,), np.nan, dtype=torch.float64, device="cuda")
    gl0 = torch.full((g.nv,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = fn(g.h, C.c_void_p(gring.data_ptr()), C.c_void_p(gl0.data_ptr()))
    assert rc == 0, g.L.dpm_last_error(g.h)
    return gring.cpu().numpy(), gl0.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_twice():
    print("no-alloc: 36-ring set, then 9-ring set", flush=True)
    big, small = prepared(36, 11), prepared(9, 12)
    assert small.nv < big.nv
    for g in (big, small):  # the second call's set is smaller than the first's
        ref_ring, ref_l0 = grads(g, noalloc=False)
        ring, l0 = grads(g, noalloc=True)
        assert np.all(np.isfinite(ref_ring)) and np.all(np.isfinite(ref_l0))
        assert np.array_equal(ring, ref_ring)
        assert np.array_equal(l0, ref_l0)
    # and back to the larger set: the buffer's leftover partials are overwritten, not reused
    ring, l0 = grads(big, noalloc=True)
    ref_ring, ref_l0 = grads(big, noalloc=False)
    assert np.array_equal(ring, ref_ring) and np.array_equal(l0, ref_l0)
    big.close()
    small.close()


def test_noalloc_refuses_sets_beyond_the_buffer():
    n = MAX_VERTICES // 16 + 1  # rings of 16 vertices: one ring past the limit
    g = DPMSet(np.full(n, 16))
    assert g.nv > MAX_VERTICES
    gring = torch.empty(3 * g.np, dtype=torch.float64, device="cuda")
    gl0 = torch.empty(g.nv, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = g.L.dpm_energy_param_grad_noalloc(g.h, C.c_void_p(gring.data_ptr()), C.c_void_p(gl0.data_ptr()))
    assert rc == DPM_ERR_ARG
    g.close()
