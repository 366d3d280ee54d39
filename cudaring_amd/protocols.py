"""Packing preparation: the compressFIRE.cpp workflow (compressFIRE.cpp:47-200 of
farceri/cudaRing) driven through a DPMSet.

Rings start as regular polygons on a dilute square lattice (no ring crosses
another), then grow about their centres in steps of delta phi (scaleVertices,
DPM2D.cu:1802-1811) with a FIRE relaxation after each step, until the preferred
packing fraction reaches the target.  Works on any object with the DPMSet
method names, for one or many systems at once.
"""
import numpy as np

from . import packing

FIRE_PARAMS = (0.2, 0.5, 1.1, 0.99)  # a_start, f_dec, f_inc, f_a (compressFIRE.cpp:106)


def fire_relax(sys_, dt, iters, chunk=500, tol=0.0):
    prm = [FIRE_PARAMS[0], FIRE_PARAMS[1], FIRE_PARAMS[2], FIRE_PARAMS[3], dt, 10 * dt, FIRE_PARAMS[0]]
    sys_.init_fire(prm, 20, 0)
    done = 0
    while done < iters:
        sys_.fire(min(chunk, iters - done))
        done += chunk
        if tol > 0 and np.max(sys_.max_unbalanced_force()) < tol:
            break


def compress(sys_, phi_start, phi_target, dphi=0.02, iters=2000, skin=0.3, progress=None):
    """Grow every ring of every system until the preferred phi reaches phi_target;
    progress(phi), if given, is called after each relaxed compression step."""
    phi = phi_start
    while phi < phi_target - 1e-12:
        nxt = min(phi + dphi, phi_target)
        sys_.calc_particle_shape()
        sys_.scale_vertices(np.sqrt(nxt / phi))
        size = 2 * float(np.mean(sys_.get("rad")))
        cut = sys_.set_displacement_cutoff(skin, size)
        sys_.calc_neighbor_list(cut)
        sys_.calc_force_energy()
        dt = 0.01 * size * 8
        fire_relax(sys_, dt, iters)
        phi = nxt
        if progress is not None:
            progress(phi)
    return phi


def prepared_state(make_set, nrings, phi, seed, calA0=1.10, costs=(1, 1, 0.01, 1), phi_start=0.45,
                   dphi=0.02, iters=2000, final=("harmonic", "vertexVertex"), final_iters=6000):
    """A relaxed bidisperse packing at preferred phi; make_set(nvp) builds the system object.

    Compression runs with harmonic vertex-vertex repulsion; the final (potential,
    interaction) pair is then relaxed by `final_iters` FIRE iterations so that a
    stiff potential (WCA) starts without deep overlaps.
    Returns (system object, state dict of the relaxed configuration)."""
    st = packing.lattice_system(nrings, phi_start, seed, calA0=calA0, perturb=0.0)
    s = make_set(st["nvp"])
    s.load_state(st)
    s.set_types("harmonic", "vertexVertex")
    s.set_energy_costs(*costs)
    compress(s, phi_start, phi, dphi=dphi, iters=iters)
    if tuple(final) != ("harmonic", "vertexVertex"):
        s.set_types(*final)
        size = 2 * float(np.mean(s.get("rad")))
        cut = s.set_displacement_cutoff(0.5, size)
        s.calc_neighbor_list(cut)
        s.calc_force_energy()
        fire_relax(s, 0.01 * size, final_iters)
    out = dict(st)
    for k in ("pos", "rad", "l0", "a0"):
        out[k] = s.get(k)
    out["theta0"] = st["theta0"]
    return s, out
