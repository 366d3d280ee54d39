"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product package
(cudaring_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

POT = {"harmonic": 0, "lj": 1, "adhesive": 2, "wca": 3}
INTER = {"vertexVertex": 0, "vertexSmooth": 1, "cellSmooth": 2, "all": 3}
MONO = {"harmonic": 0, "FENE": 1}
PART = {"passive": 0, "active": 1, "activeShape": 2}
NOISE = {"langevin": 0, "baoab": 1, "brownian": 2, "drivenBrownian": 3}
GEOM = {"normal": 0, "fixedBox": 1, "fixedSides": 2, "roundBox": 3}
FIELDS = ["pos", "vel", "force", "energy", "rad", "l0", "theta0", "theta", "length",
          "a0", "area", "perimeter", "ppos", "penergy", "pangle", "lastpos", "box", "wallforce",
          "pvel", "pangvel", "pforce", "ptorque", "vtorque", "prad"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def use_library(path):
    """Load the oracle from `path` (a -march=native build for the CPU baseline) instead of the
    in-tree liboracle.so; must be called before the first Oracle is created."""
    global _LIB_PATH, _lib
    if _lib is not None and _LIB_PATH != path:
        raise RuntimeError("oracle library already loaded from " + _LIB_PATH)
    _LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d, l, vp, i = C.c_double, C.c_long, C.c_void_p, C.c_int
        dp = C.POINTER(C.c_double)
        sig = {
            "orc_create": (vp, [l, vp]),
            "orc_destroy": (None, [vp]),
            "orc_num_vertices": (l, [vp]),
            "orc_field_size": (l, [vp, i]),
            "orc_set_field": (i, [vp, i, vp, l]),
            "orc_get_field": (i, [vp, i, vp, l]),
            "orc_set_types": (None, [vp, i, i, i, i]),
            "orc_set_energy_costs": (None, [vp, d, d, d, d]),
            "orc_set_ring_costs": (None, [vp, vp, vp, vp]),
            "orc_set_attraction_constants": (None, [vp, d, d]),
            "orc_set_lj_cutoff": (None, [vp, d]),
            "orc_set_fene_constants": (None, [vp, d, d]),
            "orc_set_time_step": (None, [vp, d]),
            "orc_set_self_propulsion": (None, [vp, d, d]),
            "orc_set_rng": (None, [vp, C.c_uint64, C.c_uint32]),
            "orc_pbc_distance": (d, [d, d, d]),
            "orc_next_id": (l, [vp, l, l]),
            "orc_previous_id": (l, [vp, l, l]),
            "orc_calc_particle_shape": (None, [vp]),
            "orc_calc_particle_positions": (None, [vp]),
            "orc_calc_shape_force_energy": (None, [vp]),
            "orc_pair_grad_multiple_and_energy": (d, [vp, d, d, dp]),
            "orc_calc_vertex_interaction": (None, [vp]),
            "orc_calc_all_to_all_interaction": (None, [vp]),
            "orc_calc_smooth_interaction": (None, [vp]),
            "orc_add_self_propulsion": (None, [vp]),
            "orc_calc_force_energy": (None, [vp]),
            "orc_set_displacement_cutoff": (d, [vp, d, d]),
            "orc_calc_neighbor_list": (None, [vp, d]),
            "orc_check_neighbors": (i, [vp]),
            "orc_neighbor_list_size": (l, [vp]),
            "orc_get_neighbors": (l, [vp, vp, vp, l]),
            "orc_get_update_count": (l, [vp]),
            "orc_reset_update_count": (None, [vp]),
            "orc_reset_last_positions": (None, [vp]),
            "orc_get_cut_distance": (d, [vp]),
            "orc_set_list_method": (None, [vp, i]),
            "orc_set_threads": (None, [vp, i]),
            "orc_init_langevin": (None, [vp, d, d, i, i]),
            "orc_set_noise_type": (None, [vp, i]),
            "orc_calc_contacts": (None, [vp]),
            "orc_calc_geometric_contacts": (None, [vp, d]),
            "orc_contact_list_size": (l, [vp]),
            "orc_get_contacts": (None, [vp, vp, vp]),
            "orc_get_hexatic_order": (d, [vp]),
            "orc_set_initial_positions": (None, [vp]),
            "orc_get_vertex_msd": (d, [vp]),
            "orc_get_particle_msd": (d, [vp]),
            "orc_get_vertex_isf": (d, [vp]),
            "orc_get_particle_isf": (d, [vp, d]),
            "orc_calc_per_particle_stress": (None, [vp, vp]),
            "orc_calc_stress_tensor": (None, [vp, vp]),
            "orc_get_pressure": (d, [vp]),
            "orc_set_geometry": (None, [vp, i]),
            "orc_set_wall_energy_scale": (None, [vp, d]),
            "orc_set_box_radius": (None, [vp, d]),
            "orc_get_box_radius": (d, [vp]),
            "orc_set_active_tension": (None, [vp, d, d]),
            "orc_update_rest_length": (None, [vp]),
            "orc_init_nve": (None, [vp, d, i, i]),
            "orc_init_nve_rescale": (None, [vp, d]),
            "orc_inject_kinetic_energy": (None, [vp, i]),
            "orc_langevin_loop": (None, [vp, l]),
            "orc_nve_loop": (None, [vp, l]),
            "orc_nve_rescale_loop": (None, [vp, l]),
            "orc_init_fire": (None, [vp, vp, l, l]),
            "orc_fire_loop": (None, [vp, l]),
            "orc_get_fire_state": (None, [vp, vp]),
            "orc_fire_minimize": (l, [vp, d, l]),
            "orc_get_potential_energy": (d, [vp]),
            "orc_get_kinetic_energy": (d, [vp]),
            "orc_get_temperature": (d, [vp]),
            "orc_get_max_unbalanced_force": (d, [vp]),
            "orc_get_total_force_magnitude": (d, [vp]),
            "orc_get_phi": (d, [vp]),
            "orc_get_preferred_phi": (d, [vp]),
            "orc_get_mean_particle_size": (d, [vp]),
            "orc_get_mean_vertex_radius": (d, [vp]),
            "orc_adjust_temperature": (None, [vp, d]),
            "orc_adjust_kinetic_energy": (d, [vp, d]),
            "orc_scale_vertices": (None, [vp, d]),
            "orc_scale_box": (None, [vp, d]),
            "orc_calc_rigid_force_energy": (None, [vp]),
            "orc_init_rigid": (None, [vp, d, d, i, i]),
            "orc_rigid_loop": (None, [vp, l]),
            "orc_init_rigid_fire": (None, [vp, vp, l, l]),
            "orc_rigid_fire_loop": (None, [vp, l]),
            "orc_get_rigid_max_unbalanced_force": (d, [vp]),
            "orc_get_rigid_kinetic_energy": (d, [vp]),
            "orc_get_particle_shapes": (None, [vp, vp]),
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_gauss_pair": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One DPM2D system held by the CPU oracle (AoS layout as in the reference)."""

    def __init__(self, nvp):
        self.nvp = np.ascontiguousarray(nvp, dtype=np.int64)
        self.L = lib()
        self.h = self.L.orc_create(len(self.nvp), _ptr(self.nvp))
        self.np = len(self.nvp)
        self.nv = int(self.nvp.sum())

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        rc = self.L.orc_set_field(self.h, FIELD_ID[name], _ptr(a), a.size)
        if rc != 0:
            raise ValueError(f"oracle set {name}: size {a.size}")

    def get(self, name):
        n = self.L.orc_field_size(self.h, FIELD_ID[name])
        a = np.empty(n, dtype=np.float64)
        self.L.orc_get_field(self.h, FIELD_ID[name], _ptr(a), n)
        return a

    def load_state(self, st):
        """st: dict with box, pos, rad, l0, theta0, a0 (+ optional vel, pangle)."""
        for k in ("box", "pos", "rad", "l0", "theta0", "a0", "vel", "pangle"):
            if k in st and st[k] is not None:
                self.set(k, st[k])

    def set_types(self, potential="harmonic", interaction="vertexVertex", monomer="harmonic",
                  particle="passive"):
        self.L.orc_set_types(self.h, POT[potential], INTER[interaction], MONO[monomer], PART[particle])

    def set_energy_costs(self, ea, el, eb, ec):
        self.L.orc_set_energy_costs(self.h, ea, el, eb, ec)

    def set_ring_costs(self, ka=None, kl=None, kb=None):
        if ka is None:
            self.L.orc_set_ring_costs(self.h, None, None, None)
            return
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (self.np,)))
                for a in (ka, kl, kb)]
        self.L.orc_set_ring_costs(self.h, *[_ptr(a) for a in arrs])

    def set_attraction(self, l1, l2):
        self.L.orc_set_attraction_constants(self.h, l1, l2)

    def set_lj_cutoff(self, c):
        self.L.orc_set_lj_cutoff(self.h, c)

    def set_dt(self, dt):
        self.L.orc_set_time_step(self.h, dt)

    def set_self_propulsion(self, driving, taup):
        self.L.orc_set_self_propulsion(self.h, driving, taup)

    def set_rng(self, seed, system_id=0):
        self.L.orc_set_rng(self.h, seed, system_id)

    def set_displacement_cutoff(self, cutoff, size):
        return self.L.orc_set_displacement_cutoff(self.h, cutoff, size)

    def calc_neighbor_list(self, cut):
        self.L.orc_calc_neighbor_list(self.h, cut)

    def set_list_method(self, cells=True):
        """Row build of calcNeighborList: cell list (default) or the reference's all-pairs loop."""
        self.L.orc_set_list_method(self.h, 1 if cells else 0)

    def set_threads(self, n):
        self.L.orc_set_threads(self.h, n)

    def check_neighbors(self):
        return self.L.orc_check_neighbors(self.h)

    def neighbors(self):
        """List of sorted neighbour lists (first `count` entries of each row) and W."""
        W = self.L.orc_neighbor_list_size(self.h)
        cap = max(W, 1)
        lst = np.empty(self.nv * cap, dtype=np.int64)
        cnt = np.empty(self.nv, dtype=np.int64)
        self.L.orc_get_neighbors(self.h, _ptr(lst), _ptr(cnt), cap)
        lst = lst.reshape(self.nv, cap)
        return [list(lst[v, :cnt[v]]) for v in range(self.nv)], W, cnt

    def calc_force_energy(self):
        self.L.orc_calc_force_energy(self.h)

    def potential_energy(self):
        return self.L.orc_get_potential_energy(self.h)

    def kinetic_energy(self):
        return self.L.orc_get_kinetic_energy(self.h)

    def temperature(self):
        return self.L.orc_get_temperature(self.h)

    def max_unbalanced_force(self):
        return self.L.orc_get_max_unbalanced_force(self.h)

    def phi(self):
        return self.L.orc_get_phi(self.h)

    def preferred_phi(self):
        return self.L.orc_get_preferred_phi(self.h)

    def mean_particle_size(self):
        return self.L.orc_get_mean_particle_size(self.h)

    def mean_vertex_radius(self):
        return self.L.orc_get_mean_vertex_radius(self.h)

    def set_geometry(self, geometry, ew=None, box_radius=None):
        self.L.orc_set_geometry(self.h, GEOM[geometry])
        if ew is not None:
            self.L.orc_set_wall_energy_scale(self.h, ew)
        if box_radius is not None:
            self.L.orc_set_box_radius(self.h, box_radius)

    # ---- observables
    def contacts(self):
        """(list [np, width] -1 padded, counts [np]) after calcContacts."""
        self.L.orc_calc_contacts(self.h)
        w = self.L.orc_contact_list_size(self.h)
        lst = np.empty(self.np * w, dtype=np.int64)
        cnt = np.empty(self.np, dtype=np.int64)
        self.L.orc_get_contacts(self.h, _ptr(lst), _ptr(cnt))
        return lst.reshape(self.np, w), cnt

    def geometric_contacts(self, gap=1.0):
        """(list [np, 8] -1 padded, counts [np]) after calcGeometricContacts."""
        self.L.orc_calc_geometric_contacts(self.h, gap)
        w = self.L.orc_contact_list_size(self.h)
        lst = np.empty(self.np * w, dtype=np.int64)
        cnt = np.empty(self.np, dtype=np.int64)
        self.L.orc_get_contacts(self.h, _ptr(lst), _ptr(cnt))
        return lst.reshape(self.np, w), cnt

    def hexatic_order(self):
        return self.L.orc_get_hexatic_order(self.h)

    def set_initial_positions(self):
        self.L.orc_set_initial_positions(self.h)

    def vertex_msd(self):
        return self.L.orc_get_vertex_msd(self.h)

    def particle_msd(self):
        return self.L.orc_get_particle_msd(self.h)

    def vertex_isf(self):
        return self.L.orc_get_vertex_isf(self.h)

    def particle_isf(self, k):
        return self.L.orc_get_particle_isf(self.h, k)

    def per_particle_stress(self):
        a = np.empty(4 * self.np)
        self.L.orc_calc_per_particle_stress(self.h, _ptr(a))
        return a.reshape(self.np, 4)

    def stress_tensor(self):
        a = np.empty(4)
        self.L.orc_calc_stress_tensor(self.h, _ptr(a))
        return a

    def pressure(self):
        return self.L.orc_get_pressure(self.h)

    def set_noise_type(self, noise):
        self.L.orc_set_noise_type(self.h, NOISE[noise])

    def set_active_tension(self, l0_tau, l0_diff):
        self.L.orc_set_active_tension(self.h, l0_tau, l0_diff)

    def init_langevin(self, T, gamma, read_state=False, zero_momentum=False):
        self.L.orc_init_langevin(self.h, T, gamma, int(read_state), int(zero_momentum))

    def init_nve(self, T, read_state=False, zero_momentum=False):
        self.L.orc_init_nve(self.h, T, int(read_state), int(zero_momentum))

    def init_nve_rescale(self, T):
        self.L.orc_init_nve_rescale(self.h, T)

    def langevin(self, n):
        self.L.orc_langevin_loop(self.h, n)

    def nve(self, n):
        self.L.orc_nve_loop(self.h, n)

    def nve_rescale(self, n):
        self.L.orc_nve_rescale_loop(self.h, n)

    def init_fire(self, params7, min_step=20, num_step=0):
        p = np.ascontiguousarray(params7, dtype=np.float64)
        self.L.orc_init_fire(self.h, _ptr(p), min_step, num_step)

    def fire(self, n):
        self.L.orc_fire_loop(self.h, n)

    def fire_state(self):
        o = np.zeros(5)
        self.L.orc_get_fire_state(self.h, _ptr(o))
        return {"dt": o[0], "alpha": o[1], "numStep": int(o[2]), "P": o[3], "updates": int(o[4])}

    def fire_minimize(self, tol, max_iter):
        return self.L.orc_fire_minimize(self.h, tol, max_iter)

    def adjust_temperature(self, T):
        self.L.orc_adjust_temperature(self.h, T)

    def adjust_kinetic_energy(self, prev_etot):
        return self.L.orc_adjust_kinetic_energy(self.h, prev_etot)

    def total_force_magnitude(self):
        return self.L.orc_get_total_force_magnitude(self.h)

    def calc_smooth_interaction(self):
        self.L.orc_calc_smooth_interaction(self.h)

    def calc_all_to_all_interaction(self):
        self.L.orc_calc_all_to_all_interaction(self.h)

    def scale_vertices(self, s):
        self.L.orc_scale_vertices(self.h, s)

    def scale_box(self, s):
        self.L.orc_scale_box(self.h, s)

    def box_radius(self):
        return self.L.orc_get_box_radius(self.h)

    # ---- rigid rings
    def calc_rigid_force_energy(self):
        self.L.orc_calc_rigid_force_energy(self.h)

    def init_rigid_langevin(self, T, gamma, read_state=False):
        self.L.orc_init_rigid(self.h, T, gamma, 1, int(read_state))

    def init_rigid_nve(self, T, read_state=False):
        self.L.orc_init_rigid(self.h, T, 0.0, 0, int(read_state))

    def rigid_loop(self, steps):
        self.L.orc_rigid_loop(self.h, steps)

    def init_rigid_fire(self, params, min_step=20, num_step=0):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self.L.orc_init_rigid_fire(self.h, _ptr(p), min_step, num_step)

    def rigid_fire(self, iterations):
        self.L.orc_rigid_fire_loop(self.h, iterations)

    def rigid_max_unbalanced_force(self):
        return self.L.orc_get_rigid_max_unbalanced_force(self.h)

    def rigid_kinetic_energy(self):
        return self.L.orc_get_rigid_kinetic_energy(self.h)

    def shapes(self):
        o = np.zeros(3 * self.np)
        self.L.orc_get_particle_shapes(self.h, _ptr(o))
        return o.reshape(self.np, 3)

    def update_count(self):
        return self.L.orc_get_update_count(self.h)

    def reset_update_count(self):
        self.L.orc_reset_update_count(self.h)

    def reset_last_positions(self):
        """resetLastPositions: the list positions become the current positions."""
        self.L.orc_reset_last_positions(self.h)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def gauss_pair(seed, system_id, tag, counter, index):
    z0, z1 = C.c_double(), C.c_double()
    lib().orc_gauss_pair(seed, system_id, tag, counter, index, C.byref(z0), C.byref(z1))
    return z0.value, z1.value


def pbc_distance(x1, x2, L):
    return lib().orc_pbc_distance(x1, x2, L)
