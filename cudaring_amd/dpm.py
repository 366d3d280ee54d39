"""Python binding of libdpmring.so (the C ABI in include/dpmring.h).

`DPMSet` holds B >= 1 independent DPM2D systems on one GPU.  Method names follow
the reference's DPM2D class (include/DPM2D.h of farceri/cudaRing); per-system
observables return numpy arrays of length B.  The library is required: if it
is missing or fails to load this module raises, there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdpmring.so")

FIELDS = ["pos", "vel", "force", "energy", "rad", "l0", "theta0", "theta", "length", "a0", "area",
          "perimeter", "ppos", "penergy", "pangle", "lastpos", "box", "wallforce",
          "pvel", "pangvel", "pforce", "ptorque", "vtorque", "prad"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
POT = {"harmonic": 0, "lj": 1, "adhesive": 2, "wca": 3}
INTER = {"vertexVertex": 0, "vertexSmooth": 1, "all": 3}
MONO = {"harmonic": 0, "FENE": 1}
PART = {"passive": 0, "active": 1, "activeShape": 2}
NOISE = {"langevin": 0, "baoab": 1, "brownian": 2, "drivenBrownian": 3}
GEOM = {"normal": 0, "fixedBox": 1, "fixedSides": 2, "roundBox": 3}
ERRORS = {-1: "bad argument", -2: "HIP error", -3: "out of memory", -4: "unsupported",
          -5: "overflow", -6: "bad state"}

# every exported symbol with (restype, argtypes); tests check this list against the header
_vp, _d, _l, _i = C.c_void_p, C.c_double, C.c_long, C.c_int
SIGNATURES = {
    "dpm_create": (_i, [C.POINTER(_vp), _i, _l, _vp, _vp]),
    "dpm_destroy": (_i, [_vp]),
    "dpm_set_stream": (_i, [_vp, _vp]),
    "dpm_last_error": (C.c_char_p, [_vp]),
    "dpm_get_sizes": (_i, [_vp, _vp, _vp, _vp]),
    "dpm_synchronize": (_i, [_vp]),
    "dpm_set_field": (_i, [_vp, _i, _vp, _l]),
    "dpm_get_field": (_i, [_vp, _i, _vp, _l]),
    "dpm_device_ptr": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "dpm_set_types": (_i, [_vp, _i, _i, _i, _i]),
    "dpm_set_energy_costs": (_i, [_vp, _d, _d, _d, _d]),
    "dpm_set_ring_costs": (_i, [_vp, _vp, _vp, _vp]),
    "dpm_set_attraction_constants": (_i, [_vp, _d, _d]),
    "dpm_set_lj_cutoff": (_i, [_vp, _d]),
    "dpm_set_fene_constants": (_i, [_vp, _d, _d]),
    "dpm_set_time_step": (_i, [_vp, _d]),
    "dpm_set_self_propulsion": (_i, [_vp, _d, _d]),
    "dpm_set_active_tension": (_i, [_vp, _d, _d]),
    "dpm_set_noise_type": (_i, [_vp, _i]),
    "dpm_set_geometry": (_i, [_vp, _i]),
    "dpm_set_initial_positions": (_i, [_vp]),
    "dpm_get_vertex_msd": (_i, [_vp, _vp]),
    "dpm_get_particle_msd": (_i, [_vp, _vp]),
    "dpm_get_vertex_isf": (_i, [_vp, _vp]),
    "dpm_get_particle_isf": (_i, [_vp, _d, _vp]),
    "dpm_calc_contacts": (_i, [_vp]),
    "dpm_calc_geometric_contacts": (_i, [_vp]),
    "dpm_get_contacts": (_i, [_vp, _vp, _vp, _l, _vp]),
    "dpm_get_hexatic_order": (_i, [_vp, _vp]),
    "dpm_get_per_particle_stress": (_i, [_vp, _vp]),
    "dpm_get_stress_tensor": (_i, [_vp, _vp]),
    "dpm_get_pressure": (_i, [_vp, _vp]),
    "dpm_set_box_energy_cost": (_i, [_vp, _d]),
    "dpm_set_box_radius": (_i, [_vp, _d]),
    "dpm_get_box_radius": (_i, [_vp, _vp]),
    "dpm_set_seed": (_i, [_vp, C.c_uint64]),
    "dpm_set_displacement_cutoff": (_i, [_vp, _d, _d, C.POINTER(_d)]),
    "dpm_calc_neighbor_list": (_i, [_vp, _d]),
    "dpm_set_cell_scale": (_i, [_vp, _d]),
    "dpm_check_neighbors": (_i, [_vp]),
    "dpm_get_neighbor_list": (_i, [_vp, _vp, _vp, _l, C.POINTER(_l)]),
    "dpm_reset_last_positions": (_i, [_vp]),
    "dpm_get_update_count": (_i, [_vp, _vp]),
    "dpm_reset_update_count": (_i, [_vp]),
    "dpm_calc_particle_shape": (_i, [_vp]),
    "dpm_calc_force_energy": (_i, [_vp]),
    "dpm_get_interaction_forces": (_i, [_vp, _vp, _l]),
    "dpm_init_langevin": (_i, [_vp, _d, _d, _i, _i]),
    "dpm_inject_kinetic_energy": (_i, [_vp, _d]),
    "dpm_update_position": (_i, [_vp, _d]),
    "dpm_update_velocity": (_i, [_vp, _d]),
    "dpm_update_thermal_vel": (_i, [_vp]),
    "dpm_conserve_momentum": (_i, [_vp]),
    "dpm_langevin_loop": (_i, [_vp, _l]),
    "dpm_init_nve": (_i, [_vp, _d, _i, _i]),
    "dpm_nve_loop": (_i, [_vp, _l]),
    "dpm_init_nve_rescale": (_i, [_vp, _d]),
    "dpm_nve_rescale_loop": (_i, [_vp, _l]),
    "dpm_adjust_temperature": (_i, [_vp, _d]),
    "dpm_adjust_kinetic_energy": (_i, [_vp, _vp, _vp]),
    "dpm_init_fire": (_i, [_vp, _vp, _l, _l]),
    "dpm_fire_loop": (_i, [_vp, _l]),
    "dpm_get_fire_state": (_i, [_vp, _vp]),
    "dpm_calc_rigid_force_energy": (_i, [_vp]),
    "dpm_init_rigid_langevin": (_i, [_vp, _d, _d, _i]),
    "dpm_rigid_langevin_loop": (_i, [_vp, _l]),
    "dpm_init_rigid_nve": (_i, [_vp, _d, _i]),
    "dpm_rigid_nve_loop": (_i, [_vp, _l]),
    "dpm_init_rigid_fire": (_i, [_vp, _vp, _l, _l]),
    "dpm_rigid_fire_loop": (_i, [_vp, _l]),
    "dpm_get_rigid_fire_state": (_i, [_vp, _vp]),
    "dpm_get_rigid_max_unbalanced_force": (_i, [_vp, _vp]),
    "dpm_get_rigid_kinetic_energy": (_i, [_vp, _vp]),
    "dpm_fire_minimize": (_i, [_vp, _d, _l, _vp]),
    "dpm_set_fire_kernel": (_i, [_vp, _i, _i]),
    "dpm_get_potential_energy": (_i, [_vp, _vp]),
    "dpm_get_kinetic_energy": (_i, [_vp, _vp]),
    "dpm_get_temperature": (_i, [_vp, _vp]),
    "dpm_get_max_unbalanced_force": (_i, [_vp, _vp]),
    "dpm_get_total_force_magnitude": (_i, [_vp, _vp]),
    "dpm_get_phi": (_i, [_vp, _vp]),
    "dpm_get_preferred_phi": (_i, [_vp, _vp]),
    "dpm_get_particle_shapes": (_i, [_vp, _vp]),
    "dpm_scale_vertices": (_i, [_vp, _vp]),
    "dpm_scale_box": (_i, [_vp, _vp]),
    "dpm_force_hvp": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "dpm_energy_param_grad": (_i, [_vp, _vp, _vp]),
    "dpm_energy_param_grad_noalloc": (_i, [_vp, _vp, _vp]),
    "dpm_nve_forward": (_i, [_vp, _l, _l]),
    "dpm_nve_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dpm_fire_forward": (_i, [_vp, _l]),
    "dpm_fire_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dpm_time_step_kernel": (_i, [_vp, _l, C.POINTER(_d), C.POINTER(_d)]),
    "dpm_set_launch_config": (_i, [_vp, _i]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libdpmring.so and declare every entry point.  Raises if it is absent."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class DPMError(RuntimeError):
    pass


class DPMSet:
    """B independent DPM2D systems on one device (B = 1 is the reference's single DPM2D)."""

    def __init__(self, nvp, rings_per_system=None, device=0):
        self.L = load_library()
        nvp = np.ascontiguousarray(nvp, dtype=np.int64)
        if rings_per_system is None:
            rings_per_system = [len(nvp)]
        rps = np.ascontiguousarray(rings_per_system, dtype=np.int64)
        h = C.c_void_p()
        rc = self.L.dpm_create(C.byref(h), device, len(rps), _p(rps), _p(nvp))
        if rc != 0:
            raise DPMError(f"dpm_create failed: {ERRORS.get(rc, rc)}")
        self.h = h
        self.nvp = nvp
        self.rps = rps
        self.nsys = len(rps)
        self.np = len(nvp)
        self.nv = int(nvp.sum())
        ring_off = np.concatenate([[0], np.cumsum(rps)])
        vfirst = np.concatenate([[0], np.cumsum(nvp)])
        self.vert_off = vfirst[ring_off]
        self.ring_off = ring_off
        self.tape_serial = 0  # number of nve_forward tapes recorded (nve_backward reads the last)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.dpm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.dpm_last_error(self.h)
            raise DPMError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    # ---------------------------------------------------------------- fields
    def field_size(self, name):
        n = {"v": self.nv, "r": self.np, "s": self.nsys}
        kind = {"pos": ("v", 2), "vel": ("v", 2), "force": ("v", 2), "lastpos": ("v", 2),
                "wallforce": ("v", 2), "ppos": ("r", 2), "box": ("s", 2), "pvel": ("r", 2), "pforce": ("r", 2)}
        if name in kind:
            k, c = kind[name]
            return n[k] * c
        if name in ("a0", "area", "perimeter", "penergy", "pangle", "pangvel", "ptorque", "prad"):
            return self.np
        return self.nv

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._chk(self.L.dpm_set_field(self.h, FIELD_ID[name], _p(a), a.size), f"set {name}")

    def get(self, name):
        a = np.empty(self.field_size(name), dtype=np.float64)
        self._chk(self.L.dpm_get_field(self.h, FIELD_ID[name], _p(a), a.size), f"get {name}")
        return a

    def device_ptr(self, name, component=0):
        p = C.c_void_p()
        self._chk(self.L.dpm_device_ptr(self.h, FIELD_ID[name], component, C.byref(p)), "device_ptr")
        return p.value

    def load_state(self, st):
        for k in ("box", "rad", "pos", "l0", "theta0", "a0", "vel", "pangle"):
            if k in st and st[k] is not None:
                self.set(k, st[k])

    # ------------------------------------------------------------ parameters
    def set_types(self, potential="harmonic", interaction="vertexVertex", monomer="harmonic",
                  particle="passive"):
        self._chk(self.L.dpm_set_types(self.h, POT[potential], INTER[interaction], MONO[monomer],
                                       PART[particle]), "set_types")
        self.particle = particle

    def set_energy_costs(self, ea, el, eb, ec):
        self._chk(self.L.dpm_set_energy_costs(self.h, ea, el, eb, ec), "set_energy_costs")

    def set_ring_costs(self, ka=None, kl=None, kb=None):
        """Per-ring ka, kl, kb (arrays of numParticles); no arguments returns to the global costs."""
        if ka is None:
            self._chk(self.L.dpm_set_ring_costs(self.h, None, None, None), "set_ring_costs")
            return
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (self.np,)))
                for a in (ka, kl, kb)]
        self._chk(self.L.dpm_set_ring_costs(self.h, *[_p(a) for a in arrs]), "set_ring_costs")

    def set_attraction(self, l1, l2):
        self._chk(self.L.dpm_set_attraction_constants(self.h, l1, l2), "set_attraction")

    def set_lj_cutoff(self, c):
        self._chk(self.L.dpm_set_lj_cutoff(self.h, c), "set_lj_cutoff")

    def set_fene(self, stiff, ext):
        self._chk(self.L.dpm_set_fene_constants(self.h, stiff, ext), "set_fene")

    def set_dt(self, dt):
        self._chk(self.L.dpm_set_time_step(self.h, dt), "set_dt")

    def set_self_propulsion(self, driving, taup):
        self._chk(self.L.dpm_set_self_propulsion(self.h, driving, taup), "set_self_propulsion")

    def set_active_tension(self, l0_tau, l0_diff):
        """activeShape rest-length dynamics; the rest lengths become the current l0."""
        self._chk(self.L.dpm_set_active_tension(self.h, l0_tau, l0_diff), "set_active_tension")

    def set_geometry(self, geometry, ew=None, box_radius=None):
        """Walls: normal (periodic), fixedBox, fixedSides or roundBox; wall energy scale ew and
        round-box radius when given."""
        self._chk(self.L.dpm_set_geometry(self.h, GEOM[geometry]), "set_geometry")
        if ew is not None:
            self._chk(self.L.dpm_set_box_energy_cost(self.h, ew), "set_box_energy_cost")
        if box_radius is not None:
            self._chk(self.L.dpm_set_box_radius(self.h, box_radius), "set_box_radius")

    def set_noise_type(self, noise):
        """Integrator of the Langevin loop: langevin, baoab, brownian or drivenBrownian."""
        self._chk(self.L.dpm_set_noise_type(self.h, NOISE[noise]), "set_noise_type")

    def set_seed(self, seed):
        self._chk(self.L.dpm_set_seed(self.h, seed), "set_seed")

    # ------------------------------------------------------------ neighbours
    def set_displacement_cutoff(self, cutoff, size):
        cd = C.c_double()
        self._chk(self.L.dpm_set_displacement_cutoff(self.h, cutoff, size, C.byref(cd)), "set_cutoff")
        return cd.value

    def set_cell_scale(self, scale):
        """Cell side as a multiple (>= 1) of the smallest one the neighbour search allows."""
        self._chk(self.L.dpm_set_cell_scale(self.h, float(scale)), "set_cell_scale")

    def calc_neighbor_list(self, cut):
        self._chk(self.L.dpm_calc_neighbor_list(self.h, cut), "calc_neighbor_list")

    def check_neighbors(self):
        self._chk(self.L.dpm_check_neighbors(self.h), "check_neighbors")

    def neighbors(self, cap=None):
        w = C.c_long()
        self._chk(self.L.dpm_get_neighbor_list(self.h, None, None, 0, C.byref(w)), "neighbors")
        cap = cap or max(w.value, 1)
        lst = np.empty(self.nv * cap, dtype=np.int64)
        cnt = np.empty(self.nv, dtype=np.int64)
        self._chk(self.L.dpm_get_neighbor_list(self.h, _p(lst), _p(cnt), cap, C.byref(w)), "neighbors")
        lst = lst.reshape(self.nv, cap)
        return [list(lst[v, :cnt[v]]) for v in range(self.nv)], w.value, cnt

    def update_count(self):
        a = np.zeros(self.nsys, dtype=np.int64)
        self._chk(self.L.dpm_get_update_count(self.h, _p(a)), "update_count")
        return a

    def reset_update_count(self):
        self._chk(self.L.dpm_reset_update_count(self.h), "reset_update_count")

    def reset_last_positions(self):
        self._chk(self.L.dpm_reset_last_positions(self.h), "reset_last_positions")

    # ---------------------------------------------------------------- forces
    def calc_particle_shape(self):
        self._chk(self.L.dpm_calc_particle_shape(self.h), "calc_particle_shape")

    def calc_force_energy(self):
        self._chk(self.L.dpm_calc_force_energy(self.h), "calc_force_energy")

    # ----------------------------------------------------------- integrators
    def init_langevin(self, T, gamma, read_state=False, zero_momentum=False):
        self._chk(self.L.dpm_init_langevin(self.h, T, gamma, int(read_state), int(zero_momentum)),
                  "init_langevin")

    def langevin(self, n):
        self._chk(self.L.dpm_langevin_loop(self.h, n), "langevin_loop")

    # the stages of Langevin::integrate, one call each (Simulator.cu:23-95)
    def inject_kinetic_energy(self, T):
        self._chk(self.L.dpm_inject_kinetic_energy(self.h, float(T)), "inject_kinetic_energy")

    def update_position(self, dt):
        self._chk(self.L.dpm_update_position(self.h, float(dt)), "update_position")

    def update_velocity(self, dt):
        self._chk(self.L.dpm_update_velocity(self.h, float(dt)), "update_velocity")

    def update_thermal_vel(self):
        self._chk(self.L.dpm_update_thermal_vel(self.h), "update_thermal_vel")

    def conserve_momentum(self):
        self._chk(self.L.dpm_conserve_momentum(self.h), "conserve_momentum")

    def init_nve(self, T, read_state=False, zero_momentum=False):
        self._chk(self.L.dpm_init_nve(self.h, T, int(read_state), int(zero_momentum)), "init_nve")

    def nve(self, n):
        self._chk(self.L.dpm_nve_loop(self.h, n), "nve_loop")

    def init_nve_rescale(self, T):
        self._chk(self.L.dpm_init_nve_rescale(self.h, T), "init_nve_rescale")

    def nve_rescale(self, n):
        self._chk(self.L.dpm_nve_rescale_loop(self.h, n), "nve_rescale_loop")

    def adjust_temperature(self, T):
        self._chk(self.L.dpm_adjust_temperature(self.h, T), "adjust_temperature")

    def adjust_kinetic_energy(self, prev_etot):
        """adjustKineticEnergy per system; returns the scales (0 where nothing was adjusted)."""
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(prev_etot, dtype=np.float64), (self.nsys,)))
        out = np.zeros(self.nsys)
        self._chk(self.L.dpm_adjust_kinetic_energy(self.h, _p(e), _p(out)), "adjust_kinetic_energy")
        return out

    # ------------------------------------------------------------------ FIRE
    def init_fire(self, params7, min_step=20, num_step=0):
        p = np.ascontiguousarray(params7, dtype=np.float64)
        self._chk(self.L.dpm_init_fire(self.h, _p(p), min_step, num_step), "init_fire")

    def fire(self, n):
        self._chk(self.L.dpm_fire_loop(self.h, n), "fire_loop")

    def fire_state(self):
        o = np.zeros(5 * self.nsys)
        self._chk(self.L.dpm_get_fire_state(self.h, _p(o)), "fire_state")
        o = o.reshape(self.nsys, 5)
        return {"dt": o[:, 0], "alpha": o[:, 1], "numStep": o[:, 2].astype(np.int64), "P": o[:, 3],
                "iters": o[:, 4].astype(np.int64)}

    # ---------------------------------------------------------- rigid rings
    def calc_rigid_force_energy(self):
        """calcRigidForceEnergy (DPM2D.cu:2128-2142): ring force, torque / inertia, energies."""
        self._chk(self.L.dpm_calc_rigid_force_energy(self.h), "calc_rigid_force_energy")

    def init_rigid_langevin(self, T, gamma, read_state=False):
        self._chk(self.L.dpm_init_rigid_langevin(self.h, T, gamma, int(read_state)), "init_rigid_langevin")

    def init_rigid_nve(self, T, read_state=False):
        self._chk(self.L.dpm_init_rigid_nve(self.h, T, int(read_state)), "init_rigid_nve")

    def rigid_langevin(self, n):
        self._chk(self.L.dpm_rigid_langevin_loop(self.h, n), "rigid_langevin_loop")

    def rigid_nve(self, n):
        self._chk(self.L.dpm_rigid_nve_loop(self.h, n), "rigid_nve_loop")

    def init_rigid_fire(self, params7, min_step=20, num_step=0):
        p = np.ascontiguousarray(params7, dtype=np.float64)
        self._chk(self.L.dpm_init_rigid_fire(self.h, _p(p), min_step, num_step), "init_rigid_fire")

    def rigid_fire(self, n):
        self._chk(self.L.dpm_rigid_fire_loop(self.h, n), "rigid_fire_loop")

    def rigid_fire_state(self):
        o = np.zeros(5 * self.nsys)
        self._chk(self.L.dpm_get_rigid_fire_state(self.h, _p(o)), "rigid_fire_state")
        o = o.reshape(self.nsys, 5)
        return {"dt": o[:, 0], "alpha": o[:, 1], "numStep": o[:, 2].astype(np.int64), "P": o[:, 3],
                "iters": o[:, 4].astype(np.int64)}

    def rigid_max_unbalanced_force(self):
        return self._per_sys(self.L.dpm_get_rigid_max_unbalanced_force, "rigid_max_unbalanced_force")

    def rigid_kinetic_energy(self):
        return self._per_sys(self.L.dpm_get_rigid_kinetic_energy, "rigid_kinetic_energy")

    def set_fire_kernel(self, persistent=True, waves=0):
        """fire_minimize on one persistent workgroup per system (waves 1..16, 0 = one per tile of
        the largest system up to 16), or on graphs of step kernels."""
        self._chk(self.L.dpm_set_fire_kernel(self.h, int(bool(persistent)), int(waves)), "set_fire_kernel")

    def fire_minimize(self, tol, max_iter):
        it = np.zeros(self.nsys, dtype=np.int64)
        self._chk(self.L.dpm_fire_minimize(self.h, tol, max_iter, _p(it)), "fire_minimize")
        return it

    # ----------------------------------------------------------- observables
    def _per_sys(self, fn, what):
        o = np.zeros(self.nsys)
        self._chk(fn(self.h, _p(o)), what)
        return o

    def potential_energy(self):
        return self._per_sys(self.L.dpm_get_potential_energy, "potential_energy")

    def kinetic_energy(self):
        return self._per_sys(self.L.dpm_get_kinetic_energy, "kinetic_energy")

    def temperature(self):
        return self._per_sys(self.L.dpm_get_temperature, "temperature")

    def max_unbalanced_force(self):
        return self._per_sys(self.L.dpm_get_max_unbalanced_force, "max_unbalanced_force")

    def total_force_magnitude(self):
        return self._per_sys(self.L.dpm_get_total_force_magnitude, "total_force_magnitude")

    def phi(self):
        return self._per_sys(self.L.dpm_get_phi, "phi")

    def preferred_phi(self):
        return self._per_sys(self.L.dpm_get_preferred_phi, "preferred_phi")

    def shapes(self):
        o = np.zeros(3 * self.np)
        self._chk(self.L.dpm_get_particle_shapes(self.h, _p(o)), "shapes")
        return o.reshape(self.np, 3)

    # ----------------------------------------------------------- observables
    def set_initial_positions(self):
        self._chk(self.L.dpm_set_initial_positions(self.h), "set_initial_positions")

    def vertex_msd(self):
        return self._per_sys(self.L.dpm_get_vertex_msd, "vertex_msd")

    def particle_msd(self):
        return self._per_sys(self.L.dpm_get_particle_msd, "particle_msd")

    def vertex_isf(self):
        return self._per_sys(self.L.dpm_get_vertex_isf, "vertex_isf")

    def particle_isf(self, k):
        o = np.zeros(self.nsys)
        self._chk(self.L.dpm_get_particle_isf(self.h, k, _p(o)), "particle_isf")
        return o

    def contacts(self):
        """calcContacts, then (rows [np, width] -1 padded, counts [np])."""
        self._chk(self.L.dpm_calc_contacts(self.h), "calc_contacts")
        w = C.c_long()
        self._chk(self.L.dpm_get_contacts(self.h, None, None, 0, C.byref(w)), "contacts")
        cap = max(w.value, 1)
        lst = np.empty(self.np * cap, dtype=np.int64)
        cnt = np.empty(self.np, dtype=np.int64)
        self._chk(self.L.dpm_get_contacts(self.h, _p(lst), _p(cnt), cap, C.byref(w)), "contacts")
        return lst.reshape(self.np, cap), cnt

    def geometric_contacts(self):
        """calcGeometricContacts (gap 1, rows of 8), then (rows [np, 8] -1 padded, counts [np])."""
        self._chk(self.L.dpm_calc_geometric_contacts(self.h), "calc_geometric_contacts")
        w = C.c_long()
        self._chk(self.L.dpm_get_contacts(self.h, None, None, 0, C.byref(w)), "contacts")
        cap = max(w.value, 1)
        lst = np.empty(self.np * cap, dtype=np.int64)
        cnt = np.empty(self.np, dtype=np.int64)
        self._chk(self.L.dpm_get_contacts(self.h, _p(lst), _p(cnt), cap, C.byref(w)), "contacts")
        return lst.reshape(self.np, cap), cnt

    def hexatic_order(self):
        return self._per_sys(self.L.dpm_get_hexatic_order, "hexatic_order")

    def per_particle_stress(self):
        o = np.zeros(4 * self.np)
        self._chk(self.L.dpm_get_per_particle_stress(self.h, _p(o)), "per_particle_stress")
        return o.reshape(self.np, 4)

    def stress_tensor(self):
        o = np.zeros(4 * self.nsys)
        self._chk(self.L.dpm_get_stress_tensor(self.h, _p(o)), "stress_tensor")
        return o.reshape(self.nsys, 4)

    def pressure(self):
        return self._per_sys(self.L.dpm_get_pressure, "pressure")

    def scale_vertices(self, scale):
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, dtype=np.float64), (self.nsys,)))
        self._chk(self.L.dpm_scale_vertices(self.h, _p(s)), "scale_vertices")

    def box_radius(self):
        r = C.c_double()
        self._chk(self.L.dpm_get_box_radius(self.h, C.byref(r)), "get_box_radius")
        return r.value

    def scale_box(self, scale):
        """scaleBox (DPM2D.cu:1693-1719) per system: box (or round-box radius) / scale."""
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, dtype=np.float64), (self.nsys,)))
        self._chk(self.L.dpm_scale_box(self.h, _p(s)), "scale_box")

    def synchronize(self):
        self._chk(self.L.dpm_synchronize(self.h), "synchronize")

    def set_stream(self, stream_ptr):
        self._chk(self.L.dpm_set_stream(self.h, C.c_void_p(stream_ptr)), "set_stream")

    def set_launch_config(self, waves_per_tile):
        self._chk(self.L.dpm_set_launch_config(self.h, waves_per_tile), "set_launch_config")

    def time_step_kernel(self, reps=200):
        us, kbar = C.c_double(), C.c_double()
        self._chk(self.L.dpm_time_step_kernel(self.h, reps, C.byref(us), C.byref(kbar)), "time_step_kernel")
        return us.value, kbar.value

    # ---------------------------------------------------------------- adjoint
    # Vectors are device pointers (ints) of float64 arrays on the set's stream; "2V SoA" = the V x
    # components followed by the V y components.
    def force_hvp(self, ux_ptr, uy_ptr, ox_ptr, oy_ptr):
        """out = H u at the current positions."""
        self._chk(self.L.dpm_force_hvp(self.h, C.c_void_p(ux_ptr), C.c_void_p(uy_ptr),
                                       C.c_void_p(ox_ptr), C.c_void_p(oy_ptr)), "force_hvp")

    def energy_param_grad(self, gring_ptr, gl0_ptr):
        """dU/d(ka, kl, kb) per ring (3R) and dU/dl0 (V) at the current shape variables."""
        self._chk(self.L.dpm_energy_param_grad(self.h, C.c_void_p(gring_ptr), C.c_void_p(gl0_ptr)),
                  "energy_param_grad")

    def nve_forward(self, steps, every):
        """NVE steps with checkpoints every `every` steps for nve_backward."""
        self._chk(self.L.dpm_nve_forward(self.h, steps, every), "nve_forward")
        self.tape_serial += 1

    def fire_forward(self, iterations):
        """FIRE iterations (no stopping rule) recorded for fire_backward."""
        self._chk(self.L.dpm_fire_forward(self.h, int(iterations)), "fire_forward")

    def fire_backward(self, ax_ptr, av_ptr, gx_ptr, gv_ptr, gring_ptr, gl0_ptr):
        """Reverse pass of the last fire_forward with its branches held fixed (2V SoA)."""
        p = [C.c_void_p(x) if x else None for x in (ax_ptr, av_ptr, gx_ptr, gv_ptr, gring_ptr, gl0_ptr)]
        self._chk(self.L.dpm_fire_backward(self.h, *p), "fire_backward")

    def nve_backward(self, ax_ptr, av_ptr, gx_ptr, gv_ptr, gring_ptr, gl0_ptr):
        """Reverse pass of the last nve_forward (2V SoA adjoints in, gradients out)."""
        p = [C.c_void_p(x) if x else None for x in (ax_ptr, av_ptr, gx_ptr, gv_ptr, gring_ptr, gl0_ptr)]
        self._chk(self.L.dpm_nve_backward(self.h, *p), "nve_backward")
