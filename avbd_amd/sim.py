"""Host-side mirror of the reference's solver API over the avbd C ABI.

Names follow the reference so a caller reads the same:
  * ``SimulationContext``  capacities + params (simulation/SimulationContext.hpp:18-60)
  * ``VBDSolver``          ctor (ctx, solver_iters, iter_acceleration) and
                           ``solve_Parallel`` (simulation/VBDSolver.hpp:37-103)
  * ``Simulation``         setup / update / _timeStep (simulation/Simulation.cpp:48-220)
The device does all per-step work; this layer only stages inputs and reads
results back.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib
from .scenes import Rod, Scene, Sphere, TetMesh, tip_body


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _uptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _fptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _bptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class SimulationContext:
    """Owns one device context (avbd_ctx): particle and energy pools, colouring,
    device buffers, stream and step graph."""

    def __init__(self, dt: float, g_accel: float = 9.81, solver_iters: int = 20,
                 iter_acceleration: float = 0.95, particle_capacity: int = 0, energy_capacity: int = 0,
                 device: int = 0, lanes_per_vertex: int = 0, block_size: int = 0, use_graph: bool = True,
                 serial_broadphase: bool = False, separate_chebyshev: bool = False, gather_batch: int = 0,
                 fp64_state: bool = False, host_coloring: bool = False):
        desc = _lib.SceneDesc()
        desc.particle_capacity = particle_capacity
        desc.energy_capacity = energy_capacity
        desc.dt = dt
        desc.g_accel = g_accel
        desc.solver_iters = solver_iters
        desc.iter_acceleration = iter_acceleration
        desc.flags = ((0 if use_graph else _lib.FLAG_NO_GRAPH) | (_lib.FLAG_SERIAL_BROADPHASE if serial_broadphase else 0)
                      | (_lib.FLAG_SEPARATE_CHEBYSHEV if separate_chebyshev else 0)
                      | (_lib.FLAG_FP64_STATE if fp64_state else 0)
                      | (_lib.FLAG_HOST_COLORING if host_coloring else 0))
        desc.device = device
        desc.lanes_per_vertex = lanes_per_vertex
        desc.block_size = block_size
        desc.gather_batch = gather_batch
        h = C.c_void_p()
        rc = lib.avbd_create(C.byref(desc), C.byref(h))
        check(rc, None)
        self.handle = h
        self.dt = dt
        self.g_accel = g_accel
        self.solver_iters = solver_iters
        self.iter_acceleration = iter_acceleration
        self.built = False

    def close(self):
        if getattr(self, "handle", None):
            lib.avbd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        check(rc, self.handle)

    # --- pools -----------------------------------------------------------
    def add_tet_mesh(self, rest: np.ndarray, tets: np.ndarray, E: float, nu: float, density: float,
                     kd: float = 0.0, with_ground: bool = False, mu_s: float = 0.4, mu_k: float = 0.2) -> int:
        rest = np.ascontiguousarray(rest, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.uint32)
        first = C.c_uint32()
        self._chk(lib.avbd_add_tet_mesh(self.handle, rest.shape[0], _dptr(rest), tets.shape[0], _uptr(tets),
                                        E, nu, density, kd, int(with_ground), mu_s, mu_k, C.byref(first)))
        return first.value

    def add_particles(self, x, v=None, v_prev=None, x_prev=None, mass=None, fixed=None) -> int:
        x = np.ascontiguousarray(x, dtype=np.float64)
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        v, v_prev, x_prev, mass = arr(v), arr(v_prev), arr(x_prev), arr(mass)
        fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        first = C.c_uint32()
        self._chk(lib.avbd_add_particles(self.handle, x.shape[0], _dptr(x), _dptr(v), _dptr(v_prev),
                                         _dptr(x_prev), _dptr(mass), _bptr(fixed), C.byref(first)))
        return first.value

    def add_energies(self, type_: int, records) -> None:
        n = len(records)
        self._chk(lib.avbd_add_energies(self.handle, type_, C.cast(records, C.c_void_p), n))

    # --- oriented particles (ParticlePool::addOrientedParticle, Rod, RigidSphere) ---
    def add_oriented_particles(self, x, inertia, q=None, v=None, w=None, mass=None, fixed=None) -> int:
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        x, inertia, q, v, w, mass = arr(x), arr(inertia), arr(q), arr(v), arr(w), arr(mass)
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        first = C.c_uint32()
        self._chk(lib.avbd_add_oriented_particles(self.handle, x.shape[0], _dptr(x), _dptr(q), _dptr(v), _dptr(w),
                                                  _dptr(mass), _dptr(inertia), _bptr(fx), C.byref(first)))
        return first.value

    def add_rod(self, r: Rod) -> int:
        v3 = lambda t: np.ascontiguousarray(t, dtype=np.float64)
        first = C.c_uint32()
        self._chk(lib.avbd_add_rod(self.handle, r.elements, r.length, _dptr(v3(r.curvature)), r.E, r.nu, r.density,
                                   r.diameter, _dptr(v3(r.pos)), _dptr(v3(r.rot)), _dptr(v3(r.vel)),
                                   int(r.base_fixed), int(r.tip_fixed), int(r.ground), C.byref(first)))
        return first.value

    def add_sphere(self, sp: Sphere) -> int:
        v3 = lambda t: np.ascontiguousarray(t, dtype=np.float64)
        first = C.c_uint32()
        self._chk(lib.avbd_add_sphere(self.handle, sp.radius, sp.density, _dptr(v3(sp.pos)), _dptr(v3(sp.rot)),
                                      int(sp.fixed), int(sp.ground), C.byref(first)))
        return first.value

    def add_mass(self, p: int, dm: float, dI=None) -> None:
        """Weld a rigid body's mass and principal inertia onto particle p."""
        dI = None if dI is None else np.ascontiguousarray(dI, dtype=np.float64)
        self._chk(lib.avbd_add_mass(self.handle, p, dm, _dptr(dI)))

    def add_rigid_ground(self, p: int, shape: int, radius: float, k_start: float, mu_s: float = 0.5,
                         mu_k: float = 0.2) -> None:
        """RigidBodyGroundCollisionEnergyPool::addEnergy for a lowest-point shape."""
        r = _lib.RigidGroundRecord()
        r.particle_index = p
        r.shape = shape
        r.radius = radius
        r.cp_rb_local[1] = -radius
        r.k_start = r.k_n = r.k_t = r.k_b = k_start
        r.mu_s, r.mu_k = mu_s, mu_k
        self.add_energies(_lib.RIGID_BODY_GROUND_COLLISION, (_lib.RigidGroundRecord * 1)(r))

    def add_joint(self, p: int, pos_offset, rot_offset, ref_position, ref_rotation) -> None:
        """OneSidedFixedJointEnergyPool::addEnergy (energy/joint/OneSidedFixedJointEnergyPool.hpp:33-50):
        k = k_start = 1e4, lambda = 0, C_prev = 0."""
        r = _lib.JointRecord()
        r.particle_index = p
        for k in range(3):
            r.position_offset[k] = pos_offset[k]
            r.ref_position[k] = ref_position[k]
        for k in range(4):
            r.rotation_offset[k] = rot_offset[k]
            r.ref_rotation[k] = ref_rotation[k]
        for k in range(6):
            r.k[k] = 1e4
        recs = (_lib.JointRecord * 1)(r)
        self.add_energies(_lib.ONE_SIDED_FIXED_JOINT, recs)

    def add_fixed_joint(self, p0: int, p1: int, pos_offsets, rot_offsets) -> None:
        """Two-sided fixed joint (new) between oriented particles p0 and p1:
        pos_offsets (2, 3) and rot_offsets (2, 4) place the joint frame on each
        body; k = 1e4 (the one-sided joint's k_start), lambda = 0, C_prev = 0."""
        r = _lib.FixedJointRecord()
        r.particle_index[0], r.particle_index[1] = p0, p1
        for i in range(2):
            for k in range(3):
                r.position_offset[i][k] = pos_offsets[i][k]
            for k in range(4):
                r.rotation_offset[i][k] = rot_offsets[i][k]
        for k in range(6):
            r.k[k] = 1e4
        self.add_energies(_lib.FIXED_JOINT, (_lib.FixedJointRecord * 1)(r))

    def add_cosserat(self, a: int, b: int, stiffness, rest_length: float, rest_curvature=(0, 0, 0)) -> None:
        r = _lib.CosseratRecord()
        r.particle_indices[0], r.particle_indices[1] = a, b
        for k in range(6):
            r.stiffness[k] = stiffness[k]
        r.rest_length = rest_length
        for k in range(3):
            r.rest_curvature[k] = rest_curvature[k]
        self.add_energies(_lib.COSSERAT_ROD, (_lib.CosseratRecord * 1)(r))

    def set_rotation_state(self, first: int, q=None, w=None) -> None:
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        q, w = arr(q), arr(w)
        n = q.shape[0] if q is not None else w.shape[0]
        self._chk(lib.avbd_set_rotation_state(self.handle, first, n, _dptr(q), _dptr(w)))

    def rotations(self):
        """Per particle: orientation (w, x, y, z) and body-frame angular velocity."""
        q = np.empty((self.num_particles, 4))
        w = np.empty((self.num_particles, 3))
        self._chk(lib.avbd_download_rotations(self.handle, _dptr(q), _dptr(w)))
        return q, w

    def rigid_ground_records(self):
        n = lib.avbd_num_energies(self.handle, _lib.RIGID_BODY_GROUND_COLLISION)
        recs = (_lib.RigidGroundRecord * n)()
        self._chk(lib.avbd_download_energies(self.handle, _lib.RIGID_BODY_GROUND_COLLISION,
                                             C.cast(recs, C.c_void_p), n))
        return recs

    def joint_records(self):
        n = lib.avbd_num_energies(self.handle, _lib.ONE_SIDED_FIXED_JOINT)
        recs = (_lib.JointRecord * n)()
        self._chk(lib.avbd_download_energies(self.handle, _lib.ONE_SIDED_FIXED_JOINT, C.cast(recs, C.c_void_p), n))
        return recs

    def fixed_joint_records(self):
        n = lib.avbd_num_energies(self.handle, _lib.FIXED_JOINT)
        recs = (_lib.FixedJointRecord * n)()
        self._chk(lib.avbd_download_energies(self.handle, _lib.FIXED_JOINT, C.cast(recs, C.c_void_p), n))
        return recs

    def set_particle_state(self, first: int, x=None, x_prev=None, v=None, v_prev=None) -> None:
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        x, x_prev, v, v_prev = arr(x), arr(x_prev), arr(v), arr(v_prev)
        n = next(a.shape[0] for a in (x, x_prev, v, v_prev) if a is not None)
        self._chk(lib.avbd_set_particle_state(self.handle, first, n, _dptr(x), _dptr(x_prev), _dptr(v),
                                              _dptr(v_prev)))

    def set_external_accel(self, first: int, a: np.ndarray) -> None:
        """Per-particle external acceleration (N/kg) added to gravity."""
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
        self._chk(lib.avbd_set_external_accel(self.handle, first, a.shape[0], _dptr(a)))

    def set_fixed(self, first: int, fixed: np.ndarray) -> None:
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
        self._chk(lib.avbd_set_fixed(self.handle, first, fixed.shape[0], _bptr(fixed)))

    def add_collision_triangles(self, tris: np.ndarray) -> None:
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        self._chk(lib.avbd_add_collision_triangles(self.handle, tris.shape[0], _uptr(tris)))

    def detect_collisions(self) -> None:
        self._chk(lib.avbd_detect_collisions(self.handle))

    # --- dynamic contacts (CollisionDetector::detectCollisionsAndRecolor) ----
    def add_collision_segments(self, nodes, radius: float) -> None:
        """The segments between consecutive rod nodes, one object."""
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        self._chk(lib.avbd_add_collision_segments(self.handle, nodes.shape[0], _uptr(nodes), float(radius)))

    def add_collision_sphere(self, p: int, radius: float) -> None:
        self._chk(lib.avbd_add_collision_sphere(self.handle, int(p), float(radius)))

    def set_dynamic_contacts(self, margin: float, capacity: int = 0) -> None:
        self._chk(lib.avbd_set_dynamic_contacts(self.handle, float(margin), int(capacity)))

    def set_self_collision(self, enable: bool = True) -> None:
        """Vertex-face contacts between collision triangles (new); before set_dynamic_contacts."""
        self._chk(lib.avbd_set_self_collision(self.handle, int(bool(enable))))

    def contacts(self, kind: int) -> np.ndarray:
        """Active contacts of one kind (0 triangle-rigid, 1 triangle-rod, 2 rigid-rigid, 3 vertex-triangle) in slot
        order, one row of 32 per contact: slot, idx[5], n, t, b, barys, s,
        cp_local, k_start, k_n, k_t, k_b, lambda_n, lambda_t, lambda_b,
        C_n_prev, use_static, unclamped."""
        cnt = C.c_uint32()
        self._chk(lib.avbd_download_contacts(self.handle, kind, None, 0, C.byref(cnt)))
        n = cnt.value
        buf = (_lib.Contact * max(n, 1))()
        self._chk(lib.avbd_download_contacts(self.handle, kind, buf, n, C.byref(cnt)))
        out = np.zeros((n, 32))
        for i in range(n):
            c = buf[i]
            out[i, 0] = c.slot
            out[i, 1:6] = list(c.idx)
            out[i, 6:9], out[i, 9:12], out[i, 12:15] = list(c.n), list(c.t), list(c.b)
            out[i, 15:18] = list(c.barys)
            out[i, 18] = c.s
            out[i, 19:22] = list(c.cp_local)
            out[i, 22:30] = [c.k_start, c.k_n, c.k_t, c.k_b, c.lambda_n, c.lambda_t, c.lambda_b, c.C_n_prev]
            out[i, 30], out[i, 31] = c.use_static, c.unclamped
        return out

    def detected(self):
        """Last detection's contacts in key order: (keys, slots, kinds)."""
        cnt = C.c_uint32()
        self._chk(lib.avbd_download_detected(self.handle, None, None, None, 0, C.byref(cnt)))
        n = cnt.value
        keys = np.zeros(max(n, 1), dtype=np.uint64)
        slots = np.zeros(max(n, 1), dtype=np.uint32)
        kinds = np.zeros(max(n, 1), dtype=np.uint8)
        self._chk(lib.avbd_download_detected(self.handle, keys.ctypes.data_as(C.POINTER(C.c_uint64)), _uptr(slots),
                                             kinds.ctypes.data_as(C.POINTER(C.c_uint8)), n, C.byref(cnt)))
        return keys[:n], slots[:n], kinds[:n]

    # --- checkpoint / resume -------------------------------------------------
    def save_checkpoint(self) -> bytes:
        n = C.c_uint64()
        self._chk(lib.avbd_checkpoint_size(self.handle, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        w = C.c_uint64()
        self._chk(lib.avbd_save_checkpoint(self.handle, buf, n.value, C.byref(w)))
        return buf.raw[:w.value]

    def load_checkpoint(self, blob: bytes) -> None:
        buf = C.create_string_buffer(blob, len(blob))
        self._chk(lib.avbd_load_checkpoint(self.handle, buf, len(blob)))

    def conflicted(self) -> np.ndarray:
        f = np.zeros(lib.avbd_num_particles(self.handle), dtype=np.uint8)
        self._chk(lib.avbd_download_conflicts(self.handle, f.ctypes.data_as(C.POINTER(C.c_uint8))))
        return f.astype(bool)

    def set_residual_target(self, rel_tol: float, max_iters: int, check_every: int = 5) -> None:
        """Iterate each step until ||r|| / ||M g|| <= rel_tol (checked every
        check_every iterations) or max_iters; rel_tol = 0: fixed count."""
        self._chk(lib.avbd_set_residual_target(self.handle, rel_tol, max_iters, check_every))

    # --- learned warm start (BF16 MFMA MLP) --------------------------------
    IN, OUT = 24, 3

    def mlp_init(self, seed: int = 0) -> None:
        self._chk(lib.avbd_mlp_init(self.handle, seed))

    def mlp_params(self) -> np.ndarray:
        n = lib.avbd_mlp_param_count()
        p = np.empty(n, dtype=np.float32)
        self._chk(lib.avbd_mlp_get_params(self.handle, _fptr(p), n))
        return p

    def mlp_set_params(self, p) -> None:
        p = _f32(p)
        self._chk(lib.avbd_mlp_set_params(self.handle, _fptr(p), p.shape[0]))

    def mlp_set_normalization(self, norm) -> None:
        norm = _f32(norm)
        assert norm.shape[0] == 2 * self.IN + 2 * self.OUT
        self._chk(lib.avbd_mlp_set_normalization(self.handle, _fptr(norm)))

    def mlp_forward_backward(self, X, T):
        X, T = _f32(X), _f32(T)
        n = lib.avbd_mlp_param_count() - 2 * self.IN - 2 * self.OUT
        g = np.empty(n, dtype=np.float32)
        loss = C.c_float()
        self._chk(lib.avbd_mlp_forward_backward(self.handle, _fptr(X), _fptr(T), X.shape[0], _fptr(g), n,
                                                C.byref(loss)))
        return g, loss.value

    def mlp_predict(self, X) -> np.ndarray:
        X = _f32(X)
        out = np.empty((X.shape[0], self.OUT), dtype=np.float32)
        self._chk(lib.avbd_mlp_predict(self.handle, _fptr(X), X.shape[0], _fptr(out)))
        return out

    def mlp_train(self, X, T, epochs: int, batch: int, lr: float, seed: int = 0) -> float:
        X, T = _f32(X), _f32(T)
        loss = C.c_float()
        self._chk(lib.avbd_mlp_train(self.handle, _fptr(X), _fptr(T), X.shape[0], epochs, batch, lr, seed,
                                     C.byref(loss)))
        return loss.value

    def record_steps(self, n_steps: int, n_items: int):
        """Run n_steps recording (features, targets) of every positional work item."""
        cap = n_steps * n_items
        X = np.empty((cap, self.IN), dtype=np.float32)
        T = np.empty((cap, self.OUT), dtype=np.float32)
        rows = C.c_uint32()
        self._chk(lib.avbd_record_steps(self.handle, n_steps, _fptr(X), _fptr(T), cap, C.byref(rows)))
        return X[:rows.value], T[:rows.value]

    def set_warm_start(self, enable: bool) -> None:
        self._chk(lib.avbd_set_warm_start(self.handle, int(enable)))

    def mlp_gemm_test(self, A, B, ta: bool, tb: bool) -> np.ndarray:
        A, B = _f32(A), _f32(B)
        M, K = A.shape
        N = B.shape[1]
        Cm = np.empty((M, N), dtype=np.float32)
        self._chk(lib.avbd_mlp_gemm_test(self.handle, M, N, K, int(ta), int(tb), _fptr(A), _fptr(B), _fptr(Cm)))
        return Cm

    def residual_history(self, worst_scene: bool = False) -> np.ndarray:
        """Relative residuals at the checks of the last step (residual-targeted
        mode): over all scenes, or (worst_scene) the largest per-scene one."""
        cap = 1024
        buf = np.empty(cap)
        n = C.c_uint32()
        fn = lib.avbd_get_scene_residual_history if worst_scene else lib.avbd_get_residual_history
        self._chk(fn(self.handle, _dptr(buf), cap, C.byref(n)))
        return buf[:min(cap, n.value)].copy()

    def warm_start_decision(self):
        """(ip_delta [scenes, 3], step [scenes]) of the last warm-started step:
        the incremental-potential change at step lengths 1, 1/2, 1/4 along
        the prediction, and the chosen length (0 = inertial start kept)."""
        n = C.c_uint32()
        self._chk(lib.avbd_get_warm_start_decision(self.handle, None, None, 0, C.byref(n)))
        ipd = np.zeros((n.value, 3))
        st = np.zeros(n.value, dtype=np.float32)
        self._chk(lib.avbd_get_warm_start_decision(self.handle, _dptr(ipd), _fptr(st), n.value, C.byref(n)))
        return ipd, st

    def scene_iterations(self) -> np.ndarray:
        """Iterations each scene ran in the last step (residual-targeted mode)."""
        n = C.c_uint32()
        self._chk(lib.avbd_get_scene_iterations(self.handle, None, 0, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint32)
        self._chk(lib.avbd_get_scene_iterations(self.handle, _uptr(buf), n.value, C.byref(n)))
        return buf

    def set_scenes(self, offsets: Sequence[int]) -> None:
        off = np.ascontiguousarray(offsets, dtype=np.uint32)
        self._chk(lib.avbd_set_scenes(self.handle, off.shape[0] - 1, _uptr(off)))

    def build_graph(self) -> None:
        self._chk(lib.avbd_build_graph(self.handle))
        self.built = True

    # --- stepping / readback ---------------------------------------------
    @property
    def num_particles(self) -> int:
        return lib.avbd_num_particles(self.handle)

    def step(self, n: int = 1) -> None:
        self._chk(lib.avbd_step(self.handle, n))

    def synchronize(self) -> None:
        self._chk(lib.avbd_synchronize(self.handle))

    def positions(self) -> np.ndarray:
        x = np.empty((self.num_particles, 3))
        self._chk(lib.avbd_download_state(self.handle, _dptr(x), None))
        return x

    def state(self):
        x = np.empty((self.num_particles, 3))
        v = np.empty((self.num_particles, 3))
        self._chk(lib.avbd_download_state(self.handle, _dptr(x), _dptr(v)))
        return x, v

    def inertial_positions(self) -> np.ndarray:
        y = np.empty((self.num_particles, 3))
        self._chk(lib.avbd_download_inertial(self.handle, _dptr(y)))
        return y

    def coloring(self):
        col = np.empty(self.num_particles, dtype=np.uint32)
        nc = C.c_uint32()
        self._chk(lib.avbd_download_coloring(self.handle, _uptr(col), C.byref(nc)))
        return col, nc.value

    def ground_records(self):
        n = lib.avbd_num_energies(self.handle, _lib.GROUND_COLLISION)
        recs = (_lib.GroundRecord * n)()
        self._chk(lib.avbd_download_energies(self.handle, _lib.GROUND_COLLISION, C.cast(recs, C.c_void_p), n))
        return recs

    def tet_records(self):
        n = lib.avbd_num_energies(self.handle, _lib.NEO_HOOKEAN)
        recs = (_lib.NeoHookeanRecord * n)()
        self._chk(lib.avbd_download_energies(self.handle, _lib.NEO_HOOKEAN, C.cast(recs, C.c_void_p), n))
        return recs

    def stats(self) -> _lib.Stats:
        st = _lib.Stats()
        self._chk(lib.avbd_get_stats(self.handle, C.byref(st)))
        return st

    def persistent_grid(self) -> int:
        g = C.c_uint32()
        self._chk(lib.avbd_get_step_mode(self.handle, C.byref(g), None))
        return g.value

    def fused_chebyshev(self) -> bool:
        f = C.c_uint32()
        self._chk(lib.avbd_get_step_mode(self.handle, None, C.byref(f)))
        return bool(f.value)

    def launch_config(self):
        """(lanes per vertex, block size, incidences gathered at once per lane)."""
        L, B, U = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._chk(lib.avbd_get_launch_config(self.handle, C.byref(L), C.byref(B), C.byref(U)))
        return L.value, B.value, U.value

    def set_timing(self, enable: bool) -> None:
        self._chk(lib.avbd_set_timing(self.handle, int(enable)))

    def calibrate_copy(self, nbytes: int, iters: int = 10) -> float:
        ms = C.c_double()
        self._chk(lib.avbd_calibrate_copy(self.handle, nbytes, iters, C.byref(ms)))
        return ms.value

    def time_solve_launches(self, reps: int = 20):
        """Average duration (ms) of one vertex-solve launch, back to back on the ctx stream."""
        ms, n = C.c_double(), C.c_uint32()
        self._chk(lib.avbd_time_solve_launches(self.handle, reps, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def broadphase_timing(self) -> float:
        ms = C.c_double()
        self._chk(lib.avbd_get_broadphase_timing(self.handle, C.byref(ms)))
        return ms.value

    def solve_timing(self):
        ms, n = C.c_double(), C.c_uint32()
        self._chk(lib.avbd_get_solve_timing(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class VBDSolver:
    """VBDSolver(ctx, solver_iters, iter_acceleration) (simulation/VBDSolver.hpp:37-42).
    The iteration count and Chebyshev rho are fixed when the context is created;
    ``solve_Parallel`` runs one full time step on the device."""

    def __init__(self, ctx: SimulationContext, solver_iters: int, iter_acceleration: float):
        if solver_iters != ctx.solver_iters or iter_acceleration != ctx.iter_acceleration:
            raise ValueError("solver_iters / iter_acceleration must match the SimulationContext")
        self._ctx = ctx

    def solve_Parallel(self, w_ctx=None) -> None:
        # every SPMD worker of the reference enters here; only worker 0 issues work
        if w_ctx is None or getattr(w_ctx, "idx", 0) == 0:
            self._ctx.step(1)


class SimulationLogger:
    """SimulationLogger (simulation/SimulationLogger.hpp:18-77,
    simulation/SimulationLogger.cpp): named quantities evaluated by callables,
    one space-separated line per logging interval after a header line."""

    def __init__(self, output_filename, logging_interval: float):
        from pathlib import Path
        path = Path(output_filename)
        path.parent.mkdir(parents=True, exist_ok=True)
        self._out = open(path, "w")
        self._interval = float(logging_interval)
        self._last = -1e10
        self._vars: List[tuple] = []
        self._logging = False

    def add_output(self, name: str, func) -> None:
        if self._logging:
            raise RuntimeError("outputs cannot be added after logging started")
        self._vars.append((name, func))

    def start_logging(self) -> None:
        assert not self._logging
        self._logging = True
        self._out.write("".join(f"{n} " for n, _ in self._vars) + "\n")

    def log_to_file(self, time: float) -> None:
        if time - self._last + 1e-12 >= self._interval:
            self._out.write("".join(f"{f()!r} " for _, f in self._vars) + "\n")
            self._last = time

    def stop_logging(self) -> None:
        self._logging = False
        self._out.close()


class Simulation:
    """Simulation shell (simulation/Simulation.cpp:22-220): builds the context from
    a Scene description, runs setup() (adjacency + colouring) and time steps.
    A logger (``logger`` argument or ``Scene.logging``) records time and, with
    ``log_residuals``, the force residual and incremental potential after
    every logged step (the quantities of the reference's log-residuals
    option, config/SimulationConfig.hpp:50-53)."""

    def __init__(self, scene: Scene, device: int = 0, lanes_per_vertex: int = 0, block_size: int = 0,
                 use_graph: bool = True, serial_broadphase: bool = False,
                 residual_target: Optional[tuple] = None, separate_chebyshev: bool = False, gather_batch: int = 0,
                 fp64_state: bool = False, host_coloring: bool = False):
        self.scene = scene
        self.ctx = SimulationContext(dt=scene.dt, g_accel=scene.g, solver_iters=scene.iters,
                                     iter_acceleration=scene.rho, device=device,
                                     lanes_per_vertex=lanes_per_vertex, block_size=block_size,
                                     use_graph=use_graph, serial_broadphase=serial_broadphase,
                                     separate_chebyshev=separate_chebyshev, gather_batch=gather_batch,
                                     fp64_state=fp64_state, host_coloring=host_coloring)
        self.solver = VBDSolver(self.ctx, scene.iters, scene.rho)
        if residual_target is not None:
            self.ctx.set_residual_target(*residual_target)
        self.time = 0.0
        self.offsets: List[int] = []
        self._setup = False
        self.logger: Optional[SimulationLogger] = None
        lg = getattr(scene, "logging", None)
        if lg:
            import datetime
            from pathlib import Path
            name = datetime.datetime.now().strftime("%Y-%m-%d_%H:%M:%S") + ".txt"
            self.logger = SimulationLogger(Path(lg.get("output_dir", "output")) / name, lg.get("interval", 1e-2))
            self.logger.add_output("time", lambda: self.time)
            if lg.get("log_residuals", False):
                self._stats_cache = None
                self.logger.add_output("residual_rel", lambda: self._stats().residual_rel)
                self.logger.add_output("residual_abs", lambda: self._stats().residual_abs)
                self.logger.add_output("incremental_potential", lambda: self._stats().incremental_potential)

    def _stats(self):
        # one device stats pass per logged line
        if self._stats_cache is None or self._stats_cache[0] != self.time:
            self._stats_cache = (self.time, self.ctx.stats())
        return self._stats_cache[1]

    def setup(self) -> None:
        sc = self.scene
        offs = [0]
        self.rod_first: List[int] = []
        self.sphere_index: List[int] = []
        self._spins: List[tuple] = []
        for kind in sc.object_order:
            if kind == "meshes":
                for m in sc.meshes:
                    first = self._add_object(m)
                    offs.append(first + m.rest.shape[0])
            elif kind == "rods":
                for r in sc.rods:
                    self.rod_first.append(self._add_rod(r))
            elif kind == "spheres":
                for sp in sc.spheres:
                    self.sphere_index.append(self._add_sphere(sp))
            else:
                raise ValueError(f"unknown object kind {kind!r}")
        self.offsets = offs
        for p, _rot, w0 in self._spins:
            if any(c != 0.0 for c in w0):
                self.ctx.set_rotation_state(p, w=np.asarray([w0], dtype=np.float64))
        if len(sc.meshes) > 1 and sc.batch:
            if sc.rods or sc.spheres:
                raise ValueError("batched tet scenes cannot be mixed with rods or spheres")
            self.ctx.set_scenes(offs)
        self.ctx.build_graph()
        if self.scene.contact_margin is not None:
            if self.scene.self_collision:
                self.ctx.set_self_collision(True)
            self.ctx.set_dynamic_contacts(self.scene.contact_margin)
        self._setup = True

    def _add_rod(self, r: Rod) -> int:
        first = self.ctx.add_rod(r)
        if self.scene.contact_margin is not None and r.collide:
            self.ctx.add_collision_segments(np.arange(first, first + r.elements + 1), 0.5 * r.diameter)
        if r.base_joint:
            # hold the first node at its initial pose
            self.ctx.add_joint(first, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), r.pos, r.rot)
        if r.tip_sphere is not None and r.tip_joint:
            rad, m, w0 = r.tip_sphere
            centre, density, offs, rots = tip_body(r)
            body = self._add_sphere(Sphere(radius=rad, density=density, pos=tuple(centre), rot=r.rot, w0=w0,
                                           ground=r.ground))
            self.ctx.add_fixed_joint(first + r.elements, body, offs, rots)
        elif r.tip_sphere is not None:
            rad, m, w0 = r.tip_sphere
            tip = first + r.elements
            self.ctx.add_mass(tip, m, np.full(3, 0.4 * m * rad * rad))
            # the node's total mass sets k_start, as for any rigid-ground contact
            node_mass = self._rod_node_mass(r, True) + m
            self.ctx.add_rigid_ground(tip, _lib.SHAPE_SPHERE, rad, node_mass / (self.scene.dt * self.scene.dt))
            self._spins.append((tip, r.rot, w0))
        return first

    @staticmethod
    def _rod_node_mass(r: Rod, end: bool) -> float:
        """Lumped node mass of the Rod ctor (simobject/Rod.cpp:47-58): half an
        element at the ends, a whole one inside."""
        area = np.pi * (r.diameter / 2) ** 2
        elem = r.length / r.elements * area * r.density
        return 0.5 * elem if end else elem

    def _add_sphere(self, sp: Sphere) -> int:
        p = self.ctx.add_sphere(sp)
        if self.scene.contact_margin is not None:
            self.ctx.add_collision_sphere(p, sp.radius)
        if any(c != 0.0 for c in sp.w0):
            self.ctx.set_rotation_state(p, q=np.asarray([sp.rot], dtype=np.float64),
                                        w=np.asarray([sp.w0], dtype=np.float64))
        if any(c != 0.0 for c in sp.v0):
            v0 = np.asarray([sp.v0], dtype=np.float64)
            self.ctx.set_particle_state(p, v=v0, v_prev=v0)
        return p

    def _add_object(self, m: TetMesh) -> int:
        first = self.ctx.add_tet_mesh(m.rest, m.tets, m.E, m.nu, m.density, m.kd, m.ground, m.mu_s, m.mu_k)
        # CollisionPrimitivePool::addObject(TetMeshObject) registers the surface
        # faces (collision/CollisionPrimitivePool.cpp:36-50); independent scenes of
        # a batch are not collided with each other, so batches skip it
        if m.triangles is not None and (len(self.scene.meshes) == 1 or not self.scene.batch):
            self.ctx.add_collision_triangles(m.triangles + first)
        if m.fixed is not None:
            self.ctx.set_fixed(first, m.fixed)
        if m.x0 is not None or m.v0 is not None:
            x0 = m.rest if m.x0 is None else m.x0
            v0 = np.zeros_like(x0) if m.v0 is None else m.v0
            self.ctx.set_particle_state(first, x=x0, x_prev=x0, v=v0, v_prev=v0)
        if getattr(m, "ext_accel", None) is not None:
            self.ctx.set_external_accel(first, m.ext_accel)
        return first

    def _timeStep(self) -> None:
        self.solver.solve_Parallel()
        self.time += self.scene.dt

    def update(self, n_steps: Optional[int] = None) -> None:
        if not self._setup:
            self.setup()
        n = self.scene.steps if n_steps is None else n_steps
        if self.logger is None:
            self.ctx.step(n)
            self.time += n * self.scene.dt
            return
        if not self.logger._logging:
            self.logger.start_logging()
        for _ in range(n):
            self.ctx.step(1)
            self.time += self.scene.dt
            self.logger.log_to_file(self.time)

    def positions(self) -> np.ndarray:
        return self.ctx.positions()
