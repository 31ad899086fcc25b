"""ctypes wrapper of the FP64 oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker / the CPU baseline --
never by the avbd_amd product code.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_U8 = C.POINTER(C.c_uint8)

_SIG = {
    "oracle_last_error": (C.c_char_p, []),
    "oracle_new": (C.c_void_p, [C.c_double, C.c_double, C.c_uint32, C.c_double]),
    "oracle_free": (None, [C.c_void_p]),
    "oracle_set_residual_target": (None, [C.c_void_p, C.c_double, C.c_uint32, C.c_uint32]),
    "oracle_last_iters": (C.c_uint32, [C.c_void_p]),
    "oracle_total_iters": (C.c_uint64, [C.c_void_p]),
    "oracle_add_tet_mesh": (C.c_int64, [C.c_void_p, C.c_uint32, _D, C.c_uint32, _U32, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_double, C.c_double]),
    "oracle_add_particles": (C.c_uint32, [C.c_void_p, C.c_uint32, _D, _D]),
    "oracle_num_particles": (C.c_uint32, [C.c_void_p]),
    "oracle_set_collision_triangles": (None, [C.c_void_p, C.c_uint32, _U32]),
    "oracle_last_candidates": (C.c_uint64, [C.c_void_p]),
    "oracle_num_tets": (C.c_uint32, [C.c_void_p]),
    "oracle_num_ground": (C.c_uint32, [C.c_void_p]),
    "oracle_set_positions": (None, [C.c_void_p, _D, C.c_int]),
    "oracle_set_velocities": (None, [C.c_void_p, _D, _D]),
    "oracle_set_fixed": (None, [C.c_void_p, _U8]),
    "oracle_build_graph": (C.c_int, [C.c_void_p]),
    "oracle_step": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "oracle_get_positions": (None, [C.c_void_p, _D]),
    "oracle_get_velocities": (None, [C.c_void_p, _D]),
    "oracle_get_inertial": (None, [C.c_void_p, _D]),
    "oracle_get_masses": (None, [C.c_void_p, _D]),
    "oracle_get_tets": (None, [C.c_void_p, _U32, _D, _D, _D, _D, _D, _D]),
    "oracle_set_tet_eprev": (None, [C.c_void_p, _D]),
    "oracle_get_ground": (None, [C.c_void_p, _U32, _D]),
    "oracle_set_ground": (None, [C.c_void_p, _D]),
    "oracle_add_ground": (None, [C.c_void_p, C.c_uint32, C.c_double, C.c_double, C.c_double]),
    "oracle_num_colors": (C.c_uint32, [C.c_void_p]),
    "oracle_get_colors": (None, [C.c_void_p, _U32]),
    "oracle_num_adj_entries": (C.c_uint32, [C.c_void_p]),
    "oracle_num_p_neighbors": (C.c_uint32, [C.c_void_p]),
    "oracle_get_adjacency": (None, [C.c_void_p, _U32, _U32, _U32, _U32]),
    "oracle_dsatur_csr": (C.c_uint32, [C.c_uint32, _U32, _U32, _U32]),
    "oracle_stats": (None, [C.c_void_p, _D]),
    "oracle_nh_single": (C.c_double, [_D, _D, _D, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_uint32, _D, _D, C.c_int]),
    "oracle_lu_solve": (None, [C.c_int, _D, _D, _D]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    for k, (r, a) in _SIG.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def up(a):
    return None if a is None else a.ctypes.data_as(_U32)


class OracleError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise OracleError(lib.oracle_last_error().decode())


class OracleSim:
    """One reference Simulation restated in FP64 (single scene)."""

    def __init__(self, dt: float, g: float, iters: int, rho: float):
        self.h = lib.oracle_new(dt, g, iters, rho)
        self.dt, self.g, self.iters, self.rho = dt, g, iters, rho

    def __del__(self):
        if getattr(self, "h", None):
            lib.oracle_free(self.h)
            self.h = None

    def add_tet_mesh(self, rest, tets, E, nu, density, kd=0.0, with_ground=False, mu_s=0.4, mu_k=0.2) -> int:
        rest = np.ascontiguousarray(rest, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.uint32)
        base = lib.oracle_add_tet_mesh(self.h, rest.shape[0], dp(rest), tets.shape[0], up(tets), E, nu, density,
                                       kd, int(with_ground), mu_s, mu_k)
        if base < 0:
            raise OracleError(lib.oracle_last_error().decode())
        return int(base)

    def add_particles(self, x, m) -> int:
        x = np.ascontiguousarray(x, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        return int(lib.oracle_add_particles(self.h, x.shape[0], dp(x), dp(m)))

    def set_collision_triangles(self, tris):
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        lib.oracle_set_collision_triangles(self.h, t.shape[0], up(t))

    def last_candidates(self) -> int:
        return int(lib.oracle_last_candidates(self.h))

    @property
    def n(self) -> int:
        return lib.oracle_num_particles(self.h)

    def set_positions(self, x, also_prev=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib.oracle_set_positions(self.h, dp(x), int(also_prev))

    def set_velocities(self, v, v_prev=None):
        v = np.ascontiguousarray(v, dtype=np.float64)
        vp = v if v_prev is None else np.ascontiguousarray(v_prev, dtype=np.float64)
        lib.oracle_set_velocities(self.h, dp(v), dp(vp))

    def set_fixed(self, fixed):
        f = np.ascontiguousarray(fixed, dtype=np.uint8)
        lib.oracle_set_fixed(self.h, f.ctypes.data_as(_U8))

    def build_graph(self):
        _chk(lib.oracle_build_graph(self.h))

    def set_residual_target(self, tol, max_iters, check_every):
        lib.oracle_set_residual_target(self.h, tol, max_iters, check_every)

    def last_iters(self) -> int:
        return int(lib.oracle_last_iters(self.h))

    def total_iters(self) -> int:
        return int(lib.oracle_total_iters(self.h))

    def step(self, n=1, threads=1):
        _chk(lib.oracle_step(self.h, n, threads))

    def positions(self):
        out = np.empty((self.n, 3))
        lib.oracle_get_positions(self.h, dp(out))
        return out

    def velocities(self):
        out = np.empty((self.n, 3))
        lib.oracle_get_velocities(self.h, dp(out))
        return out

    def inertial(self):
        out = np.empty((self.n, 3))
        lib.oracle_get_inertial(self.h, dp(out))
        return out

    def masses(self):
        out = np.empty(self.n)
        lib.oracle_get_masses(self.h, dp(out))
        return out

    def colors(self):
        out = np.empty(self.n, dtype=np.uint32)
        lib.oracle_get_colors(self.h, up(out))
        return out, lib.oracle_num_colors(self.h)

    def tets(self):
        nt = lib.oracle_num_tets(self.h)
        idx = np.empty((nt, 4), dtype=np.uint32)
        Q = np.empty((nt, 9))
        V = np.empty(nt)
        lam = np.empty(nt)
        mu = np.empty(nt)
        kd = np.empty(nt)
        Ep = np.empty((nt, 9))
        lib.oracle_get_tets(self.h, up(idx), dp(Q), dp(V), dp(lam), dp(mu), dp(kd), dp(Ep))
        return dict(idx=idx, Q=Q, V=V, lam=lam, mu=mu, kd=kd, E_prev=Ep)

    def ground(self):
        ng = lib.oracle_num_ground(self.h)
        p = np.empty(ng, dtype=np.uint32)
        rec = np.empty((ng, 14))
        lib.oracle_get_ground(self.h, up(p), dp(rec))
        return p, rec

    def set_ground(self, rec):
        """Overwrite every ground record (rows of 14, the layout of ground())."""
        r = np.ascontiguousarray(rec, dtype=np.float64)
        lib.oracle_set_ground(self.h, dp(r))

    def add_ground(self, p, k_start, mu_s=0.4, mu_k=0.2):
        """A ground contact on particle p (CollisionConstraintEnergyPool::addEnergy)."""
        lib.oracle_add_ground(self.h, p, k_start, mu_s, mu_k)

    def adjacency(self):
        n = self.n
        ne = lib.oracle_num_adj_entries(self.h)
        nn = lib.oracle_num_p_neighbors(self.h)
        eo = np.empty((n + 1, 5), dtype=np.uint32)
        ent = np.empty((ne, 3), dtype=np.uint32)
        po = np.empty(n + 1, dtype=np.uint32)
        pn = np.empty(nn, dtype=np.uint32)
        lib.oracle_get_adjacency(self.h, up(eo), up(ent), up(po), up(pn))
        return eo, ent, po, pn

    def stats(self):
        out = np.empty(3)
        lib.oracle_stats(self.h, dp(out))
        return dict(residual_abs=out[0], residual_rel=out[1], ip=out[2])


def dsatur_csr(offsets, neighbors):
    off = np.ascontiguousarray(offsets, dtype=np.uint32)
    nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
    n = off.shape[0] - 1
    col = np.empty(n, dtype=np.uint32)
    k = lib.oracle_dsatur_csr(n, up(off), up(nb), up(col))
    return col, k


def nh_single(x4, Q, Eprev, V, lam, mu, kd, dt, local, use4=False):
    x4 = np.ascontiguousarray(x4, dtype=np.float64).reshape(12)
    Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(9)
    Ep = np.ascontiguousarray(Eprev, dtype=np.float64).reshape(9)
    G = np.empty(3)
    H = np.empty(9)
    e = lib.oracle_nh_single(dp(x4), dp(Q), dp(Ep), V, lam, mu, kd, dt, local, dp(G), dp(H), int(use4))
    return e, G, H.reshape(3, 3)


def lu_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = b.shape[0]
    x = np.empty(n)
    lib.oracle_lu_solve(n, dp(A.reshape(-1)), dp(b), dp(x))
    return x


# --- dynamic contacts (collision primitives, narrow phase, lifecycle) -------
for _k, (_r, _a2) in {
    "oracle_add_collision_triangles": (C.c_uint32, [C.c_void_p, C.c_uint32, _U32]),
    "oracle_add_collision_segments": (C.c_uint32, [C.c_void_p, C.c_uint32, _U32, C.c_double]),
    "oracle_add_collision_sphere": (C.c_uint32, [C.c_void_p, C.c_uint32, C.c_double]),
    "oracle_num_prims": (C.c_uint32, [C.c_void_p]),
    "oracle_enable_dynamic_contacts": (None, [C.c_void_p, C.c_double, C.c_uint32]),
    "oracle_set_self_collision": (None, [C.c_void_p, C.c_int]),
    "oracle_detect_and_recolor": (C.c_int, [C.c_void_p]),
    "oracle_num_contacts": (C.c_uint32, [C.c_void_p, C.c_uint32]),
    "oracle_get_contacts": (C.c_uint32, [C.c_void_p, C.c_uint32, _D, C.c_uint32]),
    "oracle_get_detected": (C.c_uint32, [C.c_void_p, C.POINTER(C.c_uint64), _U32, _U8, C.c_uint32]),
    "oracle_get_conflicted": (C.c_uint32, [C.c_void_p, _U8]),
    "oracle_closest_segment_triangle": (None, [_D, _D, _D, _D, _D, _D, _D, _D]),
    "oracle_closest_point_triangle": (None, [_D, _D, _D, _D, _D]),
    "oracle_golden_section_quadratic": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "oracle_unit_orthogonal": (None, [_D, _D, _D]),
    "oracle_contact_key": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "oracle_tri_sphere_ccd": (C.c_int, [_D, _D, _D, _D, _D, _D, C.c_double, C.c_double, C.c_double, _D, _D, _D]),
    "oracle_vertex_face": (C.c_int, [_D, _D, C.c_double, _D, _D]),
    "oracle_dyn_single": (None, [C.c_uint32, _D, _D, _D, C.c_uint32, C.c_uint32, _D, _D, _D, _D, _D]),
    "oracle_slot_sequence": (C.c_int, [C.c_uint32, C.POINTER(C.c_int32), C.c_uint32, _U32]),
    "oracle_update_collision_frame": (None, [_D, _D]),
}.items():
    _f = getattr(lib, _k)
    _f.restype = _r
    _f.argtypes = _a2

CONTACT_FIELDS = dict(slot=0, idx=slice(1, 6), n=slice(6, 9), t=slice(9, 12), b=slice(12, 15), barys=slice(15, 18),
                      s=18, cpl=slice(19, 22), k_start=22, k_n=23, k_t=24, k_b=25, lambda_n=26, lambda_t=27,
                      lambda_b=28, C_n_prev=29, use_static=30, unclamped=31)


def _add_collision_triangles(self, tris):
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    return int(lib.oracle_add_collision_triangles(self.h, t.shape[0], up(t)))


def _add_collision_segments(self, nodes, radius):
    nd = np.ascontiguousarray(nodes, dtype=np.uint32)
    return int(lib.oracle_add_collision_segments(self.h, nd.shape[0], up(nd), float(radius)))


def _add_collision_sphere(self, p, radius):
    return int(lib.oracle_add_collision_sphere(self.h, int(p), float(radius)))


def _enable_dynamic_contacts(self, margin, capacity=1 << 16, self_collision=False):
    lib.oracle_set_self_collision(self.h, int(bool(self_collision)))
    lib.oracle_enable_dynamic_contacts(self.h, float(margin), int(capacity))


def _detect_and_recolor(self):
    _chk(lib.oracle_detect_and_recolor(self.h))


def _contacts(self, kind):
    """Active contacts of pool `kind` (0 triangle-rigid, 1 triangle-rod, 2 rigid-rigid) in
    slot order, one row of 32 per contact (CONTACT_FIELDS)."""
    n = lib.oracle_num_contacts(self.h, kind)
    rec = np.zeros((max(n, 1), 32))
    got = lib.oracle_get_contacts(self.h, kind, dp(rec), n)
    return rec[:got]


def _detected(self):
    cap = 1 << 16
    keys = np.zeros(cap, dtype=np.uint64)
    slots = np.zeros(cap, dtype=np.uint32)
    types = np.zeros(cap, dtype=np.uint8)
    n = lib.oracle_get_detected(self.h, keys.ctypes.data_as(C.POINTER(C.c_uint64)), up(slots),
                                types.ctypes.data_as(_U8), cap)
    return keys[:n], slots[:n], types[:n]


def _conflicted(self):
    f = np.zeros(self.n, dtype=np.uint8)
    lib.oracle_get_conflicted(self.h, f.ctypes.data_as(_U8))
    return f.astype(bool)


OracleSim.add_collision_triangles = _add_collision_triangles
OracleSim.add_collision_segments = _add_collision_segments
OracleSim.add_collision_sphere = _add_collision_sphere
OracleSim.enable_dynamic_contacts = _enable_dynamic_contacts
OracleSim.detect_and_recolor = _detect_and_recolor
OracleSim.contacts = _contacts
OracleSim.detected = _detected
OracleSim.conflicted = _conflicted


def closest_segment_triangle(p0, p1, a, b, c):
    s = C.c_double()
    d = C.c_double()
    bc = np.empty(3)
    lib.oracle_closest_segment_triangle(dp(_a(p0, 3)), dp(_a(p1, 3)), dp(_a(a, 3)), dp(_a(b, 3)), dp(_a(c, 3)),
                                        C.byref(s), dp(bc), C.byref(d))
    return s.value, bc, d.value


def closest_point_triangle(p, a, b, c):
    out = np.empty(3)
    lib.oracle_closest_point_triangle(dp(_a(p, 3)), dp(_a(a, 3)), dp(_a(b, 3)), dp(_a(c, 3)), dp(out))
    return out


def golden_section_quadratic(a, b, c):
    return lib.oracle_golden_section_quadratic(a, b, c)


def complete_basis(n):
    t, b = np.empty(3), np.empty(3)
    lib.oracle_unit_orthogonal(dp(_a(n, 3)), dp(t), dp(b))
    return t, b


def contact_key(kind, a, b, code=0):
    return int(lib.oracle_contact_key(kind, a, b, code))


def tri_sphere_ccd(tri, tri_vel, pos, q, v, w, radius, dt, margin):
    n, bc, cp = np.empty(3), np.empty(3), np.empty(3)
    hit = lib.oracle_tri_sphere_ccd(dp(_a(tri, 9)), dp(_a(tri_vel, 9)), dp(_a(pos, 3)), dp(_a(q, 4)), dp(_a(v, 3)),
                                    dp(_a(w, 3)), radius, dt, margin, dp(n), dp(bc), dp(cp))
    return bool(hit), n, bc, cp


def vertex_face(p, tri, margin):
    """Vertex-face detection (new): hit, normal, barycentrics of the closest point."""
    n, bc = np.empty(3), np.empty(3)
    hit = lib.oracle_vertex_face(dp(_a(p, 3)), dp(_a(tri, 9)), margin, dp(n), dp(bc))
    return bool(hit), n, bc


def dyn_single(kind, x5, q5, rec26, local, dof):
    """One dynamic contact over particles 0..4: C (3), raw gradients (3 x 6),
    rotational Hessian blocks (3 x 3 x 3), accumulated G and H."""
    Cv, g, h = np.empty(3), np.empty(18), np.empty(27)
    G, H = np.empty(dof), np.empty(dof * dof)
    lib.oracle_dyn_single(kind, dp(_a(x5, 15)), dp(_a(q5, 20)), dp(_a(rec26, 26)), local, dof, dp(Cv), dp(g), dp(h),
                          dp(G), dp(H))
    return Cv, g.reshape(3, 6), h.reshape(3, 3, 3), G, H.reshape(dof, dof)


def slot_sequence(ops, capacity=1 << 20):
    """TombstonePool alloc / free sequence: op < 0 allocates, op >= 0 frees that
    slot.  Returns the allocated slot per op (-1 for frees)."""
    o = np.ascontiguousarray(ops, dtype=np.int32)
    out = np.zeros(o.shape[0], dtype=np.uint32)
    _chk(lib.oracle_slot_sequence(o.shape[0], o.ctypes.data_as(C.POINTER(C.c_int32)), capacity, up(out)))
    return [(-1 if v == 0xFFFFFFFF else int(v)) for v in out]


def update_collision_frame(n, t, b, lam_t, lam_b, new_n):
    f = np.zeros(12)
    f[0:3], f[3:6], f[6:9], f[9], f[10] = n, t, b, lam_t, lam_b
    lib.oracle_update_collision_frame(dp(f), dp(_a(new_n, 3)))
    return f[0:3], f[3:6], f[6:9], f[9], f[10]


def from_scene(scene, mesh_index=0, collide=True) -> OracleSim:
    """Oracle instance of an avbd_amd.scenes.Scene (one reference Simulation):
    objects in the scene's object order; of a batch of independent meshes,
    mesh `mesh_index` alone."""
    o = OracleSim(scene.dt, scene.g, scene.iters, scene.rho)
    order = getattr(scene, "object_order", ("meshes", "rods", "spheres"))
    batch = getattr(scene, "batch", True)
    meshes = scene.meshes[mesh_index:mesh_index + 1] if (batch and scene.meshes) else scene.meshes
    dynamic = getattr(scene, "contact_margin", None) is not None
    fixed_rows, x0_rows, v0_rows, spins, field_rows = [], [], [], [], []
    for kind in order:
        if kind == "meshes":
            for m in meshes:
                first = o.add_tet_mesh(m.rest, m.tets, m.E, m.nu, m.density, m.kd, m.ground, m.mu_s, m.mu_k)
                if getattr(m, "ext_accel", None) is not None:
                    field_rows.append((first, m.ext_accel))
                if m.fixed is not None:
                    fixed_rows.append((first, m.fixed))
                if m.x0 is not None:
                    x0_rows.append((first, m.x0))
                if m.v0 is not None:
                    v0_rows.append((first, m.v0))
                if collide and m.triangles is not None and (len(scene.meshes) == 1 or not batch):
                    o.add_collision_triangles(np.asarray(m.triangles, dtype=np.int64) + first)
        elif kind == "rods":
            for r in getattr(scene, "rods", []):
                first = o.add_rod(r.elements, r.length, r.curvature, r.E, r.nu, r.density, r.diameter, r.pos, r.rot,
                                  r.vel, r.base_fixed, r.tip_fixed, r.ground)
                if dynamic and getattr(r, "collide", True):
                    o.add_collision_segments(np.arange(first, first + r.elements + 1), 0.5 * r.diameter)
                if r.base_joint:
                    o.add_joint(first, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), r.pos, r.rot)
                if getattr(r, "tip_sphere", None) is not None and getattr(r, "tip_joint", False):
                    # a sphere body of its own below the tip, held by a two-sided joint
                    rad, m, w0 = r.tip_sphere
                    if any(c != 0.0 for c in r.curvature):
                        raise ValueError("a jointed tip sphere needs a straight rod")
                    qr = np.asarray(r.rot, dtype=np.float64)
                    u = qr[1:]
                    t = 2.0 * np.cross(u, (0.0, 0.0, r.length + rad))
                    centre = np.asarray(r.pos, dtype=np.float64) + (0.0, 0.0, r.length + rad) + qr[0] * t + np.cross(u, t)
                    density = m / (4.0 / 3.0 * np.pi * rad ** 3)
                    body = o.add_sphere(rad, density, centre, r.rot, False, r.ground)
                    if dynamic:
                        o.add_collision_sphere(body, rad)
                    spins.append((body, r.rot, w0))
                    o.add_joint2(first + r.elements, body, [(0.0, 0.0, 0.0), (0.0, 0.0, -rad)],
                                 [(1.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0)])
                elif getattr(r, "tip_sphere", None) is not None:
                    rad, m, w0 = r.tip_sphere
                    tip = first + r.elements
                    dI = np.full(3, 0.4 * m * rad * rad)
                    lib.oracle_add_mass(o.h, tip, m, dp(dI))
                    # sphere ground contact, k_start from the node's total mass (mu 0.5 / 0.2)
                    mass, _ = o.mass_fixed()
                    lib.oracle_add_rground(o.h, tip, 0, rad, mass[tip] / (scene.dt * scene.dt), 0.5, 0.2)
                    spins.append((tip, r.rot, w0))
        elif kind == "spheres":
            for sp in getattr(scene, "spheres", []):
                p = o.add_sphere(sp.radius, sp.density, sp.pos, sp.rot, sp.fixed, sp.ground)
                if dynamic:
                    o.add_collision_sphere(p, sp.radius)
                if any(c != 0.0 for c in sp.w0):
                    spins.append((p, sp.rot, sp.w0))
                if any(c != 0.0 for c in getattr(sp, "v0", (0.0, 0.0, 0.0))):
                    v0_rows.append((p, np.asarray([sp.v0], dtype=np.float64)))
    if fixed_rows:
        _, f = o.mass_fixed()
        for first, arr in fixed_rows:
            f[first:first + len(arr)] = arr
        o.set_fixed(f)
    if x0_rows:
        x = o.positions()
        for first, arr in x0_rows:
            x[first:first + len(arr)] = arr
        o.set_positions(x, also_prev=True)
    if v0_rows:
        v = o.velocities()
        for first, arr in v0_rows:
            v[first:first + len(arr)] = arr
        o.set_velocities(v, v)
    if field_rows:
        a = np.zeros((o.n, 3))
        for first, arr in field_rows:
            a[first:first + len(arr)] = arr
        o.set_external_accel(a)
    if spins:
        cur = o.rotations()
        q, w = cur["q"].copy(), cur["w"].copy()
        for p, _rot, w0 in spins:
            w[p] = w0
        o.set_rotation_state(q, w)
    o.build_graph()
    if dynamic:
        o.enable_dynamic_contacts(scene.contact_margin, self_collision=getattr(scene, "self_collision", False))
    return o


# --- LBVH ------------------------------------------------------------------
_U64 = C.POINTER(C.c_uint64)
for _k, (_r, _a) in {
    "oracle_expand_bits21": (C.c_uint64, [C.c_uint64]),
    "oracle_lbvh_build": (None, [C.c_uint32, _U32, _D, _D, C.c_double, _D, _U64, _U32, _U32, _U32, _U32, _D, _U32]),
    "oracle_lbvh_self_pairs": (C.c_uint32, [C.c_uint32, _U32, _U32, _D, C.c_uint32, _U32, C.c_uint32]),
    "oracle_lbvh_from_codes": (None, [C.c_uint32, _U64, _U32, _U32, _U32, _U32, _U32]),
}.items():
    _f = getattr(lib, _k)
    _f.restype = _r
    _f.argtypes = _a


def lbvh_build(tris, x, v, dt):
    """FP64 LBVH over triangles; returns a dict of the reference's arrays."""
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = tris.shape[0]
    nn = 2 * n - 1
    out = dict(aabb=np.empty((n, 6)), morton=np.empty(n, dtype=np.uint64), sorted=np.empty(n, dtype=np.uint32),
               left=np.empty(nn, dtype=np.uint32), right=np.empty(nn, dtype=np.uint32),
               parent=np.empty(nn, dtype=np.uint32), node_box=np.empty((nn, 6)))
    root = C.c_uint32()
    lib.oracle_lbvh_build(n, up(tris), dp(x), dp(v), dt, dp(out["aabb"]), out["morton"].ctypes.data_as(_U64),
                          up(out["sorted"]), up(out["left"]), up(out["right"]), up(out["parent"]),
                          dp(out["node_box"]), C.byref(root))
    out["root"] = root.value
    return out


def lbvh_self_pairs(bvh):
    n = bvh["sorted"].shape[0]
    cap = 1 << 22
    buf = np.empty(2 * cap, dtype=np.uint32)
    cnt = lib.oracle_lbvh_self_pairs(n, up(bvh["left"]), up(bvh["right"]), dp(bvh["node_box"]), bvh["root"],
                                     up(buf), cap)
    assert cnt <= cap
    return buf[:2 * cnt].reshape(-1, 2)


def lbvh_from_codes(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint64)
    n = codes.shape[0]
    nn = 2 * n - 1
    s = np.empty(n, dtype=np.uint32)
    l, r, p = (np.empty(nn, dtype=np.uint32) for _ in range(3))
    root = C.c_uint32()
    lib.oracle_lbvh_from_codes(n, codes.ctypes.data_as(_U64), up(s), up(l), up(r), up(p), C.byref(root))
    return dict(sorted=s, left=l, right=r, parent=p, root=root.value)


def expand_bits21(v: int) -> int:
    return int(lib.oracle_expand_bits21(v))


# --- oriented (6-DOF) particles ------------------------------------------------
for _k, (_r, _a) in {
    "oracle_add_rod": (C.c_int64, [C.c_void_p, C.c_uint32, C.c_double, _D, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _D, _D, _D, C.c_int, C.c_int, C.c_int]),
    "oracle_add_sphere": (C.c_int64, [C.c_void_p, C.c_double, C.c_double, _D, _D, C.c_int, C.c_int]),
    "oracle_add_oriented": (C.c_int64, [C.c_void_p, _D, _D, C.c_double, _D]),
    "oracle_add_cosserat": (None, [C.c_void_p, C.c_uint32, C.c_uint32, _D, C.c_double, _D]),
    "oracle_add_joint": (C.c_int64, [C.c_void_p, C.c_uint32, _D, _D, _D, _D]),
    "oracle_num_cosserat": (C.c_uint32, [C.c_void_p]),
    "oracle_num_rground": (C.c_uint32, [C.c_void_p]),
    "oracle_num_joints": (C.c_uint32, [C.c_void_p]),
    "oracle_add_joint2": (C.c_int64, [C.c_void_p, C.c_uint32, C.c_uint32, _D, _D]),
    "oracle_num_joints2": (C.c_uint32, [C.c_void_p]),
    "oracle_get_joints2": (None, [C.c_void_p, _D]),
    "oracle_joint2_single": (None, [_D, _D, _D, _D, _D, _D, _D, C.c_uint32, _D, _D, _D, _D]),
    "oracle_get_rotations": (None, [C.c_void_p, _D, _D, _D, _U8]),
    "oracle_get_mass": (None, [C.c_void_p, _D, _U8]),
    "oracle_set_rotation_state": (None, [C.c_void_p, _D, _D]),
    "oracle_add_mass": (None, [C.c_void_p, C.c_uint32, C.c_double, _D]),
    "oracle_add_rground_point": (None, [C.c_void_p, C.c_uint32, _D, C.c_double, C.c_double, C.c_double]),
    "oracle_set_external_accel": (None, [C.c_void_p, _D]),
    "oracle_add_rground": (None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                  C.c_double]),
    "oracle_get_rground": (None, [C.c_void_p, _D]),
    "oracle_get_joints": (None, [C.c_void_p, _D]),
    "oracle_cosserat_single": (C.c_double, [_D, _D, _D, C.c_double, _D, C.c_uint32, _D, _D]),
    "oracle_joint_single": (None, [_D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "oracle_rground_single": (None, [_D, _D, _D, _D, _D, _D, _D, _D]),
    "oracle_exp_s3": (None, [_D, _D]),
    "oracle_log_s3": (None, [_D, _D]),
    "oracle_qmat": (None, [_D, _D]),
    "oracle_exp_so3": (None, [_D, _D]),
    "oracle_jr": (None, [_D, _D]),
    "oracle_jr_inv": (None, [_D, _D]),
}.items():
    _f = getattr(lib, _k)
    _f.restype = _r
    _f.argtypes = _a


def _a(v, n=None):
    a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    if n is not None:
        assert a.shape[0] == n
    return a


def _add_rod(self, elements, length, curvature=(0, 0, 0), E=1e7, nu=0.3, density=1000.0, diameter=0.01,
             pos=(0, 0, 0), rot=(1, 0, 0, 0), vel=(0, 0, 0), base_fixed=False, tip_fixed=False, with_ground=False):
    r = lib.oracle_add_rod(self.h, elements, length, dp(_a(curvature, 3)), E, nu, density, diameter,
                           dp(_a(pos, 3)), dp(_a(rot, 4)), dp(_a(vel, 3)), int(base_fixed), int(tip_fixed),
                           int(with_ground))
    if r < 0:
        raise OracleError(lib.oracle_last_error().decode())
    return int(r)


def _add_sphere(self, radius, density, pos, rot=(1, 0, 0, 0), fixed=False, with_ground=True):
    return int(lib.oracle_add_sphere(self.h, radius, density, dp(_a(pos, 3)), dp(_a(rot, 4)), int(fixed),
                                     int(with_ground)))


def _add_rground_point(self, p, cp_local, k_start, mu_s=0.4, mu_k=0.2):
    """A rigid-ground contact at a fixed body-frame point (box / capsule shapes)."""
    lib.oracle_add_rground_point(self.h, int(p), dp(_a(cp_local, 3)), float(k_start), float(mu_s), float(mu_k))


def _set_external_accel(self, a):
    """Per-particle external acceleration added to gravity ((n, 3); None clears it)."""
    if a is None:
        lib.oracle_set_external_accel(self.h, None)
        return
    lib.oracle_set_external_accel(self.h, dp(_a(np.asarray(a, dtype=np.float64).reshape(-1), 3 * self.n)))


def _add_oriented(self, pos, rot, mass, inertia):
    return int(lib.oracle_add_oriented(self.h, dp(_a(pos, 3)), dp(_a(rot, 4)), mass, dp(_a(inertia, 3))))


def _add_cosserat(self, a, b, stiffness, L, kappa=(0, 0, 0)):
    lib.oracle_add_cosserat(self.h, a, b, dp(_a(stiffness, 6)), L, dp(_a(kappa, 3)))


def _add_joint(self, p, pos_offset, rot_offset, ref_pos, ref_rot):
    return int(lib.oracle_add_joint(self.h, p, dp(_a(pos_offset, 3)), dp(_a(rot_offset, 4)), dp(_a(ref_pos, 3)),
                                    dp(_a(ref_rot, 4))))


def _add_joint2(self, p0, p1, pos_offsets, rot_offsets):
    """Two-sided fixed joint (new): pos_offsets (2, 3) and rot_offsets (2, 4)
    of body p0 and body p1."""
    return int(lib.oracle_add_joint2(self.h, p0, p1, dp(_a(pos_offsets, 6)), dp(_a(rot_offsets, 8))))


def _rotations(self):
    n = self.n
    q = np.empty((n, 4))
    w = np.empty((n, 3))
    inertia = np.empty((n, 3))
    o = np.empty(n, dtype=np.uint8)
    lib.oracle_get_rotations(self.h, dp(q), dp(w), dp(inertia), o.ctypes.data_as(_U8))
    return dict(q=q, w=w, inertia=inertia, oriented=o.astype(bool))


def _mass_fixed(self):
    m = np.empty(self.n)
    f = np.empty(self.n, dtype=np.uint8)
    lib.oracle_get_mass(self.h, dp(m), f.ctypes.data_as(_U8))
    return m, f


def _rground(self):
    n = lib.oracle_num_rground(self.h)
    rec = np.empty((n, 16))
    lib.oracle_get_rground(self.h, dp(rec))
    return rec


def _joints(self):
    n = lib.oracle_num_joints(self.h)
    rec = np.empty((n, 18))
    lib.oracle_get_joints(self.h, dp(rec))
    return dict(k=rec[:, :6], lam=rec[:, 6:12], C_prev=rec[:, 12:])


def _joints2(self):
    n = lib.oracle_num_joints2(self.h)
    rec = np.empty((n, 18))
    lib.oracle_get_joints2(self.h, dp(rec))
    return dict(k=rec[:, :6], lam=rec[:, 6:12], C_prev=rec[:, 12:])


OracleSim.add_rod = _add_rod
OracleSim.add_sphere = _add_sphere
OracleSim.add_oriented = _add_oriented
OracleSim.add_rground_point = _add_rground_point
OracleSim.set_external_accel = _set_external_accel
OracleSim.add_cosserat = _add_cosserat
OracleSim.add_joint = _add_joint
OracleSim.add_joint2 = _add_joint2
def _set_rotation_state(self, q, w):
    lib.oracle_set_rotation_state(self.h, dp(_a(q, 4 * self.n)), dp(_a(w, 3 * self.n)))


OracleSim.rotations = _rotations
OracleSim.set_rotation_state = _set_rotation_state
OracleSim.mass_fixed = _mass_fixed
OracleSim.rground = _rground
OracleSim.joints = _joints
OracleSim.joints2 = _joints2


def cosserat_single(x2, q2, stiffness, L, kappa, local):
    G = np.empty(6)
    H = np.empty(36)
    e = lib.oracle_cosserat_single(dp(_a(x2, 6)), dp(_a(q2, 8)), dp(_a(stiffness, 6)), L, dp(_a(kappa, 3)), local,
                                   dp(G), dp(H))
    return e, G, H.reshape(6, 6)


def joint_single(x, q, pos_offset, rot_offset, ref_pos, ref_rot, k, lam, C_prev=(0,) * 6):
    Cv, G, H = np.empty(6), np.empty(6), np.empty(36)
    lib.oracle_joint_single(dp(_a(x, 3)), dp(_a(q, 4)), dp(_a(pos_offset, 3)), dp(_a(rot_offset, 4)),
                            dp(_a(ref_pos, 3)), dp(_a(ref_rot, 4)), dp(_a(k, 6)), dp(_a(lam, 6)), dp(_a(C_prev, 6)),
                            dp(Cv), dp(G), dp(H))
    return Cv, G, H.reshape(6, 6)


def joint2_single(x2, q2, pos_offsets, rot_offsets, k, lam, side, C_prev=(0,) * 6):
    """Two-sided joint between bodies 0 and 1 accumulated for `side`:
    C (6), the gradient rows dC/d(x, theta) of that side (6, 6), G, H."""
    Cv, grad, G, H = np.empty(6), np.empty(36), np.empty(6), np.empty(36)
    lib.oracle_joint2_single(dp(_a(x2, 6)), dp(_a(q2, 8)), dp(_a(pos_offsets, 6)), dp(_a(rot_offsets, 8)),
                             dp(_a(k, 6)), dp(_a(lam, 6)), dp(_a(C_prev, 6)), side, dp(Cv), dp(grad), dp(G), dp(H))
    return Cv, grad.reshape(6, 6), G, H.reshape(6, 6)


def rground_single(x, q, rec12, cp_local, cp_ground):
    Cv, G, H = np.empty(3), np.empty(6), np.empty(36)
    lib.oracle_rground_single(dp(_a(x, 3)), dp(_a(q, 4)), dp(_a(rec12, 12)), dp(_a(cp_local, 3)),
                              dp(_a(cp_ground, 3)), dp(Cv), dp(G), dp(H))
    return Cv, G, H.reshape(6, 6)


def _v3out(fn, arg, n):
    out = np.empty(n)
    fn(dp(_a(arg)), dp(out))
    return out


def exp_s3(v):
    return _v3out(lib.oracle_exp_s3, v, 4)


def log_s3(q):
    return _v3out(lib.oracle_log_s3, q, 3)


def qmat(q):
    return _v3out(lib.oracle_qmat, q, 9).reshape(3, 3)


def exp_so3(v):
    return _v3out(lib.oracle_exp_so3, v, 9).reshape(3, 3)


def jr(v):
    return _v3out(lib.oracle_jr, v, 9).reshape(3, 3)


def jr_inv(v):
    return _v3out(lib.oracle_jr_inv, v, 9).reshape(3, 3)
