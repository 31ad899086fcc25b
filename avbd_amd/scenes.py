"""Synthetic scene generators for the solver's workloads (host-side, numpy).

Axis convention is the reference's: y-up, gravity (0, -g, 0)
(simulation/VBDSolver.hpp:51), ground plane y = 0
(energy/collision/GroundCollisionEnergySolver.hpp:22-24).

Workloads:
  * ``hanging_cube``  Kuhn 8^3 cube, top face fixed, jittered, no contact
  * ``dropped_cube``  Kuhn 32^3 cube, random rotation, lowest point at 0.5 m, ground AL
  * ``cube_batch``    512 independent Kuhn cubes (n ~ U{4..16}) dropped on the ground
All randomness comes from SplitMix64 so the inputs are reproducible bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform() uses the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniforms(self, n: int) -> np.ndarray:
        return np.array([self.uniform() for _ in range(n)], dtype=np.float64)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def kuhn_cube(n: int, edge: float = 1.0):
    """Kuhn (Freudenthal) triangulation of an n^3 grid of cubes with side `edge`.

    Vertex (i, j, k) sits at (i, j, k) * edge/n with index (i*(n+1) + j)*(n+1) + k.
    Each cell is split into 6 tets along its main diagonal, one per permutation
    of the axes: v0 = corner, v1 = v0 + e_a, v2 = v1 + e_b, v3 = far corner.
    Returns (vertices float64 [(n+1)^3, 3], tets uint32 [6 n^3, 4]).
    """
    m = n + 1
    h = edge / n
    ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    verts = np.stack([ii, jj, kk], axis=-1).reshape(-1, 3).astype(np.float64) * h

    def vid(i, j, k):
        return (i * m + j) * m + k

    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    out = np.empty((ci.size, 6, 4), dtype=np.uint32)
    for pi, perm in enumerate(perms):
        cur = [ci.copy(), cj.copy(), ck.copy()]
        ids = [vid(*cur)]
        for ax in perm:
            cur[ax] = cur[ax] + 1
            ids.append(vid(*cur))
        out[:, pi, :] = np.stack(ids, axis=-1)
    return verts, out.reshape(-1, 4)


def surface_triangles(n: int) -> np.ndarray:
    """Boundary triangles of the Kuhn cube (2 per boundary cell face): 12 n^2."""
    m = n + 1

    def vid(i, j, k):
        return (i * m + j) * m + k

    tris = []
    for axis in range(3):
        for side in (0, n):
            a, b = [ax for ax in range(3) if ax != axis]
            for u in range(n):
                for w in range(n):
                    c = [0, 0, 0]
                    c[axis] = side
                    quad = []
                    for du, dw in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        cc = list(c)
                        cc[a] = u + du
                        cc[b] = w + dw
                        quad.append(vid(*cc))
                    tris.append((quad[0], quad[1], quad[2]))
                    tris.append((quad[0], quad[2], quad[3]))
    return np.asarray(tris, dtype=np.uint32)


def random_rotation(rng: SplitMix64) -> np.ndarray:
    """Rotation about an axis uniform on S^2 by an angle uniform in [0, pi]."""
    z = 1.0 - 2.0 * rng.uniform()
    phi = 2.0 * np.pi * rng.uniform()
    r = np.sqrt(max(0.0, 1.0 - z * z))
    axis = np.array([r * np.cos(phi), r * np.sin(phi), z])
    ang = np.pi * rng.uniform()
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


@dataclass
class TetMesh:
    """One tet-mesh object: rest state plus material (TetMeshObjectConfig fields,
    config/TetMeshObjectConfig.hpp:18-32) and its initial state."""
    rest: np.ndarray            # [nv, 3] rest positions (already placed in the world)
    tets: np.ndarray            # [nt, 4] uint32
    E: float = 1e6
    nu: float = 0.4
    density: float = 1000.0
    kd: float = 0.0
    ground: bool = False
    mu_s: float = 0.4
    mu_k: float = 0.2
    x0: Optional[np.ndarray] = None      # initial positions (default: rest)
    v0: Optional[np.ndarray] = None
    fixed: Optional[np.ndarray] = None   # uint8 [nv]
    triangles: Optional[np.ndarray] = None
    ext_accel: Optional[np.ndarray] = None   # [nv, 3] external acceleration added to gravity (N/kg)


@dataclass
class Rod:
    """A Cosserat rod (RodConfig fields, config/RodConfig.hpp:18-57): elements+1
    oriented nodes from `pos` along the body z axis of `rot` (quaternion
    w, x, y, z).  base_joint holds the first node at its initial pose with a
    one-sided fixed joint (the form left commented out at simobject/Rod.cpp:79-89)."""
    elements: int
    length: float = 1.0
    E: float = 1e7
    nu: float = 0.3
    density: float = 1000.0
    diameter: float = 0.01
    pos: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    rot: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 0.0)
    vel: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    curvature: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    base_fixed: bool = False
    tip_fixed: bool = False
    ground: bool = False
    base_joint: bool = False
    # rigid sphere at the tip: (radius, mass, body-frame angular velocity).
    # Welded (tip_joint False): its mass and inertia 2/5 m r^2 are added to
    # the tip node, which gets a sphere ground contact.  Jointed (tip_joint
    # True, straight rods only): a sphere body of its own centred one radius
    # beyond the tip along the rod axis, held to the tip node by a two-sided
    # fixed joint at the tip, with the sphere's ground contact.
    tip_sphere: Optional[Tuple[float, float, Tuple[float, float, float]]] = None
    tip_joint: bool = False
    collide: bool = True   # rod segments are collision primitives when the scene has contacts


def quat_rotate(q, v) -> np.ndarray:
    """Rotate v by the unit quaternion q = (w, x, y, z)."""
    w, u = q[0], np.asarray(q[1:], dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def tip_body(r: "Rod"):
    """Centre and density of a jointed tip sphere (mass m exactly as the
    sphere constructor's 4/3 pi r^3 density) and the joint's offsets: at the
    tip node's origin, and one radius back along the sphere's body z axis."""
    if any(c != 0.0 for c in r.curvature):
        raise ValueError("a jointed tip sphere needs a straight rod")
    rad, m, _ = r.tip_sphere
    centre = np.asarray(r.pos, dtype=np.float64) + quat_rotate(r.rot, (0.0, 0.0, r.length + rad))
    density = m / (4.0 / 3.0 * np.pi * rad ** 3)
    pos_offsets = ((0.0, 0.0, 0.0), (0.0, 0.0, -rad))
    rot_offsets = ((1.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0))
    return centre, density, pos_offsets, rot_offsets


@dataclass
class Sphere:
    """A rigid sphere (RigidSphereConfig, config/RigidSphereConfig.hpp) with an
    initial angular velocity w0 (body frame)."""
    radius: float
    density: float
    pos: Tuple[float, float, float]
    rot: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 0.0)
    w0: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    fixed: bool = False
    ground: bool = True
    v0: Tuple[float, float, float] = (0.0, 0.0, 0.0)   # initial linear velocity (v and v_prev)


@dataclass
class Scene:
    """A batch of independent scenes sharing one solver configuration."""
    name: str
    meshes: List[TetMesh] = field(default_factory=list)
    rods: List[Rod] = field(default_factory=list)
    spheres: List[Sphere] = field(default_factory=list)
    dt: float = 1.0 / 60.0
    g: float = 9.81
    iters: int = 20
    # Chebyshev rho: the reference's SimulationConfig default
    # (config/SimulationConfig.hpp:148) for scenes that do not set it; its
    # YAML scenes set 0.95
    rho: float = 0.5
    steps: int = 100
    # collision margin (SimulationParams::collision_margin); when set, every
    # step runs the triangle narrow phase against rods and spheres and keeps
    # the resulting dynamic contacts (CollisionDetector::detectCollisionsAndRecolor)
    contact_margin: Optional[float] = None
    # vertex-face contacts between the collision triangles (new; needs contact_margin)
    self_collision: bool = False
    # several meshes: independent scenes of a batch (True) or objects of one
    # simulation (False, as a YAML scene with several TetMeshObjects)
    batch: bool = True
    # creation order of the object kinds (particle and primitive numbering)
    object_order: Tuple[str, ...] = ("meshes", "rods", "spheres")
    # logging (SimulationConfig logging / logging-output-folder /
    # logging-interval / log-residuals): dict(output_dir, interval, log_residuals)
    logging: Optional[dict] = None

    @property
    def n_vertices(self) -> int:
        return int(sum(m.rest.shape[0] for m in self.meshes))

    @property
    def n_tets(self) -> int:
        return int(sum(m.tets.shape[0] for m in self.meshes))

    @property
    def n_ground(self) -> int:
        return int(sum(m.rest.shape[0] for m in self.meshes if m.ground))


def hanging_cube(n: int = 8, jitter: float = 1e-3, seed: int = 1, iters: int = 20, steps: int = 100) -> Scene:
    """1 m Kuhn cube, top face (y = 1) fixed, E = 1e5, nu = 0.45, rho = 1000,
    no contact; free vertices jittered by U(-jitter, jitter) per component
    (SplitMix64(seed), vertex-index order) in both x and x_prev; v = 0."""
    verts, tets = kuhn_cube(n)
    fixed = (np.abs(verts[:, 1] - 1.0) < 1e-12).astype(np.uint8)
    rng = SplitMix64(seed)
    x0 = verts.copy()
    for i in range(verts.shape[0]):
        if fixed[i]:
            continue
        for k in range(3):
            x0[i, k] += (2.0 * rng.uniform() - 1.0) * jitter
    mesh = TetMesh(rest=verts, tets=tets, E=1e5, nu=0.45, density=1000.0, kd=0.0, ground=False,
                   x0=x0, fixed=fixed, triangles=surface_triangles(n))
    return Scene(name=f"hanging_cube_{n}", meshes=[mesh], iters=iters, steps=steps)


def dropped_cube(n: int = 32, seed: int = 2, height: float = 0.5, E: float = 1e6, nu: float = 0.4,
                 iters: int = 10, steps: int = 200, self_collision_margin: Optional[float] = None) -> Scene:
    """1 m Kuhn cube rotated about its centre by a random rotation, translated so
    its lowest point is at y = height, with ground contact on every vertex.
    self_collision_margin: the surface's vertex-face self collision through
    the dynamic-contact pipeline, with that margin."""
    verts, tets = kuhn_cube(n)
    R = random_rotation(SplitMix64(seed))
    c = verts.mean(axis=0)
    p = (verts - c) @ R.T
    p[:, 1] += height - p[:, 1].min()
    mesh = TetMesh(rest=p, tets=tets, E=E, nu=nu, density=1000.0, kd=0.0, ground=True,
                   triangles=surface_triangles(n))
    if self_collision_margin is not None:
        return Scene(name=f"dropped_cube_{n}", meshes=[mesh], iters=iters, steps=steps,
                     contact_margin=self_collision_margin, self_collision=True)
    return Scene(name=f"dropped_cube_{n}", meshes=[mesh], iters=iters, steps=steps)


def cube_batch(n_scenes: int = 512, seed: int = 4, n_min: int = 4, n_max: int = 16, iters: int = 10,
               steps: int = 50) -> Scene:
    """Independent Kuhn cubes: n ~ U{n_min..n_max}, E log-uniform [1e4, 1e6],
    nu ~ U[0.3, 0.48], random rotation, lowest point at U[0.2, 1.0] m, ground
    contact.  Draw order per scene: n, E, nu, rotation (3 draws), height."""
    rng = SplitMix64(seed)
    sc = Scene(name=f"cube_batch_{n_scenes}", iters=iters, steps=steps)
    for _ in range(n_scenes):
        n = rng.randint(n_min, n_max)
        E = 10.0 ** (4.0 + 2.0 * rng.uniform())
        nu = 0.3 + 0.18 * rng.uniform()
        R = random_rotation(rng)
        h = 0.2 + 0.8 * rng.uniform()
        verts, tets = kuhn_cube(n)
        p = (verts - verts.mean(axis=0)) @ R.T
        p[:, 1] += h - p[:, 1].min()
        sc.meshes.append(TetMesh(rest=p, tets=tets, E=E, nu=nu, density=1000.0, kd=0.0, ground=True))
    return sc


def quat_about(axis, angle: float) -> Tuple[float, float, float, float]:
    """Unit quaternion (w, x, y, z) of a rotation by `angle` about `axis`."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    s = np.sin(0.5 * angle)
    return (float(np.cos(0.5 * angle)), float(s * a[0]), float(s * a[1]), float(s * a[2]))


def rod_cantilever(elements: int = 8, length: float = 0.2, E: float = 1e8, diameter: float = 0.02,
                   iters: int = 100, dt: float = 1.0 / 600.0, steps: int = 500) -> Scene:
    """One horizontal rod along +x, base node fixed, sagging under gravity."""
    rod = Rod(elements=elements, length=length, E=E, nu=0.3, density=1000.0, diameter=diameter,
              rot=quat_about((0, 1, 0), np.pi / 2), base_fixed=True)
    return Scene("rod_cantilever", rods=[rod], dt=dt, iters=iters, rho=0.95, steps=steps)


def rods_and_spheres(n_rods: int = 4, elements: int = 16, seed: int = 3, dt: float = 1.0 / 120.0,
                     iters: int = 20, steps: int = 120) -> Scene:
    """Hanging rods held at the top by one-sided fixed joints, swinging from a
    tilted start and touching the ground, plus spinning spheres resting on the
    ground (rigid-ground contacts on every node and sphere)."""
    rng = SplitMix64(seed)
    rods, spheres = [], []
    for i in range(n_rods):
        tilt = 0.3 + 0.4 * rng.uniform()
        rot = quat_about((1.0, 0.0, 0.2 * (rng.uniform() - 0.5)), np.pi / 2 + tilt)
        rods.append(Rod(elements=elements, length=0.5, E=1e7, nu=0.3, density=1100.0, diameter=0.01,
                        pos=(0.1 * i, 0.49, 0.0), rot=rot, ground=True, base_joint=True))
    for i in range(n_rods):
        w0 = tuple(float(10.0 * (rng.uniform() - 0.5)) for _ in range(3))
        spheres.append(Sphere(radius=0.02, density=1500.0, pos=(0.1 * i, 0.02, 0.3), w0=w0))
    return Scene("rods_and_spheres", rods=rods, spheres=spheres, dt=dt, iters=iters, rho=0.95, steps=steps)


def rod_grid(nx: int = 16, nz: int = 16, segments: int = 128, seed: int = 3, dt: float = 1.0 / 120.0,
             iters: int = 20, steps: int = 120, tip_joint: bool = False,
             contact_margin: Optional[float] = None) -> Scene:
    """Rods + rigid bodies: nx x nz Cosserat rods at 5 cm spacing, 1 m long,
    radius 5 mm, E = 1e7 Pa, nu = 0.3 (G = E/2.6), density 1100 kg/m^3, hanging
    down from a one-sided fixed joint at the top node, a rigid sphere
    (r = 2 cm, m = 0.05 kg) at the tip with angular velocity U[-5, 5]^3 rad/s
    (SplitMix64(seed)); every node and sphere has a rigid-ground contact.
    Tops at y = 1.05 m.  tip_joint False: the sphere is welded into the tip
    node (centre at the tip, 3 cm above the ground); True: it is a body of its
    own below the tip (1 cm above the ground), held by a two-sided fixed
    joint.  contact_margin: dynamic contacts between the spheres (the rod
    segments are not collision primitives)."""
    rng = SplitMix64(seed)
    down = quat_about((1.0, 0.0, 0.0), np.pi / 2)   # body z -> world -y
    rods = []
    for i in range(nx):
        for k in range(nz):
            w0 = tuple(float(10.0 * rng.uniform() - 5.0) for _ in range(3))
            rods.append(Rod(elements=segments, length=1.0, E=1e7, nu=0.3, density=1100.0, diameter=0.01,
                            pos=(0.05 * i, 1.05, 0.05 * k), rot=down, ground=True, base_joint=True,
                            tip_sphere=(0.02, 0.05, w0), tip_joint=tip_joint, collide=contact_margin is None))
    return Scene("rod_grid", rods=rods, dt=dt, iters=iters, rho=0.95, steps=steps, contact_margin=contact_margin)


def soft_block(n: int, edge: float, E: float = 2e4, nu: float = 0.45, density: float = 1000.0) -> TetMesh:
    """Kuhn block [0, edge]^3 with its bottom layer (y = 0) held fixed and its
    boundary triangles as collision primitives."""
    verts, tets = kuhn_cube(n, edge)
    fixed = (np.abs(verts[:, 1]) < 1e-12).astype(np.uint8)
    return TetMesh(rest=verts, tets=tets, E=E, nu=nu, density=density, kd=0.0, ground=False, fixed=fixed,
                   triangles=surface_triangles(n))


def ball_on_block(n: int = 6, edge: float = 0.2, radius: float = 0.04, drop: float = 0.004, speed: float = 0.5,
                  density: float = 2000.0, dt: float = 1.0 / 120.0, iters: int = 20, steps: int = 60,
                  margin: float = 1e-3) -> Scene:
    """A rigid sphere falling onto a soft block: triangle-rigid contacts found
    by the continuous triangle-SDF test.  The sphere starts `drop` above the
    top face, off the face centre, moving down at `speed` and spinning."""
    block = soft_block(n, edge)
    c = 0.5 * edge
    sphere = Sphere(radius=radius, density=density, pos=(c + 0.013, edge + radius + drop, c - 0.007),
                    w0=(0.0, 2.0, 0.0), ground=False, v0=(0.0, -speed, 0.0))
    return Scene("ball_on_block", meshes=[block], spheres=[sphere], dt=dt, iters=iters, rho=0.95, steps=steps,
                 contact_margin=margin)


def block_stack(n: int = 6, edge: float = 0.2, n_top: int = 3, top_edge: float = 0.08, gap: float = 5e-4,
                E_top: float = 5e4, dt: float = 1.0 / 120.0, iters: int = 20, steps: int = 60,
                margin: float = 3e-3) -> Scene:
    """A soft cube set down on a soft block, both with their boundary
    triangles as collision primitives and vertex-face contacts on (new): the
    cube's bottom vertices meet the block's top faces and the block's top
    vertices the cube's bottom faces.  The cube starts `gap` above the block,
    off the block's vertex grid."""
    block = soft_block(n, edge)
    verts, tets = kuhn_cube(n_top, top_edge)
    off = np.array([0.5 * (edge - top_edge) + 0.013, edge + gap, 0.5 * (edge - top_edge) - 0.007])
    top = TetMesh(rest=verts + off, tets=tets, E=E_top, nu=0.4, density=1000.0, kd=0.0, ground=False,
                  triangles=surface_triangles(n_top))
    return Scene("block_stack", meshes=[block, top], dt=dt, iters=iters, rho=0.95, steps=steps,
                 contact_margin=margin, self_collision=True, batch=False)


def rod_on_block(n: int = 6, edge: float = 0.2, elements: int = 12, length: float = 0.16, diameter: float = 0.01,
                 drop: float = 0.003, dt: float = 1.0 / 120.0, iters: int = 20, steps: int = 60,
                 margin: float = 1e-3) -> Scene:
    """A Cosserat rod (E = 1e6 Pa) dropped across a soft block, as in the
    reference's tissue scene (a rod falling onto a tet mesh): triangle-rod
    contacts from the segment-triangle closest points."""
    block = soft_block(n, edge)
    c = 0.5 * edge
    rod = Rod(elements=elements, length=length, E=1e6, nu=0.4, density=1000.0, diameter=diameter,
              pos=(c - 0.5 * length, edge + 0.5 * diameter + drop, c + 0.011),
              rot=quat_about((0.05, 1.0, 0.0), np.pi / 2))
    return Scene("rod_on_block", meshes=[block], rods=[rod], dt=dt, iters=iters, rho=0.95, steps=steps,
                 contact_margin=margin)


def hail_on_block(n: int = 24, edge: float = 1.0, grid: int = 6, radius: float = 0.04, speed: float = 0.3,
                  dt: float = 1.0 / 60.0, iters: int = 10, steps: int = 60, margin: float = 1e-3) -> Scene:
    """grid x grid rigid spheres (r = 4 cm, 2000 kg/m^3, 16 cm apart, no
    spin) falling at `speed` onto a soft block (n^3 cells, E = 5e6 Pa,
    bottom layer held): many simultaneous triangle-rigid contacts.  The
    spheres stay apart over the benchmarked steps."""
    block = soft_block(n, edge, E=5e6, nu=0.4)
    spacing = 0.16
    x0 = 0.5 * edge - 0.5 * spacing * (grid - 1)
    spheres = []
    for i in range(grid):
        for k in range(grid):
            h = edge + radius + 0.004 + 0.01 * ((i + 2 * k) % 3)
            spheres.append(Sphere(radius=radius, density=2000.0, pos=(x0 + spacing * i, h, x0 + spacing * k),
                                  ground=False, v0=(0.0, -speed, 0.0)))
    return Scene("hail_on_block", meshes=[block], spheres=spheres, dt=dt, iters=iters, rho=0.95, steps=steps,
                 contact_margin=margin)


def sphere_pile(radius: float = 0.05, density: float = 1000.0, spin: float = 3.0, seed: int = 11,
                dt: float = 1.0 / 120.0, iters: int = 20, steps: int = 60, margin: float = 1e-3) -> Scene:
    """Six rigid spheres (r = 5 cm) with ground contact: three in a row just
    above the ground, two dropped into the grooves between them and one on
    top, each spinning at up to `spin` rad/s about a seeded axis, so that
    rigid-rigid (sphere-sphere) contacts with friction form and persist
    next to the rigid-ground ones."""
    rng = SplitMix64(seed)
    r = radius
    h0 = r + 0.002
    pts = [(-2.02 * r, h0, 0.0), (0.0, h0, 0.0), (2.02 * r, h0, 0.0),
           (-1.01 * r, h0 + 1.75 * r + 0.006, 0.01 * r), (1.01 * r, h0 + 1.75 * r + 0.006, -0.01 * r),
           (0.0, h0 + 3.5 * r + 0.012, 0.0)]
    spheres = []
    for pos in pts:
        w = np.array([rng.uniform() - 0.5 for _ in range(3)])
        w = spin * w / max(np.linalg.norm(w), 1e-12) * rng.uniform()
        spheres.append(Sphere(radius=r, density=density, pos=pos, w0=tuple(float(v) for v in w), ground=True))
    return Scene("sphere_pile", spheres=spheres, dt=dt, iters=iters, rho=0.95, steps=steps, contact_margin=margin)


def gaussian_force_field(n_vertices: int, sigma: float = 10.0, seed: int = 0) -> np.ndarray:
    """Seeded Gaussian per-vertex external acceleration field: every component of
    every vertex drawn independently from N(0, sigma^2) (N/kg), by Box-Muller
    over SplitMix64 uniforms in vertex-major, component order."""
    rng = SplitMix64(seed)
    out = np.empty((n_vertices, 3))
    for i in range(n_vertices):
        for k in range(3):
            u1 = max(rng.uniform(), 1e-300)
            u2 = rng.uniform()
            out[i, k] = sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return out


def algorithmic_bytes_per_step(n_v: int, n_tet: int, n_g: int, iters: int) -> int:
    """FP32 gather model of one time step: per iteration the vertex-block
    solve (40 B per vertex, 100 B per tet incidence, four per tet, 48 B per
    ground record), Chebyshev (72 B per vertex) and dual update (88 B per
    contact); per step inertial (84 B per vertex), velocity (72 B per vertex),
    E_prev (136 B per tet) and contact update (100 B per contact):
    I*(112 N_v + 400 N_tet + 136 N_g) + 156 N_v + 136 N_tet + 100 N_g."""
    return iters * (112 * n_v + 400 * n_tet + 136 * n_g) + 156 * n_v + 136 * n_tet + 100 * n_g
