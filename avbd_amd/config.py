"""Scene ingestion: the reference's YAML scene files and gmsh meshes.

* ``read_msh``: gmsh MSH 4.1 ASCII reader with the semantics of
  MeshUtils::loadTetMeshFromGmshFile (common/mesh/MeshUtils.cpp:66-191):
  nodes of every entity in entity order, 4-node tetrahedra (element type 4)
  and 3-node triangles (type 2, the collision / render surface), node tag t
  -> vertex t - 1.
* ``tet_mesh_from_config``, ``rod_from_config``, ``sphere_from_config``: the
  object set-up of TetMeshObject::setup (simobject/TetMeshObject.cpp:21-113),
  the Rod constructor (simobject/Rod.cpp:10-78) and RigidSphere /
  RigidObject_Base (simobject/rigid/RigidSphere.cpp:8-22,
  simobject/rigid/RigidObjectBase.cpp:10-20) from their YAML fields and
  defaults (config/TetMeshObjectConfig.hpp, config/RodConfig.hpp,
  config/RigidSphereConfig.hpp, config/RigidObjectConfig.hpp,
  config/ObjectConfig.hpp).
* ``scene_from_yaml``: SimulationConfig (config/SimulationConfig.hpp:40-91)
  -> a Scene, objects created in the reference's order (every RigidSphere,
  then every TetMeshObject, then every Rod: the ObjectConfigs type list of
  common/common.hpp:275-279), with the ground contacts that
  Simulation::_addObjectFromConfig adds for every object
  (simulation/Simulation.hpp:54-138).
"""
from __future__ import annotations

import math
from pathlib import Path
from typing import Dict, Optional, Tuple, Union

import numpy as np

from .scenes import Rod, Scene, Sphere, TetMesh


class MeshFormatError(ValueError):
    pass


def read_msh(path: Union[str, Path]) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(vertices [N, 3] float64, faces [F, 3] uint32, tets [T, 4] uint32)."""
    lines = Path(path).read_text().split("\n")
    sections: Dict[str, list] = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != f"$End{name}":
                j += 1
            if j >= len(lines):
                raise MeshFormatError(f"unterminated section ${name}")
            sections[name] = lines[i + 1:j]
            i = j + 1
        else:
            i += 1
    fmt = sections.get("MeshFormat")
    if not fmt:
        raise MeshFormatError("no $MeshFormat section")
    version, file_type = fmt[0].split()[:2]
    if not version.startswith("4.1") or file_type != "0":
        raise MeshFormatError(f"only MSH 4.1 ASCII is supported (got version {version}, file type {file_type})")
    if "Nodes" not in sections:
        raise MeshFormatError("no $Nodes section")

    def tokens(sec):
        for ln in sec:
            for t in ln.split():
                yield t

    # nodes: blocks of (dim, tag, parametric, count), tags, then coordinates
    it = tokens(sections["Nodes"])
    n_blocks, n_nodes = int(next(it)), int(next(it))
    next(it), next(it)
    blocks = []
    for _ in range(n_blocks):
        dim, tag, param, cnt = int(next(it)), int(next(it)), int(next(it)), int(next(it))
        tags = [int(next(it)) for _ in range(cnt)]
        ncoord = 3 + (dim if param else 0)
        coords = [[float(next(it)) for _ in range(ncoord)][:3] for _ in range(cnt)]
        blocks.append(((dim, tag), tags, coords))
    # gmsh::model::getEntities order: by dimension, then entity tag
    blocks.sort(key=lambda b: b[0])
    tag_list = [t for b in blocks for t in b[1]]
    if tag_list != list(range(1, n_nodes + 1)):
        # the reference indexes vertices by node tag - 1 over the entity-ordered
        # concatenation; other numberings would be read inconsistently there
        raise MeshFormatError("node tags must run 1..N in entity order")
    verts = np.asarray([c for b in blocks for c in b[2]], dtype=np.float64).reshape(-1, 3)

    faces, tets = [], []
    if "Elements" in sections:
        it = tokens(sections["Elements"])
        n_blocks = int(next(it))
        next(it), next(it), next(it)
        eblocks = []
        for _ in range(n_blocks):
            dim, tag, etype, cnt = int(next(it)), int(next(it)), int(next(it)), int(next(it))
            nper = {15: 1, 1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5}.get(etype)
            if nper is None:
                raise MeshFormatError(f"unsupported element type {etype}")
            rows = []
            for _ in range(cnt):
                next(it)   # element tag
                rows.append([int(next(it)) for _ in range(nper)])
            eblocks.append(((dim, tag), etype, rows))
        eblocks.sort(key=lambda b: b[0])
        for _, etype, rows in eblocks:
            if etype == 4:
                tets.extend(rows)
            elif etype == 2:
                faces.extend(rows)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3) - 1
    tets = np.asarray(tets, dtype=np.int64).reshape(-1, 4) - 1
    return verts, faces.astype(np.uint32), tets.astype(np.uint32)


def mass_center(x: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """ParticleMesh::massCenter (common/mesh/ParticleMesh.cpp:129-161): volume
    centroid of the closed surface from the signed tets (origin, face)."""
    v1, v2, v3 = x[faces[:, 0]], x[faces[:, 1]], x[faces[:, 2]]
    vol = np.einsum("ij,ij->i", v1, np.cross(v2, v3)) / 6.0
    cm = 0.25 * (v1 + v2 + v3)
    return (cm * vol[:, None]).sum(0) / vol.sum()


def rot_from_xyz_euler(deg) -> np.ndarray:
    """Math::RotMatFromXYZEulerAngles (common/Math.hpp:248-269), degrees."""
    x, y, z = (float(a) * math.pi / 180.0 for a in deg)
    cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    return np.array([[cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
                     [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
                     [-sy, sx * cy, cx * cy]])


def quat_from_xyz_euler(deg) -> Tuple[float, float, float, float]:
    """Math::QuaternionFromXYZEulerAngles (common/Math.hpp:298-318): qz qy qx,
    returned as (w, x, y, z)."""
    x, y, z = (float(a) * math.pi / 180.0 for a in deg)
    qx = (math.cos(0.5 * x), math.sin(0.5 * x), 0.0, 0.0)
    qy = (math.cos(0.5 * y), 0.0, math.sin(0.5 * y), 0.0)
    qz = (math.cos(0.5 * z), 0.0, 0.0, math.sin(0.5 * z))

    def mul(a, b):
        aw, ax, ay, az = a
        bw, bx, by, bz = b
        return (aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw)
    return mul(mul(qz, qy), qx)


def _vec(node, key, default):
    v = node.get(key, default)
    return tuple(float(a) for a in v)


def tet_mesh_from_config(node: dict, base_dir: Path, ground: bool = True) -> TetMesh:
    """TetMeshObject::setup: load, centre on the mass centre (position-by-com),
    scale, rotate (XYZ Euler degrees), translate; Lame parameters and lumped
    masses follow from E, nu, density in the pools."""
    fname = Path(node.get("filename", ""))
    if not fname.is_absolute():
        fname = base_dir / fname
    verts, faces, tets = read_msh(fname)
    x = verts.copy()
    if bool(node.get("position-by-com", True)):
        x = x - mass_center(x, faces)
    x = x * np.asarray(_vec(node, "scaling", (1, 1, 1)))
    x = x @ rot_from_xyz_euler(_vec(node, "rotation", (0, 0, 0))).T
    x = x + np.asarray(_vec(node, "position", (0, 0, 0)))
    vel = np.asarray(_vec(node, "velocity", (0, 0, 0)))
    v0 = np.repeat(vel[None], x.shape[0], axis=0) if np.any(vel != 0) else None
    return TetMesh(rest=x, tets=tets, E=float(node.get("E", 1e6)), nu=float(node.get("nu", 0.4)),
                   density=float(node.get("density", 1000.0)), kd=float(node.get("damping-coeff", 0.0)),
                   ground=ground, triangles=faces, v0=v0)


def rod_from_config(node: dict, ground: bool = True) -> Rod:
    """Rod constructor fields (config/RodConfig.hpp defaults)."""
    return Rod(elements=int(node.get("elements", 20)), length=float(node.get("length", 1.0)),
               E=float(node.get("E", 3e6)), nu=float(node.get("nu", 0.45)),
               density=float(node.get("density", 1000.0)), diameter=float(node.get("diameter", 0.1)),
               pos=_vec(node, "position", (0, 0, 0)), rot=quat_from_xyz_euler(_vec(node, "rotation", (0, 0, 0))),
               vel=_vec(node, "velocity", (0, 0, 0)), curvature=_vec(node, "curvature", (0, 0, 0)),
               base_fixed=bool(node.get("base-fixed", True)), tip_fixed=bool(node.get("tip-fixed", False)),
               ground=ground)


def sphere_from_config(node: dict, ground: bool = True) -> Sphere:
    """RigidSphere: position, XYZ-Euler rotation and fixed flag; the
    reference's rigid objects take no initial velocity from the config."""
    return Sphere(radius=float(node.get("radius", 0.5)), density=float(node.get("density", 1000.0)),
                  pos=_vec(node, "position", (0, 0, 0)), rot=quat_from_xyz_euler(_vec(node, "rotation", (0, 0, 0))),
                  fixed=bool(node.get("fixed", False)), ground=ground)


def scene_from_config(cfg: dict, base_dir: Union[str, Path] = ".", steps: Optional[int] = None) -> Scene:
    """SimulationConfig -> Scene (defaults of config/SimulationConfig.hpp:134-150)."""
    base_dir = Path(base_dir)
    dt = float(cfg.get("time-step", 1e-3))
    end = float(cfg.get("end-time", 60.0))
    sc = Scene(name=str(cfg.get("name", "yaml_scene")), dt=dt, g=float(cfg.get("g-accel", 9.81)),
               iters=int(cfg.get("solver-iters", 1)), rho=float(cfg.get("iter-acceleration", 0.5)),
               steps=steps if steps is not None else max(1, int(round(end / dt))), batch=False,
               object_order=("spheres", "meshes", "rods"))
    for obj in cfg.get("objects", []) or []:
        ty = obj.get("type")
        if ty == "TetMeshObject":
            sc.meshes.append(tet_mesh_from_config(obj, base_dir))
        elif ty == "Rod":
            sc.rods.append(rod_from_config(obj))
        elif ty == "RigidSphere":
            sc.spheres.append(sphere_from_config(obj))
        else:
            raise ValueError(f"unknown object type {ty!r}")
    if bool(cfg.get("logging", False)):
        sc.logging = dict(output_dir=str(base_dir / str(cfg.get("logging-output-folder", "../output/"))),
                          interval=float(cfg.get("logging-interval", 1e-2)),
                          log_residuals=bool(cfg.get("log-residuals", False)))
    # every primitive joins the collision pool; contacts other than the
    # no-op triangle-triangle pairs need a rod or sphere next to a mesh
    if sc.meshes and (sc.rods or sc.spheres):
        sc.contact_margin = float(cfg.get("collision-margin", 1e-3))
    return sc


def scene_from_yaml(path: Union[str, Path], steps: Optional[int] = None) -> Scene:
    import yaml
    path = Path(path)
    with open(path) as f:
        cfg = yaml.safe_load(f)
    return scene_from_config(cfg, path.parent, steps)
