"""Scene ingestion: gmsh MSH 4.1 reader, object set-up formulas and YAML
scenes in the reference's schema.  The meshes under tests/golden/meshes are
the reference's own resource files (data fixtures).  CPU only."""
from pathlib import Path

import numpy as np
import pytest

from avbd_amd import config
from oracle import oracle_py as O

HERE = Path(__file__).resolve().parent
MESHES = HERE / "golden" / "meshes"
SCENES = HERE / "golden" / "scenes"


def _boundary_faces(tets):
    """Faces that belong to exactly one tet, as sorted index triples."""
    count = {}
    for t in tets:
        for f in ((t[0], t[1], t[2]), (t[0], t[1], t[3]), (t[0], t[2], t[3]), (t[1], t[2], t[3])):
            k = tuple(sorted(int(v) for v in f))
            count[k] = count.get(k, 0) + 1
    return {k for k, c in count.items() if c == 1}


@pytest.mark.parametrize("name", ["cube2", "single", "stanford_bunny_lowpoly", "tumor"])
def test_msh_surface_is_the_tet_boundary(name):
    verts, faces, tets = config.read_msh(MESHES / f"{name}.msh")
    assert verts.shape[1] == 3 and tets.shape[1] == 4 and faces.shape[1] == 3
    assert tets.max() < verts.shape[0] and faces.max() < verts.shape[0]
    # every tet has volume; the triangle elements are exactly the boundary of the tets
    X = verts[tets[:, :3]] - verts[tets[:, 3:4]]
    assert np.all(np.abs(np.linalg.det(X)) > 1e-12 * np.abs(verts).max() ** 3)
    assert {tuple(sorted(int(v) for v in f)) for f in faces} == _boundary_faces(tets)


def test_msh_cube2_known_answers():
    verts, faces, tets = config.read_msh(MESHES / "cube2.msh")
    assert verts.shape == (27, 3)
    assert np.allclose(verts.min(0), 0) and np.allclose(verts.max(0), 1)
    # the 27 nodes are the 3 x 3 x 3 grid with spacing 0.5
    assert {tuple(v) for v in np.round(verts * 2).astype(int)} == {(i, j, k) for i in range(3) for j in range(3)
                                                                   for k in range(3)}
    X = verts[tets[:, :3]] - verts[tets[:, 3:4]]
    assert abs(np.abs(np.linalg.det(X)).sum() / 6.0 - 1.0) < 1e-12   # total volume of the unit cube
    assert len(faces) == 48                                           # 6 faces x 4 squares x 2 triangles


def test_mass_center_and_rotations():
    verts, faces, _ = config.read_msh(MESHES / "cube2.msh")
    assert np.allclose(config.mass_center(verts, faces), [0.5, 0.5, 0.5])
    from scipy.spatial.transform import Rotation
    for ang in [(30, 0, 0), (0, 45, 0), (0, 0, 60), (10, -20, 75)]:
        R = config.rot_from_xyz_euler(ang)
        # "123" convention: about x first, then y, then z (extrinsic)
        assert np.allclose(R, Rotation.from_euler("xyz", ang, degrees=True).as_matrix())
        w, x, y, z = config.quat_from_xyz_euler(ang)
        assert np.allclose(R, Rotation.from_quat([x, y, z, w]).as_matrix())


def test_tumor_rod_yaml_scene():
    sc = config.scene_from_yaml(SCENES / "tumor_rod.yaml")
    assert sc.dt == 1e-3 and sc.iters == 15 and sc.rho == 0.95 and sc.steps == 50
    assert sc.contact_margin == 1e-4 and not sc.batch
    assert sc.object_order == ("spheres", "meshes", "rods")
    m = sc.meshes[0]
    # scaled to metres and centred on the mass centre, then moved to (0, 10 mm, 0)
    assert np.allclose(config.mass_center(m.rest, m.triangles), [0, 10e-3, 0], atol=1e-12)
    assert m.E == 20e3 and m.nu == 0.45 and m.kd == 1e-6 and m.ground
    r = sc.rods[0]
    assert r.elements == 20 and r.length == 20e-3 and not r.base_fixed and r.vel == (0.0, -0.3, 0.0)
    # the oracle builds it: lumped tet masses sum to density x volume
    o = O.from_scene(sc)
    mass, fixed = o.mass_fixed()
    X = m.rest[m.tets[:, :3]] - m.rest[m.tets[:, 3:4]]
    vol = np.abs(np.linalg.det(X)).sum() / 6.0
    nm = m.rest.shape[0]
    assert abs(mass[:nm].sum() - 1000.0 * vol) < 1e-12
    assert o.n == nm + 21


def test_yaml_scenes_run_in_the_oracle_with_contacts():
    for name, steps, kind in (("tumor_rod", 30, 1), ("bunny_ball", 60, 0)):
        sc = config.scene_from_yaml(SCENES / f"{name}.yaml")
        o = O.from_scene(sc)
        seen = 0
        for _ in range(steps):
            o.step(1)
            seen += len(o.contacts(kind))
        assert np.all(np.isfinite(o.positions()))
        assert seen > 0, name


def test_reader_rejects_other_formats(tmp_path):
    p = tmp_path / "old.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    with pytest.raises(config.MeshFormatError):
        config.read_msh(p)


def test_simulation_logger_format(tmp_path):
    """SimulationLogger: header of names, then one line per interval."""
    from avbd_amd.sim import SimulationLogger
    lg = SimulationLogger(tmp_path / "sub" / "log.txt", 0.01)
    t = [0.0]
    lg.add_output("time", lambda: t[0])
    lg.add_output("twice", lambda: 2 * t[0])
    lg.start_logging()
    with pytest.raises(RuntimeError):
        lg.add_output("late", lambda: 0.0)
    for k in range(1, 11):
        t[0] = k * 0.004
        lg.log_to_file(t[0])
    lg.stop_logging()
    lines = (tmp_path / "sub" / "log.txt").read_text().splitlines()
    assert lines[0] == "time twice "
    logged = [float(l.split()[0]) for l in lines[1:]]
    # first call logs, then every time >= last + interval
    assert logged == [0.004, 0.016, 0.028, 0.04]


def test_yaml_logging_fields():
    cfg = {"time-step": 1e-3, "end-time": 0.01, "logging": True, "logging-output-folder": "out",
           "logging-interval": 0.005, "log-residuals": True, "objects": []}
    sc = config.scene_from_config(cfg, "/tmp/base")
    assert sc.logging == dict(output_dir="/tmp/base/out", interval=0.005, log_residuals=True)
