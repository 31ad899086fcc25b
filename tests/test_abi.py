"""The C ABI library loads, exports every symbol include/avbd.h declares, and its
host-side pieces (staging, setup formulas, DSATUR colouring) behave -- no GPU
needed: HIP is only touched by the device part of avbd_build_graph."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from avbd_amd import scenes

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "avbd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint32_t|uint64_t|const char\*)\s+(avbd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from avbd_amd import _lib
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(_lib.lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures out of sync with include/avbd.h"


def test_struct_layouts_match_header():
    from avbd_amd import _lib
    # 4*u32 + 3 doubles + 9 + 1 + 9 doubles
    assert C.sizeof(_lib.NeoHookeanRecord) == 16 + 8 * 22
    assert C.sizeof(_lib.GroundRecord) == 16 + 8 * 12


def make_ctx(**kw):
    from avbd_amd.sim import SimulationContext
    return SimulationContext(dt=kw.pop("dt", 1 / 60), solver_iters=kw.pop("iters", 10), **kw)


def test_create_rejects_bad_descriptors():
    from avbd_amd import _lib
    from avbd_amd.sim import SimulationContext
    with pytest.raises(_lib.AvbdError) as e:
        SimulationContext(dt=0.0)
    assert e.value.code == _lib.AVBD_E_ARG
    with pytest.raises(_lib.AvbdError):
        SimulationContext(dt=1 / 60, lanes_per_vertex=3)
    with pytest.raises(_lib.AvbdError):
        SimulationContext(dt=1 / 60, block_size=100)
    with pytest.raises(_lib.AvbdError):
        SimulationContext(dt=1 / 60, gather_batch=3)


def test_capacity_error_code():
    from avbd_amd import _lib
    v, t = scenes.kuhn_cube(2)
    ctx = make_ctx(particle_capacity=v.shape[0] - 1)
    with pytest.raises(_lib.AvbdError) as e:
        ctx.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)
    assert e.value.code == _lib.AVBD_E_CAPACITY


def test_not_built_error_code():
    from avbd_amd import _lib
    v, t = scenes.kuhn_cube(2)
    ctx = make_ctx()
    ctx.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)
    with pytest.raises(_lib.AvbdError) as e:
        ctx.step(1)
    assert e.value.code == _lib.AVBD_E_NOT_BUILT


def _planned_coloring(ctx):
    """avbd_build_graph plans on the host first; on a machine without a GPU the
    device part fails (AVBD_E_HIP) but the colouring is already available."""
    from avbd_amd import _lib
    try:
        ctx.build_graph()
    except _lib.AvbdError as e:
        assert e.code == _lib.AVBD_E_HIP
    return ctx.coloring()


@pytest.mark.parametrize("n", [3, 4, 8])
def test_product_coloring_bit_exact_vs_oracle(oracle, n):
    v, t = scenes.kuhn_cube(n)
    ctx = make_ctx()
    ctx.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)
    col, k = _planned_coloring(ctx)
    o = oracle.OracleSim(1 / 60, 9.81, 10, 0.95)
    o.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)
    o.build_graph()
    ocol, ok = o.colors()
    assert k == ok
    assert np.array_equal(col, ocol)


def test_product_coloring_bit_exact_rotated_dropped_cube(oracle):
    sc = scenes.dropped_cube(n=12)
    m = sc.meshes[0]
    ctx = make_ctx()
    ctx.add_tet_mesh(m.rest, m.tets, m.E, m.nu, m.density, with_ground=True)
    col, k = _planned_coloring(ctx)
    o = oracle.from_scene(sc)
    ocol, ok = o.colors()
    assert (k, col.tolist()) == (ok, ocol.tolist())


def test_batched_coloring_is_per_scene(oracle):
    sc = scenes.cube_batch(n_scenes=6, seed=11)
    from avbd_amd.sim import Simulation
    sim = Simulation(sc)
    # set up without touching the device: stage + plan only
    offs = [0]
    for m in sc.meshes:
        first = sim._add_object(m)
        offs.append(first + m.rest.shape[0])
    sim.ctx.set_scenes(offs)
    col, k = _planned_coloring(sim.ctx)
    kmax = 0
    for i, m in enumerate(sc.meshes):
        o = oracle.from_scene(sc, i)
        ocol, ok = o.colors()
        assert np.array_equal(col[offs[i]:offs[i + 1]], ocol)
        kmax = max(kmax, ok)
    assert k == kmax


def test_pools_frozen_after_build():
    from avbd_amd import _lib
    v, t = scenes.kuhn_cube(2)
    ctx = make_ctx()
    ctx.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)
    _planned_coloring(ctx)
    with pytest.raises(_lib.AvbdError):
        ctx.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)


def test_scene_offsets_validated():
    from avbd_amd import _lib
    v, t = scenes.kuhn_cube(2)
    ctx = make_ctx()
    ctx.add_tet_mesh(v, t, 1e5, 0.45, 1000.0)
    with pytest.raises(_lib.AvbdError):
        ctx.set_scenes([0, 5])
    # a tet spanning two scenes is rejected at build time
    ctx.set_scenes([0, 5, v.shape[0]])
    with pytest.raises(_lib.AvbdError) as e:
        ctx.build_graph()
    assert "spans two scenes" in str(e.value)


def test_record_layouts_match_c_compiler(tmp_path):
    """sizeof/offsetof of every record as the C compiler lays it out from
    include/avbd.h equals the ctypes mirror."""
    import subprocess
    from avbd_amd import _lib
    src = tmp_path / "sizes.c"
    recs = ["avbd_neo_hookean_record", "avbd_ground_record", "avbd_cosserat_record", "avbd_rigid_ground_record",
            "avbd_joint_record", "avbd_fixed_joint_record", "avbd_stats", "avbd_scene_desc", "avbd_contact"]
    body = "\n".join(f'printf("%zu\\n", sizeof({r}));' for r in recs)
    src.write_text(f'#include <stdio.h>\n#include "avbd.h"\nint main(void) {{ {body} return 0; }}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    mirror = [_lib.NeoHookeanRecord, _lib.GroundRecord, _lib.CosseratRecord, _lib.RigidGroundRecord,
              _lib.JointRecord, _lib.FixedJointRecord, _lib.Stats, _lib.SceneDesc, _lib.Contact]
    assert sizes == [C.sizeof(m) for m in mirror]


def _rod_scene_ctx(scene):
    from avbd_amd.sim import Simulation
    sim = Simulation(scene)
    for r in scene.rods:
        sim._add_rod(r)
    for sp in scene.spheres:
        sim._add_sphere(sp)
    return sim


def test_rod_and_sphere_coloring_bit_exact_vs_oracle(oracle):
    sc = scenes.rods_and_spheres(n_rods=3, elements=9)
    sim = _rod_scene_ctx(sc)
    col, k = _planned_coloring(sim.ctx)
    o = oracle.from_scene(sc)
    ocol, ok = o.colors()
    assert (k, col.tolist()) == (ok, ocol.tolist())
    assert k == 2
    from avbd_amd import _lib
    assert _lib.lib.avbd_num_energies(sim.ctx.handle, _lib.COSSERAT_ROD) == 27
    assert _lib.lib.avbd_num_energies(sim.ctx.handle, _lib.RIGID_BODY_GROUND_COLLISION) == 30 + 3
    assert _lib.lib.avbd_num_energies(sim.ctx.handle, _lib.ONE_SIDED_FIXED_JOINT) == 3


def test_jointed_tip_spheres_coloring_bit_exact_vs_oracle(oracle):
    """Tip spheres as bodies of their own, held by two-sided fixed joints: the
    joint couples the tip node and its sphere in the colouring (different
    colours), and the colouring is the oracle's."""
    from avbd_amd import _lib
    sc = scenes.rod_grid(nx=2, nz=3, segments=8, tip_joint=True)
    sim = _rod_scene_ctx(sc)
    col, k = _planned_coloring(sim.ctx)
    o = oracle.from_scene(sc)
    ocol, ok = o.colors()
    assert (k, col.tolist()) == (ok, ocol.tolist())
    assert _lib.lib.avbd_num_energies(sim.ctx.handle, _lib.FIXED_JOINT) == 6
    assert _lib.lib.avbd_num_energies(sim.ctx.handle, _lib.ONE_SIDED_FIXED_JOINT) == 6
    for r in range(6):
        tip, body = r * 10 + 8, r * 10 + 9   # 9 rod nodes, then the sphere
        assert col[tip] != col[body]


def test_fixed_joint_validated():
    """A two-sided joint needs two distinct oriented particles."""
    from avbd_amd import _lib
    ctx = make_ctx()
    ctx.add_oriented_particles(np.zeros((1, 3)), inertia=np.ones((1, 3)), mass=np.ones(1))
    ctx.add_particles(np.ones((1, 3)), mass=np.ones(1))
    ctx.add_fixed_joint(0, 1, [(0, 0, 0)] * 2, [(1, 0, 0, 0)] * 2)
    with pytest.raises(_lib.AvbdError) as e:
        ctx.build_graph()
    assert "oriented particle" in str(e.value)
    ctx = make_ctx()
    ctx.add_oriented_particles(np.zeros((2, 3)), inertia=np.ones((2, 3)), mass=np.ones(2))
    ctx.add_fixed_joint(1, 1, [(0, 0, 0)] * 2, [(1, 0, 0, 0)] * 2)
    with pytest.raises(_lib.AvbdError) as e:
        ctx.build_graph()
    assert "one particle twice" in str(e.value)


def test_energy_particle_kind_validated():
    """Neo-Hookean tets need positional particles; Cosserat, rigid-ground and
    joint energies need oriented ones (the solvers' SupportsPositional /
    SupportsOriented)."""
    from avbd_amd import _lib
    ctx = make_ctx()
    ctx.add_particles(np.zeros((2, 3)), mass=np.ones(2))
    ctx.add_cosserat(0, 1, [1.0] * 6, 0.1)
    with pytest.raises(_lib.AvbdError) as e:
        ctx.build_graph()
    assert "positional particle" in str(e.value)

    ctx = make_ctx()
    ctx.add_oriented_particles(np.eye(4)[:, :3].copy(), inertia=np.ones((4, 3)), mass=np.ones(4))
    r = _lib.NeoHookeanRecord()
    for k in range(4):
        r.particle_indices[k] = k
    r.rest_volume = 1.0
    ctx.add_energies(_lib.NEO_HOOKEAN, (_lib.NeoHookeanRecord * 1)(r))
    with pytest.raises(_lib.AvbdError) as e:
        ctx.build_graph()
    assert "oriented particle" in str(e.value)

    ctx = make_ctx()
    ctx.add_particles(np.zeros((1, 3)), mass=np.ones(1))
    ctx.add_joint(0, (0, 0, 0), (1, 0, 0, 0), (0, 0, 0), (1, 0, 0, 0))
    with pytest.raises(_lib.AvbdError) as e:
        ctx.build_graph()
    assert "oriented particle" in str(e.value)


def test_dynamic_contacts_refuse_a_batch():
    """The scenes of a batch are independent, so dynamic contacts (whose
    primitives would meet across scenes) are refused for one (AVBD_E_ARG).
    Without a device avbd_set_dynamic_contacts stops at the build check
    first; both are argument errors."""
    from avbd_amd import _lib
    from avbd_amd.sim import SimulationContext
    ctx = SimulationContext(dt=1 / 60, solver_iters=4)
    ctx.add_particles(np.zeros((6, 3)), mass=np.ones(6))
    ctx.set_scenes([0, 3, 6])
    with pytest.raises(_lib.AvbdError) as e:
        ctx.set_dynamic_contacts(1e-3)
    assert e.value.code in (-1, -7)
