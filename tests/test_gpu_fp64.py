"""The FP64 particle state (AVBD_FLAG_FP64_STATE): parity with the FP64 oracle
at tighter bounds than the FP32 state can meet, residual targets of 1e-4
reached with no step at the iteration budget, the fused / separate
Chebyshev and graph / eager forms bit-identical, the per-scene stopping rule
of a batch, checkpoints, and the non-finite state report."""
import numpy as np
import pytest

from avbd_amd import scenes

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def bbox(x):
    return float(np.max(x.max(axis=0) - x.min(axis=0)))


def test_fp64_one_step_and_100_steps_match_oracle(oracle):
    """Hanging cube n = 6: one step within 1e-7 of the bounding box (the FP32
    state's bar is 1e-5), 100 steps within 1e-5 (FP32: 1e-3).  What remains
    is the FP32 rounding of the material constants and of the velocities."""
    from avbd_amd.sim import Simulation
    banner("fp64 state vs oracle: hanging cube n=6")
    sc = scenes.hanging_cube(n=6)
    g = Simulation(sc, fp64_state=True, lanes_per_vertex=8, block_size=128)
    g.setup()
    o = oracle.from_scene(sc)
    g.update(1)
    o.step(1)
    xo = o.positions()
    e1 = np.abs(g.positions() - xo).max() / bbox(xo)
    g.update(99)
    o.step(99)
    xo = o.positions()
    e100 = np.abs(g.positions() - xo).max() / bbox(xo)
    banner(f"  1 step rel err {e1:.2e}, 100 steps {e100:.2e}")
    assert e1 <= 1e-7
    assert e100 <= 1e-5


def test_fp64_reaches_1e4_on_contact_steps(oracle):
    """Dropped cube n = 8 through its impact: every step iterates to a
    relative force residual of 1e-4 (checked every 10 iterations) with no
    step at the budget; the device residual after the step is within the
    target; the iteration counts follow the FP64 oracle's under the same
    rule (the trajectories differ by FP32 model constants, so counts may
    differ by a few checks)."""
    from avbd_amd.sim import Simulation
    banner("fp64 residual target 1e-4: dropped cube n=8, 30 steps")
    sc = scenes.dropped_cube(n=8, height=0.1)
    tol, budget, K = 1e-4, 5000, 10
    g = Simulation(sc, fp64_state=True, residual_target=(tol, budget, K))
    g.setup()
    o = oracle.from_scene(sc)
    o.set_residual_target(tol, budget, K)
    gi, oi = [], []
    for step in range(30):
        g.update(1)
        st = g.ctx.stats()
        gi.append(st.last_step_iterations)
        o.step(1)
        oi.append(o.last_iters())
        assert st.residual_rel <= tol * 1.01, (step, st.residual_rel)
    st = g.ctx.stats()
    banner(f"  iterations gpu {sum(gi)} oracle {sum(oi)}; per step gpu {gi}")
    assert st.steps_at_budget == 0
    assert max(gi) > 100                      # the contact steps are hard ones
    assert abs(sum(gi) - sum(oi)) <= 0.1 * sum(oi)
    xo = o.positions()
    assert np.abs(g.positions() - xo).max() / bbox(xo) <= 1e-4


@pytest.mark.parametrize("residual", [False, True])
def test_fp64_fused_separate_graph_eager_identical(residual):
    """The FP64 state's fused last colour + Chebyshev phase equals the
    separate launch bit for bit, and graph replay equals eager launches."""
    from avbd_amd.sim import Simulation
    banner(f"fp64 fused / separate / eager (residual={residual})")
    sc = scenes.dropped_cube(n=6, seed=3, height=0.02)
    sc.iters = 7
    kw = dict(fp64_state=True, lanes_per_vertex=8, block_size=128)
    if residual:
        kw["residual_target"] = (1e-3, 400, 4)
    runs = [Simulation(sc, **kw), Simulation(sc, separate_chebyshev=True, **kw), Simulation(sc, use_graph=False, **kw)]
    for r in runs:
        r.setup()
    assert runs[0].ctx.fused_chebyshev() and not runs[1].ctx.fused_chebyshev()
    for k in range(4):
        for r in runs:
            r.update(5)
        x0 = runs[0].positions()
        for r in runs[1:]:
            assert np.array_equal(r.positions(), x0), f"round {k}"
        if residual:
            its = [r.ctx.stats().iterations_total for r in runs]
            assert its[0] == its[1] == its[2]
    ra = runs[0].ctx.ground_records()
    for r in runs[1:]:
        assert bytes(r.ctx.ground_records()) == bytes(ra)


def test_batch_stops_on_the_worst_scene():
    """A batch of soft cubes and one stiff cube: the step ends only when every
    scene meets the target on its own.  Some check before that has the whole
    batch's relative residual below the target while the stiff scene is
    above it -- where a single criterion over the batch would have stopped.
    (The rule is the same code for both state precisions; the FP64 state is
    used so that the stiff scene can reach the target.)"""
    from avbd_amd.sim import Simulation
    from avbd_amd.scenes import Scene, TetMesh, kuhn_cube, random_rotation, SplitMix64
    banner("per-scene stopping rule")
    sc = Scene(name="mixed_batch", iters=10, steps=10)
    rng = SplitMix64(9)
    for k, E in enumerate([2e4, 2e4, 2e4, 2e4, 2e4, 1e6]):
        verts, tets = kuhn_cube(6)
        p = (verts - verts.mean(axis=0)) @ random_rotation(rng).T * 0.3
        p[:, 1] += 0.02 - p[:, 1].min()
        p[:, 0] += 0.5 * k
        sc.meshes.append(TetMesh(rest=p, tets=tets, E=E, nu=0.4, density=1000.0, kd=0.0, ground=True))
    tol = 1e-3
    sim = Simulation(sc, fp64_state=True, residual_target=(tol, 4000, 10))
    sim.setup()
    seen = False
    for _ in range(8):
        sim.update(1)
        whole = sim.ctx.residual_history()
        worst = sim.ctx.residual_history(worst_scene=True)
        st = sim.ctx.stats()
        assert len(whole) == len(worst) > 0
        assert worst[-1] <= tol
        assert np.all(worst + 1e-15 >= whole * 0)   # defined at every check
        if st.last_step_iterations < 4000:
            seen = seen or bool(np.any((whole[:-1] <= tol) & (worst[:-1] > tol)))
    banner(f"  a check with the batch below and the stiff scene above the target: {seen}")
    assert sim.ctx.stats().steps_at_budget == 0
    assert seen


@pytest.mark.parametrize("fp64", [False, True])
def test_nonfinite_state_is_reported_and_cleared(fp64):
    """A NaN position makes the next synchronising call fail with
    AVBD_E_NONFINITE (the download still fills its arrays), stepping is
    refused while it stands, and setting the state again clears it."""
    from avbd_amd.sim import Simulation
    from avbd_amd._lib import AvbdError, AVBD_E_NONFINITE
    banner(f"non-finite state report (fp64={fp64})")
    sc = scenes.hanging_cube(n=4, iters=5)
    sc.meshes[0].triangles = None      # no broad phase over NaN bounds
    sim = Simulation(sc, fp64_state=fp64)
    sim.setup()
    sim.update(2)
    x, v = sim.ctx.state()
    free = int(np.flatnonzero(sc.meshes[0].fixed == 0)[0])
    bad = x.copy()
    bad[free, 0] = np.nan
    sim.ctx.set_particle_state(0, x=bad)
    sim.ctx.step(1)
    with pytest.raises(AvbdError) as e:
        sim.ctx.synchronize()
    assert e.value.code == AVBD_E_NONFINITE
    with pytest.raises(AvbdError) as e:
        sim.ctx.step(1)
    assert e.value.code == AVBD_E_NONFINITE
    sim.ctx.set_particle_state(0, x=x, x_prev=x, v=v, v_prev=v)
    sim.ctx.step(2)
    sim.ctx.synchronize()
    assert np.isfinite(sim.positions()).all()


def test_fp64_checkpoint_resumes_bit_exact():
    """Checkpoint of an FP64-state context resumed in a fresh context: the
    continued run reproduces the original bit for bit."""
    from avbd_amd.sim import Simulation
    banner("fp64 checkpoint / resume")
    sc = scenes.dropped_cube(n=6, seed=5, height=0.05)
    kw = dict(fp64_state=True, lanes_per_vertex=8, block_size=128, residual_target=(1e-3, 600, 10))
    a = Simulation(sc, **kw)
    a.setup()
    a.update(6)
    blob = a.ctx.save_checkpoint()
    a.update(5)
    b = Simulation(sc, **kw)
    b.setup()
    b.ctx.load_checkpoint(blob)
    b.update(5)
    assert np.array_equal(a.positions(), b.positions())
    assert a.ctx.stats().iterations_total == b.ctx.stats().iterations_total


@pytest.mark.parametrize("fp64", [False, True])
def test_batch_scene_ends_its_step_as_on_its_own(fp64):
    """Residual-targeted batch of three cubes of different stiffness: each
    scene stops iterating at the first check where it meets the target, so
    its positions equal those of the same scene run on its own, bit for bit,
    with the same iteration count per step; the step runs the largest count
    and the vertex-block solves are the per-scene counts times each scene's
    work items."""
    from avbd_amd.sim import Simulation
    from avbd_amd.scenes import Scene, TetMesh, kuhn_cube, random_rotation, SplitMix64
    banner(f"batch scene = independent run (fp64={fp64})")
    rng = SplitMix64(4)
    meshes = []
    for k, E in enumerate([2e4, 6e4, 2e5]):
        verts, tets = kuhn_cube(5)
        p = (verts - verts.mean(axis=0)) @ random_rotation(rng).T * 0.3
        p[:, 1] += 0.01 - p[:, 1].min()
        p[:, 0] += 0.5 * k
        meshes.append(TetMesh(rest=p, tets=tets, E=E, nu=0.4, density=1000.0, kd=0.0, ground=True))
    kw = dict(fp64_state=fp64, lanes_per_vertex=8, block_size=128, gather_batch=1,
              residual_target=(1e-3, 2000, 10))
    batch = Simulation(Scene(name="batch3", iters=10, steps=10, meshes=list(meshes)), **kw)
    batch.setup()
    alone = [Simulation(Scene(name=f"alone{k}", iters=10, steps=10, meshes=[m]), **kw) for k, m in enumerate(meshes)]
    for a in alone:
        a.setup()
    off = batch.offsets
    spread = set()
    for step in range(8):
        batch.update(1)
        its = batch.ctx.scene_iterations()
        for k, a in enumerate(alone):
            a.update(1)
            assert its[k] == a.ctx.stats().last_step_iterations, (step, k)
            assert np.array_equal(batch.positions()[off[k]:off[k + 1]], a.positions()), (step, k)
        assert batch.ctx.stats().last_step_iterations == its.max()
        spread.add(len(set(its.tolist())))
    vs = sum(a.ctx.stats().vertex_updates for a in alone)
    banner(f"  last step iterations per scene {its.tolist()}; solves batch "
           f"{batch.ctx.stats().vertex_updates} alone {vs}")
    assert batch.ctx.stats().vertex_updates == vs
    assert max(spread) > 1          # the scenes did stop at different checks


def test_fp64_vertex_face_contacts_match_oracle(oracle):
    """The FP64 state with dynamic contacts (vertex-face, the only kind that
    takes positional particles alone): the stacked blocks per step for 10
    steps -- detected keys, slots, colours and conflicts bit-exact, positions
    within 1e-7 x bbox after one step and 1e-6 after ten."""
    from avbd_amd.sim import Simulation
    banner("fp64 state + vertex-face contacts vs oracle: block stack")
    sc = scenes.block_stack(steps=20)
    g = Simulation(sc, fp64_state=True)
    g.setup()
    o = oracle.from_scene(sc)
    bb = bbox(o.positions())
    for step in range(10):
        g.update(1)
        o.step(1)
        ko, so, to = o.detected()
        kd, sd, td = g.ctx.detected()
        assert np.array_equal(ko, kd) and np.array_equal(so, sd) and np.array_equal(to, td), step
        col_o, k_o = o.colors()
        col_d, k_d = g.ctx.coloring()
        assert k_o == k_d and np.array_equal(col_o, col_d), step
        assert np.array_equal(o.conflicted(), g.ctx.conflicted()), step
        err = np.abs(g.positions() - o.positions()).max() / bb
        if step in (0, 9):
            banner(f"  step {step}: {len(kd)} contacts, rel err {err:.2e}")
        assert err <= (1e-7 if step == 0 else 1e-6), (step, err)


def test_fp64_vertex_face_residual_target_follows_oracle(oracle):
    """Residual-targeted iteration with dynamic contacts (the steps launched
    eagerly, the work list re-laid out every detection round): the stacked
    blocks at 1e-4 for 20 steps -- every step ends within the target, the
    per-step iteration counts follow the FP64 oracle's under the same rule,
    and the positions stay within 1e-5 x bbox."""
    from avbd_amd.sim import Simulation
    banner("fp64 residual target 1e-4 with vertex-face contacts: block stack")
    sc = scenes.block_stack(steps=30)
    tol, budget, K = 1e-4, 5000, 10
    g = Simulation(sc, fp64_state=True, residual_target=(tol, budget, K))
    g.setup()
    o = oracle.from_scene(sc)
    o.set_residual_target(tol, budget, K)
    gi, oi = [], []
    for step in range(20):
        g.update(1)
        st = g.ctx.stats()
        gi.append(st.last_step_iterations)
        o.step(1)
        oi.append(o.last_iters())
        assert st.residual_rel <= tol * 1.01, (step, st.residual_rel)
    banner(f"  iterations gpu {gi}, oracle {oi}")
    assert g.ctx.stats().steps_at_budget == 0
    assert abs(sum(gi) - sum(oi)) <= 0.1 * sum(oi)
    xo = o.positions()
    assert np.abs(g.positions() - xo).max() / bbox(xo) <= 1e-5


def test_fp64_residual_cube_with_self_collision_on():
    """The headline configuration on a small cube: FP64 state, residual
    target 1e-4, the surface's vertex-face self collision on.  Through the
    fall and the impact of a convex cube no self contact is found, and the
    run equals the one without self collision (same iteration counts,
    positions within 1e-9 x bbox)."""
    from avbd_amd.sim import Simulation
    banner("fp64 residual 1e-4, self collision on vs off: dropped cube n=8")
    kw = dict(fp64_state=True, residual_target=(1e-4, 5000, 10), lanes_per_vertex=8, block_size=64, gather_batch=1)
    on = Simulation(scenes.dropped_cube(n=8, height=0.05, self_collision_margin=1e-3), **kw)
    off = Simulation(scenes.dropped_cube(n=8, height=0.05), **kw)
    on.setup()
    off.setup()
    for step in range(20):
        on.update(1)
        off.update(1)
        assert len(on.ctx.detected()[0]) == 0, step
        assert on.ctx.stats().last_step_iterations == off.ctx.stats().last_step_iterations, step
    x = off.positions()
    err = np.abs(on.positions() - x).max() / bbox(x)
    banner(f"  20 steps: {on.ctx.stats().iterations_total} iterations, on vs off rel err {err:.2e}")
    assert err <= 1e-9


def test_fp64_contact_residual_graphs_equal_eager_launches():
    """Residual mode with dynamic contacts replays head / interval / tail
    graphs captured for each detection round's layout (captured again when
    the layout, the active contacts or the adjacency buffers change): the
    stacked blocks for 12 steps give the same iteration counts and bit-equal
    positions as the same steps launched eagerly."""
    from avbd_amd.sim import Simulation
    banner("fp64 residual with vertex-face contacts: graphs vs eager")
    kw = dict(fp64_state=True, residual_target=(1e-4, 5000, 10), lanes_per_vertex=8, block_size=64, gather_batch=1)
    a = Simulation(scenes.block_stack(steps=20), use_graph=True, **kw)
    b = Simulation(scenes.block_stack(steps=20), use_graph=False, **kw)
    a.setup()
    b.setup()
    for step in range(12):
        a.update(1)
        b.update(1)
        assert a.ctx.stats().last_step_iterations == b.ctx.stats().last_step_iterations, step
        assert np.array_equal(a.ctx.detected()[0], b.ctx.detected()[0]), step
    assert np.array_equal(a.positions(), b.positions())
    banner(f"  12 steps: {a.ctx.stats().iterations_total} iterations, positions bit-equal")
