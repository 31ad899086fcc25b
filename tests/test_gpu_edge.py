"""Edge cases of the step on the device: a scene with no free vertex, and
residual targets met at the first check or never met (the interval-graph
stepping's early stop and its full budget), each against the separate
Chebyshev launch and the eager path."""
import numpy as np
import pytest

from avbd_amd import scenes

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def test_all_fixed_scene_does_not_move():
    banner("all vertices fixed")
    from avbd_amd.sim import Simulation
    sc = scenes.hanging_cube(n=4, iters=6)
    m = sc.meshes[0]
    m.fixed = np.ones(m.rest.shape[0], dtype=np.uint8)
    sim = Simulation(sc)
    sim.setup()
    x0 = sim.positions().copy()
    sim.update(5)
    x1, v1 = sim.ctx.state()
    assert np.array_equal(x0, x1)
    assert np.all(v1 == 0)


@pytest.mark.parametrize("tol,expect", [(10.0, "first"), (1e-9, "budget")])
def test_residual_target_first_check_and_budget(tol, expect):
    """tol = 10: every step stops at its first check; tol = 1e-9: no check
    converges and every step uses the whole budget.  The interval-graph replay,
    the eager launches and the separate Chebyshev launch agree bit for bit."""
    banner(f"residual target {tol:g} ({expect})")
    from avbd_amd.sim import Simulation
    K, budget = 4, 28
    sc = scenes.dropped_cube(n=5, seed=11, height=0.03)
    runs = {}
    for name, kw in (("graph", {}), ("eager", dict(use_graph=False)), ("separate", dict(separate_chebyshev=True))):
        s = Simulation(sc, lanes_per_vertex=8, block_size=64, residual_target=(tol, budget, K), **kw)
        s.setup()
        s.update(7)
        runs[name] = s
    st = runs["graph"].ctx.stats()
    want = K if expect == "first" else budget
    banner(f"  last step {st.last_step_iterations} iterations, total {st.iterations_total}")
    assert st.last_step_iterations == want
    assert st.iterations_total == 7 * want
    assert st.steps_at_budget == (0 if expect == "first" else 7)
    x = runs["graph"].positions()
    for name in ("eager", "separate"):
        assert runs[name].ctx.stats().iterations_total == st.iterations_total, name
        assert np.array_equal(runs[name].positions(), x), name
    assert runs["graph"].ctx.fused_chebyshev() and not runs["separate"].ctx.fused_chebyshev()
