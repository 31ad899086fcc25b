"""Checkpoint / resume through the C ABI: saving after k steps, running on,
then loading the checkpoint (into the same or a freshly built context) and
running the same steps again reproduces the positions bit for bit."""
import sys

import numpy as np
import pytest

from avbd_amd import scenes
from avbd_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _state(sim):
    x, v = sim.ctx.state()
    q, w = sim.ctx.rotations()
    return [x, v, q, w]


CASES = {
    "hanging_cube": lambda: scenes.hanging_cube(n=6),
    "dropped_cube": lambda: scenes.dropped_cube(n=6, height=0.05),
    "rods_and_spheres": lambda: scenes.rods_and_spheres(n_rods=2, elements=8),
    "ball_on_block": lambda: scenes.ball_on_block(),
    "block_stack": lambda: scenes.block_stack(),
    "jointed_rod_grid": lambda: scenes.rod_grid(nx=2, nz=2, segments=8, tip_joint=True),
}
# the FP64 state with vertex-face contacts, iterated to a residual target
KW = {"block_stack_fp64_residual": dict(fp64_state=True, residual_target=(1e-4, 5000, 10))}
CASES["block_stack_fp64_residual"] = lambda: scenes.block_stack()


@pytest.mark.parametrize("name", list(CASES))
def test_resume_reproduces_the_run(name):
    sim = Simulation(CASES[name](), **KW.get(name, {}))
    sim.setup()
    sim.ctx.step(6)
    blob = sim.ctx.save_checkpoint()
    sim.ctx.step(9)
    ref = _state(sim)
    # same context
    sim.ctx.load_checkpoint(blob)
    sim.ctx.step(9)
    again = _state(sim)
    # a fresh context built from the same scene
    sim2 = Simulation(CASES[name](), **KW.get(name, {}))
    sim2.setup()
    sim2.ctx.load_checkpoint(blob)
    sim2.ctx.step(9)
    fresh = _state(sim2)
    _log(f"{name}: checkpoint {len(blob)} bytes")
    for a, b, c in zip(ref, again, fresh):
        assert np.array_equal(a, b)
        assert np.array_equal(a, c)


def test_checkpoint_of_another_scene_is_refused():
    from avbd_amd._lib import AvbdError
    a = Simulation(scenes.hanging_cube(n=6))
    a.setup()
    blob = a.ctx.save_checkpoint()
    b = Simulation(scenes.hanging_cube(n=5))
    b.setup()
    with pytest.raises(AvbdError) as e:
        b.ctx.load_checkpoint(blob)
    assert e.value.code == -1
    with pytest.raises(AvbdError):
        a.ctx.load_checkpoint(blob[:100])
