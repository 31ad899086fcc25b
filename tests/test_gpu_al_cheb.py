"""GPU parity of the Chebyshev schedule and the ground contact's friction
cone on the small hand-computed cases of test_oracle_al_cheb.py: the HIP
solver through the C ABI against the oracle on the same inputs (FP32 against
FP64: positions within 1e-6 of the displacement scale, multipliers within
1e-5 relative; the friction branch flags bit for bit)."""
import ctypes as C

import numpy as np
import pytest

from avbd_amd import _lib
from tests.test_oracle_al_cheb import (FIXED, MOVE, TET_REST, WARM, DT, G, E, NU, RHO_M, two_free_vertices)

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.mark.parametrize("iters", [1, 2, 3, 5, 8])
def test_chebyshev_two_vertex_iterates_match_oracle(iters):
    from avbd_amd.sim import SimulationContext
    banner(f"chebyshev two free vertices, {iters} iterations")
    ctx = SimulationContext(dt=DT, g_accel=G, solver_iters=iters, iter_acceleration=0.95, lanes_per_vertex=4,
                            block_size=64, use_graph=False)
    ctx.add_tet_mesh(TET_REST, np.array([[0, 1, 2, 3]]), E, NU, RHO_M)
    x0 = TET_REST.copy()
    x0[:2] += MOVE
    ctx.set_particle_state(0, x=x0, x_prev=x0)
    ctx.set_fixed(0, FIXED)
    ctx.build_graph()
    ctx.step(1)
    xg = ctx.positions()
    o, _ = two_free_vertices(iters, 0.95)
    xo = o.positions()
    err = np.abs(xg - xo).max() / np.linalg.norm(MOVE)
    banner(f"  rel err {err:.2e}")
    assert err <= 1e-6
    ctx.close()


def ground_case(iters, v, lam_n0, use_static, unclamped, x0=(0.0, 0.0, 0.0)):
    """The single particle of test_oracle_al_cheb.particle_on_ground on the GPU."""
    from avbd_amd.sim import SimulationContext
    ctx = SimulationContext(dt=1.0, g_accel=1.0, solver_iters=iters, iter_acceleration=0.0, lanes_per_vertex=4,
                            block_size=64, use_graph=False)
    vv = np.asarray([v], dtype=np.float64)
    ctx.add_particles(np.asarray([x0], dtype=np.float64), v=vv, v_prev=vv, mass=np.ones(1))
    rec = (_lib.GroundRecord * 1)()
    r = rec[0]
    r.particle_index = 0
    r.use_static = int(use_static)
    r.unclamped_lambda_tb = int(unclamped)
    r.k_start = r.k_n = r.k_t = r.k_b = 1.0
    r.lambda_n = lam_n0
    r.mu_s, r.mu_k = 0.4, 0.25
    ctx.add_energies(_lib.GROUND_COLLISION, rec)
    ctx.build_graph()
    ctx.step(1)
    x = ctx.positions()[0]
    g = ctx.ground_records()[0]
    out = dict(x=x, use_static=g.use_static, unclamped=g.unclamped_lambda_tb, lambda_n=g.lambda_n,
               lambda_t=g.lambda_t, k_n=g.k_n, k_t=g.k_t, C_n_prev=g.C_n_prev)
    ctx.close()
    return out


def test_friction_static_to_kinetic_switch():
    banner("friction: static -> kinetic switch, 2 iterations")
    r = ground_case(2, (1.0, -0.5, 0.0), 1.0 / WARM, True, True)
    assert np.allclose(r["x"], [0.584375, -1.0 / 18.0, 0.0], rtol=0, atol=1e-6)
    assert r["use_static"] == 0 and r["unclamped"] == 0
    assert r["lambda_n"] == pytest.approx(13.0 / 9.0, rel=1e-5)
    assert r["k_n"] == pytest.approx(73.0 / 18.0, rel=1e-5)
    assert r["lambda_t"] == pytest.approx(13.0 / 36.0, rel=1e-5)


def test_friction_switch_back_and_sticking():
    banner("friction: kinetic -> static at step start; sticking inside the cone")
    r = ground_case(1, (0.0, 0.0, 0.0), 1.0 / WARM, False, True)
    assert r["use_static"] == 1 and r["unclamped"] == 1
    r = ground_case(1, (0.0, 0.0, 0.0), 1.0 / WARM, False, False)
    assert r["use_static"] == 0 and r["unclamped"] == 1
    r = ground_case(1, (0.2, 0.0, 0.0), 1.0 / WARM, True, True)
    assert np.allclose(r["x"], [0.1, 0.0, 0.0], atol=1e-6)
    assert r["use_static"] == 1 and r["lambda_t"] == pytest.approx(0.1, rel=1e-5)
    assert r["k_t"] == pytest.approx(2.0, rel=1e-5)


def test_normal_correction_by_previous_violation():
    banner("ground: corrected normal violation")
    r = ground_case(1, (0.0, 0.0, 0.0), 1.0 / WARM, True, True, x0=(0.0, -0.1, 0.0))
    assert np.allclose(r["x"], [0.0, -0.0975, 0.0], rtol=0, atol=1e-6)
    assert r["C_n_prev"] == pytest.approx(0.1, rel=1e-6)
    assert r["lambda_n"] == pytest.approx(1.0025, rel=1e-5)
    assert r["k_n"] == pytest.approx(1.025, rel=1e-5)
