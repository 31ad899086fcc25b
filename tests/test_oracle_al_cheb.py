"""Pins of two oracle functions that the scene-level tests leave open: the
Chebyshev weight schedule of the iteration loop and the friction-cone branch
of the ground contact's dual update.

* Chebyshev (simulation/VBDSolver.hpp:65-74,513-536): the weights are checked
  against the closed form of the Chebyshev semi-iterative method,
  W_1 = 1, W_k = 2 T_{k-1}(1/rho) / (rho T_k(1/rho)) (Golub & Varga 1961, the
  recurrence Wang 2015 uses), evaluated with numpy's Chebyshev polynomials, and
  the iterates against x_{i} = w_i (N(x_{i-1}) - x_{i-2}) + x_{i-2}, where N is
  one vertex block solve computed here from the single-incidence
  Neo-Hookean gradient and Hessian (pinned by finite differences elsewhere)
  and numpy's linear solve.  The reference runs one plain iteration before
  the schedule starts (w = 1 for iterations 0 and 1, 2/(2 - rho^2) at 2).
* Friction cone (energy/collision/CollisionConstraintEnergySolver.hpp:60-134,
  141-167; GroundCollisionEnergySolver.hpp:20-91): one particle on the
  ground with hand-computed iterates (dt = 1, m = 1, g = 1, k = 1, so every
  quantity is a short fraction).
"""
import numpy as np
from numpy.polynomial import chebyshev as cheb
import pytest

from oracle import oracle_py


# ---------------------------------------------------------------------------
# Chebyshev schedule
# ---------------------------------------------------------------------------
def chebyshev_weight(k: int, rho: float) -> float:
    """W_k of the Chebyshev semi-iterative method for a basic iteration of
    spectral radius rho: W_1 = 1, W_k = 2 T_{k-1}(1/rho) / (rho T_k(1/rho))."""
    if k <= 1:
        return 1.0
    T = lambda n: cheb.chebval(1.0 / rho, [0.0] * n + [1.0])
    return 2.0 * T(k - 1) / (rho * T(k))


def reference_weights(iters: int, rho: float):
    # iteration 0 is a plain iteration; from iteration 1 on the weights are W_i
    return [1.0] + [chebyshev_weight(i, rho) for i in range(1, iters)]


def test_chebyshev_closed_form_matches_recurrence():
    """The closed form satisfies W_{k+1} = 4 / (4 - rho^2 W_k) (the identity
    T_{k+1} = 2 x T_k - T_{k-1}), so the two published forms agree."""
    for rho in (0.5, 0.9, 0.95):
        for k in range(2, 12):
            assert chebyshev_weight(k + 1, rho) == pytest.approx(4.0 / (4.0 - rho * rho * chebyshev_weight(k, rho)),
                                                                 rel=1e-13)
        assert chebyshev_weight(2, rho) == pytest.approx(2.0 / (2.0 - rho * rho), rel=1e-14)


TET_REST = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.0, 0.0, 0.1]])
MOVE = np.array([[0.012, -0.025, 0.008], [-0.02, 0.015, 0.01]])   # vertices 0 and 1 displaced from rest
DT, G, E, NU, RHO_M = 1.0 / 60.0, 9.81, 1e5, 0.3, 1000.0
FIXED = np.array([0, 0, 1, 1], np.uint8)


def two_free_vertices(iters: int, rho: float):
    """A tet with vertices 2 and 3 fixed at rest and vertices 0 and 1
    displaced, at rest (v = v_prev = 0), stepped once with `iters` iterations.
    With one vertex moving, the energy is quadratic in it (J is linear in each
    column of F), so one block solve is exact; two coupled vertices make the
    Gauss-Seidel sweep converge linearly, which the weights act on."""
    o = oracle_py.OracleSim(DT, G, iters, rho)
    o.add_tet_mesh(TET_REST, np.array([[0, 1, 2, 3]]), E, NU, RHO_M)
    x0 = TET_REST.copy()
    x0[:2] += MOVE
    o.set_positions(x0)
    o.set_fixed(FIXED)
    o.build_graph()
    o.step(1)
    return o, x0


def predicted_iterates(iters: int, weights):
    """Iterates of the free vertices from the Chebyshev semi-iteration around
    one sweep N: the free vertices in colour order, each replaced by its
    block solve p + (m/dt^2 I + H(p))^-1 (-m/dt^2 (p - y) - g(p)) at the
    current positions."""
    o, x0 = two_free_vertices(1, 0.0)
    t = o.tets()
    Q = t["Q"][0].reshape(3, 3)
    V, lam, mu = t["V"][0], t["lam"][0], t["mu"][0]
    m = o.masses()
    color, _ = o.colors()
    free = [v for v in np.argsort(color, kind="stable") if not FIXED[v]]
    y = x0 + DT * DT * np.array([0.0, -G, 0.0])   # v = 0: y = x + dt^2 g

    def N(x4):
        x4 = x4.copy()
        for v in free:
            md = m[v] / (DT * DT)
            _, g, H = oracle_py.nh_single(x4, Q, np.zeros((3, 3)), V, lam, mu, 0.0, DT, int(v))
            x4[v] = x4[v] + np.linalg.solve(md * np.eye(3) + H, -md * (x4[v] - y[v]) - g)
        return x4

    x = x0.copy()           # adaptive initialisation: a = 0 -> x stays
    x_last = x.copy()
    x_lastlast = x.copy()
    out = []
    for i in range(iters):
        xh = N(x)
        x = weights[i] * (xh - x_lastlast) + x_lastlast
        x[FIXED == 1] = x0[FIXED == 1]
        x_lastlast, x_last = x_last, x
        out.append(x[:2].copy())
    return out


def test_chebyshev_iterates_follow_the_semi_iterative_recurrence():
    rho, n_it = 0.95, 8
    expect = predicted_iterates(n_it, reference_weights(n_it, rho))
    # a schedule off by one iteration, or with rho instead of rho^2, predicts
    # visibly different iterates: the comparison below can tell them apart
    shifted = predicted_iterates(n_it, [1.0] + [chebyshev_weight(i + 1, rho) for i in range(1, n_it)])
    wrong_sq = predicted_iterates(n_it, [1.0, 1.0] + [chebyshev_weight(i, np.sqrt(rho)) for i in range(2, n_it)])
    scale = np.linalg.norm(MOVE)
    for k in range(1, n_it + 1):
        o, _ = two_free_vertices(k, rho)
        got = o.positions()[:2]
        assert np.linalg.norm(got - expect[k - 1]) <= 1e-10 * scale, (k, got, expect[k - 1])
    assert np.linalg.norm(shifted[-1] - expect[-1]) > 1e-4 * scale
    assert np.linalg.norm(wrong_sq[-1] - expect[-1]) > 1e-4 * scale


def test_chebyshev_is_plain_gauss_seidel_for_rho_zero():
    """rho = 0 makes every weight 1: the iterates are N applied repeatedly."""
    expect = predicted_iterates(4, [1.0] * 4)
    o, _ = two_free_vertices(4, 0.0)
    assert np.allclose(o.positions()[:2], expect[-1], rtol=0, atol=1e-12)


# ---------------------------------------------------------------------------
# Friction cone of the ground contact
# ---------------------------------------------------------------------------
# ground record layout (oracle_py.OracleSim.ground): k_start, k_n, k_t, k_b,
# lambda_n, lambda_t, lambda_b, C_n_prev, mu_s, mu_k, cp_x, cp_z, use_static, unclamped
K_N, K_T, L_N, L_T, C_PREV, CP_X, STATIC, UNCL = 1, 2, 4, 5, 7, 10, 12, 13
WARM = 0.95 * 0.99          # CONSTRAINT_ALPHA * STIFFNESS_GAMMA: step-start decay of lambda_n


def particle_on_ground(iters, v, lam_n0, use_static, unclamped, mu_s=0.4, mu_k=0.25):
    """One particle of unit mass at the origin (on the ground), dt = 1, g = 1,
    one ground contact with k_start = 1 and the given multiplier state."""
    o = oracle_py.OracleSim(1.0, 1.0, iters, 0.0)
    o.add_particles(np.zeros((1, 3)), np.ones(1))
    o.add_ground(0, 1.0, mu_s, mu_k)
    _, rec = o.ground()
    rec[0, L_N] = lam_n0
    rec[0, STATIC] = float(use_static)
    rec[0, UNCL] = float(unclamped)
    o.set_ground(rec)
    vv = np.asarray([v], dtype=np.float64)
    o.set_velocities(vv, vv)
    o.build_graph()
    o.step(1)
    return o.positions()[0], o.ground()[1][0]


def test_friction_static_to_kinetic_switch_hand_computed():
    """Static friction with lambda_n = 1 after the step-start decay; the
    particle slides at v = (1, -0.5, 0).  By hand (k = 1, m/dt^2 = 1):

    step start: cp = (0, 0); C_n_prev = C_n = 0; lambda_n = 1; x0 = (1, -0.5, 0),
    y = (1, -1.5, 0).
    iteration 0: C_n = 0.5 -> lambda_n+ = 1.5; friction trial 1 > mu_s 1.5 = 0.6
      -> clamped to 0.6; grad (0.6, -1.5, 0), H = 2 I, rhs = (-0.6, 0.5, 0)
      -> x = (0.7, -0.25, 0).
      dual: C_n = 0.25 -> lambda_n+ = 1.25, k_n = 1 + 10 * 0.25 = 3.5,
      lambda_n = 1.25; |lambda_t+| = 0.7 >= mu_s * 1.25 = 0.5 -> switch to
      kinetic, bound mu_k * lambda_n+ = 0.3125 -> lambda_t = 0.3125.
    iteration 1: lambda_n+ = 3.5 * 0.25 + 1.25 = 2.125; friction trial
      0.7 + 0.3125 > mu_k * 2.125 = 0.53125 -> 0.53125; H = diag(2, 4.5, 2);
      rhs = (0.3 - 0.53125, -1.25 + 2.125, 0) -> x = (0.584375, -1/18, 0).
      dual: C_n = 1/18 -> lambda_n = 3.5/18 + 1.25 = 13/9, k_n = 73/18;
      |lambda_t+| = 0.896875 >= mu_k 13/9 -> lambda_t = 13/36 (kinetic)."""
    # after iteration 0 only
    x1, r1 = particle_on_ground(1, (1.0, -0.5, 0.0), 1.0 / WARM, True, True)
    assert np.allclose(x1, [0.7, -0.25, 0.0], rtol=0, atol=1e-7)
    assert r1[STATIC] == 0 and r1[UNCL] == 0
    assert r1[L_N] == pytest.approx(1.25, rel=1e-12)
    assert r1[K_N] == pytest.approx(3.5, rel=1e-12)
    assert r1[L_T] == pytest.approx(0.3125, rel=1e-7)      # mu_k * lambda_n+, not mu_s
    assert r1[K_T] == pytest.approx(1.0, rel=1e-12)        # no ramp on the clamped branch
    assert r1[CP_X] == 0.0 and r1[C_PREV] == 0.0
    # both iterations
    x2, r2 = particle_on_ground(2, (1.0, -0.5, 0.0), 1.0 / WARM, True, True)
    assert np.allclose(x2, [0.584375, -1.0 / 18.0, 0.0], rtol=0, atol=1e-7)
    assert r2[STATIC] == 0 and r2[UNCL] == 0
    assert r2[L_N] == pytest.approx(13.0 / 9.0, rel=1e-12)
    assert r2[K_N] == pytest.approx(73.0 / 18.0, rel=1e-12)
    assert r2[L_T] == pytest.approx(13.0 / 36.0, rel=1e-7)


def test_friction_kinetic_to_static_at_step_start():
    """A resting particle whose last iteration left the friction unclamped
    switches from kinetic back to static at the next step start (and only
    then: the unclamped branch of the dual update never sets it)."""
    x, r = particle_on_ground(1, (0.0, 0.0, 0.0), 1.0 / WARM, False, True)
    assert np.allclose(x, 0.0, atol=1e-12)
    assert r[STATIC] == 1 and r[UNCL] == 1
    assert r[L_N] == pytest.approx(1.0, rel=1e-12) and r[L_T] == 0.0
    # same state, but the last iteration had clamped: stays kinetic
    _, r = particle_on_ground(1, (0.0, 0.0, 0.0), 1.0 / WARM, False, False)
    assert r[STATIC] == 0 and r[UNCL] == 1


def test_friction_static_holds_inside_the_cone():
    """Static friction, slow slide v = (0.2, 0, 0): the trial force 0.2 is
    below mu_s * lambda_n = 0.4, so the contact sticks at the contact point:
    x = x0 - 0.2 / 2 = 0.1 (H = 2), lambda_t = 0.1, k_t = 1 + 10 * 0.1 = 2,
    unclamped."""
    x, r = particle_on_ground(1, (0.2, 0.0, 0.0), 1.0 / WARM, True, True)
    assert np.allclose(x, [0.1, 0.0, 0.0], atol=1e-12)
    assert r[STATIC] == 1 and r[UNCL] == 1
    assert r[L_T] == pytest.approx(0.1, rel=1e-12)
    assert r[K_T] == pytest.approx(2.0, rel=1e-12)


def test_normal_multiplier_corrects_by_previous_violation():
    """A resting particle 0.1 below the ground with lambda_n = 1 after the
    decay: the step start stores C_n_prev = 0.1, so the corrected violation
    is C_n - 0.95 C_n_prev.  By hand: x0 = (0, -0.1, 0), y = (0, -1.1, 0);
    iteration 0: C_corr = 0.1 - 0.095 = 0.005, lambda_n+ = 1.005,
    rhs_y = -1 + 1.005 -> x_y = -0.1 + 0.0025 = -0.0975; dual: C_corr =
    0.0975 - 0.095 = 0.0025 -> lambda_n = 1.0025, k_n = 1 + 10 * 0.0025."""
    o = oracle_py.OracleSim(1.0, 1.0, 1, 0.0)
    o.add_particles(np.array([[0.0, -0.1, 0.0]]), np.ones(1))
    o.add_ground(0, 1.0, 0.4, 0.25)
    _, rec = o.ground()
    rec[0, L_N] = 1.0 / WARM
    rec[0, STATIC] = 1.0
    rec[0, UNCL] = 1.0
    o.set_ground(rec)
    o.build_graph()
    o.step(1)
    x, r = o.positions()[0], o.ground()[1][0]
    assert np.allclose(x, [0.0, -0.0975, 0.0], rtol=0, atol=1e-12)
    assert r[C_PREV] == pytest.approx(0.1, rel=1e-14)
    assert r[L_N] == pytest.approx(1.0025, rel=1e-12)
    assert r[K_N] == pytest.approx(1.025, rel=1e-12)
