"""GPU parity of the oriented (6-DOF) path: Cosserat rods, rigid spheres,
rigid-ground contacts and one-sided fixed joints, through the C ABI, against
the FP64 oracle on the same scenes.

Tolerances (bbox = bounding-box extent of the scene):
  * positions after 1 step: max |x_gpu - x_oracle| <= 1e-5 * bbox;
    orientations after 1 step: max |q_gpu - q_oracle| <= 1e-5 (unit quaternions,
    sign-aligned)
  * contact scenes after 60 steps: positions <= 2e-3 * bbox (friction switches are
    discontinuous; FP32 and FP64 may branch differently near the cone boundary)
  * the free rigid body matches Euler's equations in closed form to FP32 rounding.
"""
import numpy as np
import pytest

from avbd_amd import scenes

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def bbox(x):
    return float(np.max(x.max(axis=0) - x.min(axis=0)))


def qerr(qa, qb):
    """Max over particles of the distance between unit quaternions, sign-aligned."""
    d1 = np.abs(qa - qb).max(axis=1)
    d2 = np.abs(qa + qb).max(axis=1)
    return float(np.minimum(d1, d2).max())


def gpu_sim(sc, **kw):
    from avbd_amd.sim import Simulation
    s = Simulation(sc, **kw)
    s.setup()
    return s


def test_free_rigid_body_euler_equations():
    banner("free rigid body: Euler's equations")
    from avbd_amd.sim import SimulationContext
    from oracle import oracle_py as O
    dt = 1e-3
    ctx = SimulationContext(dt=dt, solver_iters=5)
    inertia = np.array([[1.0, 2.0, 3.5]])
    q0 = np.array([[0.9, 0.1, -0.3, 0.2]])
    q0 /= np.linalg.norm(q0)
    w = np.array([[0.3, -1.2, 2.0]])
    ctx.add_oriented_particles(np.array([[0.0, 1.0, 0.0]]), inertia=inertia, q=q0, w=w, mass=np.array([2.0]))
    ctx.build_graph()
    ctx.step(1)
    q, wg = ctx.rotations()
    w_expect = w[0] - dt * np.cross(w[0], inertia[0] * w[0]) / inertia[0]
    assert np.abs(wg[0] - w_expect).max() < 2e-4 * np.abs(w_expect).max()
    qe = O.exp_s3(dt * w_expect)
    w1, x1, y1, z1 = q0[0]
    w2, x2, y2, z2 = qe
    q_expect = np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                         w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])
    assert qerr(q, q_expect[None]) < 1e-6
    x = ctx.positions()
    assert x[0, 1] == pytest.approx(1.0 - 9.81 * dt * dt, abs=1e-6)


def test_rod_cantilever_parity(oracle):
    banner("cantilever rod vs oracle: 1 and 100 steps")
    sc = scenes.rod_cantilever()
    g = gpu_sim(sc)
    o = oracle.from_scene(sc)
    col, k = g.ctx.coloring()
    ocol, ok = o.colors()
    assert k == ok == 2 and np.array_equal(col, ocol)
    g.update(1)
    o.step(1)
    xo = o.positions()
    scale = sc.rods[0].length
    err = np.abs(g.positions() - xo).max() / scale
    qg, _ = g.ctx.rotations()
    qo = o.rotations()["q"]
    banner(f"  1-step pos rel err {err:.3e}, q err {qerr(qg, qo):.3e}")
    assert err <= 1e-5
    assert qerr(qg, qo) <= 1e-5
    g.update(99)
    o.step(99)
    xo = o.positions()
    err = np.abs(g.positions() - xo).max() / scale
    qg, _ = g.ctx.rotations()
    banner(f"  100-step pos rel err {err:.3e}, q err {qerr(qg, o.rotations()['q']):.3e}")
    assert err <= 1e-3


def test_rods_and_spheres_contact_and_joints(oracle):
    banner("rods + spheres with ground contact and joints vs oracle: 1 and 60 steps")
    sc = scenes.rods_and_spheres(n_rods=4, elements=16)
    g = gpu_sim(sc)
    o = oracle.from_scene(sc)
    col, k = g.ctx.coloring()
    ocol, ok = o.colors()
    assert k == ok and np.array_equal(col, ocol)
    g.update(1)
    o.step(1)
    xo = o.positions()
    B = bbox(xo)
    err = np.abs(g.positions() - xo).max() / B
    qg, wg = g.ctx.rotations()
    ro = o.rotations()
    banner(f"  1-step pos rel err {err:.3e}, q err {qerr(qg, ro['q']):.3e}")
    assert err <= 1e-5
    assert qerr(qg, ro["q"]) <= 1e-5
    g.update(59)
    o.step(59)
    xo = o.positions()
    xg = g.positions()
    err = np.abs(xg - xo).max() / bbox(xo)
    banner(f"  60-step pos rel err {err:.3e}")
    assert err <= 2e-3
    # joint multipliers carry the hanging rods; compare the vertical ones
    jg = np.array([list(r.lambda_) for r in g.ctx.joint_records()])
    jo = o.joints()["lam"]
    banner(f"  joint lambda_y gpu {jg[:, 1]} oracle {jo[:, 1]}")
    assert np.allclose(jg[:, 1], jo[:, 1], rtol=2e-2, atol=1e-3)
    # sphere contacts: normal multipliers carry the spheres
    rg = g.ctx.rigid_ground_records()
    lam_g = np.array([r.lambda_n for r in rg])
    lam_o = o.rground()[:, 4]
    assert lam_g[-4:].sum() == pytest.approx(lam_o[-4:].sum(), rel=2e-2)


def test_oriented_graph_matches_eager_launches():
    banner("oriented scene: step graph replay vs eager launches (bitwise)")
    from avbd_amd.sim import Simulation
    sc = scenes.rods_and_spheres(n_rods=3, elements=12)
    a = Simulation(sc)
    b = Simulation(sc, use_graph=False)
    a.setup()
    b.setup()
    for _ in range(3):
        a.update(10)
        b.update(10)
        xa, va = a.ctx.state()
        xb, vb = b.ctx.state()
        qa, wa = a.ctx.rotations()
        qb, wb = b.ctx.rotations()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        assert np.array_equal(qa, qb) and np.array_equal(wa, wb)


def test_oriented_stats_match_oracle(oracle):
    """Residual and incremental potential of a soft rod.  (On a stiff rod the
    FP32 state itself bounds the residual: with axial stiffness EA/L the
    rounding of positions alone, ~1e-8 m, gives forces of EA/L * 1e-8 that are
    comparable to a node's weight; the soft rod keeps that floor far below
    the compared values.)"""
    banner("oriented residual / incremental potential vs oracle")
    sc = scenes.rod_cantilever(E=1e5, iters=20, dt=1.0 / 240.0)
    g = gpu_sim(sc)
    o = oracle.from_scene(sc)
    g.update(3)
    o.step(3)
    sg = g.ctx.stats()
    so = o.stats()
    banner(f"  residual gpu {sg.residual_rel:.4e} oracle {so['residual_rel']:.4e}; "
           f"ip gpu {sg.incremental_potential:.6e} oracle {so['ip']:.6e}")
    assert sg.incremental_potential == pytest.approx(so["ip"], rel=2e-2, abs=1e-9)
    assert sg.residual_rel == pytest.approx(so["residual_rel"], rel=0.1, abs=1e-3)


def test_rod_grid_with_tip_spheres(oracle):
    """The rods + rigid bodies workload at reduced size (3 x 3 rods of 16
    segments): joints at the top, spinning spheres welded at the tips, ground
    contacts on every node and sphere."""
    banner("rod grid with tip spheres vs oracle: 1 and 30 steps")
    sc = scenes.rod_grid(nx=3, nz=3, segments=16)
    g = gpu_sim(sc)
    o = oracle.from_scene(sc)
    col, k = g.ctx.coloring()
    ocol, ok = o.colors()
    assert k == ok and np.array_equal(col, ocol)
    _, wg0 = g.ctx.rotations()
    assert np.allclose(wg0, o.rotations()["w"], atol=1e-6)
    g.update(1)
    o.step(1)
    xo = o.positions()
    err = np.abs(g.positions() - xo).max() / bbox(xo)
    qg, _ = g.ctx.rotations()
    banner(f"  1-step pos rel err {err:.3e}, q err {qerr(qg, o.rotations()['q']):.3e}")
    assert err <= 1e-5 and qerr(qg, o.rotations()["q"]) <= 1e-5
    g.update(29)
    o.step(29)
    xo = o.positions()
    err = np.abs(g.positions() - xo).max() / bbox(xo)
    banner(f"  30-step pos rel err {err:.3e}")
    assert err <= 2e-3


def test_rod_grid_with_jointed_tip_spheres(oracle):
    """The tip spheres as 6-DOF bodies of their own, held to the tip nodes by
    two-sided fixed joints (new), spinning: colouring bit-exact, positions and
    orientations after one step within 1e-5 x bbox, after 30 steps within
    2e-3 x bbox, and the joints' multipliers follow the oracle's (both
    bodies' terms in the solves, the per-joint start-of-step and dual
    updates)."""
    banner("rod grid with jointed tip spheres vs oracle: 1 and 30 steps")
    sc = scenes.rod_grid(nx=3, nz=3, segments=16, tip_joint=True)
    g = gpu_sim(sc)
    o = oracle.from_scene(sc)
    col, k = g.ctx.coloring()
    ocol, ok = o.colors()
    assert k == ok and np.array_equal(col, ocol)
    _, wg0 = g.ctx.rotations()
    assert np.allclose(wg0, o.rotations()["w"], atol=1e-6)
    g.update(1)
    o.step(1)
    xo = o.positions()
    err = np.abs(g.positions() - xo).max() / bbox(xo)
    qg, _ = g.ctx.rotations()
    banner(f"  1-step pos rel err {err:.3e}, q err {qerr(qg, o.rotations()['q']):.3e}")
    assert err <= 1e-5 and qerr(qg, o.rotations()["q"]) <= 1e-5
    g.update(29)
    o.step(29)
    xo = o.positions()
    err = np.abs(g.positions() - xo).max() / bbox(xo)
    jg = np.array([list(r.lambda_) for r in g.ctx.fixed_joint_records()])
    jo = o.joints2()["lam"]
    lam_err = np.abs(jg - jo).max() / np.abs(jo).max()
    banner(f"  30-step pos rel err {err:.3e}, joint lambda rel err {lam_err:.3e}")
    assert err <= 2e-3
    assert jg.shape == jo.shape == (9, 6) and lam_err <= 2e-2
    # the joint carries the sphere: vertical multiplier ~ its weight
    assert np.all(np.abs(np.abs(jo[:, 1]) - 0.05 * 9.81) < 0.5 * 0.05 * 9.81)


def test_full_rod_grid_with_jointed_tip_spheres(oracle):
    """The rods + rigid bodies workload at its full size: 16 x 16 rods of 128
    segments with their tip spheres as bodies of their own on two-sided
    fixed joints (33,280 oriented particles, 20 iterations): colouring
    bit-exact, positions and orientations within 1e-5 x bbox after one step
    and 1e-4 after five."""
    banner("full rod grid with jointed tip spheres vs oracle")
    sc = scenes.rod_grid(tip_joint=True)
    g = gpu_sim(sc)
    o = oracle.from_scene(sc)
    col, k = g.ctx.coloring()
    ocol, ok = o.colors()
    assert k == ok and np.array_equal(col, ocol)
    for steps, tol in ((1, 1e-5), (4, 1e-4)):
        g.update(steps)
        o.step(steps)
        xo = o.positions()
        err = np.abs(g.positions() - xo).max() / bbox(xo)
        qg, _ = g.ctx.rotations()
        qe = qerr(qg, o.rotations()["q"])
        banner(f"  {steps if steps == 1 else 5} steps: pos rel err {err:.3e}, q err {qe:.3e}")
        assert err <= tol and qe <= 10 * tol


def test_box_corner_indefinite_block_matches_oracle(oracle):
    """A rigid box balanced on one corner: the 6x6 block of its vertex solve
    (inertia plus the corner's AL contact) becomes indefinite -- the
    contact's rotational Hessian at an off-centre point
    (energy/collision/RigidBodyGroundCollisionEnergySolver.hpp:111-120),
    scaled by the normal force, outweighs I/dt^2 in the twists that keep the
    corner in place.  The device solves it by LU with partial pivoting, as
    the reference's partialPivLu (simulation/VBDSolver.hpp:472), and follows
    the oracle step by step."""
    import ctypes as C
    from avbd_amd import _lib
    from avbd_amd.sim import SimulationContext
    banner("box on one corner: indefinite 6x6 block vs oracle")
    dt, m, inertia, iters = 1 / 60, 20.0, 0.02, 10
    axis = np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    ang = 0.35
    q0 = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])
    cp = np.array([0.1, -0.5, 0.1])
    w_, x_, y_, z_ = q0
    R = np.array([[1 - 2 * (y_ * y_ + z_ * z_), 2 * (x_ * y_ - w_ * z_), 2 * (x_ * z_ + w_ * y_)],
                  [2 * (x_ * y_ + w_ * z_), 1 - 2 * (x_ * x_ + z_ * z_), 2 * (y_ * z_ - w_ * x_)],
                  [2 * (x_ * z_ - w_ * y_), 2 * (y_ * z_ + w_ * x_), 1 - 2 * (x_ * x_ + y_ * y_)]])
    pos = np.array([0.0, -(R @ cp)[1] + 0.002, 0.0])
    k_start = m / (dt * dt)
    ctx = SimulationContext(dt=dt, solver_iters=iters, iter_acceleration=0.95)
    p = ctx.add_oriented_particles(pos[None], inertia=np.full((1, 3), inertia), q=q0[None], mass=np.array([m]))
    r = _lib.RigidGroundRecord()
    r.particle_index = p
    r.shape = 3   # AVBD_SHAPE_BOX: a fixed contact point in the body frame
    for k in range(3):
        r.cp_rb_local[k] = cp[k]
    r.k_start = r.k_n = r.k_t = r.k_b = k_start
    r.mu_s, r.mu_k = 0.5, 0.2
    ctx.add_energies(_lib.RIGID_BODY_GROUND_COLLISION, (_lib.RigidGroundRecord * 1)(r))
    ctx.build_graph()
    o = oracle.OracleSim(dt, 9.81, iters, 0.95)
    o.add_oriented(pos, q0, m, [inertia] * 3)
    o.add_rground_point(0, cp, k_start, 0.5, 0.2)
    o.build_graph()
    min_eig = np.inf
    # the first steps, up to touch-down: the tangential AL terms, whose
    # constraint is measured from the body's centre (cp_ground takes the
    # particle's x and z, RigidBodyGroundCollisionEnergySolver.hpp:161-170),
    # already make the block indefinite; after touch-down the restated
    # dynamics of this shape run away in the FP64 oracle as well
    for step in range(5):
        ctx.step(1)
        o.step(1)
        xg, xo = ctx.positions(), o.positions()
        qg, _ = ctx.rotations()
        qo = o.rotations()["q"]
        ex, eq = float(np.abs(xg - xo).max()), qerr(qg, qo)
        # the block at the end of the step, from the oracle's state
        rec = o.rground()[0]
        rec12 = [rec[1], rec[2], rec[3], rec[4], rec[5], rec[6], rec[7], 0.5, 0.2, rec[14], 0.0, 2]
        _, _, H = oracle.rground_single(xo[0], qo[0], rec12, rec[8:11], rec[11:14])
        H = H + np.diag([m / dt ** 2] * 3 + [inertia / dt ** 2] * 3)
        ev = float(np.linalg.eigvalsh(0.5 * (H + H.T)).min())
        min_eig = min(min_eig, ev)
        banner(f"  step {step}: |dx| {ex:.2e} m, |dq| {eq:.2e}, normal force {rec[4]:.1f} N, min eigenvalue {ev:.3g}")
        assert ex <= 2e-4 * (step + 1) and eq <= 2e-4 * (step + 1), step
    assert min_eig < 0.0    # the solves went through indefinite blocks
