"""Pins of the FP64 oracle's 6-DOF part: S3/SO3 maps, Cosserat rod, one-sided
fixed joint, rigid-body ground contact, and the oriented-particle step.

Each check is chosen so that a plausible mistake fails it:
  * maps: closed forms (scipy's rotation matrices), inverse pairs, and the
    defining first-order property of the right Jacobian;
  * energies: finite differences of the energy / constraint under the
    reference's own perturbation p += dp, q = q * Exp(dtheta) (the reference's
    test/test_cosserat_accumulate.cpp and test/test_one_sided_fixed_joint_accumulate.cpp
    use exactly these inputs and this perturbation);
  * step: Euler's rigid-body equations, exact free fall, force balance of a
    joint holding a body and of a sphere resting on the ground, the
    Euler-Bernoulli tip deflection of a cantilever rod.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle_py as O


def euler_xyz(deg):
    """Math::QuaternionFromXYZEulerAngles (common/Math.hpp:298-317): q = qz*qy*qx, as (w, x, y, z)."""
    r = Rotation.from_euler("ZYX", [deg[2], deg[1], deg[0]], degrees=True)
    x, y, z, w = r.as_quat()
    return np.array([w, x, y, z])


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def scipy_mat(q):
    return Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()


RNG = np.random.default_rng(7)


# ---------------------------------------------------------------------------
# maps
# ---------------------------------------------------------------------------
def test_qmat_matches_scipy():
    for _ in range(20):
        q = RNG.normal(size=4)
        q /= np.linalg.norm(q)
        assert np.allclose(O.qmat(q), scipy_mat(q), atol=1e-14)


@pytest.mark.parametrize("scale", [1e-8, 1e-7, 1e-3, 0.5, 2.0, 3.0])
def test_exp_log_s3_roundtrip_and_rotvec(scale):
    for _ in range(10):
        v = RNG.normal(size=3)
        v = scale * v / np.linalg.norm(v)
        q = O.exp_s3(v)
        assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-14)
        # Exp_s3 is the rotation vector map
        assert np.allclose(scipy_mat(q), Rotation.from_rotvec(v).as_matrix(), atol=1e-12)
        assert np.allclose(O.log_s3(q), v, atol=1e-12 * max(1.0, scale), rtol=1e-10)
        # Exp_so3 agrees with the quaternion route
        assert np.allclose(O.exp_so3(v), O.qmat(q), atol=1e-12)


def test_right_jacobian_first_order_property():
    """Exp(theta + d) = Exp(theta) Exp(J_r(theta) d) + O(|d|^2)."""
    for _ in range(10):
        th = RNG.normal(size=3) * 0.8
        d = RNG.normal(size=3) * 1e-6
        lhs = O.exp_so3(th + d)
        rhs = O.exp_so3(th) @ O.exp_so3(O.jr(th) @ d)
        assert np.abs(lhs - rhs).max() < 1e-11
        assert np.allclose(O.jr_inv(th) @ O.jr(th), np.eye(3), atol=1e-12)
    # small-angle branches
    th = np.array([1e-9, -2e-9, 3e-9])
    assert np.allclose(O.jr_inv(th) @ O.jr(th), np.eye(3), atol=1e-15)


# ---------------------------------------------------------------------------
# finite-difference helpers (perturbation of one oriented particle)
# ---------------------------------------------------------------------------
def perturb(x, q, d):
    return np.asarray(x) + d[:3], qmul(q, O.exp_s3(d[3:]))


def fd_grad(f, h=1e-6):
    g = np.empty(6)
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        g[i] = (f(e) - f(-e)) / (2 * h)
    return g


def fd_hess(f, h=1e-4):
    H = np.empty((6, 6))
    for i in range(6):
        for j in range(6):
            ei = np.zeros(6)
            ej = np.zeros(6)
            ei[i] = h
            ej[j] = h
            H[i, j] = (f(ei + ej) - f(ei - ej) - f(-ei + ej) + f(-ei - ej)) / (4 * h * h)
    return H


# ---------------------------------------------------------------------------
# Cosserat rod (energy/elastic/CosseratRodEnergySolver.hpp)
# ---------------------------------------------------------------------------
STIFF = np.array([2, 2, 3.5, 0.5, 0.5, 0.8])
X2 = np.array([0, 0, 0, 0.3, 0.4, 0.5])
Q2 = np.concatenate([[1, 0, 0, 0], euler_xyz([60, -30, 20])])


@pytest.mark.parametrize("local", [0, 1])
def test_cosserat_gradient_fd(local):
    """Inputs of the reference's test/test_cosserat_accumulate.cpp:14-23."""
    def energy(d):
        x = X2.copy().reshape(2, 3)
        q = Q2.copy().reshape(2, 4)
        x[local], q[local] = perturb(x[local], q[local], d)
        e, _, _ = O.cosserat_single(x, q, STIFF, 0.2, np.zeros(3), 0)
        return e

    _, G, H = O.cosserat_single(X2, Q2, STIFF, 0.2, np.zeros(3), local)
    g_fd = fd_grad(energy)
    assert np.allclose(G, g_fd, rtol=1e-6, atol=1e-7 * np.abs(g_fd).max())
    assert np.allclose(H, H.T, atol=1e-12)
    assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > -1e-12   # Gauss-Newton: PSD


@pytest.mark.parametrize("local", [0, 1])
def test_cosserat_hessian_exact_at_zero_strain(local):
    """Where the strain vanishes the Gauss-Newton Hessian is the exact one."""
    kappa = np.array([0.3, -0.2, 0.5])
    L = 0.15
    q1 = euler_xyz([10, 20, -35])
    q2 = qmul(q1, O.exp_s3(L * kappa))
    qmid = qmul(q1, O.exp_s3(0.5 * L * kappa))
    p1 = np.array([0.1, 0.2, 0.3])
    p2 = p1 + O.qmat(qmid) @ np.array([0, 0, L])
    x = np.concatenate([p1, p2])
    q = np.concatenate([q1, q2])
    e0, G, H = O.cosserat_single(x, q, STIFF, L, kappa, local)
    assert abs(e0) < 1e-20 and np.abs(G).max() < 1e-12

    def energy(d):
        xx = x.copy().reshape(2, 3)
        qq = q.copy().reshape(2, 4)
        xx[local], qq[local] = perturb(xx[local], qq[local], d)
        return O.cosserat_single(xx, qq, STIFF, L, kappa, 0)[0]

    H_fd = fd_hess(energy)
    assert np.allclose(H, H_fd, rtol=1e-4, atol=1e-5 * np.abs(H).max())


# ---------------------------------------------------------------------------
# one-sided fixed joint (energy/joint/OneSidedFixedJointEnergySolver.hpp)
# ---------------------------------------------------------------------------
JX = np.zeros(3)
JQ = np.array([1.0, 0, 0, 0])
JOFF = (np.array([0.4, 0.6, 0.2]), euler_xyz([60, -30, 20]))
JREF = (np.array([-0.3, -0.5, -0.2]), euler_xyz([-60, 50, -20]))


def test_joint_constraint_jacobian_fd():
    """With k = 0 and lambda = e_i the gradient is row i of dC/d(p, theta)."""
    x = np.array([0.05, -0.1, 0.2])
    q = euler_xyz([15, -25, 40])

    def C(d):
        xx, qq = perturb(x, q, d)
        return O.joint_single(xx, qq, *JOFF, *JREF, np.zeros(6), np.zeros(6))[0]

    for i in range(6):
        lam = np.zeros(6)
        lam[i] = 1.0
        _, G, H = O.joint_single(x, q, *JOFF, *JREF, np.zeros(6), lam)
        row = fd_grad(lambda d: C(d)[i])
        assert np.allclose(G, row, atol=1e-8), i
        assert np.abs(H).max() == 0.0


def test_joint_energy_gradient_fd():
    """Inputs of the reference's test/test_one_sided_fixed_joint_accumulate.cpp:14-22,
    with stiffness and multipliers set: E = sum 1/2 k C^2 + lambda C."""
    k = np.array([1e4, 2e4, 3e4, 5e3, 6e3, 7e3])
    lam = np.array([1.0, -2.0, 0.5, 0.2, -0.3, 0.1])

    def energy(d):
        xx, qq = perturb(JX, JQ, d)
        Cv = O.joint_single(xx, qq, *JOFF, *JREF, k, lam)[0]
        return float(np.sum(0.5 * k * Cv * Cv + lam * Cv))

    _, G, H = O.joint_single(JX, JQ, *JOFF, *JREF, k, lam)
    g_fd = fd_grad(energy)
    assert np.allclose(G, g_fd, rtol=1e-6, atol=1e-6 * np.abs(g_fd).max())
    assert np.allclose(H, H.T, atol=1e-9) and np.linalg.eigvalsh(H).min() > -1e-6


def test_joint_zero_at_reference():
    """C = 0 when the particle frame composed with the offsets hits the reference."""
    q = euler_xyz([20, 30, -10])
    off_p, off_q = JOFF
    ref_q = qmul(q, off_q)
    ref_p = np.array([1.0, 2.0, 3.0])
    x = ref_p - O.qmat(q) @ off_p
    Cv, _, _ = O.joint_single(x, q, off_p, off_q, ref_p, ref_q, np.zeros(6), np.zeros(6))
    assert np.abs(Cv).max() < 1e-14


# ---------------------------------------------------------------------------
# two-sided fixed joint (new: the reference's joint is one-sided)
# ---------------------------------------------------------------------------
J2X = np.array([[0.05, -0.1, 0.2], [0.3, 0.45, -0.15]])
J2Q = np.array([euler_xyz([15, -25, 40]), euler_xyz([-35, 10, 70])])
J2OFF = (np.array([[0.4, 0.6, 0.2], [-0.2, 0.1, 0.35]]), np.array([euler_xyz([60, -30, 20]), euler_xyz([-60, 50, -20])]))


def _perturb_side(side, d):
    x = J2X.copy()
    q = J2Q.copy()
    x[side], q[side] = perturb(x[side], q[side], d)
    return x, q


@pytest.mark.parametrize("side", [0, 1])
def test_joint2_constraint_jacobian_fd(side):
    """Rows of dC/d(x, theta) of either body by central differences under
    q [+] dtheta, and the accumulated gradient with k = 0, lambda = e_i is that
    row (a sign, a transposed offset rotation or J_r^-1(theta) in place of
    J_r^-1(-theta) on body 1 fails this)."""
    def C(d):
        x, q = _perturb_side(side, d)
        return O.joint2_single(x, q, *J2OFF, np.zeros(6), np.zeros(6), side)[0]

    Cv, grad, _, _ = O.joint2_single(J2X, J2Q, *J2OFF, np.zeros(6), np.zeros(6), side)
    assert np.abs(Cv).max() > 0.1   # far from closed: the rotation Jacobians are not identity
    for i in range(6):
        lam = np.zeros(6)
        lam[i] = 1.0
        _, _, G, H = O.joint2_single(J2X, J2Q, *J2OFF, np.zeros(6), lam, side)
        row = fd_grad(lambda d: C(d)[i])
        assert np.allclose(grad[i], row, atol=1e-8), (side, i)
        assert np.allclose(G, row, atol=1e-8), (side, i)
        assert np.abs(H).max() == 0.0


@pytest.mark.parametrize("side", [0, 1])
def test_joint2_energy_gradient_fd(side):
    """E = sum 1/2 k C^2 + lambda C: G = dE/d(x, theta) of the body, and H is
    the Gauss-Newton k grad grad^T (symmetric, positive semi-definite)."""
    k = np.array([1e4, 2e4, 3e4, 5e3, 6e3, 7e3])
    lam = np.array([1.0, -2.0, 0.5, 0.2, -0.3, 0.1])

    def energy(d):
        x, q = _perturb_side(side, d)
        Cv = O.joint2_single(x, q, *J2OFF, k, lam, side)[0]
        return float(np.sum(0.5 * k * Cv * Cv + lam * Cv))

    _, grad, G, H = O.joint2_single(J2X, J2Q, *J2OFF, k, lam, side)
    g_fd = fd_grad(energy)
    assert np.allclose(G, g_fd, rtol=1e-6, atol=1e-6 * np.abs(g_fd).max())
    assert np.allclose(H, grad.T @ np.diag(k) @ grad, rtol=1e-12, atol=1e-9)
    assert np.linalg.eigvalsh(H).min() > -1e-6


def test_joint2_zero_when_frames_meet_and_antisymmetric():
    """C = 0 when both joint frames coincide; swapping the bodies negates the
    position part and the rotation part (Log(b^-1 a) = -Log(a^-1 b))."""
    q0 = euler_xyz([20, 30, -10])
    off_p, off_q = J2OFF
    frame_q = qmul(q0, off_q[0])
    q1 = qmul(frame_q, np.array([off_q[1][0], -off_q[1][1], -off_q[1][2], -off_q[1][3]]))
    x0 = np.array([1.0, 2.0, 3.0])
    frame_p = x0 + O.qmat(q0) @ off_p[0]
    x1 = frame_p - O.qmat(q1) @ off_p[1]
    Cv = O.joint2_single([x0, x1], [q0, q1], off_p, off_q, np.zeros(6), np.zeros(6), 0)[0]
    assert np.abs(Cv).max() < 1e-14
    Ca = O.joint2_single(J2X, J2Q, *J2OFF, np.zeros(6), np.zeros(6), 0)[0]
    Cb = O.joint2_single(J2X[::-1], J2Q[::-1], J2OFF[0][::-1], J2OFF[1][::-1], np.zeros(6), np.zeros(6), 0)[0]
    assert np.allclose(Ca, -Cb, atol=1e-13)


def test_joint2_to_a_fixed_body_is_the_one_sided_joint():
    """A two-sided joint whose second body is fixed is the one-sided joint
    with the reference pose at that body's joint frame: a hanging, spinning
    body follows the same trajectory and multipliers over 120 steps."""
    dt = 1 / 60
    m, inertia = 3.0, [0.1, 0.2, 0.15]
    x0, q0 = [0.1, 1.2, -0.05], euler_xyz([10, -20, 30])
    anchor_x, anchor_q = np.array([0.3, 1.5, 0.2]), euler_xyz([-40, 25, 60])
    off_b, rot_b = np.array([-0.1, 0.2, 0.05]), euler_xyz([30, 10, -50])
    off_a, rot_a = np.array([0.05, 0.1, -0.02]), euler_xyz([5, 15, 25])
    ref_p = anchor_x + O.qmat(anchor_q) @ off_b
    ref_q = qmul(anchor_q, rot_b)

    one = O.OracleSim(dt, 9.81, 10, 0.95)
    p = one.add_oriented(x0, q0, m, inertia)
    one.add_joint(p, off_a, rot_a, ref_p, ref_q)
    one.build_graph()
    two = O.OracleSim(dt, 9.81, 10, 0.95)
    p2 = two.add_oriented(x0, q0, m, inertia)
    a = two.add_oriented(anchor_x, anchor_q, 1.0, [0.1, 0.1, 0.1])
    two.set_fixed(np.array([0, 1], dtype=np.uint8))
    two.add_joint2(p2, a, [off_a, off_b], [rot_a, rot_b])
    two.build_graph()
    w0 = np.array([[0.5, -1.0, 2.0]])
    one.set_rotation_state(np.array([q0]), w0)
    two.set_rotation_state(np.array([q0, anchor_q]), np.vstack([w0, np.zeros((1, 3))]))
    for _ in range(12):
        one.step(10)
        two.step(10)
        assert np.allclose(two.positions()[0], one.positions()[0], rtol=0, atol=1e-12)
        assert np.allclose(two.rotations()["q"][0], one.rotations()["q"][0], rtol=0, atol=1e-12)
    assert np.allclose(two.positions()[1], anchor_x, rtol=0, atol=0)
    j1, j2 = one.joints(), two.joints2()
    assert np.allclose(j2["lam"], j1["lam"], rtol=1e-9, atol=1e-9)
    assert j1["lam"][0, 1] == pytest.approx(-m * 9.81, rel=5e-2)


def test_joint2_welded_pair_falls_as_one_body():
    """Two free bodies welded by a two-sided joint and set spinning together:
    the joint stays closed (the second body's gradient enters its own solve
    with the right sign) and the pair's centre of mass follows the backward
    Euler free fall (the joint forces are internal)."""
    dt = 1 / 120
    s = O.OracleSim(dt, 9.81, 20, 0.0)
    ma, mb = 2.0, 0.5
    xa, xb = np.array([0.0, 2.0, 0.0]), np.array([0.0, 1.8, 0.0])
    qa, qb = euler_xyz([0, 0, 0]), euler_xyz([20, -10, 5])
    a = s.add_oriented(xa, qa, ma, [0.02, 0.02, 0.02])
    b = s.add_oriented(xb, qb, mb, [0.005, 0.005, 0.005])
    # the joint frame at the midpoint, with body b's offsets taken from the start pose
    mid = 0.5 * (xa + xb)
    off_a = O.qmat(qa).T @ (mid - xa)
    off_b = O.qmat(qb).T @ (mid - xb)
    rot_b = qmul(np.array([qb[0], -qb[1], -qb[2], -qb[3]]), qa)   # q_b rot_b = q_a
    s.add_joint2(a, b, [off_a, off_b], [[1.0, 0, 0, 0], rot_b])
    s.build_graph()
    s.set_rotation_state(np.array([qa, qb]), np.array([[0.0, 0.0, 0.3], [0.0, 0.0, 0.3]]))
    com0 = (ma * xa + mb * xb) / (ma + mb)
    steps = 60
    s.step(steps)
    x = s.positions()
    q = s.rotations()["q"]
    com = (ma * x[0] + mb * x[1]) / (ma + mb)
    # backward Euler free fall: y_n = y_0 - g dt^2 n (n + 1) / 2
    fall = com0 + np.array([0.0, -9.81 * dt * dt * steps * (steps + 1) / 2, 0.0])
    assert np.linalg.norm(com - fall) < 1e-4
    Cv = O.joint2_single(x, q, [off_a, off_b], [[1.0, 0, 0, 0], rot_b], np.zeros(6), np.zeros(6), 0)[0]
    assert np.abs(Cv[:3]).max() < 1e-3 and np.abs(Cv[3:]).max() < 1e-2
    assert np.linalg.norm(x[0] - x[1]) == pytest.approx(0.2, abs=1e-3)


# ---------------------------------------------------------------------------
# rigid-body ground contact (energy/collision/RigidBodyGroundCollisionEnergySolver.hpp)
# ---------------------------------------------------------------------------
def rg_rec(kn=0.0, kt=0.0, kb=0.0, ln=0.0, lt=0.0, lb=0.0, cprev=0.0, mu_s=10.0, mu_k=10.0, static=0,
           radius=0.1, shape=2):
    return np.array([kn, kt, kb, ln, lt, lb, cprev, mu_s, mu_k, static, radius, shape], dtype=np.float64)


@pytest.mark.parametrize("shape", [0, 2])
def test_rground_constraint_jacobian_fd(shape):
    """Rows of dC/d(p, theta) for normal and both tangents (k = 0, one unit
    multiplier at a time, friction cone wide open)."""
    x = np.array([0.2, 0.05, -0.1])
    q = euler_xyz([30, -40, 10])
    cpl = np.array([0.03, -0.08, 0.02])
    cpg = np.array([0.25, 0.0, -0.12])

    def C(d):
        xx, qq = perturb(x, q, d)
        return O.rground_single(xx, qq, rg_rec(shape=shape), cpl, cpg)[0]

    for i in range(3):
        lams = [0.0, 0.0, 0.0]
        lams[i] = 1.0
        if i > 0:
            lams[0] = 1.0   # a normal multiplier is needed to open the friction cone
        _, G, _ = O.rground_single(x, q, rg_rec(ln=lams[0], lt=lams[1], lb=lams[2], shape=shape), cpl, cpg)
        if i > 0:
            _, Gn, _ = O.rground_single(x, q, rg_rec(ln=1.0, shape=shape), cpl, cpg)
            G = G - Gn
        row = fd_grad(lambda d: C(d)[i])
        assert np.allclose(G, row, atol=1e-8), (shape, i)


def test_rground_rotational_hessian_fd():
    """The rotational block of lambda_t * Hess(C_t) equals the second
    derivative of lambda_t C_t(q * Exp(theta)) at theta = 0."""
    x = np.array([0.2, 0.05, -0.1])
    q = euler_xyz([30, -40, 10])
    cpl = np.array([0.03, -0.08, 0.02])
    cpg = np.array([0.25, 0.0, -0.12])
    for i, lam in [(1, (1.0, 0.7, 0.0)), (2, (1.0, 0.0, -0.4))]:
        rec = rg_rec(ln=lam[0], lt=lam[1], lb=lam[2])
        _, _, H = O.rground_single(x, q, rec, cpl, cpg)
        _, _, Hn = O.rground_single(x, q, rg_rec(ln=1.0), cpl, cpg)

        def f(d):
            xx, qq = perturb(x, q, d)
            return lam[i] * O.rground_single(xx, qq, rec, cpl, cpg)[0][i]

        Hfd = fd_hess(f)
        assert np.allclose((H - Hn)[3:, 3:], Hfd[3:, 3:], atol=1e-6)


def test_rground_sphere_normal_ignores_rotation():
    """Spheres and rods: C_n uses the lowest point p - r*y, independent of q."""
    x = np.array([0.0, 0.05, 0.0])
    for q in (np.array([1.0, 0, 0, 0]), euler_xyz([70, 10, -30])):
        Cv, _, _ = O.rground_single(x, q, rg_rec(shape=0, radius=0.1), np.zeros(3), np.zeros(3))
        assert Cv[0] == pytest.approx(0.1 - 0.05)


# ---------------------------------------------------------------------------
# the oriented step
# ---------------------------------------------------------------------------
def test_free_rigid_body_follows_euler_equations():
    """One free oriented particle, no energies: the solve lands on q_in, so
    w_new = w - dt I^-1 (w x I w) -- explicit Euler on Euler's equations
    I dw/dt + w x (I w) = 0 (body frame) -- and q_new = q Exp(dt w_new)."""
    dt = 1e-3
    inertia = np.array([1.0, 2.0, 3.5])
    q0 = euler_xyz([10, 20, 30])
    s = O.OracleSim(dt, 9.81, 5, 0.95)
    s.add_oriented([0, 1, 0], q0, 2.0, inertia)
    s.build_graph()
    # from rest nothing rotates and the body falls freely
    s.step(3)
    r = s.rotations()
    assert np.allclose(r["q"][0], q0, atol=1e-15) and np.allclose(r["w"][0], 0)
    assert s.positions()[0, 1] == pytest.approx(1 - 9.81 * dt * dt * 6, abs=1e-12)
    w = np.array([0.3, -1.2, 2.0])
    s.set_rotation_state(np.array([q0]), np.array([w]))
    s.step(1)
    r = s.rotations()
    w_expect = w - dt * np.cross(w, inertia * w) / inertia
    assert np.allclose(r["w"][0], w_expect, rtol=1e-9, atol=1e-12)
    assert np.allclose(r["q"][0], qmul(q0, O.exp_s3(dt * w_expect)), atol=1e-12)
    # a spin about a principal axis is steady
    s.set_rotation_state(np.array([q0]), np.array([[0.0, 0.0, 1.5]]))
    s.step(10)
    assert np.allclose(s.rotations()["w"][0], [0, 0, 1.5], atol=1e-12)


def test_joint_holds_body_force_balance():
    """A body hung from a one-sided fixed joint settles with the joint closed
    to within a millimetre and the translational multiplier carrying its
    weight (lambda_y ~ -m g).  The remaining offset is the fixed point of the
    per-step multiplier decay (lambda *= alpha*gamma) against the alpha-corrected
    constraint C - 0.95 C_prev, the same in the reference's update rules."""
    dt = 1 / 60
    s = O.OracleSim(dt, 9.81, 10, 0.95)
    m = 3.0
    p = s.add_oriented([0, 1, 0], [1, 0, 0, 0], m, [0.1, 0.1, 0.1])
    s.add_joint(p, [0, 0, 0], [1, 0, 0, 0], [0, 1, 0], [1, 0, 0, 0])
    s.build_graph()
    s.step(400)
    y0 = s.positions()[0, 1]
    s.step(50)
    y1 = s.positions()[0, 1]
    j = s.joints()
    assert abs(y1 - y0) < 1e-9                     # steady
    assert abs(y1 - 1.0) < 1e-3
    assert j["lam"][0, 1] == pytest.approx(-m * 9.81, rel=3e-2)
    assert np.abs(j["lam"][0, [0, 2, 3, 4, 5]]).max() < 1e-9
    assert np.abs(s.rotations()["w"][0]).max() < 1e-9


def test_sphere_rests_on_ground():
    """A sphere placed on the ground stays there; the normal multiplier
    converges to its weight, up to the fixed point of the per-step multiplier
    decay against the alpha-corrected constraint (as for the joint above)."""
    dt = 1 / 60
    s = O.OracleSim(dt, 9.81, 10, 0.95)
    r, rho = 0.2, 500.0
    s.add_sphere(r, rho, [0.0, r, 0.0], [1, 0, 0, 0], fixed=False, with_ground=True)
    s.build_graph()
    s.step(300)
    m, _ = s.mass_fixed()
    assert m[0] == pytest.approx(4 / 3 * math.pi * r ** 3 * rho)
    assert s.rotations()["inertia"][0] == pytest.approx(np.full(3, 0.4 * m[0] * r * r))
    rec = s.rground()
    y = s.positions()[0, 1]
    assert abs(y - r) < 2e-3
    assert rec[0, 4] == pytest.approx(m[0] * 9.81, rel=3e-2)
    assert np.abs(s.rotations()["w"][0]).max() < 1e-6


def test_sphere_free_fall_exact():
    dt = 1 / 60
    s = O.OracleSim(dt, 9.81, 4, 0.95)
    s.add_sphere(0.1, 1000.0, [0, 5, 0], euler_xyz([5, 6, 7]), with_ground=False)
    s.build_graph()
    for n in range(1, 5):
        s.step(1)
        assert s.positions()[0, 1] == pytest.approx(5 - 9.81 * dt * dt * n * (n + 1) / 2, abs=1e-12)


def test_rod_setup_masses_and_adjacency():
    """Rod ctor (simobject/Rod.cpp): node masses/inertias lumped half per
    element; a chain's adjacency is its path graph and DSATUR gives 2 colours."""
    s = O.OracleSim(1 / 60, 9.81, 10, 0.95)
    E, nu, rho, d, L, ne = 1e7, 0.3, 1000.0, 0.02, 1.0, 10
    first = s.add_rod(ne, L, E=E, nu=nu, density=rho, diameter=d, base_fixed=True)
    s.build_graph()
    m, fixed = s.mass_fixed()
    A = math.pi * (d / 2) ** 2
    me = A * L / ne * rho
    expect = np.full(ne + 1, me)
    expect[0] = expect[-1] = 0.5 * me
    assert np.allclose(m, expect, rtol=1e-14)
    assert fixed[first] == 1 and fixed[1:].sum() == 0
    assert m.sum() == pytest.approx(A * L * rho)
    eo, ent, po, pn = s.adjacency()
    for i in range(ne + 1):
        nb = sorted(pn[po[i]:po[i + 1]])
        want = [j for j in (i - 1, i + 1) if 0 <= j <= ne]
        assert nb == want
    col, k = s.colors()
    assert k == 2 and all(col[i] != col[i + 1] for i in range(ne))
    # initial positions along the body z axis, straight
    x = s.positions()
    assert np.allclose(x[:, 2], np.linspace(0, L, ne + 1)) and np.allclose(x[:, :2], 0)


def test_cantilever_rod_tip_deflection():
    """A horizontal cantilever (8 elements, base node fixed in position and
    orientation) under gravity settles to the Euler-Bernoulli tip deflection
    w L^4 / (8 E I), w = rho A g, within the discretisation error of lumped
    node masses (observed 0.7 %; 3 % allowed).  A stiff rod needs a small
    step for the Gauss-Seidel sweeps to converge, as in the reference's own
    rod scene (1 ms, 15 iterations)."""
    dt = 1 / 600
    s = O.OracleSim(dt, 9.81, 100, 0.95)
    E, nu, rho, d, L, ne = 1e8, 0.3, 1000.0, 0.02, 0.2, 8
    # rod body z axis along world x: rotate +90 deg about y
    s.add_rod(ne, L, E=E, nu=nu, density=rho, diameter=d, rot=euler_xyz([0, 90, 0]), base_fixed=True)
    s.build_graph()
    s.step(450)
    t0 = s.positions()[-1, 1]
    s.step(50)
    tip = s.positions()[-1, 1]
    A = math.pi * (d / 2) ** 2
    I = math.pi * (d / 2) ** 4 / 4
    delta = rho * A * 9.81 * L ** 4 / (8 * E * I)
    assert abs(tip - t0) < 1e-3 * delta
    assert -tip == pytest.approx(delta, rel=0.03)
    x = s.positions()
    assert np.abs(x[:, 2]).max() < 1e-12          # stays in its plane
