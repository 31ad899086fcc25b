"""Pins of the oracle's dynamic-contact pipeline (narrow phase, contact
lifecycle, dynamic adjacency, incremental recolouring, triangle-rigid /
triangle-rod AL contacts) against closed forms, brute force, finite
differences and invariants.  CPU only."""
import numpy as np
import pytest

from oracle import oracle_py as O
from avbd_amd import scenes

rng = np.random.default_rng(11)


def test_contact_key_bit_layout():
    # collision/DetectedCollision.hpp:43-66: 5 type | 26 idxA | 26 idxB | 7 code, (A, B) canonical
    for _ in range(50):
        t = int(rng.integers(0, 32))
        a, b = (int(v) for v in rng.integers(0, 1 << 26, size=2))
        code = int(rng.integers(0, 128))
        lo, hi = min(a, b), max(a, b)
        expect = (t << 59) | (lo << 33) | (hi << 7) | code
        assert O.contact_key(t, a, b, code) == expect
        assert O.contact_key(t, b, a, code) == expect


def test_slot_pool_is_a_lifo_stack():
    # TombstonePool: slots 0, 1, 2, ... on a stack; freed slots are reused last-in first-out
    assert O.slot_sequence([-1, -1, -1, 1, 0, -1, -1, -1]) == [0, 1, 2, -1, -1, 0, 1, 3]
    assert O.slot_sequence([-1, -1, 0, -1, -1]) == [0, 1, -1, 0, 2]
    with pytest.raises(O.OracleError):
        O.slot_sequence([-1, 0, 0])          # freeing an inactive slot
    with pytest.raises(O.OracleError):
        O.slot_sequence([-1, -1, -1], capacity=2)


def test_golden_section_known_answers():
    for a, b, c in [(0.0, 1.0, 0.3), (0.0, 1.0 / 60.0, 0.004), (-2.0, 3.0, 1.7)]:
        t = O.golden_section_quadratic(a, b, c)
        assert abs(t - c) <= 1e-3
    # minimum outside the interval: the nearer end point
    assert O.golden_section_quadratic(0.0, 1.0, 2.0) == 1.0
    assert O.golden_section_quadratic(0.0, 1.0, -1.0) == 0.0
    # an interval narrower than the tolerance is not refined: the midpoint or an end
    t = O.golden_section_quadratic(0.0, 5e-4, 2.6e-4)
    assert t in (0.0, 2.5e-4, 5e-4)


def _tri_samples(a, b, c, n=120):
    u, v = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = (u + v) <= 1
    u, v = u[m], v[m]
    return a[None] * (1 - u - v)[:, None] + b[None] * u[:, None] + c[None] * v[:, None]


def test_closest_point_triangle_brute_force():
    for _ in range(40):
        a, b, c, p = rng.normal(size=(4, 3))
        q = O.closest_point_triangle(p, a, b, c)
        s = _tri_samples(a, b, c)
        d_bf = np.min(np.linalg.norm(s - p, axis=1))
        assert np.linalg.norm(q - p) <= d_bf + 1e-9
        # q lies on the triangle
        bc = np.linalg.lstsq(np.stack([b - a, c - a], 1), q - a, rcond=None)[0]
        assert bc.min() >= -1e-9 and bc.sum() <= 1 + 1e-9


def test_closest_segment_triangle():
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 0, 1])
    # crossing the face: distance 0 and the crossing point
    s, bc, d = O.closest_segment_triangle([0.2, -1, 0.3], [0.2, 1, 0.3], a, b, c)
    assert d == 0.0 and abs(s - 0.5) < 1e-12
    assert np.allclose(bc[0] * a + bc[1] * b + bc[2] * c, [0.2, 0, 0.3])
    # a segment hovering above the face interior: the lower end point is closest
    s, bc, d = O.closest_segment_triangle([0.2, 0.05, 0.2], [0.25, 0.5, 0.2], a, b, c)
    assert s == 0.0 and abs(d - 0.05) < 1e-12 and np.allclose(bc, [0.6, 0.2, 0.2])
    # any configuration: the reported distance is a real point-pair distance, so
    # never below the brute-force minimum, and equal to it when an end point
    # over the face interior is the closest feature
    for _ in range(30):
        p0, p1 = rng.normal(size=(2, 3))
        s, bc, d = O.closest_segment_triangle(p0, p1, a, b, c)
        seg = p0[None] + np.linspace(0, 1, 400)[:, None] * (p1 - p0)[None]
        tri = _tri_samples(a, b, c, 60)
        d_bf = np.min(np.linalg.norm(seg[:, None] - tri[None], axis=2))
        assert d >= d_bf - 2e-2 * max(1.0, d_bf) - 1e-6 or d == 0.0
        assert 0.0 <= s <= 1.0


def test_complete_basis_follows_eigen_unit_orthogonal():
    # Eigen's rule: (-y, x, 0)/|(x, y)| unless x, y are negligible against z
    t, b = O.complete_basis([0, 0, 1.0])
    assert np.allclose(t, [0, -1, 0]) and np.allclose(b, np.cross([0, 0, 1.0], t))
    t, b = O.complete_basis([1.0, 0, 0])
    assert np.allclose(t, [0, 1, 0])
    for _ in range(30):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t, b = O.complete_basis(n)
        assert abs(t @ n) < 1e-12 and abs(np.linalg.norm(t) - 1) < 1e-12 and t[2] == 0.0
        assert np.allclose(b, np.cross(n, t))


def test_update_collision_frame_preserves_friction_force():
    n = np.array([0.0, 1.0, 0.0])
    t, b = O.complete_basis(n)
    lt, lb = 0.7, -0.3
    f = lt * t + lb * b
    # rotate the normal about the force direction: the force stays tangent
    axis = f / np.linalg.norm(f)
    ang = 0.4
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
    nn = R @ n
    n2, t2, b2, lt2, lb2 = O.update_collision_frame(n, t, b, lt, lb, nn)
    assert np.allclose(n2, nn)
    F = np.stack([n2, t2, b2])
    assert np.allclose(F @ F.T, np.eye(3), atol=1e-12)
    assert np.allclose(lt2 * t2 + lb2 * b2, f, atol=1e-12)
    # the new tangent is the old one projected onto the new plane
    tp = t - (t @ nn) * nn
    assert np.allclose(t2, tp / np.linalg.norm(tp))


def test_tri_sphere_ccd_closed_form():
    """A sphere falling straight onto a static horizontal triangle: contact iff
    the gap closes to within the margin during the step; normal = -y (from
    the centre towards the triangle); contact point on the sphere = (0, -r, 0)
    in the body frame; barycentrics = the centre's projection."""
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 0, 1.0]])
    tv = np.zeros((3, 3))
    r, dt, margin = 0.1, 1.0 / 60.0, 1e-3
    centre = np.array([0.25, 0.0, 0.3])
    for gap, speed, hit in [(0.0005, 0.0, True), (0.002, 0.0, False), (0.05, 3.6, True), (0.08, 3.6, False)]:
        pos = centre + [0, r + gap, 0]
        got, n, bc, cpl = O.tri_sphere_ccd(tri, tv, pos, [1, 0, 0, 0], [0, -speed, 0], [0, 0, 0], r, dt, margin)
        assert got == hit, (gap, speed)
        if hit:
            # the spatial search stops at the golden-section tolerance (1e-3 m),
            # so the contact point sits within ~1e-3 of the projection
            assert np.allclose(n, [0, -1, 0], atol=1e-2) and abs(np.linalg.norm(n) - 1) < 1e-12
            assert np.allclose(cpl, r * n, atol=1e-12)
            p = bc[0] * tri[0] + bc[1] * tri[1] + bc[2] * tri[2]
            assert np.linalg.norm(p - centre) < 3e-3
    # spin about the contact normal does not move the contact
    got, n, bc, cpl = O.tri_sphere_ccd(tri, tv, centre + [0, r + 0.0005, 0], [1, 0, 0, 0], [0, 0, 0], [0, 5.0, 0],
                                       r, dt, margin)
    assert got and np.allclose(n, [0, -1, 0], atol=1e-2)


def _rand_quat():
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _qmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _plus(q, d):
    e = O.exp_s3(d)
    w1, x1, y1, z1 = q
    w2, x2, y2, z2 = e
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _rec(n, t, b, barys, s, cpl, kn=50.0, kt=40.0, kb=30.0, ln=2.0, lt=0.1, lb=-0.2, cprev=0.0, mus=0.4, muk=0.2,
         use_static=0):
    return np.concatenate([n, t, b, barys, [s], cpl, [kn, kt, kb, ln, lt, lb, cprev, mus, muk, use_static]])


@pytest.mark.parametrize("kind", [0, 1])
def test_dynamic_contact_gradients_by_finite_differences(kind):
    """C_n gradients of both contact kinds against central differences:
    positions for the triangle vertices and the body / rod nodes, the S3
    perturbation q (+) dtheta for the orientations (test/finite_difference.hpp:35-49)."""
    x = rng.normal(size=(5, 3)) * 0.1
    q = np.zeros((5, 4))
    q[:, 0] = 1
    q[3], q[4] = _rand_quat(), _rand_quat()
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t, b = O.complete_basis(n)
    bc = np.array([0.2, 0.3, 0.5])
    rec = _rec(n, t, b, bc, 0.35, rng.normal(size=3) * 0.02)
    h = 1e-6

    def Cn(xx, qq):
        return O.dyn_single(kind, xx, qq, rec, 0, 3)[0][0]

    locals_ = [0, 1, 2, 3] if kind == 0 else [0, 1, 2, 3, 4]
    for loc in locals_:
        dof = 3 if loc < 3 else 6
        _, g, hblk, _, _ = O.dyn_single(kind, x, q, rec, loc, dof)
        fd = np.zeros(dof)
        for i in range(3):
            xp, xm = x.copy(), x.copy()
            xp[loc, i] += h
            xm[loc, i] -= h
            fd[i] = (Cn(xp, q) - Cn(xm, q)) / (2 * h)
        if dof == 6:
            for i in range(3):
                d = np.zeros(3)
                d[i] = h
                qp, qm = q.copy(), q.copy()
                qp[loc] = _plus(q[loc], d)
                qm[loc] = _plus(q[loc], -d)
                fd[3 + i] = (Cn(x, qp) - Cn(x, qm)) / (2 * h)
        if kind == 1 and dof == 6:
            # TriangleRodCollisionEnergySolver.hpp:97,125 negates R [cp]x twice
            # (R_rloc = -R [cp]x, then -n^T R_rloc dRs/dR): the rotational part
            # is the negated derivative of C_n.  Restated as the reference has it.
            fd[3:] = -fd[3:]
        assert np.allclose(g[0, :dof], fd, atol=1e-6), (kind, loc, g[0, :dof], fd)
        if dof == 6:
            # tangent gradients of the oriented side are zeroed in the reference
            assert np.all(g[1] == 0) and np.all(g[2] == 0)
            if kind == 0:
                # rotational Hessian block of C_n = symmetrised second derivative
                H = np.zeros((3, 3))
                for i in range(3):
                    for j in range(3):
                        di = np.zeros(3)
                        dj = np.zeros(3)
                        di[i] = 1e-4
                        dj[j] = 1e-4
                        vals = []
                        for si, sj in [(1, 1), (1, -1), (-1, 1), (-1, -1)]:
                            qq = q.copy()
                            qq[3] = _plus(q[3], si * di + sj * dj)
                            vals.append(Cn(x, qq))
                        H[i, j] = (vals[0] - vals[1] - vals[2] + vals[3]) / (4e-8)
                assert np.allclose(hblk[0], 0.5 * (H + H.T), atol=2e-5)
            else:
                assert np.all(hblk == 0)


def test_dynamic_contact_accumulate_al_terms():
    """Positional incidence: G = lambda+ * grad, H = k grad grad^T per direction,
    with the friction multipliers clamped to the cone mu * lambda_n+."""
    x = rng.normal(size=(5, 3)) * 0.1
    q = np.zeros((5, 4))
    q[:, 0] = 1
    n = np.array([0.0, 1.0, 0.0])
    t, b = O.complete_basis(n)
    rec = _rec(n, t, b, np.array([0.2, 0.3, 0.5]), 0.0, np.array([0.0, -0.05, 0.0]), lt=5.0, lb=5.0)
    Cv, g, _, G, H = O.dyn_single(0, x, q, rec, 1, 3)
    kn, kt, kb, ln, lt, lb, cprev, mus, muk = rec[16:25]
    ln_p = max(0.0, kn * (Cv[0] - 0.95 * cprev) + ln)
    ltp, lbp = kt * Cv[1] + lt, kb * Cv[2] + lb
    mag = np.hypot(ltp, lbp)
    if mag > muk * ln_p:
        ltp, lbp = ltp / (mag + 1e-8) * muk * ln_p, lbp / (mag + 1e-8) * muk * ln_p
    assert np.allclose(G, ln_p * g[0, :3] + ltp * g[1, :3] + lbp * g[2, :3])
    Hx = kn * np.outer(g[0, :3], g[0, :3]) + kt * np.outer(g[1, :3], g[1, :3]) + kb * np.outer(g[2, :3], g[2, :3])
    assert np.allclose(H, Hx)
    # -bary * frame vectors for a triangle vertex
    assert np.allclose(g[0, :3], -0.3 * n) and np.allclose(g[1, :3], -0.3 * t)


def _union_graph(o, sc):
    adj = o.adjacency()
    n = o.n
    nb = [set() for _ in range(n)]
    _, _, off, nbrs = adj
    for p in range(n):
        for k in range(off[p], off[p + 1]):
            nb[p].add(int(nbrs[k]))
    for kind, npe in ((0, 4), (1, 5), (2, 2), (3, 4)):
        for r in o.contacts(kind):
            idx = [int(v) for v in r[1:1 + npe]]
            for i in idx:
                for j in idx:
                    if i != j:
                        nb[i].add(j)
    return nb


@pytest.mark.parametrize("name", ["ball_on_block", "rod_on_block", "block_stack"])
def test_recolouring_invariants_and_lifecycle(name):
    """After every step's detection: the colouring is proper on the static +
    dynamic graph except between conflicted particles, and every same-colour
    adjacent pair is conflicted; contacts persist with their slots while
    detected; the detected list is key-sorted."""
    sc = getattr(scenes, name)(steps=30)
    o = O.from_scene(sc)
    seen_contacts = 0
    prev = {}
    for step in range(25):
        o.step(1)
        keys, slots, types = o.detected()
        assert np.all(keys[1:] > keys[:-1])
        cur = {int(k): int(s) for k, s in zip(keys, slots)}
        for k, s in cur.items():
            if k in prev:
                assert prev[k] == s   # an updated contact keeps its slot
        prev = cur
        seen_contacts += len(cur)
        col, _ = o.colors()
        conf = o.conflicted()
        nb = _union_graph(o, sc)
        for p in range(o.n):
            for q in nb[p]:
                if col[p] == col[q]:
                    assert conf[p] and conf[q], (step, p, q)
        # contact records: slots unique, frames orthonormal, k >= k_start
        for kind in (0, 1, 2, 3):
            rec = o.contacts(kind)
            if len(rec):
                assert len(set(rec[:, 0].astype(int))) == len(rec)
                F = rec[:, 6:15].reshape(-1, 3, 3)
                assert np.allclose(np.einsum("nij,nkj->nik", F, F), np.eye(3)[None], atol=1e-9)
                assert np.all(rec[:, 23] >= rec[:, 22] * (1 - 1e-12))
    assert seen_contacts > 0


def test_resting_ball_multipliers_carry_its_weight():
    """Sphere resting on the block: the vertical force of the triangle-rigid
    contacts balances m g (AL fixed point with the per-step multiplier decay)."""
    sc = scenes.ball_on_block(steps=200, speed=0.0, drop=0.0005)
    o = O.from_scene(sc)
    o.step(150)
    m, _ = o.mass_fixed()
    p_ball = sc.meshes[0].rest.shape[0]
    w = m[p_ball] * sc.g
    rec = o.contacts(0)
    assert len(rec) > 0
    fy = 0.0
    for r in rec:
        n, t, b = r[6:9], r[9:12], r[12:15]
        fy += -(r[26] * n[1] + r[27] * t[1] + r[28] * b[1])
    # multipliers after the step-start decay alpha*gamma of the next step
    assert abs(fy - w) < 0.1 * w, (fy, w)


def test_unsupported_pairs_raise():
    """Rod-segment-sphere candidate pairs throw as in the reference
    (collision/CollisionDetector.cpp:210-213); sphere-sphere pairs, which the
    reference also throws on (:686-687), are detected here (rigid-rigid
    contacts, new)."""
    sc = scenes.Scene("two_balls", spheres=[scenes.Sphere(0.05, 1000.0, (0, 0.5, 0), ground=False),
                                            scenes.Sphere(0.05, 1000.0, (0.09, 0.5, 0), ground=False)],
                      contact_margin=1e-3)
    o = O.from_scene(sc)
    o.detect_and_recolor()
    keys, _, types = o.detected()
    assert list(types) == [6] and int(keys[0]) == O.contact_key(6, 0, 1)
    rod = scenes.Rod(elements=2, length=0.2, E=1e7, nu=0.3, density=1000.0, diameter=0.02,
                     pos=(-0.1, 0.5, 0.0), rot=scenes.quat_about((0, 1, 0), np.pi / 2), ground=False)
    sc2 = scenes.Scene("rod_ball", rods=[rod], spheres=[scenes.Sphere(0.05, 1000.0, (0.0, 0.5, 0.055), ground=False)],
                       contact_margin=1e-3)
    o2 = O.from_scene(sc2)
    with pytest.raises(O.OracleError):
        o2.detect_and_recolor()


def test_rigid_rigid_gradients_by_finite_differences():
    """Rigid-rigid contact (new): C = dir . (x_2 + R_2 cpl_2 - x_1 - R_1 cpl_1)
    for dir = n, t, b; the gradients of both bodies (position and the S3
    perturbation q (+) dtheta) and the rotational Hessian blocks against
    central differences.  Both bodies carry all three directions."""
    x = rng.normal(size=(5, 3)) * 0.1
    q = np.zeros((5, 4))
    q[:, 0] = 1
    q[3], q[4] = _rand_quat(), _rand_quat()
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t, b = O.complete_basis(n)
    cpl1, cpl2 = rng.normal(size=3) * 0.03, rng.normal(size=3) * 0.03
    rec = _rec(n, t, b, cpl2, 0.0, cpl1)
    h = 1e-6

    def C(xx, qq):
        return O.dyn_single(2, xx, qq, rec, 0, 6)[0]

    c0 = C(x, q)
    R1, R2 = _qmat(q[3]), _qmat(q[4])
    d = x[4] + R2 @ cpl2 - x[3] - R1 @ cpl1
    assert np.allclose(c0, [n @ d, t @ d, b @ d], atol=1e-14)
    for loc in (0, 1):
        p = 3 + loc
        _, g, hblk, _, _ = O.dyn_single(2, x, q, rec, loc, 6)
        fd = np.zeros((3, 6))
        for i in range(3):
            xp, xm = x.copy(), x.copy()
            xp[p, i] += h
            xm[p, i] -= h
            fd[:, i] = (C(xp, q) - C(xm, q)) / (2 * h)
            dv = np.zeros(3)
            dv[i] = h
            qp, qm = q.copy(), q.copy()
            qp[p] = _plus(q[p], dv)
            qm[p] = _plus(q[p], -dv)
            fd[:, 3 + i] = (C(x, qp) - C(x, qm)) / (2 * h)
        assert np.allclose(g, fd, atol=1e-6), (loc, g, fd)
        for k in range(3):
            H = np.zeros((3, 3))
            for i in range(3):
                for j in range(3):
                    di, dj = np.zeros(3), np.zeros(3)
                    di[i] = 1e-4
                    dj[j] = 1e-4
                    vals = []
                    for si, sj in [(1, 1), (1, -1), (-1, 1), (-1, -1)]:
                        qq = q.copy()
                        qq[p] = _plus(q[p], si * di + sj * dj)
                        vals.append(C(x, qq)[k])
                    H[i, j] = (vals[0] - vals[1] - vals[2] + vals[3]) / (4e-8)
            assert np.allclose(hblk[k], 0.5 * (H + H.T), atol=2e-5), (loc, k)


def test_head_on_spheres_conserve_momentum_and_stop_at_contact():
    """Two equal free spheres approaching head-on without gravity: the contact
    forces are equal and opposite, so the total linear momentum stays zero and
    the motion stays symmetric about the midpoint; the spheres stop at
    contact (the AL normal force removes the approach speed, no sliding
    friction acts on a head-on contact), and never overlap by more than a
    fraction of a millimetre."""
    r, rho, v0 = 0.05, 1000.0, 0.3
    sc = scenes.Scene("head_on", spheres=[scenes.Sphere(r, rho, (-0.06, 0.5, 0), ground=False),
                                          scenes.Sphere(r, rho, (0.06, 0.5, 0), ground=False)],
                      contact_margin=1e-3, g=0.0, dt=1.0 / 120.0, iters=100)
    o = O.from_scene(sc)
    v = np.zeros((2, 3))
    v[0, 0], v[1, 0] = v0, -v0
    o.set_velocities(v, v)
    m, _ = o.mass_fixed()
    min_gap, seen = 1.0, 0
    for _ in range(40):
        o.step(1)
        x = o.positions()
        vv = o.velocities()
        p = m[:, None] * vv
        assert np.abs(p.sum(axis=0)).max() < 1e-6 * m[0] * v0
        assert np.abs(x[0, 0] + x[1, 0]) < 1e-6
        min_gap = min(min_gap, np.linalg.norm(x[1] - x[0]) - 2 * r)
        seen = max(seen, len(o.contacts(2)))
    assert seen == 1
    assert min_gap > -5e-4
    assert abs(o.velocities()[0, 0]) < 0.05 * v0


def test_resting_sphere_on_a_fixed_sphere_carries_its_weight():
    """A sphere dropped onto a fixed sphere straight below it comes to rest on
    top; the normal multiplier of the rigid-rigid contact balances its weight
    (AL fixed point with the per-step multiplier decay alpha * gamma)."""
    r, rho = 0.05, 1000.0
    sc = scenes.Scene("stack", spheres=[scenes.Sphere(r, rho, (0, 0.5, 0), ground=False, fixed=True),
                                        scenes.Sphere(r, rho, (0, 0.5 + 2 * r + 0.002, 0), ground=False)],
                      contact_margin=1e-3, dt=1.0 / 120.0, iters=50)
    o = O.from_scene(sc)
    o.step(240)
    m, _ = o.mass_fixed()
    w = m[1] * sc.g
    x = o.positions()
    assert abs(x[1, 0]) < 1e-9 and abs(x[1, 2]) < 1e-9          # stays on the axis
    assert abs(np.linalg.norm(x[1] - x[0]) - 2 * r) < 2e-3      # resting on top
    rec = o.contacts(2)
    assert len(rec) == 1
    n = rec[0, 6:9]
    assert abs(abs(n[1]) - 1.0) < 1e-9
    fy = rec[0, 26] * abs(n[1])
    assert abs(fy - w) < 0.05 * w, (fy, w)


# ---------------------------------------------------------------------------
# vertex-face contacts between triangles (new: the reference's triangle-
# triangle narrow phase is empty, collision/CollisionDetector.cpp:589-599)
# ---------------------------------------------------------------------------
def test_vertex_face_closed_form():
    """A vertex near a triangle: contact iff its distance to the triangle is
    below the margin; the normal is the face normal turned towards the face
    from the vertex's side (either side); barycentrics of the closest point,
    on an edge when the vertex is beyond it."""
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 0, 1.0]])   # (b - a) x (c - a) = -y
    margin = 2e-3
    for d, hit in [(1e-3, True), (3e-3, False)]:
        for side in (1.0, -1.0):
            p = np.array([0.25, side * d, 0.3])
            got, n, bc = O.vertex_face(p, tri, margin)
            assert got == hit
            if hit:
                assert np.allclose(n, [0, -side, 0], atol=1e-15)
                assert np.allclose(bc, [0.45, 0.25, 0.3], atol=1e-14)
    # beyond the hypotenuse: the closest point is on edge b-c
    p = np.array([0.5 + 1e-3, 0.0, 0.5 + 1e-3])
    got, n, bc = O.vertex_face(p, tri, margin)
    assert got and np.allclose(bc, [0.0, 0.5, 0.5], atol=1e-12)
    got, _, _ = O.vertex_face(np.array([0.5 + 2e-3, 0.0, 0.5 + 2e-3]), tri, margin)   # 2.8 mm away
    assert not got


def test_vertex_triangle_gradients_by_finite_differences():
    """C = dir . (x_v - sum_i barys_i x_i) for dir = n, t, b: every direction's
    gradient for the three face vertices (-barys_i dir) and the vertex (+dir)
    against central differences, and the AL accumulate G = sum lambda+ grad."""
    x = rng.normal(size=(5, 3)) * 0.1
    q = np.zeros((5, 4))
    q[:, 0] = 1
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t, b = O.complete_basis(n)
    bc = np.array([0.2, 0.3, 0.5])
    rec = _rec(n, t, b, bc, 0.0, np.zeros(3))
    h = 1e-6
    for loc in range(4):
        Cv, g, hblk, G, H = O.dyn_single(3, x, q, rec, loc, 3)
        for dirn in range(3):
            fd = np.zeros(3)
            for i in range(3):
                xp, xm = x.copy(), x.copy()
                xp[loc, i] += h
                xm[loc, i] -= h
                fd[i] = (O.dyn_single(3, xp, q, rec, 0, 3)[0][dirn] - O.dyn_single(3, xm, q, rec, 0, 3)[0][dirn]) / (2 * h)
            assert np.allclose(g[dirn, :3], fd, atol=1e-8), (loc, dirn)
        assert np.all(hblk == 0)
        w = 1.0 if loc == 3 else -bc[loc]
        assert np.allclose(g[0, :3], w * n, atol=1e-15)


def _key_pair(key, prims):
    """(vertex, face) of a vertex-face key: type 0, owner / face, corner code."""
    code = int(key) & 127
    lo = (int(key) >> 33) & ((1 << 26) - 1)
    hi = (int(key) >> 7) & ((1 << 26) - 1)
    own, face = (lo, hi) if code < 3 else (hi, lo)
    return int(prims[own][code % 3]), face


def test_vertex_face_pairs_found_once_and_match_brute_force():
    """Per step of the stacked blocks: every detected vertex-face pair is
    reported once, and the set equals the brute-force set of (vertex, face)
    pairs closer than the margin, excluding a vertex of the face and pairs
    whose vertex's owner triangle shares a particle with the face in the
    same mesh (the reference's _shouldSkip on the pair that tests it)."""
    sc = scenes.block_stack(steps=20)
    o = O.from_scene(sc)
    nb = sc.meshes[0].rest.shape[0]
    prims = np.vstack([sc.meshes[0].triangles.astype(np.int64), sc.meshes[1].triangles.astype(np.int64) + nb])
    obj = np.r_[np.zeros(len(sc.meshes[0].triangles), int), np.ones(len(sc.meshes[1].triangles), int)]
    owner = {}
    for i in range(len(prims) - 1, -1, -1):
        for v in prims[i]:
            owner[int(v)] = i
    def brute(x):
        want = set()
        for f in range(len(prims)):
            tri = x[prims[f]]
            lo, hi = tri.min(0) - sc.contact_margin, tri.max(0) + sc.contact_margin
            for v, ow in owner.items():
                if v in prims[f] or np.any(x[v] < lo) or np.any(x[v] > hi):
                    continue
                if obj[ow] == obj[f] and len(set(prims[ow]) & set(prims[f])):
                    continue
                if O.vertex_face(x[v], tri, sc.contact_margin)[0]:
                    want.add((v, f))
        return want

    seen = 0
    for step in range(12):
        want = brute(o.positions())   # the detection runs on the start-of-step positions
        o.step(1)
        keys, _, types = o.detected()
        assert np.all(types == 0)
        got = [_key_pair(k, prims) for k in keys]
        assert len(set(got)) == len(got)
        assert set(got) == want, step
        seen += len(got)
    assert seen > 0


def test_cube_resting_on_a_block_is_carried_by_vertex_face_contacts():
    """The stacked cube settles on the block: the vertical forces of the
    vertex-face contacts on the cube (vertex forces -(lambda . dir) on its
    vertices, face forces +(lambda . dir) on its faces) balance its weight
    within 2 %, and no vertex ends more than 0.5 mm behind a face."""
    sc = scenes.block_stack(steps=300)
    o = O.from_scene(sc)
    nb = sc.meshes[0].rest.shape[0]
    m, _ = o.mass_fixed()
    W = m[nb:].sum() * sc.g
    o.step(200)
    x = o.positions()
    rec = o.contacts(3)
    assert len(rec) > 8
    fy, pen = 0.0, []
    for r in rec:
        idx = r[1:5].astype(int)
        n, t, b, bc = r[6:9], r[9:12], r[12:15], r[15:18]
        f = r[26] * n + r[27] * t + r[28] * b
        pen.append(np.dot(n, x[idx[3]] - (bc[0] * x[idx[0]] + bc[1] * x[idx[1]] + bc[2] * x[idx[2]])))
        fy += -f[1] if idx[3] >= nb else f[1]
    assert abs(fy - W) < 0.02 * W, (fy, W)
    assert max(pen) < 5e-4
