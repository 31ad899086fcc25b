"""LBVH broad phase: oracle pins (closed forms, the Karras-2012 known answer,
brute force) and bit-exact GPU parity (Morton codes, sort order, tree
topology, node boxes, candidate-pair sets)."""
import numpy as np
import pytest

from avbd_amd import scenes

INV = 0xFFFFFFFF


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------
def test_expand_bits21_is_bit_interleave(oracle):
    rng = np.random.default_rng(0)
    for v in [0, 1, 2, 3, 0x1FFFFF, 0x155555, 0x0AAAAA] + list(rng.integers(0, 1 << 21, 200)):
        v = int(v)
        want = 0
        for b in range(21):
            want |= ((v >> b) & 1) << (3 * b)
        assert oracle.expand_bits21(v) == want


def test_karras_known_answer(oracle):
    """Keys of test/test_lbvh.cpp:38-47 (Karras 2012, Fig. 3).  Expected tree in
    sorted space: internal i -> global 8+i; root 0 = [0,7] split 3; node 3 =
    [0,3] split 1; node 1 = {0,1}; node 2 = {2,3}; node 4 = [4,7] -> (leaf 4,
    node 5); node 5 = [5,7] -> (node 6, leaf 7); node 6 = {5,6}."""
    keys = [0b11001, 0b00100, 0b00101, 0b00010, 0b11000, 0b00001, 0b10011, 0b11110]
    t = oracle.lbvh_from_codes(keys)
    assert t["sorted"].tolist() == [5, 3, 1, 2, 6, 4, 0, 7]
    n = 8
    want = {0: (3 + n, 4 + n), 3: (1 + n, 2 + n), 1: (0, 1), 2: (2, 3), 4: (4, 5 + n), 5: (6 + n, 7), 6: (5, 6)}
    for i, (l, r) in want.items():
        assert (t["left"][n + i], t["right"][n + i]) == (l, r), i
    assert t["root"] == n
    assert t["parent"][n] == INV


def recursive_tree(codes, idx):
    """Karras' definition written from scratch: split a sorted range at the
    highest bit where the (code, index) composite keys of its ends differ."""
    n = len(codes)
    order = sorted(range(n), key=lambda i: (codes[i], i))
    key = [(int(codes[o]) << 32) | o for o in order]
    nodes = {}

    def build(a, b):
        if a == b:
            return a  # leaf
        hb = (key[a] ^ key[b]).bit_length() - 1
        s = a
        while not ((key[s + 1] >> hb) & 1):
            s += 1
        la, lb = build(a, s), build(s + 1, b)
        nodes[(a, b)] = (la, lb)
        return ("n", a, b)

    root = build(0, n - 1)
    return order, nodes, root


def test_karras_matches_recursive_definition(oracle):
    rng = np.random.default_rng(1)
    for trial in range(40):
        n = int(rng.integers(2, 60))
        codes = rng.integers(0, 16 if trial % 2 else 1 << 40, n).astype(np.uint64)   # many ties when small
        t = oracle.lbvh_from_codes(codes)
        order, nodes, _ = recursive_tree(codes, None)
        assert t["sorted"].tolist() == order

        # collect (range -> children ranges) from the oracle tree
        def rng_of(k):
            if k < n:
                return (k, k)
            a, b = rng_of(t["left"][k]), rng_of(t["right"][k])
            return (a[0], b[1])

        for k in range(n, 2 * n - 1):
            a, b = rng_of(t["left"][k]), rng_of(t["right"][k])
            assert a[1] + 1 == b[0]
            kids = nodes[(a[0], b[1])]
            norm = lambda c: (c, c) if not isinstance(c, tuple) else (c[1], c[2])
            assert (norm(kids[0]), norm(kids[1])) == (a, b)


def test_refit_and_traversal_brute_force(oracle):
    rng = np.random.default_rng(2)
    v, tets = scenes.kuhn_cube(3)
    tris = scenes.surface_triangles(3)
    x = v + rng.normal(scale=0.05, size=v.shape)
    vel = rng.normal(size=v.shape)
    b = oracle.lbvh_build(tris, x, vel, 1 / 60)
    n = tris.shape[0]
    # each node box = union of the primitive boxes of its leaves
    def leaves(k):
        return [k] if k < n else leaves(b["left"][k]) + leaves(b["right"][k])
    for k in range(2 * n - 1):
        L = leaves(k)
        boxes = b["aabb"][b["sorted"][L]]
        assert np.array_equal(b["node_box"][k, :3], boxes[:, :3].min(axis=0))
        assert np.array_equal(b["node_box"][k, 3:], boxes[:, 3:].max(axis=0))
    # speculative box of a triangle = bounds of its vertices at x and x + dt v
    t0 = tris[0]
    pts = np.concatenate([x[t0], x[t0] + (1 / 60) * vel[t0]])
    assert np.allclose(b["aabb"][0, :3], pts.min(axis=0), rtol=0, atol=0)
    # traversal = all overlapping leaf pairs
    pairs = oracle.lbvh_self_pairs(b)
    lb = b["node_box"][:n]
    ov = np.all(lb[:, None, :3] <= lb[None, :, 3:], axis=2) & np.all(lb[:, None, 3:] >= lb[None, :, :3], axis=2)
    want = {(i, j) for i in range(n) for j in range(i + 1, n) if ov[i, j]}
    assert set(map(tuple, pairs.tolist())) == want
    assert len(pairs) == len(want)


# ---------------------------------------------------------------------------
# GPU bit-exact parity
# ---------------------------------------------------------------------------
def _gpu_bvh(ctx, n):
    import ctypes as C
    from avbd_amd._lib import lib, check
    nn = 2 * n - 1
    out = dict(morton=np.empty(n, dtype=np.uint64), sorted=np.empty(n, dtype=np.uint32),
               left=np.empty(nn, dtype=np.uint32), right=np.empty(nn, dtype=np.uint32),
               parent=np.empty(nn, dtype=np.uint32), node_box=np.empty((nn, 6)))
    root = C.c_uint32()
    U32 = C.POINTER(C.c_uint32)
    check(lib.avbd_download_bvh(ctx.handle, out["morton"].ctypes.data_as(C.POINTER(C.c_uint64)),
                                out["sorted"].ctypes.data_as(U32), out["left"].ctypes.data_as(U32),
                                out["right"].ctypes.data_as(U32), out["parent"].ctypes.data_as(U32),
                                out["node_box"].ctypes.data_as(C.POINTER(C.c_double)), C.byref(root)), ctx.handle)
    out["root"] = root.value
    cap = 64 * n
    buf = np.empty(2 * cap, dtype=np.uint32)
    cnt = C.c_uint32()
    check(lib.avbd_download_candidates(ctx.handle, buf.ctypes.data_as(U32), cap, C.byref(cnt)), ctx.handle)
    out["pairs"] = buf[:2 * cnt.value].reshape(-1, 2)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n_cells,steps", [(4, 3), (32, 25), (40, 2)])
def test_gpu_lbvh_bit_exact(oracle, n_cells, steps):
    """n=32: 12,288 triangles (LDS radix sort); n=40: 19,200 (rank sort)."""
    from avbd_amd.sim import Simulation
    from avbd_amd._lib import lib, check
    print(f"[gpu-test] lbvh dropped cube n={n_cells} after {steps} steps", flush=True)
    sc = scenes.dropped_cube(n=n_cells, iters=4)
    m = sc.meshes[0]
    sim = Simulation(sc)
    first = sim.ctx.add_tet_mesh(m.rest, m.tets, m.E, m.nu, m.density, m.kd, m.ground)
    tris = np.ascontiguousarray(m.triangles + first, dtype=np.uint32)
    import ctypes as C
    check(lib.avbd_add_collision_triangles(sim.ctx.handle, tris.shape[0],
                                           tris.ctypes.data_as(C.POINTER(C.c_uint32))), sim.ctx.handle)
    sim.ctx.build_graph()
    sim.ctx.step(steps)                      # every step rebuilds the LBVH first
    check(lib.avbd_detect_collisions(sim.ctx.handle), sim.ctx.handle)
    x, v = sim.ctx.state()                   # FP32 state, exactly representable in FP64
    g = _gpu_bvh(sim.ctx, tris.shape[0])
    o = oracle.lbvh_build(tris, x, v, sc.dt)
    assert np.array_equal(g["morton"], o["morton"])
    assert np.array_equal(g["sorted"], o["sorted"])
    assert np.array_equal(g["left"], o["left"])
    assert np.array_equal(g["right"], o["right"])
    assert np.array_equal(g["parent"], o["parent"])
    assert g["root"] == o["root"]
    assert np.array_equal(g["node_box"], o["node_box"])
    op = oracle.lbvh_self_pairs(o)
    gs = set(map(tuple, g["pairs"].tolist()))
    assert len(gs) == len(g["pairs"])        # no duplicates
    assert gs == set(map(tuple, op.tolist()))
    print(f"[gpu-test]   {tris.shape[0]} triangles, {len(gs)} candidate pairs: bit-exact", flush=True)


@pytest.mark.gpu
def test_gpu_concurrent_broadphase_matches_serial(oracle):
    """The broad phase forked onto a second stream (concurrent with the solve,
    inside the step graph) reports the same candidates and leaves the same
    particle state as the serial broad phase on the solve stream."""
    from avbd_amd.sim import Simulation
    print("[gpu-test] concurrent vs serial broad phase", flush=True)
    sc = scenes.dropped_cube(n=10, iters=5)
    a = Simulation(sc, lanes_per_vertex=8, block_size=128)
    b = Simulation(sc, lanes_per_vertex=8, block_size=128, serial_broadphase=True)
    a.setup()
    b.setup()
    for _ in range(4):
        a.update(5)
        b.update(5)
        assert np.array_equal(a.positions(), b.positions())
        n = sc.meshes[0].triangles.shape[0]
        ga, gb = _gpu_bvh(a.ctx, n), _gpu_bvh(b.ctx, n)
        assert set(map(tuple, ga["pairs"].tolist())) == set(map(tuple, gb["pairs"].tolist()))
        assert np.array_equal(ga["left"], gb["left"]) and np.array_equal(ga["node_box"], gb["node_box"])
