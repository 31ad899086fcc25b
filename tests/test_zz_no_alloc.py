"""The allocation-free broad phase (avbd_detect_collisions_noalloc) against the
context broad phase (avbd_detect_collisions + avbd_download_candidates) on the
same FP32 state: identical candidate-pair sets and counts, on two calls in a
row, the second with a smaller input, so the shared static scratch array
keeps the larger call's contents while the smaller one runs."""
import ctypes as C

import numpy as np
import pytest

from avbd_amd import scenes

pytestmark = pytest.mark.gpu


def _context_candidates(n_cells, steps):
    """Candidates of the context path after `steps` steps of a dropped cube,
    with the FP32 state they were computed from."""
    from avbd_amd._lib import lib, check
    from avbd_amd.sim import Simulation
    sc = scenes.dropped_cube(n=n_cells, iters=4)
    m = sc.meshes[0]
    sim = Simulation(sc)
    first = sim.ctx.add_tet_mesh(m.rest, m.tets, m.E, m.nu, m.density, m.kd, m.ground)
    tris = np.ascontiguousarray(m.triangles + first, dtype=np.uint32)
    U32 = C.POINTER(C.c_uint32)
    check(lib.avbd_add_collision_triangles(sim.ctx.handle, tris.shape[0], tris.ctypes.data_as(U32)), sim.ctx.handle)
    sim.ctx.build_graph()
    sim.ctx.step(steps)
    check(lib.avbd_detect_collisions(sim.ctx.handle), sim.ctx.handle)
    x, v = sim.ctx.state()                   # FP32 values, exact in FP64
    cap = 64 * tris.shape[0]
    buf = np.empty(2 * cap, dtype=np.uint32)
    cnt = C.c_uint32()
    check(lib.avbd_download_candidates(sim.ctx.handle, buf.ctypes.data_as(U32), cap, C.byref(cnt)), sim.ctx.handle)
    pairs = buf[:2 * cnt.value].reshape(-1, 2)
    sim.ctx.close()
    return tris, x, v, sc.dt, pairs


class _Hip:
    """The few HIP runtime calls the test needs (device buffers for the
    context-free entry point), straight from the HIP runtime library."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), max(a.nbytes, 16)))
        self.ptrs.append(p.value)
        self.check(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))   # host to device
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        self.check(self.lib.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2))   # device to host
        return out

    def sync(self):
        self.check(self.lib.hipDeviceSynchronize())

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(C.c_void_p(p))
        self.ptrs = []


def _noalloc_candidates(tris, x, v, dt):
    from avbd_amd._lib import lib, check
    hip = _Hip()
    try:
        x4 = np.zeros((x.shape[0], 4), np.float32)
        v4 = np.zeros((x.shape[0], 4), np.float32)
        x4[:, :3] = x
        v4[:, :3] = v
        cap = 64 * tris.shape[0]
        dx, dv = hip.upload(x4), hip.upload(v4)
        dt_ = hip.upload(np.ascontiguousarray(tris, dtype=np.uint32))
        dp = hip.upload(np.zeros(2 * cap, np.uint32))
        cnt0 = np.full(2, 12345, np.uint32)   # overwritten by the call
        dc = hip.upload(cnt0)
        check(lib.avbd_detect_collisions_noalloc(dx, dv, dt_, tris.shape[0], dt, dp, cap, dc, None))
        hip.sync()
        cnt, status = (int(c) for c in hip.download(dc, cnt0))
        assert status == 0
        assert cnt <= cap
        pairs = hip.download(dp, np.zeros(2 * cap, np.uint32))
        return pairs[:2 * cnt].reshape(-1, 2)
    finally:
        hip.free()


def test_noalloc_broad_phase_matches_context_path():
    # 12,288 triangles (32^3 cube after 20 steps), then 768 (8^3 cube)
    for n_cells, steps in ((32, 20), (8, 3)):
        print(f"[gpu-test] noalloc broad phase: dropped cube n={n_cells} after {steps} steps", flush=True)
        tris, x, v, dt, want = _context_candidates(n_cells, steps)
        got = _noalloc_candidates(tris, x, v, dt)
        gs = set(map(tuple, got.tolist()))
        assert len(gs) == len(got)           # no duplicates
        assert gs == set(map(tuple, want.tolist()))
        print(f"[gpu-test]   {tris.shape[0]} triangles, {len(gs)} pairs: identical", flush=True)


def test_noalloc_rejects_oversized_input_and_handles_empty():
    from avbd_amd._lib import lib, AVBD_E_CAPACITY
    hip = _Hip()
    try:
        cnt0 = np.full(2, 7, np.uint32)
        dc = hip.upload(cnt0)
        rc = lib.avbd_detect_collisions_noalloc(None, None, None, 16385, 1.0 / 60, None, 0, dc, None)
        assert rc == AVBD_E_CAPACITY
        assert b"static maximum" in lib.avbd_last_error(None)
        rc = lib.avbd_detect_collisions_noalloc(None, None, None, 0, 1.0 / 60, None, 0, dc, None)
        assert rc == 0
        hip.sync()
        assert hip.download(dc, cnt0).tolist() == [0, 0]
    finally:
        hip.free()
