"""pmgpu_direct_sum_noalloc (include/pmgpu.h): the direct sum without device
allocation, through the C ABI, against pmgpu_direct_sum (the same kernels
on copies the library allocates): two calls in a row, the second smaller,
isolated and periodic, and the size limits."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, fn, pos, m, tpos, eps=0.0, box=None):
    pos = np.ascontiguousarray(pos, np.float32)
    m = np.ascontiguousarray(m, np.float32)
    tpos = np.ascontiguousarray(tpos, np.float32)
    out = np.zeros((len(tpos), 3), np.float32)
    fp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))
    b = None if box is None else np.ascontiguousarray(box, np.float32)
    rc = getattr(L, fn)(len(m), fp(pos), fp(m), len(tpos), fp(tpos), float(eps), fp(b), fp(out))
    return rc, out


def _set(seed, n, nt, L=64.0):
    r = np.random.default_rng(seed)
    pos = r.uniform(0, L, (n, 3)).astype(np.float32)
    pos[: n // 3] = (L / 2 + r.standard_normal((n // 3, 3)) * 2.0).astype(np.float32) % L
    m = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    return pos, m, pos[r.choice(n, nt, replace=False)]


@pytest.mark.parametrize("box", [None, (64.0, 64.0, 64.0)])
def test_noalloc_direct_sum_matches_the_allocating_entry_point(gpu_lib, box):
    L = gpu_lib.load_library()
    print("no-alloc direct sum, box", box, flush=True)
    for seed, n, nt, eps in ((1, 300000, 8192, 0.0), (2, 5000, 300, 0.4)):
        pos, m, tpos = _set(seed, n, nt)
        rc0, want = _call(L, "pmgpu_direct_sum", pos, m, tpos, eps, box)
        rc1, got = _call(L, "pmgpu_direct_sum_noalloc", pos, m, tpos, eps, box)
        assert rc0 == rc1 == 0
        assert np.array_equal(got, want), (seed, n, nt)
        assert np.abs(want).max() > 0


def test_noalloc_direct_sum_reloads_the_table_for_a_new_box(gpu_lib):
    L = gpu_lib.load_library()
    pos, m, tpos = _set(3, 20000, 1000, L=32.0)
    for box in ((32.0, 32.0, 32.0), (32.0, 48.0, 40.0), (32.0, 32.0, 32.0)):
        rc0, want = _call(L, "pmgpu_direct_sum", pos, m, tpos, 0.0, box)
        rc1, got = _call(L, "pmgpu_direct_sum_noalloc", pos, m, tpos, 0.0, box)
        assert rc0 == rc1 == 0
        assert np.array_equal(got, want), box


def test_noalloc_direct_sum_rejects_sizes_beyond_its_array(gpu_lib):
    L = gpu_lib.load_library()
    P = gpu_lib
    big = (1 << 20) + 1
    pos = np.zeros((big, 3), np.float32)
    m = np.zeros(big, np.float32)
    rc, _ = _call(L, "pmgpu_direct_sum_noalloc", pos, m, pos[:10])
    assert rc == P.ERR_UNSUPPORTED
    rc, _ = _call(L, "pmgpu_direct_sum_noalloc", pos[:100], m[:100], np.zeros(((1 << 16) + 1, 3), np.float32))
    assert rc == P.ERR_UNSUPPORTED
    with pytest.raises(P.PMGPUError):
        P.direct_sum_noalloc(pos, m, pos[:10])
    # still usable after the rejections
    p2, m2, t2 = _set(4, 1000, 50)
    assert np.array_equal(P.direct_sum_noalloc(p2, m2, t2), P.direct_sum(p2, m2, t2))
