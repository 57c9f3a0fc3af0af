"""The C ABI library loads and exports every symbol include/pmgpu.h declares;
argument validation that happens before any device call is checked on the
CPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "pmgpu.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(pmgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_and_python_exports_agree(gpu_lib):
    assert sorted(gpu_lib.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol(gpu_lib):
    L = C.CDLL(gpu_lib.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_mlp_parameter_count(gpu_lib):
    # 4 -> 64 -> 64 -> 64 -> 1 with biases
    assert gpu_lib.mlp_num_params() == 4 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 8705


@pytest.mark.parametrize("field,value,code", [
    ("interpolation", 4, 1), ("interpolation", -1, 1), ("finite_diff", 2, 1), ("greens_function", 4, 1), ("ext_kind", 7, 1),
])
def test_create_rejects_invalid_enums(gpu_lib, field, value, code):
    P = gpu_lib
    L = P.load_library()
    p = P.PMParams()
    p.grid[:] = [32, 32, 32]
    p.interpolation, p.finite_diff, p.greens_function = 1, 0, 0
    setattr(p, field, value)
    ctx = C.c_void_p()
    assert L.pmgpu_create(C.byref(p), C.byref(ctx)) == code
    assert not ctx.value
    assert L.pmgpu_last_error(None)


def test_create_rejects_unsupported_grid(gpu_lib):
    P = gpu_lib
    L = P.load_library()
    # any side in [8, 1024] is taken (powers of two from 32 on the tiled FFT
    # path, the rest on the direct-DFT path); outside that range is refused
    for grid in ([7, 32, 32], [32, 4, 32], [32, 32, 2048], [1025, 64, 64]):
        p = P.PMParams()
        p.grid[:] = grid
        ctx = C.c_void_p()
        assert L.pmgpu_create(C.byref(p), C.byref(ctx)) == P.ERR_UNSUPPORTED
    p = P.PMParams()
    p.grid[:] = [32, 32, 32]
    p.greens_function = P.S2_OPTIMAL  # needs a particle diameter
    assert L.pmgpu_create(C.byref(p), C.byref(C.c_void_p())) == P.ERR_INVALID_ARGUMENT


def test_host_greens_value_matches_oracle(gpu_lib, oracle):
    # the device kernel's formulas, compiled for the host, against the oracle
    grid = (32, 16, 64)
    for gf, scheme, fds in ((0, 1, 0), (1, 2, 0), (2, 2, 1), (3, 0, 0), (2, 1, 0)):
        G = oracle.green_grid(gf, scheme, fds, 2.5, grid)
        for k in ((1, 0, 0), (5, 3, 60), (31, 15, 63), (16, 8, 32), (0, 8, 0), (7, 1, 2)):
            v = gpu_lib.greens_value(gf, scheme, fds, 2.5, k, grid)
            ref = G[k[2], k[1], k[0]]
            assert abs(v - ref) <= 1e-12 * max(1.0, abs(ref)), (gf, k, v, ref)
    assert np.isnan(gpu_lib.greens_value(7, 0, 0, 1.0, (1, 0, 0), grid))


def test_null_context_is_rejected(gpu_lib):
    L = gpu_lib.load_library()
    assert L.pmgpu_pm_step(None) == gpu_lib.ERR_INVALID_ARGUMENT
    assert L.pmgpu_run(None, 1, 1, 0, None) == gpu_lib.ERR_INVALID_ARGUMENT
    assert L.pmgpu_num_particles(None) == -1
