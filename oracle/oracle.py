"""ctypes front end of the CPU oracle (oracle/pm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to check
the GPU result, and by bench.py for the cpu_baseline leg.  The product library
(particlesimulation_amd) never imports this module.

All arrays are numpy; positions are (n, 3) float32 in code units (grid cells),
masses (n,) float32 in code units.  Grids are indexed [z, y, x] (x fastest in
memory), matching the reference's flat index x + y*Nx + z*Nx*Ny
(include/grid.h:52-54).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_KISS_PATH = os.path.join(_HERE, "_ref", "libkissfft.so")

NGP, CIC, TSC = 0, 1, 2
TWO_POINT, FOUR_POINT = 0, 1
DISCRETE_LAPLACIAN, S1_OPTIMAL, S2_OPTIMAL, POOR_MAN = 0, 1, 2, 3
S1, S2 = 0, 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/libkissfft.so where the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_deposit.argtypes = [C.c_int64, _f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, _f64p, C.c_int]
        L.orc_green_grid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_green_optimal.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_double, C.c_int, C.c_int]
        L.orc_green_optimal.restype = C.c_double
        L.orc_s1_fourier.argtypes = [C.c_double, C.c_double]
        L.orc_s1_fourier.restype = C.c_double
        L.orc_s2_fourier.argtypes = [C.c_double, C.c_double]
        L.orc_s2_fourier.restype = C.c_double
        L.orc_alias_sum_bruteforce.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_alias_sum_bruteforce.restype = C.c_double
        L.orc_alias_sum_closed.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_alias_sum_closed.restype = C.c_double
        L.orc_fd_d.argtypes = [C.c_int, C.c_double]
        L.orc_fd_d.restype = C.c_double
        L.orc_field.argtypes = [_f64p, C.c_int, C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_interpolate.argtypes = [C.c_int64, _f32p, _f64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_void_p, _f64p]
        L.orc_pm_accel.argtypes = [C.c_int64, _f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _f64p, C.c_int, C.c_int, C.c_void_p, _f64p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        L.orc_reference_force.argtypes = [C.c_int, C.c_float, C.c_float]
        L.orc_reference_force.restype = C.c_float
        L.orc_sr_table.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, _f32p,
                                   C.POINTER(C.c_float)]
        L.orc_sr_pair_factor.argtypes = [C.c_int, C.c_double, C.c_double, _f64p, C.c_int64, _f64p]
        L.orc_sr_table_log.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, _f32p, C.POINTER(C.c_float),
                                       C.POINTER(C.c_uint32)]
        L.orc_p3m_sr.argtypes = [C.c_int64, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float, C.c_float,
                                 C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, _f32p,
                                 C.POINTER(C.c_int64)]
        L.orc_p3m_mesh_dims.argtypes = [_f32p, C.c_float, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
        L.orc_direct_sum.argtypes = [C.c_int64, C.c_void_p, C.c_int64, _f32p, _f32p, C.c_double, C.c_double,
                                     C.c_int, _f64p]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_ewald_accel.argtypes = [_f64p, C.c_double, _f64p, _f64p]
        L.orc_ewald_accel.restype = None
        L.orc_direct_sum_periodic.argtypes = [C.c_int64, C.c_void_p, C.c_int64, _f32p, _f32p, C.c_double, _f64p,
                                              _f64p]
        L.orc_mlp_forward.argtypes = [_f64p, C.c_int64, _f64p, _f64p]
        L.orc_mlp_forward.restype = None
        L.orc_mlp_backward.argtypes = [_f64p, C.c_int64, _f64p, _f64p, _f64p]
        L.orc_mlp_backward.restype = None
        _lib = L
    return _lib


def _pos(pos):
    return np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def deposit(pos, mass, grid, scheme, wrap=True):
    """PMMethod::spreadMass (source/pmMethod.cpp:200-277) -> rho[z, y, x] float64."""
    nx, ny, nz = grid
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float32)
    rho = np.zeros(nx * ny * nz, np.float64)
    rc = lib().orc_deposit(len(mass), pos.ravel(), mass, nx, ny, nz, scheme, int(wrap), rho)
    if rc != 0:
        raise ValueError(f"orc_deposit failed ({rc})")
    return rho.reshape(nz, ny, nx)


def fft3d(a, inverse=False):
    """Unnormalised complex 3-D DFT of a[z, y, x] (sign -1 forward, +1 inverse)."""
    a = np.asarray(a, dtype=np.complex128)
    nz, ny, nx = a.shape
    re = np.ascontiguousarray(a.real).ravel().copy()
    im = np.ascontiguousarray(a.imag).ravel().copy()
    rc = lib().orc_fft3d(nx, ny, nz, re, im, +1 if inverse else -1)
    if rc != 0:
        raise ValueError("orc_fft3d: lengths must be positive")
    return (re + 1j * im).reshape(nz, ny, nx)


def green_grid(gf, scheme, fds, a, grid):
    """PMMethod::initGreensFunction (source/pmMethod.cpp:164-185) -> G[kz, ky, kx] (unfolded)."""
    nx, ny, nz = grid
    G = np.zeros(nx * ny * nz, np.float64)
    if lib().orc_green_grid(gf, scheme, fds, float(a), nx, ny, nz, G) != 0:
        raise ValueError("bad Green's function options")
    return G.reshape(nz, ny, nx)


def green_optimal(scheme, k, grid, a, cs, fds):
    return lib().orc_green_optimal(scheme, int(k[0]), int(k[1]), int(k[2]), grid[0], grid[1], grid[2],
                                   float(a), cs, fds)


def s_fourier(cs, k, a):
    return (lib().orc_s1_fourier if cs == S1 else lib().orc_s2_fourier)(float(k), float(a))


def alias_sum_bruteforce(scheme, k, lim):
    return lib().orc_alias_sum_bruteforce(scheme, float(k[0]), float(k[1]), float(k[2]), int(lim))


def alias_sum_closed(scheme, k):
    return lib().orc_alias_sum_closed(scheme, float(k[0]), float(k[1]), float(k[2]))


def fd_d(fds, k):
    """D(k)/i of the 2-point or 4-point gradient the optimal G assumes."""
    return lib().orc_fd_d(int(fds), float(k))


def field(phi, fds):
    """PMMethod::findFieldInCells (source/pmMethod.cpp:373-382) -> E[z, y, x, 3]."""
    nz, ny, nx = phi.shape
    E = np.zeros(nx * ny * nz * 3, np.float64)
    lib().orc_field(np.ascontiguousarray(phi, np.float64).ravel(), nx, ny, nz, fds, E)
    return E.reshape(nz, ny, nx, 3)


def _ext_buf(ext_prm):
    if ext_prm is None:
        return 0, None, None
    buf = np.ascontiguousarray(ext_prm, np.float64)
    assert buf.size == 8, "ext_prm = (cx, cy, cz, R, M, G, H, DT)"
    return 1, buf, buf.ctypes.data


def interpolate(pos, E, scheme, wrap=True, ext_prm=None):
    nz, ny, nx, _ = E.shape
    pos = _pos(pos)
    acc = np.zeros(pos.size, np.float64)
    ext, keep, ptr = _ext_buf(ext_prm)
    rc = lib().orc_interpolate(len(pos), pos.ravel(), np.ascontiguousarray(E).ravel(), nx, ny, nz, scheme,
                               int(wrap), ext, ptr, acc)
    if rc != 0:
        raise ValueError("orc_interpolate: particle stencil outside the grid")
    return acc.reshape(-1, 3)


def pm_accel(pos, mass, grid, scheme, fds, G, wrap=True, ext_prm=None, want_grids=False):
    """PMMethod::pmMethodStep (source/pmMethod.cpp:137-144) -> acc (n,3) float64
    [, rho, phi, E] following the reference's complex-to-complex route."""
    nx, ny, nz = grid
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float32)
    acc = np.zeros(pos.size, np.float64)
    ext, keep, ptr = _ext_buf(ext_prm)
    m = nx * ny * nz
    rho = np.zeros(m, np.float64) if want_grids else None
    phi = np.zeros(m, np.float64) if want_grids else None
    E = np.zeros(3 * m, np.float64) if want_grids else None
    rc = lib().orc_pm_accel(len(mass), pos.ravel(), mass, nx, ny, nz, scheme, fds,
                            np.ascontiguousarray(G, np.float64).ravel(), int(wrap), ext, ptr, acc,
                            rho.ctypes.data if want_grids else None,
                            phi.ctypes.data if want_grids else None,
                            E.ctypes.data if want_grids else None)
    if rc != 0:
        raise ValueError(f"orc_pm_accel failed ({rc})")
    acc = acc.reshape(-1, 3)
    if want_grids:
        return acc, rho.reshape(nz, ny, nx), phi.reshape(nz, ny, nx), E.reshape(nz, ny, nx, 3)
    return acc


def reference_force(cs, r, a):
    return lib().orc_reference_force(cs, float(r), float(a))


def sr_table(rc, a, eps, cs, cnt=500):
    t = np.zeros(cnt, np.float32)
    d2 = C.c_float(0.0)
    lib().orc_sr_table(float(rc), float(a), float(eps), cs, cnt, t, C.byref(d2))
    return t, d2.value


def sr_pair_factor(cs, a, eps, r2):
    """FP64 pair factor F of the short-range force for an array of r^2:
    a_i += m_j F (x_i - x_j) over 0 < r^2 < rc^2 (0 where r^2 <= 0)."""
    r2 = np.ascontiguousarray(r2, np.float64)
    out = np.empty_like(r2)
    lib().orc_sr_pair_factor(int(cs), float(a), float(eps), r2, r2.size, out)
    return out


def sr_table_log(rc, a, eps, cs):
    """The accurate (use_table=2) table: (entries, 2) float32 (T, S) rows over
    bit-binned s = r^2 * scale, the scale and the float bits of the first bin."""
    t = np.zeros((lib().orc_sr_table_log_entries(), 2), np.float32)
    c = C.c_float(0.0)
    lo = C.c_uint32(0)
    lib().orc_sr_table_log(float(rc), float(a), float(eps), cs, t, C.byref(c), C.byref(lo))
    return t, c.value, lo.value


def p3m_mesh_dims(box_orig, rc_orig):
    out = np.zeros(3, np.int32)
    lib().orc_p3m_mesh_dims(np.asarray(box_orig, np.float32), float(rc_orig), out)
    return tuple(int(v) for v in out)


def p3m_short_range(pos, mass, box_orig, rc_orig, H, rc, a, eps, cs, use_table=True, ysort=True,
                    sample_stride=1):
    """P3MMethod::calculateShortRangeForces + correctAccelerations
    (source/p3mMethod.cpp:168-192, 50-57) -> (acc_sr (n,3) float32, ordered pairs).
    use_table: 0/False direct formula, 1/True reference table, 2 accurate table.
    sample_stride > 1: every sample_stride-th particle of each cell's list as
    the home particle only (partial results; for timing a bounded sample)."""
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float32)
    acc = np.zeros(pos.size, np.float32)
    pairs = C.c_int64(0)
    rc_ = lib().orc_p3m_sr(len(mass), pos.ravel(), mass, np.asarray(box_orig, np.float32), float(rc_orig),
                           float(H), float(rc), float(a), float(eps), cs, int(use_table), int(ysort),
                           int(sample_stride), acc, C.byref(pairs))
    if rc_ != 0:
        raise ValueError(f"orc_p3m_sr failed ({rc_})")
    return acc.reshape(-1, 3), pairs.value


def direct_sum(pos, mass, Gc, eps=0.0, kernel=0, targets=None):
    """FP64 direct sum (source/ppMethod.cpp:23-36 semantics) on the target subset."""
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float32)
    if targets is None:
        nt, tptr, keep = len(mass), None, None
    else:
        keep = np.ascontiguousarray(targets, np.int64)
        nt, tptr = len(keep), keep.ctypes.data
    acc = np.zeros(3 * nt, np.float64)
    lib().orc_direct_sum(nt, tptr, len(mass), pos.ravel(), mass, float(Gc), float(eps), int(kernel), acc)
    return acc.reshape(-1, 3)


def ewald_accel(box, x, alpha=None):
    """Periodic acceleration of a unit mass at the origin (G = 1, neutralising background)."""
    L = np.ascontiguousarray(box, np.float64)
    a = 2.0 / L.min() if alpha is None else float(alpha)
    out = np.zeros(3)
    lib().orc_ewald_accel(L, a, np.ascontiguousarray(x, np.float64), out)
    return out


def direct_sum_periodic(pos, mass, Gc, box, targets=None):
    """FP64 periodic (Ewald) direct sum on the target subset."""
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, np.float32)
    if targets is None:
        nt, tptr = len(mass), None
    else:
        keep = np.ascontiguousarray(targets, np.int64)
        nt, tptr = len(keep), keep.ctypes.data
    acc = np.zeros(3 * nt, np.float64)
    lib().orc_direct_sum_periodic(nt, tptr, len(mass), pos.ravel(), mass, float(Gc),
                                  np.ascontiguousarray(box, np.float64), acc)
    return acc.reshape(-1, 3)


def mlp_forward(params, inputs):
    """FP64 forward pass of the correction network (csrc/mlp.hip)."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(inputs, np.float64).reshape(-1, 4)
    y = np.zeros(len(x), np.float64)
    lib().orc_mlp_forward(p, len(x), x.ravel(), y)
    return y


def mlp_backward(params, inputs, up):
    """FP64 gradient of sum_m up[m] * y(m) with respect to the parameters."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(inputs, np.float64).reshape(-1, 4)
    u = np.ascontiguousarray(up, np.float64)
    g = np.zeros(len(p), np.float64)
    lib().orc_mlp_backward(p, len(x), x.ravel(), u, g)
    return g


# --------------------------------------------------------------------------
# The reference's vendored FFT backend (kissfft via include/kissFFTAdapter.h),
# built from the reference sources into oracle/_ref/ by oracle/Makefile.
# --------------------------------------------------------------------------
_kiss = None
_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def _libc_free(p):
    _libc.free(p)


def kissfft_available() -> bool:
    return os.path.exists(_KISS_PATH)


def kiss_fftnd(a, inverse=False):
    """KissFFTAdapter<float>::fft / ifft (include/kissFFTAdapter.h:16-29): complex64
    3-D transform; the inverse is divided by the length as the adapter does."""
    global _kiss
    if _kiss is None:
        _kiss = C.CDLL(_KISS_PATH)
        _kiss.kiss_fftnd_alloc.restype = C.c_void_p
        _kiss.kiss_fftnd_alloc.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        _kiss.kiss_fftnd.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    a = np.ascontiguousarray(a, np.complex64)
    dims = (C.c_int * a.ndim)(*a.shape)  # slowest first, as the adapter receives {Nz,Ny,Nx}
    cfg = _kiss.kiss_fftnd_alloc(dims, a.ndim, int(inverse), None, None)
    out = np.zeros_like(a)
    _kiss.kiss_fftnd(cfg, a.ctypes.data, out.ctypes.data)
    _libc_free(cfg)  # kiss_fft_free is a macro for free() (kiss_fft.h:138)
    if inverse:
        out /= np.float32(a.size)
    return out
