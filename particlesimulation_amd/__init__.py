"""Python front end of libpmgpu, the MI355X PM / P3M force step.

A thin ctypes layer over the C ABI in include/pmgpu.h.  It mirrors the
reference's PMMethod / P3MMethod surface (include/pmMethod.h:12-76,
include/p3mMethod.h:11-63 of the reference) for tests and benchmarks:

    pm = PM(grid=(64, 64, 64), interpolation=CIC, finite_diff=TWO_POINT,
            greens_function=DISCRETE_LAPLACIAN, particle_diameter=0.0,
            box=(64, 64, 64))
    pm.set_particles(pos, vel, mass)     # code units, caller order
    pm.init_greens()                     # PMMethod::initGreensFunction
    pm.pm_step()                         # PMMethod::pmMethodStep
    pos, vel, acc = pm.get_particles()

The native library must be built (particlesimulation_amd.build); there is no
fallback: loading fails loudly when libpmgpu.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmgpu.so")

# enums, same numbering as the reference's enum classes (include/pmConfig.h)
NGP, CIC, TSC = 0, 1, 2
TSC_CENTRED = 3  # TSC about the nearest grid point (an extension; TSC is the reference's floor anchor)
TWO_POINT, FOUR_POINT = 0, 1
DISCRETE_LAPLACIAN, S1_OPTIMAL, S2_OPTIMAL, POOR_MAN = 0, 1, 2, 3
S1, S2 = 0, 1
EXT_NONE, EXT_SPH_RAD_DECR = 0, 1
HALF_KICK, INTEGER_STEP, KICK, DRIFT = 0, 1, 2, 3
GRID_DENSITY, GRID_POTENTIAL, GRID_FIELD, GRID_GREENS = 0, 1, 2, 3

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_NOT_READY = 2
ERR_DEVICE = 3
ERR_OUT_OF_MEMORY = 4
ERR_UNSUPPORTED = 5

# The C ABI symbols declared in include/pmgpu.h.
EXPORTS = (
    "pmgpu_create", "pmgpu_destroy", "pmgpu_last_error", "pmgpu_synchronize",
    "pmgpu_set_particles", "pmgpu_get_particles", "pmgpu_update_particles", "pmgpu_num_particles",
    "pmgpu_init_greens", "pmgpu_pm_step", "pmgpu_p3m_configure", "pmgpu_p3m_short_range",
    "pmgpu_leapfrog", "pmgpu_get_integer_step_velocities", "pmgpu_run", "pmgpu_escaped",
    "pmgpu_get_grid", "pmgpu_set_greens_correction", "pmgpu_mlp_num_params",
    "pmgpu_profile_enable", "pmgpu_profile_reset", "pmgpu_profile_read", "pmgpu_p3m_count_pairs",
    "pmgpu_greens_value", "pmgpu_diagnostics", "pmgpu_mlp_forward", "pmgpu_mlp_backward",
    "pmgpu_trainer_create", "pmgpu_trainer_destroy", "pmgpu_trainer_last_error", "pmgpu_trainer_add_snapshot",
    "pmgpu_trainer_loss_grad", "pmgpu_train", "pmgpu_direct_sum", "pmgpu_direct_sum_host", "pmgpu_direct_sum_noalloc", "pmgpu_set_tuning", "pmgpu_get_tuning", "pmgpu_p3m_pair_stats",
    "pmgpu_p3m_cell_geometry", "pmgpu_add_accelerations", "pmgpu_p3m_record_short_range",
    "pmgpu_p3m_get_short_range",
    "pmgpu_batch_create", "pmgpu_batch_destroy", "pmgpu_batch_last_error", "pmgpu_batch_size",
    "pmgpu_batch_set_particles", "pmgpu_batch_prepare", "pmgpu_batch_step", "pmgpu_batch_synchronize",
    "pmgpu_batch_time", "pmgpu_batch_member_latency", "pmgpu_batch_get_particles", "pmgpu_batch_get_tuning",
)


class PMParams(C.Structure):
    _fields_ = [
        ("grid", C.c_int32 * 3),
        ("interpolation", C.c_int32),
        ("finite_diff", C.c_int32),
        ("greens_function", C.c_int32),
        ("particle_diameter", C.c_float),
        ("box", C.c_float * 3),
        ("ext_kind", C.c_int32),
        ("ext_params", C.c_float * 6),
        ("H", C.c_float),
        ("DT", C.c_float),
    ]


class P3MParams(C.Structure):
    _fields_ = [
        ("cutoff", C.c_float),
        ("particle_diameter", C.c_float),
        ("softening", C.c_float),
        ("cloud_shape", C.c_int32),
        ("use_table", C.c_int32),
        ("box", C.c_float * 3),
    ]


class PMGPUError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pmgpu error {code}: {msg}")
        self.code = code


class InvalidArgument(PMGPUError, ValueError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libpmgpu.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m particlesimulation_amd.build`")
    L = C.CDLL(path)
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    L.pmgpu_create.argtypes = [C.POINTER(PMParams), C.POINTER(vp)]
    L.pmgpu_destroy.argtypes = [vp]
    L.pmgpu_destroy.restype = None
    L.pmgpu_last_error.argtypes = [vp]
    L.pmgpu_last_error.restype = C.c_char_p
    L.pmgpu_synchronize.argtypes = [vp]
    L.pmgpu_set_particles.argtypes = [vp, C.c_int64, fp, fp, fp]
    L.pmgpu_get_particles.argtypes = [vp, fp, fp, fp]
    L.pmgpu_add_accelerations.argtypes = [vp, fp]
    L.pmgpu_p3m_record_short_range.argtypes = [vp, C.c_int32]
    L.pmgpu_p3m_get_short_range.argtypes = [vp, fp]
    L.pmgpu_update_particles.argtypes = [vp, fp, fp, fp]
    L.pmgpu_num_particles.argtypes = [vp]
    L.pmgpu_num_particles.restype = C.c_int64
    L.pmgpu_init_greens.argtypes = [vp]
    L.pmgpu_pm_step.argtypes = [vp]
    L.pmgpu_p3m_configure.argtypes = [vp, C.POINTER(P3MParams)]
    L.pmgpu_p3m_short_range.argtypes = [vp]
    L.pmgpu_leapfrog.argtypes = [vp, C.c_int32, C.c_float]
    L.pmgpu_get_integer_step_velocities.argtypes = [vp, fp]
    L.pmgpu_run.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    L.pmgpu_escaped.argtypes = [vp, C.POINTER(C.c_int32)]
    L.pmgpu_get_grid.argtypes = [vp, C.c_int32, fp]
    L.pmgpu_set_greens_correction.argtypes = [vp, fp]
    L.pmgpu_mlp_num_params.restype = C.c_int64
    L.pmgpu_profile_enable.argtypes = [vp, C.c_int32]
    L.pmgpu_profile_reset.argtypes = [vp]
    L.pmgpu_profile_read.argtypes = [vp, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.pmgpu_p3m_count_pairs.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.pmgpu_p3m_pair_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    L.pmgpu_p3m_cell_geometry.argtypes = [vp, C.POINTER(C.c_int32)]
    L.pmgpu_greens_value.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_double] + [C.c_int32] * 6
    L.pmgpu_greens_value.restype = C.c_double
    L.pmgpu_diagnostics.argtypes = [vp, C.POINTER(C.c_double)]
    L.pmgpu_batch_create.argtypes = [C.c_int32, C.POINTER(PMParams), C.POINTER(P3MParams), C.POINTER(C.c_int32),
                                     C.c_int32, C.POINTER(vp)]
    L.pmgpu_batch_destroy.argtypes = [vp]
    L.pmgpu_batch_destroy.restype = None
    L.pmgpu_batch_last_error.argtypes = [vp]
    L.pmgpu_batch_last_error.restype = C.c_char_p
    L.pmgpu_batch_size.argtypes = [vp]
    L.pmgpu_batch_set_particles.argtypes = [vp, C.c_int32, C.c_int64, fp, fp, fp]
    L.pmgpu_batch_prepare.argtypes = [vp, C.c_int32]
    L.pmgpu_batch_step.argtypes = [vp, C.c_int32]
    L.pmgpu_batch_synchronize.argtypes = [vp]
    L.pmgpu_batch_time.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.pmgpu_batch_member_latency.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.pmgpu_batch_get_particles.argtypes = [vp, C.c_int32, fp, fp, fp]
    L.pmgpu_batch_get_tuning.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.pmgpu_set_tuning.argtypes = [vp, C.POINTER(C.c_int32)]
    L.pmgpu_get_tuning.argtypes = [vp, C.POINTER(C.c_int32)]
    L.pmgpu_mlp_forward.argtypes = [fp, C.c_int64, fp, fp]
    L.pmgpu_mlp_backward.argtypes = [fp, C.c_int64, fp, fp, fp]
    L.pmgpu_trainer_create.argtypes = [C.POINTER(PMParams), C.POINTER(vp)]
    L.pmgpu_trainer_destroy.argtypes = [vp]
    L.pmgpu_trainer_destroy.restype = None
    L.pmgpu_trainer_last_error.argtypes = [vp]
    L.pmgpu_trainer_last_error.restype = C.c_char_p
    L.pmgpu_trainer_add_snapshot.argtypes = [vp, C.c_int64, fp, fp, C.c_int64, C.POINTER(C.c_int64), fp]
    L.pmgpu_trainer_loss_grad.argtypes = [vp, fp, C.POINTER(C.c_double), fp, fp]
    L.pmgpu_train.argtypes = [C.POINTER(vp), C.c_int32, fp, C.c_int32, C.c_float, fp]
    L.pmgpu_direct_sum.argtypes = [C.c_int64, fp, fp, C.c_int64, fp, C.c_float, fp, fp]
    L.pmgpu_direct_sum_noalloc.argtypes = [C.c_int64, fp, fp, C.c_int64, fp, C.c_float, fp, fp]
    L.pmgpu_direct_sum_host.argtypes = [C.c_int64, fp, fp, C.c_int64, fp, C.c_float, fp, C.c_int32,
                                        C.POINTER(C.c_double)]
    _lib = L
    return L


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


class PM:
    """Context of one simulation: PMMethod (+ P3MMethod) on the GPU."""

    def __init__(self, grid, interpolation=CIC, finite_diff=TWO_POINT, greens_function=DISCRETE_LAPLACIAN,
                 particle_diameter=0.0, box=None, ext_field=None, H=1.0, DT=1.0):
        L = load_library()
        p = PMParams()
        p.grid[:] = [int(g) for g in grid]
        p.interpolation = int(interpolation)
        p.finite_diff = int(finite_diff)
        p.greens_function = int(greens_function)
        p.particle_diameter = float(particle_diameter)
        box = grid if box is None else box
        p.box[:] = [float(b) for b in box]
        if ext_field is None:
            p.ext_kind = EXT_NONE
        else:
            p.ext_kind = EXT_SPH_RAD_DECR
            p.ext_params[:] = [float(v) for v in ext_field]
        p.H = float(H)
        p.DT = float(DT)
        self.params = p
        self.grid = tuple(int(g) for g in grid)
        self._lib = L
        self._ctx = C.c_void_p()
        rc = L.pmgpu_create(C.byref(p), C.byref(self._ctx))
        if rc != OK:
            msg = L.pmgpu_last_error(None).decode()
            raise (InvalidArgument if rc == ERR_INVALID_ARGUMENT else PMGPUError)(rc, msg)
        self.n = 0

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc):
        if rc != OK:
            msg = self._lib.pmgpu_last_error(self._ctx).decode()
            raise (InvalidArgument if rc == ERR_INVALID_ARGUMENT else PMGPUError)(rc, msg)

    def close(self):
        if self._ctx:
            self._lib.pmgpu_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- particles --------------------------------------------------------
    def set_particles(self, pos, vel, mass):
        pos = _f32(pos, (-1, 3))
        mass = _f32(mass)
        vel = None if vel is None else _f32(vel, (-1, 3))
        self._keep = (pos, vel, mass)
        self._check(self._lib.pmgpu_set_particles(self._ctx, len(mass), _fptr(pos), _fptr(vel), _fptr(mass)))
        self.n = len(mass)

    def get_particles(self):
        pos = np.zeros((self.n, 3), np.float32)
        vel = np.zeros((self.n, 3), np.float32)
        acc = np.zeros((self.n, 3), np.float32)
        self._check(self._lib.pmgpu_get_particles(self._ctx, _fptr(pos), _fptr(vel), _fptr(acc)))
        return pos, vel, acc

    def add_accelerations(self, acc):
        """Adds host accelerations (code units, (n, 3), caller order) to those
        of the last force evaluation."""
        a = _f32(acc, (-1, 3))
        self._check(self._lib.pmgpu_add_accelerations(self._ctx, _fptr(a)))

    def update_particles(self, pos=None, vel=None, acc=None):
        a = [None if x is None else _f32(x, (-1, 3)) for x in (pos, vel, acc)]
        self._check(self._lib.pmgpu_update_particles(self._ctx, *[_fptr(x) for x in a]))

    # -- force step -------------------------------------------------------
    def init_greens(self):
        self._check(self._lib.pmgpu_init_greens(self._ctx))

    def pm_step(self):
        self._check(self._lib.pmgpu_pm_step(self._ctx))

    def p3m_configure(self, cutoff, particle_diameter, softening=0.0, cloud_shape=S2, use_table=1, box=None):
        """use_table: 0 direct formula, 1 (True) the reference's table, 2 the accurate table."""
        q = P3MParams()
        q.cutoff = float(cutoff)
        q.particle_diameter = float(particle_diameter)
        q.softening = float(softening)
        q.cloud_shape = int(cloud_shape)
        q.use_table = int(use_table)
        box = self.params.box[:] if box is None else box
        q.box[:] = [float(b) for b in box]
        self._check(self._lib.pmgpu_p3m_configure(self._ctx, C.byref(q)))

    def p3m_short_range(self):
        self._check(self._lib.pmgpu_p3m_short_range(self._ctx))

    def p3m_record_short_range(self, on=True):
        self._check(self._lib.pmgpu_p3m_record_short_range(self._ctx, int(bool(on))))

    def p3m_short_range_accelerations(self):
        """The short-range part of the last short-range pass's accelerations (n, 3)."""
        a = np.zeros((self.n, 3), np.float32)
        self._check(self._lib.pmgpu_p3m_get_short_range(self._ctx, _fptr(a)))
        return a

    def p3m_count_pairs(self, tested=False):
        v = C.c_int64(0)
        t = C.c_int64(0)
        self._check(self._lib.pmgpu_p3m_count_pairs(self._ctx, C.byref(v), C.byref(t)))
        return (v.value, t.value) if tested else v.value

    def p3m_pair_stats(self):
        """(accepted pairs, tested pairs, lane-slots issued) of the short-range pass."""
        s = (C.c_int64 * 3)()
        self._check(self._lib.pmgpu_p3m_pair_stats(self._ctx, s))
        return tuple(s)

    def p3m_cell_geometry(self):
        """(x, y, z divisions, cells) of the short-range cell grid in use."""
        g = (C.c_int32 * 4)()
        self._check(self._lib.pmgpu_p3m_cell_geometry(self._ctx, g))
        return tuple(g)

    def leapfrog(self, op, dt=1.0):
        self._check(self._lib.pmgpu_leapfrog(self._ctx, int(op), float(dt)))

    def integer_step_velocities(self):
        v = np.zeros((self.n, 3), np.float32)
        self._check(self._lib.pmgpu_get_integer_step_velocities(self._ctx, _fptr(v)))
        return v

    def run(self, iterations, initial=True, p3m=False, sync=True):
        done = C.c_int32(0)
        self._check(self._lib.pmgpu_run(self._ctx, int(iterations), int(bool(initial)), int(bool(p3m)),
                                        C.byref(done) if sync else None))
        return done.value if sync else None

    def escaped(self):
        v = C.c_int32(0)
        self._check(self._lib.pmgpu_escaped(self._ctx, C.byref(v)))
        return bool(v.value)

    def synchronize(self):
        self._check(self._lib.pmgpu_synchronize(self._ctx))

    def diagnostics(self):
        """FP64 device reductions (code units): dict of kinetic, momentum, angular_momentum,
        rho_phi, mass, mass_moment."""
        out = (C.c_double * 12)()
        self._check(self._lib.pmgpu_diagnostics(self._ctx, out))
        v = list(out)
        return {"kinetic": v[0], "momentum": v[1:4], "angular_momentum": v[4:7], "rho_phi": v[7],
                "mass": v[8], "mass_moment": v[9:12]}

    # -- grids ------------------------------------------------------------
    def get_grid(self, which):
        nx, ny, nz = self.grid
        if which == GRID_FIELD:
            out = np.zeros((nz, ny, nx, 3), np.float32)
        elif which == GRID_GREENS:
            out = np.zeros((nz, ny, nx // 2 + 1), np.float32)
        else:
            out = np.zeros((nz, ny, nx), np.float32)
        self._check(self._lib.pmgpu_get_grid(self._ctx, int(which), _fptr(out)))
        return out

    def density(self):
        return self.get_grid(GRID_DENSITY)

    def potential(self):
        return self.get_grid(GRID_POTENTIAL)

    # -- learned correction ----------------------------------------------
    def set_greens_correction(self, weights):
        w = None if weights is None else _f32(weights)
        self._check(self._lib.pmgpu_set_greens_correction(self._ctx, _fptr(w)))

    # -- launch tuning ----------------------------------------------------
    def set_tuning(self, deposit_chunk=1024, interp_chunk=2048, blocks_per_cu=32, sr_div=2, interp_kernel=0,
                   sr_chunking=0):
        k = (C.c_int32 * 6)(int(deposit_chunk), int(interp_chunk), int(blocks_per_cu), int(sr_div),
                            int(interp_kernel), int(sr_chunking))
        self._check(self._lib.pmgpu_set_tuning(self._ctx, k))

    def get_tuning(self):
        k = (C.c_int32 * 6)()
        self._check(self._lib.pmgpu_get_tuning(self._ctx, k))
        return tuple(k)

    # -- profiling --------------------------------------------------------
    def profile(self, on=True, short_range_only=False):
        """HIP events around every kernel of the step, or (short_range_only)
        around the short-range pair kernel alone; off with on=False."""
        self._check(self._lib.pmgpu_profile_enable(self._ctx, 0 if not on else (2 if short_range_only else 1)))

    def profile_reset(self):
        self._check(self._lib.pmgpu_profile_reset(self._ctx))

    def profile_read(self):
        n = 64
        names = (C.c_char_p * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        k = C.c_int32(0)
        self._check(self._lib.pmgpu_profile_read(self._ctx, n, names, ms, cnt, C.byref(k)))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k.value)}


def greens_value(gf, scheme, fds, a, k, grid) -> float:
    """G(k) of one unfolded mode, FP64, on the host (the device kernel's formulas)."""
    return float(load_library().pmgpu_greens_value(int(gf), int(scheme), int(fds), float(a), int(k[0]), int(k[1]),
                                                   int(k[2]), int(grid[0]), int(grid[1]), int(grid[2])))


def mlp_num_params() -> int:
    return int(load_library().pmgpu_mlp_num_params())


# ---------------------------------------------------------------------------
# learned influence-function correction
# ---------------------------------------------------------------------------

def _raise(rc, msg):
    raise (InvalidArgument if rc == ERR_INVALID_ARGUMENT else PMGPUError)(rc, msg)


def mlp_init(seed=0, scale=1.0):
    """Initial MLP parameters: He-scaled normal weights, zero biases and a zero
    output layer, so the correction starts as the analytic G (factor 1)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp_num_params(), np.float32)
    o = 0
    for fan_in, fan_out in ((4, 64), (64, 64), (64, 64)):
        w = rng.standard_normal((fan_out, fan_in)) * scale * np.sqrt(2.0 / fan_in)
        p[o:o + fan_in * fan_out] = w.ravel()
        o += fan_in * fan_out + fan_out
    return p


def mlp_grid_inputs(grid, scheme):
    """MLP inputs of every half-spectrum mode [z][y][kx] (n x 4)."""
    nx, ny, nz = grid
    kx = np.arange(nx // 2 + 1)
    ky = np.minimum(np.arange(ny), ny - np.arange(ny))
    kz = np.minimum(np.arange(nz), nz - np.arange(nz))
    Z, Y, X = np.meshgrid(kz, ky, kx, indexing="ij")
    x = np.stack([X / (nx // 2), Y / (ny // 2), Z / (nz // 2), np.full(X.shape, float(scheme - CIC))], axis=-1)
    return x.reshape(-1, 4).astype(np.float32)


def mlp_forward(params, inputs):
    L = load_library()
    params, inputs = _f32(params), _f32(inputs, (-1, 4))
    out = np.zeros(len(inputs), np.float32)
    rc = L.pmgpu_mlp_forward(_fptr(params), len(inputs), _fptr(inputs), _fptr(out))
    if rc != OK:
        _raise(rc, "pmgpu_mlp_forward failed")
    return out


def mlp_backward(params, inputs, up):
    L = load_library()
    params, inputs, up = _f32(params), _f32(inputs, (-1, 4)), _f32(up)
    grad = np.zeros(mlp_num_params(), np.float32)
    rc = L.pmgpu_mlp_backward(_fptr(params), len(inputs), _fptr(inputs), _fptr(up), _fptr(grad))
    if rc != OK:
        _raise(rc, "pmgpu_mlp_backward failed")
    return grad


def direct_sum(pos, mass, targets, softening=0.0, box=None):
    """Direct-sum accelerations (code units, G' = 1/(4 pi)) at target positions, on the GPU;
    periodic (Ewald) in `box` when given."""
    L = load_library()
    pos, mass, targets = _f32(pos, (-1, 3)), _f32(mass), _f32(targets, (-1, 3))
    out = np.zeros((len(targets), 3), np.float32)
    b = None if box is None else _f32(box)
    rc = L.pmgpu_direct_sum(len(mass), _fptr(pos), _fptr(mass), len(targets), _fptr(targets), float(softening),
                            _fptr(b), _fptr(out))
    if rc != OK:
        _raise(rc, "pmgpu_direct_sum failed")
    return out


DS_NOALLOC_MAX_SOURCES = 1 << 20
DS_NOALLOC_MAX_TARGETS = 1 << 16


def direct_sum_noalloc(pos, mass, targets, softening=0.0, box=None):
    """direct_sum without device allocation (a static device array holds the
    staging copies): at most DS_NOALLOC_MAX_SOURCES sources and
    DS_NOALLOC_MAX_TARGETS targets; calls that overlap on one device race."""
    L = load_library()
    pos = _f32(pos, (-1, 3))
    mass = _f32(mass)
    targets = _f32(targets, (-1, 3))
    if len(mass) > DS_NOALLOC_MAX_SOURCES or len(targets) > DS_NOALLOC_MAX_TARGETS:
        raise InvalidArgument(ERR_UNSUPPORTED, f"direct_sum_noalloc takes at most {DS_NOALLOC_MAX_SOURCES} sources "
                              f"and {DS_NOALLOC_MAX_TARGETS} targets")
    out = np.zeros((len(targets), 3), np.float32)
    b = None if box is None else _f32(box)
    rc = L.pmgpu_direct_sum_noalloc(len(mass), _fptr(pos), _fptr(mass), len(targets), _fptr(targets),
                                    float(softening), _fptr(b), _fptr(out))
    if rc != OK:
        _raise(rc, "pmgpu_direct_sum_noalloc failed")
    return out


def direct_sum_host(pos, mass, targets, softening=0.0, box=None, threads=None):
    """The same direct sum on the host CPU in FP64 (float64 (nt, 3)), over
    `threads` threads (default: the CPUs this process may run on)."""
    L = load_library()
    pos = _f32(pos, (-1, 3))
    mass = _f32(mass)
    targets = _f32(targets, (-1, 3))
    if threads is None:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    out = np.zeros((len(targets), 3), np.float64)
    b = None if box is None else _f32(box)
    rc = L.pmgpu_direct_sum_host(len(mass), _fptr(pos), _fptr(mass), len(targets), _fptr(targets),
                                 float(softening), _fptr(b), int(threads),
                                 out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != OK:
        _raise(rc, "pmgpu_direct_sum_host failed")
    return out


class Trainer:
    """Snapshots of one grid configuration for training the correction."""

    def __init__(self, grid, interpolation=TSC, finite_diff=FOUR_POINT, greens_function=DISCRETE_LAPLACIAN,
                 particle_diameter=0.0):
        L = load_library()
        p = PMParams()
        p.grid[:] = [int(g) for g in grid]
        p.interpolation = int(interpolation)
        p.finite_diff = int(finite_diff)
        p.greens_function = int(greens_function)
        p.particle_diameter = float(particle_diameter)
        p.box[:] = [0.0, 0.0, 0.0]
        p.H = p.DT = 1.0
        self._lib = L
        self.grid = tuple(int(g) for g in grid)
        self.interpolation = int(interpolation)
        self._t = C.c_void_p()
        rc = L.pmgpu_trainer_create(C.byref(p), C.byref(self._t))
        if rc != OK:
            _raise(rc, L.pmgpu_last_error(None).decode())
        self.snapshots = 0

    def _check(self, rc):
        if rc != OK:
            _raise(rc, self._lib.pmgpu_trainer_last_error(self._t).decode())

    def add_snapshot(self, pos, mass, target_index, target_acc):
        pos, mass = _f32(pos, (-1, 3)), _f32(mass)
        idx = np.ascontiguousarray(target_index, dtype=np.int64)
        acc = _f32(target_acc, (-1, 3))
        self._check(self._lib.pmgpu_trainer_add_snapshot(
            self._t, len(mass), _fptr(pos), _fptr(mass), len(idx), idx.ctypes.data_as(C.POINTER(C.c_int64)),
            _fptr(acc)))
        self.snapshots += 1

    def loss_grad(self, params=None, want_grad=True, want_grad_c=False):
        p = None if params is None else _f32(params)
        loss = C.c_double(0.0)
        g = np.zeros(mlp_num_params(), np.float32) if want_grad else None
        nx, ny, nz = self.grid
        gc = np.zeros((nz, ny, nx // 2 + 1), np.float32) if want_grad_c else None
        self._check(self._lib.pmgpu_trainer_loss_grad(self._t, _fptr(p), C.byref(loss), _fptr(g), _fptr(gc)))
        return loss.value, g, gc

    def close(self):
        if self._t:
            self._lib.pmgpu_trainer_destroy(self._t)
            self._t = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train(trainers, params, steps, lr=1e-3):
    """Adam on the joint loss of the trainers; returns (params, loss history)."""
    L = load_library()
    p = _f32(params).copy()
    hist = np.zeros(steps, np.float32)
    arr = (C.c_void_p * len(trainers))(*[t._t for t in trainers])
    rc = L.pmgpu_train(arr, len(trainers), _fptr(p), int(steps), float(lr), _fptr(hist))
    if rc != OK:
        _raise(rc, L.pmgpu_trainer_last_error(trainers[0]._t).decode())
    return p, hist


# ---------------------------------------------------------------------------
# ensembles
# ---------------------------------------------------------------------------

def pm_params(grid, interpolation=CIC, finite_diff=TWO_POINT, greens_function=DISCRETE_LAPLACIAN,
              particle_diameter=0.0, box=(0.0, 0.0, 0.0)):
    p = PMParams()
    p.grid[:] = [int(g) for g in grid]
    p.interpolation = int(interpolation)
    p.finite_diff = int(finite_diff)
    p.greens_function = int(greens_function)
    p.particle_diameter = float(particle_diameter)
    p.box[:] = [float(b) for b in box]
    p.ext_kind = EXT_NONE
    p.H = p.DT = 1.0
    return p


def p3m_params(cutoff, particle_diameter, softening=0.0, cloud_shape=S2, use_table=2, box=(0.0, 0.0, 0.0)):
    q = P3MParams()
    q.cutoff = float(cutoff)
    q.particle_diameter = float(particle_diameter)
    q.softening = float(softening)
    q.cloud_shape = int(cloud_shape)
    q.use_table = int(use_table)
    q.box[:] = [float(b) for b in box]
    return q


class Batch:
    """Independent simulations stepped together by one graph launch per two steps."""

    def __init__(self, members, lanes=8):
        """members: list of (PMParams, P3MParams or None)."""
        L = load_library()
        n = len(members)
        pa = (PMParams * n)(*[m[0] for m in members])
        qa = (P3MParams * n)(*[m[1] if m[1] is not None else P3MParams() for m in members])
        use = (C.c_int32 * n)(*[1 if m[1] is not None else 0 for m in members])
        self._lib = L
        self._b = C.c_void_p()
        rc = L.pmgpu_batch_create(n, pa, qa, use, int(lanes), C.byref(self._b))
        if rc != OK:
            _raise(rc, "pmgpu_batch_create failed")
        self.size = n
        self.grids = [tuple(m[0].grid) for m in members]

    def _check(self, rc):
        if rc != OK:
            _raise(rc, self._lib.pmgpu_batch_last_error(self._b).decode())

    def set_particles(self, i, pos, vel, mass):
        pos, mass = _f32(pos, (-1, 3)), _f32(mass)
        vel = None if vel is None else _f32(vel, (-1, 3))
        self._check(self._lib.pmgpu_batch_set_particles(self._b, int(i), len(mass), _fptr(pos), _fptr(vel),
                                                        _fptr(mass)))

    def prepare(self, autotune=True):
        self._check(self._lib.pmgpu_batch_prepare(self._b, int(bool(autotune))))

    def step(self, steps):
        self._check(self._lib.pmgpu_batch_step(self._b, int(steps)))

    def synchronize(self):
        self._check(self._lib.pmgpu_batch_synchronize(self._b))

    def time(self, steps):
        ms = C.c_double(0.0)
        self._check(self._lib.pmgpu_batch_time(self._b, int(steps), C.byref(ms)))
        return ms.value

    def member_latency(self, reps=5):
        out = (C.c_double * self.size)()
        self._check(self._lib.pmgpu_batch_member_latency(self._b, int(reps), out))
        return np.array(out[:])

    def get_particles(self, i, n):
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        acc = np.zeros((n, 3), np.float32)
        self._check(self._lib.pmgpu_batch_get_particles(self._b, int(i), _fptr(pos), _fptr(vel), _fptr(acc)))
        return pos, vel, acc

    def tuning(self, i):
        k = (C.c_int32 * 6)()
        self._check(self._lib.pmgpu_batch_get_tuning(self._b, int(i), k))
        return tuple(k)

    def close(self):
        if self._b:
            self._lib.pmgpu_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
