"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (FP32 device arithmetic vs the FP64 / FP32 oracle):
  density        max|rho_gpu - rho_oracle| / max|rho_oracle| <= 1e-5
  acceleration   RMS relative (||a_gpu - a_oracle|| / ||a_oracle||) <= 1e-4
  G(k)           relative 1e-6 (FP64 on both sides, rounded to FP32)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY_TOL = 1e-5
FORCE_TOL = 1e-4


def _rng(s):
    return np.random.default_rng(s)


def _particles(n, grid, seed, margin=2.0, clustered=False):
    r = _rng(seed)
    g = np.array(grid, np.float64)
    if clustered:
        c = g / 2
        pos = c + r.standard_normal((n, 3)) * (g.min() / 12)
        pos = np.clip(pos, margin, g - margin)
    else:
        pos = r.uniform(margin, g - margin, (n, 3))
    vel = r.standard_normal((n, 3)) * 0.05
    mass = r.uniform(0.5, 1.5, n) / n
    return pos.astype(np.float32), vel.astype(np.float32), mass.astype(np.float32)


def _make(P, grid, scheme, fds, gf, a=3.0, box=None):
    return P.PM(grid, interpolation=scheme, finite_diff=fds, greens_function=gf, particle_diameter=a,
                box=box if box is not None else grid)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


# --------------------------------------------------------------------------
# deposit
# --------------------------------------------------------------------------

@pytest.mark.parametrize("grid", [(32, 32, 32), (64, 32, 128)])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_deposit_matches_oracle(gpu_lib, oracle, grid, scheme):
    pos, vel, m = _particles(4096, grid, 1 + scheme)
    pm = _make(gpu_lib, grid, scheme, 0, 0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.pm_step()
    rho = pm.density()
    ref = oracle.deposit(pos, m, grid, scheme, wrap=False)
    assert np.max(np.abs(rho - ref)) / np.max(np.abs(ref)) <= DENSITY_TOL


@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_deposit_wraps_periodically(gpu_lib, oracle, scheme):
    grid = (32, 32, 32)
    r = _rng(3)
    pos = r.uniform(-0.9, 32.9, (3000, 3)).astype(np.float32)
    pos[:10, 0] = [0.0, 31.999, 32.0, -0.5, 15.0, 16.0, 0.5, 31.5, 1e-6, 24.0]
    m = r.uniform(0.5, 1.5, 3000).astype(np.float32)
    pm = _make(gpu_lib, grid, scheme, 0, 0, box=(0, 0, 0))
    pm.set_particles(pos, None, m)
    pm.init_greens()
    pm.pm_step()
    ref = oracle.deposit(pos, m, grid, scheme, wrap=True)
    rho = pm.density()
    assert np.max(np.abs(rho - ref)) / np.max(np.abs(ref)) <= DENSITY_TOL


def test_deposit_heavy_tiles_split_over_work_items(gpu_lib, oracle):
    # 60k particles inside two tiles: exercises the multi-work-item fixed-point path
    grid = (64, 64, 64)
    r = _rng(4)
    pos = (np.array([20.0, 12.0, 9.0]) + r.uniform(0, 1, (60000, 3)) * [14, 6, 2]).astype(np.float32)
    m = r.uniform(0.5, 1.5, 60000).astype(np.float32)
    pm = _make(gpu_lib, grid, 2, 0, 0)
    pm.set_particles(pos, None, m)
    pm.init_greens()
    pm.pm_step()
    ref = oracle.deposit(pos, m, grid, 2)
    assert np.max(np.abs(pm.density() - ref)) / np.max(np.abs(ref)) <= DENSITY_TOL


def test_deposit_is_bitwise_reproducible_and_order_independent(gpu_lib):
    grid = (64, 64, 64)
    pos, vel, m = _particles(50000, grid, 5, clustered=True)
    out = []
    for perm in (np.arange(len(m)), np.arange(len(m)), _rng(6).permutation(len(m))):
        pm = _make(gpu_lib, grid, 2, 0, 0)
        pm.set_particles(pos[perm], vel[perm], m[perm])
        pm.init_greens()
        pm.pm_step()
        out.append(pm.density())
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])


# --------------------------------------------------------------------------
# Green's function
# --------------------------------------------------------------------------

@pytest.mark.parametrize("gf,scheme,fds", [(0, 1, 0), (3, 1, 0), (1, 2, 0), (2, 2, 1), (2, 1, 0), (1, 0, 1)])
def test_greens_matches_oracle(gpu_lib, oracle, gf, scheme, fds):
    grid = (32, 64, 32)
    pm = _make(gpu_lib, grid, scheme, fds, gf, a=2.5)
    pm.init_greens()
    g = pm.get_grid(gpu_lib.GRID_GREENS)  # [kz, ky, kx<=nx/2], symmetrised
    G = oracle.green_grid(gf, scheme, fds, 2.5, grid)
    Gm = np.roll(np.flip(G, axis=(0, 1, 2)), 1, axis=(0, 1, 2))  # G(-k mod N)
    sym = 0.5 * (G + Gm)[:, :, : grid[0] // 2 + 1]
    assert np.max(np.abs(g - sym)) <= 2e-6 * np.max(np.abs(sym))


# --------------------------------------------------------------------------
# full PM step
# --------------------------------------------------------------------------

COMBOS = [(s, f, g) for s in (0, 1, 2) for f in (0, 1) for g in (0, 1, 2, 3)] + [(3, 0, 2), (3, 1, 2), (3, 1, 0)]


@pytest.mark.parametrize("scheme,fds,gf", COMBOS)
def test_pm_step_matches_oracle(gpu_lib, oracle, scheme, fds, gf):
    grid = (32, 32, 64)
    pos, vel, m = _particles(3000, grid, 10 + scheme + 3 * fds + 7 * gf, clustered=(gf % 2 == 0))
    pm = _make(gpu_lib, grid, scheme, fds, gf, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.pm_step()
    _, _, acc = pm.get_particles()
    G = oracle.green_grid(gf, scheme, fds, 3.0, grid)
    ref, rho, phi, E = oracle.pm_accel(pos, m, grid, scheme, fds, G, wrap=False, want_grids=True)
    assert _rel(acc, ref) <= FORCE_TOL
    assert _rel(pm.potential(), phi) <= FORCE_TOL
    assert _rel(pm.get_grid(gpu_lib.GRID_FIELD), E) <= FORCE_TOL


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_centred_tsc_interpolation_variants_match_oracle(gpu_lib, oracle, variant):
    # the textbook TSC through each interpolation kernel (tiled, field
    # gather, potential gather) and a run of the fused kick/drift
    grid = (64, 64, 64)
    pos, vel, m = _particles(5000, grid, 61, clustered=True)
    pm = _make(gpu_lib, grid, gpu_lib.TSC_CENTRED, 1, 2, a=3.0)
    pm.set_tuning(interp_kernel=variant)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.pm_step()
    acc = pm.get_particles()[2]
    G = oracle.green_grid(2, 3, 1, 3.0, grid)
    assert _rel(acc, oracle.pm_accel(pos, m, grid, 3, 1, G)) <= FORCE_TOL


@pytest.mark.parametrize("grid,scheme,fds,gf", [((60, 60, 30), 1, 0, 0), ((68, 68, 34), 2, 1, 2), ((36, 44, 52), 0, 0, 3),
                                               ((45, 33, 21), 2, 0, 1), ((48, 64, 32), 3, 1, 2), ((16, 16, 16), 2, 0, 0)])
def test_general_grid_matches_oracle(gpu_lib, oracle, grid, scheme, fds, gf):
    # sides that are not powers of two in [32, 1024] (the reference's accuracy
    # demo sweeps 2Ng x 2Ng x Ng, Ng = 30 .. 70, source/demos.cpp:381-408)
    # take the untiled path: global fixed-point deposit, direct DFTs, gather
    pos, vel, m = _particles(3000, grid, 70 + sum(grid), clustered=True, margin=3.0)
    pm = _make(gpu_lib, grid, scheme, fds, gf, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.pm_step()
    acc = pm.get_particles()[2]
    G = oracle.green_grid(gf, scheme, fds, 3.0, grid)
    ref, rho, phi, _ = oracle.pm_accel(pos, m, grid, scheme, fds, G, wrap=False, want_grids=True)
    assert np.max(np.abs(pm.density() - rho)) / np.max(np.abs(rho)) <= DENSITY_TOL
    assert _rel(pm.potential(), phi) <= FORCE_TOL
    assert _rel(acc, ref) <= FORCE_TOL
    # the density is bitwise reproducible and the run loop follows the oracle
    rho0 = pm.density()
    pm.pm_step()
    assert np.array_equal(pm.density(), rho0)
    pm.set_particles(pos, vel, m)
    assert pm.run(4, initial=True) == 4
    x = pm.get_particles()[0]
    xr, _, _ = _oracle_run(oracle, pos, vel, m, grid, scheme, fds, G, 4)
    assert _rel(x, xr) <= 1e-5


def test_general_grid_p3m_matches_oracle(gpu_lib, oracle):
    grid = (60, 60, 30)
    H = 0.25
    pos, vel, m = _particles(4000, grid, 77, clustered=True, margin=4.0)
    pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1, use_table=1)
    pm.pm_step()
    pm.p3m_short_range()
    acc = pm.get_particles()[2]
    G = oracle.green_grid(2, 2, 0, 3.0, grid)
    ref_sr, _ = oracle.p3m_short_range(pos, m, np.array(grid, np.float32) * H, 3 * H, H, 3.0, 3.0, 0.0, 1,
                                       use_table=1)
    assert _rel(acc, oracle.pm_accel(pos, m, grid, 2, 0, G) + ref_sr) <= FORCE_TOL


def test_external_field_matches_oracle(gpu_lib, oracle):
    grid = (32, 32, 32)
    H, DT = 0.5, 0.1
    ext = (8.0, 8.0, 8.0, 1.5, 3.0, 0.7)  # sphRadDecrField(center, R, M, G), original units
    pos, vel, m = _particles(2000, grid, 20)
    pm = gpu_lib.PM(grid, interpolation=2, finite_diff=0, greens_function=0, ext_field=ext, H=H, DT=DT)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.pm_step()
    _, _, acc = pm.get_particles()
    G = oracle.green_grid(0, 2, 0, 0.0, grid)
    ref = oracle.pm_accel(pos, m, grid, 2, 0, G, ext_prm=list(ext) + [H, DT])
    assert _rel(acc, ref) <= FORCE_TOL


# --------------------------------------------------------------------------
# P3M short range
# --------------------------------------------------------------------------

@pytest.mark.parametrize("cs,use_table,eps,a", [(1, 1, 0.0, 3.0), (0, 1, 0.0, 3.0), (1, 0, 0.0, 3.0), (1, 1, 0.3, 3.0),
                                                 (1, 2, 0.0, 3.0), (0, 2, 0.3, 3.0), (1, 2, 0.0, 2.2),
                                                 (0, 2, 0.0, 3.6)])
def test_p3m_matches_oracle(gpu_lib, oracle, cs, use_table, eps, a):
    # use_table 2 bins r^2 itself when the force vanishes inside the cutoff
    # (no softening, a <= rc) and r^2 / rc^2 otherwise (softened, or a > rc)
    grid = (64, 64, 64)
    H = 0.25
    pos, vel, m = _particles(6000, grid, 30 + cs, clustered=True)
    rc = 3.0
    gf = 1 + cs
    pm = _make(gpu_lib, grid, 2, 0, gf, a=a)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=rc, particle_diameter=a, softening=eps, cloud_shape=cs, use_table=use_table)
    pm.p3m_record_short_range(True)
    pm.pm_step()
    pm.p3m_short_range()
    _, _, acc = pm.get_particles()
    a_sr = pm.p3m_short_range_accelerations()
    G = oracle.green_grid(gf, 2, 0, a, grid)
    ref_pm = oracle.pm_accel(pos, m, grid, 2, 0, G)
    box_orig = np.array(grid, np.float32) * H
    ref_sr, pairs = oracle.p3m_short_range(pos, m, box_orig, rc * H, H, rc, a, eps, cs, use_table=use_table)
    ref = ref_pm + ref_sr
    assert _rel(acc, ref) <= FORCE_TOL
    assert _rel(acc - ref_pm, ref_sr) <= 10 * FORCE_TOL  # the short-range part alone
    assert _rel(a_sr, ref_sr) <= 10 * FORCE_TOL  # as recorded for the drop-in's Particle records
    # r^2 may round differently (FMA) for pairs within an ulp of the cutoff
    assert abs(pm.p3m_count_pairs() - pairs) <= 2 + 1e-5 * pairs


def test_p3m_clamped_table_layout_matches_oracle(gpu_lib, oracle):
    """The accurate table over r^2 comes whole (every r^2 of the clamped
    positions has its entry, no clamp in the pair loop) while the particle
    box is at most about 73 particle diameters (the table then spans at most
    SR_TABLE_EXT entries); a 128-cell box with a = rc = 1.5 cells (85
    diameters) takes the clamped absolute layout instead."""
    grid = (128, 128, 128)
    H = 0.25
    pos, vel, m = _particles(6000, grid, 43, clustered=True)
    rc = a = 1.5
    pm = _make(gpu_lib, grid, 2, 0, 2, a=a)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=rc, particle_diameter=a, cloud_shape=1, use_table=2)
    pm.p3m_record_short_range(True)
    pm.pm_step()
    pm.p3m_short_range()
    a_sr = pm.p3m_short_range_accelerations()
    ref_sr, pairs = oracle.p3m_short_range(pos, m, np.array(grid, np.float32) * H, rc * H, H, rc, a, 0.0, 1,
                                           use_table=2)
    assert _rel(a_sr, ref_sr) <= 10 * FORCE_TOL
    assert abs(pm.p3m_count_pairs() - pairs) <= 2 + 1e-5 * pairs


@pytest.mark.parametrize("common_mass", [False, True])
def test_p3m_pair_table_layouts_agree_up_to_the_mass_scale(gpu_lib, oracle, common_mass):
    # The accurate table over r^2 is read in its absolute layout (positions
    # scaled by 2^-q, masses by 2^(q+t); here the whole table, no clamp)
    # unless the scaled masses could overflow; then the same bins are read
    # wrapped and unscaled.  Masses
    # times 2^60 take the second layout.  Every product of the two differs by
    # a power of two, so the accelerations scale by 2^60 up to the order of
    # the sums (the cell sort's order within a cell varies from run to run)
    # and, with the common-mass loop, the gradual underflow of negligible
    # terms (its unit-mass sums run at about 2^-94).
    grid = (64, 64, 64)
    pos, vel, m = _particles(6000, grid, 33, clustered=True)
    if common_mass:
        m = np.full_like(m, 1.0 / len(m))
    out = []
    for scale in (1.0, 2.0 ** 60):
        pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
        pm.set_particles(pos, vel, (m * np.float32(scale)).astype(np.float32))
        pm.init_greens()
        pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1, use_table=2)
        pm.p3m_record_short_range(True)
        pm.pm_step()
        pm.p3m_short_range()
        out.append((pm.get_particles()[2].astype(np.float64) / scale, pm.p3m_short_range_accelerations().astype(np.float64) / scale))
        pm.close()
    (a1, s1), (a2, s2) = out
    assert _rel(s2, s1) <= 1e-6
    assert _rel(a2, a1) <= 1e-6
    # and both against the oracle
    H = 0.25
    ref_sr, _ = oracle.p3m_short_range(pos, m, np.array(grid, np.float32) * H, 3.0 * H, H, 3.0, 3.0, 0.0, 1,
                                       use_table=2)
    assert _rel(s1, ref_sr) <= 10 * FORCE_TOL


@pytest.mark.parametrize("div,chunking", [(1, 2), (3, 2), (4, 1), (5, 2), (6, 1)])
def test_p3m_cell_division_keeps_the_pair_set(gpu_lib, oracle, div, chunking):
    # finer short-range cells with trimmed rows visit fewer candidates but must
    # accept exactly the same pairs and give the same force
    grid = (64, 64, 64)
    H = 0.25
    pos, vel, m = _particles(6000, grid, 31, clustered=True)
    pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.set_tuning(sr_div=2)  # explicit: p3m_configure would otherwise pick the finest division
    pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1, use_table=2)
    assert pm.get_tuning()[3] == 2
    pairs2, tested2 = pm.p3m_count_pairs(tested=True)
    pm.pm_step()
    pm.p3m_short_range()
    a2 = pm.get_particles()[2]
    k = list(pm.get_tuning())
    k[3] = div
    k[5] = chunking
    pm.set_tuning(*k)
    pairs_d, tested_d = pm.p3m_count_pairs(tested=True)
    assert pairs_d == pairs2
    if div > 2:
        assert tested_d < tested2
    pm.pm_step()
    pm.p3m_short_range()
    a_d = pm.get_particles()[2]
    assert _rel(a_d, a2) <= 1e-6
    G = oracle.green_grid(2, 2, 0, 3.0, grid)
    ref_sr, _ = oracle.p3m_short_range(pos, m, np.array(grid, np.float32) * H, 3 * H, H, 3.0, 3.0, 0.0, 1,
                                       use_table=2)
    assert _rel(a_d, oracle.pm_accel(pos, m, grid, 2, 0, G) + ref_sr) <= FORCE_TOL


@pytest.mark.parametrize("use_table,n_heavy,m0", [(0, 0, 1.0), (1, 0, 1.0), (2, 0, 1.0), (0, 5, 1.0), (1, 5, 1.0),
                                                  (2, 5, 1.0), (2, 64, 1.0), (2, 65, 1.0), (1, 64, 1.0),
                                                  (2, 40, 0.0), (0, 64, 0.0)])
def test_p3m_common_mass_matches_oracle(gpu_lib, oracle, use_table, n_heavy, m0):
    # particles of one mass m0 / n (but for a few heavy ones) take the pair
    # loop without a mass multiply and a per-i correction for the heavy ones;
    # 64 heavy particles is the most the correction list holds, 65 takes the
    # per-pair mass multiply, and a zero common mass leaves only the heavy
    # particles' forces.  The reference divides the pair forces by m_i
    # (correctAccelerations), which is 0/0 for a massless particle, so that
    # case is checked against an FP64 brute-force sum of m_j F(r) d_ij
    grid = (64, 64, 64)
    pos, vel, _ = _particles(6000, grid, 41, clustered=True)
    m = np.full(len(pos), np.float32(m0 / len(pos)))
    m[_rng(42).choice(len(pos), n_heavy, replace=False)] = np.float32(20.0 / len(pos))
    pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1, use_table=use_table)
    pm.pm_step()
    _, _, a_pm = pm.get_particles()
    pm.p3m_short_range()
    acc = pm.get_particles()[2]
    H = 0.25
    if m0 > 0:
        ref_sr, _ = oracle.p3m_short_range(pos, m, np.array(grid, np.float32) * H, 3 * H, H, 3.0, 3.0, 0.0, 1,
                                           use_table=use_table)
    else:
        x, mj = pos.astype(np.float64), m.astype(np.float64)
        ref_sr = np.zeros_like(x)
        for i in range(len(x)):
            d = x[i] - x
            r2 = np.einsum("ij,ij->i", d, d)
            sel = (r2 > 0.0) & (r2 < 9.0) & (mj > 0)
            F = oracle.sr_pair_factor(1, 3.0, 0.0, r2[sel])
            ref_sr[i] = ((mj[sel] * F)[:, None] * d[sel]).sum(0)
        assert np.count_nonzero(np.any(ref_sr != 0.0, axis=1)) > 100
    assert _rel(acc - a_pm, ref_sr) <= 10 * FORCE_TOL
    G = oracle.green_grid(2, 2, 0, 3.0, grid)
    assert _rel(acc, oracle.pm_accel(pos, m, grid, 2, 0, G) + ref_sr) <= FORCE_TOL


def test_p3m_anisotropic_cells_for_a_thin_disk(gpu_lib, oracle):
    # with the division left to p3m_configure, a thin disk gets cells finer in
    # the disk's plane and one cutoff deep across it: fewer candidates than
    # cubes of the same division, the same pair set and the same force
    grid = (64, 64, 64)
    r = _rng(33)
    n = 20000
    rad = r.exponential(5.0, n)
    th = r.uniform(0.0, 2 * np.pi, n)
    pos = np.stack([32 + rad * np.cos(th), 32 + rad * np.sin(th), 32 + 0.3 * r.standard_normal(n)], 1)
    pos = np.clip(pos, 2.0, 62.0).astype(np.float32)
    vel = np.zeros_like(pos)
    m = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    out = []
    for mode in ("auto", "explicit", "configured_first"):
        pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
        if mode == "configured_first":  # the cells follow the particles set later
            pm.set_particles(pos[:100], vel[:100], m[:100])
            pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1, use_table=2)
        pm.set_particles(pos, vel, m)
        pm.init_greens()
        if mode == "explicit":
            k = list(pm.get_tuning())
            k[3] = out[0][0]
            pm.set_tuning(*k)
        if mode != "configured_first":
            pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1, use_table=2)
        pairs, tested = pm.p3m_count_pairs(tested=True)
        pm.pm_step()
        pm.p3m_short_range()
        out.append((pm.get_tuning()[3], pairs, tested, pm.get_particles()[2], pm.p3m_cell_geometry()))
    (d, pairs_a, tested_a, a_auto, geo_a), (d2, pairs_i, tested_i, a_iso, geo_i), late = out
    assert d == d2 and pairs_a == pairs_i
    assert late[4] == geo_a and late[1] == pairs_a and late[2] == tested_a
    assert _rel(late[3], a_auto) <= 1e-5
    assert geo_i[1] == geo_i[2] == d  # explicit: cubes in y and z
    assert geo_a[1] > geo_a[2]  # the disk lies in x-y: finer in y, deeper in z
    assert tested_a < 0.95 * tested_i
    # the cells order each i's j-stream differently: FP32 rounding only
    assert _rel(a_auto, a_iso) <= 1e-5
    G = oracle.green_grid(2, 2, 0, 3.0, grid)
    H = 0.25
    ref_sr, _ = oracle.p3m_short_range(pos, m, np.array(grid, np.float32) * H, 3 * H, H, 3.0, 3.0, 0.0, 1,
                                       use_table=2)
    assert _rel(a_auto, oracle.pm_accel(pos, m, grid, 2, 0, G) + ref_sr) <= FORCE_TOL


def test_p3m_short_range_conserves_momentum_at_scale(gpu_lib):
    grid = (128, 128, 128)
    pos, vel, m = _particles(200000, grid, 40, clustered=True)
    pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1)
    pm.pm_step()
    _, _, a_pm = pm.get_particles()
    pm.p3m_short_range()
    _, _, a_tot = pm.get_particles()
    p_pm = (m[:, None].astype(np.float64) * a_pm).sum(0)
    p_sr = (m[:, None].astype(np.float64) * (a_tot - a_pm)).sum(0)
    assert np.max(np.abs(p_pm)) < 1e-4 * np.abs(m[:, None] * a_pm).sum()
    assert np.max(np.abs(p_sr)) < 1e-4 * np.abs(m[:, None] * (a_tot - a_pm)).sum()


# --------------------------------------------------------------------------
# leapfrog / run loop
# --------------------------------------------------------------------------

def _oracle_run(oracle, pos, vel, m, grid, scheme, fds, G, iters, p3m=None):
    x = pos.astype(np.float64).copy()
    v = vel.astype(np.float64).copy()

    def force(xx):
        a = oracle.pm_accel(xx.astype(np.float32), m, grid, scheme, fds, G)
        if p3m is not None:
            a = a + oracle.p3m_short_range(xx.astype(np.float32), m, *p3m)[0]
        return a
    a = force(x)
    v += 0.5 * a
    for _ in range(iters):
        x += v
        a = force(x)
        v += a
    return x, v, a


@pytest.mark.parametrize("p3m", [False, True])
def test_run_matches_oracle_leapfrog(gpu_lib, oracle, p3m):
    grid = (64, 64, 64)
    pos, vel, m = _particles(4000, grid, 50, margin=8.0, clustered=True)
    m = m * 20
    pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    sr = None
    if p3m:
        pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1)
        sr = (np.array(grid, np.float32), 3.0, 1.0, 3.0, 3.0, 0.0, 1)
    done = pm.run(4, initial=True, p3m=p3m)
    assert done == 4
    x, v, a = pm.get_particles()
    G = oracle.green_grid(2, 2, 0, 3.0, grid)
    rx, rv, ra = _oracle_run(oracle, pos, vel, m, grid, 2, 0, G, 4, p3m=sr)
    # an ordering mistake (a missing half kick, a kick before the force) gives
    # O(1) errors in v; FP32 rounding through four steps stays far below these
    assert _rel(x, rx) < 1e-5
    assert _rel(v, rv) < 1e-3
    assert _rel(a, ra) < 1e-2


@pytest.mark.parametrize("p3m", [False, True])
def test_graph_replayed_run_matches_the_direct_loop(gpu_lib, p3m):
    """With the stage events off the run loop replays two iterations at a time
    from a captured graph whose kernels read the iteration number from device
    memory; with them on it launches every kernel itself.  Two calls in a row
    (7 iterations, then 5 more) start the graph on both particle buffers and
    run an odd tail outside it."""
    grid = (64, 64, 64)
    pos, vel, m = _particles(4000, grid, 51, margin=8.0, clustered=True)
    m = m * 20
    out = []
    for profiled in (True, False):
        pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
        pm.set_particles(pos, vel, m)
        pm.init_greens()
        if p3m:
            pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1)
        pm.profile(profiled)
        assert pm.run(7, initial=True, p3m=p3m) == 7
        assert pm.run(5, initial=False, p3m=p3m) == 5
        assert not pm.escaped()
        out.append(pm.get_particles())
        pm.close()
    (x0, v0, a0), (x1, v1, a1) = out
    if p3m:
        # the short-range sums run over cells filled by atomics, so their
        # order (and the last bits) differ from run to run
        assert _rel(x1, x0) < 1e-6 and _rel(v1, v0) < 1e-4 and _rel(a1, a0) < 1e-3
    else:
        # fixed-point deposit, fixed FFT and per-particle interpolation: the
        # two loops compute the same bits
        assert np.array_equal(x1, x0) and np.array_equal(v1, v0) and np.array_equal(a1, a0)


def test_profile_levels_bracket_the_named_stages(gpu_lib):
    """Level 1 times every stage; level 2 only the short-range pair kernel,
    one launch per force evaluation of the run."""
    grid = (64, 64, 64)
    pos, vel, m = _particles(4000, grid, 52, margin=8.0, clustered=True)
    pm = _make(gpu_lib, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1)
    assert pm.run(2, initial=True, p3m=True) == 2
    seen = []
    for short_range_only in (True, False):
        pm.profile(True, short_range_only=short_range_only)
        pm.profile_reset()
        assert pm.run(5, initial=False, p3m=True) == 5
        st = pm.profile_read()
        pm.profile(False)
        seen.append({k for k, v in st.items() if v[1]})
        assert st["sr_pairs"][1] == 5 and st["sr_pairs"][0] > 0
    assert seen[0] == {"sr_pairs"}
    assert {"deposit", "interp", "sr_pairs", "sr_sort"} <= seen[1]
    pm.close()


def test_leapfrog_ops_match_reference_formulas(gpu_lib):
    grid = (32, 32, 32)
    pos, vel, m = _particles(1000, grid, 60)
    pm = _make(gpu_lib, grid, 1, 0, 0)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.pm_step()
    _, _, a = pm.get_particles()
    pm.leapfrog(gpu_lib.HALF_KICK, 1.0)
    _, v1, _ = pm.get_particles()
    assert np.allclose(v1, vel + 0.5 * a, atol=1e-7)
    pm.leapfrog(gpu_lib.INTEGER_STEP, 1.0)
    assert np.allclose(pm.integer_step_velocities(), v1 + 0.5 * a, atol=1e-7)
    pm.leapfrog(gpu_lib.KICK, 2.0)
    _, v2, _ = pm.get_particles()
    assert np.allclose(v2, v1 + 2.0 * a, atol=1e-7)
    pm.leapfrog(gpu_lib.DRIFT, 0.5)
    x2, _, _ = pm.get_particles()
    assert np.allclose(x2, pos + 0.5 * v2, atol=1e-6)


# the run loop replays iterations 1-2, 3-4, ... from one captured graph: the
# escape falls in the first or the second iteration of one replay; in a P3M
# run the short-range kernel does the kick and the drift
@pytest.mark.parametrize("p3m", [False, True])
@pytest.mark.parametrize("x0,escape_at", [(30.5, 3), (30.1, 4)])
def test_escape_stops_the_run(gpu_lib, x0, escape_at, p3m):
    grid = (32, 32, 32)
    pos, vel, m = _particles(500, grid, 70, margin=8.0)
    vel[:] = 0
    pos[7] = [x0, 16.0, 16.0]
    vel[7] = [0.4, 0.0, 0.0]  # crosses x = 32 during the drift of iteration escape_at
    # no forces to speak of: the trajectory is the free drift
    m[:] = 1e-12 if p3m else 0.0
    pm = _make(gpu_lib, grid, 1, 0, 0, box=grid)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    if p3m:
        pm.p3m_configure(cutoff=3.0, particle_diameter=3.0, cloud_shape=1)
    assert not pm.escaped()
    done = pm.run(10, initial=True, p3m=p3m)
    assert done == escape_at  # iterations before it are whole; its drift leaves the box
    x, v, _ = pm.get_particles()
    assert np.isclose(x[7, 0], x0 + (escape_at + 1) * 0.4, atol=1e-5)
    assert pm.escaped()


def test_empty_and_single_particle(gpu_lib, oracle):
    grid = (32, 32, 32)
    pm = _make(gpu_lib, grid, 2, 0, 0)
    pm.init_greens()
    with pytest.raises(gpu_lib.PMGPUError):
        pm.pm_step()
    pos = np.array([[16.25, 16.5, 16.75]], np.float32)
    m = np.array([1.0], np.float32)
    pm.set_particles(pos, None, m)
    pm.pm_step()
    assert np.isclose(pm.density().sum(), 1.0, rtol=1e-6)
    _, _, a = pm.get_particles()
    ref = oracle.pm_accel(pos, m, grid, 2, 0, oracle.green_grid(0, 2, 0, 0.0, grid))
    assert np.max(np.abs(a - ref)) <= FORCE_TOL * max(np.max(np.abs(ref)), 1e-3)


# --------------------------------------------------------------------------
# full-size properties (the 1M-particle P3M workload)
# --------------------------------------------------------------------------

def test_full_size_p3m_properties(gpu_lib):
    from particlesimulation_amd import ics
    w = ics.C2_P3M_DISK
    pos, vel, m = w.code_units(seed=42)
    pm = _make(gpu_lib, w.grid, w.interpolation, w.finite_diff, w.greens_function, a=w.a_cells,
               box=w.grid)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=w.rc_cells, particle_diameter=w.a_cells, cloud_shape=1)
    pm.pm_step()
    rho1 = pm.density()
    assert abs(rho1.astype(np.float64).sum() - m.astype(np.float64).sum()) < 1e-5 * m.sum()
    _, _, a_pm = pm.get_particles()
    pm.pm_step()
    assert np.array_equal(pm.density(), rho1)
    p = (m[:, None].astype(np.float64) * a_pm).sum(0)
    assert np.max(np.abs(p)) < 1e-4 * np.abs(m[:, None] * a_pm).sum()
    pairs = pm.p3m_count_pairs()
    assert pairs > 1e9


def test_full_size_short_range_matches_brute_force(gpu_lib, oracle):
    # the short-range pass on the whole 1M-particle disk (accurate table):
    # sampled targets against an FP64 brute-force sum over every particle,
    # and the pair forces' antisymmetry through the total momentum
    from particlesimulation_amd import ics
    w = ics.C2_P3M_DISK
    pos, vel, m = w.code_units(seed=42)
    pm = _make(gpu_lib, w.grid, w.interpolation, w.finite_diff, w.greens_function, a=w.a_cells,
               box=w.grid)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    pm.p3m_configure(cutoff=w.rc_cells, particle_diameter=w.a_cells, cloud_shape=w.cloud_shape,
                     use_table=2)
    pm.pm_step()
    _, _, a_pm = pm.get_particles()
    pm.p3m_short_range()
    _, _, a_tot = pm.get_particles()
    a_sr = a_tot.astype(np.float64) - a_pm
    p_sr = (m[:, None].astype(np.float64) * a_sr).sum(0)
    assert np.max(np.abs(p_sr)) < 1e-4 * np.abs(m[:, None] * a_sr).sum()
    x = pos.astype(np.float64)
    mj = m.astype(np.float64)
    idx = _rng(77).choice(len(pos), 64, replace=False)
    ref = np.zeros((len(idx), 3))
    rc2 = float(w.rc_cells) ** 2
    for k, i in enumerate(idx):
        d = x[i] - x
        r2 = np.einsum("ij,ij->i", d, d)
        sel = (r2 > 0.0) & (r2 < rc2)
        F = oracle.sr_pair_factor(w.cloud_shape, w.a_cells, w.softening, r2[sel])
        ref[k] = ((mj[sel] * F)[:, None] * d[sel]).sum(0)
    assert np.count_nonzero(np.any(ref != 0.0, axis=1)) > 32  # most targets have neighbours
    assert _rel(a_sr[idx], ref) <= FORCE_TOL


def test_p3m_rejects_a_cutoff_too_small_for_the_box(gpu_lib):
    # the short-range cell count is checked before it is formed (no overflow)
    P = gpu_lib
    grid = (64, 64, 64)
    pos, vel, m = _particles(1000, grid, 5)
    pm = _make(P, grid, 2, 0, 2, a=3.0)
    pm.set_particles(pos, vel, m)
    with pytest.raises(P.PMGPUError, match="too many short-range cells"):
        pm.p3m_configure(cutoff=1e-4, particle_diameter=3.0, cloud_shape=1)
