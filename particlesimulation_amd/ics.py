"""Seeded initial conditions and workload definitions.

Every generator takes a numpy Generator seed and returns original-unit
positions (n, 3), velocities (n, 3) and masses (n,), float64.  Sampling
follows the distributions the reference samples in
source/plummerSampler.cpp:29-83 (Plummer positions by the inverse mass
profile, speeds from the isotropic distribution function q^2 (1-q^2)^(7/2)
times the escape speed) and an exponential disk with circular velocities, but
uses numpy's PCG64 stream instead of std::default_random_engine, whose
sequence differs between standard libraries (source/utils.cpp:7).

`to_code_units` applies include/unitConversions.h:8-46:
x' = x/H, v' = v DT/H, m' = 4 pi G DT^2 / H^3 m.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def to_code_units(pos, vel, mass, H, DT, G):
    pos_c = np.asarray(pos, np.float64) / H
    vel_c = np.asarray(vel, np.float64) * DT / H
    mass_c = np.asarray(mass, np.float64) * (DT * DT * 4.0 * np.pi * G / (H * H * H))
    return pos_c.astype(np.float32), vel_c.astype(np.float32), mass_c.astype(np.float32)


def acc_to_original_units(acc, H, DT):
    return np.asarray(acc, np.float64) * H / (DT * DT)


def _isotropic(rng, n):
    cos_t = 1.0 - 2.0 * rng.random(n)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    phi = 2.0 * np.pi * rng.random(n)
    return np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], axis=1)


def plummer(n, seed=42, a=1.0, M=1.0, G=1.0, center=(0.0, 0.0, 0.0), r_max=None):
    """Plummer sphere in isotropic equilibrium (the reference's PlummerSampler)."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    u = np.clip(u, 1e-12, 1.0 - 1e-12)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    if r_max is not None:
        r = np.minimum(r, r_max)  # plummerSampler.cpp:50-52 clamps r to rMax
    pos = np.asarray(center, np.float64) + r[:, None] * _isotropic(rng, n)
    # speed: q = v / v_esc with density q^2 (1 - q^2)^(7/2) (rejection sampling)
    q = np.empty(n)
    todo = np.arange(n)
    gmax = 0.0921  # max of q^2 (1-q^2)^3.5 is ~0.0920
    while todo.size:
        x = rng.random(todo.size)
        y = gmax * rng.random(todo.size)
        ok = y < x * x * (1.0 - x * x) ** 3.5
        q[todo[ok]] = x[ok]
        todo = todo[~ok]
    v_esc = np.sqrt(2.0 * G * M / np.sqrt(r * r + a * a))
    vel = (q * v_esc)[:, None] * _isotropic(rng, n)
    mass = np.full(n, M / n)
    return pos, vel, mass


def exponential_disk(n, seed=42, r_d=1.0, h=0.1, M=1.0, point_mass_frac=0.1, G=1.0,
                     center=(0.0, 0.0, 0.0), r_max=12.0, z_max=1.0, dispersion=0.05):
    """Exponential disk (scale length r_d, sech^2 height h) plus a central point
    mass carrying `point_mass_frac` of the total mass M; particle 0 is the point
    mass.  Circular velocities from the enclosed mass (spherical estimate) plus
    a Gaussian dispersion of `dispersion` times v_c per component."""
    rng = np.random.default_rng(seed)
    nd = n - 1
    m_pm = point_mass_frac * M
    m_disk = M - m_pm
    R = np.empty(nd)
    todo = np.arange(nd)
    while todo.size:  # surface density ~ exp(-R/r_d): R ~ Gamma(2, r_d), truncated
        s = -r_d * np.log(rng.random(todo.size) * rng.random(todo.size) + 1e-300)
        ok = s < r_max
        R[todo[ok]] = s[ok]
        todo = todo[~ok]
    z = np.empty(nd)
    todo = np.arange(nd)
    while todo.size:  # sech^2(z/h)
        s = h * np.arctanh(2.0 * rng.random(todo.size) - 1.0)
        ok = np.abs(s) < z_max
        z[todo[ok]] = s[ok]
        todo = todo[~ok]
    phi = 2.0 * np.pi * rng.random(nd)
    x, y = R * np.cos(phi), R * np.sin(phi)
    r3 = np.sqrt(R * R + z * z)
    m_enc = m_pm + m_disk * (1.0 - np.exp(-r3 / r_d) * (1.0 + r3 / r_d))
    vc = np.sqrt(G * m_enc / np.maximum(r3, 1e-6))
    vel = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(nd)], axis=1)
    vel += dispersion * vc[:, None] * rng.standard_normal((nd, 3))
    pos = np.concatenate([np.zeros((1, 3)), np.stack([x, y, z], axis=1)]) + np.asarray(center, np.float64)
    vel = np.concatenate([np.zeros((1, 3)), vel])
    mass = np.concatenate([[m_pm], np.full(nd, m_disk / nd)])
    return pos, vel, mass


def uniform_with_clumps(n, seed=42, box=1.0, clump_frac=0.3, n_clumps=64, sigma=0.01, M=1.0):
    """70 % uniform in a periodic box, 30 % in Gaussian clumps with uniform
    centres, wrapped periodically; cold start."""
    rng = np.random.default_rng(seed)
    nc = int(round(clump_frac * n))
    nu = n - nc
    pos_u = rng.random((nu, 3)) * box
    centres = rng.random((n_clumps, 3)) * box
    which = rng.integers(0, n_clumps, nc)
    pos_c = centres[which] + sigma * rng.standard_normal((nc, 3))
    pos = np.mod(np.concatenate([pos_u, pos_c]), box)
    vel = np.zeros_like(pos)
    mass = np.full(n, M / n)
    return pos, vel, mass


@dataclass
class Workload:
    """One configuration of the force step, in original units."""
    name: str
    n: int
    grid: tuple
    box: float
    interpolation: int
    finite_diff: int
    greens_function: int
    dt: float
    steps: int
    G: float = 1.0
    p3m: bool = False
    a_cells: float = 0.0       # particle diameter a in grid cells
    rc_cells: float = 0.0      # short-range cutoff in grid cells
    softening: float = 0.0
    cloud_shape: int = 1       # S2
    sr_table: int = 2          # short-range pair force: 0 direct, 1 reference table, 2 accurate table
    periodic: bool = False
    kind: str = "plummer"
    extra: dict = field(default_factory=dict)

    @property
    def H(self):
        return self.box / self.grid[0]

    def sample(self, seed=42, n=None):
        n = self.n if n is None else n
        c = self.box / 2.0
        if self.kind == "plummer":
            return plummer(n, seed, a=1.0, M=1.0, G=self.G, center=(c, c, c), r_max=0.45 * self.box)
        if self.kind == "disk":
            return exponential_disk(n, seed, G=self.G, center=(c, c, c), r_max=min(12.0, 0.4 * self.box))
        if self.kind == "clumps":
            return uniform_with_clumps(n, seed, box=self.box)
        raise ValueError(self.kind)

    def code_units(self, seed=42, n=None):
        pos, vel, mass = self.sample(seed, n)
        return to_code_units(pos, vel, mass, self.H, self.dt, self.G)


# numbering matches the configuration list the benchmark is quoted on
C1_PM_PLUMMER = Workload("pm_plummer_32k_64", 32768, (64, 64, 64), 8.0, 1, 0, 0, 0.01, 100, kind="plummer")
C2_P3M_DISK = Workload("p3m_disk_1M_128", 1048576, (128, 128, 128), 32.0, 2, 0, 2, 0.01, 50, p3m=True,
                       a_cells=3.0, rc_cells=3.0, kind="disk")
C3_PM_CLUMPS = Workload("pm_clumps_16M_512", 16777216, (512, 512, 512), 1.0, 2, 1, 0, 1e-3, 20, kind="clumps",
                        periodic=True)
# the same disk and grid at the smallest cutoff (in half cells) whose force
# error against an FP64 direct sum (65 536 seeded targets) is below 1e-3:
# the textbook (centred) TSC, the 4-point gradient and the S2-optimal G with
# a = rc = 3.5 cells (7.7e-4; 3 cells give 1.6e-3) ...
C2_P3M_DISK_ERR = Workload("p3m_disk_1M_128_err1e-3", 1048576, (128, 128, 128), 32.0, 3, 1, 2, 0.01, 50,
                           p3m=True, a_cells=3.5, rc_cells=3.5, kind="disk")
# ... and with the reference's own scheme (floor-anchored TSC, 2-point
# gradient): a = rc = 5 cells (7.9e-4; 4 cells give 1.7e-3)
C2_P3M_DISK_ERR_REF = Workload("p3m_disk_1M_128_err1e-3_reference_tsc", 1048576, (128, 128, 128), 32.0, 2, 0, 2,
                               0.01, 50, p3m=True, a_cells=5.0, rc_cells=5.0, kind="disk")
# the learned-correction workload: plain PM uses the discrete-Laplacian G, the
# learned PM multiplies it by the network's factor
C5_PM_CLUMPS_256 = Workload("pm_clumps_4M_256", 4194304, (256, 256, 256), 1.0, 2, 1, 0, 1e-3, 20,
                            kind="clumps", periodic=True)
WORKLOADS = {w.name: w for w in (C1_PM_PLUMMER, C2_P3M_DISK, C2_P3M_DISK_ERR, C2_P3M_DISK_ERR_REF, C3_PM_CLUMPS,
                                 C5_PM_CLUMPS_256)}
