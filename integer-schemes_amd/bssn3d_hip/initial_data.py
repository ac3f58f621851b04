"""Initial data for the BSSN workloads (torch tensors, any device).

* minkowski / gauge_wave / hamiltonian_bump / robust_stability: the reference's test data
  (3D/src/bssn3d/initial_data.py:28-137).
* punctures: Brill-Lindquist punctures with seeded sinusoid perturbations, the workload data
  of the benchmark configurations. The reference has no puncture data
  (initial_data.py:1: "NOT BBH ID"), so this construction is defined here:
      psi   = 1 + sum_i m_i / (2 max(r_i, h/2))
      chi   = psi^-4,  alpha = psi^-2,  gt_ij = delta_ij,  every other field 0,
  then every field f gets amp * p_f(x) with
      p_f(x) = (1/8) sum_{n=1..8} sin(2 pi k_n . (x - x_lo) / L + phi_n),
  integer k_n with 0 < |k_n|_inf <= 4 and phi_n ~ U[0, 2 pi), drawn per field from
  numpy default_rng(seed) (L = interior extent; the modes are periodic on the domain).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from .grid import Grid
from .state import (ALPHA, AT0, AT3, AT5, CHI, GT0, GT3, GT5, GTILDE0, K, NUM_VARS)


def _coords(grid: Grid, device, dtype=torch.float64):
    x = torch.as_tensor(grid.x, dtype=dtype, device=device).view(-1, 1, 1)
    y = torch.as_tensor(grid.y, dtype=dtype, device=device).view(1, -1, 1)
    z = torch.as_tensor(grid.z, dtype=dtype, device=device).view(1, 1, -1)
    return x, y, z


def minkowski(grid: Grid, device="cpu") -> torch.Tensor:
    d = torch.zeros((NUM_VARS,) + grid.shape, dtype=torch.float64, device=device)
    for f in (ALPHA, CHI, GT0, GT3, GT5):
        d[f] = 1.0
    return d


def gauge_wave(grid: Grid, amplitude: float = 0.01, wavelength: float | None = None,
               axis: int = 0, t: float = 0.0, device="cpu") -> torch.Tensor:
    """Harmonic gauge wave along ``axis`` (initial_data.py:71-137, :140-196 for t != 0)."""
    coords1d = (grid.x, grid.y, grid.z)[axis]
    if wavelength is None:
        wavelength = float(coords1d[-grid.ng - 1] - coords1d[grid.ng]) + \
            (grid.dx, grid.dy, grid.dz)[axis]
    X, Y, Z = _coords(grid, device)
    coord = (X, Y, Z)[axis].expand(grid.shape)
    k = 2.0 * math.pi / wavelength
    phase = k * (coord - t)
    A = amplitude
    H = 1.0 - A * torch.sin(phase)
    dH_dx = -A * k * torch.cos(phase)
    dH_dt = -dH_dx
    sqrtH = torch.sqrt(H)
    Kxx = -dH_dt / (2.0 * sqrtH)
    Ktrace = Kxx / H
    chi = H ** (-1.0 / 3.0)
    gt_wave = H ** (2.0 / 3.0)
    gt_trans = chi
    At_wave = chi * (Kxx - (1.0 / 3.0) * H * Ktrace)
    At_trans = chi * (0.0 - (1.0 / 3.0) * 1.0 * Ktrace)
    Gt_wave = (2.0 / 3.0) * H ** (-5.0 / 3.0) * dH_dx
    d = torch.zeros((NUM_VARS,) + grid.shape, dtype=torch.float64, device=device)
    d[ALPHA] = sqrtH
    d[CHI] = chi
    d[K] = Ktrace
    gt_diag, at_diag = (GT0, GT3, GT5), (AT0, AT3, AT5)
    for a in range(3):
        d[gt_diag[a]] = gt_wave if a == axis else gt_trans
        d[at_diag[a]] = At_wave if a == axis else At_trans
    d[GTILDE0 + axis] = Gt_wave
    return d


def gowdy(grid: Grid, t: float = 1.0, device="cpu") -> torch.Tensor:
    """Polarized Gowdy T^3 vacuum cosmology at time t as BSSN variables, inhomogeneous in z,
    periodic on [-1/2, 1/2]^3 (initial_data.py:200-291):
        ds^2 = t^(-1/2) e^(L/2) (-dt^2 + dz^2) + t (e^P dx^2 + e^-P dy^2),
        P = J0(2 pi t) cos(2 pi z),
        L = -2 pi t J0 J1 cos^2(2 pi z) + 2 pi^2 t^2 (J0^2 + J1^2) - C,
    Jn = Jn(2 pi t), C fixed by t = 1; alpha = t^(-1/4) e^(L/4), zero shift,
    K_ij = -d_t gamma_ij / (2 alpha)."""
    from scipy.special import j0, j1

    tw = 2.0 * math.pi
    pi2 = math.pi ** 2
    b0, b1 = float(j0(tw * t)), float(j1(tw * t))
    c0, c1 = float(j0(tw)), float(j1(tw))
    C = 0.5 * (tw * tw * (c0 * c0 + c1 * c1) - tw * c0 * c1)
    _, _, Z = _coords(grid, device)
    z = Z.expand(grid.shape)
    cz = torch.cos(tw * z)
    s2z = torch.sin(2.0 * tw * z)
    P = b0 * cz
    P_t = -tw * b1 * cz
    lam = -tw * t * b0 * b1 * cz ** 2 + 2.0 * pi2 * t * t * (b0 * b0 + b1 * b1) - C
    lam_z = 4.0 * pi2 * t * b0 * b1 * s2z
    lam_t = 4.0 * pi2 * t * (b0 * b0 - (b0 * b0 - b1 * b1) * cz ** 2)
    eP = torch.exp(P)
    g_xx = t * eP
    g_yy = t / eP
    g_zz = t ** (-0.5) * torch.exp(0.5 * lam)
    alpha = t ** (-0.25) * torch.exp(0.25 * lam)
    Kxx = -(eP * (1.0 + t * P_t)) / (2.0 * alpha)
    Kyy = -((1.0 / eP) * (1.0 - t * P_t)) / (2.0 * alpha)
    Kzz = -(g_zz * (0.5 * lam_t - 0.5 / t)) / (2.0 * alpha)
    Ktr = Kxx / g_xx + Kyy / g_yy + Kzz / g_zz
    chi = (g_xx * g_yy * g_zz) ** (-1.0 / 3.0)
    d = torch.zeros((NUM_VARS,) + grid.shape, dtype=torch.float64, device=device)
    d[ALPHA] = alpha
    d[CHI] = chi
    d[K] = Ktr
    d[GT0] = chi * g_xx
    d[GT3] = chi * g_yy
    d[GT5] = chi * g_zz
    d[AT0] = chi * (Kxx - g_xx * Ktr / 3.0)
    d[AT3] = chi * (Kyy - g_yy * Ktr / 3.0)
    d[AT5] = chi * (Kzz - g_zz * Ktr / 3.0)
    d[GTILDE0 + 2] = (t / 3.0) * torch.exp(-lam / 3.0) * lam_z
    return d


def hamiltonian_bump(grid: Grid, amplitude: float = 0.01, device="cpu") -> torch.Tensor:
    """Minkowski with chi + A cos(2 pi x) cos(2 pi y) cos(2 pi z) (initial_data.py:33-55): a
    smooth, coherent Hamiltonian-constraint violation with no momentum violation, periodic on
    [-1/2, 1/2]^3."""
    X, Y, Z = _coords(grid, device)
    twopi = 2.0 * math.pi
    d = minkowski(grid, device)
    d[CHI] += amplitude * torch.cos(twopi * X) * torch.cos(twopi * Y) * torch.cos(twopi * Z)
    return d


def robust_stability(grid: Grid, amp: float = 1.0e-10, seed: int = 0,
                     device="cpu") -> torch.Tensor:
    """Minkowski + uniform noise in [-amp, amp) on all 24 fields (initial_data.py:58-69).
    The noise is drawn from numpy's default_rng(seed) (field-major, C order), not from JAX's
    PRNG stream; the property that matters (bounded round-off-scale noise) is the same."""
    noise = np.random.default_rng(seed).uniform(-amp, amp, size=(NUM_VARS,) + grid.shape)
    return minkowski(grid, device) + torch.from_numpy(noise).to(device)


@dataclass
class Perturbation:
    """Per-field sinusoid draws: k (24, 8, 3) integers, phi (24, 8)."""
    k: np.ndarray
    phi: np.ndarray

    @staticmethod
    def draw(rng: np.random.Generator, nmodes: int = 8, kmax: int = 4) -> "Perturbation":
        k = np.zeros((NUM_VARS, nmodes, 3), dtype=np.int64)
        for f in range(NUM_VARS):
            for n in range(nmodes):
                while True:
                    v = rng.integers(-kmax, kmax + 1, size=3)
                    if np.max(np.abs(v)) > 0:
                        break
                k[f, n] = v
        phi = rng.uniform(0.0, 2.0 * math.pi, size=(NUM_VARS, nmodes))
        return Perturbation(k, phi)


def _mode_sum(grid: Grid, pert: Perturbation, f: int, device, domain=None):
    X, Y, Z = _coords(grid, device)
    if domain is None:
        L = (grid.Nx * grid.dx, grid.Ny * grid.dy, grid.Nz * grid.dz)
        lo = (grid.xmin, grid.ymin, grid.zmin)
    else:
        lo, L = domain
    acc = torch.zeros(grid.shape, dtype=torch.float64, device=device)
    for n in range(pert.k.shape[1]):
        kx, ky, kz = (int(v) for v in pert.k[f, n])
        arg = (2.0 * math.pi * kx / L[0]) * (X - lo[0]) + (2.0 * math.pi * ky / L[1]) * (Y - lo[1]) \
            + (2.0 * math.pi * kz / L[2]) * (Z - lo[2]) + float(pert.phi[f, n])
        acc += torch.sin(arg)
    return acc / pert.k.shape[1]


def punctures(grid: Grid, masses: Sequence[float], positions: Sequence[Sequence[float]],
              amp: float = 1e-4, seed: int = 0, device="cpu", domain=None) -> torch.Tensor:
    """Brill-Lindquist punctures + seeded perturbations (see the module docstring). ``domain``
    = (lo, L) (3 coordinates each) sets the box the modes are periodic on (default: the grid's
    own interior), so that blocks of one hierarchy sample one smooth function."""
    X, Y, Z = _coords(grid, device)
    h = min(grid.dx, grid.dy, grid.dz)
    psi = torch.ones(grid.shape, dtype=torch.float64, device=device)
    for m, (px, py, pz) in zip(masses, positions):
        r = torch.sqrt((X - px) ** 2 + (Y - py) ** 2 + (Z - pz) ** 2)
        psi += m / (2.0 * torch.clamp(r, min=0.5 * h))
    d = torch.zeros((NUM_VARS,) + grid.shape, dtype=torch.float64, device=device)
    d[ALPHA] = psi ** -2
    d[CHI] = psi ** -4
    for f in (GT0, GT3, GT5):
        d[f] = 1.0
    del psi
    if amp != 0.0:
        pert = Perturbation.draw(np.random.default_rng(seed))
        for f in range(NUM_VARS):
            d[f] += amp * _mode_sum(grid, pert, f, device, domain)
    return d


def offgrid_shift(rng: np.random.Generator, h: float) -> np.ndarray:
    """A per-axis shift in [0.25 h, 0.75 h): keeps punctures at least h/4 off every grid plane."""
    return (0.5 + rng.uniform(-0.25, 0.25, size=3)) * h


# --- the workload configurations -------------------------------------------------------------

def config_single_puncture(seed: int = 0, N: int = 48, device="cpu", ng: int = 3):
    """48^3 interior, 3 ghosts, h = 0.25 on [-6, 6): one puncture m = 1 at a seeded off-grid
    position, 1e-4 perturbations. Returns (grid, state, positions)."""
    h = 12.0 / N
    grid = Grid(N, N, N, h, h, h, ng=ng, xmin=-6.0, ymin=-6.0, zmin=-6.0)
    rng = np.random.default_rng(seed)
    pos = offgrid_shift(rng, h)
    state = punctures(grid, [1.0], [pos], amp=1e-4, seed=seed + 1000, device=device)
    return grid, state, [pos.tolist()]


def config_binary(seed: int = 0, N: int = 256, L: float = 16.0, masses=(0.5, 0.5),
                  separation: float = 6.0, device="cpu", ng: int = 3):
    """N^3 interior, ng ghosts (3 by default), h = L/N on [-L/2, L/2): two punctures separated
    by ``separation`` along x, shifted together off the grid planes by a seeded sub-h offset,
    1e-4 perturbations. The default is the 256^3, h = 1/16 equal-mass binary."""
    h = L / N
    lo = -0.5 * L
    grid = Grid(N, N, N, h, h, h, ng=ng, xmin=lo, ymin=lo, zmin=lo)
    rng = np.random.default_rng(seed)
    s = offgrid_shift(rng, h)
    p1 = np.array([-0.5 * separation, 0.0, 0.0]) + s
    p2 = np.array([0.5 * separation, 0.0, 0.0]) + s
    state = punctures(grid, list(masses), [p1, p2], amp=1e-4, seed=seed + 1000, device=device)
    return grid, state, [p1.tolist(), p2.tolist()]
