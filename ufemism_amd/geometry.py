"""Seeded ice-sheet geometries and friction fields for the benchmark workloads.

* ``dome``   -- flat bed at 0 m, radially symmetric Vialov-type dome
  H(r) = 3600 (1 - (r/750 km)^(4/3))^(3/8), uniform A and friction.
* ``marine`` -- bed b = -150 - 0.5 r[km] + seeded 50 m Gaussian bumps
  (20 km correlation); 4000 m of ice at the centre thinning to a 300 m
  floating shelf; grounding line from flotation (rho_i 910, rho_w 1028);
  log-normal friction.

Friction is expressed as a basal-friction override on the combined AaAc
points: tau_c = beta * u_th^q (regularised Coulomb law of the reference,
src/ice_dynamics_module.f90:727-779), optionally times a seeded
log-normal field exp(sigma * g).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import meshgen
from .config import DynParams, tau_c_from_beta

SEC_A = 1e-16  # Pa^-3 yr^-1


def dome_resolution(r_margin=750e3, fine=20e3, coarse=100e3, band=250e3):
    def res(x, y):
        r = np.hypot(x, y)
        return fine + (coarse - fine) * np.clip(np.abs(r - r_margin) / band, 0, 1)
    return res


def marine_resolution(r_gl=640e3, fine=5e3, coarse=40e3, band=300e3):
    def res(x, y):
        r = np.hypot(x, y)
        return fine + (coarse - fine) * np.clip(np.abs(r - r_gl) / band, 0, 1)
    return res


def gaussian_field(points, corr_len, seed, xmin=-800e3, xmax=800e3, ymin=-800e3, ymax=800e3):
    """Unit-variance Gaussian random field with a Gaussian covariance of the
    given correlation length, sampled on a regular grid by FFT filtering and
    interpolated bilinearly to ``points`` (N x 2)."""
    h = corr_len / 4.0
    nx = int(math.ceil((xmax - xmin) / h)) + 1
    ny = int(math.ceil((ymax - ymin) / h)) + 1
    pad = 8
    Nx, Ny = nx + 2 * pad, ny + 2 * pad
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((Ny, Nx))
    kx = np.fft.fftfreq(Nx, d=h) * 2 * np.pi
    ky = np.fft.fftfreq(Ny, d=h) * 2 * np.pi
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    filt = np.exp(-K2 * corr_len ** 2 / 8.0)
    g = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    g = (g - g.mean()) / g.std()
    g = g[pad:pad + ny, pad:pad + nx]
    fx = (points[:, 0] - xmin) / h
    fy = (points[:, 1] - ymin) / h
    i0 = np.clip(np.floor(fx).astype(int), 0, nx - 2)
    j0 = np.clip(np.floor(fy).astype(int), 0, ny - 2)
    tx = fx - i0
    ty = fy - j0
    return ((1 - tx) * (1 - ty) * g[j0, i0] + tx * (1 - ty) * g[j0, i0 + 1]
            + (1 - tx) * ty * g[j0 + 1, i0] + tx * ty * g[j0 + 1, i0 + 1])


@dataclasses.dataclass
class Workload:
    """A mesh, its initial ice state (Aa-mesh fields) and the parameters."""
    name: str
    mesh: meshgen.Mesh
    Hi: np.ndarray
    Hb: np.ndarray
    SL: np.ndarray
    A_flow: float
    beta: float
    friction_sigma: float
    friction_seed: int
    friction_corr: float
    params: DynParams
    SMB: np.ndarray = None
    BMB: np.ndarray = None
    dt: float = 1.0

    def tau_c_override(self, VAaAc: np.ndarray) -> np.ndarray:
        tau = np.full(VAaAc.shape[0], tau_c_from_beta(self.beta))
        if self.friction_sigma > 0:
            g = gaussian_field(VAaAc, self.friction_corr, self.friction_seed)
            tau = tau * np.exp(self.friction_sigma * g)
        return tau

    def temperature(self) -> np.ndarray:
        """Englacial temperature Ti (nV, nZ) [K] for the Arrhenius branch of
        ice_physical_properties: 253.15 K minus 6 K per km of ice at the surface,
        warming with zeta^2 towards a base 25 K per km warmer, capped 0.5 K below
        the pressure-melting point; it crosses the 263.15 K switch of the law."""
        z = np.asarray(self.params.zeta, np.float64)
        H = np.asarray(self.Hi, np.float64)
        Ts = 253.15 - 6e-3 * H
        Tb = np.minimum(Ts + 25e-3 * H, 273.16 - 8.7e-4 * H - 0.5)
        return np.asfortranarray(Ts[:, None] + (Tb - Ts)[:, None] * z[None, :] ** 2)


def dome_thickness(x, y, H0=3600.0, R=750e3):
    r = np.hypot(x, y)
    s = np.clip(r / R, 0, 1)
    return np.where(r < R, H0 * np.power(np.clip(1 - s ** (4.0 / 3.0), 0, None), 3.0 / 8.0), 0.0)


def dome(n_target=8000, seed=1, picard=10, omega=1.3) -> Workload:
    """Flat-bed dome: plain FP64 SIA+SSA single dynamics step."""
    mesh = meshgen.generate_mesh(dome_resolution(), n_target, seed=seed)
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    p = DynParams(m_enh_sia=1.0, m_enh_ssa=1.0, SSA_max_outer_loops=picard,
                  SSA_SOR_omega=omega, use_tau_c_override=True)
    nV = mesh.nV
    return Workload("dome", mesh, dome_thickness(x, y), np.zeros(nV), np.zeros(nV),
                    SEC_A, 1e4, 0.0, seed + 1000, 50e3, p,
                    SMB=np.zeros(nV), BMB=np.zeros(nV), dt=1.0)


def marine_thickness(x, y, H0=4000.0, R=650e3, H_shelf=300.0, R_front=750e3):
    r = np.hypot(x, y)
    s = np.clip(r / R, 0, 1)
    H = H0 * np.sqrt(np.clip(1 - s ** 2, 0, None))
    H = np.maximum(H, H_shelf)
    return np.where(r < R_front, H, 0.0)


def marine(n_target=200_000, seed=2, picard=30, RN_tol=1e-6, omega=1.3,
           friction_sigma=0.5, bump_amp=50.0, bump_corr=20e3) -> Workload:
    """Marine ice sheet with a floating shelf (grounding line by flotation)."""
    mesh = meshgen.generate_mesh(marine_resolution(), n_target, seed=seed)
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    r = np.hypot(x, y)
    Hb = -150.0 - 0.5 * r / 1e3 + bump_amp * gaussian_field(mesh.V, bump_corr, seed + 7)
    p = DynParams(m_enh_sia=1.0, m_enh_ssa=1.0, SSA_max_outer_loops=picard,
                  SSA_RN_tol=RN_tol, SSA_SOR_omega=omega, use_tau_c_override=True)
    nV = mesh.nV
    return Workload("marine", mesh, marine_thickness(x, y), Hb, np.zeros(nV),
                    SEC_A, 1e4, friction_sigma, seed + 2000, 50e3, p,
                    SMB=np.zeros(nV), BMB=np.zeros(nV), dt=1.0)
