"""Seeded synthetic inputs for the benchmark workloads.

Permeability is lognormal, K = K_g * exp(sigma * G), where G is a zero-mean,
unit-variance Gaussian field with exponential covariance exp(-r/lambda)
(r in cells). G is made by spectral synthesis on the periodic grid: white
noise from numpy's PCG64 generator, filtered in Fourier space by the square
root of the covariance's discrete spectrum (the FFT of the periodic
covariance), then normalised to unit variance. With lambda much smaller than
the grid the wrap-around correlation is negligible.

The field sets follow the workload definitions: hydrostatic Dirichlet
pressure on the x faces (k points up, top at k = nz-1), no flow elsewhere,
porosity 0.2, linear EOS rho = 1000 + 700*C, viscosity 1e-3 Pa s, fluid /
rock compressibility 4.5e-10 / 1e-10 1/Pa.
"""
from __future__ import annotations

import math

import numpy as np


def gaussian_field(shape, corr_len, seed, workers=-1):
    """Unit-variance Gaussian field with exponential covariance exp(-r/corr_len)."""
    import scipy.fft as sfft
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    # periodic distance to the origin, per axis
    ax = [np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
          for n in shape]
    r2 = (ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2
          + ax[2][None, None, :] ** 2)
    cov = np.exp(-np.sqrt(r2) / float(corr_len))
    del r2
    spec = sfft.rfftn(cov, workers=workers).real
    del cov
    np.maximum(spec, 0.0, out=spec)
    amp = np.sqrt(spec)
    del spec
    noise = rng.standard_normal(shape)
    f = sfft.rfftn(noise, workers=workers)
    del noise
    f *= amp
    del amp
    g = sfft.irfftn(f, s=shape, workers=workers)
    del f
    g -= g.mean()
    g /= g.std()
    return g


def lognormal_K(shape, sigma, corr_len, seed, k_geo=1e-12):
    G = gaussian_field(shape, corr_len, seed)
    np.multiply(G, sigma, out=G)
    np.exp(G, out=G)
    G *= k_geo
    return G


def hydrostatic_P(dx, dz, beta_deg, rho0=1000.0, g=9.80665, p_top=1e5,
                  ny=1):
    """Discrete hydrostatic pressure of the reference's flux form: zero qz /
    qx for uniform density (tse.py:758-777), k up, top at nz-1."""
    b = math.radians(beta_deg)
    nz = len(dz)
    col = np.empty(nz)
    col[nz - 1] = p_top + rho0 * g * math.cos(b) * 0.5 * dz[nz - 1]
    for k in range(nz - 2, -1, -1):
        col[k] = col[k + 1] + rho0 * g * math.cos(b) * 0.5 * (dz[k] + dz[k + 1])
    xs = np.cumsum(dx) - 0.5 * np.asarray(dx)
    P = col[None, None, :] - rho0 * g * math.sin(b) * xs[:, None, None]
    return np.broadcast_to(P, (len(dx), ny, nz)).copy()


class PressureProblem:
    """All inputs of one compute_P call (reference layout, float64)."""

    def __init__(self, nx, ny, nz, spacing=5.0, sigma=2.5, corr_len=16,
                 seed=3, beta=0.0, dz=None, salt_top=4):
        shp = (nx, ny, nz)
        self.shape = shp
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx = np.full(nx, float(spacing))
        self.dy = np.full(ny, float(spacing))
        self.dz = np.full(nz, float(spacing)) if dz is None else np.asarray(dz, float)
        self.g = 9.80665
        self.beta = beta
        K = lognormal_K(shp, sigma, corr_len, seed)
        self.Kx = K
        self.Ky = K
        self.Kz = K
        self.phi = np.full(shp, 0.2)
        self.c_f = np.full(shp, 4.5e-10)
        self.c_r = np.full(shp, 1e-10)
        self.nue = np.full(shp, 1e-3)
        C = np.zeros(shp)
        if salt_top:
            C[nx // 4: nx - nx // 4, :, nz - salt_top:] = 1.0
        self.C = C
        self.rho_f = 1000.0 + 700.0 * C
        self.BP = np.zeros(shp)
        self.BP[0] = 1.0
        self.BP[-1] = 1.0
        self.WP = np.zeros(shp)
        self.P = hydrostatic_P(self.dx, self.dz, beta, ny=ny)
        self.qx = np.zeros(shp)
        self.qy = np.zeros(shp)
        self.qz = np.zeros(shp)

    def grid_args(self):
        return (self.nx, self.ny, self.nz, self.dx, self.dy, self.dz, self.g,
                self.beta)

    def p_inputs(self):
        """(c_f, nue, rho_f, phi, Kx, Ky, Kz, BP, WP, c_r) in C-ABI order."""
        return dict(c_f=self.c_f, nue=self.nue, rho_f=self.rho_f, phi=self.phi,
                    Kx=self.Kx, Ky=self.Ky, Kz=self.Kz, BP=self.BP, WP=self.WP,
                    c_r=self.c_r)
