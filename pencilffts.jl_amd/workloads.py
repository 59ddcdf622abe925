"""Synthetic input fields of the benchmark workloads.

fourier_chebyshev_field: the mixed-transform workload's data, a Float32
field on an X x Y x Z grid that is periodic in x and y and Chebyshev in z,
built as a sum of separable modes
    A cos(kx x + ky y + phi) T_m(zeta),
x = 2 pi i / X, y = 2 pi j / Y on the periodic axes and the Gauss-Chebyshev
points zeta_l = cos(pi (l + 1/2) / Z) on z, where T_m(zeta_l) =
cos(pi m (l + 1/2) / Z).  Those are the grids on which the plan's
(RFFT, FFT, R2R(REDFT10)) is exact for such modes: the RFFT of the x part
and the FFT of the y part each pick one wavenumber, and REDFT10 (DCT-II) of
T_m at the Gauss-Chebyshev points is Z at index m (2Z for m = 0).  Seeded
N(0, 1) amplitudes, wavenumbers 1 <= kx < X/2, -Y/2 <= ky < Y/2,
0 <= m < Z, phases uniform in [0, 2 pi), plus white noise of standard
deviation `noise`.
"""
import math

import numpy as np
import torch


def fourier_chebyshev_modes(dims, n_modes=32, seed=0):
    """The modes of fourier_chebyshev_field: a list of (A, kx, ky, m, phi)."""
    X, Y, Z = dims
    rng = np.random.default_rng(seed)
    modes = []
    for _ in range(n_modes):
        A = float(rng.standard_normal())
        kx = int(rng.integers(1, X // 2))
        ky = int(rng.integers(-(Y // 2), (Y + 1) // 2))
        m = int(rng.integers(0, Z))
        phi = float(rng.uniform(0.0, 2.0 * math.pi))
        modes.append((A, kx, ky, m, phi))
    return modes


def fourier_chebyshev_field(dims, n_modes=32, noise=1e-3, seed=0, device="cuda",
                            dtype=torch.float32):
    """The field on the logical grid, indexed [i, j, l] (x, y, z), as a torch
    tensor of `dtype` on `device`; built in float64, rounded once."""
    X, Y, Z = dims
    f64 = torch.float64
    x = torch.arange(X, device=device, dtype=f64) * (2.0 * math.pi / X)
    y = torch.arange(Y, device=device, dtype=f64) * (2.0 * math.pi / Y)
    zc = (torch.arange(Z, device=device, dtype=f64) + 0.5) * (math.pi / Z)
    u = torch.zeros((X, Y, Z), device=device, dtype=f64)
    for A, kx, ky, m, phi in fourier_chebyshev_modes(dims, n_modes, seed):
        # cos(a + b) = cos a cos b - sin a sin b over the (x, y) plane
        ax = kx * x[:, None] + phi
        by = ky * y[None, :]
        plane = torch.cos(ax) * torch.cos(by) - torch.sin(ax) * torch.sin(by)
        u += A * plane[:, :, None] * torch.cos(m * zc)[None, None, :]
    if noise:
        g = torch.Generator(device=device).manual_seed(int(seed) + 1)
        u += noise * torch.randn((X, Y, Z), generator=g, device=device, dtype=f64)
    return u.to(dtype)


def fourier_chebyshev_spectrum(dims, n_modes=32, seed=0):
    """Closed-form output of mul! with (RFFT, FFT, R2R(REDFT10)) on the
    noiseless field: complex128 numpy array indexed [kx, ky, m] with
    shape (X/2 + 1, Y, Z)."""
    X, Y, Z = dims
    S = np.zeros((X // 2 + 1, Y, Z), dtype=np.complex128)
    for A, kx, ky, m, phi in fourier_chebyshev_modes(dims, n_modes, seed):
        cz = 2 * Z if m == 0 else Z
        # only e^{i(kx x + ky y + phi)} / 2 lands in the stored half spectrum
        S[kx, ky % Y, m] += A * X * Y * cz * 0.5 * np.exp(1j * phi)
    return S
