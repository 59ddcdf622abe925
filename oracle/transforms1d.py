"""1D transform definitions (the arithmetic PencilFFTs delegates to FFTW).

The reference calls FFTW (FFTW.jl 1.4.3 / FFTW_jll 3.3.9 in
benchmarks/Manifest.toml:85-95) at src/Transforms/c2c.jl:53-56,
r2c.jl:83-97, r2r.jl:86-93 and executes the plans at
src/operations.jl:250.  FFTW is not available here, so the transforms are
restated from FFTW's published definitions (FFTW 3.3 manual, sections 4.8.1
"The 1d Discrete Fourier Transform", 4.8.2 "The 1d Real-data DFT", 4.8.3-4.8.5
"1d Real-even / Real-odd DFTs, DHT"):

  FFT      Y_k = sum_j X_j exp(-2 pi i j k / n)
  BFFT     Y_k = sum_j X_j exp(+2 pi i j k / n)
  RFFT     first n/2+1 outputs of the FFT of real X
  BRFFT    real Y_j of length d from the Hermitian half X_0..X_{d/2}
           (imaginary parts of X_0 and, for even d, X_{d/2} ignored)
  REDFT00  Y_k = X_0 + (-1)^k X_{n-1} + 2 sum_{j=1}^{n-2} X_j cos(pi j k/(n-1))
  REDFT10  Y_k = 2 sum_j X_j cos(pi (j+1/2) k / n)
  REDFT01  Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2) / n)
  REDFT11  Y_k = 2 sum_j X_j cos(pi (j+1/2)(k+1/2) / n)
  RODFT00  Y_k = 2 sum_j X_j sin(pi (j+1)(k+1) / (n+1))
  RODFT10  Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1) / n)
  RODFT01  Y_k = (-1)^k X_{n-1} + 2 sum_{j=0}^{n-2} X_j sin(pi (j+1)(k+1/2) / n)
  RODFT11  Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1/2) / n)
  DHT      Y_k = sum_j X_j [cos(2 pi j k / n) + sin(2 pi j k / n)]

R2R transforms of complex data act on the real and imaginary parts
separately (they are real-linear).  `direct` evaluates the sums in long
double with every angle reduced exactly in integers; `fast` uses numpy.fft /
scipy.fft (pocketfft) in float64.  tests/test_oracle.py checks one against the
other and both against closed forms.
"""
import numpy as np

try:
    import scipy.fft as _sfft
except ImportError:  # pragma: no cover - scipy is part of the image
    _sfft = None

LD = np.longdouble
# 2*pi to long-double precision
_TAU = LD("6.283185307179586476925286766559005768")


def _cis(num, den):
    """exp(+2 pi i num/den) as (cos, sin) long-double arrays, exact reduction."""
    num = np.mod(np.asarray(num, dtype=np.int64), den)
    th = _TAU * num.astype(LD) / LD(den)
    return np.cos(th), np.sin(th)


def _r2r_matrix(kind, n):
    """Real n x n matrix of an FFTW r2r kind (rows k, columns j), long double."""
    j = np.arange(n, dtype=np.int64)[None, :]
    k = np.arange(n, dtype=np.int64)[:, None]
    if kind == "REDFT00":
        if n < 2:
            raise ValueError("REDFT00 needs n >= 2")
        c, _ = _cis(j * k, 2 * (n - 1))
        m = 2 * c
        m[:, 0] = 1
        m[:, n - 1] = np.where(k[:, 0] % 2 == 0, 1, -1)
        return m
    if kind == "REDFT10":
        c, _ = _cis((2 * j + 1) * k, 4 * n)
        return 2 * c
    if kind == "REDFT01":
        c, _ = _cis(j * (2 * k + 1), 4 * n)
        m = 2 * c
        m[:, 0] = 1
        return m
    if kind == "REDFT11":
        c, _ = _cis((2 * j + 1) * (2 * k + 1), 8 * n)
        return 2 * c
    if kind == "RODFT00":
        _, s = _cis((j + 1) * (k + 1), 2 * (n + 1))
        return 2 * s
    if kind == "RODFT10":
        _, s = _cis((2 * j + 1) * (k + 1), 4 * n)
        return 2 * s
    if kind == "RODFT01":
        _, s = _cis((j + 1) * (2 * k + 1), 4 * n)
        m = 2 * s
        m[:, n - 1] = np.where(k[:, 0] % 2 == 0, 1, -1)
        return m
    if kind == "RODFT11":
        _, s = _cis((2 * j + 1) * (2 * k + 1), 8 * n)
        return 2 * s
    if kind == "DHT":
        c, s = _cis(j * k, n)
        return c + s
    raise ValueError(kind)


def _apply_real_matrix(m, x, axis):
    """y = m @ x along `axis`; complex x handled by parts."""
    x = np.moveaxis(np.asarray(x), axis, 0)
    if np.iscomplexobj(x):
        xr = x.real.astype(LD)
        xi = x.imag.astype(LD)
        y = (np.tensordot(m, xr, axes=(1, 0)).astype(np.float64)
             + 1j * np.tensordot(m, xi, axes=(1, 0)).astype(np.float64))
    else:
        y = np.tensordot(m, x.astype(LD), axes=(1, 0)).astype(np.float64)
    return np.moveaxis(y, 0, axis)


def _dft_direct(x, axis, sign, n_out=None):
    """sum_j x_j exp(sign 2 pi i j k / n) in long double, complex128 result."""
    x = np.moveaxis(np.asarray(x, dtype=np.complex128), axis, 0)
    n = x.shape[0]
    n_out = n if n_out is None else n_out
    j = np.arange(n, dtype=np.int64)[None, :]
    k = np.arange(n_out, dtype=np.int64)[:, None]
    c, s = _cis(j * k, n)
    s = s * sign
    xr = x.real.astype(LD)
    xi = x.imag.astype(LD)
    yr = np.tensordot(c, xr, axes=(1, 0)) - np.tensordot(s, xi, axes=(1, 0))
    yi = np.tensordot(s, xr, axes=(1, 0)) + np.tensordot(c, xi, axes=(1, 0))
    y = yr.astype(np.float64) + 1j * yi.astype(np.float64)
    return np.moveaxis(y, 0, axis)


def _hermitian_extend(X, d, axis):
    """Full length-d spectrum from the half X_0..X_{d//2}; imaginary parts of
    X_0 and (even d) X_{d/2} dropped, as FFTW's c2r transforms do."""
    X = np.moveaxis(np.asarray(X, dtype=np.complex128), axis, 0).copy()
    h = X.shape[0]
    assert h == d // 2 + 1, (h, d)
    X[0] = X[0].real
    if d % 2 == 0:
        X[d // 2] = X[d // 2].real
    full = np.zeros((d,) + X.shape[1:], dtype=np.complex128)
    full[:h] = X
    for k in range(1, (d + 1) // 2):
        full[d - k] = np.conj(X[k])
    return np.moveaxis(full, 0, axis)


def direct(tr, x, axis=0, d=None):
    """Brute-force definition of transform `tr` (oracle name) along `axis`.
    For BRFFT the output length d is implied by the kind and input length."""
    b = tr[:-1] if tr.endswith("!") else tr
    n = np.shape(x)[axis]
    if b == "NoTransform":
        return np.array(x, copy=True)
    if b == "FFT":
        return _dft_direct(x, axis, -1)
    if b == "BFFT":
        return _dft_direct(x, axis, +1)
    if b == "RFFT":
        return _dft_direct(np.asarray(x, dtype=np.float64), axis, -1, n // 2 + 1)
    if b.startswith("BRFFT"):
        d = 2 * n - 2 if b == "BRFFT_even" else 2 * n - 1
        full = _hermitian_extend(x, d, axis)
        return _dft_direct(full, axis, +1).real
    return _apply_real_matrix(_r2r_matrix(b[4:], n), x, axis)


_DCT_TYPE = {"REDFT00": 1, "REDFT10": 2, "REDFT01": 3, "REDFT11": 4}
_DST_TYPE = {"RODFT00": 1, "RODFT10": 2, "RODFT01": 3, "RODFT11": 4}


def _r2r_fast_real(kind, x, axis, workers):
    if kind in _DCT_TYPE:
        return _sfft.dct(x, type=_DCT_TYPE[kind], axis=axis, workers=workers)
    if kind in _DST_TYPE:
        return _sfft.dst(x, type=_DST_TYPE[kind], axis=axis, workers=workers)
    if kind == "DHT":
        f = _sfft.fft(x, axis=axis, workers=workers)
        return f.real - f.imag
    raise ValueError(kind)


def fast(tr, x, axis=0, workers=1):
    """FP64 numpy/scipy evaluation of transform `tr` along `axis`."""
    b = tr[:-1] if tr.endswith("!") else tr
    n = np.shape(x)[axis]
    if b == "NoTransform":
        return np.array(x, copy=True)
    if b == "FFT":
        return _sfft.fft(x, axis=axis, workers=workers)
    if b == "BFFT":
        return _sfft.ifft(x, axis=axis, norm="forward", workers=workers)
    if b == "RFFT":
        return _sfft.rfft(x, axis=axis, workers=workers)
    if b.startswith("BRFFT"):
        d = 2 * n - 2 if b == "BRFFT_even" else 2 * n - 1
        return _sfft.irfft(x, n=d, axis=axis, norm="forward", workers=workers)
    kind = b[4:]
    if np.iscomplexobj(x):
        return (_r2r_fast_real(kind, x.real, axis, workers)
                + 1j * _r2r_fast_real(kind, x.imag, axis, workers))
    return _r2r_fast_real(kind, x, axis, workers)


def apply(tr, x, axis=0, method="auto", workers=1):
    """Dispatch: brute force for small lines, FFTs otherwise."""
    n = np.shape(x)[axis]
    if method == "direct" or (method == "auto" and n <= 48 and np.size(x) <= 1 << 16):
        return direct(tr, x, axis)
    return fast(tr, x, axis, workers)
