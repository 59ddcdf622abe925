"""Spectral operators fused into the passes (pfft_execute_op) through the host
emulator, against numpy multipliers on the oracle's spectra.

The operators are the reference's Fourier-space operations: the derivative
i k_j (test/include/FourierOperations.jl:18-86, docs/examples/gradient.jl),
the Laplacian and its inverse, and the 2/3 dealiasing rule
(docs/examples/navier_stokes.jl:331-344), with integer wavenumbers laid out
as in test/include/Grids.jl:44-59 (rfftfreq along the r2c dimension, fftfreq
elsewhere, 2 pi box).  The transposes are checked by the dot-product identity
and `accumulate` by adding into a prefilled output."""
import ctypes as C

import numpy as np
import pytest

from oracle import pipeline as OP
from tests.cases import CODE_NAMES, NP_DTYPES, expand, meta_for, random_logical, tol_for


def wavenumbers(dims, r2c):
    ks = []
    for d, n in enumerate(dims):
        if d == 0 and r2c:
            ks.append(np.arange(n // 2 + 1, dtype=float))
        else:
            ks.append(np.fft.fftfreq(n, 1.0 / n))
    return ks


def multiplier(op, dim, dims, r2c, extra=()):
    """m(k) on the logical spectrum (plus broadcast extra dims)."""
    ks = wavenumbers(dims, r2c)
    N = len(dims)
    grids = np.meshgrid(*ks, indexing="ij")
    if op == 1:
        m = 1j * grids[dim]
    elif op == 2:
        m = -sum(g ** 2 for g in grids) + 0j
    elif op == 3:
        s = sum(g ** 2 for g in grids)
        m = np.where(s > 0, -1.0 / np.where(s > 0, s, 1.0), 0.0) + 0j
    else:
        kmax = [np.max(np.abs(k)) for k in ks]
        keep = np.ones(grids[0].shape, dtype=bool)
        for g, km in zip(grids, kmax):
            keep &= np.abs(g) <= (2.0 / 3.0) * km
        m = keep.astype(float) + 0j
    return m.reshape(m.shape + (1,) * len(extra))


def _op(pf, op, dim=0, alpha=1.0, acc=0):
    return pf._lib.SpectralOp(op, dim, alpha, acc)


def _setup(pf, dims, trs, real_type, permute=True, extra=()):
    T = pf.Transforms
    meta, full = meta_for(T, dims, trs, CODE_NAMES[real_type], permute, extra)
    kinds = [t.code for t in full]
    r2c = (full[0].code & ~pf._lib.INPLACE) == pf._lib.RFFT
    return meta, kinds, r2c


OPS = [(1, 0), (1, 1), (1, 2), (2, 0), (3, 0), (4, 0)]
SHAPES = [((16, 12, 6), "RFFT", True, ()), ((15, 7, 5), "FFT", True, ()),
          ((16, 12, 6), "RFFT", False, ()), ((12, 10, 8), "RFFT", True, (3,)),
          ((11, 15), "RFFT", True, ())]


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s[0])) + "-" + s[1]
                         + ("" if s[2] else "-noperm") + ("-extra" if s[3] else ""))
@pytest.mark.parametrize("op,dim", OPS, ids=["dx", "dy", "dz", "lap", "invlap", "dealias"])
def test_forward_and_backward_with_operator(pf, emu, shape, op, dim, real_type):
    dims, tname, permute, extra = shape
    if dim >= len(dims):
        pytest.skip("no such dimension")
    T = pf.Transforms
    meta, kinds, r2c = _setup(pf, dims, (getattr(T, tname)(),), real_type, permute, extra)
    dt_in, dt_out = NP_DTYPES[meta.eltype_in], NP_DTYPES[meta.eltype_out]
    rng = np.random.default_rng(11)
    x = random_logical(rng, tuple(dims) + tuple(extra), meta.eltype_in)
    xm = OP.to_memory(x, meta.perm_in).astype(dt_in)
    xw = xm.astype(np.complex128 if np.iscomplexobj(xm) else np.float64)
    m = multiplier(op, dim, dims, r2c, extra)
    mm = OP.to_memory(np.broadcast_to(m, tuple(meta.size_out) + tuple(extra)), meta.perm_out)
    alpha = 0.75
    tol = tol_for(CODE_NAMES[real_type], dims)
    # mul! with the operator: alpha m F(x)
    out = emu.run_op(dims, kinds, real_type, 0, OP.memory_image(xm), _op(pf, op, dim, alpha),
                     permute=permute, extra=extra).view(dt_out)
    ref = alpha * mm * OP.forward(meta, xw)
    assert OP.rel_l2(out, OP.memory_image(ref)) <= tol
    # ldiv! with the operator: alpha B(m y) / N, on a spectrum of a real field
    y = OP.forward(meta, xw)
    ym = OP.memory_image(y).astype(dt_out)
    back = emu.run_op(dims, kinds, real_type, 2, ym, _op(pf, op, dim, alpha),
                      permute=permute, extra=extra).view(dt_in)
    want = alpha * OP.backward(meta, (mm * y), scaled=True)
    if not np.iscomplexobj(xm):
        want = want.real if np.iscomplexobj(want) else want
    assert OP.rel_l2(back, OP.memory_image(want)) <= tol * 10


def _inner(a, b):
    return float(np.real(np.vdot(a, b)))


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("op,dim", [(1, 0), (1, 2), (2, 0), (3, 0), (4, 0)],
                         ids=["dx", "dz", "lap", "invlap", "dealias"])
def test_transposes_with_operator(pf, emu, op, dim, real_type):
    """<y, A x> = <A^T y, x> for A = mul! and bmul! with the operator; the
    transposes take conj(m)."""
    T = pf.Transforms
    dims = (16, 12, 6)
    meta, kinds, r2c = _setup(pf, dims, (T.RFFT(),), real_type)
    dt_in, dt_out = NP_DTYPES[meta.eltype_in], NP_DTYPES[meta.eltype_out]
    rng = np.random.default_rng(3)
    x = OP.memory_image(OP.to_memory(random_logical(rng, dims, meta.eltype_in),
                                     meta.perm_in)).astype(dt_in)
    y = OP.memory_image(OP.to_memory(random_logical(rng, tuple(meta.size_out), meta.eltype_out),
                                     meta.perm_out)).astype(dt_out)
    o = _op(pf, op, dim, 1.5)
    Ax = emu.run_op(dims, kinds, real_type, 0, x, o).view(dt_out)
    ATy = emu.run_op(dims, kinds, real_type, 3, y, o).view(dt_in)
    lhs, rhs = _inner(y.astype(np.complex128), Ax.astype(np.complex128)), _inner(
        ATy.astype(np.float64), x.astype(np.float64))
    tol = 1e-12 if real_type == 1 else 3e-5
    assert abs(lhs - rhs) <= tol * max(abs(lhs), 1.0), (lhs, rhs)
    By = emu.run_op(dims, kinds, real_type, 1, y, o).view(dt_in)
    BTx = emu.run_op(dims, kinds, real_type, 4, x, o).view(dt_out)
    lhs, rhs = _inner(x.astype(np.float64), By.astype(np.float64)), _inner(
        BTx.astype(np.complex128), y.astype(np.complex128))
    assert abs(lhs - rhs) <= tol * max(abs(lhs), 1.0), (lhs, rhs)


def test_accumulate_sums_components(pf, emu):
    """divergence-like sum: sum_j i k_j F(u_j) by three accumulating
    forwards into one spectrum."""
    T = pf.Transforms
    dims = (16, 12, 6)
    meta, kinds, r2c = _setup(pf, dims, (T.RFFT(),), 1)
    rng = np.random.default_rng(8)
    us = [OP.to_memory(random_logical(rng, dims, "Float64"), meta.perm_in) for _ in range(3)]
    acc = np.zeros(int(np.prod(meta.mem_dims("out"))) * 16, dtype=np.uint8)
    for j in range(3):
        acc = emu.run_op(dims, kinds, 1, 0, OP.memory_image(us[j]), _op(pf, 1, j, 1.0, 1),
                         out=acc)
    ref = 0
    for j in range(3):
        mm = OP.to_memory(multiplier(1, j, dims, r2c), meta.perm_out)
        ref = ref + mm * OP.forward(meta, us[j])
    assert OP.rel_l2(acc.view(np.complex128), OP.memory_image(ref)) <= 1e-13


def test_operator_errors(pf, emu):
    T = pf.Transforms
    L = pf._lib
    x3 = np.zeros(16 * 12 * 6)
    _, k3, _ = _setup(pf, (16, 12, 6), (T.RFFT(),), 1)
    with pytest.raises(RuntimeError, match="accumulate"):
        emu.run_op((16, 12, 6), k3, 1, 2, np.zeros(9 * 12 * 6, dtype=complex),
                   _op(pf, 1, 0, 1.0, 1))
    with pytest.raises(RuntimeError, match="unknown spectral operator"):
        emu.run_op((16, 12, 6), k3, 1, 0, x3, _op(pf, 9))
    with pytest.raises(RuntimeError, match="derivative dimension"):
        emu.run_op((16, 12, 6), k3, 1, 0, x3, _op(pf, 1, 3))
    _, kr, _ = _setup(pf, (16, 12, 6), (T.R2R(T.REDFT10),), 1)
    with pytest.raises(RuntimeError, match="wavenumbers"):
        emu.run_op((16, 12, 6), kr, 1, 0, x3, _op(pf, 2))
    _, k1, _ = _setup(pf, (16,), (T.RFFT(),), 1)
    with pytest.raises(RuntimeError, match="c2c"):
        emu.run_op((16,), k1, 1, 0, np.zeros(16), _op(pf, 2))
    assert L.OP_DEALIAS == 4
