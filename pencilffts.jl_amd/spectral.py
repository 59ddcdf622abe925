"""Spectral operators fused into the transform passes (pfft_execute_op).

A diagonal multiplier alpha m(k) of the spectrum, with the integer
wavenumbers of test/include/Grids.jl:44-59 (rfftfreq along the r2c
dimension, fftfreq elsewhere, 2 pi box):

    DERIVATIVE     m = i k_dim
    LAPLACIAN      m = -|k|^2
    INV_LAPLACIAN  m = -1/|k|^2 (0 at k = 0)
    DEALIAS        2/3 rule (docs/examples/navier_stokes.jl:331-344)

The reference applies these in Julia loops between transforms
(test/include/FourierOperations.jl, docs/examples/gradient.jl); here the
multiplier rides on the pass that writes the spectrum (mul!, the transpose of
bmul!) or reads it (bmul!/ldiv!, the transpose of mul!), so it costs no memory
pass.  `accumulate` adds into dst on the writing side.
"""
from . import _lib as L

DERIVATIVE, LAPLACIAN, INV_LAPLACIAN, DEALIAS = (L.OP_DERIVATIVE, L.OP_LAPLACIAN,
                                                 L.OP_INV_LAPLACIAN, L.OP_DEALIAS)


def spectral_op(op, dim=0, alpha=1.0, accumulate=False):
    return L.SpectralOp(int(op), int(dim), float(alpha), int(bool(accumulate)))


def mul_(dst, p, src, op, dim=0, alpha=1.0, accumulate=False):
    """dst (+)= alpha m(k) F(src)."""
    return p._execute(L.FORWARD, src, dst, op=spectral_op(op, dim, alpha, accumulate))


def bmul_(dst, p, src, op, dim=0, alpha=1.0):
    """dst = alpha B(m(k) src), unnormalised."""
    return p._execute(L.BACKWARD, src, dst, op=spectral_op(op, dim, alpha))


def ldiv_(dst, p, src, op, dim=0, alpha=1.0):
    """dst = alpha B(m(k) src) / N: e.g. a derivative of a field from its spectrum."""
    return p._execute(L.BACKWARD_SCALED, src, dst, op=spectral_op(op, dim, alpha))


def mul_adjoint_(dst, p, src, op, dim=0, alpha=1.0):
    """dst = alpha F^T(conj(m) src): the transpose of mul_ with the operator."""
    return p._execute(L.ADJOINT_FORWARD, src, dst, op=spectral_op(op, dim, alpha))


def bmul_adjoint_(dst, p, src, op, dim=0, alpha=1.0, accumulate=False):
    """dst (+)= alpha conj(m) B^T(src): the transpose of bmul_ with the operator."""
    return p._execute(L.ADJOINT_BACKWARD, src, dst, op=spectral_op(op, dim, alpha, accumulate))
