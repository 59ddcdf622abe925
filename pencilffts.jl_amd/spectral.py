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
import ctypes as C

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


# ---- vector calculus from the fused derivative (FourierOperations.jl) ------
# `scale[d]` = 2 pi / L_d turns the integer wavenumbers of a 2 pi box into
# those of a box of length L_d along logical dimension d.

def divergence_(dst, p, us, scale=(1.0, 1.0, 1.0)):
    """dst = F(div u) = sum_j i k_j F(u_j): the spectrum of the divergence of
    the vector field us (test/include/FourierOperations.jl:18-42), by
    accumulating forwards."""
    for j, u in enumerate(us):
        mul_(dst, p, u, DERIVATIVE, dim=j, alpha=scale[j], accumulate=j > 0)
    return dst


def curl_(dsts, p, us, scale=(1.0, 1.0, 1.0)):
    """dsts[i] = F(curl u)_i: i k_j F(u_k) - i k_k F(u_j) for (i, j, k) cyclic
    (test/include/FourierOperations.jl:44-86), two accumulating forwards per
    component."""
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        mul_(dsts[i], p, us[k], DERIVATIVE, dim=j, alpha=scale[j])
        mul_(dsts[i], p, us[j], DERIVATIVE, dim=k, alpha=-scale[k], accumulate=True)
    return dsts


def gradient_(dsts, p, uh, scale=None):
    """dsts[j] = d_j u from the spectrum uh: one normalised backward transform
    per component with the derivative on its first pass
    (docs/examples/gradient.jl:254-273)."""
    n = len(dsts)
    scale = scale or (1.0,) * n
    for j in range(n):
        ldiv_(dsts[j], p, uh, DERIVATIVE, dim=j, alpha=scale[j])
    return dsts


def project_divergence_free_(uhs, p, scale=None, stream=None):
    """In place: uh_i <- uh_i - k_i (k . uh) / |k|^2 on the three output-shaped
    spectra uhs (docs/examples/navier_stokes.jl:351-366); the mean mode is
    left as is.  One elementwise kernel (pfft_project_divergence_free)."""
    import torch
    if len(uhs) != 3:
        raise L.ArgumentError(L.ERR_ARGUMENT,
                              "the divergence-free projection takes three components")
    for a in uhs:
        if a.pencil != p.pencil_output or a.eltype != p.eltype_output:
            raise L.ArgumentError(L.ERR_ARGUMENT, "unexpected dimensions of input/output data")
    sc = (C.c_double * 3)(*scale) if scale is not None else None
    s = stream if stream is not None else torch.cuda.current_stream(uhs[0].parent.device)
    L.check(L.lib.pfft_project_divergence_free(p._h, *(a.data_ptr() for a in uhs), sc,
                                               s.cuda_stream))
    return uhs
