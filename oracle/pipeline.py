"""Execution layer of PencilFFTs.jl v0.15.3, restated for one process.

Arrays are held the way the reference holds them: a PencilArray's parent is a
column-major array whose axes are the pencil's memory order, extra dims last.
Here such an array is a numpy array indexed [m_0, m_1, ..., e...] in memory
order; its flat memory image is `a.ravel(order="F")`.

forward  (mul!,  operations.jl:24-31,161-172,227-257): for stage n = 1..N,
         transpose the data into stage n's pencil (Transpositions.transpose!,
         :239-240 -- on one process a pure permuted copy), then apply the 1D
         transform along the memory axis of dim n (:250; plans.jl:531-538).
backward (bmul!/ldiv!, :34-52): the same with the plans reversed and the
         binv transforms (transform_info, :154-158); ldiv! then scales the
         result by inv(scale_factor(p)) (:50,64-66).
"""
import numpy as np

from . import plan as P
from . import transforms1d as T1


def _np_dtype(eltype):
    return {"Float32": np.float32, "Float64": np.float64,
            "ComplexF32": np.complex64, "ComplexF64": np.complex128}[eltype]


def to_memory(a_logical, perm):
    """Memory-order array of a logical-order array for a pencil permutation
    (perm[i] = logical dim stored in memory axis i, 0-based)."""
    nd = a_logical.ndim
    N = len(perm)
    return np.transpose(a_logical, tuple(perm) + tuple(range(N, nd)))


def from_memory(a_mem, perm):
    N = len(perm)
    inv = np.argsort(perm)
    return np.transpose(a_mem, tuple(inv) + tuple(range(N, a_mem.ndim)))


def memory_image(a_mem):
    """Flat column-major buffer of a memory-order array."""
    return np.ravel(a_mem, order="F")


def from_image(buf, mem_dims):
    return np.reshape(buf, mem_dims, order="F")


def _permute(a_mem, perm_from, perm_to):
    """Transpose between two pencil memory orders (permuted copy)."""
    return np.ascontiguousarray(to_memory(from_memory(a_mem, perm_from), perm_to))


def _transform(tr, a_mem, axis, method, workers):
    return T1.apply(tr, a_mem, axis=axis, method=method, workers=workers)


def _widen(x, single):
    if single:
        return x.astype(np.complex64 if np.iscomplexobj(x) else np.float32, copy=False)
    return x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)


def forward(meta, u_mem, method="auto", workers=1, single=False):
    """mul!(v, p, u): u_mem in the input pencil's memory order (unpermuted);
    returns v in the output pencil's memory order, float64/complex128
    (float32/complex64 with single=True, used only to time the CPU baseline)."""
    x = _widen(np.asarray(u_mem), single)
    perm_prev = meta.stages[0]["perm"]
    for st in meta.stages:
        if st["perm"] != perm_prev:
            x = _permute(x, perm_prev, st["perm"])
            perm_prev = st["perm"]
        x = _transform(st["tr"], x, st["mem_axis"], method, workers)
    return x


def backward(meta, v_mem, scaled=True, method="auto", workers=1, single=False):
    """ldiv!(u, p, v) (scaled) or bmul!(u, p, v): v_mem in the output pencil's
    memory order; returns u in the input pencil's memory order."""
    x = _widen(np.asarray(v_mem), single)
    perm_prev = meta.stages[-1]["perm"]
    for st in reversed(meta.stages):
        if st["perm"] != perm_prev:
            x = _permute(x, perm_prev, st["perm"])
            perm_prev = st["perm"]
        x = _transform(st["tr_bwd"], x, st["mem_axis"], method, workers)
    # the result lives in stage 1's pencil, which is the input pencil
    if scaled:
        x *= x.dtype.type(1.0 / meta.scale) if single else 1.0 / meta.scale
    return x


def forward_logical(meta, u_logical, **kw):
    """Convenience: logical-order input -> logical-order output."""
    v = forward(meta, to_memory(u_logical, meta.perm_in), **kw)
    return from_memory(v, meta.perm_out)


def rel_l2(y, y_ref):
    y = np.asarray(y, dtype=np.complex128 if np.iscomplexobj(y) or np.iscomplexobj(y_ref)
                   else np.float64)
    num = np.linalg.norm((y - y_ref).ravel())
    den = np.linalg.norm(np.asarray(y_ref).ravel())
    return float(num / den) if den > 0 else float(num)
