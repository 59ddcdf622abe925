"""Shared test cases: the reference's test matrix and helpers mapping product
transforms to oracle names."""
import numpy as np

from oracle import pipeline as OP
from oracle import plan as OM

NP_DTYPES = {"Float32": np.float32, "Float64": np.float64,
             "ComplexF32": np.complex64, "ComplexF64": np.complex128}
CODE_NAMES = {0: "Float32", 1: "Float64", 2: "ComplexF32", 3: "ComplexF64"}

# test/transforms.jl:12
DATA_DIMS = (16, 12, 6)


def oracle_name(t):
    s = repr(t)
    s = s.replace("{even}", "_even").replace("{odd}", "_odd")
    if s.startswith("R2R{"):
        s = "R2R_" + s[4:-1]
    return s


def transform_list(T, size_in):
    """test/transforms.jl:312-337 (the full list, not the FAST_TESTS subset)."""
    trs = [(T.FFT(),), (T.RFFT(),), (T.BFFT(),), (T.NoTransform(),)]
    trs += [(T.R2R(k),) for k in T.R2R_SUPPORTED_KINDS]
    trs += [(T.NoTransform(), T.RFFT(), T.FFT()),
            (T.FFT(), T.NoTransform(), T.FFT()),
            (T.FFT(), T.NoTransform(), T.NoTransform()),
            (T.BRFFT(size_in),)]
    return trs


def expand(T, trs, N):
    if len(trs) == 1 and N > 1:
        return T.expand_dims(trs[0], N)
    return tuple(trs)


def meta_for(T, dims, trs, real_name, permute=True, extra=()):
    trs = expand(T, trs, len(dims))
    return OM.PlanMeta(tuple(dims), tuple(oracle_name(t) for t in trs), real_name, permute,
                       tuple(extra)), trs


def random_logical(rng, shape, eltype):
    if eltype.startswith("Complex"):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return rng.standard_normal(shape)


def f32_tol(dims):
    """5e-7 * log2(N) with N the total transform size."""
    return 5e-7 * max(np.log2(float(np.prod(dims))), 1.0)


def tol_for(real_name, dims):
    return 1e-13 if real_name == "Float64" else f32_tol(dims)


def forward_ref(meta, x_logical):
    return OP.memory_image(OP.forward(meta, OP.to_memory(x_logical, meta.perm_in)))
