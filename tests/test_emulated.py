"""CPU parity of the generic pass with the oracle, through the host emulator.

libpencilffts_emu.so runs the very phase functions of the generic GPU kernel
(generic_pass.hpp) thread by thread on the CPU, with the same plan, pass
descriptors and precision-rounded twiddle tables as the product library.
These tests cover the kernel arithmetic and index maps on the reference's
test matrix without a GPU; the -m gpu tests repeat them on the device."""
import numpy as np
import pytest

from oracle import pipeline as OP
from oracle import plan as OM
from tests.cases import (CODE_NAMES, DATA_DIMS, NP_DTYPES, meta_for, random_logical, tol_for,
                         transform_list)


def emu_case(pf, emu, dims, trs, real_type, permute=True, extra=(), seed=0):
    T = pf.Transforms
    real_name = CODE_NAMES[real_type]
    meta, full = meta_for(T, dims, trs, real_name, permute, extra)
    kinds = [t.code for t in full]
    rng = np.random.default_rng(seed)
    is_c2r = full[-1].code in (pf._lib.BRFFT_EVEN, pf._lib.BRFFT_ODD)
    dt_in = NP_DTYPES[meta.eltype_in]
    dt_out = NP_DTYPES[meta.eltype_out]
    if is_c2r:
        y = random_logical(rng, tuple(meta.size_out) + tuple(extra), meta.eltype_out)
        ym = OP.memory_image(OP.to_memory(y, meta.perm_out)).astype(dt_out)
        xin = emu.run(dims, kinds, real_type, 2, ym, permute, extra).view(dt_in)
    else:
        x = random_logical(rng, tuple(dims) + tuple(extra), meta.eltype_in)
        xin = OP.memory_image(OP.to_memory(x, meta.perm_in)).astype(dt_in)
    xw = xin.astype(np.complex128 if np.iscomplexobj(xin) else np.float64)
    out = emu.run(dims, kinds, real_type, 0, xin, permute, extra).view(dt_out)
    xm = OP.from_image(xw, meta.mem_dims("in"))
    ref = OP.memory_image(OP.forward(meta, xm))
    err_f = OP.rel_l2(out, ref)
    back = emu.run(dims, kinds, real_type, 2, out, permute, extra).view(dt_in)
    err_b = OP.rel_l2(back, xw)
    return err_f, err_b


@pytest.mark.parametrize("extra", [(), (3,)], ids=["noextra", "extra3"])
@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("idx", range(16))
def test_transform_list(pf, emu, idx, real_type, extra):
    """test/transforms.jl:312-337 at DATA_DIMS (16, 12, 6)."""
    T = pf.Transforms
    trs = transform_list(T, DATA_DIMS)[idx]
    err_f, err_b = emu_case(pf, emu, DATA_DIMS, trs, real_type, extra=extra, seed=idx)
    tol = tol_for(CODE_NAMES[real_type], DATA_DIMS)
    assert err_f <= tol, err_f
    assert err_b <= tol, err_b


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims,permute", [((11, 15), True), ((11, 7, 13), True),
                                          ((9, 7, 5, 12), True), ((9, 7, 5, 12), False),
                                          ((42, 24, 16), True), ((41, 23, 15), True),
                                          ((13,), True)])
def test_dimensionality_and_sizes(pf, emu, dims, permute, real_type):
    """transforms.jl:414-462 and the rfft.jl sizes (42,24,16) / (41,23,15)."""
    T = pf.Transforms
    for trs in [(T.RFFT(),), (T.FFT(),)]:
        err_f, err_b = emu_case(pf, emu, dims, trs, real_type, permute=permute)
        tol = tol_for(CODE_NAMES[real_type], dims)
        assert err_f <= tol and err_b <= tol, (trs, err_f, err_b)


R2R_KINDS = ["REDFT00", "REDFT01", "REDFT10", "REDFT11", "RODFT00", "RODFT01", "RODFT10",
             "RODFT11", "DHT"]


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims", [(16, 12, 6), (15, 9, 7)], ids=["even", "odd"])
@pytest.mark.parametrize("data", ["real", "complex"])
@pytest.mark.parametrize("kind", R2R_KINDS)
def test_r2r_kinds(pf, emu, kind, data, dims, real_type):
    """The nine kinds of src/Transforms/r2r.jl:5-15 on the generic pass, real
    data ((R2R(kind),) x 3) and complex data ((FFT, R2R(kind), R2R(kind))), at
    even lengths (REDFT11 / RODFT11 as two half-length FFTs in one) and odd
    ones (their zero-padded 2n-point form)."""
    T = pf.Transforms
    k = getattr(T, kind)
    trs = (T.R2R(k),) if data == "real" else (T.FFT(), T.R2R(k), T.R2R(k))
    err_f, err_b = emu_case(pf, emu, dims, trs, real_type, seed=5)
    tol = tol_for(CODE_NAMES[real_type], dims)
    assert err_f <= tol and err_b <= tol, (err_f, err_b, tol)


def test_mixed_config3_small(pf, emu):
    """(RFFT, FFT, R2R(REDFT10)) -- the mixed workload's transform tuple -- at a
    small size, both precisions."""
    T = pf.Transforms
    trs = (T.RFFT(), T.FFT(), T.R2R(T.REDFT10))
    for rt in (0, 1):
        err_f, err_b = emu_case(pf, emu, (24, 12, 10), trs, rt)
        tol = tol_for(CODE_NAMES[rt], (24, 12, 10))
        assert err_f <= tol and err_b <= tol


def test_index_map_bit_exact(pf, emu):
    T = pf.Transforms
    for dims, permute, extra in [((16, 12, 6), True, ()), ((5, 7, 3), True, (2,)),
                                 ((9, 7, 5, 12), True, ()), ((9, 7, 5, 12), False, ())]:
        n = int(np.prod(dims)) * int(np.prod(extra))
        vals = np.arange(n, dtype=np.float64)
        meta = OM.PlanMeta(dims, ("NoTransform",), "Float64", permute, extra)
        out = emu.run(dims, [T.NoTransform().code] * len(dims), 1, 0, vals, permute,
                      extra).view(np.float64)
        want = OP.memory_image(OP.to_memory(OP.from_image(vals, tuple(dims) + tuple(extra)),
                                            meta.perm_out))
        assert np.array_equal(out, want)


@pytest.mark.parametrize("dims,trs_name,permute,extra", [
    ((16, 12, 6), "RFFT", True, ()), ((16, 12, 6), "FFT", True, (3,)),
    ((9, 7, 5, 12), "RFFT", True, ()), ((11, 7, 13), "FFT", False, ()),
    ((12, 10, 8), "RFFT", True, (3, 2))],
    ids=["16x12x6-RFFT", "16x12x6-FFT-extra3", "9x7x5x12-RFFT", "11x7x13-FFT-noperm",
         "12x10x8-RFFT-extra3x2"])
def test_tile_traversals(pf, emu, dims, trs_name, permute, extra):
    """Every tile order and direction the autotuner may give a pass
    (tile_split: line dim 0 first or last, ascending or descending) covers
    each line exactly once: outputs bit-identical to the default traversal,
    forward, backward and with a fused spectral operator."""
    T = pf.Transforms
    meta, full = meta_for(T, dims, (getattr(T, trs_name)(),), "Float64", permute, extra)
    kinds = [t.code for t in full]
    rng = np.random.default_rng(5)
    x = random_logical(rng, tuple(dims) + tuple(extra), meta.eltype_in)
    xin = OP.memory_image(OP.to_memory(x, meta.perm_in)).astype(NP_DTYPES[meta.eltype_in])
    op = pf._lib.SpectralOp(pf._lib.OP_LAPLACIAN, 0, 1.0, 0)
    runs = {}
    for trav in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        y = emu.run(dims, kinds, 1, 0, xin, permute, extra, traversal=trav)
        b = emu.run(dims, kinds, 1, 2, y.view(NP_DTYPES[meta.eltype_out]), permute, extra,
                    traversal=trav)
        o = None
        if len(dims) >= 2 and permute:
            o = emu.run_op(dims, kinds, 1, 0, xin, op, permute=permute, extra=extra,
                           traversal=trav)
        runs[trav] = (y, b, o)
    ref = runs[(0, 0)]
    for trav, got in runs.items():
        for a, r in zip(got, ref):
            if r is not None:
                assert np.array_equal(a, r), trav


@pytest.mark.parametrize("dims,trs_name,real_type", [
    ((8, 128, 4), "RFFT", 1), ((9, 256, 2), "RFFT", 0), ((5, 128, 3, 2), "RFFT", 1),
    ((4, 1024, 2), "FFT", 1), ((2, 4, 1024), "FFT", 0)],
    ids=["8x128x4-RFFT-F64", "9x256x2-RFFT-F32", "5x128x3x2-RFFT-F64", "4x1024x2-FFT-F64",
         "2x4x1024-FFT-F32"])
def test_padded_work_buffers(pf, emu, dims, trs_name, real_type):
    """Intermediate buffers with padded rows (a buffer the next pass loads as
    tiles with rows of 2 KiB and more, or rows of 8 KiB and more) give results
    bit-identical to dense rows (PFFT_FLAG_DENSE_WORK), forward and backward,
    and the padding shows in the work-buffer size."""
    T = pf.Transforms
    L = pf._lib
    rt = T.Float64 if real_type == 1 else T.Float32
    padded = pf.PencilFFTPlan(dims, getattr(T, trs_name)(), (1,) * (len(dims) - 1), None, rt)
    dense = pf.PencilFFTPlan(dims, getattr(T, trs_name)(), (1,) * (len(dims) - 1), None, rt,
                             flags=L.FLAG_DENSE_WORK)
    assert padded.work_bytes > dense.work_bytes, (padded.work_bytes, dense.work_bytes)
    meta, full = meta_for(T, dims, (getattr(T, trs_name)(),), CODE_NAMES[real_type])
    kinds = [t.code for t in full]
    rng = np.random.default_rng(11)
    x = random_logical(rng, tuple(dims), meta.eltype_in)
    xin = OP.memory_image(OP.to_memory(x, meta.perm_in)).astype(NP_DTYPES[meta.eltype_in])
    got = {}
    for flags in (0, L.FLAG_DENSE_WORK):
        y = emu.run(dims, kinds, real_type, 0, xin, flags=flags)
        b = emu.run(dims, kinds, real_type, 2, y.view(NP_DTYPES[meta.eltype_out]), flags=flags)
        got[flags] = (y, b)
    for a, r in zip(got[0], got[L.FLAG_DENSE_WORK]):
        assert np.array_equal(a, r)
    # and the padded chain is still the transform
    xw = xin.astype(np.complex128 if np.iscomplexobj(xin) else np.float64)
    ref = OP.memory_image(OP.forward(meta, OP.from_image(xw, meta.mem_dims("in"))))
    err = OP.rel_l2(got[0][0].view(NP_DTYPES[meta.eltype_out]), ref)
    assert err <= tol_for(CODE_NAMES[real_type], dims), err


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims,trs_name,bs", [
    ((1031, 6), "FFT", [4096, 0]),        # prime line: Bluestein at 4096 points
    ((2062, 4), "RFFT", [4096, 0]),       # 1031-point half-length FFT
    ((8, 134), "FFT", [0, 512]),          # 2 x 67 along dim 2
    ((67, 5), "R2R:REDFT10", [256, 0]), ((67, 5), "R2R:REDFT01", [256, 0]),
    ((67, 5), "R2R:REDFT11", [512, 0]),   # 134-point FFT
    ((66, 5), "R2R:RODFT00", [512, 0]),   # 134-point odd extension
    ((67, 5), "R2R:DHT", [256, 0]), ((67, 5), "R2R:RODFT10", [256, 0]),
    ((71, 9), "R2R:REDFT00", [0, 0]),     # 140 = 4 * 5 * 7: direct stages
])
def test_bluestein_lengths(pf, emu, dims, trs_name, bs, real_type):
    """Lines whose length has a prime factor above 64 run the generic pass in
    its Bluestein (chirp-z) form: the oracle's direct transforms at those
    lengths, forward and back (the reference accepts any length, e.g. the
    odd sizes of test/rfft.jl)."""
    T = pf.Transforms
    if trs_name.startswith("R2R:"):
        trs = (T.R2R(getattr(T, trs_name[4:])),)
    else:
        trs = (getattr(T, trs_name)(),)
    err_f, err_b = emu_case(pf, emu, dims, trs, real_type)
    tol = tol_for(CODE_NAMES[real_type], dims)
    assert err_f <= tol and err_b <= tol, (err_f, err_b, tol)
    _, full = meta_for(T, dims, trs, CODE_NAMES[real_type], True, ())
    assert emu.bluestein(dims, [t.code for t in full], real_type) == bs


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims,trs_name,npass", [
    ((16384, 3), "FFT", 3),         # 128 x 128 four-step split, then the 3-point pass
    ((5, 16384), "FFT", 3),         # split on the second (permuted) dim
    ((6, 20000), "RFFT", 3),        # r2c on dim 1, 20000 = 125 x 160 c2c split on dim 2
    ((12288, 2), "BFFT", 3),        # 96 x 128
])
def test_four_step_lengths(pf, emu, dims, trs_name, npass, real_type):
    """c2c lines too long for one LDS-resident pass run as two passes (the
    four-step split M = M1 M2, twiddle w_M^(j2 k1) between them): the
    oracle's transforms at those lengths, forward and back."""
    T = pf.Transforms
    trs = (getattr(T, trs_name)(),)
    err_f, err_b = emu_case(pf, emu, dims, trs, real_type)
    tol = tol_for(CODE_NAMES[real_type], dims)
    assert err_f <= tol and err_b <= tol, (err_f, err_b, tol)
    _, full = meta_for(T, dims, trs, CODE_NAMES[real_type], True, ())
    assert len(emu.bluestein(dims, [t.code for t in full], real_type)) == npass


def test_four_step_with_bluestein(pf, emu):
    """64 x 1031: the split's second pass runs its 1031-point lines in the
    Bluestein form."""
    T = pf.Transforms
    dims = (64 * 1031, 2)
    for rt in (1, 0):
        err_f, err_b = emu_case(pf, emu, dims, (T.FFT(),), rt)
        tol = tol_for(CODE_NAMES[rt], dims)
        assert err_f <= tol and err_b <= tol, (rt, err_f, err_b)
        _, full = meta_for(T, dims, (T.FFT(),), CODE_NAMES[rt], True, ())
        assert emu.bluestein(dims, [t.code for t in full], rt) == [0, 4096, 0]


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims,trs_name,permute,extra", [
    ((16384, 2), "FFT", False, (3,)),     # four-step, unpermuted, extra dims
    ((3, 16384), "BFFT", True, (2,)),
    ((1031, 4), "FFT", False, (2,)),      # Bluestein, unpermuted, extra dims
    ((6, 134, 4), "RFFT", False, ()),
])
def test_long_and_prime_lines_layouts(pf, emu, dims, trs_name, permute, extra, real_type):
    """The four-step split and the Bluestein form under the layout options
    of transforms.jl:414-462 (permute_dims=false) and extra dims
    (plans.jl:164-182)."""
    T = pf.Transforms
    err_f, err_b = emu_case(pf, emu, dims, (getattr(T, trs_name)(),), real_type,
                            permute=permute, extra=extra)
    tol = tol_for(CODE_NAMES[real_type], tuple(dims) + tuple(extra))
    assert err_f <= tol and err_b <= tol, (err_f, err_b, tol)
