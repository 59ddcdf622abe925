"""The C ABI on the CPU (no GPU calls): the library loads and exports every
symbol of include/pfft.h; the plan layer agrees with the oracle's restatement
of the reference's metadata and pencil chain; the error contract of the
reference (ArgumentError conditions) comes back as status codes with the
reference's messages."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import plan as OM

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "pfft.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pfft_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_header_symbol(pf):
    names = header_functions()
    assert len(names) >= 30
    lib = C.CDLL(pf._lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes binding declares each of them
    assert set(names) == set(pf._lib.SIGNATURES), set(names) ^ set(pf._lib.SIGNATURES)


def _name(t):
    """Oracle name of a product transform ("BRFFT!{odd}" -> "BRFFT_odd!")."""
    s = repr(t)
    bang = "!" if "!" in s else ""
    s = s.replace("!", "")
    if s.startswith("R2R{"):
        return "R2R_" + s[4:-1] + bang
    return s.replace("{even}", "_even").replace("{odd}", "_odd") + bang


def all_transforms(T):
    out = [T.FFT(), T.BFFT(), T.RFFT(), T.BRFFT(8), T.BRFFT(9), T.NoTransform(),
           T.FFT_(), T.BFFT_(), T.RFFT_(), T.BRFFT_(8), T.NoTransform_()]
    out += [T.R2R(k) for k in T.R2R_SUPPORTED_KINDS] + [T.R2R_(k) for k in T.R2R_SUPPORTED_KINDS]
    return out


def test_transform_metadata_matches_oracle(pf):
    T = pf.Transforms
    for t in all_transforms(T):
        o = _name(t)
        for n in range(1, 20):
            assert T.length_output(t, n) == OM.length_output(o, n), (t, n)
            assert _name(T.binv(t, n)) == OM.binv(o, n), (t, n)
            if not (o.startswith("R2R_REDFT00") and n < 2):
                assert T.scale_factor(t, (n,)) == OM.scale_factor_1d(o, n), (t, n)
        for rt, rname in ((T.Float32, "Float32"), (T.Float64, "Float64")):
            e = T.eltype_input(t, rt)
            eo = OM.eltype_input(o, rname)
            assert (None if e is None else T.ELTYPE_NAMES[e]) == eo
        for et, ename in T.ELTYPE_NAMES.items():
            try:
                want = OM.eltype_output(o, ename)
            except ValueError as err:
                with pytest.raises(pf.ArgumentError) as ei:
                    T.eltype_output(t, et)
                assert str(ei.value) == str(err)
                continue
            assert T.ELTYPE_NAMES[T.eltype_output(t, et)] == want
        for N in range(0, 5):
            assert tuple(_name(x) for x in T.expand_dims(t, N)) == OM.expand_dims(o, N)


def test_doctest_known_answers(pf):
    T = pf.Transforms
    assert repr(T.binv(T.FFT(), 42)) == "BFFT"
    assert repr(T.binv(T.BRFFT(9), 42)) == "RFFT"
    assert T.is_inplace(T.RFFT()) is False
    assert T.is_inplace(T.NoTransform_()) is True
    assert T.is_inplace(T.FFT_(), T.R2R_(T.REDFT01)) is True
    assert T.is_inplace(T.FFT(), T.R2R(T.REDFT01)) is False
    assert T.is_inplace(T.FFT(), T.R2R_(T.REDFT01)) is None
    C3 = (3, 4, 5)
    assert T.scale_factor(T.FFT(), C3) == 60
    assert T.scale_factor(T.BFFT(), C3) == 60
    assert T.scale_factor(T.BFFT(), C3, range(2, 4)) == 20
    assert T.scale_factor(T.RFFT(), C3, 2) == 4
    assert T.scale_factor(T.RFFT(), C3, range(2, 4)) == 20
    assert T.scale_factor(T.BRFFT(8), C3) == 96
    assert T.scale_factor(T.BRFFT(9), C3) == 108
    assert T.eltype_input(T.FFT(), T.Float32) == T.ComplexF32
    assert T.eltype_input(T.R2R(T.REDFT01), T.Float64) is None
    assert T.eltype_output(T.BRFFT(4), T.ComplexF32) == T.Float32
    with pytest.raises(pf.ArgumentError, match="invalid input data type for "
                                               "PencilFFTs.Transforms.FFT: Float64"):
        T.eltype_output(T.FFT(), T.Float64)
    assert [repr(t) for t in T.expand_dims(T.RFFT(), 3)] == ["RFFT", "FFT", "FFT"]
    assert [repr(t) for t in T.expand_dims(T.BRFFT(4), 3)] == ["BFFT", "BFFT", "BRFFT{even}"]
    assert [repr(t) for t in T.expand_dims(T.NoTransform(), 2)] == ["NoTransform"] * 2
    assert repr(T.R2R(T.REDFT10)) == "R2R{REDFT10}"
    for k in (T.R2HC, T.HC2R):
        with pytest.raises(pf.ArgumentError, match="unsupported r2r transform kind"):
            T.R2R(k)


PLAN_CASES = [
    ((64, 32, 128), "RFFT", True, ()),
    ((16, 12, 6), "RFFT", True, (3,)),
    ((16, 12, 6), "FFT", False, ()),
    ((11, 15), "RFFT", True, ()),
    ((9, 7, 5, 12), "RFFT", True, ()),
    ((9, 7, 5, 12), "RFFT", False, ()),
    ((512, 512, 512), "RFFT", True, ()),
    ((1024, 1024, 1024), "FFT", True, ()),
    ((384, 384, 256), ("RFFT", "FFT", "R2R_REDFT10"), True, ()),
    ((12, 13, 9), "BRFFT_odd", True, ()),
    ((7,), "RFFT", True, ()),
]


def _transforms(T, spec, N):
    def one(s):
        if s == "BRFFT_odd":
            return T.BRFFT(17)
        if s.startswith("R2R_"):
            return T.R2R(getattr(T, s[4:]))
        return getattr(T, s)()
    if isinstance(spec, tuple):
        return tuple(one(s) for s in spec), spec
    t = one(spec)
    return T.expand_dims(t, N), tuple(_name(x) for x in T.expand_dims(t, N))


@pytest.mark.parametrize("dims,spec,permute,extra", PLAN_CASES)
@pytest.mark.parametrize("real", ["Float32", "Float64"])
def test_plan_matches_oracle(pf, dims, spec, permute, extra, real):
    T = pf.Transforms
    trs, names = _transforms(T, spec, len(dims))
    rt = T.Float32 if real == "Float32" else T.Float64
    plan = pf.PencilFFTPlan(dims, trs, (1,), None, rt, extra_dims=extra, permute_dims=permute)
    m = OM.PlanMeta(dims, names, real, permute, extra)
    assert plan.size_global_in == m.size_in and plan.size_global_out == m.size_out
    assert plan.scale_factor == m.scale
    assert T.ELTYPE_NAMES[plan.eltype_input] == m.eltype_in
    assert T.ELTYPE_NAMES[plan.eltype_output] == m.eltype_out
    assert plan.pencil_input.permutation == tuple(p + 1 for p in m.perm_in)
    assert plan.pencil_output.permutation == tuple(p + 1 for p in m.perm_out)
    assert plan.pencil_output.mem_dims() == m.mem_dims("out")
    for n, st in enumerate(m.stages, start=1):
        perm, before, after, ax = plan.stage(n)
        assert perm == tuple(p + 1 for p in st["perm"])
        assert before == st["before"] and after == st["after"]
        assert ax == st["mem_axis"] + 1
    kf, kb = plan.kinds()
    assert tuple(_name(T._from_code(k)) for k in kb) == tuple(s["tr_bwd"] for s in m.stages)
    esz = {"Float32": 4, "Float64": 8, "ComplexF32": 8, "ComplexF64": 16}
    ne = int(np.prod(extra)) if extra else 1
    assert plan.in_bytes == int(np.prod(m.size_in)) * ne * esz[m.eltype_in]
    assert plan.out_bytes == int(np.prod(m.size_out)) * ne * esz[m.eltype_out]


def test_describe_matches_global_params_show(pf):
    T = pf.Transforms
    plan = pf.PencilFFTPlan((64, 32, 128), T.RFFT(), (1, 1), None, T.Float64)
    text = plan.describe()
    # global_params.jl:31-40 (doctest output)
    assert text.startswith("Transforms: (RFFT, FFT, FFT)\nInput type: Float64\n"
                           "Global dimensions: (64, 32, 128) -> (33, 32, 128)")
    plan = pf.PencilFFTPlan((16, 12, 6), T.RFFT(), (1, 1), None, T.Float32, extra_dims=(3,))
    assert "Extra dimensions: (3,)" in plan.describe()


def test_error_contract(pf):
    T = pf.Transforms
    L = pf._lib
    # global_params.jl:57-60
    with pytest.raises(pf.ArgumentError, match=r"cannot combine in-place and out-of-place "
                                                r"transforms: \(FFT, FFT!, FFT!\)"):
        pf.PencilFFTPlan((16, 12, 6), (T.FFT(), T.FFT_(), T.FFT_()), (1,), None, T.Float64)
    # Transforms.jl:263-264 via the stage chain: a second RFFT gets complex data
    with pytest.raises(pf.ArgumentError, match="invalid input data type for "
                                               "PencilFFTs.Transforms.RFFT: ComplexF64"):
        pf.PencilFFTPlan((16, 12), (T.RFFT(), T.RFFT()), (1,), None, T.Float64)
    # plans.jl:367-372
    with pytest.raises(pf.ArgumentError, match="number of transforms"):
        pf.PencilFFTPlan((16, 12, 6), (T.FFT(), T.FFT()), (1,), None, T.Float64)
    # one GPU: proc_dims must be a single process
    with pytest.raises(pf.UnsupportedError, match="one GPU"):
        pf.PencilFFTPlan((16, 12, 6), T.FFT(), (2, 1), None, T.Float64)
    # in-place chains other than the reference's ManyPencilArray forms
    # (allocate.jl:57-83): RFFT! must be followed by FFT!s, c2r is not in-place
    with pytest.raises(pf.UnsupportedError, match="in-place plans take"):
        pf.PencilFFTPlan((16, 12, 6), (T.RFFT_(), T.BFFT_(), T.FFT_()), (1,), None, T.Float64)
    with pytest.raises(pf.UnsupportedError, match="in-place plans take"):
        pf.PencilFFTPlan((16, 12, 6), T.BRFFT_(6), (1,), None, T.Float64)
    # REDFT00 needs length >= 2
    with pytest.raises(pf.ArgumentError, match="REDFT00"):
        pf.PencilFFTPlan((1, 4), T.R2R(T.REDFT00), (1,), None, T.Float64)


def test_execute_argument_checks_without_gpu(pf):
    """operations.jl:91-93: aliased arrays are rejected before any device work;
    bad directions and null pointers too."""
    L = pf._lib
    T = pf.Transforms
    plan = pf.PencilFFTPlan((16, 12, 6), T.FFT(), (1,), None, T.Float64)
    h = plan._h
    base = 1 << 40
    st = L.lib.pfft_execute(h, L.FORWARD, C.c_void_p(base), C.c_void_p(base + 64), None, 0, None)
    assert st == L.ERR_ARGUMENT
    assert L.lib.pfft_last_error().decode() == "out-of-place plan applied to aliased data"
    st = L.lib.pfft_execute(h, 7, C.c_void_p(base), C.c_void_p(base << 1), None, 0, None)
    assert st == L.ERR_ARGUMENT and "direction" in L.lib.pfft_last_error().decode()
    st = L.lib.pfft_execute(h, L.FORWARD, None, C.c_void_p(base), None, 0, None)
    assert st == L.ERR_ARGUMENT
    # python-side collection length check (operations.jl:130-137): the message
    # is length(dst) ≠ length(src), as check_compatible(dst, src) formats it
    with pytest.raises(pf.ArgumentError, match="collections have different lengths: 2 ≠ 3"):
        pf.mul_([1, 2], plan, [1, 2, 3])


def test_execute_batch_argument_checks_without_gpu(pf):
    """pfft_execute_batch: bad counts and items are refused before any device
    work; a one-plan batch applies pfft_execute's checks to each item."""
    L = pf._lib
    T = pf.Transforms
    plan = pf.PencilFFTPlan((16, 12, 6), T.FFT(), (1,), None, T.Float64)
    assert L.lib.pfft_execute_batch(0, None, None) == L.OK
    assert L.lib.pfft_execute_batch(-1, None, None) == L.ERR_ARGUMENT
    assert L.lib.pfft_execute_batch(1, None, None) == L.ERR_ARGUMENT
    items = (L.BatchItem * 2)()
    items[0].plan = plan._h
    items[0].direction = L.FORWARD
    base = 1 << 40
    items[0].inp, items[0].out = base, base << 1
    assert L.lib.pfft_execute_batch(2, items, None) == L.ERR_ARGUMENT   # items[1].plan is null
    assert "null plan" in L.lib.pfft_last_error().decode()
    items[1] = items[0]
    items[0].out = base + 64     # the first item aliases its input
    st = L.lib.pfft_execute_batch(2, items, None)
    assert st == L.ERR_ARGUMENT
    assert L.lib.pfft_last_error().decode() == "out-of-place plan applied to aliased data"


def test_tune_data_round_trip_and_refusal(pf):
    """pfft_tune_export / pfft_tune_import on the host: the export carries the
    build and device header, imports back, and data of another build, a bad
    line or a foreign format is refused with ArgumentError."""
    text = pf.export_wisdom()
    head = text.split("\n")[:3]
    assert head[0] == "pfft-tune 1" and head[1].startswith("build ") and \
        head[2].startswith("device ")
    pf.import_wisdom(text)
    assert pf.export_wisdom() == text
    with pytest.raises(pf.ArgumentError):
        pf.import_wisdom("pfft-tune 1\nbuild 0-0000000000000000\ndevice none\n")
    with pytest.raises(pf.ArgumentError):
        pf.import_wisdom("\n".join(head) + "\nnot-a-number\tkey\n")
    pf.import_wisdom("\n".join(head) + "\n7\tsome:key\n")
    assert "7\tsome:key" in pf.export_wisdom()
    pf.forget_wisdom()
    assert pf.export_wisdom() == "\n".join(head) + "\n"
