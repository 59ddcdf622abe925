"""In-place plans through the host emulator: FFT!/BFFT!/R2R!/NoTransform!
chains and RFFT! followed by FFT!s.

One buffer holds the input view (the first pencil; for RFFT! the real data in
a parent padded to 2 (n/2 + 1) along memory axis 0) and the output view (the
last pencil), both at offset 0 -- the reference's ManyPencilArray and
ManyPencilArrayRFFT! (allocate.jl:57-83, multiarrays_r2c.jl:26-79).  The
forward transform done in place must equal the oracle of the out-of-place
chain, and the in-place ldiv! must restore the input (test/transforms.jl:
187-236, test/rfft.jl:194-251)."""
import numpy as np
import pytest

from oracle import pipeline as OP
from tests.cases import CODE_NAMES, NP_DTYPES, expand, meta_for, random_logical, tol_for


def _oop(T, L, t):
    return T._from_code(t.code & ~L.INPLACE)


def inplace_case(pf, emu, dims, trs, real_type, permute=True, extra=(), seed=0):
    T, L = pf.Transforms, pf._lib
    full = expand(T, trs, len(dims))
    oop = tuple(_oop(T, L, t) for t in full)
    meta, _ = meta_for(T, dims, oop, CODE_NAMES[real_type], permute, extra)
    kinds = [t.code for t in full]
    dt_in, dt_out = NP_DTYPES[meta.eltype_in], NP_DTYPES[meta.eltype_out]
    r2c = (full[0].code & ~L.INPLACE) == L.RFFT
    nbytes, nbytes_out = emu.nbytes(dims, kinds, real_type, permute, extra)
    assert nbytes == nbytes_out
    rng = np.random.default_rng(seed)
    x = random_logical(rng, tuple(dims) + tuple(extra), meta.eltype_in)
    xm = OP.to_memory(x, meta.perm_in).astype(dt_in)
    parent = list(xm.shape)
    if r2c:
        parent[0] = 2 * (dims[0] // 2 + 1)
    P = np.zeros(parent, dtype=dt_in)
    P[tuple(slice(0, n) for n in xm.shape)] = xm
    img = OP.memory_image(P)
    buf = np.zeros(nbytes, dtype=np.uint8)
    buf[:img.nbytes] = img.view(np.uint8)
    # forward in place
    emu.run_inplace(dims, kinds, real_type, [0], buf, permute, extra)
    n_out = int(np.prod(meta.mem_dims("out")))
    out = buf.view(dt_out)[:n_out]
    xw = xm.astype(np.complex128 if np.iscomplexobj(xm) else np.float64)
    ref = OP.memory_image(OP.forward(meta, xw))
    err_f = OP.rel_l2(out, ref)
    # normalised backward in place: the input view comes back
    emu.run_inplace(dims, kinds, real_type, [2], buf, permute, extra)
    back = OP.from_image(buf.view(dt_in)[:int(np.prod(parent))], parent)
    back = back[tuple(slice(0, n) for n in xm.shape)]
    err_b = OP.rel_l2(OP.memory_image(back), OP.memory_image(xw))
    return err_f, err_b, nbytes


CHAINS = [
    ("FFT!", lambda T: (T.FFT_(),)),
    ("BFFT!", lambda T: (T.BFFT_(),)),
    ("RFFT!", lambda T: (T.RFFT_(),)),
    ("R2R!", lambda T: (T.R2R_(T.REDFT10), T.R2R_(T.RODFT01), T.R2R_(T.DHT))),
    ("mixed!", lambda T: (T.FFT_(), T.NoTransform_(), T.BFFT_())),
    ("NoTransform!", lambda T: (T.NoTransform_(),)),
]


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("name,mk", CHAINS, ids=[c[0] for c in CHAINS])
@pytest.mark.parametrize("dims", [(16, 12, 6), (15, 7, 5)], ids=["16x12x6", "15x7x5"])
def test_inplace_chains(pf, emu, dims, name, mk, real_type):
    T = pf.Transforms
    err_f, err_b, _ = inplace_case(pf, emu, dims, mk(T), real_type, seed=3)
    tol = tol_for(CODE_NAMES[real_type], dims)
    assert err_f <= tol and err_b <= tol, (err_f, err_b)


@pytest.mark.parametrize("dims,permute,extra", [((11, 15), True, ()), ((9, 7, 5, 12), True, ()),
                                                ((9, 7, 5, 12), False, ()),
                                                ((16, 12, 6), True, (3,)),
                                                ((13,), True, ()), ((12, 10), False, (2,))])
def test_inplace_shapes(pf, emu, dims, permute, extra):
    """N = 1, 2, 4, permute_dims off, extra dims (transforms.jl:414-462 forms)."""
    T = pf.Transforms
    for trs in [(T.FFT_(),), (T.RFFT_(),)]:
        for rt in (1, 0):
            err_f, err_b, _ = inplace_case(pf, emu, dims, trs, rt, permute, extra, seed=5)
            tol = tol_for(CODE_NAMES[rt], dims)
            assert err_f <= tol and err_b <= tol, (trs, rt, err_f, err_b)


def test_inplace_buffer_is_the_largest_pencil(pf, emu):
    """ManyPencilArray holds the largest pencil; ManyPencilArrayRFFT! holds
    2 (n/2 + 1) reals per first-axis line (multiarrays_r2c.jl:57)."""
    T = pf.Transforms
    nb, _ = emu.nbytes((16, 12, 6), [T.FFT_().code] * 3, 1)
    assert nb == 16 * 12 * 6 * 16
    full = T.expand_dims(T.RFFT_(), 3)
    nb, _ = emu.nbytes((15, 12, 6), [t.code for t in full], 0)
    assert nb == 2 * (15 // 2 + 1) * 12 * 6 * 4


def test_rfft_inplace_impulse(pf, emu):
    """test/rfft.jl:205-233: x[1] = 1, x[2] = 2; RFFT! then ldiv! gives back
    [1, 2, 0, ...], and the spectrum equals the out-of-place RFFT's."""
    T, L = pf.Transforms, pf._lib
    dims = (16, 12, 6)
    full = T.expand_dims(T.RFFT_(), 3)
    kinds = [t.code for t in full]
    pad = 2 * (dims[0] // 2 + 1)
    nb, _ = emu.nbytes(dims, kinds, 1)
    buf = np.zeros(nb, dtype=np.uint8)
    P = np.zeros((pad,) + dims[1:])
    P[0, 0, 0], P[1, 0, 0] = 1.0, 2.0
    img = OP.memory_image(P)
    buf[:img.nbytes] = img.view(np.uint8)
    emu.run_inplace(dims, kinds, 1, [0], buf)
    # closed form of the spectrum of the impulse pair (the rfft.jl impulse test)
    X = dims[0] // 2 + 1
    k = np.arange(X)
    spec_x = 1.0 + 2.0 * np.exp(-2j * np.pi * k / dims[0])
    full_spec = np.zeros((X,) + dims[1:], dtype=complex)
    full_spec[:] = spec_x[:, None, None]
    meta, _ = meta_for(T, dims, (T.RFFT(),), "Float64")
    want = OP.memory_image(OP.to_memory(full_spec, meta.perm_out))
    got = buf.view(np.complex128)[:want.size]
    assert np.max(np.abs(got - want)) < 1e-12
    emu.run_inplace(dims, kinds, 1, [2], buf)
    back = OP.from_image(buf.view(np.float64)[:P.size], P.shape)[:dims[0]]
    assert np.allclose(back[:3, 0, 0], [1.0, 2.0, 0.0], atol=1e-13)
    assert np.max(np.abs(back[:, 1:, :])) < 1e-13 and np.max(np.abs(back[:, :, 1:])) < 1e-13


def test_inplace_requires_one_buffer(pf, emu):
    """operations.jl:100-105: in-place plans take the same array as input and
    output."""
    T = pf.Transforms
    L = pf._lib
    import ctypes as C
    h = C.c_void_p()
    dims = (8, 6, 4)
    st = emu.lib.pfft_emu_plan_create(C.byref(h), 3, (C.c_int64 * 3)(*dims),
                                      (C.c_int32 * 3)(*[T.FFT_().code] * 3), 1, 1, 0,
                                      (C.c_int64 * 1)(), 0)
    assert st == 0
    try:
        a = np.zeros(8 * 6 * 4 * 16, dtype=np.uint8)
        b = np.zeros_like(a)
        st = emu.lib.pfft_emu_execute(h, 0, a.ctypes.data, b.ctypes.data)
        assert st == L.ERR_ARGUMENT
        assert "must be the same" in emu.lib.pfft_emu_last_error().decode()
    finally:
        emu.lib.pfft_emu_plan_destroy(h)


@pytest.mark.parametrize("real_type", [1, 0], ids=["F64", "F32"])
@pytest.mark.parametrize("dims", [(16384, 3), (4, 12288), (1031, 6)])
def test_inplace_long_and_prime_lines(pf, emu, dims, real_type):
    """In-place FFT! plans whose lines take the four-step split (16384,
    12288) or the Bluestein form (1031): the split's intermediate moves
    through the scratch buffer like any transposing pass."""
    T = pf.Transforms
    err_f, err_b, _ = inplace_case(pf, emu, dims, (T.FFT_(),), real_type, seed=5)
    tol = tol_for(CODE_NAMES[real_type], dims)
    assert err_f <= tol and err_b <= tol, (err_f, err_b)
