"""Transforms: one-dimensional transform types and their metadata.

Mirror of PencilFFTs.Transforms (src/Transforms/Transforms.jl:11-305).  Julia's
in-place names carry a `!`; here they carry a trailing underscore (FFT_ is
FFT!).  The metadata functions call the C ABI, which restates the reference's
rules (c2c.jl, r2c.jl, r2r.jl, no_transform.jl).
"""
import ctypes as C

from . import _lib as L

# FFTW r2r kind codes (FFTW.REDFT00 etc.), as in fftw3.h
R2HC, HC2R, DHT = 0, 1, 2
REDFT00, REDFT01, REDFT10, REDFT11 = 3, 4, 5, 6
RODFT00, RODFT01, RODFT10, RODFT11 = 7, 8, 9, 10
_KIND_NAMES = {R2HC: "R2HC", HC2R: "HC2R", DHT: "DHT", REDFT00: "REDFT00", REDFT01: "REDFT01",
               REDFT10: "REDFT10", REDFT11: "REDFT11", RODFT00: "RODFT00",
               RODFT01: "RODFT01", RODFT10: "RODFT10", RODFT11: "RODFT11"}
R2R_SUPPORTED_KINDS = (DHT, REDFT00, REDFT01, REDFT10, REDFT11,
                       RODFT00, RODFT01, RODFT10, RODFT11)   # r2r.jl:5-15

Float32, Float64, ComplexF32, ComplexF64 = L.F32, L.F64, L.C32, L.C64
ELTYPE_NAMES = {L.F32: "Float32", L.F64: "Float64", L.C32: "ComplexF32", L.C64: "ComplexF64"}


def kind2string(kind):
    return _KIND_NAMES[kind]


class AbstractTransform:
    """A one-dimensional FFT-based transform (Transforms.jl:31)."""

    code = None  # C ABI transform code

    def __eq__(self, other):
        return isinstance(other, AbstractTransform) and self.code == other.code

    def __hash__(self):
        return hash(self.code)

    def __repr__(self):
        buf = C.create_string_buffer(64)
        L.check(L.lib.pfft_transform_name(self.code, buf, 64))
        return buf.value.decode()

    __str__ = __repr__


def _simple(name, code, doc):
    cls = type(name, (AbstractTransform,), {"code": code, "__doc__": doc})
    return cls


FFT = _simple("FFT", L.FFT, "Complex-to-complex FFT (c2c.jl:11).")
FFT_ = _simple("FFT_", L.FFT | L.INPLACE, "In-place FFT! (c2c.jl:21).")
BFFT = _simple("BFFT", L.BFFT, "Unnormalised backward c2c FFT (c2c.jl:34).")
BFFT_ = _simple("BFFT_", L.BFFT | L.INPLACE, "In-place BFFT! (c2c.jl:45).")
RFFT = _simple("RFFT", L.RFFT, "Real-to-complex FFT (r2c.jl:12).")
RFFT_ = _simple("RFFT_", L.RFFT | L.INPLACE, "In-place RFFT! (r2c.jl:19).")
NoTransform = _simple("NoTransform", L.NOTRANSFORM, "Identity transform (no_transform.jl:8).")
NoTransform_ = _simple("NoTransform_", L.NOTRANSFORM | L.INPLACE,
                       "In-place identity transform (no_transform.jl:15).")


class _BRFFTBase(AbstractTransform):
    inplace = 0

    def __init__(self, d):
        """BRFFT(d::Integer) or BRFFT((d1, ..., dN)): only the parity of the
        output length along the **last** dimension matters (r2c.jl:66-69)."""
        if isinstance(d, bool):
            even = d
        elif isinstance(d, (tuple, list)):
            even = int(d[-1]) % 2 == 0
        else:
            even = int(d) % 2 == 0
        self.even_output = bool(even)
        self.code = (L.BRFFT_EVEN if even else L.BRFFT_ODD) | self.inplace


class BRFFT(_BRFFTBase):
    """Unnormalised inverse of RFFT (r2c.jl:48-51)."""


class BRFFT_(_BRFFTBase):
    """In-place BRFFT! (r2c.jl:58-61)."""
    inplace = L.INPLACE


class R2R(AbstractTransform):
    """Real-to-real transform of FFTW kind `kind` (r2r.jl:40-48)."""
    inplace = 0

    def __init__(self, kind):
        self.kind = int(kind)
        code = (L.R2R + self.kind) | self.inplace
        if not (0 <= self.kind <= RODFT11):
            raise L.ArgumentError(L.ERR_ARGUMENT, "unknown r2r kind %r" % (kind,))
        L.check(L.lib.pfft_transform_check(code))
        self.code = code


class R2R_(R2R):
    """In-place R2R! (r2r.jl:55-61)."""
    inplace = L.INPLACE


def _from_code(code):
    base = code & ~L.INPLACE
    ip = bool(code & L.INPLACE)
    if base == L.FFT:
        return FFT_() if ip else FFT()
    if base == L.BFFT:
        return BFFT_() if ip else BFFT()
    if base == L.RFFT:
        return RFFT_() if ip else RFFT()
    if base in (L.BRFFT_EVEN, L.BRFFT_ODD):
        return (BRFFT_ if ip else BRFFT)(base == L.BRFFT_EVEN)
    if base == L.NOTRANSFORM:
        return NoTransform_() if ip else NoTransform()
    return (R2R_ if ip else R2R)(base - L.R2R)


def binv(transform, d):
    """Backward transform of `transform`; d is the forward input length along
    the first transformed dimension (Transforms.jl:82-108)."""
    out = C.c_int32()
    L.check(L.lib.pfft_transform_binv(transform.code, int(d), C.byref(out)))
    return _from_code(out.value)


def length_output(transform, length_in):
    out = C.c_int64()
    L.check(L.lib.pfft_transform_length_output(transform.code, int(length_in), C.byref(out)))
    return out.value


def is_inplace(*transforms):
    """True/False, or None for a mixed list (Transforms.jl:139-155)."""
    arr = (C.c_int32 * len(transforms))(*[t.code for t in transforms])
    out = C.c_int32()
    L.check(L.lib.pfft_transforms_is_inplace(len(transforms), arr, C.byref(out)))
    return None if out.value < 0 else bool(out.value)


def scale_factor(transform, shape, dims=None):
    """scale_factor(transform, A, dims) with A given by its shape (1-based dims,
    default all); the c2r length applies to the last of `dims`
    (Transforms.jl:195-197, r2c.jl:104-111)."""
    shape = tuple(shape)
    if dims is None:
        dims = range(1, len(shape) + 1)
    elif isinstance(dims, int):
        dims = (dims,)
    dims = list(dims)
    base = transform.code & ~L.INPLACE
    s = 1
    for i in dims:
        n = shape[i - 1]
        if base in (L.BRFFT_EVEN, L.BRFFT_ODD) and i != dims[-1]:
            s *= n
            continue
        out = C.c_int64()
        L.check(L.lib.pfft_transform_scale_factor(transform.code, int(n), C.byref(out)))
        s *= out.value
    return s


def eltype_input(transform, real_type):
    """Input element type code for a real type, or None (Transforms.jl:199-225)."""
    out = C.c_int32()
    L.check(L.lib.pfft_transform_eltype_input(transform.code, int(real_type), C.byref(out)))
    return None if out.value == L.NOTHING else out.value


def eltype_output(transform, eltype_in):
    """Output element type code; raises ArgumentError on an invalid input type
    (Transforms.jl:227-264)."""
    out = C.c_int32()
    L.check(L.lib.pfft_transform_eltype_output(transform.code, int(eltype_in), C.byref(out)))
    return out.value


def expand_dims(transform, N):
    """Expand one transform into N (Transforms.jl:266-291, r2c.jl:114-124)."""
    arr = (C.c_int32 * max(N, 1))()
    L.check(L.lib.pfft_transform_expand_dims(transform.code, int(N), arr))
    return tuple(_from_code(arr[i]) for i in range(N))


def kind(transform):
    """FFTW kind of an R2R transform (r2r.jl:78-80)."""
    return transform.kind
