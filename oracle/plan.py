"""PLAN-layer metadata of PencilFFTs.jl v0.15.3, restated (oracle side).

Transforms are named by strings:
  "FFT", "BFFT", "RFFT", "BRFFT_even", "BRFFT_odd", "NoTransform",
  "R2R_<KIND>" with KIND one of the FFTW r2r kinds below,
and a trailing "!" marks the in-place variant (FFT!, ...).
Element types: "Float32", "Float64", "ComplexF32", "ComplexF64".
"""

R2R_KINDS = ("DHT", "REDFT00", "REDFT01", "REDFT10", "REDFT11",
             "RODFT00", "RODFT01", "RODFT10", "RODFT11")   # r2r.jl:5-15


def base(tr):
    return tr[:-1] if tr.endswith("!") else tr


def is_inplace(tr):
    """Transforms.jl:139-147 for one transform."""
    return tr.endswith("!")


def is_inplace_list(trs):
    """Tuple form (Transforms.jl:149-155): True/False, or None when mixed."""
    vals = {is_inplace(t) for t in trs}
    return vals.pop() if len(vals) == 1 else None


def check(tr):
    b = base(tr)
    if b in ("FFT", "BFFT", "RFFT", "BRFFT_even", "BRFFT_odd", "NoTransform"):
        return
    if b.startswith("R2R_"):
        k = b[4:]
        if k in ("R2HC", "HC2R"):
            raise ValueError("unsupported r2r transform kind: " + k)   # r2r.jl:42-45
        if k in R2R_KINDS:
            return
    raise ValueError("unknown transform " + tr)


def length_output(tr, n):
    """c2c.jl:49, r2c.jl:74-75, r2r.jl:82, no_transform.jl:23"""
    b = base(tr)
    if b == "RFFT":
        return n // 2 + 1
    if b == "BRFFT_even":
        return 2 * n - 2
    if b == "BRFFT_odd":
        return 2 * n - 1
    return n


def brfft(d):
    """BRFFT(d) = BRFFT(iseven(d)) (r2c.jl:66)."""
    return "BRFFT_even" if d % 2 == 0 else "BRFFT_odd"


_R2R_INV = {"REDFT00": "REDFT00", "REDFT01": "REDFT10", "REDFT10": "REDFT01",
            "REDFT11": "REDFT11", "RODFT00": "RODFT00", "RODFT01": "RODFT10",
            "RODFT10": "RODFT01", "RODFT11": "RODFT11", "DHT": "DHT"}


def binv(tr, d):
    """c2c.jl:58-61, r2c.jl:99-102, r2r.jl:103-131, no_transform.jl:21"""
    bang = "!" if is_inplace(tr) else ""
    b = base(tr)
    if b == "FFT":
        r = "BFFT"
    elif b == "BFFT":
        r = "FFT"
    elif b == "RFFT":
        r = brfft(d)
    elif b.startswith("BRFFT"):
        r = "RFFT"
    elif b == "NoTransform":
        r = "NoTransform"
    else:
        r = "R2R_" + _R2R_INV[b[4:]]
    return r + bang


def scale_factor_1d(tr, n):
    """Factor for one dimension of input length n: c2c.jl:71, r2c.jl:104-111,
    r2r.jl:96-101, no_transform.jl:25."""
    b = base(tr)
    if b == "NoTransform":
        return 1
    if b in ("FFT", "BFFT", "RFFT"):
        return n
    if b.startswith("BRFFT"):
        return length_output(b, n)
    k = b[4:]
    if k == "REDFT00":
        return 2 * (n - 1)
    if k == "RODFT00":
        return 2 * (n + 1)
    if k == "DHT":
        return n
    return 2 * n


def scale_factor_array(tr, shape, dims):
    """scale_factor(tr, A, dims) for an array of `shape` (1-based dims);
    the c2r case uses the output length along the last of `dims`."""
    b = base(tr)
    s = 1
    dims = list(dims)
    for i in dims:
        n = shape[i - 1]
        if b.startswith("BRFFT"):
            s *= length_output(b, n) if i == dims[-1] else n
        elif b.startswith("R2R_"):
            s *= scale_factor_1d(b, n)
        elif b == "NoTransform":
            s *= 1
        else:
            s *= n
    return s


def complex_of(t):
    return {"Float32": "ComplexF32", "Float64": "ComplexF64"}.get(t, t)


def real_of(t):
    return {"ComplexF32": "Float32", "ComplexF64": "Float64"}.get(t, t)


def is_complex(t):
    return t.startswith("Complex")


def eltype_input(tr, real_type):
    """c2c.jl:51, r2c.jl:80-81, r2r.jl:83; None means `nothing`."""
    b = base(tr)
    if b in ("FFT", "BFFT") or b.startswith("BRFFT"):
        return complex_of(real_type)
    if b == "RFFT":
        return real_type
    return None


_TYPE_NAMES = {"FFT": "FFT", "BFFT": "BFFT", "RFFT": "RFFT", "BRFFT_even": "BRFFT",
               "BRFFT_odd": "BRFFT", "NoTransform": "NoTransform"}


def eltype_output(tr, t):
    """c2c.jl:50, r2c.jl:77-78, r2r.jl:84; Transforms.jl:263-264 error."""
    b = base(tr)
    ok, out = True, t
    if b in ("FFT", "BFFT"):
        ok = is_complex(t)
    elif b == "RFFT":
        ok, out = not is_complex(t), complex_of(t)
    elif b.startswith("BRFFT"):
        ok, out = is_complex(t), real_of(t)
    if not ok:
        name = _TYPE_NAMES[b] + ("!" if is_inplace(tr) else "")
        raise ValueError("invalid input data type for PencilFFTs.Transforms.%s: %s" % (name, t))
    return out


def expand_dims(tr, N):
    """Transforms.jl:285-291, r2c.jl:114-124"""
    b = base(tr)
    bang = "!" if is_inplace(tr) else ""
    if b == "RFFT":
        return tuple([tr] + ["FFT" + bang] * (N - 1)) if N else ()
    if b.startswith("BRFFT"):
        return tuple(["BFFT" + bang] * (N - 1) + [tr]) if N else ()
    return (tr,) * N


def input_data_type(real_type, trs):
    """global_params.jl:81-95"""
    for t in trs:
        e = eltype_input(t, real_type)
        if e is not None:
            return e
    return real_type


def stage_perm(N, n, permute=True):
    """0-based memory permutation of stage n's pencils (plans.jl:451-521):
    stage 0 unpermuted; stage n >= 1 (n, 0..n-1, n+1..N-1); last reversed."""
    if not permute or n == 0:
        return tuple(range(N))
    if n == N - 1:
        return tuple(range(N - 1, -1, -1))
    return (n,) + tuple(range(n)) + tuple(range(n + 1, N))


class PlanMeta:
    """GlobalFFTParams (global_params.jl:50-62) + the stage chain of
    PencilFFTPlan (plans.jl:241-293, 399-553) for one process."""

    def __init__(self, size_in, transforms, real_type="Float64", permute=True, extra_dims=()):
        self.N = N = len(size_in)
        if len(transforms) == 1 and N > 1:
            transforms = expand_dims(transforms[0], N)
        assert len(transforms) == N
        for t in transforms:
            check(t)
        if is_inplace_list(transforms) is None:
            raise ValueError("cannot combine in-place and out-of-place transforms: (%s)"
                             % ", ".join(t.replace("_even", "{even}").replace("_odd", "{odd}")
                                         for t in transforms))
        self.transforms = tuple(transforms)
        self.size_in = tuple(size_in)
        self.size_out = tuple(length_output(t, n) for t, n in zip(transforms, size_in))
        self.extra_dims = tuple(extra_dims)
        self.permute = permute
        self.real_type = real_type
        self.eltype_in = input_data_type(real_type, transforms)
        self.stages = []
        T = self.eltype_in
        scale = 1.0
        for n in range(N):
            tr = transforms[n]
            To = eltype_output(tr, T)
            before = tuple(self.size_out[j] if j < n else size_in[j] for j in range(N))
            after = tuple(self.size_out[j] if j <= n else size_in[j] for j in range(N))
            perm = stage_perm(N, n, permute)
            d = size_in[n]   # input length along the transformed dim
            tb = binv(tr, d)
            sc = scale_factor_1d(tb, self.size_out[n])
            scale *= float(sc)
            self.stages.append(dict(dim=n, tr=tr, tr_bwd=tb, T_in=T, T_out=To, perm=perm,
                                    before=before, after=after, mem_axis=perm.index(n),
                                    scale=sc))
            T = To
        self.eltype_out = T
        self.scale = scale

    @property
    def perm_in(self):
        return self.stages[0]["perm"]

    @property
    def perm_out(self):
        return self.stages[-1]["perm"]

    def mem_dims(self, which):
        st = self.stages[-1] if which == "out" else self.stages[0]
        dims = self.size_out if which == "out" else self.size_in
        return tuple(dims[p] for p in st["perm"]) + self.extra_dims
