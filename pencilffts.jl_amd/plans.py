"""PencilFFTPlan and the arrays it transforms, on one GPU.

Mirror of src/plans.jl, src/allocate.jl and src/operations.jl of
PencilFFTs.jl v0.15.3.  A plan wraps a C-ABI plan (include/pfft.h); arrays are
torch tensors on the GPU holding the column-major memory image of a
PencilArray.  Julia's `!` functions are spelled with a trailing underscore:
mul_ = mul!, ldiv_ = ldiv!, bmul_ = PencilFFTs.bmul!.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import transforms as Tr

_TORCH_DTYPES = {L.F32: torch.float32, L.F64: torch.float64,
                 L.C32: torch.complex64, L.C64: torch.complex128}
_NAMES = Tr.ELTYPE_NAMES


def _real_type_code(t):
    if t in (L.F32, L.F64):
        return t
    if t in (np.float32, torch.float32, "Float32", "float32"):
        return L.F32
    if t in (np.float64, torch.float64, float, "Float64", "float64", None):
        return L.F64
    raise L.ArgumentError(L.ERR_ARGUMENT, "real_type must be Float32 or Float64, got %r" % (t,))


class Pencil:
    """Layout of one array of a plan: logical global size, memory permutation
    (1-based, memory axis i holds logical dim perm[i]) and extra dims."""

    def __init__(self, size_global, permutation, extra_dims=()):
        self.size_global = tuple(int(s) for s in size_global)
        self.permutation = tuple(int(p) for p in permutation)
        self.extra_dims = tuple(int(e) for e in extra_dims)

    def mem_dims(self):
        return tuple(self.size_global[p - 1] for p in self.permutation) + self.extra_dims

    def __eq__(self, other):
        return (isinstance(other, Pencil) and self.size_global == other.size_global
                and self.permutation == other.permutation
                and self.extra_dims == other.extra_dims)

    def __repr__(self):
        return "Pencil(size_global=%s, permutation=%s, extra_dims=%s)" % (
            self.size_global, self.permutation, self.extra_dims)


class PencilArray:
    """A GPU array on a pencil.

    `parent` is a torch tensor with the memory dims in reverse (torch is
    row-major), so `parent.reshape(-1)` is the column-major memory image of
    the Julia array.  It is contiguous, except for the real input view of an
    in-place RFFT! plan, a slice of its padded parent.  `logical()` is a view
    indexed in logical order (i1, i2, ..., e1, ...), like indexing a
    PencilArray.
    """

    def __init__(self, pencil, eltype, parent):
        self.pencil = pencil
        self.eltype = eltype
        self.parent = parent

    @classmethod
    def empty(cls, pencil, eltype, device="cuda"):
        dims = pencil.mem_dims()
        t = torch.empty(tuple(reversed(dims)), dtype=_TORCH_DTYPES[eltype], device=device)
        return cls(pencil, eltype, t)

    def logical(self):
        perm = self.pencil.permutation
        N = len(perm)
        ne = len(self.pencil.extra_dims)
        nd = N + ne
        # torch axis of memory axis i is nd-1-i; logical dim d is memory axis perm.index(d+1)
        order = [nd - 1 - perm.index(d + 1) for d in range(N)]
        order += [nd - 1 - (N + e) for e in range(ne)]
        return self.parent.permute(*order)

    def size_global(self):
        return self.pencil.size_global

    @property
    def permutation(self):
        return self.pencil.permutation

    def data_ptr(self):
        return self.parent.data_ptr()

    def nbytes(self):
        return self.parent.numel() * self.parent.element_size()

    def similar(self):
        return PencilArray(self.pencil, self.eltype, torch.empty_like(self.parent))

    def copy(self):
        return PencilArray(self.pencil, self.eltype, self.parent.clone())

    def __repr__(self):
        return "PencilArray{%s}(%s)" % (_NAMES[self.eltype], self.pencil)


class ManyPencilArray:
    """One GPU buffer viewed on the pencils of an in-place plan: PencilArrays'
    ManyPencilArray, or ManyPencilArrayRFFT! for RFFT! (allocate.jl:57-83,
    multiarrays_r2c.jl:26-79).

    `first()` is the input view (the first pencil; for RFFT! real data whose
    parent is padded to 2 (n/2 + 1) along memory axis 0, so it is a strided
    view) and `last()` the output view (the last pencil); both start at the
    buffer's first byte, and transforming in place moves the data from one to
    the other."""

    def __init__(self, plan, device="cuda"):
        self.data = torch.empty(max(plan.in_bytes, 1), dtype=torch.uint8, device=device)
        self.arrays = (self._view(plan.pencil_input, plan.eltype_input, plan._parent_dims(0)),
                       self._view(plan.pencil_output, plan.eltype_output, plan._parent_dims(1)))

    def _view(self, pencil, eltype, parent_dims):
        dims = pencil.mem_dims()
        t = self.data.view(_TORCH_DTYPES[eltype])[:int(np.prod(parent_dims))]
        t = t.view(tuple(reversed(parent_dims)))
        if tuple(parent_dims) != tuple(dims):
            t = t[(Ellipsis, slice(0, dims[0]))]   # memory axis 0 is torch's last
        return PencilArray(pencil, eltype, t)

    def first(self):
        return self.arrays[0]

    def last(self):
        return self.arrays[-1]

    def data_ptr(self):
        return self.data.data_ptr()

    def nbytes(self):
        return self.data.numel()

    def __repr__(self):
        return "ManyPencilArray(%s -> %s)" % (self.arrays[0], self.arrays[-1])


def _mightalias(a, b):
    pa, pb = a.data_ptr(), b.data_ptr()
    return pa < pb + b.nbytes() and pb < pa + a.nbytes()


class PencilFFTPlan:
    """Plan for an N-dimensional transform on one GPU (plans.jl:204-320).

    PencilFFTPlan(dims_global, transforms, proc_dims=(1,), comm=None,
                  real_type=Float64, extra_dims=(), permute_dims=True, ...)
    PencilFFTPlan(A::PencilArray, transforms, ...)

    `transforms` is one transform (expanded with expand_dims, plans.jl:316-320)
    or one per dimension.  proc_dims must describe a single process (one GPU);
    fftw_timelimit and transpose_method are accepted and ignored, as the
    reference does for GPU arrays (plans.jl:325-332).  fftw_flags chooses
    how kernels are picked, as FFTW's planner flags do (plans.jl:241-250):
    FFTW_ESTIMATE (or FFTW_WISDOM_ONLY) times nothing and takes tuned
    choices already known to the process (see import_wisdom), else the
    first candidate of each pass; FFTW_MEASURE / PATIENT / EXHAUSTIVE, and
    None (not given), tune the plan on its first execute.
    """

    def __init__(self, dims_global, transforms, proc_dims=None, comm=None, real_type=None, *,
                 extra_dims=(), permute_dims=True, timer=None, fftw_flags=None,
                 fftw_timelimit=None, transpose_method=None, flags=0, device="cuda"):
        if isinstance(dims_global, PencilArray):
            A = dims_global
            if real_type is None:
                real_type = L.F32 if A.eltype in (L.F32, L.C32) else L.F64
            extra_dims = A.pencil.extra_dims
            if A.pencil.permutation != tuple(range(1, len(A.pencil.size_global) + 1)):
                raise L.ArgumentError(L.ERR_ARGUMENT,
                                      "dimensions of input array must be unpermuted")
            dims_global = A.pencil.size_global
            input_eltype = A.eltype
        else:
            input_eltype = None
        dims_global = tuple(int(d) for d in dims_global)
        N = len(dims_global)
        if proc_dims is not None and int(np.prod(proc_dims)) != 1:
            raise L.UnsupportedError(L.ERR_UNSUPPORTED,
                                     "this build runs on one GPU: proc_dims must multiply to 1,"
                                     " got %s" % (tuple(proc_dims),))
        if isinstance(transforms, Tr.AbstractTransform):
            transforms = Tr.expand_dims(transforms, N)
        transforms = tuple(transforms)
        if len(transforms) != N:
            raise L.ArgumentError(L.ERR_ARGUMENT,
                                  "number of transforms (%d) must be equal to number of "
                                  "transformable dimensions in array (`ndims(pencil(A))`)"
                                  % len(transforms))
        self.transforms = transforms
        self.real_type = _real_type_code(real_type)
        self.permute_dims = bool(permute_dims)
        self.extra_dims = tuple(int(e) for e in extra_dims)
        self.timer = timer
        self.device = device
        self.fftw_flags = fftw_flags
        flags = int(flags)
        if fftw_flags is not None and int(fftw_flags) & (L.FFTW_ESTIMATE | L.FFTW_WISDOM_ONLY):
            flags |= L.FLAG_ESTIMATE
        h = C.c_void_p()
        dims_arr = (C.c_int64 * N)(*dims_global)
        kinds = (C.c_int32 * N)(*[t.code for t in transforms])
        ne = len(self.extra_dims)
        ext = (C.c_int64 * max(ne, 1))(*self.extra_dims)
        L.check(L.lib.pfft_plan_create(C.byref(h), N, dims_arr, kinds, self.real_type,
                                       int(self.permute_dims), ne, ext, int(flags)))
        self._h = h
        self._lib = L.lib
        self.N = N
        # queries
        ei, eo = C.c_int32(), C.c_int32()
        L.check(L.lib.pfft_plan_eltypes(h, C.byref(ei), C.byref(eo)))
        self.eltype_input, self.eltype_output = ei.value, eo.value
        if input_eltype is not None and input_eltype != self.eltype_input:
            raise L.ArgumentError(L.ERR_ARGUMENT, "wrong input datatype %s, expected %s\n%s" % (
                _NAMES[input_eltype], _NAMES[self.eltype_input], self.describe()))
        sc = C.c_double()
        L.check(L.lib.pfft_plan_scale_factor(h, C.byref(sc)))
        self.scale_factor = sc.value
        nd = N + ne
        sin_, sout = (C.c_int64 * nd)(), (C.c_int64 * nd)()
        L.check(L.lib.pfft_plan_size_global(h, 0, sin_))
        L.check(L.lib.pfft_plan_size_global(h, 1, sout))
        pin, pout = (C.c_int32 * N)(), (C.c_int32 * N)()
        L.check(L.lib.pfft_plan_permutation(h, 0, pin))
        L.check(L.lib.pfft_plan_permutation(h, 1, pout))
        self.size_global_in = tuple(sin_[:N])
        self.size_global_out = tuple(sout[:N])
        self.pencil_input = Pencil(self.size_global_in, tuple(pin), self.extra_dims)
        self.pencil_output = Pencil(self.size_global_out, tuple(pout), self.extra_dims)
        bi, bo, bw = C.c_size_t(), C.c_size_t(), C.c_size_t()
        L.check(L.lib.pfft_plan_nbytes(h, C.byref(bi), C.byref(bo), C.byref(bw)))
        self.in_bytes, self.out_bytes, self.work_bytes = bi.value, bo.value, bw.value
        ip = C.c_int32()
        L.check(L.lib.pfft_plan_is_inplace(h, C.byref(ip)))
        self.inplace = bool(ip.value)
        self._work = None

    # ---- reference accessors (plans.jl:555-591) ----
    def size(self):
        """Global dims of the input (plans.jl:588)."""
        return self.size_global_in

    def is_inplace(self):
        return self.inplace

    def _parent_dims(self, which):
        nd = self.N + len(self.extra_dims)
        d = (C.c_int64 * nd)()
        L.check(L.lib.pfft_plan_parent_dims(self._h, which, d))
        return tuple(d)

    def describe(self):
        buf = C.create_string_buffer(4096)
        L.check(L.lib.pfft_plan_describe(self._h, buf, 4096))
        return buf.value.decode()

    __repr__ = describe

    def stage(self, n):
        """Pencil chain of 1-based stage n: (perm, dims_before, dims_after, mem_axis)."""
        N = self.N
        pi, po = (C.c_int32 * N)(), (C.c_int32 * N)()
        di, do = (C.c_int64 * N)(), (C.c_int64 * N)()
        ax = C.c_int32()
        L.check(L.lib.pfft_plan_stage(self._h, n - 1, pi, di, po, do, C.byref(ax)))
        return tuple(pi), tuple(di), tuple(do), ax.value

    def kinds(self):
        N = self.N
        f, b = (C.c_int32 * N)(), (C.c_int32 * N)()
        L.check(L.lib.pfft_plan_kinds(self._h, f, b))
        return tuple(f), tuple(b)

    def kernels(self, direction=L.FORWARD):
        """Kernel of each pass of a direction, in launch order."""
        buf = C.create_string_buffer(4096)
        L.check(L.lib.pfft_plan_kernels(self._h, direction, buf, 4096))
        return buf.value.decode().split(";")

    def prepare(self):
        """Upload tables and pick kernels now (pfft_plan_prepare) rather than at
        the first execute -- needed before capturing executes in a graph."""
        L.check(L.lib.pfft_plan_prepare(self._h, None))
        self._workspace()

    def set_timing(self, on=True):
        L.check(L.lib.pfft_plan_set_timing(self._h, int(on)))

    def stage_times(self):
        n = self.N
        ms = (C.c_float * n)()
        L.check(L.lib.pfft_plan_stage_times(self._h, n, ms))
        return list(ms)

    # ---- execution (operations.jl:24-62) ----
    def _workspace(self):
        if self.work_bytes and (self._work is None or self._work.numel() < self.work_bytes):
            self._work = torch.empty(self.work_bytes, dtype=torch.uint8, device=self.device)
        return self._work

    def _check_arrays(self, inp, out, inp_pencil, out_pencil, inp_type, out_type):
        # operations.jl:87-127
        for a in (inp, out):
            if not isinstance(a, PencilArray):
                raise L.ArgumentError(L.ERR_ARGUMENT, "array types (%s, %s) incompatible with "
                                      "out-of-place plans" % (type(inp).__name__,
                                                              type(out).__name__))
        if _mightalias(inp, out):
            raise L.ArgumentError(L.ERR_ARGUMENT, "out-of-place plan applied to aliased data")
        if inp.pencil != inp_pencil or inp.eltype != inp_type:
            raise L.ArgumentError(L.ERR_ARGUMENT, "unexpected dimensions of input data")
        if out.pencil != out_pencil or out.eltype != out_type:
            raise L.ArgumentError(L.ERR_ARGUMENT, "unexpected dimensions of output data")
        for a in (inp, out):
            if not a.parent.is_cuda or not a.parent.is_contiguous():
                raise L.ArgumentError(L.ERR_ARGUMENT,
                                      "arrays must be contiguous GPU tensors")

    def _execute(self, direction, src, dst, stream=None, op=None):
        """Run one direction (include/pfft.h PFFT_FORWARD ...); `op` is an
        optional _lib.SpectralOp fused into the passes (pfft_execute_op)."""
        if self.inplace:
            return self._execute_inplace(direction, src, dst, stream, op)
        # mul! and the transpose of bmul! read the input pencil; bmul!, ldiv!
        # and the transpose of mul! read the output pencil
        if direction in (L.FORWARD, L.ADJOINT_BACKWARD):
            self._check_arrays(src, dst, self.pencil_input, self.pencil_output,
                               self.eltype_input, self.eltype_output)
        else:
            self._check_arrays(src, dst, self.pencil_output, self.pencil_input,
                               self.eltype_output, self.eltype_input)
        s = stream if stream is not None else torch.cuda.current_stream(src.parent.device)
        self._launch(direction, src.data_ptr(), dst.data_ptr(), s, op)
        return dst

    def _launch(self, direction, pin, pout, s, op):
        w = self._workspace()
        wp = w.data_ptr() if w is not None else None
        if op is None:
            L.check(L.lib.pfft_execute(self._h, direction, pin, pout, wp, self.work_bytes,
                                       s.cuda_stream))
        else:
            L.check(L.lib.pfft_execute_op(self._h, direction, C.byref(op), pin, pout, wp,
                                          self.work_bytes, s.cuda_stream))

    def _execute_inplace(self, direction, src, dst, stream=None, op=None):
        # operations.jl:100-116: the same ManyPencilArray in and out
        if not isinstance(src, ManyPencilArray) or not isinstance(dst, ManyPencilArray):
            raise L.ArgumentError(L.ERR_ARGUMENT, "array types (%s, %s) incompatible with "
                                  "in-place plans" % (type(src).__name__, type(dst).__name__))
        if src is not dst:
            raise L.ArgumentError(L.ERR_ARGUMENT,
                                  "input and output arrays for in-place plan must be the same")
        if src.first().pencil != self.pencil_input or src.first().eltype != self.eltype_input:
            raise L.ArgumentError(L.ERR_ARGUMENT, "unexpected dimensions of input data")
        if src.last().pencil != self.pencil_output or src.last().eltype != self.eltype_output:
            raise L.ArgumentError(L.ERR_ARGUMENT, "unexpected dimensions of output data")
        if src.nbytes() < self.in_bytes or not src.data.is_cuda:
            raise L.ArgumentError(L.ERR_ARGUMENT, "buffer too small for the in-place plan")
        s = stream if stream is not None else torch.cuda.current_stream(src.data.device)
        self._launch(direction, src.data_ptr(), src.data_ptr(), s, op)
        return dst

    def _maybe_allocate(self, allocator, src):
        # operations.jl:73-85
        if self.inplace and isinstance(src, ManyPencilArray):
            return src
        if not self.inplace and isinstance(src, PencilArray):
            return allocator(self)
        raise L.ArgumentError(L.ERR_ARGUMENT, "input array type %s incompatible with %s plans"
                              % (type(src).__name__,
                                 "in-place" if self.inplace else "out-of-place"))

    def __mul__(self, src):
        """plan * u (operations.jl:54-57); in place for in-place plans."""
        if isinstance(src, (tuple, list)):
            return type(src)(self * u for u in src)
        return mul_(self._maybe_allocate(allocate_output, src), self, src)

    def ldiv(self, src):
        """plan \\ v (operations.jl:59-62); in place for in-place plans."""
        if isinstance(src, (tuple, list)):
            return type(src)(self.ldiv(v) for v in src)
        return ldiv_(self._maybe_allocate(allocate_input, src), self, src)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.pfft_plan_destroy(h)
            except Exception:
                pass
            self._h = None


def _collection(x):
    return isinstance(x, (tuple, list))


def _check_compatible(a, b):
    # operations.jl:130-137
    if len(a) != len(b):
        raise L.ArgumentError(L.ERR_ARGUMENT, "collections have different lengths: %d ≠ %d"
                              % (len(a), len(b)))


def mul_(dst, p, src):
    """mul!(dst, p, src): forward transform (operations.jl:24-31)."""
    if _collection(dst) or _collection(src):
        _check_compatible(dst, src)
        for d, s in zip(dst, src):
            mul_(d, p, s)
        return dst
    return p._execute(L.FORWARD, src, dst)


def bmul_(dst, p, src):
    """PencilFFTs.bmul!(dst, p, src): unnormalised backward (operations.jl:34-41)."""
    if _collection(dst) or _collection(src):
        _check_compatible(dst, src)
        for d, s in zip(dst, src):
            bmul_(d, p, s)
        return dst
    return p._execute(L.BACKWARD, src, dst)


def ldiv_(dst, p, src):
    """ldiv!(dst, p, src): normalised backward (operations.jl:44-52)."""
    if _collection(dst) or _collection(src):
        _check_compatible(dst, src)
        for d, s in zip(dst, src):
            ldiv_(d, p, s)
        return dst
    return p._execute(L.BACKWARD_SCALED, src, dst)


def allocate_input(p, *dims):
    """allocate_input(p[, dims...]) / allocate_input(p, Val(N)) with an int
    count in a 1-tuple (allocate.jl:44-54,85-86)."""
    if dims:
        return _allocate_many(allocate_input, p, dims)
    if p.inplace:
        return ManyPencilArray(p, p.device)
    return PencilArray.empty(p.pencil_input, p.eltype_input, p.device)


def allocate_output(p, *dims):
    """allocate_output(p[, dims...]) (allocate.jl:100-110,115-122)."""
    if dims:
        return _allocate_many(allocate_output, p, dims)
    if p.inplace:
        return ManyPencilArray(p, p.device)
    return PencilArray.empty(p.pencil_output, p.eltype_output, p.device)


def _allocate_many(f, p, dims):
    if len(dims) == 1 and isinstance(dims[0], Val):
        return tuple(f(p) for _ in range(dims[0].n))
    n = int(np.prod(dims))
    arr = np.empty(n, dtype=object)
    for i in range(n):
        arr[i] = f(p)
    return arr.reshape(dims, order="F")


class Val:
    """Val(N): ask allocate_input/allocate_output for an N-tuple."""

    def __init__(self, n):
        self.n = int(n)


def scale_factor(p):
    return p.scale_factor


def export_wisdom(path=None):
    """The tuned kernel choices of this process as text (pfft_tune_export),
    written to `path` when given: the counterpart of FFTW.export_wisdom."""
    need = C.c_size_t()
    L.check(L.lib.pfft_tune_export(None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    L.check(L.lib.pfft_tune_export(buf, need.value, C.byref(need)))
    text = buf.value.decode()
    if path is not None:
        with open(path, "w") as f:
            f.write(text)
    return text


def import_wisdom(path_or_text):
    """Merge tuned choices exported by export_wisdom (a file path or the text
    itself) into this process (pfft_tune_import): FFTW.import_wisdom."""
    text = path_or_text
    if not text.startswith("pfft-tune"):
        with open(path_or_text) as f:
            text = f.read()
    L.check(L.lib.pfft_tune_import(text.encode()))


def forget_wisdom():
    """Drop every tuned choice of this process: FFTW.forget_wisdom."""
    L.check(L.lib.pfft_tune_forget())


def execute_batch(items, stream=None):
    """Transforms of several out-of-place plans, of different shapes, in one
    C call (pfft_execute_batch): `items` holds (plan, direction, src, dst) or
    (plan, direction, src, dst, op) tuples, op a _lib.SpectralOp (see
    spectral.py).  Items of different plans may run at the same time on the
    library's streams, forked from and joined back into `stream` (the
    current stream by default); items of one plan run in order."""
    items = list(items)
    if not items:
        return
    arr = (L.BatchItem * len(items))()
    keep = []
    for k, it in enumerate(items):
        p, direction, src, dst = it[:4]
        op = it[4] if len(it) > 4 else None
        if p.inplace:
            raise L.UnsupportedError(L.ERR_UNSUPPORTED, "execute_batch takes out-of-place plans")
        if direction in (L.FORWARD, L.ADJOINT_BACKWARD):
            p._check_arrays(src, dst, p.pencil_input, p.pencil_output, p.eltype_input,
                            p.eltype_output)
        else:
            p._check_arrays(src, dst, p.pencil_output, p.pencil_input, p.eltype_output,
                            p.eltype_input)
        w = p._workspace()
        arr[k].plan = p._h
        arr[k].direction = direction
        if op is not None:
            keep.append(op)
            arr[k].op = C.pointer(op)
        arr[k].inp = src.data_ptr()
        arr[k].out = dst.data_ptr()
        arr[k].work = w.data_ptr() if w is not None else None
        arr[k].work_bytes = p.work_bytes
    s = stream if stream is not None else torch.cuda.current_stream(items[0][2].parent.device)
    L.check(L.lib.pfft_execute_batch(len(items), arr, s.cuda_stream))


def execute_many(direction, p, dsts, srcs):
    """All arrays of a collection in one C call on one stream
    (pfft_execute_many): the batched form of the collection broadcast."""
    _check_compatible(dsts, srcs)
    n = len(srcs)
    for s, d in zip(srcs, dsts):
        if direction == L.FORWARD:
            p._check_arrays(s, d, p.pencil_input, p.pencil_output, p.eltype_input,
                            p.eltype_output)
        else:
            p._check_arrays(s, d, p.pencil_output, p.pencil_input, p.eltype_output,
                            p.eltype_input)
    ins = (C.c_void_p * n)(*[s.data_ptr() for s in srcs])
    outs = (C.c_void_p * n)(*[d.data_ptr() for d in dsts])
    w = p._workspace()
    st = torch.cuda.current_stream(srcs[0].parent.device)
    L.check(L.lib.pfft_execute_many(p._h, direction, n, ins, outs,
                                    w.data_ptr() if w is not None else None, p.work_bytes,
                                    st.cuda_stream))
    return dsts
