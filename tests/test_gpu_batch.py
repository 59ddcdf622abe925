"""pfft_execute_batch: transforms of plans of different shapes in one call,
the plans on forked streams.  Each item's result equals that of its own
pfft_execute / pfft_execute_op call (bit-identical: the same kernels), items
of one plan keep their order, and a captured batch replays the eager
result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _fields(pf):
    T = pf.Transforms
    specs = [((64, 48, 32), T.RFFT(), T.Float32), ((96, 64, 40), T.FFT(), T.Float32),
             ((32, 32, 32), T.RFFT(), T.Float64), ((128, 64, 16), T.RFFT(), T.Float32)]
    out = []
    g = torch.Generator(device="cuda").manual_seed(3)
    for dims, tr, rt in specs:
        p = pf.PencilFFTPlan(dims, tr, (1, 1), None, rt)
        u = pf.allocate_input(p)
        if u.parent.is_complex():
            u.parent.copy_(torch.randn(u.parent.shape, generator=g, device="cuda",
                                       dtype=torch.float32 if rt == T.Float32 else torch.float64)
                           .to(u.parent.dtype))
        else:
            u.parent.uniform_(-1, 1, generator=g)
        out.append((p, u))
    return out


def _bytes(a):
    return a.parent.detach().cpu().contiguous().numpy().tobytes()


def test_batch_matches_single_calls(pf, gpu):
    L = pf._lib
    S = pf.spectral
    fields = _fields(pf)
    # single calls: forward, then the derivative along dim 1 back (ldiv!)
    ref = []
    for p, u in fields:
        v = pf.allocate_output(p)
        w = pf.allocate_input(p)
        pf.mul_(v, p, u)
        S.ldiv_(w, p, v, S.DERIVATIVE, dim=1)
        ref.append((_bytes(v), _bytes(w)))
    torch.cuda.synchronize()
    # one batch: per plan, the forward and then the dependent ldiv (in order)
    outs = []
    items = []
    for p, u in fields:
        v = pf.allocate_output(p)
        w = pf.allocate_input(p)
        outs.append((v, w))
        items.append((p, L.FORWARD, u, v))
    for (p, u), (v, w) in zip(fields, outs):
        items.append((p, L.BACKWARD_SCALED, v, w, S.spectral_op(S.DERIVATIVE, 1)))
    pf.execute_batch(items)
    torch.cuda.synchronize()
    for (v, w), (rv, rw) in zip(outs, ref):
        assert _bytes(v) == rv
        assert _bytes(w) == rw


def test_batch_graph_replays(pf, gpu):
    """A batch captured in a hipGraph (forked streams as parallel branches),
    replayed three times with eager work and synchronisations between the
    replays: every replay gives the eager result."""
    L = pf._lib
    fields = _fields(pf)
    outs = [pf.allocate_output(p) for p, _ in fields]
    backs = [pf.allocate_input(p) for p, _ in fields]
    items = [(p, L.FORWARD, u, v) for (p, u), v in zip(fields, outs)]
    items += [(p, L.BACKWARD_SCALED, v, b) for (p, _), v, b in zip(fields, outs, backs)]
    pf.execute_batch(items)          # eager (creates the library's streams)
    torch.cuda.synchronize()
    want = [(_bytes(v), _bytes(b)) for v, b in zip(outs, backs)]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        pf.execute_batch(items)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        pf.execute_batch(items)
    torch.cuda.synchronize()
    for r in range(3):
        for v, b in zip(outs, backs):
            v.parent.zero_()
            b.parent.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [(_bytes(v), _bytes(b)) for v, b in zip(outs, backs)]
        assert got == want, r
        # the round trip itself
        for (p, u), b in zip(fields, backs):
            x = u.parent.double() if not u.parent.is_complex() else u.parent.cdouble()
            y = b.parent.double() if not b.parent.is_complex() else b.parent.cdouble()
            err = float(torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x))
            assert err < 1e-5, err
