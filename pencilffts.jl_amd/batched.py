"""Many fields of uneven sizes through the pipeline at once, differentiable.

The reference transforms collections of arrays one at a time
(operations.jl:139-152: `mul!.(dst, p, src)`), one plan per field shape.
Here every field shape becomes one *group*:
* all vector components of all fields of that shape are one array, whose
  trailing extra dims (PencilFFTPlan's `extra_dims`, plans.jl:164-182) batch
  them, so each pass of the group is one kernel launch;
* all groups' executions go to the library in one call (pfft_execute_batch),
  which runs the groups on forked streams so that they overlap;
* those launches are captured in a single hipGraph, so one graph launch runs
  the whole batch.

`GradientEnergy` evaluates, for periodic vector fields u on [0, 2pi)^3,

    L(u) = sum_fields sum_c sum_j sum_x (d_j u_c)^2,
    d_j u = BRFFT(i k_j RFFT(u)) / N        (spectral derivative),

with wavenumbers laid out as in test/include/Grids.jl:44-59 (rfftfreq along
the r2c dimension, fftfreq elsewhere), and its gradient by the hand-written
adjoint chain:

    g_j     = ldiv!(i k_j u^)                 ldiv_ with the derivative operator
    u^_bar  = sum_j (2/N) conj(i k_j) B^T g_j  transpose of bmul! with the operator,
                                              accumulated over j
    u_bar   = (mul!)^T u^_bar                 PFFT_ADJOINT_FORWARD

The multipliers i k_j and their conjugates ride on the passes that read or
write the spectrum (pfft_execute_op), so no elementwise kernel runs between
transforms.  No reduction runs inside the captured graph: the loss sum is
taken after a replay.
"""
import math

import torch

from . import _lib as L
from . import spectral as S
from . import transforms as Tr
from .plans import PencilArray, PencilFFTPlan, execute_batch


class _Group:
    def __init__(self, edge, nfields, real_type, device):
        self.edge = edge
        self.n = nfields
        dims = (edge, edge, edge)
        # u[field, component, z, y, x] (memory order x, y, z, c, f)
        self.plan = PencilFFTPlan(dims, Tr.RFFT(), (1,), None, real_type,
                                  extra_dims=(3, nfields), device=device)
        p = self.plan
        L.check(L.lib.pfft_plan_prepare(p._h, None))
        p._workspace()
        new_in = lambda: PencilArray.empty(p.pencil_input, p.eltype_input, device)
        new_out = lambda: PencilArray.empty(p.pencil_output, p.eltype_output, device)
        # static buffers (the graph's inputs and outputs)
        self.u, self.uh, self.uhb, self.ub = new_in(), new_out(), new_out(), new_in()
        self.g = [new_in() for _ in range(3)]     # d_j u, j = x, y, z
        self.alpha = 2.0 / p.scale_factor

    def items(self):
        """The same executions as run(), as pfft_execute_batch items (in
        order: one plan, so they run in this order)."""
        p = self.plan
        it = [(p, L.FORWARD, self.u, self.uh)]
        for j in range(3):
            it.append((p, L.BACKWARD_SCALED, self.uh, self.g[j], S.spectral_op(S.DERIVATIVE, j)))
        for j in range(3):
            it.append((p, L.ADJOINT_BACKWARD, self.g[j], self.uhb,
                       S.spectral_op(S.DERIVATIVE, j, self.alpha, j > 0)))
        it.append((p, L.ADJOINT_FORWARD, self.uhb, self.ub))
        return it

    def run(self):
        """Enqueue loss terms and gradient on the current stream."""
        p = self.plan
        p._execute(L.FORWARD, self.u, self.uh)                          # u^ = RFFT(u)
        for j in range(3):
            S.ldiv_(self.g[j], p, self.uh, S.DERIVATIVE, dim=j)         # g_j = ldiv(i k_j u^)
        for j in range(3):                                              # dL/du^
            S.bmul_adjoint_(self.uhb, p, self.g[j], S.DERIVATIVE, dim=j, alpha=self.alpha,
                            accumulate=j > 0)
        p._execute(L.ADJOINT_FORWARD, self.uhb, self.ub)                # (mul!)^T

    def loss(self):
        return sum(torch.sum(g.parent.double() ** 2) for g in self.g)

    def grad_bytes(self):
        """Algorithmic HBM bytes of the plan executions of one run(): eight
        three-pass executions over 3n components (R + 5C each) and the reads
        of the two accumulating stores."""
        e = self.edge
        rb = self.u.parent.element_size()
        R = e ** 3 * rb
        C = (e // 2 + 1) * e * e * 2 * rb
        return (3 * self.n) * (8 * (R + 5 * C) + 2 * C)


class GradientEnergy:
    """sum |grad u|^2 of many periodic vector fields of uneven cube sizes, and
    its gradient, through grouped plans and one hipGraph."""

    def __init__(self, edges, real_type=L.F32, device="cuda", batch=True):
        self.edges = list(edges)
        self.real_type = real_type
        # batch: every group's executions in one pfft_execute_batch call, the
        # groups on forked streams (they overlap); else group after group
        self.batch = batch
        order = sorted(set(self.edges))
        self.groups = {e: _Group(e, self.edges.count(e), real_type, device) for e in order}
        # field i -> (group, index within group)
        self.slot = []
        seen = {e: 0 for e in order}
        for e in self.edges:
            self.slot.append((self.groups[e], seen[e]))
            seen[e] += 1
        self.graph = None

    def field(self, i):
        """Logical view u[x, y, z, c] of field i (a view into the group buffer)."""
        g, j = self.slot[i]
        return g.u.parent[j].permute(3, 2, 1, 0)

    def derivative(self, i, j):
        """Logical view d_j u[x, y, z, c] of field i after run()."""
        g, k = self.slot[i]
        return g.g[j].parent[k].permute(3, 2, 1, 0)

    def grad(self, i):
        g, j = self.slot[i]
        return g.ub.parent[j].permute(3, 2, 1, 0)

    def run(self):
        if self.batch:
            items = []
            for g in self.groups.values():
                items += g.items()
            execute_batch(items)
        else:
            for g in self.groups.values():
                g.run()

    def loss(self):
        return sum(g.loss() for g in self.groups.values())

    def capture(self, warmup=2):
        """Capture the whole batched evaluation in one hipGraph."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.run()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run()
        torch.cuda.synchronize()
        return self.graph

    def replay(self):
        self.graph.replay()

    def algorithmic_bytes(self):
        return sum(g.grad_bytes() for g in self.groups.values())


def taylor_green_fields(model, seed=0, u0=1.0, k0=1.0, noise=0.01):
    """Fill every field with the Taylor-Green velocity of
    test/taylor_green.jl:29-44 on [0, 2pi)^3 plus N(0, noise^2) noise."""
    gen = torch.Generator(device="cpu").manual_seed(seed)
    for i, e in enumerate(model.edges):
        x = torch.arange(e, dtype=torch.float64) * (2 * math.pi / e)
        X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
        f = torch.stack([u0 * torch.sin(k0 * X) * torch.cos(k0 * Y) * torch.cos(k0 * Z),
                         -u0 * torch.cos(k0 * X) * torch.sin(k0 * Y) * torch.cos(k0 * Z),
                         torch.zeros_like(X)], dim=-1)
        f = f + noise * torch.randn(f.shape, generator=gen, dtype=torch.float64)
        dst = model.field(i)
        dst.copy_(f.to(dst.dtype))


def draw_edges(n=12, seed=42, choices=(64, 96, 128, 192, 256, 320)):
    gen = torch.Generator(device="cpu").manual_seed(seed)
    idx = torch.randint(len(choices), (n,), generator=gen)
    return [choices[int(i)] for i in idx]
