"""Many fields of uneven sizes through the pipeline at once, differentiable.

The reference transforms collections of arrays one at a time
(operations.jl:139-152: `mul!.(dst, p, src)`), one plan per field shape.
Here every field shape becomes one *group*:
* all vector components of all fields of that shape are one array, whose
  trailing extra dims (PencilFFTPlan's `extra_dims`, plans.jl:164-182) batch
  them, so each pass of the group is one kernel launch;
* the groups' launches are captured in a single hipGraph, so one graph launch
  runs the whole batch.

`GradientEnergy` evaluates, for periodic vector fields u on [0, 2pi)^3,

    L(u) = sum_fields sum_c sum_j sum_x (d_j u_c)^2,
    d_j u = BRFFT(i k_j RFFT(u)) / N        (spectral derivative),

with wavenumbers laid out as in test/include/Grids.jl:44-59 (rfftfreq along
the r2c dimension, fftfreq elsewhere), and its gradient by the hand-written
adjoint chain:

    g_bar   = 2 g
    gh_bar  = (ldiv!)^T g_bar            PFFT_ADJOINT_BACKWARD, / N
    uh_bar  = sum_j conj(i k_j) gh_bar_j
    u_bar   = (mul!)^T uh_bar            PFFT_ADJOINT_FORWARD

The constant factors of the chain (i, 2, 1/N, -i) are folded into two sets
of complex multipliers, c_j = i k_j for the derivative and
cb_j = -2 i k_j / N for the gradient, so each elementwise step is one
torch kernel.  No reduction runs inside the captured graph: the loss sum is
taken after a replay.
"""
import math

import torch

from . import _lib as L
from . import transforms as Tr
from .plans import PencilArray, PencilFFTPlan

_REAL = {L.F32: torch.float32, L.F64: torch.float64}
_CPLX = {L.F32: torch.complex64, L.F64: torch.complex128}


class _Group:
    def __init__(self, edge, nfields, real_type, device):
        self.edge = edge
        self.n = nfields
        dims = (edge, edge, edge)
        # input: u[field, component, z, y, x] (memory order x, y, z, c, f)
        self.plan_u = PencilFFTPlan(dims, Tr.RFFT(), (1,), None, real_type,
                                    extra_dims=(3, nfields), device=device)
        # derivatives: g[field, component, direction, ...] (memory x, y, z, j, c, f)
        self.plan_g = PencilFFTPlan(dims, Tr.RFFT(), (1,), None, real_type,
                                    extra_dims=(3, 3, nfields), device=device)
        for p in (self.plan_u, self.plan_g):
            L.check(L.lib.pfft_plan_prepare(p._h, None))
            p._workspace()
        # wavenumbers on the output layout: parent dims (f, [c,] [j,] x', y, z)
        X = edge // 2 + 1
        kx = torch.arange(X, dtype=_REAL[real_type], device=device)
        kf = torch.fft.fftfreq(edge, 1.0 / edge, dtype=_REAL[real_type], device=device)
        self.k = [kx.view(X, 1, 1), kf.view(1, edge, 1), kf.view(1, 1, edge)]
        ct = _CPLX[real_type]
        N = self.plan_g.scale_factor
        self.c = [(1j * k).to(ct) for k in self.k]                   # d/dx_j
        self.cb = [(-2j / N * k).to(ct) for k in self.k]            # gradient chain
        # static buffers (the graph's inputs and outputs)
        self.u = PencilArray.empty(self.plan_u.pencil_input, self.plan_u.eltype_input, device)
        self.uh = PencilArray.empty(self.plan_u.pencil_output, self.plan_u.eltype_output, device)
        self.gh = PencilArray.empty(self.plan_g.pencil_output, self.plan_g.eltype_output, device)
        self.g = PencilArray.empty(self.plan_g.pencil_input, self.plan_g.eltype_input, device)
        self.ghb = PencilArray.empty(self.plan_g.pencil_output, self.plan_g.eltype_output,
                                     device)
        self.uhb = PencilArray.empty(self.plan_u.pencil_output, self.plan_u.eltype_output,
                                     device)
        self.ub = PencilArray.empty(self.plan_u.pencil_input, self.plan_u.eltype_input, device)

    def run(self):
        """Enqueue loss terms and gradient on the current stream."""
        pu, pg = self.plan_u, self.plan_g
        pu._execute(L.FORWARD, self.u, self.uh)                         # u^ = RFFT(u)
        uh = self.uh.parent                                             # (f, c, x', y, z)
        gh = self.gh.parent                                             # (f, c, j, x', y, z)
        for j in range(3):
            torch.mul(uh, self.c[j], out=gh[:, :, j])                   # i k_j u^
        pg._execute(L.BACKWARD_SCALED, self.gh, self.g)                 # g = ldiv(i k u^)
        # dL/dg = 2 g; (ldiv!)^T = (bmul!)^T / N; the 2 / N ride in cb_j
        pg._execute(L.ADJOINT_BACKWARD, self.g, self.ghb)
        ghb = self.ghb.parent
        uhb = self.uhb.parent
        torch.mul(ghb[:, :, 0], self.cb[0], out=uhb)                    # sum_j conj(i k_j) .
        uhb.addcmul_(ghb[:, :, 1], self.cb[1]).addcmul_(ghb[:, :, 2], self.cb[2])
        pu._execute(L.ADJOINT_FORWARD, self.uhb, self.ub)               # (mul!)^T

    def loss(self):
        return torch.sum(self.g.parent.double() ** 2)

    def grad_bytes(self):
        """Algorithmic HBM bytes of the six plan executions of one run()."""
        e = self.edge
        R = e ** 3 * self.g.parent.element_size()
        C = (e // 2 + 1) * e * e * 2 * self.g.parent.element_size()
        per_dir = R + 5 * C           # r2c/c2r pass + two c2c passes
        return per_dir * (3 * self.n) * 2 + per_dir * (9 * self.n) * 2


class GradientEnergy:
    """sum |grad u|^2 of many periodic vector fields of uneven cube sizes, and
    its gradient, through grouped plans and one hipGraph."""

    def __init__(self, edges, real_type=L.F32, device="cuda"):
        self.edges = list(edges)
        self.real_type = real_type
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

    def grad(self, i):
        g, j = self.slot[i]
        return g.ub.parent[j].permute(3, 2, 1, 0)

    def run(self):
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
