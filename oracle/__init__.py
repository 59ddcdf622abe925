"""CPU oracle for the PencilFFTs transform pipeline.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this package: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
only as the checker (or, for cpu_baseline, as the timed CPU restatement).

Contents
  plan.py         -- the PLAN-layer metadata of PencilFFTs.jl v0.15.3 restated
                     in Python (transform sizes, binv, scale factors, eltypes,
                     expand_dims, the pencil permutation chain).
  transforms1d.py -- the 1D transform definitions FFTW implements (the
                     arithmetic the reference delegates to FFTW), twice:
                     brute-force O(n^2) sums in long double with exact angle
                     reduction, and fast FP64 numpy/scipy (pocketfft) calls.
  pipeline.py     -- the execution layer (src/operations.jl:24-62,161-257):
                     per stage a transpose into the stage's pencil memory
                     order, then the 1D transform along the fastest axis;
                     backward in reverse with binv transforms, then the
                     ldiv! scaling.

Pinning (each function is checked by tests/test_oracle.py):
  plan.py         doctest known-answer values of the reference
                  (Transforms.jl:100-106,121-137,162-187,222-259,273-283;
                  global_params.jl:31-40) and transforms.jl:24-37.
  transforms1d.py brute force vs fast path on many small sizes; closed forms
                  (impulse, single mode, DC), inverse-pair identities from the
                  FFTW manual (REDFT10 <-> REDFT01 etc. times the logical size),
                  Parseval, and the reference's numerical KATs (rfft.jl:215-231,
                  brfft.jl:22-44, taylor_green.jl:125-150).
  pipeline.py     the permutation chain against the closed-form index map
                  out[z + Z (y + Y x)] = in[x + X (y + Y z)] on integer data,
                  separability (a 3D transform equals numpy's n-D transform),
                  and round trips.
"""
