"""Write the golden fixtures of tests/golden/ from the FP64 oracle.

The reference cannot run here (no Julia, no FFTW), so the fixtures are the
oracle's outputs on seeded inputs at the reference's own test shapes, stored
in the memory layout of the reference's pencils (column-major images):

  g1_transform_list.npz  test/transforms.jl:312-337 list at DATA_DIMS (16,12,6)
  g2_brfft_single_mode.npz  test/brfft.jl:17-45, odd and even real length
  g3_rfft_impulse.npz    test/rfft.jl:215-231 at (41,23,15) and (42,24,16)
  g4_taylor_green.npz    RFFT of the test/taylor_green.jl:29-44 field (16,8,16)
  g5_dimensionality.npz  test/transforms.jl:414-462 shapes, RFFT and FFT
  g6_metadata.json       sizes, permutations, binv and scale of every case

Run:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pipeline as OP  # noqa: E402
from oracle import plan as OM  # noqa: E402

DATA_DIMS = (16, 12, 6)
R2R = ["R2R_" + k for k in OM.R2R_KINDS]
LIST = ([("FFT",), ("RFFT",), ("BFFT",), ("NoTransform",)] + [(k,) for k in R2R] +
        [("NoTransform", "RFFT", "FFT"), ("FFT", "NoTransform", "FFT"),
         ("FFT", "NoTransform", "NoTransform"), ("BRFFT_even",)])


def case_name(trs, dims, permute=True, extra=()):
    s = "-".join(trs) + "_" + "x".join(map(str, dims))
    if not permute:
        s += "_noperm"
    if extra:
        s += "_e" + "x".join(map(str, extra))
    return s


def meta_json(m):
    return {"size_in": m.size_in, "size_out": m.size_out, "transforms": m.transforms,
            "eltype_in": m.eltype_in, "eltype_out": m.eltype_out, "scale": m.scale,
            "perm_in": m.perm_in, "perm_out": m.perm_out, "extra": m.extra_dims,
            "permute": m.permute, "mem_in": m.mem_dims("in"), "mem_out": m.mem_dims("out"),
            "stages": [{"perm": s["perm"], "before": s["before"], "after": s["after"],
                        "tr": s["tr"], "tr_bwd": s["tr_bwd"], "mem_axis": s["mem_axis"]}
                       for s in m.stages]}


def random_input(rng, m, shape):
    x = rng.standard_normal(shape)
    if m.eltype_in.startswith("Complex"):
        x = x + 1j * rng.standard_normal(shape)
    return x


def add_case(store, meta, name, m, x_logical):
    xin = OP.to_memory(x_logical, m.perm_in)
    y = OP.forward(m, xin, method="direct")
    store[name + "/in"] = OP.memory_image(xin)
    store[name + "/out"] = OP.memory_image(y)
    meta[name] = meta_json(m)


def main():
    meta = {}
    # G1
    g1 = {}
    for i, trs in enumerate(LIST):
        rng = np.random.default_rng(42 + i)
        m = OM.PlanMeta(DATA_DIMS, trs, "Float64", True)
        if trs[-1].startswith("BRFFT"):
            # input consistent with a c2r (transforms.jl:266-272): the inverse
            # transform of random (real) output data
            y = rng.standard_normal(m.size_out)
            x = OP.from_memory(OP.backward(m, OP.to_memory(y, m.perm_out)), m.perm_in)
        else:
            x = random_input(rng, m, DATA_DIMS)
        add_case(g1, meta, case_name(m.transforms, DATA_DIMS), m, x)
    np.savez_compressed(os.path.join(HERE, "g1_transform_list.npz"), **g1)
    # G2
    g2 = {}
    for odd in (0, 1):
        m = OM.PlanMeta((12, 13, 9), ("BRFFT_odd" if odd else "BRFFT_even",))
        c = np.zeros((12, 13, 9), complex)
        c[1, 3, 2] = 1 + 2j
        add_case(g2, meta, case_name(m.transforms, (12, 13, 9)), m, c)
    np.savez_compressed(os.path.join(HERE, "g2_brfft_single_mode.npz"), **g2)
    # G3
    g3 = {}
    for dims in ((41, 23, 15), (42, 24, 16)):
        m = OM.PlanMeta(dims, ("RFFT",))
        x = np.zeros(dims)
        x[0, 0, 0], x[1, 0, 0] = 1.0, 2.0
        add_case(g3, meta, case_name(m.transforms, dims), m, x)
    np.savez_compressed(os.path.join(HERE, "g3_rfft_impulse.npz"), **g3)
    # G4
    g4 = {}
    dims = (16, 8, 16)
    L = (4 * np.pi, 2 * np.pi, 2 * np.pi)
    xs = [np.arange(n) * (l / n) for n, l in zip(dims, L)]
    x, y, z = np.meshgrid(*xs, indexing="ij")
    u0, k0 = 3.0, 2.0
    comps = [u0 * np.sin(k0 * x) * np.cos(k0 * y) * np.cos(k0 * z),
             -u0 * np.cos(k0 * x) * np.sin(k0 * y) * np.cos(k0 * z)]
    m = OM.PlanMeta(dims, ("RFFT",))
    for i, f in enumerate(comps):
        add_case(g4, meta, case_name(m.transforms, dims) + "_u%d" % (i + 1), m, f)
    np.savez_compressed(os.path.join(HERE, "g4_taylor_green.npz"), **g4)
    # G5
    g5 = {}
    for dims, permute in [((11, 15), True), ((11, 7, 13), True), ((9, 7, 5, 12), True),
                          ((9, 7, 5, 12), False)]:
        for tr in ("RFFT", "FFT"):
            rng = np.random.default_rng(7)
            m = OM.PlanMeta(dims, (tr,), "Float64", permute)
            add_case(g5, meta, case_name(m.transforms, dims, permute), m,
                     random_input(rng, m, dims))
    np.savez_compressed(os.path.join(HERE, "g5_dimensionality.npz"), **g5)
    # G6: metadata of the workloads as well
    for dims, trs in [((64, 64, 64), ("RFFT",)), ((512, 512, 512), ("RFFT",)),
                      ((512, 512, 512), ("FFT",)), ((256, 256, 256), ("RFFT",)),
                      ((384, 384, 256), ("RFFT", "FFT", "R2R_REDFT10")),
                      ((1024, 1024, 1024), ("FFT",))]:
        m = OM.PlanMeta(dims, trs)
        meta["workload_" + case_name(m.transforms, dims)] = meta_json(m)
    with open(os.path.join(HERE, "g6_metadata.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", len(meta), "cases")


if __name__ == "__main__":
    main()
