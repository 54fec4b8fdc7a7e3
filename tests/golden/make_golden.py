"""Write the golden fixtures in tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

The reference ships no tests, fixtures or golden vectors for its gravity path
and cannot be run here (GF needs nvcc/CUDA; TFM's gravity lives inside FM::FM
next to hours of SET table generation), so these vectors are the oracle's,
whose own correctness is pinned by tests/test_oracle.py. They freeze three
cases so that the GPU tests and any later oracle change are checked against
fixed numbers, and they record checksums of the seeded inputs so a change of
the workload generators is caught too.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ztm_amd import workloads as W  # noqa: E402


def checksum(a):
    a = np.asarray(a, dtype=np.float64).ravel()
    return {"n": int(a.size), "sum": float(a.sum()), "sumsq": float((a * a).sum()),
            "first": float(a[0]), "last": float(a[-1])}


def main():
    man = {}
    # 1. pm_cpu_small forward (32x32x16 cells, 1,024 stations), SI units, FP64
    w = W.pm_cpu_small()
    g = oracle.pm_forward(w.grid, w.kappa, w.rho, w.stations)
    np.save(os.path.join(HERE, "pm_cpu_small_g.npy"), g)
    man["pm_cpu_small"] = {"rho": checksum(w.rho), "stations": checksum(w.stations),
                           "file": "pm_cpu_small_g.npy"}
    # 2. the reference's own call shape: GF on TFM's grid, index units (GFM_GPU.cu:16-22)
    rho, s = W.tfm_shaped()
    gg = oracle.gf(10, 61, 61, 33, 1, 2, rho, s)
    np.save(os.path.join(HERE, "tfm_shaped_gf.npy"), gg)
    man["tfm_shaped"] = {"rho": checksum(rho), "s": checksum(s), "file": "tfm_shaped_gf.npy",
                         "geometry": {"bfr": 10, "di": 61, "dj": 61, "dk": 33, "fh": 1, "fv": 2}}
    # 3. adjoint_small: G^T r
    w = W.adjoint_small()
    a = oracle.pm_adjoint(w.grid, w.kappa, w.stations, w.extra["residual"])
    np.save(os.path.join(HERE, "adjoint_small_gtr.npy"), a)
    man["adjoint_small"] = {"r": checksum(w.extra["residual"]), "file": "adjoint_small_gtr.npy"}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1)
    print("wrote", ", ".join(v["file"] for v in man.values()))


if __name__ == "__main__":
    main()
