"""Writes the golden vectors under tests/golden/ from the hand-built reference library.

Run in a container that has the reference tree (it builds oracle/_ref/libmoao_ref.so from the
reference's own CPU sources with oracle/Makefile.ref):

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

For each system (the reference's datafile/check nx456 fixtures, copied under
tests/golden/systems/, and the seeded "small" workload) it stores, in <name>.npz:
  * the Tomo_struct vectors matcov reads (X, Y, u, v, sspSizeL, L0diff, cn2, noise*,
    lgsExt, lgsTheta), target 0 applied;
  * 3000 sampled entries (row, col, value) of each of Cmm (lower), Ctm, Ctt, plus every column
    sum of each matrix (so a wrong element anywhere changes a checked number);
  * compareMatrices(A, B) outputs on two fixed pairs, for the metric restatement.
Only data is written; no reference source is copied.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

FIELDS = ["X", "Y", "u", "v", "sspSizeL", "L0diff", "cn2", "noiseNGS", "noiseLGSxx",
          "noiseLGSyy", "noiseLGSxy", "lgsExt", "lgsTheta", "Nsubap", "Nssp"]
NSAMPLE = 3000


def ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libmoao_ref.so")
    lib = ctypes.CDLL(path)
    lib.ref_compare_matrices.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                         ctypes.c_long, ctypes.c_char, ctypes.c_char,
                                         ctypes.c_void_p, ctypes.c_void_p]
    return lib


def field(lib, name):
    n = lib.ref_tomo_field(name.encode(), None, 0)
    a = np.zeros(n)
    lib.ref_tomo_field(name.encode(), a.ctypes.data_as(ctypes.c_void_p), n)
    return a


def tile(lib, r, c, typ):
    a = np.zeros((r, c), order="F")
    assert lib.ref_matcov_tile(a.ctypes.data_as(ctypes.c_void_p), r, c, 0, 0, r, typ) == 0
    return a


def sample(a, rng, lower=False):
    r, c = a.shape
    i = rng.integers(0, r, NSAMPLE)
    j = rng.integers(0, c, NSAMPLE)
    if lower:
        i, j = np.maximum(i, j), np.minimum(i, j)
    return np.stack([i, j]).astype(np.int32), a[i, j]


def one(lib, name, sys_file, atm_file, seed):
    assert lib.ref_tomo_load(sys_file.encode(), atm_file.encode()) == 0
    assert lib.ref_tomo_set_target(0) == 0
    dims = (ctypes.c_long * 8)()
    lib.ref_tomo_dims(dims)
    n, m = dims[5], dims[6]
    out = {"dims": np.array(list(dims))}
    for f in FIELDS:
        out["f_" + f] = field(lib, f)
    rng = np.random.default_rng(seed)
    cmm = tile(lib, n, n, 1)
    ctm = tile(lib, m, n, 3)
    ctt = tile(lib, m, m, 4)
    for key, a, low in (("cmm", cmm, True), ("ctm", ctm, False), ("ctt", ctt, False)):
        idx, val = sample(a, rng, low)
        out[key + "_idx"], out[key + "_val"] = idx, val
        out[key + "_colsum"] = (np.tril(a) if low else a).sum(axis=0)
        out[key + "_fro"] = np.array([np.linalg.norm(np.tril(a) if low else a)])
    # compareMatrices outputs on (Ctt, perturbed Ctt) for uplo N/L/U and trans N/T
    pert = np.asfortranarray(ctt * (1 + 1e-9 * rng.standard_normal(ctt.shape)))
    cmp = []
    for uplo in "NLU":
        for tr in "NT":
            me, fr = ctypes.c_double(), ctypes.c_double()
            lib.ref_compare_matrices(ctt.ctypes.data, pert.ctypes.data, m, m,
                                     uplo.encode(), tr.encode(), ctypes.byref(me),
                                     ctypes.byref(fr))
            cmp.append([me.value, fr.value])
    out["cmp_pert_seed"] = np.array([seed])
    out["cmp"] = np.array(cmp)
    out["cmp_pert"] = pert
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"wrote {name}.npz  n={n} m={m}")


def main():
    from moao_amd import systems
    lib = ref_lib()
    sysdir = os.path.join(HERE, "systems")
    for k, name in enumerate(sorted(os.listdir(sysdir))):
        d = os.path.join(sysdir, name)
        one(lib, name, os.path.join(d, "sys-params.txt"),
            os.path.join(d, "prof-1-atmos-night0.txt"), 100 + k)


if __name__ == "__main__":
    main()
