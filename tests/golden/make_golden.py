"""Writes tests/golden/golden.json from the reference's own host helpers.

Run in a container where the reference tree is mounted and oracle/_ref is
built (`make -C oracle ref`): it loads oracle/_ref/libref_*.so (the
reference's read_mtx / change2tran / get_x_b / matrix_layer /
matrix_transposition / matrix_warp4 / serial executor, compiled from their
own sources by oracle/Makefile) and records inputs and outputs for every
fixture in tests/golden/mtx/.  Floating-point values are stored as
float.hex() strings so comparisons are bit-exact.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

MTX = os.path.join(HERE, "mtx")
OUT = os.path.join(HERE, "golden.json")


def hexs(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64)]


def ints(a):
    return [int(v) for v in a]


def lower_with_values(rp, ci, seed):
    """A real-valued, diagonally dominant lower-triangular CSR with L's pattern."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, size=len(ci))
    for i in range(len(rp) - 1):
        d = rp[i + 1] - 1
        v[d] = 1.0 + np.abs(v[rp[i]:d]).sum()
    return v


def main():
    if not oracle.ref_available("d64"):
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` where the reference tree is mounted")
    cases = {}
    for name in sorted(os.listdir(MTX)):
        path = os.path.join(MTX, name)
        ret, m, n, nnzA, rpA, ciA, vA = oracle.ref_read_mtx(path, "d64")
        c = {"read_mtx": {"ret": ret}}
        cases[name] = c
        if ret != 0:
            continue
        c["read_mtx"].update(m=m, n=n, nnz=nnzA, row_ptr=ints(rpA), col_idx=ints(ciA), val=hexs(vA))
        mm = min(m, n)  # leading square block (CUDA/ourWrtFst/main.cu:66-71)
        nnzL, rpL, ciL, vL = oracle.ref_change2tran(mm, nnzA, rpA[:mm + 1], ciA, vA, "d64")
        c["change2tran"] = dict(m=mm, nnzL=nnzL, row_ptr=ints(rpL), col_idx=ints(ciL), val=hexs(vL))
        x_ref, b = oracle.ref_get_x_b(mm, mm, rpL, ciL, vL, "d64")
        c["get_x_b"] = dict(x_ref=hexs(x_ref), b=hexs(b))
        layer, par = oracle.ref_matrix_layer(mm, rpL, ciL, vL, "d64")
        c["matrix_layer"] = dict(layer=layer, parallelism=float(par).hex())
        cr, cp, cv = oracle.ref_transpose(mm, mm, nnzL, rpL, ciL, vL, "d64")
        c["matrix_transposition"] = dict(csc_row=ints(cr), csc_colptr=ints(cp), csc_val=hexs(cv))
        w = {}
        for variant, chunk in (("d64", 64), ("d32", 32)):
            for border in (1, 2, 3):
                ln, wn, wo, eo = oracle.ref_matrix_warp4(mm, nnzL, rpL, ciL, vL, border, variant)
                w[f"{chunk}_{border}"] = dict(len=ln, warp_num=ints(wn), warp_occupy=float(wo).hex(),
                                              element_occupy=float(eo).hex())
        c["matrix_warp4"] = w
        # serial executor on a real-valued L with the same pattern
        vR = lower_with_values(rpL, ciL, seed=len(name))
        crR, cpR, cvR = oracle.ref_transpose(mm, mm, nnzL, rpL, ciL, vR, "d64")
        indeg = oracle.ref_analyser(crR, mm, nnzL, "d64")
        ex = dict(val=hexs(vR), indeg=ints(indeg))
        for rhs in (1, 3):
            bb = np.random.default_rng(rhs).uniform(-1, 1, size=mm * rhs)
            ex[f"b_{rhs}"] = hexs(bb)
            ex[f"forward_{rhs}"] = hexs(oracle.ref_csc_executor(cpR, crR, cvR, indeg, mm, mm, 0, rhs, bb, "d64"))
            # backward mode on the CSR of L read as the CSC of L^T (the adjoint)
            indegT = np.bincount(ciL, minlength=mm).astype(np.int32)
            ex[f"backward_T_{rhs}"] = hexs(oracle.ref_csc_executor(rpL, ciL, vR, indegT, mm, mm, 1, rhs, bb, "d64"))
        # single precision executor
        v32 = vR.astype(np.float32)
        crS, cpS, cvS = oracle.ref_transpose(mm, mm, nnzL, rpL, ciL, v32, "s64")
        b32 = np.random.default_rng(7).uniform(-1, 1, size=mm).astype(np.float32)
        ex["b32"] = hexs(b32)
        ex["forward32"] = hexs(oracle.ref_csc_executor(cpS, crS, cvS, indeg, mm, mm, 0, 1, b32, "s64"))
        c["executor"] = ex
    with open(OUT, "w") as f:
        json.dump({"source": "reference host helpers via oracle/_ref", "cases": cases}, f, indent=0,
                  sort_keys=True)
    print(f"wrote {OUT} ({len(cases)} fixtures)")


if __name__ == "__main__":
    main()
