"""CPU tests of the library's serial solves (host_prep.cpp), no GPU calls.

The column-driven executor must give the bits of the reference's own serial
executor (openCL/SynFree_csc/sptrsv_syncfree_serialref.h:25-108, compiled
from its sources into oracle/_ref) forward and backward, for 1-5 right-hand
sides, in FP64 and FP32; the dense triangular solve pins the result where
_ref is absent.  Inputs on which the reference would spin forever or index
out of range return error codes here (each test runs under a time limit).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg

import oracle
from capellinisptrsv_amd import api
from capellinisptrsv_amd._native import CapError
from conftest import ROOT, random_lower, upper_from_lower

need_ref = pytest.mark.skipif(not oracle.ref_available("d64"), reason="oracle/_ref not built (no reference tree)")
MTX = os.path.join(ROOT, "tests", "golden", "mtx")


def dense_lower(rp, ci, v):
    n = len(rp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            A[i, ci[j]] += v[j]
    return A


def csc_of(rp, ci, v):
    n = len(rp) - 1
    cr, cp, cv = api.matrix_transposition(n, n, len(ci), rp, ci, v)
    return cp, cr, cv


def cases():
    out = []
    for seed, n, off, band in ((0, 1, 0, 1), (1, 7, 3, 4), (2, 200, 5, 40), (3, 1500, 8, 300)):
        out.append((f"rand{n}", *random_lower(n, max_off=off, band=band, seed=seed)))
    out.append(("dup", *random_lower(300, max_off=6, band=30, seed=9, dup=True)))
    rp, ci, v = api.make_config(0, seed=0, scale=0.02)
    out.append(("band2k", rp, ci, v))
    return out


CASES = cases()


@pytest.mark.timeout(120)
@pytest.mark.parametrize("rhs", [1, 2, 3, 5])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_executor_forward_matches_dense_solve(case, rhs):
    _, rp, ci, v = case
    n = len(rp) - 1
    cp, cr, cv = csc_of(rp, ci, v)
    deg = api.serial_analyser(cr, n, len(ci))
    assert np.array_equal(deg, np.diff(rp))
    b = api.uniform(11, n * rhs)
    x = api.serial_executor(cp, cr, cv, deg, n, n, api.FORWARD, rhs, b).reshape(n, rhs)
    ref = scipy.linalg.solve_triangular(dense_lower(rp, ci, v), b.reshape(n, rhs), lower=True)
    assert np.max(np.abs(x - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@need_ref
@pytest.mark.timeout(120)
@pytest.mark.parametrize("variant,dtype", [("d64", np.float64), ("s64", np.float32)])
@pytest.mark.parametrize("rhs", [1, 3, 5])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_executor_bit_exact_vs_reference_build(case, rhs, variant, dtype):
    """Forward on the CSC of L, backward on the CSR of L read as the CSC of
    L^T (diagonal last in each column): the reference's bits both ways."""
    _, rp, ci, v = case
    n = len(rp) - 1
    v = v.astype(dtype)
    b = api.uniform(12, n * rhs).astype(dtype)
    cp, cr, cv = csc_of(rp, ci, v)
    for sub, (P, R, V) in ((api.FORWARD, (cp, cr, cv)), (api.BACKWARD, (rp, ci, v))):
        deg = api.serial_analyser(R, n, len(R))
        assert np.array_equal(deg, oracle.ref_analyser(R, n, len(R), variant=variant))
        x = api.serial_executor(P, R, V, deg, n, n, sub, rhs, b)
        y = oracle.ref_csc_executor(P, R, V, deg, n, n, sub, rhs, b, variant=variant)
        assert x.dtype == dtype
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (sub, rhs)


@pytest.mark.timeout(120)
def test_executor_backward_is_the_transposed_solve():
    """Backward substitution on the CSR of L read as CSC solves L^T y = g."""
    rp, ci, v = random_lower(400, max_off=5, band=60, seed=4)
    n = len(rp) - 1
    g = api.uniform(13, n)
    y = api.serial_executor(rp, ci, v, api.serial_analyser(ci, n, len(ci)), n, n, api.BACKWARD, 1, g)
    ref = scipy.linalg.solve_triangular(dense_lower(rp, ci, v).T, g, lower=False)
    assert np.max(np.abs(y - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.timeout(120)
@pytest.mark.parametrize("name", ["general_real.mtx", "symmetric_real.mtx", "pattern_general.mtx",
                                  "integer_symmetric.mtx"])
def test_executor_on_fixture_pipeline_gives_ones(name):
    """read_mtx -> change2tran -> get_x_b -> CSC -> executor: the reference's
    pipeline, whose exact answer on the unit-valued L is all ones."""
    ret, m, n, nnz, rp, ci, v = api.read_mtx(os.path.join(MTX, name))
    assert ret == 0
    k = min(m, n)
    nnzL, rpL, ciL, vL = api.change2tran(k, int(rp[k]), rp, ci, v)
    x_ref, b = api.get_x_b(k, k, rpL, ciL, vL)
    cp, cr, cv = csc_of(rpL, ciL, vL)
    ret, x = api.sptrsv_syncfree_serialref(cp, cr, cv, k, k, nnzL, api.FORWARD, 1, b, x_ref)
    assert ret == 0
    assert np.array_equal(x, x_ref)


@pytest.mark.timeout(60)
def test_serialref_wrapper_reports_like_the_reference(capfd):
    libc = ctypes.CDLL(None)

    def printed():
        libc.fflush(None)  # the report goes through C stdio, buffered when stdout is not a terminal
        return capfd.readouterr().out

    rp, ci, v = random_lower(100, seed=5)
    n = len(rp) - 1
    cp, cr, cv = csc_of(rp, ci, v)
    x_ref = np.ones(n)
    b = dense_lower(rp, ci, v) @ x_ref
    ret, x = api.sptrsv_syncfree_serialref(cp, cr, cv, n, n, len(ci), api.FORWARD, 1, b, x_ref)
    out = printed()
    assert ret == 0 and "SpTRSV Serial executor passed!" in out and "analyser on L used" in out
    # a wrong x_ref fails the relative L1 check but, as in the reference, returns 0
    ret, _ = api.sptrsv_syncfree_serialref(cp, cr, cv, n, n, len(ci), api.FORWARD, 1, b, 2 * x_ref)
    assert ret == 0 and "_NOT_ passed!" in printed()
    # non-square: -1 and the reference's message
    ret, _ = api.sptrsv_syncfree_serialref(cp, cr, cv, n, n + 1, len(ci), api.FORWARD, 1, b, x_ref)
    assert ret == -1 and "not a square matrix" in printed()


@pytest.mark.timeout(60)
def test_malformed_inputs_return_codes_instead_of_spinning():
    rp, ci, v = random_lower(50, max_off=3, band=10, seed=6)
    n = len(rp) - 1
    cp, cr, cv = csc_of(rp, ci, v)
    deg = api.serial_analyser(cr, n, len(cr))
    b = np.ones(n)

    def run(P, R, V, d=deg, sub=api.FORWARD):
        with pytest.raises(CapError) as e:
            api.serial_executor(P, R, V, d, n, n, sub, 1, b)
        return e.value.code

    # backward on the forward layout: every column's last entry is off-diagonal
    assert run(cp, cr, cv, sub=api.BACKWARD) == -5
    # an entry above the diagonal in a forward column (the reference waits forever)
    c = next(j for j in range(n) if cp[j + 1] - cp[j] >= 2)
    bad = cr.copy()
    bad[cp[c] + 1] = max(c - 1, 0) if c > 0 else 0
    assert run(cp, bad, cv) in (-4, -5)
    # a row index past m (the reference writes outside left_sum)
    bad = cr.copy()
    bad[cp[c] + 1] = n + 7
    assert run(cp, bad, cv) == -2
    # an in-degree table that does not match the matrix
    d2 = deg.copy()
    d2[-1] += 1
    assert run(cp, cr, cv, d=d2) == -2
    # decreasing column pointers
    cp2 = cp.copy()
    cp2[n // 2] = cp2[n // 2 + 1] + 1
    assert run(cp2, cr, cv) == -3
    # the wrapper checks that the column pointers end at its nnz
    ret, _ = api.sptrsv_syncfree_serialref(cp, cr, cv, n, n, len(cr) - 1, api.FORWARD, 1, b, b)
    assert ret == -3
    # non-square
    with pytest.raises(CapError) as e:
        api.serial_executor(cp, cr, cv, deg, n, n - 1, api.FORWARD, 1, b[:-1])
    assert e.value.code == -1
    # the analyser refuses an out-of-range row index too
    with pytest.raises(CapError):
        api.serial_analyser(np.array([0, 3], np.int32), 2, 2)


@pytest.mark.timeout(120)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("rhs", [1, 4])
def test_csr_serial_gives_the_executor_and_fused_bits(dtype, rhs):
    """Sorted rows: the row-ordered form in reference arithmetic equals the
    column-driven executor bit for bit; in fused arithmetic it equals the
    fused oracle (the GPU kernels' bits).  Shuffled rows keep stored order."""
    for seed in range(3):
        rp, ci, v = random_lower(800, max_off=7, band=100, seed=20 + seed, dtype=dtype)
        n = len(rp) - 1
        b = api.uniform(14 + seed, n * rhs).astype(dtype)
        x_ref_arith = api.serial_csr_solve(rp, ci, v, b, rhs=rhs, arith="reference")
        cp, cr, cv = csc_of(rp, ci, v)
        x_exec = api.serial_executor(cp, cr, cv, api.serial_analyser(cr, n, len(cr)), n, n, api.FORWARD, rhs, b)
        assert np.array_equal(x_ref_arith.view(np.uint8), x_exec.view(np.uint8))
        x_fused = api.serial_csr_solve(rp, ci, v, b, rhs=rhs, arith="fused")
        assert np.array_equal(x_fused, oracle.csr_forward(rp, ci, v, b, rhs=rhs, fused=True))
        rs, cs, vs = random_lower(800, max_off=7, band=100, seed=20 + seed, dtype=dtype, shuffle=True)
        for arith, fused in (("reference", False), ("fused", True)):
            assert np.array_equal(api.serial_csr_solve(rs, cs, vs, b, rhs=rhs, arith=arith),
                                  oracle.csr_forward(rs, cs, vs, b, rhs=rhs, fused=fused))


@pytest.mark.timeout(60)
def test_csr_serial_refuses_what_validation_refuses():
    rp, ci, v = random_lower(20, seed=7)
    b = np.ones(20)
    urp, uci, uv = upper_from_lower(rp, ci, v)
    with pytest.raises(CapError):
        api.serial_csr_solve(urp, uci, uv, b)  # upper layout: diagonal first
    v0 = v.copy()
    v0[rp[5 + 1] - 1] = 0.0
    with pytest.raises(CapError) as e:
        api.serial_csr_solve(rp, ci, v0, b)
    assert e.value.code == -6
    with pytest.raises(ValueError):
        api.serial_csr_solve(rp, ci, v, b, arith="reference", rhs=2)  # b holds n values, not 2n


@pytest.mark.timeout(120)
@pytest.mark.parametrize("exe,dtype", [("sptrsv_main", np.float64), ("sptrsv_main_f32", np.float32)])
def test_driver_serial_mode_and_sidecar(tmp_path, exe, dtype):
    """./sptrsv_main A.mtx -a serial (no GPU call): the reference's 10
    result.csv columns, and one result.jsonl object per run whose error is
    measured against the FP64 serial solve of the same L and b."""
    import json
    path = os.path.join(ROOT, "capellinisptrsv_amd", exe)
    names = ["general_real.mtx", "symmetric_real.mtx", "pattern_general.mtx"]
    for name in names:
        r = subprocess.run([path, os.path.join(MTX, name), "-a", "serial", "--values"], cwd=tmp_path,
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (name, r.stdout, r.stderr)
    rows = (tmp_path / "result.csv").read_text().splitlines()
    assert rows[0::2] == ["filename,m,nnzL,nnzL/m,layer,parallelism,pre_time,solve_time,gflops,bandwith"] * 3
    side = [json.loads(line) for line in (tmp_path / "result.jsonl").read_text().splitlines()]
    assert len(side) == 3
    for name, row, s in zip(names, rows[1::2], side):
        ret, m, n, nnzA, rp, ci, v = oracle.read_mtx(os.path.join(MTX, name), dtype=dtype)
        k = min(m, n)
        nnzL, rpL, ciL, vL = oracle.change2tran(k, int(rp[k]), rp, ci, v)
        f = [c.strip() for c in row.split(",")]
        assert int(f[1]) == k and int(f[2]) == nnzL and float(f[7]) >= 0
        assert s["algo"] == "serial" and s["device"] == "cpu" and s["roofline"] is None and s["passed"]
        assert (s["m"], s["nnzL"], s["layer"]) == (k, nnzL, int(f[4]))
        assert s["value_type"] == ("double" if dtype == np.float64 else "float")
        assert s["x_from"] == "timed solve"
        if dtype == np.float64:
            assert s["rel_err_vs_fp64_serial"] == 0.0 and s["bit_identical_to_fp64_serial"]
        else:
            assert s["rel_err_vs_fp64_serial"] <= 1e-5
    # --no-sidecar leaves result.jsonl alone
    r = subprocess.run([path, os.path.join(MTX, names[0]), "-a", "serial", "--no-sidecar"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert len((tmp_path / "result.jsonl").read_text().splitlines()) == 3


@pytest.mark.timeout(120)
def test_sweep_runs_the_serial_mode(tmp_path):
    """The run.sh-style sweep with -a serial runs on a machine without a GPU:
    one result.csv row and one result.jsonl object per valid fixture."""
    import json

    from capellinisptrsv_amd import sweep
    good = [f for f in sorted(os.listdir(MTX)) if f.endswith(".mtx") and not f.startswith("err_")]
    mats = tmp_path / "m"
    mats.mkdir()
    for f in good:
        (mats / f).write_bytes(open(os.path.join(MTX, f), "rb").read())
    out = tmp_path / "out"
    out.mkdir()
    (out / "result.jsonl").write_text("stale\n")
    assert sweep.main([str(mats), "-a", "serial", "--out-dir", str(out)]) == 0
    rows = (out / "result.csv").read_text().splitlines()[1::2]
    side = [json.loads(line) for line in (out / "result.jsonl").read_text().splitlines()]
    assert len(rows) == len(side) == len(good)
    assert all(s["passed"] and s["rel_err_vs_fp64_serial"] == 0.0 for s in side)
