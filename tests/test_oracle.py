"""Pins of the CPU oracle (CPU only).

Each oracle function is checked against at least one of:
  [R] the reference's own host helpers compiled from its sources (oracle/_ref;
      skipped where that build is absent),
  [G] committed golden vectors recorded from those helpers (tests/golden),
  [C] closed forms / invariants, [D] dense triangular solves, [B] brute force.
A mutated oracle input (a dropped term, a transposed operand) must fail the
same checks -- see test_planted_mistakes_are_caught.

  read_mtx             [R][G]   error codes [R][G]
  change2tran          [R][G][C]
  get_x_b              [R][G][C]
  matrix_layer         [R][G][C][B]
  matrix_transposition [R][G][C]
  matrix_warp4         [R][G][C]  (chunk 64 and 32)
  csc executor         [R][G][D]  forward/backward, 1 and 3 RHS, fp64/fp32
  csr_forward          [C] bit-equal to the executor (unfused), [D]
  adjoint / grad_val   [R][G] (backward executor on L^T), [D]
  residual, datasize   [C]
  refine2              [C] x0 bit-equal to the FP32 forward solve (division form;
                       the reciprocal form equals it on power-of-two diagonals),
                       exact inputs give d = 0, per-RHS columns; [D] x within 1e-12
                       of the dense FP64 solve, d within 1e-5 of the dense
                       solve of the FP64 residual
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

import oracle
from conftest import ROOT, random_lower, rel_maxnorm, upper_from_lower, drop_diagonal

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))["cases"]
MTX = os.path.join(ROOT, "tests", "golden", "mtx")
need_ref = pytest.mark.skipif(not oracle.ref_available("d64"), reason="oracle/_ref not built (no reference tree)")


def f64(hexes):
    return np.array([float.fromhex(h) for h in hexes], dtype=np.float64)


def dense_lower(rp, ci, v):
    n = len(rp) - 1
    L = np.zeros((n, n))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            L[i, ci[j]] += v[j]
    return L


def random_A(n, seed, per_row=5):
    """Random square CSR with a full diagonal (the reference's change2tran
    allocates only nnzA entries for L, so every row needs an entry >= row)."""
    rng = np.random.default_rng(seed)
    rp, ci, v = [0], [], []
    for i in range(n):
        cols = list(rng.integers(0, n, size=per_row)) + [i]
        rng.shuffle(cols)
        ci += [int(c) for c in cols]
        v += list(rng.uniform(-1, 1, size=len(cols)))
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


# ------------------------------------------------------------------- goldens
@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_matches_golden(name):
    g = GOLD[name]
    ret, m, n, nnz, rp, ci, v = oracle.read_mtx(os.path.join(MTX, name))
    assert ret == g["read_mtx"]["ret"]
    if ret != 0:
        return
    r = g["read_mtx"]
    assert (m, n, nnz) == (r["m"], r["n"], r["nnz"])
    assert rp.tolist() == r["row_ptr"] and ci.tolist() == r["col_idx"] and np.array_equal(v, f64(r["val"]))
    mm = min(m, n)
    nnzL, rpL, ciL, vL = oracle.change2tran(mm, nnz, rp[:mm + 1], ci, v)
    c = g["change2tran"]
    assert nnzL == c["nnzL"] and rpL.tolist() == c["row_ptr"] and ciL.tolist() == c["col_idx"]
    assert np.array_equal(vL, f64(c["val"]))
    x, b = oracle.get_x_b(mm, mm, rpL, ciL, vL)
    assert np.array_equal(x, f64(g["get_x_b"]["x_ref"])) and np.array_equal(b, f64(g["get_x_b"]["b"]))
    layer, par = oracle.matrix_layer(mm, rpL, ciL)
    assert layer == g["matrix_layer"]["layer"] and par == float.fromhex(g["matrix_layer"]["parallelism"])
    cr, cp, cv = oracle.transpose(mm, mm, nnzL, rpL, ciL, vL)
    t = g["matrix_transposition"]
    assert cr.tolist() == t["csc_row"] and cp.tolist() == t["csc_colptr"] and np.array_equal(cv, f64(t["csc_val"]))
    for key, w in g["matrix_warp4"].items():
        chunk, border = map(int, key.split("_"))
        ln, wn, wo, eo = oracle.matrix_warp4(mm, nnzL, rpL, border, chunk)
        assert ln == w["len"] and wn.tolist() == w["warp_num"]
        assert wo == float.fromhex(w["warp_occupy"]) and eo == float.fromhex(w["element_occupy"])
    ex = g["executor"]
    vR = f64(ex["val"])
    crR, cpR, cvR = oracle.transpose(mm, mm, nnzL, rpL, ciL, vR)
    indeg = oracle.csc_analyser(crR, mm, nnzL)
    assert indeg.tolist() == ex["indeg"]
    for rhs in (1, 3):
        bb = f64(ex[f"b_{rhs}"])
        fw = oracle.csc_executor(cpR, crR, cvR, indeg, mm, mm, oracle.FORWARD, rhs, bb)
        assert np.array_equal(fw, f64(ex[f"forward_{rhs}"]))
        assert np.array_equal(oracle.csr_forward(rpL, ciL, vR, bb, rhs=rhs, fused=False), fw)
        indegT = np.bincount(ciL, minlength=mm).astype(np.int32)
        bw = oracle.csc_executor(rpL, ciL, vR, indegT, mm, mm, oracle.BACKWARD, rhs, bb)
        assert np.array_equal(bw, f64(ex[f"backward_T_{rhs}"]))
        if rhs == 1:
            assert np.array_equal(oracle.adjoint(rpL, ciL, vR, bb, fused=False), bw)
    v32 = vR.astype(np.float32)
    crS, cpS, cvS = oracle.transpose(mm, mm, nnzL, rpL, ciL, v32)
    b32 = f64(ex["b32"]).astype(np.float32)
    fw32 = oracle.csc_executor(cpS, crS, cvS, indeg, mm, mm, oracle.FORWARD, 1, b32)
    assert np.array_equal(fw32.astype(np.float64), f64(ex["forward32"]))


# ------------------------------------------------------- against oracle/_ref
@need_ref
@pytest.mark.parametrize("seed", range(6))
def test_helpers_match_reference_random(seed):
    n = [1, 2, 64, 65, 257, 3000][seed]
    rp, ci, v = random_A(n, seed)
    a = oracle.change2tran(n, len(ci), rp, ci, v)
    r = oracle.ref_change2tran(n, len(ci), rp, ci, v)
    assert a[0] == r[0] and np.array_equal(a[1], r[1]) and np.array_equal(a[2], r[2]) and np.array_equal(a[3], r[3])
    nnzL, rpL, ciL, vL = a
    for x, y in zip(oracle.get_x_b(n, n, rpL, ciL, vL), oracle.ref_get_x_b(n, n, rpL, ciL, vL)):
        assert np.array_equal(x, y)
    assert oracle.matrix_layer(n, rpL, ciL) == oracle.ref_matrix_layer(n, rpL, ciL, vL)
    for x, y in zip(oracle.transpose(n, n, nnzL, rpL, ciL, vL), oracle.ref_transpose(n, n, nnzL, rpL, ciL, vL)):
        assert np.array_equal(x, y)
    for variant, chunk in (("d64", 64), ("d32", 32)):
        for border in (1, 2, 3, 4):
            x = oracle.matrix_warp4(n, nnzL, rpL, border, chunk)
            y = oracle.ref_matrix_warp4(n, nnzL, rpL, ciL, vL, border, variant)
            assert x[0] == y[0] and np.array_equal(x[1], y[1]) and x[2] == y[2] and x[3] == y[3]


@need_ref
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("rhs", [1, 2, 5])
def test_executor_matches_reference(seed, rhs):
    rp, ci, v = random_lower(800 + 300 * seed, max_off=6, band=100, seed=seed, dup=seed % 2 == 1)
    n = len(rp) - 1
    b = np.random.default_rng(seed).uniform(-1, 1, size=n * rhs)
    cr, cp, cv = oracle.transpose(n, n, len(ci), rp, ci, v)
    indeg = oracle.csc_analyser(cr, n, len(ci))
    assert np.array_equal(indeg, oracle.ref_analyser(cr, n, len(ci)))
    for sub in (oracle.FORWARD, oracle.BACKWARD):
        cpp, rii, vv, dg = (cp, cr, cv, indeg) if sub == oracle.FORWARD else (
            rp, ci, v, np.bincount(ci, minlength=n).astype(np.int32))
        x = oracle.csc_executor(cpp, rii, vv, dg, n, n, sub, rhs, b)
        y = oracle.ref_csc_executor(cpp, rii, vv, dg, n, n, sub, rhs, b)
        assert np.array_equal(x, y)
    # single precision
    v32, b32 = v.astype(np.float32), b.astype(np.float32)
    cr, cp, cv = oracle.transpose(n, n, len(ci), rp, ci, v32)
    x = oracle.csc_executor(cp, cr, cv, indeg, n, n, oracle.FORWARD, rhs, b32)
    y = oracle.ref_csc_executor(cp, cr, cv, indeg, n, n, oracle.FORWARD, rhs, b32, variant="s64")
    assert np.array_equal(x, y)


@need_ref
def test_read_mtx_matches_reference_on_fixtures():
    for name in sorted(os.listdir(MTX)):
        a = oracle.read_mtx(os.path.join(MTX, name))
        r = oracle.ref_read_mtx(os.path.join(MTX, name))
        assert a[:4] == r[:4], name
        if a[0] == 0:
            for x, y in zip(a[4:], r[4:]):
                assert np.array_equal(x, y), name


# ------------------------------------------------- closed forms and invariants
def test_change2tran_closed_form():
    rp, ci, v = random_A(500, 11)
    nnzL, rpL, ciL, vL = oracle.change2tran(500, len(ci), rp, ci, v)
    for i in range(500):
        want = sorted(int(c) for c in ci[rp[i]:rp[i + 1]] if c < i) + [i]
        assert ciL[rpL[i]:rpL[i + 1]].tolist() == want
    assert np.all(vL == 1.0)


def test_unit_lower_known_answer():
    """x = 1 exactly: b_i is the row length and every partial sum is a small integer."""
    rp, ci, v = random_A(2000, 12)
    nnzL, rpL, ciL, vL = oracle.change2tran(2000, len(ci), rp, ci, v)
    x_ref, b = oracle.get_x_b(2000, 2000, rpL, ciL, vL)
    assert np.array_equal(b, np.diff(rpL).astype(np.float64))
    for fused in (False, True):
        assert np.array_equal(oracle.csr_forward(rpL, ciL, vL, b, fused=fused), x_ref)


def test_levels_closed_forms_and_brute_force():
    n = 50
    diag = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32))
    assert oracle.matrix_layer(n, *diag) == (1, float(n))
    bi_rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    bi_ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    assert oracle.matrix_layer(n, bi_rp, bi_ci) == (n, 1.0)
    # brute force: levels = 1 + length (in edges) of the longest dependency path
    for seed in range(5):
        rp, ci, _ = random_lower(14, max_off=3, band=14, seed=seed)
        m = len(rp) - 1
        deps = [[int(c) for c in ci[rp[i]:rp[i + 1] - 1]] for i in range(m)]

        def longest(i):
            return 1 + max((longest(c) for c in deps[i]), default=0)
        assert oracle.matrix_layer(m, rp, ci)[0] == max(longest(i) for i in range(m))
        assert oracle.row_levels(m, rp, ci).tolist() == [longest(i) for i in range(m)]


def test_transpose_twice_is_identity():
    rp, ci, v = random_lower(700, max_off=5, band=80, seed=3)
    n = len(rp) - 1
    cr, cp, cv = oracle.transpose(n, n, len(ci), rp, ci, v)
    rr, rpp, vv = oracle.transpose(n, n, len(ci), cp, cr, cv)
    assert np.array_equal(rpp, rp) and np.array_equal(rr, ci) and np.array_equal(vv, v)


def test_warp4_hand_computed():
    # chunk 4, border 2: rows have 1,1,1,1 | 3,3,3,3 | 1,2 entries
    rp = np.cumsum([0, 1, 1, 1, 1, 3, 3, 3, 3, 1, 2]).astype(np.int32)
    ln, wn, wo, eo = oracle.matrix_warp4(10, int(rp[-1]), rp, 2, 4)
    assert ln == 7 and wn.tolist() == [0, 4, 5, 6, 7, 8, 10]
    assert wo == 4 / 10 and eo == 12 / 19


def test_forward_against_dense_and_fused_vs_unfused():
    for seed in range(3):
        rp, ci, v = random_lower(1500, max_off=8, band=200, seed=seed)
        b = np.random.default_rng(seed).uniform(-1, 1, size=1500)
        L = dense_lower(rp, ci, v)
        xd = scipy.linalg.solve_triangular(L, b, lower=True)
        x0 = oracle.csr_forward(rp, ci, v, b, fused=False)
        x1 = oracle.csr_forward(rp, ci, v, b, fused=True)
        assert rel_maxnorm(x0, xd) <= 1e-12 and rel_maxnorm(x1, xd) <= 1e-12
        # multi-RHS interleave: column k equals the single-RHS solve of b[:, k]
        B = np.random.default_rng(seed + 9).uniform(-1, 1, size=(1500, 3))
        X = oracle.csr_forward(rp, ci, v, B.reshape(-1), rhs=3, fused=True).reshape(1500, 3)
        for k in range(3):
            assert np.array_equal(X[:, k], oracle.csr_forward(rp, ci, v, B[:, k].copy(), fused=True))


def test_adjoint_against_dense():
    rp, ci, v = random_lower(1200, max_off=6, band=100, seed=4)
    n = 1200
    L = dense_lower(rp, ci, v)
    g = np.random.default_rng(1).uniform(-1, 1, size=n)
    y = oracle.adjoint(rp, ci, v, g)
    assert rel_maxnorm(y, scipy.linalg.solve_triangular(L.T, g, lower=False)) <= 1e-12
    x = oracle.csr_forward(rp, ci, v, np.ones(n))
    gv = oracle.grad_val(rp, ci, y, x)
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.array_equal(gv, -y[rows] * x[ci])


def test_refine2_pins():
    rp, ci, v = random_lower(1500, max_off=8, band=200, seed=11)
    n = 1500
    b = np.random.default_rng(3).uniform(-1, 1, size=n)
    x0, d, x = oracle.refine2(rp, ci, v, b, rcp=False)
    # x0 is the FP32 solve on the rounded inputs
    assert np.array_equal(x0, oracle.csr_forward(rp, ci, v.astype(np.float32), b.astype(np.float32), fused=True))
    L = dense_lower(rp, ci, v)
    xd = scipy.linalg.solve_triangular(L, b, lower=True)
    assert rel_maxnorm(x0.astype(np.float64), xd) > 1e-10  # FP32 alone is not enough
    assert rel_maxnorm(x, xd) <= 1e-12
    assert np.array_equal(x, x0.astype(np.float64) + d.astype(np.float64))
    r = b - L @ x0.astype(np.float64)
    dd = scipy.linalg.solve_triangular(L, r, lower=True)
    assert rel_maxnorm(d.astype(np.float64), dd) <= 1e-5
    # exactly representable system: x0 = 1, residual 0, no correction
    w = np.where(np.arange(len(ci)) == np.repeat(rp[1:] - 1, np.diff(rp)), 1.0, 0.5)
    bw = dense_lower(rp, ci, w) @ np.ones(n)
    x0, d, x = oracle.refine2(rp, ci, w, bw)
    assert np.all(x0 == 1.0) and np.all(d == 0.0) and np.all(x == 1.0)
    # interleaved right-hand sides: column k is the single-RHS result
    B = np.random.default_rng(4).uniform(-1, 1, size=(n, 3))
    for rcp in (False, True):
        X0, D, X = (a.reshape(n, 3) for a in oracle.refine2(rp, ci, v, B.reshape(-1), rhs=3, rcp=rcp))
        for k in range(3):
            a0, ad, ax = oracle.refine2(rp, ci, v, B[:, k].copy(), rcp=rcp)
            assert np.array_equal(X0[:, k], a0) and np.array_equal(D[:, k], ad) and np.array_equal(X[:, k], ax)


def test_refine2_reciprocal_pins():
    """The reciprocal form (rows end with a product by fl(1/diag)): with
    power-of-two diagonals the product is the division exactly, so every
    output equals the division form; otherwise x0 differs in the last bits
    but the two steps still reach the dense FP64 solution."""
    rp, ci, v = random_lower(1500, max_off=8, band=200, seed=11)
    n = 1500
    b = np.random.default_rng(3).uniform(-1, 1, size=n)
    diag = rp[1:] - 1
    v2 = v.copy()
    v2[diag] = 2.0 ** np.random.default_rng(5).integers(-3, 4, size=n) * np.sign(v[diag])
    for a, e in zip(oracle.refine2(rp, ci, v2, b, rcp=True), oracle.refine2(rp, ci, v2, b, rcp=False)):
        assert np.array_equal(a, e)
    x0, d, x = oracle.refine2(rp, ci, v, b, rcp=True)
    x0d, _, _ = oracle.refine2(rp, ci, v, b, rcp=False)
    assert not np.array_equal(x0, x0d)  # the rounded reciprocal is in effect
    assert rel_maxnorm(x0.astype(np.float64), x0d.astype(np.float64)) <= 1e-5
    L = dense_lower(rp, ci, v)
    xd = scipy.linalg.solve_triangular(L, b, lower=True)
    assert rel_maxnorm(x, xd) <= 1e-12
    assert np.array_equal(x, x0.astype(np.float64) + d.astype(np.float64))
    dd = scipy.linalg.solve_triangular(L, b - L @ x0.astype(np.float64), lower=True)
    assert rel_maxnorm(d.astype(np.float64), dd) <= 1e-5


def test_residual_and_datasize():
    rp, ci, v = random_lower(300, seed=5)
    b = np.random.default_rng(2).uniform(-1, 1, size=300)
    x = oracle.csr_forward(rp, ci, v, b)
    r = oracle.residual(rp, ci, v, b, x)
    assert np.max(np.abs(r)) <= 1e-14
    assert np.allclose(oracle.residual(rp, ci, v, b, np.zeros(300)), b, rtol=0, atol=0)
    # spts_syncfree_cuda_csr_wrt.h:208 with V = 8: 4(n+1) + 4(nnz+m) + 8 nnz + 16 n
    assert oracle.datasize(10, 10, 37, 8) == 4 * 11 + 4 * 47 + 8 * 37 + 16 * 10
    assert oracle.datasize(10, 10, 37, 4, 5) == 4 * 11 + 4 * 47 + 4 * 37 + 8 * 10 + 20


def test_planted_mistakes_are_caught():
    """The checks above fail for plausible oracle mistakes (planted here, on the CPU)."""
    rp, ci, v = random_lower(400, max_off=6, band=50, seed=8)
    n = 400
    b = np.random.default_rng(3).uniform(-1, 1, size=n)
    L = dense_lower(rp, ci, v)
    xd = scipy.linalg.solve_triangular(L, b, lower=True)
    # dropped term: lose one off-diagonal entry
    k = next(j for j in range(rp[10], rp[11] - 1))
    v_drop = v.copy()
    v_drop[k] = 0.0
    assert rel_maxnorm(oracle.csr_forward(rp, ci, v_drop, b), xd) > 1e-6
    # transposed operand: solving with L^T's pattern instead of L
    assert rel_maxnorm(oracle.adjoint(rp, ci, v, b), xd) > 1e-6
    # wrong diagonal: first entry instead of last
    v_sw = v.copy()
    for i in range(n):
        a, d = rp[i], rp[i + 1] - 1
        if d > a:
            v_sw[a], v_sw[d] = v_sw[d], v_sw[a]
    assert rel_maxnorm(oracle.csr_forward(rp, ci, v_sw, b), xd) > 1e-6


@pytest.mark.parametrize("dup", [False, True])
def test_reversed_transpose_forward_equals_backward_executor(dup):
    """The gather adjoint's identity: a forward solve on the reversed transpose
    (rows of L visited in descending order, entries in stored order) is the
    executor's BACKWARD mode on L read as CSC of L^T, bit for bit."""
    rp, ci, v = random_lower(2500, max_off=5, band=12, seed=31, dup=dup)
    n, nnz = len(rp) - 1, len(ci)
    cnt = np.bincount(n - 1 - ci, minlength=n)
    rT = np.r_[0, np.cumsum(cnt)].astype(np.int32)
    fill = rT[:-1].astype(np.int64).copy()
    cT, vT = np.zeros(nnz, np.int32), np.zeros(nnz)
    for i in range(n - 1, -1, -1):
        for j in range(rp[i], rp[i + 1]):
            o = fill[n - 1 - ci[j]]
            fill[n - 1 - ci[j]] += 1
            cT[o], vT[o] = n - 1 - i, v[j]
    g = np.random.default_rng(4).uniform(-1, 1, size=n)
    for fused in (False, True):
        y = oracle.csr_forward(rT, cT, vT, g[::-1].copy(), fused=fused)[::-1]
        assert np.array_equal(y, oracle.adjoint(rp, ci, v, g, fused=fused))


# ------------------------------------------------ upper / unit-diagonal solves
def _dense(n, rp, ci, v):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            A[i, ci[j]] += v[j]
    return A


def test_csr_solve_lower_is_the_forward_substitution():
    rp, ci, v = random_lower(500, max_off=6, band=50, seed=21, dup=True)
    b = np.random.default_rng(1).uniform(-1, 1, 500)
    for fused in (True, False):
        assert np.array_equal(oracle.csr_solve(rp, ci, v, b, fused=fused), oracle.csr_forward(rp, ci, v, b, fused=fused))


@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("upper", [False, True])
def test_csr_solve_against_dense_triangular_solve(upper, unit):
    from scipy.linalg import solve_triangular
    n = 400
    rp, ci, v = random_lower(n, max_off=5, band=60, seed=22)
    if upper:
        rp, ci, v = upper_from_lower(rp, ci, v)
    if unit:
        rp, ci, v = drop_diagonal(rp, ci, v, upper=upper)
    b = np.random.default_rng(2).uniform(-1, 1, (n, 3))
    x = oracle.csr_solve(rp, ci, v, b.reshape(-1), rhs=3, upper=upper, unit=unit).reshape(n, 3)
    A = _dense(n, rp, ci, v)
    if unit:
        A += np.eye(n)
    ref = solve_triangular(A, b, lower=not upper, unit_diagonal=unit)
    assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) <= 1e-12


def test_csr_solve_closed_forms():
    n = 200
    # unit upper bidiagonal with -1 on the superdiagonal, b = 1: x_i = n - i
    rp = np.r_[np.arange(0, n), n - 1].astype(np.int32)  # rows 0..n-2 one entry, row n-1 none
    ci = np.arange(1, n).astype(np.int32)
    v = -np.ones(n - 1)
    x = oracle.csr_solve(rp, ci, v, np.ones(n), upper=True, unit=True)
    assert np.array_equal(x, (n - np.arange(n)).astype(float))
    # lower with stored diagonal 2 and -1 below, b = 1: x_i = 1 - 2^-(i+1)
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    v = np.r_[2.0, np.tile([-1.0, 2.0], n - 1)]
    x = oracle.csr_solve(rp, ci, v, np.ones(n))
    assert np.allclose(x, 1 - 0.5 ** (np.arange(n) + 1), rtol=0, atol=1e-15)


def test_upper_solve_planted_diagonal_mistake_is_caught():
    """A restatement that takes the upper row's diagonal from its last entry
    (the lower convention) fails the dense check."""
    from scipy.linalg import solve_triangular
    n = 120
    rp, ci, v = upper_from_lower(*random_lower(n, max_off=4, band=20, seed=23))
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x = np.zeros(n)
    for i in range(n - 1, -1, -1):
        d = rp[i + 1] - 1  # wrong: the diagonal is rp[i]
        s = sum(v[j] * x[ci[j]] for j in range(rp[i], rp[i + 1]) if j != d)
        x[i] = (b[i] - s) / v[d]
    ref = solve_triangular(_dense(n, rp, ci, v), b, lower=False)
    assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) > 1e-6
