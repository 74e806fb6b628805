"""GPU parity of the forward solve kernels (through the C ABI) against the oracle.

Bar: bit-exact wherever the kernel's accumulation order is the oracle's
sequential fused order (thread-per-row and mix paths in both readiness
protocols, every right-hand-side count); FP64 within 1e-12 relative max-norm
where the order differs (CSC: atomic scatter); FP32 within 1e-5 of the FP64
oracle.
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import random_lower, rel_maxnorm

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _solve(cuda, rp, ci, v, b, algo, sync, k=1, border=0, lds=False, bpc=0, readiness="auto", threads=0):
    import torch
    from capellinisptrsv_amd import Plan
    plan = Plan(torch.from_numpy(rp).to(cuda), torch.from_numpy(ci).to(cuda), torch.from_numpy(v).to(cuda),
                algo=algo, sync=sync, border=border, lds_window=lds, blocks_per_cu=bpc, readiness=readiness,
                row_threads=threads)
    bt = torch.from_numpy(np.ascontiguousarray(b.reshape(-1, k) if k > 1 else b)).to(cuda)
    x = plan.solve(bt)
    torch.cuda.synchronize()
    assert plan.status() == 0
    info = plan.info()
    out = x.cpu().numpy().reshape(-1)
    plan.close()
    return out, info


def _matrices():
    yield "band", random_lower(3000, max_off=5, band=40, seed=1)
    yield "dup", random_lower(2000, max_off=6, band=30, seed=2, dup=True)
    yield "cfg0", api.make_config(0, seed=0)
    yield "cfg1_small", api.make_config(1, seed=0, scale=0.05)  # 10 levels x 10,000 rows
    yield "cfg2_small", api.make_config(2, seed=0, scale=0.02)  # n = 10,000, ~38 nnz/row
    yield "chain", (np.r_[0, 1, np.arange(3, 2 * 4000, 2)].astype(np.int32),
                    np.r_[0, np.c_[np.arange(0, 3999), np.arange(1, 4000)].reshape(-1)].astype(np.int32),
                    np.r_[1.0, np.tile([0.5, 2.0], 3999)])


MATS = list(_matrices())


@pytest.mark.parametrize("algo,sync,lds,bpc,threads", [
    ("wrtfst", "value", False, 1, 0), ("wrtfst", "value", True, 1, 0), ("wrtfst", "value", True, 4, 0),
    ("wrtfst", "flag", False, 2, 0), ("wrtfst", "value", False, 1, 128), ("wrtfst", "value", True, 2, 128),
    ("wrtfst", "flag", False, 1, 128), ("twopart", "value", False, 1, 0), ("twopart", "flag", False, 1, 0),
    ("local", "value", False, 1, 0)])
@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_thread_per_row_bit_exact_fp64(cuda, algo, sync, lds, bpc, threads, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(7, n)
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    x, info = _solve(cuda, rp, ci, v, b, algo, sync, lds=lds, bpc=bpc, threads=threads)
    assert info["algo"] == algo and info["sync"] == sync
    if algo == "wrtfst":
        assert info["row_threads"] == (threads or 256)
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), rel_maxnorm(x, ref)
    # and within tolerance of the reference's unfused serial arithmetic
    assert rel_maxnorm(x, oracle.csr_forward(rp, ci, v, b, fused=False)) <= FP64_TOL


@pytest.mark.parametrize("algo,sync", [("wrtfst", "value"), ("wrtfst", "flag"), ("twopart", "value"),
                                       ("local", "value")])
@pytest.mark.parametrize("name,mat", MATS[:4], ids=[m[0] for m in MATS[:4]])
def test_fp32_bit_exact_and_within_1e5(cuda, algo, sync, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(11, n)
    v32, b32 = v.astype(np.float32), b.astype(np.float32)
    x, _ = _solve(cuda, rp, ci, v32, b32, algo, sync)
    ref32 = oracle.csr_forward(rp, ci, v32, b32, fused=True)
    assert np.array_equal(x.view(np.uint32), ref32.view(np.uint32))
    assert rel_maxnorm(x, oracle.csr_forward(rp, ci, v, b, fused=True)) <= FP32_TOL


@pytest.mark.parametrize("border", [4, 16, 10**6])
@pytest.mark.parametrize("sync", ["value", "flag"])
@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_mix(cuda, border, sync, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(3, n)
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    x, info = _solve(cuda, rp, ci, v, b, "mix", sync, border=border)
    ln, wn, wo, eo = oracle.matrix_warp4(n, len(ci), rp, border, 64)
    assert info["mix_len"] == ln and abs(info["warp_occupy"] - wo) < 1e-15
    # thread rows and wavefront rows both accumulate in the sequential fused order
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), rel_maxnorm(x, ref)


@pytest.mark.parametrize("name,mat", [MATS[2], MATS[3], MATS[4]], ids=["cfg0", "cfg1_small", "cfg2_small"])
def test_mix_threshold_from_rows_per_level(cuda, name, mat):
    """mix without a border takes its chunk threshold from the matrix's rows
    per level: 4 where levels hold at least 200 rows, else 32; the solve
    stays bit-exact and the chunk table is the oracle's at that border."""
    rp, ci, v = mat
    n = len(rp) - 1
    levels = oracle.matrix_layer(n, rp, ci)[0]
    want = 4 if n >= 200 * levels else 32
    b = api.uniform(9, n)
    x, info = _solve(cuda, rp, ci, v, b, "mix", "value", border=0)
    assert info["border"] == want, (name, n / levels)
    ln, _, wo, _ = oracle.matrix_warp4(n, len(ci), rp, want, 64)
    assert info["mix_len"] == ln and abs(info["warp_occupy"] - wo) < 1e-15
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=True))


@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_csc_syncfree(cuda, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(5, n)
    x, info = _solve(cuda, rp, ci, v, b, "csc", "value")
    assert info["algo"] == "csc"
    assert rel_maxnorm(x, oracle.csr_forward(rp, ci, v, b, fused=False)) <= FP64_TOL


@pytest.mark.parametrize("k", [2, 3, 8, 16, 32, 64])
@pytest.mark.parametrize("algo,sync,lds,threads", [("wrtfst", "value", False, 0), ("wrtfst", "value", True, 0),
                                                   ("wrtfst", "value", True, 128), ("wrtfst", "flag", False, 0),
                                                   ("local", "value", False, 0)])
def test_multi_rhs_bit_exact(cuda, k, algo, sync, lds, threads):
    rp, ci, v = api.make_config(0, seed=3, scale=0.2)
    n = len(rp) - 1
    b = api.uniform(100 + k, n * k)
    ref = oracle.csr_forward(rp, ci, v, b, rhs=k, fused=True)
    x, _ = _solve(cuda, rp, ci, v, b, algo, sync, k=k, lds=lds, threads=threads)
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name,mat", [MATS[0], MATS[3], MATS[5]], ids=["band", "cfg1_small", "chain"])
def test_rhs_per_lane_bit_exact(cuda, name, mat, dtype):
    """1024-thread workgroups with 1, 2, 4 (8 for FP32) right-hand sides per
    lane, at right-hand-side counts that do and do not fill the lanes (a lane
    owning fewer than R of them): bit-exact against the fused oracle, and in
    the reference arithmetic against the unfused one."""
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = mat
    v = v.astype(dtype)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    rpls = (1, 2, 4, 8) if dtype == np.float32 else (1, 2, 4)
    for k in (2, 3, 5, 16, 33, 64):
        b = api.uniform(200 + k, n * k).astype(dtype)
        bt = torch.from_numpy(b.reshape(n, k)).to(cuda)
        for arith, fused in (("fused", True), ("reference", False)):
            if arith == "reference" and k not in (3, 16):
                continue
            ref = oracle.csr_forward(rp, ci, v, b, rhs=k, fused=fused)
            for r in rpls:
                plan = Plan(*d, algo="wrtfst", sync="value", rhs_per_lane=r, arith=arith)
                x = plan.solve(bt).cpu().numpy().reshape(-1)
                assert plan.status() == 0
                plan.close()
                assert np.array_equal(x, ref), (name, k, r, arith)


@pytest.mark.parametrize("lds,threads", [(False, 0), (True, 0), (False, 128), (True, 128)])
@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_row_tiles_bit_exact(cuda, lds, threads, name, mat):
    """Tickets' col_idx/val ranges staged in LDS: bit-exact against the fused
    oracle at one and several right-hand sides, in FP64 and FP32.  cfg2_small
    (~38 nnz/row) has ranges above the 2048-entry tile at 256 rows a ticket,
    so its tickets read global memory; the others run from LDS."""
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = mat
    n = len(rp) - 1
    for dtype in (np.float64, np.float32):
        vv = v.astype(dtype)
        d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, vv)]
        for k in (1, 3, 16):
            b = api.uniform(300 + k, n * k).astype(dtype)
            bt = torch.from_numpy(b.reshape(n, k) if k > 1 else b).to(cuda)
            plan = Plan(*d, algo="wrtfst", sync="value", lds_window=lds, row_threads=threads, row_tiles=True)
            info = plan.info()
            assert info["row_tiles"] and info["row_threads"] == (threads or 256)
            x = plan.solve(bt).cpu().numpy().reshape(-1)
            assert plan.status() == 0
            plan.close()
            ref = oracle.csr_forward(rp, ci, vv, b, rhs=k, fused=True)
            assert np.array_equal(x, ref), (name, dtype, k, rel_maxnorm(x, ref))


@pytest.mark.parametrize("border", [4, 16, 10**6])
@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_mix_row_tiles_bit_exact(cuda, border, name, mat):
    """mix with the ticket's rows staged in LDS (thread-per-row units read
    the tile, single-row units global memory): bit-exact in FP64 and FP32."""
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = mat
    n = len(rp) - 1
    for dtype in (np.float64, np.float32):
        vv = v.astype(dtype)
        b = api.uniform(3, n).astype(dtype)
        plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, vv)), algo="mix", sync="value",
                    border=border, row_tiles=True)
        assert plan.info()["row_tiles"]
        x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
        assert plan.status() == 0
        plan.close()
        ref = oracle.csr_forward(rp, ci, vv, b, fused=True)
        assert np.array_equal(x, ref), (name, dtype, rel_maxnorm(x, ref))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name,mat", [MATS[0], MATS[3], MATS[4], MATS[5]],
                         ids=["band", "cfg1_small", "cfg2_small", "chain"])
def test_row_tiles_1024_threads_bit_exact(cuda, name, mat, dtype):
    """The 1024-thread shape with one right-hand side per lane and row tiles
    (the batch tuner's pick on many matrices): bit-exact against the fused
    oracle at several right-hand-side counts; cfg2_small's blocks overflow
    the tile and read global memory."""
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = mat
    v = v.astype(dtype)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    plan = Plan(*d, algo="wrtfst", sync="value", rhs_per_lane=1, row_tiles=True)
    for k in (2, 3, 16, 64):
        b = api.uniform(400 + k, n * k).astype(dtype)
        x = plan.solve(torch.from_numpy(b.reshape(n, k)).to(cuda)).cpu().numpy().reshape(-1)
        assert plan.status() == 0
        assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, rhs=k, fused=True)), (name, k)
    plan.close()


@pytest.mark.parametrize("upper,unit", [(True, False), (False, True), (True, True)])
def test_row_tiles_triangular_kinds(cuda, upper, unit):
    """Row tiles on upper (reversed rows) and implicit-unit-diagonal solves."""
    import torch
    from capellinisptrsv_amd import Plan
    from conftest import drop_diagonal, upper_from_lower
    rp, ci, v = random_lower(3000, max_off=6, band=200, seed=4, dup=True)
    if upper:
        rp, ci, v = upper_from_lower(rp, ci, v)
    if unit:
        rp, ci, v = drop_diagonal(rp, ci, v, upper=upper)
    n = len(rp) - 1
    b = api.uniform(41, n)
    plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), algo="wrtfst", sync="value",
                upper=upper, unit_diagonal=unit, row_tiles=True)
    x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
    assert plan.status() == 0
    plan.close()
    ref = oracle.csr_solve(rp, ci, v, b, upper=upper, unit=unit, fused=True)
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), rel_maxnorm(x, ref)


def test_local_deep_band_bit_exact(cuda):
    """One workgroup walks a deep banded matrix (n = 20,000, bandwidth 128,
    ~38 nnz/row, thousands of levels) and matches the oracle bit for bit."""
    rp, ci, v = api.make_config(2, seed=5, scale=0.04)
    n = len(rp) - 1
    b = api.uniform(21, n)
    x, info = _solve(cuda, rp, ci, v, b, "local", "value")
    assert info["algo"] == "local"
    assert np.array_equal(x.view(np.uint64), oracle.csr_forward(rp, ci, v, b, fused=True).view(np.uint64))


def test_unit_lower_known_answer(cuda):
    """change2tran's all-ones L with b = L*1 has the exact solution x = 1 (read_mtx.h:260-278)."""
    rng = np.random.default_rng(4)
    n = 20000
    rows, cols = [], []
    for i in range(n):
        for c in rng.integers(0, n, size=5):
            rows.append(i)
            cols.append(int(c))
        rows.append(i)
        cols.append(i)
    order = np.lexsort((np.arange(len(rows)), rows))
    rpA = np.searchsorted(np.array(rows)[order], np.arange(n + 1)).astype(np.int32)
    ciA = np.array(cols, np.int32)[order]
    vA = np.ones(len(ciA))
    nnzL, rp, ci, v = api.change2tran(n, len(ciA), rpA, ciA, vA)
    x_ref, b = api.get_x_b(n, n, rp, ci, v)
    for algo in ("wrtfst", "twopart", "mix", "csc"):
        x, _ = _solve(cuda, rp, ci, v, b, algo, "value", border=4)
        assert np.array_equal(x, x_ref), algo


def test_diagonal_and_tiny(cuda):
    for n in (1, 2, 63, 64, 65):
        rp = np.arange(n + 1, dtype=np.int32)
        ci = np.arange(n, dtype=np.int32)
        v = np.linspace(1, 2, n)
        b = np.linspace(-1, 1, n)
        for algo in ("wrtfst", "twopart", "mix", "csc"):
            x, _ = _solve(cuda, rp, ci, v, b, algo, "value")
            assert np.array_equal(x, b / v), (n, algo)


def test_unsorted_offdiagonals_follow_stored_order(cuda):
    rp, ci, v = random_lower(3000, max_off=6, band=100, seed=9, shuffle=True)
    b = api.uniform(1, len(rp) - 1)
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    x, _ = _solve(cuda, rp, ci, v, b, "wrtfst", "value")
    assert np.array_equal(x, ref)


def test_nan_input_does_not_hang(cuda):
    """A NaN in b is published as an ordinary (quiet) NaN; dependents finish."""
    rp, ci, v = random_lower(5000, max_off=3, band=20, seed=5)
    b = api.uniform(2, len(rp) - 1)
    b[100] = np.nan
    b[200] = np.frombuffer(np.uint64(0xFFF0000000000001).tobytes(), np.float64)[0]  # the sentinel's bits
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    x, _ = _solve(cuda, rp, ci, v, b, "wrtfst", "value")
    assert np.array_equal(np.isnan(x), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(x[ok], ref[ok])


def test_invalid_inputs_rejected_before_launch(cuda):
    import torch
    from capellinisptrsv_amd import Plan, CapError
    rp, ci, v = random_lower(500, seed=6)
    r = next(i for i in range(40, 500) if rp[i + 1] - rp[i] >= 2)  # a row with an off-diagonal
    upper = ci.copy()
    upper[rp[r]] = r + 3
    no_diag = ci.copy()
    no_diag[rp[r + 1] - 1] = 0
    zero_diag = v.copy()
    zero_diag[rp[r + 1] - 1] = 0.0
    bad_rp = rp.copy()
    bad_rp[r + 1] = bad_rp[r] - 1  # row r ends before it starts
    cases = {"upper": (rp, upper, v, -4), "no_diag": (rp, no_diag, v, -5), "zero_diag": (rp, ci, zero_diag, -6),
             "rowptr": (bad_rp, ci, v, -3)}
    for name, (rr, c, vv, code) in cases.items():
        with pytest.raises(CapError) as e:
            Plan(torch.from_numpy(rr).to(cuda), torch.from_numpy(c).to(cuda), torch.from_numpy(vv).to(cuda),
                 algo="wrtfst")
        assert e.value.code == code, name


def test_watchdog_forced_from_host(cuda):
    """A 1-poll spin limit makes waits give up: rows publish NaN, the grid drains, status reports it."""
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.make_config(1, seed=0, scale=0.02)
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    plan = Plan(*d, algo="wrtfst", sync="value", spin_limit=1)
    b = torch.ones(len(rp) - 1, dtype=torch.float64, device=cuda)
    x = plan.solve(b)
    torch.cuda.synchronize()
    assert plan.status() == -13
    assert plan.status() == 0  # cleared
    assert torch.isnan(x).any()


@pytest.mark.parametrize("algo", ["wrtfst", "mix", "local"])
def test_watchdog_wall_time_bound(cuda, algo):
    """No poll-count bound, a 1 ms wall-time bound: a dependency that can never
    be published (row 10 of a chain also reads row 4000, which depends on row
    10; layout validation off) makes its wait give up once 1 ms has passed
    without progress, the row publishes NaN, and every kernel drains --
    the bound that holds whatever a poll costs."""
    import time
    import torch
    from capellinisptrsv_amd import Plan
    n = 5000
    rp, ci, v = [0], [], []
    for i in range(n):
        cols = ([i - 1] if i else []) + ([4000] if i == 10 else [])
        ci += cols + [i]
        v += [0.5] * len(cols) + [2.0]
        rp.append(len(ci))
    rp, ci, v = np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    # a poll-count bound far beyond the time bound stays as a backstop (tens
    # of seconds of polls): the solve must end well before it could fire
    plan = Plan(*d, algo=algo, sync="value", validate=False, spin_limit=1 << 26, wait_limit_ms=1, border=1)
    b = torch.ones(n, dtype=torch.float64, device=cuda)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = plan.solve(b)
    assert plan.status() == -13
    elapsed = time.perf_counter() - t0
    assert elapsed < 2.0, elapsed
    x = x.cpu().numpy()
    assert not np.isnan(x[:10]).any() and np.isnan(x[10:]).all()
    plan.close()


def test_autotuner_picks_a_correct_path(cuda):
    # deep band (few rows per level): the one-workgroup kernel and, the band
    # being narrow, the partitioned solve are candidates
    rp, ci, v = api.make_config(2, seed=1, scale=0.01)
    b = api.uniform(9, len(rp) - 1)
    x, info = _solve(cuda, rp, ci, v, b, "auto", "auto")
    assert info["algo"] in ("wrtfst", "mix", "twopart", "local", "partition")
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    if info["algo"] == "partition":  # reassociated sums: the FP64 tolerance
        assert rel_maxnorm(x, ref) <= FP64_TOL
    else:
        assert np.array_equal(x.view(np.uint64), ref.view(np.uint64))
    # wide levels: it is not (n / levels far above the threshold)
    rp, ci, v = api.make_config(1, seed=1, scale=0.02)
    b = api.uniform(9, len(rp) - 1)
    x, info = _solve(cuda, rp, ci, v, b, "auto", "auto")
    assert info["algo"] in ("wrtfst", "mix")
    assert np.array_equal(x.view(np.uint64), oracle.csr_forward(rp, ci, v, b, fused=True).view(np.uint64))


@pytest.mark.parametrize("algo,lds,k", [("wrtfst", False, 1), ("wrtfst", True, 1), ("wrtfst", False, 8),
                                        ("twopart", False, 1), ("mix", False, 1)])
@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_staged_and_in_place_readiness_bit_exact(cuda, algo, lds, k, name, mat):
    """The staged readiness array (uncached memory, results also stored into
    x) and polling x in place give the oracle's bits, and the plan reports
    the choice it was given."""
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(11, n * k)
    ref = np.stack([oracle.csr_forward(rp, ci, v, b.reshape(n, k)[:, q].copy(), fused=True)
                    for q in range(k)], axis=1).reshape(-1)
    for rd in ("staged", "in_place"):
        x, info = _solve(cuda, rp, ci, v, b, algo, "value", k=k, border=16, lds=lds, readiness=rd)
        assert info["readiness"] == rd
        assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), (rd, name)


def test_staged_readiness_reused_across_plans_and_sizes(cuda):
    """Plans of growing and shrinking size share pool blocks one after the
    other; every solve is still the oracle's, and x is fully written."""
    import torch
    from capellinisptrsv_amd import Plan
    for n, seed in ((5000, 1), (20000, 2), (3000, 3), (20000, 4), (40000, 5)):
        rp, ci, v = random_lower(n, max_off=4, band=300, seed=seed)
        b = api.uniform(seed, n)
        plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), algo="wrtfst", sync="value",
                    readiness="staged")
        out = torch.full((n,), 12345.0, dtype=torch.float64, device=cuda)
        for _ in range(2):
            plan.solve(torch.from_numpy(b).to(cuda), out=out)
        torch.cuda.synchronize()
        assert plan.status() == 0
        assert np.array_equal(out.cpu().numpy(), oracle.csr_forward(rp, ci, v, b, fused=True))
        plan.close()


@pytest.mark.parametrize("k", [1, 4])
def test_local_long_rows_use_global_mode_bit_exact(cuda, k):
    """Rows with more than 63 off-diagonal entries do not fit the one-load row
    staging of the ring mode; such a matrix runs in the global mode."""
    rp, ci, v = random_lower(3000, max_off=150, band=400, seed=11)
    assert int(np.max(np.diff(rp))) - 1 > 63
    n = len(rp) - 1
    b = api.uniform(31, n * k)
    ref = oracle.csr_forward(rp, ci, v, b, rhs=k, fused=True)
    x, info = _solve(cuda, rp, ci, v, b, "local", "value", k=k)
    assert info["algo"] == "local"
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("lds", [False, True])
def test_poll_counter_build(cuda, lds):
    """The counting build of the write-first solve gives the same bits and
    counts at least one poll per off-diagonal entry (every dependency is read
    at least once); reading clears the count; plans without it refuse."""
    import torch
    from capellinisptrsv_amd import CapError, Plan
    rp, ci, v = api.make_config(1, seed=0, scale=0.05)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    b = api.uniform(71, n)
    plan = Plan(*d, algo="wrtfst", sync="value", lds_window=lds, count_polls=True)
    for _ in range(2):
        x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
        assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=True))
        polls = plan.poll_count()
        assert polls >= len(ci) - n, polls
        assert polls < 1000 * len(ci), polls
        assert plan.poll_count() == 0
    plan.close()
    plain = Plan(*d, algo="wrtfst", sync="value")
    with pytest.raises(CapError):
        plain.poll_count()
    plain.close()


@pytest.mark.parametrize("cfg,scale", [(1, 0.05), (2, 0.02), (0, 1.0)])
def test_rhs_block_tuned_per_count(cuda, cfg, scale):
    """An autotuned plan times its RHS block on the first solve with each new
    right-hand-side count (thread-per-row: workgroups per CU, threads per
    workgroup and right-hand sides per lane; partitioned: right-hand sides per
    wavefront); every solve, tuning or not, gives the oracle's results.  A
    plan whose single-RHS pick is mix (the 100k band) solves several with the
    write-first kernel."""
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.make_config(cfg, seed=4, scale=scale)
    n = len(rp) - 1
    plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)))
    algo = plan.info()["algo"]
    if algo in ("mix", "twopart", "csc"):
        algo = "wrtfst"
    for k in (8, 16, 8):
        assert plan.rhs_block(k) is None or k == 8
        b = api.uniform(80 + k, n * k)
        x = plan.solve(torch.from_numpy(b.reshape(n, k)).to(cuda)).cpu().numpy().reshape(-1)
        assert plan.status() == 0
        ref = oracle.csr_forward(rp, ci, v, b, rhs=k, fused=True)
        if algo == "partition":
            assert rel_maxnorm(x, ref) <= FP64_TOL
        else:
            assert np.array_equal(x, ref), (algo, k)
        pick = plan.rhs_block(k)
        assert pick is not None
        if algo == "partition":
            assert pick["rhs_per_wave"] in (1, 2, 4) and k % pick["rhs_per_wave"] == 0
        elif algo == "wrtfst":
            assert pick["blocks_per_cu"] in (1, 2, 4) and pick["row_threads"] in (128, 256, 1024)
            assert pick["rhs_per_lane"] in (1, 2, 4, 8) and pick["rhs_per_lane"] <= k
            assert pick["readiness"] in ("staged", "in_place")
            # several right-hand sides per lane (1024 threads) have no tiled form
            assert not pick["row_tiles"] or not (pick["row_threads"] == 1024 and pick["rhs_per_lane"] > 1), pick
    plan.close()
