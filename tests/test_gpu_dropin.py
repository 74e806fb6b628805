"""The benchmark-compatible wrappers and the ./sptrsv_main drop-in driver on the GPU.

They build x_ref = 1, b = L*1 internally and validate every repetition with
the reference's own tests (elementwise 1 % for the CSR wrappers, relative L1
< 1e-4 for the CSC one): a passing run reports solve_time > 0.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import ROOT

pytestmark = pytest.mark.gpu


def _unit_L(n=30000, seed=0):
    rp, ci, v = api.make_config(0, seed=seed, scale=n / 100000)
    nnzL, rpL, ciL, vL = api.change2tran(len(rp) - 1, len(ci), rp, ci, v)
    return rpL, ciL, vL


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_wrappers(cuda, dtype):
    rp, ci, v = _unit_L()
    v = v.astype(dtype)
    n = len(rp) - 1
    for fn in (api.spts_syncfree_hip_csr_wrt, api.spts_syncfree_hip_csr):
        ret, t, gf, bw = fn(rp, ci, v, n, n, len(ci))
        assert ret == 0 and t > 0 and gf > 0 and bw > 0, fn.__name__
        assert abs(bw - oracle.datasize(n, n, len(ci), v.itemsize) / (1e6 * t)) <= 1e-9 * bw
        assert abs(gf - 2 * len(ci) / (1e6 * t)) <= 1e-9 * gf
    ret, t, gf, bw, pre, wo, eo = api.spts_syncfree_hip_mix(rp, ci, v, n, n, len(ci))
    # the chunk threshold follows rows per level: 4 from 200 rows per level, else 32
    levels = oracle.matrix_layer(n, rp, ci)[0]
    ln, _, wo_ref, eo_ref = oracle.matrix_warp4(n, len(ci), rp, 4 if n >= 200 * levels else 32, 64)
    assert ret == 0 and t > 0 and pre >= 0 and wo == wo_ref and eo == eo_ref
    assert abs(bw - oracle.datasize(n, n, len(ci), v.itemsize, ln) / (1e6 * t)) <= 1e-9 * bw


def test_not_square_returns_minus_one(cuda):
    rp, ci, v = _unit_L(1000)
    n = len(rp) - 1
    assert api.spts_syncfree_hip_csr_wrt(rp, ci, v, n, n + 1, len(ci))[0] == -1


def test_csc_wrapper(cuda):
    rp, ci, v = _unit_L()
    n = len(rp) - 1
    cr, cp, cv = api.matrix_transposition(n, n, len(ci), rp, ci, v)
    x_ref, b = api.get_x_b(n, n, rp, ci, v)
    ret, x, pre, t, gf, bw = api.sptrsv_syncfree_hip_csc(cp, cr, cv, n, n, len(ci), b, x_ref)
    assert ret == 0 and t > 0 and pre > 0
    assert np.array_equal(x, x_ref)


def test_driver_appends_result_csv(cuda, tmp_path):
    mtx = tmp_path / "tri.mtx"
    rng = np.random.default_rng(1)
    n = 5000
    lines = ["%%MatrixMarket matrix coordinate real general", "% generated", f"{n} {n} {4 * n}"]
    for i in range(n):
        for c in rng.integers(0, n, size=3):
            lines.append(f"{i + 1} {int(c) + 1} {rng.uniform(-1, 1):.6f}")
        lines.append(f"{i + 1} {i + 1} 4.0")
    mtx.write_text("\n".join(lines) + "\n")
    env = dict(os.environ)
    for algo, exe in (("wrtfst", "sptrsv_main"), ("mix", "sptrsv_main"), ("2part", "sptrsv_main_f32"),
                      ("csc", "sptrsv_main")):
        r = subprocess.run([os.path.join(ROOT, "capellinisptrsv_amd", exe), str(mtx), "-a", algo], cwd=tmp_path,
                           env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (algo, r.stdout, r.stderr)
    rows = (tmp_path / "result.csv").read_text().splitlines()
    assert len(rows) == 8
    assert rows[0] == "filename,m,nnzL,nnzL/m,layer,parallelism,pre_time,solve_time,gflops,bandwith"
    assert rows[2].startswith("filename,m,nnzL,nnzL/m,layer,parallelism,pre_time,solve_time,gflops,bandwith,"
                              "warp_occupy,element_occupy")
    ret, m, nn, nnzA, rpA, ciA, vA = oracle.read_mtx(str(mtx))
    nnzL, rpL, ciL, vL = oracle.change2tran(m, nnzA, rpA, ciA, vA)
    layer, par = oracle.matrix_layer(m, rpL, ciL)
    for row in rows[1::2]:
        f = [s.strip() for s in row.split(",")]
        assert f[0] == str(mtx) and int(f[1]) == m and int(f[2]) == nnzL and int(f[4]) == layer
        assert f[3] == f"{nnzL / m:4.2f}" and f[5] == f"{par:4.2f}"
        assert float(f[7]) >= 0 and f[7] != "-1.00"  # -1 would mean the validation failed
    # the mix row's occupancies: the chunk threshold follows rows per level
    # (4 from 200 rows per level, else 32)
    _, _, wo, eo = oracle.matrix_warp4(m, nnzL, rpL, 4 if m >= 200 * layer else 32, 64)
    f = [c.strip() for c in rows[3].split(",")]
    assert f[10] == f"{wo:4.2f}" and f[11] == f"{eo:4.2f}", (f[10:], wo, eo)
    # result.jsonl: one object per run; x against the FP64 serial solve of the
    # same L and b (FP64 kernels: order kept, fused terms, so within 1e-12;
    # the FP32 run within the FP32 bar), roofline figures from the wrapper
    import json
    side = [json.loads(line) for line in (tmp_path / "result.jsonl").read_text().splitlines()]
    assert [s["algo"] for s in side] == ["wrtfst", "mix", "2part", "csc"]
    for s, row in zip(side, rows[1::2]):
        f = [c.strip() for c in row.split(",")]
        assert s["device"] == "gpu" and s["passed"] and (s["m"], s["nnzL"], s["layer"]) == (m, nnzL, layer)
        assert f"{s['solve_time_ms']:4.2f}" == f[7] and f"{s['bandwidth_gbs']:4.2f}" == f[9]
        assert s["x_from"] == ("timed solve" if s["algo"] == "csc" else "separate solve, same options")
        bar = 1e-5 if s["value_type"] == "float" else 1e-12
        assert s["rel_err_vs_fp64_serial"] <= bar, s
        rf = s["roofline"]
        assert rf["peak_gbs"] == 8000 and abs(rf["frac"] - s["bandwidth_gbs"] / 8000) <= 1e-12
        assert abs(rf["bytes"] - s["bandwidth_gbs"] * 1e6 * s["solve_time_ms"]) <= 1e-6 * rf["bytes"]
        assert abs(s["us_per_level"] - 1e3 * s["solve_time_ms"] / layer) <= 1e-9 * s["us_per_level"]


def test_driver_autotuned_and_deep_matrix_solvers(cuda, tmp_path):
    """-a auto / local / partition / level (this library's additions to the driver):
    the same result.csv columns as wrtfst, none failing the validation, on a
    diagonally dominant band of 3000 rows (bandwidth 100, values kept)."""
    mtx = tmp_path / "band.mtx"
    rng = np.random.default_rng(2)
    n = 3000
    lines = ["%%MatrixMarket matrix coordinate real general", "% generated", f"{n} {n} {4 * n}"]
    for i in range(n):
        for c in rng.integers(max(0, i - 100), max(1, i), size=3):
            lines.append(f"{i + 1} {int(c) + 1} {rng.uniform(-1, 1):.6f}")
        lines.append(f"{i + 1} {i + 1} 4.0")
    mtx.write_text("\n".join(lines) + "\n")
    for algo in ("auto", "local", "partition", "level"):
        r = subprocess.run([os.path.join(ROOT, "capellinisptrsv_amd", "sptrsv_main"), str(mtx), "-a", algo, "--values"],
                           cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (algo, r.stdout, r.stderr)
    rows = (tmp_path / "result.csv").read_text().splitlines()
    assert len(rows) == 8
    for row in rows[1::2]:
        f = [s.strip() for s in row.split(",")]
        assert f[0] == str(mtx) and int(f[1]) == n and float(f[7]) > 0, row


def test_sweep_over_a_matrix_directory(cuda, tmp_path):
    """The run.sh-style sweep: every valid fixture under two algorithms, one
    result.csv row (after its header line) per run, none failing validation."""
    from capellinisptrsv_amd import sweep
    src = os.path.join(ROOT, "tests", "golden", "mtx")
    mats = tmp_path / "matrices"
    mats.mkdir()
    good = [f for f in sorted(os.listdir(src)) if f.endswith(".mtx") and not f.startswith("err_")]
    for f in good:
        (mats / f).write_bytes(open(os.path.join(src, f), "rb").read())
    out = tmp_path / "out"
    assert sweep.main([str(mats), "-a", "wrtfst,mix", "--out-dir", str(out)]) == 0
    rows = (out / "result.csv").read_text().splitlines()
    data = rows[1::2]
    assert len(data) == 2 * len(good)
    for row in data:
        f = [s.strip() for s in row.split(",")]
        assert os.path.basename(f[0]) in good and f[7] != "-1.00"
