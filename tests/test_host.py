"""CPU tests of the library's host side against the oracle, plus the C-ABI surface.

No GPU calls: the shared library is loaded (it must export every symbol that
include/capellini_sptrsv.h declares) and its host-only helpers are compared
with the oracle on the committed fixtures and random inputs.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api, _native
from conftest import ROOT, random_lower

MTX = os.path.join(ROOT, "tests", "golden", "mtx")
HEADER = os.path.join(ROOT, "include", "capellini_sptrsv.h")
LIB = os.path.join(ROOT, "capellinisptrsv_amd", "libcapellini_sptrsv.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", text, flags=re.M))
    inline = set(re.findall(r"static inline \w+ (\w+)\s*\(", text))
    return sorted(n for n in names - inline if n not in {"if", "return", "sizeof"} and not n.startswith("CAP_"))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # nothing beyond the C ABI leaks out
    extra = [e for e in exported if not e.startswith(("cap_", "spts_syncfree_hip_", "sptrsv_"))]
    assert not extra, extra
    # and the ctypes binding covers the whole surface
    assert set(names) <= set(_native.SIGNATURES), set(names) - set(_native.SIGNATURES)
    _native.lib()  # loads, binds every signature


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#define VALUE_TYPE float\n#include "capellini_sptrsv.h"\n'
                   "int main(void){ double t,g,b; return spts_syncfree_hip_csr_wrt(0,0,0,1,2,0,&t,&g,&b) == -1 ? 0 : 1; }\n")
    for cc in (["gcc", "-std=c99"], ["g++", "-x", "c++"]):
        subprocess.run([*cc, "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT}/include", str(src)], check=True)


@pytest.mark.parametrize("name", sorted(os.listdir(MTX)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_read_mtx_and_preprocessing_match_oracle(name, dtype):
    path = os.path.join(MTX, name)
    a = api.read_mtx(path, dtype=dtype)
    o = oracle.read_mtx(path, dtype=dtype)
    assert a[0] == o[0], (a[0], o[0])
    if a[0] != 0:
        return
    assert a[1:4] == o[1:4]
    for x, y in zip(a[4:], o[4:]):
        assert np.array_equal(x, y)
    _, m, n, nnz, rp, ci, v = a
    mm = min(m, n)
    La = api.change2tran(mm, nnz, rp[:mm + 1], ci, v)
    Lo = oracle.change2tran(mm, nnz, rp[:mm + 1], ci, v)
    assert La[0] == Lo[0] and all(np.array_equal(x, y) for x, y in zip(La[1:], Lo[1:]))
    nnzL, rpL, ciL, vL = La
    for x, y in zip(api.get_x_b(mm, mm, rpL, ciL, vL), oracle.get_x_b(mm, mm, rpL, ciL, vL)):
        assert np.array_equal(x, y)
    lay = api.matrix_layer(mm, mm, nnzL, rpL, ciL, vL, return_levels=True)
    assert lay[:2] == oracle.matrix_layer(mm, rpL, ciL)
    assert np.array_equal(lay[2], oracle.row_levels(mm, rpL, ciL))
    for x, y in zip(api.matrix_transposition(mm, mm, nnzL, rpL, ciL, vL),
                    oracle.transpose(mm, mm, nnzL, rpL, ciL, vL)):
        assert np.array_equal(x, y)
    for chunk in (32, 64):
        for border in (1, 2, 3):
            x = api.matrix_warp4(mm, mm, nnzL, rpL, border=border, chunk=chunk)
            y = oracle.matrix_warp4(mm, nnzL, rpL, border, chunk)
            assert x[0] == y[0] and np.array_equal(x[1], y[1]) and x[2:] == y[2:]


def test_read_mtx_missing_file():
    assert api.read_mtx("/nonexistent/file.mtx")[0] == -1


def test_read_mtx_malformed_entry(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n4 1 1.0\n")
    assert api.read_mtx(str(p))[0] == _native.CAP_ERR_IO  # row 4 of a 3x3 matrix
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    assert api.read_mtx(str(p))[0] == _native.CAP_ERR_IO  # truncated


def test_read_mtx_oversized_header_counts(tmp_path):
    """A header whose entry count the file cannot hold is malformed (no
    allocation of that size, no int32 row-pointer overflow from mirrored
    symmetric entries); index values beyond int32 are malformed too."""
    p = tmp_path / "big.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2000000000\n2 1\n3 1\n")
    assert api.read_mtx(str(p))[0] == _native.CAP_ERR_IO
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 4294967297 1.0\n")
    assert api.read_mtx(str(p))[0] == _native.CAP_ERR_IO
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n2 1\n3 1\n")
    ret, m, n, nnz, rp, ci, v = api.read_mtx(str(p))
    assert ret == 0 and nnz == 4 and list(rp) == [0, 2, 3, 4]


@pytest.mark.parametrize("seed", range(4))
def test_preprocessing_random_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    n = [1, 63, 1000, 20000][seed]
    rows = np.repeat(np.arange(n), 6)
    cols = rng.integers(0, n, size=6 * n)
    order = np.argsort(rows, kind="stable")
    rp = np.searchsorted(rows[order], np.arange(n + 1)).astype(np.int32)
    ci = cols[order].astype(np.int32)
    v = rng.uniform(-1, 1, size=len(ci))
    La, Lo = api.change2tran(n, len(ci), rp, ci, v), oracle.change2tran(n, len(ci), rp, ci, v)
    assert La[0] == Lo[0] and all(np.array_equal(x, y) for x, y in zip(La[1:], Lo[1:]))
    nnzL, rpL, ciL, vL = La
    assert api.matrix_layer(n, n, nnzL, rpL, ciL, vL) == oracle.matrix_layer(n, rpL, ciL)
    for x, y in zip(api.get_x_b(n, n, rpL, ciL, vL), oracle.get_x_b(n, n, rpL, ciL, vL)):
        assert np.array_equal(x, y)


def test_change2tran_keep_values_mode():
    rp, ci, v = random_lower(300, max_off=4, seed=3)
    n = 300
    nnzL, rpL, ciL, vL = api.change2tran(n, len(ci), rp, ci, v, keep_values=True)
    assert np.array_equal(rpL, rp) and np.array_equal(ciL, ci)
    for i in range(n):
        a, d = rpL[i], rpL[i + 1] - 1
        assert np.array_equal(vL[a:d], v[a:d]) and vL[d] == 1.0 + np.abs(v[a:d]).sum()


def test_validate_codes():
    rp, ci, v = random_lower(200, seed=1)
    n = 200
    assert api.validate_csr(n, len(ci), rp, ci, v) == (0, -1)
    r = next(i for i in range(50, n) if rp[i + 1] - rp[i] >= 2)
    c = ci.copy(); c[rp[r]] = r + 1
    assert api.validate_csr(n, len(ci), rp, c, v) == (_native.CAP_ERR_UPPER, r)
    c = ci.copy(); c[rp[r + 1] - 1] = r - 1
    assert api.validate_csr(n, len(ci), rp, c, v) == (_native.CAP_ERR_NO_DIAG, r)
    w = v.copy(); w[rp[r + 1] - 1] = 0.0
    assert api.validate_csr(n, len(ci), rp, ci, w) == (_native.CAP_ERR_ZERO_DIAG, r)
    p = rp.copy(); p[r + 1] = p[r] - 1
    assert api.validate_csr(n, len(ci), p, ci, v)[0] == _native.CAP_ERR_ROWPTR
    p = rp.copy(); p[r + 1] = p[r]
    assert api.validate_csr(n, len(ci), p, ci, v) == (_native.CAP_ERR_NO_DIAG, r)
    assert api.validate_csr(n, len(ci) + 1, rp, ci, v)[0] == _native.CAP_ERR_ROWPTR


def test_algorithm_wrapper_refuses_bad_algorithms():
    """The CSR wrapper with a chosen algorithm returns CAP_ERR_ARG for mix
    (which has its own wrapper) and for codes outside CAP_ALGO_*, before any
    device work; the Python mirror raises for names it does not take."""
    rp, ci, v = random_lower(200, max_off=3, band=20, seed=3)
    n = len(rp) - 1
    for algo in (_native.ALGO["mix"], -1, max(_native.ALGO.values()) + 1):
        ret = api._bench("spts_syncfree_hip_csr_algo", rp, ci, v, n, n, len(ci), mix=False, algo=algo)[0]
        assert ret == _native.CAP_ERR_ARG, algo
    with pytest.raises(ValueError):
        api.spts_syncfree_hip_csr_algo(rp, ci, v, n, n, len(ci), algo="mix")


def test_datasize_formula():
    assert api.datasize(10, 10, 37, 8) == oracle.datasize(10, 10, 37, 8)
    assert api.datasize(10, 10, 37, 4, 5) == oracle.datasize(10, 10, 37, 4, 5)


# ------------------------------------------------------------- generators
def _check_layout(rp, ci, v):
    n = len(rp) - 1
    assert rp[0] == 0 and np.all(np.diff(rp) >= 1)
    assert api.validate_csr(n, len(ci), rp, ci, v)[0] == 0
    for i in range(0, n, max(1, n // 500)):
        off = ci[rp[i]:rp[i + 1] - 1]
        assert np.all(np.diff(off) > 0)  # strictly ascending, no duplicates
        assert v[rp[i + 1] - 1] == pytest.approx(1.0 + np.abs(v[rp[i]:rp[i + 1] - 1]).sum(), rel=1e-15)


def test_generator_band_exact_size():
    rp, ci, v = api.make_config(0)
    assert len(rp) - 1 == 100_000 and len(ci) == 399_994  # 3n - 6 off-diagonals + n
    _check_layout(rp, ci, v)
    for i in (5, 999, 1000, 50_000, 99_999):
        off = ci[rp[i]:rp[i + 1] - 1]
        assert len(off) == min(3, i) and off.min() >= max(0, i - 1000)
    assert np.all(np.abs(v[np.r_[rp[1:]] - 1]) >= 1.0)


def test_generator_level_structure():
    rp, ci, v = api.generate("level", n=60_000, levels=6, width=10_000, depth=3, offdiag=2, seed=4)
    _check_layout(rp, ci, v)
    lvl_of = np.arange(60_000) // 10_000
    cnt = np.diff(rp) - 1
    assert np.all(cnt[:10_000] == 0) and np.all((cnt[10_000:] >= 1) & (cnt[10_000:] <= 2))
    rows = np.repeat(np.arange(60_000), np.diff(rp))
    off = ci != rows
    d = lvl_of[rows[off]] - lvl_of[ci[off]]
    assert d.min() >= 1 and d.max() <= 3
    assert api.matrix_layer(60_000, 60_000, len(ci), rp, ci, v)[0] <= 6


def test_generator_bernoulli_and_rowcount():
    rp, ci, v = api.generate("bernoulli", n=5000, band=128, p=0.3, seed=1)
    _check_layout(rp, ci, v)
    mean_off = (len(ci) - 5000) / 5000
    assert 36.0 < mean_off < 39.5  # 0.3 * 128 minus the short leading rows
    rp, ci, v = api.generate("band_rowcount", n=5000, offdiag=16, band=64, seed=2)
    _check_layout(rp, ci, v)
    assert np.diff(rp).max() <= 16 and np.diff(rp).min() >= 1


def test_generator_is_deterministic_and_seeded():
    a = api.make_config(0, seed=3, scale=0.1)
    b = api.make_config(0, seed=3, scale=0.1)
    c = api.make_config(0, seed=4, scale=0.1)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1], c[1])
    f = api.make_config(0, seed=3, scale=0.1, dtype=np.float32)
    off = np.ones(len(a[1]), bool)
    off[a[0][1:] - 1] = False  # single precision: same pattern, off-diagonals rounded from the FP64 draw
    assert np.array_equal(f[0], a[0]) and np.array_equal(f[1], a[1])
    assert np.array_equal(f[2][off], a[2][off].astype(np.float32))
    u = api.uniform(5, 1000)
    assert np.array_equal(u, api.uniform(5, 1000)) and u.min() >= -1 and u.max() < 1


def test_error_strings():
    L = _native.lib()
    for code in range(0, -14, -1):
        assert L.cap_strerror(code) not in (b"", b"unknown error"), code
    assert L.cap_version().startswith(b"capellini")


def test_plan_create_refuses_bad_grid_options():
    # argument checks run before any device call
    from ctypes import byref, c_void_p
    rp, ci, v = random_lower(50, seed=3)
    h = c_void_p()
    for bpc, thr in ((5, 0), (8, 0), (-1, 0), (1, 64), (1, 512), (1, 192)):
        want = _native.CAP_ERR_ARG
        o = _native.default_opts(host_inputs=1, blocks_per_cu=bpc, row_threads=thr)
        rc = _native.lib().cap_plan_create(byref(h), 50, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                           _native.CAP_F64, byref(o), None)
        assert rc == want, (bpc, rc)
        assert not h.value


def test_plan_create_checks_rhs_per_lane():
    """rhs_per_lane takes 0, 1, 2, 4 or 8 (8 in FP32 only), with the
    write-first algorithm in the value protocol and neither the LDS window
    nor the poll counter; batches refuse it -- all before any device call."""
    from ctypes import byref, c_void_p
    rp, ci, v = random_lower(50, seed=3)
    h = c_void_p()
    cases = ((3, {}, _native.CAP_ERR_ARG), (-1, {}, _native.CAP_ERR_ARG), (16, {}, _native.CAP_ERR_ARG),
             (2, dict(algo="auto"), _native.CAP_ERR_UNSUPPORTED), (2, dict(algo="mix"), _native.CAP_ERR_UNSUPPORTED),
             (2, dict(sync="flag"), _native.CAP_ERR_UNSUPPORTED), (2, dict(lds_window=1), _native.CAP_ERR_UNSUPPORTED),
             (2, dict(count_polls=1), _native.CAP_ERR_UNSUPPORTED), (8, {}, _native.CAP_ERR_UNSUPPORTED))
    for r, kw, want in cases:
        kw = dict(dict(algo="wrtfst", sync="value"), **kw)
        o = _native.default_opts(host_inputs=1, rhs_per_lane=r, **kw)
        rc = _native.lib().cap_plan_create(byref(h), 50, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                           _native.CAP_F64, byref(o), None)
        assert rc == want, (r, kw, rc)
        assert not h.value
    ro = np.array([0, 50], np.int32)
    o = _native.default_opts(host_inputs=1, rhs_per_lane=2)
    rc = _native.lib().cap_batch_create(byref(h), 1, ro.ctypes.data, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                        _native.CAP_F64, 4, byref(o), None)
    assert rc == _native.CAP_ERR_UNSUPPORTED and not h.value



def test_plan_create_checks_row_tiles():
    """row_tiles takes 0 or 1, with the write-first or mix algorithm in the
    value protocol, without the poll counter or several right-hand sides per lane;
    batches refuse it -- all before any device call."""
    from ctypes import byref, c_void_p
    rp, ci, v = random_lower(50, seed=3)
    h = c_void_p()
    cases = ((2, {}, _native.CAP_ERR_ARG), (-1, {}, _native.CAP_ERR_ARG),
             (1, dict(algo="auto"), _native.CAP_ERR_UNSUPPORTED), (1, dict(algo="twopart"), _native.CAP_ERR_UNSUPPORTED),
             (1, dict(algo="local"), _native.CAP_ERR_UNSUPPORTED),
             (1, dict(sync="flag"), _native.CAP_ERR_UNSUPPORTED), (1, dict(count_polls=1), _native.CAP_ERR_UNSUPPORTED),
             (1, dict(rhs_per_lane=2), _native.CAP_ERR_UNSUPPORTED))
    for t, kw, want in cases:
        kw = dict(dict(algo="wrtfst", sync="value"), **kw)
        o = _native.default_opts(host_inputs=1, row_tiles=t, **kw)
        rc = _native.lib().cap_plan_create(byref(h), 50, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                           _native.CAP_F64, byref(o), None)
        assert rc == want, (t, kw, rc)
        assert not h.value
    ro = np.array([0, 50], np.int32)
    o = _native.default_opts(host_inputs=1, row_tiles=1)
    rc = _native.lib().cap_batch_create(byref(h), 1, ro.ctypes.data, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                        _native.CAP_F64, 4, byref(o), None)
    assert rc == _native.CAP_ERR_UNSUPPORTED and not h.value

def test_plan_create_checks_the_arithmetic_option():
    """arith takes CAP_ARITH_FUSED or CAP_ARITH_REFERENCE; the reference
    arithmetic is refused by the algorithms whose order is not the sequential
    one (CSC, partition) -- all before any device call."""
    from ctypes import byref, c_void_p
    rp, ci, v = random_lower(50, seed=3)
    h = c_void_p()
    for arith, algo, want in ((2, "auto", _native.CAP_ERR_ARG), (-1, "auto", _native.CAP_ERR_ARG),
                              (1, "csc", _native.CAP_ERR_UNSUPPORTED), (1, "partition", _native.CAP_ERR_UNSUPPORTED)):
        o = _native.default_opts(host_inputs=1, arith=arith, algo=algo)
        rc = _native.lib().cap_plan_create(byref(h), 50, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                           _native.CAP_F64, byref(o), None)
        assert rc == want, (arith, algo, rc)
        assert not h.value


def test_sweep_lists_matrices_and_commands(tmp_path):
    """The sweep driver (run.sh equivalent): directory and list-file sources,
    lookup by name under --root, one command per (matrix, algorithm)."""
    from capellinisptrsv_amd import sweep
    d = os.path.join(ROOT, "tests", "golden", "mtx")
    mats = sweep.matrices(d)
    assert mats == sorted(mats) and len(mats) == len([f for f in os.listdir(d) if f.endswith(".mtx")])
    lst = tmp_path / "matrix_name.txt"
    lst.write_text("/nowhere/general_real/general_real.mtx\n\n# comment\n" + mats[0] + "\n")
    got = sweep.matrices(str(lst), root=d)
    assert got == [os.path.join(d, "general_real.mtx"), mats[0]]
    cmds = sweep.commands(got, ["wrtfst", "mix"], f32=True, values=True)
    assert len(cmds) == 4 and cmds[0][0].endswith("sptrsv_main_f32")
    assert cmds[1][1:] == [got[0], "-a", "mix", "--values"]
    assert sweep.main([d, "-a", "wrtfst,csc", "--dry-run"]) == 0


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), (["--gpus", "8"], {}), ([], {"WORLD_SIZE": "4"})])
def test_bench_refuses_more_than_one_gpu(argv, env):
    """bench.py is a one-GPU benchmark: N > 1 exits with a message before any
    GPU work (exit status 2)."""
    import subprocess
    import sys
    e = dict(os.environ, **env)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv], capture_output=True, text=True,
                       env=e, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_host_preprocessing_under_address_and_ub_sanitizers(tmp_path):
    """host_prep.cpp built with -fsanitize=address,undefined (aborting on the
    first report, leaks checked) and driven over every fixture, hostile
    Matrix Market headers, layout violations and every generator kind
    (tests/sanitize_host.cpp)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "sanitize_host")
    build = subprocess.run(
        [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-I" + os.path.join(root, "include"),
         os.path.join(root, "tests", "sanitize_host.cpp"),
         os.path.join(root, "capellinisptrsv_amd", "csrc", "host_prep.cpp"), "-o", exe],
        stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert build.returncode == 0, build.stdout
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    run = subprocess.run([exe, os.path.join(root, "tests", "golden", "mtx"), str(tmp_path)], env=env,
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert run.returncode == 0, run.stdout[-4000:]
    assert "0 failure(s)" in run.stdout
