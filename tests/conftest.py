import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def cuda():
    """GPU tests fail (not skip) when no GPU is visible: they are meant to run with -m gpu on a machine with one."""
    import torch
    assert torch.cuda.is_available(), "GPU test collected but no GPU is visible"
    import capellinisptrsv_amd as cap
    cap.lib()  # the HIP library must be the thing under test; no fallback
    return torch.device("cuda:0")


def rel_maxnorm(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.max(np.abs(ref)) if ref.size else 1.0
    return float(np.max(np.abs(x - ref)) / (den if den > 0 else 1.0)) if ref.size else 0.0


def random_lower(n, max_off=4, band=50, seed=0, dtype=np.float64, dup=False, shuffle=False):
    """Small random lower-triangular CSR (diagonal last, diagonally dominant)."""
    rng = np.random.default_rng(seed)
    rp = [0]
    ci, v = [], []
    for i in range(n):
        lo = max(0, i - band)
        k = min(int(rng.integers(0, max_off + 1)), i - lo)
        cols = sorted(rng.choice(np.arange(lo, i), size=k, replace=False).tolist()) if k else []
        if dup and cols and rng.random() < 0.3:
            cols.insert(int(rng.integers(0, len(cols))), cols[0])
            cols.sort()
        if shuffle:
            rng.shuffle(cols)
        vals = rng.uniform(-1, 1, size=len(cols))
        ci += cols + [i]
        v += vals.tolist() + [1.0 + float(np.abs(vals).sum())]
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v, dtype)


def upper_from_lower(rp, ci, v):
    """U[i][j] = L[n-1-i][n-1-j]: an upper CSR with the diagonal first and
    ascending columns (rows of L reversed, entries reversed)."""
    n = len(rp) - 1
    urp, uci, uv = [0], [], []
    for i in range(n):
        r = n - 1 - i
        cols = [n - 1 - int(c) for c in ci[rp[r]:rp[r + 1]]][::-1]
        vals = [float(x) for x in v[rp[r]:rp[r + 1]]][::-1]
        uci += cols
        uv += vals
        urp.append(len(uci))
    return np.array(urp, np.int32), np.array(uci, np.int32), np.array(uv)


def drop_diagonal(rp, ci, v, upper=False):
    n = len(rp) - 1
    keep = np.ones(len(ci), bool)
    for i in range(n):
        keep[rp[i] if upper else rp[i + 1] - 1] = False
    counts = np.diff(rp) - 1
    return np.r_[0, np.cumsum(counts)].astype(np.int32), ci[keep].astype(np.int32), v[keep]
