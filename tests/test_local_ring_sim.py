"""CPU model of the one-workgroup ring protocol (k_local, ring mode).

Waves of `lanes` lanes run the kernel's loop in lockstep; waves interleave
in a random order at the kernel's phase boundaries (every poll of an
iteration, then every publish -- values before tags -- then the wave's
frontier update).  Small rings force wrap-around and throttling.  Every run
must finish, and with the oracle's bits: a publish that overwrote a slot a
reader still needed, a reader accepting a stale slot, or a throttle that
blocks the lowest unfinished row would show up here as a wrong value or a
deadlock.
"""
import numpy as np
import pytest

import oracle
from conftest import random_lower

WIN = 4
SENT = None


class Lane:
    def __init__(self, rl, rows_if, lo, hi, rp, ci, v):
        self.row, self.rows_if, self.hi = lo + rl, rows_if, hi
        self.rp, self.ci, self.v = rp, ci, v
        self.done = self.row >= hi
        self.front = lo
        if not self.done:
            self.start()

    def start(self):
        self.j, self.d, self.s = self.rp[self.row], self.rp[self.row + 1] - 1, 0.0
        self.x = [SENT] * WIN

    def window(self, u):
        return self.j + u if self.j + u < self.d else None


def run(rp, ci, v, b, R, lanes=8, waves=4, seed=0, max_steps=200_000, throttle=True):
    n = len(rp) - 1
    rows_if = lanes * waves
    bw = max(int(i - c) for i in range(n) for c in ci[rp[i]:rp[i + 1]])
    assert bw + 2 * rows_if <= R, "ring mode needs bandwidth + 2 rows in flight <= R"
    rng = np.random.default_rng(seed)
    tags, vals = [-1] * R, [0.0] * R
    wmin = [0] * waves
    out = np.full(n, np.nan)
    W = [[Lane(w * lanes + l, rows_if, 0, n, rp, ci, v) for l in range(lanes)] for w in range(waves)]
    pending = {}  # wave -> list of publishes of its current iteration
    phase = [0] * waves
    for _ in range(max_steps):
        live = [w for w in range(waves) if any(not ln.done for ln in W[w]) or phase[w]]
        if not live:
            return out
        w = int(rng.choice(live))
        if phase[w] == 0:  # polls of every lane (tag, then value)
            for ln in W[w]:
                if ln.done:
                    continue
                for u in range(WIN):
                    jj = ln.window(u)
                    if jj is not None and ln.x[u] is SENT:
                        c, sl = int(ci[jj]), int(ci[jj]) % R
                        if tags[sl] == c:
                            ln.x[u] = vals[sl]
            phase[w] = 1
        elif phase[w] == 1:  # consume, throttle, publish (values, then tags)
            pubs = []
            for ln in W[w]:
                if ln.done:
                    continue
                used = 0
                while used < WIN and ln.window(used) is not None and ln.x[used] is not SENT:
                    jj = ln.window(used)
                    ln.s = ln.s + v[jj] * ln.x[used]  # unfused: compared with the oracle's unfused order
                    used += 1
                ln.j += used
                ln.x = ln.x[used:] + [SENT] * used
                if ln.j >= ln.d:
                    if ln.row - R + bw >= ln.front:
                        ln.front = min(wmin)
                    if not throttle or ln.row - R + bw < ln.front:
                        pubs.append((ln, (b[ln.row] - ln.s) / v[ln.d]))
            for ln, xi in pubs:
                vals[ln.row % R] = xi
            for ln, xi in pubs:
                tags[ln.row % R] = ln.row
                out[ln.row] = xi
                ln.row += rows_if
                if ln.row >= n:
                    ln.done = True
                else:
                    ln.start()
            phase[w] = 2 if pubs else 0
        else:  # frontier update after an advance
            wmin[w] = min(n if ln.done else ln.row for ln in W[w])
            phase[w] = 0
    return None


def _ref(rp, ci, v, b):
    return oracle.csr_forward(rp, ci, v, b, fused=False)


@pytest.mark.parametrize("lanes,waves", [(4, 4), (1, 16)])  # (1, 16): one row per wavefront, as k_local runs
@pytest.mark.parametrize("seed", range(6))
def test_ring_protocol_finishes_with_the_oracle_bits(seed, lanes, waves):
    n, bw = 400, [12, 30, 5, 40, 20, 8][seed]
    rp, ci, v = random_lower(n, max_off=[2, 4, 1, 6, 3, 2][seed], band=bw, seed=seed)
    b = np.random.default_rng(seed).uniform(-1, 1, n)
    R = 1 << int(np.ceil(np.log2(bw + 2 * lanes * waves + 1)))  # smallest ring allowed
    x = run(rp, ci, v, b, R, lanes=lanes, waves=waves, seed=seed)
    assert x is not None, "deadlock"
    assert np.array_equal(x, _ref(rp, ci, v, b))


@pytest.mark.parametrize("lanes,waves", [(4, 4), (1, 16)])
def test_ring_protocol_chain_wraps_many_times(lanes, waves):
    n = 600  # bidiagonal chain: every row depends on the previous one
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    v = np.r_[1.0, np.tile([0.5, 2.0], n - 1)]
    b = np.linspace(-1, 1, n)
    x = run(rp, ci, v, b, R=64, lanes=lanes, waves=waves, seed=1)
    assert x is not None
    assert np.array_equal(x, _ref(rp, ci, v, b))


def test_model_detects_a_missing_throttle():
    """Planted mistake: without the throttle, lanes whose rows have no
    dependencies run laps ahead and overwrite ring slots that slower readers
    still need -- the model must see a wrong value or a deadlock."""
    n, bw = 600, 20
    rp, ci, v = random_lower(n, max_off=1, band=bw, seed=7)
    b = np.random.default_rng(7).uniform(-1, 1, n)
    ref = _ref(rp, ci, v, b)
    bad = 0
    for seed in range(8):
        x = run(rp, ci, v, b, R=64, lanes=4, waves=4, seed=seed, throttle=False, max_steps=100_000)
        bad += x is None or not np.array_equal(x, ref)
        assert np.array_equal(run(rp, ci, v, b, R=64, lanes=4, waves=4, seed=seed), ref)
    assert bad > 0
