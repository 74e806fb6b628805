"""CPU lockstep simulation of the write-first schedule (forward progress, no GPU).

Models the kernel's execution: workgroups take row tickets in the order they
start; the 64 lanes of a wavefront execute the polling loop in lockstep
(every load of an iteration happens before its stores); wavefronts of the
device interleave in a random order; at most `resident` wavefronts run at
once.  Two loop shapes are simulated:

  uniform  -- the shipped shape: a lane publishes inside the loop, the loop
              exits when every lane of the wavefront has published;
  sunk     -- the shape the compiler produced from a per-lane loop exit: the
              publishing store executes only after every lane has left the
              loop.

The uniform shape must always finish with the oracle's values; the sunk
shape must deadlock as soon as a row depends on another row of its own
wavefront (the bug the library's kernels are structured against).
"""
import numpy as np
import pytest

import oracle
from conftest import random_lower

SENT = object()


class Wave:
    def __init__(self, rows, rp, ci, v, b, sunk):
        self.rows = rows
        self.rp, self.ci, self.v, self.b = rp, ci, v, b
        self.j = {i: rp[i] for i in rows}
        self.s = {i: 0.0 for i in rows}
        self.done = {i: False for i in rows}
        self.pending = {}  # sunk shape: values computed but not yet stored
        self.sunk = sunk

    def step(self, x):
        """One lockstep iteration; returns True if any lane made progress."""
        reads = {}
        for i in self.rows:  # loads first (the whole wavefront)
            if not self.done[i] and self.j[i] < self.rp[i + 1] - 1:
                reads[i] = x[self.ci[self.j[i]]]
        progress = False
        stores = {}
        for i in self.rows:
            if self.done[i]:
                continue
            d = self.rp[i + 1] - 1
            if i in reads and reads[i] is not SENT:
                p = self.v[self.j[i]] * reads[i]
                self.s[i] = self.s[i] + p
                self.j[i] += 1
                progress = True
            if self.j[i] >= d:
                xi = (self.b[i] - self.s[i]) / self.v[d]
                self.done[i] = True
                progress = True
                if self.sunk:
                    self.pending[i] = xi
                else:
                    stores[i] = xi
        if self.sunk and all(self.done.values()):
            stores.update(self.pending)
            self.pending = {}
        for i, val in stores.items():
            x[i] = val
        return progress

    def finished(self):
        return all(self.done.values()) and not self.pending


def simulate(rp, ci, v, b, sunk=False, resident=6, rows_per_ticket=64, seed=0, max_idle=20000):
    rng = np.random.default_rng(seed)
    n = len(rp) - 1
    x = [SENT] * n
    next_row = 0
    running = []
    idle = 0
    while next_row < n or running:
        while len(running) < resident and next_row < n:  # a resident slot takes the next ticket
            rows = list(range(next_row, min(n, next_row + rows_per_ticket)))
            running.append(Wave(rows, rp, ci, v, b, sunk))
            next_row += rows_per_ticket
        w = running[int(rng.integers(len(running)))]
        idle = 0 if w.step(x) else idle + 1
        if w.finished():
            running.remove(w)
        if idle > max_idle:
            return None  # no wavefront can progress: deadlock
    return np.array(x, dtype=np.float64)


@pytest.mark.parametrize("seed", range(4))
def test_uniform_exit_always_completes(seed):
    rp, ci, v = random_lower(700, max_off=4, band=90, seed=seed)
    b = np.random.default_rng(seed).uniform(-1, 1, size=700)
    x = simulate(rp, ci, v, b, sunk=False, resident=[1, 2, 5, 11][seed], seed=seed)
    assert x is not None
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=False))


def test_uniform_exit_chain_single_resident():
    # a bidiagonal chain: every row depends on the previous lane of its wavefront
    n = 300
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    v = np.r_[1.0, np.tile([0.5, 2.0], n - 1)]
    b = np.linspace(-1, 1, n)
    x = simulate(rp, ci, v, b, resident=1)
    assert x is not None and np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=False))


def test_sunk_store_deadlocks_on_intra_wavefront_dependency():
    n = 300
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    v = np.r_[1.0, np.tile([0.5, 2.0], n - 1)]
    b = np.linspace(-1, 1, n)
    assert simulate(rp, ci, v, b, sunk=True, resident=4, max_idle=2000) is None


def test_sunk_store_survives_without_intra_wavefront_dependencies():
    # level structure with every dependency >= one ticket back: no deadlock even sunk
    rp, ci, v = random_lower(640, max_off=3, band=64, seed=3)
    keep_rp, keep_ci, keep_v = [0], [], []
    for i in range(640):
        cols = [int(c) for c in ci[rp[i]:rp[i + 1] - 1] if c < (i // 64) * 64]
        vals = [float(vv) for c, vv in zip(ci[rp[i]:rp[i + 1] - 1], v[rp[i]:rp[i + 1] - 1]) if c < (i // 64) * 64]
        keep_ci += cols + [i]
        keep_v += vals + [1.0 + sum(abs(t) for t in vals)]
        keep_rp.append(len(keep_ci))
    rp2, ci2, v2 = np.array(keep_rp, np.int32), np.array(keep_ci, np.int32), np.array(keep_v)
    b = np.ones(640)
    x = simulate(rp2, ci2, v2, b, sunk=True, resident=3)
    assert x is not None and np.array_equal(x, oracle.csr_forward(rp2, ci2, v2, b, fused=False))
