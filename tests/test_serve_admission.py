"""The server's admission rule on the CPU: the solve and the basis cache are
replaced by stand-ins that record which jobs run at once, so the scheduling
loop itself (heavy slots, light jobs passing a waiting heavy job, every
request answered once) runs without a GPU."""
import threading
import time

import pytest

serve = pytest.importorskip("ed_code_amd.serve")


class _Stand:
    def __init__(self, heavy_rows):
        self.lock = threading.Lock()
        self.running = []          # (rid, heavy) of the jobs in flight
        self.max_heavy = 0
        self.max_jobs = 0
        self.starts = {}
        self.heavy_rows = heavy_rows

    def solve(self, basis, reqs, iterations):
        heavy = serve.estimate_rows(reqs[0].lattice, serve.sector_key(reqs[0])[1]) >= self.heavy_rows
        with self.lock:
            self.running.append((reqs[0].rid, heavy))
            self.max_heavy = max(self.max_heavy, sum(h for _, h in self.running))
            self.max_jobs = max(self.max_jobs, len(self.running))
            for r in reqs:
                self.starts[r.rid] = time.perf_counter()
        time.sleep(0.15 if heavy else 0.02)
        with self.lock:
            self.running.remove((reqs[0].rid, heavy))
        now = time.perf_counter()
        for r in reqs:
            r.evals, r.errors, r.done, r.batch = [0.0], [0.0], now, len(reqs)


def _req(rid, lattice, arrival, kx=0):
    return serve.Request(rid, lattice, 0.0, kx, 0, 0.3, 1.0, 0.1, 1, 1e-10, arrival)


@pytest.mark.parametrize("slots", [1, 2])
def test_heavy_slots_and_light_jobs_pass(monkeypatch, slots):
    heavy_rows = 1 << 22
    assert serve.estimate_rows("tri6x5", 15) >= heavy_rows        # 30-site torus: heavy
    assert serve.estimate_rows("chain24", 12) < heavy_rows        # 24-site chain: light
    st = _Stand(heavy_rows)
    monkeypatch.setattr(serve, "solve", st.solve)
    srv = serve.Server(workers=4, hbm_budget=1 << 50, iterations=10, heavy_rows=heavy_rows, heavy_slots=slots)
    monkeypatch.setattr(srv.cache, "get", lambda key: None)
    # four heavy jobs on distinct sectors at once, then light ones behind them
    reqs = [_req(i, "tri6x5", 0.0, kx=i) for i in range(4)]
    reqs += [_req(4 + i, "chain24", 0.01, kx=i) for i in range(4)]
    t0, wall = srv.run(reqs)
    assert all(r.done is not None for r in reqs)
    assert st.max_heavy == slots
    assert st.max_jobs <= 4
    # the light jobs start while heavy ones still wait for a slot
    last_heavy_start = max(st.starts[i] for i in range(4))
    assert all(st.starts[4 + i] < last_heavy_start for i in range(4))
    assert srv.heavy == 0 and srv.running == 0 and srv.in_use == 0
