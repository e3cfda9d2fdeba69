"""Request stream, admission, batching, quarantine and cancellation of the
serving loop, on the CPU (host-side fake slots stand in for the GPU ones)."""
import math

import numpy as np

from xlumina_amd import serving as S


def test_request_mix_statistics():
    reqs = S.make_requests(20000, rate=1000.0, seed=7)
    n = np.array([r.n for r in reqs])
    for size, p in ((512, 0.5), (1024, 0.3), (2048, 0.2)):
        assert abs((n == size).mean() - p) < 0.015
    assert abs(np.mean([r.vectorial for r in reqs]) - 0.3) < 0.015
    assert abs(np.mean([r.grad for r in reqs]) - 0.2) < 0.015
    z = np.array([r.z for r in reqs])
    assert z.min() >= 5e3 and z.max() <= 1e5
    t = np.array([r.arrival for r in reqs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert abs(np.diff(t).mean() - 1e-3) < 5e-5          # Poisson rate 1000 / s


def test_requests_are_seeded():
    a = S.make_requests(50, rate=10.0, seed=3)
    b = S.make_requests(50, rate=10.0, seed=3)
    assert a == b
    assert all(r.arrival == 0.0 for r in S.make_requests(5, None))


def test_bucket_key():
    r = S.Request(1, 1024, True, 1e4, False, 5)
    assert S.bucket_of(r) == S.Bucket(1024, True, False)


# ---- admission, batching, quarantine and cancel with host-side fake slots (no GPU) ----

class FakeSlot:
    """A slot with the GPU slot's interface: a batch finishes after `polls`
    completion checks; a member is non-finite when its distance is NaN."""

    def __init__(self, server, b, capacity, polls=2):
        self.bucket, self.capacity = b, capacity
        self.bytes = server._size(b, capacity)
        self.requests, self.last_used, self.graphs = [], 0.0, {}
        self.polls, self._left = polls, 0
        self.launched = []
        server.created.append(self)

    def launch(self, reqs, keep):
        self.requests = list(reqs)
        self.launched.append([r.rid for r in reqs])
        self._left = self.polls
        return [(r.rid, None) for r in reqs] if keep else None

    def done(self):
        self._left -= 1
        return self._left <= 0

    def finite(self):
        return [not math.isnan(r.z) for r in self.requests]


def fake_server(budget, size=lambda b, cap: cap * b.n * b.n, **kw):
    srv = S.Server(hbm_budget=budget, slot_factory=FakeSlot, slot_size=size, **kw)
    srv.created = []
    return srv


def test_same_bucket_arrivals_are_batched_up_to_capacity():
    srv = fake_server(1 << 40, capacity={512: 4}, slots_per_bucket=1)
    reqs = [S.Request(i, 512, False, 1e4, False, i) for i in range(10)]
    rep = srv.run(reqs)
    assert rep.counts() == {"ok": 10, "nonfinite": 0, "rejected": 0, "cancelled": 0}
    assert [len(b) for b in srv.created[0].launched] == [4, 4, 2]
    assert rep.batches == 3 and rep.completed == 10


def test_bucket_larger_than_budget_is_rejected_not_raised():
    # a 2048^2 slot needs 4 * 2048^2 "bytes" here; the budget holds only 512^2 slots
    srv = fake_server(4 * 512 * 512 + 10, capacity={512: 4, 2048: 1}, slots_per_bucket=1)
    reqs = [S.Request(0, 2048, False, 1e4, False, 0), S.Request(1, 512, False, 1e4, False, 1),
            S.Request(2, 2048, True, 1e4, True, 2)]
    rep = srv.run(reqs)
    assert rep.status == {0: "rejected", 1: "ok", 2: "rejected"}
    assert all(s.bucket.n == 512 for s in srv.created)
    assert srv.used_bytes <= srv.hbm_budget


def test_capped_budget_evicts_idle_slots_and_serves_every_bucket():
    # room for one slot at a time: buckets take turns, idle slots are released
    size = lambda b, cap: 100
    srv = fake_server(150, size=size, capacity={512: 2, 1024: 2})
    reqs = [S.Request(i, (512, 1024)[i % 2], bool(i % 3 == 0), 1e4, False, i) for i in range(12)]
    rep = srv.run(reqs)
    assert rep.counts()["ok"] == 12
    assert srv.evictions >= 2
    assert srv.used_bytes <= srv.hbm_budget


def test_nonfinite_member_is_quarantined_and_mates_complete():
    srv = fake_server(1 << 40, capacity={512: 4}, slots_per_bucket=1)
    reqs = [S.Request(i, 512, False, float("nan") if i == 2 else 1e4, False, i) for i in range(4)]
    rep = srv.run(reqs, keep_results=True)
    assert [len(b) for b in srv.created[0].launched] == [4]          # one batch holds all four
    assert rep.status == {0: "ok", 1: "ok", 2: "nonfinite", 3: "ok"}
    assert sorted(srv.results) == [0, 1, 3]                           # the poisoned result is dropped


def test_cancel_queued_and_mid_batch():
    srv = fake_server(1 << 40, capacity={512: 2}, slots_per_bucket=1)
    reqs = [S.Request(i, 512, False, 1e4, False, i) for i in range(5)]
    cancelled = []

    def on_launch(batch):
        if not cancelled:
            srv.cancel(batch[1].rid)      # on the GPU already: its batch finishes, its result is dropped
            srv.cancel(4)                 # still queued: leaves the queue without running
            cancelled.append(True)

    rep = srv.run(reqs, keep_results=True, on_launch=on_launch)
    assert rep.status == {0: "ok", 1: "cancelled", 2: "ok", 3: "ok", 4: "cancelled"}
    ran = [rid for b in srv.created[0].launched for rid in b]
    assert 4 not in ran and 1 in ran
    assert sorted(srv.results) == [0, 2, 3]


def test_waiting_bucket_is_not_starved_by_a_busy_one():
    # budget fits one slot; bucket A keeps arriving, bucket B's single request must still run
    size = lambda b, cap: 100
    srv = fake_server(150, size=size, capacity={512: 1, 1024: 1}, slots_per_bucket=1)
    reqs = [S.Request(i, 512, False, 1e4, False, i, arrival=0.0) for i in range(6)]
    reqs.append(S.Request(99, 1024, False, 1e4, False, 99, arrival=0.0))
    rep = srv.run(reqs)
    assert rep.status[99] == "ok" and rep.counts()["ok"] == 7
