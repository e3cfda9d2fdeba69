"""Request stream and admission arithmetic of the serving loop (CPU only)."""
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
