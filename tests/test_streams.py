"""Request streams of the serving form and the admission estimate (CPU only:
nothing here touches a GPU)."""
import numpy as np


def test_request_stream_is_seeded_and_in_range():
    from vpfhf_amd import serve
    a = [serve.spec_request(i, seed=2) for i in range(40)]
    b = [serve.spec_request(i, seed=2) for i in range(40)]
    assert [s["n"] for s, _ in a] == [s["n"] for s, _ in b]
    assert all(24 <= s["n"] <= 96 and 1 <= len(s["cracks"]) <= 3 and 1 <= am <= 20 for s, am in a)
    assert all(5e-4 <= s["pressure"][1] <= 5e-3 for s, _ in a)
    t = serve.arrival_times(2000, 100.0, seed=2)
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / 2000 - 0.01) < 0.001  # mean gap 1/rate


def test_request_bytes_estimate():
    from vpfhf_amd.host import load
    L = load()
    nn, ne = 97 ** 3, 96 ** 3
    vq = 7 * 7 * 96 * 27 * 256 * 8  # tile-blocked V coupling table: 7 x 7 tiles of 16 x 16 slots
    assert L.vfh_request_bytes(97, 97, 97) == 600 * nn + 8 * ne + vq + (5 << 20)
    assert L.vfh_request_bytes(25, 25, 25) < L.vfh_request_bytes(26, 25, 25)
