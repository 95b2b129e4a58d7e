"""The float16-with-level-scale storage restatement (CPU): pins for
oracle.f16s_round, which the GPU parity tests feed to the oracle."""
import struct

import numpy as np

from oracle.oracle import f16s_round, f16s_scales


def _half_rne(x: float) -> float:
    """IEEE binary16, round to nearest even, by the struct module's 'e'
    format (independent of numpy's conversion)."""
    return struct.unpack("<e", struct.pack("<e", x))[0]


def test_scales_are_powers_of_two_bracketing_the_row_maximum():
    rng = np.random.default_rng(1)
    rows = [rng.normal(0, 10.0 ** rng.uniform(-9, 3), 500).astype(np.float32) for _ in range(40)]
    rows.append(np.zeros(7, np.float32))
    rows.append(np.array([np.nan, 2.0, -np.inf], np.float32))
    s = f16s_scales(rows)
    for r, sc in zip(rows, s):
        m, e = np.frexp(sc)
        assert m == 0.5                               # a power of two
        a = np.abs(r[np.isfinite(r)])
        if a.size and a.max() > 0:
            assert 2.0 ** 14 <= a.max() / sc < 2.0 ** 15
        else:
            assert sc == 1.0


def test_rounding_matches_independent_half_conversion():
    rng = np.random.default_rng(2)
    K, N = 5, 300
    u = (rng.normal(0, 0.3, (K - 1, N)) * np.geomspace(1, 1e-3, K - 1)[:, None]).astype(np.float32)
    v = (rng.normal(0, 0.2, (K - 1, N))).astype(np.float32)
    w = (rng.normal(0, 1e-5, (K, N))).astype(np.float32)
    w[0] = 0.0
    uq, vq, wq = f16s_round(u, v, w)
    suv = np.maximum(f16s_scales(list(u)), f16s_scales(list(v)))
    sw = f16s_scales(list(w))
    for L in range(K - 1):
        for j in range(0, N, 37):
            for a, aq in ((u, uq), (v, vq)):
                want = _half_rne(float(a[L, j]) / float(suv[L])) * float(suv[L])
                assert float(aq[L, j]) == want
    for L in range(K):
        for j in range(0, N, 41):
            want = _half_rne(float(w[L, j]) / float(sw[L])) * float(sw[L])
            assert float(wq[L, j]) == want
    assert np.all(wq[0] == 0.0)


def test_relative_error_bounded_by_half_ulp_of_the_row_scale():
    """|value| / scale < 2^15 and half values carry 11 significant bits, so
    the spacing of halves below 2^15 is at most 2^5 and a rounded value is
    off by at most scale * 2^4; values in the top binade keep a relative
    error of at most 2^-11."""
    rng = np.random.default_rng(3)
    u = rng.normal(0, 1.0, (3, 1000)).astype(np.float32)
    v = rng.normal(0, 1.0, (3, 1000)).astype(np.float32)
    uq, vq = f16s_round(u, v)
    s = np.maximum(f16s_scales(list(u)), f16s_scales(list(v)))
    err = np.abs(uq.astype(np.float64) - u)
    assert np.all(err <= s[:, None] * 2.0 ** 4)
    big = np.abs(u) >= s[:, None] * 2.0 ** 14
    assert np.all(err[big] <= np.abs(u[big]) * 2.0 ** -11)
