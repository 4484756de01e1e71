"""Alias tables of the env-map sampler: the host construction (no GPU). The
distribution a table draws from follows from the table alone (entry k is kept
with probability prob[k] and receives 1 - prob[m] from every entry m aliased to
it, each slot chosen with probability 1/n), and must equal the weights."""
import numpy as np
import pytest


def _table(tt, w):
    from trenchanttracer_amd import _native as N
    w = np.ascontiguousarray(w, np.float64)
    out = np.empty(2 * len(w), np.float32)
    assert N.lib.tt_debug_alias_table(w.ctypes.data, len(w), out.ctypes.data) == 0
    return out[0::2].astype(np.float64), out[1::2].astype(np.int64)


def _drawn(prob, alias):
    p = prob.copy()
    np.add.at(p, alias, 1.0 - prob)
    return p / len(prob)


def _cases():
    rng = np.random.default_rng(5)
    sky = rng.gamma(1.0, 1.0, 2048)
    sky[700] = 1e6                      # one texel holding almost all the weight (a sun)
    sparse = np.zeros(1000)
    sparse[[3, 500, 999]] = [1.0, 2.0, 3.0]
    return {
        "gamma_4096": rng.gamma(0.5, 1.0, 4096),
        "sun_row": sky,
        "zeros_between": np.where(rng.random(777) < 0.3, 0.0, rng.random(777)),
        "sparse": sparse,
        "uniform": np.ones(64),
        "single": np.array([2.5]),
        "two": np.array([1.0, 3.0]),
        "all_zero": np.zeros(16),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_alias_table_draws_the_weights(tt, name):
    w = _cases()[name]
    n = len(w)
    prob, alias = _table(tt, w)
    assert ((prob >= 0) & (prob <= 1)).all()
    assert ((alias >= 0) & (alias < n)).all()
    expected = w / w.sum() if w.sum() > 0 else np.full(n, 1.0 / n)
    drawn = _drawn(prob, alias)
    # float probabilities: each entry off by at most a few float ulps, over n
    assert np.allclose(drawn, expected, rtol=1e-6, atol=4e-7 / n)
    # zero weight is never drawn: kept with probability 0 and nobody's alias
    assert (drawn[w == 0] == 0).all() if w.sum() > 0 else True


def test_alias_table_rejects_bad_input(tt):
    from trenchanttracer_amd import _native as N
    out = np.empty(4, np.float32)
    w = np.array([1.0, -1.0])
    assert N.lib.tt_debug_alias_table(w.ctypes.data, 2, out.ctypes.data) != 0
    w = np.array([1.0, np.nan])
    assert N.lib.tt_debug_alias_table(w.ctypes.data, 2, out.ctypes.data) != 0
    assert N.lib.tt_debug_alias_table(w.ctypes.data, 0, out.ctypes.data) != 0
