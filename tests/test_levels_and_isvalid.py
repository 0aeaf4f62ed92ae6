"""CPU pins for two functions that no golden fixture covers.

* pk_sample_levels (the library's level sampler, include/util.hpp:13-64):
  a fresh uniform per level against cum[l] = 1 - M^-(l+1) gives
  P(level >= l) = M^(-l(l+1)/2), not the textbook M^-l.
* orc_layer_invalid_count (the oracle's IsValid, hnsw_layer.hpp:278-295):
  hand-built rows whose violations are known.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi


def _closed_form(M, l):
    return float(M) ** (-l * (l + 1) / 2)


@pytest.mark.parametrize("M,n", [(4, 400_000), (8, 400_000), (32, 2_000_000)])
def test_library_sampler_closed_form(M, n):
    lv = abi.sample_levels(n, M, seed=11)
    assert lv.dtype == np.uint8 and lv.shape == (n,)
    for l in (1, 2, 3):
        expect = _closed_form(M, l)
        if expect * n < 5:
            # too rare to see in n draws: at most a handful may appear
            assert np.sum(lv >= l) <= max(5, 10 * expect * n), (M, l)
            continue
        frac = float(np.mean(lv >= l))
        sd = np.sqrt(expect * (1 - expect) / n)
        assert abs(frac - expect) < 5 * sd, (M, l, frac, expect)
    # the textbook sampler gives P(level >= 2) = M^-2, far above M^-3
    if M <= 8:
        assert np.mean(lv >= 2) < 0.5 * float(M) ** -2


def test_library_sampler_chi_square():
    """Level histogram for M=4 against P(level = l) = M^(-l(l+1)/2) - M^(-(l+1)(l+2)/2)."""
    M, n = 4, 1_000_000
    lv = abi.sample_levels(n, M, seed=5)
    obs = np.array([np.sum(lv == 0), np.sum(lv == 1), np.sum(lv == 2), np.sum(lv >= 3)], np.float64)
    p = np.array([_closed_form(M, l) - _closed_form(M, l + 1) for l in range(3)] + [_closed_form(M, 3)])
    exp = p * n
    chi2 = float(((obs - exp) ** 2 / exp).sum())
    # 3 degrees of freedom: P(chi2 > 16.27) = 0.001
    assert chi2 < 16.27, (obs, exp, chi2)


def test_library_sampler_seeding():
    a = abi.sample_levels(100_000, 4, seed=1)
    b = abi.sample_levels(100_000, 4, seed=1)
    c = abi.sample_levels(100_000, 4, seed=2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert abi.sample_levels(0, 4, seed=1).shape == (0,)


def test_library_sampler_rejects_bad_arguments():
    with pytest.raises(abi.PickleError):
        abi.sample_levels(10, 1, seed=0)


def _layer_with_rows(rows):
    """A one-layer oracle graph whose layer 0 holds `rows`: a list of
    (heur, adj, dist) with len(adj) = degree <= D = 4."""
    n = len(rows)
    base = np.zeros((n, 4), np.uint8)
    g = O.Graph(base, O.L2, 2, 8, np.zeros(n, np.uint8))
    D = 4
    adj = np.full((n + 1, D), -1, np.int32)
    dist = np.full((n + 1, D), np.finfo(np.float32).max, np.float32)
    deg = np.zeros(n, np.int16)
    heur = np.zeros(n, np.int16)
    for v, (h, a, d) in enumerate(rows):
        adj[v, :len(a)] = a
        dist[v, :len(d)] = d
        deg[v] = len(a)
        heur[v] = h
    layer = {"cap": n, "D": D, "next_id": n, "adj": adj, "dist": dist, "deg": deg, "heur": heur,
             "ext": np.arange(n, dtype=np.int32), "data": np.zeros((n, 4), np.uint8),
             "has_data": np.ones(n, np.uint8)}
    g.set_layer(0, layer)
    return g


# (heur, adj, dist, valid) -- the reference's IsValid asserts, case by case
ISVALID_ROWS = [
    (2, [1, 2, 3, 4], [1.0, 2.0, 1.0, 3.0], True),    # each segment sorted; the pruned part may restart lower
    (2, [0, 0], [1.0, 1.0], False),                   # repeated id inside the heuristic part
    (3, [4, 5, 6], [1.0, 3.0, 2.0], False),           # heuristic distances decrease
    (1, [4, 5, -1], [1.0, 2.0, 3.0], False),          # empty id inside the pruned part
    (1, [-1, 1, 2], [0.0, 1.0, 2.0], True),           # slot 0 of a segment is never checked
    (2, [1, 2, 2, 3], [1.0, 2.0, 2.0, 3.0], True),    # a repeat across the segment boundary is not checked
    (3, [1, 2], [1.0, 2.0], False),                   # heuristic count above the degree
    (0, [], [], True),                                # empty row
    (2, [3, 1], [5.0, 5.0], True),                    # equal distances are allowed
    (0, [6, 7, 6, 7], [1.0, 2.0, 3.0, 4.0], True),    # only adjacent repeats are checked
    (0, [6, 7, 7], [1.0, 2.0, 2.0], False),           # adjacent repeat in the pruned part
    (1, [6, 7, 5], [1.0, 3.0, 2.0], False),           # pruned distances decrease
]


def test_oracle_isvalid_known_rows():
    g = _layer_with_rows([(h, a, d) for h, a, d, _ in ISVALID_ROWS])
    assert g.invalid_count(0) == sum(1 for *_, ok in ISVALID_ROWS if not ok)


@pytest.mark.parametrize("i", range(len(ISVALID_ROWS)))
def test_oracle_isvalid_each_row(i):
    h, a, d, ok = ISVALID_ROWS[i]
    g = _layer_with_rows([(h, a, d)])
    assert g.invalid_count(0) == (0 if ok else 1), ISVALID_ROWS[i]


def test_oracle_isvalid_on_a_single_threaded_build():
    """A single-threaded build never writes an invalid row while rows stay
    within D (capped here by M large enough that no row overflows)."""
    rng = np.random.default_rng(3)
    base = rng.integers(0, 256, (600, 16)).astype(np.uint8)
    g = O.Graph(base, O.L2, 8, 32, O.sample_levels(600, 8, seed=2))
    g.build(np.arange(600, dtype=np.int32))
    for l in range(g.nlayers):
        L = g.layer(l)
        if (L["deg"] <= L["D"]).all():
            assert g.invalid_count(l) == 0
