"""Exhaustive ground truth on the matrix cores (pk_brute_force) against the
oracle's FP64 brute force.

u8 / i8: the int8 MFMA tiles give exact int32 distances, so the 32 returned
ids and distances equal the oracle's bit for bit, ties ranked by the lower
row id.  f16 / f32: the FP64 MFMA tiles give distances within FP64 rounding
of the oracle's direct sums; ids must match and distances agree to FP32
rounding.

Shapes cover ragged dims (not a multiple of the 32- or 64-byte K block),
row counts that are not a multiple of the tile, fewer rows than the 32
returned (the rest reads -1 / inf), duplicate rows (distance ties), and
enough rows that a query's results come from the largest number of split
lists (64).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pickle_amd import abi

pytestmark = pytest.mark.gpu

K = abi.BRUTE_K


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def data(dt, n, nq, dim, seed, dup=False):
    rng = np.random.default_rng(seed)
    if np.dtype(dt).kind in "ui":
        info = np.iinfo(dt)
        base = rng.integers(info.min, info.max + 1, (n, dim)).astype(dt)
        q = rng.integers(info.min, info.max + 1, (nq, dim)).astype(dt)
    else:
        base = rng.normal(size=(n, dim)).astype(dt)
        q = rng.normal(size=(nq, dim)).astype(dt)
    if dup:  # every row twice (rows 2i and 2i+1): every distance is a tie
        base[1::2] = base[0::2][: n // 2]
        q[: nq // 2] = base[: nq // 2]  # some queries sit on a base row
    return base, q


def check(dt, metric, n, nq, dim, seed, dup=False):
    say("brute force", np.dtype(dt).name, "L2" if metric == abi.L2 else "IP", f"n={n} nq={nq} dim={dim} dup={dup}")
    base, q = data(dt, n, nq, dim, seed, dup)
    idx = abi.Index(dt, metric, dim, 8, 10, n)
    idx.set_base(base)
    ids, ds = idx.brute_force(q)
    idx.close()
    kk = min(K, n)
    gt_ids, gt_d = O.brute_force(base, q, kk, O.L2 if metric == abi.L2 else O.IP)
    if np.dtype(dt).kind in "ui":
        np.testing.assert_array_equal(ids[:, :kk], gt_ids)
        np.testing.assert_array_equal(ds[:, :kk], gt_d.astype(np.float32))
    else:
        np.testing.assert_array_equal(ids[:, :kk], gt_ids)
        np.testing.assert_allclose(ds[:, :kk], gt_d.astype(np.float32), rtol=2e-6, atol=1e-6)
    if kk < K:
        assert (ids[:, kk:] == -1).all() and np.isinf(ds[:, kk:]).all()


@pytest.mark.parametrize("dt", [np.uint8, np.int8, np.float32, np.float16])
@pytest.mark.parametrize("metric", [abi.L2, abi.IP])
def test_brute_force_shapes(dt, metric):
    check(dt, metric, 3000, 70, 100, 11)      # ragged dim, ragged tile, partial query block
    check(dt, metric, 1000, 131, 128, 12)     # exact K blocks
    check(dt, metric, 45, 5, 37, 13)          # one split, odd dim
    check(dt, metric, 180, 5, 37, 15)         # six splits: fewer tiles than a multiple of the 8 XCDs
    check(dt, metric, 20, 3, 16, 14)          # fewer rows than returned


@pytest.mark.parametrize("dt", [np.uint8, np.int8])
def test_brute_force_ties_rank_lower_row_first(dt):
    check(dt, abi.L2, 2000, 100, 64, 21, dup=True)
    check(dt, abi.IP, 2000, 100, 64, 22, dup=True)


def test_brute_force_many_split_lists():
    # 3 query blocks: rows split 32 ways (u8) / 16 ways (f32), 64 lists per query
    check(np.uint8, abi.L2, 50_000, 300, 128, 31)
    check(np.float32, abi.L2, 20_000, 130, 128, 32)
    check(np.float16, abi.IP, 20_000, 130, 96, 33)


def test_brute_force_many_ties_across_splits():
    """600 copies of one row, spread over a 150,000-row base, and queries
    equal to it: every copy is at distance 0 (L2).  The base is split 32
    ways, so all 64 lists of a query hold some of the copies, and the merge
    must return the 32 lowest-numbered ones."""
    rng = np.random.default_rng(44)
    n, dim = 150_000, 128
    base = rng.integers(0, 256, (n, dim)).astype(np.uint8)
    where = np.sort(rng.choice(n, 600, replace=False))
    base[where] = base[where[0]]
    q = np.repeat(base[where[:1]], 70, axis=0)
    q[35:] = rng.integers(0, 256, (35, dim)).astype(np.uint8)
    say("brute force many ties: 600 copies, 70 queries")
    idx = abi.Index(np.uint8, abi.L2, dim, 8, 10, n)
    idx.set_base(base)
    ids, ds = idx.brute_force(q)
    idx.close()
    np.testing.assert_array_equal(ids[:35], np.broadcast_to(where[:K], (35, K)))
    assert (ds[:35] == 0).all()
    gt_ids, gt_d = O.brute_force(base, q, K, O.L2)
    np.testing.assert_array_equal(ids, gt_ids)
    np.testing.assert_array_equal(ds, gt_d.astype(np.float32))
