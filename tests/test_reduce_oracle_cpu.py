"""Pins of the oracle's PROX contact reduction (prox_collision_detection.h:
340-345 over prox::ContactPoint, prox_contact_point.h:91-190), independent of
its code: hand-worked known answers at the precision boundary and for the
last-kept rule of std::unique (tests/prox_cases.py), and, on contacts whose
floats lie on a coarse grid (where equal within the precision means equal),
numpy's stable lexsort by the reference's key order followed by the first of
each run of equal rows.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from prox_cases import DTYPE, cases


def _reduce(a):
    return O.prox_reduce(a["body_i"], a["body_j"], a["p"], a["n"], a["depth"])


@pytest.mark.parametrize("name,recs,kept", cases(), ids=[c[0] for c in cases()])
def test_known_answers(name, recs, kept):
    assert _reduce(recs).tolist() == kept


def test_empty():
    assert _reduce(np.zeros(0, DTYPE)).tolist() == []


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_contacts_against_lexsort_unique(seed):
    rng = np.random.default_rng(seed)
    n = 20000
    a = np.zeros(n, DTYPE)
    a["body_i"] = rng.integers(0, 3, n)
    a["body_j"] = rng.integers(0, 3, n)
    g = lambda shape: (rng.integers(-1, 2, shape) / 8.0).astype(np.float32)  # 3 values, 1/8 apart
    a["p"] = g((n, 3))
    a["n"] = g((n, 3))
    a["depth"] = g(n)
    # reference key order: body i, body j, p.y, p.x, p.z, n.y, n.x, n.z, depth (lexsort: last key primary)
    cols = [a["body_i"], a["body_j"], a["p"][:, 1], a["p"][:, 0], a["p"][:, 2], a["n"][:, 1], a["n"][:, 0], a["n"][:, 2],
            a["depth"]]
    order = np.lexsort(cols[::-1])
    rows = np.stack([c.astype(np.float64) for c in cols], axis=1)[order]
    first = np.ones(n, bool)
    first[1:] = np.any(rows[1:] != rows[:-1], axis=1)
    want = order[first]
    got = _reduce(a)
    assert n // 4 < got.shape[0] < n - n // 8  # many duplicates, many distinct rows
    assert got.tolist() == want.tolist()
