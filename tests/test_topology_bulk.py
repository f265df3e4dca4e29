"""The bulk (numpy) topology check gives exactly the per-record canonical
records, and declines every input the per-record path would reject, so that
path raises its own message (python/beads/_topology.py:53-129, 203-224)."""

from __future__ import annotations

import numpy as np
import pytest

from beads_amd._topology import _canonical, _canonical_bulk, make_topology


def _random_records(rng, n, m, width, n_types):
    out = []
    for _ in range(m):
        tags = rng.choice(np.arange(1, n + 1), size=width - 1, replace=False)
        out.append(tuple(int(v) for v in tags) + (int(rng.integers(1, n_types + 1)),))
    # dense types: make sure every type id occurs
    for t in range(1, n_types + 1):
        r = out[t - 1]
        out[t - 1] = r[:-1] + (t,)
    # drop duplicate canonical keys so the set is valid
    seen, valid = set(), []
    for r in out:
        tags = r[:-1]
        key = (min(tags), max(tags)) if width == 3 else (
            tags if tags[0] < tags[-1] else tuple(reversed(tags)))
        if key not in seen:
            seen.add(key)
            valid.append(r)
    return valid


@pytest.mark.parametrize("width", [3, 4, 5])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bulk_equals_per_record(width, seed):
    rng = np.random.default_rng(seed)
    n = 300
    recs = _random_records(rng, n, 400, width, 3)
    known = frozenset(range(1, n + 1))
    slow = _canonical("x" * 5 + "s", recs, width, known)
    bulk = _canonical_bulk(np.asarray(recs), width, n)
    assert bulk is not None
    assert tuple(map(tuple, bulk.tolist())) == slow


BAD = {
    "duplicate": [(1, 2, 1), (2, 1, 1)],
    "repeated tag": [(1, 1, 1)],
    "unknown tag": [(1, 400, 1)],
    "zero tag": [(0, 2, 1)],
    "type gap": [(1, 2, 1), (2, 3, 3)],
    "zero type": [(1, 2, 0)],
}


@pytest.mark.parametrize("case", list(BAD))
def test_bulk_declines_what_per_record_rejects(case):
    recs = BAD[case]
    n = 300
    assert _canonical_bulk(np.asarray(recs), 3, n) is None
    tags = np.arange(1, n + 1)
    with pytest.raises(ValueError) as bulk_err:
        make_topology(bonds=np.asarray(recs), tags=tags)
    with pytest.raises(ValueError) as slow_err:
        _canonical("bonds", recs, 3, frozenset(range(1, n + 1)))
    assert str(bulk_err.value) == str(slow_err.value)


def test_non_integer_arrays_take_the_per_record_path():
    tags = np.arange(1, 11)
    with pytest.raises(TypeError, match="must be an integer"):
        make_topology(bonds=np.array([[1.0, 2.0, 1.0]]), tags=tags)
    with pytest.raises(TypeError, match="must be an integer"):
        make_topology(bonds=np.array([[True, True, True]]), tags=tags)


def test_unsupported_angle_message_and_order():
    tags = np.arange(1, 11)
    with pytest.raises(ValueError, match="angles require supporting topology bonds"):
        make_topology(bonds=[(1, 2, 1)], angles=np.array([[1, 2, 3, 1]]), tags=tags)
    # the angle support check runs before the dihedral records are read
    with pytest.raises(ValueError, match="angles require supporting topology bonds"):
        make_topology(bonds=[(1, 2, 1)], angles=[(1, 2, 3, 1)], dihedrals=[(1, 1, 1, 1, 1)],
                      tags=tags)


def test_large_type_id_is_declined_without_allocating():
    tags = np.arange(1, 11)
    assert _canonical_bulk(np.array([[1, 2, 2**31 - 1]]), 3, 10) is None
    with pytest.raises(ValueError, match="dense one-based"):
        make_topology(bonds=np.array([[1, 2, 2**31 - 1]]), tags=tags)
