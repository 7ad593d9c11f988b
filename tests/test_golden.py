"""The oracle reproduces the committed golden fixtures bit for bit, and the
fixtures satisfy the invariants that pin them (ascending lists, the <= rule,
self-exclusion)."""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = ["plummer_n64_seed1.npz", "plummer_n1024_seed1.npz"]


@pytest.mark.parametrize("name", FILES)
def test_oracle_reproduces_golden(name):
    g = np.load(os.path.join(HERE, name))
    acc, adot, nnb, lists = oracle.regf(g["x"], g["v"], g["m"], g["r2"], g["targets"],
                                        int(g["cap"]))
    assert np.array_equal(acc, g["acc"])
    assert np.array_equal(adot, g["adot"])
    assert np.array_equal(nnb, g["nnb"])
    assert np.array_equal(lists, g["lists"])


@pytest.mark.parametrize("name", FILES)
def test_golden_lists_follow_the_rule(name):
    g = np.load(os.path.join(HERE, name))
    x, r2 = g["x"], g["r2"]
    for k, i in enumerate(g["targets"]):
        c = g["nnb"][k]
        lst = g["lists"][k, :c]
        assert np.all(np.diff(lst) > 0) and i not in lst
        d = x - x[i]
        rr = np.sum(d * d, axis=1)
        want = np.nonzero(rr <= r2[i])[0]
        want = want[want != i]
        assert np.array_equal(lst, want)
