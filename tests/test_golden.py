"""The oracle against the committed fixed-scene vectors (tests/golden/sphere16.npz, written by
tests/golden/make_golden.py): any change of the oracle's arithmetic shows up here."""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_oracle_reproduces_golden_vectors():
    gold = np.load(os.path.join(HERE, "golden", "sphere16.npz"))
    now = load_maker().compute()
    assert set(now) == set(gold.files)
    for k in gold.files:
        assert np.array_equal(now[k], gold[k]), k
    assert (gold["live_w"] > 0).sum() > 200 and gold["mc_live"].shape[0] > 100
