"""The oracle reproduces the committed golden digests (regression pins)."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402


def test_oracle_matches_golden_digests():
    golden = json.load(open(os.path.join(HERE, "golden", "golden.json")))
    now = make_golden.cases()
    assert set(golden) == set(now)
    for k in golden:
        assert now[k] == golden[k], k
