"""The committed oracle digests (tests/golden/oracle_parity.json) are what
the oracle computes today, for every pinned run; tests/golden/make_oracle_golden.py
rewrites them deliberately."""
import json

from .golden.make_oracle_golden import OUT, compute


def test_oracle_matches_committed_digests():
    with open(OUT) as f:
        want = json.load(f)
    got = compute()
    assert set(got) == set(want)
    bad = [k for k in want if got[k] != want[k]]
    assert not bad, f"{len(bad)} pinned oracle runs changed, first {bad[:3]}"
