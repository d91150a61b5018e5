"""CPU tests of the autotuner's table plumbing (zkmetal_amd/tune.py): a sweep
result (as saved to JSON, keys as strings) turns into the 33-entry arrays the
engine takes, with 0 wherever a class was not tuned."""
import json

from zkmetal_amd import tune


def _result():
    res = {"plain": {10: 8, 22: 16}, "batch": {12: 9},
           "schedule": {"plain": {13: [0, 4], 26: [3, 0]}, "batch": {28: [4, 16]}}}
    return json.loads(json.dumps(res))  # as loaded from a saved file


def test_window_tables():
    plain, batch = tune.tables(_result())
    assert len(plain) == len(batch) == 33
    assert plain[10] == 8 and plain[22] == 16 and batch[12] == 9
    assert sum(1 for c in plain if c) == 2 and sum(1 for c in batch if c) == 1


def test_schedule_tables():
    st = tune.schedule_tables(_result())
    pt, pr = st["plain"]
    assert len(pt) == len(pr) == 33
    assert (pt[13], pr[13]) == (0, 4) and (pt[26], pr[26]) == (3, 0)
    bt, br = st["batch"]
    assert (bt[28], br[28]) == (4, 16)
    assert sum(pt) == 3 and sum(pr) == 4 and sum(bt) == 4 and sum(br) == 16


def test_schedule_tables_absent():
    st = tune.schedule_tables({"plain": {}, "batch": {}})
    assert st["plain"] == ([0] * 33, [0] * 33) and st["batch"] == ([0] * 33, [0] * 33)
