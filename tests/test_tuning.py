"""The tuning table (launch shapes and thresholds that never change a
result): rmetis_amd/tuning.json, written by the autotuner, is what the
library loads on first use, and out-of-range entries are refused.  Host
code only."""
import ctypes as C
import json
import os

import pytest

from rmetis_amd import _lib

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(HERE, "rmetis_amd", "tuning.json")
NT = 26      # 24 level buckets (4 sizes x 6 mean degrees), then two thresholds


def _get(L):
    v = (C.c_int32 * NT)()
    assert L.rmetis_hip_get_tuning(v, NT) == NT
    return list(v)


def test_committed_table_is_the_loaded_one():
    with open(TABLE) as fh:
        t = json.load(fh)
    want = t["suitor_shape_by_level_bucket"] + [t["incremental_min_n"], t["spec_min_per_part"]]
    L = _lib.lib()
    assert _get(L) == want
    assert os.path.samefile(L.rmetis_hip_tuning_source().decode(), TABLE)


@pytest.mark.parametrize("pos, value", [
    (5, 4),            # shape 4 does not exist
    (17, -1),          # negative shape
    (24, 0),           # thresholds must be positive
    (25, 0),
])
def test_out_of_range_entries_are_refused(pos, value):
    L = _lib.lib()
    before = _get(L)
    bad = list(before)
    bad[pos] = value
    assert L.rmetis_hip_set_tuning((C.c_int32 * NT)(*bad), NT) == -2
    assert _get(L) == before


def test_set_and_restore():
    L = _lib.lib()
    before = _get(L)
    new = [i % 4 for i in range(24)] + [1 << 17, 8]
    assert L.rmetis_hip_set_tuning((C.c_int32 * NT)(*new), NT) == 1
    assert _get(L) == new
    assert L.rmetis_hip_set_tuning((C.c_int32 * NT)(*before), NT) == 1
    assert _get(L) == before
