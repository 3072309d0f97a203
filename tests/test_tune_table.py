"""The autotuned defaults table (avx-512-sort_amd/s512_tune.py writes it):
every workload kind is covered for every size with no gap or overlap, every
value is one the library accepts, and the compiled-in table matches the JSON
record of the tuning run."""
import json
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "avx-512-sort_amd"))
from s512_tune import scatter_columns  # noqa: E402
INC = os.path.join(ROOT, "avx-512-sort_amd", "csrc", "s512_tune_table.inc")
JSON = os.path.join(ROOT, "avx-512-sort_amd", "tune_table.json")


def _rows():
    rows = []
    for line in open(INC):
        m = re.match(r"\s*\{([-\d,\s]+)\},", line)
        if m:
            rows.append([int(v) for v in m.group(1).split(",")])
    return rows


def test_table_covers_every_kind_and_size():
    rows = _rows()
    for kind in range(5):
        ranges = sorted((r[1], r[2]) for r in rows if r[0] == kind)
        assert ranges, f"kind {kind} missing"
        assert ranges[0][0] == 0 and ranges[-1][1] == 64, ranges
        for (a0, a1), (b0, b1) in zip(ranges, ranges[1:]):
            assert a1 == b0 and a0 < a1, ranges


def test_table_values_valid():
    for kind, lo, hi, cnt, sc, ids, tile, leaf, logf in _rows():
        assert 0 <= cnt < 5 and 0 <= sc < 6 and ids in (0, 1) and 0 <= tile < 3, (kind, lo)
        assert sc != 5 or ids == 1, (kind, lo)  # the full-line scatter reads stored ids
        assert 64 <= leaf <= 1 << 16 and 1 <= logf <= 9, (kind, lo)


def test_table_matches_tuning_record():
    doc = json.load(open(JSON))
    kinds = {"i32": 0, "f64": 1, "kv": 2, "seg_i32": 3, "seg_f64": 4}
    rows = _rows()
    assert len(rows) == len(doc["classes"])
    for r, c in zip(rows, doc["classes"]):
        o = c["options"]
        assert r[:3] == [kinds[c["kind"]], c["lg_lo"], c["lg_hi"]]
        assert r[3] == o["count_variant"] - 1 and r[6] == o["tile_variant"] - 1
        assert (r[4], r[5]) == scatter_columns(o["scatter_variant"])
        assert r[7] == o["target_leaf"] and r[8] == o["max_fanout_log2"]
