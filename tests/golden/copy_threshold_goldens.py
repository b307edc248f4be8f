"""Copies the reference's threshold golden vectors (thresholds and
surviving-pixel counts of four seeded frames per method,
tests/golden/threshold_{adaptive,percentile,fixed}.json of the reference,
written by its tests/test_char_threshold.py) into this directory as data.
Run from the repository root with the reference checked out at
/root/reference; the copies are committed, nothing reads the reference at
test time."""
import json
import os
import sys

SRC = "/root/reference/tests/golden"
DST = os.path.dirname(os.path.abspath(__file__))

for name in ("threshold_adaptive", "threshold_percentile", "threshold_fixed"):
    with open(os.path.join(SRC, name + ".json")) as f:
        data = json.load(f)
    with open(os.path.join(DST, "ref_" + name + ".json"), "w") as f:
        json.dump(data, f, indent=2, sort_keys=True)
        f.write("\n")
    print("wrote", "ref_" + name + ".json", file=sys.stderr)
