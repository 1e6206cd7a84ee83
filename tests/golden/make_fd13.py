"""Writes fd13_coeffs.json: the centred rows of the reference's 13-point
Bickley derivative tables (M=13, K=1 and K=2), read as text from
src/mates.f:786-879 (BLOCK DATA BLKDS) of the reference checkout.

newder(ICON=1) returns row 7 of the 13x13 matrix as table entries 79..85
(c_-6..c_0) and mirrors the rest with sign (-1)^K (newder.f:99-118).
Run in a container where the reference is mounted; the JSON is committed.
"""
import json
import os
import re
import sys

REF = os.environ.get("EFFPOT_REFERENCE", "/root/reference")


def block(lines, k):
    start = next(i for i, l in enumerate(lines) if re.search(r"M=13,K=\s*%d\s*$" % k, l))
    vals = []
    for l in lines[start + 1:]:
        if re.search(r"M=\d+,K=", l):
            break
        vals += [float(v.replace("D", "E")) for v in re.findall(r"[-+]?\d\.\d+D[-+]\d+", l)]
    return vals


def main(out):
    with open(os.path.join(REF, "src", "mates.f")) as f:
        lines = f.readlines()
    res = {"source": "mates.f BLOCK DATA BLKDS, M=13 (K=1 at mates.f:786-832, K=2 at :833-879)"}
    for k in (1, 2):
        v = block(lines, k)
        assert len(v) == 85, len(v)
        half = v[78:85]                      # c_-6 .. c_0
        sig = -1.0 if k == 1 else 1.0
        row = half + [sig * x for x in half[-2::-1]]
        res["K%d" % k] = row
        res["K%d_table_entries" % k] = v
    with open(out, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "fd13_coeffs.json"))
