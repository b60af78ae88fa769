"""Writes tests/golden/mc_tables.json: SHA-256 digests of the reference's marching-cubes
tables, edgeTable[256] (KillingFusion/include/MarchingCubes.h:26-59) and triTable[256][16]
(MarchingCubes.h:61-318), a digest per triTable case (its first 16 hex digits) and one per table.

    python tests/golden/make_mc_digest.py [--reference /root/reference]

Run where the reference checkout exists; the digests are committed so that the CPU tests
pin the oracle's table (oracle/kf_mc_table.h) and libkf's case strings
(killingfusion_amd/csrc/kf_mc_cases.h) to the reference's without reading it.  A case row
is digested as its 16 integers in decimal, comma-separated (-1 padding included); an edge
mask as its decimal value.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "mc_tables.json")
REL = os.path.join("KillingFusion", "include", "MarchingCubes.h")


def row_digest(values) -> str:
    return hashlib.sha256(",".join(str(int(v)) for v in values).encode()).hexdigest()


def table_digest(rows) -> str:
    return hashlib.sha256("\n".join(",".join(str(int(v)) for v in r) for r in rows).encode()).hexdigest()


def parse_tables(text: str):
    """edgeTable as 256 ints, triTable as 256 rows of 16 ints, from the header's initialisers."""
    m = re.search(r"int\s+edgeTable\s*\[\s*256\s*\]\s*=\s*\{(.*?)\};", text, re.S)
    edge = [int(t, 0) for t in re.findall(r"0x[0-9a-fA-F]+|\d+", m.group(1))]
    m = re.search(r"int\s+triTable\s*\[\s*256\s*\]\s*\[\s*16\s*\]\s*=\s*\{(.*)\};", text, re.S)
    rows = re.findall(r"\{([^{}]*)\}", m.group(1))
    tri = [[int(t) for t in r.replace(" ", "").split(",") if t] for r in rows]
    assert len(edge) == 256 and len(tri) == 256 and all(len(r) == 16 for r in tri)
    return edge, tri


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args(argv)
    path = os.path.join(a.reference, REL)
    with open(path) as fh:
        edge, tri = parse_tables(fh.read())
    out = {
        "source": "KillingFusion/include/MarchingCubes.h:26-59 (edgeTable), :61-318 (triTable)",
        "row_format": "16 decimal integers joined by ',' (triTable); one decimal integer (edgeTable)",
        "edge_table_sha256": table_digest([[e] for e in edge]),
        "tri_table_sha256": table_digest(tri),
        "tri_case_sha256_16": [row_digest(r)[:16] for r in tri],
        "tri_counts": [sum(1 for v in r if v >= 0) // 3 for r in tri],
    }
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
