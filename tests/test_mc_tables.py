"""Marching-cubes tables (CPU): the oracle's own tables (oracle/kf_mc_table.h) and libkf's case
strings (killingfusion_amd/csrc/kf_mc_cases.h) against digests of the reference's edgeTable and
triTable (KillingFusion/include/MarchingCubes.h:26-59, 61-318; tests/golden/mc_tables.json,
written by tests/golden/make_mc_digest.py), case by case, triangle by triangle and vertex by
vertex.  The triangle soup's order and orientation come from these tables, so a swapped or
flipped triangle in either one must fail here."""
import hashlib
import json
import os
import re

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "mc_tables.json")))
CASES_H = os.path.join(ROOT, "killingfusion_amd", "csrc", "kf_mc_cases.h")


def row_digest(row):
    return hashlib.sha256(",".join(str(int(v)) for v in row).encode()).hexdigest()


def table_digest(rows):
    return hashlib.sha256("\n".join(",".join(str(int(v)) for v in r) for r in rows).encode()).hexdigest()


def oracle_rows():
    return [O.mc_case(ci) for ci in range(256)]


def libkf_rows():
    """libkf's KF_MC_CASES (three hex edge ids per triangle) as -1-padded rows of 16."""
    text = open(CASES_H).read()
    body = re.search(r"KF_MC_CASES\[256\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    cases = re.findall(r'"([0-9a-f]*)"', body)
    assert len(cases) == 256
    return [[int(c, 16) for c in s] + [-1] * (16 - len(s)) for s in cases]


def libkf_edge_corners():
    text = open(CASES_H).read()
    body = re.search(r"KF_MC_EDGE_CORNERS\[12\]\[2\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    pairs = [tuple(int(v) for v in p.split(",")) for p in re.findall(r"\{\s*(\d+\s*,\s*\d+)\s*\}", body)]
    assert len(pairs) == 12
    return pairs


def mismatched_cases(rows):
    """Cases whose row digest differs from the reference's."""
    return [ci for ci, r in enumerate(rows) if row_digest(r)[:16] != GOLD["tri_case_sha256_16"][ci]]


def test_oracle_tri_table_equals_reference():
    rows = oracle_rows()
    assert mismatched_cases(rows) == []
    assert table_digest(rows) == GOLD["tri_table_sha256"]
    assert [sum(1 for v in r if v >= 0) // 3 for r in rows] == GOLD["tri_counts"]


def test_oracle_edge_table_equals_reference():
    assert table_digest([[O.mc_edge_mask(ci)] for ci in range(256)]) == GOLD["edge_table_sha256"]


def test_libkf_cases_equal_reference_and_oracle():
    rows = libkf_rows()
    assert mismatched_cases(rows) == []
    assert table_digest(rows) == GOLD["tri_table_sha256"]
    assert rows == oracle_rows()


def test_libkf_edge_masks_equal_reference():
    # libkf derives each case's cut-edge mask from its corner signs (kf_mc_init_tables); it
    # must reproduce edgeTable exactly, with the edges numbered as Polygonise's VertexInterp
    # calls (MarchingCubes.h:386-409).
    corners = libkf_edge_corners()
    assert corners == [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    masks = [sum(1 << e for e, (a, b) in enumerate(corners) if ((ci >> a) & 1) != ((ci >> b) & 1)) for ci in range(256)]
    assert table_digest([[m] for m in masks]) == GOLD["edge_table_sha256"]


@pytest.mark.parametrize("mistake", ["swap_triangles", "flip_triangle", "rotate_vertices", "wrong_edge"])
def test_planted_table_mistake_is_caught(mistake):
    # A mistake planted in a copy of the oracle's table (on the CPU) must fail the comparison
    # the tests above make: the digests see triangle order, orientation and vertex order.
    rows = [list(r) for r in oracle_rows()]
    ci = next(c for c, r in enumerate(rows) if sum(1 for v in r if v >= 0) >= 6)
    r = rows[ci]
    if mistake == "swap_triangles":
        r[0:3], r[3:6] = r[3:6], r[0:3]
    elif mistake == "flip_triangle":
        r[1], r[2] = r[2], r[1]
    elif mistake == "rotate_vertices":
        r[0], r[1], r[2] = r[1], r[2], r[0]
    else:
        r[0] = (r[0] + 1) % 12
    assert r != oracle_rows()[ci]
    assert mismatched_cases(rows) == [ci]
