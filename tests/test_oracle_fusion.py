"""Pins the oracle's fusion (EvaluateTipFusion, EvaluateInteriorFusion,
CleanupFusion) on scenes whose outcome is fixed by geometry.

A tip whose free end sits 1e-5 from the side of an interior segment at 30 %
of its length fuses (probability 1): the segment splits at that point into
0.3 L and 0.7 L pieces meeting the tip, the pieces take over the old bonds by
end, the neighbour beyond the split drops its bond to the old segment (and
is told which end to repair), and the tip becomes interior with bond scale
growing from 0.  Nothing happens with probability 0, when the gap exceeds
max_fusion_separation, when the two share a bonded neighbour, or when the tip's
closest point is not its free end.  Decisions happen at the requested rate.
"""
import math

import numpy as np
import pytest

# hyphal force parameters (W_A 5e-5, wall far below, no gravity, drag 200)
FPAR = [5.0e-5, 1.0e6, -1.0, 1e-4, 1e5, 0.0, 1e-2, 5e-4, 200.0]
FUS = [1.0, 6.0e-4, 1.0e-3, 35e-4]
L = 30e-4


def seg(recs, i, x, length, theta, phi, rid, pos, r=2e-4):
    recs["pos"][i] = x
    rd = recs["rdata"][i]
    rd[0], rd[1], rd[2], rd[3] = r, length, theta, phi
    rd[5] = math.pi * r * r * length
    rd[4] = 2 * math.pi * r * length + 2 * math.pi * r * r
    rd[24] = 1.0
    ip = recs["idata"][i]
    ip[5] = 1
    ip[19] = pos
    ip[33] = rid
    ip[23] = ip[24] = -1
    ip[26:32] = -1
    recs["id"][i] = rid


def bond(recs, i, j, si, sj):
    for a, b, s in ((i, j, si), (j, i, sj)):
        ip = recs["idata"][a]
        k = ip[6]
        ip[7 + k] = recs["idata"][b][33]
        ip[11 + k] = 0
        ip[15 + k] = s
        ip[6] = k + 1


def bonds(rec):
    ip = rec["idata"]
    return {(int(ip[7 + k]), int(ip[15 + k])) for k in range(ip[6])}


def scene(oracle, gap=1e-5, x_t=-6e-4, tip_len=20e-4, bonded=False, cross=False):
    """j (10) along x at the origin between a (11) and m (12); tip i (20)
    along +y ending `gap` below j at x_t, bonded at site 1 to k (21); with
    `cross` the tip passes 3e-4 above j instead (its middle is closest)."""
    X, Y = (math.pi / 2, 0.0), (math.pi / 2, math.pi / 2)
    recs = np.zeros(5, dtype=oracle.RECORD_DTYPE)
    seg(recs, 0, (0, 0, 0), L, *X, 10, 3)
    seg(recs, 1, (-L, 0, 0), L, *X, 11, 3)
    seg(recs, 2, (L, 0, 0), L, *X, 12, 3)
    yi, zi = (0.0, 3e-4) if cross else (-gap - 0.5 * tip_len, 0.0)
    seg(recs, 3, (x_t, yi, zi), tip_len, *Y, 20, 0)
    seg(recs, 4, (x_t, yi - 0.5 * tip_len - 10e-4, zi), 20e-4, *Y, 21, 3)
    bond(recs, 0, 1, 1, 2)
    bond(recs, 0, 2, 2, 1)
    bond(recs, 3, 4, 1, 2)
    if bonded:               # j and the tip share a bonded neighbour (k)
        bond(recs, 0, 4, 2, 1)
    return recs


def run(oracle, recs, fus=FUS, seed=1, call=0):
    g = oracle.make_geom((-1, -1, -1), (1, 1, 1), (0, 0, 0), (4, 4, 4))
    return oracle.fusion(recs, g, FPAR, oracle.NBR_RADIUS, 0.05, fus, seed, call, 100, 0)


def test_tip_fuses_into_the_side_of_a_segment(oracle):
    recs = scene(oracle)
    out, fused = run(oracle, recs)
    assert fused and len(out) == 6
    j, a, m, i, k, n = out
    tau = (-6e-4 + 0.5 * L) / L
    assert j["rdata"][1] == pytest.approx(tau * L, rel=1e-12, abs=0)
    assert n["rdata"][1] == pytest.approx((1 - tau) * L, rel=1e-12, abs=0)
    np.testing.assert_allclose(j["pos"], [-0.5 * L + 0.5 * tau * L, 0, 0], atol=1e-15)
    np.testing.assert_allclose(n["pos"], [-0.5 * L + tau * L + 0.5 * (1 - tau) * L, 0, 0],
                               atol=1e-15)
    assert n["id"] == 100 and n["idata"][33] == 100
    # bonds by end: j keeps a at its site 1 and meets n and the tip at site 2
    assert bonds(j) == {(11, 1), (100, 2), (20, 2)}
    assert bonds(n) == {(10, 1), (20, 1), (12, 2)}
    assert j["idata"][21] == 1 and n["idata"][22] == 1 and n["idata"][32] == 2
    assert list(j["idata"][[26, 28, 30]]) == [2, 12, 0]     # m recorded as deleted
    # m drops the stale bond to j and is told to repair its site 1
    assert bonds(m) == set() and m["idata"][32] == 1
    # the tip became interior, bonded to j at its free end, bond scale rising
    assert i["idata"][19] == 3 and i["idata"][20] == 1 and i["idata"][25] == 2
    assert bonds(i) == {(21, 1), (10, 2)}
    assert i["rdata"][24] == pytest.approx(FUS[2], rel=1e-15, abs=0)
    assert np.array_equal(a["idata"], recs["idata"][1])      # untouched
    # volumes follow the lengths
    for s in (j, n):
        assert s["rdata"][5] == pytest.approx(math.pi * 4e-8 * s["rdata"][1], rel=1e-14,
                                              abs=0)


@pytest.mark.parametrize("kw,fus", [
    ({}, [0.0, 6e-4, 1e-3, 35e-4]),            # probability 0
    ({"gap": 7e-4}, FUS),                     # beyond max_fusion_separation
    ({"bonded": True}, FUS),                  # a common bonded neighbour
    ({"cross": True}, FUS),                   # the tip's middle is closest
])
def test_no_fusion(oracle, kw, fus):
    recs = scene(oracle, **kw)
    out, fused = run(oracle, recs, fus=fus)
    assert not fused
    assert np.array_equal(out, recs)


def test_fusion_rate(oracle):
    """Independent tip/segment pairs 0.2 apart fuse at the requested rate."""
    n = 400
    recs = np.zeros(5 * n, dtype=oracle.RECORD_DTYPE)
    for c in range(n):
        recs[5 * c:5 * c + 5] = scene(oracle)
    for c in range(n):
        recs["pos"][5 * c:5 * c + 5] += (0.2 * (c % 9) - 0.8, 0.2 * ((c // 9) % 9) - 0.8,
                                          0.2 * (c // 81) - 0.8)
        recs["idata"][5 * c:5 * c + 5, 33] += 100 * c
        recs["id"][5 * c:5 * c + 5] += 100 * c
        for q in range(5):
            ip = recs["idata"][5 * c + q]
            ip[7:7 + ip[6]] += 100 * c
    out, fused = run(oracle, recs, fus=[0.4, 6e-4, 1e-3, 35e-4], seed=9)
    k = len(out) - len(recs)
    assert fused and abs(k - 0.4 * n) < 4.5 * math.sqrt(n * 0.4 * 0.6)


def two_tip_scene(oracle, b_first=False):
    """The scene above plus a second tip b (30) along +x whose free end sits
    1e-5 from the middle of tip i's side, bonded at its site 1 to an interior
    segment (31).  Every segment is within reach of every other.  At the
    start of the pass i is a tip, so b has no target: only i fuses (into j).
    A pass that let b see i after i turned interior would fuse b into i when
    i comes first in record order; `b_first` puts b before i."""
    X = (math.pi / 2, 0.0)
    base = scene(oracle)
    recs = np.zeros(7, dtype=oracle.RECORD_DTYPE)
    recs[:5] = base
    tip_len, blen = 20e-4, 20e-4
    yb = -1e-5 - 0.5 * tip_len            # the middle of i along y
    xb = -6e-4 - 1e-5 - 0.5 * blen        # b's +x end 1e-5 short of i's axis
    seg(recs, 5, (xb, yb, 0.0), blen, *X, 30, 0)
    seg(recs, 6, (xb - blen, yb, 0.0), blen, *X, 31, 3)
    bond(recs, 5, 6, 1, 2)
    if b_first:
        recs = recs[[5, 6, 0, 1, 2, 3, 4]].copy()
    return recs


@pytest.mark.parametrize("b_first", [False, True])
def test_tips_see_the_records_as_the_pass_began(oracle, b_first):
    recs = two_tip_scene(oracle, b_first)
    out, fused = run(oracle, recs)
    assert fused and len(out) == 8           # j split once: only tip i fused
    by_id = {int(r["idata"][33]): r for r in out}
    i, b = by_id[20], by_id[30]
    assert i["idata"][19] == 3 and i["idata"][20] == 1 and bonds(i) == {(21, 1), (10, 2)}
    # b stays a tip bonded only to 31, with no fusion recorded
    assert b["idata"][19] == 0 and b["idata"][20] == 0 and bonds(b) == {(31, 1)}
    assert b["idata"][25] == 0 and b["rdata"][24] == 1.0
