"""GPU fusion against the oracle.

The scenes of test_oracle_fusion.py (a tip ending beside a segment, and 400
such pairs fusing with probability 0.4) run through bmx_fusion and through
the oracle with the same rows (pair cutoff lists over the context's bins):
the records, including the split segments and the repaired bond lists, must
match bit for bit, and so must the positions of the mechanics steps after.
"""
import numpy as np
import pytest

from test_oracle_fusion import FPAR, FUS, scene

pytestmark = pytest.mark.gpu

SKIN = 1e-3
LO, HI = (-0.06, -0.06, -0.06), (0.06, 0.06, 0.06)


def pair(bmx, oracle, recs, block=None):
    ctx = bmx.Context(0)
    if block is not None:
        ctx.set_launch_params(block, 1)
    ctx.set_force_params(FPAR)
    ctx.set_geometry(LO, HI, (0, 0, 0))
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=1, skin=SKIN,
                            max_neighbors=256)
    ctx.bind(recs.copy())
    g = oracle.make_geom(LO, HI, (0, 0, 0), ctx.bin_dims())
    return ctx, g


def spread(oracle, n):
    recs = np.zeros(5 * n, dtype=oracle.RECORD_DTYPE)
    for c in range(n):
        recs[5 * c:5 * c + 5] = scene(oracle)
        recs["pos"][5 * c:5 * c + 5] += (0.012 * (c % 8) - 0.042,
                                         0.012 * ((c // 8) % 8) - 0.042,
                                         0.012 * (c // 64) - 0.042)
        recs["idata"][5 * c:5 * c + 5, 33] += 100 * c
        recs["id"][5 * c:5 * c + 5] += 100 * c
        for q in range(5):
            ip = recs["idata"][5 * c + q]
            ip[7:7 + ip[6]] += 100 * c
    return recs


@pytest.mark.parametrize("block", [64, 256])  # one-wave blocks are the default
@pytest.mark.parametrize("n,prob", [(1, 1.0), (400, 0.4)])
def test_fusion_matches_oracle(bmx, oracle, gpu, n, prob, block):
    recs = scene(oracle) if n == 1 else spread(oracle, n)
    ctx, g = pair(bmx, oracle, recs, block)
    fus = [prob] + FUS[1:]
    nid = 10**6
    grecs, fused, added = ctx.fusion(fus, 5, 3, nid, 0)
    out, ofused = oracle.fusion(recs, g, FPAR, oracle.NBR_PAIR, SKIN, fus, 5, 3, nid, 0)
    assert fused and ofused
    assert added == len(out) - len(recs) > 0
    assert np.array_equal(grecs, out)
    # the mechanics go on from the fused network
    st = oracle.State(out, g, np.array(FPAR), oracle.NBR_PAIR, SKIN, 1)
    ctx.evolve(0.01, 7)
    st.evolve(0.01, 7)
    assert np.array_equal(ctx.positions(), out["pos"])
    ctx.close()


def test_no_fusion_changes_nothing(bmx, oracle, gpu):
    recs = scene(oracle, gap=7e-4)
    ctx, g = pair(bmx, oracle, recs)
    before = ctx._recs.copy()
    grecs, fused, added = ctx.fusion(FUS, 5, 3, 10**6, 0)
    assert not fused and added == 0
    assert np.array_equal(grecs, before)
    ctx.close()


@pytest.mark.parametrize("b_first", [False, True])
def test_two_tips_match_oracle(bmx, oracle, gpu, b_first):
    """Two tips in reach of each other and of one segment: decisions are taken
    against the records as the pass began, on the GPU as in the oracle."""
    from test_oracle_fusion import two_tip_scene
    recs = two_tip_scene(oracle, b_first)
    ctx, g = pair(bmx, oracle, recs)
    grecs, fused, added = ctx.fusion(FUS, 5, 3, 10**6, 0)
    out, ofused = oracle.fusion(recs, g, FPAR, oracle.NBR_PAIR, SKIN, FUS, 5, 3, 10**6, 0)
    assert fused and ofused and added == 1 and len(out) == len(recs) + 1
    assert np.array_equal(grecs, out)
    ctx.close()


def test_fusion_mid_period_keeps_skin_check_clean(bmx, oracle, gpu):
    """Verlet lists every 10 steps: 9 steps, a fusion call that fuses nothing
    (it rebuilds the rows), 1 step.  The displacement record restarts at that
    rebuild, so the call succeeds, and the positions follow the oracle."""
    recs = scene(oracle, gap=7e-4)
    ctx = bmx.Context(0)
    ctx.set_force_params(FPAR)
    ctx.set_geometry(LO, HI, (0, 0, 0))
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=10, skin=SKIN,
                            max_neighbors=256)
    ctx.bind(recs.copy())
    g = oracle.make_geom(LO, HI, (0, 0, 0), ctx.bin_dims())
    ctx.evolve(0.01, 9)
    _, fused, added = ctx.fusion(FUS, 5, 3, 10**6, 0)
    assert not fused and added == 0
    ctx.evolve(0.01, 1)
    orecs = recs.copy()
    st = oracle.State(orecs, g, np.array(FPAR), oracle.NBR_PAIR, SKIN, 10)
    st.evolve(0.01, 10)
    np.testing.assert_allclose(ctx.positions(), orecs["pos"], rtol=0, atol=1e-16)
    ctx.close()
