"""Batched ensembles: many colonies of uneven size in one context, advanced by
one captured graph per call, with launch shape and bin edge from the
autotuner.

Every colony has its own periodic box and bins; the pair rows of one colony
never reach another's particles, so a colony in the batch follows exactly the
trajectory it follows alone with the same bin edge (launch shape does not
enter the arithmetic).  The checks are bit-exact against separate runs and,
for one colony, against the oracle.
"""
import os

import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ens():
    return W.c4_ensemble(ncol=24, seed=99, nmin=300, nmax=6000)


def batch_ctx(bmx, cols, K=10):
    recs, starts, geoms = W.concat_colonies(cols)
    ctx = bmx.Context(0)
    ctx.set_force_params(cols[0].fpar)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=cols[0].skin)
    ctx.bind_colonies(recs, starts, geoms)
    return ctx, starts


def single_ctx(bmx, col, pct, K=10, block=256, remap=1):
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=col.skin,
                            bin_factor_pct=pct)
    ctx.set_launch_params(block, remap)
    ctx.bind(col.recs.copy())
    return ctx


def test_autotune_leaves_state_and_uses_cache(bmx, gpu, ens, tmp_path):
    cache = str(tmp_path / "tune.jsonl")
    ctx, _ = batch_ctx(bmx, ens)
    x0 = ctx.positions()
    t = ctx.autotune(W.DT, 20, cache=cache)
    assert t["candidates"] >= 8 and not t["from_cache"]
    assert t["block_size"] in (64, 128, 256, 512) and t["bin_factor_pct"] in (100, 150, 200)
    assert t["us_per_step"] > 0
    assert np.array_equal(ctx.positions(), x0)          # tuning did not advance
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and '"key": "gfx' in lines[0]

    # the tuned context steps exactly like a fresh one with those settings
    ctx.evolve(W.DT, 30)
    ref, _ = batch_ctx(bmx, ens)
    ref.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=10, skin=ens[0].skin,
                            bin_factor_pct=t["bin_factor_pct"])
    ref.set_launch_params(t["block_size"], t["xcd_remap"])
    ref.evolve(W.DT, 30)
    assert np.array_equal(ctx.positions(), ref.positions())

    again, _ = batch_ctx(bmx, ens)
    t2 = again.autotune(W.DT, 20, cache=cache)
    assert t2["from_cache"] == 1
    assert (t2["block_size"], t2["xcd_remap"], t2["bin_factor_pct"]) == \
        (t["block_size"], t["xcd_remap"], t["bin_factor_pct"])
    assert len(open(cache).read().splitlines()) == 1
    for c in (ctx, ref, again):
        c.close()


def test_graph_ensemble_matches_separate_runs(bmx, oracle, gpu, ens):
    ctx, starts = batch_ctx(bmx, ens)
    t = ctx.autotune(W.DT, 20)
    pct = t["bin_factor_pct"]
    ctx.evolve(W.DT, 10)                 # into the steady rebuild phase
    for _ in range(3):                   # first call captures, later calls replay
        ctx.evolve_graph(W.DT, 100)
    got = ctx.positions()
    exact = 0
    for k, col in enumerate(ens):
        one = single_ctx(bmx, col, pct)
        one.evolve(W.DT, 310)
        mine = got[starts[k]:starts[k + 1]]
        if one.bin_dims() == ctx.bin_dims(k):
            # same bins: same rows in the same order
            assert np.array_equal(mine, one.positions()), k
            exact += 1
        else:
            # the batch's bin edge comes from the largest radius of all
            # colonies: same pairs, another summation order
            np.testing.assert_allclose(mine, one.positions(), rtol=0, atol=1e-12)
        one.close()
    assert exact > 0
    # three colonies against the oracle with the batch's bins
    sizes = [len(c.recs) for c in ens]
    for k in (int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]),
              int(np.argmax(sizes))):
        col = ens[k]
        recs = col.recs.copy()
        g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, ctx.bin_dims(k))
        st = oracle.State(recs, g, col.fpar, oracle.NBR_PAIR, col.skin, 10)
        st.evolve(W.DT, 310)
        assert np.array_equal(got[starts[k]:starts[k + 1]], recs["pos"]), k
        st.close()
    ctx.close()
