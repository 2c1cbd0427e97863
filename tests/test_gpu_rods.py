"""GPU parity of the segment (hyphae) path against the CPU oracle.

Segment pairs go through sin/cos (orientations) and acos (the orientation
update).  Both sides use the same IEEE-only implementation of those
(bmx_math.h, within 2 ulp of glibc), so forces, orientations, trajectories
and the integer state (bond topology, positions in the hyphae, the reset
fields) must match bit for bit.  The 1e-12 / 1e-10 tolerance helpers below
are the bar against a glibc build; the tests assert equality.
"""
import math
import os

import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def small_hyphae(n=3000, nroots=30, seed=11):
    return W.c3_hyphae(n=n, nroots=nroots, seed=seed)


def make_pair(bmx, oracle, col, mode=0, K=10, hood=1.0, maxn=0, block=None):
    ctx = bmx.Context(0)
    if block is not None:
        ctx.set_launch_params(block, 1)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    ctx.set_neighbor_params(mode=mode, rebuild_every=K, skin=col.skin,
                            neighborhood=hood, max_neighbors=maxn)
    grecs = col.recs.copy()
    ctx.bind(grecs)
    nb = ctx.bin_dims() if mode == 0 else col.n_cell
    orecs = col.recs.copy()
    g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, nb)
    if mode == 0:
        st = oracle.State(orecs, g, col.fpar, oracle.NBR_PAIR, col.skin, K)
    else:
        st = oracle.State(orecs, g, col.fpar, oracle.NBR_STENCIL, hood, 1)
    return ctx, grecs, st, orecs


def close_fc(fg, fo, rel=1e-12):
    assert np.array_equal(fg, fo)
    for blk in (slice(0, 3), slice(3, 6)):
        scale = np.abs(fo[blk]).max()
        err = np.abs(fg[blk] - fo[blk]).max()
        assert err <= rel * scale + 1e-300, (err, scale)


def close_traj(ctx, orecs, box, rel=1e-10):
    assert np.array_equal(ctx.positions(), orecs["pos"])
    assert np.array_equal(ctx.orientations(), orecs["rdata"][:, 2:4])
    err = np.abs(ctx.positions() - orecs["pos"]).max()
    assert err <= rel * box, err
    ori = ctx.orientations()
    oo = orecs["rdata"][:, 2:4]
    dphi = np.abs(ori - oo)
    dphi[:, 1] = np.minimum(dphi[:, 1], 2 * math.pi - dphi[:, 1])
    assert dphi.max() <= rel * 2 * math.pi, dphi.max()


@pytest.mark.parametrize("block", [64, 256])  # one-wave blocks are the default
@pytest.mark.parametrize("mode", [0, 1])
def test_hyphae_init_bonds_and_steps(bmx, oracle, gpu, mode, block):
    col = small_hyphae()
    hood = 1.1 * (2 * 5.47e-4 + 1e-2)  # fungi deck: 1.1 (2 r(max_vol) + width)
    ctx, grecs, st, orecs = make_pair(bmx, oracle, col, mode=mode,
                                      K=10 if mode == 0 else 1, hood=hood, maxn=256,
                                      block=block)
    ctx.init_bonds()
    assert st.init_bonds() == 0
    gi = ctx.int_fields()
    assert np.array_equal(gi[:, 6:19], orecs["idata"][:, 6:19])   # bond slots
    nb = orecs["idata"][:, 6]
    # chains are bonded end to end; the reference's mesh stencil (2e-3 cm
    # cells, 3e-3 cm segments) misses bonded pairs two cells apart
    assert nb.sum() > (1.5 if mode == 0 else 1.0) * len(orecs)
    box = (col.prob_hi - col.prob_lo).max()
    ctx.evolve(col.dt, 1)
    fo = st.evolve(col.dt, 1)
    close_fc(ctx.fc(), fo)
    assert np.array_equal(ctx.int_fields(), orecs["idata"])
    for _ in range(3):
        ctx.evolve(col.dt, 3)
        fo = st.evolve(col.dt, 3)
        close_fc(ctx.fc(), fo)
    close_traj(ctx, orecs, box)
    ctx.sync()
    assert np.array_equal(grecs["idata"], orecs["idata"])
    np.testing.assert_allclose(grecs["rdata"][:, 6:12], orecs["rdata"][:, 6:12],
                               rtol=0, atol=1e-12 * np.abs(orecs["rdata"][:, 6:12]).max())


def test_mixed_spheres_and_segments(bmx, oracle, gpu):
    """Spheres and segments never interact (r_lm = 0, K.H:39) but share the
    lists, the wall and the update."""
    hy = small_hyphae(n=1500, nroots=15, seed=3)
    L = hy.prob_hi - hy.prob_lo
    sp, _, _ = bmx.generate_spheres(1500, 1.5e-4, 2.5e-4, 0.3, 5, aspect=(L[0], L[1], 4e-3),
                                    periodic=(1, 1, 0), overlap_free=False)
    sp["pos"][:, :2] *= L[:2] / sp["pos"][:, :2].max(axis=0)
    sp["pos"][:, 2] = np.clip(sp["pos"][:, 2], 0, L[2] * 0.9)
    both = np.concatenate([hy.recs, sp]).astype(bmx.RECORD_DTYPE)
    col = W.Colony(both, hy.prob_lo, hy.prob_hi, hy.periodic,
                   hy.fpar, n_cell=hy.n_cell, skin=hy.skin)
    col.recs["id"] = np.arange(1, len(col.recs) + 1)
    col.recs["idata"][:, 33] = col.recs["id"]
    # overlapping random spheres move fast: rebuild every step
    ctx, grecs, st, orecs = make_pair(bmx, oracle, col, K=1, maxn=256)
    ctx.init_bonds()
    st.init_bonds()
    for _ in range(4):
        ctx.evolve(col.dt, 3)
        fo = st.evolve(col.dt, 3)
        close_fc(ctx.fc(), fo)
    close_traj(ctx, orecs, L.max())
    assert np.array_equal(ctx.int_fields(), orecs["idata"])


@pytest.mark.parametrize("block", [64, 256])
def test_bond_repair_and_promotion(bmx, oracle, gpu, block):
    """One-sided bond tables (a bond dropped from one end, as division and
    fusion leave them) and second segments: the force pass repairs the
    tables and promotes the segments as the row loop does (on the GPU a
    tile with a repair is recomputed row by row), bit for bit."""
    col = small_hyphae()
    hood = 1.1 * (2 * 5.47e-4 + 1e-2)
    recs = col.recs.copy()
    g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    st0 = oracle.State(recs, g, col.fpar, oracle.NBR_STENCIL, hood, 1)
    assert st0.init_bonds() == 0
    ip = recs["idata"]
    rng = np.random.default_rng(5)
    bonded = np.flatnonzero(ip[:, 6] >= 1)
    drop = rng.choice(bonded, size=len(bonded) // 10, replace=False)
    ip[drop, 6] -= 1                        # the last bond forgotten on one side
    inner = np.flatnonzero(ip[:, 19] == 3)
    sec = rng.choice(inner, size=60, replace=False)
    ip[sec, 19] = rng.integers(1, 3, size=60)       # SECOND_1 / SECOND_2
    col.recs = recs.copy()
    before = col.recs["idata"].copy()
    ctx, grecs, st, orecs = make_pair(bmx, oracle, col, mode=1, K=1, hood=hood,
                                      maxn=256, block=block)
    for _ in range(3):
        ctx.evolve(col.dt, 1)
        fo = st.evolve(col.dt, 1)
        close_fc(ctx.fc(), fo)
    close_traj(ctx, orecs, (col.prob_hi - col.prob_lo).max())
    gi = ctx.int_fields()
    assert np.array_equal(gi, orecs["idata"])
    assert (gi[drop, 6] > before[drop, 6]).sum() > len(drop) // 2   # repaired
    assert (gi[sec, 19] == 3).sum() > 10                             # promoted


def test_segment_colonies_match_separate_runs(bmx, gpu):
    """Three hyphae colonies bound as one context (one box each) step
    exactly as three separate contexts: the flat segment pass with one box
    per particle gives the single-box results bit for bit."""
    cols = [small_hyphae(n=1500, nroots=15, seed=s) for s in (3, 4, 5)]
    recs, starts, geoms = W.concat_colonies(cols)
    ctx = bmx.Context(0)
    ctx.set_force_params(cols[0].fpar)
    ctx.set_neighbor_params(mode=0, rebuild_every=10, skin=cols[0].skin,
                            max_neighbors=256)
    ctx.bind_colonies(recs, starts, geoms)
    ctx.init_bonds()
    ctx.evolve(cols[0].dt, 20)
    pos, ori, ints = ctx.positions(), ctx.orientations(), ctx.int_fields()
    for k, col in enumerate(cols):
        one = bmx.Context(0)
        one.set_force_params(col.fpar)
        one.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
        one.set_neighbor_params(mode=0, rebuild_every=10, skin=col.skin,
                                max_neighbors=256)
        one.bind(col.recs.copy())
        one.init_bonds()
        one.evolve(col.dt, 20)
        sl = slice(starts[k], starts[k + 1])
        assert np.array_equal(pos[sl], one.positions()), k
        assert np.array_equal(ori[sl], one.orientations()), k
        assert np.array_equal(ints[sl], one.int_fields()), k
        assert (ints[sl][:, 6] > 0).any()


@pytest.mark.parametrize("fname", ["fungi_init_cfg.dat", "wall_test_fungi_init_cfg.dat"])
def test_reference_single_segment_decks(bmx, oracle, gpu, fname):
    """The reference's own fungi and wall_test inputs: one segment settling on
    the growth surface (fluid.surface_location = 0.048, fungi deck forces)."""
    recs = bmx.read_particles_ascii(os.path.join(GOLDEN, "ref_inputs", fname))
    lo, hi = np.zeros(3), np.array([0.128, 0.128, 0.064])
    col = W.Colony(recs, lo, hi, (1, 1, 0), W.hyphae_fpar(z_wall=0.048, k_bond=1.0),
                   n_cell=(64, 64, 32), skin=2e-4)
    ctx, _, st, orecs = make_pair(bmx, oracle, col)
    for _ in range(5):
        ctx.evolve(0.25, 20)
        fo = st.evolve(0.25, 20)
        close_fc(ctx.fc(), fo)
    close_traj(ctx, orecs, 0.128)
    assert orecs["pos"][0, 2] != recs["pos"][0, 2]   # it moved
