"""GPU parity at the full sizes of the standard workloads.

  * 200,000 hyphal segments (2,000 roots, branching chains on the growth
    surface) in the reference's policy (mesh-cell bins, stencil list every
    step): InitBonds, then 5 steps; bond slots, forces, positions,
    orientations and every integer field bit-exact against the oracle.
  * 256 sphere colonies (500..50,000 cells, log-uniform; volume fraction
    0.1..0.35), autotuned, advanced 100 steps by one captured graph: every
    colony against the same colony run alone (bit-exact where the bins agree,
    1e-12 where the batch's larger bin edge changes the summation order), and
    the smallest, median and largest colonies bit-exact against the oracle
    with the batch's bins.
  * The adjoint of 200 steps of 100,000 spheres with checkpoints every 20:
    the gradient against the oracle's discrete adjoint (dL/dr to 1e-12 of its
    largest entry, the three scalar gradients to 1e-9 relative: sums over
    all particles in another association).
"""
import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_hyphae_200k_reference_policy(bmx, oracle, gpu):
    col = W.c3_hyphae()
    n = len(col.recs)
    assert n == 200_000
    hood = 1.1 * (2.0 * 2.0e-4 + 1.0e-2)   # DEM::neighborhood (bmx_chem.cpp:124)
    print(f"hyphae {n} segments, mesh {col.n_cell}", flush=True)
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    ctx.set_neighbor_params(mode=bmx.NBR_STENCIL, rebuild_every=1, neighborhood=hood,
                            max_neighbors=128)
    grecs = col.recs.copy()
    ctx.bind(grecs)
    orecs = col.recs.copy()
    g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    st = oracle.State(orecs, g, col.fpar, oracle.NBR_STENCIL, hood, 1)
    ctx.init_bonds()
    assert st.init_bonds() == 0
    gi = ctx.int_fields()
    assert np.array_equal(gi[:, 6:19], orecs["idata"][:, 6:19])
    assert orecs["idata"][:, 6].sum() > n          # chains are bonded
    for k in (1, 4):
        ctx.evolve(col.dt, k)
        fo = st.evolve(col.dt, k)
        assert np.array_equal(ctx.fc(), fo)
        assert np.array_equal(ctx.int_fields(), orecs["idata"])
    assert np.abs(fo[:3]).max() > 0
    assert np.array_equal(ctx.positions(), orecs["pos"])
    assert np.array_equal(ctx.orientations(), orecs["rdata"][:, 2:4])
    ctx.close()
    st.close()


def test_ensemble_256_colonies(bmx, oracle, gpu):
    cols = W.c4_ensemble()
    K = cols[0].rebuild_every
    recs, starts, geoms = W.concat_colonies(cols)
    print(f"ensemble {len(cols)} colonies, {len(recs)} spheres, K={K}", flush=True)
    ctx = bmx.Context(0)
    ctx.set_force_params(cols[0].fpar)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=cols[0].skin,
                            max_neighbors=64)
    ctx.bind_colonies(recs, starts, geoms)
    t = ctx.autotune(W.DT, 4 * K)
    steps = 10 * K
    ctx.evolve_graph(W.DT, steps)
    got = ctx.positions()
    assert np.isfinite(got).all()
    exact = 0
    for k, col in enumerate(cols):
        one = bmx.Context(0)
        one.set_force_params(col.fpar)
        one.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
        one.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=col.skin,
                                max_neighbors=64, bin_factor_pct=t["bin_factor_pct"])
        one.bind(col.recs.copy())
        one.evolve(W.DT, steps)
        mine = got[starts[k]:starts[k + 1]]
        if one.bin_dims() == ctx.bin_dims(k):
            assert np.array_equal(mine, one.positions()), k
            exact += 1
        else:
            np.testing.assert_allclose(mine, one.positions(), rtol=0, atol=1e-12)
        one.close()
    print(f"ensemble: {exact} of {len(cols)} colonies compared bit for bit", flush=True)
    assert exact >= len(cols) // 4
    sizes = np.array([len(c.recs) for c in cols])
    for k in (int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]),
              int(np.argmax(sizes))):
        col = cols[k]
        orecs = col.recs.copy()
        g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, ctx.bin_dims(k))
        st = oracle.State(orecs, g, col.fpar, oracle.NBR_PAIR, col.skin, K)
        st.evolve(W.DT, steps)
        assert np.array_equal(got[starts[k]:starts[k + 1]], orecs["pos"]), k
        st.close()
    ctx.close()


def test_adjoint_100k_matches_oracle_gradient(bmx, oracle, gpu):
    n, steps = 100_000, 200
    col = W.sphere_colony(n, 0.3, 11, W_A=6e-6, overlap_free=False, skin=5e-5)
    target = col.recs["pos"] + np.random.default_rng(8).normal(size=(n, 3)) * 1e-5
    fp = col.fpar.copy()
    ctx = bmx.Context(0)
    ctx.set_force_params(fp)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=10, skin=col.skin)
    ctx.bind(col.recs.copy())
    nb = ctx.bin_dims()
    loss, g3, gr, tm = ctx.adjoint(W.DT, steps, target, checkpoint_every=20)
    print(f"adjoint {n} x {steps}: forward {tm.forward_ms:.2f} ms, "
          f"backward {tm.backward_ms:.2f} ms", flush=True)
    orecs = col.recs.copy()
    g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, nb)
    oloss, og3, ogr = oracle.adjoint(orecs, g, fp, 1.0, W.DT, steps, 10, col.skin,
                                     target)
    assert np.array_equal(ctx.positions(), orecs["pos"])
    assert loss == pytest.approx(oloss, rel=1e-13)
    assert np.all(np.abs(g3) > 0)
    np.testing.assert_allclose(g3, og3, rtol=1e-9)
    assert np.abs(gr - ogr).max() <= 1e-12 * np.abs(ogr).max()
    ctx.close()
