"""GPU parity of the sphere path against the CPU oracle, through the C ABI.

Sphere arithmetic is +, -, *, / and sqrt only, rounded as written on both
sides (no contraction), and both sides sum each particle's pair terms in the
same canonical order, so forces, neighbor lists, bins and trajectories must
match bit for bit.  Across neighbor policies (Verlet list vs the reference's
per-step stencil list) only the summation order differs: forces then agree
to 1e-12 of the largest force and 10-step trajectories to 1e-10.
"""
import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_ctx(bmx, col, mode=0, K=10, n_cell=None, neighborhood=1.0, maxn=0,
            block=256, remap=1):
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, n_cell or col.n_cell)
    ctx.set_neighbor_params(mode=mode, rebuild_every=K, skin=col.skin,
                            neighborhood=neighborhood, max_neighbors=maxn)
    ctx.set_launch_params(block, remap)
    recs = col.recs.copy()
    ctx.bind(recs)
    return ctx, recs


def oracle_state(oracle, col, nbin, mode=None, K=10, radius=None):
    """The oracle with the library's policy: mode None is the Verlet policy
    (per-pair cutoff r_lm + skin, NBR_PAIR with radius = skin)."""
    recs = col.recs.copy()
    g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, nbin)
    if mode is None:
        mode, radius = oracle.NBR_PAIR, col.skin
    return oracle.State(recs, g, col.fpar, mode, radius, K), recs


def rows(off, idx):
    return [list(idx[off[i]:off[i + 1]]) for i in range(len(off) - 1)]


@pytest.fixture(scope="module")
def c1():
    return W.c1_small()


def test_c1_first_step_bit_exact(bmx, oracle, gpu, c1):
    ctx, _ = gpu_ctx(bmx, c1)
    ctx.evolve(W.DT, 1)
    nb = ctx.bin_dims()
    st, orecs = oracle_state(oracle, c1, nb)
    fo = st.evolve(W.DT, 1)
    fg = ctx.fc()
    assert np.array_equal(fg, fo)
    assert np.abs(fo[:3]).max() > 0
    assert not fg[3:].any()
    assert np.array_equal(ctx.positions(), orecs["pos"])
    og, ig = ctx.neighbors()
    oo, io = st.neighbors()
    assert rows(og, ig) == rows(oo, io)          # same pairs, same order
    bins = ctx.bins()
    g = oracle.make_geom(c1.prob_lo, c1.prob_hi, c1.periodic, nb)
    # bins of the rebuild, i.e. of the initial positions
    expect = [oracle.bin_index(p, g)[0] for p in c1.recs["pos"]]
    assert np.array_equal(bins, expect)


def test_c1_trajectory_100_steps_bit_exact(bmx, oracle, gpu, c1):
    ctx, recs = gpu_ctx(bmx, c1)
    nb = ctx.bin_dims()
    st, orecs = oracle_state(oracle, c1, nb)
    for chunk in (1, 9, 40, 50):
        ctx.evolve(W.DT, chunk)
        fo = st.evolve(W.DT, chunk)
        assert np.array_equal(ctx.positions(), orecs["pos"])
        assert np.array_equal(ctx.fc(), fo)
    ctx.sync()
    assert np.array_equal(recs["pos"], orecs["pos"])
    assert np.array_equal(recs["rdata"][:, 6:12], orecs["rdata"][:, 6:12])
    assert np.array_equal(recs["idata"], orecs["idata"])
    moved = np.abs(orecs["pos"] - c1.recs["pos"]).max()
    assert moved > 0


def test_stencil_mode_bit_exact(bmx, oracle, gpu, c1):
    """The reference's own policy: mesh-cell bins, +-1 stencil, every pair
    with |dx|^2 <= neighborhood, rebuilt every step."""
    L = c1.prob_hi - c1.prob_lo
    n_cell = tuple(int(L[d] // 2.0e-4) for d in range(3))
    hood = 1.1 * (2 * 7e-5 + 1e-2)   # DEM::neighborhood (bmx_chem.cpp:124)
    ctx, _ = gpu_ctx(bmx, c1, mode=1, K=1, n_cell=n_cell, neighborhood=hood, maxn=160)
    st, orecs = oracle_state(oracle, c1, n_cell, mode=1, K=1, radius=hood)
    for _ in range(3):
        ctx.evolve(W.DT, 5)
        fo = st.evolve(W.DT, 5)
        assert np.array_equal(ctx.fc(), fo)
        assert np.array_equal(ctx.positions(), orecs["pos"])
    og, ig = ctx.neighbors()
    oo, io = st.neighbors()
    assert rows(og, ig) == rows(oo, io)


def test_verlet_matches_reference_policy(bmx, gpu, c1):
    L = c1.prob_hi - c1.prob_lo
    n_cell = tuple(int(L[d] // 2.0e-4) for d in range(3))
    a, _ = gpu_ctx(bmx, c1)
    b, _ = gpu_ctx(bmx, c1, mode=1, K=1, n_cell=n_cell, neighborhood=1.0, maxn=160)
    for _ in range(10):
        a.evolve(W.DT, 1)
        b.evolve(W.DT, 1)
        fa, fb = a.fc(), b.fc()
        np.testing.assert_allclose(fa, fb, rtol=0, atol=1e-12 * np.abs(fb).max())
    np.testing.assert_allclose(a.positions(), b.positions(), rtol=0, atol=1e-10)


def test_graph_replay_matches_eager(bmx, gpu, c1):
    a, _ = gpu_ctx(bmx, c1)
    b, _ = gpu_ctx(bmx, c1)
    a.evolve(W.DT, 30)
    for _ in range(3):
        b.evolve_graph(W.DT, 10)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.fc(), b.fc())


@pytest.mark.parametrize("block,remap", [(64, 0), (128, 1), (512, 1)])
def test_launch_shape_does_not_change_results(bmx, gpu, c1, block, remap):
    a, _ = gpu_ctx(bmx, c1)
    b, _ = gpu_ctx(bmx, c1, block=block, remap=remap)
    a.evolve(W.DT, 12)
    b.evolve(W.DT, 12)
    assert np.array_equal(a.positions(), b.positions())


def test_colonies_match_separate_runs(bmx, oracle, gpu):
    cols = [W.sphere_colony(n, phi, s, W_A=2e-5) for n, phi, s in
            [(700, 0.15, 1), (1500, 0.3, 2), (520, 0.35, 3), (3000, 0.1, 4)]]
    recs, starts, geoms = W.concat_colonies(cols)
    ctx = bmx.Context(0)
    ctx.set_force_params(cols[0].fpar)
    ctx.set_neighbor_params(mode=0, rebuild_every=10, skin=cols[0].skin)
    ctx.bind_colonies(recs, starts, geoms)
    ctx.evolve(W.DT, 25)
    pos = ctx.positions()
    for k, c in enumerate(cols):
        st, orecs = oracle_state(oracle, c, ctx.bin_dims(k))
        st.evolve(W.DT, 25)
        assert np.array_equal(pos[starts[k]:starts[k + 1]], orecs["pos"]), k


def test_edge_cases(bmx, oracle, gpu, c1):
    # a single particle: wall term only
    one = W.sphere_colony(1, 0.01, 3)
    ctx, _ = gpu_ctx(bmx, one)
    ctx.evolve(W.DT, 5)
    st, orecs = oracle_state(oracle, one, ctx.bin_dims())
    st.evolve(W.DT, 5)
    assert np.array_equal(ctx.positions(), orecs["pos"])
    # particles above the box top (bins clamp) and below the surface
    col = W.c1_small()
    col.recs["pos"][::50, 2] += 2 * (col.prob_hi[2] - col.prob_lo[2])
    col.recs["pos"][1::50, 2] = -1e-4
    ctx, _ = gpu_ctx(bmx, col)
    ctx.evolve(W.DT, 12)
    st, orecs = oracle_state(oracle, col, ctx.bin_dims())
    st.evolve(W.DT, 12)
    assert np.array_equal(ctx.positions(), orecs["pos"])
    # no particles
    empty = W.c1_small()
    empty.recs = empty.recs[:0].copy()
    ctx, _ = gpu_ctx(bmx, empty)
    ctx.evolve(W.DT, 3)
    assert ctx.positions().shape == (0, 3)


def test_dense_rows_bit_exact(bmx, oracle, gpu):
    """Long rows where nearly every entry interacts (W_A = 1e-4 cm, skin
    1e-9 cm, lists rebuilt every step): a wave's entries span many windows
    of 128 and a window packs more than 64 interacting pairs."""
    col = W.sphere_colony(4000, 0.3, 11, W_A=1.0e-4, skin=1.0e-9)
    ctx, _ = gpu_ctx(bmx, col, K=1)          # lists rebuilt every step
    ctx.evolve(W.DT, 3)
    st, orecs = oracle_state(oracle, col, ctx.bin_dims(), K=1)
    fo = st.evolve(W.DT, 3)
    assert np.array_equal(ctx.fc(), fo)
    assert np.array_equal(ctx.positions(), orecs["pos"])
    off, idx = ctx.neighbors()
    assert (off[-1] / len(col.recs)) > 8
    p, r = orecs["pos"], orecs["rdata"][:, 0]
    i = np.repeat(np.arange(len(r)), np.diff(off))
    d = p[i] - p[idx]
    L = col.prob_hi - col.prob_lo
    for a in range(3):
        if col.periodic[a]:
            d[:, a] -= L[a] * np.round(d[:, a] / L[a])
    inter = np.sqrt((d * d).sum(1)) <= r[i] + r[idx] + col.fpar[0]
    assert inter.mean() > 0.6


@pytest.mark.parametrize("n,phi", [(3000, 0.001), (2000, 0.003), (64, 0.05), (1000, 0.3)])
def test_max_neighbors_at_the_longest_row(bmx, oracle, gpu, n, phi):
    """max_neighbors equal to the longest row of the run: each tile's region
    of the flat rows is as small as it can be (down to 64 entries), so the
    force pass's first window reads into the next tile's region and, for the
    last tile, into the padding; results stay bit-exact."""
    col = W.sphere_colony(n, phi, 5, W_A=2e-5)
    probe, _ = gpu_ctx(bmx, col, K=1)
    nb = probe.bin_dims()
    st, _ = oracle_state(oracle, col, nb, K=1)
    longest = 1
    for _ in range(4):                   # the list each step uses
        st.evolve(W.DT, 1)
        off, _ = st.neighbors()
        longest = max(longest, int(np.diff(off).max()))
    ctx, _ = gpu_ctx(bmx, col, K=1, maxn=longest)
    st, orecs = oracle_state(oracle, col, nb, K=1)
    for _ in range(4):
        ctx.evolve(W.DT, 1)
        fo = st.evolve(W.DT, 1)
        assert np.array_equal(ctx.fc(), fo)
    assert np.array_equal(ctx.positions(), orecs["pos"])


def test_errors_are_reported(bmx, gpu, c1):
    col = W.c1_small()
    col.recs["pos"][1] = col.recs["pos"][0]          # coincident pair
    ctx, _ = gpu_ctx(bmx, col)
    with pytest.raises(bmx.BMXError) as ei:
        ctx.evolve(W.DT, 1)
    assert ei.value.code == 4
    ctx, _ = gpu_ctx(bmx, c1, maxn=2)                 # rows overflow
    with pytest.raises(bmx.BMXError) as ei:
        ctx.evolve(W.DT, 1)
    assert ei.value.code == 5
    fast = W.c1_small()
    fast.fpar[1] = 1e12                               # very stiff: moves > skin/2
    fast.skin = 1e-9
    ctx, _ = gpu_ctx(bmx, fast, K=10)
    with pytest.raises(bmx.BMXError) as ei:
        ctx.evolve(W.DT, 10)
    assert ei.value.code == 6
    with pytest.raises(bmx.BMXError):
        ctx.set_neighbor_params(mode=1, rebuild_every=5)
    bad = W.c1_small()
    bad.recs["rdata"][3, 0] = -1.0
    with pytest.raises(bmx.BMXError) as ei:
        gpu_ctx(bmx, bad)
    assert ei.value.code == 1


@pytest.mark.parametrize("block", [256, 64])
def test_one_million_spheres_10_steps_bit_exact(bmx, oracle, gpu, block):
    """The 1,000,000-sphere workload at volume fraction 0.3: 10 steps (one
    Verlet rebuild) on the GPU and on the CPU oracle, with 256-thread blocks
    and with the one-wave blocks the benchmark launches."""
    col = W.c2_large()
    ctx, _ = gpu_ctx(bmx, col, block=block)
    ctx.evolve(W.DT, 10)
    st, orecs = oracle_state(oracle, col, ctx.bin_dims())
    fo = st.evolve(W.DT, 10)
    assert np.array_equal(ctx.fc(), fo)
    assert np.array_equal(ctx.positions(), orecs["pos"])
    og, ig = ctx.neighbors()
    oo, io = st.neighbors()
    assert np.array_equal(og, oo) and np.array_equal(ig, io)


def test_failed_rebind_leaves_context_unbound(bmx, gpu, c1):
    """A bind that fails (here: a record with a non-finite position, found
    after the arrays were resized for it) leaves no particles bound: evolve
    then reports STATE instead of running kernels over rows built for the
    previous, smaller particle set.  A later good bind works again."""
    ctx, _ = gpu_ctx(bmx, c1)
    ctx.evolve(W.DT, 3)
    big = np.concatenate([c1.recs, c1.recs, c1.recs]).astype(bmx.RECORD_DTYPE)
    big["pos"][-1, 0] = np.nan
    with pytest.raises(bmx.BMXError) as e:
        ctx.bind(big)
    assert e.value.code == 1                          # BAD_ARG
    for call in (lambda: ctx.evolve(W.DT, 1), lambda: ctx.evolve_graph(W.DT, 2),
                 lambda: ctx.fc()):
        with pytest.raises(bmx.BMXError) as e:
            call()
        assert e.value.code == 7                      # STATE
    # bad geometry fails before any state is read, and also unbinds
    ctx2, _ = gpu_ctx(bmx, c1)
    ctx2.evolve(W.DT, 1)
    with pytest.raises(bmx.BMXError):
        ctx2.bind_colonies(c1.recs.copy(), [0, len(c1.recs)],
                           [bmx.make_geometry(c1.prob_hi, c1.prob_lo, c1.periodic)])
    with pytest.raises(bmx.BMXError) as e:
        ctx2.evolve(W.DT, 1)
    assert e.value.code == 7
    ctx2.close()
    ctx.bind(c1.recs.copy())
    ctx.evolve(W.DT, 1)
    ctx.close()


def test_guard_takes_violating_periods_again(bmx, gpu):
    """A skin of 1e-9 cm is far too small for lists kept 10 steps: unguarded,
    the call fails with BMX_ERR_SKIN; guarded, every such period is taken
    again with a fresh list before each step, so the run advances and equals
    a run that rebuilds every step (the same lists: bit for bit) and the
    reference's stencil list every step (same pairs, another order: 1e-12)."""
    col = W.c1_small()
    col.skin = 1e-9
    plain, _ = gpu_ctx(bmx, col, K=10)
    with pytest.raises(bmx.BMXError) as e:
        plain.evolve(W.DT, 10)
    assert e.value.code == 6
    plain.close()
    g, _ = gpu_ctx(bmx, col, K=10)
    g.set_guard(True)
    g.evolve(W.DT, 7)
    g.evolve(W.DT, 23)                 # a period spanning two calls
    t = g.timing()
    assert t.guard_redone >= 3 and t.steps_total == 30
    every, _ = gpu_ctx(bmx, col, K=1)
    every.evolve(W.DT, 30)
    assert np.array_equal(g.positions(), every.positions())
    L = col.prob_hi - col.prob_lo
    n_cell = tuple(int(v) for v in np.floor(L / 2.0e-4))
    sten, _ = gpu_ctx(bmx, col, mode=1, K=1, n_cell=n_cell, neighborhood=1.0)
    sten.evolve(W.DT, 30)
    np.testing.assert_allclose(g.positions(), sten.positions(), rtol=0,
                               atol=1e-12 * L.max())
    moved = np.abs(g.positions() - col.recs["pos"]).max()
    assert moved > 1e-9                # the skin was exceeded many times over
    for c in (g, every, sten):
        c.close()
