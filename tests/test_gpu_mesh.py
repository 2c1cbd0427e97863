"""GPU particle <-> mesh coupling against the oracle.

Deposition (volume fraction, fluid increments, particle counts; one-to-one
and trilinear) sums each cell's particles in ascending particle order on both
sides, so the meshes must match bit for bit.  The mesh -> particle update
(interpolation, gradient, chemistry and growth over the substeps) is
per-particle arithmetic with the shared pow, so the records must match bit
for bit too, and the mechanics steps after it (rows rebuilt for the grown
sizes) must give the oracle's positions.
"""
import math

import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W
from conftest import HipArray

pytestmark = pytest.mark.gpu

# fungi deck (exec/fungi/input_fungi): k1 k2 k3 kr1 kr2 kr3 kg kv, max_seg_radius,
# max_seg_length
CHEMPAR = [2.0e-5, 4.0, 2.0e-5, 2.0e-5, 6.0e-4, 2.0e-5, 4.0, 4.0e4, 2.5e-4, 35.0e-4]


def spheres(seed=3):
    col = W.c1_small(seed=seed)
    col.n_cell = (16, 16, 8)
    r = col.recs["rdata"]
    rng = np.random.default_rng(seed)
    r[:, 5] = 4.0 / 3.0 * math.pi * r[:, 0] ** 3
    r[:, 4] = 4.0 * math.pi * r[:, 0] ** 2
    r[:, 25:28] = rng.uniform(0.0, 2e-5, (len(r), 3))
    r[:, 31:34] = rng.uniform(-1e-12, 1e-12, (len(r), 3))
    return col


def hyphae(n=3000, seed=17):
    col = W.c3_hyphae(n=n, nroots=30, seed=seed)
    rng = np.random.default_rng(seed)
    r = col.recs["rdata"]
    r[:, 5] = math.pi * r[:, 0] ** 2 * r[:, 1]
    r[:, 25:28] = rng.uniform(0.0, 2e-5, (n, 3))
    r[:, 31:34] = rng.uniform(-1e-12, 1e-12, (n, 3))
    # a few tips at the radius limit and some interior segments at it
    r[::97, 0] = CHEMPAR[8]
    return col


def setup(bmx, oracle, col, mode=0):
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    if mode == 0:
        ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=10, skin=col.skin,
                                max_neighbors=384)
    else:
        ctx.set_neighbor_params(mode=bmx.NBR_STENCIL, rebuild_every=1,
                                neighborhood=1.1 * (2 * 2e-4 + 1e-2), max_neighbors=384)
    grecs = col.recs.copy()
    ctx.bind(grecs)
    g = oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    return ctx, grecs, g


def mesh_inputs(col, seed=5):
    """X_k (3 species) and vf with one ghost layer, positive."""
    nx, ny, nz = col.n_cell
    rng = np.random.default_rng(seed)
    X = np.empty((3, nz + 2, ny + 2, nx + 2))
    X[0] = rng.uniform(1.0e-5, 3.0e-5, X[0].shape)
    X[1] = rng.uniform(0.0, 1.0e-7, X[1].shape)
    X[2] = rng.uniform(1.0e-6, 3.0e-6, X[2].shape)
    vf = rng.uniform(0.2, 1.0, (nz + 2, ny + 2, nx + 2))
    return X, vf


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("kind", ["spheres", "hyphae"])
def test_deposition_matches_oracle(bmx, oracle, gpu, scheme, kind):
    col = spheres() if kind == "spheres" else hyphae()
    ctx, grecs, g = setup(bmx, oracle, col)
    if kind == "hyphae":
        ctx.init_bonds()
    ctx.evolve(col.dt, 7)            # deposit at moved positions
    vf = ctx.volume_fraction(scheme)
    rhs = ctx.txfr_fluid(scheme)
    cnt = ctx.count_per_cell()
    ctx.sync()
    ovf, out = oracle.volume_fraction(grecs, g, scheme)
    assert out == 0
    orhs, _ = oracle.deposit(grecs, g, scheme, 31, 3)
    ocnt, _ = oracle.count_per_cell(grecs, g)
    assert vf.min() < 0.999 and np.abs(rhs).max() > 0
    assert np.array_equal(vf, ovf)
    assert np.array_equal(rhs, orhs)
    assert np.array_equal(cnt, ocnt)
    ctx.close()


def test_deposition_into_device_memory(bmx, oracle, gpu):
    col = spheres(seed=4)
    ctx, grecs, g = setup(bmx, oracle, col)
    nx, ny, nz = col.n_cell
    out = HipArray((nz, ny, nx))
    ctx.volume_fraction(1, out=out)
    ctx.sync()
    ovf, _ = oracle.volume_fraction(grecs, g, 1)
    assert np.array_equal(out.download(), ovf)
    out.free()
    ctx.close()


def test_particle_outside_the_mesh_is_an_error(bmx, oracle, gpu):
    col = spheres(seed=5)
    col.recs["pos"][17, 2] = col.prob_hi[2] + 1e-4   # above a wall-bounded top
    ctx, _, _ = setup(bmx, oracle, col)
    with pytest.raises(bmx.BMXError) as e:
        ctx.volume_fraction(0)
    assert e.value.code == 11
    ctx.close()


@pytest.mark.parametrize("cnc,vfs", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("kind", ["spheres", "hyphae"])
def test_mesh_to_particle_matches_oracle(bmx, oracle, gpu, kind, cnc, vfs):
    col = spheres() if kind == "spheres" else hyphae()
    mode = 0 if kind == "spheres" else 1
    ctx, grecs, g = setup(bmx, oracle, col, mode)
    if kind == "hyphae":
        ctx.init_bonds()
    ctx.evolve(col.dt, 3)
    ctx.sync()
    orecs = grecs.copy()
    X, vf = mesh_inputs(col)
    ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4, cnc, vfs)
    ctx.sync()
    err = oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, cnc, vfs)
    assert err == 0
    assert np.abs(grecs["rdata"][:, 5] - col.recs["rdata"][:, 5]).max() > 0  # grew
    assert np.array_equal(grecs["rdata"], orecs["rdata"])
    assert np.array_equal(grecs["idata"], orecs["idata"])
    grown = orecs.copy()
    # the mechanics go on with the new sizes
    if kind == "spheres":
        st = oracle.State(orecs, oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic,
                                                  ctx.bin_dims()),
                          col.fpar, oracle.NBR_PAIR, col.skin, 10)
    else:
        st = oracle.State(orecs, g, col.fpar, oracle.NBR_STENCIL,
                          1.1 * (2 * 2e-4 + 1e-2), 1)
    ctx.evolve(col.dt, 5)
    st.evolve(col.dt, 5)
    assert np.array_equal(ctx.positions(), orecs["pos"])
    if kind == "spheres":
        # The rows of the rebuild after the growth hold every pair within the
        # grown reaches: the bins are sized from the largest radius the call
        # reports, and one too small would leave pairs out of the stencil.
        og, ig = ctx.neighbors()
        ob, ib = oracle.build_neighbors(grown, g, oracle.NBR_PAIR, col.skin, brute=True,
                                        W_A=col.fpar[0])
        assert ob[-1] > len(grown)
        for i in range(len(grown)):
            assert sorted(ig[og[i]:og[i + 1]]) == list(ib[ob[i]:ob[i + 1]])
    ctx.close()


@pytest.mark.parametrize("kind", ["spheres", "hyphae"])
def test_chemistry_sequence_without_sync(bmx, oracle, gpu, kind):
    """Steps, exchange (segments), mesh -> particle, the deposits that read
    what it wrote, mesh -> particle again, more steps (the lists rebuilt
    for the new sizes, the device chemistry moved with the new order), and
    mesh -> particle once more, with no sync in between: each call sees the
    device chemistry state and the bond state the steps left."""
    col = spheres(seed=12) if kind == "spheres" else hyphae(seed=19)
    mode = 0 if kind == "spheres" else 1
    ctx, grecs, g = setup(bmx, oracle, col, mode)
    orecs = col.recs.copy()
    hood = 1.1 * (2 * 2e-4 + 1e-2)
    xpar = [2e-5 * 50, 2e-5 * 70, 2e-5 * 30]

    def oracle_state(bins):
        if kind == "hyphae":
            return oracle.State(orecs, g, col.fpar, oracle.NBR_STENCIL, hood, 1)
        return oracle.State(orecs, oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic,
                                                    bins),
                            col.fpar, oracle.NBR_PAIR, col.skin, 10)
    st = oracle_state(ctx.bin_dims())
    if kind == "hyphae":
        ctx.init_bonds()
        st.init_bonds()
    ctx.evolve(col.dt, 3)
    st.evolve(col.dt, 3)
    if kind == "hyphae":
        ctx.exchange(xpar, 1.0)
        st.exchange(xpar, 1.0)
    X, vf = mesh_inputs(col, seed=13)
    ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4, 1, 1)
    rhs = ctx.txfr_fluid(1)
    vfn = ctx.volume_fraction(0)
    ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4, 0, 0)
    ctx.evolve(col.dt, 3)
    bins = ctx.bin_dims()           # of these steps (the next growth changes them)
    if kind == "hyphae":
        ctx.exchange(xpar, 1.0)
    ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4, 1, 0)
    ctx.sync()
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 1, 1) == 0
    orhs, _ = oracle.deposit(orecs, g, 1, 31, 3)
    ovf, _ = oracle.volume_fraction(orecs, g, 0)
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 0, 0) == 0
    st.close()
    st = oracle_state(bins)         # the grown sizes: rows rebuilt, as on the GPU
    st.evolve(col.dt, 3)
    if kind == "hyphae":
        st.exchange(xpar, 1.0)
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 1, 0) == 0
    assert np.abs(rhs).max() > 0
    assert np.array_equal(rhs, orhs)
    assert np.array_equal(vfn, ovf)
    assert np.array_equal(grecs["pos"], orecs["pos"])
    assert np.array_equal(grecs["rdata"], orecs["rdata"])
    assert np.array_equal(grecs["idata"], orecs["idata"])
    st.close()
    ctx.close()


@pytest.mark.parametrize("tune,guard", [(False, False), (False, True), (True, False)])
def test_chemistry_state_across_guarded_steps_and_autotune(bmx, oracle, gpu, tune, guard):
    """The passes that snapshot the sorted arrays (guarded steps, the
    autotuner's trials) write the device chemistry back first: mesh ->
    particle, autotune, guarded steps, mesh -> particle, then one sync,
    against the oracle."""
    col = spheres(seed=21)
    ctx, grecs, g = setup(bmx, oracle, col)
    orecs = col.recs.copy()
    X, vf = mesh_inputs(col, seed=22)
    ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4, 0, 0)
    if tune:
        ctx.autotune(col.dt, trial_steps=10)
    ctx.set_guard(guard)
    ctx.evolve(col.dt, 12)
    bins = ctx.bin_dims()           # of these steps (the next growth changes them)
    ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4, 1, 1)
    ctx.sync()
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 0, 0) == 0
    st = oracle.State(orecs, oracle.make_geom(col.prob_lo, col.prob_hi, col.periodic, bins),
                      col.fpar, oracle.NBR_PAIR, col.skin, 10)
    st.evolve(col.dt, 12)
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 1, 1) == 0
    assert np.array_equal(grecs["pos"], orecs["pos"])
    bad = np.where((grecs["rdata"] != orecs["rdata"]).any(axis=0))[0]
    rows = np.where((grecs["rdata"] != orecs["rdata"]).any(axis=1))[0]
    assert bad.size == 0, f"rdata columns {bad.tolist()} differ in {rows.size} records"
    st.close()
    ctx.close()


def test_mesh_to_particle_from_device_memory(bmx, oracle, gpu):
    col = spheres(seed=8)
    ctx, grecs, g = setup(bmx, oracle, col)
    ctx.sync()
    orecs = grecs.copy()
    X, vf = mesh_inputs(col, seed=9)
    dX, dvf = HipArray.upload(X), HipArray.upload(vf)
    ctx.txfr_particle(CHEMPAR, dX, dvf, 0.25, 4, 1, 1)
    ctx.sync()
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 1, 1) == 0
    assert np.array_equal(grecs["rdata"], orecs["rdata"])
    dX.free()
    dvf.free()
    ctx.close()


def test_device_meshes_through_dlpack(bmx, oracle, gpu):
    """Device meshes handed over through DLPack (kDLROCM) only: the binding
    reads address, size and dtype from the capsule."""
    col = spheres(seed=10)
    ctx, grecs, g = setup(bmx, oracle, col)
    ctx.sync()
    orecs = grecs.copy()
    X, vf = mesh_inputs(col, seed=11)
    dX, dvf = HipArray.upload(X), HipArray.upload(vf)
    nx, ny, nz = col.n_cell
    out = HipArray((nz, ny, nx))
    views = [a.dlpack() for a in (dX, dvf, out)]
    for a, v in zip((dX, dvf, out), views):
        assert not hasattr(v, "__cuda_array_interface__")
        assert bmx._device_view(v) == (a.ptr.value, int(np.prod(a.shape)))
    ctx.volume_fraction(1, out=views[2])
    ovf, _ = oracle.volume_fraction(grecs, g, 1)
    ctx.txfr_particle(CHEMPAR, views[0], views[1], 0.25, 4, 1, 1)
    ctx.sync()
    assert np.array_equal(out.download(), ovf)
    assert oracle.mesh_to_particle(orecs, g, X, vf, CHEMPAR, 0.25, 4, 1, 1) == 0
    assert np.array_equal(grecs["rdata"], orecs["rdata"])
    for a in (dX, dvf, out):
        a.free()
    ctx.close()


def test_negative_fluid_concentration_is_an_error(bmx, oracle, gpu):
    col = spheres(seed=6)
    ctx, _, _ = setup(bmx, oracle, col)
    X, vf = mesh_inputs(col)
    X[0] = -1e-3
    with pytest.raises(bmx.BMXError) as e:
        ctx.txfr_particle(CHEMPAR, X, vf, 0.25, 4)
    assert e.value.code == 12
    ctx.close()
