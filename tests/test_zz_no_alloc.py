"""The allocation-free evolve (bmx_evolve_aos_noalloc) against the context
path it stands for: bind the same device records, evolve, sync.  Two calls
in a row, the second with fewer particles (the scratch keeps the first
call's contents), spheres under the Verlet policy and segments under the
reference's stencil policy: the records must match bit for bit.  Sizes past
the scratch's limits are refused before any GPU work."""
import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W
from conftest import HipArray

pytestmark = pytest.mark.gpu


def context_run(bmx, col, recs, nsteps, **nb):
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    ctx.set_neighbor_params(**nb)
    out = recs.copy()
    ctx.bind(out)
    ctx.evolve(col.dt, nsteps)
    ctx.sync()
    ctx.close()
    return out


def noalloc_run(bmx, col, recs, nsteps, **nb):
    d = HipArray.upload(recs.view(np.float64))
    try:
        bmx.evolve_aos_noalloc(d.ptr.value, len(recs), col.geometry(), col.fpar,
                               col.dt, nsteps, **nb)
        return np.ascontiguousarray(d.download()).view(bmx.RECORD_DTYPE).reshape(-1)
    finally:
        d.free()


def test_noalloc_matches_context_path(bmx, gpu):
    big = W.sphere_colony(20000, 0.3, 5, W_A=6e-6, skin=4e-6)
    small = W.sphere_colony(5000, 0.25, 6, W_A=6e-6, skin=4e-6)
    nb = dict(mode=bmx.NBR_VERLET, rebuild_every=10, skin=4e-6, max_neighbors=64)
    for col in (big, small):            # the second call is the smaller one
        print(f"noalloc: {len(col.recs)} spheres", flush=True)
        want = context_run(bmx, col, col.recs, 25, **nb)
        got = noalloc_run(bmx, col, col.recs, 25, **nb)
        assert np.array_equal(got, want)
        assert not np.array_equal(got["pos"], col.recs["pos"])   # it moved


def test_noalloc_segments_reference_policy(bmx, gpu):
    hy = W.c3_hyphae(n=3000, nroots=30, seed=11)
    hood = 1.1 * (2 * 5.47e-4 + 1e-2)
    nb = dict(mode=bmx.NBR_STENCIL, rebuild_every=1, neighborhood=hood, max_neighbors=64)
    print(f"noalloc: {len(hy.recs)} segments", flush=True)
    want = context_run(bmx, hy, hy.recs, 5, **nb)
    got = noalloc_run(bmx, hy, hy.recs, 5, **nb)
    assert np.array_equal(got, want)


def test_noalloc_limits(bmx, gpu):
    col = W.sphere_colony(1000, 0.3, 7, W_A=6e-6, skin=4e-6)
    d = HipArray.upload(col.recs.view(np.float64))
    try:
        for n, kw in ((bmx.BMX_NOALLOC_MAX_PARTICLES + 1, {}),
                      (1000, {"max_neighbors": 65}),
                      (1000, {"rebuild_every": 65})):
            with pytest.raises(bmx.BMXError) as e:
                bmx.evolve_aos_noalloc(d.ptr.value, n, col.geometry(), col.fpar,
                                       col.dt, 1, skin=4e-6, **kw)
            assert e.value.code == 1
        back = np.ascontiguousarray(d.download()).view(bmx.RECORD_DTYPE).reshape(-1)
        assert np.array_equal(back, col.recs)             # nothing ran
    finally:
        d.free()
