"""Restart from the particle checkpoints the reference ships.

exec/rst_test/chk20000 (238 spheres) and exec/fungi/chk357900 (410 bonded
segments) are read into Particle<34,35> records and continued with their
decks' geometry (finest level mesh) and force parameters, in the
reference's per-step stencil policy, without InitBonds (a restart keeps its
bonds): EvolveParticles, then ParticleExchange, every step.  The oracle runs
the same steps from the same records and must give the same bits.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_decks import parse_deck

pytestmark = pytest.mark.gpu

# (checkpoint, deck); the 23,522-bacteria monolayer of
# exec/nlev_large_test/dbg/chk300000 lies on level 2 (the refined slab
# around the growth surface), whose 128 x 128 x 64 mesh sets its bins
CASES = [("rst_test_chk20000", "rst_test/input_real_nlev"),
         ("fungi_chk357900", "fungi/input_fungi"),
         ("nlev_large_chk300000", "nlev_large_test/input_real_nlev")]


def deck_setup(path):
    kv = parse_deck(path)

    def f(k, i=0, default=None):
        if k not in kv and default is not None:
            return default
        return float(kv[k][i])
    lev = int(kv.get("amr.max_level", ["0"])[0])
    n_cell = [int(kv["amr.n_cell"][d]) * 2 ** lev for d in range(3)]
    lo = [f("geometry.prob_lo", d) for d in range(3)]
    hi = [f("geometry.prob_hi", d) for d in range(3)]
    per = [int(kv["geometry.is_periodic"][d]) for d in range(3)]
    fpar = [f("cell_force.boundary_width"), f("cell_force.stiffness"),
            f("fluid.surface_location", default=0.0), f("cell_force.wall_boundary_width"),
            f("cell_force.wall_stiffness"), f("cell_force.gravity", default=0.0),
            f("cell_force.bond_strength", default=0.0), f("cell_force.bond_cutoff", default=0.0),
            f("cell_force.viscous_drag", default=0.0)]
    vol = f("chem_species.max_vol", default=0.0)
    # the rst_test deck predates the required cell_force.neighbor_width
    # (bmx_chem.cpp:121); the value of every other deck, 1.0e-2, stands in
    width = f("cell_force.neighbor_width", default=1.0e-2)
    hood = 1.1 * (2 * (3 * vol / (4 * math.pi)) ** (1 / 3) + width)
    xpar = [f("chem_species.mass_transfer_" + s, default=0.0) for s in "ABC"]
    return lo, hi, per, n_cell, fpar, hood, xpar, f("bmx.fixed_dt")


@pytest.mark.parametrize("ckpt,deck", CASES)
def test_restart_continues_like_the_oracle(bmx, oracle, gpu, ckpt, deck):
    lo, hi, per, n_cell, fpar, hood, xpar, dt = deck_setup(
        os.path.join(GOLDEN, "ref_decks", deck))
    path = os.path.join(GOLDEN, "ref_checkpoints", ckpt)
    # the particles' level (the checkpoint's finest) sets the mesh cells
    lev = bmx.checkpoint_info(path)["finest_level"]
    kv = parse_deck(os.path.join(GOLDEN, "ref_decks", deck))
    n_cell = [int(kv["amr.n_cell"][d]) * 2 ** lev for d in range(3)]
    recs = bmx.read_checkpoint(path)
    print(f"restart {ckpt}: {len(recs)} particles, level {lev}, mesh {n_cell}", flush=True)
    ctx = bmx.Context(0)
    ctx.set_force_params(fpar)
    ctx.set_geometry(lo, hi, per, n_cell)
    ctx.set_neighbor_params(mode=bmx.NBR_STENCIL, rebuild_every=1, neighborhood=hood,
                            max_neighbors=256)
    grecs = recs.copy()
    ctx.bind(grecs)
    orecs = recs.copy()
    st = oracle.State(orecs, oracle.make_geom(lo, hi, per, n_cell), np.array(fpar),
                      oracle.NBR_STENCIL, hood, 1)
    for _ in range(20):
        ctx.evolve(dt, 1)
        ctx.exchange(xpar, dt)
        st.evolve(dt, 1)
        st.exchange(xpar, dt)
    ctx.sync()
    moved = np.abs(grecs["pos"] - recs["pos"]).max()
    assert moved > 0
    assert np.array_equal(grecs["pos"], orecs["pos"])
    assert np.array_equal(grecs["rdata"][:, :31], orecs["rdata"][:, :31])
    assert np.array_equal(grecs["idata"], orecs["idata"])
    ctx.close()
