"""The reference's own example decks through the C++ host layer.

bmx_run reads a BMX input deck (exec/fungi, exec/wall_test,
exec/nlev_large_test, copied under tests/golden/ref_decks), builds
BMXParticleContainer on the finest level mesh with the reference's per-step
stencil list, calls InitParticlesAscii, InitBonds at time 0 and
EvolveParticles, split_particles, ParticleExchange and the fusion steps each
step, and writes the particles
back in the input format (which holds no species: those are compared by
test_gpu_exchange.py).  The same run on the oracle must give the same
bits.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

DECKS = os.path.join(GOLDEN, "ref_decks")
RUN = os.path.join(ROOT, "boltzmannmfx_amd", "bmx_run")


def parse_deck(path):
    kv = {}
    for line in open(path):
        line = line.split("#", 1)[0]
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        kv[k.strip()] = v.replace('"', " ").split()
    return kv


@pytest.mark.parametrize("deck,steps", [("fungi/input_fungi", 40),
                                        ("wall_test/input_fungi", 40),
                                        ("nlev_large_test/input_real_nlev", 40)])
def test_reference_deck_runs_match_oracle(bmx, oracle, gpu, tmp_path, deck, steps):
    path = os.path.join(DECKS, deck)
    out = str(tmp_path / "out.dat")
    res = subprocess.run([RUN, path, str(steps), out], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    assert res.stdout.count("Time per step") == steps
    # the deck's particles stay inside its refined region (the driver warns
    # otherwise: the reference would move them to a coarser level)
    assert "Warning" not in res.stdout

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
    # defaults of BMXCellInteraction::setParams (src/chemistry/bmx_cell_interaction.cpp:59-60)
    fpar = [f("cell_force.boundary_width"), f("cell_force.stiffness"),
            f("fluid.surface_location", default=0.0), f("cell_force.wall_boundary_width"),
            f("cell_force.wall_stiffness"), f("cell_force.gravity", default=0.0),
            f("cell_force.bond_strength"), f("cell_force.bond_cutoff"),
            f("cell_force.viscous_drag")]
    vol = f("chem_species.max_vol", default=0.0)
    hood = 1.1 * (2 * (3 * vol / (4 * math.pi)) ** (1 / 3) + f("cell_force.neighbor_width"))
    recs = bmx.read_particles_ascii(os.path.join(os.path.dirname(path),
                                                 kv["particles.input_file"][0]))
    # species start at fluid.init_conc_species (bmx_pc_init.cpp:87-91)
    conc = [float(v) for v in kv.get("fluid.init_conc_species", [])][:3]
    recs["rdata"][:, 25:25 + len(conc)] = conc
    # chem_species.mass_transfer_A/B/C (bmx_chem.cpp:100-105)
    xpar = [f("chem_species.mass_transfer_" + s, default=0.0) for s in "ABC"]
    # split_particles limits (bmx_chem_species_parms.cpp:55-66, bmx_chem.cpp:110)
    sp = [vol, f("chem_species.max_seg_length", default=0.0035),
          f("chem_species.max_seg_radius", default=0.00025),
          f("chem_species.branching_probability", default=0.0001),
          f("chem_species.splitting_probability", default=0.1), 0.2,
          f("chem_species.seg_split_length", default=0.0030)]
    # fusion parameters (bmx_chem.cpp:106-112)
    fus = [f("chem_species.fusion_probability", default=0.0),
           f("chem_species.max_fusion_separation", default=5.0e-4),
           f("chem_species.bond_scaling_increment", default=0.001),
           f("chem_species.max_seg_length", default=0.0035)]
    g = oracle.make_geom(lo, hi, per, n_cell)
    st = oracle.State(recs, g, np.array(fpar), oracle.NBR_STENCIL, hood, 1)
    st.init_bonds()
    dt = f("bmx.fixed_dt")
    next_id = int(recs["id"].max()) + 1
    for step in range(steps):  # EvolveParticles, split_particles, ParticleExchange
        st.evolve(dt, 1)
        grown = oracle.split_particles(recs, g, sp, 0, step, next_id, 0)
        if len(grown) > len(recs):
            st.close()
            next_id += len(grown) - len(recs)
            recs = grown
            st = oracle.State(recs, g, np.array(fpar), oracle.NBR_STENCIL, hood, 1)
        st.exchange(xpar, dt)
        _, fused = oracle.fusion(recs, g, fpar, oracle.NBR_STENCIL, hood, fus,
                                 0 ^ 0x66757369, step, next_id, 0)
        assert not fused     # one seed particle per deck: nothing to fuse into
    got = bmx.read_particles_ascii(out)
    assert len(got) == len(recs)
    assert np.array_equal(got["pos"], recs["pos"])
    assert np.array_equal(got["rdata"][:, :23], recs["rdata"][:, :23])
