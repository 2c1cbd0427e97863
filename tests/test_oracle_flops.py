"""Algorithmic FP64 operation counts of the segment step, counted on the
oracle (tests/flops/flop_count.cpp compiles it with a counting double) over
the 200,000-segment workload after InitBonds, and checked against the
figures the segment roofline uses (profiles/force_general_200k_pmc.json).

Two formulations of the same arithmetic:
  * the reference's (K.H:277-993 per pair: the sines and cosines of both
    segments' angles, both segments' two-site mobility tensors and both
    sides' site velocities, every pair);
  * the segment kernel's: angles' sines and cosines and the mobility tensors
    once per segment, only particle i's side of each pair (the reference's j
    side lands in a copy and is dropped, PCI:188, 251-261).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from boltzmannmfx_amd import workloads as W
from conftest import ROOT


def counts(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.fail("g++ is needed to build the operation counter")
    exe = str(tmp_path / "flop_count")
    src = os.path.join(ROOT, "tests", "flops", "flop_count.cpp")
    subprocess.run([gxx, "-std=c++17", "-O2", "-ffp-contract=off", "-fpermissive", "-w",
                    "-o", exe, src], check=True)
    col = W.c3_hyphae()
    recs = str(tmp_path / "recs.bin")
    col.recs.tofile(recs)
    hood = 1.1 * (2.0 * 2.0e-4 + 1.0e-2)  # DEM::neighborhood (bmx_chem.cpp:124)
    args = [exe, recs, str(len(col.recs))]
    args += [repr(float(v)) for v in col.prob_lo] + [repr(float(v)) for v in col.prob_hi]
    args += [str(v) for v in col.periodic] + [str(v) for v in col.n_cell]
    args += [repr(float(v)) for v in col.fpar] + [repr(hood)]
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def formulations(c):
    trig_pair = c["trig_calls_per_pair"] // 2 * c["flop_sin_plus_cos"]
    per_particle = c["flop_accept"] + c["flop_wall"] + c["flop_update"]
    reference = c["flop_pair"] + per_particle
    kernel_pair = (c["flop_pair"] - c["pairs"] * (trig_pair + 2 * c["flop_mobility"] +
                                                 c["flop_site_velocities"]))
    kernel = (kernel_pair + per_particle + c["particles"] *
              (c["flop_mobility"] + 2 * c["flop_sin_plus_cos"]))
    return reference, kernel


def test_segment_step_operation_counts(tmp_path):
    c = counts(tmp_path)
    assert c["particles"] == 200_000 and c["pairs"] > 10 * c["particles"]
    # pieces counted on their own
    assert c["flop_sin_plus_cos"] > 20 and c["flop_mobility"] > 100
    reference, kernel = formulations(c)
    assert kernel < reference
    with open(os.path.join(ROOT, "profiles", "force_general_200k_pmc.json")) as f:
        prof = json.load(f)
    assert prof["interacting_pairs_per_launch"] == c["pairs"]
    assert prof["fp64_flop_per_launch_reference_formulation"] == reference
    assert prof["fp64_flop_per_launch"] == kernel
