#!/usr/bin/env python3
"""Benchmark of the MI355X particle-mechanics path.

Default workload: 1,000,000 FP64 spheres, radii uniform in [0.5e-4, 0.7e-4]
cm, random sequential insertion at volume fraction 0.3 in a 2:2:1 box
periodic in x/y with the growth surface at the bottom, cubic contact law
with adhesion band 0.1 * mean radius, dt = 0.25, Verlet neighbor list (skin
4e-6 cm) rebuilt every 10 steps.  A step is one pass of the hot path
(binning + list build when due, fused force + overdamped update) over all
particles, with the particles already resident in HBM.  The timed window
starts at a rebuild, so `--steps K` holds ceil(K / 10) list builds.

Prints one JSON line with the throughput (particle-steps/s), the roofline of
the dominant kernel (fused force + update: HIP events on the library's own
stream around each run of force launches between two rebuilds, in a second
window of the same length) and the CPU baseline: the oracle (the C
restatement of the reference's CPU path, 448-byte AoS records, stencil list
rebuilt every step) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "particle-steps/s at 1M FP64 spheres; batched ensemble speedup; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6  # MI355X FP64 vector


def algorithmic_bytes_per_particle(mean_row: float) -> float:
    """Compulsory HBM bytes per particle per force launch: read x,y,z,r
    (32 B), the row length (4 B) and mean_row 4-byte neighbor indices; write
    x,y,z,r (32 B).  Neighbor coordinates are gathered from cache.  The flat
    row layout also reads one owner byte per entry and two 4-byte bounds per
    64 particles; those are layout overhead and not counted."""
    return 32.0 + 4.0 + 4.0 * mean_row + 32.0


def spec_bytes_per_particle_step(mean_row: float, K: int) -> float:
    """Compulsory bytes per particle-step of a sphere step with rebuilds every
    K steps, as the workload's specification counts them: step 84 + 4(L+1)
    (x, y, z, r and type read; x, y, z and velocity written; row offset and
    L indices read), rebuild (112 + 4(L+1)) / K."""
    return 84.0 + 4.0 * (mean_row + 1.0) + (112.0 + 4.0 * (mean_row + 1.0)) / K


def host_cpu() -> dict:
    """Model name (lscpu / /proc/cpuinfo) and logical CPUs of the host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def cpu_baseline(col, budget_s: float = 12.0, min_steps_1: int = 3):
    """The oracle in the reference's configuration (AoS records, per-step
    stencil list over mesh-cell bins, per-particle loop), timed on the host:
    as many whole steps of the workload as fit in budget_s / 2 (at least
    min_steps_1 on one thread), on one thread and on all threads."""
    import numpy as np
    import oracle_bmx as orc

    L = col.prob_hi - col.prob_lo
    dx = 2.0e-4  # mesh cell no smaller than the largest r_lm (1.46e-4 cm)
    n_cell = [max(4, int(L[d] // dx)) for d in range(3)]
    hood = 1.1 * (2.0 * 7e-5 + 1e-2)  # DEM::neighborhood (bmx_chem.cpp:124)
    results = {}
    lib = orc.lib()
    threads_all = lib.orc_num_threads()
    for label, nt in (("1", 1), ("all", threads_all)):
        lib.orc_set_num_threads(nt)
        recs = col.recs.copy()
        g = orc.make_geom(col.prob_lo, col.prob_hi, col.periodic, n_cell)
        st = orc.State(recs, g, col.fpar, orc.NBR_STENCIL, hood, 1)
        steps = 0
        t0 = time.perf_counter()
        while True:
            st.evolve(col.dt, 1)
            steps += 1
            el = time.perf_counter() - t0
            if (el > budget_s / 2 and (nt > 1 or steps >= min_steps_1)) or steps >= 50:
                break
        results[label] = (len(recs) * steps / el, steps, el, nt)
        st.close()
    lib.orc_set_num_threads(threads_all)
    v1 = results["1"]
    va = results["all"]
    return {
        "value": va[0], "unit": "particle-steps/s", "cores": va[3], "kind": "port",
        "sample": f"{va[1]} steps of the {len(col.recs)}-sphere workload in the "
                  f"reference configuration (stencil list every step, AoS records) on "
                  f"{va[3]} OpenMP threads; 1 thread: {v1[0]:.4g} particle-steps/s "
                  f"over {v1[1]} steps",
        "value_1core": v1[0],
        "steps_1core": v1[1],
        "host": host_cpu(),
    }


def ensemble_bench(bmx, W, steps: int, seq: bool = True):
    """256 colonies (cell counts log-uniform in [500, 50000], volume fraction
    uniform in [0.1, 0.35], each in its own periodic box) advanced `steps`
    steps by one captured-graph launch after autotuning, against the same
    colonies run one after another through the plain per-step path."""
    import numpy as np
    cols = W.c4_ensemble()
    recs, starts, geoms = W.concat_colonies(cols)
    n = len(recs)
    K = cols[0].rebuild_every
    ctx = bmx.Context(0)
    ctx.set_force_params(cols[0].fpar)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=cols[0].skin,
                            max_neighbors=64)
    ctx.bind_colonies(recs, starts, geoms)
    ctx.evolve(W.DT, K)                      # into the steady rebuild phase
    # trials of 4 rebuild intervals from that state (2 intervals from the
    # freshly bound records picked launch shapes that ran slower later)
    t0 = time.perf_counter()
    tune = ctx.autotune(W.DT, 4 * K)
    tune["seconds"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.evolve_graph(W.DT, steps)            # capture + first replay
    capture_s = time.perf_counter() - t0
    # The replay right after a capture has measured up to 25 % slower per
    # step than later replays of the same graph (a 1000-step graph: 102.9
    # and 88.0 us against 80.5 and 77.8): one more untimed replay first.
    ctx.evolve_graph(W.DT, steps)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.evolve_graph(W.DT, steps)            # one graph launch, timed
    batched_s = time.perf_counter() - t0
    ctx.close()
    out = {
        "workload": f"{len(cols)} sphere colonies, {n} spheres, sizes "
                    f"{min(len(c.recs) for c in cols)}..{max(len(c.recs) for c in cols)}, "
                    f"phi U[0.1,0.35], {steps} steps, Verlet list every {K} steps",
        "colonies": len(cols), "particles": n, "steps": steps,
        "batched_particle_steps_per_s": n * steps / batched_s,
        "batched_ms_per_step": batched_s * 1e3 / steps,
        "graph_launches_per_call": 1,
        "capture_and_first_run_s": capture_s,
        "tuning": tune,
    }
    if seq:
        total = 0.0
        for col in cols:
            c = bmx.Context(0)
            c.set_force_params(col.fpar)
            c.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
            c.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=col.skin,
                                  max_neighbors=64)
            c.bind(col.recs.copy())
            c.synchronize()
            t0 = time.perf_counter()
            c.evolve(W.DT, steps)
            total += time.perf_counter() - t0
            c.close()
        out["sequential_particle_steps_per_s"] = n * steps / total
        out["sequential_s"] = total
        out["speedup_vs_sequential"] = total / batched_s
    return out


def adjoint_bench(bmx, W, n: int = 100_000, steps: int = 200):
    """Checkpointed adjoint (checkpoints every 20 steps) of `steps` sphere
    steps on n spheres at volume fraction 0.3, against the plain forward
    evolve of the same steps."""
    import numpy as np
    col = W.sphere_colony(n, 0.3, 11, W_A=6e-6, overlap_free=False, skin=5e-5)
    target = col.recs["pos"] + np.random.default_rng(8).normal(size=(n, 3)) * 1e-5

    def ctx_for():
        c = bmx.Context(0)
        c.set_force_params(col.fpar)
        c.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
        c.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=10, skin=col.skin)
        c.bind(col.recs.copy())
        return c

    c = ctx_for()
    c.adjoint(W.DT, 20, target)                          # warm up
    c.close()
    c = ctx_for()
    c.synchronize()
    t0 = time.perf_counter()
    c.evolve(W.DT, steps)
    fwd_s = time.perf_counter() - t0
    c.close()
    c = ctx_for()
    c.synchronize()
    t0 = time.perf_counter()
    loss, g3, gr, tm = c.adjoint(W.DT, steps, target, checkpoint_every=20)
    adj_s = time.perf_counter() - t0
    c.close()
    return {
        "workload": f"{n} spheres, phi=0.3, {steps} steps, checkpoints every 20 steps, "
                    f"loss = mean squared distance to a perturbed target",
        "plain_forward_ms": fwd_s * 1e3,
        "adjoint_forward_ms": tm.forward_ms,
        "adjoint_backward_ms": tm.backward_ms,
        "adjoint_call_ms": adj_s * 1e3,
        "backward_over_forward": tm.backward_ms / (fwd_s * 1e3),
        "adjoint_call_over_forward": adj_s / fwd_s,
    }


def rods_bench(bmx, W, steps: int = 500, cpu_steps: int = 10, block: int = 64):
    """200,000 hyphal segments (2,000 roots, branch probability 0.05 per
    segment at 30-60 degrees, 30e-4 cm long, radius 2e-4 cm, bond springs
    k = 1e-2) on the growth surface, in the reference's policy: bins are the
    level mesh cells (2e-3 cm) and the +-1 stencil list is rebuilt every
    step; InitBonds first, then `steps` timed steps."""
    col = W.c3_hyphae()
    hood = 1.1 * (2.0 * 2.0e-4 + 1.0e-2)  # DEM::neighborhood (bmx_chem.cpp:124)
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    ctx.set_neighbor_params(mode=bmx.NBR_STENCIL, rebuild_every=1, neighborhood=hood,
                            max_neighbors=128)
    ctx.set_launch_params(block, 1)
    ctx.bind(col.recs.copy())
    ctx.init_bonds()
    ctx.evolve(col.dt, 10)
    ctx.set_timing(True)
    ctx.reset_timing()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.evolve(col.dt, steps)
    secs = time.perf_counter() - t0
    tm = ctx.timing()
    ctx.close()
    n = len(col.recs)
    cpu = None
    if cpu_steps > 0:
        # the oracle on the same segments (InitBonds, then whole steps), all
        # host threads
        import oracle_bmx as orc
        recs = col.recs.copy()
        g = orc.make_geom(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
        st = orc.State(recs, g, col.fpar, orc.NBR_STENCIL, hood, 1)
        st.init_bonds()
        t0 = time.perf_counter()
        st.evolve(col.dt, cpu_steps)
        el = time.perf_counter() - t0
        st.close()
        nt = orc.lib().orc_num_threads()
        cpu = {"value": n * cpu_steps / el, "unit": "particle-steps/s", "cores": nt,
               "kind": "port", "sample": f"{cpu_steps} steps of the {n}-segment workload "
                                         f"on {nt} OpenMP threads"}
    force_ms = tm.force_ms / max(tm.force_launches, 1)
    roof = None
    pmc_path = os.path.join(ROOT, "profiles", "force_general_200k_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            prof = json.load(f)
        flop = prof.get("fp64_flop_per_launch")
        if flop:
            # algorithmic FP64 operations counted on the oracle over this
            # workload (the kernel's formulation; tests/test_oracle_flops.py)
            ach = flop / (force_ms * 1e-3) / 1e12
            ref_flop = prof.get("fp64_flop_per_launch_reference_formulation")
            cnt = prof.get("counters_per_launch", {})
            roof = {"bound": "fp64-valu", "achieved": ach, "peak": FP64_PEAK_TFS,
                    "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFS,
                    "traffic": None,
                    "flop_per_launch": flop,
                    "interacting_pairs_per_launch": prof.get("interacting_pairs_per_launch"),
                    "reference_formulation_flop_per_launch": ref_flop,
                    "reference_formulation_equivalent_tflops":
                        ref_flop / (force_ms * 1e-3) / 1e12 if ref_flop else None,
                    "valu_utilization_pct": cnt.get("VALUUtilization"),
                    "valu_busy_pct": cnt.get("VALUBusy")}
    return {
        "workload": f"{n} hyphal segments (2,000 roots), mesh bins {col.n_cell}, "
                    f"reference stencil list every step, {steps} steps",
        "cpu_baseline": cpu,
        "particle_steps_per_s": n * steps / secs,
        "ms_per_step": secs * 1e3 / steps,
        "force_kernel": "k_force_general_flat",
        "force_kernel_avg_ms": force_ms,
        "rebuild_avg_ms": tm.rebuild_ms / max(tm.rebuilds, 1),
        "mean_row": tm.neighbor_entries / n,
        "bound": "fp64-valu",
        "roofline": roof,
    }


# exec/fungi/input_fungi: k1 k2 k3 kr1 kr2 kr3 kg kv, max_seg_radius, max_seg_length
MESH_CHEMPAR = [2.0e-5, 4.0, 2.0e-5, 2.0e-5, 6.0e-4, 2.0e-5, 4.0, 4.0e4, 2.5e-4, 35.0e-4]


def mesh_bench(bmx, W, n: int = 1_000_000, n_cell=(64, 64, 32), calls: int = 20,
               cpu: bool = True):
    """Particle <-> mesh coupling of the 1,000,000-sphere colony on a
    64 x 64 x 32 level mesh (about 7.6 spheres per cell), after 10
    mechanics steps: the volume fraction (one-to-one), the fluid increments
    (one-to-one) and the mesh -> particle chemistry (4 substeps, the fungi
    deck's rates), each from and into device memory, `calls` calls each
    after one warm-up."""
    import ctypes as C
    import numpy as np
    hip = _hip_runtime()
    col = W.c2_large(n)
    col.n_cell = tuple(n_cell)
    r = col.recs["rdata"]
    rng = np.random.default_rng(11)
    r[:, 5] = 4.0 / 3.0 * np.pi * r[:, 0] ** 3
    r[:, 4] = 4.0 * np.pi * r[:, 0] ** 2
    r[:, 25:28] = rng.uniform(0.0, 2e-5, (n, 3))
    nx, ny, nz = n_cell
    ng = (nx + 2) * (ny + 2) * (nz + 2)
    X = np.empty((3, ng))
    X[0] = rng.uniform(1.0e-5, 3.0e-5, ng)
    X[1] = rng.uniform(0.0, 1.0e-7, ng)
    X[2] = rng.uniform(1.0e-6, 3.0e-6, ng)
    vfin = rng.uniform(0.2, 1.0, ng)

    class Dev:
        def __init__(self, a):
            a = np.ascontiguousarray(a, np.float64)
            self.shape, self.p = a.shape, C.c_void_p()
            assert hip.hipMalloc(C.byref(self.p), C.c_size_t(a.nbytes)) == 0
            assert hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p),
                                 C.c_size_t(a.nbytes), 1) == 0

        @property
        def __cuda_array_interface__(self):
            return {"shape": self.shape, "typestr": "<f8", "data": (self.p.value, False),
                    "version": 2, "strides": None}

    dX, dvf = Dev(X), Dev(vfin)
    dout = Dev(np.zeros(3 * nx * ny * nz))
    dvfo = Dev(np.zeros(nx * ny * nz))
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=col.rebuild_every,
                            skin=col.skin)
    recs = col.recs.copy()
    ctx.bind(recs)
    ctx.evolve(col.dt, 10)   # the particles in the order of a rebuild, as in a run
    dt = 1.0e-3
    res = {}
    for name, fn in (("volume_fraction", lambda: ctx.volume_fraction(0, out=dvfo)),
                     ("txfr_fluid", lambda: ctx.txfr_fluid(0, out=dout)),
                     ("txfr_particle", lambda: ctx.txfr_particle(MESH_CHEMPAR, dX, dvf,
                                                                  dt, 4))):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            fn()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / calls
        res[name] = {"ms_per_call": ms, "particles_per_s": n / (ms * 1e-3)}
    ctx.close()
    for d in (dX, dvf, dout, dvfo):
        hip.hipFree(d.p)
    # HBM bytes a particle needs in the chemistry pass: sorted position and
    # rod (64), type and original index (8), the record fields read (radius,
    # length, area, vol, dadt, dvdt, species A-C, position, n_bnds: 80) and
    # written (those sizes, the gradient and rdata 25-33: 144), new radius,
    # length and area (24), its cell's start (8); the mesh is L2-resident
    bytes_pp = 64 + 8 + 80 + 144 + 24 + 8
    tp = res["txfr_particle"]
    tp["algorithmic_bytes_per_particle"] = bytes_pp
    tp["gbs_at_wall_time"] = bytes_pp * n / (tp["ms_per_call"] * 1e-3) / 1e9
    out = {"workload": f"{n} spheres, level mesh {tuple(n_cell)}, one-to-one, 4 substeps",
           **res}
    if cpu:
        # the oracle on the same records, one host thread, one call each
        import oracle_bmx as orc
        g = orc.make_geom(col.prob_lo, col.prob_hi, col.periodic, col.n_cell)
        orecs = col.recs.copy()
        orc.lib()
        t0 = time.perf_counter()
        orc.volume_fraction(orecs, g, 0)
        t_vf = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.mesh_to_particle(orecs, g, X.reshape(3, nz + 2, ny + 2, nx + 2),
                             vfin.reshape(nz + 2, ny + 2, nx + 2), MESH_CHEMPAR, dt, 4, 0, 0)
        t_mp = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "volume_fraction_particles_per_s": n / t_vf,
            "txfr_particle_particles_per_s": n / t_mp,
            "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"one call of each on the same {n} records, one thread"}
    return out


def dropin_bench(bmx, col, calls: int = 40) -> dict:
    """The reference's cadence: bmx::Evolve calls EvolveParticles once per
    step (src/timestepping/bmx_evolve.cpp:97), on particles that live in
    device memory (AMReX GPU build).  Per call, three ways: (a) bind the
    records again (copied), evolve 1 step, sync; (b) records bound once in
    place, evolve 1 step and sync per call; (c) bound once in place, evolve
    1 step per call, records synced every 10 calls; (d) as (c) with the
    calls enqueued without waiting (bmx_evolve_async), completed every 10
    calls.  Each from the state after a warm-up, timed over whole rebuild
    periods."""
    import ctypes as C
    import numpy as np
    hip = _hip_runtime()
    recs = col.recs.copy()
    nbytes = recs.nbytes
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dev, recs.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
    K = col.rebuild_every
    n = len(recs)
    out = {"workload": f"{n} spheres resident in device memory, {calls} calls of one "
                       f"step each, Verlet list every {K} steps"}

    def make():
        ctx = bmx.Context(0)
        ctx.set_force_params(col.fpar)
        ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
        ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K, skin=col.skin,
                                max_neighbors=64)
        return ctx

    try:
        ctx = make()
        for _ in range(3):                      # (a) warm-up
            ctx.bind_device(dev.value, n)
            ctx.evolve(col.dt, 1)
            ctx.sync()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            ctx.bind_device(dev.value, n)
            ctx.evolve(col.dt, 1)
            ctx.sync()
        out["rebind_evolve_sync_ms_per_call"] = (time.perf_counter() - t0) * 1e3 / calls
        ctx.close()

        ctx = make()
        ctx.bind_device(dev.value, n, in_place=True)
        for _ in range(2 * K):                  # (b) warm-up
            ctx.evolve(col.dt, 1)
            ctx.sync()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            ctx.evolve(col.dt, 1)
            ctx.sync()
        out["resident_evolve_sync_ms_per_call"] = (time.perf_counter() - t0) * 1e3 / calls
        ctx.synchronize()
        t0 = time.perf_counter()
        for k in range(calls):                  # (c)
            ctx.evolve(col.dt, 1)
            if k % 10 == 9:
                ctx.sync()
        out["resident_evolve_ms_per_call_sync_every_10"] = \
            (time.perf_counter() - t0) * 1e3 / calls
        ctx.synchronize()
        t0 = time.perf_counter()
        for k in range(calls):                  # (d)
            ctx.evolve_async(col.dt, 1)
            if k % 10 == 9:
                ctx.wait()
                ctx.sync()
        out["resident_async_ms_per_call_wait_sync_every_10"] = \
            (time.perf_counter() - t0) * 1e3 / calls
        ctx.close()
    finally:
        hip.hipFree(dev)
    return out


def _hip_runtime():
    """The HIP runtime the library links (/opt/rocm), not another copy a
    framework may have loaded under the same name."""
    import ctypes as C
    try:
        return C.CDLL("/opt/rocm/lib/libamdhip64.so")
    except OSError:
        return C.CDLL("libamdhip64.so")


def copy_probe(sizes_mib=(64, 1024), reps: int = 20) -> dict:
    """Device-to-device copy rate of this GPU (read + write bytes per second,
    wall clock over `reps` back-to-back copies): the HBM rate a streaming
    kernel can reach here, and at 64 MiB what the 256 MB Infinity Cache adds
    when the data fits in it."""
    import ctypes as C
    hip = _hip_runtime()
    out = {}
    for mib in sizes_mib:
        nbytes = mib << 20
        a, b = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
        assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
        for _ in range(3):
            assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0
        assert hip.hipDeviceSynchronize() == 0
        t0 = time.perf_counter()
        for _ in range(reps):
            assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
        assert hip.hipDeviceSynchronize() == 0
        el = time.perf_counter() - t0
        out[f"copy_{mib}MiB_GBs"] = 2.0 * nbytes * reps / el / 1e9
        hip.hipFree(a)
        hip.hipFree(b)
    return out


def calibration_copy(nbytes: int) -> None:
    """One device-to-device copy of nbytes (a plain copy kernel of known
    traffic: nbytes read, nbytes written) for scaling the PMC byte counters."""
    import ctypes as C
    hip = _hip_runtime()
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--skin", type=float, default=0.0,
                    help="Verlet skin in cm (default: the workload's)")
    ap.add_argument("--rebuild-every", type=int, default=0,
                    help="steps between list builds (default: the workload's, 10)")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of one graph replay")
    # one wave per block: no block-wide barriers in the list build or the
    # force pass (1M spheres, 200 steps: 31.5 us per step at 64 against 31.8,
    # 32.0 and 32.9 at 128, 256 and 512)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--remap", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-sequential", action="store_true",
                    help="skip the 256 one-colony runs the ensemble is compared with")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-rods", action="store_true")
    ap.add_argument("--rods-steps", type=int, default=500)
    ap.add_argument("--rods-only", action="store_true",
                    help="run only the 200,000-segment workload and print its "
                         "object (for profiling that kernel)")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the 100,000-sphere adjoint against its forward "
                         "and print its object")
    ap.add_argument("--mesh-only", action="store_true",
                    help="run only the particle <-> mesh calls on the 1,000,000 "
                         "spheres and print their object (for profiling)")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the one-step-per-call cadence legs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (the "
                         "reference kernel for scaling FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path",
              file=sys.stderr)
        return 2

    import numpy as np
    import boltzmannmfx_amd as bmx
    from boltzmannmfx_amd import workloads as W

    if args.calibrate_copy > 0:
        calibration_copy(args.calibrate_copy << 20)
    if args.rods_only:
        print(json.dumps(rods_bench(bmx, W, steps=args.rods_steps,
                                    cpu_steps=0 if args.no_cpu else 10,
                                    block=args.block)))
        return 0
    if args.adjoint_only:
        print(json.dumps(adjoint_bench(bmx, W)))
        return 0
    if args.mesh_only:
        print(json.dumps(mesh_bench(bmx, W, n=args.n, cpu=not args.no_cpu)))
        return 0

    probe = copy_probe()
    col = W.c2_large(args.n)
    if args.skin > 0:
        col.skin = args.skin
    if args.rebuild_every > 0:
        col.rebuild_every = args.rebuild_every
    K = col.rebuild_every
    ctx = bmx.Context(0)
    ctx.set_force_params(col.fpar)
    ctx.set_geometry(col.prob_lo, col.prob_hi, col.periodic)
    ctx.set_neighbor_params(mode=bmx.NBR_VERLET, rebuild_every=K,
                            skin=col.skin, max_neighbors=64)
    ctx.set_launch_params(args.block, args.remap)
    recs = col.recs.copy()
    ctx.bind(recs)
    # Warm-up: at least `warmup` steps, in whole rebuild periods (the first
    # step after a bind rebuilds, then every K steps), so the timed window
    # starts with a rebuild and holds ceil(steps / K) of them.
    warm = -(-args.warmup // K) * K
    ctx.evolve(col.dt, warm)
    # and whole periods more until the device has been busy for 0.3 s, so
    # the window does not run at idle clocks
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < 0.3:
        ctx.evolve(col.dt, 10 * K)
        warm += 10 * K
    # The timed steps replay one captured HIP graph (the steps' launches,
    # rebuilds included, submitted at once).  Its capture and a first replay
    # are warm-up; a window of whole rebuild periods starts at the same list
    # phase every time, so the timed call replays the captured graph.
    graph = args.steps % K == 0 and not args.eager
    if graph:
        # capture and two replays (the replay right after a capture can run
        # slower than later ones)
        for _ in range(2):
            ctx.evolve_graph(col.dt, args.steps)
        warm += 2 * args.steps
    ctx.synchronize()
    rb0 = ctx.timing().rebuilds_total
    t0 = time.perf_counter()
    if graph:
        ctx.evolve_graph(col.dt, args.steps)  # synchronizes, checks skin/overflow/NaN
    else:
        ctx.evolve(col.dt, args.steps)
    t1 = time.perf_counter()
    rebuilds = ctx.timing().rebuilds_total - rb0
    secs = t1 - t0
    n = len(recs)
    value = n * args.steps / secs
    ms_per_step = secs * 1e3 / args.steps

    # Kernel split, in eager launches (events cannot sit inside a graph
    # replay): HIP events on the library's stream around every rebuild and
    # around every run of force launches between two rebuilds, so the force
    # time per launch carries no per-launch event overhead and cannot exceed
    # a step.  At least 40 rebuild periods, starting at a rebuild.
    ev_steps = max(args.steps, 40 * K)
    ev_steps = -(-ev_steps // K) * K
    ctx.set_timing(True)
    ctx.reset_timing()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.evolve(col.dt, ev_steps)
    secs_timed = time.perf_counter() - t0
    tm = ctx.timing()
    ctx.set_timing(False)

    mean_row = tm.neighbor_entries / max(n, 1)
    per_launch_ms = tm.force_ms / max(tm.force_launches, 1)
    rebuild_ms = tm.rebuild_ms / max(tm.rebuilds, 1)
    alg_bytes = algorithmic_bytes_per_particle(mean_row) * n
    achieved = alg_bytes / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "force_sphere_1M_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "1M FP64 spheres, r~U[0.5e-4,0.7e-4] cm, phi=0.3, 2:2:1 box "
                        "periodic x/y with growth surface, cubic law W_A=6e-6, dt=0.25, "
                        f"Verlet list (skin {col.skin:g}) rebuilt every {K} steps",
            "n": n,
            "rebuild_every": K,
            "skin": col.skin,
            "mean_row": mean_row,
            "warmup_steps_run": warm,
            "launch": "one captured HIP graph per call" if graph else "eager launches",
            "rebuilds": rebuilds,
            "rebuilds_in_events_window": tm.rebuilds,
            "block": args.block,
            "xcd_remap": args.remap,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_force_sphere_flat",
            "kernel_avg_ms": per_launch_ms,
            "kernel_launches_timed": tm.force_launches,
            "kernel_share_of_step": per_launch_ms / ms_per_step,
            "rebuild_avg_ms": rebuild_ms,
            "events_window_steps": ev_steps,
            "events_window_ms_per_step": secs_timed * 1e3 / ev_steps,
            "device_ms_per_step": (tm.force_ms + tm.rebuild_ms) / ev_steps,
            "algorithmic_bytes_per_particle": algorithmic_bytes_per_particle(mean_row),
            "algorithmic_bytes_per_launch": alg_bytes,
            "spec_bytes_per_particle_step": spec_bytes_per_particle_step(mean_row, K),
            "copy_probe": probe,
        },
    }
    ctx.close()
    if not args.no_dropin:
        out["dropin"] = dropin_bench(bmx, col)
        print("# drop-in cadence done", file=sys.stderr, flush=True)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(bmx, W, args.ensemble_steps,
                                         seq=not args.no_sequential)
        print(f"# ensemble done: {out['ensemble'].get('speedup_vs_sequential')}",
              file=sys.stderr, flush=True)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(bmx, W)
        print("# adjoint done", file=sys.stderr, flush=True)
    if not args.no_rods:
        out["rods"] = rods_bench(bmx, W, steps=args.rods_steps,
                                 cpu_steps=0 if args.no_cpu else 10, block=args.block)
        print("# rods done", file=sys.stderr, flush=True)
    if not args.no_mesh:
        out["mesh"] = mesh_bench(bmx, W, cpu=not args.no_cpu)
        print("# mesh done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        cb = cpu_baseline(col, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
