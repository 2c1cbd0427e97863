"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, refuses GPU work without a device instead of falling back,
and its host utilities (the InitParticlesAscii reader, the writer, the
synthetic generators) behave."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(bmx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(bmx):
    lib = bmx.load_library()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.bmx_abi_version() == 1


def test_default_layout_matches_record_dtype(bmx):
    L = bmx.load_library().bmx_default_layout()
    dt = bmx.RECORD_DTYPE
    assert L.stride_bytes == dt.itemsize == 448
    assert (L.pos_offset, L.rdata_offset, L.id_offset, L.cpu_offset, L.idata_offset) == \
        tuple(dt.fields[k][1] for k in ("pos", "rdata", "id", "cpu", "idata"))


def test_no_device_is_an_error_not_a_fallback(bmx):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(bmx.BMXError) as ei:
        bmx.Context(0)
    assert ei.value.code == 10


def test_reads_reference_input_decks(bmx):
    """The 27-field files shipped with the reference's example decks."""
    recs = bmx.read_particles_ascii(os.path.join(GOLDEN, "ref_inputs", "fungi_init_cfg.dat"))
    assert len(recs) == 1
    r = recs[0]
    assert tuple(r["pos"]) == (0.06401, 0.06401, 0.04820)
    assert r["rdata"][bmx.R.radius] == 2.5e-4 and r["rdata"][bmx.R.c_length] == 3.0e-3
    assert r["rdata"][bmx.R.theta] == 1.5708 and r["rdata"][bmx.R.area] == 5.105e-6
    assert r["idata"][bmx.I.cell_type] == bmx.FUNGI
    assert r["rdata"][bmx.R.bond_scale] == 1.0 and r["idata"][bmx.I.fuse_tip] == 0
    assert r["id"] == 1 and r["idata"][bmx.I.id] == 1
    recs = bmx.read_particles_ascii(
        os.path.join(GOLDEN, "ref_inputs", "nlev_large_abc_real_input.dat"))
    assert recs[0]["idata"][bmx.I.cell_type] == bmx.YEAST
    assert recs[0]["rdata"][bmx.R.vol] == 5.3e-13
    w = bmx.read_particles_ascii(
        os.path.join(GOLDEN, "ref_inputs", "wall_test_fungi_init_cfg.dat"))
    assert w[0]["rdata"][bmx.R.theta] == 0.7854 and w[0]["rdata"][bmx.R.phi] == 1.5708


def test_ascii_round_trip(bmx, tmp_path):
    recs, lo, hi = bmx.generate_spheres(300, 5e-5, 7e-5, 0.2, 11)
    path = str(tmp_path / "p.dat")
    bmx.write_particles_ascii(path, recs)
    back = bmx.read_particles_ascii(path)
    assert np.array_equal(back["pos"], recs["pos"])
    assert np.array_equal(back["rdata"][:, :23], recs["rdata"][:, :23])
    assert np.array_equal(back["idata"][:, 5], recs["idata"][:, 5])


def test_ascii_errors(bmx, tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("3\n1 2 3\n")
    with pytest.raises(bmx.BMXError) as ei:
        bmx.read_particles_ascii(str(bad))
    assert ei.value.code == 8
    with pytest.raises(bmx.BMXError):
        bmx.read_particles_ascii(str(tmp_path / "missing.dat"))
    empty = tmp_path / "empty.dat"
    empty.write_text("")
    with pytest.raises(bmx.BMXError):
        bmx.read_particles_ascii(str(empty))


def test_sphere_generator(bmx):
    n = 2000
    recs, lo, hi = bmx.generate_spheres(n, 5e-5, 7e-5, 0.2, 21893)
    r = recs["rdata"][:, 0]
    assert r.min() >= 5e-5 and r.max() <= 7e-5
    vol = np.prod(hi - lo)
    phi = (4 / 3 * np.pi * r ** 3).sum() / vol
    assert abs(phi - 0.2) < 0.01
    # box of the 2,000-sphere workload: 2:2:1, ~2.65e-3 cm
    assert hi[0] == pytest.approx(2.65e-3, rel=0.01) and hi[2] == pytest.approx(hi[0] / 2)
    # random sequential insertion: no overlaps (periodic in x/y)
    p = recs["pos"]
    L = hi - lo
    d = p[:, None, :] - p[None, :, :]
    d[..., :2] -= L[:2] * np.round(d[..., :2] / L[:2])
    dist = np.sqrt((d ** 2).sum(-1)) + np.eye(n) * 1e9
    assert (dist >= (r[:, None] + r[None, :]) * (1 - 1e-12)).all()
    assert (p[:, 2] >= r).all() and (p[:, 2] <= hi[2] - r).all()
    again, _, _ = bmx.generate_spheres(n, 5e-5, 7e-5, 0.2, 21893)
    assert np.array_equal(again, recs)


def test_hyphae_generator(bmx):
    lo, hi = np.zeros(3), np.array([0.1, 0.1, 0.02])
    recs = bmx.generate_hyphae(2000, 20, 3e-3, 2e-4, 0.05, 30, 60, 5, lo, hi, 0.0)
    assert (recs["idata"][:, bmx.I.cell_type] == bmx.FUNGI).all()
    assert (recs["rdata"][:, bmx.R.c_length] == 3e-3).all()
    # every non-root segment starts where an earlier segment ends
    th, ph = recs["rdata"][:, 2], recs["rdata"][:, 3]
    n = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    ends = recs["pos"] + 1.5e-3 * n
    starts = recs["pos"] - 1.5e-3 * n
    L = hi - lo
    linked = 0
    for i in range(20, 2000, 97):
        d = starts[i] - ends[:i]
        d[:, :2] -= L[:2] * np.round(d[:, :2] / L[:2])
        linked += np.linalg.norm(d, axis=1).min() < 1e-9
    assert linked == len(range(20, 2000, 97))


# ---------------------------------------------------------- checkpoints --
CKPT = os.path.join(GOLDEN, "ref_checkpoints")


def test_shipped_checkpoint_headers(bmx):
    a = bmx.checkpoint_info(os.path.join(CKPT, "rst_test_chk20000"))
    assert (a["version"], a["n_real"], a["n_int"], a["finest_level"], a["n_particles"]) == \
        (20, 31, 5, 1, 238)
    b = bmx.checkpoint_info(os.path.join(CKPT, "fungi_chk357900"))
    assert (b["version"], b["n_real"], b["n_int"], b["finest_level"], b["n_particles"]) == \
        (21, 33, 33, 2, 410)


def test_read_shipped_sphere_checkpoint(bmx):
    recs = bmx.read_checkpoint(os.path.join(CKPT, "rst_test_chk20000"))
    assert len(recs) == 238
    r = recs["rdata"][:, 0]
    assert r.min() > 4e-5 and r.max() < 7e-5
    # a sphere's area and volume are those of its radius
    np.testing.assert_allclose(recs["rdata"][:, 4], 4 * np.pi * r ** 2, rtol=1e-12)
    np.testing.assert_allclose(recs["rdata"][:, 5], 4 / 3 * np.pi * r ** 3, rtol=1e-12)
    assert (recs["idata"][:, 5] == bmx.YEAST).all()
    assert (recs["idata"][:, 0:5] == [3, 0, 9, 0, 6]).all()     # chemistry header
    ids = list(zip(recs["id"], recs["cpu"]))
    assert len(set(ids)) == len(ids)
    lo, hi = np.zeros(3), np.full(3, 0.0064)                    # the rst_test deck
    assert ((recs["pos"] >= lo) & (recs["pos"] < hi)).all()


def test_read_shipped_segment_checkpoint(bmx):
    recs = bmx.read_checkpoint(os.path.join(CKPT, "fungi_chk357900"))
    assert len(recs) == 410
    assert (recs["idata"][:, 5] == bmx.FUNGI).all()
    assert np.allclose(recs["rdata"][:, 0], 2.5e-4)
    # segment lengths up to chem_species.max_seg_length (tips still growing)
    cl = recs["rdata"][:, 1]
    assert cl.min() > 0 and cl.max() <= 35e-4 and np.isclose(cl, 3e-3).sum() > 200
    # the packed id words of Two_Dot_One decode to the ids the record holds
    # in its own int components
    assert (recs["idata"][:, 33] == recs["id"]).all()
    assert (recs["idata"][:, 34] == recs["cpu"]).all()
    # every bond names a segment of the file (id, cpu), and the bonds are
    # reciprocal but for 3 of 999 (one-sided bonds the reference's own
    # fusion left behind)
    key = {(int(i), int(c)): k for k, (i, c) in enumerate(zip(recs["id"], recs["cpu"]))}
    nb = recs["idata"][:, 6]
    assert nb.max() <= 4 and (nb > 0).sum() == 410
    bonds = recip = 0
    for k in range(len(recs)):
        me = (int(recs["id"][k]), int(recs["cpu"][k]))
        for b in range(nb[k]):
            j = key[(int(recs["idata"][k, 7 + b]), int(recs["idata"][k, 11 + b]))]
            back = {(int(recs["idata"][j, 7 + c]), int(recs["idata"][j, 11 + c]))
                    for c in range(nb[j])}
            bonds += 1
            recip += me in back
    assert (bonds, recip) == (999, 996)


def test_checkpoint_round_trip(bmx, tmp_path):
    from boltzmannmfx_amd import workloads as W
    col = W.c3_hyphae(n=500, nroots=5, seed=3)
    recs = col.recs.copy()
    recs["rdata"][:, 25:34] = np.random.default_rng(4).normal(size=(500, 9))
    recs["idata"][:, 26:32] = -1
    geom = col.geometry()
    d = str(tmp_path / "particles")
    bmx.write_checkpoint(d, recs, geom, level=2, next_id=777)
    info = bmx.checkpoint_info(d)
    assert (info["version"], info["n_real"], info["n_int"], info["finest_level"],
            info["n_particles"], info["next_id"]) == (21, 34, 35, 2, 500, 777)
    head = open(os.path.join(d, "Header")).read().split()
    assert head[:3] == ["Version_Two_Dot_One_double", "3", "34"]
    back = bmx.read_checkpoint(d)

    def same(a, b):  # every field (the 4 tail padding bytes are not data)
        return all(np.array_equal(a[f], b[f]) for f in a.dtype.names)
    assert same(back, recs)
    # a shipped checkpoint through the current layout and back
    old = bmx.read_checkpoint(os.path.join(CKPT, "fungi_chk357900"))
    d2 = str(tmp_path / "p2")
    bmx.write_checkpoint(d2, old, geom)
    assert same(bmx.read_checkpoint(d2), old)


def test_checkpoint_errors(bmx, tmp_path):
    with pytest.raises(bmx.BMXError):
        bmx.checkpoint_info(str(tmp_path / "missing"))
    d = tmp_path / "bad"
    d.mkdir()
    (d / "Header").write_text("Version_Nine\n3\n")
    with pytest.raises(bmx.BMXError):
        bmx.checkpoint_info(str(d))
    # an older layout without a map
    with pytest.raises(bmx.BMXError):
        bmx.read_checkpoint(os.path.join(CKPT, "fungi_chk357900"), real_map=list(range(34)))


def test_dlpack_view_reads_address_and_size(bmx):
    """The DLPack reader that device meshes go through, on host arrays
    (DLPack device type 1): address, element count, dtype and contiguity."""
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    assert bmx._dlpack_view(a) is None  # a host array is not a device mesh
    addr, count = bmx._dlpack_view(a, device_type=1)
    assert addr == a.ctypes.data and count == 24
    sub = a[1]
    addr, count = bmx._dlpack_view(sub, device_type=1)
    assert addr == a.ctypes.data + 12 * 8 and count == 12
    with pytest.raises(AssertionError):
        bmx._dlpack_view(a[:, :, ::2], device_type=1)
    with pytest.raises(AssertionError):
        bmx._dlpack_view(a.astype(np.float32), device_type=1)
