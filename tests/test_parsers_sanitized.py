"""The two readers of untrusted files -- the InitParticlesAscii reader
(csrc/bmx_io.cpp; src/des/bmx_pc_init.cpp:17-132 in the reference) and the
particle checkpoint reader (csrc/bmx_checkpoint.cpp; the Restart format of
src/io/bmx_restart.cpp:104-107) -- built on the host with AddressSanitizer
and UndefinedBehaviorSanitizer and run over the shipped inputs and
checkpoints and over corrupted copies of them: truncations, inflated counts
and offsets, junk tokens.  Every run must end without a sanitizer report;
a corrupted file must give an error status or a clean read, never a read
outside its buffers."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CSRC = os.path.join(ROOT, "boltzmannmfx_amd", "csrc")
ENV = dict(os.environ,
           ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=24")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.fail("g++ is needed for the sanitized parser build")
    exe = str(tmp_path_factory.mktemp("san") / "parse_harness")
    subprocess.run([gxx, "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-o", exe, os.path.join(ROOT, "tests", "sanitize", "parse_harness.cpp"),
                    os.path.join(CSRC, "bmx_io.cpp"), os.path.join(CSRC, "bmx_checkpoint.cpp")],
                   check=True)
    return exe


def run(harness, *args):
    p = subprocess.run([harness, *map(str, args)], env=ENV, capture_output=True,
                       text=True, timeout=300)
    bad = ("AddressSanitizer" in p.stderr or "runtime error" in p.stderr or
           "LeakSanitizer" in p.stderr)
    assert p.returncode == 0 and not bad, f"{args}\n{p.stdout}\n{p.stderr[-3000:]}"
    return p.stdout


ASCII = sorted(os.listdir(os.path.join(GOLDEN, "ref_inputs")))


@pytest.mark.parametrize("name", ASCII)
def test_ascii_reader_on_shipped_and_corrupted_inputs(harness, tmp_path, name):
    src = os.path.join(GOLDEN, "ref_inputs", name)
    data = open(src, "rb").read()
    out = run(harness, "ascii", src, tmp_path / "w.dat")
    assert "ascii-reread status=0" in out
    first, _, rest = data.partition(b"\n")
    cases = {
        "empty": b"",
        "count_only": first,
        "inflated_count": b"%d\n" % (int(first) + 7) + rest,
        "huge_count": b"9223372036854775807\n" + rest,
        "negative_count": b"-3\n" + rest,
        "junk_token": data.replace(b" ", b" x", 1) if b" " in data else data + b"x",
        "nan_fields": data.replace(b"0.", b"nan", 3),
        "no_newlines": data.replace(b"\n", b" "),
        "binary": bytes(range(256)) * 4,
    }
    for frac in (0.01, 0.1, 0.37, 0.5, 0.93, 0.999):
        cases[f"cut_{frac}"] = data[: int(len(data) * frac)]
    for tag, blob in cases.items():
        f = tmp_path / f"{tag}.dat"
        f.write_bytes(blob)
        run(harness, "ascii", f, tmp_path / f"{tag}_w.dat")


CHECKPOINTS = sorted(os.listdir(os.path.join(GOLDEN, "ref_checkpoints")))


def maps_for(bmx, path):
    info = bmx.checkpoint_info(path)
    key = (info["n_real"], info["n_int"])
    if key == (34, 35):
        return []
    rm, im = bmx.LEGACY_CHECKPOINT_MAPS[key]
    return [",".join(str(int(v)) for v in rm), ",".join(str(int(v)) for v in im)]


def level_dir(path):
    return [d for d in os.listdir(path) if d.startswith("Level_")][0]


@pytest.mark.parametrize("name", CHECKPOINTS)
def test_checkpoint_reader_on_shipped_and_corrupted_checkpoints(bmx, harness, tmp_path, name):
    src = os.path.join(GOLDEN, "ref_checkpoints", name)
    maps = maps_for(bmx, src)
    out = run(harness, "chk", src, tmp_path / "w", *maps)
    assert "chk-reread status=0" in out
    lev = level_dir(src)
    header = open(os.path.join(src, "Header"), "rb").read()
    ph = open(os.path.join(src, lev, "Particle_H"), "rb").read()
    data_files = sorted(f for f in os.listdir(os.path.join(src, lev)) if f.startswith("DATA_"))
    hl = header.split(b"\n")

    def variant(tag, header_bytes=None, ph_bytes=None, data_edit=None):
        d = tmp_path / tag
        shutil.copytree(src, d)
        if header_bytes is not None:
            (d / "Header").write_bytes(header_bytes)
        if ph_bytes is not None:
            (d / lev / "Particle_H").write_bytes(ph_bytes)
        if data_edit is not None:
            f = d / lev / data_files[0]
            f.write_bytes(data_edit(f.read_bytes()))
        run(harness, "chk", d, tmp_path / f"{tag}_w", *maps)

    for frac in (0.05, 0.3, 0.6, 0.95):
        variant(f"header_cut_{frac}", header_bytes=header[: int(len(header) * frac)])
        variant(f"data_cut_{frac}", data_edit=lambda b, fr=frac: b[: int(len(b) * fr)])
    variant("ph_cut", ph_bytes=ph[: len(ph) // 2])
    variant("ph_junk", ph_bytes=b"(((((\n")
    variant("data_empty", data_edit=lambda b: b"")
    # the last Header lines are (file, count, offset) per grid
    tail = hl[-2] if hl[-1] == b"" else hl[-1]
    parts = tail.split()
    if len(parts) == 3:
        for tag, rep in (("count_inflated", [parts[0], b"99999999", parts[2]]),
                         ("count_negative", [parts[0], b"-5", parts[2]]),
                         ("offset_past_end", [parts[0], parts[1], b"123456789"]),
                         ("offset_negative", [parts[0], parts[1], b"-8"]),
                         ("file_missing", [b"77", parts[1], parts[2]])):
            variant(tag, header_bytes=header.replace(tail, b" ".join(rep)))
    # component counts: second line names the dimension, third the reals
    variant("n_real_huge", header_bytes=b"\n".join(hl[:2] + [b"100000"] + hl[3:]))
    variant("n_real_negative", header_bytes=b"\n".join(hl[:2] + [b"-1"] + hl[3:]))
    variant("dim_bad", header_bytes=b"\n".join(hl[:1] + [b"7"] + hl[2:]))
    variant("version_bad", header_bytes=b"Version_Two_Dot_Zero_single\n" +
            b"\n".join(hl[1:]))
