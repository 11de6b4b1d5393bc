"""CPU checks of the C++ host side (gpugraph500_amd/host) and of the 2D block
bounds; no GPU call is made."""
import os
import subprocess

import pytest

import gpugraph500_amd as g5
from gpugraph500_amd import build as gbuild

HERE = os.path.dirname(os.path.abspath(__file__))


def _parts(n_pad, parts):
    return [g5.block_bounds(n_pad, parts, parts, i, i)[:2] for i in range(parts)]


@pytest.mark.parametrize("nglobal", [1, 63, 64, 65, 1000, 4096, 65535, 1 << 20, 67108862])
@pytest.mark.parametrize("R", [1, 2, 3, 4, 7])
def test_block_bounds_partition(nglobal, R):
    """Row parts tile [0, padded n) in 64-vertex units, contiguous, longer
    parts first, lengths differing by at most one unit
    (cpu_2d/distmatrix2d.hh:807-815, ALG = 64, vertex count padded to R)."""
    n_pad = nglobal + (-nglobal) % R
    units = -(-n_pad // 64)
    parts = [g5.block_bounds(nglobal, R, R, i, 0)[:2] for i in range(R)]
    pos = 0
    lens = []
    for start, length in parts:
        assert start == pos and start % 64 == 0 and length % 64 == 0
        pos += length
        lens.append(length // 64)
    assert pos == units * 64
    assert max(lens) - min(lens) <= 1 and lens == sorted(lens, reverse=True)
    assert lens.count(max(lens)) == (units % R or R)


def test_block_bounds_known_cases():
    # scale 26, 4 x 4: 67,108,862 vertices pad to 2^26 -> 2^24 per part
    assert g5.block_bounds(67108862, 4, 4, 1, 2) == (1 << 24, 1 << 24, 2 << 24, 1 << 24)
    # 1000 vertices, 3 x 3: pad to 1002 -> 16 units -> 6, 5, 5
    assert g5.block_bounds(1000, 3, 3, 2, 0) == (704, 320, 0, 384)
    assert g5.block_bounds(1000, 3, 3, 0, 1) == (0, 384, 384, 320)


def test_block_bounds_rejects_bad_arguments():
    for args in [(-1, 1, 1, 0, 0), (10, 0, 1, 0, 0), (10, 2, 2, 2, 0), (10, 2, 2, 0, -1), (10, 65, 1, 0, 0)]:
        with pytest.raises(g5.G5Error):
            g5.block_bounds(*args)


def test_host_unit_cpp(tmp_path):
    """Compression schemas' host-only behaviour (gate, receiver rule,
    configuration errors, factory names)."""
    lib_dir = os.path.dirname(g5.LIB_PATH)
    exe = str(tmp_path / "host_unit")
    cmd = [os.environ.get("CXX", "g++")] + gbuild.HOST_FLAGS + [os.path.join(HERE, "host_unit.cpp"), "-o", exe,
                                                                "-L" + lib_dir, "-lg5hip", "-Wl,-rpath," + lib_dir]
    subprocess.run(cmd, check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "all checks passed" in r.stdout


@pytest.mark.parametrize("name", ["g500_hip", "g500_serve"])
def test_host_sources_compile_cleanly(tmp_path, name):
    """The host side and each driver build warning-free against the C ABI alone."""
    lib_dir = os.path.dirname(g5.LIB_PATH)
    main = os.path.join(gbuild.HOST, name + ".cpp")
    exe = str(tmp_path / name)
    cmd = [os.environ.get("CXX", "g++")] + gbuild.HOST_FLAGS + ["-Werror"] + gbuild.host_sources() + [
        main, "-o", exe, "-L" + lib_dir, "-lg5hip", "-Wl,-rpath," + lib_dir]
    subprocess.run(cmd, check=True)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: the driver would run")
def test_driver_fails_loudly_without_gpu():
    r = subprocess.run([gbuild.DRIVER, "-s", "8"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "g5_init failed" in r.stderr


def test_driver_rejects_bad_options():
    for args in (["-x", "wire"], ["-n", "0"], ["-R", "2", "-C", "2", "-share", "2", "0"], ["-bogus"]):
        r = subprocess.run([gbuild.DRIVER] + args, capture_output=True, text=True)
        assert r.returncode == 1 and "g500_hip:" in r.stderr
    for args in (["-rate", "0"], ["-maxroots", "9", "-lanes", "8"], ["-lanes", "65"], ["-g", ""], ["-x"]):
        r = subprocess.run([gbuild.SERVE] + args, capture_output=True, text=True)
        assert r.returncode == 1 and "g500_serve:" in r.stderr


def test_persisted_shapes_round_trip(tmp_path, monkeypatch):
    """Tuned shapes are kept per graph key in a JSON file and read back."""
    from gpugraph500_amd import shapes
    monkeypatch.setattr(shapes, "PATH", str(tmp_path / "tuned.json"))
    assert shapes.load("s20-h1.2.3") is None
    rec = {"params": {"thread_max": 64, "wave_max": 2048}, "best_ms": 1.5}
    shapes.save("s20-h1.2.3", rec)
    shapes.save("s21-h4.5", {"params": {"thread_max": 32}})
    assert shapes.load("s20-h1.2.3") == rec
    assert shapes.load("s21-h4.5")["params"] == {"thread_max": 32}
    # a record without one of the current keys is tuned again
    assert shapes.load("s20-h1.2.3", ("thread_max", "wave_max")) == rec
    assert shapes.load("s21-h4.5", ("thread_max", "split_max")) is None
    (tmp_path / "tuned.json").write_text("not json")
    assert shapes.load("s20-h1.2.3") is None
