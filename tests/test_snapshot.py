"""GALX v1 snapshots (galax-io/src/lib.rs:5-92) through the product library's
C ABI (galax_snapshot_write / _info / _read), host code only. Pinned by the
three snapshots the reference holds (runs/demo/snap_0010.bin,
runs/test/snap_0005.bin, runs/viz-demo/snap_0010.bin, copied under
tests/golden/ as galx_*.bin), parsed here independently with struct / numpy,
and by the reference's own tests (round trip, bad magic:
galax-io/src/lib.rs:95-133)."""
import os
import struct

import numpy as np
import pytest

from galax_amd import fmm, ic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = ["galx_demo_0010.bin", "galx_test_0005.bin", "galx_viz_demo_0010.bin"]


def _parse(path):
    """Independent reader: header <4sIQd, then 7 little-endian f64 arrays."""
    raw = open(path, "rb").read()
    magic, ver, n, ts = struct.unpack("<4sIQd", raw[:24])
    arrs = np.frombuffer(raw[24:24 + 56 * n], dtype="<f8").reshape(7, n)
    return magic, ver, n, ts, arrs


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_snapshots_read_and_rewrite_byte_for_byte(name, tmp_path):
    path = os.path.join(GOLDEN, name)
    magic, ver, n, ts, arrs = _parse(path)
    assert magic == b"GALX" and ver == 1 and n == 1000
    assert fmm.snapshot_info(path) == (n, ts)
    b = fmm.read_snapshot(path)
    for got, want in zip((b.x, b.y, b.vx, b.vy, b.mass, b.ax, b.ay), arrs):
        np.testing.assert_array_equal(got, want)
    out = tmp_path / "rewrite.bin"
    fmm.write_snapshot(out, b, timestamp=ts)
    assert out.read_bytes() == open(path, "rb").read()


def test_round_trip_random_bodies(tmp_path):  # galax-io/src/lib.rs:97-125, bit-exact here
    r = ic.SplitMix64(77)
    n = 100
    b = fmm.BodiesSoA(*(r.uniform(n) * 100 - 50 for _ in range(7)), None)
    p = tmp_path / "s.bin"
    fmm.write_snapshot(p, b, timestamp=123.5)
    assert fmm.snapshot_info(p) == (n, 123.5)
    c = fmm.read_snapshot(p)
    for k in ("x", "y", "vx", "vy", "mass", "ax", "ay"):
        np.testing.assert_array_equal(getattr(c, k), getattr(b, k))
    assert os.path.getsize(p) == 24 + 56 * n


def test_empty_snapshot(tmp_path):
    p = tmp_path / "e.bin"
    fmm.write_snapshot(p, fmm.BodiesSoA.new(0, with_phi=False), timestamp=0.0)
    assert os.path.getsize(p) == 24
    assert len(fmm.read_snapshot(p)) == 0


def test_bad_magic_is_a_format_error(tmp_path):  # galax-io/src/lib.rs:127-132
    p = tmp_path / "bad.bin"
    p.write_bytes(b"BADXsomejunkdata")
    with pytest.raises(fmm.GalaxError) as e:
        fmm.read_snapshot(p)
    assert e.value.code == -9 and "magic" in str(e.value)


def test_unsupported_version_and_short_files(tmp_path):
    raw = open(os.path.join(GOLDEN, FIXTURES[0]), "rb").read()
    p = tmp_path / "v2.bin"
    p.write_bytes(raw[:4] + struct.pack("<I", 2) + raw[8:])
    with pytest.raises(fmm.GalaxError) as e:
        fmm.read_snapshot(p)
    assert e.value.code == -6 and "version: 2" in str(e.value)
    for cut in (2, 10, 20, 24 + 8 * 1000 * 3 + 5, len(raw) - 1):
        p = tmp_path / f"cut{cut}.bin"
        p.write_bytes(raw[:cut])
        with pytest.raises(fmm.GalaxError) as e:
            fmm.read_snapshot(p)
        assert e.value.code == -9, cut


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(fmm.GalaxError) as e:
        fmm.read_snapshot(tmp_path / "nope.bin")
    assert e.value.code == -8
