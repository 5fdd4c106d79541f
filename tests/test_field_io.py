"""Restart / physical / EKT files in the reference's formats, and the seeded initial field.

Each test follows one of the reference's FieldIO and initial_bc tests (tests/test_field_io.cpp,
tests/test_initial_bc.cpp); the mesh is the reference's examples/config.toml (32^3).  The
transforms behind to_physical and initialize_velocity are the oracle's coarse transforms here;
tests/test_gpu_io.py runs the same compositions on the GPU.
"""
import math
import os
import struct

import numpy as np
import pytest

from cdns import config, field_io, initial_bc
import pyoracle

CFG = config.reference_example()


def _source():
    io = field_io.FieldIO(CFG, "/nonexistent")
    n = int(np.prod(io.velocity_shape()))
    i = np.arange(n, dtype=np.float64)
    v = np.empty(n, np.complex128)
    v.real, v.imag = i * 0.125, -i * 0.25  # {i 0.125, -i 0.25}: entry 0 has imaginary part -0.0
    return v.reshape(io.velocity_shape())


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_restart_round_trip_is_bit_exact(tmp_path, fmt):
    """tests/test_field_io.cpp:19-37, 56-75."""
    io = field_io.FieldIO(CFG, tmp_path, output_format=fmt)
    src = _source()
    io.write_restart(17, 0.75, src)
    nt, t, restored = io.read_restart()
    assert nt == 17 and t == 0.75
    assert np.array_equal(restored.view(np.float64), src.view(np.float64))
    assert sorted(os.listdir(tmp_path)) == ["field_S_ 0.bin" if fmt == "binary" else "field_S_ 0.dat"]


def test_binary_header_layout(tmp_path):
    """Header fields at their byte offsets (field_io.cpp:21-27, 121-139) and the payload order
    [iz][iy][kx][c] (re, im) (field_io.cpp:299-307)."""
    io = field_io.FieldIO(CFG, tmp_path, output_format="binary")
    src = _source()
    io.write_restart(7, 0.25, src)
    b = open(tmp_path / "field_S_ 0.bin", "rb").read()
    assert b[:8] == b"CHDNS\x00\x00\x00"
    assert struct.unpack_from("<II", b, 8) == (1, 1)
    assert struct.unpack_from("<q", b, 16) == (7,) and struct.unpack_from("<d", b, 24) == (0.25,)
    m = CFG
    assert struct.unpack_from("<12q", b, 32) == (32, 33, 32, 3, 33, 17, 0, 1, 0, 17, 0, 31)
    assert m["NZS"] == 16 and m["NXS"] == 16 and m["NYS"] == 30
    assert len(b) == 128 + 33 * 31 * 17 * 3 * 16
    # entry (iz=2, iy=5, kx=3, c=1)
    k = ((2 * 31 + 5) * 17 + 3) * 3 + 1
    assert struct.unpack_from("<dd", b, 128 + 16 * k) == (k * 0.125, -k * 0.25)


def test_text_restart_layout(tmp_path):
    """First line "ntpit time", then "(re,im)" per entry with 17 significant digits
    (field_io.cpp:270-280)."""
    io = field_io.FieldIO(CFG, tmp_path)
    io.write_restart(7, 0.25, _source())
    lines = open(tmp_path / "field_S_ 0.dat").read().splitlines()
    assert lines[0] == "7 2.50000000000000000e-01"
    assert lines[1] == "(0.00000000000000000e+00,-0.00000000000000000e+00)"
    assert lines[2] == "(1.25000000000000000e-01,-2.50000000000000000e-01)"
    assert len(lines) == 1 + 33 * 31 * 17 * 3


def test_ekt_appends(tmp_path):
    """tests/test_field_io.cpp:39-55: EKT is the sum of v^2 over the snapshot."""
    io = field_io.FieldIO(CFG, tmp_path)
    snap = {"velocity": [np.zeros(4), np.full(4, 2.0), np.zeros(4)], "pressure": np.zeros(4)}
    io.append_ekt(1, 0.1, snap)
    io.append_ekt(2, 0.2, snap)
    t1, e1, t2, e2 = map(float, open(tmp_path / "EKT 0.dat").read().split())
    assert (t1, e1, t2, e2) == (0.1, 16.0, 0.2, 16.0)


def test_physical_text_uses_component_first_order(tmp_path):
    """tests/test_field_io.cpp:78-96."""
    io = field_io.FieldIO(CFG, tmp_path)
    snap = {"velocity": [np.array([10.0 + c, 20.0 + c]) for c in range(3)], "pressure": np.array([30.0, 31.0])}
    io.write_physical(9, 0.5, snap)
    vals = [float(x) for x in open(tmp_path / "field_P_ 0.dat").read().split()]
    assert vals == [9, 0.5, 10, 11, 12, 20, 21, 22, 30, 31]


def test_physical_binary_header_and_planes(tmp_path):
    io = field_io.FieldIO(CFG, tmp_path, output_format="binary")
    n = CFG["NZ"] * (CFG["NY"] + 1) * CFG["NX"]
    snap = {"velocity": [np.full(n, 1.0 + c) for c in range(3)], "pressure": np.full(n, 9.0)}
    io.write_physical(3, 1.5, snap)
    b = open(tmp_path / "field_P_ 0.bin", "rb").read()
    assert struct.unpack_from("<II", b, 8) == (1, 2)
    assert struct.unpack_from("<12q", b, 32) == (32, 33, 32, 4, 0, 0, 0, 1, 0, 32, 0, 33)
    body = np.frombuffer(b, "<f8", offset=128).reshape(4, n)
    assert np.array_equal(body[:, 0], [1.0, 2.0, 3.0, 9.0])


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_truncated_restart_raises(tmp_path, fmt):
    """tests/test_field_io.cpp:98-121: a truncated file is rejected (the caller's state is the
    reader's return value, so nothing changes)."""
    io = field_io.FieldIO(CFG, tmp_path, output_format=fmt)
    io.write_restart(17, 0.75, np.full(io.velocity_shape(), 3.0 + 4.0j))
    path = tmp_path / ("field_S_ 0.bin" if fmt == "binary" else "field_S_ 0.dat")
    with open(path, "r+b") as f:
        f.truncate(64)
    with pytest.raises(RuntimeError):
        io.read_restart()


def test_both_modes_fallback_and_old_binary_dat_files(tmp_path):
    """tests/test_field_io.cpp:138-163."""
    src = _source()
    field_io.FieldIO(CFG, tmp_path, output_format="both").write_restart(7, 0.25, src)
    assert (tmp_path / "field_S_ 0.dat").exists() and (tmp_path / "field_S_ 0.bin").exists()
    for fmt in field_io.FORMATS:
        nt, t, v = field_io.FieldIO(CFG, tmp_path, output_format=fmt).read_restart()
        assert (nt, t) == (7, 0.25) and np.array_equal(v, src)
    os.remove(tmp_path / "field_S_ 0.dat")
    field_io.FieldIO(CFG, tmp_path).read_restart()  # ascii reader falls back to the .bin
    os.rename(tmp_path / "field_S_ 0.bin", tmp_path / "field_S_ 0.dat")
    nt, t, v = field_io.FieldIO(CFG, tmp_path, output_format="binary").read_restart()
    assert nt == 7 and np.array_equal(v, src)


def test_output_failure_leaves_no_temporary_files(tmp_path):
    """tests/test_field_io.cpp:185-195."""
    io = field_io.FieldIO(CFG, tmp_path / "missing", output_format="both")
    with pytest.raises(RuntimeError):
        io.write_restart(1, 0.1, _source())
    snap = {"velocity": [np.zeros(2)] * 3, "pressure": np.zeros(2)}
    with pytest.raises(RuntimeError):
        io.write_physical(1, 0.1, snap)
    with pytest.raises(RuntimeError):
        io.append_ekt(1, 0.1, snap)
    assert os.listdir(tmp_path) == []


def test_text_parser_handles_d_exponents_commas_and_rejects_corrupt_payloads(tmp_path):
    """tests/test_field_io.cpp:197-219."""
    io = field_io.FieldIO(CFG, tmp_path)
    src = _source()
    io.write_restart(7, 0.25, src)
    path = tmp_path / "field_S_ 0.dat"
    original = path.read_text()
    variant = original.replace("e", "D").replace("7 ", "7, ", 1)
    path.write_text(variant)
    nt, t, v = io.read_restart()
    assert (nt, t) == (7, 0.25) and np.array_equal(v, src)
    for corrupt in (original + "1\n", original.replace("(", "["), original.replace(",", ";"), "-1 0\n",
                    original.replace("(", "(nan,")):
        path.write_text(corrupt)
        with pytest.raises(RuntimeError):
            io.read_restart()


@pytest.mark.parametrize("offset", [0, 7, 8, 12, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120])
def test_corrupt_binary_header_is_rejected(tmp_path, offset):
    """tests/test_field_io.cpp:221-237: flipping bit 6 of the magic, version, kind or any
    dimension is detected (the step and the time are free values, as in the reference's list)."""
    io = field_io.FieldIO(CFG, tmp_path, output_format="binary")
    io.write_restart(7, 0.25, _source())
    path = tmp_path / "field_S_ 0.bin"
    b = bytearray(path.read_bytes())
    b[offset] ^= 0x40
    path.write_bytes(bytes(b))
    with pytest.raises(RuntimeError):
        io.read_restart()


def test_non_finite_binary_payload_is_rejected(tmp_path):
    io = field_io.FieldIO(CFG, tmp_path, output_format="binary")
    src = _source()
    src[1, 2, 3, 0] = complex(math.inf, 0.0)
    io.write_restart(7, 0.25, src)
    with pytest.raises(RuntimeError):
        io.read_restart()


def test_rank_filename_pads_to_width_two():
    assert field_io.rank_filename("field_S_", 0, ".bin") == "field_S_ 0.bin"
    assert field_io.rank_filename("EKT", 12, ".dat") == "EKT12.dat"
    assert field_io.rank_filename("EKT", 123, ".dat") == "EKT123.dat"


def test_to_physical_preserves_inputs_and_separates_components(tmp_path):
    """tests/test_field_io.cpp:165-183, with the oracle's coarse transforms."""
    o = pyoracle.Oracle(CFG)
    io = field_io.FieldIO(CFG, tmp_path, transforms=o)
    vel = np.zeros(io.velocity_shape(), np.complex128)
    vel[CFG["NZS"], 0, 0, :] = [2.0, 3.0, 4.0]
    pr = np.zeros(io.velocity_shape()[:3], np.complex128)
    pr[CFG["NZS"], 0, 0] = 7.0
    v0, p0 = vel.copy(), pr.copy()
    snap = io.to_physical(vel, pr, finite=True)
    n = CFG["NZ"] * (CFG["NY"] + 1) * CFG["NX"]
    for c in range(3):
        assert snap["velocity"][c].size == n
        assert np.abs(snap["velocity"][c] - (2.0 + c)).max() < 1e-13
    assert np.abs(snap["pressure"] - 7.0).max() < 1e-13 and snap["finite"]
    assert np.array_equal(vel, v0) and np.array_equal(pr, p0)
    zero = io.to_physical(np.zeros_like(vel), np.zeros_like(pr))
    assert all((c == 0).all() for c in zero["velocity"]) and (zero["pressure"] == 0).all()


# ------------------------------------------------------------- initial field --
def test_engine_is_std_mt19937_64():
    """The 10000th output of a default-seeded (5489) std::mt19937_64 is 9981545732273789042
    (the C++ standard's check value); the stream equals the oracle's restatement."""
    e = initial_bc.DisturbanceEngine(5489)
    assert int(e.raw(10000)[-1]) == 9981545732273789042
    e, r = initial_bc.DisturbanceEngine(12345), pyoracle.MT19937_64(12345)
    assert [int(x) for x in e.raw(1000)] == [r() for _ in range(1000)]


def test_centered_uniform_mapping_and_moments():
    """initial_bc.cpp:15-18 on the raw stream; tests/test_initial_bc.cpp:68-93 moments."""
    raw = initial_bc.DisturbanceEngine(987654321).raw(1000)
    cu = initial_bc.DisturbanceEngine(987654321).centered_uniform(1000)
    assert np.array_equal(cu, (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 - 0.5)
    x = initial_bc.DisturbanceEngine(987654321).centered_uniform(200000)
    assert x.min() >= -0.5 and x.max() < 0.5 and x.min() < -0.499 and x.max() > 0.499
    mean = x.sum() / x.size
    assert abs(mean) < 2e-3
    assert abs((x * x).sum() / x.size - mean * mean - 1.0 / 12.0) < 8e-4
    assert abs((x[:-1] * x[1:]).sum() / (x.size - 1)) < 8e-4


def test_engine_continues_across_draws():
    a = initial_bc.DisturbanceEngine(77)
    first, second = a.raw(5), a.raw(5)
    assert np.array_equal(np.concatenate([first, second]), initial_bc.DisturbanceEngine(77).raw(10))


def test_wall_normal_points():
    """mesh.cpp:100-110: cos(pi j/NY) mirrored, exactly 0 in the middle."""
    yc = initial_bc.wall_normal_points(32)
    assert yc[0] == 1.0 and yc[32] == -1.0 and yc[16] == 0.0
    assert np.array_equal(yc[::-1], -yc)
    assert np.abs(yc - np.cos(np.pi * np.arange(33) / 32)).max() < 1e-15


def test_initialize_velocity_builds_the_parabolic_profile_and_boundary():
    """tests/test_initial_bc.cpp:13-46: u = 1.5 (1 - y^2), v = V_Microblow0, w = 0 on the coarse
    grid (INDEX_GRAD_PRE = 0), to 1e-12; the boundary carries V_Microblow0 on v of the mean mode."""
    o = pyoracle.Oracle(CFG)
    vel = initial_bc.initialize_velocity(o)
    yc = initial_bc.wall_normal_points(CFG["NY"])
    phys = o.coarse_phys()
    for c, expect in enumerate((1.5 * (1.0 - yc * yc), np.full_like(yc, CFG["V_Microblow0"]), np.zeros_like(yc))):
        o.xyzfft("B", phys, vel.copy(), 3, c)
        assert np.abs(phys - expect[None, :, None]).max() < 1e-12
    # Chebyshev form: 1 - y^2 = (T0 - T2) / 2
    assert abs(vel[CFG["NZS"], 0, 0, 0] - 0.75) < 1e-14 and abs(vel[CFG["NZS"], 2, 0, 0] + 0.75) < 1e-14
    b = initial_bc.velocity_boundary(CFG)
    nz = np.argwhere(b != 0)
    assert [tuple(r) for r in nz] == [(CFG["NZS"], 0, 0, 1), (CFG["NZS"], 1, 0, 1)]
    assert (b[CFG["NZS"], :, 0, 1] == CFG["V_Microblow0"]).all()


def test_seeded_disturbance_is_reproducible_and_matches_the_draws():
    """tests/test_initial_bc.cpp:48-66 (same seed, same field; the field differs from the
    laminar one), and the physical increment is exactly the scaled draws in [iz][iy][ix][u,v,w]
    order, up to the round trip of the transforms."""
    o = pyoracle.Oracle(CFG)
    base = initial_bc.initialize_velocity(o)
    d1 = initial_bc.initialize_velocity(o, initial_bc.DisturbanceEngine(12345), eps_init=0.1)
    d2 = initial_bc.initialize_velocity(o, initial_bc.DisturbanceEngine(12345), eps_init=0.1)
    assert np.array_equal(d1, d2)
    assert np.abs(d1 - base).max() > 1e-6
    # the increment, synthesised, against the draws projected the same way
    NX, NY, NZ = CFG["NX"], CFG["NY"], CFG["NZ"]
    yc = initial_bc.wall_normal_points(NY)
    draws = initial_bc.DisturbanceEngine(12345).centered_uniform(3 * NZ * (NY + 1) * NX).reshape(NZ, NY + 1, NX, 3)
    scale = (1.5 * (1.0 - yc * yc) * 0.1)[None, :, None]
    for c in range(3):
        noise = scale * draws[..., c] * (CFG.get("INDEX_disturb_Z", 1) if c == 2 else 1.0)
        proj = o.vector()
        o.xyzfft("F", noise.copy(), proj, 3, c)
        assert np.abs((d1 - base)[..., c] - proj[..., c]).max() < 1e-13


def test_zero_disturbance_preserves_field_and_engine():
    """tests/test_initial_bc.cpp:95-105."""
    o = pyoracle.Oracle(CFG)
    vel = np.full((2 * CFG["NZS"] + 1, CFG["NYS"] + 1, CFG["NXS"] + 1, 3), 3.0 + 4.0j)
    e, untouched = initial_bc.DisturbanceEngine(123), initial_bc.DisturbanceEngine(123)
    initial_bc.disturb_velocity(o, vel, 0.0, e)
    assert (vel == 3.0 + 4.0j).all()
    assert e() == untouched()
