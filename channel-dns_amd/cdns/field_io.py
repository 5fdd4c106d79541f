"""Restart, physical-snapshot and EKT files in the reference's formats (serial run, rank 0).

  FieldIO::write_restart / read_restart   src/io/field_io.cpp:265-427
  FieldIO::write_physical / append_ekt     src/io/field_io.cpp:222-263
  FieldIO::to_physical                     src/io/field_io.cpp:199-220
  binary header                            src/io/field_io.cpp:21-27, 113-156
  rank_filename                            src/core/rank_filename.hpp:8-14

Binary files: 128-byte little-endian header (magic "CHDNS\\0\\0\\0", version 1, kind 1 restart /
2 physical, ntpit int64, time binary64, 12 int64 dimensions), then binary64 payload.  Restart
payload: [(iz+NZS)][iy][kx][c] (re, im); physical payload: u, v, w, pressure planes
[iz][iy][ix].  Text files: "ntpit time" then one value per line, 17 significant digits.
Every file is written to `<name>.tmp` and renamed, so a failed write leaves nothing behind.
A failed read raises RuntimeError and returns nothing, so the caller's state is unchanged.
"""
import math
import os
import re
import struct

import numpy as np

MAGIC = b"CHDNS\x00\x00\x00"
VERSION = 1
KIND_RESTART, KIND_PHYSICAL = 1, 2
HEADER_BYTES = 128
FORMATS = ("ascii", "binary", "both")

# what std::stod accepts as one whole token (decimal or hexadecimal floating literal; "inf" and
# "nan" forms parse and are then rejected as non-finite)
_DEC = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?\Z")
_HEX = re.compile(r"[+-]?0[xX]([0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)([pP][+-]?\d+)?\Z")
_WS = " \t\n\v\f\r"  # the C locale's isspace
# one payload entry of the fast path: optional comma, "(re,im)" with decimal tokens (d/D exponents
# allowed), whitespace as the stream extraction of the reference skips it
_DTOK = r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?"
_ENTRY = re.compile(r"[ \t\n\v\f\r]*,?[ \t\n\v\f\r]*\([ \t\n\v\f\r]*(" + _DTOK + r")[ \t\n\v\f\r]*,"
                    r"[ \t\n\v\f\r]*(" + _DTOK + r")[ \t\n\v\f\r]*\)")
_NONFINITE = re.compile(r"[+-]?(inf|infinity|nan(\([0-9A-Za-z_]*\))?)\Z", re.IGNORECASE)


def rank_filename(prefix, rank, suffix):
    """prefix + rank right-aligned in a field of width 2 + suffix ("field_S_ 0.bin")."""
    return f"{prefix}{rank:2d}{suffix}"


def _fmt(v):
    """std::scientific with precision 17."""
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    if math.isinf(v):
        return "-inf" if v < 0 else "inf"
    return "%.17e" % v


def _atomic_write(path, data):
    tmp = path + ".tmp"
    try:
        with open(tmp, "wb") as f:
            f.write(data)
        os.replace(tmp, path)
    except OSError as e:
        try:
            os.remove(tmp)
        except OSError:
            pass
        raise RuntimeError(f"cannot write {path}: {e}") from None


def _header(kind, ntpit, time, dims):
    return MAGIC + struct.pack("<IIqd12q", VERSION, kind, int(ntpit), float(time), *dims)


class FieldIO:
    """Files of one serial run in `run_dir`; `cfg` is a config dict (cdns.config.make), `transforms`
    an object with the coarse-grid xyzfft of a Channel (needed for to_physical only)."""

    def __init__(self, cfg, run_dir, transforms=None, output_format="ascii"):
        if output_format not in FORMATS:
            raise ValueError(f"output_format must be one of {FORMATS}")
        self.cfg, self.run_dir, self.transforms, self.format = cfg, str(run_dir), transforms, output_format
        self.my_id, self.nproc = 0, 1

    # ------------------------------------------------------------------ shapes --
    def velocity_shape(self):
        m = self.cfg
        return (2 * m["NZS"] + 1, m["NYS"] + 1, m["NXS"] + 1, 3)

    def _restart_dims(self):
        m = self.cfg
        return (m["NX"], m["NY"] + 1, m["NZ"], 3, 2 * m["NZS"] + 1, m["NXS"] + 1, self.my_id, self.nproc, 0,
                m["NXS"] + 1, 0, m["NYS"] + 1)

    def _physical_dims(self):
        m = self.cfg
        return (m["NX"], m["NY"] + 1, m["NZ"], 4, 0, 0, self.my_id, self.nproc, 0, m["NX"], 0, m["NY"] + 1)

    def _path(self, prefix, suffix):
        return os.path.join(self.run_dir, rank_filename(prefix, self.my_id, suffix))

    # ---------------------------------------------------------------- restart --
    def write_restart(self, ntpit, time, velocity):
        """field_S_ 0.dat (text, unless the format is binary) and field_S_ 0.bin (unless ascii)."""
        v = np.ascontiguousarray(velocity, dtype=np.complex128)
        if v.shape != self.velocity_shape():
            raise ValueError(f"velocity must have shape {self.velocity_shape()}")
        if self.format != "binary":
            flat = v.reshape(-1)
            lines = [f"{int(ntpit)} {_fmt(time)}\n"]
            lines += [f"({_fmt(z.real)},{_fmt(z.imag)})\n" for z in flat.tolist()]
            _atomic_write(self._path("field_S_", ".dat"), "".join(lines).encode())
        if self.format != "ascii":
            payload = v.view(np.float64).astype("<f8", copy=False).tobytes()
            _atomic_write(self._path("field_S_", ".bin"),
                          _header(KIND_RESTART, ntpit, time, self._restart_dims()) + payload)

    def read_restart(self):
        """(ntpit, time, velocity).  field_S_ 0.dat is read unless it is missing, or the format is
        not ascii and field_S_ 0.bin exists; a file that starts with the magic is binary whatever
        its name.  Any mismatch, truncation, non-finite value or stray text raises RuntimeError."""
        path = self._path("field_S_", ".dat")
        bpath = self._path("field_S_", ".bin")
        if not os.path.exists(path) or (self.format != "ascii" and os.path.exists(bpath)):
            path = bpath
        try:
            with open(path, "rb") as f:
                data = f.read()
        except OSError:
            raise RuntimeError(f"cannot open restart: {path}") from None
        if data[:8] == MAGIC:
            return self._read_binary(data)
        return self._read_text(data)

    def _read_binary(self, data):
        if len(data) < 16:
            raise RuntimeError("truncated field header")
        (version,) = struct.unpack_from("<I", data, 8)
        if version != VERSION:
            raise RuntimeError("unsupported field version")
        if len(data) < HEADER_BYTES:
            raise RuntimeError("truncated field header")
        kind, ntpit, time = struct.unpack_from("<Iqd", data, 12)
        dims = struct.unpack_from("<12q", data, 32)
        if kind != KIND_RESTART or ntpit < 0 or not math.isfinite(time) or dims != self._restart_dims():
            raise RuntimeError("restart header does not match this rank and mesh")
        n = int(np.prod(self.velocity_shape())) * 2
        if len(data) != HEADER_BYTES + 8 * n:
            raise RuntimeError("restart payload length mismatch")
        vals = np.frombuffer(data, dtype="<f8", count=n, offset=HEADER_BYTES)
        if not np.isfinite(vals).all():
            raise RuntimeError("restart contains non-finite data")
        vel = vals.astype(np.float64).view(np.complex128).reshape(self.velocity_shape()).copy()
        return int(ntpit), float(time), vel

    def _read_text(self, data):
        try:
            text = data.decode("ascii")
        except UnicodeDecodeError:
            raise RuntimeError("invalid text restart data") from None
        pos, end = 0, len(text)

        def skip_ws():
            nonlocal pos
            while pos < end and text[pos] in _WS:
                pos += 1

        def number():
            nonlocal pos
            skip_ws()
            start = pos
            while pos < end and text[pos] not in _WS and text[pos] not in "),":
                pos += 1
            token = text[start:pos].replace("d", "e").replace("D", "e")
            if _DEC.match(token):
                value = float(token)
            elif _HEX.match(token):
                value = float.fromhex(token)
            elif _NONFINITE.match(token):
                value = math.inf
            else:
                raise RuntimeError("invalid text restart number")
            if not math.isfinite(value):
                raise RuntimeError("invalid text restart number")
            return value

        def punctuation(expected):
            nonlocal pos
            skip_ws()
            if pos >= end or text[pos] != expected:
                raise RuntimeError("invalid text complex syntax")
            pos += 1

        skip_ws()
        m = re.compile(r"[+-]?\d+").match(text, pos)
        if not m:
            raise RuntimeError("invalid restart step")
        ntpit = int(m.group(0))
        if ntpit < 0 or ntpit > 2 ** 63 - 1:
            raise RuntimeError("invalid restart step")
        pos = m.end()
        skip_ws()
        if pos < end and text[pos] == ",":
            pos += 1
        time = number()
        n = int(np.prod(self.velocity_shape()))
        fast = self._read_entries_fast(text, pos, n)
        if fast is not None:
            return ntpit, time, fast
        vals = np.empty(2 * n)
        for i in range(n):
            skip_ws()
            if pos < end and text[pos] == ",":
                pos += 1
            punctuation("(")
            vals[2 * i] = number()
            punctuation(",")
            vals[2 * i + 1] = number()
            punctuation(")")
        if text[pos:].strip(_WS):
            raise RuntimeError("extra text restart data")
        return ntpit, time, vals.view(np.complex128).reshape(self.velocity_shape()).copy()

    def _read_entries_fast(self, text, pos, n):
        """The payload when every entry is a plain decimal "(re,im)" (what the writers produce),
        parsed by one regular expression per entry and one vectorised conversion; None when the
        text needs the general parser (hexadecimal or non-finite tokens, or any error, which the
        general parser then reports)."""
        toks = []
        m = None
        for _ in range(n):
            m = _ENTRY.match(text, pos)
            if m is None:
                return None
            toks.append(m.group(1))
            toks.append(m.group(2))
            pos = m.end()
        if text[pos:].strip(_WS):
            return None
        vals = np.array([t.replace("d", "e").replace("D", "e") for t in toks], dtype=np.float64)
        if not np.isfinite(vals).all():
            return None
        return vals.view(np.complex128).reshape(self.velocity_shape()).copy()

    # --------------------------------------------------------------- physical --
    def to_physical(self, velocity, pressure, finite=False):
        """Coarse-grid synthesis of u, v, w and the pressure: {"velocity": [u, v, w], "pressure": p},
        each [iz][iy][ix]; with finite=True also {"finite": all spectral inputs finite}."""
        if self.transforms is None:
            raise RuntimeError("to_physical needs transforms")
        m = self.cfg
        shape = (m["NZ"], m["NY"] + 1, m["NX"])
        vel = np.ascontiguousarray(velocity, dtype=np.complex128)
        pr = np.ascontiguousarray(pressure, dtype=np.complex128)
        out = {"velocity": [np.zeros(shape) for _ in range(3)], "pressure": np.zeros(shape)}
        for c in range(3):
            self.transforms.xyzfft("B", out["velocity"][c], vel, 3, c)
        self.transforms.xyzfft("B", out["pressure"], pr, 1, 0)
        if finite:
            out["finite"] = bool(np.isfinite(vel.view(np.float64)).all() and np.isfinite(pr.view(np.float64)).all())
        return out

    def write_physical(self, ntpit, time, snapshot):
        """field_P_ 0.dat (unless binary): per point the three components, then the pressure;
        field_P_ 0.bin (unless ascii): the u, v, w and pressure planes."""
        u = [np.asarray(c, np.float64).reshape(-1) for c in snapshot["velocity"]]
        p = np.asarray(snapshot["pressure"], np.float64).reshape(-1)
        if self.format != "binary":
            lines = [f"{int(ntpit)} {_fmt(time)}\n"]
            uvw = np.stack([c[:p.size] for c in u], axis=1).reshape(-1)
            lines += [_fmt(x) + "\n" for x in uvw.tolist()]
            lines += [_fmt(x) + "\n" for x in p.tolist()]
            _atomic_write(self._path("field_P_", ".dat"), "".join(lines).encode())
        if self.format == "ascii":
            return
        payload = b"".join(np.ascontiguousarray(a).astype("<f8").tobytes() for a in u + [p])
        _atomic_write(self._path("field_P_", ".bin"), _header(KIND_PHYSICAL, ntpit, time, self._physical_dims()) + payload)

    def append_ekt(self, ntpit, time, snapshot):
        """Appends "time EKT" to EKT 0.dat, EKT = sum of v^2 over the physical points."""
        v = np.asarray(snapshot["velocity"][1], np.float64).reshape(-1)
        ekt = 0.0
        for x in v.tolist():  # sequential sum, as the reference accumulates it
            ekt += x * x
        try:
            with open(self._path("EKT", ".dat"), "a") as f:
                f.write(f"{_fmt(time)} {_fmt(ekt)}\n")
        except OSError:
            raise RuntimeError("cannot append EKT output") from None
        return ekt
