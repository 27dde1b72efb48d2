"""CPU restatement of the reference's .vibebin classic layout (numpy + struct).

Test infrastructure only (tests/ import it as the checker of libvibe_mi.so's
vibe_mi_vibebin_* and vibe_mi_restart_* entry points); nothing in the library
or its Python package uses it.

Follows:
  header and record bytes   src/io/binary_writer.cpp:286-301 (BinaryWriter::close)
  name -> shape rule        src/io/binary_writer.cpp:117-138
  reader checks             src/io/binary_writer.cpp:333-402 (BinaryReader::open)
  restart record sequence   src/io/field_io.cpp:207-227 (write_restart) and the
                            species table field_io.cpp:44-57

Pinned by the golden files tests/golden/vibebin_*.vibebin, written by the
reference's own BinaryWriter (tests/golden/make_vibebin_golden.py), and by the
reference's unit-test expectations (tests/unit/test_binary_io.cpp).
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"VIBEBIN1"
NAME_BYTES = 32
HEADER_BYTES = 40
RESTART_INFO = "__restart_info__"
SPECIES = ("u", "v", "w", "rho", "theta", "pi", "qv", "qc", "qr", "qi", "qs", "qg")


def shape_of(name: str, nx: int, ny: int, nz: int):
    """(ni, nj, nk) from the record name (binary_writer.cpp:117-138)."""
    if name == RESTART_INFO:
        return (2, 1, 1)
    if name in ("u", "u_facex"):
        return (nx + 1, ny, nz)
    if name in ("v", "v_facey"):
        return (nx, ny + 1, nz)
    if name in ("w", "zeta", "w_facez"):
        return (nx, ny, nz + 1)
    return (nx, ny, nz)


def encode(nx, ny, nz, records, time=0.0, real_kind=0) -> bytes:
    """File bytes for [(name, values)] in order (binary_writer.cpp:286-301)."""
    out = bytearray(MAGIC)
    out += struct.pack("<iiii", nx, ny, nz, real_kind)
    out += struct.pack("<d", time)
    out += struct.pack("<ii", len(records), nz)
    for name, data in records:
        raw = name.encode("ascii")[:NAME_BYTES - 1]
        out += raw + b"\0" * (NAME_BYTES - len(raw))
        ni, nj, nk = shape_of(raw.decode("ascii"), nx, ny, nz)
        arr = np.asarray(data, dtype="<f8").reshape(-1)
        assert arr.size == ni * nj * nk, (name, arr.size, (ni, nj, nk))
        out += arr.tobytes()
    return bytes(out)


def decode(buf: bytes):
    """(header dict, [(name, flat array)]) of a classic-layout file; raises
    ValueError where BinaryReader::open throws (binary_writer.cpp:333-402) and
    on a record that runs past the end of the data."""
    if len(buf) < HEADER_BYTES or buf[:8] != MAGIC:
        raise ValueError("not a .vibebin file")
    nx, ny, nz, real_kind = struct.unpack_from("<iiii", buf, 8)
    (time,) = struct.unpack_from("<d", buf, 24)
    nvars, nlevels = struct.unpack_from("<ii", buf, 32)
    if min(nx, ny, nz) <= 0 or nvars < 0 or nlevels < 0 or real_kind not in (0, 1):
        raise ValueError("bad header")
    off = HEADER_BYTES
    recs = []
    for _ in range(nvars):
        if off + NAME_BYTES > len(buf):
            raise ValueError("truncated name")
        name = buf[off:off + NAME_BYTES - 1].split(b"\0", 1)[0].decode("ascii")
        off += NAME_BYTES
        ni, nj, nk = shape_of(name, nx, ny, nz)
        n = ni * nj * nk
        if off + 8 * n > len(buf):
            raise ValueError("truncated record " + name)
        recs.append((name, np.frombuffer(buf, dtype="<f8", count=n, offset=off).copy()))
        off += 8 * n
    hdr = dict(nx=nx, ny=ny, nz=nz, real_kind=real_kind, time=time, nvars=nvars, nlevels=nlevels)
    return hdr, recs


def restart_records(nx, ny, nz, state, rho0, theta0, pi0, time, step, perturbations=True):
    """Record list of a dry restart: write_restart's sequence
    (field_io.cpp:207-227) on the model state of dycore.h (perturbations, u
    and v without the periodic face), totals for rho/theta/pi, zero moisture,
    then the exact perturbations."""
    def faces_x(a):
        a = np.asarray(a, dtype=np.float64).reshape(nz, ny, nx)
        return np.concatenate([a, a[:, :, :1]], axis=2)

    def faces_y(a):
        a = np.asarray(a, dtype=np.float64).reshape(nz, ny, nx)
        return np.concatenate([a, a[:, :1, :]], axis=1)

    cell = nx * ny * nz
    recs = [(RESTART_INFO, np.array([time, float(step)])),
            ("u", faces_x(state["u"])), ("v", faces_y(state["v"])),
            ("w", np.asarray(state["w"], dtype=np.float64))]
    for name, base in (("rho", rho0), ("theta", theta0), ("pi", pi0)):
        recs.append((name, np.asarray(base, dtype=np.float64).reshape(-1) +
                     np.asarray(state[name], dtype=np.float64).reshape(-1)))
    for name in SPECIES[6:]:
        recs.append((name, np.zeros(cell)))
    if perturbations:
        for name in ("rho", "theta", "pi"):
            recs.append((name + "_p", np.asarray(state[name], dtype=np.float64)))
    return recs
