"""SEG-Y model ingest: read (and write) SEG-Y rev 1 files into [nz][nx] model planes.

The reference loads the elastic Marmousi model with SegyIO's ``segy_read`` and uses
``Float32.(block.data)`` (src/common/read_model.jl:4-6); SegyIO is not a dependency of its
Project.toml and no version is pinned. This module restates the published SEG-Y rev 1 layout
(SEG Technical Standards Committee, 2002) that such a reader implements:

* a 3200-byte textual header (EBCDIC or ASCII);
* a 400-byte binary header, big-endian: sample interval in microseconds at bytes 3217-3218,
  samples per trace at 3221-3222, data sample format code at 3225-3226 (1 = 4-byte IBM
  hexadecimal float, 2 = int32, 3 = int16, 5 = 4-byte IEEE float, 8 = int8), the number of
  extended textual headers at 3505-3506;
* per trace a 240-byte trace header (samples in this trace at bytes 115-116, sample interval
  at 117-118; source/receiver/CDP coordinates at 73-88 and 181-188, scaler at 71-72) and the
  samples.

``segy_read`` returns a block with ``data`` of shape (samples, traces) like SegyIO's
SeisBlock. A velocity or density model stored one trace per x position (the Marmousi
convention) therefore arrives as an (nz, nx) array, which is this package's model layout.
Files written little-endian are detected by the format code.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

TEXT_BYTES = 3200
BIN_BYTES = 400
TRACE_HEADER_BYTES = 240

_FORMATS = {1: ("ibm", 4), 2: (">i4", 4), 3: (">i2", 2), 5: (">f4", 4), 8: ("i1", 1)}

# (name, 1-based first byte, size) of the trace-header words returned by segy_read
TRACE_HEADER_FIELDS = [
    ("trace_sequence_line", 1, 4), ("trace_sequence_file", 5, 4), ("field_record", 9, 4),
    ("trace_number", 13, 4), ("cdp", 21, 4), ("scalar_coord", 71, 2),
    ("source_x", 73, 4), ("source_y", 77, 4), ("group_x", 81, 4), ("group_y", 85, 4),
    ("ns", 115, 2), ("dt", 117, 2), ("cdp_x", 181, 4), ("cdp_y", 185, 4),
    ("inline", 189, 4), ("crossline", 193, 4),
]


def ibm_to_ieee(words):
    """4-byte IBM hexadecimal floats (as uint32) to float64:
    (-1)^s * 0.f * 16^(e - 64) with s the top bit, e the next 7 bits, f 24 fraction bits."""
    w = np.asarray(words, dtype=np.uint32)
    sign = np.where((w >> np.uint32(31)) != 0, -1.0, 1.0)
    expo = ((w >> np.uint32(24)) & np.uint32(0x7F)).astype(np.int64) - 64
    frac = (w & np.uint32(0x00FFFFFF)).astype(np.float64) / float(1 << 24)
    return sign * frac * np.power(16.0, expo.astype(np.float64))


def ieee_to_ibm(values):
    """float values to 4-byte IBM hexadecimal floats (uint32), fraction truncated to 24
    bits after normalising to a hexadecimal exponent (the value is exactly representable
    when it is an IEEE single with at most 21 significant bits, e.g. a small integer)."""
    v = np.asarray(values, dtype=np.float64)
    out = np.zeros(v.shape, dtype=np.uint32)
    nz = v != 0
    a = np.abs(v[nz])
    # smallest e with a < 16^e, i.e. 0.0625 <= a / 16^e < 1
    e = np.floor(np.log2(a) / 4.0).astype(np.int64) + 1
    m = a / np.power(16.0, e.astype(np.float64))
    hi = m >= 1.0
    e[hi] += 1
    m[hi] /= 16.0
    lo = m < 0.0625
    e[lo] -= 1
    m[lo] *= 16.0
    if np.any(e + 64 > 127) or np.any(e + 64 < 0):
        raise ValueError("value outside the IBM float range")
    frac = np.floor(m * float(1 << 24)).astype(np.uint64)
    word = ((np.signbit(v[nz]).astype(np.uint64) << np.uint64(31)) |
            ((e + 64).astype(np.uint64) << np.uint64(24)) | frac)
    out[nz] = word.astype(np.uint32)
    return out


@dataclass
class SegyBlock:
    """Result of segy_read: ``data`` (samples, traces), the sample interval ``dt_us`` in
    microseconds, the binary-header ``format_code``, the textual header, the raw binary
    header and selected trace-header words (one array per field)."""
    data: np.ndarray
    dt_us: int
    format_code: int
    text_header: str
    binary_header: bytes
    trace_headers: dict = field(default_factory=dict)
    little_endian: bool = False

    @property
    def ns(self):
        return self.data.shape[0]

    @property
    def ntraces(self):
        return self.data.shape[1]


def _decode_text(raw: bytes) -> str:
    """EBCDIC when EBCDIC blanks (0x40) outnumber ASCII blanks (0x20), else ASCII."""
    if raw.count(0x40) > raw.count(0x20):
        return raw.decode("cp500", errors="replace")
    return raw.decode("ascii", errors="replace")


def segy_read(path) -> SegyBlock:
    """Read a SEG-Y rev 0/1 file whose traces all have the binary header's sample count."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < TEXT_BYTES + BIN_BYTES:
        raise ValueError(f"{path}: too short for a SEG-Y file")
    binh = buf[TEXT_BYTES:TEXT_BYTES + BIN_BYTES]
    little = False
    fmt = int.from_bytes(binh[24:26], "big", signed=False)
    if fmt not in _FORMATS:
        fmt_le = int.from_bytes(binh[24:26], "little", signed=False)
        if fmt_le not in _FORMATS:
            raise ValueError(f"{path}: unsupported data sample format code {fmt}")
        fmt, little = fmt_le, True
    order = "little" if little else "big"
    dt_us = int.from_bytes(binh[16:18], order, signed=False)
    ns = int.from_bytes(binh[20:22], order, signed=False)
    n_ext = int.from_bytes(binh[304:306], order, signed=True)
    start = TEXT_BYTES + BIN_BYTES + max(n_ext, 0) * TEXT_BYTES
    kind, size = _FORMATS[fmt]
    if ns == 0 and len(buf) >= start + TRACE_HEADER_BYTES:  # fall back to the first trace
        ns = int.from_bytes(buf[start + 114:start + 116], order, signed=False)
    rec = TRACE_HEADER_BYTES + ns * size
    ntr = (len(buf) - start) // rec
    if ntr <= 0 or ns <= 0:
        raise ValueError(f"{path}: no traces")
    body = np.frombuffer(buf, dtype=np.uint8, count=ntr * rec, offset=start).reshape(ntr, rec)
    samples = np.ascontiguousarray(body[:, TRACE_HEADER_BYTES:])
    if kind == "ibm":
        words = samples.view(">u4" if not little else "<u4").astype(np.uint32)
        data = ibm_to_ieee(words)
    else:
        dt = np.dtype(kind)
        if little and dt.itemsize > 1:
            dt = dt.newbyteorder("<")
        data = samples.view(dt).astype(np.float64)
    heads = {}
    th = np.ascontiguousarray(body[:, :TRACE_HEADER_BYTES])
    for name, first, nbytes in TRACE_HEADER_FIELDS:
        raw = np.ascontiguousarray(th[:, first - 1:first - 1 + nbytes])
        dt = np.dtype((">" if not little else "<") + ("i4" if nbytes == 4 else "i2"))
        heads[name] = raw.view(dt).ravel().astype(np.int64)
    return SegyBlock(data=np.ascontiguousarray(data.T.astype(np.float32)), dt_us=dt_us,
                     format_code=fmt, text_header=_decode_text(buf[:TEXT_BYTES]),
                     binary_header=bytes(binh), trace_headers=heads, little_endian=little)


def segy_write(path, data, dt_us=1000, format_code=5, text="", little_endian=False,
               trace_headers=None):
    """Write ``data`` (samples, traces) as SEG-Y rev 1 (format 1 = IBM float, 5 = IEEE
    float, 2 = int32, 3 = int16, 8 = int8) with a textual header (EBCDIC), the binary
    header and per-trace headers (sequence numbers, ns, dt and any ``trace_headers`` words
    named as in TRACE_HEADER_FIELDS)."""
    a = np.asarray(data, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("data must be (samples, traces)")
    ns, ntr = a.shape
    if format_code not in _FORMATS:
        raise ValueError("unsupported format code")
    order = "little" if little_endian else "big"
    e = "<" if little_endian else ">"
    lines = (text or "C 1 SEG-Y written by viscoelasticinversion_amd").splitlines()
    card = "".join(f"{ln[:80]:<80}" for ln in lines)[:TEXT_BYTES]
    texth = f"{card:<{TEXT_BYTES}}".encode("cp500")
    binh = bytearray(BIN_BYTES)
    binh[16:18] = int(dt_us).to_bytes(2, order)
    binh[20:22] = int(ns).to_bytes(2, order)
    binh[24:26] = int(format_code).to_bytes(2, order)
    binh[300:302] = (0x0100).to_bytes(2, order)  # revision 1.0
    binh[302:304] = (1).to_bytes(2, order)        # fixed-length traces
    kind, size = _FORMATS[format_code]
    if kind == "ibm":
        samples = ieee_to_ibm(a.T).astype(e + "u4")
    else:
        dt = np.dtype(kind)
        if dt.itemsize > 1:
            dt = dt.newbyteorder(e)
        samples = np.rint(a.T).astype(dt) if dt.kind == "i" else a.T.astype(dt)
    th = np.zeros((ntr, TRACE_HEADER_BYTES), dtype=np.uint8)
    words = {"trace_sequence_line": np.arange(1, ntr + 1), "trace_sequence_file": np.arange(1, ntr + 1),
             "trace_number": np.arange(1, ntr + 1), "ns": np.full(ntr, ns),
             "dt": np.full(ntr, dt_us)}
    words.update(trace_headers or {})
    spec = {n: (b, s) for n, b, s in TRACE_HEADER_FIELDS}
    for name, vals in words.items():
        first, nbytes = spec[name]
        dt = np.dtype(e + ("i4" if nbytes == 4 else "i2"))
        th[:, first - 1:first - 1 + nbytes] = (
            np.asarray(vals, dtype=np.int64).astype(dt).reshape(ntr, 1).view(np.uint8))
    body = np.concatenate([th, np.ascontiguousarray(samples).view(np.uint8).reshape(ntr, ns * size)], axis=1)
    with open(path, "wb") as f:
        f.write(texth)
        f.write(bytes(binh))
        f.write(body.tobytes())
    return os.path.getsize(path)


def model_from_segy(path, step=1, window=None):
    """Load a model plane (one trace per x position, samples along depth) as an (nz, nx)
    float64 array: the reference's ``Float32.(segy_read(path).data)``. ``step`` keeps every
    step-th sample and trace (e.g. 4 turns the 1.25 m Marmousi grid into 5 m cells);
    ``window`` = (z0, z1, x0, x1) crops after the decimation."""
    a = segy_read(path).data.astype(np.float64)
    if step > 1:
        a = a[::step, ::step]
    if window is not None:
        z0, z1, x0, x1 = window
        a = a[z0:z1, x0:x1]
    return np.ascontiguousarray(a)


def isotropic_from_segy(vp_path, vs_path, rho_path=None, step=1, window=None,
                        rho_scale=1.0):
    """(rho, lam, mu) model dict from P- and S-velocity SEG-Y files and a density file
    (Gardner 310 vp^0.25 when absent; ``rho_scale`` converts g/cm^3 files to kg/m^3)."""
    vp = model_from_segy(vp_path, step, window)
    vs = model_from_segy(vs_path, step, window)
    rho = (model_from_segy(rho_path, step, window) * rho_scale if rho_path
           else 310.0 * np.power(vp, 0.25))
    mu = rho * vs ** 2
    return dict(rho=rho, lam=rho * vp ** 2 - 2.0 * mu, mu=mu, vp=vp, vs=vs)
