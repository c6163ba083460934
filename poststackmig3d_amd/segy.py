"""Minimal SEG-Y (rev 0/1, big-endian) reader and writer for post-stack cubes.

The layout follows what the reference reads and writes (src/segy_utils.cpp):
3200-byte text header, 400-byte binary header with the sample interval in
microseconds at byte 16, samples per trace at byte 20 and the format code at
byte 24 (:59-92); fixed-length traces of a 240-byte header and n_t 4-byte
samples; INLINE_3D / CROSSLINE_3D as big-endian int32 at trace-header bytes
188 / 192 (:177-185); format 1 = IBM float, 5 = IEEE float (:243-318).
"""
from __future__ import annotations

import numpy as np

TEXT = 3200
BINARY = 400
TRACE_HEADER = 240


def ieee_to_ibm(x: np.ndarray) -> np.ndarray:
    """float32 -> IBM single (uint32), truncating the fraction as
    segy_utils.cpp:257-270 does (repeated /16 or *16 in float32)."""
    x = np.asarray(x, np.float32).ravel()
    out = np.zeros(x.size, np.uint32)
    for n, v in enumerate(x):
        v = np.float32(v)
        if v == 0.0:
            continue
        sign = np.uint32(0x80000000) if v < 0 else np.uint32(0)
        if v < 0:
            v = np.float32(-v)
        e = 64
        while v < np.float32(1.0):
            v = np.float32(v * np.float32(16.0))
            e -= 1
        while v >= np.float32(1.0):
            v = np.float32(v / np.float32(16.0))
            e += 1
        frac = np.uint32(int(np.float32(v * np.float32(16777216.0)))) & np.uint32(0x00FFFFFF)
        out[n] = sign | np.uint32((e & 0xFF) << 24) | frac
    return out


def ibm_to_ieee(u: np.ndarray) -> np.ndarray:
    """IBM single (uint32) -> float32 (segy_utils.cpp:243-255)."""
    u = np.asarray(u, np.uint32)
    sign = np.where((u >> 31) & 1, -1.0, 1.0)
    expo = ((u >> 24) & 0x7F).astype(np.int64) - 64
    frac = (u & 0x00FFFFFF).astype(np.float64) / float(0x01000000)
    val = sign * np.ldexp(frac, 4 * expo)
    val[u == 0] = 0.0
    return val.astype(np.float32)


def write(path, traces, inlines, crosslines, dt_us=2000, fmt=5, text=None, order=None,
          extra_headers=None):
    """Writes traces [n][n_t] with labels inlines[n], crosslines[n].
    order: optional permutation of the traces in the file."""
    traces = np.asarray(traces, np.float32)
    n, n_t = traces.shape
    th = bytearray(TEXT)
    msg = (text or "C 1 synthetic post-stack cube").encode("ascii")[:TEXT]
    th[:len(msg)] = msg
    bh = bytearray(BINARY)
    bh[16:18] = int(dt_us).to_bytes(2, "big")
    bh[20:22] = int(n_t).to_bytes(2, "big")
    bh[24:26] = int(fmt).to_bytes(2, "big")
    idx = np.arange(n) if order is None else np.asarray(order)
    if fmt == 5 and extra_headers is None:
        # IEEE samples and plain headers: the whole trace block at once
        body = np.zeros((n, TRACE_HEADER + 4 * n_t), np.uint8)
        body[:, 114:116] = np.frombuffer(int(n_t).to_bytes(2, "big"), np.uint8)
        body[:, 116:118] = np.frombuffer(int(dt_us).to_bytes(2, "big"), np.uint8)
        body[:, 188:192] = np.asarray(inlines, ">i4")[idx].view(np.uint8).reshape(n, 4)
        body[:, 192:196] = np.asarray(crosslines, ">i4")[idx].view(np.uint8).reshape(n, 4)
        body[:, TRACE_HEADER:] = traces[idx].astype(">f4").view(np.uint8).reshape(n, 4 * n_t)
        with open(path, "wb") as f:
            f.write(bytes(th))
            f.write(bytes(bh))
            body.tofile(f)
        return
    with open(path, "wb") as f:
        f.write(bytes(th))
        f.write(bytes(bh))
        for k in idx:
            h = bytearray(TRACE_HEADER)
            h[114:116] = int(n_t).to_bytes(2, "big")
            h[116:118] = int(dt_us).to_bytes(2, "big")
            h[188:192] = int(inlines[k]).to_bytes(4, "big", signed=True)
            h[192:196] = int(crosslines[k]).to_bytes(4, "big", signed=True)
            if extra_headers is not None:
                for off, val in extra_headers(k):
                    h[off:off + 4] = int(val).to_bytes(4, "big", signed=True)
            f.write(bytes(h))
            if fmt == 1:
                f.write(ieee_to_ibm(traces[k]).astype(">u4").tobytes())
            else:
                f.write(traces[k].astype(">f4").tobytes())


def write_cube(path, cube, il_labels=None, xl_labels=None, dt_us=2000, fmt=5, **kw):
    """Writes a regular cube [n_il][n_xl][n_t] (labels 1..n by default)."""
    n_il, n_xl, n_t = cube.shape
    ill = np.arange(1, n_il + 1) if il_labels is None else np.asarray(il_labels)
    xll = np.arange(1, n_xl + 1) if xl_labels is None else np.asarray(xl_labels)
    il = np.repeat(ill, n_xl)
    xl = np.tile(xll, n_il)
    write(path, cube.reshape(n_il * n_xl, n_t), il, xl, dt_us=dt_us, fmt=fmt, **kw)


def read(path):
    """Returns dict(dt_us, n_t, fmt, text, binary, headers [n][240] uint8,
    inlines, crosslines, traces [n][n_t] float32) in file order."""
    raw = np.fromfile(path, np.uint8)
    text = raw[:TEXT].tobytes()
    binary = raw[TEXT:TEXT + BINARY].tobytes()
    dt_us = int.from_bytes(binary[16:18], "big")
    n_t = int.from_bytes(binary[20:22], "big")
    fmt = int.from_bytes(binary[24:26], "big")
    tr_len = TRACE_HEADER + 4 * n_t
    n = (raw.size - TEXT - BINARY) // tr_len
    body = raw[TEXT + BINARY:TEXT + BINARY + n * tr_len].reshape(n, tr_len)
    hdr = body[:, :TRACE_HEADER]
    samples = body[:, TRACE_HEADER:].copy().view(">u4").astype(np.uint32)
    if fmt == 1:
        traces = ibm_to_ieee(samples.ravel()).reshape(n, n_t)
    else:
        traces = samples.view(np.float32).reshape(n, n_t)
    il = hdr[:, 188:192].copy().view(">i4").astype(np.int32).ravel()
    xl = hdr[:, 192:196].copy().view(">i4").astype(np.int32).ravel()
    return dict(dt_us=dt_us, n_t=n_t, fmt=fmt, text=text, binary=binary, headers=hdr,
                inlines=il, crosslines=xl, traces=traces)


def to_cube(seg):
    """Regularises a read file into ([n_il][n_xl][n_t], il_labels, xl_labels),
    dropping traces whose il or xl label is 0 and zero-filling missing traces
    (segy_utils.cpp:177-214, 364-367).  Later duplicates win."""
    il, xl, tr = seg["inlines"], seg["crosslines"], seg["traces"]
    keep = (il != 0) & (xl != 0)
    ill = np.unique(il[keep])
    xll = np.unique(xl[keep])
    cube = np.zeros((ill.size, xll.size, seg["n_t"]), np.float32)
    ii = np.searchsorted(ill, il[keep])
    xx = np.searchsorted(xll, xl[keep])
    cube[ii, xx] = tr[keep]
    return cube, ill.astype(np.int32), xll.astype(np.int32)
