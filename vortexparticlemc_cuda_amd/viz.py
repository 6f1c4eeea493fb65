"""Output formats of the reference's driver: the vorticity image and the
velocity-arrow drawing (include/viz_utils.hpp:26-132).

dump_vorticity_png    8-bit grey PNG of sum_i |w_i| W(|z - x_i|) at the pixel
                      centres ((i + 1/2)/res, (j + 1/2)/res) of the unit box,
                      pixel = clamp(128 + int(value * 127 / max|value|), 0, 255),
                      image row j = y index (viz_utils.hpp:59-80).  The field is
                      evaluated on the GPU by the library's grid kernel
                      (vpmc_grid_vorticity) instead of the reference's host
                      loop over all particles per pixel.
dump_velocity_svg     one line per particle, centred on it along its velocity,
                      of length vel_scale |v|, ends wrapped into the box,
                      y flipped (viz_utils.hpp:84-132).

encode_png / decode_png are a minimal grey-scale PNG writer and reader
(zlib, no external image library).
"""
import struct
import zlib

import numpy as np
import torch


def _chunk(tag, data):
    c = struct.pack(">I", len(data)) + tag + data
    return c + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def encode_png(img):
    """bytes of an 8-bit grey PNG of img (H x W uint8), row 0 first"""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))   # filter 0 per row
    return (b"\x89PNG\r\n\x1a\n" +
            _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0)) +
            _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b""))


def decode_png(data):
    """H x W uint8 array of an 8-bit grey, non-interlaced PNG with filter 0
    rows (what encode_png writes)"""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 0
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, w + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].copy()


def vorticity_image(pos, w, h_w, res=512, L=1.0, cells_per_hw=8):
    """the reference's 8-bit image of the |w|-weighted B-spline field (res x res)"""
    from . import adjoint
    pos = torch.as_tensor(pos, dtype=torch.float32).contiguous()
    w = torch.as_tensor(w, dtype=torch.float32).contiguous()
    if not pos.is_cuda:
        raise ValueError("device tensors required")
    with torch.no_grad():
        val = adjoint.grid_vorticity(pos, w.abs(), h_w, res, res, L, cells_per_hw)
    v = val.cpu().numpy().reshape(res, res)          # [i = x index, j = y index]
    m = float(np.abs(v).max())
    inv = np.float32(127.0 / m) if m > 0 else np.float32(1.0)
    q = 128 + np.trunc(v * inv).astype(np.int64)     # int() truncates toward zero
    return np.clip(q, 0, 255).astype(np.uint8).T     # row j = y index


def dump_vorticity_png(sim, path, res=512):
    """vorticity image of a Simulation's current state written to path"""
    img = vorticity_image(sim.pos, sim.omega, float(sim.P.h_w), res, float(sim.P.Lx),
                          sim.cells_per_hw)
    with open(path, "wb") as f:
        f.write(encode_png(img))
    return img


def velocity_svg(pos, vel, vel_scale=0.2, periodic=True):
    """SVG text of the reference's velocity arrows (unit box, y up)"""
    pos = np.asarray(pos, np.float32)
    vel = np.asarray(vel, np.float32)
    out = ['<?xml version="1.0" standalone="no"?>',
           '<svg xmlns="http://www.w3.org/2000/svg" width="600" height="600" '
           'viewBox="0 0 1 1">',
           '<style>.arrow{stroke:#ff4d00;stroke-width:0.0015;marker-end:url(#head);}</style>',
           '<defs><marker id="head" orient="auto" markerWidth="4" markerHeight="4" '
           'refX="0.1" refY="2"><path d="M0,0 V4 L4,2 Z" fill="#ff4d00"/></marker></defs>']
    ln = np.sqrt((vel.astype(np.float32) ** 2).sum(1)).astype(np.float32)
    for p, v, l in zip(pos, vel, ln):
        if l < 1e-8:
            continue
        d = v / l
        half = np.float32(0.5 * vel_scale) * l
        a = p - half * d
        b = p + half * d
        if periodic:
            a = a - np.floor(a)
            b = b - np.floor(b)
        out.append('<line class="arrow" x1="%g" y1="%g" x2="%g" y2="%g"/>'
                   % (a[0], 1 - a[1], b[0], 1 - b[1]))
    out.append("</svg>")
    return "\n".join(out) + "\n"


def dump_velocity_svg(sim, path, vel_scale=0.2, periodic=True):
    """velocity arrows of a Simulation's last step (its vel buffer)"""
    text = velocity_svg(sim.pos.cpu().numpy(), sim.vel[:sim.N].cpu().numpy(), vel_scale,
                        periodic)
    with open(path, "w") as f:
        f.write(text)
    return text
