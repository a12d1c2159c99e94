"""On-disk formats on either side of the solver (wos/io.py): 2D / 3D OBJ
meshes with per-vertex values (`v`, `l` / `f`, `c` lines), OpenEXR images
and sRGB transfer functions.

The reference reads and writes EXR through imageio's FreeImage plugin, which
it downloads at import time (wos/io.py:10); nothing is downloaded here: EXR
files are written and read by this module itself -- single-part scanline
files, uncompressed, FLOAT or HALF channels (OpenEXR file layout: magic,
version, attribute header, line offset table, one block per scanline with
the channels in alphabetical order).
"""
from __future__ import annotations

import struct

import numpy as np

# ---- OBJ (wos/io.py:12-85) --------------------------------------------------


def write_obj(vertices, indices, filename):
    with open(filename, "w") as f:
        for v in np.asarray(vertices):
            f.write("v %f %f %f\n" % (v[0], v[1], v[2]))
        for i in np.asarray(indices):
            f.write("f %d %d %d\n" % (i[0] + 1, i[1] + 1, i[2] + 1))


def read_2d_obj(filename, flip_orientation=False):
    vertices, indices, values = [], [], []
    with open(filename) as f:
        for line in f:
            if line.startswith("v "):
                vertices.append([float(x) for x in line.split()[1:]][:2])
            elif line.startswith("l "):
                l = [int(x) - 1 for x in line.split()[1:]]
                indices.append((l[::-1] if flip_orientation else l)[:2])
            elif line.startswith("c "):
                values.append(float(line.split()[1]))
    return np.array(vertices), np.array(indices), np.array(values)


def write_2d_obj(filename, vertices, indices, values):
    with open(filename, "w") as f:
        for v in np.asarray(vertices):
            f.write("v %f %f\n" % (v[0], v[1]))
        for i in np.asarray(indices):
            f.write("l %d %d\n" % (i[0] + 1, i[1] + 1))
        for v in np.asarray(values).reshape(-1):
            f.write("c %f\n" % v)


def read_3d_obj(filename):
    vertices, indices, values = [], [], []
    with open(filename) as f:
        for line in f:
            if line.startswith("v "):
                vertices.append([float(x) for x in line.split()[1:]])
            elif line.startswith("f "):
                indices.append([int(x.split("/")[0]) - 1 for x in line.split()[1:]])
            elif line.startswith("c "):
                values.append(float(line.split()[1]))
    return np.array(vertices), np.array(indices), np.array(values)


def write_3d_obj(filename, vertices, indices, values):
    with open(filename, "w") as f:
        for v in np.asarray(vertices):
            f.write("v %f %f %f\n" % (v[0], v[1], v[2]))
        for i in np.asarray(indices):
            f.write("f %d %d %d\n" % (i[0] + 1, i[1] + 1, i[2] + 1))
        for v in np.asarray(values).reshape(-1):
            f.write("c %f\n" % v)


# ---- colour (wos/io.py:87-116) -------------------------------------------------


def linear_to_srgb(l):
    l = np.asarray(l, dtype=np.float64)
    return np.where(l <= 0.0031308, 12.92 * l, 1.055 * np.power(np.maximum(l, 0.0031308), 1.0 / 2.4) - 0.055)


def srgb_to_linear(s):
    s = np.asarray(s, dtype=np.float64)
    return np.where(s <= 0.04045, s / 12.92, np.power((np.maximum(s, 0.04045) + 0.055) / 1.055, 2.4))


def to_srgb(image):
    return np.clip(linear_to_srgb(to_numpy(image)), 0.0, 1.0)


def to_linear(image):
    return srgb_to_linear(to_numpy(image))


def to_numpy(data):
    try:
        import torch
        if isinstance(data, torch.Tensor):
            return data.detach().cpu().numpy()
    except ImportError:   # pragma: no cover
        pass
    return np.asarray(data)


# ---- OpenEXR -----------------------------------------------------------------
_MAGIC = 20000630
_HALF, _FLOAT = 1, 2


def _attr(name, typ, payload):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def write_exr(exr_path, image, half=False):
    """Writes an [H, W], [H, W, 1], [H, W, 3] (R, G, B) or [H, W, 4] (R, G, B,
    A) float image as an uncompressed scanline OpenEXR file."""
    img = to_numpy(image).astype(np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    h, w, c = img.shape
    names = {1: ["Y"], 3: ["R", "G", "B"], 4: ["R", "G", "B", "A"]}.get(c)
    if names is None:
        raise ValueError(f"unsupported channel count {c}")
    order = sorted(range(c), key=lambda k: names[k])
    ptype = _HALF if half else _FLOAT
    chl = b"".join(names[k].encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for k in order) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (_attr("channels", "chlist", chl) + _attr("compression", "compression", b"\0") +
              _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box) +
              _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
              _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    start = 8 + len(header) + 8 * h
    dt = np.float16 if half else np.float32
    line_bytes = w * c * np.dtype(dt).itemsize
    blocks, offsets = [], []
    for y in range(h):
        offsets.append(start + y * (8 + line_bytes))
        data = b"".join(img[y, :, k].astype("<" + np.dtype(dt).str[1:]).tobytes() for k in order)
        blocks.append(struct.pack("<ii", y, len(data)) + data)
    with open(exr_path, "wb") as f:
        f.write(struct.pack("<ii", _MAGIC, 2))
        f.write(header)
        f.write(struct.pack("<%dQ" % h, *offsets))
        f.write(b"".join(blocks))


def read_exr(exr_path):
    """Reads a single-part, uncompressed scanline OpenEXR file into an
    [H, W, C] float32 array with the channels in (R, G, B, A) / Y order."""
    with open(exr_path, "rb") as f:
        buf = f.read()
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != _MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0x200 or version & 0x1000:
        raise ValueError("tiled or multi-part OpenEXR files are not supported")
    pos, attrs = 8, {}
    while buf[pos] != 0:
        e = buf.index(b"\0", pos)
        name = buf[pos:e].decode()
        e2 = buf.index(b"\0", e + 1)
        typ = buf[e + 1:e2].decode()
        (size,) = struct.unpack_from("<i", buf, e2 + 1)
        attrs[name] = (typ, buf[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    if attrs["compression"][1][0] != 0:
        raise ValueError("compressed OpenEXR files are not supported")
    chans, cp, raw = [], 0, attrs["channels"][1]
    while raw[cp] != 0:
        e = raw.index(b"\0", cp)
        nm = raw[cp:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", raw, e + 1)
        chans.append((nm, ptype))
        cp = e + 1 + 16
    x0, y0, x1, y1 = struct.unpack_from("<iiii", attrs["dataWindow"][1], 0)
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from("<%dQ" % h, buf, pos)
    out = {}
    for nm, _ in chans:
        out[nm] = np.zeros((h, w), np.float32)
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        p = off + 8
        for nm, ptype in chans:
            dt = np.dtype("<f2") if ptype == _HALF else np.dtype("<f4")
            if ptype not in (_HALF, _FLOAT):
                raise ValueError("only HALF and FLOAT channels are supported")
            out[nm][y - y0] = np.frombuffer(buf, dt, w, p).astype(np.float32)
            p += w * dt.itemsize
    names = [n for n in ("R", "G", "B", "A", "Y") if n in out] or sorted(out)
    return np.stack([out[n] for n in names], axis=-1)


def write_image(image_path, image, is_srgb=None):
    """wos/io.py:214-252 for the formats available here (.exr, .npy)."""
    path = str(image_path)
    img = to_numpy(image)
    if path.endswith(".exr"):
        write_exr(path, img)
    elif path.endswith(".npy"):
        np.save(path, img)
    else:
        raise ValueError(f"unsupported image format: {path} (use .exr or .npy)")


def read_image(image_path, is_srgb=None):
    path = str(image_path)
    if path.endswith(".exr"):
        return read_exr(path)
    if path.endswith(".npy"):
        return np.load(path)
    raise ValueError(f"unsupported image format: {path} (use .exr or .npy)")
