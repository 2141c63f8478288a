"""HDR images for environment textures: Radiance RGBE (.hdr) and OpenEXR (.exr) scanline files, read into
linear float32 RGBA arrays (row 0 at the top).

The reference loads its environment images as R16G16B16A16_SFLOAT textures after converting them with
the Khronos `ktx create` tool (pythonScripts/yapt/texture_loader_ktx.py:95-175, resources.py:29-43);
neither the tool nor an EXR library is part of this build, so the two common HDR containers are parsed
here from their published layouts:

* Radiance RGBE (Greg Ward, "Real Pixels", Graphics Gems II): a text header ending in a blank line, a
  resolution line "-Y H +X W", then scanlines of 4-byte RGBE texels, flat or in the adaptive run-length
  form (a 2, 2, width marker, then each channel as runs: count > 128 repeats one byte count - 128
  times, else count literal bytes). Value = mantissa * 2^(E - 136).
* OpenEXR (the OpenEXR file layout document): magic 20000630, a header of attributes (channels,
  compression, dataWindow, lineOrder), a table of chunk offsets, then chunks of scanlines with the
  channels stored one after another, in the order the channel list names them. Compression NONE, RLE,
  ZIPS and ZIP (zlib, then a delta predictor and two interleaved halves) with HALF or FLOAT channels;
  other compressions (PIZ, PXR24, B44, DWA) and tiled files are refused with an error naming them.

    rgba = hdr_images.read("studio.hdr")     # H x W x 4 float32, alpha 1 unless the file has it
"""
import os
import re
import struct
import zlib

import numpy as np

EXR_MAGIC = 20000630
EXR_COMPRESSION = {0: "NONE", 1: "RLE", 2: "ZIPS", 3: "ZIP", 4: "PIZ", 5: "PXR24", 6: "B44", 7: "B44A", 8: "DWAA", 9: "DWAB"}
EXR_LINES_PER_CHUNK = {0: 1, 1: 1, 2: 1, 3: 16}
EXR_PIXEL = {0: (np.uint32, 4), 1: (np.float16, 2), 2: (np.float32, 4)}


# ---------------------------------------------------------------- Radiance RGBE
def _rgbe_to_float(rgbe):
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    out = np.ones(rgbe.shape[:-1] + (4,), np.float32)
    out[..., :3] = rgbe[..., :3].astype(np.float64) * scale[..., None]
    return out


def decode_hdr(data):
    data = bytes(data)
    if not (data.startswith(b"#?RADIANCE") or data.startswith(b"#?RGBE")):
        raise ValueError("not a Radiance HDR file")
    pos = 0
    while True:   # header lines up to the blank line
        end = data.index(b"\n", pos)
        line = data[pos:end].strip()
        pos = end + 1
        if line.startswith(b"FORMAT=") and line != b"FORMAT=32-bit_rle_rgbe":
            raise ValueError("Radiance format %s is not supported" % line[7:].decode(errors="replace"))
        if not line:
            break
    end = data.index(b"\n", pos)
    m = re.match(rb"-Y (\d+) \+X (\d+)", data[pos:end].strip())
    if not m:
        raise ValueError("only the standard -Y H +X W orientation is supported")
    h, w = int(m.group(1)), int(m.group(2))
    pos = end + 1
    buf = np.frombuffer(data, np.uint8)
    img = np.empty((h, w, 4), np.uint8)
    for y in range(h):
        if pos + 4 > len(buf):
            raise ValueError("Radiance HDR scanline %d is cut short" % y)
        rle = 8 <= w < 32768 and buf[pos] == 2 and buf[pos + 1] == 2 and (int(buf[pos + 2]) << 8 | int(buf[pos + 3])) == w
        if not rle:
            n = w * 4
            if pos + n > len(buf):
                raise ValueError("Radiance HDR scanline %d is cut short" % y)
            img[y] = buf[pos:pos + n].reshape(w, 4)
            pos += n
            continue
        pos += 4
        for c in range(4):
            x = 0
            row = img[y, :, c]
            while x < w:
                count = int(buf[pos])
                pos += 1
                if count > 128:
                    count -= 128
                    if x + count > w:
                        raise ValueError("Radiance HDR run overflows scanline %d" % y)
                    row[x:x + count] = buf[pos]
                    pos += 1
                else:
                    if count == 0 or x + count > w:
                        raise ValueError("Radiance HDR run overflows scanline %d" % y)
                    row[x:x + count] = buf[pos:pos + count]
                    pos += count
                x += count
    return _rgbe_to_float(img)


# ---------------------------------------------------------------- OpenEXR
def _exr_header(data):
    if struct.unpack_from("<i", data, 0)[0] != EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    version = struct.unpack_from("<I", data, 4)[0]
    if version & 0x200:
        raise ValueError("tiled OpenEXR files are not supported")
    if version & 0x1800:
        raise ValueError("multi-part or deep OpenEXR files are not supported")
    pos, attrs = 8, {}
    while data[pos] != 0:
        name_end = data.index(b"\0", pos)
        name = data[pos:name_end].decode()
        type_end = data.index(b"\0", name_end + 1)
        kind = data[name_end + 1:type_end].decode()
        size = struct.unpack_from("<i", data, type_end + 1)[0]
        value = data[type_end + 5:type_end + 5 + size]
        attrs[name] = (kind, value)
        pos = type_end + 5 + size
    return attrs, pos + 1


def _exr_channels(value):
    out, pos = [], 0
    while value[pos] != 0:
        end = value.index(b"\0", pos)
        name = value[pos:end].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", value, end + 1)
        if xs != 1 or ys != 1:
            raise ValueError("sub-sampled OpenEXR channels are not supported")
        out.append((name, ptype))
        pos = end + 1 + 16
    return out


def _exr_unpredict(raw):
    """undo the ZIP / RLE preprocessing: running sum of (byte - 128), then the two interleaved halves"""
    t = np.frombuffer(raw, np.uint8).astype(np.int64)
    if len(t) == 0:
        return b""
    d = np.concatenate([t[:1], t[1:] - 128])
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    half = (len(t) + 1) // 2
    out = np.empty(len(t), np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _exr_rle(raw, expected):
    out, pos, src = bytearray(), 0, raw
    while pos < len(src) and len(out) < expected:
        count = struct.unpack_from("<b", src, pos)[0]
        pos += 1
        if count < 0:
            out += src[pos:pos - count]
            pos -= count
        else:
            out += bytes([src[pos]]) * (count + 1)
            pos += 1
    return bytes(out)


def decode_exr(data):
    data = bytes(data)
    attrs, pos = _exr_header(data)
    for need in ("channels", "compression", "dataWindow"):
        if need not in attrs:
            raise ValueError("OpenEXR header without %s" % need)
    comp = attrs["compression"][1][0]
    if comp not in EXR_LINES_PER_CHUNK:
        raise ValueError("OpenEXR compression %s is not supported" % EXR_COMPRESSION.get(comp, comp))
    chans = _exr_channels(attrs["channels"][1])
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    if w <= 0 or h <= 0:
        raise ValueError("empty OpenEXR data window")
    lines = EXR_LINES_PER_CHUNK[comp]
    nchunks = (h + lines - 1) // lines
    offsets = struct.unpack_from("<%dQ" % nchunks, data, pos)
    planes = {name: np.zeros((h, w), np.float32) for name, _ in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        raw = data[off + 8:off + 8 + size]
        ny = min(lines, y1 - y + 1)
        expected = ny * w * sum(EXR_PIXEL[p][1] for _, p in chans)
        if comp in (2, 3) and size < expected:
            raw = _exr_unpredict(zlib.decompress(raw))
        elif comp == 1 and size < expected:
            raw = _exr_unpredict(_exr_rle(raw, expected))
        if len(raw) != expected:
            raise ValueError("OpenEXR chunk at line %d holds %d bytes, %d expected" % (y, len(raw), expected))
        p = 0
        for r in range(ny):
            for name, ptype in chans:
                dt, sz = EXR_PIXEL[ptype]
                planes[name][y - y0 + r] = np.frombuffer(raw, dt, w, p).astype(np.float32)
                p += w * sz
    out = np.ones((h, w, 4), np.float32)
    names = {n.split(".")[-1].upper(): n for n, _ in chans}
    for k, c in enumerate("RGBA"):
        if c in names:
            out[..., k] = planes[names[c]]
        elif k < 3 and "Y" in names:      # luminance-only image
            out[..., k] = planes[names["Y"]]
    return out     # each chunk names its first line, so any line order lands top-down


def read(path):
    """H x W x 4 float32 linear RGBA of a .hdr or .exr file (or a .npy array of that shape)"""
    ext = os.path.splitext(str(path))[1].lower()
    if ext == ".npy":
        a = np.load(path, allow_pickle=False).astype(np.float32)
        if a.ndim != 3 or a.shape[2] not in (3, 4):
            raise ValueError("a .npy environment is H x W x 3 or 4")
        return a if a.shape[2] == 4 else np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], 2)
    with open(path, "rb") as f:
        data = f.read()
    if ext == ".exr" or data[:4] == struct.pack("<i", EXR_MAGIC):
        return decode_exr(data)
    return decode_hdr(data)


def encode_hdr(rgb, rle=True):
    """a Radiance RGBE file of an H x W x 3 (or 4) float image; adaptive run-length scanlines when rle"""
    rgb = np.asarray(rgb, np.float64)[..., :3]
    h, w = rgb.shape[:2]
    m = rgb.max(axis=2)
    e = np.zeros((h, w), np.int32)
    mant, ex = np.frexp(m)
    ok = m > 1e-32
    e[ok] = ex[ok]
    scale = np.where(ok, 256.0 / np.ldexp(1.0, e), 0.0)
    out = np.zeros((h, w, 4), np.uint8)
    out[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0)
    body = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
    for y in range(h):
        if not rle or not 8 <= w < 32768:
            body += out[y].tobytes()
            continue
        body += bytes([2, 2, w >> 8, w & 0xFF])
        for c in range(4):
            row, x = out[y, :, c], 0
            while x < w:     # runs of equal bytes of 4 or more, literals otherwise
                run = 1
                while x + run < w and run < 127 and row[x + run] == row[x]:
                    run += 1
                if run >= 4:
                    body += bytes([128 + run, row[x]])
                    x += run
                    continue
                n = 1
                while x + n < w and n < 128:
                    r = 1
                    while x + n + r < w and r < 4 and row[x + n + r] == row[x + n]:
                        r += 1
                    if r >= 4:
                        break
                    n += 1
                body += bytes([n]) + row[x:x + n].tobytes()
                x += n
    return bytes(body)
