"""KTX 2.0 and KTX 1.1 textures with uncompressed 8-bit texels, read into the RGBA8 arrays the
library uploads; KTX 2.0 float textures and cube maps, read into the linear float arrays of an
environment (decode_float).

The reference converts its images with the Khronos `ktx create` tool to R8G8B8A8_SRGB or
R8G8B8A8_UNORM KTX2 files and loads them with pyktx (pythonScripts/yapt/texture_loader_ktx.py:95-140,
resources.py:29-35, 82-145); neither is part of this build, so the container is parsed here from
the KTX 2.0 specification (header, level index, mip level 0). Supported: vkFormat R8, R8G8, R8G8B8,
B8G8R8, R8G8B8A8 and B8G8R8A8 in UNORM and SRGB, no supercompression, one layer and one face; KTX 1.1
with GL_UNSIGNED_BYTE RGBA, RGB, RG or RED (with an SRGB8 or SRGB8_ALPHA8 internal format for sRGB).
Block-compressed formats and supercompressed (BasisLZ, Zstandard) files are refused with an error
that names the format.

    rgba8, srgb = ktx.read("albedo.ktx2")
"""
import struct

import numpy as np

KTX2_IDENTIFIER = b"\xabKTX 20\xbb\r\n\x1a\n"
KTX1_IDENTIFIER = b"\xabKTX 11\xbb\r\n\x1a\n"

# vkFormat -> (channels, bgr order, srgb)
VK_FORMATS = {
    9: (1, False, False), 15: (1, False, True),            # R8_UNORM, R8_SRGB
    16: (2, False, False), 22: (2, False, True),           # R8G8_UNORM, R8G8_SRGB
    23: (3, False, False), 29: (3, False, True),           # R8G8B8_UNORM, R8G8B8_SRGB
    30: (3, True, False), 36: (3, True, True),             # B8G8R8_UNORM, B8G8R8_SRGB
    37: (4, False, False), 43: (4, False, True),           # R8G8B8A8_UNORM, R8G8B8A8_SRGB
    44: (4, True, False), 50: (4, True, True),             # B8G8R8A8_UNORM, B8G8R8A8_SRGB
}
SUPERCOMPRESSION = {1: "BasisLZ", 2: "Zstandard", 3: "ZLIB"}
# float vkFormats (for HDR environments) -> (numpy type, channels)
VK_FLOAT_FORMATS = {
    76: (np.float16, 1), 83: (np.float16, 2), 90: (np.float16, 3), 97: (np.float16, 4),   # R16.._SFLOAT
    100: (np.float32, 1), 103: (np.float32, 2), 106: (np.float32, 3), 109: (np.float32, 4),  # R32.._SFLOAT
}

# KTX 1: glFormat -> channels, and the sRGB internal formats
GL_RED, GL_RG, GL_RGB, GL_RGBA, GL_UNSIGNED_BYTE = 0x1903, 0x8227, 0x1907, 0x1908, 0x1401
GL_CHANNELS = {GL_RED: 1, GL_RG: 2, GL_RGB: 3, GL_RGBA: 4}
GL_SRGB8, GL_SRGB8_ALPHA8 = 0x8C41, 0x8C43


def _expand(texels, channels):
    """R -> (R, R, R, 1), RG -> (R, G, 0, 1), RGB -> (R, G, B, 1) as a sampler reads them"""
    h, w = texels.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    out[..., 3] = 255
    if channels == 1:
        out[..., 0] = out[..., 1] = out[..., 2] = texels[..., 0]
    else:
        out[..., :channels] = texels
    return out


def _read_ktx2(data):
    (vk_format, type_size, width, height, depth, layers, faces, levels,
     scheme) = struct.unpack_from("<9I", data, 12)
    if scheme:
        raise ValueError("KTX2 supercompression %s is not supported" % SUPERCOMPRESSION.get(scheme, scheme))
    if vk_format not in VK_FORMATS:
        raise ValueError("KTX2 vkFormat %d is not supported (uncompressed 8-bit formats only)" % vk_format)
    if depth > 1 or layers > 1 or faces != 1:
        raise ValueError("only 2D KTX2 textures are supported (depth %d, layers %d, faces %d)" % (depth, layers, faces))
    if width == 0 or height == 0:
        raise ValueError("KTX2 texture without texels")
    channels, bgr, srgb = VK_FORMATS[vk_format]
    # level index after the 80-byte header: (byteOffset, byteLength, uncompressedByteLength) per level
    offset, length, _ = struct.unpack_from("<3Q", data, 80)
    row = width * channels
    if length < row * height or offset + row * height > len(data):
        raise ValueError("KTX2 level 0 is shorter than %d x %d texels" % (width, height))
    texels = np.frombuffer(data, np.uint8, row * height, offset).reshape(height, width, channels)
    if bgr:
        texels = texels[..., [2, 1, 0] + ([3] if channels == 4 else [])]
    return _expand(texels, channels), srgb


def _srgb_to_linear(c):
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def decode_float(data):
    """(texels, cube) from the bytes of a KTX2 file, for an environment: linear float32 H x W x 4, or
    6 x N x N x 4 when the file holds the six faces of a cube map (faceCount 6, in the order +X, -X, +Y,
    -Y, +Z, -Z). _SFLOAT formats keep their values; 8-bit formats are decoded (sRGB or unorm) to
    linear. Mip level 0 only."""
    data = bytes(data)
    if data[:12] != KTX2_IDENTIFIER:
        raise ValueError("not a KTX2 file")
    (vk_format, type_size, width, height, depth, layers, faces, levels,
     scheme) = struct.unpack_from("<9I", data, 12)
    if scheme:
        raise ValueError("KTX2 supercompression %s is not supported" % SUPERCOMPRESSION.get(scheme, scheme))
    if depth > 1 or layers > 1 or faces not in (1, 6):
        raise ValueError("only 2D and cube-map KTX2 textures are supported (depth %d, layers %d, faces %d)" % (depth, layers, faces))
    if width == 0 or height == 0:
        raise ValueError("KTX2 texture without texels")
    offset, length, _ = struct.unpack_from("<3Q", data, 80)
    if vk_format in VK_FLOAT_FORMATS:
        dt, channels = VK_FLOAT_FORMATS[vk_format]
        size = np.dtype(dt).itemsize
        n = faces * height * width * channels
        if length < n * size or offset + n * size > len(data):
            raise ValueError("KTX2 level 0 is shorter than %d face(s) of %d x %d texels" % (faces, width, height))
        t = np.frombuffer(data, dt, n, offset).astype(np.float32).reshape(faces, height, width, channels)
        out = np.zeros((faces, height, width, 4), np.float32)
        out[..., 3] = 1.0
        if channels == 1:
            out[..., :3] = t
        else:
            out[..., :channels] = t
    elif vk_format in VK_FORMATS:
        channels, bgr, srgb = VK_FORMATS[vk_format]
        n = faces * height * width * channels
        if length < n or offset + n > len(data):
            raise ValueError("KTX2 level 0 is shorter than %d face(s) of %d x %d texels" % (faces, width, height))
        t = np.frombuffer(data, np.uint8, n, offset).reshape(faces, height, width, channels)
        if bgr:
            t = t[..., [2, 1, 0] + ([3] if channels == 4 else [])]
        out = np.stack([_expand(f, channels) for f in t]).astype(np.float64) / 255.0
        if srgb:
            out[..., :3] = _srgb_to_linear(out[..., :3])
        out = out.astype(np.float32)
    else:
        raise ValueError("KTX2 vkFormat %d is not supported (uncompressed 8-bit, 16- and 32-bit float formats)" % vk_format)
    if faces == 6:
        if width != height:
            raise ValueError("cube-map faces must be square")
        return out, True
    return out[0], False


def read_float(path):
    with open(path, "rb") as f:
        return decode_float(f.read())


def _read_ktx1(data):
    endian = struct.unpack_from("<I", data, 12)[0]
    fmt = "<" if endian == 0x04030201 else ">"
    (gl_type, type_size, gl_format, internal, base_internal, width, height, depth, elements, faces, levels,
     kv_bytes) = struct.unpack_from(fmt + "12I", data, 16)
    if gl_type != GL_UNSIGNED_BYTE or gl_format not in GL_CHANNELS:
        raise ValueError("KTX1 glType 0x%x glFormat 0x%x is not supported (uncompressed bytes only)" % (gl_type, gl_format))
    if depth > 1 or elements > 1 or faces != 1:
        raise ValueError("only 2D KTX1 textures are supported")
    channels = GL_CHANNELS[gl_format]
    off = 64 + kv_bytes
    size = struct.unpack_from(fmt + "I", data, off)[0]
    row = width * channels
    pitch = (row + 3) & ~3                       # rows are padded to 4 bytes
    if size < pitch * (height - 1) + row or off + 4 + size > len(data):
        raise ValueError("KTX1 level 0 is shorter than %d x %d texels" % (width, height))
    raw = np.frombuffer(data, np.uint8, size, off + 4)
    texels = np.stack([raw[y * pitch: y * pitch + row] for y in range(height)]).reshape(height, width, channels)
    return _expand(texels, channels), internal in (GL_SRGB8, GL_SRGB8_ALPHA8)


def decode(data):
    """(H x W x 4 uint8, srgb) from the bytes of a KTX2 or KTX1 file: mip level 0, row 0 first"""
    data = bytes(data)
    if data[:12] == KTX2_IDENTIFIER:
        return _read_ktx2(data)
    if data[:12] == KTX1_IDENTIFIER:
        return _read_ktx1(data)
    raise ValueError("not a KTX file")


def read(path):
    with open(path, "rb") as f:
        return decode(f.read())


def is_ktx(data):
    return bytes(data[:12]) in (KTX2_IDENTIFIER, KTX1_IDENTIFIER)
