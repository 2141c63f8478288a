"""Generates the constant tables compiled into the path-tracing library.

* ``sobol_matrices.inc``: 513 Sobol generator matrices of 32 rows (uint32). Dimension 0 is the
  bit-reversal (van der Corput) matrix; dimensions 1..512 are built from the Joe & Kuo primitive
  polynomials and initial direction numbers in ``joe_kuo_first513.txt`` with the Bratley-Fox
  recurrence  v_i = v_{i-s} ^ (v_{i-s} >> s) ^ XOR_k a_k v_{i-k}  (the construction the reference
  documents in Tools/generate_sobol_matrices.py:37-90 and ships as Source/Common/inc/Math/SobolMatrices.h).
* ``cie_tables.inc``: CIE 1931 2-degree colour matching functions and the D65 illuminant on the
  380..700 nm, 1 nm grid the reference samples (spectralConstants.h:19-26; SpectralUtility.cpp:97-134),
  as float32, plus the RGB<->XYZ matrices.

Usage: ``python -m yapt_amd.data.gen_tables`` (writes into ``yapt_amd/csrc/data``).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "..", "csrc", "data")

SOBOL_BITS = 32
SOBOL_DIMS = 512  # plus the implicit first dimension

CIE_OFFSET = 20      # 380 nm in the 360..830 table
CIE_RES = 321        # 380..700 nm


def load_joe_kuo(path):
    rows = []
    with open(path) as f:
        header_seen = False
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if not header_seen:
                if t[:4] != ["d", "s", "a", "m_i"]:
                    raise ValueError("unexpected Joe-Kuo header: %r" % line)
                header_seen = True
                continue
            d, s, a = int(t[0]), int(t[1]), int(t[2])
            m = [int(x) for x in t[3:3 + s]]
            rows.append((d, s, a, m))
    return rows


def direction_numbers(s, a, m, bits=SOBOL_BITS):
    """Direction numbers v_0..v_{bits-1} (left aligned) of one Sobol dimension."""
    mask = (1 << bits) - 1
    v = [0] * bits
    for i in range(min(s, bits)):
        v[i] = (m[i] << (bits - 1 - i)) & mask
    for i in range(s, bits):
        x = v[i - s] ^ (v[i - s] >> s)
        for k in range(1, s):
            if (a >> (s - 1 - k)) & 1:
                x ^= v[i - k]
        v[i] = x & mask
    return v


def sobol_matrices(joe_kuo_path, dims=SOBOL_DIMS, bits=SOBOL_BITS):
    rows = load_joe_kuo(joe_kuo_path)
    if len(rows) < dims:
        raise ValueError("direction file holds %d dimensions, %d requested" % (len(rows), dims))
    mats = [[1 << (bits - 1 - j) for j in range(bits)]]
    for d, s, a, m in rows[:dims]:
        mats.append(direction_numbers(s, a, m, bits))
    return np.array(mats, dtype=np.uint32)


def write_sobol_inc(mats, path):
    with open(path, "w") as f:
        f.write("// generated by yapt_amd/data/gen_tables.py from joe_kuo_first513.txt; do not edit\n")
        f.write("#define YPT_SOBOL_DIMS %d\n#define YPT_SOBOL_BITS %d\n" % mats.shape)
        f.write("static const unsigned int YPT_SOBOL_MATRICES[%d] = {\n" % mats.size)
        for row in mats:
            f.write("  " + ",".join("0x%08xu" % int(v) for v in row) + ",\n")
        f.write("};\n")


def cie_tables(npz_path):
    d = np.load(npz_path)
    lam = d["lambda"][CIE_OFFSET:CIE_OFFSET + CIE_RES]
    assert lam[0] == 380 and lam[-1] == 700
    xyz = np.stack([d["x"], d["y"], d["z"]], axis=1)[CIE_OFFSET:CIE_OFFSET + CIE_RES].astype(np.float32)
    d65 = d["d65"][CIE_OFFSET:CIE_OFFSET + CIE_RES].astype(np.float32)
    mats = {k: d[k].astype(np.float32) for k in ("srgb_xyz_to_rgb", "srgb_rgb_to_xyz",
                                                  "rec2020_xyz_to_rgb", "rec2020_rgb_to_xyz")}
    return xyz, d65, mats


def _flt(v):
    return "%.9ef" % float(v)


def write_cie_inc(npz_path, path):
    xyz, d65, mats = cie_tables(npz_path)
    d = np.load(npz_path)
    sl = slice(CIE_OFFSET, CIE_OFFSET + CIE_RES)
    with open(path, "w") as f:
        f.write("// generated by yapt_amd/data/gen_tables.py from cie_1931_2deg_d65.npz; do not edit\n")
        f.write("#define YPT_CIE_RES %d\n" % CIE_RES)
        f.write("static const float YPT_CIE_XYZ[%d] = {\n" % (CIE_RES * 4))
        for r in xyz:
            f.write("  %s,%s,%s,0.0f,\n" % tuple(_flt(v) for v in r))
        f.write("};\nstatic const float YPT_CIE_D65[%d] = {\n" % CIE_RES)
        for i in range(0, CIE_RES, 8):
            f.write("  " + ",".join(_flt(v) for v in d65[i:i + 8]) + ",\n")
        f.write("};\n")
        for k, m in mats.items():
            f.write("static const float YPT_%s[9] = {%s};\n" % (k.upper(), ",".join(_flt(v) for v in m.ravel())))
        # double-precision copies for the host RGB->spectrum fit (SpectralUtility.cpp works in double)
        for key in ("x", "y", "z", "d65"):
            f.write("static const double YPT_CIE_%s_D[%d] = {\n" % (key.upper(), CIE_RES))
            vals = d[key][sl]
            for i in range(0, CIE_RES, 6):
                f.write("  " + ",".join(repr(float(v)) for v in vals[i:i + 6]) + ",\n")
            f.write("};\n")
        for k in ("srgb_xyz_to_rgb", "srgb_rgb_to_xyz"):
            f.write("static const double YPT_%s_D[9] = {%s};\n" % (k.upper(), ",".join(repr(float(v)) for v in d[k].ravel())))


def main():
    os.makedirs(OUT, exist_ok=True)
    mats = sobol_matrices(os.path.join(HERE, "joe_kuo_first513.txt"))
    write_sobol_inc(mats, os.path.join(OUT, "sobol_matrices.inc"))
    write_cie_inc(os.path.join(HERE, "cie_1931_2deg_d65.npz"), os.path.join(OUT, "cie_tables.inc"))
    print("tables written to", os.path.normpath(OUT))


if __name__ == "__main__":
    main()
