"""Constant tables against the data the reference holds (SobolMatrices.h, SpectralData.h, spectralConstants.h)."""
import json
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _inc_uints(path, name):
    text = open(path).read()
    body = re.search(name + r"\[\d+\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    return np.array([int(t.rstrip("u"), 16) for t in re.split(r"[\s,]+", body) if t], dtype=np.uint64).astype(np.uint32)


def _inc_floats(path, name):
    text = open(path).read()
    body = re.search(name + r"\[\d+\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    return np.array([float(t.rstrip("f")) for t in re.split(r"[\s,]+", body) if t], dtype=np.float32)


def test_sobol_generator_reproduces_reference_table():
    from yapt_amd.data.gen_tables import sobol_matrices
    ours = sobol_matrices(os.path.join(ROOT, "yapt_amd", "data", "joe_kuo_first513.txt"))
    ref = np.load(os.path.join(GOLDEN, "sobol_matrices_ref.npy"))
    assert ours.shape == (513, 32)
    assert np.array_equal(ours, ref)


def test_compiled_sobol_table_matches_generator():
    ref = np.load(os.path.join(GOLDEN, "sobol_matrices_ref.npy"))
    inc = _inc_uints(os.path.join(ROOT, "yapt_amd", "csrc", "data", "sobol_matrices.inc"), "YPT_SOBOL_MATRICES")
    assert np.array_equal(inc.reshape(513, 32), ref)


def test_sobol_dimension0_is_van_der_corput():
    ref = np.load(os.path.join(GOLDEN, "sobol_matrices_ref.npy"))
    assert [int(v) for v in ref[0]] == [1 << (31 - j) for j in range(32)]


def test_sobol_points_match_matrix_xor(oracle):
    ref = np.load(os.path.join(GOLDEN, "sobol_matrices_ref.npy")).astype(np.uint64)
    for idx in (0, 1, 2, 3, 7, 64, 1000, 123457):
        exp = np.zeros(256, np.uint64)
        for d in range(256):
            v = 0
            for bit in range(32):
                if (idx >> bit) & 1:
                    v ^= int(ref[d][bit])
            exp[d] = v
        assert np.array_equal(oracle.sobol_point(idx).astype(np.uint64), exp), idx


def test_sobol_first_points_are_stratified(oracle):
    # a (0, m, 2) net in every pair of the first dimensions: 2^m points fill the 2^m elementary intervals
    pts = np.array([oracle.sobol_point(i)[:8] for i in range(64)], dtype=np.float64) / 2.0 ** 32
    for d in range(8):
        assert sorted(np.floor(pts[:, d] * 64).astype(int)) == list(range(64))


def test_cie_sums_match_reference_constants():
    consts = json.load(open(os.path.join(GOLDEN, "spectral_constants.json")))
    d = np.load(os.path.join(ROOT, "yapt_amd", "data", "cie_1931_2deg_d65.npz"))
    off, n = int(consts["CIE_LUT_ARRAY_OFFSET"]), int(consts["CIE_LUT_RESOLUTION"])
    assert d["lambda"][off] == consts["CIE_LUT_LAMBDA_MIN"] and d["lambda"][off + n - 1] == consts["CIE_LUT_LAMBDA_MAX"]
    assert abs(d["y"][off:off + n].sum() - consts["CIE_Y_SUM"]) < 1e-6 * consts["CIE_Y_SUM"]
    assert abs(d["d65"][off:off + n].sum() - consts["CIE_D65_SUM"]) < 1e-6 * consts["CIE_D65_SUM"]
    assert abs(1.0 / consts["CIE_Y_SUM"] - consts["CIE_Y_SUM_INV"]) < 1e-8
    assert abs(1.0 / consts["CIE_D65_SUM"] - consts["CIE_D65_SUM_INV"]) < 1e-9


def test_compiled_cie_tables_are_float32_of_the_data():
    d = np.load(os.path.join(ROOT, "yapt_amd", "data", "cie_1931_2deg_d65.npz"))
    inc = os.path.join(ROOT, "yapt_amd", "csrc", "data", "cie_tables.inc")
    xyz = _inc_floats(inc, "YPT_CIE_XYZ").reshape(321, 4)
    assert np.array_equal(xyz[:, 0], d["x"][20:341].astype(np.float32))
    assert np.array_equal(xyz[:, 1], d["y"][20:341].astype(np.float32))
    assert np.array_equal(xyz[:, 2], d["z"][20:341].astype(np.float32))
    assert np.array_equal(_inc_floats(inc, "YPT_CIE_D65"), d["d65"][20:341].astype(np.float32))


def test_srgb_matrices_are_inverse():
    d = np.load(os.path.join(ROOT, "yapt_amd", "data", "cie_1931_2deg_d65.npz"))
    assert np.allclose(d["srgb_xyz_to_rgb"] @ d["srgb_rgb_to_xyz"], np.eye(3), atol=2e-3)
