"""Extracts the numeric fixtures the reference holds for the sampler and the spectral tables.

Run in a container that has the reference checked out (default path below). It writes data only.
The two inputs of the table generator go to ``yapt_amd/data/``; the expected outputs stay here:

* ``joe_kuo_first513.txt``  - the first 513 rows of the Joe & Kuo "new-joe-kuo-6.21201" direction-number
  file (Tools/new-joe-kuo-6.21201), the input the reference's Sobol table was generated from
  (Tools/generate_sobol_matrices.py:37-90).
* ``sobol_matrices_ref.npy`` - the reference's 513x32 uint32 Sobol generator matrices as printed in
  Source/Common/inc/Math/SobolMatrices.h:5-7 (expected output of the table generator).
* ``cie_1931_2deg_d65.npz`` - CIE 1931 2-degree colour matching functions, the D65 illuminant and the
  wavelength grid, 360..830 nm at 1 nm (Source/Renderer/inc/Renderer/Shared/Utility/SpectralData.h:10-268),
  plus the RGB<->XYZ matrices (SpectralData.h:271-297). These are CIE standard data.
* ``spectral_constants.json`` - the sums and range constants of
  Source/Renderer/shaders/sharedIncludes/spectralConstants.h:19-29.
* ``material_constants.json`` - the material presets of MaterialProxy::setFromMaterialPreset
  (Source/Renderer/src/Shared/MaterialProxy.cpp:15-52) evaluated as the constructor arguments of
  MaterialParameters (glm::pow in float32), the defaults of MaterialParameters()
  (Source/Renderer/inc/Renderer/Shared/MaterialInternal.h:41-75), the GLambda sheen fit coefficients
  (Source/Renderer/shaders/common/miscBrdf.hlsl:155-166) and the average-Fresnel fit
  (Source/Renderer/shaders/common/multiScatter.hlsl:95-108), each as the numbers the text holds.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "..", "..", "yapt_amd", "data")


def _array_body(text, name):
    m = re.search(name + r"\s*\[[^\]]*\]\s*=\s*\{(.*?)\};", text, re.S)
    if m is None:
        raise RuntimeError("array %s not found" % name)
    return m.group(1)


def _numbers(body):
    return [t for t in re.split(r"[\s,]+", body) if t]


def main(ref_root="/root/reference"):
    jk = os.path.join(ref_root, "Tools", "new-joe-kuo-6.21201")
    with open(jk) as f:
        lines = f.readlines()
    with open(os.path.join(DATA, "joe_kuo_first513.txt"), "w") as f:
        f.writelines(lines[:514])  # header + 513 dimension rows

    sob = os.path.join(ref_root, "Source", "Common", "inc", "Math", "SobolMatrices.h")
    text = open(sob).read()
    body = _array_body(text, "SOBOL_MATRICES")
    vals = [int(t.rstrip("u"), 16) for t in _numbers(body)]
    arr = np.array(vals, dtype=np.uint32).reshape(-1, 32)
    np.save(os.path.join(HERE, "sobol_matrices_ref.npy"), arr)

    sd = os.path.join(ref_root, "Source", "Renderer", "inc", "Renderer", "Shared", "Utility", "SpectralData.h")
    text = open(sd).read()
    out = {}
    for key, name in [("x", "c_cieDeg2X"), ("y", "c_cieDeg2Y"), ("z", "c_cieDeg2Z"), ("d65", "c_cieD65StandardIllum")]:
        out[key] = np.array([float(t) for t in _numbers(_array_body(text, name))], dtype=np.float64)
    out["lambda"] = np.array([int(t) for t in _numbers(_array_body(text, "c_cieLambda"))], dtype=np.int32)
    for key, name in [("rec2020_xyz_to_rgb", "c_rec2020XYZToRGB"), ("rec2020_rgb_to_xyz", "c_rec2020RGBToXYZ"),
                      ("srgb_xyz_to_rgb", "c_srgbXYZToRGB"), ("srgb_rgb_to_xyz", "c_srgbRGBToXYZ")]:
        out[key] = np.array([float(t) for t in _numbers(_array_body(text, name))], dtype=np.float64).reshape(3, 3)
    np.savez(os.path.join(DATA, "cie_1931_2deg_d65.npz"), **out)

    sc = os.path.join(ref_root, "Source", "Renderer", "shaders", "sharedIncludes", "spectralConstants.h")
    consts = {}
    for line in open(sc):
        m = re.match(r"#define\s+(\w+)\s+([-0-9.eE]+)\s*$", line.strip())
        if m:
            consts[m.group(1)] = float(m.group(2))
    with open(os.path.join(HERE, "spectral_constants.json"), "w") as f:
        json.dump(consts, f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "material_constants.json"), "w") as f:
        json.dump(material_constants(ref_root), f, indent=1, sort_keys=True)
    print("wrote fixtures to", HERE)


def _float(tok):
    return float(tok.strip().rstrip("fF"))


def _split_top(args):
    """splits a C++ argument list at the commas outside parentheses"""
    out, depth, cur = [], 0, ""
    for ch in args:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _value(expr):
    """a constructor argument: a float literal, vec3p(a, b, c), vec3p(a) or glm::pow(vec3p(...), vec3p(p))"""
    expr = expr.strip()
    m = re.fullmatch(r"glm::pow\(\s*vec3p\(([^)]*)\)\s*,\s*vec3p\(([^)]*)\)\s*\)", expr)
    if m:
        base = np.array([_float(t) for t in m.group(1).split(",")], np.float32)
        p = np.float32(_float(m.group(2)))
        return [float(v) for v in np.power(base, p)]
    m = re.fullmatch(r"vec3p\(([^)]*)\)", expr)
    if m:
        v = [_float(t) for t in m.group(1).split(",")]
        return v * 3 if len(v) == 1 else v
    return _float(expr)


def material_constants(ref_root):
    shared = os.path.join(ref_root, "Source", "Renderer")
    text = open(os.path.join(shared, "src", "Shared", "MaterialProxy.cpp")).read()
    fields = ["albedo", "transparency", "specular", "metalness", "absorption", "dielectricIOR", "emissive", "roughness",
              "anisotropy", "anisotropyRotation"]
    presets = {}
    for m in re.finditer(r"case\s+YAPT::MaterialPreset::(\w+):\s*m_materialParams\s*=\s*MaterialParameters\((.*?)\);\s*break;",
                         text, re.S):
        args = _split_top(m.group(2))
        presets[m.group(1)] = dict(zip(fields, [_value(a) for a in args])) if args else {}
    text = open(os.path.join(shared, "inc", "Renderer", "Shared", "MaterialInternal.h")).read()
    ctor = re.search(r"MaterialParameters\(\)\s*:(.*?)\{\}", text, re.S).group(1)
    defaults = {}
    for m in re.finditer(r"(\w+)\(([^()]*)\)", ctor):
        if m.group(2).strip() == "TEX_UNBOUND_INDEX":
            continue
        vals = [_float(t) for t in m.group(2).split(",")]
        defaults[m.group(1)] = vals if len(vals) > 1 else vals[0]
    text = open(os.path.join(shared, "shaders", "common", "miscBrdf.hlsl")).read()
    body = re.search(r"float GLambdaSheen\(.*?\{(.*?)float t =", text, re.S).group(1)
    sheen = {m.group(1): _float(m.group(2)) for m in re.finditer(r"float\s+([a-e][01])\s*=\s*([-0-9.eE]+f?)\s*;", body)}
    text = open(os.path.join(shared, "shaders", "common", "multiScatter.hlsl")).read()
    body = re.search(r"float getAvgFresnel\(float etaR\)\s*\{(.*?)return fAvg;", text, re.S).group(1)
    lo = re.search(r"fAvg\s*=\s*([-0-9.]+)f\s*\+\s*([-0-9.]+)f\s*\*\s*etaR\s*-\s*([-0-9.]+)f\s*\*\s*etaR\s*\*\s*etaR\s*-\s*"
                   r"([-0-9.]+)f\s*\*\s*etaR\s*\*\s*etaR\s*\*\s*etaR", body)
    hi = re.search(r"fAvg\s*=\s*\(etaR\s*-\s*1\.f\)\s*/\s*\(([-0-9.]+)f\s*\+\s*([-0-9.]+)f\s*\*\s*etaR\)", body)
    avg_fresnel = {"below_1": [float(lo.group(1)), float(lo.group(2)), -float(lo.group(3)), -float(lo.group(4))],
                   "above_1_denominator": [float(hi.group(1)), float(hi.group(2))]}
    return {"presets": presets, "defaults": defaults, "glambda_sheen": sheen, "avg_fresnel": avg_fresnel}


if __name__ == "__main__":
    main(*sys.argv[1:])
