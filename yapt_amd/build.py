"""In-tree build of the gfx950 libraries: the path tracer (libyapt_pt.so) and the denoiser kernels
(libyapt_denoise.so).

Flags: -ffp-contract=off and IEEE-correct fp32 division/sqrt so device arithmetic matches the host
build bit for bit; packed-fp32 instructions disabled; -O3 for gfx950 only.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "ypt_lib.hip")
OUT = os.path.join(HERE, "libyapt_pt.so")
DENOISE_SRC = os.path.join(HERE, "csrc", "denoise.hip")
DENOISE_OUT = os.path.join(HERE, "libyapt_denoise.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    out = []
    for d, _, files in os.walk(os.path.join(HERE, "csrc")):
        for f in files:
            if f.endswith((".h", ".hip", ".inc")):
                out.append(os.path.join(d, f))
    out.append(os.path.join(ROOT, "include", "yapt_pt.h"))
    return out


def _compile(src, out, verbose):
    cmd = [HIPCC] + HIP_FLAGS + ["-o", out + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in res.stderr.splitlines() if "not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (res.returncode, errs))
    if verbose and errs.strip():
        print(errs)
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False):
    _compile(DENOISE_SRC, DENOISE_OUT, verbose)
    return _compile(SRC, OUT, verbose)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
