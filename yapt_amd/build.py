"""In-tree build of the gfx950 libraries: the path tracer (libyapt_pt.so), the denoiser kernels
(libyapt_denoise.so) and the pybind11 binding of the path tracer's C ABI (_yapt_native, g++, host
code only).

Flags: -ffp-contract=off and IEEE-correct fp32 division/sqrt so device arithmetic matches the host
build bit for bit; packed-fp32 instructions disabled; -O3 for gfx950 only.

The path tracer is compiled as several objects in parallel: the host unit (ypt_lib.hip, with the
non-template kernels) and one object per group of template kernel instances (csrc/instances.hip
with -DYPT_INST_GROUP=n), linked into one shared library.
"""
import os
import re
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SRC = os.path.join(CSRC, "ypt_lib.hip")
INSTANCES = os.path.join(CSRC, "instances.hip")
OUT = os.path.join(HERE, "libyapt_pt.so")
DENOISE_SRC = os.path.join(CSRC, "denoise.hip")
DENOISE_OUT = os.path.join(HERE, "libyapt_denoise.so")
NATIVE_SRC = os.path.join(CSRC, "py_native.cpp")


def native_module_path():
    """the pybind11 module yapt_amd/_yapt_native<EXT_SUFFIX>"""
    import sysconfig
    return os.path.join(HERE, "_yapt_native" + sysconfig.get_config_var("EXT_SUFFIX"))


def _native_cmd(out):
    """g++ line of the pybind11 binding (host code only), linked against the in-tree libyapt_pt.so"""
    import sysconfig
    import pybind11
    return ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
            "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"], NATIVE_SRC,
            "-L" + HERE, "-lyapt_pt", "-Wl,-rpath,$ORIGIN", "-o", out]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    out = []
    for d, _, files in os.walk(CSRC):
        for f in files:
            if f.endswith((".h", ".hip", ".inc")):
                out.append(os.path.join(d, f))
    out.append(os.path.join(ROOT, "include", "yapt_pt.h"))
    return out


def instance_groups():
    text = open(os.path.join(CSRC, "kernels", "instances.h")).read()
    return int(re.search(r"YPT_INST_GROUPS\s*=\s*(\d+)", text).group(1))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in res.stderr.splitlines() if "not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (res.returncode, errs))
    if verbose and errs.strip():
        print(errs)


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _cost(cmd):
    """rough compile time of a unit in seconds: orders the parallel compilations"""
    g = [a for a in cmd if a.startswith("-DYPT_INST_GROUP=")]
    if not g:
        return 20
    return {4: 50, 5: 80, 6: 80, 7: 80, 8: 80, 9: 50}.get(int(g[0].split("=")[1]), 20)


def build(verbose=False):
    ngroups = instance_groups()
    with tempfile.TemporaryDirectory(prefix="yapt_build_") as tmp:
        units = [(SRC, os.path.join(tmp, "ypt_lib.o"), [])]
        units += [(INSTANCES, os.path.join(tmp, "inst%d.o" % g), ["-DYPT_INST_GROUP=%d" % g]) for g in range(ngroups)]
        cmds = [[HIPCC] + HIP_FLAGS + extra + ["-c", "-o", obj, src] for src, obj, extra in units]
        cmds.append([HIPCC] + HIP_FLAGS + ["-shared", "-o", DENOISE_OUT + ".tmp", DENOISE_SRC])
        # the largest units first, so the pool's tail is short
        cmds.sort(key=lambda c: -_cost(c))
        with ThreadPoolExecutor(max_workers=_jobs()) as pool:
            for f in [pool.submit(_run, c, verbose) for c in cmds]:
                f.result()
        _run([HIPCC, "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", OUT + ".tmp"] + [o for _, o, _ in units], verbose)
    os.replace(DENOISE_OUT + ".tmp", DENOISE_OUT)
    os.replace(OUT + ".tmp", OUT)
    native = native_module_path()
    _run(_native_cmd(native + ".tmp"), verbose)
    os.replace(native + ".tmp", native)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
