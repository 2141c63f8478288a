"""The library's host-side code (BVH builder, scene preparation, LUT and spectrum-fit code) and the CPU
checker's multithreaded renders under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer: tests/native/host_checks.cpp built with g++ (host code only, no GPU) and run on the
Cornell box and a small random-shapes scene with punctual lights."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "host_checks.cpp")
FLAGS = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread"]
SANITIZERS = {
    "address_undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "thread": ["-fsanitize=thread"],
}


def _blob(scene, frame, path):
    """the scene description arrays and frame constants in the layout host_checks.cpp reads"""
    d = scene.desc
    nv, nt, ni, nm, nl = d.vertex_count, d.triangle_count, d.instance_count, d.material_count, d.punctual_light_count
    assert d.texture_count == 0 and not d.environment
    has_n, has_t, has_uv = bool(d.normals), bool(d.tangents), bool(d.uvs)
    hdr = struct.pack("<12I", 0x59505448, ctypes.sizeof(_lib.Material), ctypes.sizeof(_lib.PunctualLight),
                      ctypes.sizeof(_lib.FrameDesc), nv, nt, ni, nm, nl, int(has_n), int(has_t), int(has_uv))
    parts = [hdr, bytes(frame), ctypes.string_at(ctypes.addressof(d.materials.contents), nm * ctypes.sizeof(_lib.Material))]
    if nl:
        parts.append(ctypes.string_at(ctypes.addressof(d.punctual_lights.contents), nl * ctypes.sizeof(_lib.PunctualLight)))
    parts.append(scene.positions.tobytes())
    if has_n:
        parts.append(scene.normals.tobytes())
    if has_t:
        parts.append(scene.tangents.tobytes())
    if has_uv:
        parts.append(scene.uvs.tobytes())
    parts += [scene.indices.tobytes(), scene.inst_first.tobytes(), scene.inst_count.tobytes(), scene.inst_material.tobytes()]
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def _supported(tmp_path, flags):
    src = tmp_path / "probe.cpp"
    src.write_text("int main() { return 0; }\n")
    exe = tmp_path / "probe"
    r = subprocess.run(FLAGS + flags + ["-o", str(exe), str(src)], capture_output=True, text=True)
    return r.returncode == 0 and subprocess.run([str(exe)], capture_output=True).returncode == 0


@pytest.mark.parametrize("kind", list(SANITIZERS))
def test_host_code_and_checker_under_sanitizers(tmp_path, kind):
    flags = SANITIZERS[kind]
    if not _supported(tmp_path, flags):
        pytest.skip("g++ cannot build and run %s-sanitized programs here" % kind)
    exe = tmp_path / ("host_checks_" + kind)
    r = subprocess.run(FLAGS + flags + ["-o", str(exe), SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    cornell, ccam = scenes.cornell_box()
    shapes, scam = scenes.random_shapes(n_shapes=40, seed=7, cube=8.0, n_lights=2, light_emission=20.0)
    punct, pcam, env_scale = scenes.punctual_lights("mixed", sphere_res=(10, 10))
    for k, (scene, cam, envs) in enumerate([(cornell, ccam, 0.0), (shapes, scam, 0.0), (punct, pcam, env_scale)]):
        blob = tmp_path / ("scene%d.bin" % k)
        _blob(scene, camera.make_frame(cam, 24, 16, max_depth=5, env_intensity_scale=envs, light_path_vertices=5), blob)
        r = subprocess.run([str(exe), str(blob)], capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0 and "host checks ok" in r.stdout, (k, r.stdout[-2000:], r.stderr[-6000:])
