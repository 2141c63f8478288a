"""Command-line autotuner for the benchmark workloads.

    python -m yapt_amd.autotune [--out PATH] [--history PATH] [--width 1920 --height 1080 --depth 12 --shapes 2000]

Coordinate-descent search over the launch options (yapt_amd.tuning.SPACE) for the random-shapes
workload; stores the fastest configuration (yapt_amd/tuned.json by default).
"""
import argparse

from . import _lib, camera, scenes, tuning


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=None)
    p.add_argument("--history", default=None, help="JSON file for the timing of every configuration tried")
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--depth", type=int, default=12)
    p.add_argument("--shapes", type=int, default=2000)
    p.add_argument("--integrator", default="nee", choices=["nee", "bpt", "bdpt"])
    p.add_argument("--passes", type=int, default=16)
    p.add_argument("--rounds", type=int, default=2)
    a = p.parse_args()
    bdpt = a.integrator == "bdpt"
    if bdpt and (a.width, a.height, a.depth) == (1920, 1080, 12):
        a.width, a.height, a.depth = 1280, 720, 8   # the bidirectional workload
    scene, cam = scenes.random_shapes(n_shapes=a.shapes, seed=2, emissive_triangles=8 if bdpt else 0)
    frame = camera.make_frame(cam, a.width, a.height, max_depth=a.depth, light_path_vertices=9 if bdpt else 16)
    integ = {"nee": _lib.YPT_INTEGRATOR_NEE, "bpt": _lib.YPT_INTEGRATOR_BPT, "bdpt": _lib.YPT_INTEGRATOR_BDPT}[a.integrator]
    it = _lib.Integrator(integ, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    key = "%s-shapes%d-%dx%d-d%d" % (a.integrator, a.shapes, a.width, a.height, a.depth)
    best = tuning.autotune(it, key, passes=a.passes, rounds=a.rounds, out=a.out, history=a.history,
                           space=tuning.SPACE_BDPT if bdpt else None)
    print("best", key, best)
    it.close()


if __name__ == "__main__":
    main()
