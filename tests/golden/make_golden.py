"""Regenerates the committed golden vectors of tests/golden/ (run here, where hipcc and the
reference tree exist; the tests only read the JSON files).

  thrust_kat.json : rocThrust minstd_rand / uniform distributions (tests/golden/thrust_kat.hip)
  pcg32_kat.json  : pcg32 from the reference's vendored libs/pcg (oracle/_ref/pcg_kat)
  scenes/shipped_*.scene : input data: the reference's scene files that name no mesh file, copied
                    byte for byte, rendered by the GPU tests (extended_primitives names a JPEG
                    texture that is not kept: its material samples an empty image)
"""
import json
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def thrust():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "kat")
        subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "--offload-arch=gfx950", os.path.join(HERE, "thrust_kat.hip"),
                        "-o", exe], check=True)
        out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout
    data = json.loads(out)
    with open(os.path.join(HERE, "thrust_kat.json"), "w") as f:
        json.dump(data, f)


def pcg():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "pcg_kat")
    cases = []
    for st, sq in [(42, 54), (0, 0), (1, 1), (0xdeadbeefcafebabe, 7), (2580, 1 << 40)]:
        out = subprocess.run([exe, str(st), str(sq), "32"], check=True, stdout=subprocess.PIPE, text=True).stdout
        cases.append({"initstate": st, "initseq": sq, "out": [int(x) for x in out.split()]})
    with open(os.path.join(HERE, "pcg32_kat.json"), "w") as f:
        json.dump({"source": "libs/pcg/pcg_random.hpp pcg32", "cases": cases}, f)


SHIPPED = ("all_lights_types.scene", "simple_spot_light.scene", "extended_primitives.scene")


def shipped_scenes():
    for name in SHIPPED:
        shutil.copyfile(os.path.join("/root/reference/data/scenes", name), os.path.join(HERE, "scenes", "shipped_" + name))


if __name__ == "__main__":
    thrust()
    pcg()
    shipped_scenes()
