"""Extract the iiwa 14 model data (joint origins/axes/damping, link inertials)
from the reference's utils/iiwa_14.urdf into tests/golden/iiwa_14_urdf.json.

Run once in the build container (the reference is not on the GPU host):
    python tests/golden/make_urdf_fixture.py [path/to/iiwa_14.urdf]
The JSON is data only: numbers read from the URDF's XML attributes.
"""
import json
import os
import sys
import xml.etree.ElementTree as ET

HERE = os.path.dirname(os.path.abspath(__file__))


def floats(s):
    return [float(v) for v in s.split()]


def main(path):
    root = ET.parse(path).getroot()
    links = {}
    for link in root.findall("link"):
        inert = link.find("inertial")
        if inert is None:
            continue
        o = inert.find("origin")
        i = inert.find("inertia")
        links[link.get("name")] = {
            "mass": float(inert.find("mass").get("value")),
            "com": floats(o.get("xyz")),
            "com_rpy": floats(o.get("rpy")),
            "inertia": {k: float(i.get(k)) for k in ("ixx", "ixy", "ixz", "iyy", "iyz", "izz")},
        }
    joints = []
    for j in root.findall("joint"):
        if j.get("type") != "revolute":
            continue
        o = j.find("origin")
        d = j.find("dynamics")
        joints.append({
            "name": j.get("name"),
            "parent": j.find("parent").get("link"),
            "child": j.find("child").get("link"),
            "xyz": floats(o.get("xyz")),
            "rpy": floats(o.get("rpy")),
            "axis": floats(j.find("axis").get("xyz")),
            "damping": float(d.get("damping")) if d is not None else 0.0,
        })
    out = {"source": "utils/iiwa_14.urdf (plancherb1/fast-rbd-gradients)", "joints": joints,
           "links": {j["child"]: links[j["child"]] for j in joints}}
    with open(os.path.join(HERE, "iiwa_14_urdf.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(joints), "joints")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/utils/iiwa_14.urdf")
