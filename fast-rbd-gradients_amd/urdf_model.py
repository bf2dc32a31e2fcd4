"""Robot model header from a URDF: the compile-time tables the kernels are
specialised on (csrc/iiwa_model.hpp is the one for the KUKA iiwa 14).

    python fast-rbd-gradients_amd/urdf_model.py ROBOT.urdf NAME
        -> fast-rbd-gradients_amd/csrc/model_NAME.hpp
    make -C fast-rbd-gradients_amd MODEL=NAME URDF=ROBOT.urdf
        -> the header, then librbdgrad_NAME.so built on it

The kernels keep the reference's structure: a fixed-base serial chain of
NUM_POS = 7 revolute joints with S = e_2 in each joint frame.  So the URDF must
describe such a chain: exactly seven revolute or continuous joints, each one's
child link the next one's parent.  Fixed joints and links outside the chain are
ignored, as the reference's tables ignore the iiwa's base and end-effector
links.  The base frame is the first joint's parent link, with gravity along
its -z.

A joint axis other than +z is handled by re-framing: link k's frame is rotated
by the R_a(k) that takes z to the axis, so the joint turns about the new frame's
z.  The joint-space results (qdd and its gradients) do not depend on the link
frames; the external forces of the fext entry points are then given in the
re-framed link frames.  For an axis of +z, R_a = 1 and the frames are the URDF's.
A link inertial given in a rotated frame (<inertial><origin rpy>) is rotated
into the link frame.  <dynamics damping> is not read: as in the reference,
VEL_DAMPING adds 0.5 qd on every joint.

Rotation entries within 1e-12 of 0 or +-1 are snapped to those values, so that
structural zeros of the transforms stay exact zeros the compiler folds away.
"""
import argparse
import json
import math
import os
import sys
import xml.etree.ElementTree as ET

NQ = 7
HERE = os.path.dirname(os.path.abspath(__file__))


class ModelError(ValueError):
    pass


def _floats(s, n=3):
    v = [float(x) for x in (s or "").split()]
    if not v:
        v = [0.0] * n
    if len(v) != n:
        raise ModelError(f"expected {n} numbers, got {s!r}")
    return v


def rpy_matrix(roll, pitch, yaw):
    """URDF fixed-axis rpy: R = Rz(yaw) Ry(pitch) Rx(roll)."""
    cr, sr = math.cos(roll), math.sin(roll)
    cp, sp = math.cos(pitch), math.sin(pitch)
    cy, sy = math.cos(yaw), math.sin(yaw)
    return [[cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
            [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
            [-sp, cp * sr, cp * cr]]


def snap(x, eps=1e-12):
    for v in (0.0, 1.0, -1.0):
        if abs(x - v) < eps:
            return v
    return x


def _mm(A, B):
    return [[sum(A[i][l] * B[l][j] for l in range(3)) for j in range(3)] for i in range(3)]


def _mt(A):
    return [[A[j][i] for j in range(3)] for i in range(3)]


def _mv(A, x):
    return [sum(A[i][l] * x[l] for l in range(3)) for i in range(3)]


_EYE = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
MOVING = ("revolute", "continuous")


def read_urdf(path):
    """The chain of a URDF: {'joints': [7 x {'name', 'rot', 'xyz', 'axis'}],
    'links': [7 x {'mass', 'com', 'com_rpy', 'inertia'}]}.  Joint k's frame is
    rot Rot(axis, q) at xyz in the frame of chain link k-1 (the base for k = 0).
    A .json file in the format of tests/golden/iiwa_14_urdf.json is read too."""
    if path.endswith(".json"):
        with open(path) as f:
            d = json.load(f)
        joints = [dict(j, type="revolute") for j in d["joints"]]
        return build(joints, {k: dict(v) for k, v in d["links"].items()})
    root = ET.parse(path).getroot()
    links = {}
    for link in root.findall("link"):
        inert = link.find("inertial")
        if inert is None:
            links[link.get("name")] = None
            continue
        o = inert.find("origin")
        i = inert.find("inertia")
        if inert.find("mass") is None or i is None:
            raise ModelError(f"link {link.get('name')}: inertial without mass or inertia")
        links[link.get("name")] = {
            "mass": float(inert.find("mass").get("value")),
            "com": _floats(o.get("xyz") if o is not None else None),
            "com_rpy": _floats(o.get("rpy") if o is not None else None),
            "inertia": {k: float(i.get(k, "0")) for k in ("ixx", "ixy", "ixz", "iyy", "iyz", "izz")},
        }
    joints = []
    for j in root.findall("joint"):
        o = j.find("origin")
        ax = j.find("axis")
        joints.append({
            "name": j.get("name"),
            "type": j.get("type"),
            "parent": j.find("parent").get("link"),
            "child": j.find("child").get("link"),
            "xyz": _floats(o.get("xyz") if o is not None else None),
            "rpy": _floats(o.get("rpy") if o is not None else None),
            "axis": _floats(ax.get("xyz") if ax is not None else "1 0 0"),
        })
    return build(joints, links)


def _lump(parts):
    """One rigid body from parts [(R, p, inertial)] posed in a common frame."""
    if len(parts) == 1 and parts[0][0] == _EYE and parts[0][1] == [0.0, 0.0, 0.0]:
        return parts[0][2]
    m = sum(float(L["mass"]) for _, _, L in parts)
    if m <= 0.0:
        raise ModelError("a link of the chain has no mass")
    cs, Is = [], []
    for R, p, L in parts:
        c = [p[i] + v for i, v in enumerate(_mv(R, [float(x) for x in L["com"]]))]
        I = L["inertia"]
        Ic = [[I["ixx"], I["ixy"], I["ixz"]], [I["ixy"], I["iyy"], I["iyz"]], [I["ixz"], I["iyz"], I["izz"]]]
        Rc = _mm(R, rpy_matrix(*L.get("com_rpy", [0.0, 0.0, 0.0])))
        cs.append(c)
        Is.append(_mm(_mm(Rc, Ic), _mt(Rc)))
    com = [sum(float(L["mass"]) * c[i] for (_, _, L), c in zip(parts, cs)) / m for i in range(3)]
    tot = [[0.0] * 3 for _ in range(3)]
    for (_, _, L), c, Ii in zip(parts, cs, Is):
        d = [c[i] - com[i] for i in range(3)]
        dd = sum(x * x for x in d)
        for a in range(3):
            for b in range(3):
                tot[a][b] += Ii[a][b] + float(L["mass"]) * ((dd if a == b else 0.0) - d[a] * d[b])
    return {"mass": m, "com": com, "com_rpy": [0.0, 0.0, 0.0],
            "inertia": {"ixx": tot[0][0], "ixy": tot[0][1], "ixz": tot[0][2],
                        "iyy": tot[1][1], "iyz": tot[1][2], "izz": tot[2][2]}}


def build(joints, links):
    """Order the revolute joints into the chain and fold fixed joints away:
    a link hanging off chain link k through fixed joints is lumped into link k
    (its inertia moved into link k's frame), a revolute joint whose parent is
    such a link gets the fixed offsets composed into its origin, and links fixed
    to the tree's root belong to the base and are dropped.  The base frame is
    the root link's frame (gravity along its -z)."""
    for j in joints:
        if j["type"] not in MOVING + ("fixed",):
            raise ModelError(f"joint {j['name']}: type {j['type']} -- the kernels take revolute joints only")
    by_child = {}
    for j in joints:
        if j["child"] in by_child:
            raise ModelError(f"link {j['child']} has two parent joints")
        by_child[j["child"]] = j

    def body(link):
        """(link of the body `link` is rigidly part of, R, p of `link` in that body's frame)."""
        R, p = _EYE, [0.0, 0.0, 0.0]
        while link in by_child and by_child[link]["type"] == "fixed":
            J = by_child[link]
            Rj = rpy_matrix(*J["rpy"])
            R, p = _mm(Rj, R), [J["xyz"][i] + v for i, v in enumerate(_mv(Rj, p))]
            link = J["parent"]
        return link, R, p

    moving = [j for j in joints if j["type"] in MOVING]
    if len(moving) != NQ:
        raise ModelError(f"the kernels take a chain of {NQ} revolute joints; the URDF has {len(moving)}")
    parent_body = {j["name"]: body(j["parent"])[0] for j in moving}
    order, prev_child = [], None
    for _ in range(NQ):
        if prev_child is None:
            nxt = [j for j in moving if parent_body[j["name"]] not in by_child]
        else:
            nxt = [j for j in moving if parent_body[j["name"]] == prev_child]
        if len(nxt) != 1:
            raise ModelError("the revolute joints do not form one serial chain")
        order.append(nxt[0])
        prev_child = nxt[0]["child"]
    out_joints, out_links = [], []
    chain_links = [j["child"] for j in order]
    for j in order:
        if sum(float(a) ** 2 for a in j["axis"]) < 1e-24:
            raise ModelError(f"joint {j['name']}: zero axis")
        _, Rp, pp = body(j["parent"])
        rot = _mm(Rp, rpy_matrix(*j["rpy"]))
        xyz = [pp[i] + v for i, v in enumerate(_mv(Rp, [float(x) for x in j["xyz"]]))]
        out_joints.append({"name": j["name"], "rot": rot, "xyz": xyz, "axis": [float(a) for a in j["axis"]]})
    for c in chain_links:
        parts = []
        for name, L in links.items():
            if L is None:
                continue
            b, R, p = body(name)
            if b == c:
                parts.append((R, p, L))
        if not parts:
            raise ModelError(f"link {c} has no inertial")
        # the chain link's own inertial first, so a link without fixed children stays as written
        parts.sort(key=lambda t: t[0] != _EYE or t[1] != [0.0, 0.0, 0.0])
        out_links.append(_lump(parts))
    return {"joints": out_joints, "links": out_links}


def axis_frame(axis):
    """A rotation R_a with R_a z = axis/|axis|: its columns are x', y' = a x x'
    and a, with x' the unit part of e_x (e_y for axes near x) orthogonal to a.
    R_a = 1 exactly for +z; exact to rounding for every other axis."""
    n = math.sqrt(sum(float(v) ** 2 for v in axis))
    a = [float(v) / n for v in axis]
    e = [1.0, 0.0, 0.0] if abs(a[0]) < 0.9 else [0.0, 1.0, 0.0]
    d = sum(e[i] * a[i] for i in range(3))
    x = [e[i] - d * a[i] for i in range(3)]
    xn = math.sqrt(sum(v * v for v in x))
    x = [v / xn for v in x]
    y = [a[1] * x[2] - a[2] * x[1], a[2] * x[0] - a[0] * x[2], a[0] * x[1] - a[1] * x[0]]
    return [[x[i], y[i], a[i]] for i in range(3)]


def tables(model):
    """The model as the kernels take it, after re-framing for the joint axes:
    links (7 x [m, cx, cy, cz, ixx, ixy, ixz, iyy, iyz, izz], inertia about the
    COM), rot (7 x 3x3: the joint frame is rot Rz(q) at offset in the parent
    frame; entries snapped) and offset (7 x 3)."""
    links, rots, offsets = [], [], []
    Ra_prev = _EYE
    for j, L in zip(model["joints"], model["links"]):
        Ra = axis_frame(j["axis"])
        R = _mm(_mm(_mt(Ra_prev), j["rot"]), Ra)
        rots.append([[snap(x) for x in row] for row in R])
        offsets.append(_mv(_mt(Ra_prev), j["xyz"]))
        I = L["inertia"]
        Ic = [[I["ixx"], I["ixy"], I["ixz"]], [I["ixy"], I["iyy"], I["iyz"]], [I["ixz"], I["iyz"], I["izz"]]]
        Rc = _mm(_mt(Ra), rpy_matrix(*L.get("com_rpy", [0.0, 0.0, 0.0])))
        Il = _mm(_mm(Rc, Ic), _mt(Rc))            # about the COM, in the (re-framed) link frame
        c = _mv(_mt(Ra), [float(x) for x in L["com"]])
        links.append([float(L["mass"]), *c, Il[0][0], Il[0][1], Il[0][2], Il[1][1], Il[1][2], Il[2][2]])
        Ra_prev = Ra
    return {"links": links, "rot": rots, "offset": offsets}


def header(tab, name, source):
    def num(x):
        return repr(float(x))
    lines = [
        f"// Robot model '{name}' for the device code, generated by urdf_model.py from",
        f"// {source}.  Same layout as iiwa_model.hpp: link k is the child of the",
        "// k-th revolute joint of the chain, the joint frame is R Rz(q) at offset r",
        "// from the parent, S = e_2, and kFixedRot[k] = R^T (entries within 1e-12 of",
        "// 0 or +-1 snapped).",
        "#pragma once",
        "",
        "namespace rbd {",
        "",
        f'constexpr const char* kModelName = "{name}";',
        f"constexpr int NQ = {NQ};",
        "constexpr double kGravity = 9.81;",
        "",
        "struct LinkInertial { double m, cx, cy, cz, ixx, ixy, ixz, iyy, iyz, izz; };",
        "",
        "constexpr LinkInertial kLinks[NQ] = {",
    ]
    for row in tab["links"]:
        lines.append("  {" + ", ".join(num(v) for v in row) + "},")
    lines += ["};", "", "constexpr double kFixedRot[NQ][3][3] = {"]
    for R in tab["rot"]:
        F = _mt(R)
        lines.append("  {" + ", ".join("{" + ", ".join(num(v) for v in row) + "}" for row in F) + "},")
    lines += ["};", "", "constexpr double kJointOffset[NQ][3] = {"]
    for r in tab["offset"]:
        lines.append("  {" + ", ".join(num(v) for v in r) + "},")
    lines += [
        "};",
        "",
        "// E = Rz(q)^T R^T: E(i,j) = ec(i,j) cos q + es(i,j) sin q + e1(i,j)",
        "constexpr double ecoef_c(int k, int i, int j) {",
        "  return i == 0 ? kFixedRot[k][0][j] : (i == 1 ? kFixedRot[k][1][j] : 0.0);",
        "}",
        "constexpr double ecoef_s(int k, int i, int j) {",
        "  return i == 0 ? kFixedRot[k][1][j] : (i == 1 ? -kFixedRot[k][0][j] : 0.0);",
        "}",
        "constexpr double ecoef_1(int k, int i, int j) { return i == 2 ? kFixedRot[k][2][j] : 0.0; }",
        "",
        "// spatial inertia about the joint frame origin (parallel-axis shift)",
        "constexpr double inertia_entry(int k, int r, int c) {",
        "  const LinkInertial L = kLinks[k];",
        "  const double com[3] = {L.cx, L.cy, L.cz};",
        "  if (r < 3 && c < 3) {",
        "    const double ibar[3][3] = {{L.ixx, L.ixy, L.ixz}, {L.ixy, L.iyy, L.iyz}, {L.ixz, L.iyz, L.izz}};",
        "    const double cc = L.cx * L.cx + L.cy * L.cy + L.cz * L.cz;",
        "    return ibar[r][c] + L.m * ((r == c ? cc : 0.0) - com[r] * com[c]);",
        "  }",
        "  if (r >= 3 && c >= 3) return r == c ? L.m : 0.0;",
        "  const int i = r < 3 ? r : c, j = r < 3 ? c - 3 : r - 3;",
        "  const double mc[3] = {L.m * L.cx, L.m * L.cy, L.m * L.cz};",
        "  const double skew[3][3] = {{0.0, -mc[2], mc[1]}, {mc[2], 0.0, -mc[0]}, {-mc[1], mc[0], 0.0}};",
        "  return skew[i][j];",
        "}",
        "",
        "}  // namespace rbd",
        "",
    ]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("urdf")
    ap.add_argument("name")
    ap.add_argument("--out", help="output header (default csrc/model_NAME.hpp next to this script)")
    args = ap.parse_args(argv)
    if not args.name.replace("_", "").isalnum():
        raise SystemExit("model name: letters, digits and _ only")
    try:
        model = read_urdf(args.urdf)
    except (ModelError, OSError, ET.ParseError) as e:
        raise SystemExit(f"{args.urdf}: {e}")
    out = args.out or os.path.join(HERE, "csrc", f"model_{args.name}.hpp")
    text = header(tables(model), args.name, os.path.basename(args.urdf))
    if not (os.path.exists(out) and open(out).read() == text):   # keep the timestamp when unchanged
        with open(out + ".tmp", "w") as f:
            f.write(text)
        os.replace(out + ".tmp", out)
    print(out)


if __name__ == "__main__":
    sys.exit(main())
