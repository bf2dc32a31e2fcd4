"""Shared helpers for the test suite: seeded knot-point inputs, the error
metric, and independent textbook spatial-algebra models used to pin the
oracle: of the iiwa, built from the URDF data in tests/golden/iiwa_14_urdf.json,
and of any URDF tree (UrdfTree), read straight from the XML.
"""
import contextlib
import ctypes
import fcntl
import json
import os
import subprocess
import sys
import xml.etree.ElementTree as ET

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fast-rbd-gradients_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HOSTCHECK = os.path.join(ROOT, "tests", "hostcheck")

NQ = 7
BLOCKS = {"qdd": slice(0, 7), "dqdd_dq": slice(7, 56), "dqdd_dqd": slice(56, 105)}
REL_TOL = 1e-4   # max relative error per output block vs the FP64 oracle (the target bar)


def knots(n, seed, near_singular_frac=0.0, dtype=np.float32):
    """n knot points [q | qd | u]: q ~ U(-pi, pi), qd ~ U(-2, 2) rad/s, u ~ U(-50, 50) N m
    per joint (numpy PCG64 with the given seed).  A fraction of the knots is put
    at the stretched-arm pose: joints 2, 4, 6 (0-based q[1], q[3], q[5]) within
    +-1e-3 rad of 0.  Returned in `dtype`; the oracle is run on exactly these
    values (upcast), so both sides see the same inputs."""
    rng = np.random.default_rng(seed)
    q = rng.uniform(-np.pi, np.pi, (n, NQ))
    qd = rng.uniform(-2.0, 2.0, (n, NQ))
    u = rng.uniform(-50.0, 50.0, (n, NQ))
    if near_singular_frac > 0:
        m = rng.random(n) < near_singular_frac
        for j in (1, 3, 5):
            q[m, j] = rng.uniform(-1e-3, 1e-3, int(m.sum()))
    return np.concatenate([q, qd, u], axis=1).astype(dtype)


def block_errors(out, ref, floor=1e-6):
    """Per knot and block: max|out - ref| / max(max|ref|, floor) (normwise per block)."""
    out = np.atleast_2d(np.asarray(out, dtype=np.float64))
    ref = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    res = {}
    for name, sl in BLOCKS.items():
        if sl.stop > ref.shape[1]:
            continue
        d = np.max(np.abs(out[:, sl] - ref[:, sl]), axis=1)
        s = np.maximum(np.max(np.abs(ref[:, sl]), axis=1), floor)
        res[name] = d / s
    return res


def dqdd_block_errors(out98, ref98, floor=1e-6):
    """Same metric for the reference-compatible 98-value output."""
    pad = np.zeros((np.atleast_2d(out98).shape[0], 7))
    return {k: v for k, v in block_errors(np.hstack([pad, out98]), np.hstack([pad, ref98]), floor).items()
            if k != "qdd"}


def max_block_errors(out, ref):
    return {k: float(v.max()) for k, v in block_errors(out, ref).items()}


# ---------------------------------------------------------------------------
# Independent textbook model (Featherstone, Rigid Body Dynamics Algorithms,
# ch. 2 and Table 5.1) from the URDF fields.  Generic dense 6x6 spatial algebra,
# no shared code with the oracle or the kernels.
# ---------------------------------------------------------------------------
def _rpy(r, p, y):
    cr, sr, cp, sp, cy, sy = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


class UrdfModel:
    def __init__(self, path=os.path.join(GOLDEN, "iiwa_14_urdf.json")):
        with open(path) as f:
            d = json.load(f)
        self.joints = d["joints"]
        self.links = [d["links"][j["child"]] for j in self.joints]

    def transform(self, k, q):
        """Motion transform parent -> link k frame: X = [E 0; -E r^ E],
        E = (R_rpy Rz(q))^T (joint axis z), r = joint origin."""
        j = self.joints[k]
        assert j["axis"] == [0.0, 0.0, 1.0]
        R = _rpy(*j["rpy"]) @ _rpy(0.0, 0.0, q)
        E = R.T
        r = np.array(j["xyz"])
        X = np.zeros((6, 6))
        X[:3, :3] = E
        X[3:, 3:] = E
        X[3:, :3] = -E @ _skew(r)
        return X

    def inertia(self, k):
        """Spatial inertia about the link frame origin (URDF inertia is about the COM)."""
        L = self.links[k]
        assert L["com_rpy"] == [0.0, 0.0, 0.0]
        m = L["mass"]
        c = np.array(L["com"])
        i = L["inertia"]
        Ic = np.array([[i["ixx"], i["ixy"], i["ixz"]], [i["ixy"], i["iyy"], i["iyz"]], [i["ixz"], i["iyz"], i["izz"]]])
        C = _skew(c)
        I = np.zeros((6, 6))
        I[:3, :3] = Ic + m * C @ C.T
        I[:3, 3:] = m * C
        I[3:, :3] = m * C.T
        I[3:, 3:] = m * np.eye(3)
        return I

    def rnea(self, q, qd, qdd, gravity=9.81):
        S = np.array([0, 0, 1.0, 0, 0, 0])
        a0 = np.array([0, 0, 0, 0, 0, gravity])   # fictitious base acceleration
        v_prev, a_prev = np.zeros(6), a0
        Xs, fs = [], []
        for k in range(NQ):
            X = self.transform(k, q[k])
            v = X @ v_prev + S * qd[k]
            a = X @ a_prev + S * qdd[k] + _crm(v) @ S * qd[k]
            I = self.inertia(k)
            fs.append(I @ a + _crf(v) @ I @ v)
            Xs.append(X)
            v_prev, a_prev = v, a
        tau = np.zeros(NQ)
        for k in range(NQ - 1, -1, -1):
            tau[k] = S @ fs[k]
            if k > 0:
                fs[k - 1] = fs[k - 1] + Xs[k].T @ fs[k]
        return tau

    def link_velocities(self, q, qd):
        """Spatial velocity [omega; v] of each link in its own frame, base at rest."""
        S = np.array([0, 0, 1.0, 0, 0, 0])
        v, out = np.zeros(6), []
        for k in range(NQ):
            v = self.transform(k, q[k]) @ v + S * qd[k]
            out.append(v)
        return np.array(out)

    def com_heights(self, q):
        """World z of each link's COM (base frame z up)."""
        R = np.eye(3)
        p = np.zeros(3)
        zs = []
        for k in range(NQ):
            j = self.joints[k]
            p = p + R @ np.array(j["xyz"])
            R = R @ _rpy(*j["rpy"]) @ _rpy(0.0, 0.0, q[k])
            zs.append((p + R @ np.array(self.links[k]["com"]))[2])
        return np.array(zs)

    def potential(self, q, gravity=9.81):
        return gravity * sum(L["mass"] * z for L, z in zip(self.links, self.com_heights(q)))


def _crm(v):
    w, u = v[:3], v[3:]
    M = np.zeros((6, 6))
    M[:3, :3] = _skew(w)
    M[3:, :3] = _skew(u)
    M[3:, 3:] = _skew(w)
    return M


def _crf(v):
    return -_crm(v).T


# ---------------------------------------------------------------------------
# Textbook RNEA over a whole URDF tree, in the URDF's own frames: every link is
# a body, a fixed joint moves nothing (S = 0), a revolute joint turns about its
# (normalised) axis.  Shares no code with fast-rbd-gradients_amd/urdf_model.py
# (which folds fixed joints into the chain and re-frames links for the axes) or
# with the oracle, so agreement in joint space pins both.
# ---------------------------------------------------------------------------
def _rot_axis(a, q):
    """Rodrigues: rotation by q about unit axis a."""
    K = _skew(a)
    return np.eye(3) + np.sin(q) * K + (1.0 - np.cos(q)) * K @ K


class UrdfTree:
    def __init__(self, path):
        root = ET.parse(path).getroot()

        def vec(el, attr, default="0 0 0"):
            return np.array([float(v) for v in (el.get(attr) if el is not None and el.get(attr) else default).split()])

        self.inertial = {}
        for link in root.findall("link"):
            inr = link.find("inertial")
            if inr is None:
                continue
            o = inr.find("origin")
            i = inr.find("inertia")
            Ic = np.array([[float(i.get("ixx")), float(i.get("ixy", "0")), float(i.get("ixz", "0"))],
                           [float(i.get("ixy", "0")), float(i.get("iyy")), float(i.get("iyz", "0"))],
                           [float(i.get("ixz", "0")), float(i.get("iyz", "0")), float(i.get("izz"))]])
            self.inertial[link.get("name")] = (float(inr.find("mass").get("value")), vec(o, "xyz"),
                                               _rpy(*vec(o, "rpy")), Ic)
        self.joints = []
        for j in root.findall("joint"):
            o = j.find("origin")
            a = vec(j.find("axis"), "xyz", "1 0 0")
            self.joints.append({"name": j.get("name"), "type": j.get("type"),
                                "parent": j.find("parent").get("link"), "child": j.find("child").get("link"),
                                "xyz": vec(o, "xyz"), "R": _rpy(*vec(o, "rpy")), "axis": a / np.linalg.norm(a)})
        children = {j["child"] for j in self.joints}
        roots = {j["parent"] for j in self.joints} - children
        assert len(roots) == 1
        self.root = roots.pop()
        # parents before children
        order, frontier = [], [self.root]
        while frontier:
            nxt = [j for j in self.joints if j["parent"] in frontier]
            order += nxt
            frontier = [j["child"] for j in nxt]
        self.joints = order
        depth = {self.root: 0}
        for j in self.joints:
            depth[j["child"]] = depth[j["parent"]] + (j["type"] != "fixed")
        moving = [j for j in self.joints if j["type"] != "fixed"]
        self.qindex = {j["name"]: depth[j["child"]] - 1 for j in moving}
        assert sorted(self.qindex.values()) == list(range(NQ))

    def _x(self, j, q):
        R = j["R"] @ (_rot_axis(j["axis"], q) if j["type"] != "fixed" else np.eye(3))
        E = R.T
        X = np.zeros((6, 6))
        X[:3, :3] = E
        X[3:, 3:] = E
        X[3:, :3] = -E @ _skew(j["xyz"])
        return X

    def _inertia(self, link):
        if link not in self.inertial:
            return np.zeros((6, 6))
        m, c, Rc, Ic = self.inertial[link]
        C = _skew(c)
        I = np.zeros((6, 6))
        I[:3, :3] = Rc @ Ic @ Rc.T + m * C @ C.T
        I[:3, 3:] = m * C
        I[3:, :3] = m * C.T
        I[3:, 3:] = m * np.eye(3)
        return I

    def rnea(self, q, qd, qdd, gravity=9.81):
        v = {self.root: np.zeros(6)}
        a = {self.root: np.array([0, 0, 0, 0, 0, gravity])}
        f, X, S = {}, {}, {}
        for j in self.joints:
            c = j["child"]
            k = self.qindex.get(j["name"])
            s = np.concatenate([j["axis"], np.zeros(3)]) if k is not None else np.zeros(6)
            qk, qdk, qddk = (q[k], qd[k], qdd[k]) if k is not None else (0.0, 0.0, 0.0)
            X[c] = self._x(j, qk)
            S[c] = s
            v[c] = X[c] @ v[j["parent"]] + s * qdk
            a[c] = X[c] @ a[j["parent"]] + s * qddk + _crm(v[c]) @ s * qdk
            I = self._inertia(c)
            f[c] = I @ a[c] + _crf(v[c]) @ I @ v[c]
        tau = np.zeros(NQ)
        for j in reversed(self.joints):
            c = j["child"]
            k = self.qindex.get(j["name"])
            if k is not None:
                tau[k] = S[c] @ f[c]
            if j["parent"] in f:
                f[j["parent"]] = f[j["parent"]] + X[c].T @ f[c]
        return tau

    def mass_matrix(self, q):
        z = np.zeros(NQ)
        g0 = self.rnea(q, z, z)
        return np.stack([self.rnea(q, z, e) - g0 for e in np.eye(NQ)], axis=1)

    def potential(self, q, gravity=9.81):
        """Sum of m g z_com over the bodies, z in the root frame."""
        pose = {self.root: (np.eye(3), np.zeros(3))}
        U = 0.0
        for j in self.joints:
            k = self.qindex.get(j["name"])
            R0, p0 = pose[j["parent"]]
            Rj = j["R"] @ (_rot_axis(j["axis"], q[k]) if k is not None else np.eye(3))
            pose[j["child"]] = (R0 @ Rj, p0 + R0 @ j["xyz"])
            if j["child"] in self.inertial:
                m, c, _, _ = self.inertial[j["child"]]
                R, p = pose[j["child"]]
                U += gravity * m * (p + R @ c)[2]
        return U


# ---------------------------------------------------------------------------
# Robot models from a URDF (fast-rbd-gradients_amd/urdf_model.py) and the CPU
# build of the kernel arithmetic (tests/hostcheck) on them.
# ---------------------------------------------------------------------------
def model_tables(urdf):
    """The kernel tables of a URDF as the oracle takes them: links (7, 10) and
    joints (7, 12) = [rot row-major | offset]."""
    import urdf_model
    tab = urdf_model.tables(urdf_model.read_urdf(urdf))
    joints = np.concatenate([np.array(tab["rot"]).reshape(NQ, 9), np.array(tab["offset"])], axis=1)
    return np.array(tab["links"]), joints


@contextlib.contextmanager
def _build_lock():
    """One test process at a time writes generated headers and test libraries
    (pytest-xdist workers would otherwise race on the same files)."""
    with open(os.path.join(HOSTCHECK, ".build.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        try:
            yield
        finally:
            fcntl.flock(lk, fcntl.LOCK_UN)


def generate_model(urdf, name):
    """Write fast-rbd-gradients_amd/csrc/model_NAME.hpp; returns model_tables(urdf)."""
    import urdf_model
    with _build_lock():
        urdf_model.main([urdf, name])
    return model_tables(urdf)


def hostcheck_lib(model=None):
    """Build and load tests/hostcheck's libknot_host[_MODEL].so."""
    args = ["make", "-s", "-C", HOSTCHECK] + ([f"MODEL={model}"] if model else [])
    with _build_lock():
        subprocess.run(args, check=True)
    L = ctypes.CDLL(os.path.join(HOSTCHECK, f"libknot_host_{model}.so" if model else "libknot_host.so"))
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    L.host_fdgrad_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.host_fdgrad_refmode_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.host_model_inertia.argtypes = [ctypes.c_int, dp]
    L.host_fd_f32.argtypes = [ctypes.c_long, fp, fp, fp, ctypes.c_int]
    L.host_fdgrad_minv_f32.argtypes = [ctypes.c_long, fp, fp, fp, ctypes.c_int, ctypes.c_int]
    L.host_model_transform.argtypes = [ctypes.c_int, ctypes.c_double, dp]
    L.host_fdgrad_fext_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_long, fp, ctypes.c_int, ctypes.c_int]
    L.host_fd_fext_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_long, fp, fp, ctypes.c_int]
    L.host_id_gradient_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.host_vaf_f32.argtypes = [ctypes.c_long, fp, ctypes.c_int]
    L.host_finish_f32.argtypes = [ctypes.c_long, fp, fp, fp]
    L.host_fdgrad_two_wave_f32.argtypes = [ctypes.c_long, fp, fp, ctypes.c_int, ctypes.c_int]
    L.host_model_name.restype = ctypes.c_char_p
    L.host_minv_columns.argtypes = [ctypes.c_long, dp, dp, dp]
    L.host_fdgrad_wafr_f32.argtypes = [ctypes.c_long, fp, fp, fp, ctypes.c_int, ctypes.c_int]
    return L
