"""Write the golden fixtures under tests/golden/ from the verified FP64 oracle.

The reference repository holds no golden vectors, known-answer tests or
recorded outputs (its only check is TEST_FOR_EQUIVALENCE printing values at
q = qd = u = 0.1, utils/experiment_helpers.h:11-16), so the expected outputs
here come from the oracle after it has been pinned by tests/test_oracle.py
(complex-step derivatives, Minv M = I, RNEA(FD) = u, the URDF-built textbook
model).  Inputs are stored as float32 and the oracle runs on exactly those
values; outputs are float64.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import common  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def main():
    sets = {}
    # config 0: one trajectory of 16 knots; config 1: 128 knots (seeded)
    sets["c0_n16"] = (common.knots(16, seed=100), 0)
    sets["c1_n128"] = (common.knots(128, seed=101), 0)
    # stretched-arm poses: joints 2, 4, 6 within 1e-3 rad of zero
    sets["near_singular_n256"] = (common.knots(256, seed=102, near_singular_frac=1.0), 0)
    # the reference's TEST_FOR_EQUIVALENCE inputs: every q, qd, u = 0.1 (time_CPU.cpp:68-71)
    sets["equivalence_0p1"] = (np.full((1, 21), 0.1, dtype=np.float32), 0)
    # model options
    sets["vel_damping_n64"] = (common.knots(64, seed=103), O.VEL_DAMPING)
    sets["no_gravity_n64"] = (common.knots(64, seed=104), O.NO_GRAVITY)
    sets["damping_no_gravity_n64"] = (common.knots(64, seed=105), O.VEL_DAMPING | O.NO_GRAVITY)
    # zero velocity: dqdd/dqd must be exactly -Minv * 0 ... = 0 without damping
    z = common.knots(32, seed=106)
    z[:, 7:14] = 0.0
    sets["zero_velocity_n32"] = (z, 0)
    arrays = {}
    for name, (x, flags) in sets.items():
        arrays[name + "__in"] = x.astype(np.float32)
        arrays[name + "__flags"] = np.array(flags, dtype=np.int32)
        arrays[name + "__out"] = O.fdgrad(x.astype(np.float64), flags)
    # reference-compatible mode: q, qd, u, qdd, Minv (dense) -> dqdd
    x = common.knots(128, seed=107).astype(np.float64)
    vol = np.zeros((128, 77))
    for i in range(128):
        qdd, minv = O.fd(x[i, :7], x[i, 7:14], x[i, 14:])
        vol[i, :21] = x[i]
        vol[i, 21:28] = qdd
        vol[i, 28:] = minv.T.reshape(-1)          # column-major
    vol = vol.astype(np.float32)
    arrays["refmode_n128__in"] = vol
    arrays["refmode_n128__flags"] = np.array(0, dtype=np.int32)
    arrays["refmode_n128__out"] = O.fdgrad_refmode(vol.astype(np.float64))
    np.savez_compressed(os.path.join(HERE, "fdgrad_golden.npz"), **arrays)
    print("wrote", sorted(arrays))


if __name__ == "__main__":
    main()
