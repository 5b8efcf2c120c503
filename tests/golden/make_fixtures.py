"""Writes the oracle regression fixtures of tests/golden/ (run in this repo;
needs oracle/liboracle.so).  These are OUTPUTS OF THE ORACLE, not of Simbody:
the reference is not buildable here (it needs SimTKcommon's CMake build and an
external LAPACK), so they pin the oracle against silent change and give the
GPU tests fixed expected values; the oracle itself is pinned by
tests/test_oracle.py.

  pendulum10.json : 10-link planar Pin chain, seeded state, udot,
                    and q,u after RK-Merson accuracy 1e-8 to t = 1 with the
                    accept/reject counts.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT); sys.path.insert(0, os.path.join(ROOT, "oracle"))

from simbody_amd import models as M  # noqa: E402
from oracle_ctypes import OracleModel  # noqa: E402


def pendulum10():
    spec = M.pendulum_chain(10)
    o = OracleModel(spec)
    rng = np.random.default_rng(20261202)
    q = rng.uniform(-np.pi, np.pi, 10); u = rng.normal(size=10)
    ud, qd, qdd, _ = o.realize_acceleration(q, u)
    y, st = o.integrate(np.concatenate([q, u]), 1.0, 1e-8)
    return {"what": "10-link planar Pin chain (simbody_amd.models.pendulum_chain(10)); "
                    "oracle udot at (q,u) and RK-Merson accuracy 1e-8 to t=1",
            "q": q.tolist(), "u": u.tolist(), "udot": ud.tolist(),
            "q1": y[:10].tolist(), "u1": y[10:].tolist(),
            "attempts": st["attempts"], "accepted": st["accepted"], "rejected": st["rejected"]}


if __name__ == "__main__":
    with open(os.path.join(HERE, "pendulum10.json"), "w") as f:
        json.dump(pendulum10(), f, indent=1)
    print("wrote pendulum10.json")
