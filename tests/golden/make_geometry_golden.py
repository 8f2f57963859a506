"""Writes geometry_golden.json: planogram sampling produced by the reference's own src/planogram.cpp
(compiled from the reference sources into oracle/_ref by `make -C oracle ref`) for the add_planogram
calls the benchmark configurations and demos make.  Run in a container that has the reference sources."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

CALLS = [
    # psi, D, L, H, T, v_m0, v_m1 (None = PETLserver.add_planogram defaults)
    (0.0, 100.0, 64.0, 64.0, 2.0, None, None),
    (0.0, 200.0, 256.0, 256.0, 2.0, None, None),
    (90.0, 200.0, 256.0, 256.0, 2.0, None, None),
    (0.0, 200.0, 128.0, 128.0, 2.0, None, None),
    (90.0, 200.0, 96.0, 96.0, 2.0, None, None),
    (0.0, 200.0, 192.0, 192.0, 2.0, None, None),
    (30.0, 180.0, 192.0, 144.0, 2.0, float(np.tan(15.0 * np.pi / 180)), float(np.tan(15.0 * np.pi / 180))),
    (0.0, 101.0, 251.52, 156.72, 1.6, None, None),
    (90.0, 266.6, 125.12, 156.72, 1.6, None, None),
    (0.0, 60.0, 32.0, 32.0, 2.0, None, None),
]
FIELDS = ["psi", "R", "L", "H", "v_m0", "v_m1", "T_u0", "T_u1", "T_v0", "T_v1", "u0_0", "u1_0", "v0_0", "v1_0",
          "cos_psi", "sin_psi"]


def main():
    O.build()
    R = O.ref_lib()
    if R is None:
        raise SystemExit("reference geometry library not built (needs the reference sources)")
    out = []
    for psi, D, L, H, T, vm0, vm1 in CALLS:
        v_m0 = 1.0 if vm0 is None else vm0
        v_m1 = min(1.0, (H - 1.5 * T) / D) if vm1 is None else vm1
        shape = np.zeros(4, np.int32)
        vals = np.zeros(16, np.float32)
        assert R.ref_plan_init(psi, D, L, H, v_m0, v_m1, T, shape, vals)
        out.append({"call": [psi, D, L, H, T, v_m0, v_m1], "shape": [int(s) for s in shape],
                    "fields": {k: float(v) for k, v in zip(FIELDS, vals)},
                    "bits": {k: int(b) for k, b in zip(FIELDS, vals.view(np.int32))}})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "geometry_golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
