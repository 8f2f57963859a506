"""Pins the oracle's (and the product library's) planogram geometry against the reference's own
src/planogram.cpp: the committed golden table was produced by compiling that file from the reference
sources (tests/golden/make_geometry_golden.py); where the reference build is present, it is also
called directly."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "geometry_golden.json")
FIELDS = ["psi", "R", "L", "H", "v_m0", "v_m1", "T_u0", "T_u1", "T_v0", "T_v1", "u0_0", "u1_0", "v0_0", "v1_0",
          "cos_psi", "sin_psi"]


def load_golden():
    with open(GOLDEN) as fh:
        return json.load(fh)


@pytest.mark.parametrize("entry", load_golden(), ids=lambda e: "_".join(str(x) for x in e["call"][:5]))
def test_oracle_geometry_bit_exact(oracle, entry):
    psi, D, L, H, T, v_m0, v_m1 = entry["call"]
    p = oracle.planogram(psi, D, L, H, T, v_m0, v_m1)
    assert list(p.shape) == entry["shape"]
    mine = np.array([getattr(p, k) for k in FIELDS], np.float32).view(np.int32)
    want = np.array([entry["bits"][k] for k in FIELDS], np.int32)
    assert np.array_equal(mine, want), {k: (getattr(p, k), entry["fields"][k]) for k in FIELDS}


@pytest.mark.parametrize("entry", load_golden(), ids=lambda e: "_".join(str(x) for x in e["call"][:5]))
def test_library_planogram_shape(petl_lib, entry):
    from petl_amd.petl_server import PETLserver
    psi, D, L, H, T, v_m0, v_m1 = entry["call"]
    s = PETLserver()
    s.add_planogram(psi, D, L, H, T, v_m0, v_m1)
    assert list(s.get_planogramSize(0)) == entry["shape"]


def test_reference_build_matches_golden(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference geometry build not present")
    for e in load_golden():
        psi, D, L, H, T, v_m0, v_m1 = e["call"]
        shape = np.zeros(4, np.int32)
        vals = np.zeros(16, np.float32)
        assert R.ref_plan_init(psi, D, L, H, v_m0, v_m1, T, shape, vals)
        assert list(shape) == e["shape"]
        assert np.array_equal(vals.view(np.int32), np.array([e["bits"][k] for k in FIELDS], np.int32))


def test_volume_grid(oracle):
    # src/parameters.cpp:271-284: voxel-0 centre -(n-1)/2 * T + offset
    v = oracle.volume(128, 64, 32, 2.0, 3.0, 1.0, -2.0, 0.5)
    assert (v.x0, v.y0, v.z0) == (-127.0 + 1.0, -63.0 - 2.0, -46.5 + 0.5)
