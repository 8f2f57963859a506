"""libpetl.so on the CPU: it loads, exports every symbol include/petl_c_interface.h declares, and its
host-side parameter API behaves as the reference's (src/petl_c_interface.cpp, src/parameters.cpp).
No kernel is launched here."""
import os
import re

import numpy as np
import pytest

from petl_amd.petl_server import PETLserver

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "petl_c_interface.h")


def declared_symbols():
    text = open(HEADER).read()
    return re.findall(r"PETL_EXTERN\s+[\w\s\*]+?\b(\w+)\s*\(", text)


def test_every_declared_symbol_is_exported(petl_lib):
    names = declared_symbols()
    assert len(names) >= 60
    missing = [n for n in names if not hasattr(petl_lib, n)]
    assert not missing


def test_reference_abi_names_present(petl_lib):
    # the full export list of the reference's src/petl_c_interface.h:14-97
    ref = ("set_model create_new_model about copy_parameters clearAll print_parameters volume_defined "
           "geometry_defined get_numX set_numX get_numY set_numY get_numZ set_numZ get_voxelWidth set_voxelWidth "
           "get_voxelHeight set_voxelHeight get_offsetX set_offsetX get_offsetY set_offsetY get_offsetZ set_offsetZ "
           "get_numPlanograms get_planogramSize add_planogram remove_planogram keep_only_planogram "
           "set_default_volume set_volume bin simulate_scatter project backproject FBP PFDR add_object add_module "
           "clear_phantom clear_modules scale_phantom voxelize ray_trace stopping_power set_solid_angle_correction "
           "apply_corrections multiply3D multiply4D divide3D divide4D rdivide3D rdivide4D reciprocal3D reciprocal4D "
           "inner_product3D inner_product4D TV_cost TV_gradient TV_quadForm diffuse TV_denoise "
           "relativeDifferences_gradient gaussian_filter").split()
    assert len(ref) == 64
    for n in ref:
        assert hasattr(petl_lib, n), n


def test_parameter_api():
    s = PETLserver()
    assert not s.all_defined()
    s.add_planogram(0.0, 200.0, 256.0, 256.0, 2.0)
    s.add_planogram(90.0, 200.0, 256.0, 256.0, 2.0)
    assert s.get_numPlanograms() == 2
    assert list(s.get_planogramSize(1)) == [101, 101, 128, 228]
    s.set_volume(128, 128, 128, 2.0)
    assert s.all_defined()
    assert (s.numX(), s.numY(), s.numZ(), s.voxelWidth(), s.voxelHeight()) == (128, 128, 128, 2.0, 2.0)
    # src/petl_c_interface.cpp:154-195: offsets must be positive
    assert s.offsetX(-1.0) is None and s.offsetX() == 0.0
    assert s.offsetX(3.0) == 3.0
    s.keep_only_planogram(1)
    assert s.get_numPlanograms() == 1
    s.remove_planogram(0)
    assert not s.geometry_defined()
    with pytest.raises(ValueError):
        s.add_planogram(0.0, -1.0, 10.0, 10.0, 2.0)
    assert s.allocate_projections() is None


def test_models_are_independent():
    a, b = PETLserver(), PETLserver()
    a.add_planogram(0.0, 100.0, 64.0, 64.0, 2.0)
    assert b.get_numPlanograms() == 0
    assert a.get_numPlanograms() == 1
    b.set_model(a.param_id)
    assert b.libprojectors.copy_parameters(a.param_id)


def test_set_default_volume():
    """src/parameters.cpp:205-248 for the PET/X preset (src/petx_server.py:8-11)."""
    s = PETLserver()
    s.add_planogram(psi=0.0, D=101.0, L=251.52, H=156.72, T=1.6)
    s.add_planogram(psi=90.0, D=266.6, L=125.12, H=156.72, T=1.6)
    s.set_default_volume()
    # X = min(L0, |L1 cos90| + |R1 sin90|) = min(125.76, ~133.3); Y = min(R0, L1) = min(50.5, 62.56)
    f32 = np.float32
    X = min(f32(125.76), abs(f32(62.56) * np.cos(f32(np.pi / 2), dtype=f32)) + abs(f32(133.3)))
    assert s.numX() == 2 * int(f32(X) / f32(1.6))
    assert s.numY() == 2 * int(f32(50.5) / f32(1.6))
    assert s.numZ() == 2 * int(f32(78.36) / f32(1.6))


def test_gpu_entry_points_fail_loudly_without_gpu():
    import subprocess
    import sys
    code = ("import numpy as np; from petl_amd.petl_server import PETLserver; s = PETLserver(); "
            "s.add_planogram(0, 60, 32, 32, 2); s.set_volume(8, 8, 8, 2.0); "
            "s.project(s.allocate_projections(), np.ones((8, 8, 8), np.float32))")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       cwd=os.path.dirname(os.path.dirname(HEADER)))
    assert r.returncode != 0 and "project failed" in r.stderr
