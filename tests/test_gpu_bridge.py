"""The C++ bridge (semo_mesh_amd/bridge) linked into SEMO.

oracle/_ref/semo_gpu_eq and semo_gpu_step are the reference's own objects
(compiled in place from the reference checkout by oracle/ref.mk) and the
test driver oracle/ref_driver.cpp, with the bridge linked in place of
SEMO's equation files (eq: SEMO's own incompressiblePressureBased runs and
its calcMomentumEqs / calcPressureEq / calcVolfracEq calls go to the GPU) or
of those and pressureBased.cpp (step: the whole step on the GPU, fields
resident).  Each runs a golden case as SEMO would and must end at the
fixture's fields (the compiled reference's own run, AMGCL to 1e-9): relative
1e-7, the bar of the GPU path's own full-run test.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, golden_case, load_golden

pytestmark = pytest.mark.gpu

REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def _run(binary, name, env_extra):
    path = os.path.join(REF_DIR, binary)
    if not os.path.exists(path):
        pytest.skip(f"{binary} is built only where the reference checkout is present (oracle/ref.mk)")
    d, mesh = load_golden(name)
    case = golden_case(d)
    with tempfile.TemporaryDirectory() as tmp:
        mesh.write_raw(os.path.join(tmp, "mesh.raw"))
        ip = None
        if "init" in d:
            ip = os.path.join(tmp, "init.f64")
            np.ascontiguousarray(d["init"], dtype=np.float64).tofile(ip)
        case.write_ref_case(os.path.join(tmp, "case.txt"), steps=int(d["steps"]),
                            correctors=int(d["correctors"]), init=ip)
        if case.vf_points is not None:
            case.write_vf_csv(os.path.join(tmp, "VF.csv"))
        env = dict(os.environ, **env_extra)
        r = subprocess.run([path, os.path.join(tmp, "mesh.raw"), os.path.join(tmp, "case.txt"),
                            os.path.join(tmp, "out")], capture_output=True, text=True, timeout=300,
                           cwd=tmp, env=env)
        assert "SEMO_REF_DONE" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
        out = {f: np.fromfile(os.path.join(tmp, "out", f"final_{f}.f64")) for f in ("P", "U", "V", "W", "VF")}
    return d, out


@pytest.mark.parametrize("binary", ["semo_gpu_eq", "semo_gpu_step"])
@pytest.mark.parametrize("name", ["box654_jitter_2x2", "vof_box654_2x2", "bc_vof_box654_2x1"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_semo_with_the_gpu_bridge_reaches_the_reference_fields(binary, name, prec):
    env = {"SEMO_GPU_PRECISION": prec, "SEMO_GPU_TOL": "1e-13" if prec == "fp64" else "1e-10"}
    d, out = _run(binary, name, env)
    vel = np.sqrt(sum(np.linalg.norm(d["final_" + f]) ** 2 for f in "UVW"))
    for f in ("P", "U", "V", "W", "VF"):
        a, b = out[f], d["final_" + f]
        scale = np.linalg.norm(b) if f in ("P", "VF") else vel
        err = np.linalg.norm(a - b) / scale
        print(f"{binary} {name} {prec} {f}: rel {err:.2e}")
        assert err <= 1e-7, f
