"""SEMO's initial-field region initialisers (semo_mesh_amd.initial), CPU only.

Pinned by tests/golden/initial_cube6.npz: SEMO's own initial-field utility
(src/utility/initial.cpp, compiled as it is into oracle/_ref/semo_initial)
run on a jittered 6^3 cube for three sets of initialField blocks covering
pressure boxToCell / fixedValue, velocities cylinderToCell / boxToCell /
fixedValue, temperature boxToCell / fixedValue and the first species'
sphereToCell / cylinderToCell / boxToCell: every field the utility saved
equals ours exactly.  (Velocities sphereToCell is implemented from the scalar
form but not compared: the reference's parser throws on every input.)"""
import os

import numpy as np
import pytest

from conftest import load_golden
from semo_mesh_amd.initial import initial_fields, read_initial_field, region_field, vertex_average

SPECS = ("box_cyl_box_sphere", "fixed_box_fixed_cyl", "box_fixed_box_box")


def _write_case(tmp_path, spec):
    const = os.path.join(tmp_path, "constant")
    os.makedirs(const, exist_ok=True)
    for fname, body in zip(("pressure", "velocities", "temperature", "water"), spec):
        with open(os.path.join(const, fname), "w") as fh:
            fh.write("initialField\n{\n" + "".join(f"\t{ln}\n" for ln in str(body).split("\n"))
                     + "}\n\nboundaryField\n{\n}\n")


@pytest.mark.parametrize("name", SPECS)
def test_initial_fields_match_semo_utility(tmp_path, name):
    d, m = load_golden("initial_cube6")
    _write_case(tmp_path, d[name + "_spec"])
    f = initial_fields(str(tmp_path), m)
    assert np.array_equal(f["P"], d[name + "_P"])
    uvw = d[name + "_UVW"]
    for k, c in enumerate("UVW"):
        assert np.array_equal(f[c], uvw[:, k]), c
    assert np.array_equal(f["T"], d[name + "_T"])
    assert np.array_equal(f["VF"], d[name + "_VF"])
    assert np.array_equal(f["VF"], d[name + "_MF"])
    # the regions cut the cube: both values occur in every region field
    for key, arr in (("P", d[name + "_P"]), ("VF", d[name + "_VF"])):
        if "fixed" not in name.split("_")[0 if key == "P" else 3]:
            assert len(np.unique(arr)) == 2, key


def test_vertex_average_is_the_point_mean():
    _, m = load_golden("initial_cube6")
    c = vertex_average(m)
    i = 17
    pts = m.points[m.cell_pts[m.cell_ptr[i]:m.cell_ptr[i + 1]]]
    acc = np.zeros(3)
    for p in pts:
        acc += p
    assert np.array_equal(c[i], acc / len(pts))


def test_parser_value_order(tmp_path):
    p = os.path.join(tmp_path, "velocities")
    with open(p, "w") as fh:
        fh.write("initialField\n{\n\ttype cylinderToCell;\n\tvalue (1 2 3);\n\tpoints (0 0 0) (1 0 0);\n"
                 "\tradius 0.5;\n\tvalue (4 5 6);\n}\n")
    t, v = read_initial_field(p, vector=True)
    assert t == "cylinderToCell" and v == [1, 2, 3, 0, 0, 0, 1, 0, 0, 0.5, 4, 5, 6]


def test_region_rules_on_points():
    """Closed forms of the three region tests: inclusive box bounds, the
    cylinder's open ends and radius, the sphere's closed ball."""
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0000001, 0.5, 0.5],
                    [0.5, 0.2, 0.0], [0.0, 0.1, 0.0], [0.5, 0.3, 0.0], [0.5, 0.0, 0.4]])
    box = region_field("boxToCell", [7, 0, 0, 0, 1, 1, 1, 9], xyz, 1)[:, 0]
    assert box.tolist() == [9, 9, 7, 9, 9, 9, 9]
    cyl = region_field("cylinderToCell", [7, 0, 0, 0, 1, 0, 0, 0.25, 9], xyz, 1)[:, 0]
    # on an end plane (x = 0) is outside; radius 0.25 from the x axis
    assert cyl.tolist() == [7, 7, 7, 9, 7, 7, 7]
    sph = region_field("sphereToCell", [7, 0.5, 0.0, 0.0, 0.4, 9], xyz, 1)[:, 0]
    assert sph.tolist() == [7, 7, 7, 9, 7, 9, 9]


def test_function_and_unknown_types():
    xyz = np.array([[0.5, 0.25, 1.0]])
    v = region_field("function", [], xyz, 1, func=lambda t, x, y, z: x + 2 * y + 3 * z)
    assert v[0, 0] == 4.0
    with pytest.raises(ValueError):
        region_field("function", [], xyz, 1)
    with pytest.raises(ValueError):
        region_field("coneToCell", [0], xyz, 1)
