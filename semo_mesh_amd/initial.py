"""SEMO's initial-field utility (src/utility/initial.cpp:48-540) for the
path's ``Mesh``: the ``initialField`` blocks of a case's ``constant/``
files turned into the cell fields a run starts from.

Files and types (initial.cpp:144-530):

* ``constant/pressure``, ``constant/temperature``: ``fixedValue``,
  ``boxToCell``;
* ``constant/velocities``: ``fixedValue``, ``boxToCell``,
  ``cylinderToCell``, ``sphereToCell`` (vector values ``(u v w)``);
* ``constant/<first species>``: ``fixedValue``, ``boxToCell``,
  ``cylinderToCell``, ``sphereToCell``; the value goes to that species'
  volume and mass fractions both.

A block is read line by line as initialConditionSetting /
initialConditionSettingVel do (initial.cpp:695-1118): the line holding
``type`` names it and the lines after it hold, in order, the default value,
the region (``(x0 y0 z0) (x1 y1 z1)`` box corners or cylinder axis ends, or
``(x y z)`` sphere centre), the radius (cylinder and sphere) and the value
inside; the first word of each line is a key and is ignored.

Regions test the cell's vertex average (the mean of its points, summed in
the cell's point order, initial.cpp:88-105), not its centroid: inclusive
box bounds; a cylinder holds a point whose projections on the axis fall
strictly between its ends and whose distance from the axis is below the
radius; a sphere holds a point at squared distance not above the squared
radius.

``function`` (a compiled ``constant/initialFunctions.so`` in SEMO) takes a
Python callable f(t, x, y, z) here, evaluated at t = 0 and each cell's
vertex average (SEMO passes cell.x, y, z, which that utility leaves unset);
it is not compared with the reference.
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import Mesh

SCALAR_TYPES = ("fixedValue", "boxToCell", "function")
VECTOR_TYPES = ("fixedValue", "boxToCell", "cylinderToCell", "sphereToCell", "function")
SPECIES_TYPES = ("fixedValue", "boxToCell", "cylinderToCell", "sphereToCell", "function")


def _block(path: str):
    """Lines of the initialField block (from the line after the one naming
    it up to and including the first line holding '}'), initial.cpp:706-722."""
    out, reading = [], False
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if reading:
                if "}" in line:
                    reading = False
                out.append(line)
            elif "initialField" in line:
                reading = True
    return out


def _num(tok: str) -> float:
    return float(tok.replace("(", "").replace(")", "").replace(";", ""))


def _nums(line: str, count: int):
    """The `count` numbers after the line's first word (parentheses and the
    closing ';' dropped, as the reference erases them)."""
    toks = line.split()[1:1 + count]
    if len(toks) < count:
        raise ValueError(f"initialField: expected {count} numbers in {line!r}")
    return [_num(t) for t in toks]


def read_initial_field(path: str, vector: bool = False):
    """(type, values) of a file's initialField block, in the reference's
    value order (initial.cpp:695-1118)."""
    lines = _block(path)
    for i, line in enumerate(lines):
        if "type" not in line:
            continue
        typ = line.split()[1].replace(";", "")
        nd = 3 if vector else 1
        if typ == "fixedValue":
            return typ, _nums(lines[i + 1], nd)
        if typ == "boxToCell":
            return typ, _nums(lines[i + 1], nd) + _nums(lines[i + 2], 6) + _nums(lines[i + 3], nd)
        if typ == "cylinderToCell":
            return typ, (_nums(lines[i + 1], nd) + _nums(lines[i + 2], 6) + _nums(lines[i + 3], 1)
                         + _nums(lines[i + 4], nd))
        if typ == "sphereToCell":
            return typ, (_nums(lines[i + 1], nd) + _nums(lines[i + 2], 3) + _nums(lines[i + 3], 1)
                         + _nums(lines[i + 4], nd))
        return typ, []
    raise ValueError(f"{path}: no type in the initialField block")


def vertex_average(mesh: Mesh) -> np.ndarray:
    """[n, 3] mean of each cell's points, summed in the cell's point order
    then divided by the count (initial.cpp:88-105)."""
    ptr = np.asarray(mesh.cell_ptr, dtype=np.int64)
    cnt = np.diff(ptr)
    acc = np.zeros((mesh.n_cells, 3))
    pts = np.asarray(mesh.points, dtype=np.float64)
    cp = np.asarray(mesh.cell_pts, dtype=np.int64)
    for k in range(int(cnt.max()) if len(cnt) else 0):
        has = cnt > k
        acc[has] += pts[cp[ptr[:-1][has] + k]]
    return acc / cnt[:, None].astype(np.float64)


def _in_box(c, lo, hi):
    return ((c[:, 0] >= lo[0]) & (c[:, 0] <= hi[0]) & (c[:, 1] >= lo[1]) & (c[:, 1] <= hi[1])
            & (c[:, 2] >= lo[2]) & (c[:, 2] <= hi[2]))


def _in_cylinder(c, p0, p1, radius):
    """initial.cpp:232-266 (velocity) / 425-460 (species)."""
    pv = np.asarray(p1) - np.asarray(p0)
    v0 = c - np.asarray(p0)
    v1 = c - np.asarray(p1)
    pmag = np.sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2])
    m0 = np.sqrt(v0[:, 0] * v0[:, 0] + v0[:, 1] * v0[:, 1] + v0[:, 2] * v0[:, 2])
    m1 = np.sqrt(v1[:, 0] * v1[:, 0] + v1[:, 1] * v1[:, 1] + v1[:, 2] * v1[:, 2])
    cos0 = (v0[:, 0] * pv[0] + v0[:, 1] * pv[1] + v0[:, 2] * pv[2]) / pmag / (m0 + 1.e-200)
    cos1 = -(v1[:, 0] * pv[0] + v1[:, 1] * pv[1] + v1[:, 2] * pv[2]) / pmag / (m1 + 1.e-200)
    u = pv / pmag
    inside = (cos0 > 0.0) & (cos1 > 0.0)
    proj = u[0] * v0[:, 0] + u[1] * v0[:, 1] + u[2] * v0[:, 2]
    with np.errstate(invalid="ignore"):
        dist = np.sqrt(m0 * m0 - proj * proj)
    return inside & ~(dist >= radius)


def _in_sphere(c, o, radius):
    """initial.cpp:280-296 / 465-480: outside when the squared distance
    exceeds the squared radius."""
    d = c - np.asarray(o)
    return ~(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2] > radius * radius)


def region_field(typ: str, values, xyz: np.ndarray, ncomp: int, func=None) -> np.ndarray:
    """[n, ncomp] field of one initialField block over the cell positions."""
    n = len(xyz)
    v = list(values)
    d = np.asarray(v[:ncomp], dtype=np.float64) if v else None
    out = np.empty((n, ncomp))
    if typ == "fixedValue":
        out[:] = d
    elif typ == "boxToCell":
        inside = _in_box(xyz, v[ncomp:ncomp + 3], v[ncomp + 3:ncomp + 6])
        out[:] = d
        out[inside] = np.asarray(v[ncomp + 6:ncomp + 6 + ncomp])
    elif typ == "cylinderToCell":
        inside = _in_cylinder(xyz, v[ncomp:ncomp + 3], v[ncomp + 3:ncomp + 6], v[ncomp + 6])
        out[:] = d
        out[inside] = np.asarray(v[ncomp + 7:ncomp + 7 + ncomp])
    elif typ == "sphereToCell":
        inside = _in_sphere(xyz, v[ncomp:ncomp + 3], v[ncomp + 3])
        out[:] = d
        out[inside] = np.asarray(v[ncomp + 4:ncomp + 4 + ncomp])
    elif typ == "function":
        if func is None:
            raise ValueError("initialField type 'function' needs a callable f(t, x, y, z)")
        for i in range(n):
            out[i] = func(0.0, xyz[i, 0], xyz[i, 1], xyz[i, 2])
    else:
        raise ValueError(f"initialField type {typ!r} is not one SEMO's initial utility knows")
    return out


def initial_fields(case_dir: str, mesh: Mesh, species: str = "water", functions: dict | None = None) -> dict:
    """P, U, V, W, T and VF (= the first species' volume and mass fraction)
    from case_dir/constant/{pressure, velocities, temperature, <species>},
    as SEMO's initial utility sets them (initial.cpp:144-530).  functions:
    {"pressure" | "velocities" | "temperature" | species: f(t, x, y, z)}."""
    functions = functions or {}
    xyz = vertex_average(mesh)
    const = os.path.join(case_dir, "constant")
    out = {}
    for fname, key, allowed, vec in (("pressure", "P", SCALAR_TYPES, False),
                                     ("velocities", "UVW", VECTOR_TYPES, True),
                                     ("temperature", "T", SCALAR_TYPES, False),
                                     (species, "VF", SPECIES_TYPES, False)):
        typ, vals = read_initial_field(os.path.join(const, fname), vector=vec)
        if typ not in allowed:
            raise ValueError(f"constant/{fname}: initialField type {typ!r} is not read for this field")
        f = region_field(typ, vals, xyz, 3 if vec else 1, functions.get(fname))
        if vec:
            out["U"], out["V"], out["W"] = f[:, 0].copy(), f[:, 1].copy(), f[:, 2].copy()
        else:
            out[key] = f[:, 0].copy()
    return out
