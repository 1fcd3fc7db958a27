"""Case descriptions for the pressure-based incompressible path.

A case is the physics and boundary conditions that SEMO reads from
``constant/*`` and ``system/*`` dictionaries (src/controls/build.cpp:7-82,
87-623; src/load/loadFiles.cpp:12-190): two constant-density phases
(index 0 = the phase tracked by the volume fraction), gravity, surface
tension, a fixed time step, under-relaxation factors and per-patch boundary
conditions.
"""
from __future__ import annotations

import numpy as np

from dataclasses import dataclass, field

# boundary-condition type codes shared with the C ABI (include/semo_gpu.h)
BC_ZERO_GRADIENT = 0
BC_FIXED_VALUE = 1
BC_NO_SLIP = 2
BC_SLIP = 3
BC_SURFACE_NORMAL_FIXED_VALUE = 4
BC_INLET_OUTLET = 5
BC_SWITCH = 6             # pressure: owner value above Mach 1, else fixed (reconIncom.cpp:618-631)
BC_FUNCTION = 7           # volume fraction: nearest point of ./VF.csv (reconIncom.cpp:929-1019)
BC_NAMES = {BC_ZERO_GRADIENT: "zeroGradient", BC_FIXED_VALUE: "fixedValue",
            BC_NO_SLIP: "noSlip", BC_SLIP: "slip",
            BC_SURFACE_NORMAL_FIXED_VALUE: "surfaceNormalFixedValue",
            BC_INLET_OUTLET: "inletOutlet", BC_SWITCH: "switch", BC_FUNCTION: "function"}


@dataclass
class PatchBC:
    p_type: int = BC_ZERO_GRADIENT
    p_value: float = 0.0
    u_type: int = BC_NO_SLIP
    u_value: tuple = (0.0, 0.0, 0.0)
    vf_type: int = BC_ZERO_GRADIENT
    vf_value: float = 0.0


@dataclass
class Case:
    rho: tuple = (1000.0, 1000.0)
    mu: tuple = (1e-3, 1e-3)
    sigma: float = 0.0
    g: tuple = (0.0, 0.0, 0.0)
    dt: float = 0.01
    urf_mom: float = 0.3      # momVelURF  (eqMomentum.cpp:1040)
    urf_p: float = 0.3        # prePreURF  (eqPressure.cpp:827)
    urf_pu: float = 0.3       # preVelURF  (eqPressure.cpp:832)
    limit: float = 1e10       # |P|,|U|,|V|,|W| clip bounds (fvSolution limiter)
    urf_vf: float = 1.0       # vofVofURF  (eqVolfrac.cpp:776)
    n_vof: int = 0            # nSpeciesEq: VOF iterations per time step (pressureBased.cpp:104-127)
    adjust_dt: bool = False   # adjustTimeStep: dt = maxCo * min time scale (timestep.cpp:260-330)
    max_co: float = 1.0       # maxCo
    max_dt: float = 1.0       # maxTimeStep
    bcs: dict = field(default_factory=dict)   # patch name -> PatchBC
    # points (VF, x, y, z) of the volume-fraction "function" boundary: the
    # rows of SEMO's ./VF.csv, columns 3-6 (reconIncom.cpp:929-1019)
    vf_points: np.ndarray | None = None
    # turbType LES with this turbLESModel (constant/turbulenceProperties,
    # controls/build.cpp:100-115): "WALE", "smagorinsky" or
    # "dynamicSmagorinsky"; None: laminar
    les_model: str | None = None

    def write_vf_csv(self, path: str) -> None:
        """./VF.csv as reconIncomZeroOrder reads it: a header line, then
        comma-separated rows whose columns 3-6 are VF, x, y, z."""
        with open(path, "w") as fh:
            fh.write("a,b,c,VF,x,y,z\n")
            for k, (vf, x, y, z) in enumerate(np.asarray(self.vf_points, dtype=np.float64).reshape(-1, 4)):
                fh.write(f"{k},0,0,{float(vf)!r},{float(x)!r},{float(y)!r},{float(z)!r}\n")

    def write_ref_case(self, path: str, steps: int = 1, correctors: int = 1,
                       init: str | None = None) -> None:
        """The case file of oracle/ref_driver.cpp (numbers as shortest
        round-trip decimals, read back exactly by operator>>)."""
        r = lambda x: repr(float(x))  # noqa: E731
        with open(path, "w") as fh:
            fh.write(f"dt {r(self.dt)}\nrho {r(self.rho[0])} {r(self.rho[1])}\n")
            fh.write(f"mu {r(self.mu[0])} {r(self.mu[1])}\nsigma {r(self.sigma)}\n")
            fh.write("g %s %s %s\n" % tuple(r(x) for x in self.g))
            fh.write(f"urf {r(self.urf_mom)} {r(self.urf_p)} {r(self.urf_pu)}\n")
            fh.write(f"limit {r(self.limit)}\nsteps {int(steps)}\ncorrectors {int(correctors)}\n")
            if self.n_vof:
                fh.write(f"vof {int(self.n_vof)}\nurfvf {r(self.urf_vf)}\n")
            if self.les_model:
                fh.write(f"turb {self.les_model}\n")
            if self.adjust_dt:
                fh.write(f"adjustdt 1 {r(self.max_co)} {r(self.max_dt)}\n")
            if init:
                fh.write(f"init {init}\n")
            for name, b in self.bcs.items():
                fh.write("bc %s %s %s %s %s %s %s %s %s\n" % (
                    name, BC_NAMES[b.p_type], r(b.p_value), BC_NAMES[b.u_type],
                    r(b.u_value[0]), r(b.u_value[1]), r(b.u_value[2]),
                    BC_NAMES[b.vf_type], r(b.vf_value)))


def cavity_case(dt: float = 0.01, lid=(1.0, 0.0, 0.0)) -> Case:
    """Lid-driven cavity: lid (zmax) fixedValue U, other walls noSlip,
    pressure zeroGradient everywhere except the one-face ``pref`` patch
    (fixedValue p = 0).  Single phase run through the two-phase path:
    alpha = 1, both phases rho = 1000, mu = 1e-3, sigma = 0, g = 0."""
    wall = PatchBC()
    bcs = {n: wall for n in ("xmin", "xmax", "ymin", "ymax", "zmin")}
    bcs["pref"] = PatchBC(p_type=BC_FIXED_VALUE, p_value=0.0)
    bcs["zmax"] = PatchBC(u_type=BC_FIXED_VALUE, u_value=tuple(lid))
    return Case(dt=dt, bcs=bcs)


def dam_break_case(dt: float = 1e-4, n_vof: int = 1, urf_vf: float = 1.0) -> Case:
    """Two-phase VOF dam break: water (phase 0: rho 1000, mu 1e-3) and air
    (rho 1.2, mu 1.8e-5), gravity -9.81 in z.  Walls (xmin, xmax, ymin, ymax,
    zmin): noSlip, zeroGradient p and alpha.  Top (zmax): open, fixedValue
    p = 0, zeroGradient U, inletOutlet alpha = 0 (air enters)."""
    wall = PatchBC()
    bcs = {n: wall for n in ("xmin", "xmax", "ymin", "ymax", "zmin")}
    bcs["zmax"] = PatchBC(p_type=BC_FIXED_VALUE, p_value=0.0, u_type=BC_ZERO_GRADIENT,
                          vf_type=BC_INLET_OUTLET, vf_value=0.0)
    return Case(rho=(1000.0, 1.2), mu=(1e-3, 1.8e-5), g=(0.0, 0.0, -9.81), dt=dt,
                urf_vf=urf_vf, n_vof=n_vof, bcs=bcs)


def channel_case(u_in: float, dt: float, rho: float = 1.177, mu: float = 1.85e-5) -> Case:
    """Channel flow: inlet (xmin) fixedValue U = (u_in, 0, 0), outlet (xmax)
    fixedValue p = 0 with zeroGradient U, the four side walls noSlip;
    single phase (air at 300 K by default)."""
    wall = PatchBC()
    bcs = {n: wall for n in ("ymin", "ymax", "zmin", "zmax")}
    bcs["xmin"] = PatchBC(u_type=BC_FIXED_VALUE, u_value=(u_in, 0.0, 0.0))
    bcs["xmax"] = PatchBC(p_type=BC_FIXED_VALUE, p_value=0.0, u_type=BC_ZERO_GRADIENT)
    return Case(rho=(rho, rho), mu=(mu, mu), dt=dt, bcs=bcs)


def channel_workload(nx: int = 256, ny: int = 64, nz: int = 64):
    """The block-coupled channel configuration: a 4 x 1 x 1 box of nx x ny x
    nz jittered hexes (seed 20261204) spanning x in [1, 5] -- clear of the
    reference's live inlet UDF, which rewrites fixedValue-U faces within
    1.6e-4 of x = 0.0006 k, k = 0..23 (reconIncom.cpp:707-727) -- air
    (rho 1.177, mu 1.85e-5), inlet U = Mach 0.3 at 300 K (gamma 1.4) with a
    seeded sinusoidal perturbation of 2 % amplitude, outlet p = 0, noSlip
    walls, dt at Courant 1 on the inlet velocity, under-relaxation 0.7 (u)
    and 0.3 (p).  Returns (mesh, case, u_in, per-face inlet values)."""
    from .mesh import hex_mesh
    m = hex_mesh(nx, ny, nz, lo=(1.0, 0.0, 0.0), hi=(5.0, 1.0, 1.0), jitter=0.1, seed=20261204,
                 ref_face=False)
    u_in = 0.3 * (1.4 * 287.05 * 300.0) ** 0.5
    case = channel_case(u_in, dt=(4.0 / nx) / u_in)
    case.urf_mom, case.urf_p, case.urf_pu = 0.7, 0.3, 0.3
    return m, case, u_in, inlet_profile(m, "xmin", u_in, amplitude=0.02)


def inlet_profile(mesh, patch: str, u_in: float, amplitude: float = 0.02, seed: int = 20261206):
    """Per-face fixedValue-U values ([3][n_boundary_faces], for
    GpuSolver.set_boundary_velocity) with a seeded sinusoidal perturbation
    on one inlet patch: u = u_in (1 + amplitude sin(2 pi (y/Ly + a)) sin(2 pi
    (z/Lz + b))) along +x, a and b seeded phases; every other boundary face
    gets (u_in, 0, 0) (ignored unless it is fixedValue-U)."""
    nif = mesh.n_internal_faces
    nb = mesh.n_faces - nif
    u = np.zeros((3, nb))
    u[0] = u_in
    b = mesh.patch(patch)
    f0, nf = int(mesh.patch_start[b]), int(mesh.patch_nfaces[b])
    ptr = np.asarray(mesh.face_ptr[f0:f0 + nf + 1], dtype=np.int64)
    pts = mesh.points[np.asarray(mesh.face_pts[ptr[0]:ptr[-1]])]
    cen = np.add.reduceat(pts, ptr[:-1] - ptr[0], axis=0) / np.diff(ptr)[:, None]
    lo, hi = mesh.points.min(axis=0), mesh.points.max(axis=0)
    ph = np.random.default_rng(seed).random(2)
    y = (cen[:, 1] - lo[1]) / (hi[1] - lo[1])
    z = (cen[:, 2] - lo[2]) / (hi[2] - lo[2])
    u[0, f0 - nif:f0 - nif + nf] = u_in * (1.0 + amplitude * np.sin(2 * np.pi * (y + ph[0]))
                                           * np.sin(2 * np.pi * (z + ph[1])))
    return u

