"""A subset of the reference's ``modelinput`` file: what the implicit NR step needs.

The reference reads the file section by section (``ModelInputObject::readParameters``,
ModelInputObject.cpp:28-253): a line holding only ``Name:`` opens a section, and each section
reader takes ``Key: value ...`` tokens in a fixed order.  This module reads the same layout and
keeps the keys of the sections that drive the implicit step:

- ``InputMeshParameters`` (ModelInputObject.cpp:709-748): MeshInputMode 4 (a ``.mesh`` file);
- ``TimeParameters`` (:1320-1344): TimeStep, SimulationLength;
- ``SaveOptions`` (:1436-1491): SaveData, DataSaveInterval;
- ``NodeFixingOptions`` (:866-1025): surface and circumference fixes, with or without the fixing
  viscosity (``Simulation::checkForNodeFixing``, Simulation.cpp:3695-3800, ``fixX`` :3416-3430);
- ``PysicalProperties`` (:1346-1434): E by placement, Poisson's ratio, internal viscosity;
- ``ExternalViscositySetup`` (:750-813): disc-proper apical and basal external viscosity;
- ``GrowthOptions`` (:268-441, :443-707): growth functions of type 1 (uniform, optionally at an
  angle), 2 (ring) and 3 (grid from a rate file), applied to the columnar layer, with the grid
  options (pinned on the initial mesh, pinning update times, step or linear interpolation); their
  per-step increments are evaluated by :mod:`growth`.

Anything else the file switches on, and the step cannot reproduce, is reported by
:meth:`ModelInput.unsupported` and refused by :func:`build_tissue`: explicit ECM, symmetry
planes, peripodial membrane, growth of the peripodial membrane or the ECM, node binding by
ellipses, adhesion and collapse, noise on the physical properties.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import configs
from . import growth as G
from .mesh import APICAL, BASAL, TissueMesh, read_mesh


def _sections(path: str) -> dict[str, list[tuple[str, list[str]]]]:
    """{section: [(key, values), ...]} in file order."""
    out: dict[str, list[tuple[str, list[str]]]] = {}
    cur = None
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            toks = line.split()
            if not toks:
                continue
            if len(toks) == 1 and toks[0].endswith(":") and not line[0].isspace():
                cur = toks[0][:-1]
                out.setdefault(cur, [])
                continue
            if cur is None:
                continue
            key, vals = None, []
            for t in toks:
                if t.endswith(":") and not _is_number(t[:-1]):
                    if key is not None:
                        out[cur].append((key, vals))
                    key, vals = t[:-1], []
                else:
                    vals.append(t)
            if key is not None:
                out[cur].append((key, vals))
    return out


def _is_number(s: str) -> bool:
    try:
        float(s)
        return True
    except ValueError:
        return False


@dataclass
class GrowthFunction:
    kind: int
    t0: float
    t1: float
    columnar: bool
    peripodial: bool
    basal_ecm: bool
    lateral_ecm: bool
    rate: tuple[float, float, float]   # per second (DV, AP, AB)
    angle: float                       # radians
    centre: tuple[float, float] = (0.0, 0.0)   # ring
    radii: tuple[float, float] = (0.0, 0.0)    # ring: inner, outer
    filename: str = ""                          # grid: rate file
    zrange: tuple[float, float] = (0.0, 1.0)    # grid


@dataclass
class ModelInput:
    path: str
    mesh_mode: int = 4
    mesh_file: str = ""
    symmetric: tuple[bool, bool] = (False, False)
    dt: float = 1.0
    length: float = 0.0
    save_data: bool = False
    save_interval: float = 0.0
    fixing_viscosity: tuple[float, float, float] = (0.0, 0.0, 0.0)
    fixes: dict = field(default_factory=dict)        # name -> ((x, y, z), with_viscosity)
    E: tuple[float, float, float] = (1000.0, 1000.0, 1000.0)   # apical, basal, mid
    nu: float = 0.3
    noise: list = field(default_factory=list)
    eta_int: tuple[float, float, float] = (0.0, 0.0, 0.0)      # apical, basal, mid
    ext_extend: bool = False
    ext_visc: tuple[float, float] = (0.0, 0.0)                 # disc proper apical, basal
    growth: list[GrowthFunction] = field(default_factory=list)
    grid_pinned: bool = True                                   # GridGrowthsPinnedOnInitialMesh
    pin_times: tuple = ()                                      # PinningUpdateTimes (s)
    grid_interpolation: int = 1                                # 0 = step, 1 = linear
    flags: dict = field(default_factory=dict)                  # other switches, by name

    def unsupported(self) -> list[str]:
        bad = []
        if self.mesh_mode != 4:
            bad.append(f"MeshInputMode {self.mesh_mode} (only 4, a .mesh file)")
        if any(self.symmetric):
            bad.append("symmetry planes (symmetricInX/Y)")
        if any(float(n) != 0 for n in self.noise):
            bad.append("noise on the physical properties")
        for name in ("ThereIsExplicitECM", "AddPeripodialMembrane", "ThereIsAdhesion", "ThereIsNodeCollapse",
                     "bindCircumferenceXYToBasal", "bindEllipses", "AddCurvature", "AddSoftPeriphery",
                     "AddRandomForce"):
            if self.flags.get(name):
                bad.append(name)
        for k, g in enumerate(self.growth):
            if g.kind not in (1, 2, 3):
                bad.append(f"growth function {k}: type {g.kind} (types 1, 2 and 3 only)")
            elif g.peripodial or g.basal_ecm or g.lateral_ecm:
                bad.append(f"growth function {k}: applied to peripodial membrane or ECM")
            elif g.kind == 3 and g.columnar and not os.path.exists(self._grid_path(g)):
                bad.append(f"growth function {k}: grid rate file {g.filename} not found")
        return bad

    def _grid_path(self, g: GrowthFunction) -> str:
        return g.filename if os.path.isabs(g.filename) else os.path.join(os.path.dirname(self.path), g.filename)

    def needs_increments(self) -> bool:
        """True when some columnar growth function is not a diagonal rate: ring, grid, or an angle."""
        return any(g.columnar and (g.kind != 1 or g.angle != 0.0) for g in self.growth)

    def growth_functions(self) -> list:
        """The columnar growth functions as :mod:`growth` objects, in the file's order."""
        out = []
        for g in self.growth:
            if not g.columnar:
                continue
            if g.kind == 1:
                out.append(G.UniformGrowth(g.t0, g.t1, g.rate, g.angle))
            elif g.kind == 2:
                out.append(G.RingGrowth(g.t0, g.t1, g.centre, g.radii[0], g.radii[1], g.rate, g.angle))
            elif g.kind == 3:
                rates, angles = G.read_grid_file(self._grid_path(g))
                out.append(G.GridGrowth(g.t0, g.t1, rates, angles, g.zrange[0], g.zrange[1]))
        return out

    def rates_at(self, mesh: TissueMesh, time_sec: float) -> np.ndarray:
        """Per-element growth rates (per second) of the uniform functions active at time_sec:
        the step's increments multiply, so diagonal rates add."""
        r = np.zeros((mesh.n_elements, 3))
        for g in self.growth:
            if g.kind == 1 and g.columnar and g.t0 <= time_sec < g.t1:
                r += np.asarray(g.rate)
        return r


def _f32(v) -> float:
    return float(np.float32(float(v)))


def _f32_rate(v) -> float:
    """A rate read into a float and scaled to per second in place (ModelInputObject.cpp:412-420)."""
    return float(np.float32(float(np.float32(float(v))) * (1.0 / 3600.0)))


def _b(v):
    return bool(int(float(v)))


def read(path: str) -> ModelInput:
    S = _sections(path)
    mi = ModelInput(path=path)
    get = lambda sec, key: next((v for k, v in S.get(sec, []) if k.startswith(key)), None)  # noqa: E731
    v = get("InputMeshParameters", "MeshInputMode")
    if v:
        mi.mesh_mode = int(v[0])
    v = get("InputMeshParameters", "MeshFile")
    if v:
        mi.mesh_file = v[0]
    mi.symmetric = (_b((get("InputMeshParameters", "symmetricInX") or ["0"])[0]),
                    _b((get("InputMeshParameters", "symmetricInY") or ["0"])[0]))
    mi.dt = float((get("TimeParameters", "TimeStep") or ["1"])[0])
    mi.length = float((get("TimeParameters", "SimulationLength") or ["0"])[0])
    mi.save_data = _b((get("SaveOptions", "SaveData") or ["0"])[0])
    mi.save_interval = float((get("SaveOptions", "DataSaveInterval") or ["0"])[0])
    fx = S.get("NodeFixingOptions", [])
    for k, vals in fx:
        if k.startswith("FixingViscosity"):
            mi.fixing_viscosity = tuple(float(x) for x in vals[:3])
    pending = None
    for k, vals in fx:
        name = k.split("(")[0]
        if name.endswith("Fix") and len(vals) >= 3:
            pending = name
            mi.fixes[name] = (tuple(_b(x) for x in vals[:3]), False)
        elif name.startswith("Fix") and pending is not None and vals:
            mi.fixes[pending] = (mi.fixes[pending][0], _b(vals[0]))
            pending = None
    pp = S.get("PysicalProperties", [])
    d = {k: vals for k, vals in pp}
    mi.E = (float(d["YoungsModulusApical"][0]), float(d["YoungsModulusBasal"][0]), float(d["YoungsModulusMid"][0]))
    mi.nu = min(max(float(d["PoissonsRatio"][0]), 0.0), 0.5)   # clamped to [0, 0.5] (Prism.cpp:491-516)
    mi.noise = [vals[0] for k, vals in pp if k.startswith("Noise")]
    mi.eta_int = (float(d.get("ApicalViscosity", ["0"])[0]), float(d.get("BasalViscosity", ["0"])[0]),
                  float(d.get("MidLineViscosity", ["0"])[0]))
    ev = {k: vals for k, vals in S.get("ExternalViscositySetup", [])}
    mi.ext_extend = _b(ev.get("ExtendToWholeTissue", ["0"])[0])
    mi.ext_visc = (float(ev.get("DiscProperApicalExternalViscosity", ["0"])[0]),
                   float(ev.get("DiscProperBasalExternalViscosity", ["0"])[0]))
    for sec in ("PeripodialMembraneParameters", "ExplicitECMOptions", "AdhesionOptions", "NodeCollapseOptions",
                "NodeBindingOptions", "Manipulations"):
        for k, vals in S.get(sec, []):
            name = k.split("(")[0]
            if vals and name in ("ThereIsExplicitECM", "AddPeripodialMembrane", "ThereIsAdhesion",
                                 "ThereIsNodeCollapse", "bindCircumferenceXYToBasal", "bindEllipses",
                                 "AddCurvature", "AddSoftPeriphery", "AddRandomForce"):
                mi.flags[name] = _b(vals[0])
    # growth functions: the keys repeat per function, in the reader's order
    g = S.get("GrowthOptions", [])
    opts = {k.split("(")[0]: vals for k, vals in g if not k.startswith("GrowthFunctionType")}
    if "GridGrowthsPinnedOnInitialMesh" in opts:
        mi.grid_pinned = _b(opts["GridGrowthsPinnedOnInitialMesh"][0])
    if "PinningUpdateTimes" in opts and opts["PinningUpdateTimes"]:
        n = int(opts["PinningUpdateTimes"][0])
        mi.pin_times = tuple(float(x) for x in opts["PinningUpdateTimes"][1:1 + n])
    if "GridGrowthsInterpolationType" in opts:
        mi.grid_interpolation = int(opts["GridGrowthsInterpolationType"][0])
    i = 0
    while i < len(g):
        k, vals = g[i]
        if k.startswith("GrowthFunctionType"):
            kind = int(vals[0])
            fields = {}
            j = i + 1
            while j < len(g) and not g[j][0].startswith("GrowthFunctionType"):
                fields[g[j][0].split("(")[0]] = g[j][1]
                j += 1
            common = (_f32(fields["InitialTime"][0]), _f32(fields["FinalTime"][0]),
                      _b(fields["ApplyToColumnarLayer"][0]), _b(fields["ApplyToPeripodialMembrane"][0]),
                      _b(fields["ApplyToBasalECM"][0]), _b(fields["ApplyToLateralECM"][0]))
            if kind == 1:
                r = [_f32_rate(x) for x in fields["MaxValue"][:3]]
                ang = float(np.float32(float(np.float32(float(fields["Angle"][0]))) * (math.pi / 180.0)))
                mi.growth.append(GrowthFunction(1, *common, (r[0], r[1], r[2]), ang))
            elif kind == 2:
                r = [_f32_rate(x) for x in fields["MaxValue"][:3]]
                mi.growth.append(GrowthFunction(2, *common, (r[0], r[1], r[2]),
                                                float(fields["Angle"][0]) * (math.pi / 180.0),
                                                centre=(_f32(fields["Centre"][0]), _f32(fields["Centre"][1])),
                                                radii=(_f32(fields["InnerRadius"][0]), _f32(fields["OuterRadius"][0]))))
            elif kind == 3:
                z = fields.get("zRange", ["0", "1"])
                mi.growth.append(GrowthFunction(3, *common, (0.0, 0.0, 0.0), 0.0, filename=fields["Filename"][0],
                                                zrange=(float(z[0]), float(z[1]))))
            else:
                mi.growth.append(GrowthFunction(kind, 0.0, 0.0, False, False, False, False, (0.0, 0.0, 0.0), 0.0))
            i = j
        else:
            i += 1
    return mi


def build_tissue(mi: ModelInput, mesh: TissueMesh | None = None, name: str | None = None) -> configs.Tissue:
    """The tissue a model input describes (mesh from ``mesh`` or the file's MeshFile)."""
    bad = mi.unsupported()
    if bad:
        raise NotImplementedError("model input uses features outside the implicit NR step: " + "; ".join(bad))
    if mesh is None:
        path = mi.mesh_file if os.path.isabs(mi.mesh_file) else os.path.join(os.path.dirname(mi.path), mi.mesh_file)
        mesh = read_mesh(path)
    E, lam, mu = configs.assign_material(mesh, mi.E[0], mi.E[1], mi.E[2], mi.nu)
    pl, _ = mesh.element_placement()
    eta = np.full(mesh.n_elements, mi.eta_int[2])
    eta[pl == BASAL] = mi.eta_int[1]
    eta[pl == APICAL] = mi.eta_int[0]
    eta_ext = configs.assign_external_viscosity(mesh, mi.ext_visc[0], mi.ext_visc[1], mi.ext_extend).reshape(-1, 3)
    fixed = np.zeros((mesh.n_nodes, 3), np.uint8)
    plc, circ = mesh.node_placement, mesh.at_circumference
    groups = {"ApicSurfaceFix": plc == APICAL, "BasalSurfaceFix": plc == BASAL, "CircumferenceFix": circ,
              "ApicCircumFix": circ & (plc == APICAL), "BasalCircumFix": circ & (plc == BASAL)}
    for nm, ((fxx, fxy, fxz), with_visc) in mi.fixes.items():
        if nm not in groups:
            if any((fxx, fxy, fxz)):
                raise NotImplementedError(f"node fixing option {nm}")
            continue
        sel = groups[nm]
        for dim, on in enumerate((fxx, fxy, fxz)):
            if not on:
                continue
            if with_visc:   # fix by a high external viscosity instead of a fixed DoF
                eta_ext[sel, dim] = mi.fixing_viscosity[dim]
            else:
                fixed[sel, dim] = 1
    rates = mi.rates_at(mesh, 0.0)
    nu = np.full(mesh.n_elements, mi.nu)
    n_steps = int(round(mi.length / mi.dt)) if mi.dt > 0 else 0
    return configs.Tissue(mesh, lam, mu, eta, eta_ext.ravel(), fixed.ravel(), rates, mesh.exposed_faces(),
                          mi.dt, n_steps, name or os.path.basename(mi.path), E, nu)


def run(mi: ModelInput, tissue: configs.Tissue, solver, n_steps: int | None = None, save_dir: str | None = None):
    """Advances ``solver`` (a NewtonRaphsonSolver on ``tissue``) step by step: growth rates of
    the functions active at the step's start time, one implicit timestep, and a saved frame every
    DataSaveInterval when ``save_dir`` is given.  Returns the per-step statistics."""
    from .savefile import FrameWriter

    steps = tissue.n_steps if n_steps is None else n_steps
    writer = FrameWriter(save_dir, tissue.mesh) if save_dir else None
    every = max(1, int(round(mi.save_interval / mi.dt))) if mi.save_interval > 0 else 0
    stats, last = [], None
    tracker = GrowthTracker(mi, tissue.mesh, solver) if mi.needs_increments() else None
    try:
        for k in range(steps):
            t = k * mi.dt
            if tracker is None:
                r = mi.rates_at(tissue.mesh, t)
                if last is None or not np.array_equal(r, last):
                    solver.set_growth_rates(r)
                    last = r
                factors = np.exp(r * mi.dt)
            else:
                tracker.before_step(t)
                factors = None
            stats.append(solver.runOneStep(mi.dt))
            if writer is not None and every and (k + 1) % every == 0:
                fg, _ = solver.get_fg()
                if factors is None:  # the diagonal of the step's increments, evaluated on the device
                    inc = solver.growth_increments()
                    factors = np.stack([inc[:, 0, 0], inc[:, 1, 1], inc[:, 2, 2]], axis=1)
                writer.write((k + 1) * mi.dt, solver.positions, fg, factors)
    finally:
        if writer is not None:
            writer.close()
    return stats


class GrowthTracker:
    """A model input's growth functions on the device (``tf_set_growth_functions``): set once from
    the initial mesh, the handle's time set before every step, and the grids' relative positions
    re-pinned at every PinningUpdateTime for grids pinned on the mesh
    (Simulation::checkForPinningPositionsUpdate, Simulation.cpp:10425-10443)."""

    def __init__(self, mi: ModelInput, mesh: TissueMesh, solver):
        self.mi = mi
        self.solver = solver
        self.functions = mi.growth_functions()
        geo = G.GrowthGeometry(mesh)
        solver.set_growth_functions(self.functions, geo.initial_z, geo.rel_src, mi.grid_pinned,
                                    mi.grid_interpolation)
        self.pins_done = [False] * len(mi.pin_times)

    def before_step(self, time_sec: float):
        self.solver.set_time(time_sec)
        if self.mi.grid_pinned:
            for k, tp in enumerate(self.mi.pin_times):
                if time_sec >= tp and not self.pins_done[k]:
                    self.solver.pin_growth_positions()
                    self.pins_done[k] = True
