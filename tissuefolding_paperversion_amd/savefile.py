"""Save files of a running simulation in the reference's layout, and continuing from them.

``Simulation::saveStep`` (Simulation.cpp:9265-9280) appends one frame per save to a directory:

- ``Save_Frame`` (text, ``writeSaveFileStepHeader`` / ``writeNodes`` / ``writeElements`` /
  ``writeSaveFileStepFooter``, Simulation.cpp:9282-9355): a header line with the time, the node
  count, one line per node (x, y, z; placement, tissue type, circumference flag), the element
  count, one line per element (shape type, ablation flag, node ids, reference positions), and a
  footer line;
- ``Save_Growth`` (binary, ``writeGrowth``, Simulation.cpp:9433-9452): Fg of every element,
  row-major, as native doubles;
- ``Save_GrowthRate`` (binary, same function): the three per-step growth factors exp(r dt) of
  every element (``ShapeBase::setGrowthRate``, ShapeBase.cpp:3141-3146).

Continuing from a save (``readNodeDataToContinueFromSave``, ``updateGrowthFromSave``,
Simulation.cpp:2027-2052, 2410-2435) restores positions from the text frame (10 significant
digits) and Fg bit-exactly from the binary file.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .mesh import TissueMesh

PRISM_SHAPE_TYPE = 1   # ShapeBase::setShapeType("Prism"), ShapeBase.cpp:42-43
_HEAD_FILL = "==================================================="
_FOOT_FILL = "============================================"


def _g(x, width, prec):
    """C++ stream output with precision(prec), width(width) and the default float format."""
    return f"{x:{width}.{prec}g}"


@dataclass
class Frame:
    time: float
    positions: np.ndarray          # (N, 3)
    placement: np.ndarray          # (N,)
    tissue_type: np.ndarray        # (N,)
    at_circumference: np.ndarray   # (N,)
    conn: np.ndarray               # (E, 6)
    ablated: np.ndarray            # (E,)
    ref_positions: np.ndarray      # (E, 18)
    fg: np.ndarray | None = None   # (E, 3, 3)
    growth_rate: np.ndarray | None = None  # (E, 3) per-step growth factors


class FrameWriter:
    """Appends frames to Save_Frame, Save_Growth and Save_GrowthRate in ``directory``."""

    def __init__(self, directory: str, mesh: TissueMesh):
        os.makedirs(directory, exist_ok=True)
        self.mesh = mesh
        self.frame = open(os.path.join(directory, "Save_Frame"), "w")
        self.growth = open(os.path.join(directory, "Save_Growth"), "wb")
        self.rate = open(os.path.join(directory, "Save_GrowthRate"), "wb")

    def write(self, time_sec: float, positions, fg, growth_factor):
        m = self.mesh
        x = np.asarray(positions, float).reshape(-1, 3)
        if x.shape[0] != m.n_nodes:
            raise ValueError("positions do not match the mesh")
        out = [f"=============== TIME: {_g(time_sec, 10, 6)}{_HEAD_FILL}\n", f"{m.n_nodes}\n"]
        for i in range(m.n_nodes):
            out.append(_g(x[i, 0], 20, 10) + _g(x[i, 1], 20, 10) + _g(x[i, 2], 20, 10) +
                       f"{int(m.node_placement[i]):4d}{int(m.node_type[i]):4d}{int(bool(m.at_circumference[i])):4d}\n")
        out.append(f"{m.n_elements}\n")
        for e in range(m.n_elements):
            line = f"{PRISM_SHAPE_TYPE:4d}{0:4d}" + "".join(f"{int(n):9d}" for n in m.conn[e])
            line += "".join(_g(v, 15, 4) for v in m.xref[e])
            out.append(line + "\n")
        out.append(f"=============== END OF TIME: {_g(time_sec, 10, 6)}{_FOOT_FILL}\n")
        self.frame.write("".join(out))
        self.frame.flush()
        np.ascontiguousarray(np.asarray(fg, float).reshape(-1, 9)).tofile(self.growth)
        np.ascontiguousarray(np.asarray(growth_factor, float).reshape(-1, 3)).tofile(self.rate)
        self.growth.flush()
        self.rate.flush()

    def close(self):
        for f in (self.frame, self.growth, self.rate):
            f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_frames(directory: str) -> list[Frame]:
    """Every frame of a save directory (binary growth data attached when present)."""
    frames = []
    with open(os.path.join(directory, "Save_Frame")) as f:
        tok = f.read().split("\n")
    k = 0
    while k < len(tok):
        line = tok[k].strip()
        k += 1
        if not line.startswith("=============== TIME:"):
            continue
        t = float(line[len("=============== TIME:"):].split("=")[0])
        n = int(tok[k]); k += 1
        nodes = [tok[k + i].split() for i in range(n)]
        k += n
        pos = np.array([[float(v) for v in r[:3]] for r in nodes])
        plc = np.array([int(r[3]) for r in nodes], np.int8)
        typ = np.array([int(r[4]) for r in nodes], np.int8)
        circ = np.array([int(r[5]) != 0 for r in nodes])
        ne = int(tok[k]); k += 1
        els = [tok[k + i].split() for i in range(ne)]
        k += ne
        if any(int(r[0]) != PRISM_SHAPE_TYPE for r in els):
            raise ValueError("only prism elements are supported")
        abl = np.array([int(r[1]) for r in els], np.int8)
        conn = np.array([[int(v) for v in r[2:8]] for r in els], np.int32)
        ref = np.array([[float(v) for v in r[8:26]] for r in els])
        if not tok[k].strip().startswith("=============== END OF TIME:"):
            raise ValueError("malformed Save_Frame: missing footer")
        k += 1
        frames.append(Frame(t, pos, plc, typ, circ, conn, abl, ref))
    gpath, rpath = os.path.join(directory, "Save_Growth"), os.path.join(directory, "Save_GrowthRate")
    if frames and os.path.exists(gpath):
        E = frames[0].conn.shape[0]
        g = np.fromfile(gpath, dtype=np.float64).reshape(-1, E, 3, 3)
        r = np.fromfile(rpath, dtype=np.float64).reshape(-1, E, 3) if os.path.exists(rpath) else None
        for i, fr in enumerate(frames[:g.shape[0]]):
            fr.fg = g[i]
            fr.growth_rate = None if r is None else r[i]
    return frames


def continue_from_save(solver, frame: Frame, mesh: TissueMesh):
    """Restores a solver's state from a saved frame (Simulation.cpp:2027-2052, 2410-2435): the
    node and element data must match the mesh, as the reference checks."""
    if frame.positions.shape[0] != mesh.n_nodes or frame.conn.shape[0] != mesh.n_elements:
        raise ValueError("save file and model are not consistent: node or element count differs")
    if (np.any(frame.placement != mesh.node_placement) or np.any(frame.at_circumference != mesh.at_circumference)
            or np.any(frame.conn != mesh.conn)):
        raise ValueError("save file and model are not consistent: node or element properties differ")
    solver.positions = frame.positions.ravel()
    if frame.fg is not None:
        solver.set_fg(frame.fg)
