"""VTK output of a simulation (src/IO/visualization_beams.jl): write_VTK's two PolyData files per
output time — the beams as polylines of `intermediate_points` points each (positions with the
bending offset, Stress, Strain, Displacement, Velocity; eb_output_points computes them on the
device) and the nodes with the beams as two-point lines (Displacement, In Contact, Contact Force,
Contact Distance) — and the ParaView collection simulation.pvd that lists them by time.

Files are VTK XML PolyData written in ASCII. The node contact fields are written as the
reference writes them: its solver never sets nodes.incontact / contactforce / contactdistance
(src/beams/def_node_beam.jl:45-47), so they are zero.
"""
from __future__ import annotations

import os
from typing import Dict, List, Sequence, Tuple

import numpy as np


def _array(name, a, ncomp=1, kind="Float64"):
    a = np.asarray(a)
    flat = a.reshape(-1)
    if kind == "Float64":
        body = " ".join(repr(float(v)) for v in flat)
    else:
        body = " ".join(str(int(v)) for v in flat)
    nm = f' Name="{name}"' if name else ""
    return f'<DataArray type="{kind}"{nm} NumberOfComponents="{ncomp}" format="ascii">{body}</DataArray>'


def write_vtp(path: str, points: np.ndarray, lines: Sequence[Sequence[int]], point_data: Dict[str, np.ndarray],
              field_data: Dict[str, Tuple[str, float]]) -> None:
    """One VTK XML PolyData file: `points` (n × 3), polylines given by 0-based point indices,
    point arrays (n or n × 3) and field data {name: (VTK type, value)}."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    conn = np.concatenate([np.asarray(l, dtype=np.int64) for l in lines]) if len(lines) else np.zeros(0, np.int64)
    offsets = np.cumsum([len(l) for l in lines], dtype=np.int64)
    out = ['<?xml version="1.0"?>',
           '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">', "<PolyData>",
           "<FieldData>"]
    for name, (kind, value) in field_data.items():
        out.append(_array(name, np.array([value]), 1, kind).replace("NumberOfComponents", 'NumberOfTuples="1" NumberOfComponents'))
    out.append("</FieldData>")
    out.append(f'<Piece NumberOfPoints="{len(pts)}" NumberOfVerts="0" NumberOfLines="{len(lines)}" '
               f'NumberOfStrips="0" NumberOfPolys="0">')
    out += ["<Points>", _array(None, pts, 3), "</Points>", "<PointData>"]
    for name, a in point_data.items():
        a = np.asarray(a)
        kind = "Int64" if np.issubdtype(a.dtype, np.integer) else "Float64"
        out.append(_array(name, a, 3 if a.ndim == 2 else 1, kind))
    out += ["</PointData>", "<Lines>", _array("connectivity", conn, 1, "Int64"), _array("offsets", offsets, 1, "Int64"),
            "</Lines>", "</Piece>", "</PolyData>", "</VTKFile>"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


class VTKDataBeams:
    """VTKDataBeams(conf, output_dir, intermediate_points = 10) (visualization_beams.jl:28-72)
    and write_VTK_beams (:75-128)."""

    def __init__(self, conf, output_dir: str, intermediate_points: int = 10):
        self.conf = conf
        self.output_dir = output_dir
        self.n = int(intermediate_points)
        E = len(conf.beams)
        self.lines = [list(range(b * self.n, (b + 1) * self.n)) for b in range(E)]
        self.lines_nodes = [[int(a) - 1, int(b) - 1] for a, b in conf.beams.conn]
        self.collection: List[Tuple[float, str]] = []
        # clean_folders (src/inizialization.jl:2-14): the output folder starts empty of VTK files
        os.makedirs(output_dir, exist_ok=True)
        for f in os.listdir(output_dir):
            if f.endswith((".vtp", ".pvd")):
                os.remove(os.path.join(output_dir, f))

    def write(self, write_counter: int, step: int, t: float, engine) -> None:
        pts = engine.output_points(self.n)
        fd = {"time": ("Float64", t), "step": ("Int64", step)}
        name = f"beams_{write_counter}.vtp"
        write_vtp(os.path.join(self.output_dir, name), pts["position"], self.lines,
                  {"Stress": pts["stress"], "Strain": pts["strain"], "Displacement": pts["displacement"],
                   "Velocity": pts["velocity"]}, fd)
        self.collection.append((t, name))
        N = len(self.conf.nodes)
        u = engine.get("U").reshape(N, 3)
        name = f"beams_nodes_{write_counter}.vtp"
        write_vtp(os.path.join(self.output_dir, name), self.conf.nodes.X0 + u, self.lines_nodes,
                  {"Displacement": u, "In Contact": np.zeros(N, np.int64), "Contact Force": np.zeros((N, 3)),
                   "Contact Distance": np.zeros(N)}, fd)
        self.collection.append((t, name))
        self.save()

    def save(self) -> None:
        """simulation.pvd: every file written so far, by time."""
        rows = [f'<DataSet timestep="{t!r}" part="0" file="{f}"/>' for t, f in self.collection]
        with open(os.path.join(self.output_dir, "simulation.pvd"), "w") as fh:
            fh.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">\n'
                     "<Collection>\n" + "\n".join(rows) + "\n</Collection>\n</VTKFile>\n")
