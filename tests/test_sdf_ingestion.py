"""SDF ingestion (src/IO/read_input_files.jl:154-194, src/interactions/def_interactions.jl:67-82):
a legacy ASCII VTK structured-points file read by read_vtk_sdf gives the grid, the domain and the
samples x-fastest; DiscreteSignedDistanceField.from_vtk keeps or flips the sign (`inside`); the
interpolant built from the file passes through the samples."""
import numpy as np
import pytest

import oracle as O
from endobeams_amd.model import (BeamElementSurface, Beams, BeamsConfiguration, DiscreteSignedDistanceField,
                                 InteractionProperties, NodesBeams, RigidInteraction, SimulationParams, read_vtk_sdf)


def write_legacy_vtk(path, n, spacing, origin, values, per_line=7):
    lines = ["# vtk DataFile Version 3.0", "sdf written by the test", "ASCII", "DATASET STRUCTURED_POINTS",
             f"DIMENSIONS {n[0]} {n[1]} {n[2]}", f"SPACING {spacing[0]} {spacing[1]} {spacing[2]}",
             f"ORIGIN {origin[0]} {origin[1]} {origin[2]}", f"POINT_DATA {values.size}",
             "SCALARS sdf double 1", "LOOKUP_TABLE default"]
    flat = values.reshape(-1)
    for i in range(0, flat.size, per_line):   # uneven line lengths, as VTK writers wrap them
        lines.append(" ".join(repr(float(v)) for v in flat[i:i + per_line]))
    path.write_text("\n".join(lines) + "\n")


@pytest.fixture
def sphere_file(tmp_path):
    n, h, x0 = (9, 7, 6), (0.05, 0.04, 0.06), (-0.2, -0.12, -0.15)
    axes = [x0[a] + h[a] * np.arange(n[a]) for a in range(3)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    vals = 0.1 - np.sqrt(X * X + Y * Y + Z * Z)
    p = tmp_path / "sdf.vtk"
    write_legacy_vtk(p, n, h, x0, vals)
    return p, n, h, x0, vals


def test_read_vtk_sdf_grid_and_samples(sphere_file):
    p, n, h, x0, vals = sphere_file
    nx, ny, nz, dx, dy, dz, dom, sdf = read_vtk_sdf(str(p))
    assert (nx, ny, nz) == n and (dx, dy, dz) == h
    np.testing.assert_allclose(dom, (x0[0], x0[0] + 8 * h[0], x0[1], x0[1] + 6 * h[1], x0[2], x0[2] + 5 * h[2]), rtol=1e-15)
    np.testing.assert_array_equal(sdf, vals.reshape(-1))


def test_from_vtk_sign_and_interpolation(sphere_file):
    p, n, h, x0, vals = sphere_file
    inside = DiscreteSignedDistanceField.from_vtk(str(p), inside=True)
    outside = DiscreteSignedDistanceField.from_vtk(str(p), inside=False)
    np.testing.assert_array_equal(outside.samples, -inside.samples)
    nodes = NodesBeams(np.array([[0.0, 0, 0], [1.0, 0, 0]]))
    conf = BeamsConfiguration(nodes, Beams(nodes, [[1, 2]], 1e5, 0.3, 1.0, 0.05, 0.0), None, None)
    eng = O.OracleEngine(conf, SimulationParams.hht(-0.05),
                         RigidInteraction(inside, BeamElementSurface([1]), InteractionProperties(1.0, 0.1, 1.0, 0.0)))
    for (i, j, k) in [(1, 1, 1), (4, 3, 2), (7, 5, 4)]:
        pt = np.array([x0[0] + i * h[0], x0[1] + j * h[1], x0[2] + k * h[2]])
        assert eng.sdf_eval(pt)[0] == pytest.approx(vals[k, j, i], abs=1e-12)
