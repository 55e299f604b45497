"""The VTK writer (stand-in for the reference's Silo output): a valid legacy
unstructured grid whose hexahedra are the O-grid cells (the mean of each
cell's 8 corners is the mean of its 6 face mid-points)."""
import numpy as np

from nim_amd import host


def test_vtk_grid(tmp_path):
    case = host.reference_case()
    geo = host.build_geometry(case)
    T = np.arange(case.ncells, dtype=np.float64) * 0.5
    path = tmp_path / "fields.vtk"
    host.write_vtk(path, case, geo, {"T": T})
    lines = open(path).read().split("\n")
    assert lines[0].startswith("# vtk DataFile") and lines[3] == "DATASET UNSTRUCTURED_GRID"
    npts = int(lines[4].split()[1])
    nt, nr, nz = case.dims
    assert npts == nt * (nr + 1) * (nz + 1)
    c = lines.index("CELLS %d %d" % (case.ncells, 9 * case.ncells))
    conn = np.array([list(map(int, l.split())) for l in lines[c + 1:c + 1 + case.ncells]])
    assert np.all(conn[:, 0] == 8) and conn[:, 1:].max() == npts - 1
    xyz = geo["node_xyz"].reshape(-1, 3)
    np.testing.assert_allclose(xyz[conn[:, 1:]].mean(axis=1), host.cell_centroids(geo), atol=1e-12)
    t = lines.index("CELL_TYPES %d" % case.ncells)
    assert set(lines[t + 1:t + 1 + case.ncells]) == {"12"}
    d = lines.index("SCALARS T double 1")
    np.testing.assert_array_equal(np.array(lines[d + 2:d + 2 + case.ncells], dtype=float), T)
