"""The oracle on polyhedral cells (two jittered hexahedra merged into one 10-face cell):
each cell of the unjittered mesh closes and the merged cells hold the volume of the two they
replace, and on the jittered mesh the least-squares gradient stays exact for linear fields,
as it is on any mesh."""
import numpy as np

import prism_amd as P
from polymesh import merged_box


def test_merged_cells_geometry_and_gradient(ref):
    data = merged_box(P, 6, 5, 4, jitter=0.2, seed=3)
    assert np.diff(data.cell_face_off).max() == 10
    bcs = {s: {"T": ("fixed", 0.0)} for s in data.patch_names}
    m = P.Mesh(ref, data, bcs, [("T", "scalar", "leastSquares")])
    g = m.geometry()
    # without jitter every face is planar, so each cell closes, and the merged cells hold
    # two unit volumes
    flat = merged_box(P, 6, 5, 4)
    mf = P.Mesh(ref, flat, {s: {"T": ("fixed", 0.0)} for s in flat.patch_names}, [("T", "scalar", "leastSquares")])
    gf = mf.geometry()
    acc = np.zeros((mf.n_cells, 3))
    np.add.at(acc, gf["owner"], gf["face_area"])
    inner = gf["neighbor"] >= 0
    np.add.at(acc, gf["neighbor"][inner], -gf["face_area"][inner])
    assert np.abs(acc).max() < 1e-15
    nf = np.diff(flat.cell_face_off)
    vol = gf["cell_volume"] * 120
    assert (nf == 10).sum() == 20
    assert np.abs(vol - np.where(nf == 10, 2.0, 1.0)).max() < 1e-12
    c, f = g["cell_center"], g["face_center"]
    lin = lambda p: 0.7 + 1.3 * p[:, 0] - 2.1 * p[:, 1] + 0.4 * p[:, 2]
    T = m.scalar("T", values=lin(c))
    T.set_face_values(lin(f))
    assert np.abs(T.grad() - np.array([1.3, -2.1, 0.4])).max() < 1e-12
