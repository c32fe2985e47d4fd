"""Matrix Market and VTK output (io/src/matrix_market.rs, io/src/vtk.rs) on
the CPU: round trips, the reference's own known-answer tests, and SciPy's
reader / writer as an independent check."""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fem-rs_amd"))
from femgpu import io as fio  # noqa: E402


def _lap1d(n):
    rows, cols, vals = [], [], []
    for i in range(n):
        rows.append(i); cols.append(i); vals.append(2.0)
        if i > 0: rows.append(i); cols.append(i - 1); vals.append(-1.0)
        if i < n - 1: rows.append(i); cols.append(i + 1); vals.append(-1.0)
    return fio._coo_to_csr(n, np.array(rows), np.array(cols), np.array(vals))


def test_matrix_market_roundtrip(tmp_path):
    # matrix_market.rs: roundtrip_matrix_market (1-D Laplacian, n = 5)
    rp, ci, v = _lap1d(5)
    p = tmp_path / "a.mtx"
    fio.write_matrix_market(p, rp, ci, v)
    rp2, ci2, v2, shape = fio.read_matrix_market(p)
    assert shape == (5, 5)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(v, v2)
    # SciPy reads the same matrix
    A = scipy.io.mmread(str(p)).tocsr()
    B = sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=(5, 5))
    assert (A != B).nnz == 0


def test_matrix_market_known_answer(tmp_path):
    # matrix_market.rs: read_mtx_from_string
    p = tmp_path / "b.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 4.0\n1 2 1.0\n2 1 1.0\n3 3 9.0\n")
    rp, ci, v, shape = fio.read_matrix_market(p)
    assert shape == (3, 3)
    assert rp.tolist() == [0, 2, 3, 4] and ci.tolist() == [0, 1, 0, 2] and v.tolist() == [4.0, 1.0, 1.0, 9.0]


def test_matrix_market_symmetric_duplicates_and_exact_values(tmp_path):
    rng = np.random.default_rng(0)
    M = sp.random(40, 40, density=0.1, random_state=1) + sp.eye(40)
    S = (M + M.T).tocoo()
    p = tmp_path / "s.mtx"
    scipy.io.mmwrite(str(p), S, symmetry="symmetric", precision=17)
    rp, ci, v, shape = fio.read_matrix_market(p)
    B = sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=shape)
    assert abs(B - S.tocsr()).max() == 0
    # duplicates are summed, values survive a write / read exactly
    q = tmp_path / "d.mtx"
    vals = [float(t) for t in rng.standard_normal(3)]
    q.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n"
                 f"2 1 {vals[0]!r}\n1 1 {vals[1]!r}\n2 1 {vals[2]!r}\n")
    rp, ci, v, _ = fio.read_matrix_market(q)
    assert ci.tolist() == [0, 0] and v[0] == vals[1] and v[1] == vals[0] + vals[2]
    r = tmp_path / "e.mtx"
    fio.write_matrix_market(r, rp, ci, v)
    assert np.array_equal(fio.read_matrix_market(r)[2], v)
    with pytest.raises(ValueError):
        (tmp_path / "bad.mtx").write_text("%%MatrixMarket matrix array real general\n1 1\n1.0\n")
        fio.read_matrix_market(tmp_path / "bad.mtx")


def test_vtu_writer(tmp_path):
    coords = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    conn = np.array([0, 1, 2, 1, 3, 2])
    w = fio.VtkWriter(coords, conn, 2, "tri3")
    w.add_point_data(fio.DataArray.scalars("u", [0.0, 1.0, 2.0, 3.5]))
    w.add_cell_data(fio.DataArray.vectors("g", 2, [1.0, 2.0, 3.0, 4.0]))
    p = tmp_path / "m.vtu"
    w.write_file(p)
    root = ET.parse(p).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == "4" and piece.get("NumberOfCells") == "2"
    arrays = {a.get("Name"): a for a in root.iter("DataArray")}
    assert arrays["connectivity"].text.split() == ["0", "1", "2", "1", "3", "2"]
    assert arrays["offsets"].text.split() == ["3", "6"] and arrays["types"].text.split() == ["5", "5"]
    assert np.allclose([float(t) for t in arrays["u"].text.split()], [0.0, 1.0, 2.0, 3.5])
    pts = np.array([float(t) for t in piece.find("Points/DataArray").text.split()]).reshape(-1, 3)
    assert np.array_equal(pts[:, :2].ravel(), coords) and not pts[:, 2].any()
    with pytest.raises(ValueError):
        w.add_point_data(fio.DataArray.scalars("bad", [1.0]))


def test_vti_writer(tmp_path):
    u = np.arange(3 * 4 * 5, dtype=float)
    p = tmp_path / "g.vti"
    fio.write_vti(p, (3, 4, 5), {"u": u})
    root = ET.parse(p).getroot()
    img = root.find("ImageData")
    assert img.get("WholeExtent") == "0 2 0 3 0 4"
    vals = np.array([float(t) for t in root.find(".//DataArray").text.split()])
    assert np.array_equal(vals, u)
