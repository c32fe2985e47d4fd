"""On-disk formats on either side of the solve path (host code).

- Matrix Market coordinate files (io/src/matrix_market.rs:54-78):
  read_matrix_market -> (row_ptr u64, col_idx u32, values f64) in CsrMatrix
  layout (duplicates summed, columns sorted; symmetric files expanded), and
  write_matrix_market of a CsrMatrix (device, read back) or of the three
  arrays, as "coordinate real general" with 1-based indices and 17
  significant digits, so a write / read round trip is exact.
- VtkWriter (io/src/vtk.rs:60-240): ASCII VTK UnstructuredGrid (.vtu) of a
  simplex or hex mesh with point and cell data, the same XML layout as the
  reference writer.
- write_vti: the DOF grid of a structured hex operator as VTK ImageData, one
  point per DOF, so hex solution vectors can be viewed without a cell list.
"""
import numpy as np

__all__ = ["read_matrix_market", "write_matrix_market", "DataArray", "VtkWriter", "write_vti",
           "VTK_CELL_TYPES"]


def _coo_to_csr(nrows, rows, cols, vals):
    """Sort by (row, col), keeping input order among duplicates, and sum them."""
    order = np.lexsort((np.arange(len(rows)), cols, rows))
    r, c, v = rows[order], cols[order], vals[order]
    if len(r):
        head = np.ones(len(r), bool)
        head[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
        starts = np.flatnonzero(head)
        v = np.add.reduceat(v, starts)  # sums each run in order
        r, c = r[starts], c[starts]
    rp = np.zeros(nrows + 1, np.uint64)
    np.add.at(rp, r.astype(np.int64) + 1, 1)
    rp = np.cumsum(rp).astype(np.uint64)
    return rp, c.astype(np.uint32), np.ascontiguousarray(v, np.float64)


def read_matrix_market(path):
    """read_matrix_market (io/src/matrix_market.rs:54-60): a coordinate file
    (real / integer / pattern; general / symmetric) -> (row_ptr, col_idx,
    values, (nrows, ncols))."""
    with open(path) as f:
        header = f.readline().split()
        if len(header) < 5 or header[0].lower() != "%%matrixmarket" or header[1].lower() != "matrix":
            raise ValueError(f"{path}: not a Matrix Market matrix file")
        fmt, field, sym = header[2].lower(), header[3].lower(), header[4].lower()
        if fmt != "coordinate":
            raise ValueError(f"{path}: only coordinate format is supported")
        if field not in ("real", "integer", "pattern", "double"):
            raise ValueError(f"{path}: unsupported field {field}")
        if sym not in ("general", "symmetric"):
            raise ValueError(f"{path}: unsupported symmetry {sym}")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        nrows, ncols, nnz = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2) if nnz else np.zeros((0, 3))
    if data.shape[0] != nnz:
        raise ValueError(f"{path}: expected {nnz} entries, found {data.shape[0]}")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else data[:, 2].astype(np.float64)
    if nnz and (rows.min() < 0 or cols.min() < 0 or rows.max() >= nrows or cols.max() >= ncols):
        raise ValueError(f"{path}: index out of range")
    if sym == "symmetric":
        off = rows != cols
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, vals[off]]))
    rp, ci, v = _coo_to_csr(nrows, rows, cols, vals)
    return rp, ci, v, (nrows, ncols)


def write_matrix_market(path, a, col_idx=None, values=None, ncols=None):
    """write_matrix_market (io/src/matrix_market.rs:75-78) of a femgpu
    CsrMatrix, or of (row_ptr, col_idx, values)."""
    if col_idx is None:
        rp, ci = a.pattern()
        v = a.values().astype(np.float64)
        ncols = a.ncols
    else:
        rp, ci, v = np.asarray(a, np.uint64), np.asarray(col_idx, np.uint32), np.asarray(values, np.float64)
    nrows = len(rp) - 1
    ncols = nrows if ncols is None else ncols
    rows = np.repeat(np.arange(nrows), np.diff(rp.astype(np.int64)))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{nrows} {ncols} {len(v)}\n")
        for r, c, x in zip(rows, ci, v):
            f.write(f"{int(r) + 1} {int(c) + 1} {float(x)!r}\n")


# VTK cell type ids: Tri3, Quad4, Tet4, Hex8, Tri6, Tet10
VTK_CELL_TYPES = {"tri3": 5, "quad4": 9, "tet4": 10, "hex8": 12, "tri6": 22, "tet10": 24}


class DataArray:
    """DataArray (io/src/vtk.rs:60-83): named point or cell values with a component count."""

    def __init__(self, name, values, n_components=1):
        self.name, self.n_components = str(name), int(n_components)
        self.values = np.ascontiguousarray(values, np.float64).ravel()
        if self.values.size % self.n_components:
            raise ValueError("values are not a whole number of tuples")

    @classmethod
    def scalars(cls, name, values):
        return cls(name, values, 1)

    @classmethod
    def vectors(cls, name, n_components, values):
        return cls(name, values, n_components)


class VtkWriter:
    """VtkWriter (io/src/vtk.rs:85-240): ASCII UnstructuredGrid of a mesh with
    coords f64[n_nodes * dim] (dim 2 or 3) and conn u32[n_elem * nodes per
    element] of one cell type."""

    def __init__(self, coords, conn, dim, cell_type):
        self.dim = int(dim)
        if self.dim not in (2, 3):
            raise ValueError("VtkWriter: dim must be 2 or 3")
        self.coords = np.ascontiguousarray(coords, np.float64).ravel()
        self.cell = VTK_CELL_TYPES[cell_type] if isinstance(cell_type, str) else int(cell_type)
        npe = {5: 3, 9: 4, 10: 4, 12: 8, 22: 6, 24: 10}[self.cell]
        self.conn = np.ascontiguousarray(conn, np.int64).ravel()
        self.npe = npe
        self.n_nodes = self.coords.size // self.dim
        self.n_elems = self.conn.size // npe
        self.point_data, self.cell_data = [], []

    def add_point_data(self, arr):
        if arr.values.size != self.n_nodes * arr.n_components:
            raise ValueError(f"point data {arr.name!r}: expected {self.n_nodes} tuples")
        self.point_data.append(arr)
        return self

    def add_cell_data(self, arr):
        if arr.values.size != self.n_elems * arr.n_components:
            raise ValueError(f"cell data {arr.name!r}: expected {self.n_elems} tuples")
        self.cell_data.append(arr)
        return self

    def to_string(self):
        out = ['<?xml version="1.0"?>',
               '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
               "  <UnstructuredGrid>",
               f'    <Piece NumberOfPoints="{self.n_nodes}" NumberOfCells="{self.n_elems}">',
               "      <Points>",
               '        <DataArray type="Float64" NumberOfComponents="3" format="ascii">']
        xyz = self.coords.reshape(-1, self.dim)
        for p in xyz:
            out.append("          " + " ".join(repr(float(t)) for t in p) + (" 0" if self.dim == 2 else ""))
        out += ["        </DataArray>", "      </Points>", "      <Cells>",
                '        <DataArray type="Int32" Name="connectivity" format="ascii">']
        for row in self.conn.reshape(-1, self.npe):
            out.append("          " + " ".join(str(int(t)) for t in row))
        out += ["        </DataArray>", '        <DataArray type="Int32" Name="offsets" format="ascii">',
                "          " + " ".join(str((i + 1) * self.npe) for i in range(self.n_elems)),
                "        </DataArray>", '        <DataArray type="UInt8" Name="types" format="ascii">',
                "          " + " ".join([str(self.cell)] * self.n_elems), "        </DataArray>", "      </Cells>"]
        for tag, arrs in (("PointData", self.point_data), ("CellData", self.cell_data)):
            if not arrs:
                continue
            out.append(f"      <{tag}>")
            for a in arrs:
                out.append(f'        <DataArray type="Float64" Name="{a.name}" '
                           f'NumberOfComponents="{a.n_components}" format="ascii">')
                for t in a.values.reshape(-1, a.n_components):
                    out.append("          " + " ".join(f"{float(x):.10e}" for x in t))
                out.append("        </DataArray>")
            out.append(f"      </{tag}>")
        out += ["    </Piece>", "  </UnstructuredGrid>", "</VTKFile>", ""]
        return "\n".join(out)

    def write_file(self, path):
        with open(path, "w") as f:
            f.write(self.to_string())


def write_vti(path, shape, point_data, spacing=None, origin=(0.0, 0.0, 0.0)):
    """VTK ImageData with one point per DOF of a structured grid of shape
    (Nx, Ny, Nz) (DOF index gx + Nx (gy + Ny gz)); point_data: {name: values}.
    spacing defaults to 1 / (N - 1) per axis (the unit cube)."""
    nx, ny, nz = (int(s) for s in shape)
    if spacing is None:
        spacing = tuple(1.0 / max(s - 1, 1) for s in (nx, ny, nz))
    spacing = tuple(float(t) for t in spacing)
    origin = tuple(float(t) for t in origin)
    out = ['<?xml version="1.0"?>',
           '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">',
           f'  <ImageData WholeExtent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}" '
           f'Origin="{origin[0]!r} {origin[1]!r} {origin[2]!r}" '
           f'Spacing="{spacing[0]!r} {spacing[1]!r} {spacing[2]!r}">',
           f'    <Piece Extent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}">', "      <PointData>"]
    for name, vals in point_data.items():
        v = np.ascontiguousarray(vals, np.float64).ravel()
        if v.size != nx * ny * nz:
            raise ValueError(f"point data {name!r}: expected {nx * ny * nz} values")
        out.append(f'        <DataArray type="Float64" Name="{name}" format="ascii">')
        out.append("          " + " ".join(f"{float(x):.10e}" for x in v))
        out.append("        </DataArray>")
    out += ["      </PointData>", "    </Piece>", "  </ImageData>", "</VTKFile>", ""]
    with open(path, "w") as f:
        f.write("\n".join(out))
