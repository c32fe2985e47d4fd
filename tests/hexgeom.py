"""Geometry of the structured hex Q_p meshes, written independently of the
oracle and the library (test helper only): physical positions of the DOF
nodes under the trilinear map, and exact integrals of linear fields over the
unit cube.  A linear field a.x + c is in every Q_p space on a trilinear mesh
(a.x(xi) is trilinear in the reference coordinates), which is what the patch
tests below rest on."""
import numpy as np


def dof_coords(nx, ny, nz, p, vertices):
    """Physical coordinates (N, 3) of the lexicographic DOFs
    dof = gx + Nx (gy + Ny gz), Nd = n p + 1: each node is the trilinear
    interpolation of its element's 8 vertices at the equispaced reference node
    xi = -1 + 2 i / p (on shared faces every element gives the same point).
    The trilinear interpolation is done as three linear ones, along x on the
    vertex rows, then y, then z one layer of nodes at a time."""
    V = np.asarray(vertices, np.float64).reshape(nz + 1, ny + 1, nx + 1, 3)
    Nx, Ny, Nz = nx * p + 1, ny * p + 1, nz * p + 1

    def lin(n, N):
        g = np.arange(N)
        e = np.minimum(g // p, n - 1)
        return e, ((g - e * p) / p)[:, None]  # element and local coordinate in [0, 1]

    ex, tx = lin(nx, Nx)
    ey, ty = lin(ny, Ny)
    ez, tz = lin(nz, Nz)
    Vx = (1.0 - tx) * V[:, :, ex] + tx * V[:, :, ex + 1]          # (nz+1, ny+1, Nx, 3)
    Vxy = (1.0 - ty[:, :, None]) * Vx[:, ey] + ty[:, :, None] * Vx[:, ey + 1]  # (nz+1, Ny, Nx, 3)
    out = np.empty((Nz, Ny, Nx, 3))
    for gz in range(Nz):
        out[gz] = (1.0 - tz[gz, 0]) * Vxy[ez[gz]] + tz[gz, 0] * Vxy[ez[gz] + 1]
    return out.reshape(-1, 3)


def boundary_mask(nx, ny, nz, p):
    """True on the DOFs of the boundary of the unit cube."""
    Nx, Ny, Nz = nx * p + 1, ny * p + 1, nz * p + 1
    gz, gy, gx = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    m = (gx == 0) | (gx == Nx - 1) | (gy == 0) | (gy == Ny - 1) | (gz == 0) | (gz == Nz - 1)
    return m.ravel()


def int_linear_product(a, c, b, d):
    """Exact integral over [0,1]^3 of (a.x + c)(b.x + d):
    int x_i x_j = 1/3 (i = j), 1/4 (i != j); int x_i = 1/2."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    quad = sum(a[i] * b[j] * (1 / 3 if i == j else 1 / 4) for i in range(3) for j in range(3))
    return quad + 0.5 * (c * b.sum() + d * a.sum()) + c * d
