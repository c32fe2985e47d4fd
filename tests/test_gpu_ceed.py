"""FemCeed::apply_mass_2d / apply_poisson_2d on the HIP backend against the
oracle's simplex PA (itself pinned against the assembled matrix) and closed
forms (sum of M 1 = area, K 1 = 0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _p2_mesh(fr, n):
    coords, conn = fr.unit_square_tri(n)
    nv = len(coords) // 2
    nd, ed = fr.simplex_dofs(2, 2, nv, conn)
    ed = np.asarray(ed, np.uint32).reshape(-1, 6)
    c = np.zeros((nd, 2))
    c[:nv] = np.asarray(coords).reshape(-1, 2)
    for e in ed:  # edge nodes at edge midpoints (edges (0,1), (1,2), (0,2))
        for k, (a, b) in enumerate(((0, 1), (1, 2), (0, 2))):
            c[e[3 + k]] = 0.5 * (c[e[a]] + c[e[b]])
    return coords, conn, nd, ed, c


def test_femceed_resource_and_apply(fg, fr, ctx):
    from femgpu.ceed import FemCeed, CeedBackend, SimplexMesh2, SizeMismatch
    ceed, rep = FemCeed.from_backend_resource("/gpu/hip")
    assert ceed.backend() is CeedBackend.HIP
    assert ceed.effective_resource() == rep.effective_resource == "/gpu/hip"
    assert ceed.backend_note()
    coords, conn = fr.unit_square_tri(8)
    mesh = SimplexMesh2(np.asarray(coords).reshape(-1, 2), np.asarray(conn, np.uint32).reshape(-1, 3))
    x = np.random.default_rng(0).standard_normal(mesh.n_nodes())
    m = ceed.apply_mass_2d(mesh, 1, 3, x)
    ref = fr.simplex_pa_apply(2, 1, coords, conn, conn, 1, 1.0, 2, x)
    assert _rel(m, ref) < 1e-12
    assert abs(ceed.apply_mass_2d(mesh, 1, 3, np.ones(mesh.n_nodes())).sum() - 1.0) < 1e-13  # area
    k = ceed.apply_poisson_2d(mesh, 1, 3, x)
    assert _rel(k, fr.simplex_pa_apply(2, 1, coords, conn, conn, 2, 1.0, 2, x)) < 1e-12
    assert np.abs(ceed.apply_poisson_2d(mesh, 1, 3, np.ones(mesh.n_nodes()))).max() < 1e-12
    with pytest.raises(SizeMismatch):
        ceed.apply_mass_2d(mesh, 1, 3, x[:-1])


def test_femceed_p2(fg, fr, ctx):
    from femgpu.ceed import FemCeed, SimplexMesh2
    coords, conn, nd, ed, c = _p2_mesh(fr, 6)
    ceed = FemCeed()
    mesh = SimplexMesh2(c, ed)
    x = np.random.default_rng(1).standard_normal(nd)
    ref = fr.simplex_pa_apply(2, 2, coords, conn, ed.ravel(), 1, 1.0, 4, x)
    assert _rel(ceed.apply_mass_2d(mesh, 2, 6, x), ref) < 1e-12
    assert abs(ceed.apply_mass_2d(mesh, 2, 6, np.ones(nd)).sum() - 1.0) < 1e-13
    refk = fr.simplex_pa_apply(2, 2, coords, conn, ed.ravel(), 2, 1.0, 4, x)
    assert _rel(ceed.apply_poisson_2d(mesh, 2, 6, x), refk) < 1e-12
