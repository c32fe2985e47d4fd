"""Host mirror of the reference's operator-backend selectors for this path.

- OperatorBackend (assembly/src/backend.rs:10-36): "native" / "reed" plus the
  new "hip" (this library), parsed case-insensitively.
- FemCeed (assembly/src/reed/context.rs:41-300): libCEED-style mass and
  Poisson applies on 2-D triangle meshes, selected by a backend resource
  string.  This build runs them on the GPU only: the resource "/gpu/hip"
  (and sub-resources "/gpu/hip/...") resolves to CeedBackend.HIP; every other
  resource is refused with a ReedError, since there is no CPU or WGPU
  implementation behind it here.

apply_poisson_2d computes K = |det J| w J^-1 J^-T (the stiffness matrix of
PADiffusionOperator with kappa = 1).  The reference's FemCeed::apply_poisson_2d
uses J^-T J^-1 (context.rs:282-284), which differs when J is not normal; this
mirror keeps the correct form and takes parity from PADiffusionOperator.
"""
from dataclasses import dataclass
from enum import Enum

import numpy as np

from . import DIFFUSION, MASS, PADiffusionOperator, PAMassOperator, default_context


class OperatorBackend(Enum):
    NATIVE = "native"
    REED = "reed"
    HIP = "hip"

    @classmethod
    def parse(cls, s):
        """OperatorBackend::parse: case-insensitive name -> backend, None if unknown."""
        try:
            return cls(str(s).lower())
        except ValueError:
            return None

    def as_str(self):
        return self.value


class CeedBackend(Enum):
    REED_CPU = "reed-cpu"          # reference backends, listed for completeness
    REED_GPU_WGPU = "reed-gpu-wgpu"
    HIP = "hip"                    # this library


class FemCeedError(RuntimeError):
    pass


class ReedError(FemCeedError):
    """FemCeedError::Reed: backend initialisation or basis errors."""


class SizeMismatch(FemCeedError):
    """FemCeedError::SizeMismatch { expected, got }."""

    def __init__(self, expected, got):
        super().__init__(f"input size mismatch: expected {expected}, got {got}")
        self.expected, self.got = expected, got


@dataclass
class BackendSelectionReport:
    requested_resource: str
    effective_resource: str
    note: str


@dataclass
class SimplexMesh2:
    """2-D triangle mesh: coords f64[n_nodes, 2]; elem_nodes u32[n_elem, 3] (P1)
    or [n_elem, 6] (P2: vertices, then the nodes of edges (0,1), (1,2), (0,2),
    the TriP2 local order of element/src/lagrange/tri.rs)."""
    coords: np.ndarray
    elem_nodes: np.ndarray

    def n_nodes(self):
        return len(self.coords)

    def n_elems(self):
        return len(self.elem_nodes)


# number of points of reed's triangle rules -> the tri_rule order with the same
# exactness (q = 6 is reed's degree-4 rule; the 7-point degree-5 rule
# integrates the same P2 mass integrand exactly)
_Q_TO_ORDER = {1: 1, 3: 2, 6: 4, 7: 5, 12: 6}


class FemCeed:
    """FemCeed on the HIP backend.  Operators are built per (mesh, poly, q)
    and cached, so repeated applies reuse the device quadrature data."""

    def __init__(self, backend=CeedBackend.HIP, ctx=None):
        if backend != CeedBackend.HIP:
            raise ReedError(f"backend {backend.value} is not available in this build (GPU/HIP only)")
        self._backend = backend
        self._resource = "/gpu/hip"
        self._note = "FemCeed backend: HIP partial assembly on the GPU."
        self._ctx = ctx or default_context()
        self._cache = {}

    @classmethod
    def new(cls):
        return cls()

    @classmethod
    def with_backend(cls, backend):
        return cls(backend)

    @classmethod
    def from_backend_resource(cls, resource):
        """-> (FemCeed, BackendSelectionReport); ReedError for any resource
        other than /gpu/hip[/...]."""
        r = str(resource)
        if r == "/gpu/hip" or r.startswith("/gpu/hip/"):
            ceed = cls()
            ceed._resource = r
            return ceed, BackendSelectionReport(r, r, ceed._note)
        raise ReedError(f"backend resource {r!r} is not available in this build (use /gpu/hip)")

    def backend(self):
        return self._backend

    def effective_resource(self):
        return self._resource

    def backend_note(self):
        return self._note

    def _op(self, kind, mesh, poly, q):
        if poly not in (1, 2):
            raise ReedError(f"unsupported polynomial order {poly} (1 or 2)")
        if q not in _Q_TO_ORDER:
            raise ReedError(f"unsupported number of quadrature points {q}")
        en = np.ascontiguousarray(mesh.elem_nodes, np.uint32)
        npe = 3 if poly == 1 else 6
        if en.ndim != 2 or en.shape[1] < npe:
            raise ReedError(f"P{poly} needs {npe} nodes per element, mesh has {en.shape}")
        key = (kind, id(mesh), poly, q)
        if key not in self._cache:
            coords = np.ascontiguousarray(mesh.coords, np.float64).ravel()
            dofs = np.ascontiguousarray(en[:, :npe]).ravel()  # flat, as the C ABI takes them
            verts = np.ascontiguousarray(en[:, :3]).ravel()
            cls = PAMassOperator if kind == MASS else PADiffusionOperator
            self._cache[key] = (cls(2, poly, coords, verts, dofs, mesh.n_nodes(), 1.0, _Q_TO_ORDER[q],
                                    ctx=self._ctx), mesh)
        return self._cache[key][0]

    def _apply(self, kind, mesh, poly, q, inp):
        inp = np.ascontiguousarray(inp, np.float64)
        if inp.shape != (mesh.n_nodes(),):
            raise SizeMismatch(mesh.n_nodes(), inp.size)
        return self._op(kind, mesh, poly, q).matvec(inp)

    def apply_mass_2d(self, mesh, poly, q, inp):
        """output = M input, M_ij = sum_e sum_q w |det J| phi_i phi_j (context.rs:126-205)."""
        return self._apply(MASS, mesh, poly, q, inp)

    def apply_poisson_2d(self, mesh, poly, q, inp):
        """output = K input with K_ij = int grad phi_i . grad phi_j (see the module note)."""
        return self._apply(DIFFUSION, mesh, poly, q, inp)
