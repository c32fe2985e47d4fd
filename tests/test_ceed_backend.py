"""OperatorBackend::parse and FemCeed backend-resource selection
(assembly/src/backend.rs:21-36, 90-112; assembly/src/reed/context.rs:83-100,
327-354).  Host logic only: no GPU needed."""
import pytest


def test_operator_backend_parse():
    from femgpu.ceed import OperatorBackend as B
    assert B.parse("native") is B.NATIVE
    assert B.parse("REED") is B.REED
    assert B.parse("Hip") is B.HIP
    assert B.parse("other") is None
    for b in B:
        assert B.parse(b.as_str()) is b


def test_femceed_unknown_or_unavailable_resources():
    from femgpu.ceed import FemCeed, ReedError, CeedBackend
    for r in ("/solver/unknown", "/cpu/self", "/cpu/self/ref", "/gpu/wgpu", "/gpu/hipx"):
        with pytest.raises(ReedError):
            FemCeed.from_backend_resource(r)
    with pytest.raises(ReedError):
        FemCeed.with_backend(CeedBackend.REED_CPU)
