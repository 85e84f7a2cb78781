"""rbh_sdf_build_noalloc (staging in the library's fixed per-GPU array, no
device allocation) against rbh_sdf_build, through the C ABI.

Two calls in a row on one context, the second with a smaller mesh and grid:
each equals rbh_sdf_build's grid bit for bit (both run the same kernel on the
same triangle corners).  A second context on the same GPU shares the array.
A call whose staging exceeds the array fails with RBH_E_CAPACITY before any
device work."""
import numpy as np
import pytest

from rainbow_amd import mesh as MESH
from rainbow_amd import native

pytestmark = pytest.mark.gpu


def grid_box(V, res, margin=0.1):
    lo = V.min(axis=0) - margin
    hi = V.max(axis=0) + margin
    return lo, hi, res, res, res


def test_noalloc_sdf_equals_allocating_build():
    rng = np.random.default_rng(7)
    big = MESH.create_noisy_icosphere(rng, 22)            # ~4.8k vertices, ~9.7k triangles
    small = MESH.create_box(1.0, 2.0, 3.0)                # 8 vertices, 12 triangles
    ctx = native.Context(0, "fp32")
    try:
        for (V, T), res in ((big, 48), (small, 20)):
            V = np.asarray(V, dtype=np.float64)
            T = np.asarray(T, dtype=np.int32)
            lo, hi, I, J, K = grid_box(V, res)
            want = ctx.sdf_build(V, T, lo, hi, I, J, K)
            got = ctx.sdf_build_noalloc(V, T, lo, hi, I, J, K)
            print("noalloc SDF: %d triangles, %d^3 nodes, max |diff| %g" % (len(T), res, np.abs(got - want).max()))
            assert got.shape == want.shape and np.isfinite(got).all()
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
            assert (got < 0).any() and (got > 0).any()
        # a second context on the same GPU uses the same array
        ctx2 = native.Context(0, "fp64")
        try:
            V, T = (np.asarray(a) for a in small)
            lo, hi, I, J, K = grid_box(np.asarray(V, dtype=np.float64), 12)
            assert np.array_equal(ctx2.sdf_build_noalloc(V, T, lo, hi, I, J, K).view(np.uint64),
                                  ctx.sdf_build(V, T, lo, hi, I, J, K).view(np.uint64))
        finally:
            ctx2.close()
    finally:
        ctx.close()


def test_noalloc_sdf_rejects_oversize_and_bad_input():
    V, T = (np.asarray(a) for a in MESH.create_box(1.0, 1.0, 1.0))
    V = V.astype(np.float64)
    ctx = native.Context(0, "fp32")
    try:
        lo, hi = V.min(axis=0) - 0.1, V.max(axis=0) + 0.1
        with pytest.raises(native.RbhError) as ei:
            ctx.sdf_build_noalloc(V, T, lo, hi, 256, 256, 256)      # 128 MiB of grid
        assert ei.value.code == native.RBH_E_CAPACITY
        with pytest.raises(native.RbhError) as ei:
            ctx.sdf_build_noalloc(V, np.array([[0, 1, 99]]), lo, hi, 8, 8, 8)
        assert ei.value.code == native.RBH_E_ARG
        # the context still works after the rejected calls
        assert np.array_equal(ctx.sdf_build_noalloc(V, T, lo, hi, 8, 8, 8), ctx.sdf_build(V, T, lo, hi, 8, 8, 8))
    finally:
        ctx.close()
