"""GPU parity: HIP kernels (through the C ABI) against the FP64 oracle.

Tolerances:
* FP64 kernels follow the reference operation order with contraction off,
  so interpolation and positions are compared bit for bit (``assert_array_equal``)
  and the FTLE (device ``log``/``sqrt``) to 1e-13 relative.
* FP32 kernels: FTLE max abs error <= 1e-3 of the FP64 field range.
"""

import ctypes

import numpy as np
import pytest

from oracle import numpy_oracle as O
from oracle import native as ONat
from tests import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import _native

    _native.load()
    print("[gpu] libpyftle_hip loaded", flush=True)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------- ginterp --

@pytest.mark.parametrize("shape", [(37, 23), (2, 2), (128, 64)])
def test_interp2d_vec_bit_exact(shape, rng):
    from pyftle_amd import _native as N

    v = rng.standard_normal(shape)
    pts = rng.uniform(-2, max(shape) + 2, (5000, 2))
    pts[:6] = [[shape[0] - 1, 0], [0, shape[1] - 1], [shape[0] - 1.5, shape[1] - 1.5],
               [np.nan, 1], [1e12, 3], [-0.0, -0.0]]
    out = torch.empty(len(pts), dtype=torch.float64, device="cuda")
    dv, dp = dev(v), dev(pts)
    N.call("pyftle_interp2d_vec", P(dv), shape[0], shape[1], P(dp), len(pts), P(out), stream())
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got, O.interp2d_vec(v, pts))
    ref = ONat.reference_ginterp()
    if ref is not None:
        exp = np.empty(len(pts))
        ref.interp2d_vec_inplace(v, pts, exp)
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("shape", [(9, 11, 13), (2, 2, 2)])
def test_interp3d_vec_bit_exact(shape, rng):
    from pyftle_amd import _native as N

    v = rng.standard_normal(shape)
    pts = rng.uniform(-1, max(shape) + 1, (5000, 3))
    out = torch.empty(len(pts), dtype=torch.float64, device="cuda")
    dv, dp = dev(v), dev(pts)
    N.call("pyftle_interp3d_vec", P(dv), *shape, P(dp), len(pts), P(out), stream())
    np.testing.assert_array_equal(out.cpu().numpy(), O.interp3d_vec(v, pts))


def test_interp_golden_vectors():
    """Vectors made by the compiled reference GridInterp.cpp (tests/golden)."""
    import os

    from pyftle_amd import _native as N

    path = os.path.join(os.path.dirname(__file__), "golden", "ginterp_vectors.npz")
    g = np.load(path)
    out = torch.empty(len(g["p2"]), dtype=torch.float64, device="cuda")
    dv, dp = dev(g["v2"]), dev(g["p2"])
    N.call("pyftle_interp2d_vec", P(dv), *g["v2"].shape, P(dp), len(g["p2"]), P(out), stream())
    np.testing.assert_array_equal(out.cpu().numpy(), g["o2"])
    out = torch.empty(len(g["p3"]), dtype=torch.float64, device="cuda")
    dv, dp = dev(g["v3"]), dev(g["p3"])
    N.call("pyftle_interp3d_vec", P(dv), *g["v3"].shape, P(dp), len(g["p3"]), P(out), stream())
    np.testing.assert_array_equal(out.cpu().numpy(), g["o3"])


# ------------------------------------------------------- field sampling --

def test_grid_interpolator_2d_matches_oracle(rng):
    from pyftle_amd.interpolate import HighPerfInterpolator

    vels, pts = H.double_gyre_window(nx=50, ny=26, n_snap=1)
    ref = O.GridField((50, 26), vels[0], pts)
    q = rng.uniform([-0.1, -0.1], [2.1, 1.1], (4000, 2))
    gi = HighPerfInterpolator((50, 26), precision="float64")
    gi.update(vels[0], pts)
    np.testing.assert_array_equal(gi.interpolate(q), ref.interpolate(q))
    g32 = HighPerfInterpolator((50, 26), precision="float32")
    g32.update(vels[0], pts)
    np.testing.assert_allclose(g32.interpolate(q), ref.interpolate(q), atol=2e-6)


def test_grid_interpolator_3d_matches_oracle(rng):
    from pyftle_amd.interpolate import HighPerfInterpolator

    vels, pts = H.abc_window(n=12)
    ref = O.GridField((12, 12, 12), vels[0], pts)
    q = rng.uniform(-0.5, 6.5, (4000, 3))
    gi = HighPerfInterpolator((12, 12, 12), precision="float64")
    gi.update(vels[0], pts)
    np.testing.assert_array_equal(gi.interpolate(q), ref.interpolate(q))


# ------------------------------------------------------- FTLE tensors --

def test_ftle_tensor_kats(jacobians_2x2, jacobians_3x3):
    """tests/test_ftle.py:17-92 known answers, FP64 on device."""
    from pyftle_amd import ftle as F

    cg2 = F.compute_cauchy_green_2x2(jacobians_2x2)
    np.testing.assert_allclose(cg2, np.einsum("...ij,...kj->...ik", jacobians_2x2, jacobians_2x2), rtol=1e-12)
    cg3 = F.compute_cauchy_green_3x3(jacobians_3x3)
    np.testing.assert_allclose(cg3, jacobians_3x3 @ np.transpose(jacobians_3x3, (0, 2, 1)), rtol=1e-12)
    np.testing.assert_allclose(F.max_eigenvalue_2x2(cg2), np.max(np.linalg.eigvals(cg2), axis=1).real, rtol=1e-10)
    np.testing.assert_allclose(F.max_eigenvalue_3x3(cg3), np.max(np.linalg.eigvals(cg3), axis=1).real, rtol=1e-10)
    np.testing.assert_allclose(F.max_eigenvalue_2x2(cg2), [4.0, 4.0], rtol=1e-15)
    np.testing.assert_allclose(F.compute_ftle_2x2(jacobians_2x2, 1.0), [np.log(2.0)] * 2, rtol=1e-14)
    exp3 = 0.5 * np.log(np.max(np.linalg.eigvals(cg3).real, axis=1)) / 2.0
    np.testing.assert_allclose(F.compute_ftle_3x3(jacobians_3x3, 2.0), exp3, rtol=1e-8, atol=1e-15)
    eye = np.repeat(np.eye(2)[None], 3, axis=0)
    np.testing.assert_allclose(F.compute_ftle_2x2(eye, 1.0), np.zeros(3), atol=1e-12)
    assert F.compute_ftle_2x2(2 * np.eye(2)[None], 1.0)[0] > F.compute_ftle_2x2(np.eye(2)[None], 1.0)[0]


def test_ftle_random_symmetric_3x3(rng):
    from pyftle_amd import ftle as F

    J = rng.standard_normal((2000, 3, 3)) * rng.uniform(0.1, 30, (2000, 1, 1))
    C = J @ np.transpose(J, (0, 2, 1))
    np.testing.assert_allclose(F.max_eigenvalue_3x3(C), np.linalg.eigvalsh(C)[:, -1], rtol=1e-11)
    J2 = rng.standard_normal((2000, 2, 2)) * rng.uniform(0.1, 30, (2000, 1, 1))
    C2 = O.cauchy_green_2x2(J2)
    np.testing.assert_allclose(F.max_eigenvalue_2x2(C2), np.linalg.eigvalsh(C2)[:, -1], rtol=1e-12)


def test_flow_map_jacobian_linear_map(rng):
    """A linear flow map x -> M x gives J = M exactly (cauchy_green.py:70-74,140-151)."""
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.cauchy_green import compute_flow_map_jacobian_2x2, compute_flow_map_jacobian_3x3

    for dim, fn in ((2, compute_flow_map_jacobian_2x2), (3, compute_flow_map_jacobian_3x3)):
        M = rng.standard_normal((dim, dim))
        c = rng.uniform(-1, 1, (50, dim))
        p = NeighboringParticles(O.make_neighbor_positions(c, 0.125))
        p.positions[...] = p.positions @ M.T
        J = fn(p)
        np.testing.assert_allclose(J, np.broadcast_to(M, J.shape), rtol=1e-12, atol=1e-13)
        np.testing.assert_array_equal(J, O.flow_map_jacobian(p.positions, p.initial_positions))


def test_pinned_staged_pack_equals_device_pack():
    """Host snapshots go through the pinned double-buffered stager: slot contents equal
    a pack from device tensors byte for byte, also when the host array is reused."""
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=9)
    geom = GridGeometry.from_reference((64, 32), pts)
    fa = DeviceField(geom, dtype="float32", layout=N.LAYOUT_PAIR, ring_slots=9, n_snapshots=9, t0=0.0, dt_snap=0.1)
    fb = DeviceField(geom, dtype="float32", layout=N.LAYOUT_PAIR, ring_slots=9, n_snapshots=9, t0=0.0, dt_snap=0.1)
    scratch = np.empty_like(vels[0])
    for k, v in enumerate(vels):
        scratch[...] = v  # the same host array for every snapshot: staging copies it at once
        fa.pack(k, scratch)
        fb.pack(k, torch.as_tensor(v, device="cuda"))
    torch.cuda.synchronize()
    a = fa._buf[fa._offset:fa._offset + fa.nbytes]
    b = fb._buf[fb._offset:fb._offset + fb.nbytes]
    assert torch.equal(a, b)
