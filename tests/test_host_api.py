"""Host-side logic of the drop-in API (no GPU): factories, layouts, geometry, readers."""

import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import flows
from oracle import numpy_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_integrator_factory_names_and_errors():
    """integrate.py:350-387 (tests/test_integrate.py:40-63)."""
    from pyftle_amd.integrate import (AdamsBashforth2Integrator, EulerIntegrator,
                                      RungeKutta4Integrator, create_integrator)
    from pyftle_amd.interpolate import HighPerfInterpolator

    interp = HighPerfInterpolator((4, 4))
    assert isinstance(create_integrator("ab2", interp), AdamsBashforth2Integrator)
    assert isinstance(create_integrator("EULER", interp), EulerIntegrator)
    assert isinstance(create_integrator("rK4", interp), RungeKutta4Integrator)
    with pytest.raises(ValueError, match="Invalid integrator name 'invalid'.*"):
        create_integrator("invalid", interp)
    with pytest.raises(ValueError, match="Invalid integrator name ''.*"):
        create_integrator("", interp)


def test_interpolator_factory():
    from pyftle_amd.interpolate import (AnalyticalInterpolator, ExecutableNotFoundError,
                                        HighPerfInterpolator, create_interpolator)

    assert isinstance(create_interpolator("GRID", (8, 8)), HighPerfInterpolator)
    assert isinstance(create_interpolator("analytical", velocity_fn=lambda t, x: x), AnalyticalInterpolator)
    with pytest.raises(ExecutableNotFoundError):
        create_interpolator("analytical")
    with pytest.raises(ValueError, match="Invalid interpolation type"):
        create_interpolator("spline")
    # with a grid shape the SciPy names are RegularGridInterpolator strategies (interpolate.py:895)
    from pyftle_amd.interpolate import GridInterpolator

    for name, mode in (("cubic", 3), ("Linear", 1), ("nearest", 2)):
        g = create_interpolator(name, (8, 8))
        assert isinstance(g, GridInterpolator) and g.mode == mode
    assert GridInterpolator((8, 8), "slinear").mode == 1
    with pytest.raises(ValueError, match="Invalid method"):
        GridInterpolator((8, 8), "spline")
    with pytest.raises(NotImplementedError):
        GridInterpolator((8, 8), "quintic")
    with pytest.raises(ValueError, match="at least  4 points"):
        GridInterpolator((8, 3), "cubic")
    # without one they are the scattered-data strategies (interpolate.py:114-294)
    from pyftle_amd.interpolate import CubicInterpolator, LinearInterpolator, NearestNeighborInterpolator

    assert isinstance(create_interpolator("cubic"), CubicInterpolator)
    assert isinstance(create_interpolator("LINEAR"), LinearInterpolator)
    assert isinstance(create_interpolator("nearest"), NearestNeighborInterpolator)
    with pytest.raises(ValueError, match="Unsupported grid dimensionality"):
        HighPerfInterpolator((8,))


def test_rgi_geometry_reproduces_linspace_nodes():
    """GridGeometry.from_rgi: node k = k*spacing + origin and the last node = max reproduce
    numpy.linspace bit for bit (the device evaluates nodes this way), C-order strides, and
    the reference's point-count check (interpolate.py:611-617)."""
    from pyftle_amd.field import GridGeometry

    rng = np.random.default_rng(3)
    for n in (4, 17, 256, 1000):
        lo, hi = rng.uniform(-5, 0), rng.uniform(0.1, 7)
        ax = np.linspace(lo, hi, n)
        other = np.linspace(0.0, 1.0, 5)
        X, Y = np.meshgrid(ax, other, indexing="ij")
        g = GridGeometry.from_rgi((n, 5), np.stack([X.ravel(), Y.ravel()]))
        k = np.arange(n - 1, dtype=np.float64)
        np.testing.assert_array_equal(k * g.spacing[0] + g.origin[0], ax[:-1])
        assert g.upper[0] == ax[-1]
        assert g.src_strides == (5 * n, 5, 1, 0)
    g3 = GridGeometry.from_rgi((4, 5, 6), np.zeros((3, 120)) + np.arange(120.0))
    assert g3.n == (4, 5, 6) and g3.src_strides == (120, 30, 6, 1)
    with pytest.raises(ValueError, match="implies 20 points, but got 21"):
        GridGeometry.from_rgi((4, 5), np.zeros((2, 21)))


def test_analytical_interpolator_time_is_frozen_at_zero():
    """interpolate.py:788,812: velocity_fn(0.0, points) whatever the step."""
    from pyftle_amd.interpolate import AnalyticalInterpolator

    seen = []
    ai = AnalyticalInterpolator(lambda t, x: (seen.append(t), x)[1])
    ai.update(None)
    ai.interpolate(np.zeros((3, 2)))
    assert seen == [0.0]


def test_neighboring_particles_reference_fixture():
    """tests/test_neighboring_particles.py:8-162 values."""
    from pyftle_amd import NeighboringParticles

    p = NeighboringParticles(np.array([[0.0, 0.0], [1.0, 0.0], [0.5, 1.0], [0.5, 0.0]]))
    assert len(p) == 1 and p.num_neighbors == 4
    np.testing.assert_array_equal(p.initial_delta_right_left, [[1.0, 0.0]])
    np.testing.assert_array_equal(p.initial_delta_top_bottom, [[0.0, 1.0]])
    np.testing.assert_array_equal(p.initial_centroid, [[0.5, 0.25]])
    p.positions[1] = [2.0, 0.0]
    np.testing.assert_array_equal(p.delta_right_left, [[2.0, 0.0]])
    np.testing.assert_array_equal(p.initial_delta_right_left, [[1.0, 0.0]])
    p3 = NeighboringParticles(np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0],
                                        [0.5, 0, 1], [0.5, 0, -1.0]]))
    np.testing.assert_array_equal(p3.initial_delta_front_back, [[0.0, 0.0, 2.0]])


def test_neighbor_positions_match_oracle(rng):
    from pyftle_amd.particles import neighbor_positions

    c = rng.uniform(0, 1, (7, 3))
    np.testing.assert_array_equal(neighbor_positions(c, 1e-3), O.make_neighbor_positions(c, 1e-3))


def test_grid_geometry_matches_reference_layout():
    from pyftle_amd.field import GridGeometry

    x, y = np.linspace(0.0, 2.0, 9), np.linspace(0.0, 1.0, 5)
    vel, pts = flows.sample_2d(flows.double_gyre, 0.3, x, y)
    g = GridGeometry.from_reference((9, 5), pts)
    gf = O.GridField((9, 5), vel, pts)
    assert g.n == (9, 5) and list(g.origin) == gf.mins and list(g.spacing) == gf.deltas
    # element (c, ix, iy) through the strides equals the oracle's component array
    s0, s1, s2, _ = g.src_strides
    for ix, iy in ((3, 2), (8, 4), (0, 1)):
        assert vel.ravel()[1 * s0 + ix * s1 + iy * s2] == gf.v[1][ix, iy]
    a = np.linspace(0, 1, 4)
    vel3, pts3 = flows.sample_3d(lambda t, q: q * [1, 2, 3], 0.0, a, a, a)
    g3 = GridGeometry.from_reference((4, 4, 4), pts3)
    gf3 = O.GridField((4, 4, 4), vel3, pts3)
    s0, s1, s2, s3 = g3.src_strides
    for i, j, k in ((1, 2, 3), (3, 0, 1)):
        for c in range(3):
            assert vel3.ravel()[c * s0 + i * s1 + j * s2 + k * s3] == gf3.v[c][i, j, k]


def test_file_readers_roundtrip(tmp_path):
    """file_readers.py:35-182 keys and block concatenation order L, R, T, B(, F, Bk)."""
    from scipy.io import savemat

    from pyftle_amd.file_readers import read_coordinate, read_seed_particles_coordinates, read_velocity

    savemat(tmp_path / "v.mat", {"velocity_x": np.arange(6.0).reshape(2, 3), "velocity_y": np.ones((2, 3))})
    v = read_velocity(str(tmp_path / "v.mat"))
    assert v.shape == (2, 6) and v[0, 4] == 4.0
    savemat(tmp_path / "c.mat", {"coordinate_x": np.zeros(4), "coordinate_y": np.ones(4), "coordinate_z": 2 * np.ones(4)})
    assert read_coordinate(str(tmp_path / "c.mat")).shape == (3, 4)
    blocks = {k: np.full((2, 2), i, float) for i, k in enumerate(["left", "right", "top", "bottom"])}
    savemat(tmp_path / "p.mat", blocks)
    p = read_seed_particles_coordinates(str(tmp_path / "p.mat"))
    np.testing.assert_array_equal(p.positions[:, 0], [0, 0, 1, 1, 2, 2, 3, 3])
    savemat(tmp_path / "bad.mat", {"left": np.zeros((1, 2))})
    with pytest.raises(ValueError):
        read_seed_particles_coordinates(str(tmp_path / "bad.mat"))


def test_file_batch_source_reuses_coordinates(tmp_path):
    """data_source.py:140-211: coordinates read on step 0 only when one file is reused."""
    from scipy.io import savemat

    from pyftle_amd.data_source import FileBatchSource

    savemat(tmp_path / "c.mat", {"coordinate_x": np.zeros(4), "coordinate_y": np.ones(4)})
    for k in range(3):
        savemat(tmp_path / f"v{k}.mat", {"velocity_x": np.zeros(4), "velocity_y": np.ones(4) * k})
    src = FileBatchSource([str(tmp_path / f"v{k}.mat") for k in range(3)], [str(tmp_path / "c.mat")] * 3,
                          "p.mat", 0.1, 0.2)
    assert src.num_steps == 3 and src.timestep == 0.1 and src.id == "v0"
    assert src.get_data_for_step(0)[1] is not None
    assert src.get_data_for_step(2)[1] is None
    # read-ahead hands out exactly what read_velocity returns, in any order, repeatedly
    from pyftle_amd.file_readers import read_velocity

    for k in (1, 0, 2, 2, 1):
        np.testing.assert_array_equal(src.get_data_for_step(k)[0],
                                      read_velocity(str(tmp_path / f"v{k}.mat")))


def test_prefetcher_order_and_errors(tmp_path):
    from scipy.io import savemat

    from pyftle_amd.file_readers import Prefetcher, read_velocity

    paths = []
    for k in range(7):
        savemat(tmp_path / f"s{k}.mat", {"velocity_x": np.full((3, 2), k), "velocity_y": -np.arange(6.0)})
        paths.append(str(tmp_path / f"s{k}.mat"))
    got = list(Prefetcher(paths, read_velocity, depth=3))
    assert len(got) == 7
    for k, g in enumerate(got):
        np.testing.assert_array_equal(g, read_velocity(paths[k]))
    pf = Prefetcher(paths + [str(tmp_path / "missing.mat")], read_velocity, depth=8)
    np.testing.assert_array_equal(pf.get(0)[0], np.zeros(6))
    with pytest.raises(FileNotFoundError):
        pf.get(7)
    with pytest.raises(IndexError):
        pf.get(8)


def test_analytical_solver_windows():
    """api.py:105,125-155: p = int(T/|dt|)+1 snapshots, windows start at k*dt."""
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.api import AnalyticalSolver

    parts = NeighboringParticles(O.make_neighbor_positions(np.array([[1.0, 0.5]]), 1e-3))
    s = AnalyticalSolver(flows.double_gyre, parts, 0.1, 0.3, 3, "rk4")
    assert s.num_snapshots == int(0.3 / 0.1) + 1 == 3
    b = s._create_batches()
    assert [x.num_steps for x in b] == [3, 3, 3]
    np.testing.assert_allclose(b[2].times, [0.2, 0.3, 0.4])
    assert b[1].id == f"{0.1:06f}"
    assert b[0].get_particles() is not parts


def test_launch_config_to_opts():
    from pyftle_amd import _native as N
    from pyftle_amd import window as W

    o = W.make_opts("ab2", "linear", 7, -0.05, first=3, t_start=1.5,
                    cfg=W.LaunchConfig(groups_per_block=32, lane_order=1, xcd_remap=0))
    assert (o.method, o.time_mode, o.n_steps, o.first_snapshot) == (N.AB2, N.TIME_LINEAR, 7, 3)
    assert (o.h, o.t_start, o.groups_per_block, o.lane_order, o.xcd_remap) == (-0.05, 1.5, 32, 1, 0)


def test_tuning_shape_classes():
    from pyftle_amd import tuning

    assert tuning.shape_class(3, 128**3, "steady") == "d3-steady-f32-s21"
    assert tuning.shape_class(3, 128**3, "steady", records="cell") == "d3-steady-f32-s21-cell"
    cands = tuning.candidates(3)
    assert all(c.groups_per_block * 6 <= 512 for c in cands)
    assert tuning.best_config("no-such-class").groups_per_block == 64


def test_library_missing_raises(tmp_path, monkeypatch):
    from pyftle_amd import _native

    with pytest.raises(RuntimeError, match="not built"):
        _native.load(str(tmp_path / "nope.so"))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_parse_tuple_rules():
    from pyftle_amd.experiment import parse_tuple

    assert parse_tuple("10,20,30") == (10, 20, 30)
    assert parse_tuple("128, 128 , 256") == (128, 128, 256)
    assert parse_tuple("10,20,1") == (10, 20)
    for bad in ("10", "1,2,3,4", "0,5", "-1,4"):
        with pytest.raises(ValueError):
            parse_tuple(bad)


def test_load_config_keys_and_required(tmp_path):
    from pyftle_amd.experiment import load_config

    cfg_file = tmp_path / "config.yaml"
    cfg_file.write_text(
        'experiment_name: "exp"\nlist_velocity_files: "v.txt"\nlist_coordinate_files: "c.txt"\n'
        'list_particle_files: "p.txt"\nsnapshot_timestep: -0.1\nflow_map_period: 5\nintegrator: rk4\n'
        'interpolator: grid\nnum_processes: 8\noutput_format: "mat"\nflow_grid_shape: 128, 128 , 256\n'
        'particles_grid_shape: 64, 32, 1\nnew_key: 3\n')
    c = load_config(str(cfg_file))
    assert c.flow_grid_shape == (128, 128, 256) and c.particles_grid_shape == (64, 32)
    assert c.snapshot_timestep == -0.1 and c.flow_map_period == 5.0 and c.num_processes == 8
    assert c.extra == {"new_key": 3}
    assert load_config(str(cfg_file), integrator="euler").integrator == "euler"
    (tmp_path / "short.yaml").write_text("experiment_name: x\n")
    with pytest.raises(ValueError, match="missing required"):
        load_config(str(tmp_path / "short.yaml"))


def _write_lists(tmp_path, n_snap, n_coord=1, n_part=1):
    vs = [str(tmp_path / f"v{k:03d}.mat") for k in range(n_snap)]
    cs = [str(tmp_path / f"c{k}.mat") for k in range(n_coord)]
    ps = [str(tmp_path / f"p{k}.mat") for k in range(n_part)]
    for name, items in (("v.txt", vs), ("c.txt", cs), ("p.txt", ps)):
        (tmp_path / name).write_text("\n".join(items) + "\n")
    return vs, cs, ps


def test_file_windows_follow_the_reference_construction(tmp_path):
    from pyftle_amd.experiment import FileFTLEManager, load_config

    vs, cs, ps = _write_lists(tmp_path, 7, n_coord=2, n_part=3)
    base = dict(experiment_name="e", list_velocity_files=str(tmp_path / "v.txt"),
                list_coordinate_files=str(tmp_path / "c.txt"), list_particle_files=str(tmp_path / "p.txt"),
                flow_map_period=0.75, integrator="rk4", interpolator="grid", flow_grid_shape="4,4")
    m = FileFTLEManager(load_config(None, snapshot_timestep=0.25, **base), write=False)
    bs = m.create_batches()
    assert m.window_length() == 4 and len(bs) == 4
    assert [b.snapshot_files for b in bs] == [vs[i:i + 4] for i in range(4)]
    coords = (cs * 10)[:7 + 4 - 1]
    assert [b.coordinate_files for b in bs] == [coords[i:i + 4] for i in range(4)]
    assert [b.particle_file for b in bs] == [ps[i % 3] for i in range(4)]
    back = FileFTLEManager(load_config(None, snapshot_timestep=-0.25, **base), write=False)
    assert back.create_batches()[0].snapshot_files == vs[::-1][:4]


def test_mat_writer_and_factory(tmp_path):
    from scipy.io import loadmat

    from pyftle_amd.file_writers import MatWriter, create_writer

    w = create_writer("mat", str(tmp_path / "out"), (3, 2))
    assert isinstance(w, MatWriter)
    cen = np.arange(12.0).reshape(6, 2)
    w.write("ftle_a", np.arange(6.0), cen)
    d = loadmat(str(tmp_path / "out" / "ftle_a.mat"))
    # 2D grids are stored as (nx, ny, 1), as the reference's writer does
    np.testing.assert_array_equal(d["ftle"], np.arange(6.0).reshape(3, 2, 1))
    np.testing.assert_array_equal(d["x"], cen[:, 0].reshape(3, 2, 1))
    create_writer("mat", str(tmp_path / "flat")).write("f", np.arange(4.0), np.ones((4, 3)))
    assert set(loadmat(str(tmp_path / "flat" / "f.mat"))) >= {"ftle", "x", "y", "z"}
    with pytest.raises(ValueError, match="Unsupported output format"):
        create_writer("csv", str(tmp_path))


def _read_vtk(path):
    """Arrays of a VTK XML file with inline binary (UInt64 header) data."""
    import base64
    import xml.etree.ElementTree as ET

    root = ET.parse(path).getroot()
    assert root.get("header_type") == "UInt64" and root.get("byte_order") == "LittleEndian"
    dt = {"Float64": "<f8", "Float32": "<f4", "Int64": "<i8", "Int32": "<i4"}
    out = {}
    for el in root.iter("DataArray"):
        raw = base64.b64decode(el.text)
        n = int(np.frombuffer(raw[:8], "<u8")[0])
        a = np.frombuffer(raw[8:8 + n], dt[el.get("type")])
        k = int(el.get("NumberOfComponents", "1"))
        out[el.get("Name") or "points"] = a.reshape(-1, k) if k > 1 else a
    piece = next(root.iter("Piece"))
    return out, piece.attrib


def test_vtk_writer_arrays_follow_the_reference_layout(tmp_path):
    """file_writers.py:223-264 with pyvista's StructuredGrid(x, y, z) point order
    (point i + nx*(j + ny*k) = (x, y, z)[i, j, k]) and PolyData vertex cells."""
    from pyftle_amd.file_writers import VTKWriter, create_writer

    # 2D (nx, ny) = (3, 2): Fortran reshape, z = 0, ftle in seed order
    w = create_writer("vtk", str(tmp_path / "v2"), (3, 2))
    assert isinstance(w, VTKWriter)
    cen = np.stack([np.arange(6.0), 10 + np.arange(6.0)], 1)
    w.write("f2", np.arange(6.0) * 0.5, cen)
    a, piece = _read_vtk(str(tmp_path / "v2" / "f2.vts"))
    assert piece["Extent"] == "0 2 0 1 0 0"
    np.testing.assert_array_equal(a["points"], np.column_stack([cen, np.zeros(6)]))
    np.testing.assert_array_equal(a["ftle"], np.arange(6.0) * 0.5)
    # 3D (nx, ny, nz) = (2, 3, 2): C reshape of the centroids, ftle with x/y swapped
    nx, ny, nz = 2, 3, 2
    cen3 = np.random.default_rng(0).random((nx * ny * nz, 3))
    f3 = np.arange(12.0)
    w3 = VTKWriter(str(tmp_path / "v3"), (nx, ny, nz))
    w3.write("f3", f3, cen3)
    a, piece = _read_vtk(str(tmp_path / "v3" / "f3.vts"))
    assert piece["Extent"] == "0 1 0 2 0 1"
    c = cen3.reshape(nx, ny, nz, 3)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                p = i + nx * (j + ny * k)
                np.testing.assert_array_equal(a["points"][p], c[i, j, k])
    # the reference swaps the first two axes of the FTLE block before the Fortran ravel
    m = f3.reshape(nx, ny, nz)
    want = [m[b, a_, c_] for c_ in range(nz) for b in range(nx) for a_ in range(ny)]
    np.testing.assert_array_equal(a["ftle"], want)
    # point cloud: .vtp with one vertex per point
    VTKWriter(str(tmp_path / "pc")).write("p", np.arange(4.0), np.arange(8.0).reshape(4, 2))
    a, piece = _read_vtk(str(tmp_path / "pc" / "p.vtp"))
    assert piece["NumberOfPoints"] == "4" and piece["NumberOfVerts"] == "4"
    np.testing.assert_array_equal(a["points"][:, 2], 0.0)
    np.testing.assert_array_equal(a["connectivity"], np.arange(4))
    np.testing.assert_array_equal(a["offsets"], np.arange(1, 5))
    with pytest.raises(ValueError, match="Invalid grid_shape"):
        VTKWriter(str(tmp_path / "bad"), (2, 2, 2, 2)).write("x", np.zeros(16), np.zeros((16, 3)))


@pytest.mark.parametrize("dim", [2, 3])
def test_scatter_host_structures(dim):
    """Bucket grids of the scattered-data modes: every point sits in exactly one bucket,
    and every simplex that contains a query point is listed in that point's bucket (so
    the device's first-fit search over the bucket cannot miss it)."""
    from scipy.spatial import Delaunay

    from pyftle_amd.scatter import build_host

    rng = np.random.default_rng(dim)
    pts = rng.uniform(-1.0, 2.0, (dim, 400 if dim == 2 else 300))
    h = build_host(pts)
    ids = np.sort(h["p_ids"])
    np.testing.assert_array_equal(ids, np.arange(pts.shape[1]))
    assert h["p_off"][-1] == pts.shape[1] and h["s_off"][-1] == h["s_ids"].size
    tri = Delaunay(pts.T)
    np.testing.assert_array_equal(h["simplices"], tri.simplices)
    q = rng.uniform(-1.0, 2.0, (2000, dim))
    s = tri.find_simplex(q)
    for x, si in zip(q[s >= 0], s[s >= 0]):
        k = np.clip(np.floor((x - h["s_org"]) / h["s_size"]).astype(int), 0, h["s_dims"] - 1)
        b = k[0] + h["s_dims"][0] * (k[1] + (h["s_dims"][1] * k[2] if dim == 3 else 0))
        assert si in h["s_ids"][h["s_off"][b]:h["s_off"][b + 1]]
    # hull flags: simplices with a neighbour missing
    np.testing.assert_array_equal(h["hull"], (tri.neighbors < 0).any(axis=1))


def test_clough_tocher_factors_and_patch_restatement():
    """The Clough-Tocher construction the device evaluates (Bezier ordinates from vertex
    values and gradients, cross-edge derivative toward the neighbour's centroid, -1/2 on
    the hull; Farin 1986), restated here in NumPy from the factors of ``ct_factors``,
    reproduces SciPy's CloughTocher2DInterpolator with SciPy's own gradients."""
    from scipy.interpolate import CloughTocher2DInterpolator
    from scipy.spatial import Delaunay

    from pyftle_amd.scatter import ct_factors

    rng = np.random.default_rng(0)
    P = rng.uniform(0, 1, (80, 2))
    f = np.sin(3 * P[:, 0]) + np.cos(2 * P[:, 1]) + P[:, 0] * P[:, 1]
    tri = Delaunay(P)
    ct = CloughTocher2DInterpolator(tri, f)
    G = np.asarray(ct.grad).reshape(-1, 2)
    g = ct_factors(P, tri.simplices, tri.transform, tri.neighbors)
    xs = rng.uniform(0.1, 0.9, (300, 2))
    ss = tri.find_simplex(xs)
    worst = 0.0
    for x, s_ in zip(xs[ss >= 0], ss[ss >= 0]):
        v = tri.simplices[s_]
        (P1, P2, P3), (f1, f2, f3), (g1, g2, g3) = P[v], f[v], G[v]
        e12, e13, e23 = P2 - P1, P3 - P1, P3 - P2
        c3000, c0300, c0030 = f1, f2, f3
        c2100, c2010 = (g1 @ e12 + 3 * c3000) / 3, (g1 @ e13 + 3 * c3000) / 3
        c1200, c0210 = (-(g2 @ e12) + 3 * c0300) / 3, (g2 @ e23 + 3 * c0300) / 3
        c1020, c0120 = (-(g3 @ e13) + 3 * c0030) / 3, (-(g3 @ e23) + 3 * c0030) / 3
        c2001, c0201, c0021 = (c2100 + c2010 + c3000) / 3, (c1200 + c0300 + c0210) / 3, (c1020 + c0120 + c0030) / 3
        ga = g[s_]
        c0111 = (ga[0] * (-c0300 + 3 * c0210 - 3 * c0120 + c0030) + (-c0300 + 2 * c0210 - c0120 + c0021 + c0201)) / 2
        c1011 = (ga[1] * (-c0030 + 3 * c1020 - 3 * c2010 + c3000) + (-c0030 + 2 * c1020 - c2010 + c2001 + c0021)) / 2
        c1101 = (ga[2] * (-c3000 + 3 * c2100 - 3 * c1200 + c0300) + (-c3000 + 2 * c2100 - c1200 + c2001 + c0201)) / 2
        c1002, c0102, c0012 = (c1101 + c1011 + c2001) / 3, (c1101 + c0111 + c0201) / 3, (c1011 + c0111 + c0021) / 3
        c0003 = (c1002 + c0102 + c0012) / 3
        T = tri.transform[s_]
        cc = T[:2] @ (x - T[2])
        b = np.array([cc[0], cc[1], 1 - cc.sum()])
        mn = b.min()
        b1, b2, b3 = b - mn
        b4 = 3 * mn
        val = (b1**3 * c3000 + 3 * b1**2 * b2 * c2100 + 3 * b1**2 * b3 * c2010 + 3 * b1**2 * b4 * c2001
               + 3 * b1 * b2**2 * c1200 + 6 * b1 * b2 * b4 * c1101 + 3 * b1 * b3**2 * c1020
               + 6 * b1 * b3 * b4 * c1011 + 3 * b1 * b4**2 * c1002 + b2**3 * c0300 + 3 * b2**2 * b3 * c0210
               + 3 * b2**2 * b4 * c0201 + 3 * b2 * b3**2 * c0120 + 6 * b2 * b3 * b4 * c0111
               + 3 * b2 * b4**2 * c0102 + b3**3 * c0030 + 3 * b3**2 * b4 * c0021 + 3 * b3 * b4**2 * c0012
               + b4**3 * c0003)
        worst = max(worst, abs(val - ct(x[None])[0]))
    assert worst < 1e-13


def test_surrogate_probe_tiles_cover_the_seeds():
    """The probe plan's tiling (CPU tensors): tile ids in range, probe indices distinct, every
    tile of a seed grid probed, and fewer tiles per axis when a tile would hold under 32
    probed seeds."""
    import torch

    from pyftle_amd.surrogate_ftle import SurrogateFTLE

    xs = (torch.arange(2048, dtype=torch.float64) + 0.5) / 1024.0
    ys = (torch.arange(1024, dtype=torch.float64) + 0.5) / 1024.0
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1)], 1)
    # the default spacing (every 256th seed): 8192 probes, 16 x 16 tiles of 32
    plan = SurrogateFTLE(None, None, 10.0, 0.05)._tiles_of(seeds)
    assert plan.n_tiles == 256 and plan.probe.numel() == seeds.shape[0] // 256
    assert int(torch.bincount(plan.tile[plan.probe], minlength=256).min()) >= 20
    sf = SurrogateFTLE(None, None, 10.0, 0.05, probe=64)
    plan = sf._tiles_of(seeds)
    tid, idx, n_tiles = plan.tile, plan.probe, plan.n_tiles
    assert n_tiles == 256 and int(plan.tile_seeds.sum()) == seeds.shape[0] and int(tid.min()) == 0 and int(tid.max()) == 255
    assert idx.numel() == seeds.shape[0] // 64 and torch.unique(idx).numel() == idx.numel()
    per_tile = torch.bincount(tid[idx], minlength=n_tiles)
    assert int(per_tile.min()) >= 100 and int(per_tile.max()) <= 160
    # 256 x 256 seeds: 1024 probes, 5 tiles per axis (about 41 probes each)
    small = seeds.reshape(1024, 2048, 2)[:256, :256].reshape(-1, 2).contiguous()
    plan = sf._tiles_of(small)
    tid, idx, n_tiles = plan.tile, plan.probe, plan.n_tiles
    assert n_tiles == 25 and int(torch.bincount(tid[idx], minlength=25).min()) >= 20
    # non-finite coordinates do not break the tiling (they go to tile 0 of their axis)
    bad = small.clone()
    bad[5, 0] = float("nan")
    bad[9, 1] = float("inf")
    plan = sf._tiles_of(bad)
    assert int(plan.tile.min()) >= 0 and int(plan.tile.max()) < plan.n_tiles
    assert torch.equal(plan.tile[10:], sf._tiles_of(small).tile[10:])
    with pytest.raises(ValueError):
        SurrogateFTLE(None, None, 10.0, 0.05, tile_fallback_max=1.5)
