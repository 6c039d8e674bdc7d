"""The geometry part of the CPU restatement, pinned without a GPU.

* The ray-parity predicate: closed forms on a triangulated cube, including rays through the shared
  diagonals of its faces (the tie rule must count them once), and orientation independence.
* The seeded noisy icosphere (host code of the HIP library): topology (closed, consistently
  oriented, Euler characteristic 2), enclosed volume, radii and noise statistics, seeding.
* The STL reader against the reference's own geometry inputs (input/geometry/*.stl).
* Voxelization against an independent inside test, the generalized winding number.
* The near-wall hierarchy against invariants of the marks (mesh_voxelizer_refine.cu:23-414).
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from agal_amd import capi
from oracle_lib import OracleMesh, oracle_library, velset
from parity import body, sphere_params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "geometry")


def ray_count(tri, p):
    t = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
    return oracle_library().oracle_ray_count(t.ctypes.data_as(C.POINTER(C.c_float)), t.shape[0], *map(float, p))


def cube_triangles(lo=0.0, hi=1.0):
    """The 12 triangles of the box [lo, hi]^3, outward facing, each face split along a diagonal."""
    v = np.array([[x, y, z] for z in (lo, hi) for y in (lo, hi) for x in (lo, hi)], dtype=np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, c, d in quads:
        tris += [(a, b, c), (a, c, d)]
    return v[np.array(tris)]


def winding_number(tri, pts):
    """Generalized winding number (sum of signed solid angles / 4 pi, Van Oosterom & Strackee)."""
    t = tri.astype(np.float64)
    out = np.zeros(len(pts))
    for s in range(0, len(pts), 256):
        p = pts[s:s + 256, None, None, :]
        a, b, c = (t[None, :, k, :] - p[:, :, 0, :] for k in range(3))
        la, lb, lc = (np.linalg.norm(x, axis=-1) for x in (a, b, c))
        num = np.einsum("...i,...i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("...i,...i", a, b) * lc + np.einsum("...i,...i", a, c) * lb \
            + np.einsum("...i,...i", b, c) * la
        out[s:s + 256] = (2.0 * np.arctan2(num, den)).sum(axis=1) / (4.0 * np.pi)
    return out


# ---- predicate -----------------------------------------------------------------------------------
def test_ray_parity_on_a_cube_counts_shared_diagonals_once():
    cube = cube_triangles()
    # The +x ray from inside crosses the x = 1 face once; rays along y == z and y == 1 - z run
    # through the diagonals shared by that face's two triangles (and the x = 0 face's).
    for y in (0.25, 0.5, 0.75):
        for z in (0.25, 0.5, 0.75):
            assert ray_count(cube, (0.5, y, z)) == 1, (y, z)
            assert ray_count(cube, (-0.5, y, z)) == 2, (y, z)
            assert ray_count(cube, (1.5, y, z)) == 0, (y, z)
    for y, z in ((-0.5, 0.5), (0.5, 1.5), (2.0, 2.0)):
        assert ray_count(cube, (0.5, y, z)) == 0
    # Flipping every triangle's orientation changes nothing (the predicate reorients in (y, z)).
    assert ray_count(cube[:, ::-1, :], (0.5, 0.5, 0.5)) == 1
    assert ray_count(cube[:, ::-1, :], (-0.5, 0.5, 0.5)) == 2


def test_ray_parity_ignores_degenerate_projections():
    # A triangle seen edge-on from the x axis (zero (y, z) area) is never crossed.
    edge_on = np.array([[[0, 0, 0], [1, 1, 0], [2, 0.5, 0]]], dtype=np.float32)
    assert ray_count(edge_on, (-1.0, 0.25, 0.0)) == 0


# ---- icosphere -----------------------------------------------------------------------------------
def _topology(tri):
    verts, inv = np.unique(tri.reshape(-1, 3), axis=0, return_inverse=True)
    f = inv.reshape(-1, 3)
    directed = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    return verts, f, directed


@pytest.mark.parametrize("s", [0, 1, 2, 3, 4])
def test_icosphere_is_a_closed_oriented_sphere(s):
    r = 0.25
    tri = capi.icosphere((0.5, 0.5, 0.5), r, s, 0.0, 0)
    assert tri.shape == (20 * 4 ** s, 3, 3)
    verts, f, directed = _topology(tri)
    assert len(verts) == 10 * 4 ** s + 2
    # Every directed edge once, and its reverse once: closed and consistently oriented.
    d = set(map(tuple, directed.tolist()))
    assert len(d) == len(directed)
    assert all((b, a) in d for a, b in d)
    assert len(verts) - len(d) // 2 + len(f) == 2
    rad = np.linalg.norm(verts.astype(np.float64) - 0.5, axis=1)
    np.testing.assert_allclose(rad, r, rtol=2e-6)
    # Outward normals: positive enclosed volume approaching the ball's.
    v = tri.astype(np.float64) - 0.5
    vol = np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum() / 6.0
    ball = 4.0 / 3.0 * np.pi * r ** 3
    assert 0 < vol < ball
    if s == 4:
        assert vol > 0.99 * ball


def test_icosphere_noise_is_seeded_uniform_and_shared_by_faces():
    r, eps = 0.08, 0.05
    a = capi.icosphere((0.3, 0.5, 0.5), r, 4, eps, 123)
    b = capi.icosphere((0.3, 0.5, 0.5), r, 4, eps, 123)
    c = capi.icosphere((0.3, 0.5, 0.5), r, 4, eps, 124)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    verts, f, directed = _topology(a)
    assert len(verts) == 2562  # a noisy vertex keeps one position in every face that uses it
    d = set(map(tuple, directed.tolist()))
    assert all((y, x) in d for x, y in d)
    rad = np.linalg.norm(verts.astype(np.float64) - np.array([0.3, 0.5, 0.5]), axis=1) / r
    assert rad.min() >= 1 - eps - 1e-5 and rad.max() <= 1 + eps + 1e-5
    assert rad.min() < 1 - 0.95 * eps and rad.max() > 1 + 0.95 * eps
    assert abs(rad.mean() - 1.0) < 0.004
    # Uniform on [1 - eps, 1 + eps): variance eps^2 / 3.
    assert abs(rad.var() / (eps ** 2 / 3.0) - 1.0) < 0.1


# ---- STL -----------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["test.stl", "monkey.stl"])
def test_stl_reader_matches_reference_inputs(name):
    path = os.path.join(GOLD, name)
    text = open(path).read()
    vals = [np.float32(x) for x in re.findall(r"vertex\s+(\S+)\s+(\S+)\s+(\S+)", text) for x in x]
    ref = np.array(vals, dtype=np.float32).reshape(-1, 3, 3)
    keep = ~((ref[:, 0] == ref[:, 1]).all(1) | (ref[:, 0] == ref[:, 2]).all(1) | (ref[:, 1] == ref[:, 2]).all(1))
    tri = capi.load_stl(path)
    assert tri.shape[0] == keep.sum() == text.count("facet normal") - (~keep).sum()
    np.testing.assert_array_equal(tri, ref[keep])


# ---- voxelization and marks ----------------------------------------------------------------------
def _cell_centres(m):
    st = m.state()
    X = st["coords"]
    act = st["ref"] != capi.REF_INACTIVE
    lev = st["level"]
    dx = 1.0 / m.params.Nx / (2.0 ** lev)
    t = np.arange(64)
    I, J, K = t % 4, (t // 4) % 4, t // 16
    cx = X[0][:, None] + (I[None, :] + 0.5) * dx[:, None]
    cy = X[1][:, None] + (J[None, :] + 0.5) * dx[:, None]
    cz = X[2][:, None] + (K[None, :] + 0.5) * dx[:, None]
    return st, act, np.stack([cx, cy, cz], axis=-1)


def test_voxelization_matches_the_winding_number():
    tri = body(center=(0.5, 0.5, 0.5), radius=0.3, subdivisions=3, noise=0.05, seed=5)
    m = OracleMesh(sphere_params(vs=capi.D3Q19, N=16, levels=1, collision=capi.BGK), 128)
    m.set_geometry(tri)
    m.M_Advance_RefineNearWall()
    st, act, pts = _cell_centres(m)
    leaf = act & (st["nbr_child"][0] < 0)
    mask = st["cell_mask"].reshape(-1, 64)[leaf]
    wn = winding_number(tri, pts[leaf].reshape(-1, 3)).reshape(-1, 64)
    assert np.all((np.abs(wn) < 1e-6) | (np.abs(wn - 1) < 1e-6))
    inside = wn > 0.5
    assert inside.sum() > 400  # 4/3 pi (0.3 * 16)^3 = 463 cells
    np.testing.assert_array_equal(mask == capi.CELL_SOLID, inside)
    # Blocks holding a solid cell carry a solid block mask (SOLID, SOLIDB).
    blk = st["blk_mask"][leaf]
    assert np.all(np.isin(blk[inside.any(1)], [capi.BLOCKMASK_SOLID, capi.BLOCKMASK_SOLIDB]))
    assert np.all(blk[~inside.any(1)] != capi.BLOCKMASK_SOLID)


def _dense(m):
    """Leaf cells of a uniform-level mesh as a dense grid of masks."""
    st, act, pts = _cell_centres(m)
    n = m.params.Nx
    g = np.full((n, n, n), 99, dtype=np.int64)
    idx = np.floor(pts[act] * n).astype(np.int64)
    g[idx[..., 0], idx[..., 1], idx[..., 2]] = st["cell_mask"].reshape(-1, 64)[act]
    return st, g


def test_boundary_cells_are_fluid_cells_next_to_solid():
    tri = body(center=(0.5, 0.5, 0.5), radius=0.3, subdivisions=3, noise=0.05, seed=5)
    m = OracleMesh(sphere_params(vs=capi.D3Q19, N=16, levels=1, collision=capi.BGK), 128)
    m.set_geometry(tri)
    m.M_Advance_RefineNearWall()
    st, g = _dense(m)
    solid = g == capi.CELL_SOLID
    near = np.zeros_like(solid)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                near |= np.roll(solid, (dx, dy, dz), axis=(0, 1, 2))
    expect = near & ~solid
    got = g == capi.CELL_BOUNDARY
    assert got.sum() > 500
    assert not np.any(got & ~expect)
    # Cells the rule would flag but the marks leave out lie in interior SOLID blocks only.
    blk_of = np.zeros_like(g)
    _, act, pts = _cell_centres(m)
    idx = np.floor(pts[act] * 16).astype(np.int64)
    blk_of[idx[..., 0], idx[..., 1], idx[..., 2]] = np.repeat(st["blk_mask"][act][:, None], 64, axis=1)
    assert np.all(blk_of[expect & ~got] == capi.BLOCKMASK_SOLID)


def test_near_wall_hierarchy_resolves_the_surface_on_the_finest_level():
    tri = body(center=(0.5, 0.5, 0.5), radius=0.2, subdivisions=3, noise=0.05, seed=2)
    levels = 3
    m = OracleMesh(sphere_params(vs=capi.D3Q19, N=16, levels=levels, collision=capi.BGK), 8192)
    m.set_geometry(tri)
    assert m.M_Advance_RefineNearWall() == capi.AGAL_OK
    st, act, pts = _cell_centres(m)
    leaf = act & (st["nbr_child"][0] < 0)
    masks = st["cell_mask"].reshape(-1, 64)
    lev = st["level"]
    surface = leaf & (masks == capi.CELL_SOLID).any(1) & (masks != capi.CELL_SOLID).any(1)
    assert surface.sum() > 50
    # Every leaf block the surface passes through is on the finest level ...
    assert np.all(lev[surface] == levels - 1)
    # ... and so is every leaf block next to one (the exterior marks), on every level.
    nb = st["nbr"][:, surface]
    nb = nb[nb >= 0]
    assert np.all((lev[nb] == levels - 1) | ~leaf[nb])
    # Wall refinement is frozen against coarsening; boundary cells exist only on the finest level.
    refined = act & (st["nbr_child"][0] >= 0)
    assert np.all(np.isin(st["ref"][refined], [capi.REF_REFINED_PERM, capi.REF_REFINED_WCHILD_PERM]))
    bnd_blocks = (masks == capi.CELL_BOUNDARY).any(1)
    assert bnd_blocks.sum() > 0 and np.all(lev[bnd_blocks] == levels - 1)
    # The solid cells of every level agree with the winding number.
    sel = np.nonzero(leaf)[0][::7]
    wn = winding_number(tri, pts[sel].reshape(-1, 3)).reshape(-1, 64)
    np.testing.assert_array_equal(masks[sel] == capi.CELL_SOLID, wn > 0.5)


def test_near_wall_band_width_follows_the_parameter():
    """A wider near-wall distance gives more propagation passes and more refined blocks."""
    tri = body(center=(0.5, 0.5, 0.5), radius=0.15, subdivisions=2, noise=0.0)
    counts = []
    for d in (0.05, 0.4):
        m = OracleMesh(sphere_params(vs=capi.D3Q19, N=32, levels=2, collision=capi.BGK, G_NEAR_WALL_DISTANCE=d), 16384)
        m.set_geometry(tri)
        m.M_Advance_RefineNearWall()
        counts.append(m.sizes.n_ids[1])
    assert counts[1] > counts[0] > 0


