"""Pins the restatement's step, interpolation, averaging and criterion with closed forms,
invariants and brute force (the reference ships no golden outputs for this path)."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
import oracle_lib as O
from oracle_lib import OracleMesh
from parity import ldc_params

import json as _json
import os as _os

GOLD_MRT = _json.load(open(_os.path.join(_os.path.dirname(__file__), "golden", "reference_tables.json")))["mrt_d3q19_legacy"]


def feq_np(Q, rho, u):
    """Second-order equilibrium, f_q = w_q rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u)."""
    c, w, _, _ = O.velset(Q)
    c, w = c[:, :Q].astype(float), w[:Q]
    cu = np.einsum("aq,an->qn", c, u)
    uu = np.sum(u * u, axis=0)
    return w[:, None] * rho[None, :] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * uu[None, :])


def moments_np(Q, f):
    c, _, _, _ = O.velset(Q)
    c = c[:, :Q].astype(float)
    rho = f.sum(axis=0)
    return rho, (c @ f) / rho


def cell_positions(m):
    """Integer cell coordinates (level-0 units) of every cell of every block up to id_max."""
    s = m.sizes
    dim, cpb = s.n_dim, s.cells_per_block
    X = m.coords().astype(np.float64)
    lev = m.block_ints(capi.BLK_LEVEL)
    dx0 = 1.0 / m.params.Nx
    t = np.arange(cpb)
    ijk = np.stack([t % 4, (t // 4) % 4, t // 16])
    out = np.zeros((3, X.shape[1] * cpb))
    for a in range(dim):
        h = dx0 / (2.0 ** lev)
        out[a] = ((X[a][:, None] + (ijk[a][None, :] + 0.5) * h[:, None]) / dx0).ravel()
    return out


# ---- BGK collision (solver_lbm_kernels.cu:204-289) ---------------------------------------------
@pytest.mark.parametrize("Q", [9, 19, 27])
def test_collision_fixed_point_conservation_relaxation(Q):
    p = ldc_params(vs=Q, N=8 if Q != 9 else 16, Re=20.0, prec=capi.FP64, bc_preset=capi.BC_NONE)
    m = OracleMesh(p, 64)
    m.S_SetIC(0)
    nc = m.populations().shape[1]
    rng = np.random.default_rng(Q)
    dim = 2 if Q == 9 else 3
    rho = 1 + 0.1 * rng.uniform(-1, 1, nc)
    u = np.zeros((3, nc))
    u[:dim] = 0.08 * rng.uniform(-1, 1, (dim, nc))
    fe = feq_np(Q, rho, u)
    _, _, pb, _ = O.velset(Q)
    # equilibrium is a fixed point
    m.set_populations(fe)
    m.collide(0)
    post = m.populations()[pb[:Q]]  # after collide, slot q holds f*_q; the download shows slot pb(q)
    assert np.max(np.abs(post - fe) / fe) < 1e-14
    # conservation and relaxation towards equilibrium with omega = dx/tau
    f = fe * (1 + 0.2 * rng.uniform(-1, 1, fe.shape))
    m.set_populations(f)
    m.collide(0)
    post = m.populations()[pb[:Q]]
    r0, u0 = moments_np(Q, f)
    r1, u1 = moments_np(Q, post)
    assert np.max(np.abs(r1 - r0)) < 1e-14 and np.max(np.abs(u1 - u0)) < 1e-14
    dxf, tau = m.level_params()
    nu = p.v0
    omega = dxf[0] / (3 * nu + 0.5 * dxf[0])
    assert abs(omega - dxf[0] / tau[0]) < 1e-15
    assert np.max(np.abs((post - feq_np(Q, r0, u0)) - (1 - omega) * (f - feq_np(Q, r0, u0)))) < 1e-13


# ---- swap stream (solver_lbm_kernels.cu:532-633): brute-force permutation ------------------------
@pytest.mark.parametrize("Q,N", [(9, 16), (19, 8), (27, 8)])
@pytest.mark.parametrize("periodic", [1, 0])
def test_stream_is_the_reference_permutation(Q, N, periodic):
    dim = 2 if Q == 9 else 3
    p = ldc_params(vs=Q, N=N, prec=capi.FP64, bc_preset=capi.BC_NONE,
                   PERIODIC_X=periodic, PERIODIC_Y=periodic, PERIODIC_Z=periodic)
    m = OracleMesh(p, 64)
    m.S_SetIC(0)
    f = m.populations()
    nc = f.shape[1]
    c, _, pb, _ = O.velset(Q)
    # Unique positive value per (slot, cell): slot s of a cell is download row pb(s).
    slots = 1.0 + np.arange(Q * nc, dtype=np.float64).reshape(Q, nc) / (Q * nc)
    m.set_populations(slots[pb[:Q]])
    m.stream(0)
    after = m.populations()[pb[:Q]]
    pos = np.rint(cell_positions(m) - 0.5).astype(int)
    n = np.array([N, N, N if dim == 3 else 1])
    lin = pos[0] + n[0] * (pos[1] + n[1] * pos[2])
    where = np.empty(n.prod(), dtype=int)
    where[lin] = np.arange(nc)
    expect = slots.copy()
    for s in range(1, Q):
        tgt = pos + c[:, s][:, None]
        inside = np.all((tgt >= 0) & (tgt < n[:, None]), axis=0)
        if periodic:
            tgt = tgt % n[:, None]
            inside[:] = True
        src = np.nonzero(inside)[0]
        tl = tgt[0, src] + n[0] * (tgt[1, src] + n[1] * tgt[2, src])
        # the value in slot s at x moves to slot pb(s) at x + c_s, and vice versa
        expect[pb[s], where[tl]] = slots[s, src]
    assert np.array_equal(after, expect)
    assert np.array_equal(np.sort(after.ravel()), np.sort(slots.ravel()))


# ---- moving wall + half-way bounce-back: plane Couette flow is exact ---------------------------
@pytest.mark.parametrize("Q", [9, 19])
def test_couette_closed_form(Q):
    dim = 2 if Q == 9 else 3
    N, U = (16, 0.01) if dim == 2 else (8, 0.01)
    nu_lat = 1.0 / 6.0  # tau = 1
    p = ldc_params(vs=Q, N=N, prec=capi.FP64, PERIODIC_X=1, PERIODIC_Y=1 if dim == 3 else 0,
                   v0=nu_lat / N, bc_velocity=U)
    m = OracleMesh(p, 64)
    m.S_SetIC(0)
    for _ in range(4000 if dim == 2 else 1500):
        m.S_Advance(0)
    m.S_RefreshVariables(0)
    mac = m.macros()
    pos = cell_positions(m)
    h = pos[dim - 1]  # wall-normal coordinate in cell units (centres at j + 0.5)
    exact = U * h / N
    assert np.max(np.abs(mac[1] - exact)) < 1e-9
    assert np.max(np.abs(mac[2:2 + dim - 1])) < 1e-12


def two_level_mesh(Q=19, N=8, **kw):
    p = ldc_params(vs=Q, N=N, levels=2, prec=capi.FP64, **kw)
    m = OracleMesh(p, 256)
    m.S_SetIC(0)
    return m, p


def zero_moment_perturbation(Q, nc, rng, scale=1e-3):
    """Random non-equilibrium parts with zero mass and zero momentum per cell."""
    c, _, _, _ = O.velset(Q)
    A = np.vstack([np.ones(Q), c[:, :Q].astype(float)])
    d = scale * rng.standard_normal((Q, nc))
    return d - np.linalg.pinv(A) @ (A @ d)


# ---- interpolation (solver_lbm_kernels_comm.cu:13-138) -----------------------------------------
def test_interpolation_rescales_non_equilibrium_by_tau_ratio():
    m, p = two_level_mesh()
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    nc = m.populations().shape[1]
    rng = np.random.default_rng(1)
    rho0, u0 = 1.02, np.array([[0.03], [-0.01], [0.02]])
    fe = feq_np(19, np.full(nc, rho0), np.repeat(u0, nc, axis=1))
    delta = zero_moment_perturbation(19, 1, rng)
    m.set_populations(fe + delta)
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.S_Interpolate(0, capi.INTERP_ADDED)
    f = m.populations()
    fc = m.block_ints(capi.BLK_NBR_CHILD)[0, 3]
    kids = f[:, fc * 64:(fc + 8) * 64]
    dx0 = 1.0 / p.Nx
    ratio = (3 * p.v0 + 0.25 * dx0) / (3 * p.v0 + 0.5 * dx0)
    assert np.max(np.abs(kids - (fe[:, :1] + delta * ratio))) < 1e-15


def test_interpolation_stencil_is_trilinear_in_parent_sub_cube():
    """Children equal the trilinear interpolant (with extrapolation at block edges) through the
    2x2x2 parent cells of their octant, evaluated in Lagrange form."""
    m, p = two_level_mesh()
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    nc = m.populations().shape[1]
    rng = np.random.default_rng(2)
    rho = np.ones(nc)
    u = np.zeros((3, nc))
    base = feq_np(19, rho, u)
    f = base * (1 + 0.3 * rng.uniform(-1, 1, base.shape))
    m.set_populations(f)
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.S_Interpolate(0, capi.INTERP_ADDED)
    g = m.populations()
    # rescaled parent values, s = feq + (f - feq) * tau_{1}/tau_{0}
    dxf, tau = m.level_params()
    par = f[:, 3 * 64:4 * 64]
    r, uu = moments_np(19, par)
    fe = feq_np(19, r, uu)
    s = fe + (par - fe) * (tau[1] / tau[0])
    fc = m.block_ints(capi.BLK_NBR_CHILD)[0, 3]
    t = np.arange(64)
    ijk = np.stack([t % 4, (t // 4) % 4, t // 16])
    for c in range(8):
        oc = np.array([c % 2, (c // 2) % 2, c // 4])
        # child cell centres in parent-cell units relative to the octant's first parent cell
        xc = 2 * oc[:, None] + (ijk + 0.5) / 2 - 0.5 - 2 * oc[:, None]
        val = np.zeros((19, 64))
        for corner in range(8):
            k = np.array([corner % 2, (corner // 2) % 2, corner // 4])
            wgt = np.prod(np.where(k[:, None] == 1, xc, 1 - xc), axis=0)
            pc = (2 * oc + k)
            pidx = pc[0] + 4 * pc[1] + 16 * pc[2]
            val += s[:, pidx][:, None] * wgt[None, :]
        assert np.max(np.abs(g[:, (fc + c) * 64:(fc + c + 1) * 64] - val)) < 1e-13, c


@pytest.mark.parametrize("Q", [9, 19])
def test_cubic_interpolation_reproduces_tricubic_fields(Q):
    """Cu_Interpolate_Cubic (solver_lbm_kernels_comm.cu:406-638): with a uniform macroscopic state
    and non-equilibrium parts that are cubic per axis in the cell position, the rescaled parent
    values are a tensor-product cubic, which the Lagrange interpolant through the 4^d parent cells
    reproduces exactly; so every child equals feq + tau ratio * d(child centre)."""
    dim = 2 if Q == 9 else 3
    p = ldc_params(vs=Q, N=16 if dim == 2 else 8, levels=2, prec=capi.FP64, interp_order=capi.INTERP_CUBIC)
    m = OracleMesh(p, 256)
    m.S_SetIC(0)
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    nc = m.populations().shape[1]
    rng = np.random.default_rng(11)
    u0 = np.array([[0.02], [0.01], [-0.03 if dim == 3 else 0.0]])
    fe = feq_np(Q, np.full(nc, 0.99), np.repeat(u0, nc, axis=1))
    pos = cell_positions(m) / 4.0  # block units keep the monomials O(1)

    def cubic_field(x):
        # sum over a few random monomials x^i y^j z^k with i, j, k <= 3, each with a zero-moment
        # direction vector
        out = np.zeros((Q, x.shape[1]))
        for _ in range(6):
            e = rng.integers(0, 4, 3)
            if dim == 2:
                e[2] = 0
            vec = zero_moment_perturbation(Q, 1, rng, scale=2e-3)
            out += vec * (x[0] ** e[0] * x[1] ** e[1] * x[2] ** e[2])[None, :]
        return out, None

    state = rng.bit_generator.state
    d, _ = cubic_field(pos)
    m.set_populations(fe + d)
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.S_Interpolate(0, capi.INTERP_ADDED)
    g = m.populations()
    dxf, tau = m.level_params()
    ratio = tau[1] / tau[0]
    rng.bit_generator.state = state
    fc = m.block_ints(capi.BLK_NBR_CHILD)[0, 3]
    cpb = 64 if dim == 3 else 16
    kids = slice(fc * cpb, (fc + (8 if dim == 3 else 4)) * cpb)
    dk, _ = cubic_field(cell_positions(m)[:, kids] / 4.0)
    assert np.max(np.abs(g[:, kids] - (fe[:, kids] + ratio * dk))) < 1e-12
    # a linear interpolant cannot do this: the same field through the linear kernel misses
    p_lin = ldc_params(vs=Q, N=16 if dim == 2 else 8, levels=2, prec=capi.FP64)
    m2 = OracleMesh(p_lin, 256)
    m2.S_SetIC(0)
    m2.mark_blocks([3], capi.REF_MARK_REFINE)
    m2.M_RefineAndCoarsenBlocks()
    m2.set_populations(fe + d)
    m2.mark_blocks([3], capi.REF_MARK_REFINE)
    m2.S_Interpolate(0, capi.INTERP_ADDED)
    assert np.max(np.abs(m2.populations()[:, kids] - (fe[:, kids] + ratio * dk))) > 1e-6


# ---- averaging (solver_lbm_kernels_comm.cu:174-347) --------------------------------------------
def interpolate_average_error(plant=0):
    """Uniform macroscopic state with non-equilibrium parts linear in x, y and z: interpolation to
    new children followed by a grid average returns the parent values (tau ratios cancel); the max
    deviation."""
    m, p = two_level_mesh()
    m.plant(plant)
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    nc = m.populations().shape[1]
    rng = np.random.default_rng(4)
    fe = feq_np(19, np.full(nc, 0.99), np.repeat(np.array([[0.02], [0.01], [-0.03]]), nc, axis=1))
    pos = cell_positions(m)
    d0 = zero_moment_perturbation(19, 1, rng)
    d1 = zero_moment_perturbation(19, 3, rng)
    f = fe + d0 + d1 @ pos * 0.1
    m.set_populations(f)
    m.mark_blocks([3], capi.REF_MARK_REFINE)
    m.S_Interpolate(0, capi.INTERP_ADDED)
    m.S_Average(0, capi.AVERAGE_GRID)
    g = m.populations()
    return float(np.max(np.abs(g[:, 3 * 64:4 * 64] - f[:, 3 * 64:4 * 64])))


def test_interpolate_then_average_restores_parent():
    assert interpolate_average_error() < 1e-14


def test_average_of_uniform_children():
    m, p = two_level_mesh()
    m.mark_blocks([5], capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    nc = m.populations().shape[1]
    rng = np.random.default_rng(5)
    fe = feq_np(19, np.full(nc, 1.01), np.repeat(np.array([[0.01], [0.02], [0.0]]), nc, axis=1))
    delta = zero_moment_perturbation(19, 1, rng)
    m.set_populations(fe + delta)
    m.S_Average(0, capi.AVERAGE_GRID)
    g = m.populations()
    dxf, tau = m.level_params()
    # A parent cell takes the average of its 2x2x2 child group when the group's first cell is not
    # a ghost (masks < 2 count as interface in the GRID mode).
    fc = m.block_ints(capi.BLK_NBR_CHILD)[0, 5]
    cm = m.cell_masks()
    t = np.arange(64)
    I, J, K = t % 4, (t // 4) % 4, t // 16
    child = (I // 2) + 2 * (J // 2) + 4 * (K // 2)
    c0 = 2 * (I % 2) + 8 * (J % 2) + 32 * (K % 2)
    averaged = cm[(fc + child) * 64 + c0] < 2
    assert averaged.any() and (~averaged).any()
    got = g[:, 5 * 64:6 * 64]
    assert np.max(np.abs(got[:, averaged] - (fe[:, :1] + delta * tau[0] / tau[1]))) < 1e-15
    assert np.array_equal(got[:, ~averaged], (fe + delta)[:, 5 * 64:6 * 64][:, ~averaged])


# ---- refinement criterion (solver_lbm_criterion.cu:438-592) ------------------------------------
def criterion_case(field, crit=0, N=16, levels=3):
    p = ldc_params(vs=19, N=N, levels=levels, prec=capi.FP64, S_CRITERION=crit,
                   N_REFINE_START=-2.0, N_REFINE_INC=1.0, N_REFINE_MAX=1.0)
    m = OracleMesh(p, 1024)
    m.S_SetIC(0)
    pos = cell_positions(m) / N  # physical coordinates of cell centres
    u, v, w = field(*pos)
    rho = np.ones_like(u)
    m.set_macros(np.vstack([rho, u, v, w]))
    m.S_ComputeRefCriteria()
    return m, pos


def expected_marks(W_cell, level, MLI=3, NRS=-2, NRI=1, NRM=1):
    Wb = np.minimum(W_cell.reshape(-1, 64).max(axis=1), NRM)
    Ld = np.full(Wb.shape, MLI - 1)
    for pp in range(1, MLI):
        Ld = np.where(Wb < NRS - NRI * pp, MLI - 1 - pp, Ld)
    return np.where(Ld > level, capi.REF_MARK_REFINE, capi.REF_UNREFINED)


@pytest.mark.parametrize("a,expect", [(0.3, capi.REF_MARK_REFINE), (0.1, capi.REF_MARK_REFINE), (0.01, capi.REF_UNREFINED)])
def test_criterion_linear_shear(a, expect):
    m, _ = criterion_case(lambda x, y, z: (a * y, 0 * x, 0 * x))
    ref = m.block_ints(capi.BLK_REF)
    assert np.all(ref == expect)


def test_criterion_cubic_field_extrapolated_halo():
    """u = a y^3: the 4th-order halo extrapolation is exact for cubics and central differences give
    a (3y^2 + dx^2), so the block maxima and marks follow in closed form."""
    N, a = 16, 0.05
    m, pos = criterion_case(lambda x, y, z: (a * y ** 3, 0 * x, 0 * x), N=N)
    dx = 1.0 / N
    y = pos[1]
    W = np.floor(np.log2(a * (3 * y ** 2 + dx ** 2)))
    ref = m.block_ints(capi.BLK_REF)
    assert np.array_equal(ref, expected_marks(W, np.zeros(ref.size)))
    assert (ref == capi.REF_MARK_REFINE).any() and (ref == capi.REF_UNREFINED).any()


@pytest.mark.parametrize("A", [((0.3, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
                               ((0.01, -0.025, 0.005), (0.035, 0.0, 0.015), (-0.02, 0.01, -0.01)),
                               ((0.0, 0.15, 0.0), (-0.15, 0.0, 0.0), (0.0, 0.0, 0.0))])
def test_criterion_gradient_norm_of_linear_fields(A):
    """S_CRITERION 2 on u = A x: central differences and the halo extrapolation are exact for
    linear fields, so every cell sees |grad u| = ||A||_F; the vorticity criterion on the same field
    sees |curl| instead (a pure strain has none), which tells the two apart."""
    A = np.array(A)
    field = lambda x, y, z: tuple(A[i, 0] * x + A[i, 1] * y + A[i, 2] * z for i in range(3))
    m, _ = criterion_case(field, crit=2)
    ref = m.block_ints(capi.BLK_REF)
    W = np.floor(np.log2(np.sqrt(np.sum(A ** 2))))
    assert np.all(ref == expected_marks(np.full(ref.size * 64, W), np.zeros(ref.size)))
    m0, _ = criterion_case(field, crit=0)
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    nc = np.sqrt(np.sum(curl ** 2))
    W0 = np.floor(np.log2(nc)) if nc > 0 else -np.inf
    assert np.all(m0.block_ints(capi.BLK_REF) == expected_marks(np.full(ref.size * 64, W0), np.zeros(ref.size)))


def test_criterion_q_invariant():
    a, b = 0.5, 0.25
    m, _ = criterion_case(lambda x, y, z: (a * y, b * x, 0 * x), crit=1)
    W = np.floor(np.log2(abs(a * b)))  # Q-like = 0.5*(2 u_y v_x)
    ref = m.block_ints(capi.BLK_REF)
    assert np.all(ref == expected_marks(np.full(ref.size * 64, W), np.zeros(ref.size)))


# ---- moments (Cu_RefreshVariables, solver_lbm_kernels.cu:110-159) --------------------------------
@pytest.mark.parametrize("Q", [9, 19, 27])
def test_refresh_returns_the_moments_of_equilibria(Q):
    """Equilibria built from random (rho, u) have exactly those zeroth and first moments."""
    dim = 2 if Q == 9 else 3
    p = ldc_params(vs=Q, N=8, prec=capi.FP64, bc_preset=capi.BC_NONE)
    m = OracleMesh(p, 64)
    m.S_SetIC(0)
    nc = m.populations().shape[1]
    rng = np.random.default_rng(100 + Q)
    rho = 1 + 0.2 * rng.uniform(-1, 1, nc)
    u = np.zeros((3, nc))
    u[:dim] = 0.1 * rng.uniform(-1, 1, (dim, nc))
    m.set_populations(feq_np(Q, rho, u))
    m.S_RefreshVariables(0)
    mac = m.macros()
    assert np.max(np.abs(mac[0] - rho)) < 1e-14
    assert np.max(np.abs(mac[1:1 + dim] - u[:dim])) < 1e-14


# ---- inflow / outflow presets of the boundary hook (inc/custom.h:87-97,107-117) ------------------
@pytest.mark.parametrize("Q,preset", [(9, capi.BC_FPSC2D), (19, capi.BC_FPSC3D), (27, capi.BC_FPSC3D)])
def test_inflow_outflow_presets_one_step_closed_form(Q, preset):
    """One step from a uniform flow u' = (0.03, 0, 0) with inflow velocity U = 0.05, rho = 1.

    Collision keeps the equilibrium; links inside the domain swap equal values; a link leaving the
    domain keeps its own post-collision value in slot p after the hook: the moving wall (faces
    -1, -3, -4 and in 3D -5, -6) gives feq_p - 6 w_p c_px U, the anti-bounce-back outlet (face -2)
    gives -feq_p + 2 w_p (1 + 4.5 (c_p.u')^2 - 1.5 u'.u'), which equals feq_pb(u').  The face of a
    link leaving through an edge or corner is its x face, else its y face (x before y before z).
    Expected values come from this rule applied cell by cell to the positions of the cells."""
    dim = 2 if Q == 9 else 3
    N, U, Up = (16, 0.05, 0.03) if dim == 2 else (8, 0.05, 0.03)
    p = ldc_params(vs=Q, N=N, prec=capi.FP64, bc_preset=preset, bc_velocity=U, ic_u=Up, Re=50.0)
    m = OracleMesh(p, 256)
    m.S_SetIC(0)
    f0 = m.populations().copy()
    m.S_Advance(0)
    f1 = m.populations()
    c, w, pb, _ = O.velset(Q)
    pos = np.floor(cell_positions(m)).astype(int)[:, : f0.shape[1]]
    ncell = f0.shape[1]
    uvec = np.zeros((3, ncell))
    uvec[0] = Up
    fe = feq_np(Q, np.ones(ncell), uvec)
    expect = f0.copy()
    n_touched = 0
    for q in range(1, Q):
        tgt = pos + c[:, q][:, None]
        face = np.zeros(ncell, dtype=int)
        for a in reversed(range(dim)):  # x has precedence: assigned last
            lo, hi = tgt[a] < 0, tgt[a] >= N
            face[lo] = -(2 * a + 1)
            face[hi] = -(2 * a + 2)
        out = face < 0
        if not out.any():
            continue
        n_touched += out.sum()
        wall = out & (face != -2)
        outlet = face == -2
        val = np.empty(ncell)
        val[wall] = fe[q][wall] - 6.0 * w[q] * c[0, q] * U
        cu = c[0, q] * Up
        val[outlet] = -fe[q][outlet] + 2.0 * w[q] * (1 + 4.5 * cu * cu - 1.5 * Up * Up)
        # slot q holds direction pb(q): the download shows it as row pb(q)
        expect[pb[q]][out] = val[out]
    assert n_touched > 0
    assert np.max(np.abs(f1 - expect)) < 1e-14
    # the outlet really reproduces the equilibrium of the opposite direction
    assert np.max(np.abs(f1 - f0)) > 1e-3 * U


# ---- MRT (legacy D3Q19 definition: src/solver_old/solver_lbm_collide_mrt_d3q19.cu:14-1460) -------
def mrt_basis(Q):
    m = (O.C.c_int * (Q * Q))()
    r = (O.C.c_int * Q)()
    assert O.oracle_library().oracle_mrt_basis(Q, m, r) == 0
    return np.array(list(m)).reshape(Q, Q), np.array(list(r))


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_mrt_basis_orthogonal_integer_complete(Q):
    M, rate = mrt_basis(Q)
    G = M @ M.T
    assert np.array_equal(G, np.diag(np.diag(G)))
    assert np.linalg.matrix_rank(M) == Q
    c, _, _, _ = O.velset(Q)
    dim = 2 if Q == 9 else 3
    cons = M[rate == 0]
    # the conserved rows span exactly {1, c_x, c_y(, c_z)}
    basis = np.vstack([np.ones(Q), c[:dim, :Q]])
    assert cons.shape[0] == dim + 1
    assert np.linalg.matrix_rank(np.vstack([cons, basis])) == dim + 1
    # the viscous (omega) rows span the traceless second-order tensors
    sec = [c[a, :Q] * c[b, :Q] for a in range(dim) for b in range(a, dim)]
    trace = sum(c[a, :Q] ** 2 for a in range(dim))
    dev = [s for s in sec]
    visc = M[rate == 1]
    assert visc.shape[0] == (2 if dim == 2 else 5)
    # every viscous row is a second-order polynomial orthogonal to the trace
    span = np.vstack([np.ones(Q)] + dev)
    assert np.linalg.matrix_rank(np.vstack([span, visc])) == span.shape[0]
    assert np.allclose(visc @ (trace - trace.mean()), 0)


def test_mrt_d3q19_is_the_legacy_definition():
    M, rate = mrt_basis(19)
    leg = GOLD_MRT["rows"]
    assert np.array_equal(M, np.array(leg))
    sym = {"s_1": 2, "s_2": 3, "s_3": 4, "s_4": 5, "s_5": 6, "s_6": 7, "omeg": 1}
    for k in range(19):
        expect = sym[GOLD_MRT["rates"][str(k)]] if str(k) in GOLD_MRT["rates"] else 0
        assert rate[k] == expect, k


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_mrt_relaxes_each_moment_at_its_rate_and_equals_bgk_at_one_rate(Q):
    dim = 2 if Q == 9 else 3
    rates = dict(s1=1.19, s2=1.4, s3=1.2, s4=1.1, s5=1.3, s6=1.98)
    p = ldc_params(vs=Q, N=8 if dim == 3 else 16, Re=20.0, prec=capi.FP64, bc_preset=capi.BC_NONE,
                   collision=capi.MRT, **rates)
    m = OracleMesh(p, 64)
    m.S_SetIC(0)
    nc = m.populations().shape[1]
    rng = np.random.default_rng(40 + Q)
    rho = 1 + 0.1 * rng.uniform(-1, 1, nc)
    u = np.zeros((3, nc))
    u[:dim] = 0.08 * rng.uniform(-1, 1, (dim, nc))
    fe = feq_np(Q, rho, u)
    f = fe * (1 + 0.2 * rng.uniform(-1, 1, fe.shape))
    _, _, pb, _ = O.velset(Q)
    m.set_populations(f)
    m.collide(0)
    post = m.populations()[pb[:Q]]
    r0, u0 = moments_np(Q, f)
    fe0 = feq_np(Q, r0, u0)
    M, rate = mrt_basis(Q)
    dxf, tau = m.level_params()
    omega = dxf[0] / tau[0]
    S = np.array([0.0, omega] + [rates[f"s{i}"] for i in range(1, 7)])[rate]
    lhs = M @ (post - fe0)
    rhs = (1 - S)[:, None] * (M @ (f - fe0))
    assert np.max(np.abs(lhs - rhs)) < 1e-13
    # all rates equal to omega: BGK
    p1 = ldc_params(vs=Q, N=8 if dim == 3 else 16, Re=20.0, prec=capi.FP64, bc_preset=capi.BC_NONE,
                    collision=capi.MRT, **{f"s{i}": omega for i in range(1, 7)})
    p2 = ldc_params(vs=Q, N=8 if dim == 3 else 16, Re=20.0, prec=capi.FP64, bc_preset=capi.BC_NONE)
    a, b = OracleMesh(p1, 64), OracleMesh(p2, 64)
    for mm in (a, b):
        mm.S_SetIC(0)
        mm.set_populations(f)
        mm.collide(0)
    assert np.max(np.abs(a.populations() - b.populations())) < 1e-14


def test_mrt_couette_closed_form():
    """Steady Couette flow depends only on the shear viscosity, which MRT sets through omega."""
    N, U = 8, 0.01
    p = ldc_params(vs=19, N=N, prec=capi.FP64, PERIODIC_X=1, PERIODIC_Y=1, v0=(1.0 / 6.0) / N, bc_velocity=U,
                   collision=capi.MRT)
    m = OracleMesh(p, 64)
    m.S_SetIC(0)
    for _ in range(4000):
        m.S_Advance(0)
    m.S_RefreshVariables(0)
    mac = m.macros()
    h = cell_positions(m)[2]
    # the steady profile is linear to within O(U^2) effects of the relaxed non-hydrodynamic
    # moments (~3e-9 here, against U h / N ~ 1e-2)
    assert np.max(np.abs(mac[1] - U * h / N)) < 1e-8
