"""Pins the restatement's tables and rules against the reference's own text (tests/golden)."""
from __future__ import annotations

import json
import os
from fractions import Fraction
from itertools import product

import numpy as np
import pytest

from agal_amd import capi
import oracle_lib as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tables.json")))


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_velocity_set_matches_reference_text(Q):
    c, w, pb, cm = O.velset(Q)
    dim = 2 if Q == 9 else 3
    conn = np.array(GOLD["V_CONN_ID_2D" if dim == 2 else "V_CONN_ID_3D"]).reshape(3, 27)
    qmax = c.shape[1]
    assert np.array_equal(c, conn[:, :qmax])
    assert np.array_equal(pb, GOLD["V_CONN_ID_PB_2D" if dim == 2 else "V_CONN_ID_PB_3D"][:qmax])
    assert np.array_equal(cm, GOLD["V_CONN_MAP_2D" if dim == 2 else "V_CONN_MAP_3D"][:qmax])
    name = {9: "D2Q9", 19: "D3Q19", 27: "D3Q27"}[Q]
    wf = [Fraction(a, b) for a, b in GOLD["LBMw_" + name]]
    assert np.array_equal(w[:Q], [float(x) for x in wf])
    assert np.array_equal(pb[:Q], GOLD["LBMpb_" + name])


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_isotropy_closed_forms(Q):
    c, w, pb, _ = O.velset(Q)
    c, w = c[:, :Q].astype(float), w[:Q]
    dim = 2 if Q == 9 else 3
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(c @ w, 0, atol=1e-16)
    second = np.einsum("aq,bq,q->ab", c, c, w)[:dim, :dim]
    assert np.allclose(second, np.eye(dim) / 3, atol=1e-15)
    assert abs(np.sum(w * c[0] ** 4) - 1 / 3) < 1e-15
    assert abs(np.sum(w * c[0] ** 2 * c[1] ** 2) - 1 / 9) < 1e-15
    # opposite directions
    assert np.array_equal(c[:, pb[:Q]], -c)


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_half_set_pairs_every_link_once(Q):
    _, _, pb, _ = O.velset(Q)
    dim = 2 if Q == 9 else 3
    for p in range(1, Q):
        assert O.in_half_set(dim, p) != O.in_half_set(dim, int(pb[p])), p


@pytest.mark.parametrize("dim", [2, 3])
def test_child_connectivity_matches_reference_table(dim):
    Q = 9 if dim == 2 else 27
    table = GOLD["child_connectivity"][str(dim)]
    for p, per_child in table.items():
        for c, (src, off) in per_child.items():
            assert O.child_neighbor_rule(Q, int(c), int(p)) == (src, off), (p, c)


@pytest.mark.parametrize("dim", [2, 3])
def test_ghost_rule_matches_reference_conditions(dim):
    Q = 9 if dim == 2 else 27
    qmax = 9 if dim == 2 else 27
    conds = GOLD["ghost_conditions"][str(dim)]
    cells = 16 if dim == 2 else 64
    rng = np.random.default_rng(0)
    patterns = [[p] for p in range(1, qmax)] + [sorted(rng.choice(np.arange(1, qmax), 3, replace=False)) for _ in range(40)]
    for pat in patterns:
        skip = [0] * qmax
        for p in pat:
            skip[p] = 1
        for t in range(cells):
            ijk = {"I": t % 4, "J": (t // 4) % 4, "K": t // 16}
            exp = 1
            for p in pat:
                ok = all((ijk[a] >= 2) if op == ">=" else (ijk[a] < 2) for a, op in conds[str(p)])
                if ok:
                    exp = 2
            assert O.cell_mask_rule(Q, skip, t) == exp, (pat, t)
    assert O.cell_mask_rule(Q, [0] * qmax, 5) == 0


def test_parameter_defaults_match_reference_parser():
    p = capi.default_params()
    ints = GOLD["parser_defaults"]["int"]
    dbls = GOLD["parser_defaults"]["double"]
    for k in ("MAX_LEVELS_WALL", "MAX_LEVELS_INTERIOR", "N_LEVEL_START", "Nx", "PERIODIC_X", "PERIODIC_Y",
              "PERIODIC_Z", "S_CRITERION", "P_REFINE"):
        assert getattr(p, k) == ints[k], k
    for k in ("L_c", "L_fy", "L_fz", "v0", "N_REFINE_START", "N_REFINE_INC", "N_REFINE_MAX",
              "s1", "s2", "s3", "s4", "s5", "s6", "G_NEAR_WALL_DISTANCE"):
        assert getattr(p, k) == dbls[k], k


def _lagrange_monomial_matrix():
    """M[i][a]: coefficient of x^i in the cubic Lagrange basis polynomial of node a/3, exactly."""
    nodes = [Fraction(a, 3) for a in range(4)]
    M = [[Fraction(0)] * 4 for _ in range(4)]
    for a in range(4):
        poly = [Fraction(1)]  # coefficients, lowest degree first
        den = Fraction(1)
        for b in range(4):
            if b == a:
                continue
            poly = [Fraction(0)] + poly  # times x
            for k in range(len(poly) - 1):
                poly[k] -= nodes[b] * poly[k + 1]
            den *= nodes[a] - nodes[b]
        for i in range(4):
            M[i][a] = poly[i] / den
    return M


@pytest.mark.parametrize("dim", [2, 3])
def test_cubic_interpolation_tables_match_reference_text(dim):
    """Cu_Interp2D/3D (inc/util_interp.h): every listed term is the exact tensor product of the
    Lagrange monomial matrix for nodes {0,1/3,2/3,1}, except (3D only) terms whose exact value is
    zero, which the reference lists with |C| < 1e-10; every exactly nonzero product is listed;
    terms run in ascending s_F index.  The restatement sums the same terms in the same order."""
    M = _lagrange_monomial_matrix()
    lib = O.oracle_library()
    lib.oracle_cubic_terms.restype = O.C.c_int
    lib.oracle_cubic_terms.argtypes = [O.C.c_int, O.C.c_int, O.C.POINTER(O.C.c_int), O.C.POINTER(O.C.c_double)]
    n = 4 ** dim
    gold = GOLD["interp_cubic"][str(dim)]
    assert len(gold) == n
    n_tiny = 0
    for idx in range(n):
        i, j, k = idx % 4, (idx // 4) % 4, idx // 16
        exact = {}
        for q in range(n):
            a, b, c = q % 4, (q // 4) % 4, q // 16
            exact[q] = M[i][a] * M[j][b] * (M[k][c] if dim == 3 else 1)
        terms = gold[str(idx)]
        qs = [q for q, _ in terms]
        assert qs == sorted(qs) and len(set(qs)) == len(qs)
        for q, coef in terms:
            if exact[q] == 0:
                assert dim == 3 and abs(coef) < 1e-10
                n_tiny += 1
            else:
                assert coef == float(exact[q]), (idx, q)
        assert {q for q in range(n) if exact[q] != 0} <= set(qs)
        qq = (O.C.c_int * 64)()
        cc = (O.C.c_double * 64)()
        m = lib.oracle_cubic_terms(dim, idx, qq, cc)
        assert [[qq[t], cc[t]] for t in range(m)] == terms
    assert n_tiny == (432 if dim == 3 else 0)
