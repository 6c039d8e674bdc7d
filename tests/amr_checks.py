"""Size-independent checks of a block-pool state (dict from LbmMesh.state()).

Neighbour and neighbour-child tables are reconstructed by brute force from block levels and
coordinates with a hash map, and the reference's own validators (src/mesh/mesh_amr.cu:1925-2047)
are applied as predicates, plus pool bookkeeping and cell/block mask rules.
"""
from __future__ import annotations

import numpy as np

from agal_amd import capi

N_SKIPID = capi.N_SKIPID


def velocities(dim):
    if dim == 2:
        return np.array([[0, 1, 0, -1, 0, 1, -1, -1, 1], [0, 0, 1, 0, -1, 1, 1, -1, -1], [0] * 9])
    return np.array([
        [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1],
        [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1],
        [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1]])


def check_state(st: dict, dim: int, n_root: tuple, periodic=(0, 0, 0), n_max: int | None = None) -> list:
    """Returns a list of violated invariants (empty when the state is consistent).

    n_root: root grid size in blocks per axis (bx, by, bz)."""
    bad = []
    c = velocities(dim)
    qmax = 9 if dim == 2 else 27
    nch = 4 if dim == 2 else 8
    ref, level, nbr, nbc = st["ref"], st["level"], st["nbr"], st["nbr_child"]
    idm = st["id_max"]
    active = ref != capi.REF_INACTIVE
    ids_active = np.nonzero(active)[0]
    # -- ID sets and gap stack -------------------------------------------------------------------
    seen = np.zeros(idm, dtype=bool)
    for L, s in enumerate(st["id_sets"]):
        if s.size and not np.all(np.diff(s) > 0):
            bad.append(f"id_set[{L}] not strictly ascending")
        if s.size and (np.any(~active[s]) or np.any(level[s] != L)):
            bad.append(f"id_set[{L}] holds inactive blocks or blocks of another level")
        seen[s] = True
    if not np.array_equal(seen, active):
        bad.append("id sets do not cover exactly the active blocks")
    gaps = st["gap_set"]
    if n_max is not None:
        if len(gaps) + active.sum() != n_max:
            bad.append("gaps + active != pool size")
        allids = np.concatenate([gaps, ids_active])
        if np.unique(allids).size != n_max or allids.min() < 0 or allids.max() >= n_max:
            bad.append("gap stack and active set do not partition the pool")
    # -- integer coordinates ----------------------------------------------------------------------
    X = st["coords"][:, :idm].astype(np.float64)
    h0 = None
    roots = np.nonzero(active & (level == 0))[0]
    if roots.size:
        xs = np.unique(X[0, roots])
        h0 = xs[1] - xs[0] if xs.size > 1 else 1.0
    key = {}
    icoord = {}
    for b in ids_active:
        L = int(level[b])
        h = h0 / (1 << L)
        ijk = tuple(int(round(X[a, b] / h)) for a in range(dim)) + ((0,) if dim == 2 else ())
        icoord[b] = ijk
        if (L, ijk) in key:
            bad.append(f"two active blocks at level {L} {ijk}")
        key[(L, ijk)] = b
    # -- neighbours by brute force ---------------------------------------------------------------
    def expected_nbr(b, p):
        L = int(level[b])
        ext = [n_root[a] << L for a in range(3)]
        if dim == 2:
            ext[2] = 1
        loc = [icoord[b][a] + int(c[a, p]) for a in range(3)]
        # Precedence of face IDs on edges/corners: x over y over z (mesh_init.cu:406-460).
        for a, (lo, hi) in enumerate(((-1, -2), (-3, -4), (-5, -6))):
            if a >= dim:
                break
            if loc[a] < 0 or loc[a] >= ext[a]:
                if periodic[a]:
                    loc[a] %= ext[a]
                else:
                    return lo if loc[a] < 0 else hi
        return key.get((L, tuple(loc)), N_SKIPID)

    n_nbr_bad = 0
    for b in ids_active:
        for p in range(qmax):
            e = expected_nbr(b, p)
            if nbr[p, b] != e:
                n_nbr_bad += 1
                if n_nbr_bad <= 3:
                    bad.append(f"nbr[{p}][{b}] = {nbr[p, b]}, brute force {e}")
            # neighbour-child: first child of an existing refined neighbour, else the
            # neighbour's negative code, else N_SKIPID
            nb = nbr[p, b]
            if nb >= 0:
                fc = nbc[0, nb]
                ec = fc if fc >= 0 else N_SKIPID
            else:
                ec = nb
            if p == 0:
                ec = nbc[0, b]
            if nbc[p, b] != ec:
                n_nbr_bad += 1
                if n_nbr_bad <= 3:
                    bad.append(f"nbr_child[{p}][{b}] = {nbc[p, b]}, expected {ec}")
        # symmetry
        for p in range(1, qmax):
            nb = nbr[p, b]
            if nb >= 0:
                pb = [q for q in range(qmax) if all(c[a, q] == -c[a, p] for a in range(3))][0]
                if nbr[pb, nb] != b:
                    bad.append(f"asymmetric neighbours {b} -> {nb} (p={p})")
                    break
    # -- children ---------------------------------------------------------------------------------
    for b in ids_active:
        fc = nbc[0, b]
        r = ref[b]
        if fc >= 0:
            kids = np.arange(fc, fc + nch)
            if np.any(~active[kids]) or np.any(level[kids] != level[b] + 1):
                bad.append(f"children of {b} not active at level+1")
                continue
            for k, kid in enumerate(kids):
                off = (k % 2, (k // 2) % 2, (k // 4) % 2)
                exp = tuple(2 * icoord[b][a] + off[a] for a in range(3)) if dim == 3 else \
                    (2 * icoord[b][0] + off[0], 2 * icoord[b][1] + off[1], 0)
                if icoord[kid] != exp:
                    bad.append(f"child {kid} of {b} at {icoord[kid]} expected {exp}")
                    break
            kid_ref = ref[kids]
            has_ref_kid = np.any(np.isin(kid_ref, [capi.REF_REFINED, capi.REF_REFINED_WCHILD,
                                                   capi.REF_REFINED_PERM, capi.REF_REFINED_WCHILD_PERM]))
            if r == capi.REF_REFINED and has_ref_kid:
                bad.append(f"block {b} REFINED but has refined children")
            if r == capi.REF_REFINED_WCHILD and not has_ref_kid:
                bad.append(f"block {b} REFINED_WCHILD without refined children")
        # Cu_RefinementValidator_1
        if r == capi.REF_REFINED and fc < 0:
            bad.append(f"refined block {b} without children")
        if r == capi.REF_UNREFINED and fc > -1:
            bad.append(f"unrefined block {b} with children")
        # Cu_RefinementValidator_4
        if fc >= idm:
            bad.append(f"child id beyond id_max at {b}")
    # Cu_RefinementValidator_2
    refined_states = [capi.REF_REFINED, capi.REF_REFINED_WCHILD, capi.REF_REFINED_PERM,
                      capi.REF_REFINED_WCHILD_PERM, capi.REF_MARK_REFINE]
    for b in ids_active:
        for p in range(qmax):
            nb = nbr[p, b]
            if nb > -1 and ref[nb] in refined_states and nbc[p, b] == N_SKIPID:
                bad.append(f"validator 2: block {b} direction {p}")
    # -- masks ------------------------------------------------------------------------------------
    cpb = 16 if dim == 2 else 64
    cm = st["cell_mask"].reshape(-1, cpb)
    t = np.arange(cpb)
    ijk = np.stack([t % 4, (t // 4) % 4, t // 16])
    for b in ids_active:
        skip = nbr[:, b] == N_SKIPID
        exp = np.zeros(cpb, dtype=np.int32)
        if skip[1:].any():
            exp[:] = 1
        for p in np.nonzero(skip)[0]:
            if p == 0:
                continue
            outer = np.ones(cpb, dtype=bool)
            for a in range(dim):
                if c[a, p] == 1:
                    outer &= ijk[a] >= 2
                if c[a, p] == -1:
                    outer &= ijk[a] < 2
            exp[outer] = 2
        keep = cm[b] < 0
        if not np.array_equal(np.where(keep, cm[b], exp), cm[b]):
            bad.append(f"cell masks of block {b} do not follow the ghost rule")
            break
    bm = st["blk_mask"]
    for b in ids_active:
        if bm[b] < 0:
            continue
        e = 0
        if nbc[0, b] > -1 and np.any(nbc[1:, b] == N_SKIPID):
            e = 1
        if bm[b] != e:
            bad.append(f"block mask of {b} = {bm[b]}, expected {e}")
            break
        # Cu_RefinementValidator_3
        fc = nbc[0, b]
        if fc > -1 and np.any(cm[fc:fc + nch] > 0) and bm[b] != 1:
            bad.append(f"validator 3: block {b}")
            break
    # -- 2:1 balance: every neighbour position of a level-L block (L >= 2) lies in a refined
    # level L-2 block or outside the domain.
    for b in ids_active:
        L = int(level[b])
        if L < 2:
            continue
        ext = [n_root[a] << L for a in range(3)]
        for p in range(1, qmax):
            loc = [icoord[b][a] + int(c[a, p]) for a in range(3)]
            out = False
            for a in range(dim):
                if loc[a] < 0 or loc[a] >= ext[a]:
                    if periodic[a]:
                        loc[a] %= ext[a]
                    else:
                        out = True
            if out:
                continue
            gp = tuple(v >> 2 for v in loc)
            g = key.get((L - 2, gp if dim == 3 else (gp[0], gp[1], 0)))
            if g is None or nbc[0, g] < 0:
                bad.append(f"2:1 balance violated next to block {b} (level {L})")
                break
    return bad


def check_state_fast(st: dict, n_root: tuple, n_max: int | None = None) -> list:
    """check_state's invariants for a 3D, non-periodic pool, vectorized over blocks (for pools of
    hundreds of thousands of blocks): ID sets and gap stack, unique lattice positions per level,
    the neighbour table against a position lookup (domain faces by x, y, z precedence, missing
    same-level neighbours N_SKIPID), the neighbour-child rule, children of refined blocks in place,
    the reference's validators 1, 2 and 4, and 2:1 balance."""
    bad = []
    c = velocities(3)
    ref, level, nbr, nbc = st["ref"], st["level"], st["nbr"], st["nbr_child"]
    idm = st["id_max"]
    active = ref != capi.REF_INACTIVE
    ids = np.nonzero(active)[0]
    # ID sets and gaps.
    seen = np.zeros(idm, dtype=bool)
    for L, s in enumerate(st["id_sets"]):
        if s.size and not np.all(np.diff(s) > 0):
            bad.append(f"id_set[{L}] not strictly ascending")
        if s.size and (np.any(~active[s]) or np.any(level[s] != L)):
            bad.append(f"id_set[{L}] holds inactive blocks or blocks of another level")
        seen[s] = True
    if not np.array_equal(seen, active):
        bad.append("id sets do not cover exactly the active blocks")
    if n_max is not None:
        allids = np.concatenate([st["gap_set"], ids])
        if allids.size != n_max or np.unique(allids).size != n_max or allids.min() < 0 or allids.max() >= n_max:
            bad.append("gap stack and active set do not partition the pool")
    # Integer positions: block coordinate / block width of its level.
    X = st["coords"][:, :idm].astype(np.float64)
    roots = ids[level[ids] == 0]
    xs = np.unique(X[0, roots])
    h0 = xs[1] - xs[0]
    lev = level[ids].astype(np.int64)
    ijk = np.rint(X[:, ids] / (h0 / (1 << lev))[None, :]).astype(np.int64)
    span = np.int64(1) << 16  # positions + 1 per axis stay below 2^16 (256 root blocks x 2^7)

    def keyof(L, a):
        return ((L * span + (a[0] + 1)) * span + (a[1] + 1)) * span + (a[2] + 1)

    keys = keyof(lev, ijk)
    order = np.argsort(keys)
    ks = keys[order]
    if np.any(ks[1:] == ks[:-1]):
        bad.append("two active blocks share a position")

    def lookup(L, a):
        k = keyof(L, a)
        pos = np.searchsorted(ks, k)
        pos = np.minimum(pos, ks.size - 1)
        hit = ks[pos] == k
        return np.where(hit, ids[order[pos]], N_SKIPID)

    ext = np.stack([np.int64(n_root[a]) << lev for a in range(3)])
    for p in range(27):
        loc = ijk + c[:, p][:, None]
        exp = lookup(lev, loc)
        # Faces: x over y over z (mesh_init.cu:406-460).
        for a, (lo, hi) in reversed(list(enumerate(((-1, -2), (-3, -4), (-5, -6))))):
            exp = np.where(loc[a] < 0, lo, np.where(loc[a] >= ext[a], hi, exp))
        got = nbr[p, ids]
        if not np.array_equal(got, exp):
            k = np.nonzero(got != exp)[0][:3]
            bad.append(f"nbr[{p}] differs at blocks {ids[k].tolist()}: {got[k].tolist()} vs {exp[k].tolist()}")
        # neighbour-child rule
        ec = np.where(got >= 0, np.where(nbc[0, np.maximum(got, 0)] >= 0, nbc[0, np.maximum(got, 0)], N_SKIPID), got)
        if p == 0:
            ec = nbc[0, ids]
        if not np.array_equal(nbc[p, ids], ec):
            bad.append(f"nbr_child[{p}] breaks the first-child rule")
        # validator 2
        refined = np.isin(ref[np.maximum(got, 0)], [capi.REF_REFINED, capi.REF_REFINED_WCHILD, capi.REF_REFINED_PERM,
                                                    capi.REF_REFINED_WCHILD_PERM, capi.REF_MARK_REFINE])
        if np.any((got > -1) & refined & (nbc[p, ids] == N_SKIPID)):
            bad.append(f"validator 2 fails in direction {p}")
    # Children in place; validators 1 and 4.
    fc = nbc[0, ids]
    par = ids[fc >= 0]
    if par.size:
        f0 = nbc[0, par]
        if f0.max() + 7 >= idm:
            bad.append("child id beyond id_max")
        else:
            pij = ijk[:, np.searchsorted(ids, par)]
            for k in range(8):
                kid = f0 + k
                off = np.array([k % 2, (k // 2) % 2, (k // 4) % 2])[:, None]
                if np.any(~active[kid]) or np.any(level[kid] != level[par] + 1):
                    bad.append("children not active one level down")
                    break
                kij = ijk[:, np.searchsorted(ids, kid)]
                if not np.array_equal(kij, 2 * pij + off):
                    bad.append(f"child {k} of some block out of place")
                    break
    r = ref[ids]
    if np.any((r == capi.REF_REFINED) & (fc < 0)) or np.any((r == capi.REF_UNREFINED) & (fc > -1)):
        bad.append("validator 1 fails")
    # 2:1 balance: every in-domain neighbour position of a level-L block (L >= 2) lies in a refined
    # block of level L-2.
    sel = lev >= 2
    if sel.any():
        for p in range(1, 27):
            loc = ijk[:, sel] + c[:, p][:, None]
            inside = np.all((loc >= 0) & (loc < ext[:, sel]), axis=0)
            g = lookup(lev[sel] - 2, loc >> 2)
            ok = ~inside | ((g >= 0) & (nbc[0, np.maximum(g, 0)] >= 0))
            if not np.all(ok):
                bad.append(f"2:1 balance violated (direction {p})")
                break
    return bad
