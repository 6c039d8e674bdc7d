"""Extracts the reference's constant tables from its source text into reference_tables.json.

Reads (as text, nothing is compiled or run):
  inc/util.h:94-101                          V_CONN_ID / V_CONN_ID_PB / V_CONN_MAP (2D and 3D)
  src/solver_lbm/solver_lbm_init.cu:79-88    LBMw and LBMpb for D2Q9, D3Q19, D3Q27
  src/mesh/mesh_amr.cu:154-944               child-neighbour table of Cu_UpdateConnectivity
  src/mesh/mesh_amr.cu:981-1297              ghost-cell conditions of Cu_UpdateMasks_2
  inc/parser.h:62-144                        input.txt keys and defaults
  inc/util_interp.h:13-191                   coefficient tables of the cubic interpolation
  src/solver_old/solver_lbm_vel2mom_mrt_d3q19.cu:118-136, solver_lbm_collide_mrt_d3q19.cu:132-146
                                             legacy D3Q19 MRT moment rows and rate assignment

Run from the repository root: python tests/golden/gen_reference_tables.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys
from fractions import Fraction

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def int_array(src, name):
    m = re.search(r"int\s+" + re.escape(name) + r"\[\d+\]\s*=\s*\{([^}]*)\}", src)
    return [int(x) for x in m.group(1).split(",")]


def frac_array(src, name):
    m = re.search(r"double\s+" + re.escape(name) + r"\[\d+\]\s*=\s*\{([^}]*)\}", src)
    out = []
    for tok in m.group(1).split(","):
        a, b = tok.strip().split("/")
        out.append([int(float(a)), int(float(b))])
    return out


def connectivity(src):
    """(dim, direction p, child c) -> (parent-level direction of the source pointer, offset)."""
    start = src.index("void Cu_UpdateConnectivity")
    end = src.index("void Cu_UpdateMasks_1")
    body = src[start:end]
    i2 = body.index("if (N_DIM==2)")
    i3 = body.index("\nelse\n", i2)
    table = {"2": {}, "3": {}}
    for dim, text in (("2", body[i2:i3]), ("3", body[i3:])):
        pending = {}
        for line in text.splitlines():
            m = re.search(r"s_ID_child_nbrs\[(\d+) \+ threadIdx\.x\*N_CHILDREN\] = id_nbr_(\d+)_child<0 \? "
                          r"id_nbr_\d+_child:id_nbr_\d+_child \+ (\d+);", line)
            if m:
                pending[int(m.group(1))] = [int(m.group(2)), int(m.group(3))]
                continue
            m = re.search(r"cblock_ID_nbr\[ID_child_q \+ t_id_incrementor \+ (\d+)\*n_maxcblocks\]", line)
            if m:
                table[dim][m.group(1)] = {str(c): v for c, v in sorted(pending.items())}
                pending = {}
    return table


def ghost_conditions(src):
    """Per direction p: the per-axis conditions under which a cell becomes a ghost (mask 2)."""
    start = src.index("void Cu_UpdateMasks_2")
    end = src.index("Cu_RefineCells_Prep_S1")
    body = src[start:end]
    i3 = body.index("else", body.index("if (N_DIM==2)"))
    out = {"2": {}, "3": {}}
    for dim, text in (("2", body[:i3]), ("3", body[i3:])):
        for m in re.finditer(r"cblock_ID_nbr\[i_kap_b \+ (\d+)\*n_maxcblocks\]==N_SKIPID\)and\([^\n]*\)\)\s*\{\s*"
                             r"if \(([^\n]*)\)\s*cell_mask = 2;", text):
            conds = re.findall(r"\(([IJK])_kap ([<>]=?) 2\)", m.group(2))
            out[dim][m.group(1)] = [[axis, op] for axis, op in conds]
    return out


def parser_defaults(src):
    a, b, c = src.index("params_int = std::map"), src.index("params_dbl = std::map"), src.index("params_str = std::map")
    ints = dict(re.findall(r'\{"(\w+)",\s*(-?\d+(?:\*\d+)?)\}', src[a:b]))
    dbls = dict(re.findall(r'\{"(\w+)",\s*(-?[\d.eE+-]+)\}', src[b:c]))
    ival = {k: eval(v, {"__builtins__": {}}) for k, v in ints.items()}  # products like 10*64
    return {"int": ival, "double": {k: float(v) for k, v in dbls.items()}}


def cubic_tables(src):
    """Cu_Interp2D / Cu_Interp3D (inc/util_interp.h): case i -> [[q, coefficient], ...] in the
    order of the source expression."""
    out = {}
    for name, key in (("Cu_Interp2D", "2"), ("Cu_Interp3D", "3")):
        start = src.index(name)
        nxt = src.find("template", start)
        body = src[start:nxt if nxt > 0 else len(src)]
        table = {}
        for m in re.finditer(r"case\s+(\d+):\s*return\s+([^;]*);", body):
            terms = re.findall(r"static_cast<ufloat_t>\(([-0-9.eE+]+)\)\*s_F\[(\d+)\]", m.group(2))
            table[int(m.group(1))] = [[int(q), float(c)] for c, q in terms]
        out[key] = table
    return out


def mrt_d3q19(vel2mom, collide):
    """Legacy D3Q19 MRT: the moment rows m_k = sum_j C f_j (solver_lbm_vel2mom_mrt_d3q19.cu:118-136)
    and, per moment, the rate symbol its collision uses (solver_lbm_collide_mrt_d3q19.cu:132-146)."""
    rows = []
    for k in range(19):
        m = re.search(r"\tm_%d = (.*);" % k, vel2mom)
        row = [0.0] * 19
        for c, j in re.findall(r"N_Pf\(([-0-9.]+)\)\*f_(\d+)", m.group(1)):
            row[int(j)] = float(c)
        rows.append(row)
    rates = {}
    for k, rate in re.findall(r"\tm_(\d+) = m_\d+ - \(?(s_\d|omeg)\)?\*", collide):
        rates[int(k)] = rate
    return {"rows": rows, "rates": rates}


def main():
    util = read("inc/util.h")
    mrt = mrt_d3q19(read("src/solver_old/solver_lbm_vel2mom_mrt_d3q19.cu"),
                    read("src/solver_old/solver_lbm_collide_mrt_d3q19.cu"))
    interp = read("inc/util_interp.h")
    init = read("src/solver_lbm/solver_lbm_init.cu")
    amr = read("src/mesh/mesh_amr.cu")
    parser = read("inc/parser.h")
    out = {
        "source": "KhodrJ/AGAL (text of inc/util.h, util_interp.h, solver_lbm_init.cu, mesh_amr.cu, parser.h)",
        "V_CONN_ID_2D": int_array(util, "V_CONN_ID_2D"),
        "V_CONN_ID_3D": int_array(util, "V_CONN_ID_3D"),
        "V_CONN_ID_PB_2D": int_array(util, "V_CONN_ID_PB_2D"),
        "V_CONN_ID_PB_3D": int_array(util, "V_CONN_ID_PB_3D"),
        "V_CONN_MAP_2D": int_array(util, "V_CONN_MAP_2D"),
        "V_CONN_MAP_3D": int_array(util, "V_CONN_MAP_3D"),
        "LBMw_D2Q9": frac_array(init, "LBMw_D2Q9"),
        "LBMw_D3Q19": frac_array(init, "LBMw_D3Q19"),
        "LBMw_D3Q27": frac_array(init, "LBMw_D3Q27"),
        "LBMpb_D2Q9": int_array(init, "LBMpb_D2Q9"),
        "LBMpb_D3Q19": int_array(init, "LBMpb_D3Q19"),
        "LBMpb_D3Q27": int_array(init, "LBMpb_D3Q27"),
        "child_connectivity": connectivity(amr),
        "ghost_conditions": ghost_conditions(amr),
        "parser_defaults": parser_defaults(parser),
        "interp_cubic": cubic_tables(interp),
        "mrt_d3q19_legacy": mrt,
    }
    assert len(mrt["rates"]) == 15, mrt["rates"]
    assert len(out["interp_cubic"]["2"]) == 16 and len(out["interp_cubic"]["3"]) == 64
    n2 = sum(len(v) for v in out["child_connectivity"]["2"].values())
    n3 = sum(len(v) for v in out["child_connectivity"]["3"].values())
    assert n2 == 9 * 4 and n3 == 27 * 8, (n2, n3)
    assert len(out["ghost_conditions"]["2"]) == 8 and len(out["ghost_conditions"]["3"]) == 26
    for k in ("LBMw_D2Q9", "LBMw_D3Q19", "LBMw_D3Q27"):
        assert sum(Fraction(a, b) for a, b in out[k]) == 1, k
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
