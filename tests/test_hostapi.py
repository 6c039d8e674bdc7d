"""The header-library host API (include/agal/*.hpp) on the CPU: the input.txt parser against the
reference's own input file and parser defaults, the geometry front-end, and the whole driver
(Parser -> Mesh -> Solver_LBM -> M_AdvanceLoop with rebuilds, probes, .vthb output and restart)
running on the CPU restatement through the same C ABI (the oracle library's oracle_* entries).

* Parser: keys, defaults and `A*Nx` values of input/input.txt (copied to tests/golden/input/), the
  defaults against the table parsed from the reference's parser.h (tests/golden/
  reference_tables.json), unknown keys reported and skipped, the Nx pass before the others.
* Geometry: the reference's monkey.stl (968 facets); long-face refinement keeps the area and
  bounds every edge; the primitive list's SPHERE line gives 2 (N1-1)(N2-1) faces of area ~4 pi R^2.
* Restart: a run of 60 coarse steps equals, bit for bit, a run of 30 that writes the reference's
  restart file and a second run that loads it and continues (rebuilds on both sides).
* Output: the .vthb lists one DataSet per active block of each level with the block's AMR box;
  each .vti holds the reference's cell arrays; density and velocity are the moments of the run's
  final populations; vorticity and the Q-criterion of prescribed linear fields are exact.
"""
from __future__ import annotations

import json
import math
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from agal_amd import capi
import hostapi_util as H
import oracle_lib as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return H.build(str(tmp_path_factory.mktemp("hostapi")))


def test_parser_reads_reference_input(exe):
    r = H.run(exe, "parser", os.path.join(GOLD, "input"))
    assert r["found"] and r["invalid"] == []
    i, d, s = r["int"], r["dbl"], r["str"]
    # values written in input/input.txt, "A*Nx" expanded with Nx 64
    assert i["Nx"] == 64 and i["MAX_LEVELS_WALL"] == 5 and i["MAX_LEVELS_INTERIOR"] == 5
    assert i["N_ITER_TOTAL"] == 100 * 64 and i["N_PROBE_F_START"] == 50 * 64 and i["P_OUTPUT"] == 100 * 64
    assert i["VOL_I_MAX"] == 64 and i["N_OUTPUT_START"] == 0 and i["G_BIN_DENSITY"] == 16 and i["G_BIN_SPEC"] == 7
    assert d["v0"] == 7.8125e-5 and d["s3"] == 1.7 and d["N_REFINE_START"] == -2 and d["S_FORCEVOLUME_Xm"] == 0.1
    assert s["G_FILENAME"] == "xyzrgbdragon_ascii.stl" and s["P_DIR_NAME"] == "../out/COMPCHECK/"
    assert s["I_DIR_NAME"] == "../input/"
    # a key the file does not set keeps its default
    assert i["N_OUTPUT"] == 10


def test_parser_defaults_are_the_reference_table(exe):
    gold = json.load(open(os.path.join(GOLD, "reference_tables.json")))["parser_defaults"]
    r = H.run(exe, "text", "")
    assert r["int"] == gold["int"]
    assert r["dbl"] == gold["double"]
    assert set(r["str"]) == {"I_DIR_NAME", "P_DIR_NAME", "G_FILENAME"}
    assert r["str"]["G_FILENAME"] == "bunny_ascii.stl" and r["str"]["P_DIR_NAME"] == "../out/TEST_UPLOAD/"


def test_parser_syntax(exe):
    text = "# comment line\nN_ITER_TOTAL 3*Nx\nVOL_K_MAX 1*Nx\nNx 32\nNOT_A_KEY 5\nv0 2.5e-3\n\nP_DIR_NAME out/here/\nP_REFINE 7\n"
    r = H.run(exe, "text", text)
    assert r["int"]["Nx"] == 32
    assert r["int"]["N_ITER_TOTAL"] == 96 and r["int"]["VOL_K_MAX"] == 32  # Nx read before the other keys
    assert r["int"]["P_REFINE"] == 7 and r["dbl"]["v0"] == 2.5e-3 and r["str"]["P_DIR_NAME"] == "out/here/"
    assert r["invalid"] == ["NOT_A_KEY"]


def test_geometry_front_end(exe):
    l_spec = 0.05
    r = H.run(exe, "geometry", os.path.join(GOLD, "geometry", "monkey.stl"), os.path.join(GOLD, "input", "geometry.txt"),
              l_spec)
    assert r["stl_faces"] == 968
    assert r["stl_max_edge"] > l_spec >= r["refined_max_edge"]
    assert r["refined_faces"] > r["stl_faces"]
    assert abs(r["refined_area"] / r["stl_area"] - 1) < 1e-6
    # each face splits into 4^k pieces: the count grows by multiples of 3
    assert (r["refined_faces"] - r["stl_faces"]) % 3 == 0
    # geometry.txt: "SPHERE 64 64 0.33 0.5 0.5 7.8125e-3"
    R = 7.8125e-3
    assert r["primitive_faces"] == 2 * 63 * 63
    assert abs(r["primitive_area"] / (4 * math.pi * R * R) - 1) < 2e-3
    assert r["box_faces"] == 12 and abs(r["box_area"] - 2 * (0.4 * 0.4 + 0.4 * 0.6 + 0.4 * 0.6)) < 1e-6
    assert r["sphere_faces"] == 2 * 7 * 11 and abs(r["sphere_area"] / (4 * math.pi * 0.0625) - 1) < 0.1


def moments(f: np.ndarray, Q: int):
    c, _, _, _ = O.velset(Q)
    c = c[:, :Q].astype(np.float64)
    rho = f.sum(axis=0)
    return rho, (c @ f) / rho


def check_vthb(outdir: str, name: str, n_ids: list, f: np.ndarray | None = None, Q: int = 19):
    top = ET.parse(os.path.join(outdir, name + ".vthb")).getroot()
    assert top.get("type") == "vtkOverlappingAMR"
    amr = top.find("vtkOverlappingAMR")
    blocks = amr.findall("Block")
    assert [len(b.findall("DataSet")) for b in blocks] == [n for n in n_ids if n > 0][:len(blocks)]
    for b in blocks:
        L = int(b.get("level"))
        h = [float(x) for x in b.get("spacing").split()]
        assert abs(h[0] - (1.0 / 16) / 2 ** L) < 1e-15
        for ds in b.findall("DataSet"):
            box = [int(x) for x in ds.get("amr_box").split()]
            assert box[1] - box[0] == 3 and box[3] - box[2] == 3 and box[5] - box[4] == 3
            assert os.path.exists(os.path.join(outdir, ds.get("file")))
    if f is None:
        return
    rho, u = moments(f, Q)
    # every piece of the last level checked against the moments of the populations
    vti = ET.parse(os.path.join(outdir, blocks[0].find("DataSet").get("file"))).getroot()
    assert vti.get("type") == "ImageData"
    arrs = H.read_vti_arrays(os.path.join(outdir, blocks[0].find("DataSet").get("file")))
    assert set(arrs) == {"Ref ID", "On Boundary?", "Cell Mask", "Block Mask", "Density", "Velocity", "Vorticity",
                         "Q-Criterion"}
    # block 0 of level 0 is block ID 0 (root grid, x fastest)
    assert np.allclose(arrs["Density"], rho[:64], rtol=0, atol=1e-6)
    assert np.allclose(arrs["Velocity"], u[:, :64].T, rtol=0, atol=1e-6)


def test_loop_restart_and_vthb_on_the_restatement(exe, tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    ra = H.run(exe, "run", "oracle", H.ORACLE, a, "A", 60, 0, capi.FP32)
    rb = H.run(exe, "run", "oracle", H.ORACLE, b, "B", 30, 0, capi.FP32)
    rc = H.run(exe, "run", "oracle", H.ORACLE, b, "C", 30, 1, capi.FP32)
    assert ra["roundtrip"] and rb["roundtrip"] and rc["roundtrip"]
    assert ra["last_iter"] == 59 and rb["last_iter"] == 29 and rc["last_iter"] == 59
    assert ra["n_ids"][1] > 0 and rb["n_ids"][1] > 0  # rebuilds happened on both sides of the restart
    fa = np.fromfile(a / "A_f.bin", dtype=np.float32)
    fc = np.fromfile(b / "C_f.bin", dtype=np.float32)
    assert fa.size == 19 * ra["n_blocks"] * 64
    assert np.array_equal(fa, fc), "resumed run differs from the continuous run"
    assert (b / "restart.bin").exists()
    # the outputs at iteration 30 of both sequences are the same files
    for piece in sorted(os.listdir(a / "out_30")):
        assert (a / "out_30" / piece).read_bytes() == (b / "out_30" / piece).read_bytes(), piece
    check_vthb(str(a), "out_60", ra["n_ids"], fa.reshape(19, -1).astype(np.float64))


def test_restart_file_layout(exe, tmp_path):
    """The restart file follows the reference's layout field by field (mesh_restart.cu:8-40)."""
    r = H.run(exe, "run", "oracle", H.ORACLE, tmp_path, "A", 30, 0, capi.FP32)
    raw = (tmp_path / "restart.bin").read_bytes()
    ML, nb = 2, r["n_blocks"]
    off = 0

    def ints(n):
        nonlocal off
        v = np.frombuffer(raw, dtype="<i4", count=n, offset=off)
        off += 4 * n
        return v

    n_max = ints(1)[0]
    n_ids = ints(ML + 1)
    n_gaps = ints(1)[0]
    id_max = ints(ML + 1)
    cb = ints(1)[0]
    cells = np.frombuffer(raw, dtype="<i8", count=1, offset=off)[0]
    off += 8
    n_probed, it = ints(2)
    assert n_max == 4096 and list(n_ids[:ML]) == r["n_ids"] and n_gaps == r["n_gaps"]
    assert cb == id_max[ML] == nb and cells == nb * 64 and it == 29 and n_probed == r["probes"]
    size = (off + 4 * sum(r["n_ids"]) + 4 * n_gaps + 4 * nb * 64 + 4 * 3 * nb + 4 * nb + 4 * 2 * 27 * nb + 4 * 3 * nb
            + 4 * 19 * nb * 64 + 4 * 3 * n_probed * 64 + 4 * 4 * r["n_ids"][0] * 64)
    assert len(raw) == size


def test_output_properties_closed_form(exe, tmp_path):
    r = H.run(exe, "props", "oracle", H.ORACLE, tmp_path)
    assert r["err_vorticity"] < 1e-12 and r["err_q"] < 1e-12
    arrs = H.read_vti_arrays(str(tmp_path / "props" / "props_0_5.vti"))  # ascii pieces
    assert np.allclose(arrs["Vorticity"][:, 2], r["wz"], atol=1e-12)
    assert np.allclose(arrs["Q-Criterion"], r["q"], atol=1e-12)
