"""Build and run tests/cpp/hostapi_check.cpp, the test program of the header-library host API."""
from __future__ import annotations

import base64
import json
import os
import struct
import subprocess
import xml.etree.ElementTree as ET

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "hostapi_check.cpp")
LIBDIR = os.path.join(ROOT, "agal_amd", "lib")
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")


def build(outdir: str) -> str:
    exe = os.path.join(outdir, "hostapi_check")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L", LIBDIR, "-lagal", f"-Wl,-rpath,{LIBDIR}", "-ldl"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("building the host API driver failed:\n" + res.stdout + res.stderr)
    return exe


def run(exe: str, *args, timeout=600) -> dict:
    res = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    if res.returncode != 0 or not lines:
        raise RuntimeError(f"{args[0]} failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    return json.loads(lines[-1])


def read_vti_arrays(path: str) -> dict:
    """Cell arrays of one image-data piece (inline binary with UInt64 headers, or ascii)."""
    root = ET.parse(path).getroot()
    out = {}
    for da in root.iter("DataArray"):
        comps = int(da.get("NumberOfComponents", "1"))
        txt = (da.text or "").strip()
        if da.get("format") == "binary":
            raw = base64.b64decode(txt)
            (n,) = struct.unpack("<Q", raw[:8])
            vals = np.frombuffer(raw[8:8 + n], dtype="<f8")
        else:
            vals = np.array([float(x) for x in txt.split()]) if txt else np.zeros(0)
        out[da.get("Name")] = vals.reshape(-1, comps) if comps > 1 else vals
    return out
