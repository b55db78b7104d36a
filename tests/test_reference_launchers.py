"""The reference-launcher binding (include/fsr_reference_launchers.h) and the
POD layouts of the boundary, checked without a GPU.

* The shim compiles against the reference's types and links all 19 launchers of
  ReconstructionCUDA.cuh:17-248 to libfsr.so.
* Where the reference tree is present (this container), each launcher's
  parameter types, in order, equal the reference declaration's.
* Every field of fsr_sim_param, fsr_grid_info, fsr_index_range and fsr_matrix
  sits at the offset of the reference's field of the same position
  (GPU/Defines.h:18-68), in C (offsetof, compiled) and in the ctypes binding.
"""
import os
import re
import shutil
import subprocess

import pytest

from fluidsurfacereconstructiongpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "include", "fsr_reference_launchers.h")
REF_CUH = "/root/reference/FluidSurfaceReconstruction/GPU/ReconstructionCUDA.cuh"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

LAUNCHERS = [
    "bindTextures", "launchExtractionOfSurfaceParticles", "launchEstimateSurfaceVertices",
    "launchUpdateScalarGridValues", "launchCompactSurfaceVertex",
    "launchUpdateScalarGridValuesCompacted", "launchDetectValidSurfaceCubes",
    "launchCompactValidSurfaceCubes", "launchGenerateTriangles",
    "launchThrustExclusivePrefixSumScan", "launchExtractionOfSurfaceAndInvolveParticles",
    "launchEstimationOfSurfaceVertices", "launchCalculationOfMeanAndSmoothParticles",
    "launchCompactationOfSurfaceVerticesAndParticles",
    "launchCalculationOfTransformMatricesForParticles", "launchComputationOfScalarFieldGrid",
    "launchExtractionOfSurfaceParticlesForAkinci", "launchDetectionOfValidSurfaceCubesForAkinci",
    "launchSpatialGridBuilding",
]


def _param_types(text: str, name: str) -> list[str]:
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", text)
    assert m, name
    out = []
    for p in m.group(1).split(","):
        p = " ".join(p.replace("*", "* ").split())
        if not p:
            continue
        toks = p.split()
        # "Type name" -> Type; a lone token is a type
        if len(toks) >= 2 and re.fullmatch(r"[A-Za-z_]\w*", toks[-1]):
            toks = toks[:-1]
        out.append("".join(toks))
    return out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc absent")
def test_shim_compiles_and_links_every_launcher(tmp_path):
    exe = tmp_path / "shim_program"
    cmd = [HIPCC, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "tests", "shim"),
           os.path.join(ROOT, "tests", "shim", "shim_program.cpp"), "-o", str(exe),
           "-L", os.path.dirname(_lib.LIB_PATH), "-lfsr", "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    text = open(SHIM).read()
    for name in LAUNCHERS:
        assert re.search(r"\binline\s+\w+\s+" + name + r"\s*\(", text), name


@pytest.mark.skipif(not os.path.exists(REF_CUH), reason="reference tree absent")
def test_shim_signatures_equal_reference_declarations():
    ref = open(REF_CUH, encoding="latin-1").read()
    shim = open(SHIM).read()
    for name in LAUNCHERS:
        a, b = _param_types(ref, name), _param_types(shim, name)
        assert a == b, (name, a, b)


# (field, offset) in the reference's declaration order: every field is 4 bytes
# (float / uint) or a float3 / uint3 of 12 bytes (GPU/Defines.h:18-68).
LAYOUTS = {
    "fsr_sim_param": ["particle_radius", "particle_mass", "smoothing_radius",
                      "smoothing_radius_sq", "smoothing_radius_inv", "lambda_for_smoothed",
                      "anisotropic_radius", "iso_value", "sc_sp_grid_res_ratio", "expand_extent",
                      "min_num_neighbors", "spatial_cell_size_scale"],
    "fsr_grid_info": [("min_pos", 12), ("max_pos", 12), ("cell_size", 4), ("resolution", 12)],
    "fsr_index_range": ["start", "end"],
    "fsr_matrix": ["a11", "a12", "a13", "a21", "a22", "a23", "a31", "a32", "a33"],
}


def _expected(fields):
    off, out = 0, []
    for f in fields:
        name, size = f if isinstance(f, tuple) else (f, 4)
        out.append((name, off))
        off += size
    return out, off


def test_pod_field_offsets_in_c(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "fsr.h"', "int main(void) {"]
    for st, fields in LAYOUTS.items():
        src.append(f'  printf("{st} sizeof %zu\\n", sizeof({st}));')
        for name, _ in _expected(fields)[0]:
            src.append(f'  printf("{st} {name} %zu\\n", offsetof({st}, {name}));')
    src += ["  return 0;", "}"]
    c = tmp_path / "offsets.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "offsets"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(c), "-o",
                    str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        st, name, val = line.split()
        got[(st, name)] = int(val)
    for st, fields in LAYOUTS.items():
        exp, size = _expected(fields)
        assert got[(st, "sizeof")] == size, st
        for name, off in exp:
            assert got[(st, name)] == off, (st, name)


def test_pod_field_offsets_in_ctypes():
    pairs = {"fsr_sim_param": _lib.SimParam, "fsr_grid_info": _lib.GridInfo}
    for st, cls in pairs.items():
        exp, size = _expected(LAYOUTS[st])
        assert [f[0] for f in cls._fields_] == [e[0] for e in exp], st
        for name, off in exp:
            assert getattr(cls, name).offset == off, (st, name)
        import ctypes

        assert ctypes.sizeof(cls) == size
