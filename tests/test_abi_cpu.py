"""The C ABI without a GPU: the library loads, exports every function that
include/*.h declares, and the ctypes mirror matches the C struct layouts."""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

from fvmic2d_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in os.listdir(INC):
        if f.endswith(".h"):
            txt = open(os.path.join(INC, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"^\s*(?:int|void|size_t|const char \*|void \*)\s*\**\s*(fvm_\w+)\s*\(",
                                    txt, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    h = C.CDLL(lib.LIB_PATH)
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing
    # and the ctypes binding covers all of them
    assert set(names) <= set(lib.exported_symbols()), set(names) - set(lib.exported_symbols())


COMPAT_NAMES = ["findMarkerCells", "projectMarkersNodesAll2", "projectMarkerFieldToNodes",
                "projectMarkersNodesFromScalar", "projectNodalFieldToMarkersS", "advectMarkersRK",
                "checkPlasticYielding", "limitViscosity", "subgridTemperatureChanges",
                "subgridStressChanges", "updateMarkerStress", "updateMarkerStrainPressure",
                "updateGlobalStrainRateResidual", "getGlobalStrainRateResidual", "checkConvergence"]


def test_reference_named_wrappers_are_exported():
    txt = open(os.path.join(INC, "fvm_compat.h")).read()
    for n in COMPAT_NAMES:
        assert re.search(r"\b%s\s*\(" % n, txt), n
    h = C.CDLL(lib.LIB_PATH)
    assert all(hasattr(h, n) for n in COMPAT_NAMES)


def test_binding_loads_without_gpu():
    L = lib.load()
    assert L.fvm_last_error(None) == b"no context"


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "fvm_abi.h"
#define P(s, f) printf("\"%s.%s\": %zu,", #s, #f, offsetof(s, f))
int main(void) {
  printf("{");
  printf("\"Marker\": %zu, \"Materials\": %zu, \"Options\": %zu, \"BoundaryValues\": %zu,",
         sizeof(Marker), sizeof(Materials), sizeof(Options), sizeof(BoundaryValues));
  printf("\"GridData\": %zu, \"NodalArrays\": %zu, \"fvm_markers\": %zu, \"fvm_grid\": %zu,"
         "\"fvm_projection\": %zu, \"MarkerSet\": %zu,",
         sizeof(GridData), sizeof(NodalArrays), sizeof(fvm_markers), sizeof(fvm_grid),
         sizeof(fvm_projection), sizeof(MarkerSet));
  P(Marker, T); P(Marker, s); P(Marker, e); P(Marker, E); P(Marker, Eii); P(Marker, cellX);
  P(Marker, strainRateResidual); P(Marker, isYielding);
  P(Materials, hasPlasticity); P(Materials, gamma); P(Materials, hasEtaT); P(Materials, Tref);
  P(Options, gy); P(Options, mechBCTop); P(Options, fractionalEtamin); P(Options, subgridStressDiffusivity);
  P(Options, slabAngle); P(Options, decoupleEtaScaler);
  P(fvm_grid, x); P(fvm_markers, X); P(fvm_markers, Mat); P(fvm_projection, stagx);
  printf("\"end\": 0}");
  return 0;
}
"""


def test_struct_layouts_match_ctypes(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I" + INC, str(src), "-o", str(exe)], check=True)
    got = json.loads(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert got["Marker"] == 216 == C.sizeof(abi.Marker)
    for name, cls in (("Materials", abi.Materials), ("Options", abi.Options),
                      ("BoundaryValues", abi.BoundaryValues), ("GridData", abi.GridData),
                      ("NodalArrays", abi.NodalArrays), ("MarkerSet", abi.MarkerSet),
                      ("fvm_markers", lib.fvm_markers), ("fvm_grid", lib.fvm_grid),
                      ("fvm_projection", lib.fvm_projection)):
        assert got[name] == C.sizeof(cls), name
    for key, off in got.items():
        if "." not in key:
            continue
        s, f = key.split(".")
        cls = {"Marker": abi.Marker, "Materials": abi.Materials, "Options": abi.Options,
               "fvm_grid": lib.fvm_grid, "fvm_markers": lib.fvm_markers,
               "fvm_projection": lib.fvm_projection}[s]
        assert getattr(cls, f).offset == off, key
    # numpy AoS dtype agrees with the C struct
    assert abi.MARKER_DTYPE.itemsize == 216
    assert abi.FIELD_OFFSET["cellX"] == got["Marker.cellX"]
    assert abi.FIELD_OFFSET["Eii"] == got["Marker.Eii"]
