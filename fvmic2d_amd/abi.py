"""ctypes / numpy mirror of include/structures.h and include/fvm_abi.h.

The layouts are the reference's own (fdcode.h:13-238): a Marker is 216 bytes
with the field offsets asserted below, Materials/Options/BoundaryValues keep
the reference field order.  Nothing here touches a GPU.
"""
import ctypes as C

import numpy as np

MAXMAT = 8

# Projection methods, markerProjection.h:3-8
ARITHMETIC, ARITHMETIC_LOCAL, HARMONIC, GEOMETRIC = 0, 1, 2, 3


class Tensor33s(C.Structure):
    _fields_ = [("T11", C.c_double), ("T22", C.c_double), ("T12", C.c_double)]


class Marker(C.Structure):
    """fdcode.h:97-129 (no TEXTURE)."""
    _fields_ = [
        ("X", C.c_double), ("Y", C.c_double), ("VX", C.c_double), ("VY", C.c_double),
        ("Mat", C.c_int32), ("T", C.c_double), ("Tdot", C.c_double), ("s", Tensor33s),
        ("eta", C.c_double), ("mu", C.c_double), ("rho", C.c_double), ("rhodot", C.c_double),
        ("p", C.c_double), ("e", Tensor33s), ("E", Tensor33s), ("Eii", C.c_double),
        ("wxy", C.c_double), ("cellX", C.c_int32), ("cellY", C.c_int32), ("cpu", C.c_int32),
        ("strainRateResidual", C.c_double), ("isYielding", C.c_int32),
    ]


assert C.sizeof(Marker) == 216

# numpy view of the AoS block: same offsets as Marker
MARKER_DTYPE = np.dtype({
    "names": ["X", "Y", "VX", "VY", "Mat", "T", "Tdot", "sxx", "syy", "sxy", "eta", "mu", "rho",
              "rhodot", "p", "exx", "eyy", "exy", "Exx", "Eyy", "Exy", "Eii", "wxy", "cellX",
              "cellY", "cpu", "strainRateResidual", "isYielding"],
    "formats": ["f8", "f8", "f8", "f8", "i4", "f8", "f8", "f8", "f8", "f8", "f8", "f8", "f8",
                "f8", "f8", "f8", "f8", "f8", "f8", "f8", "f8", "f8", "f8", "i4", "i4", "i4",
                "f8", "i4"],
    "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144,
                152, 160, 168, 176, 184, 188, 192, 200, 208],
    "itemsize": 216,
})

# byte offsets used to name a marker field by pointer (markerProjection.c:99-102)
FIELD_OFFSET = {n: MARKER_DTYPE.fields[n][1] for n in MARKER_DTYPE.names}
assert Marker.X.offset == 0 and Marker.cellX.offset == 184 and Marker.isYielding.offset == 208
assert Marker.s.offset == 56 and Marker.e.offset == 120 and Marker.E.offset == 144

# Device SoA column order (fvm_abi.h FVM_COL_*): doubles first, then int32
DCOLS = ["X", "Y", "VX", "VY", "T", "Tdot", "sxx", "syy", "sxy", "eta", "mu", "rho", "rhodot",
         "p", "exx", "eyy", "exy", "Exx", "Eyy", "Exy", "Eii", "wxy", "strainRateResidual"]
ICOLS = ["Mat", "cellX", "cellY", "cpu", "isYielding"]


class MarkerSet(C.Structure):
    _fields_ = [("nMark", C.c_int32), ("maxMark", C.c_int32), ("nOut", C.c_int32),
                ("markers", C.c_void_p)]


class GridData(C.Structure):
    _fields_ = [("da", C.c_void_p), ("vda", C.c_void_p), ("LX", C.c_double), ("LY", C.c_double),
                ("NX", C.c_int32), ("NY", C.c_int32), ("mp", C.c_int32), ("np", C.c_int32),
                ("x", C.c_void_p), ("y", C.c_void_p), ("xc", C.c_void_p), ("yc", C.c_void_p),
                ("xperiodic", C.c_int32), ("yperiodic", C.c_int32)]


class Materials(C.Structure):
    _fields_ = [
        ("nMaterials", C.c_int32),
        ("materialEta", C.c_double * MAXMAT), ("materialRho", C.c_double * MAXMAT),
        ("materialkThermal", C.c_double * MAXMAT), ("materialCp", C.c_double * MAXMAT),
        ("materialAlpha", C.c_double * MAXMAT), ("materialMu", C.c_double * MAXMAT),
        ("materialFriction", C.c_double * MAXMAT), ("materialCohesion", C.c_double * MAXMAT),
        ("hasPlasticity", C.c_int32 * MAXMAT),
        ("gamma", (C.c_double * 2) * MAXMAT), ("C", (C.c_double * 2) * MAXMAT),
        ("F", (C.c_double * 2) * MAXMAT),
        ("hasEtaT", C.c_int32 * MAXMAT), ("QonR", C.c_double * MAXMAT),
        ("Tref", C.c_double * MAXMAT),
    ]


class BCInfo(C.Structure):
    _fields_ = [("type", C.c_int32 * 3), ("value", C.c_double * 3)]


class BoundaryValues(C.Structure):
    _fields_ = [("mechBCLeft", BCInfo), ("mechBCRight", BCInfo), ("mechBCTop", BCInfo),
                ("mechBCBottom", BCInfo)]


class Options(C.Structure):
    _fields_ = [
        ("NX", C.c_int32), ("NY", C.c_int32), ("LX", C.c_double), ("LY", C.c_double),
        ("NMX", C.c_int32), ("NMY", C.c_int32), ("gridRatio", C.c_double), ("nTime", C.c_int32),
        ("restartStep", C.c_int32), ("totalTime", C.c_double), ("maxNumPlasticity", C.c_int32),
        ("plasticDelay", C.c_int32), ("Tperturb", C.c_double), ("shearHeating", C.c_int32),
        ("adiabaticHeating", C.c_int32), ("gy", C.c_double), ("gx", C.c_double), ("gz", C.c_double),
        ("mechBCLeft", BCInfo), ("mechBCRight", BCInfo), ("mechBCTop", BCInfo),
        ("mechBCBottom", BCInfo), ("thermalBCLeft", BCInfo), ("thermalBCRight", BCInfo),
        ("thermalBCTop", BCInfo), ("thermalBCBottom", BCInfo),
        ("fractionalEtamin", C.c_double), ("etamin", C.c_double), ("etamax", C.c_double),
        ("dtMax", C.c_double), ("displacementStepLimit", C.c_double), ("maxTempChange", C.c_double),
        ("saveInterval", C.c_int32), ("markerSaveInterval", C.c_int32),
        ("subgridStressDiffusivity", C.c_double), ("subgridTemperatureDiffusivity", C.c_double),
        ("maxMarkFactor", C.c_double), ("minMarkers", C.c_int32), ("maxMarkers", C.c_int32),
        ("grainSize", C.c_double), ("overridingPlateThickness", C.c_double),
        ("slabAge", C.c_double), ("slabAngle", C.c_double), ("slabVelocity", C.c_double),
        ("rootThickness", C.c_double), ("rootCenter", C.c_double), ("rootWidth", C.c_double),
        ("staticVelocity", C.c_int32), ("minDecoupleDepth", C.c_double),
        ("maxDecoupleDepth", C.c_double), ("decoupleThickness", C.c_double),
        ("decoupleEtaScaler", C.c_double),
    ]


NODAL_NAMES = ["rho", "rhodot", "etaS", "etaN", "thisT", "lastT", "kThermal", "Cp", "edotxx",
               "edotxy", "edotyy", "soxx", "soxy", "soyy", "dsxx", "dsyy", "dsxy", "muS", "muN",
               "vx", "vy", "p", "wxy", "strainRateResidual", "eii", "sii"]


class NodalArrays(C.Structure):
    """structures.h NodalArrays: one pointer per nodal field (host or device)."""
    _fields_ = [(n, C.c_void_p) for n in NODAL_NAMES]


def default_options():
    """options.c:146-229 defaults for the fields the marker path reads."""
    o = Options()
    o.fractionalEtamin = 1e-4
    o.etamin = 1.0e14
    o.etamax = 1.0e25
    o.subgridStressDiffusivity = 1.0
    o.subgridTemperatureDiffusivity = 1.0
    o.slabAngle = 45.0 / 180.0 * np.pi  # options.c:263 converts degrees to radians
    o.minDecoupleDepth = 5.1e4
    o.maxDecoupleDepth = 8.0e4
    o.decoupleThickness = 141.0
    o.decoupleEtaScaler = 1.0
    return o


def default_materials():
    """options.c:195-215 defaults (material 0)."""
    m = Materials()
    m.nMaterials = 1
    m.materialEta[0] = 1.0e21
    m.materialRho[0] = 4000.0
    m.materialkThermal[0] = 4.0
    m.materialCp[0] = 1250.0
    m.materialAlpha[0] = 0.0
    m.materialMu[0] = 1.0e99
    m.materialCohesion[0] = 1.0e99
    m.materialFriction[0] = 1.0
    m.Tref[0] = 273.0
    for i in range(MAXMAT):
        m.C[i][0] = m.C[i][1] = 1.0e99
        m.F[i][0] = m.F[i][1] = 1.0e99
        m.gamma[i][0] = 1e99
        m.gamma[i][1] = 1e98
    return m
