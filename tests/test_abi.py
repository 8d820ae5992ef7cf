"""The C-ABI library: it loads without a GPU, exports every entry point that
include/lbmib.h declares, validates arguments before touching the device and
refuses to run without one (no CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "lbmib.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbmib_\w+)\s*\(", text)))


def header_defines():
    out = {}
    for name, val in re.findall(r"#define\s+(LBMIB_\w+)\s+(-?\d+)", open(HEADER).read()):
        out[name] = int(val)
    return out


def test_library_exports_every_declared_function(lbmib_mod):
    lib = lbmib_mod.load()
    names = declared_functions()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(lbmib_mod.EXPORTS) == names


def test_binding_constants_match_the_header(lbmib_mod):
    d = header_defines()
    assert d["LBMIB_PREC_FP64"] == lbmib_mod.PREC_FP64
    assert d["LBMIB_PREC_FP32DEV"] == lbmib_mod.PREC_FP32DEV
    assert d["LBMIB_NTIMINGS"] == lbmib_mod.NTIMINGS
    assert d["LBMIB_E_ARG"] == -1 and d["LBMIB_E_HIP"] == -2
    assert d["LBMIB_GEOMETRY_BATCH"] == lbmib_mod.GEOMETRY_BATCH
    for k in ("AUTO", "CHAIN", "SEPARATOR"):
        assert d["LBMIB_SOLVER_" + k] == getattr(lbmib_mod, "SOLVER_" + k)


def test_fortran_module_constants_match_the_header():
    d = header_defines()
    src = open(os.path.join(ROOT, "fortran_loadfile_mitiwing_amd", "fortran", "lbmib_mod.f90")).read()
    found = dict((n, int(v)) for n, v in re.findall(r"parameter, public :: (LBMIB_\w+) = (-?\d+)", src))
    assert len(found) >= 12
    for name, val in found.items():
        assert d[name] == val, name


def test_launch_shape_struct_layout(lbmib_mod):
    """The ctypes and Fortran mirrors of lbmib_launch_shape have the header's
    field order and size (8 int32 fields)."""
    text = open(HEADER).read()
    body = text[text.index("typedef struct lbmib_launch_shape"):text.index("} lbmib_launch_shape;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"int32_t\s+(\w+)(?:\[(\d+)\])?;", body)
    names = [f for f, _ in fields]
    count = sum(int(n) if n else 1 for _, n in fields)
    assert [f[0] for f in lbmib_mod.LaunchShape._fields_] == names
    assert C.sizeof(lbmib_mod.LaunchShape) == 4 * count


def test_params_struct_layout(lbmib_mod):
    # 2 ints, 8 doubles, 9 doubles, 7 ints, 5 reserved ints
    P = lbmib_mod.Params
    assert P.ox.offset == 8
    assert P.s.offset == 8 + 8 * 8
    assert P.smooth.offset == 8 + 17 * 8
    assert C.sizeof(P) == 8 + 17 * 8 + 12 * 4


def test_dynamics_struct_layout(lbmib_mod):
    """The ctypes mirror of lbmib_dynamics has the header's field order and
    size: 4 int32 fields, then 5 doubles."""
    text = open(HEADER).read()
    body = text[text.index("typedef struct lbmib_dynamics"):text.index("} lbmib_dynamics;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(?:int32_t|double)\s+([\w, ]+);", body)
    names = [n.strip() for group in names for n in group.split(",")]
    D = lbmib_mod.DynamicsParams
    assert [f[0] for f in D._fields_] == names
    assert D.solid_den.offset == 16 and C.sizeof(D) == 16 + 5 * 8


def test_version_and_last_error(lbmib_mod):
    lib = lbmib_mod.load()
    assert lbmib_mod.version().startswith("lbmib")
    assert isinstance(lib.lbmib_last_error(), bytes)


def _params(lbmib_mod, **kw):
    from fortran_loadfile_mitiwing_amd import cases
    case = cases.config(0, n_steps=10, scale_down=4)
    p = lbmib_mod.make_params(case.params)
    for k, v in kw.items():
        setattr(p, k, v)
    return case, p


@pytest.mark.parametrize("field,value,msg", [
    ("nx", 4, "at least 8x8"),
    ("n_wings", 0, "wing_count"),
    ("n_wings", 9, "wing_count"),
    ("pts_per_wing", 3, "boundary points"),
    ("scale", 0.0, "scale"),
    ("smooth", 2, "smooth"),
])
def test_create_rejects_bad_arguments_before_touching_the_device(lbmib_mod, field, value, msg):
    lib = lbmib_mod.load()
    case, p = _params(lbmib_mod, **{field: value})
    xb = np.ascontiguousarray(case.xb0)
    yb = np.ascontiguousarray(case.yb0)
    h = C.c_void_p()
    rc = lib.lbmib_create(C.byref(p), lbmib_mod._d(xb), lbmib_mod._d(yb), C.byref(h))
    assert rc == -1
    assert msg in lib.lbmib_last_error().decode()
    assert not h.value


def test_null_arguments_are_rejected(lbmib_mod):
    lib = lbmib_mod.load()
    h = C.c_void_p()
    assert lib.lbmib_create(None, None, None, C.byref(h)) == -1
    assert lib.lbmib_step(None, 1) == -1
    assert lib.lbmib_destroy(None) in (0, -1)
    # post-processing and body dynamics: a null handle is an argument error, checked before any device call
    buf = (C.c_double * 64)()
    n = C.c_int32()
    assert lib.lbmib_vorticity(None, 0, None, None) == -1
    assert lib.lbmib_surface_pressure(None, buf, 64) == -1
    assert lib.lbmib_surface_pressure_noalloc(None, buf, 64) == -1
    assert lib.lbmib_lattice_window(None, C.byref(n), C.byref(n), C.byref(n), C.byref(n)) == -1
    assert lib.lbmib_set_dynamics(None, None) == -1
    assert lib.lbmib_get_body_state(None, buf, 64) == -1
    assert lib.lbmib_set_body_state(None, buf, 12) == -1
    assert lib.lbmib_get_motion_rows(None, buf, 1, C.byref(n)) == -1
    assert lib.lbmib_get_edge_f0(None, buf, buf) == -1
    assert lib.lbmib_set_edge_f0(None, buf, buf) == -1


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a GPU")
def test_no_device_is_an_error_not_a_fallback(lbmib_mod):
    case, _ = _params(lbmib_mod)
    with pytest.raises(lbmib_mod.LBMIBError) as e:
        lbmib_mod.LBMIB(case)
    assert e.value.code == -2
