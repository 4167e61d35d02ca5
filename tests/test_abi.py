"""The C ABI library on the host: it loads, exports every entry point that
include/mgfvm.h declares, parses the reference's input files like the
reference, and rejects bad inputs with status codes (these paths run before
any device call, so no GPU is needed)."""
import ctypes
import os
import re

import pytest

import mgfvm
from mgfvm import _lib
from mgfvm.inputs import Problem, channel_128_inline, parse

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mgfvm.h")

ERR = {"ARG": -1, "IO": -2, "PARSE": -3, "CONFIG": -4, "GEOMETRY": -5, "LIMIT": -6,
       "UNSUPPORTED": -9}


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mgf_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_and_library_exports_everything():
    names = declared()
    assert len(names) >= 25
    L = mgfvm.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(set(mgfvm.EXPORTS)) == names


def test_adjoint_header_matches_binding():
    """include/mgfvm_adjoint.h's entry points are exported by libmgfvm.so
    and declared by the Python binding."""
    from mgfvm import adjoint
    names = _header_names("mgfvm_adjoint.h")
    assert sorted(adjoint.EXPORTS) == names
    L = mgfvm.lib()
    assert all(hasattr(L, n) for n in names)


def test_abi_version():
    assert mgfvm.lib().mgf_abi_version() == 1


def _create(prob: Problem, tmp_path, strict=False, **optkw):
    pin, ppat = prob.write(str(tmp_path))
    p = _lib.read_problem(pin, ppat)
    try:
        opt = _lib.Options()
        mgfvm.lib().mgf_options_default(ctypes.byref(opt))
        opt.strict_limits = 1 if strict else 0
        for k, v in optkw.items():
            setattr(opt, k, v)
        h = ctypes.c_void_p()
        rc = mgfvm.lib().mgf_create(p, ctypes.byref(opt), ctypes.byref(h))
        msg = mgfvm.lib().mgf_error_string().decode()
        if rc == 0:
            mgfvm.lib().mgf_destroy(h)
        return rc, msg
    finally:
        _lib.free_problem(p)


def test_problem_read_matches_writer(tmp_path):
    prob = channel_128_inline().replace(
        inlets=[(1, 4, 0.5, 0.1, 2.0)], outlets=[(2, 3, 0.7, 0.0, 0.3)], ig_inlt=2, ig_outlt=3,
        xl=2.0, sc_obs=0.25, ibc=(2, 4, 3, 1))
    pin, ppat = prob.write(str(tmp_path))
    p = _lib.read_problem(pin, ppat).contents
    assert (p.nfx, p.nfy, p.ngrid, p.ig_obs, p.nvcycle, p.nswpm, p.nswpp, p.isolve) == \
        (prob.nfx, prob.nfy, prob.ngrid, prob.ig_obs, prob.nvcycle, prob.nswpm, prob.nswpp,
         prob.isolve)
    assert (p.xl, p.yl, p.amuf, p.prsc, p.relxm, p.relxp) == \
        (prob.xl, prob.yl, prob.amuf, prob.prsc, prob.relxm, prob.relxp)
    assert tuple(p.ibc) == prob.ibc
    assert tuple(p.bc[0]) == prob.bc_xm and tuple(p.guess) == prob.guess
    assert p.nobs == len(prob.obstacles)
    assert [tuple(p.obs[4 * k + i] for i in range(4)) for k in range(p.nobs)] == prob.obstacles
    assert (p.ninlt, p.noutlt, p.sc_obs) == (1, 1, 0.25)
    assert (p.inlt_j[0], p.inlt_j[1], p.inlt_v[0], p.inlt_v[2]) == (1, 4, 0.5, 2.0)
    assert parse(pin, ppat) == prob


def test_read_errors(tmp_path):
    L = mgfvm.lib()
    out = ctypes.POINTER(_lib.ProblemStruct)()
    assert L.mgf_problem_read(b"/nonexistent/InputData", b"x", ctypes.byref(out)) == ERR["IO"]
    prob = channel_128_inline()
    pin, ppat = prob.write(str(tmp_path))
    assert L.mgf_problem_read(pin.encode(), b"/nonexistent/pattern", ctypes.byref(out)) == ERR["IO"]
    # truncated first line (V6:288-292 exits there)
    with open(pin, "w") as f:
        f.write("128 128 4\n")
    assert L.mgf_problem_read(pin.encode(), ppat.encode(), ctypes.byref(out)) == ERR["PARSE"]
    # short obstacle line (V6:2385-2389)
    prob.write(str(tmp_path))
    with open(ppat, "w") as f:
        f.write("2\n1 1 1 1\n3 3\n")
    assert L.mgf_problem_read(pin.encode(), ppat.encode(), ctypes.byref(out)) == ERR["PARSE"]


@pytest.mark.parametrize("bad,code", [
    (dict(obstacles=[(17, 17, 1, 1)]), "GEOMETRY"),          # outside the 16x16 base
    (dict(obstacles=[(0, 0, 1, 1)]), "GEOMETRY"),
    (dict(ig_obs=5), "CONFIG"),                               # > ngrid (V6:471 writes out of bounds)
    (dict(nfx=100), "GEOMETRY"),                              # levels do not nest 2:1
    (dict(ngrid=9), "CONFIG"),                                # no interior cells left
    (dict(relxm=0.0), "CONFIG"),
])
def test_create_rejects(tmp_path, bad, code):
    rc, msg = _create(channel_128_inline().replace(**bad), tmp_path)
    assert rc == ERR[code], msg
    assert msg


def test_strict_limits(tmp_path):
    many = [(i, i, j, j) for i in range(1, 17) for j in range(1, 17)] * 4   # 1024 rectangles
    rc, msg = _create(channel_128_inline().replace(obstacles=many), tmp_path, strict=True)
    assert rc == ERR["LIMIT"] and "1000" in msg


def test_batch_create_rejects_null_and_reports_member(tmp_path):
    L = mgfvm.lib()
    h = ctypes.c_void_p()
    assert L.mgf_batch_create(None, 0, None, ctypes.byref(h)) == ERR["ARG"]
    good = channel_128_inline()
    bad = good.replace(obstacles=[(40, 40, 1, 1)])
    pa = _lib.read_problem(*good.write(str(tmp_path / "a")))
    pb = _lib.read_problem(*bad.write(str(tmp_path / "b")))
    arr = (ctypes.POINTER(_lib.ProblemStruct) * 2)(pa, pb)
    rc = L.mgf_batch_create(arr, 2, None, ctypes.byref(h))
    assert rc == ERR["GEOMETRY"]
    assert L.mgf_error_string().decode().startswith("problem 1")
    _lib.free_problem(pa)
    _lib.free_problem(pb)


def test_driver_reports_input_errors(tmp_path):
    import subprocess
    exe = os.path.join(ROOT, "multigrid-fvm_amd", "mgfvm_run")
    assert os.path.exists(exe)
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 1 and "InputData" in r.stderr


def test_sensitivity_library_exports():
    """libmgfvm_tangent.so exports the whole mgfvm.h ABI and the calls of
    mgfvm_tangent.h."""
    from mgfvm import tangent
    L = tangent.tlib()
    txt = open(os.path.join(ROOT, "include", "mgfvm_tangent.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    extra = sorted(set(re.findall(r"\b(mgf_[a-z_0-9]+)\s*\(", txt)))
    assert extra == sorted(tangent.EXPORTS)
    missing = [n for n in declared() + extra if not hasattr(L, n)]
    assert not missing, missing


def _dynamic_symbols(path):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return sorted(line.split()[-1] for line in out.splitlines() if line.strip())


def _header_names(*headers):
    names = set()
    for h in headers:
        txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", h)).read(), flags=re.S)
        names.update(re.findall(r"\b(mgf_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


@pytest.mark.parametrize("lib,headers", [
    ("libmgfvm.so", ("mgfvm.h", "mgfvm_adjoint.h")),
    ("libmgfvm_tangent.so", ("mgfvm.h", "mgfvm_tangent.h")),
])
def test_library_exports_only_its_c_abi(lib, headers):
    """Each library exports exactly the C entry points its headers declare
    (no C++ launcher, set-up helper or registration symbol) and binds its own
    references locally (DF_SYMBOLIC), so the FP64 and the sensitivity
    library, which share names, can be loaded into one process without one
    calling into the other."""
    import subprocess
    path = os.path.join(ROOT, "multigrid-fvm_amd", lib)
    assert _dynamic_symbols(path) == _header_names(*headers)
    dyn = subprocess.run(["readelf", "-d", path], capture_output=True, text=True,
                         check=True).stdout
    assert "SYMBOLIC" in dyn


def test_options_defaults():
    """The defaults a caller gets: FP64, graphs, red-black, compaction on,
    automatic lanes and launch shapes."""
    opt = _lib.Options()
    mgfvm.lib().mgf_options_default(ctypes.byref(opt))
    assert (opt.precision, opt.use_graph, opt.device, opt.strict_limits) == (0, 1, -1, 0)
    assert (opt.sweeps_per_launch, opt.sweep_threads, opt.split_assembly) == (0, 0, 0)
    assert (opt.gs_order, opt.compact, opt.lanes) == (0, 1, 0)


@pytest.mark.parametrize("optkw,code", [
    (dict(gs_order=2), "ARG"),                           # neither red-black nor lexicographic
    (dict(gs_order=1, precision=1), "UNSUPPORTED"),      # the variant is FP64 only
])
def test_create_rejects_options(tmp_path, optkw, code):
    """Option errors are reported before any device work (no GPU needed)."""
    prob = channel_128_inline().replace(isolve=2)
    rc, msg = _create(prob, tmp_path, **optkw)
    assert rc == ERR[code], msg
    assert msg


def test_tuning_table_errors(tmp_path):
    """A malformed tuning table is rejected with MGF_ERR_PARSE before any
    device work."""
    prob = channel_128_inline().replace(isolve=2)
    rc, msg = _create(prob, tmp_path, tuning=b"level 130 130 7 1 512\n")
    assert rc == -3 and "tuning" in msg
    rc, msg = _create(prob, tmp_path, tuning=b"speed 11\n")
    assert rc == -3


def test_lexicographic_variant_row_limit(tmp_path):
    prob = Problem(nfx=64, nfy=2048, ngrid=2, ig_obs=2, isolve=2, obstacles=[])
    rc, msg = _create(prob, tmp_path, gs_order=1)
    assert rc == ERR["UNSUPPORTED"] and "1024" in msg
