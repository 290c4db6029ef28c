"""The C-ABI libraries load and export every function their headers
declare (CPU only: nothing is called that needs a GPU)."""
import ctypes as C
import os
import re

import pytest

import csparks as cs

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {
    "csparks.h": "libcsparks.so",
    "cs_host.h": "libcshost.so",
    "cs_io.h": "libcsio.so",
    "cs_model.h": "libcshost.so",
}


def declared_functions(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef struct[^;]*?\{.*?\}[^;]*;", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(cs_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", sorted(HEADERS.items()))
def test_library_exports_every_declared_function(header, lib):
    names = declared_functions(header)
    assert names, header
    dll = C.CDLL(os.path.join(cs.BUILD, lib))
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_host_struct_layouts_match_python_mirror():
    h = cs.host_lib()
    for i, st in enumerate([cs.CGeometry, cs.CSpeciesModel, cs.COptions, cs.CProblem, cs.CSetup, cs.CFluxDesc]):
        assert h.cs_host_sizeof(i) == C.sizeof(st), st.__name__


def test_error_paths_without_gpu():
    """Argument checks fire before any device call."""
    lib = cs.device_lib()
    ctx = C.c_void_p()
    assert lib.cs_create(None, C.byref(ctx)) == -1
    assert b"NULL" in lib.cs_last_error()
    assert lib.cs_step(None, 1, 1) == -1


def test_ensemble_rejects_mixed_models_without_gpu():
    """cs_batch_create checks that every realisation shares the first one's
    species model before it touches the device."""
    import dyad
    lib = cs.device_lib()
    b = C.c_void_p()
    assert lib.cs_batch_create(None, 0, C.byref(b)) == -1
    d1, d2 = dyad.small_dyad(), dyad.small_dyad(n=(16, 16, 12), h=3.0)
    o = cs.Options(seed=1, open={"ryr": [0]})
    s1, s2 = cs.Setup(d1.geometry, d1.model, o), cs.Setup(d2.geometry, d2.model, o)
    arr = (cs.CProblem * 2)(s1.s.problem, s2.s.problem)
    assert lib.cs_batch_create(arr, 2, C.byref(b)) == -1
    assert b"realisation 1" in lib.cs_last_error() and b"differs" in lib.cs_last_error()
    assert lib.cs_batch_step(None, 1, 1) == -1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 ends with status 2 and says
    why, before it touches a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU only" in r.stderr
