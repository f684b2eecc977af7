"""CPU-tier checks of the drop-in boundary (no GPU calls).

* libcilregcuda.so loads and exports every function include/*.h declares,
  including the reference's five entry points and the PDTV/SB_TV aliases.
* The Python layer mirrors the reference's ccpi.filters surface: the five
  wrappers, device dispatch (KeyError for an unknown device), argument order.
* Without the reference's own CPU library next to the package, device='cpu'
  raises instead of silently falling back to anything.
"""
import ctypes
import glob
import inspect
import os
import re

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "ccpi-regularisation-toolkit_amd", "ccpi", "filters", "libcilregcuda.so")
REFERENCE_ENTRY_POINTS = ["TV_ROF_GPU_main", "TV_FGP_GPU_main", "TV_PD_GPU_main", "TV_SB_GPU_main", "TGV_GPU_main",
                          "PDTV_GPU_main", "SB_TV_GPU_main", "dTV_FGP_GPU_main", "LLT_ROF_GPU_main",
                          "NonlDiff_GPU_main", "Diffus4th_GPU_main"]


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"^CILREG_API\s+(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, re.M))
    return sorted(names)


def test_headers_declare_reference_entry_points():
    names = declared_symbols()
    for n in REFERENCE_ENTRY_POINTS:
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcilregcuda.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_per_method_headers_match_reference_names():
    # the reference's C users include these names (regularisers_GPU/*.h)
    for h in ["TV_ROF_GPU_core.h", "TV_FGP_GPU_core.h", "TV_PD_GPU_core.h", "TV_SB_GPU_core.h", "TGV_GPU_core.h",
              "dTV_FGP_GPU_core.h", "LLT_ROF_GPU_core.h", "NonlDiff_GPU_core.h", "Diffus_4thO_GPU_core.h"]:
        assert os.path.exists(os.path.join(ROOT, "include", h)), h


def test_python_surface_signatures():
    from ccpi.filters import TV
    # argument order of the reference's GPU wrappers (TV.py:477-733)
    assert list(inspect.signature(TV.TV_ROF_GPU).parameters)[:6] == [
        "inputData", "regularisation_parameter", "iterationsNumb", "marching_step_parameter", "tolerance_param",
        "gpu_device"]
    assert list(inspect.signature(TV.TV_FGP_GPU).parameters)[:7] == [
        "inputData", "lambdaPar", "iterationsNumb", "epsil", "methodTV", "nonneg", "gpu_device"]
    assert list(inspect.signature(TV.PDTV_GPU).parameters)[:10] == [
        "Input", "lambdaPar", "iter", "epsil", "lipschitz_const", "methodTV", "nonneg", "gpu_device", "Output",
        "infovector"]
    assert list(inspect.signature(TV.SB_TV_GPU).parameters)[:6] == [
        "inputData", "lambdaPar", "iterationsNumb", "epsil", "methodTV", "gpu_device"]
    assert list(inspect.signature(TV.TGV_GPU).parameters)[:8] == [
        "inputData", "lambdaPar", "alpha1", "alpha0", "iterationsNumb", "L2", "epsil", "gpu_device"]
    # TV.py:650, diffusion.py:89-90, :133
    assert list(inspect.signature(TV.LLT_ROF_GPU).parameters) == [
        "inputData", "lambdaROF", "lambdaLLT", "iterationsNumb", "tau", "epsil", "gpu_device", "out", "infovector"]
    from ccpi.filters import diffusion
    assert list(inspect.signature(diffusion.NDF_GPU).parameters) == [
        "inputData", "lambdaPar", "sigmaPar", "iterationsNumb", "tau", "penaltytype", "epsil", "gpu_device", "out",
        "infovector"]
    assert list(inspect.signature(diffusion.Diffus4th_GPU).parameters) == [
        "inputData", "lambdaPar", "sigmaPar", "iterationsNumb", "tau", "epsil", "gpu_device", "out", "infovector"]
    # TV.py:735-736
    assert list(inspect.signature(TV.dTV_FGP_GPU).parameters) == [
        "inputData", "inputRef", "lambdaPar", "iterationsNumb", "epsil", "eta", "methodTV", "nonneg", "gpu_device",
        "out", "infovector"]


def test_device_dispatch(monkeypatch):
    from ccpi.filters import regularisers as R
    calls = []

    def cpu(*a, **k):
        calls.append(("cpu", a, k))
        return "cpu"

    def gpu(*a, **k):
        calls.append(("gpu", a, k))
        return "gpu"

    w = R.create_wrapper(cpu, gpu)
    assert w(1, 2, device="cpu") == "cpu"
    assert w(1, 2, device="gpu") == "gpu" and calls[-1][2]["gpu_device"] == 0
    if R.cilregcuda is not None:
        assert w(1, 2, device=3) == "gpu" and calls[-1][2]["gpu_device"] == 3
    with pytest.raises(KeyError):
        w(1, 2, device="tpu")


def test_cpu_device_needs_the_reference_library():
    from ccpi.filters import regularisers as R
    from ccpi.filters import utils
    if utils.cilreg is not None:
        pytest.skip("a user-supplied libcilreg.so is present")
    x = np.zeros((4, 4), dtype=np.float32)
    with pytest.raises(RuntimeError):
        R.FGP_TV(x, 0.02, 5, 0.0, 0, 0, device="cpu")


def test_opts_struct_matches_header():
    from ccpi.filters.device import CilregOpts
    text = open(os.path.join(ROOT, "include", "cilreg_ext.h")).read()
    body = re.search(r"typedef struct cilreg_opts \{(.*?)\} cilreg_opts;", text, re.S).group(1)
    nints = sum(int(m.group(1)) if m.group(1) else 1
                for m in re.finditer(r"int\s+\w+(?:\[(\d+)\])?;", body))
    assert ctypes.sizeof(CilregOpts) == 4 * nints


def test_job_struct_matches_header(tmp_path):
    # the ctypes mirror of cilreg_job against the C compiler's layout
    import shutil
    import subprocess
    from ccpi.filters.batch import CilregJob
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f[0] for f in CilregJob._fields_]
    cnames = ["lambda" if f == "lambda_" else f for f in fields]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cilreg_ext.h"\nint main(void){\n'
                   '  printf("%zu\\n", sizeof(cilreg_job));\n'
                   + "".join(f'  printf("%zu\\n", offsetof(cilreg_job, {c}));\n' for c in cnames) + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(CilregJob)
    assert got[1:] == [getattr(CilregJob, f).offset for f in fields]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcilregcuda.so not built")
def test_null_pointers_rejected_before_device_work():
    """Null host or device pointers return -1 with a message, before any HIP
    call (so this runs without a GPU)."""
    lib = ctypes.CDLL(LIB)
    lib.cilreg_last_error.restype = ctypes.c_char_p
    f = ctypes.c_float
    buf = np.zeros(64, dtype=np.float32)
    info = np.zeros(2, dtype=np.float32)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    pi = info.ctypes.data_as(ctypes.c_void_p)
    lib.TV_FGP_GPU_main.argtypes = [ctypes.c_void_p] * 3 + [f, ctypes.c_int, f] + [ctypes.c_int] * 6
    for args, what in [((None, p, pi), b"input"), ((p, None, pi), b"output"), ((p, p, None), b"infovector")]:
        assert lib.TV_FGP_GPU_main(*args, f(0.02), 5, f(0.0), 0, 0, 0, 8, 8, 1) == -1
        assert what in lib.cilreg_last_error()
    lib.dTV_FGP_GPU_main.argtypes = [ctypes.c_void_p] * 4 + [f, ctypes.c_int, f, f] + [ctypes.c_int] * 6
    assert lib.dTV_FGP_GPU_main(p, None, p, pi, f(0.02), 5, f(0.0), f(0.01), 0, 0, 0, 8, 8, 1) == -1
    assert b"reference" in lib.cilreg_last_error()
    lib.cilreg_fgp_dev.argtypes = [ctypes.c_void_p] * 3 + [f, ctypes.c_int, f] + [ctypes.c_int] * 5 + \
        [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.cilreg_fgp_dev(None, p, pi, f(0.02), 5, f(0.0), 0, 0, 8, 8, 1, None, None) == -1
    assert b"null device pointer" in lib.cilreg_last_error()
