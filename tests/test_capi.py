"""The C-ABI library builds, loads without a GPU and exports every entry
point declared in include/optimet_mi355x.h (CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

from optimet_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "optimet_mi355x.h")).read()
    return sorted(set(re.findall(r"\b(omx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declarations_match_python_list():
    assert _declared() == sorted(solver.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(solver.LIB_PATH):
        pytest.skip("library not built (run python -m optimet_amd.build)")
    lib = ctypes.CDLL(solver.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_gpu_fails_loudly_not_silently():
    """Without a device the library reports an error instead of computing."""
    if not os.path.exists(solver.LIB_PATH):
        pytest.skip("library not built")
    import subprocess, sys
    code = ("import os; os.environ['HIP_VISIBLE_DEVICES']='-1'\n"
            "from optimet_amd import solver\n"
            "try:\n    solver.Context()\n    print('CREATED')\n"
            "except solver.OmxError as e:\n    print('ERR', e.code)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert "ERR -2" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("compiler,lang", [("gcc", "c"), ("g++", "c++")])
def test_header_compiles(tmp_path, compiler, lang):
    """include/optimet_mi355x.h is plain C and C++ (what the adapter includes)."""
    import shutil
    import subprocess
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not found")
    src = tmp_path / ("t." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "optimet_mi355x.h"\nint main(void) { omx_stats s; (void)s; return OMX_OK; }\n')
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    r = subprocess.run([compiler, "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(src)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
