"""The C-ABI library loads (no GPU needed to load it) and exports every symbol the header
declares; the oracle exports the same set under its own prefix."""
import ctypes
import os
import re

import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "include", "endobeams_hip.h")
LEARNED = os.path.join(REPO, "include", "endobeams_learned.h")
LIB = os.path.join(REPO, "endobeams.jl_amd", "lib", "libendobeams_hip.so")
ORACLE = os.path.join(REPO, "oracle", "liboracle.so")


def declared(path=HEADER):
    text = open(path).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(eb_\w+)\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("eb_create", "eb_assemble", "eb_assemble_contact", "eb_corrector", "eb_predictor",
              "eb_solve_free_dofs", "eb_update_converged"):
        assert n in names
    assert len(names) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() + declared(LEARNED) if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_the_same_abi():
    import oracle as O
    lib = O.load()
    missing = [n for n in declared() if not hasattr(lib, "ebo_" + n[3:])]
    assert not missing, missing


def test_python_bindings_cover_the_abi():
    import sys
    sys.path.insert(0, os.path.join(REPO, "endobeams.jl_amd"))
    from endobeams_amd import abi
    bound = {"eb_" + k for k in abi.PROTOTYPES}
    assert bound == set(declared())
    assert {"eb_" + k for k in abi.LEARNED_PROTOTYPES} == set(declared(LEARNED))
    assert len(declared(LEARNED)) == 15


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_a_second_build_in_one_process_is_refused(tmp_path):
    # two builds export the same kernel and launcher symbols; the loader refuses the second path
    # instead of silently binding it to the first (run in a child so this process stays clean)
    import shutil
    import subprocess
    import sys
    other = tmp_path / "libendobeams_hip.so"
    shutil.copy(LIB, other)
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from endobeams_amd.engine import load_hip_library, EBError\n"
            "a = load_hip_library(); assert load_hip_library() is a\n"
            "try:\n    load_hip_library(%r)\nexcept EBError as e:\n    print('refused', e)\n") % (
        os.path.join(REPO, "endobeams.jl_amd"), str(other))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "refused" in out.stdout
