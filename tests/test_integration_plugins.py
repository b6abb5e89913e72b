"""The reference-side binding: the plugin sources INTEGRATION.md gives a maintainer
(a macproject3_interface projection plugin and an RCMatrix_solver_interface linear
solver plugin over the C-ABI) compile against the reference's own headers, with the
flags its build passes (-DNUM_THREAD, -DSHKZ_SUFFIX; wscript:75), so their overrides
match macproject3_interface.h:79-93 and RCMatrix_solver.h:77 as declared there.
Syntax-only: the reference's libraries (Boost, GSL) are absent, nothing is linked."""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_INCLUDE = Path("/root/reference/include")

pytestmark = pytest.mark.skipif(not (REF_INCLUDE / "shiokaze").is_dir() or shutil.which("g++") is None,
                                reason="needs the reference's headers and g++")


def _plugins():
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    return [b for b in blocks if "create_instance" in b]


def test_integration_gives_both_plugins():
    srcs = _plugins()
    assert len(srcs) == 2
    assert any("public macproject3_interface" in s and "shk_project(" in s for s in srcs)
    assert any("public RCMatrix_solver_interface" in s and "shk_solve_csr(" in s for s in srcs)


@pytest.mark.parametrize("idx", [0, 1])
def test_plugin_compiles_against_reference_headers(tmp_path, idx):
    src = tmp_path / f"plugin{idx}.cpp"
    src.write_text(_plugins()[idx])
    # every method is marked override, so a signature that does not match the reference
    # interface is an error
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Werror=return-type",
           "-DNUM_THREAD=8", "-DSHKZ_SUFFIX=", f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
