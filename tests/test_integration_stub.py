"""The reference-side binding classes in INTEGRATION.md compile against the
reference's own headers (syntax-only; skipped where the reference tree is absent)."""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = "/root/reference"


def _blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return [b for b in re.findall(r"```cpp\n(.*?)```", text, flags=re.S) if "#pragma once" in b]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
@pytest.mark.parametrize("idx", [0, 1])
def test_binding_compiles_against_reference_headers(idx):
    blocks = _blocks()
    assert len(blocks) == 2
    inc = ["src", "src/Energy", "src/Utils", "src/LinSysSolver", "ext/libigl/include",
           "ext/libigl/external/eigen"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "binding.cpp")
        with open(src, "w") as f:
            f.write(blocks[idx].replace("#pragma once", ""))
            if idx == 1:
                f.write("\ntemplate class OptCuts::GPULinSysSolver<Eigen::VectorXi, Eigen::VectorXd>;\n")
            else:
                f.write("\nstatic OptCuts::GPUSymDirichletEnergy* make() { return new OptCuts::GPUSymDirichletEnergy(); }\n")
        cmd = ["g++", "-std=c++14", "-fsyntax-only", "-w", "-I", os.path.join(ROOT, "include")]
        for i in inc:
            cmd += ["-I", os.path.join(REF, i)]
        r = subprocess.run(cmd + [src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
