"""The C++ that INTEGRATION.md shows a maintainer adding to the reference
(StabFEM::solveFullyImplicit rewritten against include/sfem.h) compiles against the
header and links against the built libsfem_hip.so.  The reference's own members the
snippet uses are declared by a small stand-in class with the reference's names
(StabFEM.h:35-154); nothing is run (no GPU here)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STANDIN = r'''
#include <cstdio>
#include <iostream>
#include <vector>
#define MPI_COMM_WORLD 0
#define PetscPrintf(comm, ...) std::printf(__VA_ARGS__)
struct SolutionData { std::vector<double> soln; };
class StabFEM {
 public:
  int ndim = 2, npElem = 4, nNode_global = 0, nElem_global = 0, nDBC = 0, tis = 2, stepsMax = 1;
  double rhoInf = 0.0, dt = 0.1, timeFinal = 1.0;
  double elemData[11] = {1.0, 0.01, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0};
  std::vector<std::vector<double>> node_coords, DirichletBCs;
  std::vector<std::vector<int>> elemConn;
  SolutionData SolnData;
  int solveFullyImplicit();
};
'''


def _snippet():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    code = [b for b in blocks if "int StabFEM::solveFullyImplicit()" in b]
    assert len(code) == 1, "INTEGRATION.md holds one rewritten solveFullyImplicit"
    return code[0]


def test_integration_snippet_compiles_and_links(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    src = tmp_path / "snippet.cpp"
    body = _snippet().replace('#include "sfem.h"', "")
    src.write_text('#include "sfem.h"\n' + STANDIN + body +
                   "\nint main() { StabFEM f; return f.solveFullyImplicit() == 0 ? 0 : 1; }\n")
    obj = tmp_path / "snippet.o"
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Wno-unused-variable", "-I", os.path.join(ROOT, "include"),
                        "-c", str(src), "-o", str(obj)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = os.path.join(ROOT, "cfdsfemmpi_amd", "libsfem_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libsfem_hip.so not built")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    r = subprocess.run([gxx, str(obj), "-o", str(tmp_path / "snippet"), "-L", os.path.dirname(lib), "-lsfem_hip",
                        "-Wl,-rpath," + os.path.dirname(lib), "-L", os.path.join(rocm, "lib"),
                        "-Wl,-rpath-link," + os.path.join(rocm, "lib")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
