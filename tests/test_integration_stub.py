"""The CVODE-callback stub of INTEGRATION.md compiles as C against
include/nsim.h and links against the library (no GPU needed).

The SUNDIALS 2.3.0 headers are not installed here, so the few declarations
the stub uses are given by two small headers written by this test: the
scalar and vector types and the four callback types, with the signatures of
the reference's own wrappers (src/sundials_sp_stubs.c:264-267, 352-356,
473-476, 579-581).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "nmag-src_amd", "nmag")

NVECTOR_H = """
#ifndef NVECTOR_PARALLEL_STANDIN_H
#define NVECTOR_PARALLEL_STANDIN_H
typedef double realtype;
typedef int booleantype;
typedef struct { long local_length; realtype* data; } N_VectorContent_Parallel_standin;
typedef struct { N_VectorContent_Parallel_standin* content; } * N_Vector;
#define NV_DATA_P(v) ((v)->content->data)
#endif
"""

CVSPILS_H = """
#ifndef CVODE_SPILS_STANDIN_H
#define CVODE_SPILS_STANDIN_H
#include <nvector/nvector_parallel.h>
typedef int (*CVRhsFn)(realtype t, N_Vector y, N_Vector ydot, void* f_data);
typedef int (*CVSpilsJacTimesVecFn)(N_Vector v, N_Vector Jv, realtype t, N_Vector y,
                                    N_Vector fy, void* jac_data, N_Vector tmp);
typedef int (*CVSpilsPrecSetupFn)(realtype t, N_Vector y, N_Vector fy, booleantype jok,
                                  booleantype* jcurPtr, realtype gamma, void* P_data,
                                  N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);
typedef int (*CVSpilsPrecSolveFn)(realtype t, N_Vector y, N_Vector fy, N_Vector r,
                                  N_Vector z, realtype gamma, realtype delta, int lr,
                                  void* P_data, N_Vector tmp);
#endif
"""


def stub_source():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", txt, flags=re.S)
    assert blocks, "no C block in INTEGRATION.md"
    return blocks[0]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_cvode_stub_compiles_against_the_header_and_links(tmp_path):
    inc = tmp_path / "inc"
    (inc / "nvector").mkdir(parents=True)
    (inc / "cvode").mkdir(parents=True)
    (inc / "nvector" / "nvector_parallel.h").write_text(NVECTOR_H)
    (inc / "cvode" / "cvode_spils.h").write_text(CVSPILS_H)
    src = tmp_path / "nsim_cvode_stub.c"
    src.write_text(stub_source())
    obj = tmp_path / "stub.o"
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-fPIC",
                    "-I", os.path.join(ROOT, "include"), "-I", str(inc), "-c", str(src),
                    "-o", str(obj)], check=True, capture_output=True, text=True)
    lib = os.path.join(LIB_DIR, "libnsim_hip.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    so = tmp_path / "stub.so"
    subprocess.run(["gcc", "-shared", "-o", str(so), str(obj), "-L", LIB_DIR, "-lnsim_hip",
                    "-Wl,--no-undefined", "-Wl,-rpath," + LIB_DIR], check=True,
                   capture_output=True, text=True)
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(so)], check=True,
                        capture_output=True, text=True).stdout
    for name in ("nsim_rhs", "nsim_jtimes", "nsim_psetup", "nsim_psolve"):
        assert name in nm, name
