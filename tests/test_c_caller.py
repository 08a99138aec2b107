"""The C ABI from a C caller, as the reference's main() (or a cgo / JNI binding)
would use it: a C99 program compiled with gcc against include/turbchannel.h and
linked against the built library calls the entry points that need no device
and prints what they return; the values are checked against the oracle. The
library must be built (turbchannel_amd/build.py)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "turbchannel_amd")

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "turbchannel.h"

int main(void) {
    int ns = 0, type = -1;
    double be[42], bi[42], d[6], atm[2];
    int rc = tc_imex_table(TC_IMEX_ARS_343, &ns, &type, be, bi, d, atm);
    printf("imex %d %d %d\n", rc, ns, type);
    printf("E");
    for (int k = 0; k < ns * (ns + 1); k++) printf(" %.17g", be[k]);
    printf("\nI");
    for (int k = 0; k < ns * (ns + 1); k++) printf(" %.17g", bi[k]);
    printf("\nparams %d\n", tc_mlp_param_count());
    double zz[17];
    for (int i = 0; i <= 16; i++) zz[i] = ZZ_VALUE;
    tc_params P;
    memset(&P, 0, sizeof P);
    P.scheme = TC_SCHEME_VV;
    P.precision = TC_FP64;
    P.nx = 16; P.ny = 12; P.nz = 16;
    P.visc = 1.0 / 5600.0; P.hx = HX; P.hy = HY;
    P.zz = zz;
    double dz[16];
    int n = tc_mesh_array(&P, "DZ", dz, 16);
    printf("DZ %d", n);
    for (int i = 0; i < n; i++) printf(" %.17g", dz[i]);
    printf("\n");
    tc_engine* e = NULL;
    int r2 = tc_create(NULL, 0, &e);
    printf("create_null %d %s\n", r2, tc_last_error());
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_c_program_calls_the_abi(tmp_path):
    lib = os.path.join(LIBDIR, "libturbchannel_amd.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    zz = cases.tanh_faces(16, 1.8)
    hx, hy = 2 * math.pi / 16, math.pi / 12
    # the faces as a C expression of i (the same doubles as zz)
    zexpr = "(1.0 + tanh(1.8 * (2.0 * i / 16 - 1.0)) / tanh(1.8))"
    src = PROGRAM.replace("ZZ_VALUE", f"(i == 0 ? 0.0 : i == 16 ? 2.0 : {zexpr})")
    src = src.replace("HX", repr(hx)).replace("HY", repr(hy))
    src = "#include <math.h>\n" + src
    c = tmp_path / "caller.c"
    c.write_text(src)
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), str(c),
                    "-o", str(exe), "-L", LIBDIR, "-lturbchannel_amd", f"-Wl,-rpath,{LIBDIR}", "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout.splitlines()
    rows = {ln.split()[0]: ln.split()[1:] for ln in out}
    # IMEX tables: the oracle's restatement of collocated_SRK/imex.cpp
    rc, ns, typ = map(int, rows["imex"])
    ref = oracle.imex_table(2)  # ARS(3,4,3)
    assert rc == 0 and ns == ref["s"] and typ == ref["type"]
    np.testing.assert_allclose(np.array(rows["E"], float), ref["E"].ravel(), rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.array(rows["I"], float), ref["I"].ravel(), rtol=0, atol=1e-15)
    assert int(rows["params"][0]) == 9414
    # mesh: the oracle's DZ of the same faces
    n = int(rows["DZ"][0])
    sol = oracle.Solver(1, 16, 12, 16, zz, visc=1.0 / 5600.0, hx=hx, hy=hy)
    np.testing.assert_allclose(np.array(rows["DZ"][1:], float), sol.mesh("DZ")[:n], rtol=1e-15, atol=0)
    sol.close()
    # errors come back as a status and a message, nothing exits
    assert rows["create_null"][0] == "1" and "NULL" in " ".join(rows["create_null"][1:])
