"""The C-ABI boundary, checked without a GPU.

* libpluto_amd.so loads and exports every function that include/*.h
  declares;
* the struct mirrors of include/pluto_abi.h have the reference's layout
  (sizeof/offsetof measured from Src/structs.h, tests/golden/abi_layout.json).
"""
import ctypes
import json
import os
import re
import subprocess

import pytest

from conftest import GOLDEN, ROOT

INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")


def declared_functions():
    names = set()
    for f in sorted(os.listdir(INCLUDE)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text)
        text = re.sub(r"(struct|typedef)[^{;]*\{[^}]*\}[^;]*;", "", text, flags=re.S)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{)]*\)\s*;", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("AdvanceStep", "UpdateStage", "HLLD_Solver", "States", "pgpu_run",
                 "pgpu_domain_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    # UserDefBoundary is the host program's problem-file hook, which the
    # library calls (a weak reference) and does not define
    hooks = {"UserDefBoundary"}
    missing = [n for n in declared_functions() if n not in hooks and not hasattr(lib, n)]
    assert not missing, missing
    assert "UserDefBoundary" in declared_functions()


def test_library_is_hip_gfx950():
    out = subprocess.run(["strings", LIB], capture_output=True, text=True).stdout
    assert "gfx950" in out


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "pluto_abi.h"
#define F(T, m) printf("\"%s.%s\": %zu,\n", #T, #m, offsetof(T, m))
#define S(T) printf("\"sizeof.%s\": %zu,\n", #T, sizeof(T))
int main(void){
  printf("{\n");
  S(Data); S(Grid); S(timeStep); S(State); S(Sweep); S(EMF); S(RBox);
  F(RBox,ibeg); F(RBox,vpos); F(RBox,n); F(RBox,bend);
  F(Data,Vc); F(Data,Uc); F(Data,Vs); F(Data,flag); F(Data,emf); F(Data,Dts);
  F(Data,fill);
  F(Grid,xbeg); F(Grid,x); F(Grid,xr); F(Grid,xl); F(Grid,dx); F(Grid,inv_dx);
  F(Grid,np_tot); F(Grid,np_int); F(Grid,nghost); F(Grid,lbound); F(Grid,rbound);
  F(Grid,beg); F(Grid,end); F(Grid,lbeg); F(Grid,lend); F(Grid,level); F(Grid,fill);
  F(timeStep,cmax); F(timeStep,invDt_hyp); F(timeStep,cfl); F(timeStep,Nrkl);
  F(State,v); F(State,u); F(State,flux); F(State,prs); F(State,a2); F(State,Lp);
  F(State,fill);
  F(Sweep,vn); F(Sweep,flux); F(Sweep,rhs); F(Sweep,press); F(Sweep,bn); F(Sweep,SL);
  F(Sweep,SR); F(Sweep,flag); F(Sweep,stateL); F(Sweep,stateR); F(Sweep,stateC);
  F(EMF,ezi); F(EMF,svx); F(EMF,ibeg); F(EMF,SxL); F(EMF,ex); F(EMF,dbx_dy);
  printf("\"EMF.dvx_dx\": %zu, \"EMF.dvz_dz\": %zu\n}\n", offsetof(EMF, dvx_dx), offsetof(EMF, dvz_dz));
  return 0;
}
"""


def test_struct_layout_matches_reference(tmp_path):
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text(PROBE)
    subprocess.run(["gcc", "-w", str(src), "-I" + INCLUDE, "-o", str(exe)], check=True)
    mine = json.loads(subprocess.run([str(exe)], capture_output=True, text=True,
                                     check=True).stdout)
    ref = json.load(open(os.path.join(GOLDEN, "abi_layout.json")))
    for k, v in mine.items():
        assert k in ref, k
        assert v == ref[k], "%s: mirror %d, reference %d" % (k, v, ref[k])
