#!/usr/bin/env python3
"""Compile the reference's own hot-path routines into oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY.  The Fortran sources are compiled where they lie
under /root/reference (nothing is copied); objects, module files and the
shared library go to oracle/_ref/ (git-ignored).  Flags follow the reference
release build for the DP version: -DMYREAL8, -O2, -ffp-contract=off
(engine/CMake_Compilers/cmake_linux64_gf.txt:134,163), compiled with ROCm's
amdflang because gfortran is absent from the image.

Modules that the leaves USE are compiled on demand (a module missing at
compile time is located by its MODULE statement and compiled first).  Routines
outside the restated branch (message printing, other formulations) are not
compiled: the library leaves those symbols unresolved, it is loaded with lazy
binding, and no test drives a branch that would call them.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "obj")
FC = "/opt/rocm/llvm/bin/amdflang"
INC = [f"-I{REF}/engine/share/includes", f"-I{REF}/engine/share/spe_inc", f"-I{REF}/engine/share/r8",
       f"-I{REF}/common_source/includes", f"-I{REF}/engine/share/message_module", f"-I{OBJ}"]
FLAGS = ["-cpp", "-DMYREAL8", "-DCPP_mach=CPP_p4linux964", "-DCPP_rel=00", "-ffixed-line-length=132",
         "-O2", "-ffp-contract=off", "-fPIC", f"-J{OBJ}"]

LEAVES = [
    "engine/source/elements/solid/solide/sderi3.F",
    "engine/source/elements/solid/solide4/schkjabt3.F",
    "engine/source/output/message/message.F",  # ANCMSG, called by SCHKJABT3 on a negative Jacobian
    "engine/source/elements/solid/solide/sdlen3.F",
    "engine/source/elements/solid/solide/slen.F",
    "engine/source/elements/solid/solide/sdefo3.F",
    "engine/source/elements/solid/solide/srho3.F",
    "engine/source/elements/solid/solide/srota3.F",
    "engine/source/elements/solid/solide/smallb3.F",
    "engine/source/elements/solid/solide/shvis3.F",
    "engine/source/elements/solid/solide/sfint3.F",
    "engine/source/elements/solid/solide/scumu3p.F",
    "engine/source/materials/mat/mat002/m2law.F",
    "engine/source/materials/mat_share/mqviscb.F",
    "engine/source/materials/mat_share/mstrain_rate.F",
    "engine/source/assembly/accele.F",
    "engine/source/assembly/velocity.F",
    "engine/source/assembly/displacement.F",
    "engine/source/elements/shell/coque/ccoor3.F",
    "engine/source/elements/shell/coque/cnvec3.F",
    "engine/source/elements/shell/coque/cortdir3.F",
    "engine/source/elements/shell/coque/cderi3.F",
    "engine/source/elements/shell/coque/ccoef3.F",
    "engine/source/elements/shell/coque/cdlen3.F",
    "engine/source/elements/shell/coque/cdefo3.F",
    "engine/source/elements/shell/coque/ccurv3.F",
    "engine/source/elements/shell/coque/cstra3.F",
    "engine/source/elements/shell/coque/chvis3.F",
    "engine/source/elements/shell/coque/cdt3.F",
    "engine/source/elements/shell/coque/cfint3.F",
    "engine/source/elements/shell/coque/cupdt3.F",
    "engine/source/elements/shell/coqini.F",
    "engine/source/materials/mat/mat002/sigeps02c.F",
    "engine/source/materials/mat/mat002/m2cplr.F",
    "engine/source/elements/solid/solide/sbilan.F",
    "engine/source/elements/shell/coque/cbilan.F",
    "engine/source/tools/sensor/sensor_energy_bilan.F",
]
DRIVERS = ["ref_driver.F", "ref_shell.F", "ref_bilan.F"]

_MOD = re.compile(r"^\s*module\s+([a-z0-9_]+)\s*(!.*)?$", re.I)


def module_index():
    idx = {}
    for root in (f"{REF}/common_source", f"{REF}/engine/source", f"{REF}/engine/share"):
        for d, _, fs in os.walk(root):
            for f in fs:
                if f.endswith((".F", ".F90")):
                    p = os.path.join(d, f)
                    with open(p, errors="ignore") as fh:
                        for line in fh:
                            m = _MOD.match(line)
                            if m and m.group(1).lower() != "procedure":
                                idx.setdefault(m.group(1).lower(), p)
    return idx


def main():
    if not os.path.isdir(REF):
        print("reference tree absent: oracle/_ref not built")
        return 0
    if os.path.isdir(OBJ):
        shutil.rmtree(OBJ)
    os.makedirs(OBJ, exist_ok=True)
    idx = module_index()
    done: dict[str, str] = {}

    def compile_one(path: str, depth: int = 0) -> bool:
        if path in done:
            return True
        obj = os.path.join(OBJ, os.path.relpath(path, REF).replace("/", "__") + ".o")
        for _ in range(60):
            r = subprocess.run([FC, *FLAGS, *INC, "-c", path, "-o", obj], capture_output=True, text=True)
            if r.returncode == 0:
                done[path] = obj
                return True
            m = re.search(r"module '([a-z0-9_]+)'", r.stderr, re.I)
            dep = idx.get(m.group(1).lower()) if m else None
            if dep and dep != path and depth < 40 and dep not in done:
                if not compile_one(dep, depth + 1):
                    print(f"FAIL dependency {dep} of {path}")
                    return False
                continue
            print(f"FAIL {path}\n{r.stderr[-2000:]}")
            return False
        return False

    ok = all(compile_one(os.path.join(REF, leaf)) for leaf in LEAVES)
    if not ok:
        return 1
    drv_objs = []
    for name in DRIVERS:
        drv = os.path.join(HERE, name)
        drv_obj = os.path.join(OBJ, name.replace(".F", ".o"))
        for _ in range(40):
            r = subprocess.run([FC, *FLAGS, *INC, "-c", drv, "-o", drv_obj], capture_output=True, text=True)
            if r.returncode == 0:
                break
            m = re.search(r"module '([a-z0-9_]+)'", r.stderr, re.I)
            dep = idx.get(m.group(1).lower()) if m else None
            if not (dep and compile_one(dep)):
                print(r.stderr[-3000:])
                return 1
        drv_objs.append(drv_obj)
    lib = os.path.join(OUT, "libref.so")
    r = subprocess.run([FC, "-shared", "-o", lib, *drv_objs, *done.values()], capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stderr[-3000:])
        return 1
    with open(os.path.join(OUT, "MANIFEST.txt"), "w") as fh:
        ver = subprocess.run([FC, "--version"], capture_output=True, text=True).stdout.splitlines()[0]
        fh.write(f"compiler: {ver}\nflags: {' '.join(FLAGS)}\n")
        for p in sorted(done):
            fh.write(os.path.relpath(p, REF) + "\n")
    print(f"oracle/_ref/libref.so built from {len(done)} reference files")
    return 0


if __name__ == "__main__":
    sys.exit(main())
