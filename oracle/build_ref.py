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

# The Starter's own element initialisation routines that fix the Engine's inputs on this path:
# the solid volume (SDERI3), the solid corner masses (SMASS3), the shell area (CEVECI), the shell
# mass and inertia (CINMAS) and the nodal sums in element-id order (SPMD_MSIN, with the C sort
# MY_ORDERS it calls).  They build a second library with the Starter's include tree and modules.
STARTER_LEAVES = [
    "starter/source/elements/solid/solide/sderi3.F",
    "starter/source/elements/solid/solide/smass3.F",
    "starter/source/elements/shell/coque/ceveci.F",
    "starter/source/elements/shell/coque/cinmas.F",
    "starter/source/elements/initia/spmd_msin.F",
]
STARTER_C = ["common_source/tools/sort/my_orders.c", "common_source/tools/memory/cpp_report_alloc.cpp"]
STARTER_DRIVERS = ["ref_starter.F"]
STARTER_INC_DIRS = ["starter/share/includes", "starter/share/spe_inc", "starter/share/r8", "common_source/includes",
                    "starter/share/message_module"]

_MOD = re.compile(r"^\s*module\s+([a-z0-9_]+)\s*(!.*)?$", re.I)


def module_index(roots):
    idx = {}
    for root in roots:
        for d, _, fs in os.walk(os.path.join(REF, root)):
            for f in fs:
                if f.endswith((".F", ".F90")):
                    p = os.path.join(d, f)
                    with open(p, errors="ignore") as fh:
                        for line in fh:
                            m = _MOD.match(line)
                            if m and m.group(1).lower() != "procedure":
                                idx.setdefault(m.group(1).lower(), p)
    return idx


def build_library(name, leaves, drivers, roots, inc_dirs, objdir, c_sources=()):
    """Compile `leaves` (and the modules they use, found under `roots`) and our `drivers` into
    oracle/_ref/<name>; returns the number of reference files compiled, or -1."""
    if os.path.isdir(objdir):
        shutil.rmtree(objdir)
    os.makedirs(objdir, exist_ok=True)
    inc = [f"-I{REF}/{d}" for d in inc_dirs] + [f"-I{objdir}"]
    flags = [f if not f.startswith("-J") else f"-J{objdir}" for f in FLAGS]
    idx = module_index(roots)
    done: dict[str, str] = {}

    def compile_one(path: str, depth: int = 0) -> bool:
        if path in done:
            return True
        obj = os.path.join(objdir, os.path.relpath(path, REF).replace("/", "__") + ".o")
        for _ in range(60):
            r = subprocess.run([FC, *flags, *inc, "-c", path, "-o", obj], capture_output=True, text=True)
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

    if not all(compile_one(os.path.join(REF, leaf)) for leaf in leaves):
        return -1
    objs = list(done.values())
    for c in c_sources:
        obj = os.path.join(objdir, c.replace("/", "__") + ".o")
        cc = "g++" if c.endswith(".cpp") else "gcc"
        r = subprocess.run([cc, "-O2", "-fPIC", "-c", os.path.join(REF, c), "-o", obj], capture_output=True,
                           text=True)
        if r.returncode != 0:
            print(r.stderr[-2000:])
            return -1
        objs.append(obj)
        done[os.path.join(REF, c)] = obj
    drv_objs = []
    for dname in drivers:
        drv = os.path.join(HERE, dname)
        drv_obj = os.path.join(objdir, dname.replace(".F", ".o"))
        for _ in range(40):
            r = subprocess.run([FC, *flags, *inc, "-c", drv, "-o", drv_obj], capture_output=True, text=True)
            if r.returncode == 0:
                break
            m = re.search(r"module '([a-z0-9_]+)'", r.stderr, re.I)
            dep = idx.get(m.group(1).lower()) if m else None
            if not (dep and compile_one(dep)):
                print(r.stderr[-3000:])
                return -1
        drv_objs.append(drv_obj)
    objs = list(dict.fromkeys(list(done.values()) + objs))
    lib = os.path.join(OUT, name)
    extra = ["-lstdc++"] if any(c.endswith(".cpp") for c in c_sources) else []
    r = subprocess.run([FC, "-shared", "-o", lib, *drv_objs, *objs, *extra], capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stderr[-3000:])
        return -1
    with open(os.path.join(OUT, name.replace(".so", "_MANIFEST.txt")), "w") as fh:
        ver = subprocess.run([FC, "--version"], capture_output=True, text=True).stdout.splitlines()[0]
        fh.write(f"compiler: {ver}\nflags: {' '.join(flags)}\n")
        for p in sorted(done):
            fh.write(os.path.relpath(p, REF) + "\n")
    return len(done)


def build_ieee() -> int:
    """The time-history file's number encoders (engine/source/output/tools/ieee.cpp, one C++ file)
    as oracle/_ref/libref_ieee.so."""
    src = "engine/source/output/tools/ieee.cpp"
    lib = os.path.join(OUT, "libref_ieee.so")
    # the machine and precision macros of the reference's own gfortran build (cmake_linux64_gf.txt:16-17, 134)
    defs = ["-DCPP_mach=CPP_p4linux964", "-DCPP_rel=80", "-DMYREAL8", "-std=c++14"]
    r = subprocess.run(["g++", "-O2", "-fPIC", "-shared", *defs, f"-I{REF}/engine/share/spe_inc",
                        os.path.join(REF, src), "-o", lib], capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stderr[-2000:])
        return -1
    with open(os.path.join(OUT, "libref_ieee_MANIFEST.txt"), "w") as fh:
        fh.write(f"compiler: g++ -O2 {' '.join(defs)}\n{src}\n")
    return 1


def main():
    if not os.path.isdir(REF):
        print("reference tree absent: oracle/_ref not built")
        return 0
    os.makedirs(OUT, exist_ok=True)
    if build_ieee() < 0:
        return 1
    print("oracle/_ref/libref_ieee.so built from engine/source/output/tools/ieee.cpp")
    n = build_library("libref.so", LEAVES, DRIVERS, ("common_source", "engine/source", "engine/share"),
                      ("engine/share/includes", "engine/share/spe_inc", "engine/share/r8", "common_source/includes",
                       "engine/share/message_module"), OBJ)
    if n < 0:
        return 1
    print(f"oracle/_ref/libref.so built from {n} reference files")
    n = build_library("libref_starter.so", STARTER_LEAVES, STARTER_DRIVERS,
                      ("common_source", "starter/source", "starter/share"), STARTER_INC_DIRS, OBJ + "_starter",
                      STARTER_C)
    if n < 0:
        return 1
    print(f"oracle/_ref/libref_starter.so built from {n} reference files")
    return 0


if __name__ == "__main__":
    sys.exit(main())
