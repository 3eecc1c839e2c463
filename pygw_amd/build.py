"""Build libpygw_hip.so (gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpygw_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["tetra_weights.hip", "capi_tetra.hip", "chi0.hip", "chi0_mixed.hip", "eps_inv.hip", "capi_chain.hip", "synth.hip", "mwm.hip", "mwm_mixed.hip", "selfc.hip", "capi_selfc.hip", "gamma.hip", "dos.hip", "coulomb.hip", "minm_is.hip", "pw_basis.hip", "mixed_coulomb.hip", "capi_coulomb.hip", "capi_noalloc.hip", "omega_basis.hip", "mt_coulomb.hip", "minm_mt.hip", "lapw.hip", "head.hip", "vhalf_mixed.hip"]

HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def build(verbose=False, jobs=8):
    os.makedirs(OBJDIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, procs = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJDIR, s.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *HIPFLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp", *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def _wait(procs):
    err = []
    for name, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            err.append(f"--- {name}\n{out}")
        elif out.strip() and "warning: feature flag" not in out:
            sys.stderr.write(out)
    procs.clear()
    if err:
        raise RuntimeError("hipcc failed:\n" + "\n".join(err))


if __name__ == "__main__":
    print(build(verbose=True))
