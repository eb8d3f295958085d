"""Build ``libpof_hip.so`` in-tree with hipcc for gfx950."""
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
SOURCES = ["pof_inst_d1.hip", "pof_inst_d2.hip", "pof_inst_d3.hip", "pof_inst_d4.hip", "pof_registry.hip",
           "pof_large.hip", "pof_abi.hip"]
ROCM_LIB = "/opt/rocm/lib"
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found")
    return p


def build(verbose=False, jobs=None):
    """Compile every HIP source for gfx950 and link pof/libpof_hip.so."""
    out = PKG / "libpof_hip.so"
    objdir = PKG.parent / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        objs.append(obj)
        cmd = [cc, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for cmd, p in procs:
        outp, _ = p.communicate()
        text = "\n".join(l for l in outp.decode(errors="replace").splitlines() if "not a recognized feature" not in l)
        if p.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{text}")
        if verbose and text.strip():
            print(text)
    tmp = out.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           f"-Wl,-rpath,{ROCM_LIB}"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout.decode(errors='replace')}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
