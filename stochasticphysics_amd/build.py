"""Build recipe of the gfx950 engine library ``stochasticphysics_amd/libstophys_gpu.so``.

Every translation unit under csrc/ is compiled by hipcc for gfx950 (in
parallel), objects go to ``_build/`` at the repository root, and the shared
library is linked in-tree next to this file so that it travels with the
repository snapshot.

    python -m stochasticphysics_amd.build        # or __graft_entry__.build()
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libstophys_gpu.so"
APP = PKG / "stophys_gpu_app"
ARCH = "gfx950"

# packed FP32 instructions are disabled for every HIP translation unit
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]

# (source, extra flags, object name); a source compiled several times gets one object per flag set
SOURCES = [
    "sp_kernels_neel.hip", "sp_kernels_uc.hip",
    # BrownAndNeel: FP64 and ImplicitMidpoint; mixed precision (see sp_kernels_bn.hip)
    ("sp_kernels_bn.hip", ["-DSP_BN_PART=1"], "sp_kernels_bn64.o"),
    ("sp_kernels_bn.hip", ["-DSP_BN_PART=2"], "sp_kernels_bn.o"),
    # the FP64 explicit NeelSpherical classes without machine LICM (see the head of sp_kernels_sph.hip)
    ("sp_kernels_sph.hip", ["-DSP_SPH_PART=1", "-mllvm", "-disable-machine-licm"], "sp_kernels_sph64.o"),
    ("sp_kernels_sph.hip", ["-DSP_SPH_PART=2"], "sp_kernels_sph.o"),
    ("sp_kernels_sph.hip", ["-DSP_SPH_PART=3"], "sp_kernels_sph1w.o"),
    ("sp_kernels_bnes.hip", ["-DSP_BNES_ANISO=1"], "sp_kernels_bnes_a1.o"),
    ("sp_kernels_bnes.hip", ["-DSP_BNES_ANISO=2"], "sp_kernels_bnes_a2.o"),
    ("sp_kernels_bnes.hip", ["-DSP_BNES_ANISO=4"], "sp_kernels_bnes_a4.o"),
    "sp_kernels_common.hip", "sp_kernels_provider.hip", "sp_noalloc.hip",
    "sp_params.cpp", "sp_enums.cpp", "sp_runtime.cpp", "sp_server.cpp",
    "frontend/ini.cpp", "frontend/settings.cpp",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _compile(entry):
    src, extra, obj_name = (entry, [], Path(entry).stem + ".o") if isinstance(entry, str) else entry
    obj = BUILD / obj_name
    cmd = [hipcc(), *HIP_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass warns that it does not know the packed-fp32 target feature
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj


def build(verbose=False, jobs=None):
    """Compile all sources and link the shared library.  Always recompiles."""
    BUILD.mkdir(exist_ok=True)
    srcs = list(SOURCES)
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # command-line front-end (StoPhysApp_MultiArch options), linked against the library
    app_tmp = APP.with_suffix(".tmp")
    cmd = [hipcc(), "-O2", "-std=c++17", str(CSRC / "frontend" / "stophys_gpu_app.cpp"), "-o", str(app_tmp),
           f"-L{PKG}", "-lstophys_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"app link failed:\n{r.stderr}")
    os.replace(app_tmp, APP)
    if verbose:
        print(f"built {LIB} and {APP}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
