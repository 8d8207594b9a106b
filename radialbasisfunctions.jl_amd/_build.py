"""Build librbfx.so (the gfx950 C-ABI library) in-tree with hipcc.

Every .hip file under csrc/ is compiled to an object in _build/ (in parallel), then
linked into radialbasisfunctions.jl_amd/librbfx.so.
"""
import concurrent.futures as cf
import os
import pathlib
import re
import subprocess
import sys
import time

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "librbfx.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", str(ROOT / "include"),
    "-I", str(CSRC),
    "-Rpass-analysis=kernel-resource-usage",
]


def _compile(src, verbose):
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, "-c", *FLAGS, "-o", str(obj), str(src)]
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{p.stderr[-6000:]}")
    (OBJ / (src.stem + ".remarks")).write_text(p.stderr)
    if verbose:
        print(f"  compiled {src.name} in {time.time() - t0:.1f}s", flush=True)
    return obj


def resource_usage():
    """Per-kernel register use from the compiler's remarks of the last build:
    {kernel: {"VGPRs", "ScratchSize", "VGPRs Spill", "Occupancy"}}."""
    out = {}
    for f in sorted(OBJ.glob("*.remarks")):
        cur = None
        for line in f.read_text().splitlines():
            m = re.search(r"remark: +([A-Za-z ]+?)(?: \[bytes/lane\]| \[waves/SIMD\])?: (\S+)", line)
            if not m:
                continue
            key, val = m.group(1).strip(), m.group(2)
            if key == "Function Name":
                cur = out.setdefault(val, {})
            elif cur is not None and key in ("VGPRs", "ScratchSize", "VGPRs Spill", "Occupancy"):
                cur[key] = int(val)
    return out


def write_resource_table(path):
    rows = resource_usage()
    lines = ["kernel  VGPRs  scratch_bytes_per_lane  vgpr_spills  waves_per_SIMD"]
    for name, r in sorted(rows.items()):
        if name.startswith("_ZN4rbfx"):
            lines.append(f"{name}  {r.get('VGPRs')}  {r.get('ScratchSize')}  {r.get('VGPRs Spill')}  {r.get('Occupancy')}")
    pathlib.Path(path).write_text("\n".join(lines) + "\n")
    return rows


def build_library(jobs=None, verbose=True):
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr[-6000:]}")
    os.replace(tmp, LIB)
    write_resource_table(OBJ / "resource_usage.txt")
    if verbose:
        print(f"  linked {LIB}; per-kernel registers in {OBJ / 'resource_usage.txt'}", flush=True)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
    sys.exit(0)
