"""Build recipe for the in-tree native libraries.

libvt.so         HIP kernels + C-ABI (include/vt.h), gfx950 code objects.
libvt_scenes.so  host C generators of the synthetic benchmark worlds.

Both land in ``unnamed-voxel-tracer_amd/lib/``, with ``libvt.build.json``:
the SHA-256 of the sources and flags libvt.so was built from and of the
library itself, so a run can tell whether the library it loads is the build
of the sources beside it (:func:`provenance`).  Run ``python build.py`` or
call :func:`build`.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")

HIP_SOURCES = ["vt_api.hip", "vt_render.hip", "vt_world.hip", "vt_batch.hip", "vt_lambert.hip",
               "vt_present.hip", "vt_vox.hip", "vt_edit.hip", "vt_procgen.hip"]

# -ffp-contract=off: every binary32 multiply and add rounds separately (the
#   traversal must reproduce the CPU restatement bit for bit).
# correctly rounded divide/sqrt: 1/d, normalize() and length() bit-exact.
# -munsafe-fp-atomics: float atomicAdd as one global_atomic_add_f32 (device
#   memory from hipMalloc is coarse-grained), not a CAS loop.
# -packed-fp32-ops: no v_pk_*_f32 in the device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _run(cmd, verbose):
    if verbose:
        print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in r.stdout.splitlines()
        if "packed-fp32-ops' is not a recognized feature" not in line
    )
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def _sha256_file(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def sources_digest() -> str:
    """SHA-256 over libvt's sources (csrc/, include/vt.h) and its flags."""
    h = hashlib.sha256()
    files = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC))
    files.append(os.path.join(ROOT, "include", "vt.h"))
    for path in files:
        h.update(os.path.relpath(path, ROOT).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    h.update(" ".join(HIP_FLAGS).encode())
    return h.hexdigest()


def provenance(lib_dir: str = None) -> dict:
    """The build record of the libvt.so in lib_dir (default: the in-tree one)
    checked against that library and today's sources:
    built_from_current_sources is True when both match the record."""
    lib_dir = lib_dir or LIB
    lib = os.path.join(lib_dir, "libvt.so")
    out = {"lib": os.path.relpath(lib, ROOT), "record": None,
           "built_from_current_sources": False}
    try:
        with open(os.path.join(lib_dir, "libvt.build.json")) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return out
    lib_sha = _sha256_file(lib) if os.path.exists(lib) else None
    src_sha = sources_digest()
    out.update(record={k: rec.get(k) for k in ("sources_sha256", "libvt_sha256", "hipcc")},
               lib_sha256_matches=lib_sha == rec.get("libvt_sha256"),
               sources_sha256_matches=src_sha == rec.get("sources_sha256"))
    out["built_from_current_sources"] = bool(out["lib_sha256_matches"] and
                                             out["sources_sha256_matches"])
    return out


def build(verbose: bool = True) -> None:
    """Compiles every HIP source to an object in parallel (one hipcc each),
    then links libvt.so; each translation unit holds its own kernels."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([hipcc, *compile_flags, "-c", "-o", obj, os.path.join(CSRC, src)])
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    src_sha = sources_digest()
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o",
          os.path.join(LIB, "libvt.so"), *objs], verbose)
    ver = subprocess.run([hipcc, "--version"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True).stdout.splitlines()
    with open(os.path.join(LIB, "libvt.build.json"), "w") as f:
        json.dump({"sources_sha256": src_sha,
                   "libvt_sha256": _sha256_file(os.path.join(LIB, "libvt.so")),
                   "hipcc": ver[0] if ver else None, "flags": HIP_FLAGS}, f, indent=1)
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          "-o", os.path.join(LIB, "libvt_scenes.so"),
          os.path.join(PKG, "scenes", "vt_scenes.c"), "-lm", "-lpthread"], verbose)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
