"""Build the HIP library ``libwsb.so`` in-tree for gfx950 (hipcc cross-compiles without a GPU).

Each source is compiled to its own object in parallel, then the objects are linked into one shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["api.hip", "geom2d.hip", "moments2d.hip", "batch2d.hip", "tri2d.hip", "sphere3d.hip", "mesh3d.hip", "disk2d.hip"]
OUT = os.path.join(HERE, "libwsb.so")
OBJDIR = os.path.join(HERE, "_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=fast-honor-pragmas"]


def _clean(text):
    return "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip())


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    r = subprocess.run([HIPCC, *FLAGS, "-c", "-o", obj + ".tmp", os.path.join(CSRC, src)], cwd=CSRC,
                       capture_output=True, text=True)
    if r.returncode == 0:
        os.replace(obj + ".tmp", obj)
    return src, obj, r.returncode, _clean(r.stdout + r.stderr)


def build(verbose=False, jobs=None):
    sys.path.insert(0, os.path.dirname(HERE))
    from warpsphboundaries_amd.csrc import gen_tables
    gen_tables.main()
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    errors = [(s, m) for s, _, rc, m in results if rc != 0]
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(f"--- {s}\n{m}" for s, m in errors))
    if verbose:
        for s, _, _, m in results:
            if m:
                print(f"--- {s}\n{m}")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp",
                        *[o for _, o, _, _ in results]], cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed ({r.returncode}):\n{_clean(r.stdout + r.stderr)}")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
