"""Build libhfgpu.so (HIP for gfx950) in-tree, plus the CPU oracle used by the tests.

    python -m hartreefock_amd.build [-j N] [--force]

Steps: generate the per-class ERI code (csrc/gen_eri.py), compile every .hip with
hipcc --offload-arch=gfx950 and every host .cpp, link hartreefock_amd/libhfgpu.so
against rocBLAS/rocSOLVER, then `make -C oracle`.  Objects go to csrc/_build/;
a translation unit is rebuilt when it or any shared header is newer than its object.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import re
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(CSRC, "generated")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(PKG, "libhfgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HF_OFFLOAD_ARCH", "gfx950")
LMAX = 2

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-variable", "-Wno-unused-function",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _newer(src_paths, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in src_paths if os.path.exists(p))


def _compile(src, obj, hip, deps, force):
    if not force and not _newer([src] + deps, obj):
        return src, 0.0, ""
    cmd = [HIPCC] + COMMON + (HIP_FLAGS + ["-x", "hip"] if hip else ["-D__HIP_PLATFORM_AMD__"]) + ["-c", src, "-o", obj]
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    return src, time.time() - t0, p.stderr


def build(jobs=None, force=False, verbose=False, oracle=True):
    jobs = jobs or min(16, os.cpu_count() or 4)
    os.makedirs(OBJ, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_eri.py"), GEN, str(LMAX)], check=True)
    kdeps = [os.path.join(CSRC, "hf_device.h"), os.path.join(CSRC, "wave_reduce.h"), os.path.join(CSRC, "eri_kernel.hpp"),
             os.path.join(CSRC, "coop_kernel.hpp"), os.path.join(GEN, "eri_classes.inc")]
    kdeps += sorted(glob.glob(os.path.join(GEN, "coop_*.inc")))
    hdeps = [os.path.join(CSRC, "hf_device.h"), os.path.join(CSRC, "wave_reduce.h"), os.path.join(CSRC, "hf_internal.h"),
             os.path.join(CSRC, "gershgorin.h"), os.path.join(CSRC, "scf_run.h"), os.path.join(CSRC, "dense_tile.h"),
             os.path.join(GEN, "dispatch.inc"), os.path.join(ROOT, "include", "hfgpu.h")]
    for stale in glob.glob(os.path.join(OBJ, "k_????.hip.o")):
        os.remove(stale)                 # objects of the former one-file-per-class layout
    units = []
    for src in sorted(glob.glob(os.path.join(GEN, "k_*.hip"))):
        units.append((src, True, kdeps))
    for src in sorted(glob.glob(os.path.join(GEN, "d_*.hip"))):
        units.append((src, True, [os.path.join(CSRC, "hf_device.h")]))
    for name in ("context.hip", "scf.hip", "batch.hip", "one_electron.hip", "dense.hip", "incore.hip", "post.hip"):
        units.append((os.path.join(CSRC, name), True, hdeps))
    for name in ("basis.cpp", "boys.cpp", "serve.cpp"):
        units.append((os.path.join(CSRC, name), False, hdeps))
    objs = []
    # biggest kernels first so the parallel build ends sooner
    units.sort(key=lambda u: ("2222" not in u[0], "222" not in u[0], u[0]))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for src, hip, deps in units:
            obj = os.path.join(OBJ, os.path.basename(src) + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, src, obj, hip, deps, force))
        for f in cf.as_completed(futs):
            src, dt, err = f.result()
            if verbose and dt > 0:
                print("  %-40s %6.1fs" % (os.path.basename(src), dt), flush=True)
    if force or _newer(objs, LIB):
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + \
              ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-pthread", "-Wl,-rpath,/opt/rocm/lib"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (p.stdout, p.stderr))
        os.replace(LIB + ".tmp", LIB)
    if oracle:
        # test harness: generated class code compiled for the host (CPU checks of the kernel math)
        hsrc = os.path.join(ROOT, "tests", "cpu_kernel", "quartet_host.hip")
        hlib = os.path.join(ROOT, "tests", "cpu_kernel", "libquartet_host.so")
        hdeps = kdeps + [hsrc, os.path.join(GEN, "host_eval.inc"), os.path.join(CSRC, "boys.cpp")]
        if force or _newer(hdeps, hlib):
            cmd = [HIPCC] + COMMON[1:] + ["-O1", "-x", "hip", "--offload-arch=" + ARCH, "-shared", "-o", hlib + ".tmp",
                                          hsrc, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                                          os.path.join(CSRC, "boys.cpp")]
            p = subprocess.run(cmd, capture_output=True, text=True)
            if p.returncode != 0:
                raise RuntimeError("host harness build failed:\n%s%s" % (p.stdout, p.stderr))
            os.replace(hlib + ".tmp", hlib)
        p = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("oracle build failed:\n%s%s" % (p.stdout, p.stderr))
    return LIB


def _usage_of(src):
    """Device compile of one source with the compiler's kernel-resource remarks: per kernel VGPRs, AGPRs, SGPRs,
    scratch bytes per lane, LDS bytes per block and the occupancy (waves per SIMD) the compiler derives."""
    cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-x", "hip", "--offload-device-only", "-c", src, "-o", os.devnull,
                                          "-Rpass-analysis=kernel-resource-usage"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    rows, cur = [], None
    for line in p.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = {"kernel": m.group(1), "file": os.path.basename(src)}
            rows.append(cur)
            continue
        m = re.search(r"remark:\s+(TotalSGPRs|VGPRs Spill|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|"
                      r"Occupancy \[waves/SIMD\]|LDS Size \[bytes/block\]): (\d+)", line)
        if m and cur is not None:
            key = {"TotalSGPRs": "SGPRs", "VGPRs Spill": "Spill"}.get(m.group(1), m.group(1).split(" ")[0])
            cur[key] = int(m.group(2))
    return rows


def resource_usage(jobs=None):
    """Print the kernel-resource table of every device source (after build(), which generates the class sources)."""
    srcs = sorted(glob.glob(os.path.join(GEN, "k_*.hip"))) + \
        [os.path.join(CSRC, n) for n in ("incore.hip", "dense.hip", "scf.hip", "one_electron.hip", "post.hip")]
    with cf.ThreadPoolExecutor(max_workers=jobs or min(16, os.cpu_count() or 4)) as ex:
        rows = [r for rs in ex.map(_usage_of, srcs) for r in rs]
    print("# static LDS only (the J/K kernels' exchange-row slots and Boys table are dynamic LDS set at launch)")
    print("%-18s %5s %5s %5s %6s %8s %7s %6s  %s" % ("file", "VGPR", "AGPR", "SGPR", "spill", "scratch", "LDS", "waves",
                                                    "kernel"))
    for r in rows:
        print("%-18s %5s %5s %5s %6s %8s %7s %6s  %s" % (r["file"], r.get("VGPRs"), r.get("AGPRs"), r.get("SGPRs"),
                                                       r.get("Spill"), r.get("ScratchSize"), r.get("LDS"),
                                                       r.get("Occupancy"), r["kernel"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--resource-usage", action="store_true",
                    help="after building, print every kernel's registers, scratch, LDS and occupancy")
    a = ap.parse_args()
    t0 = time.time()
    build(jobs=a.j, force=a.force, verbose=True)
    print("built %s in %.1fs" % (LIB, time.time() - t0))
    if a.resource_usage:
        resource_usage(a.j)


if __name__ == "__main__":
    main()
