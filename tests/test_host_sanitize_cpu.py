"""The library's host C code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/capi/host_check.c drives every entry point that needs no device (rect
chains, scene and palette files -- the gallery's and damaged copies of them
--, PNG writing, error paths, mdzg_create without a device) through the host
sources compiled by gcc with -fsanitize=address,undefined, linked with the
device objects as the library build left them (their code is not
instrumented and is not run here).  A sanitizer report, a leak or a crash
fails the test."""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "mdz_amd", "_build")
DEVICE_OBJS = ["launch.o", "noalloc.o", "graph.o", "k_fp64.o", "k_gmp.o", "k_colour.o", "k_probe.o"] + \
    [f"k_rn{n}.o" for n in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)]


def test_host_code_under_asan_and_ubsan(tmp_path):
    objs = [os.path.join(BUILD, o) for o in DEVICE_OBJS]
    if not all(os.path.exists(o) for o in objs):
        pytest.skip("device objects not built (run __graft_entry__.build())")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    mp = os.environ.get("MDZG_MP_PREFIX", "/opt/conda")
    syslib = os.environ.get("MDZG_SYS_LIBDIR", "/usr/lib/x86_64-linux-gnu")
    exe = str(tmp_path / "host_check")
    host = sorted(glob.glob(os.path.join(ROOT, "mdz_amd", "host", "*.c")))
    cmd = ["gcc", "-std=gnu99", "-g", "-O1", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fPIC", "-ffp-contract=off", "-pthread",
           "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include", f"-I{mp}/include",
           "-I", os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(ROOT, "tests", "capi", "host_check.c"), *host, *objs,
           f"-L{rocm}/lib", f"-Wl,-rpath,{rocm}/lib", "-lamdhip64", "-lstdc++",
           f"-L{syslib}", "-l:libmpfr.so.6", "-l:libgmp.so.10", "-l:libz.so.1", "-lm", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    gal = os.path.join(ROOT, "tests", "golden", "mdz")
    scenes = sorted(glob.glob(os.path.join(gal, "*.mdz")))
    maps = sorted(glob.glob(os.path.join(gal, "*.map")))
    assert scenes and maps
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    r = subprocess.run([exe, str(tmp_path), *scenes, "--maps", *maps], capture_output=True,
                       text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "ERROR: LeakSanitizer" not in out, out[-4000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("host_check calls")]
    assert line, out[-2000:]
    calls, errors = int(line[0].split()[2]), int(line[0].split()[4])
    assert calls > 2000 and 0 < errors < calls
