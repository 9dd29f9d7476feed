"""Product checks that need no GPU: the C-ABI library, its host-side chains,
and the multi-word arithmetic header built for the host against MPFR."""
import ctypes
import os
import re
import subprocess

import pytest

from tests import rn_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "mdzg.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdzg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import mdz_amd
    L = mdz_amd.lib()
    names = _declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.mdzg_abi_version() == 1


def test_strerror_covers_codes():
    import mdz_amd
    L = mdz_amd.lib()
    for code, text in mdz_amd.ERRORS.items():
        assert L.mdzg_strerror(code).decode() == text


@pytest.mark.parametrize("args", [
    ("-0.75", "0", "3.0", 53, 240, 180),
    ("-0.743643887037158704752191506114774", "0.131825904205311970493132056385139",
     "1e-25", 128, 640, 480),
    ("-0.75", "0", "1e-290", 1024, 1920, 1080),
    ("0.25", "-0.5", "0.001", 176, 100, 300),
])
def test_rect_from_center_matches_oracle_and_exact(oracle, args):
    import mdz_amd
    got = mdz_amd.rect_from_center(*args)
    assert got == oracle.rect_from_center(*args)
    assert [rn_ref.parse(s) for s in got] == list(rn_ref.rect_from_center(*args))


@pytest.mark.parametrize("args", [
    ("-0.75", "0", "3.0", 128, 240, 180),
    ("-0.743643887037158704752191506114774", "0.131825904205311970493132056385139", "1e-25",
     100, 640, 480),
    ("-1.768578453196147181888758097455841394335", "9.168e-4", "1.7911e-37", 176, 180, 240),
])
def test_rect_at_coords_precision_matches_oracle_and_exact(oracle, args):
    """The GMP path's rect (coords' own values, src/coords.c:317-324)."""
    import mdz_amd
    got = mdz_amd.rect_from_center(*args, coords=True)
    assert got == oracle.rect_from_center(*args, coords=True)
    assert [rn_ref.parse(s) for s in got] == list(rn_ref.rect_from_center(*args, coords=True))


def test_gmp_limb_precision_rule_matches_libgmp():
    """The GMP path's kernel choice uses mpf_init2's limb precision; compare it
    with what libgmp itself allocates (_mp_prec of an mpf_init2'ed value)."""
    import mdz_amd
    g = ctypes.CDLL("/opt/conda/lib/libgmp.so")

    class Mpf(ctypes.Structure):
        _fields_ = [("prec", ctypes.c_int), ("size", ctypes.c_int), ("exp", ctypes.c_long),
                    ("d", ctypes.c_void_p)]
    buckets = (4, 6, 10, 18)
    for p in (2, 53, 54, 64, 80, 100, 128, 129, 192, 193, 256, 320, 500, 512, 640, 1000, 1024):
        x = Mpf()
        g.__gmpf_init2(ctypes.byref(x), ctypes.c_ulong(p))
        lr = x.prec + 1
        g.__gmpf_clear(ctypes.byref(x))
        kid = mdz_amd.kernel_id(mdz_amd.PATH_GMP, p)
        first = mdz_amd.kernel_id(mdz_amd.PATH_GMP, 2)
        assert buckets[kid - first] >= lr and (kid == first or buckets[kid - first - 1] < lr), p


def test_rect_rejects_bad_input():
    import mdz_amd
    with pytest.raises(mdz_amd.MdzgError) as e:
        mdz_amd.rect_from_center("nonsense", "0", "1", 128, 10, 10)
    assert e.value.code == -6
    with pytest.raises(mdz_amd.MdzgError) as e:
        mdz_amd.rect_from_center("0", "0", "1", 128, 0, 10)
    assert e.value.code == -1


def _gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    import mdz_amd
    with pytest.raises(mdz_amd.MdzgError) as e:
        mdz_amd.Context(0)
    assert e.value.code == -3


def test_workload_specs_deterministic():
    from mdz_amd import workloads as WL
    a = WL.config3_specs()
    b = WL.config3_specs()
    assert a == b and len(a) == 256
    sides = {s["side"] for s in a}
    assert sides == {64, 128, 256, 512}
    fr = [s["fractal"] for s in a]
    assert 0.45 < fr.count(0) / 256 < 0.75
    for s in a:
        assert s["precision"] in WL.PREC_CLASSES
        assert 1000 <= s["depth"] <= 50000


def test_arithmetic_header_matches_mpfr(tmp_path):
    """mdz_amd/csrc/mp_rn.h compiled for the host, every op vs MPFR_RNDN."""
    exe = tmp_path / "mp_selftest"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I/opt/conda/include", "-o", str(exe),
                    os.path.join(ROOT, "tests", "mp_selftest.cpp"), "-L/opt/conda/lib",
                    "-Wl,-rpath,/opt/conda/lib", "-lmpfr", "-lgmp"], check=True)
    for seed in (11, 12):
        r = subprocess.run([str(exe), str(seed), "3000"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "OK fails=0" in r.stdout
        # the short products' exact fallback was exercised, not only the fast path
        nfb = int(r.stdout.split("short-product fallbacks=")[1].split()[0])
        assert nfb > 0


def test_mpf_header_matches_gmp(tmp_path):
    """mdz_amd/csrc/mpf_gmp.h compiled for the host, every operation of the
    GMP loops vs libgmp's mpf layer, limb for limb."""
    exe = tmp_path / "mpf_selftest"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I/opt/conda/include", "-o", str(exe),
                    os.path.join(ROOT, "tests", "mpf_selftest.cpp"), "-L/opt/conda/lib",
                    "-Wl,-rpath,/opt/conda/lib", "-lgmp"], check=True)
    for seed in (21, 22):
        r = subprocess.run([str(exe), str(seed), "8000"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "OK fails=0" in r.stdout


def test_device_op_check_library_loads():
    """tests/gpu_arith/libarith_check.so (built by __graft_entry__.build) loads
    and exports its entry point; calling it needs a GPU."""
    import ctypes
    from tests.gpu_arith import build as AB
    if not os.path.exists(AB.LIB):
        pytest.skip("not built")
    L = ctypes.CDLL(AB.LIB)
    assert hasattr(L, "arith_check")
