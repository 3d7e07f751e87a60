"""The C ABI library builds for gfx950, loads, and exports exactly what
include/painty_hip.h declares (CPU only: no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "painty_hip.h")
LIB = os.path.join(ROOT, "painty_amd", "libpainty_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pnt_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("pnt_ctx_create", "pnt_canvas_create", "pnt_compose", "pnt_canvas_dry",
                 "pnt_brush_create", "pnt_imprint", "pnt_imprint_stamps", "pnt_mix_single",
                 "pnt_layer_compose_onto_host", "pnt_dev_compose2"):
        assert must in names
    assert len(names) >= 50


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libpainty_hip.so first (python -m painty_amd.build)"
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from painty_amd import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_library_is_gfx950_code(tmp_path):
    fatbin = str(tmp_path / "fatbin.bin")
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, fatbin],
                   check=True)
    out = subprocess.run(["/opt/rocm/llvm/bin/clang-offload-bundler", "--list", "--type=o",
                          "--input=" + fatbin], capture_output=True, text=True, check=True)
    assert "hipv4-amdgcn-amd-amdhsa--gfx950" in out.stdout.split()


def test_host_only_entry_points_without_gpu():
    """Mixer entry points are host code in the reference and here."""
    from painty_amd import hip as H
    m = H.PaintMixer([([0.1, 0.2, 0.3], [1.0, 1.0, 1.0]), ([0.3, 0.2, 0.1], [0.0, 2.0, 4.0])])
    K, S = m.mixSinglePaint([0.25, 0.75])
    np.testing.assert_allclose(K, [0.25, 0.2, 0.15])
    np.testing.assert_allclose(S, [0.25, 1.75, 3.25])
    K, S = m.mixSinglePaint([1.0, 3.0])  # normalised by the weight sum
    np.testing.assert_allclose(K, [0.25, 0.2, 0.15])
    try:
        m.mixSinglePaint([1.0])
        raise AssertionError("size mismatch must raise")
    except H.InvalidArgument as e:
        assert "Palette size does not match" in str(e)
    K, S = H.PaintMixer.mixed(([1, 2, 3], [4, 5, 6]), 1.0, ([3, 2, 1], [0, 0, 0]), 3.0)
    np.testing.assert_allclose(K, [2.5, 2.0, 1.5])


def test_fp32_thresholds_match_fp64_comparisons():
    """The kernels compare FP32 values against FP32 constants that must decide
    exactly like the reference's FP64 comparisons."""
    thr = np.finfo(np.float64).eps * 1e4          # KubelkaMunk.hxx:31, :38
    thr_f = np.float32(float.fromhex("0x1.388p-39"))
    assert float(thr_f) == thr
    xs = [thr_f]
    for _ in range(4):
        xs.append(np.nextafter(xs[-1], np.float32(1)))
    ys = [thr_f]
    for _ in range(4):
        ys.append(np.nextafter(ys[-1], np.float32(0)))
    for x in xs + ys:
        assert (x < thr_f) == (float(x) < thr)
        assert (x > thr_f) == (float(x) > thr)
    mv = np.float32(0.001)                         # FootprintBrush.hxx:26, x > 0.001
    for x in (mv, np.nextafter(mv, np.float32(0)), np.nextafter(mv, np.float32(1))):
        assert (x >= mv) == (float(x) > 0.001)


def test_cpp_dropin_headers_compile():
    """The C++ drop-in classes compile cleanly against the C ABI header."""
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "painty_amd", "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_dropin_program_links():
    """The drop-in test program is built and starts (usage path, no device call)."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin")
    assert os.path.exists(exe), "build with make -C tests/cpp"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_calls_during_process_exit_are_safe():
    """A caller's static destructor calling the library after thread-local
    objects are gone (process exit) must not write to freed memory."""
    exe = os.path.join(ROOT, "tests", "cpp", "exit_order")
    assert os.path.exists(exe), "build with make -C tests/cpp"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "exit ok" in r.stdout, (r.returncode, r.stdout, r.stderr)
