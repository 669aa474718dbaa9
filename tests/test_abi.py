"""C ABI checks that need no GPU: the library loads, exports every entry point
include/lfmm.h declares, reports errors without a device, and the C++
drop-in header compiles and links against it."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lfmm.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lfmm_\w+)\s*\(", txt)))


def test_header_lists_entry_points(lfmm):
    names = declared_functions()
    assert "lfmm_fmm3d" in names and "lfmm_create" in names
    assert set(names) == set(lfmm.EXPORTS)


def test_library_exports_every_declared_symbol(lfmm):
    L = ctypes.CDLL(lfmm.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_status_strings_and_version(lfmm):
    L = lfmm.lib()
    assert L.lfmm_strerror(0) == b"ok"
    assert L.lfmm_strerror(8) == b"unsupported option"
    assert b"gfx950" in L.lfmm_version()


def test_invalid_arguments_are_rejected_without_gpu(lfmm):
    L = lfmm.lib()
    # a null context is an invalid argument before any device work
    pts = lfmm.Points3d(10, None, 3, None, 1)
    assert L.lfmm_fmm3d(None, ctypes.byref(pts), None, 1e-8, 5, 3, 0, 0, 0, None, None, None) == 1
    assert L.lfmm_set_stream(None, None) == 1
    assert L.lfmm_tree3d(None, None, 3, 0, None, None, None, None, None, None) == 1
    assert L.lfmm_direct3d(None, ctypes.byref(pts), None, 0, 0, None, None) == 1
    assert L.lfmm_direct2d(None, ctypes.byref(pts), None, 0, 0, None, None) == 1


def test_cpp_dropin_header_compiles_and_links(lfmm, tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "fmm/multipole3d.h"
#include "fmm/multipole.h"
#include <cstdio>
int main() {
    std::vector<fmm::particle3d> p(3);
    for (int i = 0; i < 3; ++i) { p[i].center = fmm::Vector3d(0.1 * i, 0.2, 0.3); p[i].q = 1.0; }
    std::vector<fmm::particle2d> p2(3);
    for (int i = 0; i < 3; ++i) { p2[i].center = fmm::point2d(0.1 * i, 0.2); p2[i].q = 1.0; }
    fmm::detail::InitMathConstants(18);
    fmm::detail::cudaCopyMathConstants();
    try {
        fmm::gpu::FastMultipole3d<double> f(p, 1e-8, 8, 2);
        std::printf("%zu %zu\n", f.potentials.size(), f.forces.size());
    } catch (const std::exception& e) { std::printf("error: %s\n", e.what()); }
    try {
        fmm::gpu::FastMultipole3d g(p, {8, 64});           // (p, leaf_max), deduced <double>
        fmm::gpu::FastMultipole3d h(p, p, fmm::Options(8, 64, false));
        fmm::Options o; o.adaptive = true; o.mixed = true;
        fmm::gpu::FastMultipole3d<double> k(p, o);
        std::printf("%zu %zu %zu\n", g.potentials.size(), h.potentials.size(), k.potentials.size());
        fmm::gpu::FastMultipole f3(p2, {8, 64, false});
        std::printf("%zu\n", f3.potentials.size());
    } catch (const std::exception& e) { std::printf("error: %s\n", e.what()); }
    try {
        fmm::gpu::FastMultipole f2(p2, 1e-8, 8, 3);
        std::printf("%zu %zu\n", f2.potentials.size(), f2.forces.size());
    } catch (const std::exception& e) { std::printf("error: %s\n", e.what()); }
    return 0;
}
''')
    exe = tmp_path / "use"
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        lfmm.LIB_PATH, "-Wl,-rpath," + os.path.dirname(lfmm.LIB_PATH)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
