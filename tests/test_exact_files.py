"""The reference's checking path on the host: exact-result files and the
ReadError figures (src/common/simple_math.h:218-402, simple_math.cu:103-129).

Pinned by:
  text format  the reference writes `fout.precision(20); fout << x` -- C
               printf "%.20g" for a double, "(re,im)" for std::complex
               (operator<< of <complex>), "x y z" for Vector3 (utils.h:108-109)
  read back    bit-exact (20 significant digits round-trip any double)
  statistics   a hand computation and the oracle's restatement of ReadError
No GPU: these entry points take host memory and no context.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def g20(x):
    return "%.20g" % x


@pytest.fixture
def vals():
    rng = np.random.default_rng(5)
    v = rng.standard_normal(40) * 10.0 ** rng.integers(-300, 300, 40)
    return np.concatenate([v, [0.0, -0.0, 1.0 / 3.0, 1e-5, 123456789.0, 2.0 ** -1074]])


def test_file_names_follow_the_reference(lfmm):
    assert [lfmm.exact_filename(k) for k in range(4)] == ["potential2d_exact.txt", "force2d_exact.txt",
                                                          "potential3d_exact.txt", "force3d_exact.txt"]
    assert lfmm.lib().lfmm_exact_filename(4) is None


def test_potential_file_text_and_roundtrip(lfmm, tmp_path, vals):
    p = tmp_path / "potential3d_exact.txt"
    lfmm.write_exact(p, lfmm.POTENTIAL3D, vals)
    assert p.read_text() == "".join(g20(x) + "\n" for x in vals)
    n, back = lfmm.read_exact(p, lfmm.POTENTIAL3D, len(vals))
    assert n == len(vals) and np.array_equal(back, vals)
    assert np.array_equal(np.signbit(back), np.signbit(vals))


def test_force_files_text_and_roundtrip(lfmm, tmp_path, vals):
    f2 = vals[:46].reshape(-1, 2)
    p = tmp_path / "f2.txt"
    lfmm.write_exact(p, lfmm.FORCE2D, f2)
    assert p.read_text() == "".join("(%s,%s)\n" % (g20(a), g20(b)) for a, b in f2)
    n, back = lfmm.read_exact(p, lfmm.FORCE2D, len(f2))
    assert n == len(f2) and np.array_equal(back, f2)
    f3 = vals[:45].reshape(-1, 3)
    p = tmp_path / "f3.txt"
    lfmm.write_exact(p, lfmm.FORCE3D, f3)
    assert p.read_text() == "".join("%s %s %s\n" % tuple(g20(c) for c in r) for r in f3)
    n, back = lfmm.read_exact(p, lfmm.FORCE3D, len(f3))
    assert n == len(f3) and np.array_equal(back, f3)


def test_read_counts_short_files_and_rejects_bad_arguments(lfmm, tmp_path):
    p = tmp_path / "short.txt"
    lfmm.write_exact(p, lfmm.POTENTIAL3D, np.array([1.5, 2.5]))
    n, back = lfmm.read_exact(p, lfmm.POTENTIAL3D, 5)
    assert n == 2 and list(back) == [1.5, 2.5, 0.0, 0.0, 0.0]
    L = lfmm.lib()
    assert L.lfmm_write_exact(str(p).encode(), 7, 1, np.zeros(1).ctypes.data) == 1
    assert L.lfmm_read_exact(str(tmp_path / "missing.txt").encode(), 2, 1, np.zeros(1).ctypes.data) < 0
    assert L.lfmm_error_stats(9, 1, None, None, None, None, None) == 1


def test_error_stats_by_hand(lfmm):
    exact = np.array([2.0, -4.0, 1.0])
    res = np.array([2.5, -4.0, 0.0])
    # |e| = .5, 0, 1; |x| = 2, 4, 1
    mx, l2, rel = lfmm.error_stats(lfmm.POTENTIAL3D, exact, res)
    assert mx == 1.0
    assert l2 == pytest.approx(np.sqrt(((0.5 / 2) ** 2 + 0 + 1.0) / 3), rel=1e-15)
    assert rel == pytest.approx(1.5 / 7.0, rel=1e-15)
    # vectors: Euclidean norms; complex: modulus
    e3 = np.array([[3.0, 0.0, 4.0]])
    mx, l2, rel = lfmm.error_stats(lfmm.FORCE3D, e3, np.zeros((1, 3)))
    assert (mx, l2, rel) == (5.0, 1.0, 1.0)
    mx, l2, rel = lfmm.error_stats(lfmm.FORCE2D, np.array([[3.0, 4.0]]), np.array([[0.0, 4.0]]))
    assert (mx, l2, rel) == (3.0, 0.6, 0.6)


@pytest.mark.parametrize("kind,width", [(0, 1), (1, 2), (2, 1), (3, 3)])
def test_error_stats_match_oracle(lfmm, orc, kind, width):
    rng = np.random.default_rng(kind)
    exact = rng.standard_normal((500, width)) + 3.0
    res = exact * (1 + 1e-6 * rng.standard_normal((500, width)))
    if width == 1:
        exact, res = exact[:, 0], res[:, 0]
    got = lfmm.error_stats(kind, exact, res)
    want = orc.read_error(exact, res)
    assert np.allclose(got, want, rtol=1e-12, atol=0)


def test_cpp_checking_header_compiles(lfmm, tmp_path):
    src = tmp_path / "chk.cpp"
    src.write_text(r'''
#include "fmm/multipole3d.h"
#include "fmm/multipole.h"
int main() {
    std::vector<fmm::particle3d> p(4);
    std::vector<fmm::particle2d> p2(4);
    std::vector<double> r(4);
    std::vector<fmm::Vector3d> f(4);
    std::vector<std::complex<double>> f2(4);
    if (p.size() == 99) {  // instantiated, not run
        fmm::gpu::ComputeExact<fmm::InteractionType::Potential3d>(p);
        fmm::gpu::ComputeExact<fmm::InteractionType::Force3d>(p, p);
        fmm::ComputeExact<fmm::InteractionType::Force2d>(p2);
        fmm::ReadError<fmm::InteractionType::Potential3d>(r);
        fmm::ReadError<fmm::InteractionType::Force3d>(f);
        fmm::ReadError<fmm::InteractionType::Force2d>(f2);
    }
    return 0;
}
''')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
