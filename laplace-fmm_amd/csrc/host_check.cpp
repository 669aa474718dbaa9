// host_check.cpp -- exercises the host-side C++ of the library without a
// GPU: the FP64 table builders (3D and 2D), the dense and split-fp16 M2L
// operators, the exact-file reader and writer, ReadError's statistics, and
// the argument checks every entry point runs before device work.  Built
// together with the library's host sources under AddressSanitizer and
// UndefinedBehaviorSanitizer (make sanitize); any report aborts the run.
// Prints "host_check ok" and exits 0 when every check holds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lfmm.h"
#include "lfmm_internal.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what, int detail = 0)
{
    if (!ok) {
        std::printf("FAILED: %s (%d)\n", what, detail);
        ++failures;
    }
}

// d^j(beta) is orthogonal: d d^T = I.
void check_wigner()
{
    for (int j = 0; j <= 24; ++j) {
        const int n = 2 * j + 1;
        std::vector<double> d((size_t)n * n);
        expect(lfmm_host_wigner_d(j, 0.37 + 0.1 * j, d.data()) == LFMM_OK, "wigner status", j);
        double worst = 0.0;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += d[(size_t)a * n + k] * d[(size_t)b * n + k];
                worst = std::fmax(worst, std::fabs(s - (a == b ? 1.0 : 0.0)));
            }
        expect(worst < 1e-11, "wigner orthogonality", j);
    }
    expect(lfmm_host_wigner_d(-1, 0.1, nullptr) == LFMM_ERR_INVALID, "wigner bad j");
    expect(lfmm_host_wigner_d(65, 0.1, nullptr) == LFMM_ERR_INVALID, "wigner j cap");
}

// The tables of the expansion sizes at the ends of the range and of the
// sizes the solves use most, built once per size; the C entry points that
// copy them out are exercised at one size each (every call rebuilds).
void check_tables_3d()
{
    for (int N : {1, 2, 3, 9, 13, LFMM_MAX_TERMS_3D}) {
        lfmm::OpTables T;
        lfmm::build_op_tables(N, T);
        for (const std::vector<double>* t : {&T.rot, &T.ephi, &T.trans, &T.knm}) {
            expect(!t->empty(), "op table size", N);
            for (double v : *t) expect(std::isfinite(v), "op table finite", N);
        }
        for (int d : T.dir_theta) expect(d >= 0 && d < T.num_theta, "dir_theta", N);
        for (int d : T.class_dirs) expect(d >= 0 && d < lfmm::kNumM2LDirs, "class_dirs", N);
        std::vector<double> K((size_t)N * N * N * N);
        for (int dir : {0, 57, 188, lfmm::kNumM2LDirs - 1}) {
            lfmm::m2l_dense_matrix(T, dir, K.data());
            for (double v : K) expect(std::isfinite(v), "m2l dense finite", N);
        }
        std::vector<uint16_t> pk;
        std::vector<int> kpre;
        double kinv = 0.0;
        lfmm::pack_m2l16(T, pk, kinv, kpre);
        const int tiles = (N * N + 31) / 32;
        expect(!pk.empty() && kinv > 0.0, "m2l16 pack", N);
        expect(kpre.size() == (size_t)lfmm::kNumM2LDirs * tiles * 2, "m2l16 prefix count", N);
        for (size_t i = 0; i + 1 < kpre.size(); i += 2)
            expect(kpre[i] >= 0 && kpre[i + 1] >= 0 && kpre[i + 1] <= kpre[i] + 32 * tiles, "m2l16 prefixes", N);
    }
    const int N = 5;
    for (int which = 0; which < 8; ++which) {
        const int64_t count = lfmm_host_op_tables(N, which, nullptr, 0);
        expect(count > 0, "op table count", which);
        if (count <= 0) continue;
        std::vector<double> buf((size_t)count);  // 8-byte slots hold either element size
        expect(lfmm_host_op_tables(N, which, buf.data(), count) == count, "op table copy", which);
    }
    std::vector<double> K((size_t)N * N * N * N);
    expect(lfmm_host_m2l_dense(N, 7, K.data()) == LFMM_OK, "m2l dense entry");
    double kinv = 0.0;
    const int64_t np = lfmm_host_m2l16_pack(N, nullptr, 0, &kinv, nullptr);
    expect(np > 0, "m2l16 pack entry size");
    std::vector<uint16_t> pk((size_t)(np > 0 ? np : 1));
    std::vector<int32_t> kpre((size_t)lfmm::kNumM2LDirs * 2);
    expect(lfmm_host_m2l16_pack(N, pk.data(), np, &kinv, kpre.data()) == np, "m2l16 pack entry");
    expect(lfmm_host_op_tables(0, 0, nullptr, 0) < 0, "op tables N=0");
    expect(lfmm_host_op_tables(19, 0, nullptr, 0) < 0, "op tables N=19");
    expect(lfmm_host_op_tables(5, 8, nullptr, 0) < 0, "op tables bad kind");
    expect(lfmm_host_m2l_dense(5, lfmm::kNumM2LDirs, nullptr) != LFMM_OK, "m2l dense bad dir");
}

void check_tables_2d()
{
    for (int N = 1; N <= LFMM_MAX_TERMS_2D; ++N) {
        lfmm::OpTables2d T;
        lfmm::build_op_tables_2d(N, T);
        expect(T.m2l.size() == (size_t)lfmm::kNumM2LDirs2d * N * N * 2, "2d m2l size", N);
        for (double v : T.m2l) expect(std::isfinite(v), "2d m2l finite", N);
        for (double v : T.m2m) expect(std::isfinite(v), "2d m2m finite", N);
        for (double v : T.l2l) expect(std::isfinite(v), "2d l2l finite", N);
    }
}

void check_exact_files(const std::string& dir)
{
    const int widths[4] = {1, 2, 1, 3};
    for (int kind = 0; kind < 4; ++kind) {
        const int64_t n = 257;
        const int w = widths[kind];
        std::vector<double> v((size_t)(n * w)), back((size_t)(n * w), 0.0);
        for (int64_t i = 0; i < n * w; ++i) v[(size_t)i] = std::sin(0.7 * (double)i) * std::pow(10.0, (double)(i % 9) - 4);
        const std::string path = dir + "/" + lfmm_exact_filename(kind);
        expect(lfmm_write_exact(path.c_str(), kind, n, v.data()) == LFMM_OK, "write exact", kind);
        expect(lfmm_read_exact(path.c_str(), kind, n, back.data()) == n, "read exact count", kind);
        for (int64_t i = 0; i < n * w; ++i) expect(back[(size_t)i] == v[(size_t)i], "exact round trip", kind);
        // asking for more values than the file holds returns what is there
        std::vector<double> more((size_t)((n + 10) * w), 0.0);
        expect(lfmm_read_exact(path.c_str(), kind, n + 10, more.data()) == n, "short file", kind);
        double mx = -1.0, l2 = -1.0, rel = -1.0;
        expect(lfmm_error_stats(kind, n, v.data(), v.data(), &mx, &l2, &rel) == LFMM_OK, "error stats", kind);
        expect(mx == 0.0 && rel == 0.0, "error stats of equal data", kind);
        std::remove(path.c_str());
    }
    expect(lfmm_read_exact((dir + "/missing.txt").c_str(), 0, 4, nullptr) < 0, "missing file");
    expect(lfmm_write_exact(nullptr, 0, 1, nullptr) != LFMM_OK, "write null path");
    expect(lfmm_error_stats(7, 1, nullptr, nullptr, nullptr, nullptr, nullptr) != LFMM_OK, "stats bad kind");
}

// Entry points given a null context fail with LFMM_ERR_INVALID before any
// device work.
void check_null_context()
{
    double buf[16] = {};
    lfmm_points3d pts{4, buf, 3, buf, 1};
    const int bad = LFMM_ERR_INVALID;
    expect(lfmm_fmm3d(nullptr, &pts, nullptr, 1e-8, 5, 3, 0, 0, 0, buf, buf, nullptr) == bad, "fmm3d null ctx");
    expect(lfmm_fmm2d(nullptr, &pts, nullptr, 1e-8, 5, 3, 0, 0, 0, buf, buf, nullptr) == bad, "fmm2d null ctx");
    expect(lfmm_fmm3d_vec(nullptr, &pts, nullptr, 1e-8, 5, 3, 0, 0, 0, buf, nullptr) == bad, "fmm3d_vec null ctx");
    expect(lfmm_direct3d(nullptr, &pts, nullptr, 0, 0u, buf, buf) == bad, "direct3d null ctx");
    expect(lfmm_direct2d(nullptr, &pts, nullptr, 0, 0u, buf, buf) == bad, "direct2d null ctx");
    expect(lfmm_direct3d_vec(nullptr, &pts, nullptr, 0, buf) == bad, "direct3d_vec null ctx");
    expect(lfmm_direct3d_noalloc(nullptr, &pts, nullptr, 0u, buf, buf) == bad, "direct3d_noalloc null ctx");
    expect(lfmm_set_stream(nullptr, nullptr) == bad, "set_stream null ctx");
    expect(lfmm_trim(nullptr) == bad, "trim null ctx");
    expect(lfmm_tree3d(nullptr, &pts, 3, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == bad, "tree3d");
    expect(lfmm_tree2d(nullptr, &pts, 3, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == bad, "tree2d");
    const int64_t off[2] = {0, 4};
    expect(lfmm_fmm3d_batch(nullptr, 1, off, buf, buf, 1e-8, 5, 3, 0, 0, 0u, 1, buf, buf, nullptr) == bad,
           "batch null ctx");
    expect(lfmm_set_order_model(nullptr, nullptr, 0.9, 1e-6, 16) == bad, "set_order_model null ctx");
    lfmm_destroy(nullptr);
    for (int s = -1; s <= 10; ++s) expect(lfmm_strerror(s) != nullptr, "strerror", s);
    expect(lfmm_version() != nullptr, "version");
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    check_wigner();
    check_tables_3d();
    check_tables_2d();
    check_exact_files(dir);
    check_null_context();
    if (failures) {
        std::printf("host_check: %d failures\n", failures);
        return 1;
    }
    std::printf("host_check ok\n");
    return 0;
}
