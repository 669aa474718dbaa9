// tables2d.cpp -- host operator tables of the 2D (complex-series) FMM.
//
// Expansions are stored normalised per level, with h the box edge:
//   multipole  at_k = a_k / h^k,   local  bt_k = b_k h^k,
// where a, b are the reference's coefficients (src/cpu/multipole.cpp:22-97:
// a_0 = sum q, a_k = -sum q (z - c)^k / k; the local expansion is evaluated
// as Re sum b_k (z - c)^k).  With this scaling every translation depends only
// on an integer box offset d (complex, in box units):
//   M2M (child class c, delta = child centre - parent centre in child edges)
//     bt_j = 2^-j [ sum_{s=1..j} at_s delta^(j-s) C(j-1, s-1) - at_0 delta^j / j ],  bt_0 = at_0
//   M2L (d = source centre - target centre, in box edges)
//     bt_0 = at_0 (log h + conj(log(-d))) + sum_{j>=1} (-1)^j at_j / d^j
//     bt_l = [ -at_0 / l + sum_{j>=1} (-1)^j at_j C(l+j-1, j-1) / d^j ] / d^l
//   L2L (parent -> child class c): bt_k = sum_{j>=k} at_j 2^-j C(j, k) delta^(j-k)
// (reference M2M :38-53, M2L :66-84, L2L :86-97).  The level term at_0 log h
// of the M2L is added in the kernel.  All tables are computed in long double.
#include "lfmm_internal.h"

#include <cmath>
#include <complex>

namespace lfmm {

using ld = long double;
using cld = std::complex<ld>;

static ld binom_l(int n, int k)
{
    if (k < 0 || k > n) return 0;
    ld r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

static cld cpow_l(cld z, int e)
{
    cld r = 1;
    for (int i = 0; i < e; ++i) r *= z;
    return r;
}

// child class c = (x bit << 1) | y bit of the child's Morton key
static cld child_offset(int c) { return cld((c & 2) ? 0.5L : -0.5L, (c & 1) ? 0.5L : -0.5L); }

void build_op_tables_2d(int N, OpTables2d& T)
{
    T.N = N;
    T.dir_vec.clear();
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy) {
            if (std::abs(dx) <= 1 && std::abs(dy) <= 1) continue;
            T.dir_vec.push_back(dx);
            T.dir_vec.push_back(dy);
        }
    const int ND = (int)T.dir_vec.size() / 2;  // 40
    T.class_dirs.assign(4 * kDirsPerClass2d, -1);
    for (int c = 0; c < 4; ++c) {
        const int cx = (c >> 1) & 1, cy = c & 1;
        int cnt = 0;
        for (int d = 0; d < ND; ++d) {
            const int x = T.dir_vec[2 * d], y = T.dir_vec[2 * d + 1];
            if (x < -2 - cx || x > 3 - cx || y < -2 - cy || y > 3 - cy) continue;
            T.class_dirs[c * kDirsPerClass2d + cnt++] = d;
        }
    }
    auto put = [](std::vector<double>& v, size_t i, cld z) {
        v[2 * i] = (double)z.real();
        v[2 * i + 1] = (double)z.imag();
    };
    T.m2l.assign((size_t)ND * N * N * 2, 0.0);
    for (int d = 0; d < ND; ++d) {
        const cld z((ld)T.dir_vec[2 * d], (ld)T.dir_vec[2 * d + 1]);
        for (int l = 0; l < N; ++l)
            for (int j = 0; j < N; ++j) {
                cld v;
                const ld sj = (j & 1) ? -1.0L : 1.0L;
                if (l == 0 && j == 0) v = std::conj(std::log(-z));
                else if (l == 0) v = sj / cpow_l(z, j);
                else if (j == 0) v = -1.0L / ((ld)l * cpow_l(z, l));
                else v = sj * binom_l(l + j - 1, j - 1) / cpow_l(z, l + j);
                put(T.m2l, ((size_t)d * N + l) * N + j, v);
            }
    }
    T.m2m.assign((size_t)4 * N * N * 2, 0.0);
    T.l2l.assign((size_t)4 * N * N * 2, 0.0);
    for (int c = 0; c < 4; ++c) {
        const cld dl = child_offset(c);
        for (int j = 0; j < N; ++j) {
            const ld sc = ldexpl(1.0L, -j);
            for (int s = 0; s < N; ++s) {
                cld v = 0;
                if (j == 0 && s == 0) v = 1;
                else if (j >= 1 && s == 0) v = -sc * cpow_l(dl, j) / (ld)j;
                else if (s >= 1 && s <= j) v = sc * cpow_l(dl, j - s) * binom_l(j - 1, s - 1);
                put(T.m2m, ((size_t)c * N + j) * N + s, v);
            }
        }
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j) {
                const cld v = (j >= k) ? ldexpl(1.0L, -j) * binom_l(j, k) * cpow_l(dl, j - k) : cld(0);
                put(T.l2l, ((size_t)c * N + k) * N + j, v);
            }
    }
}

}  // namespace lfmm
