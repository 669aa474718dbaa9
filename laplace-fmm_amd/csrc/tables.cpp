// tables.cpp -- host FP64 operator tables for the rotation-based 3D FMM.
//
// Everything here is computed on the host in long double and rounded once to
// double: the Wigner rotations, the axial translation coefficients and the
// harmonic normalisation.  The GPU kernels only read these tables.
//
// Rotation convention (matches the reference's RotateY, src/cpu/
// multipole3d.cpp:79-113, whose table is built by ComputeWignerD,
// src/common/special_functions.cpp:211-351): for degree n the reference uses
//   D_n(m, k) = eps(k) d^n_{k,m}(theta),  eps(k) = (-1)^k for k < 0, else 1,
// where d^j_{m'm}(beta) = <j m'| exp(-i beta J_y) |j m> is the standard small
// Wigner matrix.  We compute d^j by scaling-and-squaring the exponential of
// the tridiagonal generator (no recurrences in theta), then fold the
// conjugate-symmetric negative orders into two real (n+1)x(n+1) matrices:
//   P(m,k) = D(m,k) + D(m,-k),  Q(m,k) = D(m,k) - D(m,-k)   (k >= 1)
//   P(m,0) = Q(m,0) = D(m,0)
// acting on Re and Im of the coefficients respectively.  The inverse rotation
// D(-theta) is (-1)^(m-k) D(theta) and is applied with that sign in-kernel.
#include "lfmm_internal.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <tuple>
#include <utility>

namespace lfmm {

using ld = long double;

static ld factl(int n)
{
    ld r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

// K_nm = sqrt((n-m)!/(n+m)!)  (reference knm_wrapper, special_functions.h:56-85)
static ld knm_l(int n, int m) { return sqrtl(factl(n - m) / factl(n + m)); }
// A_nm = (-1)^n / sqrt((n-m)!(n+m)!)  (reference anm_wrapper, :87-116)
static ld anm_l(int n, int m) { return ((n & 1) ? -1.0L : 1.0L) / sqrtl(factl(n - m) * factl(n + m)); }
// m2l(j,k,n) = A_nk A_jk (-1)^(n+k) / A_{j+n,0}  (m2lcoef_wrapper, :118-138)
static ld m2l_l(int j, int k, int n)
{
    return anm_l(n, k) * anm_l(j, k) * (((n + k) & 1) ? -1.0L : 1.0L) / anm_l(j + n, 0);
}

double host_knm(int n, int m) { return (double)knm_l(n, m); }
double host_anm(int n, int m) { return (double)anm_l(n, m); }

// d^j_{m'm}(beta) for m', m in [-j, j], row-major [m'+j][m+j].
static void wigner_d_l(int j, ld beta, std::vector<ld>& out)
{
    const int D = 2 * j + 1;
    std::vector<ld> A(D * D, 0.0L), E(D * D, 0.0L), T(D * D), W(D * D);
    for (int a = 0; a + 1 < D; ++a) {
        const int m = a - j;
        const ld v = sqrtl((ld)j * (j + 1) - (ld)m * (m + 1));  // <m+1|J+|m>
        A[(a + 1) * D + a] = -0.5L * beta * v;                   // -(beta/2) J+
        A[a * D + (a + 1)] = 0.5L * beta * v;                    // +(beta/2) J-
    }
    ld nrm = 0;
    for (int r = 0; r < D; ++r) {
        ld s = 0;
        for (int c = 0; c < D; ++c) s += fabsl(A[r * D + c]);
        nrm = std::max(nrm, s);
    }
    int sq = 0;
    while (nrm > 0.125L) { nrm *= 0.5L; ++sq; }
    const ld scl = ldexpl(1.0L, -sq);
    for (auto& x : A) x *= scl;
    // Horner Taylor: E = I + A(I + A/2(I + A/3(...)))
    for (int r = 0; r < D; ++r) E[r * D + r] = 1.0L;
    for (int t = 24; t >= 1; --t) {
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) {
                ld s = 0;
                for (int k = 0; k < D; ++k) s += A[r * D + k] * E[k * D + c];
                T[r * D + c] = s / t;
            }
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) E[r * D + c] = T[r * D + c] + (r == c ? 1.0L : 0.0L);
    }
    for (int s = 0; s < sq; ++s) {
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) {
                ld v = 0;
                for (int k = 0; k < D; ++k) v += E[r * D + k] * E[k * D + c];
                W[r * D + c] = v;
            }
        E.swap(W);
    }
    out = E;
}

void host_wigner_small_d(int j, double beta, double* out)
{
    std::vector<ld> d;
    wigner_d_l(j, (ld)beta, d);
    for (size_t i = 0; i < d.size(); ++i) out[i] = (double)d[i];
}

static int gcd3(int a, int b, int c)
{
    return std::gcd(std::gcd(std::abs(a), std::abs(b)), std::abs(c));
}

void build_op_tables(int N, OpTables& T)
{
    T.N = N;
    // ---- directions ----
    T.dir_vec.assign(kNumDirs * 4, 0);
    int nd = 0;
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dz = -3; dz <= 3; ++dz) {
                if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                T.dir_vec[4 * nd] = dx; T.dir_vec[4 * nd + 1] = dy; T.dir_vec[4 * nd + 2] = dz;
                ++nd;
            }
    for (int s = 0; s < kNumDiagDirs; ++s) {
        // class bit layout of a Morton child: x << 2 | y << 1 | z
        T.dir_vec[4 * (kDiagBase + s)] = (s & 4) ? 1 : -1;
        T.dir_vec[4 * (kDiagBase + s) + 1] = (s & 2) ? 1 : -1;
        T.dir_vec[4 * (kDiagBase + s) + 2] = (s & 1) ? 1 : -1;
    }
    // ---- unique polar angles (from the gcd-reduced direction) ----
    std::map<std::tuple<int, int, int>, int> dir_key;  // reduced (|xy|^2, z, norm2) -> theta index
    std::vector<ld> thetas;
    T.dir_theta.assign(kNumDirs, 0);
    T.dir_trans.assign(kNumDirs, -1);
    std::vector<ld> phis(kNumDirs);
    std::map<int, int> dist_class;  // |d|^2 -> class
    for (int d = 0; d < kNumDirs; ++d) {
        const int x = T.dir_vec[4 * d], y = T.dir_vec[4 * d + 1], z = T.dir_vec[4 * d + 2];
        const int g = gcd3(x, y, z);
        const int rx = x / g, ry = y / g, rz = z / g;
        const int rxy = rx * rx + ry * ry;
        auto key = std::make_tuple(rxy, rz, rxy + rz * rz);
        auto it = dir_key.find(key);
        if (it == dir_key.end()) {
            const ld r = sqrtl((ld)(rxy + rz * rz));
            thetas.push_back(acosl((ld)rz / r));
            it = dir_key.emplace(key, (int)thetas.size() - 1).first;
        }
        T.dir_theta[d] = it->second;
        phis[d] = (x == 0 && y == 0) ? 0.0L : atan2l((ld)y, (ld)x);
        if (d < kNumM2LDirs) {
            const int n2 = x * x + y * y + z * z;
            auto dc = dist_class.find(n2);
            if (dc == dist_class.end()) dc = dist_class.emplace(n2, (int)dist_class.size()).first;
            T.dir_trans[d] = dc->second;
        }
    }
    T.num_theta = (int)thetas.size();
    // ---- rotation pairs ----
    const int RS = rot_size(N);
    T.rot.assign((size_t)T.num_theta * RS * 2, 0.0);
    std::vector<ld> dm;
    for (int t = 0; t < T.num_theta; ++t)
        for (int n = 0; n < N; ++n) {
            wigner_d_l(n, thetas[t], dm);
            const int D = 2 * n + 1;
            auto dd = [&](int mp, int m) { return dm[(mp + n) * D + (m + n)]; };
            for (int m = 0; m <= n; ++m)
                for (int k = 0; k <= n; ++k) {
                    ld P, Q;
                    if (k == 0) {
                        P = Q = dd(0, m);
                    } else {
                        const ld a = dd(k, m);                              // D(m,k)
                        const ld b = ((k & 1) ? -1.0L : 1.0L) * dd(-k, m);  // D(m,-k) = eps(-k) d_{-k,m}
                        P = a + b;
                        Q = a - b;
                    }
                    const size_t o = ((size_t)t * RS + rot_base(n) + m * (n + 1) + k) * 2;
                    T.rot[o] = (double)P;
                    T.rot[o + 1] = (double)Q;
                }
        }
    // ---- azimuthal factors ----
    T.ephi.assign((size_t)kNumDirs * N * 2, 0.0);
    for (int d = 0; d < kNumDirs; ++d)
        for (int k = 0; k < N; ++k) {
            T.ephi[((size_t)d * N + k) * 2] = (double)cosl(k * phis[d]);
            T.ephi[((size_t)d * N + k) * 2 + 1] = (double)sinl(k * phis[d]);
        }
    // ---- translations (normalised; see lfmm_internal.h) ----
    const int TS = tr_size(N);
    T.trans.assign((size_t)kNumTrans * TS, 0.0);
    for (auto& kv : dist_class) {
        const ld D = sqrtl((ld)kv.first);
        double* tt = &T.trans[(size_t)kv.second * TS];
        for (int k = 0; k < N; ++k)
            for (int j = k; j < N; ++j)
                for (int i = k; i < N; ++i)
                    tt[tr_base(N, k) + (j - k) * (N - k) + (i - k)] = (double)(m2l_l(j, k, i) * powl(D, -(ld)(j + i + 1)));
    }
    const ld rr = sqrtl(3.0L) / 4.0L;  // |child centre - parent centre| / parent edge
    {
        double* tt = &T.trans[(size_t)kTransM2M * TS];  // reference M2M, multipole3d.cpp:137-147
        for (int k = 0; k < N; ++k)
            for (int j = k; j < N; ++j)
                for (int i = k; i <= j; ++i)
                    tt[tr_base(N, k) + (j - k) * (N - k) + (i - k)] =
                        (double)(anm_l(j - i, 0) * anm_l(i, k) / anm_l(j, k) * powl(rr, (ld)(j - i)) * ldexpl(1.0L, -i));
    }
    {
        double* tt = &T.trans[(size_t)kTransL2L * TS];  // reference L2L, multipole3d.cpp:213-223
        for (int k = 0; k < N; ++k)
            for (int j = k; j < N; ++j)
                for (int i = j; i < N; ++i)
                    tt[tr_base(N, k) + (j - k) * (N - k) + (i - k)] =
                        (double)(anm_l(i - j, 0) * anm_l(j, k) * (((i + j) & 1) ? -1.0L : 1.0L) / anm_l(i, k) *
                                 powl(rr, (ld)(i - j)) * ldexpl(1.0L, -(j + 1)));
    }
    // ---- per child-class M2L direction lists ----
    // Source-class-major order, parent offsets in a serpentine within a
    // class: consecutive steps of the M2L kernels read the same packed class
    // plane shifted by one parent box, so most of their source boxes are
    // still cache-resident (1-3% faster than lexicographic offset order on
    // the 4M workloads; ordering by rotation angle measured 10% slower).
    T.class_dirs.assign(8 * kDirsPerClass, -1);
    for (int c = 0; c < 8; ++c) {
        const int cx = (c >> 2) & 1, cy = (c >> 1) & 1, cz = c & 1;
        int cnt = 0;
        std::vector<std::pair<std::tuple<int, int, int, int>, int>> order;
        for (int d = 0; d < kNumM2LDirs; ++d) {
            const int x = T.dir_vec[4 * d], y = T.dir_vec[4 * d + 1], z = T.dir_vec[4 * d + 2];
            if (x < -2 - cx || x > 3 - cx || y < -2 - cy || y > 3 - cy || z < -2 - cz || z > 3 - cz) continue;
            const int ax = cx + x, ay = cy + y, az = cz + z;
            const int s = ((ax & 1) << 2) | ((ay & 1) << 1) | (az & 1);
            const int px = ax >> 1, py = ay >> 1, pz = az >> 1;
            const int ky = (px & 1) ? -py : py, kz = ((px + py) & 1) ? -pz : pz;
            order.push_back({std::make_tuple(s, px, ky, kz), d});
        }
        std::sort(order.begin(), order.end());
        for (auto& o : order) T.class_dirs[c * kDirsPerClass + cnt++] = o.second;
    }
    // ---- harmonic normalisation ----
    T.knm.assign(tri(N + 1), 0.0);
    for (int n = 0; n <= N; ++n)
        for (int m = 0; m <= n; ++m) T.knm[tri(n) + m] = (double)knm_l(n, m);
}

}  // namespace lfmm
