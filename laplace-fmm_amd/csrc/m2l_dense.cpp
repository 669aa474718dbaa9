// m2l_dense.cpp -- dense real form of the normalised M2L operators, and their
// split-FP16 packing for the matrix-core M2L kernel.
//
// The rotation-based M2L of one direction d (reference M2L, src/cpu/
// multipole3d.cpp:152-190: Rz(phi), Ry(theta), axial translation, Ry(-theta),
// Rz(-phi)) is a real-linear map on the N^2 real coordinates of an expansion:
//   index(n, 0)      = n^2          (Re; Im of m = 0 vanishes for real charges)
//   index(n, m, Re)  = n^2 + 2m - 1,   index(n, m, Im) = n^2 + 2m   (m >= 1)
// K_d is obtained column by column by applying the FP64 operator tables of
// build_op_tables (the same tables the FP64 kernel reads) to unit vectors in
// long double.  For the matrix cores K_d is scaled by one power of two and
// split into fp16 hi + lo parts (hi = fp16(x), lo = fp16(x - hi)), laid out
// as the A operand fragments of v_mfma_f32_32x32x16_f16: per (row tile of 32,
// K chunk of 16, part), lane l holds row l % 32 and the 8 consecutive K
// entries starting at 8 * (l / 32).
#include "lfmm_internal.h"

#include <cmath>
#include <cstring>
#include <complex>

namespace lfmm {

using ld = long double;
using cld = std::complex<ld>;

// Applies the normalised FP64 rotate-translate-rotate operator of direction
// `dir` with translation table `tsel` (tables of T) to one expansion, in long
// double, in the order of the FP64 kernel's phases.  M2L: dir < kNumM2LDirs
// and its own distance class; M2M from the child of class c: dir =
// kDiagBase + c and the M2M table (reference M2M, multipole3d.cpp:115-150).
static void apply_rtr(const OpTables& T, int dir, int tsel, const cld* in, cld* out)
{
    const int N = T.N;
    const int RS = rot_size(N), TS = tr_size(N);
    const double* rot = &T.rot[(size_t)T.dir_theta[dir] * RS * 2];
    const double* eph = &T.ephi[(size_t)dir * N * 2];
    const double* tt = &T.trans[(size_t)tsel * TS];
    std::vector<cld> s(tri(N));
    // Rz(phi) then Ry(theta), per degree: (P, Q) act on (Re, Im)
    for (int n = 0; n < N; ++n) {
        std::vector<cld> av(n + 1);
        for (int k = 0; k <= n; ++k) av[k] = (k == 0) ? in[tri(n)] : in[tri(n) + k] * cld(eph[2 * k], eph[2 * k + 1]);
        for (int m = 0; m <= n; ++m) {
            ld re = 0, im = 0;
            for (int k = 0; k <= n; ++k) {
                const double* pq = rot + (rot_base(n) + m * (n + 1) + k) * 2;
                re += (ld)pq[0] * av[k].real();
                im += (ld)pq[1] * av[k].imag();
            }
            s[tri(n) + m] = cld(re, im);
        }
    }
    // axial translation, per order m
    std::vector<cld> t2(tri(N));
    for (int m = 0; m < N; ++m) {
        const int L = N - m;
        for (int j = 0; j < L; ++j) {
            cld y = 0;
            for (int i = 0; i < L; ++i) y += (ld)tt[tr_base(N, m) + j * L + i] * s[tri(m + i) + m];
            t2[tri(m + j) + m] = y;
        }
    }
    // Ry(-theta): (-1)^(mu - m) (P, Q), then Rz(-phi)
    for (int j = 0; j < N; ++j)
        for (int mu = 0; mu <= j; ++mu) {
            ld re = 0, im = 0;
            for (int m = 0; m <= j; ++m) {
                const double* pq = rot + (rot_base(j) + mu * (j + 1) + m) * 2;
                const ld sg = ((mu - m) & 1) ? -1.0L : 1.0L;
                re += sg * (ld)pq[0] * t2[tri(j) + m].real();
                im += sg * (ld)pq[1] * t2[tri(j) + m].imag();
            }
            cld z(re, im);
            if (mu > 0) z *= std::conj(cld(eph[2 * mu], eph[2 * mu + 1]));
            out[tri(j) + mu] = z;
        }
}

int m2l_real_index(int n, int m, int part) { return n * n + (m == 0 ? 0 : 2 * m - 1 + part); }

static void rtr_dense_matrix(const OpTables& T, int dir, int tsel, double* K)
{
    const int N = T.N, NR = N * N;
    std::vector<cld> in(tri(N)), out(tri(N));
    for (int col = 0; col < NR; ++col) {
        std::fill(in.begin(), in.end(), cld(0, 0));
        const int n = (int)std::sqrt((double)col + 0.5);
        const int r = col - n * n;
        const int m = (r + 1) / 2;
        in[tri(n) + m] = (r == 0 || (r & 1)) ? cld(1, 0) : cld(0, 1);
        apply_rtr(T, dir, tsel, in.data(), out.data());
        for (int j = 0; j < N; ++j)
            for (int mu = 0; mu <= j; ++mu) {
                K[(size_t)m2l_real_index(j, mu, 0) * NR + col] = (double)out[tri(j) + mu].real();
                if (mu > 0) K[(size_t)m2l_real_index(j, mu, 1) * NR + col] = (double)out[tri(j) + mu].imag();
            }
    }
}

void m2l_dense_matrix(const OpTables& T, int dir, double* K) { rtr_dense_matrix(T, dir, T.dir_trans[dir], K); }

void m2m_dense_matrix(const OpTables& T, int cls, double* K) { rtr_dense_matrix(T, kDiagBase + cls, kTransM2M, K); }

static uint16_t half_bits(float x)
{
    const _Float16 h = (_Float16)x;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

// Splits nd dense operators K (nd x NR x NR) into the fragment layout and
// block prefixes of the matrix-core kernel, under one power-of-two scale.
static void pack16(int N, int nd, const std::vector<double>& K, std::vector<uint16_t>& out, double& kinv,
                   std::vector<int>& kpre)
{
    const int NR = N * N;
    const int RT = (NR + 31) / 32, KC = (NR + 15) / 16;
    double kmax = 0.0;
    for (double v : K) kmax = std::max(kmax, std::fabs(v));
    int e = 0;
    std::frexp(kmax, &e);  // kmax < 2^e
    const double sk = std::ldexp(1.0, 14 - e);
    kinv = 1.0 / sk;
    out.assign((size_t)nd * RT * KC * 2 * 64 * 8, 0);
    for (int d = 0; d < nd; ++d)
        for (int rt = 0; rt < RT; ++rt)
            for (int kc = 0; kc < KC; ++kc)
                for (int lane = 0; lane < 64; ++lane)
                    for (int s = 0; s < 8; ++s) {
                        const int row = rt * 32 + lane % 32, k = kc * 16 + 8 * (lane / 32) + s;
                        const double x = (row < NR && k < NR) ? K[((size_t)d * NR + row) * NR + k] * sk : 0.0;
                        const float xf = (float)x;
                        const _Float16 hi = (_Float16)xf;
                        const float lo = (float)(x - (double)(float)hi);
                        const size_t base = ((((size_t)d * RT + rt) * KC + kc) * 2) * 64 * 8;
                        out[base + lane * 8 + s] = half_bits((float)hi);
                        out[base + 64 * 8 + lane * 8 + s] = half_bits(lo);
                    }
    // block prefixes: a 32 x 16 block whose largest |entry| is below 2^-11 of
    // kmax keeps hi*hi only (its lo part and the hi*lo product are below
    // 2^-22 kmax max|X|); below 2^-22 kmax it contributes nothing measurable
    kpre.assign((size_t)nd * RT * 2, 0);
    const double t3 = std::ldexp(kmax, -11), t1 = std::ldexp(kmax, -22);
    for (int d = 0; d < nd; ++d)
        for (int rt = 0; rt < RT; ++rt) {
            int k3 = 0, k1 = 0;
            for (int kc = 0; kc < KC; ++kc) {
                double bm = 0.0;
                for (int i = rt * 32; i < std::min(NR, rt * 32 + 32); ++i)
                    for (int k = kc * 16; k < std::min(NR, kc * 16 + 16); ++k)
                        bm = std::max(bm, std::fabs(K[((size_t)d * NR + i) * NR + k]));
                if (bm >= t3) k3 = kc + 1;
                if (bm >= t1) k1 = kc + 1;
            }
            kpre[((size_t)d * RT + rt) * 2] = k3;
            kpre[((size_t)d * RT + rt) * 2 + 1] = std::max(k1, k3);
        }
}

void pack_m2l16(const OpTables& T, std::vector<uint16_t>& out, double& kinv, std::vector<int>& kpre)
{
    const int NR = T.N * T.N;
    std::vector<double> K((size_t)kNumM2LDirs * NR * NR);
    for (int d = 0; d < kNumM2LDirs; ++d) m2l_dense_matrix(T, d, &K[(size_t)d * NR * NR]);
    pack16(T.N, kNumM2LDirs, K, out, kinv, kpre);
}

void pack_m2m16(const OpTables& T, std::vector<uint16_t>& out, double& kinv, std::vector<int>& kpre)
{
    const int NR = T.N * T.N;
    std::vector<double> K((size_t)8 * NR * NR);
    for (int c = 0; c < 8; ++c) m2m_dense_matrix(T, c, &K[(size_t)c * NR * NR]);
    pack16(T.N, 8, K, out, kinv, kpre);
}

}  // namespace lfmm
