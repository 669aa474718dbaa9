// lfmm_internal.h -- shared declarations of the MI355X Laplace FMM library.
//
// Expansion storage convention (3D): coefficient (n, m), 0 <= m <= n < N, at
// index n(n+1)/2 + m as a double2 (re, im), boxes contiguous per level
// ([box][coef]).  Coefficients are stored NORMALISED per level:
//   multipole  Mt_nm = M_nm / h^n,   local  Lt_nm = L_nm * h^(n+1),
// where h is the box edge at that level and M, L are the reference's
// coefficients (src/cpu/multipole3d.cpp:20-49 defines M; the local expansion
// is evaluated as sum (2-delta_m0) Re(L_nm r^n K_nm P_n^m e^{i m phi}),
// :1152-1208).  With this scaling every translation operator depends only on
// the integer box offset, never on the level.
#pragma once
#include <cstdint>
#include <vector>

namespace lfmm {

constexpr int kMaxTerms3d = 18;
constexpr int kNumM2LDirs = 316;   // offsets in [-3,3]^3 minus [-1,1]^3
constexpr int kNumDiagDirs = 8;    // (+-1, +-1, +-1): child <-> parent
constexpr int kNumDirs = kNumM2LDirs + kNumDiagDirs;
constexpr int kDiagBase = kNumM2LDirs;
constexpr int kDirsPerClass = 189; // M2L offsets reachable from one child class
constexpr int kNumTransM2L = 16;   // index = distance class (|d|^2 classes)
constexpr int kTransM2M = kNumTransM2L;
constexpr int kTransL2L = kNumTransM2L + 1;
constexpr int kNumTrans = kNumTransM2L + 2;

constexpr int tri(int n) { return n * (n + 1) / 2; }
// sum_{n' < n} (n'+1)^2: start of degree n in a per-degree (n+1)x(n+1) table
constexpr int rot_base(int n) { return n * (n + 1) * (2 * n + 1) / 6; }
constexpr int rot_size(int N) { return rot_base(N); }
// translation table: per order k an (N-k)x(N-k) block, rows = output degree
constexpr int tr_base(int N, int k) { return rot_base(N) - rot_base(N - k); }
constexpr int tr_size(int N) { return rot_base(N); }

// Algorithmic FP64 flops of one rotation-based M2L with N terms: Rz and Ry
// forward and back, 2 (6 tri(N) + sum_{n<N} (n+1)(4n+2)), plus the axial
// translation, 4 sum_{t=1..N} t^2.
constexpr int64_t m2l_flops_per_op(int N)
{
    int64_t rot = 0, tr = 0;
    for (int n = 0; n < N; ++n) rot += (int64_t)(n + 1) * (4 * n + 2);
    for (int t = 1; t <= N; ++t) tr += (int64_t)t * t;
    return 2 * (6 * (int64_t)tri(N) + rot) + 4 * tr;
}

// Host-built FP64 operator tables for one expansion size N.
struct OpTables {
    int N = 0;
    int num_theta = 0;
    std::vector<double> rot;       // [theta][rot_size][2]: (P, Q) rotation pair
    std::vector<double> ephi;      // [dir][N][2]: (cos k phi, sin k phi)
    std::vector<double> trans;     // [kNumTrans][tr_size]
    std::vector<int> dir_theta;    // [kNumDirs]
    std::vector<int> dir_trans;    // [kNumDirs] (M2L class; -1 for diagonals)
    std::vector<int> dir_vec;      // [kNumDirs][4]: dx, dy, dz, 0
    std::vector<int> class_dirs;   // [8][kDirsPerClass]
    std::vector<double> knm;       // [tri(N + 1)]
};

void build_op_tables(int N, OpTables& T);

// Host-built FP64 operator tables of the 2D FMM for N terms (complex values
// as (re, im) pairs; see tables2d.cpp for the normalisation).
constexpr int kMaxTerms2d = 30;
constexpr int kNumM2LDirs2d = 40;   // offsets in [-3,3]^2 minus [-1,1]^2
constexpr int kDirsPerClass2d = 27; // M2L offsets reachable from one child class
struct OpTables2d {
    int N = 0;
    std::vector<int> dir_vec;      // [40][2]
    std::vector<int> class_dirs;   // [4][27]
    std::vector<double> m2l;       // [40][N][N] complex: out l, in j
    std::vector<double> m2m;       // [4][N][N] complex (child class c)
    std::vector<double> l2l;       // [4][N][N] complex (child class c)
};
void build_op_tables_2d(int N, OpTables2d& T);

// Dense real form of the M2L operator of direction `dir` (< kNumM2LDirs):
// K[row * N^2 + col], rows and columns in m2l_real_index order.
int m2l_real_index(int n, int m, int part);
void m2l_dense_matrix(const OpTables& T, int dir, double* K);
// All M2L operators scaled by 1/kinv (a power of two) and split into fp16
// hi/lo A-operand fragments of the matrix-core M2L kernel, and per
// (direction, row tile of 32) the block prefixes (k3, k1): K chunks below k3
// need all three split products, below k1 the hi*hi product alone.
void pack_m2l16(const OpTables& T, std::vector<uint16_t>& out, double& kinv, std::vector<int>& kpre);
// The same for the M2M from the child of class c (c = 0..7, the 8
// "directions"; reference M2M, multipole3d.cpp:115-150), under its own scale.
void m2m_dense_matrix(const OpTables& T, int cls, double* K);
void pack_m2m16(const OpTables& T, std::vector<uint16_t>& out, double& kinv, std::vector<int>& kpre);

// Small host helpers exposed for tests through the C ABI debug entry points.
double host_knm(int n, int m);
double host_anm(int n, int m);
void host_wigner_small_d(int j, double beta, double* out); // (2j+1)^2, [mp][m]

}  // namespace lfmm
