/*
 * fmm_oracle.h -- CPU restatement of the reference laplace-fmm evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (laplace-fmm_amd/csrc) never links it and
 * never falls back to it.
 *
 * It restates, in plain C99 + OpenMP, the algorithm of the reference's CPU
 * path (the reference itself cannot be built here: it needs oneTBB headers and
 * libcu++ <cuda/std/complex>, neither of which exists in this image):
 *   src/common/special_functions.{h,cpp}  Legendre, Knm, Anm, m2l coefficients,
 *                                         Wigner d, direction tables
 *   src/common/simple_math.h              DecToSph, pair kernels, direct sum,
 *                                         ReadError statistics
 *   src/cpu/morton_tree.cpp               extent, Morton keys, stable counting
 *                                         sort, levels, partition, lists
 *   src/cpu/multipole3d.cpp               P2M, M2M, M2L, L2L, L2P, P2P, Solve
 *   src/cpu/multipole.cpp                 2D complex-series FMM
 * Each function in fmm_oracle.c cites the reference file:line it follows.
 *
 * Conventions: N = number of expansion terms (order p = N-1); 3D coefficient
 * (n,m), 0<=m<=n<N, lives at n(n+1)/2+m as interleaved (re,im) doubles.
 */
#ifndef LFMM_ORACLE_H
#define LFMM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_TERMS_3D 18
#define ORC_MAX_TERMS_2D 30
#define ORC_AUTO 2147483647
#define ORC_MAX_DEPTH 21

/* ---- math tables (special_functions.h:12-138, special_functions.cpp:194-381) ---- */
void orc_init(void);                                   /* InitMathConstants(18) */
void orc_knm(int N, double* out);                      /* N(N+1)/2 */
void orc_anm(int N, double* out);                      /* 2N(2N+1)/2 */
double orc_m2lcoef(int j, int k, int n);
void orc_pnm(int N, double theta, double* out);        /* N(N+1)/2 */
void orc_wignerd(int N, double theta, double* out);    /* N(N+1)(4N-1)/6 */
void orc_dec_to_sph(const double* v, double* out);     /* r, theta, phi */
/* Radius offset of DecToSph: 1e-12 reproduces the reference (default); 0
 * gives the exact geometry (used to pin exactness properties tightly). */
void orc_set_r_eps(double eps);
double orc_get_r_eps(void);
int  orc_doublehash(double v);
int  orc_dmatrix_lookup(double theta, int N, double* out); /* via the hash map; 0 = found */

/* ---- single operators (multipole3d.cpp:20-226, 1152-1208) ---- */
void orc_p2m3d(int N, int64_t n, const double* xyz, const double* q, const double* centre, double* M);
void orc_m2m3d(int N, const double* a, const double* p0, double* b);   /* b += */
void orc_m2l3d(int N, const double* a, const double* p0, double* b);   /* b += */
void orc_l2l3d(int N, const double* a, const double* p0, double* b);   /* b += */
void orc_l2p3d(int N, const double* L, const double* centre, int64_t n, const double* xyz,
               double* pot, double* field);                            /* += */

/* ---- tree (morton_tree.cpp:48-493) ---- */
typedef struct orc_tree3d orc_tree3d;
orc_tree3d* orc_tree3d_build(int64_t n, const double* xyz, const double* q, int depth);
orc_tree3d* orc_tree2d_build(int64_t n, const double* xy, const double* q, int depth);
void orc_tree3d_free(orc_tree3d* t);
int  orc_tree3d_depth(const orc_tree3d* t);
void orc_tree3d_info(const orc_tree3d* t, int64_t* level_sizes, double* shift_scale, int64_t* targets_num);
void orc_tree3d_particles(const orc_tree3d* t, uint64_t* keys, int64_t* pos_map, double* xyzq);
void orc_tree3d_level(const orc_tree3d* t, int level, uint64_t* keys, int64_t* src_range,
                      int64_t* tgt_range, int64_t* parent, double* centres);
int64_t orc_tree3d_list_size(const orc_tree3d* t, int level, int far);
void orc_tree3d_lists(const orc_tree3d* t, int level, int far, int64_t* offsets, int64_t* idx);

/* ---- whole solves ---- */
/* Reference FastMultipole3d<double>: n_tgt == 0 selects the single-set
 * constructor (multipole3d.cpp:1296-1302), otherwise the source/target one
 * (:1311-1319; targets are given q = 0).  N / depth may be ORC_AUTO with eps.
 * timings (may be NULL) receives [tree, p2m, m2m, m2l+l2l, leaves, total] s. */
int orc_fmm3d(int64_t n_src, const double* src_xyz, const double* src_q,
              int64_t n_tgt, const double* tgt_xyz,
              double eps, int N, int depth, double* pot, double* field, double* timings);
/* The same with per-box truncation orders requested per output particle
 * (pterms, input order; see fmm3d_impl in fmm_oracle.c). */
int orc_fmm3d_orders(int64_t n_src, const double* src_xyz, const double* src_q, int64_t n_tgt, const double* tgt_xyz,
                     const int32_t* pterms, int depth, double* pot_out, double* field_out);
int orc_fmm2d(int64_t n_src, const double* src_xy, const double* src_q,
              int64_t n_tgt, const double* tgt_xy,
              double eps, int N, int depth, double* pot, double* field, double* timings);

/* FP64 direct sums (simple_math.h:119-139, 218-281).  include_self != 0
 * reproduces the reference's clamp-including self term; 0 skips i == j
 * (only meaningful when the target set is the source set, tgt == NULL). */
void orc_direct3d(int64_t n_src, const double* src_xyz, const double* src_q,
                  int64_t n_tgt, const int64_t* tgt_index, const double* tgt_xyz,
                  int include_self, double* pot, double* field);
void orc_direct2d(int64_t n_src, const double* src_xy, const double* src_q,
                  int64_t n_tgt, const int64_t* tgt_index, const double* tgt_xy,
                  int include_self, double* pot, double* field);

/* ReadError statistics (simple_math.h:342-402): out = [max, l2 (rms of
 * relative), sum|e|/sum|x|].  dim = 1 (scalar) or 3 (vector norms) or 2 (complex). */
void orc_read_error(int64_t n, int dim, const double* exact, const double* res, double* out);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
