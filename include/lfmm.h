/*
 * lfmm.h -- C ABI of the MI355X-native Laplace FMM evaluation path.
 *
 * The reference exposes this path as C++ classes that compute on
 * construction; there is no foreign-function layer in the reference, so the
 * entry points below are exactly what a binding of those constructors needs:
 *
 *   lfmm_fmm3d  replaces fmm::gpu::FastMultipole3d<double>(particles, eps, N, tree_depth)
 *               src/cuda/multipole3d.h:16 (and its CPU twin src/cpu/multipole3d.h:19-20),
 *               and, with targets, the (sources, targets, ...) constructor
 *               src/cuda/multipole3d.h:17 / src/cpu/multipole3d.h:21.
 *               Outputs match the public members `potentials` and `forces`
 *               (src/cuda/multipole3d.h:20-21): potential and
 *               field = sum q_j (r_i - r_j)/|r_i - r_j|^3 in input order.
 *
 * The C++ drop-in classes in include/fmm/ wrap these calls with the
 * reference's constructor signatures.
 *
 * Threading: one context per host thread (calls on one context serialise on
 * its stream).  No C++ exceptions cross this boundary; every entry point
 * returns an lfmm_status.
 */
#ifndef LFMM_H
#define LFMM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same value as the reference's fmm::FMM_AUTO (src/common/defs.h:22). */
#define LFMM_AUTO 2147483647

/* Expansion-term caps of the reference (src/common/defs.h:31-32). */
#define LFMM_MAX_TERMS_3D 18
#define LFMM_MAX_TERMS_2D 30

typedef enum lfmm_status {
    LFMM_OK = 0,
    LFMM_ERR_INVALID = 1,      /* null pointer, negative count, bad stride */
    LFMM_ERR_TERMS = 2,        /* expansion terms outside [1, cap]         */
    LFMM_ERR_DEPTH = 3,        /* tree depth outside the supported range   */
    LFMM_ERR_NOMEM = 4,        /* device allocation failed                 */
    LFMM_ERR_HIP = 5,          /* a HIP runtime call failed                */
    LFMM_ERR_NONFINITE = 6,    /* NaN or Inf in positions or charges       */
    LFMM_ERR_NODEVICE = 7,     /* no usable GPU                            */
    LFMM_ERR_UNSUPPORTED = 8,  /* option combination not implemented       */
    LFMM_ERR_TOO_LARGE = 9     /* input beyond a fixed scratch size         */
} lfmm_status;

typedef enum lfmm_memory {
    LFMM_MEM_HOST = 0,   /* all array arguments are host pointers   */
    LFMM_MEM_DEVICE = 1  /* all array arguments are device pointers */
} lfmm_memory;

/* Flags (bitwise or). */
#define LFMM_FLAG_EXCLUDE_SELF 1u  /* skip the i == j pair; the reference includes
                                      q_i / FORCE_EPS in every potential
                                      (src/common/simple_math.h:130-133) */
#define LFMM_FLAG_TIMINGS 2u       /* fill the per-stage timings of the report
                                      (adds stream synchronisations) */
#define LFMM_FLAG_COUNT_OPS 4u     /* count M2L and near-field pair operations
                                      into the report (extra kernels) */
#define LFMM_FLAG_MIXED 8u         /* M2L at levels >= 4 on the matrix cores in
                                      split FP16 (22-bit operands, FP32 sums);
                                      levels 2-3, L2L, P2M, M2M, L2P, P2P stay
                                      FP64.  Adds ~1e-7 relative error.      */
#define LFMM_FLAG_ADAPTIVE 16u     /* 3D: adaptive octree.  A box is split only
                                      while it holds more than leaf_max
                                      particles (64 when leaf_max <= 0), down
                                      to the uniform depth of that bound (or
                                      tree_depth); leaves sit on any level and
                                      the solve adds P2L (X lists) and M2P
                                      (W lists) to M2L and P2P.  The reference
                                      has uniform trees only.                 */

#define LFMM_FLAG_LEARNED_ORDER 32u /* 3D: per-box expansion orders chosen by
                                      the context's order network
                                      (lfmm_set_order_model) on the adaptive
                                      octree (implies LFMM_FLAG_ADAPTIVE).  A
                                      sampled FP64 direct sum checks every run;
                                      above the model's tolerance the solve is
                                      rerun with p = 16 (report->fallback).
                                      terms and eps are then ignored.        */

typedef struct lfmm_ctx lfmm_ctx;

/* Strided particle view.  Particle i has position
 * (pos[i*pos_stride], pos[i*pos_stride+1], pos[i*pos_stride+2]) and charge
 * q[i*q_stride] (strides in doubles).  Packed SoA uses strides 3 and 1; the
 * reference's 40-byte particle {Vector3d center; double q; size_t morton}
 * (src/common/utils.h:26-32) uses pos = &p[0].center[0], q = &p[0].q and
 * strides 5 and 5. */
typedef struct lfmm_points3d {
    int64_t n;
    const double* pos;
    int64_t pos_stride;
    const double* q;      /* may be NULL for targets */
    int64_t q_stride;
} lfmm_points3d;

/* 2D particle view: the same layout with two coordinates per point,
 * (pos[i*pos_stride], pos[i*pos_stride+1]).  The reference's 32-byte particle
 * {std::complex<double> center; double q; size_t morton} (src/common/utils.h:
 * 26-32, 149-151) uses pos = &p[0].center, q = &p[0].q and strides 4 and 4. */
typedef lfmm_points3d lfmm_points2d;

typedef struct lfmm_report {
    int32_t terms;          /* N: expansion terms used (order p = N - 1)  */
    int32_t tree_depth;     /* number of levels; leaves at tree_depth - 1 */
    int64_t num_leaves;
    int64_t max_leaf;       /* largest leaf occupancy                     */
    int64_t m2l_ops;        /* M2L translations performed (with LFMM_FLAG_COUNT_OPS, else -1) */
    int64_t p2p_pairs;      /* near-field pair evaluations (with LFMM_FLAG_COUNT_OPS, else -1) */
    double t_tree_ms, t_upward_ms, t_downward_ms, t_leaf_ms, t_total_ms;
    /* With LFMM_FLAG_TIMINGS: summed device durations (HIP events on the
     * solve's stream) of all M2L kernels (FP64 M2L+L2L, and with
     * LFMM_FLAG_MIXED the matrix-core M2L), their launch count, and of the
     * leaf step (L2P and the near-field pair kernels). */
    double t_m2l_kernels_ms;
    int32_t m2l_launches;
    int32_t m2l_mixed_launches;  /* matrix-core M2L launches (LFMM_FLAG_MIXED) */
    double t_leaf_kernel_ms;
    double t_m2l_mixed_ms;       /* matrix-core M2L kernels alone (timings)    */
    int64_t m2l_ops_mixed;       /* M2L translations on the matrix cores (count_ops) */
    int64_t num_boxes;           /* boxes of all levels                        */
    int64_t xw_pairs;            /* LFMM_FLAG_ADAPTIVE: X-list (P2L) pairs, equal
                                    to the W-list (M2P) pairs                  */
    /* Per-box orders (LFMM_FLAG_LEARNED_ORDER, lfmm_fmm3d_orders): terms is
     * the largest order of any box, terms_min the smallest leaf order. */
    int32_t terms_min;
    int32_t fallback;            /* 1: the learned orders missed the sampled
                                    check and the solve was rerun at p = 16   */
    double check_rel_l2;         /* sampled check of a learned-order solve:   */
    double check_rel_l2_field;   /* rel-L2 of potential and field vs FP64 direct
                                    sums (self pair excluded)                 */
    double m2l_flops;            /* with LFMM_FLAG_COUNT_OPS: algorithmic FP64
                                    flops of all M2L operations at their orders */
    /* With LFMM_FLAG_TIMINGS, summed device durations per kernel kind: P2M,
     * M2M, the standalone L2L of the matrix-core levels, P2L (X lists), L2P
     * (with the W lists' M2P) and the near-field pair step. */
    double t_p2m_ms, t_m2m_ms, t_l2l_ms, t_p2l_ms, t_l2p_ms, t_p2p_ms;
    int64_t p2p_source_reads;    /* with LFMM_FLAG_COUNT_OPS: particles of every
                                    leaf's near list, summed over the leaves  */
    double m2m_bytes, l2l_bytes; /* algorithmic HBM bytes of the M2M and the
                                    standalone L2L launches (every expansion
                                    read once, written once)                  */
} lfmm_report;

/* Create a context bound to a GPU (device index as HIP numbers them). */
int lfmm_create(lfmm_ctx** ctx, int device);
void lfmm_destroy(lfmm_ctx* ctx);
const char* lfmm_strerror(int status);
/* Use a caller-owned hipStream_t (NULL restores the context's own stream;
 * LFMM_STREAM_DEFAULT selects the device's null stream, the handle 0 that
 * e.g. torch reports for its default stream).  Work queued on the previous
 * stream comes before anything the context queues afterwards. */
#define LFMM_STREAM_DEFAULT ((void*)1)
int lfmm_set_stream(lfmm_ctx* ctx, void* hip_stream);
/* Release cached device workspace (it is otherwise reused across calls). */
int lfmm_trim(lfmm_ctx* ctx);

/*
 * 3D Laplace FMM.  targets == NULL (or targets->n == 0) is the single-set
 * solve: every source is also a target and outputs are per source.
 *   eps, terms, tree_depth: the reference's constructor arguments; terms
 *     and tree_depth accept LFMM_AUTO and then follow the reference's
 *     formulas (src/cpu/multipole3d.cpp:1328-1337).  terms is clamped to 18.
 *   leaf_max > 0 (used only when tree_depth == LFMM_AUTO): pick the smallest
 *     uniform depth whose fullest leaf holds at most leaf_max particles.
 *   potential: n_out doubles; field: n_out*3 doubles (x,y,z per particle);
 *     either may be NULL.  n_out = targets ? targets->n : sources->n.
 *   report: optional.
 * With targets (charges ignored), the solve follows the reference's
 * source/target constructor: targets are treated as zero-charge particles of
 * one tree (src/cpu/multipole3d.cpp:1311-1319) and get outputs; sources get
 * none.  In a single-set solve every particle gets an output, also when some
 * charges are exactly zero (the reference then evaluates only the q = 0
 * particles, src/cpu/multipole3d.cpp:878; deliberate deviation).
 */
int lfmm_fmm3d(lfmm_ctx* ctx, const lfmm_points3d* sources, const lfmm_points3d* targets,
               double eps, int terms, int tree_depth, int leaf_max,
               int memory, unsigned flags,
               double* potential, double* field, lfmm_report* report);

/*
 * The Morton octree of a particle set, as built inside lfmm_fmm3d.  Replaces
 * the reference's MortonTree<3, point3d, double> (src/cpu/morton_tree.h:24-45,
 * GPU twin src/cuda/morton_tree.h): positions_map, level_sizes and the cell
 * keys / ranges of every level.  Inputs are host pointers; with charges,
 * q = 0 particles follow the others inside each leaf (morton_tree.cpp:426-468).
 * Outputs are host arrays, any of which may be NULL:
 *   depth_out      number of levels (leaves at depth-1)
 *   level_sizes    [32] boxes per level
 *   geom           [4] {shift_x, shift_y, shift_z, scale} (morton_tree.cpp:293-301)
 *   pos_map        [n] input index -> sorted position (stable key order)
 *   keys           [sum level_sizes] Morton keys, level 0 first
 *   begins         [sum (level_sizes+1)] per level: the start of each box's
 *                  range (particles at the leaves, children above) plus the end
 */
int lfmm_tree3d(lfmm_ctx* ctx, const lfmm_points3d* particles, int tree_depth, int leaf_max,
                int32_t* depth_out, int64_t* level_sizes, double* geom, int64_t* pos_map,
                uint64_t* keys, int64_t* begins);

/*
 * Per-box expansion orders (no reference counterpart: the reference uses one
 * expansion size per solve, src/cpu/multipole3d.cpp:1328-1331).
 *
 * lfmm_fmm3d_orders: lfmm_fmm3d on the adaptive octree with the order of
 *   every leaf set to the largest request of its particles:
 *   particle_terms[i] (terms N = p + 1 for target i: n_out entries, memory as
 *   the outputs), clamped to [1, 18].  A box with children takes the largest
 *   order of its children; from level 4 on the M2L and L2L into a box run
 *   at that box's own order (levels 2-3, P2M and M2M use the largest order,
 *   the sources of a translation are read up to the target's order).
 *   report->terms is the largest order, report->terms_min the smallest leaf
 *   order.
 *
 * The order network: 8 per-leaf features -> 64 -> 64 -> 13 classes (FP32),
 * class c = order p = 4 + c.  Its parameter block holds LFMM_ORDER_MODEL_SIZE
 * floats: feature normalisation mu[8], sd[8], then W1[64][8], b1[64],
 * W2[64][64], b2[64], W3[13][64], b3[13] (row-major, out x in).  Features of a
 * leaf: log2(1 + particles), level / 16, log10 sum|q|, log10(|dipole| /
 * (|monopole| + |dipole|)) about the leaf centre in box units, V-list size /
 * 189, log10(V sources' sum|q| / the leaf's), log2(1 + near-field
 * particles), log2(1 + |W| + |X|).  A leaf takes the smallest order whose
 * cumulative class probability reaches `confidence`.
 *
 * lfmm_set_order_model: copies the model to the device (NULL removes it);
 *   check_tol: sampled rel-L2 potential error above which a learned-order
 *   solve is rerun at p = 16; check_samples targets (evenly spaced in input
 *   order) are summed directly.
 * lfmm_order_features: the adaptive octree of a particle set (as built by a
 *   solve with the same tree_depth and leaf_max) with every leaf's features
 *   (features [leaves * 8]) and every particle's leaf (particle_leaf [n],
 *   input order, leaves numbered in Morton order).  *nleaf receives the leaf
 *   count; call with NULL arrays to size them.
 * lfmm_order_predict: class probabilities (probs [n * 13]) and/or chosen
 *   terms (terms [n]) of n feature rows.
 * lfmm_order_loss_grad: mean cross-entropy of n labelled rows (labels =
 *   classes 0..12) and its gradient (grad [LFMM_ORDER_MODEL_SIZE]; zero for
 *   the normalisation entries).
 * lfmm_order_train: Adam on the mean cross-entropy, `epochs` passes over the
 *   rows in batches of `batch`, shuffled by `seed`; model is updated in
 *   place; loss_history (optional) receives the mean loss of each epoch.
 * All order-network arrays are host pointers; the work runs on the GPU.
 */
#define LFMM_ORDER_FEATURES 8
#define LFMM_ORDER_CLASSES 13
#define LFMM_ORDER_MIN_TERMS 5
#define LFMM_ORDER_MODEL_SIZE 5597
int lfmm_fmm3d_orders(lfmm_ctx* ctx, const lfmm_points3d* sources, const lfmm_points3d* targets,
                      const int32_t* particle_terms, int tree_depth, int leaf_max, int memory, unsigned flags,
                      double* potential, double* field, lfmm_report* report);
int lfmm_set_order_model(lfmm_ctx* ctx, const float* model, double confidence, double check_tol, int check_samples);
int lfmm_order_features(lfmm_ctx* ctx, const lfmm_points3d* particles, int tree_depth, int leaf_max, int64_t* nleaf,
                        float* features, int32_t* particle_leaf);
int lfmm_order_predict(lfmm_ctx* ctx, const float* model, const float* features, int64_t n, double confidence,
                       float* probs, int32_t* terms);
int lfmm_order_loss_grad(lfmm_ctx* ctx, const float* model, const float* features, const int32_t* labels, int64_t n,
                         float* loss, float* grad);
int lfmm_order_train(lfmm_ctx* ctx, float* model, const float* features, const int32_t* labels, int64_t n,
                     int epochs, int batch, double lr, uint64_t seed, float* loss_history);

/*
 * The adaptive octree of a 3D particle set and its interaction lists, as
 * lfmm_fmm3d builds them with LFMM_FLAG_ADAPTIVE (no reference counterpart:
 * the reference's MortonTree, src/cpu/morton_tree.h:24-45, is uniform-depth).
 * Host pointers.  counts[4] receives {boxes, leaves, U entries, X pairs};
 * every other output may be NULL (call once with NULLs to size the arrays;
 * the build is deterministic).  Boxes are numbered level by level, each level
 * in key order; leaves in particle (Morton) order.
 *   box_level [boxes], box_key [boxes], box_range [2 * boxes] (sorted
 *   particle range), leaf_box [leaves] (box of each leaf);
 *   U: u_off [leaves + 1], u_leaf [U entries]   leaves touching each leaf
 *   X: x_off [boxes + 1],  x_leaf [X pairs]     coarser leaves whose particles
 *                                               feed the box's local (P2L)
 *   W: w_off [leaves + 1], w_box [X pairs]      finer boxes whose multipoles
 *                                               act on the leaf (M2P)
 */
int lfmm_tree3d_adaptive(lfmm_ctx* ctx, const lfmm_points3d* particles, int tree_depth, int leaf_max,
                         int64_t* counts, int32_t* depth_out, double* geom, int32_t* box_level, uint64_t* box_key,
                         int32_t* box_range, int32_t* leaf_box, int64_t* u_off, int32_t* u_leaf, int64_t* x_off,
                         int32_t* x_leaf, int64_t* w_off, int32_t* w_box);

/*
 * 2D Laplace FMM with complex-series expansions; replaces the constructors of
 * fmm::gpu::FastMultipole (src/cuda/multipole.h:11-12; CPU twin
 * src/cpu/multipole.h:13-15).  Arguments as lfmm_fmm3d, except:
 *   terms: LFMM_AUTO follows src/cpu/multipole.cpp:478-481, clamped to 30;
 *   tree_depth: LFMM_AUTO follows src/cpu/multipole.cpp:483-485 (at least 3);
 *   potential: n_out doubles, sum q log|z - z_j| (self pair clamped at 1e-4);
 *   field: n_out*2 doubles, (re, im) of the complex force sum q / (z - z_j),
 *     the reference's std::complex<double> forces;
 *   LFMM_FLAG_MIXED: the M2L of levels >= 4 forms each direction's sum in
 *     FP32 (FP32 operator, sources rounded on load) and adds it to the FP64
 *     local; levels 2-3, L2L, P2M, M2M, L2P and P2P stay FP64.
 */
int lfmm_fmm2d(lfmm_ctx* ctx, const lfmm_points2d* sources, const lfmm_points2d* targets,
               double eps, int terms, int tree_depth, int leaf_max,
               int memory, unsigned flags,
               double* potential, double* field, lfmm_report* report);

/* The Morton quadtree of a 2D particle set, as lfmm_tree3d (reference
 * MortonTree<2, point2d, double>); keys interleave x << 1 | y. */
int lfmm_tree2d(lfmm_ctx* ctx, const lfmm_points2d* particles, int tree_depth, int leaf_max,
                int32_t* depth_out, int64_t* level_sizes, double* geom, int64_t* pos_map,
                uint64_t* keys, int64_t* begins);

/*
 * FP64 direct sums over all sources for every target: replaces
 * fmm::gpu::ComputeExact<InteractionType> (src/common/simple_math.cu:24-130;
 * CPU twin fmm::ComputeExact, src/common/simple_math.h:218-345), which the
 * reference's drivers call to check an FMM run (src/examples/main.cpp:129-130,
 * 183-184).  Pair terms as the FMM's (src/common/simple_math.h:120-151):
 *   3D  potential q / max(|r|, 1e-4), field q r / max(|r|, 1e-4)^3;
 *   2D  potential q * 0.5 log max(|dz|^2, 1e-8), force q conj(dz) / max(|dz|^2, 1e-8),
 * r = x_i - x_j, summed over j in input order.  The self pair is included (as
 * in the reference) unless LFMM_FLAG_EXCLUDE_SELF is given in a single-set
 * call (targets NULL); other flags are ignored.  Outputs as lfmm_fmm3d /
 * lfmm_fmm2d (field n_out*3 in 3D, n_out*2 = complex force in 2D).  Cost is
 * n_out * n_sources pair terms: a checking tool, not a solver.
 */
int lfmm_direct3d(lfmm_ctx* ctx, const lfmm_points3d* sources, const lfmm_points3d* targets, int memory,
                  unsigned flags, double* potential, double* field);
/* lfmm_direct3d_noalloc: lfmm_direct3d with device memory that allocates
 * nothing (no reference counterpart): for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  Its
 * packed inputs go to a fixed 32 MiB array per GPU (1,048,576 points:
 * sources plus targets; more returns LFMM_ERR_TOO_LARGE), held until the
 * process ends.  All calls on one GPU share that array: calls that overlap
 * in time on one GPU race.  Runs on ctx's stream and returns when done. */
int lfmm_direct3d_noalloc(lfmm_ctx* ctx, const lfmm_points3d* sources, const lfmm_points3d* targets,
                          unsigned flags, double* potential, double* field);
int lfmm_direct2d(lfmm_ctx* ctx, const lfmm_points2d* sources, const lfmm_points2d* targets, int memory,
                  unsigned flags, double* potential, double* field);

/*
 * Vector charges: replaces fmm::gpu::FastMultipole3d<Vector3d>
 * (src/cuda/multipole3d.h:7-42 with value = Vector3d; kernels
 * src/cuda/multipole3d.cu:633-839, CPU twin src/cpu/multipole3d.cpp:414-853).
 * Each particle carries a charge vector (q[i*q_stride], q[i*q_stride+1],
 * q[i*q_stride+2]); the reference's 56-byte particle3d3 {Vector3d center;
 * Vector3d q; size_t morton} uses pos = &p[0].center[0], q = &p[0].q[0] and
 * strides 7 and 7.  Output: force n_out*3 doubles,
 *   F_i = sum_j q_j x (r_i - r_j) / max(|r_i - r_j|, 1e-4)^3
 * (src/common/simple_math.h:148-152); no potentials, as in the reference.
 * Computed as three scalar solves (one per charge component, the same tree)
 * whose fields are combined by a cross product, as the reference's
 * forces_cross (src/cuda/multipole3d.cu:765-776).  Other arguments as
 * lfmm_fmm3d (the report sums the timings of the three solves).
 * lfmm_direct3d_vec: the same sums directly, in FP64 (reference
 * ComputeExact<InteractionType::Force3d> on particle3d3).
 */
int lfmm_fmm3d_vec(lfmm_ctx* ctx, const lfmm_points3d* sources, const lfmm_points3d* targets,
                   double eps, int terms, int tree_depth, int leaf_max,
                   int memory, unsigned flags, double* force, lfmm_report* report);
int lfmm_direct3d_vec(lfmm_ctx* ctx, const lfmm_points3d* sources, const lfmm_points3d* targets, int memory,
                      double* force);

/*
 * Many independent single-set 3D systems in one call.  The reference has one
 * system per fmm::gpu::FastMultipole3d construction (src/cuda/multipole3d.h:16,
 * on the legacy default stream); its drivers construct one object per
 * particle set (src/examples/main.cpp:190, 194).  This call replaces a loop
 * of such constructions over many systems: system s is
 * particles offsets[s] .. offsets[s+1]-1 of the packed arrays
 *   pos      [3 * offsets[n_sys]]  x, y, z per particle
 *   q        [offsets[n_sys]]
 * and its outputs go to the same rows of potential [offsets[n_sys]] and
 * field [3 * offsets[n_sys]] (either may be NULL).  offsets (n_sys + 1
 * entries, non-decreasing) is always a host array; the others follow
 * `memory`.  eps, terms, tree_depth, leaf_max and flags apply to every system
 * as in lfmm_fmm3d.  reports: NULL or n_sys entries.
 * workers: worker contexts (own stream, workspace, host thread) that take
 * the systems largest first; 0 = LFMM_BATCH_DEFAULT_WORKERS, at most
 * LFMM_BATCH_MAX_WORKERS.  They are kept in ctx until lfmm_destroy.  With
 * device memory the workers first wait for work queued on ctx's stream.  The
 * call returns when every system is done; the first failing status is
 * returned and the remaining systems are not started.
 * With LFMM_FLAG_ADAPTIVE every system is solved on its own adaptive octree
 * (leaf_max as lfmm_fmm3d) and all of them in one solve on ctx's stream: the
 * systems are mapped into separate cells of one cube, each by its own extent
 * (the reference's rule), with empty cells between them, so that no
 * interaction list crosses from one system to another; each system's results
 * equal its own adaptive lfmm_fmm3d solve up to rounding (FP64: rel-L2
 * ~1e-12).  workers and tree_depth are then ignored, every report holds the
 * one solve's figures, and LFMM_FLAG_LEARNED_ORDER is not supported.
 */
#define LFMM_BATCH_DEFAULT_WORKERS 8
#define LFMM_BATCH_MAX_WORKERS 16
int lfmm_fmm3d_batch(lfmm_ctx* ctx, int n_sys, const int64_t* offsets, const double* pos, const double* q,
                     double eps, int terms, int tree_depth, int leaf_max, int memory, unsigned flags, int workers,
                     double* potential, double* field, lfmm_report* reports);

/* Result kinds, in the order of the reference's fmm::InteractionType
 * (src/common/utils.h:19-24). */
typedef enum lfmm_interaction {
    LFMM_POTENTIAL2D = 0,  /* one double per value                  */
    LFMM_FORCE2D = 1,      /* complex: (re, im)                     */
    LFMM_POTENTIAL3D = 2,  /* one double per value                  */
    LFMM_FORCE3D = 3       /* vector: (x, y, z)                     */
} lfmm_interaction;

/* The reference's exact-result files ("potential3d_exact.txt", ...; names
 * from src/common/simple_math.cu:103-119).  lfmm_write_exact writes n values
 * one per line with 20 significant digits, a complex as "(re,im)" and a
 * vector as "x y z" (src/common/simple_math.cu:121-129, utils.h:108-109);
 * lfmm_read_exact reads them back (src/common/simple_math.h:374-376) and
 * returns the count read or a negative status.  Host memory; no context. */
const char* lfmm_exact_filename(int kind);
int lfmm_write_exact(const char* path, int kind, int64_t n, const double* values);
int64_t lfmm_read_exact(const char* path, int kind, int64_t n, double* values);
/* The figures of the reference's ReadError (src/common/simple_math.h:378-401):
 * max |e|, the "l2 error" sqrt(mean((|e_i| / |exact_i|)^2)) and the
 * "relative error" sum |e_i| / sum |exact_i|, |.| the absolute value, complex
 * modulus or Euclidean norm.  Any output pointer may be NULL. */
int lfmm_error_stats(int kind, int64_t n, const double* exact, const double* result, double* max_err,
                     double* l2_err, double* rel_err);

/* Inspection hooks of the host FP64 table builder (no reference
 * counterpart; used by the CPU test suite to pin the tables without a GPU).
 * lfmm_host_wigner_d: standard small Wigner matrix d^j(beta), row-major
 * [m'+j][m+j].  lfmm_host_op_tables: copies table `which` (0 rot, 1 ephi,
 * 2 trans, 3 dir_theta, 4 dir_trans, 5 dir_vec, 6 class_dirs, 7 knm) for N
 * terms into out (if capacity suffices) and returns its element count, or a
 * negative status. */
int lfmm_host_wigner_d(int j, double beta, double* out);
int64_t lfmm_host_op_tables(int N, int which, void* out, int64_t capacity);
/* lfmm_host_m2l_dense: the dense real M2L operator of direction `dir`
 * (0..315, the order of table 5) for N terms, N^2 x N^2 row-major doubles.
 * lfmm_host_m2l16_pack: its split-fp16 packing for all directions as the
 * matrix-core kernel reads it (LFMM_FLAG_MIXED); returns the element count
 * (uint16 bit patterns), the scale *kinv and, if kpre is not NULL, the block
 * prefixes [316][ceil(N^2/32)][2] = (k3, k1), or a negative status. */
int lfmm_host_m2l_dense(int N, int dir, double* out);
int64_t lfmm_host_m2l16_pack(int N, uint16_t* out, int64_t capacity, double* kinv, int32_t* kpre);
/* lfmm_host_m2m_dense: the dense real M2M operator from the child of class
 * `cls` (0..7, bits x y z of its position in the parent) to the parent, on
 * the normalised coefficients of both (box edges 1/2 and 1), N^2 x N^2
 * row-major doubles; the matrix-core M2M of LFMM_FLAG_MIXED uses it. */
int lfmm_host_m2m_dense(int N, int cls, double* out);

/* Library version string. */
const char* lfmm_version(void);

#ifdef __cplusplus
}
#endif
#endif
