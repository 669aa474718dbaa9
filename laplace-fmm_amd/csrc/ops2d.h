// ops2d.h -- argument blocks and launchers of the 2D operator kernels (ops2d.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

// Device copies of OpTables2d (complex values as double2).
struct Tables2d {
    const double2* m2l;        // [40][N][N]
    const float2* m2l32;       // the same rounded to FP32 (mixed M2L)
    const double2* m2m;        // [4][N][N]
    const double2* l2l;        // [4][N][N]
    const int* dir_vec;        // [40][2]
    const int* class_dirs;     // [4][27]
};

struct P2M2dArgs {
    const double4* src;        // sorted (x, y, 0, q)
    const int32_t* leaf_begin;
    const uint64_t* leaf_key;
    int nleaf;
    int level;
    const double* geom;        // {shift x, y, -, scale}
    int N;
    double2* outer;
};

struct M2M2dArgs {
    int N;
    int num_parents;           // boxes at the parent level
    const int32_t* child_begin;
    const uint64_t* child_keys;
    const double2* src;        // multipoles of the child level
    double2* dst;              // multipoles of the parent level
};

struct M2L2dArgs {
    int N;
    int level;                 // target level
    int num_parents;           // boxes at level - 1
    const int32_t* child_begin;
    const uint64_t* child_keys;   // keys at `level`
    const int32_t* lut;        // dense key -> box table at `level`, or null
    const uint64_t* keys;
    int size;                  // boxes at `level`
    double log_h;              // log of the box edge at `level`
    const double2* src;        // multipoles at `level`
    const double2* parent_local;  // locals at level - 1 (null: no L2L)
    double2* dst;              // locals at `level` (written)
    bool f32;                  // M2L products and per-direction sums in FP32
};

struct Leaf2dArgs {
    const double4* src;
    const int32_t* leaf_of;
    const uint64_t* leaf_key;
    const int32_t* leaf_begin;
    int nleaf;
    int level;
    const double* geom;
    const int32_t* lut;
    const double2* local;      // normalised locals of the leaves, or null
    int N;
    const uint32_t* perm;
    int64_t n;
    int64_t n_out;
    int exclude_self;
    double* pot;
    double* field;             // (re, im) of the complex force per output
};

hipError_t launch_p2m2d(const P2M2dArgs& a, hipStream_t st);
hipError_t launch_m2m2d(const M2M2dArgs& a, const Tables2d& T, hipStream_t st);
hipError_t launch_m2l2d(const M2L2dArgs& a, const Tables2d& T, hipStream_t st);
hipError_t launch_leaf2d(const Leaf2dArgs& a, hipStream_t st);

}  // namespace lfmm
