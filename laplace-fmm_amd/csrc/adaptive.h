// adaptive.h -- adaptive octree: device view shared by the tree, list and
// operator kernels (adaptive.hip, ops3d_kernels.h).
//
// The reference builds uniform-depth trees only (src/cpu/morton_tree.cpp:
// 371-423; its near/far lists :147-208 and the theta-hash lookup of
// src/cpu/multipole3d.cpp:122-125 assume every leaf on one level).  Here a
// box is split only while it holds more than leaf_max particles, and the
// interactions follow the adaptive FMM's four lists:
//   U(b)  leaf b: every leaf that touches b (any level, b itself included)  -> P2P
//   V(b)  any box: boxes of b's level, children of the neighbours of b's
//         parent, not touching b                                      -> M2L (key lookup, as uniform)
//   W(b)  leaf b: boxes c finer than b, not touching b, whose parent touches b
//                                                                     -> M2P (multipole of c at b's particles)
//   X(b)  any box: leaves c with b in W(c)                            -> P2L (c's particles into b's local)
// A uniform input (every leaf on one level) has empty W and X lists and the
// U lists are the 27-neighbourhoods, so the solve reduces to the uniform one.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

constexpr int kAdMaxLevels = 16;

// Device view of an adaptive octree with levels 0 .. depth-1.  Boxes are
// numbered globally level by level: level l holds gids [off[l], off[l+1]),
// each level in Morton key order.
struct AdLevels {
    int depth;
    int off[kAdMaxLevels + 1];
    const uint64_t* keys;              // [total] key of each box at its level
    const int32_t* lut[kAdMaxLevels];  // dense key -> index at the level, or null (binary search)
    const int32_t* box2leaf;           // [total] leaf index, -1 for a box with children
    const int2* range;                 // [total] particle range of each box
};

}  // namespace lfmm
