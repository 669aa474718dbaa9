// order.hip -- per-box expansion orders on the adaptive octree.
//
// The reference picks one expansion size for a whole solve from eps
// (src/cpu/multipole3d.cpp:1328-1331, GPU src/cuda/multipole3d.cu:888-891).
// Here every leaf gets its own order, either requested by the caller (per
// particle: the leaf takes the largest request of its particles) or chosen
// by a small network from per-leaf features:
//   features (8)  log2(1 + particles), level / 16, log10 sum|q|,
//                 moment-norm decay log10(|dipole| / (|monopole| + |dipole|)),
//                 V-list size / 189, log10 of the V sources' sum|q| over the
//                 leaf's, log2(1 + near-field particles), log2(1 + |W| + |X|)
//   network       8 -> 64 (ReLU) -> 64 (ReLU) -> 13 classes, FP32; class c
//                 is order p = 4 + c (terms 5 .. 17)
//   decision      the smallest order whose cumulative class probability
//                 reaches the confidence level
// The network is trained on the GPU by the kernels below (forward, backward
// with deterministic per-chunk gradient sums, Adam).  A box with children
// takes the largest order of its children; the M2L and L2L into a box run
// at the box's own order (the kernels take lists of (parent, child class),
// one launch per (level, order) bucket), so a leaf's local expansion has
// exactly the terms the leaf asked for and a box's never fewer than any of
// its descendants.
#include "api_internal.h"
#include "device_common.h"
#include "direct.h"
#include "order.h"
#include "tree.h"

namespace lfmm {

namespace {

inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

__device__ inline int od_find(const AdLevels& A, int l, uint64_t key)
{
    return find_box(key, A.lut[l], A.keys + A.off[l], A.off[l + 1] - A.off[l]);
}

// sum |q| of every box of one level: a leaf from its particles, a box with
// children from theirs (levels processed deepest first)
__global__ void __launch_bounds__(256) k_box_absq(const double4* __restrict__ src, const int2* __restrict__ range,
                                                  const int32_t* __restrict__ box2leaf,
                                                  const int32_t* __restrict__ child, int g0, int count, int gnext,
                                                  double* __restrict__ absq)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int g = g0 + i;
    double s = 0.0;
    if (box2leaf[g] >= 0) {
        const int2 r = range[g];
        for (int p = r.x; p < r.y; ++p) s += fabs(src[p].w);
    } else {
        for (int c = child[i]; c < child[i + 1]; ++c) s += absq[gnext + c];
    }
    absq[g] = s;
}

__global__ void __launch_bounds__(256) k_leaf_features(AdLevels A, const int32_t* __restrict__ leaf_gid,
                                                       const int32_t* __restrict__ leaf_lvl,
                                                       const int32_t* __restrict__ leaf_begin, int nleaf,
                                                       const double4* __restrict__ src, const double* __restrict__ geom,
                                                       const double* __restrict__ absq,
                                                       const int* __restrict__ class_dirs,
                                                       const int* __restrict__ dir_vec,
                                                       const uint32_t* __restrict__ u_off, const int2* __restrict__ u_rng,
                                                       const uint32_t* __restrict__ w_off,
                                                       const uint32_t* __restrict__ x_off, int g2,
                                                       float* __restrict__ feat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int g = leaf_gid[i], l = leaf_lvl[i];
    const uint64_t key = A.keys[g];
    const uint32_t cx = compact3(key >> 2), cy = compact3(key >> 1), cz = compact3(key);
    const double scale = geom[3];
    const double ox = box_centre(cx, l, scale, geom[0]), oy = box_centre(cy, l, scale, geom[1]),
                 oz = box_centre(cz, l, scale, geom[2]);
    const double ih = ldexp(1.0, l) / scale;
    const int b = leaf_begin[i], e = leaf_begin[i + 1];
    double q0 = 0.0, qa = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    for (int p = b; p < e; ++p) {
        const double4 s = src[p];
        q0 += s.w;
        qa += fabs(s.w);
        dx += s.w * (s.x - ox) * ih;
        dy += s.w * (s.y - oy) * ih;
        dz += s.w * (s.z - oz) * ih;
    }
    const double m1 = sqrt(dx * dx + dy * dy + dz * dz), m0 = fabs(q0);
    int nv = 0;
    double qv = 0.0;
    if (l >= 2) {
        const int c = (int)(key & 7u);
        const int side = 1 << l;
        for (int d = 0; d < 189; ++d) {
            const int dir = class_dirs[c * 189 + d];
            const int sx = (int)cx + dir_vec[4 * dir], sy = (int)cy + dir_vec[4 * dir + 1],
                      sz = (int)cz + dir_vec[4 * dir + 2];
            if (sx < 0 || sy < 0 || sz < 0 || sx >= side || sy >= side || sz >= side) continue;
            const int j = od_find(A, l, morton3(sx, sy, sz));
            if (j >= 0) {
                ++nv;
                qv += absq[A.off[l] + j];
            }
        }
    }
    int64_t nu = 0;
    for (uint32_t k = u_off[i]; k < u_off[i + 1]; ++k) nu += u_rng[k].y - u_rng[k].x;
    int nxw = (int)(w_off[i + 1] - w_off[i]);
    if (g >= g2 && x_off) nxw += (int)(x_off[g - g2 + 1] - x_off[g - g2]);
    const double tiny = 1e-30;
    float* f = feat + (int64_t)i * kOrderFeatures;
    f[0] = (float)log2(1.0 + (double)(e - b));
    f[1] = (float)l / 16.0f;
    f[2] = (float)log10(qa + tiny);
    f[3] = (float)log10((m1 + tiny) / (m0 + m1 + tiny));
    f[4] = (float)nv / 189.0f;
    f[5] = (float)log10((qv + tiny) / (qa + tiny));
    f[6] = (float)log2(1.0 + (double)nu);
    f[7] = (float)log2(1.0 + (double)nxw);
}

// ---- the network
__device__ inline float relu(float x) { return x > 0.0f ? x : 0.0f; }

// forward of one sample from weights W (global or LDS); h1, h2 hold the
// hidden activations, z the class scores
template <typename WP>
__device__ inline void mlp_forward(WP W, const float* f, float* xh, float* h1, float* h2, float* z)
{
    for (int k = 0; k < kOrderFeatures; ++k) xh[k] = (f[k] - W[kOffMu + k]) / W[kOffSd + k];
    for (int j = 0; j < kOrderHidden; ++j) {
        float a = W[kOffB1 + j];
        for (int k = 0; k < kOrderFeatures; ++k) a = fmaf(W[kOffW1 + j * kOrderFeatures + k], xh[k], a);
        h1[j] = relu(a);
    }
    for (int j = 0; j < kOrderHidden; ++j) {
        float a = W[kOffB2 + j];
        for (int k = 0; k < kOrderHidden; ++k) a = fmaf(W[kOffW2 + j * kOrderHidden + k], h1[k], a);
        h2[j] = relu(a);
    }
    for (int c = 0; c < kOrderClasses; ++c) {
        float a = W[kOffB3 + c];
        for (int k = 0; k < kOrderHidden; ++k) a = fmaf(W[kOffW3 + c * kOrderHidden + k], h2[k], a);
        z[c] = a;
    }
}

__device__ inline void softmax(float* z)
{
    float m = z[0];
    for (int c = 1; c < kOrderClasses; ++c) m = fmaxf(m, z[c]);
    float s = 0.0f;
    for (int c = 0; c < kOrderClasses; ++c) {
        z[c] = expf(z[c] - m);
        s += z[c];
    }
    for (int c = 0; c < kOrderClasses; ++c) z[c] /= s;
}

// class probabilities and/or the order of each sample (smallest order whose
// cumulative probability reaches conf)
__global__ void __launch_bounds__(128) k_mlp_predict(const float* __restrict__ W, const float* __restrict__ feat, int64_t n,
                                                     float conf, float* __restrict__ probs, int32_t* __restrict__ terms)
{
    __shared__ float w[kOrderModelSize];
    for (int k = threadIdx.x; k < kOrderModelSize; k += blockDim.x) w[k] = W[k];
    __syncthreads();
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xh[kOrderFeatures], h1[kOrderHidden], h2[kOrderHidden], z[kOrderClasses];
    mlp_forward(w, feat + s * kOrderFeatures, xh, h1, h2, z);
    softmax(z);
    if (probs)
        for (int c = 0; c < kOrderClasses; ++c) probs[s * kOrderClasses + c] = z[c];
    if (terms) {
        float cum = 0.0f;
        int c = 0;
        for (; c < kOrderClasses - 1; ++c) {
            cum += z[c];
            if (cum >= conf) break;
        }
        terms[s] = kOrderMinTerms + c;
    }
}

constexpr int kGradChunk = 64;  // samples per workgroup

// Loss and gradient of one chunk of samples (mean cross-entropy over the
// whole batch, `inv_batch` = 1 / batch size).  Activations and deltas of
// the chunk are kept in LDS; thread t then owns weight entries t, t + 256,
// ... and sums their gradient over the chunk in sample order
// (deterministic).  Writes the chunk's partial gradient and loss.
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ W, const float* __restrict__ feat,
                                                  const int32_t* __restrict__ label, const int32_t* __restrict__ idx,
                                                  int64_t n, float inv_batch, float* __restrict__ gpart,
                                                  float* __restrict__ lpart)
{
    __shared__ float w[kOrderModelSize];
    __shared__ float sx[kGradChunk][kOrderFeatures];
    __shared__ float s1[kGradChunk][kOrderHidden];
    __shared__ float s2[kGradChunk][kOrderHidden];
    __shared__ float d1[kGradChunk][kOrderHidden];
    __shared__ float d2[kGradChunk][kOrderHidden];
    __shared__ float dzs[kGradChunk][kOrderClasses];
    __shared__ float ls[kGradChunk];
    for (int k = threadIdx.x; k < kOrderModelSize; k += blockDim.x) w[k] = W[k];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < kGradChunk) {
        const int64_t s = blockIdx.x * (int64_t)kGradChunk + t;
        float h1[kOrderHidden], h2[kOrderHidden], z[kOrderClasses], xh[kOrderFeatures];
        if (s < n) {
            const int64_t r = idx ? idx[s] : s;
            mlp_forward(w, feat + r * kOrderFeatures, xh, h1, h2, z);
            softmax(z);
            const int y = label[r];
            ls[t] = -logf(fmaxf(z[y], 1e-30f));
            for (int c = 0; c < kOrderClasses; ++c) z[c] = (z[c] - (c == y ? 1.0f : 0.0f)) * inv_batch;
        } else {
            for (int k = 0; k < kOrderFeatures; ++k) xh[k] = 0.0f;
            for (int j = 0; j < kOrderHidden; ++j) h1[j] = h2[j] = 0.0f;
            for (int c = 0; c < kOrderClasses; ++c) z[c] = 0.0f;
            ls[t] = 0.0f;
        }
        for (int k = 0; k < kOrderFeatures; ++k) sx[t][k] = xh[k];
        for (int c = 0; c < kOrderClasses; ++c) dzs[t][c] = z[c];
        for (int j = 0; j < kOrderHidden; ++j) {
            s1[t][j] = h1[j];
            s2[t][j] = h2[j];
        }
        // back through layer 3 and the second ReLU
        for (int j = 0; j < kOrderHidden; ++j) {
            float a = 0.0f;
            for (int c = 0; c < kOrderClasses; ++c) a = fmaf(w[kOffW3 + c * kOrderHidden + j], z[c], a);
            d2[t][j] = h2[j] > 0.0f ? a : 0.0f;
        }
        // back through layer 2 and the first ReLU
        for (int j = 0; j < kOrderHidden; ++j) {
            float a = 0.0f;
            for (int k = 0; k < kOrderHidden; ++k) a = fmaf(w[kOffW2 + k * kOrderHidden + j], d2[t][k], a);
            d1[t][j] = h1[j] > 0.0f ? a : 0.0f;
        }
    }
    __syncthreads();
    float* gp = gpart + (int64_t)blockIdx.x * kOrderModelSize;
    for (int e = t; e < kOrderModelSize; e += blockDim.x) {
        float g = 0.0f;
        if (e >= kOffW1 && e < kOffB1) {
            const int j = (e - kOffW1) / kOrderFeatures, k = (e - kOffW1) % kOrderFeatures;
            for (int s = 0; s < kGradChunk; ++s) g = fmaf(d1[s][j], sx[s][k], g);
        } else if (e >= kOffB1 && e < kOffW2) {
            for (int s = 0; s < kGradChunk; ++s) g += d1[s][e - kOffB1];
        } else if (e >= kOffW2 && e < kOffB2) {
            const int j = (e - kOffW2) / kOrderHidden, k = (e - kOffW2) % kOrderHidden;
            for (int s = 0; s < kGradChunk; ++s) g = fmaf(d2[s][j], s1[s][k], g);
        } else if (e >= kOffB2 && e < kOffW3) {
            for (int s = 0; s < kGradChunk; ++s) g += d2[s][e - kOffB2];
        } else if (e >= kOffW3 && e < kOffB3) {
            const int c = (e - kOffW3) / kOrderHidden, k = (e - kOffW3) % kOrderHidden;
            for (int s = 0; s < kGradChunk; ++s) g = fmaf(dzs[s][c], s2[s][k], g);
        } else if (e >= kOffB3) {
            for (int s = 0; s < kGradChunk; ++s) g += dzs[s][e - kOffB3];
        }
        gp[e] = g;  // the normalisation entries (e < kOffW1) get 0
    }
    if (t == 0) {
        float l = 0.0f;
        for (int s = 0; s < kGradChunk; ++s) l += ls[s];
        lpart[blockIdx.x] = l;
    }
}

// grad[e] = sum over chunks in order; loss likewise
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ gpart, const float* __restrict__ lpart,
                                                    int parts, float* __restrict__ grad, float* __restrict__ loss)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kOrderModelSize) {
        float g = 0.0f;
        for (int p = 0; p < parts; ++p) g += gpart[(int64_t)p * kOrderModelSize + e];
        grad[e] = g;
    }
    if (e == 0) {
        float l = 0.0f;
        for (int p = 0; p < parts; ++p) l += lpart[p];
        *loss = l;
    }
}

__global__ void k_accum(const float* __restrict__ x, float* __restrict__ acc) { *acc += *x; }

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ W, const float* __restrict__ grad, float* __restrict__ m,
                                              float* __restrict__ v, float lr, float c1, float c2)
{
    const int e = kOffW1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kOrderModelSize) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const float g = grad[e];
    const float mm = b1 * m[e] + (1.0f - b1) * g;
    const float vv = b2 * v[e] + (1.0f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    W[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + 1e-8f);
}

// ---- orders of the boxes
// leaf order = the largest request of its particles (requests in input order)
__global__ void __launch_bounds__(256) k_leaf_terms_req(const int32_t* __restrict__ req, const uint32_t* __restrict__ order,
                                                        const int32_t* __restrict__ leaf_of, int64_t n, int64_t n_req,
                                                        uint32_t* __restrict__ leaf_terms)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = order[s];
    if (o >= n_req) return;  // sources of a target solve ask for nothing
    atomicMax(&leaf_terms[leaf_of[s]], (uint32_t)req[o]);
}

__global__ void __launch_bounds__(256) k_clamp_terms(uint32_t* __restrict__ t, int64_t n, uint32_t lo, uint32_t hi)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) t[i] = min(max(t[i], lo), hi);
}

__global__ void __launch_bounds__(256) k_terms_minmax(const uint32_t* __restrict__ t, int64_t n, uint32_t* __restrict__ mm)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    uint32_t lo = 0xffffffffu, hi = 0u;
    if (i < n) lo = hi = t[i];
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, off));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

// box order: a leaf's own, else the largest of its children (deepest level first)
__global__ void __launch_bounds__(256) k_box_terms(const int32_t* __restrict__ box2leaf, const uint32_t* __restrict__ leaf_terms,
                                                   const int32_t* __restrict__ child, int g0, int count, int gnext,
                                                   uint32_t* __restrict__ box_terms)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int g = g0 + i;
    const int lf = box2leaf[g];
    uint32_t t = 0;
    if (lf >= 0) {
        t = leaf_terms[lf];
    } else {
        for (int c = child[i]; c < child[i + 1]; ++c) t = max(t, box_terms[gnext + c]);
    }
    box_terms[g] = t;
}

// launch buckets: every box g0 .. g1-1 (levels >= the first bucketed level)
// under its parent: key ((parent level) * 32 + the box's terms) * 8 + the
// box's class, value the parent's index at its level.  A (level, terms)
// bucket is then the boxes of that order as (parent, class) entries sorted
// by class.
__global__ void __launch_bounds__(256) k_bucket_keys(const uint32_t* __restrict__ box_terms, int g0, int g1, AdLevels A,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                     uint32_t* __restrict__ hist)
{
    const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g1) return;
    int l = 0;
    while (l + 1 < A.depth && g >= A.off[l + 1]) ++l;
    const uint64_t k = A.keys[g];
    const int p = od_find(A, l - 1, k >> 3);  // the parent is a kept box
    const uint32_t key = ((uint32_t)(l - 1) * 32u + box_terms[g]) * 8u + (uint32_t)(k & 7u);
    keys[g - g0] = key;
    vals[g - g0] = (uint32_t)p;
    atomicAdd(&hist[key], 1u);
}

// promoted orders: box_terms[g] = prom[(level - 1) * 32 + box_terms[g]] for
// the boxes of the bucketed levels
__global__ void __launch_bounds__(256) k_promote_terms(uint32_t* __restrict__ box_terms, int g0, int g1, AdLevels A,
                                                       const int32_t* __restrict__ prom)
{
    const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g1) return;
    int l = 0;
    while (l + 1 < A.depth && g >= A.off[l + 1]) ++l;
    box_terms[g] = (uint32_t)prom[(l - 1) * 32 + (int)box_terms[g]];
}

// algorithmic M2L flops of one level's targets at their orders: per box,
// (sources present among its 189 offsets) x flops(its terms)
struct FlopTable {
    long long f[kMaxTerms3d + 1];
};
__global__ void __launch_bounds__(256) k_m2l_flops(AdLevels A, int level, const uint32_t* __restrict__ box_terms,
                                                   int fixed_terms, FlopTable F, unsigned long long* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = A.off[level + 1] - A.off[level];
    unsigned long long c = 0;
    if (i < cnt) {
        const uint64_t k = A.keys[A.off[level] + i];
        const int x = (int)compact3(k >> 2), y = (int)compact3(k >> 1), z = (int)compact3(k);
        const int side = 1 << level;
        const int bx = (x & ~1) - 2, by = (y & ~1) - 2, bz = (z & ~1) - 2;
        unsigned long long ops = 0;
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b)
                for (int d = 0; d < 6; ++d) {
                    const int sx = bx + a, sy = by + b, sz = bz + d;
                    if (sx < 0 || sy < 0 || sz < 0 || sx >= side || sy >= side || sz >= side) continue;
                    if (abs(sx - x) <= 1 && abs(sy - y) <= 1 && abs(sz - z) <= 1) continue;
                    if (od_find(A, level, morton3(sx, sy, sz)) >= 0) ++ops;
                }
        const int t = box_terms ? (int)box_terms[A.off[level] + i] : fixed_terms;
        c = ops * (unsigned long long)F.f[t];
    }
    for (int off = 32; off > 0; off >>= 1) c += (unsigned long long)__shfl_xor((long long)c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// ---- sampled direct-sum check
// sample k: input index idx = k * n_out / S; its record from the staged inputs
__global__ void __launch_bounds__(256) k_check_gather(const double* __restrict__ pos, int64_t ps,
                                                      const double* __restrict__ q, int64_t qs, int64_t n_out, int S,
                                                      int has_q, double4* __restrict__ rec, int64_t* __restrict__ at)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    const int64_t i = (int64_t)k * n_out / S;
    rec[k] = make_double4(pos[i * ps], pos[i * ps + 1], pos[i * ps + 2], has_q ? q[i * qs] : 0.0);
    at[k] = i;
}

// per sample: (potential error, potential, field error^2, field^2) without the self pair
__global__ void __launch_bounds__(256) k_check_diff(const double4* __restrict__ rec, const int64_t* __restrict__ at, int S,
                                                    const double* __restrict__ dpot, const double* __restrict__ dfield,
                                                    const double* __restrict__ pot, const double* __restrict__ field,
                                                    int fmm_has_self, double4* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    const int64_t i = at[k];
    const double self = rec[k].w * 1e4;  // q / max(0, 1e-4): the reference's clamped self pair
    const double ref = dpot[k] - self;
    const double got = pot[i] - (fmm_has_self ? self : 0.0);
    double fe = 0.0, fr = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double d = field[3 * i + a] - dfield[3 * k + a];
        fe += d * d;
        fr += dfield[3 * k + a] * dfield[3 * k + a];
    }
    out[k] = make_double4(got - ref, ref, fe, fr);
}

}  // namespace

}  // namespace lfmm

namespace lfmm::host {

using namespace lfmm;

int order_features(lfmm_ctx* ctx, const Tree& T, const DevTables* DT, float* feat)
{
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    const int total = T.total();
    double* absq = take<double>(A, total);
    if (!absq) return LFMM_ERR_NOMEM;
    for (int l = T.depth - 1; l >= 0; --l) {
        const int cnt = T.loff[l + 1] - T.loff[l];
        const int gnext = l + 1 < T.depth ? T.loff[l + 1] : 0;
        hipLaunchKernelGGL(k_box_absq, dim3(nblk(cnt, 256)), dim3(256), 0, st, T.src, T.range_all, T.box2leaf,
                           T.lbegin[l], T.loff[l], cnt, gnext, absq);
    }
    const int g2 = T.depth > 2 ? T.loff[2] : total;
    hipLaunchKernelGGL(k_leaf_features, dim3(nblk(T.nleaf, 256)), dim3(256), 0, st, T.view(), T.leaf_gid, T.leaf_lvl,
                       T.leaf_begin, (int)T.nleaf, T.src, T.geom, absq, DT->rt.class_dirs, DT->rt.dir_vec, T.u_off,
                       T.u_rng, T.w_off, T.depth > 2 ? T.x_off : nullptr, g2, feat);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int order_predict(hipStream_t st, const float* model, const float* feat, int64_t n, float conf, float* probs,
                  int32_t* terms)
{
    if (n > 0)
        hipLaunchKernelGGL(k_mlp_predict, dim3(nblk(n, 128)), dim3(128), 0, st, model, feat, n, conf, probs, terms);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int64_t order_grad_parts(int64_t n) { return std::max(1, nblk(n, kGradChunk)); }

int order_grad(hipStream_t st, const float* model, const float* feat, const int32_t* label, const int32_t* idx,
               int64_t n, float inv_batch, float* grad, float* loss, float* gpart, float* lpart)
{
    const int parts = (int)order_grad_parts(n);
    hipLaunchKernelGGL(k_mlp_grad, dim3(parts), dim3(256), 0, st, model, feat, label, idx, n, inv_batch, gpart, lpart);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(nblk(kOrderModelSize, 256)), dim3(256), 0, st, gpart, lpart, parts, grad, loss);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int order_accum(hipStream_t st, const float* loss, float* acc)
{
    hipLaunchKernelGGL(k_accum, dim3(1), dim3(1), 0, st, loss, acc);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int order_adam(hipStream_t st, float* model, const float* grad, float* m, float* v, float lr, int step)
{
    const float c1 = 1.0f / (1.0f - powf(0.9f, (float)step)), c2 = 1.0f / (1.0f - powf(0.999f, (float)step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(kOrderModelSize - kOffW1, 256)), dim3(256), 0, st, model, grad, m, v, lr, c1, c2);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int order_leaf_terms_from_requests(hipStream_t st, const Tree& T, const int32_t* req, int64_t n_req, int lo, int hi,
                                   uint32_t* leaf_terms)
{
    LFMM_HIP(fill_u32(leaf_terms, T.nleaf, 0u, st));
    hipLaunchKernelGGL(k_leaf_terms_req, dim3(nblk(T.n, 256)), dim3(256), 0, st, req, T.order, T.leaf_of, T.n, n_req,
                       leaf_terms);
    hipLaunchKernelGGL(k_clamp_terms, dim3(nblk(T.nleaf, 256)), dim3(256), 0, st, leaf_terms, T.nleaf, (uint32_t)lo,
                       (uint32_t)hi);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int order_box_terms(lfmm_ctx* ctx, const Tree& T, const uint32_t* leaf_terms, int first_bucket_level, BoxOrders& O,
                    bool promote)
{
    constexpr int64_t kMinBucket = 1024;
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    const int total = T.total();
    O.box_terms = take<uint32_t>(A, total);
    if (!O.box_terms) return LFMM_ERR_NOMEM;
    for (int l = T.depth - 1; l >= 0; --l) {
        const int cnt = T.loff[l + 1] - T.loff[l];
        const int gnext = l + 1 < T.depth ? T.loff[l + 1] : 0;
        hipLaunchKernelGGL(k_box_terms, dim3(nblk(cnt, 256)), dim3(256), 0, st, T.box2leaf, leaf_terms, T.lbegin[l],
                           T.loff[l], cnt, gnext, O.box_terms);
    }
    // buckets of the boxes of levels >= first_bucket_level, by parent
    const int depth = T.depth;
    const int nkeys = 256 * depth;
    O.seg.assign((size_t)nkeys + 1, 0);
    O.max_terms = 0;
    O.min_leaf_terms = 0;
    std::vector<uint32_t> th(nkeys, 0);
    uint32_t* hist = take<uint32_t>(A, nkeys);
    if (!hist) return LFMM_ERR_NOMEM;
    LFMM_HIP(fill_u32(hist, nkeys, 0u, st));
    // largest and smallest leaf order
    uint32_t* mm = take<uint32_t>(A, 2);
    if (!mm) return LFMM_ERR_NOMEM;
    LFMM_HIP(fill_u32(mm, 1, 0xffffffffu, st));
    LFMM_HIP(fill_u32(mm + 1, 1, 0u, st));
    hipLaunchKernelGGL(k_terms_minmax, dim3(nblk(T.nleaf, 256)), dim3(256), 0, st, leaf_terms, T.nleaf, mm);
    if (first_bucket_level < depth && first_bucket_level >= 1) {
        const int g0 = T.loff[first_bucket_level], g1 = T.loff[depth];
        const int64_t m = g1 - g0;
        uint64_t* keys = take<uint64_t>(A, m);
        uint32_t* vals = take<uint32_t>(A, m);
        void* ws = A.alloc((size_t)sort_ws_bytes(m));
        if (!keys || !vals || !ws) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_bucket_keys, dim3(nblk(m, 256)), dim3(256), 0, st, O.box_terms, g0, g1, T.view(), keys,
                           vals, hist);
        LFMM_HIP(hipMemcpyAsync(th.data(), hist, sizeof(uint32_t) * nkeys, hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipStreamSynchronize(st));
        // A bucket costs two launches whatever its size (tens of us); a
        // bucket of fewer than kMinBucket targets is dominated by that cost
        // and runs with the next larger order present at its level instead,
        // which only adds terms.
        std::vector<int32_t> prom((size_t)32 * depth);
        bool any = false;
        for (int lp = 0; promote && lp < depth; ++lp) {
            int64_t count[32] = {};
            for (int t = 0; t < 32; ++t)
                for (int c = 0; c < 8; ++c) count[t] += th[(size_t)(lp * 32 + t) * 8 + c];
            for (int t = 0; t < 32; ++t) prom[(size_t)lp * 32 + t] = t;
            for (int t = 0; t < 32; ++t) {
                if (count[t] == 0 || count[t] >= kMinBucket) continue;
                int u = t + 1;
                while (u < 32 && count[u] == 0) ++u;
                if (u >= 32) continue;  // the level's largest order stays
                count[u] += count[t];
                count[t] = 0;
                for (int v = 0; v < 32; ++v)
                    if (prom[(size_t)lp * 32 + v] == t) prom[(size_t)lp * 32 + v] = u;
                any = true;
            }
        }
        if (any) {
            int32_t* dprom = take<int32_t>(A, 32 * depth);
            if (!dprom) return LFMM_ERR_NOMEM;
            LFMM_HIP(hipMemcpyAsync(dprom, prom.data(), sizeof(int32_t) * 32 * depth, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_promote_terms, dim3(nblk(m, 256)), dim3(256), 0, st, O.box_terms, g0, g1, T.view(),
                               dprom);
            LFMM_HIP(fill_u32(hist, nkeys, 0u, st));
            hipLaunchKernelGGL(k_bucket_keys, dim3(nblk(m, 256)), dim3(256), 0, st, O.box_terms, g0, g1, T.view(),
                               keys, vals, hist);
            LFMM_HIP(hipMemcpyAsync(th.data(), hist, sizeof(uint32_t) * nkeys, hipMemcpyDeviceToHost, st));
        }
        uint64_t* sk = keys;
        uint32_t* sv = vals;
        LFMM_HIP(radix_sort(keys, vals, m, 16, ws, st, &sk, &sv));
        O.plist = (int32_t*)sv;
    }
    uint32_t hm[2] = {0, 0};
    LFMM_HIP(hipMemcpyAsync(hm, mm, sizeof(hm), hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    O.min_leaf_terms = (int)hm[0];
    O.max_terms = (int)hm[1];
    int64_t run = 0;
    for (int k = 0; k < nkeys; ++k) {
        O.seg[k] = run;
        run += th[k];
    }
    O.seg[nkeys] = run;
    return LFMM_OK;
}

int order_m2l_flops(lfmm_ctx* ctx, const Tree& T, const BoxOrders& O, int N, int first_bucket_level, double* flops)
{
    hipStream_t st = ctx->stream;
    unsigned long long* ws = take<unsigned long long>(ctx->arena, 1);
    if (!ws) return LFMM_ERR_NOMEM;
    LFMM_HIP(fill_u32((uint32_t*)ws, 2, 0u, st));
    FlopTable F{};
    for (int t = 0; t <= kMaxTerms3d; ++t) F.f[t] = t > 0 ? m2l_flops_per_op(t) : 0;
    const AdLevels V = T.view();
    for (int l = 2; l < T.depth; ++l) {
        const int cnt = T.loff[l + 1] - T.loff[l];
        hipLaunchKernelGGL(k_m2l_flops, dim3(nblk(cnt, 256)), dim3(256), 0, st, V, l,
                           l >= first_bucket_level ? O.box_terms : nullptr, N, F, ws);
    }
    unsigned long long h = 0;
    LFMM_HIP(hipMemcpyAsync(&h, ws, 8, hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    *flops = (double)h;
    return LFMM_OK;
}

int order_check_start(lfmm_ctx* ctx, const DevPoints& DP, const double4* src, int64_t n, int64_t n_out, bool has_q,
                      int samples, OrderCheck& C)
{
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    C = OrderCheck{};
    C.S = (int)std::min<int64_t>(samples, n_out);
    if (C.S <= 0) return LFMM_OK;
    const int S = C.S;
    C.rec = take<double4>(A, S);
    C.at = take<int64_t>(A, S);
    C.dp = take<double>(A, S);
    C.df = take<double>(A, 3 * S);
    C.out = take<double4>(A, S);
    const int splits = direct3d_split_count(S, n);
    double4* part = take<double4>(A, (int64_t)splits * S);
    if (!C.rec || !C.at || !C.dp || !C.df || !C.out || !part) return LFMM_ERR_NOMEM;
    if (!ctx->side_stream) {
        LFMM_HIP(hipStreamCreateWithFlags(&ctx->side_stream, hipStreamNonBlocking));
        LFMM_HIP(hipEventCreateWithFlags(&ctx->side_events[0], hipEventDisableTiming));
        LFMM_HIP(hipEventCreateWithFlags(&ctx->side_events[1], hipEventDisableTiming));
    }
    hipStream_t ss = ctx->side_stream;
    // the sampled direct sums need only the staged inputs: they run on the
    // side stream while the solve continues on the main one
    LFMM_HIP(hipEventRecord(ctx->side_events[0], st));
    LFMM_HIP(hipStreamWaitEvent(ss, ctx->side_events[0], 0));
    hipLaunchKernelGGL(k_check_gather, dim3(nblk(S, 256)), dim3(256), 0, ss, DP.pos, DP.ps, DP.q, DP.qs, n_out, S,
                       has_q ? 1 : 0, C.rec, C.at);
    LFMM_HIP(launch_direct3d_split(C.rec, S, src, n, C.dp, C.df, part, splits, ss));
    LFMM_HIP(hipEventRecord(ctx->side_events[1], ss));
    C.started = true;
    return LFMM_OK;
}

int order_check_finish(lfmm_ctx* ctx, const OrderCheck& C, const double* pot, const double* field, bool fmm_has_self,
                       double* rel_pot, double* rel_field)
{
    hipStream_t st = ctx->stream;
    *rel_pot = *rel_field = 0.0;
    if (!C.started || C.S <= 0) return LFMM_OK;
    const int S = C.S;
    LFMM_HIP(hipStreamWaitEvent(st, ctx->side_events[1], 0));
    hipLaunchKernelGGL(k_check_diff, dim3(nblk(S, 256)), dim3(256), 0, st, C.rec, C.at, S, C.dp, C.df, pot, field,
                       fmm_has_self ? 1 : 0, C.out);
    std::vector<double4> h(S);
    LFMM_HIP(hipMemcpyAsync(h.data(), C.out, sizeof(double4) * S, hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    double e = 0, r = 0, fe = 0, fr = 0;
    for (const double4& v : h) {
        e += v.x * v.x;
        r += v.y * v.y;
        fe += v.z;
        fr += v.w;
    }
    *rel_pot = r > 0 ? std::sqrt(e / r) : 0.0;
    *rel_field = fr > 0 ? std::sqrt(fe / fr) : 0.0;
    return LFMM_OK;
}

}  // namespace lfmm::host
