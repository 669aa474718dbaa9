// api_direct.cpp -- C ABI of the direct sums and of the reference's
// exact-result files:
//   lfmm_direct3d / lfmm_direct2d   fmm::gpu::ComputeExact (src/common/simple_math.cu:24-130;
//                                   CPU twin src/common/simple_math.h:218-345)
//   lfmm_write_exact / lfmm_read_exact
//                                   the "<kind>_exact.txt" files ComputeExact writes and
//                                   ReadError reads (src/common/simple_math.cu:103-126,
//                                   src/common/simple_math.h:353-376)
//   lfmm_error_stats                the three figures ReadError prints
//                                   (src/common/simple_math.h:378-401)
#include <cmath>
#include <complex>
#include <cstring>
#include <fstream>
#include <mutex>

#include "api_internal.h"
#include "direct.h"
#include "tree.h"

using namespace lfmm;
using namespace lfmm::host;

namespace {

int direct_solve(lfmm_ctx* c, int dim, const lfmm_points3d* S, const lfmm_points3d* Tg, int memory, unsigned flags,
                 double* pot, double* field)
{
    hipStream_t st = c->stream;
    Arena& A = c->arena;
    A.reset();
    DevPoints s, t;
    if (memory == LFMM_MEM_HOST) {
        int rc = upload_view(c, S, dim, s);
        if (rc != LFMM_OK) return rc;
        if (Tg) {
            rc = upload_view(c, Tg, dim, t);
            if (rc != LFMM_OK) return rc;
        }
    } else {
        s = DevPoints{S->pos, S->pos_stride, S->q, S->q_stride, S->n, dim};
        if (Tg) t = DevPoints{Tg->pos, Tg->pos_stride, nullptr, 1, Tg->n, dim};
    }
    double4* src = take<double4>(A, s.n);
    if (!src) return LFMM_ERR_NOMEM;
    LFMM_HIP(pack_points(dim, s.pos, s.ps, s.q, s.qs, s.n, src, st));
    const double4* tgt = src;
    int64_t nt = s.n;
    if (Tg) {
        double4* tb = take<double4>(A, t.n);
        if (!tb) return LFMM_ERR_NOMEM;
        LFMM_HIP(pack_points(dim, t.pos, t.ps, nullptr, 1, t.n, tb, st));
        tgt = tb;
        nt = t.n;
    }
    double* dpot = pot;
    double* dfield = field;
    if (memory == LFMM_MEM_HOST) {
        dpot = pot ? take<double>(A, nt) : nullptr;
        dfield = field ? take<double>(A, dim * nt) : nullptr;
        if ((pot && !dpot) || (field && !dfield)) return LFMM_ERR_NOMEM;
    }
    const bool exclude = !Tg && (flags & LFMM_FLAG_EXCLUDE_SELF);
    LFMM_HIP(launch_direct(dim, tgt, nt, src, s.n, exclude, dpot, dfield, st));
    if (memory == LFMM_MEM_HOST) {
        if (pot) LFMM_HIP(hipMemcpyAsync(pot, dpot, sizeof(double) * (size_t)nt, hipMemcpyDeviceToHost, st));
        if (field)
            LFMM_HIP(hipMemcpyAsync(field, dfield, sizeof(double) * (size_t)(dim * nt), hipMemcpyDeviceToHost, st));
    }
    LFMM_HIP(hipStreamSynchronize(st));
    return LFMM_OK;
}

int direct_entry(lfmm_ctx* c, int dim, const lfmm_points3d* sources, const lfmm_points3d* targets, int memory,
                 unsigned flags, double* pot, double* field)
{
    if (!c || !valid_points(sources, true, dim)) return LFMM_ERR_INVALID;
    if (memory != LFMM_MEM_HOST && memory != LFMM_MEM_DEVICE) return LFMM_ERR_INVALID;
    if (targets && !valid_points(targets, false, dim)) return LFMM_ERR_INVALID;
    const lfmm_points3d* tg = (targets && targets->n > 0) ? targets : nullptr;
    if ((tg ? tg->n : sources->n) == 0) return LFMM_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    if (sources->n == 0) {  // no sources: every output is zero
        const size_t nt = (size_t)tg->n;
        if (memory == LFMM_MEM_HOST) {
            if (pot) std::memset(pot, 0, nt * sizeof(double));
            if (field) std::memset(field, 0, dim * nt * sizeof(double));
        } else {
            if (pot && hipMemsetAsync(pot, 0, nt * sizeof(double), c->stream) != hipSuccess) return LFMM_ERR_HIP;
            if (field && hipMemsetAsync(field, 0, dim * nt * sizeof(double), c->stream) != hipSuccess)
                return LFMM_ERR_HIP;
            if (hipStreamSynchronize(c->stream) != hipSuccess) return LFMM_ERR_HIP;
        }
        return LFMM_OK;
    }
    try {
        return direct_solve(c, dim, sources, tg, memory, flags, pot, field);
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

// doubles per value of each kind (potential: 1; 2D force: complex; 3D force: vector)
int kind_width(int kind)
{
    switch (kind) {
    case LFMM_POTENTIAL2D:
    case LFMM_POTENTIAL3D: return 1;
    case LFMM_FORCE2D: return 2;
    case LFMM_FORCE3D: return 3;
    default: return 0;
    }
}

}  // namespace

extern "C" {

int lfmm_direct3d(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, int memory, unsigned flags,
                  double* potential, double* field)
{
    return direct_entry(c, 3, sources, targets, memory, flags, potential, field);
}

int lfmm_direct2d(lfmm_ctx* c, const lfmm_points2d* sources, const lfmm_points2d* targets, int memory, unsigned flags,
                  double* potential, double* field)
{
    return direct_entry(c, 2, sources, targets, memory, flags, potential, field);
}

const char* lfmm_exact_filename(int kind)
{
    switch (kind) {
    case LFMM_POTENTIAL2D: return "potential2d_exact.txt";
    case LFMM_FORCE2D: return "force2d_exact.txt";
    case LFMM_POTENTIAL3D: return "potential3d_exact.txt";
    case LFMM_FORCE3D: return "force3d_exact.txt";
    default: return nullptr;
    }
}

// One value per line, 20 significant digits, as the reference's
// `fout.precision(20); fout << x << "\n"`: a double as is, a 2D force as a
// std::complex<double> "(re,im)", a 3D force as "x y z" (src/common/utils.h:108-109).
int lfmm_write_exact(const char* path, int kind, int64_t n, const double* values)
{
    const int w = kind_width(kind);
    if (!path || w == 0 || n < 0 || (n > 0 && !values)) return LFMM_ERR_INVALID;
    try {
        std::ofstream fout(path);
        if (!fout) return LFMM_ERR_INVALID;
        fout.precision(20);
        for (int64_t i = 0; i < n; ++i) {
            const double* v = values + (size_t)i * w;
            if (kind == LFMM_FORCE2D) fout << std::complex<double>(v[0], v[1]) << "\n";
            else if (kind == LFMM_FORCE3D) fout << v[0] << " " << v[1] << " " << v[2] << "\n";
            else fout << v[0] << "\n";
        }
        fout.flush();
        return fout ? LFMM_OK : LFMM_ERR_INVALID;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

// Reads n values written as above (the reference's `fin >> x`).  Returns the
// number of values read, or a negative status.
int64_t lfmm_read_exact(const char* path, int kind, int64_t n, double* values)
{
    const int w = kind_width(kind);
    if (!path || w == 0 || n < 0 || (n > 0 && !values)) return -LFMM_ERR_INVALID;
    try {
        std::ifstream fin(path);
        if (!fin) return -LFMM_ERR_INVALID;
        int64_t i = 0;
        for (; i < n; ++i) {
            double* v = values + (size_t)i * w;
            if (kind == LFMM_FORCE2D) {
                std::complex<double> z;
                if (!(fin >> z)) break;
                v[0] = z.real();
                v[1] = z.imag();
            } else if (kind == LFMM_FORCE3D) {
                if (!(fin >> v[0] >> v[1] >> v[2])) break;
            } else if (!(fin >> v[0])) {
                break;
            }
        }
        return i;
    } catch (...) {
        return -LFMM_ERR_INVALID;
    }
}

// ReadError's figures over n values: max |e|, sqrt(mean((|e| / |exact|)^2))
// (the reference's "l2 error") and sum |e| / sum |exact| ("relative error"),
// |.| the absolute value, the complex modulus or the Euclidean norm.
int lfmm_error_stats(int kind, int64_t n, const double* exact, const double* result, double* max_err, double* l2_err,
                     double* rel_err)
{
    const int w = kind_width(kind);
    if (w == 0 || n < 0 || (n > 0 && (!exact || !result))) return LFMM_ERR_INVALID;
    double mx = 0.0, e1 = 0.0, e2 = 0.0, l2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* a = exact + (size_t)i * w;
        const double* b = result + (size_t)i * w;
        double err, mag;
        if (w == 1) {
            err = std::abs(a[0] - b[0]);
            mag = std::abs(a[0]);
        } else if (w == 2) {
            err = std::abs(std::complex<double>(a[0] - b[0], a[1] - b[1]));
            mag = std::abs(std::complex<double>(a[0], a[1]));
        } else {
            const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
            err = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            mag = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        }
        mx = std::max(mx, err);
        e1 += err;
        e2 += mag;
        l2 += (err / mag) * (err / mag);
    }
    if (max_err) *max_err = mx;
    if (l2_err) *l2_err = n > 0 ? std::sqrt(l2 / (double)n) : 0.0;
    if (rel_err) *rel_err = e1 / e2;
    return LFMM_OK;
}

// The allocation-free variant of lfmm_direct3d (device memory): the packed
// records go to the fixed per-GPU scratch array of direct.hip instead of the
// context's workspace.  Calls overlapping on one GPU share that array and race.
int lfmm_direct3d_noalloc(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, unsigned flags,
                          double* potential, double* field)
{
    if (!c || !valid_points(sources, true, 3)) return LFMM_ERR_INVALID;
    if (targets && !valid_points(targets, false, 3)) return LFMM_ERR_INVALID;
    const lfmm_points3d* tg = (targets && targets->n > 0) ? targets : nullptr;
    const int64_t ns = sources->n, nt = tg ? tg->n : ns;
    if (nt == 0) return LFMM_OK;
    if (ns + (tg ? nt : 0) > (int64_t)(kNoAllocScratchBytes / sizeof(double4))) return LFMM_ERR_TOO_LARGE;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    double4* scr = noalloc_scratch(c->device);
    if (!scr) return LFMM_ERR_HIP;
    hipStream_t st = c->stream;
    if (ns == 0) {  // no sources: every output is zero
        if (potential) LFMM_HIP(fill_u32((uint32_t*)potential, 2 * nt, 0u, st));
        if (field) LFMM_HIP(fill_u32((uint32_t*)field, 6 * nt, 0u, st));
        LFMM_HIP(hipStreamSynchronize(st));
        return LFMM_OK;
    }
    LFMM_HIP(pack_points(3, sources->pos, sources->pos_stride, sources->q, sources->q_stride, ns, scr, st));
    const double4* tgt = scr;
    if (tg) {
        LFMM_HIP(pack_points(3, tg->pos, tg->pos_stride, nullptr, 1, nt, scr + ns, st));
        tgt = scr + ns;
    }
    const bool exclude = !tg && (flags & LFMM_FLAG_EXCLUDE_SELF);
    LFMM_HIP(launch_direct(3, tgt, nt, scr, ns, exclude, potential, field, st));
    LFMM_HIP(hipStreamSynchronize(st));
    return LFMM_OK;
}

}  // extern "C"
