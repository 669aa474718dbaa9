// ops3d.cpp -- runtime dispatch of the operator launchers over the expansion
// size N (1..18); each size is compiled in its own object (ops3d_inst.hip).
#include "ops3d.h"

namespace lfmm {

template <int N> hipError_t do_p2m(const P2MArgs& a, hipStream_t st);
template <int N> hipError_t do_p2l(const P2MArgs& a, hipStream_t st);
template <int N> hipError_t do_rtr(int kind, const RtrTables& T, const RtrLevel& L, hipStream_t st);
template <int N> hipError_t do_l2p(const LeafArgs& a, hipStream_t st);
template <int N> hipError_t do_m2l16_pack(const M2L16Args& a, hipStream_t st);
template <int N> hipError_t do_m2l16(const M2L16Args& a, hipStream_t st);

#define LFMM_FOR_TERMS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)

hipError_t launch_p2m(int N, const P2MArgs& a, hipStream_t st)
{
    switch (N) {
#define X(n) case n: return do_p2m<n>(a, st);
        LFMM_FOR_TERMS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_p2l(int N, const P2MArgs& a, hipStream_t st)
{
    switch (N) {
#define X(n) case n: return do_p2l<n>(a, st);
        LFMM_FOR_TERMS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rtr(int N, int kind, const RtrTables& T, const RtrLevel& L, hipStream_t st)
{
    switch (N) {
#define X(n) case n: return do_rtr<n>(kind, T, L, st);
        LFMM_FOR_TERMS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_l2p(int N, const LeafArgs& a, hipStream_t st)
{
    switch (N) {
#define X(n) case n: return do_l2p<n>(a, st);
        LFMM_FOR_TERMS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_m2l16_pack(int N, const M2L16Args& a, hipStream_t st)
{
    switch (N) {
#define X(n) case n: return do_m2l16_pack<n>(a, st);
        LFMM_FOR_TERMS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_m2l16(int N, const M2L16Args& a, hipStream_t st)
{
    switch (N) {
#define X(n) case n: return do_m2l16<n>(a, st);
        LFMM_FOR_TERMS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

}  // namespace lfmm
