// ops3d_inst.hip -- instantiates the operator kernels for one expansion size
// (LFMM_N, set by the build), so the 18 sizes compile in parallel.
#include "ops3d_kernels.h"
#include "m2l16_kernels.h"

#ifndef LFMM_N
#error "build with -DLFMM_N=<terms>"
#endif

namespace lfmm {
template hipError_t do_p2m<LFMM_N>(const P2MArgs&, hipStream_t);
template hipError_t do_p2l<LFMM_N>(const P2MArgs&, hipStream_t);
template hipError_t do_rtr<LFMM_N>(int, const RtrTables&, const RtrLevel&, hipStream_t);
template hipError_t do_l2p<LFMM_N>(const LeafArgs&, hipStream_t);
template hipError_t do_m2l16_pack<LFMM_N>(const M2L16Args&, hipStream_t);
template hipError_t do_m2l16<LFMM_N>(const M2L16Args&, hipStream_t);
}  // namespace lfmm
