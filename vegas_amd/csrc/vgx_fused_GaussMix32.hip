// Instantiates the fused sampler for the GaussMix32 integrand: dimensions 1-12
// specialised (unrolled, register-resident x) plus a generic runtime-dim path.
#include "vgx_fused.h"

namespace vgx {

template <>
int launch_fused<GaussMix32>(const IterGeom& g, const FusedPtrs& io, const GaussMix32& fcn, int blocks,
                       size_t lds, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    switch (g.dim) {
        case 1: return launch_fused_dim<1>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 2: return launch_fused_dim<2>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 3: return launch_fused_dim<3>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 4: return launch_fused_dim<4>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 5: return launch_fused_dim<5>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 6: return launch_fused_dim<6>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 7: return launch_fused_dim<7>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 8: return launch_fused_dim<8>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 9: return launch_fused_dim<9>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 10: return launch_fused_dim<10>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 11: return launch_fused_dim<11>(g, io, fcn, blocks, lds, st, ev0, ev1);
        case 12: return launch_fused_dim<12>(g, io, fcn, blocks, lds, st, ev0, ev1);
        default: return launch_fused_dim<0>(g, io, fcn, blocks, lds, st, ev0, ev1);
    }
}

}  // namespace vgx
