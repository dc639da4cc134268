// k_greeks_bb.hip -- instantiations of the price + adjoint-Greeks kernel for noise NZ_SOBOL_BB.
#include "k_greeks.inl"

namespace fdmc {

template <int LB>
static const void* pick_greeks_bb(int mk, int pk, int cv, int cpk) {
#define X(MKv, PKv, CVv, CPKv) \
    if (mk == MKv && pk == PKv && cv == CVv && cpk == CPKv) return (const void*)&k_greeks<MKv, PKv, CVv, CPKv, NZ_SOBOL_BB, LB>;
#define XM(MKv) X(MKv, PK_EUROPEAN, CV_NONE, 0) X(MKv, PK_ASIAN, CV_NONE, 0) X(MKv, PK_GEOM_ASIAN, CV_NONE, 0) \
    X(MKv, PK_LOOKBACK, CV_NONE, 0) X(MKv, PK_ASIAN, CV_SHARED, PK_GEOM_ASIAN)
    XM(MK_GBM) XM(MK_DUPIRE) XM(MK_LVQ)
#undef XM
#undef X
    return nullptr;
}

const void* greeks_kernel_bb(int mk, int pk, int cv, int cpk, int lb) {
    return lb == 0 ? pick_greeks_bb<0>(mk, pk, cv, cpk) : pick_greeks_bb<4>(mk, pk, cv, cpk);
}

}  // namespace fdmc
