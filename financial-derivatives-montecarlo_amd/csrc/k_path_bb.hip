// k_path_bb.hip -- instantiations of the fused path kernel for noise NZ_SOBOL_BB.
#include "k_path.inl"
#include "path_combos.inc"

namespace fdmc {

template <int LB>
static const void* pick_bb(int mk, int pk, int cv, int cpk) {
#define X(MKv, PKv, CVv, CPKv) \
    if (mk == MKv && pk == PKv && cv == CVv && cpk == CPKv) return (const void*)&k_path<MKv, PKv, CVv, CPKv, NZ_SOBOL_BB, LB>;
    FDMC_ONE_FACTOR_COMBOS(X)
#undef X
    return nullptr;
}

const void* path_kernel_bb(int mk, int pk, int cv, int cpk, int lb) {
    return lb == 0 ? pick_bb<0>(mk, pk, cv, cpk) : pick_bb<4>(mk, pk, cv, cpk);
}

}  // namespace fdmc
