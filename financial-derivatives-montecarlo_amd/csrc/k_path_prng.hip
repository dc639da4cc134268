// k_path_prng.hip -- instantiations of the fused path kernel for noise NZ_PRNG.
#include "k_path.inl"
#include "path_combos.inc"

namespace fdmc {

template <int LB>
static const void* pick_prng(int mk, int pk, int cv, int cpk) {
#define X(MKv, PKv, CVv, CPKv) \
    if (mk == MKv && pk == PKv && cv == CVv && cpk == CPKv) return (const void*)&k_path<MKv, PKv, CVv, CPKv, NZ_PRNG, LB>;
    FDMC_ONE_FACTOR_COMBOS(X)
    FDMC_HESTON_COMBOS(X)
#undef X
    return nullptr;
}

const void* path_kernel_prng(int mk, int pk, int cv, int cpk, int lb) {
    (void)lb;
    return pick_prng<0>(mk, pk, cv, cpk);
}

}  // namespace fdmc
