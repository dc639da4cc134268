// k_path_bb.hip -- instantiations of the fused path kernel for noise NZ_SOBOL_BB
// (direction records in LDS) and NZ_SOBOL_BB_GLOBAL (more than SOBOL_LDS_MAX_D dates).
#include "k_path.inl"
#include "path_combos.inc"

namespace fdmc {

template <int NZ, int LB>
static const void* pick_bb(int mk, int pk, int cv, int cpk) {
#define X(MKv, PKv, CVv, CPKv) \
    if (mk == MKv && pk == PKv && cv == CVv && cpk == CPKv) return (const void*)&k_path<MKv, PKv, CVv, CPKv, NZ, LB>;
    FDMC_ONE_FACTOR_COMBOS(X)
#undef X
    return nullptr;
}

const void* path_kernel_bb(int mk, int pk, int cv, int cpk, int lb, bool global_records) {
    if (global_records) return lb == 4 ? pick_bb<NZ_SOBOL_BB_GLOBAL, 4>(mk, pk, cv, cpk) : nullptr;
    return lb == 0 ? pick_bb<NZ_SOBOL_BB, 0>(mk, pk, cv, cpk) : pick_bb<NZ_SOBOL_BB, 4>(mk, pk, cv, cpk);
}

}  // namespace fdmc
