// k_levels_raw.hip -- instantiations of the multilevel kernel for noise NZ_SOBOL_RAW.
#include "k_levels.inl"

namespace fdmc {

const void* levels_kernel_raw(int mk, int pk) {
#define X(MKv, PKv) \
    if (mk == MKv && pk == PKv) return (const void*)&k_levels<MKv, PKv, NZ_SOBOL_RAW, 0>;
#define XP(MKv) X(MKv, PK_EUROPEAN) X(MKv, PK_ASIAN) X(MKv, PK_GEOM_ASIAN) X(MKv, PK_LOOKBACK) X(MKv, PK_BARRIER)
    XP(MK_GBM) XP(MK_DUPIRE) XP(MK_LVQ)
    XP(MK_HESTON) XP(MK_HESTON_MIL)
#undef XP
#undef X
    return nullptr;
}

}  // namespace fdmc
