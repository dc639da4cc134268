// k_path_rows.hip -- instantiations of the fused path kernel for increments
// precomputed by the PCA GEMM (NZ_ROWS).
#include "k_path.inl"
#include "path_combos.inc"

namespace fdmc {

const void* path_kernel_rows(int mk, int pk, int cv, int cpk, int lb) {
    (void)lb;
#define X(MKv, PKv, CVv, CPKv) \
    if (mk == MKv && pk == PKv && cv == CVv && cpk == CPKv) return (const void*)&k_path<MKv, PKv, CVv, CPKv, NZ_ROWS, 0>;
    FDMC_ONE_FACTOR_COMBOS(X)
#undef X
    return nullptr;
}

}  // namespace fdmc
