// k_pca.hip -- PCA noise transform on gfx950 matrix cores.
//
// The reference casts Z to FP16 in one kernel and calls cublasGemmEx (FP16 x
// FP16 -> FP32, tensor cores) for dW = Z M^T (methods_cuda.cu:721-724,
// 1151-1166).  Here:
//   k_gen_z   : scrambled-Sobol normals of R replicas written point-major
//               (row = replica * n + point, D contiguous coordinates) directly in
//               BF16 or FP16 -- the GEMM's A operand, no FP32 round trip;
//   k_gemm_nt : C[m][n] = sum_k A[m][k] Bt[n][k] with v_mfma_f32_32x32x16_{bf16,f16},
//               FP32 accumulation.  Bt is M itself stored row-major ([D][D],
//               K contiguous), so both operands stream along K.
//               128 x 128 workgroup tiles, 4 waves of 64 x 64, BK = 32, LDS
//               double buffering; tiles are assigned so the 8 column tiles of
//               one row block run on the same XCD (its L2 keeps the A rows).
//   k_gemm_nt256: the same with 256 x 256 tiles (4 waves of 128 x 128), used
//               whenever the shape allows it; with EPI = true it stores no C
//               and instead multiplies the Euler factors of a GBM basket per
//               row, asset and 128-column slab (ProdEpi), so dW never reaches
//               memory.
#include "k_path.inl"

namespace fdmc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <typename T>
__device__ __forceinline__ T to_half(float x);
template <>
__device__ __forceinline__ __bf16 to_half<__bf16>(float x) { return (__bf16)x; }
template <>
__device__ __forceinline__ _Float16 to_half<_Float16>(float x) { return (_Float16)x; }

// ---- normals, point-major ----------------------------------------------------
// Grid (workgroups, R); a workgroup runs `tpw` consecutive 256-point tiles of
// replica r (lane = point) one 64-dimension chunk at a time: the chunk's direction
// records (64 x 20 words, 5 KB) stay in LDS while the tiles are generated
// (tile words updated incrementally, two XORs per dimension and tile), and the
// normals are staged through LDS so that global stores are 16-byte and
// row-contiguous.  49 KB of LDS: three workgroups per CU.  Columns D..ld-1
// are written as zeros.
constexpr int GZ_TW = 72;   // staged row: 64 values + 8 of padding

template <typename T>
__global__ __launch_bounds__(256) void k_gen_z(const uint32_t* __restrict__ sdir, int D, int ld, int64_t first,
                                               int64_t count, int tpw, T* __restrict__ out) {
    __shared__ __align__(16) float4 itab[FDMC_ICDF_TAB_ROWS];
    __shared__ __align__(16) uint32_t rec[64 * 20];
    __shared__ __align__(16) T tile[256 * GZ_TW];
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = blockIdx.y;
    icdf_tab_load(itab, tid, 256);
    const uint32_t* grec = sdir + (int64_t)r * D * SDIR_WORDS;
    SobolLane sl;
    const uint32_t g = (uint32_t)(lane ^ (lane >> 1));
    sl.rec = rec; sl.grec = nullptr; sl.ga = g & 7u; sl.gb = g >> 3; sl.w = (uint32_t)(tid >> 6);
    sl.hi_gray = 0; sl.lds_mode = true;
    const int64_t tfirst = (first >> 8) + (int64_t)blockIdx.x * tpw;
    const int64_t tlast = (first + count - 1) >> 8;
    const int64_t tend = tfirst + tpw - 1 < tlast ? tfirst + tpw - 1 : tlast;
    for (int d0 = 0; d0 < ld; d0 += 64) {
        const int dn = D - d0 < 64 ? D - d0 : 64;     // <= 0 on padding chunks
        const uint32_t* gchunk = grec + (int64_t)d0 * SDIR_WORDS;
        __syncthreads();                              // the previous chunk's records are no longer read
        for (int i = tid; i < dn * 16; i += 256) rec[(i >> 4) * 20 + (i & 15)] = gchunk[(int64_t)(i >> 4) * SDIR_WORDS + (i & 15)];
        for (int64_t t = tfirst; t <= tend; t++) {
            __syncthreads();                          // tile words and staging free again
            if (dn > 0) sobol_tile_words(rec, gchunk, dn, (uint32_t)t, t == tfirst, tid);
            __syncthreads();
#pragma unroll
            for (int j0 = 0; j0 < 64; j0 += 8) {
                union { T h[8]; uint4 u; } v;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const float z = (j0 + j < dn) ? icdf_u32_tab(sobol_coord_lds(sl, j0 + j), itab) : 0.0f;
                    v.h[j] = to_half<T>(z);
                }
                *(uint4*)(tile + tid * GZ_TW + j0) = v.u;
            }
            __syncthreads();
            // 256 rows x 64 values = 2048 chunks of 8 values
            const int64_t base = t << 8;
            for (int c = tid; c < 2048; c += 256) {
                const int row = c >> 3, q = (c & 7) * 8;
                const int64_t p = base + row;
                if (p >= first && p < first + count && d0 + q < ld) {
                    const int64_t orow = (int64_t)r * count + (p - first);
                    *(uint4*)(out + orow * ld + d0 + q) = *(const uint4*)(tile + row * GZ_TW + q);
                }
            }
        }
    }
}

// ---- NT GEMM on MFMA ----------------------------------------------------------
constexpr int GB_M = 128, GB_N = 128, GB_K = 32, G_PAD = 8, G_LD = GB_K + G_PAD;

template <typename T>
__device__ __forceinline__ f32x16 mfma32(const T* a, const T* b, f32x16 c);
template <>
__device__ __forceinline__ f32x16 mfma32<__bf16>(const __bf16* a, const __bf16* b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*(const bf16x8*)a, *(const bf16x8*)b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ f32x16 mfma32<_Float16>(const _Float16* a, const _Float16* b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(*(const f16x8*)a, *(const f16x8*)b, c, 0, 0, 0);
}

template <typename T>
struct Frag;
template <>
struct Frag<__bf16> { typedef bf16x8 type; };
template <>
struct Frag<_Float16> { typedef f16x8 type; };

template <typename T>
__device__ __forceinline__ f32x16 mfma32v(typename Frag<T>::type a, typename Frag<T>::type b, f32x16 c) {
    if constexpr (std::is_same<T, __bf16>::value) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// Requires Mrows % 128 == 0, N % 128 == 0, K % 32 == 0 (checked on the host).
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_nt(const T* __restrict__ A, const T* __restrict__ Bt,
                                                 float* __restrict__ C, int Mrows, int N, int K) {
    __shared__ __align__(16) T As[2][GB_M * G_LD];
    __shared__ __align__(16) T Bs[2][GB_N * G_LD];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // XCD-aware tile order: workgroup i runs on XCD i % 8; give each XCD a
    // contiguous range of tiles, column tiles fastest.
    const int ntn = N / GB_N, ntm = Mrows / GB_M, ntiles = ntn * ntm;
    const int id = blockIdx.x;
    const int per = (ntiles + 7) / 8;
    const int t = (id % 8) * per + id / 8;
    if (t >= ntiles) return;
    const int m0 = (t / ntn) * GB_M, n0 = (t % ntn) * GB_N;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    // global -> LDS: each operand tile is 128 rows x 32 k = 512 chunks of 8;
    // thread tid moves chunks tid and tid + 256 (rows tid / 4 and tid / 4 + 64),
    // staged in named registers (an array here ends up in scratch memory)
    const int grow = tid >> 2, gq = (tid & 3) * 8;
    const T* ga = A + (int64_t)(m0 + grow) * K + gq;
    const T* gb = Bt + (int64_t)(n0 + grow) * K + gq;
    const int64_t gs = (int64_t)64 * K;
    uint4 ra0, ra1, rb0, rb1;
    auto gload = [&](int k0) {
        ra0 = *(const uint4*)(ga + k0);
        ra1 = *(const uint4*)(ga + gs + k0);
        rb0 = *(const uint4*)(gb + k0);
        rb1 = *(const uint4*)(gb + gs + k0);
    };
    auto lstore = [&](int buf) {
        *(uint4*)(&As[buf][grow * G_LD + gq]) = ra0;
        *(uint4*)(&As[buf][(grow + 64) * G_LD + gq]) = ra1;
        *(uint4*)(&Bs[buf][grow * G_LD + gq]) = rb0;
        *(uint4*)(&Bs[buf][(grow + 64) * G_LD + gq]) = rb1;
    };
    gload(0);
    lstore(0);
    __syncthreads();
    const int nk = K / GB_K;
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    for (int kt = 0; kt < nk; kt++) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * GB_K);
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 16) {
            const T* a0 = &As[buf][(wm + lr) * G_LD + kk + lk];
            const T* a1 = &As[buf][(wm + 32 + lr) * G_LD + kk + lk];
            const T* b0 = &Bs[buf][(wn + lr) * G_LD + kk + lk];
            const T* b1 = &Bs[buf][(wn + 32 + lr) * G_LD + kk + lk];
            acc[0][0] = mfma32<T>(a0, b0, acc[0][0]);
            acc[0][1] = mfma32<T>(a0, b1, acc[0][1]);
            acc[1][0] = mfma32<T>(a1, b0, acc[1][0]);
            acc[1][1] = mfma32<T>(a1, b1, acc[1][1]);
        }
        if (kt + 1 < nk) {
            lstore(buf ^ 1);
        }
        __syncthreads();
    }
    // D layout of 32x32 MFMA: lane l holds column l % 32, element e at row
    // (e / 4) * 8 + (l / 32) * 4 + e % 4.
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = m0 + wm + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                const int col = n0 + wn + j * 32 + lr;
                C[(int64_t)row * N + col] = acc[i][j][e];
            }
}

// ---- NT GEMM, 256 x 256 workgroup tiles -------------------------------------------
// Each of the 4 waves owns a 128 x 128 block (4 x 4 MFMA 32x32 tiles, 256
// accumulators per lane): per k16 step a wave reads 8 fragments (8 KB of LDS)
// for 16 MFMAs, half the LDS bytes per flop of the 128 x 128 kernel, whose
// fragment reads pace its MFMAs.  One wave per SIMD, so a k tile is 64 deep
// (2048 MFMA cycles per wave between barriers, enough to cover the next tile's
// HBM loads); LDS double buffered (2 x 72 KB, rows padded to 72 elements).
// Loads run two k tiles ahead.  Requires Mrows % 256 == 0, N % 256 == 0,
// K % 128 == 0 (checked on the host).
constexpr int G2_M = 256, G2_N = 256, G2_K = 64, G2_LD = G2_K + 8;
constexpr size_t G2_LDS_BYTES = (size_t)2 * (G2_M + G2_N) * G2_LD * 2;

template <typename T, bool EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_gemm_nt256(const T* __restrict__ A, const T* __restrict__ Bt,
                                                    float* __restrict__ C, int Mrows, int N, int K, ProdEpi E) {
    extern __shared__ __align__(16) unsigned char gsm[];
    T* As = (T*)gsm;                          // [2][G2_M * G2_LD]
    T* Bs = As + 2 * G2_M * G2_LD;            // [2][G2_N * G2_LD]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // XCD-aware order: each XCD takes a contiguous range of tiles, column tiles
    // fastest, so the column tiles of one row block share that XCD's L2.
    const int ntn = N / G2_N, ntm = Mrows / G2_M, ntiles = ntn * ntm;
    const int id = blockIdx.x;
    const int per = (ntiles + 7) / 8;
    const int t = (id % 8) * per + id / 8;
    if (t >= ntiles) return;
    const int m0 = (t / ntn) * G2_M, n0 = (t % ntn) * G2_N;
    const int wm = (w >> 1) * 128, wn = (w & 1) * 128;

    f32x16 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    // global -> LDS: each operand tile is 256 rows x 64 k = 2048 chunks of 8;
    // thread tid moves chunks tid + 256 s (rows tid / 8 + 32 s, chunk tid % 8).
    // Two named register sets (arrays here end up in scratch memory) hold the
    // next two k tiles: tile kt + 2 is requested while tile kt is multiplied,
    // and tile kt + 1, requested an iteration earlier, is then stored to LDS.
    const int grow = tid >> 3, gq = (tid & 7) * 8;
    const T* ga = A + (int64_t)(m0 + grow) * K + gq;
    const T* gb = Bt + (int64_t)(n0 + grow) * K + gq;
    const int64_t gs = (int64_t)32 * K;                 // 32 rows further
    const int lo = grow * G2_LD + gq, ls = 32 * G2_LD;
#define FDMC_G2_SET(p) uint4 p##a0, p##a1, p##a2, p##a3, p##a4, p##a5, p##a6, p##a7, \
                             p##b0, p##b1, p##b2, p##b3, p##b4, p##b5, p##b6, p##b7
#define FDMC_G2_LOAD(p, k0)                                     \
    do {                                                        \
        const int k_ = (k0);                                    \
        p##a0 = *(const uint4*)(ga + k_);                       \
        p##a1 = *(const uint4*)(ga + gs + k_);                  \
        p##a2 = *(const uint4*)(ga + 2 * gs + k_);              \
        p##a3 = *(const uint4*)(ga + 3 * gs + k_);              \
        p##a4 = *(const uint4*)(ga + 4 * gs + k_);              \
        p##a5 = *(const uint4*)(ga + 5 * gs + k_);              \
        p##a6 = *(const uint4*)(ga + 6 * gs + k_);              \
        p##a7 = *(const uint4*)(ga + 7 * gs + k_);              \
        p##b0 = *(const uint4*)(gb + k_);                       \
        p##b1 = *(const uint4*)(gb + gs + k_);                  \
        p##b2 = *(const uint4*)(gb + 2 * gs + k_);              \
        p##b3 = *(const uint4*)(gb + 3 * gs + k_);              \
        p##b4 = *(const uint4*)(gb + 4 * gs + k_);              \
        p##b5 = *(const uint4*)(gb + 5 * gs + k_);              \
        p##b6 = *(const uint4*)(gb + 6 * gs + k_);              \
        p##b7 = *(const uint4*)(gb + 7 * gs + k_);              \
    } while (0)
#define FDMC_G2_STORE(p, buf)                                   \
    do {                                                        \
        T* a_ = As + (buf) * G2_M * G2_LD + lo;                 \
        T* b_ = Bs + (buf) * G2_N * G2_LD + lo;                 \
        *(uint4*)(a_) = p##a0;                                  \
        *(uint4*)(a_ + ls) = p##a1;                             \
        *(uint4*)(a_ + 2 * ls) = p##a2;                         \
        *(uint4*)(a_ + 3 * ls) = p##a3;                         \
        *(uint4*)(a_ + 4 * ls) = p##a4;                         \
        *(uint4*)(a_ + 5 * ls) = p##a5;                         \
        *(uint4*)(a_ + 6 * ls) = p##a6;                         \
        *(uint4*)(a_ + 7 * ls) = p##a7;                         \
        *(uint4*)(b_) = p##b0;                                  \
        *(uint4*)(b_ + ls) = p##b1;                             \
        *(uint4*)(b_ + 2 * ls) = p##b2;                         \
        *(uint4*)(b_ + 3 * ls) = p##b3;                         \
        *(uint4*)(b_ + 4 * ls) = p##b4;                         \
        *(uint4*)(b_ + 5 * ls) = p##b5;                         \
        *(uint4*)(b_ + 6 * ls) = p##b6;                         \
        *(uint4*)(b_ + 7 * ls) = p##b7;                         \
    } while (0)
    FDMC_G2_SET(x);
    FDMC_G2_SET(y);
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    using V8 = typename Frag<T>::type;
    // the eight fragments of one k16 step (rows wm + 32 i + l % 32 of A,
    // wn + 32 j + l % 32 of B, k offset 8 (l / 32))
    auto ldfrag = [&](int buf, int kk, V8* fa, V8* fb) {
        const T* Ab = As + buf * G2_M * G2_LD + kk + lk;
        const T* Bb = Bs + buf * G2_N * G2_LD + kk + lk;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            fa[i] = *(const V8*)(Ab + (wm + 32 * i + lr) * G2_LD);
            fb[i] = *(const V8*)(Bb + (wn + 32 * i + lr) * G2_LD);
        }
    };
    auto mma = [&](const V8* fa, const V8* fb) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma32v<T>(fa[i], fb[j], acc[i][j]);
    };
#define FDMC_G2_STQ(p, buf, u, v)                               \
    do {                                                        \
        T* a_ = As + (buf) * G2_M * G2_LD + lo;                 \
        T* b_ = Bs + (buf) * G2_N * G2_LD + lo;                 \
        *(uint4*)(a_ + u * ls) = p##a##u;                       \
        *(uint4*)(a_ + v * ls) = p##a##v;                       \
        *(uint4*)(b_ + u * ls) = p##b##u;                       \
        *(uint4*)(b_ + v * ls) = p##b##v;                       \
    } while (0)
    // One k tile from LDS buffer `cur`: the fragments of step kk + 16 are read
    // while the MFMAs of step kk run, and the register set `p` (the next tile,
    // loaded from global memory an iteration earlier) is written to buffer
    // `nb` a quarter per step, so LDS stores overlap the MFMAs too.  Each step
    // is its own scheduling region with the order fixed: MFMAs alternating
    // with the eight fragment reads, then with the four stores.
#define FDMC_G2_ORDER(nr)                                                            \
    do {                                                                             \
        for (int q_ = 0; q_ < (nr); q_++) {                                          \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                       \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                       \
        }                                                                            \
        for (int q_ = 0; q_ < 4; q_++) {                                             \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                       \
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);                       \
        }                                                                            \
        __builtin_amdgcn_sched_group_barrier(0x008, 16 - 4 - (nr), 0);               \
        __builtin_amdgcn_sched_barrier(0);                                           \
    } while (0)
#define FDMC_G2_TILE(cur, p, nb)                                \
    do {                                                        \
        V8 f0a[4], f0b[4], f1a[4], f1b[4];                      \
        ldfrag(cur, 0, f0a, f0b);                               \
        __builtin_amdgcn_sched_barrier(0);                      \
        ldfrag(cur, 16, f1a, f1b);                              \
        mma(f0a, f0b);                                          \
        FDMC_G2_STQ(p, nb, 0, 1);                               \
        FDMC_G2_ORDER(8);                                       \
        ldfrag(cur, 32, f0a, f0b);                              \
        mma(f1a, f1b);                                          \
        FDMC_G2_STQ(p, nb, 2, 3);                               \
        FDMC_G2_ORDER(8);                                       \
        ldfrag(cur, 48, f1a, f1b);                              \
        mma(f0a, f0b);                                          \
        FDMC_G2_STQ(p, nb, 4, 5);                               \
        FDMC_G2_ORDER(8);                                       \
        mma(f1a, f1b);                                          \
        FDMC_G2_STQ(p, nb, 6, 7);                               \
        FDMC_G2_ORDER(0);                                       \
    } while (0)
    const int nk = K / G2_K;                 // even (checked on the host)
    const int klast = (nk - 1) * G2_K;
    FDMC_G2_LOAD(x, 0);
    FDMC_G2_STORE(x, 0);
    FDMC_G2_LOAD(x, G2_K);                   // tile 1 in flight
    __syncthreads();
    // two k tiles per trip so that the register sets alternate; loads past the
    // last tile are clamped to it (their LDS copy is never read).  The global
    // loads are issued first in each half (scheduling barrier), so they have a
    // whole tile of MFMAs to arrive.
    for (int kt = 0; kt < nk; kt += 2) {
        FDMC_G2_LOAD(y, min((kt + 2) * G2_K, klast));
        __builtin_amdgcn_sched_barrier(0);
        FDMC_G2_TILE(0, x, 1);               // tile kt; stores tile kt + 1
        __syncthreads();
        FDMC_G2_LOAD(x, min((kt + 3) * G2_K, klast));
        __builtin_amdgcn_sched_barrier(0);
        FDMC_G2_TILE(1, y, 0);               // tile kt + 1; stores tile kt + 2
        __syncthreads();
    }
#undef FDMC_G2_TILE
#undef FDMC_G2_ORDER
#undef FDMC_G2_STQ
#undef FDMC_G2_SET
#undef FDMC_G2_LOAD
#undef FDMC_G2_STORE
    // D layout of 32x32 MFMA: lane l holds column l % 32, element e at row
    // (e / 4) * 8 + (l / 32) * 4 + e % 4.
    if constexpr (EPI) {
        // the lane's columns n0 + wn + 32 j + l % 32 all belong to asset
        // a = l % n; lanes l ^ n, l ^ 2n, ... < 32 apart hold the same rows and
        // asset at other dates
        const int a = lr & (E.n - 1);
        float sg = 0.0f;
#pragma unroll
        for (int k = 0; k < BASKET_MAX; k++) sg = (k == a) ? E.sig[k] : sg;
        float* Pp = E.P + (int64_t)((n0 + wn) >> 7) * Mrows * E.n;
        float prod[4][16];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                float p = 1.0f;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int col = n0 + wn + j * 32 + lr;
                    p *= (col < E.D) ? fmaf(sg, acc[i][j][e], E.c0) : 1.0f;
                }
                prod[i][e] = p;
            }
        // all 64 products of one shuffle distance are exchanged back to back, so
        // the cross-lane round trips overlap (one wait per distance instead of
        // one per product: 5 % of the launch)
        for (int sft = E.n; sft < 32; sft <<= 1) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int e = 0; e < 16; e++) prod[i][e] *= __shfl_xor(prod[i][e], sft);
        }
        if (lr < E.n) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int row = m0 + wm + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                    Pp[(int64_t)row * E.n + a] = prod[i][e];
                }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int row = m0 + wm + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                    const int col = n0 + wn + j * 32 + lr;
                    __builtin_nontemporal_store(acc[i][j][e], C + (int64_t)row * N + col);   // streamed once
                }
    }
}

// ---- host launchers ------------------------------------------------------------
hipError_t launch_gen_z(const uint32_t* sdir, int D, int ld, int R, int64_t first, int64_t count, bool bf16,
                        void* out, hipStream_t st) {
    if (ld % 64 || ld < D) return hipErrorInvalidValue;
    const int64_t tiles = ((first + count - 1) >> 8) - (first >> 8) + 1;
    if (R > 65535) return hipErrorInvalidValue;
    // a few tiles per workgroup (records loaded and tile words formed once per
    // chunk), with enough workgroups for three per CU several times over
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t want = (int64_t)ncu * 12;
    int64_t tpw = (tiles * R + want - 1) / want;
    tpw = tpw < 1 ? 1 : (tpw > 64 ? 64 : tpw);
    const int64_t nwg = (tiles + tpw - 1) / tpw;
    if (nwg > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (bf16)
        hipLaunchKernelGGL(k_gen_z<__bf16>, dim3((unsigned)nwg, R), dim3(256), 0, st, sdir, D, ld, first, count,
                           (int)tpw, (__bf16*)out);
    else
        hipLaunchKernelGGL(k_gen_z<_Float16>, dim3((unsigned)nwg, R), dim3(256), 0, st, sdir, D, ld, first, count,
                           (int)tpw, (_Float16*)out);
    return hipGetLastError();
}

hipError_t launch_gemm_nt(const void* A, const void* Bt, float* C, int Mrows, int N, int K, bool bf16,
                          hipStream_t st, const ProdEpi* epi) {
    if (Mrows % G2_M == 0 && N % G2_N == 0 && K % (2 * G2_K) == 0) {
        const int64_t tiles = (int64_t)(Mrows / G2_M) * (N / G2_N);
        const int64_t grid = (tiles + 7) / 8 * 8;
        if (grid > 0x7FFFFFFF) return hipErrorInvalidValue;
        ProdEpi E{};
        if (epi) {
            E = *epi;
            if (E.n < 2 || E.n > BASKET_MAX || (E.n & (E.n - 1)) || !E.P) return hipErrorInvalidValue;
        }
        const void* fn = bf16 ? (epi ? (const void*)&k_gemm_nt256<__bf16, true> : (const void*)&k_gemm_nt256<__bf16, false>)
                              : (epi ? (const void*)&k_gemm_nt256<_Float16, true> : (const void*)&k_gemm_nt256<_Float16, false>);
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G2_LDS_BYTES);
        if (e != hipSuccess) return e;
        void* args[] = {(void*)&A, (void*)&Bt, (void*)&C, (void*)&Mrows, (void*)&N, (void*)&K, (void*)&E};
        e = hipLaunchKernel(fn, dim3((unsigned)grid), dim3(256), args, G2_LDS_BYTES, st);
        if (e != hipSuccess) return e;
        return hipGetLastError();
    }
    if (epi) return hipErrorInvalidValue;   // the product epilogue needs the 256 x 256 tiling
    if (Mrows % GB_M || N % GB_N || K % GB_K) return hipErrorInvalidValue;
    const int64_t tiles = (int64_t)(Mrows / GB_M) * (N / GB_N);
    const int64_t grid = (tiles + 7) / 8 * 8;
    if (grid > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (bf16)
        hipLaunchKernelGGL(k_gemm_nt<__bf16>, dim3((unsigned)grid), dim3(256), 0, st, (const __bf16*)A,
                           (const __bf16*)Bt, C, Mrows, N, K);
    else
        hipLaunchKernelGGL(k_gemm_nt<_Float16>, dim3((unsigned)grid), dim3(256), 0, st, (const _Float16*)A,
                           (const _Float16*)Bt, C, Mrows, N, K);
    return hipGetLastError();
}

}  // namespace fdmc
