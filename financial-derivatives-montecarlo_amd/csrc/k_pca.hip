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
//               whenever the shape allows it;
//   k_gemm_prod: 256 x 256 tiles that store no C: they multiply the Euler
//               factors of a GBM basket per row, asset and 128-column slab
//               (ProdEpi), so dW never reaches memory.  One persistent
//               workgroup per CU, operands streamed straight into LDS
//               (global_load_lds_dwordx4) four 32-deep k stages ahead of use.
#include "k_path.inl"
#include <algorithm>
#include <type_traits>

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
// replica r (lane = point) one 64-dimension chunk at a time.  At a chunk's start
// its direction records and the first tile's words go to LDS and each lane forms
// the scrambled Sobol words of its point in the chunk's 64 dimensions
// (sobol_coord_lds); they stay in registers for the chunk's tiles: from tile n - 1
// to tile n, gray(256 n) changes in bits 7 and 8 + ctz(n) only, so every word
// takes one XOR with v_7 ^ v_{8+ctz(n)} of its dimension, read from a table of
// these rows that replaces the records in LDS once the words are formed (the
// same for all lanes: broadcast reads).  The only other LDS read per
// coordinate is its inverse-CDF table row, eight of them in flight.  The
// normals are staged through LDS so that global stores are 16-byte and
// row-contiguous.  49 KB of LDS: three workgroups per CU.  Columns D..ld-1 are
// written as zeros.
constexpr int GZ_TW = 72;   // staged row: 64 values + 8 of padding

// icdf_u32_tab(s) with the sign applied by one v_bitop3: the table's magnitude
// p >= 0 takes the sign bit where s has none (p | (~s & sign), truth table 0xF2
// for sources p, s, sign), bit-identical to negating the scaled form.  As an
// intrinsic the sign step is not merged across coordinates by the SLP
// vectorizer, whose two-wide form of the negation the backend cannot select
// when packed FP32 instructions are disabled.
__device__ __forceinline__ float icdf_u32_tab_sgn(uint32_t s, const float4* tab) {
    const uint32_t b = __float_as_uint(fmaf((float)(int32_t)s, 0x1p-32f, 0x1p-33f));
    const float4 c = icdf_row<false>(tab, b);
    const float v = __uint_as_float((b & 0x7FFFFu) | 0x3F800000u);
    const float p = fmaf(fmaf(fmaf(c.w, v, c.z), v, c.y), v, c.x);
    return __uint_as_float(__builtin_amdgcn_bitop3_b32(__float_as_uint(p), s, 0x80000000u, 0xF2));
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gen_z(const uint32_t* __restrict__ sdir, int D, int ld, int64_t first,
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
        __syncthreads();                              // the previous chunk's records and staging are no longer read
        for (int i = tid; i < dn * 16; i += 256) rec[(i >> 4) * 20 + (i & 15)] = gchunk[(int64_t)(i >> 4) * SDIR_WORDS + (i & 15)];
        if (dn > 0) sobol_tile_words(rec, gchunk, dn, (uint32_t)tfirst, true, tid);
        __syncthreads();
        uint32_t cur[64];
#pragma unroll
        for (int j = 0; j < 64; j++) cur[j] = (j < dn) ? sobol_coord_lds(sl, j) : 0u;
        // The records are read; their LDS now holds the tile-to-tile updates
        // dx[z][j] = v_7 ^ v_{8+z} of dimension d0 + j for z = ctz(n) < 16 (0 for
        // the padding dimensions j >= dn), rows of 64 words that every lane
        // reads alike (broadcast ds_read_b128); row 16 is filled when a tile
        // has ctz(n) >= 16, row 17 stays 0 for the chunk's first tile.  One
        // unconditional update per tile keeps cur[] in the same registers.
        __syncthreads();
        for (int i = tid; i < 18 * 64; i += 256) {
            const int z = i >> 6, j = i & 63;
            rec[i] = (j < dn && z < 16) ? gchunk[j * SDIR_WORDS + 32 + 7] ^ gchunk[j * SDIR_WORDS + 32 + 8 + z] : 0u;
        }
        __syncthreads();
        for (int64_t t = tfirst; t <= tend; t++) {
            int zr = 17;
            if (t > tfirst) {
                const int z = __builtin_ctz((uint32_t)t);   // t > tfirst >= 0, t < 2^24: (uint32_t)t != 0
                zr = z < 16 ? z : 16;
                if (z >= 16) {                            // uniform; row 16's last readers passed
                    if (tid < 64)                         // the previous tile's barriers
                        rec[16 * 64 + tid] = (tid < dn) ? gchunk[tid * SDIR_WORDS + 32 + 7] ^
                                                              gchunk[tid * SDIR_WORDS + 32 + 8 + z]
                                                        : 0u;
                    __syncthreads();
                }
            }
            {
                const uint4* row = (const uint4*)(rec + zr * 64);
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const uint4 x = row[q];
                    cur[4 * q] ^= x.x;
                    cur[4 * q + 1] ^= x.y;
                    cur[4 * q + 2] ^= x.z;
                    cur[4 * q + 3] ^= x.w;
                }
            }
            __syncthreads();                          // staging free again
            // a full chunk takes a branch-free copy of the loop: under a
            // per-coordinate (j < dn) test each table read sat in its own branch
            // and was waited for alone; unconditional, the reads of a row of 8
            // coordinates are in flight together
            auto stage_row = [&](auto full) {
#pragma unroll
                for (int j0 = 0; j0 < 64; j0 += 8) {
                    union { T h[8]; uint4 u; } v;
                    if constexpr (decltype(full)::value) {
                        float4 c[8];
                        float vv[8];
#pragma unroll
                        for (int j = 0; j < 8; j++) {
                            const uint32_t b = __float_as_uint(fmaf((float)(int32_t)cur[j0 + j], 0x1p-32f, 0x1p-33f));
                            c[j] = icdf_row<false>(itab, b);
                            vv[j] = __uint_as_float((b & 0x7FFFFu) | 0x3F800000u);
                        }
#pragma unroll
                        for (int j = 0; j < 8; j++) {
                            const float p = fmaf(fmaf(fmaf(c[j].w, vv[j], c[j].z), vv[j], c[j].y), vv[j], c[j].x);
                            v.h[j] = to_half<T>(__uint_as_float(
                                __builtin_amdgcn_bitop3_b32(__float_as_uint(p), cur[j0 + j], 0x80000000u, 0xF2)));
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < 8; j++) {
                            const float z = (j0 + j < dn) ? icdf_u32_tab_sgn(cur[j0 + j], itab) : 0.0f;
                            v.h[j] = to_half<T>(z);
                        }
                    }
                    *(uint4*)(tile + tid * GZ_TW + j0) = v.u;
                }
            };
            if (dn == 64) stage_row(std::true_type{});
            else stage_row(std::false_type{});
            __syncthreads();
            // 256 rows x 64 values = 2048 chunks of 8 values: thread tid stores
            // chunk (tid & 7) of rows tid / 8 + 32 k (d0 + 64 <= ld: no column
            // test).  A tile inside [first, first + count) takes a copy of the
            // loop without the row test, whose reads and stores are not held
            // in per-row branches.
            const int64_t rel = (t << 8) - first;        // the tile's first row in the output
            const int row0 = tid >> 3, q = (tid & 7) * 8;
            T* ob = out + ((int64_t)r * count + rel + row0) * ld + d0 + q;
            auto store_rows = [&](auto inner) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int64_t pr = rel + row0 + 32 * k;
                    if (decltype(inner)::value || (pr >= 0 && pr < count))
                        *(uint4*)(ob + (int64_t)32 * k * ld) = *(const uint4*)(tile + (row0 + 32 * k) * GZ_TW + q);
                }
            };
            if (rel >= 0 && rel + 256 <= count) store_rows(std::true_type{});
            else store_rows(std::false_type{});
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

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_gemm_nt256(const T* __restrict__ A, const T* __restrict__ Bt,
                                                    float* __restrict__ C, int Mrows, int N, int K) {
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


// ---- GEMM with the basket-product epilogue: persistent, operands DMA'd to LDS --
// dW = Z M^T for 256 x 256 tiles (4 waves of 128 x 128, as k_gemm_nt256), but
// each tile ends in the Euler-factor products of ProdEpi instead of a C store,
// and the kernel is shaped around that:
//   - one workgroup per CU walks its XCD's contiguous tile range with a stride of
//     the XCD's workgroup count (the 4 column tiles of 8 row blocks in flight on
//     an XCD share its L2), so a tile's epilogue and stores overlap the next
//     tile's loads instead of holding the CU's LDS until the stores drain;
//   - operands go from global memory straight into LDS with
//     global_load_lds_dwordx4 (no staging registers, no LDS store instructions)
//     in stages of BK = 32: stage = 256 A rows + 256 B rows of 64 bytes, 16-byte
//     chunks XOR-swizzled by ((row >> 1) ^ (row >> 2)) & 3 (conflict-free
//     fragment reads), four stages in four LDS objects (128 KB);
//   - the DMA is one stream of stages across tiles: stage s + 2's B half is
//     issued during stage s's first k16 step, its A half (stage s + 3) after the
//     stage's barrier, spaced one instruction per four MFMAs; a tile's last
//     stages already load the next tile's first ones;
//   - per stage one barrier: each wave waits for its own loads of stage s + 1
//     (s_waitcnt vmcnt), then s_barrier makes stage s + 1 visible and releases
//     stage s - 1's buffer; fragments are read one k16 step ahead, across stage
//     boundaries;
//   - the epilogue runs band by band (32 rows), so it needs few registers inside
//     the tile loop; for 16 assets one v_permlane16_swap pairs the two half-row
//     partial products of two rows at once (NA = 16), other asset counts
//     combine lanes with shuffles.
// Accumulation order is that of k_gemm_nt256 (k ascending, one MFMA per k16
// step and 32 x 32 block), so the products are bit-identical to forming them
// from k_gemm_nt256's dW.  Requires Mrows % 256 == 0, N % 256 == 0,
// K % 128 == 0 and a grid that is a multiple of 8 (checked on the host).
constexpr int GP_BK = 32, GP_RB = 64, GP_STAGE = 512 * GP_RB, GP_NS = 4;

__device__ __forceinline__ int gp_swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// s_waitcnt vmcnt(N) with expcnt and lgkmcnt left free (gfx9 encoding)
template <int N_>
__device__ __forceinline__ void gp_wait_vm() {
    static_assert(N_ >= 0 && N_ < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N_ & 15) | ((N_ >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <typename T, int NA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_gemm_prod(
    const T* __restrict__ A, const T* __restrict__ Bt, int Mrows, int N, int K, ProdEpi E) {
    __shared__ __align__(1024) unsigned char b0[GP_STAGE], b1[GP_STAGE], b2[GP_STAGE], b3[GP_STAGE];
    // stage buffers are distinct objects, so the compiler can tell a DMA into one
    // from fragment reads of another (no blanket vmcnt(0) before LDS reads)
    auto stage = [&](int j) -> unsigned char* { return j == 0 ? b0 : j == 1 ? b1 : j == 2 ? b2 : b3; };
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ntn = N / G2_N, ntiles = ntn * (Mrows / G2_M);
    const int nx = gridDim.x >> 3;                      // workgroups per XCD
    const int per = (ntiles + 7) / 8;
    const int tend = min(ntiles, (int)(blockIdx.x & 7) * per + per);
    int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= tend) return;
    const int wm = (w >> 1) * 128, wn = (w & 1) * 128;
    // DMA: a wave instruction fills one 1 KB block of 16 rows; lane l writes
    // bytes 16 l .. 16 l + 15 of it = row l / 4, physical chunk l % 4, i.e.
    // logical chunk (l % 4) ^ swz(row).  Wave w fills blocks w + 4 q of A and B.
    const int drow = lane >> 2;
    const uint32_t voff = (uint32_t)(drow * K + (((lane & 3) ^ gp_swz(drow)) * 8)) * (uint32_t)sizeof(T);
    const int64_t bstep = (int64_t)64 * K * (int64_t)sizeof(T);         // 4 blocks further
    const int nk = K / GP_BK;
    auto baseA = [&](int tt) { return (const char*)(A + (int64_t)((tt / ntn) * G2_M + 16 * w) * K); };
    auto baseB = [&](int tt) { return (const char*)(Bt + (int64_t)((tt % ntn) * G2_N + 16 * w) * K); };
    const char* gA = baseA(t);
    const char* gB = baseB(t);
    int tn = t + nx;
    const char* gAn = tn < tend ? baseA(tn) : gA;       // after the last tile: reloads never read
    const char* gBn = tn < tend ? baseB(tn) : gB;
    auto dma_half = [&](const char* p, int ks, unsigned char* st) {
        const int64_t k0 = (int64_t)ks * GP_BK * (int64_t)sizeof(T);
#pragma unroll
        for (int q = 0; q < 4; q++)
            __builtin_amdgcn_global_load_lds((const void*)(p + (q * bstep + k0) + voff),
                                             (void __attribute__((address_space(3)))*)(st + (w + 4 * q) * 1024), 16, 0, 0);
    };
    // fragments: lane l reads row 32 i + l % 32 at logical chunk 2 step + l / 32
    const int lr = lane & 31, hi = lane >> 5, fs = gp_swz(lr);
    const int fo0 = lr * GP_RB + (((0 + hi) ^ fs) << 4), fo1 = lr * GP_RB + (((2 + hi) ^ fs) << 4);
    using V8 = typename Frag<T>::type;
    auto ldfrag = [&](const unsigned char* st, int f, V8* fa, V8* fb) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            fa[i] = *(const V8*)(st + (wm + 32 * i) * GP_RB + f);
            fb[i] = *(const V8*)(st + 256 * GP_RB + (wn + 32 * i) * GP_RB + f);
        }
    };
    f32x16 acc[4][4];
    auto mma = [&](const V8* fa, const V8* fb, bool zero) {
        const f32x16 z = {};
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = mfma32v<T>(fa[i], fb[j], zero ? z : acc[i][j]);
    };
    // 16 MFMAs with 8 fragment reads between them and `nv` DMA instructions
    // spaced evenly; nothing crosses the group
#define FDMC_GP_ORDER(nv)                                                            \
    do {                                                                             \
        for (int m_ = 0; m_ < 16; m_++) {                                            \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                       \
            if (m_ % 2 == 0) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);      \
            if (m_ % (16 / (nv)) == 16 / (nv) - 1) __builtin_amdgcn_sched_group_barrier(0x010, 1, 0); \
        }                                                                            \
        __builtin_amdgcn_sched_barrier(0);                                           \
    } while (0)
    V8 fa[2][4], fb[2][4];
    // one stage s in buffer j: step 0 (B half of stage s + 2 issued, step 1's
    // fragments read), wait for stage s + 1 (own loads: 4 B + 4 A instructions
    // of stage s + 2 may still be in flight), barrier, A half of stage s + 3
    // issued, step 1 (step 0 fragments of stage s + 1 read)
    auto run_stage = [&](int s, int j, bool zero) {
        {
            const int g = s + 2;
            const bool cur = g < nk;
            dma_half(cur ? gB : gBn, cur ? g : g - nk, stage((j + 2) % GP_NS) + 256 * GP_RB);
        }
        ldfrag(stage(j), fo1, fa[1], fb[1]);
        mma(fa[0], fb[0], zero);
        FDMC_GP_ORDER(4);
        gp_wait_vm<8>();
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        {
            const int g = s + 3;
            const bool cur = g < nk;
            dma_half(cur ? gA : gAn, cur ? g : g - nk, stage((j + 3) % GP_NS));
        }
        ldfrag(stage((j + 1) % GP_NS), fo0, fa[0], fb[0]);
        mma(fa[1], fb[1], false);
        FDMC_GP_ORDER(4);
    };
    // prologue: stages 0 and 1 whole, stage 2's A half
    dma_half(gA, 0, stage(0));
    dma_half(gB, 0, stage(0) + 256 * GP_RB);
    dma_half(gA, 1, stage(1));
    dma_half(gB, 1, stage(1) + 256 * GP_RB);
    dma_half(gA, 2, stage(2));
    bool first = true;
    const int a_n = lr & ((NA ? NA : E.n) - 1);        // this lane's asset
    float sg = 0.0f;
#pragma unroll
    for (int k = 0; k < BASKET_MAX; k++) sg = (k == a_n) ? E.sig[k] : sg;
    while (true) {
        const int m0 = (t / ntn) * G2_M, n0 = (t % ntn) * G2_N;
        // stage 0 loaded in every wave; after an epilogue its stores are in
        // flight too, and stores and loads may complete out of order: wait for all
        if (first) gp_wait_vm<12>(); else gp_wait_vm<0>();
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        ldfrag(stage(0), fo0, fa[0], fb[0]);
        // the first four stages peeled: the first MFMA of a tile starts from 0
#pragma unroll
        for (int j = 0; j < GP_NS; j++) run_stage(j, j, j == 0);
        for (int sb = GP_NS; sb < nk; sb += GP_NS) {
#pragma unroll
            for (int j = 0; j < GP_NS; j++) run_stage(sb + j, j, false);
        }
        // epilogue.  D layout of 32x32 MFMA: lane l holds column l % 32, element
        // e at row (e / 4) * 8 + (l / 32) * 4 + e % 4.  The lane's columns
        // n0 + wn + 32 j + l % 32 all belong to asset l % n; lanes l ^ n,
        // l ^ 2n, ... < 32 apart hold the same rows and asset at other dates.
        float* Pp = E.P + (int64_t)((n0 + wn) >> 7) * Mrows * (NA ? NA : E.n);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            __builtin_amdgcn_sched_barrier(0);
            // this band's accumulators moved to VGPRs here (left to itself the
            // compiler copies all 256 before the first band and spills)
            float x[4][16];
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 16; e++)
                    asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(x[j][e]) : "a"(acc[i][j][e]));
            float prod[16];
#pragma unroll
            for (int e = 0; e < 16; e++) {
                float p = 1.0f;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int col = n0 + wn + j * 32 + lr;
                    p *= (col < E.D) ? fmaf(sg, x[j][e], E.c0) : 1.0f;   // padding columns count as 1
                }
                prod[e] = p;
            }
            if constexpr (NA == 16) {
                // one swap of the odd 16-lane rows of prod[e] with the even rows of
                // prod[e + 1] puts e's two partials side by side in lanes 0-15 and
                // e + 1's in lanes 16-31: one product gives lanes 0-15 row e and
                // lanes 16-31 row e + 1 (one row further), which they store
                float* pb = Pp + (int64_t)(m0 + wm + i * 32 + hi * 4 + (lr >> 4)) * 16 + a_n;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int e = 2 * q;
                    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(prod[e]), __float_as_uint(prod[e + 1]),
                                                              false, false);
                    pb[((e >> 2) * 8 + (e & 3)) * 16] = __uint_as_float(r[0]) * __uint_as_float(r[1]);
                }
            } else {
                for (int sft = E.n; sft < 32; sft <<= 1) {
#pragma unroll
                    for (int e = 0; e < 16; e++) prod[e] *= __shfl_xor(prod[e], sft);
                }
                if (lr < E.n) {
#pragma unroll
                    for (int e = 0; e < 16; e++) {
                        const int row = m0 + wm + i * 32 + (e >> 2) * 8 + hi * 4 + (e & 3);
                        Pp[(int64_t)row * E.n + a_n] = prod[e];
                    }
                }
            }
        }
        if (tn >= tend) break;
        t = tn;
        tn += nx;
        gA = gAn;
        gB = gBn;
        gAn = tn < tend ? baseA(tn) : gA;
        gBn = tn < tend ? baseB(tn) : gB;
        first = false;
    }
    gp_wait_vm<0>();   // no DMA into LDS outlives the workgroup
#undef FDMC_GP_ORDER
}

// ---- host launchers ------------------------------------------------------------
hipError_t launch_gen_z(const uint32_t* sdir, int D, int ld, int R, int64_t first, int64_t count, bool bf16,
                        void* out, hipStream_t st) {
    if (ld % 64 || ld < D) return hipErrorInvalidValue;
    const int64_t tiles = ((first + count - 1) >> 8) - (first >> 8) + 1;
    if (R > 65535) return hipErrorInvalidValue;
    // several tiles per workgroup (records loaded and tile words formed once per
    // chunk), as many workgroups as are resident at once (three per CU): one
    // round with no tail (several rounds of smaller workgroups measured 9 %
    // slower on the C3 pass)
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t want = (int64_t)ncu * 3;
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
    if (epi) {
        // the product epilogue: persistent k_gemm_prod, one workgroup per CU
        // (a multiple of 8, for the per-XCD tile ranges)
        if (Mrows % G2_M || N % G2_N || K % (GP_NS * GP_BK)) return hipErrorInvalidValue;
        ProdEpi E = *epi;
        if (E.n < 2 || E.n > BASKET_MAX || (E.n & (E.n - 1)) || !E.P) return hipErrorInvalidValue;
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        const int64_t tiles = (int64_t)(Mrows / G2_M) * (N / G2_N);
        const int64_t grid = std::max<int64_t>(8, std::min<int64_t>((ncu / 8) * 8, (tiles + 7) / 8 * 8));
        const void* fn = bf16 ? (E.n == 16 ? (const void*)&k_gemm_prod<__bf16, 16> : (const void*)&k_gemm_prod<__bf16, 0>)
                              : (E.n == 16 ? (const void*)&k_gemm_prod<_Float16, 16> : (const void*)&k_gemm_prod<_Float16, 0>);
        void* args[] = {(void*)&A, (void*)&Bt, (void*)&Mrows, (void*)&N, (void*)&K, (void*)&E};
        hipError_t e = hipLaunchKernel(fn, dim3((unsigned)grid), dim3(256), args, 0, st);
        if (e != hipSuccess) return e;
        return hipGetLastError();
    }
    if (Mrows % G2_M == 0 && N % G2_N == 0 && K % (2 * G2_K) == 0) {
        const int64_t tiles = (int64_t)(Mrows / G2_M) * (N / G2_N);
        const int64_t grid = (tiles + 7) / 8 * 8;
        if (grid > 0x7FFFFFFF) return hipErrorInvalidValue;
        const void* fn = bf16 ? (const void*)&k_gemm_nt256<__bf16> : (const void*)&k_gemm_nt256<_Float16>;
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G2_LDS_BYTES);
        if (e != hipSuccess) return e;
        void* args[] = {(void*)&A, (void*)&Bt, (void*)&C, (void*)&Mrows, (void*)&N, (void*)&K};
        e = hipLaunchKernel(fn, dim3((unsigned)grid), dim3(256), args, G2_LDS_BYTES, st);
        if (e != hipSuccess) return e;
        return hipGetLastError();
    }
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
