// Steady-state surrogate for warm starts: an MLP n_in -> hidden x layers -> 1 trained and run on the
// GPU by the kernels below.  Weights and activations are bfloat16, every product accumulates in
// FP32 on the MFMA units (v_mfma_f32_16x16x32_bf16), and the optimiser (Adam) keeps FP32 master
// weights.  The reference has no counterpart: it starts every sensor at its configured initial
// temperature (SensorController::t_steady_ -> getInitTemp, psim/include/psim/sensorController.h:63-67);
// the surrogate only proposes those initial temperatures, the Monte Carlo result stays the answer.
//
// Layout in HBM: activations row-major [batch][features] (bf16), weights [in][out] (bf16 copy of
// the FP32 master), pre-activations [batch][out] (FP32, kept for the ReLU backward).  The output
// layer is padded to PAD_OUT columns with weights that stay zero, so every product is a GEMM on
// 16 x 16 MFMA tiles.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <atomic>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "psim_gpu.h"
#include "psim_surrogate.h"

namespace {

constexpr int GEMM_THREADS = 256;  // 4 waves
constexpr int BATCH_ROUND = 64;  // batch buffers are sized in multiples of this many rows
// Rows of one call: GEMM grids put row blocks of 64 on grid.y, which holds at most 65535 blocks.
constexpr int MAX_ROWS = 65535 * 64;
constexpr int PAD_OUT = 16;  // output-layer columns (1 used)
constexpr int EW_BLOCK = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (NaN kept quiet).  On the device this is one
// v_cvt_pk_bf16_f32 (round to nearest even in the default mode, denormals kept), which gives the
// host routine's bits for every finite value; the host routine costs about seven VALU operations.
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) { return (uint16_t)((u >> 16) | 0x40u); }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
#endif
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two floats -> two bf16 in one word (lo in the low half), as f32_to_bf16 rounds each: one
// v_cvt_pk_bf16_f32 for the pair.
__device__ inline uint32_t pack_bf16x2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{ lo, hi }, bf16x2));
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// C[M][N] = op(A)[M][K] x op(B)[K][N] in FP32.  Stored operands (bf16 bits):
//   A: TA ? [K][M] : [M][K];  B: TB ? [N][K] : [K][N].
// A block of 256 threads (4 waves) computes a BM x BN tile; each wave a WM x WN sub-tile of
// (WM / 16) x (WN / 16) MFMA 16x16x32 accumulators.  Per K step of 32, every thread moves 16-byte
// vectors of both operands from HBM into registers while the waves run the MFMAs of the previous
// step, then into the other of two LDS buffers, K-contiguous (As[m][k], Bs[n][k]), so a lane's
// eight K values of its MFMA fragment are one 16-byte LDS read.  One barrier per step: a buffer is
// refilled only after every wave has passed the barrier that ends the step reading it.
// Vectors that cross the matrix edge (or operands whose rows are not 16-byte aligned: vec_a /
// vec_b = 0) load element by element; out-of-range elements load as zero.
constexpr int BK = 32;
constexpr int LDK = BK + 8;  // LDS row pitch in bf16 (80 bytes: 16-byte aligned, staggers banks)

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <bool TA, bool TB, int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(GEMM_THREADS) gemm_bf16_kernel(const uint16_t* __restrict__ A,
    const uint16_t* __restrict__ B, float* __restrict__ C, int M, int N, int K, int vec_a, int vec_b, int k_split) {
    constexpr int WAVES_N = BN / WN;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int AV = BM * BK / 8 / GEMM_THREADS;  // 16-byte vectors per thread per step
    constexpr int BV = BN * BK / 8 / GEMM_THREADS;
    static_assert((BM / WM) * (BN / WN) == GEMM_THREADS / 64, "one sub-tile per wave");
    static_assert(AV >= 1 && BV >= 1, "tile too small for the thread count");
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN][LDK];
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wm0 = (wave / WAVES_N) * WM, wn0 = (wave % WAVES_N) * WN;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    u32x4 ra[AV], rb[BV];

    // One operand tile (rows r0 .. r0 + R of the output dimension, K step k0) into registers.
    //   kmajor: stored [row][K] (K contiguous) -> vector v covers row v / 4, k (v % 4) * 8 .. + 8
    //   else:   stored [K][row] (rows contiguous) -> vector v covers k v / (R / 8), rows (v % (R / 8)) * 8 .. + 8
    auto load = [&](const uint16_t* __restrict__ X, int rows, int r0, int k0, bool kmajor, int vec, u32x4* reg,
                    int count, int R) {
        for (int i = 0; i < count; ++i) {
            const int v = tid + i * GEMM_THREADS;
            int r, k;
            if (kmajor) {
                r = r0 + v / (BK / 8);
                k = k0 + (v % (BK / 8)) * 8;
            } else {
                k = k0 + v / (R / 8);
                r = r0 + (v % (R / 8)) * 8;
            }
            const bool full = kmajor ? (r < rows && k + 8 <= K) : (k < K && r + 8 <= rows);
            if (vec && full) {
                const int64_t off = kmajor ? (int64_t)r * K + k : (int64_t)k * rows + r;
                reg[i] = *reinterpret_cast<const u32x4*>(X + off);
            } else {
                uint16_t e[8];
                for (int j = 0; j < 8; ++j) {
                    const int rr = kmajor ? r : r + j, kk = kmajor ? k + j : k;
                    e[j] = (rr < rows && kk < K) ? X[kmajor ? (int64_t)rr * K + kk : (int64_t)kk * rows + rr] : 0;
                }
                std::memcpy(&reg[i], e, 16);
            }
        }
    };
    auto store = [&](uint16_t (*S)[LDK], bool kmajor, const u32x4* reg, int count, int R) {
        for (int i = 0; i < count; ++i) {
            const int v = tid + i * GEMM_THREADS;
            if (kmajor) {
                std::memcpy(&S[v / (BK / 8)][(v % (BK / 8)) * 8], &reg[i], 16);
            } else {
                uint16_t e[8];
                std::memcpy(e, &reg[i], 16);
                const int k = v / (R / 8), r = (v % (R / 8)) * 8;
                for (int j = 0; j < 8; ++j) { S[r + j][k] = e[j]; }
            }
        }
    };

    f32x4 acc[TM][TN];
    for (int i = 0; i < TM; ++i) {
        for (int j = 0; j < TN; ++j) { acc[i][j] = f32x4{ 0.f, 0.f, 0.f, 0.f }; }
    }
    // K range of this block: split z of a product split over K (k_split a multiple of BK), whose
    // partial product goes to its own M x N slice of C
    const int kb = blockIdx.z * k_split;
    const int nk = (min(K, kb + k_split) - kb + BK - 1) / BK;
    C += (int64_t)blockIdx.z * M * N;
    load(A, M, m0, kb, !TA, vec_a, ra, AV, BM);
    load(B, N, n0, kb, TB, vec_b, rb, BV, BN);
    store(As[0], !TA, ra, AV, BM);
    store(Bs[0], TB, rb, BV, BN);
    __syncthreads();
    const int kq = (lane / 16) * 8;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) {
            load(A, M, m0, kb + (kt + 1) * BK, !TA, vec_a, ra, AV, BM);
            load(B, N, n0, kb + (kt + 1) * BK, TB, vec_b, rb, BV, BN);
        }
        bf16x8 fa[TM], fb[TN];
        for (int i = 0; i < TM; ++i) { std::memcpy(&fa[i], &As[buf][wm0 + i * 16 + lane % 16][kq], 16); }
        for (int j = 0; j < TN; ++j) { std::memcpy(&fb[j], &Bs[buf][wn0 + j * 16 + lane % 16][kq], 16); }
        for (int i = 0; i < TM; ++i) {
            for (int j = 0; j < TN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
        }
        if (kt + 1 < nk) {
            store(As[buf ^ 1], !TA, ra, AV, BM);
            store(Bs[buf ^ 1], TB, rb, BV, BN);
        }
        __syncthreads();
    }
    // D fragment: lane holds column lane % 16, rows 4 (lane / 16) + e
    for (int i = 0; i < TM; ++i) {
        for (int j = 0; j < TN; ++j) {
            const int n = n0 + wn0 + j * 16 + lane % 16;
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wm0 + i * 16 + (lane / 16) * 4 + e;
                if (m < M && n < N) { C[(int64_t)m * N + n] = acc[i][j][e]; }
            }
        }
    }
}

// Tile shape of a product: 128 x 128 blocks (64 x 64 per wave) when they fill the chip several
// times over, else 64 x 64 blocks (32 x 32 per wave), which give the batch-sized products of
// training enough blocks.  A product with few output blocks and a long K (the weight gradients,
// K = batch rows) is split over K into partial products in a workspace, then summed in a fixed
// order (sum_splits_kernel), when the caller provides one.
constexpr int BIG_TILE_MIN_BLOCKS = 1024;
constexpr int SPLIT_TARGET_BLOCKS = 512;
constexpr int SPLIT_MIN_K = 512;  // per split
constexpr int MAX_SPLITS = 16;

// C[i] = sum over z of ws[z][i], z in order.
__global__ void sum_splits_kernel(const float* __restrict__ ws, float* __restrict__ C, int splits, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) { return; }
    float s = 0.f;
    for (int z = 0; z < splits; ++z) { s += ws[(int64_t)z * n + i]; }
    C[i] = s;
}

unsigned ew_blocks(int64_t n) { return (unsigned)((n + EW_BLOCK - 1) / EW_BLOCK); }

// Number of K splits of a product on 64 x 64 tiles (1: none); ws_elems floats of workspace.
int gemm_splits(int M, int N, int K, size_t ws_elems) {
    const int64_t blocks = (int64_t)((M + 63) / 64) * ((N + 63) / 64);
    if (blocks >= SPLIT_TARGET_BLOCKS || K < 2 * SPLIT_MIN_K) { return 1; }
    int splits = (int)std::min<int64_t>(MAX_SPLITS, (SPLIT_TARGET_BLOCKS + blocks - 1) / blocks);
    splits = std::min(splits, K / SPLIT_MIN_K);
    while (splits > 1 && (size_t)splits * (size_t)M * (size_t)N > ws_elems) { --splits; }
    return std::max(1, splits);
}

template <bool TA, bool TB>
void launch_gemm_t(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K, hipStream_t st, float* ws,
    size_t ws_elems) {
    const auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    const int vec_a = aligned(A) && ((TA ? M : K) % 8 == 0);
    const int vec_b = aligned(B) && ((TB ? K : N) % 8 == 0);
    const int64_t big = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
    if (big >= BIG_TILE_MIN_BLOCKS) {
        const dim3 grid((unsigned)((N + 127) / 128), (unsigned)((M + 127) / 128));
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, 128, 128, 64, 64>), grid, dim3(GEMM_THREADS), 0, st, A, B, C,
            M, N, K, vec_a, vec_b, K);
        return;
    }
    const int splits = ws ? gemm_splits(M, N, K, ws_elems) : 1;
    if (splits == 1) {
        const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, 64, 64, 32, 32>), grid, dim3(GEMM_THREADS), 0, st, A, B, C, M,
            N, K, vec_a, vec_b, K);
        return;
    }
    const int k_split = ((K + splits - 1) / splits + BK - 1) / BK * BK;
    const int used = (K + k_split - 1) / k_split;
    const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)used);
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, 64, 64, 32, 32>), grid, dim3(GEMM_THREADS), 0, st, A, B, ws, M, N, K,
        vec_a, vec_b, k_split);
    const int64_t n = (int64_t)M * N;
    hipLaunchKernelGGL(sum_splits_kernel, dim3(ew_blocks(n)), dim3(EW_BLOCK), 0, st, ws, C, used, n);
}

// ---------------------------------------------------------------------------------------------
// K-major products, C[M][N] = A[M][K] x B[N][K]^T: the layout of every product the network runs
// (forward: activations x transposed weights; input gradient: gradient x weights; weight gradient:
// the transposed activation copy x the transposed gradient copy, both written by the kernels that
// produce them).
//
// A block of WGM x WGN waves computes a BM x BN tile, each wave (BM / WGM) x (BN / WGN) on
// 16 x 16 x 32 MFMA accumulators.  K advances 64 per pipeline step: every thread loads its 16-byte
// vectors of the A and B tiles two steps ahead into registers while the waves run the current
// step's MFMAs from LDS (two 32-deep halves), and stores the next step's into the other LDS
// buffer; one barrier per step.  The epilogue stages the FP32 tile in LDS (over the pipeline buffers) and writes it with
// 16-byte stores, row-major and, when asked, column-major (the transposed copy a later product
// reads K-major).
// Blocks are dealt to the 8 XCDs in turn; tiles are numbered so that XCD x takes the x-th
// contiguous range of tiles, N tiles fastest, so the tiles of one BM-row band of A share an L2.
// A product split over K writes partial tiles to a workspace; the last block of a tile to finish
// (a per-tile arrival counter) sums the partials in split order, so the result does not depend on
// which block finishes last, and resets the counter to zero for the next product.
constexpr int KB = 64;  // K per pipeline step
constexpr int KP = KB + 8;  // LDS row pitch of the operand tiles (bf16): 144 bytes
constexpr int XCDS = 8;
constexpr int NT_COUNTERS = 4096;  // arrival counters: tiles of one split product at most

enum : int { EPI_F32 = 0, EPI_FWD = 1, EPI_BACK = 2 };

struct NtArgs {
    const uint16_t* A;  // [M][K] bf16
    const uint16_t* B;  // [N][K] bf16
    int M, N, K;
    int vec_a, vec_b;  // operand rows may be read as 16-byte vectors (K % 8 == 0, aligned bases)
    int vec_out, vec_out_t;  // 16-byte stores of result rows (N % 8 == 0) / transposed rows (M % 8 == 0)
    int k_split;  // K per split, a multiple of KB (K rounded up when the product is not split)
    float* c;  // EPI_F32: C [M][N]; EPI_FWD: pre-activation z [M][N] or null
    float* ws;  // split products: partial tiles [splits][M][N]
    unsigned* counters;  // split products summed in the kernel: one arrival counter per tile, zero
                         // between products (null: the partials are left for sum_splits_kernel)
    const float* bias;  // EPI_FWD: [N]
    int relu;  // EPI_FWD: ReLU after the bias
    const uint16_t* mask;  // EPI_BACK: the bf16 activation [M][N] whose positive entries pass the gradient
    uint16_t* out;  // EPI_FWD / EPI_BACK: bf16 result [M][N], or null
    uint16_t* out_t;  // EPI_FWD / EPI_BACK: bf16 result transposed [N][M], or null
};

template <int BM, int BN, int WGM, int WGN, int OCC, int EPI, bool VEC>
__global__ void __launch_bounds__(64 * WGM * WGN, OCC) gemm_nt_kernel(NtArgs p) {
    constexpr int NT_THREADS = 64 * WGM * WGN;
    constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 16, TN = WN / 16;
    constexpr int AV = BM * KB / 8 / NT_THREADS, BV = BN * KB / 8 / NT_THREADS;
    constexpr int CP = BN + 4;  // staging pitch (floats): the 4 row groups of an MFMA result fall on 2 bank halves
    constexpr int LOOP_BYTES = 2 * (BM + BN) * KP * 2;
    constexpr int STAGE_BYTES = BM * CP * 4;
    constexpr int SMEM = LOOP_BYTES > STAGE_BYTES ? LOOP_BYTES : STAGE_BYTES;
    constexpr int CH = BN / 8;  // 8-column chunks of a tile row
    static_assert(AV >= 1 && BV >= 1 && TM >= 1 && TN >= 1, "tile too small for the thread count");
    __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
    __shared__ int last_block;
    // operand tiles as 16-byte vectors of 8 bf16: [2][rows][KP / 8], the last vector of a row padding
    auto As = reinterpret_cast<u32x4 (*)[BM][KP / 8]>(smem);
    auto Bs = reinterpret_cast<u32x4 (*)[BN][KP / 8]>(smem + 2 * BM * KP * 2);
    auto Cs = reinterpret_cast<float (*)[CP]>(smem);  // [BM][CP], after the main loop

    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
    const int tiles_n = (p.N + BN - 1) / BN;
    const int tiles = tiles_n * ((p.M + BM - 1) / BM);
    int t = blockIdx.x;
    if (tiles % XCDS == 0) { t = (t % XCDS) * (tiles / XCDS) + t / XCDS; }
    const int m0 = (t / tiles_n) * BM, n0 = (t % tiles_n) * BN;
    const int split = blockIdx.y;
    const int kb = split * p.k_split, ke = min(p.K, kb + p.k_split);
    const int nk = (ke - kb + KB - 1) / KB;

    // rows r0 .. r0 + R of an operand, K step k0 .. k0 + KB, into registers: vector v covers row
    // v / 8, K columns (v % 8) * 8 .. + 8; what lies beyond the rows or this split's K loads as zero.
    // VEC (K % 8 == 0, aligned operands): raw buffer loads over this block's rows of the operand,
    // whose record count ends at the matrix's last row, so the hardware returns zeros past it; a
    // vector past this split's K gets an offset beyond the records.  No branch and no use of the
    // loaded value, so the loads of a step stay in flight until their LDS stores.
    u32x4 ra[AV], rb[BV];
    const __amdgpu_buffer_rsrc_t rs_a = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t*>(p.A + (int64_t)m0 * p.K), (short)0, min(BM, p.M - m0) * p.K * 2, 0x00020000);
    const __amdgpu_buffer_rsrc_t rs_b = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t*>(p.B + (int64_t)n0 * p.K), (short)0, min(BN, p.N - n0) * p.K * 2, 0x00020000);
    auto load = [&](__amdgpu_buffer_rsrc_t rs, const uint16_t* __restrict__ X, int rows, int r0, int k0, u32x4* reg,
                    int count) {
        for (int i = 0; i < count; ++i) {
            const int v = tid + i * NT_THREADS;
            const int r = r0 + v / (KB / 8), k = k0 + (v % (KB / 8)) * 8;
            if (VEC) {
                const int off = k < ke ? ((v / (KB / 8)) * p.K + k) * 2 : 0x7ffffff0;
                reg[i] = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
            } else {
                uint16_t e[8];
                for (int j = 0; j < 8; ++j) { e[j] = (r < rows && k + j < ke) ? X[(int64_t)r * p.K + k + j] : (uint16_t)0; }
                std::memcpy(&reg[i], e, 16);
            }
        }
    };
    auto store = [&](u32x4 (*S)[KP / 8], const u32x4* reg, int count) {
        for (int i = 0; i < count; ++i) {
            const int v = tid + i * NT_THREADS;
            S[v / (KB / 8)][v % (KB / 8)] = reg[i];
        }
    };

    f32x4 acc[TM][TN];
    for (int i = 0; i < TM; ++i) {
        for (int j = 0; j < TN; ++j) { acc[i][j] = f32x4{ 0.f, 0.f, 0.f, 0.f }; }
    }
    const int kq = lane / 16;  // the lane's 8-wide K vector within a 32-deep MFMA step
    auto compute = [&](int buf) {
        for (int h = 0; h < KB / 32; ++h) {
            bf16x8 fa[TM], fb[TN];
            for (int i = 0; i < TM; ++i) { fa[i] = __builtin_bit_cast(bf16x8, As[buf][wm0 + i * 16 + lane % 16][h * 4 + kq]); }
            for (int j = 0; j < TN; ++j) { fb[j] = __builtin_bit_cast(bf16x8, Bs[buf][wn0 + j * 16 + lane % 16][h * 4 + kq]); }
            for (int i = 0; i < TM; ++i) {
                for (int j = 0; j < TN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
                }
            }
        }
    };
    // Two register sets, loaded two steps ahead of the MFMAs: step s is loaded into set s % 2 during
    // step s - 2 and stored into LDS buffer s % 2 at the end of step s - 1 (the loop is unrolled by
    // two so the sets are named statically).
    u32x4 ra1[AV], rb1[BV];
    load(rs_a, p.A, p.M, m0, kb, ra, AV);
    load(rs_b, p.B, p.N, n0, kb, rb, BV);
    if (nk > 1) {
        load(rs_a, p.A, p.M, m0, kb + KB, ra1, AV);
        load(rs_b, p.B, p.N, n0, kb + KB, rb1, BV);
    }
    store(As[0], ra, AV);
    store(Bs[0], rb, BV);
    __syncthreads();
    for (int kt = 0; kt < nk; kt += 2) {
        if (kt + 2 < nk) {
            load(rs_a, p.A, p.M, m0, kb + (kt + 2) * KB, ra, AV);
            load(rs_b, p.B, p.N, n0, kb + (kt + 2) * KB, rb, BV);
        }
        compute(0);
        if (kt + 1 < nk) {
            store(As[1], ra1, AV);
            store(Bs[1], rb1, BV);
        }
        __syncthreads();
        if (kt + 1 >= nk) { break; }
        if (kt + 3 < nk) {
            load(rs_a, p.A, p.M, m0, kb + (kt + 3) * KB, ra1, AV);
            load(rs_b, p.B, p.N, n0, kb + (kt + 3) * KB, rb1, BV);
        }
        compute(1);
        if (kt + 2 < nk) {
            store(As[0], ra, AV);
            store(Bs[0], rb, BV);
        }
        __syncthreads();
    }

    // D fragment: lane holds column lane % 16, rows 4 (lane / 16) + e
    for (int i = 0; i < TM; ++i) {
        for (int j = 0; j < TN; ++j) {
            for (int e = 0; e < 4; ++e) { Cs[wm0 + i * 16 + (lane / 16) * 4 + e][wn0 + j * 16 + lane % 16] = acc[i][j][e]; }
        }
    }
    __syncthreads();

    // FP32 tile rows to X [M][N] (8-column chunks)
    auto write_rows = [&](float* __restrict__ X) {
        for (int c = tid; c < BM * CH; c += NT_THREADS) {
            const int r = c / CH, cc = (c % CH) * 8, m = m0 + r, n = n0 + cc;
            if (m >= p.M || n >= p.N) { continue; }
            float* dst = X + (int64_t)m * p.N + n;
            if (p.vec_out && n + 8 <= p.N) {
                std::memcpy(dst, &Cs[r][cc], 16);
                std::memcpy(dst + 4, &Cs[r][cc + 4], 16);
            } else {
                for (int j = 0; j < 8 && n + j < p.N; ++j) { dst[j] = Cs[r][cc + j]; }
            }
        }
    };

    if (EPI == EPI_F32) {
        if (gridDim.y == 1) {
            write_rows(p.c);
            return;
        }
        write_rows(p.ws + (int64_t)split * p.M * p.N);
        if (!p.counters) { return; }  // the partials are summed by sum_splits_kernel
        // every thread's partial stores have completed at the barrier; one thread then makes them
        // visible device-wide (release), counts the arrival and, in the last block, acquires the
        // other blocks' partials for the whole block
        __syncthreads();
        if (tid == 0) {
            __threadfence();
            last_block = atomicAdd(&p.counters[t], 1u) == gridDim.y - 1;
            if (last_block) { __threadfence(); }
        }
        __syncthreads();
        if (!last_block) { return; }
        const int64_t slice = (int64_t)p.M * p.N;
        for (int c = tid; c < BM * CH; c += NT_THREADS) {
            const int r = c / CH, cc = (c % CH) * 8, m = m0 + r, n = n0 + cc;
            if (m >= p.M || n >= p.N) { continue; }
            const int64_t at = (int64_t)m * p.N + n;
            if (p.vec_out && n + 8 <= p.N) {
                f32x4 s0 = f32x4{ 0.f, 0.f, 0.f, 0.f }, s1 = s0;
                for (unsigned q = 0; q < gridDim.y; ++q) {
                    f32x4 v0, v1;
                    std::memcpy(&v0, p.ws + q * slice + at, 16);
                    std::memcpy(&v1, p.ws + q * slice + at + 4, 16);
                    s0 += v0;
                    s1 += v1;
                }
                std::memcpy(p.c + at, &s0, 16);
                std::memcpy(p.c + at + 4, &s1, 16);
            } else {
                for (int j = 0; j < 8 && n + j < p.N; ++j) {
                    float s = 0.f;
                    for (unsigned q = 0; q < gridDim.y; ++q) { s += p.ws[q * slice + at + j]; }
                    p.c[at + j] = s;
                }
            }
        }
        if (tid == 0) { atomicExch(&p.counters[t], 0u); }
        return;
    }

    // EPI_FWD: z = acc + bias, a = bf16(relu ? max(z, 0) : z).  EPI_BACK: d = mask > 0 ? acc : 0,
    // rounded to bf16.  Pass 1 writes rows and leaves z (or d) in the staging tile for pass 2,
    // which writes the transposed copy from its columns.
    for (int c = tid; c < BM * CH; c += NT_THREADS) {
        const int r = c / CH, cc = (c % CH) * 8, m = m0 + r, n = n0 + cc;
        if (m >= p.M || n >= p.N) { continue; }
        const int64_t at = (int64_t)m * p.N + n;
        float v[8];
        std::memcpy(v, &Cs[r][cc], 16);
        std::memcpy(v + 4, &Cs[r][cc + 4], 16);
        uint16_t o[8];
        if (p.vec_out && n + 8 <= p.N) {  // whole chunk: vector loads and stores
            if (EPI == EPI_FWD) {
                float b[8];
                std::memcpy(b, p.bias + n, 16);
                std::memcpy(b + 4, p.bias + n + 4, 16);
                for (int j = 0; j < 8; ++j) { v[j] += b[j]; }
                if (p.c) {
                    std::memcpy(p.c + at, v, 16);
                    std::memcpy(p.c + at + 4, v + 4, 16);
                }
                // ReLU in place: the transposed pass below applies it again, to the same effect
                if (p.relu) {
                    for (int j = 0; j < 8; ++j) { v[j] = fmaxf(v[j], 0.f); }
                }
            } else {
                uint16_t mk[8];
                std::memcpy(mk, p.mask + at, 16);
                for (int j = 0; j < 8; ++j) { v[j] = bf16_to_f32(mk[j]) > 0.f ? v[j] : 0.f; }
            }
            if (p.out) {
                const u32x4 w = { pack_bf16x2(v[0], v[1]), pack_bf16x2(v[2], v[3]), pack_bf16x2(v[4], v[5]),
                    pack_bf16x2(v[6], v[7]) };
                *reinterpret_cast<u32x4*>(p.out + at) = w;
            }
        } else {
            const int nn = min(8, p.N - n);
            for (int j = 0; j < 8; ++j) {
                if (EPI == EPI_FWD) {
                    v[j] = j < nn ? v[j] + p.bias[n + j] : 0.f;
                    o[j] = f32_to_bf16(p.relu ? fmaxf(v[j], 0.f) : v[j]);
                } else {
                    v[j] = (j < nn && bf16_to_f32(p.mask[at + j]) > 0.f) ? v[j] : 0.f;
                    o[j] = f32_to_bf16(v[j]);
                }
            }
            #pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j >= nn) { continue; }
                if (EPI == EPI_FWD && p.c) { p.c[at + j] = v[j]; }
                if (p.out) { p.out[at + j] = o[j]; }
            }
        }
        if (p.out_t) {
            std::memcpy(&Cs[r][cc], v, 16);
            std::memcpy(&Cs[r][cc + 4], v + 4, 16);
        }
    }
    if (!p.out_t) { return; }
    __syncthreads();
    constexpr int RCH = BM / 8;  // 8-row chunks of a tile column
    for (int c = tid; c < BN * RCH; c += NT_THREADS) {
        const int col = c % BN, r8 = (c / BN) * 8, n = n0 + col, m = m0 + r8;
        if (n >= p.N || m >= p.M) { continue; }
        float y[8];
        for (int j = 0; j < 8; ++j) {
            const float x = Cs[r8 + j][col];
            y[j] = EPI == EPI_FWD && p.relu ? fmaxf(x, 0.f) : x;
        }
        const u32x4 w = { pack_bf16x2(y[0], y[1]), pack_bf16x2(y[2], y[3]), pack_bf16x2(y[4], y[5]),
            pack_bf16x2(y[6], y[7]) };
        uint16_t* dst = p.out_t + (int64_t)n * p.M + m;
        if (p.vec_out_t && m + 8 <= p.M) {
            *reinterpret_cast<u32x4*>(dst) = w;
        } else {
            #pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (m + j < p.M) { dst[j] = (uint16_t)(w[j / 2] >> (16 * (j % 2))); }
            }
        }
    }
}

// 4 x 4 transpose of 16-bit values within each quad of lanes: lane 4a + b passes row b of its
// quad's block (elements 0..3, two per word, low half first) and gets column b back.  Two DPP quad
// permutations: lanes b and b ^ 2 swap the 2 x 2 blocks each needs, then lanes b and b ^ 1 swap
// halves within them.  Every lane of the wave must be active.
__device__ inline u32x2 quad_transpose16(u32x2 w, int lane) {
    const int b = lane & 3;
    const uint32_t give = b < 2 ? w[1] : w[0];
    const uint32_t got = (uint32_t)__builtin_amdgcn_mov_dpp((int)give, 0x4E, 0xF, 0xF, false);  // quad_perm 2, 3, 0, 1
    const uint32_t A = b < 2 ? w[0] : got;  // rows b & 1 and (b & 1) + 2 of columns 2 (b / 2), + 1
    const uint32_t B = b < 2 ? got : w[1];
    const bool even = (b & 1) == 0;
    const uint32_t send = even ? ((A >> 16) | (B & 0xffff0000u)) : ((A & 0xffffu) | (B << 16));
    const uint32_t r = (uint32_t)__builtin_amdgcn_mov_dpp((int)send, 0xB1, 0xF, 0xF, false);  // quad_perm 1, 0, 3, 2
    return even ? u32x2{ (A & 0xffffu) | (r << 16), (B & 0xffffu) | (r & 0xffff0000u) }
                : u32x2{ (r & 0xffffu) | (A & 0xffff0000u), (r >> 16) | (B & 0xffff0000u) };
}

// K-major products with the weight slice resident in LDS (N a multiple of 128, K a multiple of 64
// or 256 up to 512: the network's forward and input-gradient products).  The tiles above stage
// both operands through LDS at every K step, and at 64 x 32 per wave they read about one 16-byte
// LDS vector per MFMA.  Here a block loads rows n0 .. n0 + 127 of B (all of K, 128 KB) into LDS
// once and keeps it: each of its 4 or 8 waves computes a 64 x 128 tile (32 accumulators) of a band
// of 64 rows per wave, reads its A fragments straight from global memory into a register ring
// (chunks of 32 K ahead), and its B fragments from LDS (one vector per 4 MFMAs); no barrier after
// the weights are in.  Blocks are persistent over bands: a block keeps one column slice and steps
// through the bands dealt to it, the slices of one band on the same XCD, so A is fetched from HBM
// once and shared in L2.  Each output element sums the same 32-deep chunks, in the same order and
// the same lane layout, as in gemm_nt_kernel; the MFMA takes B as its first operand, so a lane
// holds 4 consecutive columns of one row (16-byte FP32, 8-byte bf16 row stores; the transposed
// copy goes through a lane-quad transpose into 8-byte pieces of its rows).
constexpr int BR_BN = 128;  // columns of the resident weight slice
constexpr int BR_WROWS = 64;  // rows per wave and band
constexpr int BR_KMAX = 512;
constexpr int BR_PITCH = BR_KMAX / 8;  // LDS row pitch in 16-byte vectors (1024 bytes), slots swizzled

// WAVES 4: one wave per SIMD (512 registers), eight chunks of A ahead, B fragments read a chunk
// ahead.  WAVES 8: two waves per SIMD (256 registers each), two chunks ahead, each wave's waits
// covered by the other's MFMAs (a four-chunk ring spilled 20 bytes per lane in the forward and
// took 58.2 us against 56.6 us on 65536 x 512 x 512).  A band is WAVES x 64 rows.
template <int WAVES>
struct BresShape {
    static constexpr int band = WAVES * BR_WROWS;
    static constexpr int ahead = WAVES == 4 ? 8 : 2;  // 32-deep chunks of A in flight per wave
    static constexpr bool b_ahead = WAVES == 4;  // B fragments double-buffered
};

template <int EPI, int WAVES>
__global__ void __launch_bounds__(64 * WAVES, WAVES / 4) gemm_nt_bres_kernel(NtArgs p, int groups) {
    constexpr int BR_BAND = BresShape<WAVES>::band, BR_AHEAD = BresShape<WAVES>::ahead, THREADS = 64 * WAVES;
    __shared__ u32x4 Bs[BR_BN][BR_PITCH];
    // the wave index is uniform: read it as a scalar, so the row bases of the A loads built from it
    // stay in scalar registers (a VGPR-held base wrapped every load in a readfirstlane loop)
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64, kq = lane / 16;
    // block -> (column slice, band group): the host launches XCDS x slices x groups blocks; blocks
    // are dealt to the XCDs in turn, and the slices of a band group sit on one XCD
    const int slices = p.N / BR_BN;
    const int x = (int)blockIdx.x % XCDS, q = (int)blockIdx.x / XCDS;
    const int n0 = (q % slices) * BR_BN;
    const int gi = x * groups + q / slices, stride = XCDS * groups;
    const int bands = (p.M + BR_BAND - 1) / BR_BAND;
    if (gi >= bands) { return; }  // the whole block
    const int my_bands = (bands - gi + stride - 1) / stride;
    const int nc = p.K / 32;  // chunks per band (a multiple of BR_AHEAD)

    // A fragments of a chunk: lane holds row (i * 16 + lane % 16) of the wave's 64, K columns
    // chunk * 32 + kq * 8 .. + 8.  One descriptor over all of A (M K 2 < 2^31 - 2^20, so no byte
    // offset wraps): rows past M, and every row of a chunk past the block's last band (row base M),
    // lie past its records and load zeros, with no branch around the loads.
    u32x4 ra[BR_AHEAD][4];
    const __amdgpu_buffer_rsrc_t rs_a = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t*>(p.A), (short)0, p.M * p.K * 2, 0x00020000);
    const auto issue = [&](int it, int c, u32x4* r) {
        const int row0 = it < my_bands ? (gi + it * stride) * BR_BAND + wave * BR_WROWS : p.M;
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            r[i] = __builtin_amdgcn_raw_buffer_load_b128(rs_a, ((row0 + i * 16 + lane % 16) * p.K + c * 32 + kq * 8) * 2, 0, 0);
        }
    };
    // the first BR_AHEAD chunks (band iteration 0) are in flight while the weights load
    #pragma unroll
    for (int u = 0; u < BR_AHEAD; ++u) { issue(0, u, ra[u]); }

    // The weight slice in LDS: 16-byte vector q of row r at slot q ^ (r & 15) of the row, so the 16
    // consecutive rows one fragment read touches (same q) fall on 16 different slots of a 256-byte
    // span: distinct banks, with no padding.
    {  // vector v of the slice is bytes 16 v .. of its rows, row v / (K / 8)
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint16_t*>(p.B + (int64_t)n0 * p.K), (short)0, BR_BN * p.K * 2, 0x00020000);
        const int kv = p.K / 8, total = BR_BN * kv;
        constexpr int VB = 4096 / THREADS;  // vectors per thread per batch
        for (int v0 = 0; v0 < total; v0 += THREADS * VB) {
            u32x4 r[VB];
            #pragma unroll
            for (int u = 0; u < VB; ++u) {
                const int v = v0 + u * THREADS + tid;
                r[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, v < total ? v * 16 : 0x7ffffff0, 0, 0);
            }
            #pragma unroll
            for (int u = 0; u < VB; ++u) {
                const int v = v0 + u * THREADS + tid;
                if (v < total) { Bs[v / kv][(v % kv) ^ ((v / kv) & 15)] = r[u]; }
            }
        }
    }
    __syncthreads();

    f32x4 acc[4][8];
    #pragma unroll
    for (int i = 0; i < 4; ++i) {
        #pragma unroll
        for (int j = 0; j < 8; ++j) { acc[i][j] = f32x4{ 0.f, 0.f, 0.f, 0.f }; }
    }
    // B fragments of a chunk (the same for every band), from LDS: one chunk ahead when double-buffered
    constexpr bool B2 = BresShape<WAVES>::b_ahead;
    bf16x8 fb[B2 ? 2 : 1][8];
    const auto read_b = [&](int c, bf16x8* f) {
        #pragma unroll
        for (int j = 0; j < 8; ++j) { f[j] = __builtin_bit_cast(bf16x8, Bs[j * 16 + lane % 16][(c * 4 + kq) ^ (lane % 16)]); }
    };
    if constexpr (B2) { read_b(0, fb[0]); }
    int lit = 0, lc = BR_AHEAD;  // band iteration and chunk of the next group of loads
    if (lc == nc) {
        lc = 0;
        ++lit;
    }
    int it = 0, cc = 0;  // band iteration and chunk being computed
    while (it < my_bands) {
        #pragma unroll
        for (int u = 0; u < BR_AHEAD; ++u) {
            const int c = cc + u;
            if constexpr (B2) {
                read_b(c + 1 == nc ? 0 : c + 1, fb[(u + 1) % 2]);
            } else {
                read_b(c, fb[0]);
            }
            const int fcur = B2 ? u % 2 : 0;
            #pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bf16x8 fa = __builtin_bit_cast(bf16x8, ra[u][i]);
                #pragma unroll
                for (int j = 0; j < 8; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[fcur][j], fa, acc[i][j], 0, 0, 0);
                }
            }
            issue(lit, lc + u, ra[u]);
            // keep the refill here: hoisted below later chunks' MFMAs, the loads would go out in
            // bursts and leave the ring nearly empty in between
            __builtin_amdgcn_sched_barrier(0);
        }
        lc += BR_AHEAD;
        if (lc == nc) {
            lc = 0;
            ++lit;
        }
        cc += BR_AHEAD;
        if (cc < nc) { continue; }

        // The band is done: lane holds rows n = 4 kq + e of each 16 x 16 result, column m = lane % 16.
        // Every lane runs the epilogue (the transposed copy's lane exchange needs all of them);
        // rows past M only skip their loads and stores.
        const int row0 = (gi + it * stride) * BR_BAND + wave * BR_WROWS;
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = row0 + i * 16 + lane % 16;
            const bool in_rows = m < p.M;
            #pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int n = n0 + j * 16 + 4 * kq;
                const int64_t at = (int64_t)m * p.N + n;
                f32x4 v = acc[i][j];
                if (EPI == EPI_F32) {
                    if (in_rows) { *reinterpret_cast<f32x4*>(p.c + at) = v; }
                    continue;
                }
                if (EPI == EPI_FWD) {
                    v += *reinterpret_cast<const f32x4*>(p.bias + n);
                    if (p.c && in_rows) { *reinterpret_cast<f32x4*>(p.c + at) = v; }
                    if (p.relu) {
                        #pragma unroll
                        for (int e = 0; e < 4; ++e) { v[e] = fmaxf(v[e], 0.f); }
                    }
                } else {
                    const u32x2 mk = in_rows ? *reinterpret_cast<const u32x2*>(p.mask + at) : u32x2{ 0u, 0u };
                    #pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        v[e] = bf16_to_f32((uint16_t)(mk[e / 2] >> (16 * (e % 2)))) > 0.f ? v[e] : 0.f;
                    }
                }
                const u32x2 w = { pack_bf16x2(v[0], v[1]), pack_bf16x2(v[2], v[3]) };
                if (p.out && in_rows) { *reinterpret_cast<u32x2*>(p.out + at) = w; }
                if (p.out_t) {
                    // after the quad transpose, lane 4a + b holds column n + b (of the quad's 4 x 4
                    // block) for rows m4 .. m4 + 3: one 8-byte piece of a transposed row
                    const u32x2 t = quad_transpose16(w, lane);
                    const int m4 = row0 + i * 16 + (lane % 16 & ~3);
                    uint16_t* dst = p.out_t + (int64_t)(n + (lane & 3)) * p.M + m4;
                    if (p.vec_out_t && m4 + 4 <= p.M) {
                        *reinterpret_cast<u32x2*>(dst) = t;
                    } else {
                        #pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            if (m4 + e < p.M) { dst[e] = (uint16_t)(t[e / 2] >> (16 * (e % 2))); }
                        }
                    }
                }
            }
        }
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            #pragma unroll
            for (int j = 0; j < 8; ++j) { acc[i][j] = f32x4{ 0.f, 0.f, 0.f, 0.f }; }
        }
        cc = 0;
        ++it;
    }
}

// Tile shapes of the K-major kernel (block tile and its waves); which one runs a product of a given
// shape is timed once (nt_choose).  The kernel's register budget follows the waves per SIMD its
// LDS allows (4 for the 128 x 128 and 64 x 64 tiles, 2 for the others).  The tile does not change
// a result: every output element sums its K range in the same order whatever the tile.
struct NtTile {
    int bm, bn, wgm, wgn;
};
constexpr NtTile NT_TILES[] = { { 256, 128, 4, 2 }, { 128, 128, 2, 4 }, { 128, 64, 2, 2 }, { 64, 128, 2, 2 },
    { 64, 64, 2, 2 } };
constexpr int NT_NUM_TILES = 5;

int nt_tiles_of(int tile, int M, int N) {
    return ((M + NT_TILES[tile].bm - 1) / NT_TILES[tile].bm) * ((N + NT_TILES[tile].bn - 1) / NT_TILES[tile].bn);
}

// Splits over K of a K-major product: fixed by its shape alone (never by timing), so a product's
// summation order, and its result, is the same in every run.  Products with few 64 x 64 output
// tiles and a long K (the weight gradients, K = batch rows) are split into parts of at least
// SPLIT_MIN_K, up to NT_SPLIT_TARGET_BLOCKS tiles x splits.  The 512 x 512 hidden layer's
// weight gradient gets 8 splits at 4096 rows (K / SPLIT_MIN_K; 16 splits of 256 rows took 16.6 us
// against 13.4, their partials doubling the summed bytes) and 16 from 8192 rows on (at 65536 rows
// 63 us on the 128 x 64 tile, against 86 us with 8 splits).
constexpr int NT_SPLIT_TARGET_BLOCKS = 1024;
int nt_splits(int M, int N, int K) {
    const int64_t tiles = (int64_t)((M + 63) / 64) * ((N + 63) / 64);
    if (tiles >= NT_SPLIT_TARGET_BLOCKS || K < 2 * SPLIT_MIN_K) { return 1; }
    int s = (int)std::min<int64_t>(MAX_SPLITS, (NT_SPLIT_TARGET_BLOCKS + tiles - 1) / tiles);
    return std::max(1, std::min(s, K / SPLIT_MIN_K));
}

template <int EPI>
void launch_nt_tile(int tile, const NtArgs& a, int splits, hipStream_t st) {
    const dim3 grid((unsigned)nt_tiles_of(tile, a.M, a.N), (unsigned)splits);
    const NtTile& t = NT_TILES[tile];
    const dim3 block((unsigned)(64 * t.wgm * t.wgn));
    if (!(a.vec_a && a.vec_b)) {  // element loads (ragged K or unaligned operands): one tile shape
        hipLaunchKernelGGL((gemm_nt_kernel<64, 64, 2, 2, 2, EPI, false>), grid, block, 0, st, a);
        return;
    }
    switch (tile) {
    case 0: hipLaunchKernelGGL((gemm_nt_kernel<256, 128, 4, 2, 2, EPI, true>), grid, block, 0, st, a); break;
    case 1: hipLaunchKernelGGL((gemm_nt_kernel<128, 128, 2, 4, 4, EPI, true>), grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((gemm_nt_kernel<128, 64, 2, 2, 2, EPI, true>), grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((gemm_nt_kernel<64, 128, 2, 2, 2, EPI, true>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((gemm_nt_kernel<64, 64, 2, 2, 4, EPI, true>), grid, block, 0, st, a); break;
    }
    if (EPI == EPI_F32 && splits > 1 && !a.counters) {
        const int64_t n = (int64_t)a.M * a.N;
        hipLaunchKernelGGL(sum_splits_kernel, dim3(ew_blocks(n)), dim3(EW_BLOCK), 0, st, a.ws, a.c, splits, n);
    }
}

// Choice before any timing: the largest tile whose blocks fill the chip twice over (2 per CU), else
// the smallest; split products summed in the kernel.
int nt_default_choice(int M, int N, int splits) {
    for (int tile = 0; tile < NT_NUM_TILES; ++tile) {
        if ((int64_t)nt_tiles_of(tile, M, N) * splits >= 512) { return tile; }
    }
    return NT_NUM_TILES - 1;
}

// How a K-major product runs: a tile shape (NT_TILES index) and, for a product split over K,
// where its partials are summed: in the kernel by each tile's last block (choice < NT_NUM_TILES),
// or by sum_splits_kernel after it (choice - NT_NUM_TILES).  Both sum the partials in split order,
// so every choice gives the same result.  Choice NT_BRES is gemm_nt_bres_kernel (unsplit products
// of the shapes it takes); it reports tile index NT_NUM_TILES.
struct NtChoice {
    int tile;
    bool summed_in_kernel;
};

constexpr int NT_BRES = 2 * NT_NUM_TILES;  // 4 waves; NT_BRES + 1: 8 waves

NtChoice nt_decode(int choice) {
    if (choice >= NT_BRES) { return { NT_NUM_TILES + choice - NT_BRES, true }; }
    return { choice % NT_NUM_TILES, choice < NT_NUM_TILES };
}

bool aligned16(const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; }

// Products gemm_nt_bres_kernel takes: unsplit, vector operands and outputs, N a multiple of the
// 128-column slice, K a multiple of the chunk group (256 with 4 waves, 64 with 8) up to 512, and A
// within 32-bit byte offsets of one buffer descriptor.
template <int EPI>
bool nt_bres_ok(const NtArgs& a, int splits, int waves) {
    const int group = 32 * (waves == 4 ? BresShape<4>::ahead : BresShape<8>::ahead);
    return splits == 1 && a.vec_a && a.vec_b && a.vec_out && a.N % BR_BN == 0 && a.K % group == 0 && a.K <= BR_KMAX
           && (int64_t)a.M * a.K * 2 <= (int64_t)INT32_MAX - (1 << 20) && (EPI != EPI_FWD || aligned16(a.bias));
}

int device_cus(int device) {
    static std::atomic<int> cus[64];
    if (device < 0 || device >= 64) { return 256; }
    int n = cus[device].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) { n = 256; }
        cus[device].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Persistent grid of gemm_nt_bres_kernel: XCDS x slices x groups blocks, about one per CU, and no
// more band groups than bands.
template <int EPI, int WAVES>
void launch_nt_bres(int device, const NtArgs& a, hipStream_t st) {
    const int slices = a.N / BR_BN;
    const int bands = (a.M + BresShape<WAVES>::band - 1) / BresShape<WAVES>::band;
    const int groups = std::max(1, std::min(device_cus(device) / (XCDS * slices), (bands + XCDS - 1) / XCDS));
    hipLaunchKernelGGL((gemm_nt_bres_kernel<EPI, WAVES>), dim3((unsigned)(XCDS * slices * groups)), dim3(64 * WAVES), 0,
        st, a, groups);
}

template <int EPI>
void launch_nt_choice(int device, int choice, NtArgs a, int splits, hipStream_t st) {
    if (choice >= NT_BRES) {
        const int waves = choice == NT_BRES ? 4 : 8;
        if (nt_bres_ok<EPI>(a, splits, waves)) {
            if (waves == 4) {
                launch_nt_bres<EPI, 4>(device, a, st);
            } else {
                launch_nt_bres<EPI, 8>(device, a, st);
            }
            return;
        }
        choice = nt_default_choice(a.M, a.N, splits);
    }
    const NtChoice c = nt_decode(choice);
    if (!c.summed_in_kernel) { a.counters = nullptr; }
    launch_nt_tile<EPI>(c.tile, a, splits, st);
}

// Autotuner: the first product of each (device, epilogue, outputs, M, N, K, splits) times every
// choice (one launch to warm, then HIP events around three) on the caller's stream with the
// product's own arguments, and keeps the fastest for the rest of the process.  Off while the stream
// is being captured into a graph and for callers that must not create events (may_tune false):
// those take a choice already tuned, else the default.
std::mutex nt_tune_mutex;
std::map<std::array<int64_t, 7>, int> nt_tuned;

template <int EPI>
int nt_choose(int device, const NtArgs& a, int splits, hipStream_t st, bool may_tune) {
    const int64_t outputs = (a.c ? 1 : 0) | (a.out ? 2 : 0) | (a.out_t ? 4 : 0);
    const std::array<int64_t, 7> key = { device, EPI, outputs, a.M, a.N, a.K, splits };
    if (!(a.vec_a && a.vec_b)) { return NT_NUM_TILES - 1; }  // element loads: one tile shape
    // PSIM_GEMM_CHOICE forces a choice (tests run every choice and compare the results)
    const bool bres4 = nt_bres_ok<EPI>(a, splits, 4), bres8 = nt_bres_ok<EPI>(a, splits, 8);
    if (const char* forced = std::getenv("PSIM_GEMM_CHOICE")) {
        const int c = std::atoi(forced);
        if ((c >= 0 && c < (splits > 1 ? 2 * NT_NUM_TILES : NT_NUM_TILES)) || (c == NT_BRES && bres4)
            || (c == NT_BRES + 1 && bres8)) {
            return c;
        }
    }
    std::lock_guard<std::mutex> lock(nt_tune_mutex);
    const auto it = nt_tuned.find(key);
    if (it != nt_tuned.end()) { return it->second; }
    const int fallback = nt_default_choice(a.M, a.N, splits);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (!may_tune || hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
        return fallback;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) { (void)hipEventDestroy(e0); }
        return fallback;
    }
    std::vector<int> candidates;
    for (int c = 0; c < (splits > 1 ? 2 * NT_NUM_TILES : NT_NUM_TILES); ++c) { candidates.push_back(c); }
    if (bres4) { candidates.push_back(NT_BRES); }
    if (bres8) { candidates.push_back(NT_BRES + 1); }
    int best = fallback;
    float best_ms = 0.f;
    bool ok = true;
    for (size_t ci = 0; ci < candidates.size() && ok; ++ci) {
        const int choice = candidates[ci];
        launch_nt_choice<EPI>(device, choice, a, splits, st);
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 3; ++r) { launch_nt_choice<EPI>(device, choice, a, splits, st); }
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        ok = hipGetLastError() == hipSuccess && hipEventSynchronize(e1) == hipSuccess
             && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok && (ci == 0 || ms < best_ms)) {
            best = choice;
            best_ms = ms;
        }
        if (ok && std::getenv("PSIM_TUNE_LOG")) {
            const NtChoice c = nt_decode(choice);
            if (choice >= NT_BRES) {
                const int waves = choice == NT_BRES ? 4 : 8;
                std::fprintf(stderr, "psim tune: epilogue %d outputs %d %d x %d x %d, weights resident, %d x 128 bands (%d waves): %.2f us\n",
                    EPI, (int)outputs, a.M, a.N, a.K, waves * BR_WROWS, waves, ms * 1e3f / 3.f);
            } else {
                std::fprintf(stderr, "psim tune: epilogue %d outputs %d %d x %d x %d, %d splits%s, tile %d x %d (%d waves): %.2f us\n",
                    EPI, (int)outputs, a.M, a.N, a.K, splits, splits > 1 ? (c.summed_in_kernel ? " summed in kernel" : " summed after") : "",
                    NT_TILES[c.tile].bm, NT_TILES[c.tile].bn, NT_TILES[c.tile].wgm * NT_TILES[c.tile].wgn, ms * 1e3f / 3.f);
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ok) { nt_tuned[key] = best; }
    return ok ? best : fallback;
}

// Fills the shape-derived fields of a K-major product and launches it as tuned.  Split products
// need ws (splits x M x N floats) and counters (NT_COUNTERS, zero); without them, or when the tiles
// outnumber the counters, the product runs unsplit.  Returns the tile index used.
template <int EPI>
int launch_nt(int device, NtArgs a, hipStream_t st, bool may_tune, size_t ws_elems) {
    // buffer loads address a block's rows with 32-bit byte offsets: rows x K x 2 < 2^31 for the
    // tallest tile (256 rows), else element loads
    const bool fits = (int64_t)256 * a.K * 2 < (int64_t)INT32_MAX;
    a.vec_a = fits && aligned16(a.A) && a.K % 8 == 0;
    a.vec_b = fits && aligned16(a.B) && a.K % 8 == 0;
    a.vec_out = a.N % 8 == 0 && aligned16(a.c) && aligned16(a.ws) && aligned16(a.out) && aligned16(a.mask);
    a.vec_out_t = a.M % 8 == 0 && aligned16(a.out_t);
    int splits = EPI == EPI_F32 && a.ws && a.counters ? nt_splits(a.M, a.N, a.K) : 1;
    int k_split = (a.K + KB - 1) / KB * KB;
    if (splits > 1) {
        k_split = ((a.K + splits - 1) / splits + KB - 1) / KB * KB;
        splits = (a.K + k_split - 1) / k_split;
        if ((size_t)splits * (size_t)a.M * (size_t)a.N > ws_elems || nt_tiles_of(NT_NUM_TILES - 1, a.M, a.N) > NT_COUNTERS) {
            splits = 1;
            k_split = (a.K + KB - 1) / KB * KB;
        }
    }
    a.k_split = k_split;
    int choice = nt_choose<EPI>(device, a, splits, st, may_tune);
    if (choice >= NT_BRES && !nt_bres_ok<EPI>(a, splits, choice == NT_BRES ? 4 : 8)) {
        choice = nt_default_choice(a.M, a.N, splits);
    }
    launch_nt_choice<EPI>(device, choice, a, splits, st);
    return nt_decode(choice).tile;
}

// C = op(A) x op(B) in FP32 (operands as stored: A [K][M] when ta, B [N][K] when tb).  The
// K-major layout (A [M][K], B [N][K]) runs on the network's kernel, split over K when ws and
// counters are given and the shape asks for it; the others on gemm_bf16_kernel.
void launch_gemm(int device, bool ta, bool tb, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
    hipStream_t st, float* ws, size_t ws_elems, unsigned* counters, bool may_tune) {
    if (!ta && tb) {
        NtArgs a{};
        a.A = A;
        a.B = B;
        a.M = M;
        a.N = N;
        a.K = K;
        a.c = C;
        a.ws = ws;
        a.counters = counters;
        launch_nt<EPI_F32>(device, a, st, may_tune, ws_elems);
    } else if (ta && tb) {
        launch_gemm_t<true, true>(A, B, C, M, N, K, st, ws, ws_elems);
    } else if (ta) {
        launch_gemm_t<true, false>(A, B, C, M, N, K, st, ws, ws_elems);
    } else {
        launch_gemm_t<false, false>(A, B, C, M, N, K, st, ws, ws_elems);
    }
}

// a = bf16(x) [rows][n] and, when a_t is given, its transposed copy a_t [n][rows].
__global__ void input_kernel(const float* __restrict__ x, uint16_t* __restrict__ a, uint16_t* __restrict__ a_t,
    int64_t rows, int n) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= rows * n) { return; }
    const uint16_t v = f32_to_bf16(x[i]);
    a[i] = v;
    if (a_t) { a_t[(i % n) * rows + i / n] = v; }
}

// Output-layer gradient of the mean squared error on column 0: dz[r][0] = 2 (y - t) / rows, other
// columns 0 (bf16, row-major and transposed), and the loss terms (y - t)^2 / rows per row.
__global__ void loss_grad_kernel(const float* __restrict__ z, const float* __restrict__ t, uint16_t* __restrict__ dz,
    uint16_t* __restrict__ dz_t, float* __restrict__ loss_terms, int rows, int cols) {
    const int r = blockIdx.x * EW_BLOCK + threadIdx.x;
    if (r >= rows) { return; }
    const float d = z[(int64_t)r * cols] - t[r];
    loss_terms[r] = d * d / (float)rows;
    for (int c = 0; c < cols; ++c) {
        const uint16_t v = f32_to_bf16(c == 0 ? 2.f * d / (float)rows : 0.f);
        dz[(int64_t)r * cols + c] = v;
        dz_t[(int64_t)c * rows + r] = v;
    }
}

// db[c] = sum over rows of dz_t[c][r] (one block per column reading its contiguous row of the
// transposed gradient; FP32, thread-strided partial sums then a fixed tree).
__global__ void bias_grad_kernel(const uint16_t* __restrict__ dz_t, float* __restrict__ db, int rows) {
    __shared__ float part[EW_BLOCK];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += EW_BLOCK) { s += bf16_to_f32(dz_t[(int64_t)c * rows + r]); }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = EW_BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) { part[threadIdx.x] += part[threadIdx.x + off]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { db[c] = part[0]; }
}

// Where each weight matrix lies in the flat parameter vector, for kernels that write its bf16
// copies: [in][out] at w_off in pb, [out][in] at w_off in pbT.
constexpr int MAX_WEIGHTS = 17;
struct LayerTable {
    int count;
    int64_t w_off[MAX_WEIGHTS];
    int in[MAX_WEIGHTS], out[MAX_WEIGHTS];
};

__device__ inline void store_bf16_copies(const LayerTable& lt, int64_t i, uint16_t b, uint16_t* __restrict__ pb,
    uint16_t* __restrict__ pbT) {
    pb[i] = b;
    for (int l = 0; l < lt.count; ++l) {
        const int64_t o = i - lt.w_off[l];
        if (o >= 0 && o < (int64_t)lt.in[l] * lt.out[l]) {
            pbT[lt.w_off[l] + (o % lt.out[l]) * lt.in[l] + o / lt.out[l]] = b;
            return;
        }
    }
}

// Adam on the FP32 master parameters, then their bf16 copies (masked entries stay 0).  The step's
// learning rate and bias corrections (lr, 1 - b1^t, 1 - b2^t) are read from device memory, so a
// captured step replays with the values copied in before each launch.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
    float* __restrict__ m2, const float* __restrict__ mask, uint16_t* __restrict__ pb, uint16_t* __restrict__ pbT,
    LayerTable lt, int64_t n, const float* __restrict__ step_consts, float b1, float b2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) { return; }
    const float lr = step_consts[0], c1 = step_consts[1], c2 = step_consts[2];
    const float gi = g[i] * (mask ? mask[i] : 1.f);
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    store_bf16_copies(lt, i, f32_to_bf16(v), pb, pbT);
}

// bf16 copies of the FP32 parameters.
__global__ void params_bf16_kernel(const float* __restrict__ p, uint16_t* __restrict__ pb, uint16_t* __restrict__ pbT,
    LayerTable lt, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i < n) { store_bf16_copies(lt, i, f32_to_bf16(p[i]), pb, pbT); }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i < n) { y[i] = f32_to_bf16(x[i]); }
}

// y[r] = z[r][0] for a [rows][cols] matrix.
__global__ void column0_kernel(const float* __restrict__ z, float* __restrict__ y, int rows, int cols) {
    const int r = blockIdx.x * EW_BLOCK + threadIdx.x;
    if (r < rows) { y[r] = z[(int64_t)r * cols]; }
}

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;

}  // namespace

// Device memory of psim_mlp_forward_noalloc (input staging, the bf16 activations, one FP32
// pre-activation buffer and the output column), carved from this array instead of allocated.
// Every GPU the library runs on holds its own copy from the moment the library's code is loaded
// there until the process ends.  In host code `scratch` names a host-side stand-in, not device
// memory: host code uses only the per-device address that hipGetSymbolAddress returns for it.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct psim_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int n_in = 0, hidden = 0, layers = 0;  // layers = hidden layers; weight matrices = layers + 1
    std::vector<int> dims;  // n_in, hidden..., PAD_OUT
    std::vector<int64_t> w_off, b_off;  // offsets in the flat parameter vector
    int64_t n_params = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr, *mask = nullptr;
    uint16_t* pb = nullptr;  // bf16 copy, layout of p (weights [in][out]; biases unused)
    uint16_t* pbT = nullptr;  // bf16 weights transposed, [out][in] per layer (K-major for the forward)
    float* ws = nullptr;  // partial products of weight gradients split over K
    size_t ws_elems = 0;
    unsigned* counters = nullptr;  // per-tile arrival counters of split products (zero between products)
    int64_t step = 0;
    // batch buffers (cap rows)
    int cap = 0;
    std::vector<uint16_t*> a;  // bf16 input of every layer: a[0] = the network's input, a[l] = output of layer l - 1
    float* zout = nullptr;  // output layer's FP32 pre-activation [cap][PAD_OUT]
    float* xin = nullptr;
    float* t = nullptr;
    float* loss = nullptr;
    // training buffers (train_cap rows): transposed copies, [features][rows]
    int train_cap = 0;
    std::vector<uint16_t*> aT;  // aT[l]: a[l] transposed
    uint16_t *dz = nullptr, *dz2 = nullptr;  // bf16 gradient of the current layer's pre-activation [rows][features]
    uint16_t *dzT = nullptr, *dz2T = nullptr;  // the same transposed
    // device address of this GPU's `scratch` array (psim_mlp_forward_noalloc), looked up at create
    unsigned char* scratch = nullptr;
    // training steps replayed from a HIP graph: the step's kernels (forward, backward, Adam) captured
    // once per batch size after one eager step at that size has tuned the products' tiles
    bool use_graph = true;
    float* step_consts = nullptr;  // device: lr, 1 - b1^t, 1 - b2^t of the current step
    float* h_step_consts = nullptr;  // pinned host staging of the same
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_rows = 0;  // batch size of the captured step (0: none)
    int64_t graph_gen = 0;  // alloc_gen when it was captured
    int warm_rows = 0;  // batch size of the last eager step
    int64_t alloc_gen = 0;  // bumped when batch or training buffers move
};

namespace {

int mlp_fail(psim_mlp* h, int code, const std::string& msg) {
    if (h) { h->err = msg; }
    return code;
}

#define MLP_CHECK(h, expr)                                                                                   \
    do {                                                                                                     \
        const hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) { return mlp_fail((h), PSIM_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); } \
    } while (0)

template <typename T>
T* mlp_alloc(size_t n) {
    T* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) { return nullptr; }
    return p;
}

void free_train(psim_mlp* h) {
    for (uint16_t* q : h->aT) { (void)hipFree(q); }
    h->aT.clear();
    uint16_t* rest[] = { h->dz, h->dz2, h->dzT, h->dz2T };
    for (uint16_t* q : rest) {
        if (q) { (void)hipFree(q); }
    }
    h->dz = h->dz2 = h->dzT = h->dz2T = nullptr;
    h->train_cap = 0;
}

void free_batch(psim_mlp* h) {
    for (uint16_t* q : h->a) { (void)hipFree(q); }
    h->a.clear();
    float* rest[] = { h->zout, h->xin, h->t, h->loss };
    for (float* q : rest) {
        if (q) { (void)hipFree(q); }
    }
    h->zout = h->xin = h->t = h->loss = nullptr;
    h->cap = 0;
}

int widest_dim(const psim_mlp* h) {
    int widest = 0;
    for (int d : h->dims) { widest = std::max(widest, d); }
    return widest;
}

int ensure_batch(psim_mlp* h, int rows) {
    if (rows <= h->cap) { return PSIM_OK; }
    free_batch(h);
    const int cap = ((rows + BATCH_ROUND - 1) / BATCH_ROUND) * BATCH_ROUND;
    const int L = (int)h->dims.size() - 1;
    for (int l = 0; l < L; ++l) { h->a.push_back(mlp_alloc<uint16_t>((size_t)cap * h->dims[l])); }
    h->zout = mlp_alloc<float>((size_t)cap * PAD_OUT);
    h->xin = mlp_alloc<float>((size_t)cap * h->n_in);
    h->t = mlp_alloc<float>((size_t)cap);
    h->loss = mlp_alloc<float>((size_t)cap);
    bool ok = h->zout && h->xin && h->t && h->loss;
    for (uint16_t* q : h->a) { ok = ok && q; }
    if (!ok) {
        free_batch(h);
        return mlp_fail(h, PSIM_E_DEVICE, "hipMalloc failed for the surrogate batch buffers");
    }
    h->cap = cap;
    h->alloc_gen++;
    return PSIM_OK;
}

int ensure_train(psim_mlp* h, int rows) {
    int rc;
    if ((rc = ensure_batch(h, rows)) != 0) { return rc; }
    if (rows <= h->train_cap) { return PSIM_OK; }
    free_train(h);
    const int cap = ((rows + BATCH_ROUND - 1) / BATCH_ROUND) * BATCH_ROUND;
    const int L = (int)h->dims.size() - 1;
    const size_t wide = (size_t)cap * widest_dim(h);
    for (int l = 0; l < L; ++l) { h->aT.push_back(mlp_alloc<uint16_t>((size_t)cap * h->dims[l])); }
    h->dz = mlp_alloc<uint16_t>(wide);
    h->dz2 = mlp_alloc<uint16_t>(wide);
    h->dzT = mlp_alloc<uint16_t>(wide);
    h->dz2T = mlp_alloc<uint16_t>(wide);
    bool ok = h->dz && h->dz2 && h->dzT && h->dz2T;
    for (uint16_t* q : h->aT) { ok = ok && q; }
    if (!ok) {
        free_train(h);
        return mlp_fail(h, PSIM_E_DEVICE, "hipMalloc failed for the surrogate training buffers");
    }
    h->train_cap = cap;
    h->alloc_gen++;
    return PSIM_OK;
}

LayerTable layer_table(const psim_mlp* h) {
    LayerTable lt{};
    lt.count = (int)h->dims.size() - 1;
    for (int l = 0; l < lt.count; ++l) {
        lt.w_off[l] = h->w_off[l];
        lt.in[l] = h->dims[l];
        lt.out[l] = h->dims[l + 1];
    }
    return lt;
}

// Forward pass over `rows` inputs at xin: layer l reads a[l] (a[0] is written here from xin) and
// writes its bf16 activation into a[l + 1]; the output layer writes its FP32 pre-activation to zout.
// With aT (training) every layer's input is also written transposed, aT[l], for the weight gradients.
int forward_pass(psim_mlp* h, int rows, const float* xin, uint16_t* const* a, uint16_t* const* aT, float* zout,
    bool may_tune) {
    const int L = (int)h->dims.size() - 1;
    hipLaunchKernelGGL(input_kernel, dim3(ew_blocks((int64_t)rows * h->n_in)), dim3(EW_BLOCK), 0, h->stream, xin, a[0],
        aT ? aT[0] : nullptr, (int64_t)rows, h->n_in);
    MLP_CHECK(h, hipGetLastError());
    for (int l = 0; l < L; ++l) {
        const bool last = l + 1 == L;
        NtArgs f{};
        f.A = a[l];
        f.B = h->pbT + h->w_off[l];  // W^T stored [out][in]: both operands K-major
        f.M = rows;
        f.N = h->dims[l + 1];
        f.K = h->dims[l];
        f.c = last ? zout : nullptr;
        f.bias = h->p + h->b_off[l];
        f.relu = last ? 0 : 1;
        f.out = last ? nullptr : a[l + 1];
        f.out_t = (aT && !last) ? aT[l + 1] : nullptr;
        launch_nt<EPI_FWD>(h->device, f, h->stream, may_tune, 0);
        MLP_CHECK(h, hipGetLastError());
    }
    return PSIM_OK;
}

int forward(psim_mlp* h, int rows, bool train) {
    return forward_pass(h, rows, h->xin, h->a.data(), train ? h->aT.data() : nullptr, h->zout, true);
}

// Loss and gradients of every parameter (into h->g) for targets in h->t, after forward(train).
// Layer l: dW_l = a_l^T dz_l on the transposed copies (split over the batch rows), db_l = row sums of
// dz_l^T, and dz_{l-1} = (dz_l W_l^T) where a_l > 0: the ReLU of layer l - 1 passed a positive
// pre-activation exactly where its bf16 output a_l is positive (a pre-activation below 2^-134
// would round to a zero activation and drop its gradient).
int backward(psim_mlp* h, int rows) {
    const int L = (int)h->dims.size() - 1;
    hipLaunchKernelGGL(loss_grad_kernel, dim3(ew_blocks(rows)), dim3(EW_BLOCK), 0, h->stream, h->zout, h->t, h->dz,
        h->dzT, h->loss, rows, h->dims[L]);
    MLP_CHECK(h, hipGetLastError());
    uint16_t *dz = h->dz, *dzT = h->dzT, *sp = h->dz2, *spT = h->dz2T;
    for (int l = L - 1; l >= 0; --l) {
        const int K = h->dims[l], N = h->dims[l + 1];
        NtArgs w{};
        w.A = h->aT[l];
        w.B = dzT;
        w.M = K;
        w.N = N;
        w.K = rows;
        w.c = h->g + h->w_off[l];
        w.ws = h->ws;
        w.counters = h->counters;
        launch_nt<EPI_F32>(h->device, w, h->stream, true, h->ws_elems);
        MLP_CHECK(h, hipGetLastError());
        hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)N), dim3(EW_BLOCK), 0, h->stream, dzT, h->g + h->b_off[l],
            rows);
        MLP_CHECK(h, hipGetLastError());
        if (l == 0) { break; }
        NtArgs d{};
        d.A = dz;
        d.B = h->pb + h->w_off[l];  // W stored [in][out]: K-major for this product
        d.M = rows;
        d.N = K;
        d.K = N;
        d.mask = h->a[l];
        d.out = l > 1 ? sp : nullptr;  // layer 0's gradient is read only transposed
        d.out_t = spT;
        launch_nt<EPI_BACK>(h->device, d, h->stream, true, 0);
        MLP_CHECK(h, hipGetLastError());
        std::swap(dz, sp);
        std::swap(dzT, spT);
    }
    return PSIM_OK;
}

int load_batch(psim_mlp* h, const float* x, const float* t, int rows) {
    int rc;
    if ((rc = t ? ensure_train(h, rows) : ensure_batch(h, rows)) != 0) { return rc; }
    MLP_CHECK(h, hipMemcpyAsync(h->xin, x, sizeof(float) * (size_t)rows * h->n_in, hipMemcpyHostToDevice, h->stream));
    if (t) { MLP_CHECK(h, hipMemcpyAsync(h->t, t, sizeof(float) * (size_t)rows, hipMemcpyHostToDevice, h->stream)); }
    return PSIM_OK;
}

// Device address of this GPU's scratch array, looked up once per device with that device current.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mutex;
unsigned char* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(psim_mlp* h) {
    if (h->device < 0 || h->device >= NOALLOC_MAX_DEVICES) {
        return mlp_fail(h, PSIM_E_INVALID, "device index beyond the scratch table");
    }
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_ptr[h->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            return mlp_fail(h, PSIM_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        noalloc_ptr[h->device] = static_cast<unsigned char*>(ptr);
    }
    h->scratch = noalloc_ptr[h->device];
    return PSIM_OK;
}

// Layout of the no-allocation forward pass in the scratch array for `rows` rows (a multiple of
// BATCH_ROUND): every piece starts on a 256-byte boundary, as hipMalloc'd buffers do, so the
// kernels take the same vector paths as in psim_mlp_forward.
struct NoallocLayout {
    size_t xin, a0, act0, act1, zout, y, end;
};

NoallocLayout noalloc_layout(const psim_mlp* h, int64_t rows) {
    const auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t wide = (size_t)std::max(h->hidden, PAD_OUT);
    NoallocLayout l{};
    size_t at = 0;
    l.xin = at;
    at = up(at + (size_t)rows * h->n_in * sizeof(float));
    l.a0 = at;
    at = up(at + (size_t)rows * h->n_in * sizeof(uint16_t));
    l.act0 = at;
    at = up(at + (size_t)rows * wide * sizeof(uint16_t));
    l.act1 = at;
    at = up(at + (size_t)rows * wide * sizeof(uint16_t));
    l.zout = at;
    at = up(at + (size_t)rows * PAD_OUT * sizeof(float));
    l.y = at;
    at = up(at + (size_t)rows * sizeof(float));
    l.end = at;
    return l;
}

// The most rows one no-allocation call takes for this network.
int noalloc_max_rows(const psim_mlp* h) {
    int64_t lo = 0, hi = MAX_ROWS / BATCH_ROUND;  // in units of BATCH_ROUND rows
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (noalloc_layout(h, mid * BATCH_ROUND).end <= NOALLOC_SCRATCH_BYTES) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return (int)(lo * BATCH_ROUND);
}

int read_loss(psim_mlp* h, int rows, float* loss) {
    std::vector<float> terms((size_t)rows);
    MLP_CHECK(h, hipMemcpyAsync(terms.data(), h->loss, sizeof(float) * (size_t)rows, hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    double s = 0.;
    for (float v : terms) { s += v; }
    *loss = (float)s;
    return PSIM_OK;
}

}  // namespace

extern "C" {

int psim_mlp_create(psim_mlp** out, int device, int32_t n_in, int32_t hidden, int32_t layers) {
    if (!out || n_in <= 0 || hidden <= 0 || layers <= 0 || n_in > 4096 || hidden > 8192 || layers > MAX_WEIGHTS - 1) {
        return PSIM_E_INVALID;
    }
    *out = nullptr;
    auto* h = new psim_mlp();
    h->device = device;
    h->n_in = n_in;
    h->hidden = hidden;
    h->layers = layers;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
        delete h;
        return PSIM_E_DEVICE;
    }
    h->dims.push_back(n_in);
    for (int l = 0; l < layers; ++l) { h->dims.push_back(hidden); }
    h->dims.push_back(PAD_OUT);
    int64_t off = 0;
    for (size_t l = 0; l + 1 < h->dims.size(); ++l) {
        h->w_off.push_back(off);
        off += (int64_t)h->dims[l] * h->dims[l + 1];
        h->b_off.push_back(off);
        off += h->dims[l + 1];
    }
    h->n_params = off;
    h->p = mlp_alloc<float>((size_t)off);
    h->g = mlp_alloc<float>((size_t)off);
    h->m1 = mlp_alloc<float>((size_t)off);
    h->m2 = mlp_alloc<float>((size_t)off);
    h->mask = mlp_alloc<float>((size_t)off);
    h->pb = mlp_alloc<uint16_t>((size_t)off);
    h->pbT = mlp_alloc<uint16_t>((size_t)off);
    size_t widest = 0;
    for (size_t l = 0; l + 1 < h->dims.size(); ++l) { widest = std::max(widest, (size_t)h->dims[l] * h->dims[l + 1]); }
    h->ws_elems = (size_t)MAX_SPLITS * widest;
    h->ws = mlp_alloc<float>(h->ws_elems);
    h->counters = mlp_alloc<unsigned>(NT_COUNTERS);
    h->step_consts = mlp_alloc<float>(4);
    if (hipHostMalloc(reinterpret_cast<void**>(&h->h_step_consts), 4 * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        h->h_step_consts = nullptr;
    }
    if (!h->p || !h->g || !h->m1 || !h->m2 || !h->mask || !h->pb || !h->pbT || !h->ws || !h->counters || !h->step_consts
        || !h->h_step_consts) {
        psim_mlp_destroy(h);
        return PSIM_E_DEVICE;
    }
    // mask: 1 for every parameter except the padded output columns (weights and biases)
    std::vector<float> mask((size_t)off, 1.f);
    const int L = (int)h->dims.size() - 1;
    for (int k = 0; k < h->dims[L - 1]; ++k) {
        for (int n = 1; n < PAD_OUT; ++n) { mask[(size_t)(h->w_off[L - 1] + (int64_t)k * PAD_OUT + n)] = 0.f; }
    }
    for (int n = 1; n < PAD_OUT; ++n) { mask[(size_t)(h->b_off[L - 1] + n)] = 0.f; }
    if (hipMemcpy(h->mask, mask.data(), sizeof(float) * mask.size(), hipMemcpyHostToDevice) != hipSuccess
        || hipMemset(h->p, 0, sizeof(float) * (size_t)off) != hipSuccess
        || hipMemset(h->m1, 0, sizeof(float) * (size_t)off) != hipSuccess
        || hipMemset(h->m2, 0, sizeof(float) * (size_t)off) != hipSuccess
        || hipMemset(h->pb, 0, sizeof(uint16_t) * (size_t)off) != hipSuccess
        || hipMemset(h->pbT, 0, sizeof(uint16_t) * (size_t)off) != hipSuccess
        || hipMemset(h->counters, 0, sizeof(unsigned) * NT_COUNTERS) != hipSuccess) {
        psim_mlp_destroy(h);
        return PSIM_E_DEVICE;
    }
    if (noalloc_scratch(h) != PSIM_OK) {
        psim_mlp_destroy(h);
        return PSIM_E_DEVICE;
    }
    *out = h;
    return PSIM_OK;
}

void psim_mlp_destroy(psim_mlp* h) {
    if (!h) { return; }
    (void)hipSetDevice(h->device);
    if (h->stream) { (void)hipStreamSynchronize(h->stream); }
    if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); }
    if (h->graph) { (void)hipGraphDestroy(h->graph); }
    if (h->h_step_consts) { (void)hipHostFree(h->h_step_consts); }
    free_train(h);
    free_batch(h);
    void* ps[] = { h->p, h->g, h->m1, h->m2, h->mask, h->pb, h->pbT, h->ws, h->counters, h->step_consts };
    for (void* q : ps) {
        if (q) { (void)hipFree(q); }
    }
    if (h->stream) { (void)hipStreamDestroy(h->stream); }
    delete h;
}

const char* psim_mlp_last_error(const psim_mlp* h) { return h ? h->err.c_str() : "null surrogate"; }

int64_t psim_mlp_num_params(const psim_mlp* h) { return h ? h->n_params : 0; }

int psim_mlp_set_params(psim_mlp* h, const float* params) {
    if (!h || !params) { return PSIM_E_INVALID; }
    MLP_CHECK(h, hipSetDevice(h->device));
    std::vector<float> p(params, params + h->n_params);
    // padded output columns are forced to zero
    const int L = (int)h->dims.size() - 1;
    for (int k = 0; k < h->dims[L - 1]; ++k) {
        for (int n = 1; n < PAD_OUT; ++n) { p[(size_t)(h->w_off[L - 1] + (int64_t)k * PAD_OUT + n)] = 0.f; }
    }
    for (int n = 1; n < PAD_OUT; ++n) { p[(size_t)(h->b_off[L - 1] + n)] = 0.f; }
    MLP_CHECK(h, hipMemcpy(h->p, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(params_bf16_kernel, dim3(ew_blocks(h->n_params)), dim3(EW_BLOCK), 0, h->stream, h->p, h->pb,
        h->pbT, layer_table(h), h->n_params);
    MLP_CHECK(h, hipGetLastError());
    MLP_CHECK(h, hipMemsetAsync(h->m1, 0, sizeof(float) * (size_t)h->n_params, h->stream));
    MLP_CHECK(h, hipMemsetAsync(h->m2, 0, sizeof(float) * (size_t)h->n_params, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    h->step = 0;
    return PSIM_OK;
}

int psim_mlp_get_params(psim_mlp* h, float* params) {
    if (!h || !params) { return PSIM_E_INVALID; }
    MLP_CHECK(h, hipSetDevice(h->device));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    MLP_CHECK(h, hipMemcpy(params, h->p, sizeof(float) * (size_t)h->n_params, hipMemcpyDeviceToHost));
    return PSIM_OK;
}

int psim_mlp_forward(psim_mlp* h, const float* x, int32_t rows, float* y) {
    if (!h || !x || !y || rows <= 0) { return PSIM_E_INVALID; }
    if (rows > MAX_ROWS) { return mlp_fail(h, PSIM_E_INVALID, "more than 4194240 rows in one call"); }
    MLP_CHECK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = load_batch(h, x, nullptr, rows)) != 0) { return rc; }
    if ((rc = forward(h, rows, false)) != 0) { return rc; }
    std::vector<float> z((size_t)rows * PAD_OUT);
    MLP_CHECK(h, hipMemcpyAsync(z.data(), h->zout, sizeof(float) * z.size(), hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    for (int r = 0; r < rows; ++r) { y[r] = z[(size_t)r * PAD_OUT]; }
    return PSIM_OK;
}

int psim_mlp_forward_noalloc(psim_mlp* h, const float* x, int32_t rows, float* y) {
    if (!h || !x || !y || rows <= 0) { return PSIM_E_INVALID; }
    const int max_rows = noalloc_max_rows(h);
    if (rows > max_rows) {
        return mlp_fail(h, PSIM_E_INVALID,
            "psim_mlp_forward_noalloc: " + std::to_string(rows) + " rows exceed the " + std::to_string(max_rows)
                + " this network fits in the 64 MiB scratch array; split the call or use psim_mlp_forward");
    }
    if (!h->scratch) { return mlp_fail(h, PSIM_E_STATE, "scratch address unknown"); }
    MLP_CHECK(h, hipSetDevice(h->device));
    const int64_t cap = ((int64_t)rows + BATCH_ROUND - 1) / BATCH_ROUND * BATCH_ROUND;
    const NoallocLayout lay = noalloc_layout(h, cap);
    float* xin = reinterpret_cast<float*>(h->scratch + lay.xin);
    uint16_t* act0 = reinterpret_cast<uint16_t*>(h->scratch + lay.act0);
    uint16_t* act1 = reinterpret_cast<uint16_t*>(h->scratch + lay.act1);
    float* zout = reinterpret_cast<float*>(h->scratch + lay.zout);
    float* yd = reinterpret_cast<float*>(h->scratch + lay.y);
    // the kernels of forward(), with two activation buffers used in turn; tile shapes already
    // tuned by an earlier call of the same shape, else the untimed default (a tile never changes a
    // result), so no events are created here
    const int L = (int)h->dims.size() - 1;
    std::vector<uint16_t*> a((size_t)L);
    a[0] = reinterpret_cast<uint16_t*>(h->scratch + lay.a0);
    for (int l = 1; l < L; ++l) { a[(size_t)l] = (l % 2) ? act0 : act1; }
    MLP_CHECK(h, hipMemcpyAsync(xin, x, sizeof(float) * (size_t)rows * h->n_in, hipMemcpyHostToDevice, h->stream));
    int rc;
    if ((rc = forward_pass(h, rows, xin, a.data(), nullptr, zout, false)) != 0) { return rc; }
    hipLaunchKernelGGL(column0_kernel, dim3(ew_blocks(rows)), dim3(EW_BLOCK), 0, h->stream, zout, yd, rows, PAD_OUT);
    MLP_CHECK(h, hipGetLastError());
    MLP_CHECK(h, hipMemcpyAsync(y, yd, sizeof(float) * (size_t)rows, hipMemcpyDeviceToHost, h->stream));
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return PSIM_OK;
}

int psim_mlp_gradients(psim_mlp* h, const float* x, const float* t, int32_t rows, float* grad, float* loss) {
    if (!h || !x || !t || rows <= 0) { return PSIM_E_INVALID; }
    if (rows > MAX_ROWS) { return mlp_fail(h, PSIM_E_INVALID, "more than 4194240 rows in one call"); }
    MLP_CHECK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = load_batch(h, x, t, rows)) != 0) { return rc; }
    if ((rc = forward(h, rows, true)) != 0) { return rc; }
    if ((rc = backward(h, rows)) != 0) { return rc; }
    if (loss && (rc = read_loss(h, rows, loss)) != 0) { return rc; }
    if (grad) {
        MLP_CHECK(h, hipStreamSynchronize(h->stream));
        MLP_CHECK(h, hipMemcpy(grad, h->g, sizeof(float) * (size_t)h->n_params, hipMemcpyDeviceToHost));
    }
    return PSIM_OK;
}

int psim_mlp_train_step(psim_mlp* h, const float* x, const float* t, int32_t rows, float lr, float* loss) {
    if (!h || !x || !t || rows <= 0 || !(lr > 0.f)) { return PSIM_E_INVALID; }
    if (rows > MAX_ROWS) { return mlp_fail(h, PSIM_E_INVALID, "more than 4194240 rows in one call"); }
    MLP_CHECK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = load_batch(h, x, t, rows)) != 0) { return rc; }
    ++h->step;
    const float b1 = 0.9f, b2 = 0.999f;
    h->h_step_consts[0] = lr;
    h->h_step_consts[1] = 1.f - std::pow(b1, (float)h->step);
    h->h_step_consts[2] = 1.f - std::pow(b2, (float)h->step);
    MLP_CHECK(h, hipMemcpyAsync(h->step_consts, h->h_step_consts, 3 * sizeof(float), hipMemcpyHostToDevice, h->stream));
    // The step's kernels: forward, backward, Adam (all on h->stream, no host synchronisation).
    const auto enqueue = [&]() -> int {
        int r;
        if ((r = forward(h, rows, true)) != 0) { return r; }
        if ((r = backward(h, rows)) != 0) { return r; }
        hipLaunchKernelGGL(adam_kernel, dim3(ew_blocks(h->n_params)), dim3(EW_BLOCK), 0, h->stream, h->p, h->g, h->m1,
            h->m2, h->mask, h->pb, h->pbT, layer_table(h), h->n_params, h->step_consts, b1, b2, 1e-8f);
        MLP_CHECK(h, hipGetLastError());
        return PSIM_OK;
    };
    if (h->graph_rows && (h->alloc_gen != h->graph_gen || h->graph_rows != rows)) {  // buffers moved or another size
        if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); }
        if (h->graph) { (void)hipGraphDestroy(h->graph); }
        h->graph_exec = nullptr;
        h->graph = nullptr;
        h->graph_rows = 0;
    }
    if (!h->use_graph || (h->graph_rows != rows && h->warm_rows != rows)) {
        // eager: the first step at this batch size (its products' tiles are tuned here)
        if ((rc = enqueue()) != 0) { return rc; }
        h->warm_rows = rows;
    } else {
        if (h->graph_rows != rows) {  // capture the step once, then replay it
            MLP_CHECK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(h->stream, &g);
            if (rc != 0) {
                if (g) { (void)hipGraphDestroy(g); }
                return rc;
            }
            if (e != hipSuccess) { return mlp_fail(h, PSIM_E_DEVICE, std::string("capturing a training step: ") + hipGetErrorString(e)); }
            h->graph = g;
            MLP_CHECK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
            h->graph_rows = rows;
            h->graph_gen = h->alloc_gen;
        }
        MLP_CHECK(h, hipGraphLaunch(h->graph_exec, h->stream));
    }
    if (loss) { return read_loss(h, rows, loss); }
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return PSIM_OK;
}

int psim_mlp_set_graph(psim_mlp* h, int32_t enable) {
    if (!h) { return PSIM_E_INVALID; }
    h->use_graph = enable != 0;
    return PSIM_OK;
}

int psim_mlp_time_product(psim_mlp* h, int32_t kind, int32_t rows, int32_t reps, float* ms_per_launch, int32_t* tile) {
    if (!h || !ms_per_launch || rows <= 0 || reps <= 0 || kind < 0 || kind > 3 || h->layers < 2) {
        return PSIM_E_INVALID;
    }
    if (rows > MAX_ROWS) { return mlp_fail(h, PSIM_E_INVALID, "more than 4194240 rows in one call"); }
    MLP_CHECK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = kind == 0 ? ensure_batch(h, rows) : ensure_train(h, rows)) != 0) { return rc; }
    // layer 1 (hidden x hidden) of the network, on its own buffers (contents as left by earlier calls)
    const int l = 1, K = h->dims[l], N = h->dims[l + 1];
    NtArgs a{};
    if (kind <= 1) {
        a.A = h->a[l];
        a.B = h->pbT + h->w_off[l];
        a.M = rows;
        a.N = N;
        a.K = K;
        a.bias = h->p + h->b_off[l];
        a.relu = 1;
        a.out = h->a[l + 1];
        a.out_t = kind == 1 ? h->aT[l + 1] : nullptr;
    } else if (kind == 2) {
        a.A = h->aT[l];
        a.B = h->dzT;
        a.M = K;
        a.N = N;
        a.K = rows;
        a.c = h->g + h->w_off[l];
        a.ws = h->ws;
        a.counters = h->counters;
    } else {
        a.A = h->dz;
        a.B = h->pb + h->w_off[l];
        a.M = rows;
        a.N = K;
        a.K = N;
        a.mask = h->a[l];
        a.out = h->dz2;
        a.out_t = h->dz2T;
    }
    const auto launch = [&]() {
        if (kind <= 1) { return launch_nt<EPI_FWD>(h->device, a, h->stream, true, 0); }
        if (kind == 2) { return launch_nt<EPI_F32>(h->device, a, h->stream, true, h->ws_elems); }
        return launch_nt<EPI_BACK>(h->device, a, h->stream, true, 0);
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MLP_CHECK(h, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return mlp_fail(h, PSIM_E_DEVICE, "hipEventCreate failed");
    }
    launch();  // the first product of this shape tunes it
    const int used = launch();
    (void)hipEventRecord(e0, h->stream);
    for (int r = 0; r < reps; ++r) { launch(); }
    (void)hipEventRecord(e1, h->stream);
    float ms = 0.f;
    const bool ok = hipGetLastError() == hipSuccess && hipEventSynchronize(e1) == hipSuccess
                    && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) { return mlp_fail(h, PSIM_E_DEVICE, "timing the product failed"); }
    *ms_per_launch = ms / (float)reps;
    if (tile) {
        *tile = used >= NT_NUM_TILES ? -((used == NT_NUM_TILES ? 4 : 8) * BR_WROWS * 1000 + BR_BN)
                                     : NT_TILES[used].bm * 1000 + NT_TILES[used].bn;
    }
    return PSIM_OK;
}

// C = op(A) x op(B) for bf16 operands given as FP32 (rounded to bf16 here); exposed so the MFMA
// lane layout and the transposed loads can be checked against a host product.
int psim_mlp_gemm_check(int device, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float* A,
    const float* B, float* C) {
    if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || M > MAX_ROWS || N > MAX_ROWS || K > MAX_ROWS) {
        return PSIM_E_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) { return PSIM_E_DEVICE; }
    const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
    std::vector<uint16_t> ha(na), hb(nb);
    for (size_t i = 0; i < na; ++i) { ha[i] = f32_to_bf16(A[i]); }
    for (size_t i = 0; i < nb; ++i) { hb[i] = f32_to_bf16(B[i]); }
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *ws = nullptr;
    unsigned* counters = nullptr;
    const size_t ws_elems = (size_t)MAX_SPLITS * nc;  // products split over K are checked too
    int rc = PSIM_OK;
    if (hipMalloc(&da, na * 2) != hipSuccess || hipMalloc(&db, nb * 2) != hipSuccess
        || hipMalloc(&dc, nc * 4) != hipSuccess || hipMalloc(&ws, ws_elems * 4) != hipSuccess
        || hipMalloc(&counters, NT_COUNTERS * sizeof(unsigned)) != hipSuccess
        || hipMemset(counters, 0, NT_COUNTERS * sizeof(unsigned)) != hipSuccess) {
        rc = PSIM_E_DEVICE;
    } else if (hipMemcpy(da, ha.data(), na * 2, hipMemcpyHostToDevice) != hipSuccess
               || hipMemcpy(db, hb.data(), nb * 2, hipMemcpyHostToDevice) != hipSuccess) {
        rc = PSIM_E_DEVICE;
    } else {
        launch_gemm(device, ta != 0, tb != 0, da, db, dc, M, N, K, 0, ws, ws_elems, counters, true);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess
            || hipMemcpy(C, dc, nc * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = PSIM_E_DEVICE;
        }
    }
    void* bufs[] = { da, db, dc, ws, counters };
    for (void* q : bufs) {
        if (q) { (void)hipFree(q); }
    }
    return rc;
}

// Average duration of one product op(A)[M][K] x op(B)[K][N] with an FP32 result (a product split
// over K includes its ordered sum of the partials), from HIP events around `reps` products on a
// stream of its own (operands filled on the device).
int psim_mlp_time_gemm(int device, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, int32_t reps,
    float* ms_per_launch) {
    if (!ms_per_launch || M <= 0 || N <= 0 || K <= 0 || reps <= 0 || M > MAX_ROWS || N > MAX_ROWS || K > MAX_ROWS) {
        return PSIM_E_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) { return PSIM_E_DEVICE; }
    const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *src = nullptr, *ws = nullptr;
    unsigned* counters = nullptr;
    const size_t ws_elems = (size_t)MAX_SPLITS * nc;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PSIM_OK;
    const size_t nsrc = std::max(na, nb);
    if (hipMalloc(&da, na * 2) != hipSuccess || hipMalloc(&db, nb * 2) != hipSuccess
        || hipMalloc(&dc, nc * 4) != hipSuccess || hipMalloc(&src, nsrc * 4) != hipSuccess
        || hipMalloc(&ws, ws_elems * 4) != hipSuccess || hipMalloc(&counters, NT_COUNTERS * sizeof(unsigned)) != hipSuccess
        || hipMemset(counters, 0, NT_COUNTERS * sizeof(unsigned)) != hipSuccess || hipStreamCreate(&st) != hipSuccess
        || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = PSIM_E_DEVICE;
    } else {
        std::vector<float> h(nsrc);
        for (size_t i = 0; i < nsrc; ++i) { h[i] = (float)((i * 2654435761u) % 1000) * 1e-3f - 0.5f; }
        auto launch = [&] { launch_gemm(device, ta != 0, tb != 0, da, db, dc, M, N, K, st, ws, ws_elems, counters, true); };
        float ms = 0.f;
        if (hipMemcpy(src, h.data(), nsrc * 4, hipMemcpyHostToDevice) != hipSuccess) {
            rc = PSIM_E_DEVICE;
        } else {
            hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_blocks((int64_t)na)), dim3(EW_BLOCK), 0, st, src, da, (int64_t)na);
            hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_blocks((int64_t)nb)), dim3(EW_BLOCK), 0, st, src, db, (int64_t)nb);
            launch();
            launch();
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps; ++r) { launch(); }
            (void)hipEventRecord(e1, st);
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess
                || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = PSIM_E_DEVICE;
            }
        }
        *ms_per_launch = ms / (float)reps;
    }
    if (e0) { (void)hipEventDestroy(e0); }
    if (e1) { (void)hipEventDestroy(e1); }
    if (st) { (void)hipStreamDestroy(st); }
    void* bufs[] = { da, db, dc, src, ws, counters };
    for (void* q : bufs) {
        if (q) { (void)hipFree(q); }
    }
    return rc;
}

}  // extern "C"
