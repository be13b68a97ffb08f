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
#include <cstring>
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

// float -> bfloat16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) { return (uint16_t)((u >> 16) | 0x40u); }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
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

void launch_gemm(bool ta, bool tb, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
    hipStream_t st, float* ws = nullptr, size_t ws_elems = 0) {
    if (ta && tb) {
        launch_gemm_t<true, true>(A, B, C, M, N, K, st, ws, ws_elems);
    } else if (ta) {
        launch_gemm_t<true, false>(A, B, C, M, N, K, st, ws, ws_elems);
    } else if (tb) {
        launch_gemm_t<false, true>(A, B, C, M, N, K, st, ws, ws_elems);
    } else {
        launch_gemm_t<false, false>(A, B, C, M, N, K, st, ws, ws_elems);
    }
}

// dst[n][k] = src[k][n] for a K x N bf16 matrix, through a 32 x 32 LDS tile.
__global__ void transpose_bf16_kernel(const uint16_t* __restrict__ src, uint16_t* __restrict__ dst, int K, int N) {
    __shared__ uint16_t tile[32][33];
    const int n0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 256 threads: 8 rows per pass
    for (int r = ty; r < 32; r += 8) {
        const int k = k0 + r, n = n0 + tx;
        tile[r][tx] = (k < K && n < N) ? src[(int64_t)k * N + n] : (uint16_t)0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int n = n0 + r, k = k0 + tx;
        if (n < N && k < K) { dst[(int64_t)n * K + k] = tile[tx][r]; }
    }
}

// z += bias (per column); a = relu ? max(z, 0) : z, rounded to bf16.
__global__ void bias_act_kernel(float* __restrict__ z, const float* __restrict__ bias, uint16_t* __restrict__ a,
    int64_t rows, int cols, int relu) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= rows * cols) { return; }
    const float v = z[i] + bias[i % cols];
    z[i] = v;
    a[i] = f32_to_bf16(relu ? fmaxf(v, 0.f) : v);
}

// Output-layer gradient of the mean squared error on column 0: dz[r][0] = 2 (y - t) / rows, other
// columns 0 (bf16), and the loss terms (y - t)^2 / rows per row.
__global__ void loss_grad_kernel(const float* __restrict__ z, const float* __restrict__ t, uint16_t* __restrict__ dz,
    float* __restrict__ loss_terms, int rows, int cols) {
    const int r = blockIdx.x * EW_BLOCK + threadIdx.x;
    if (r >= rows) { return; }
    const float d = z[(int64_t)r * cols] - t[r];
    loss_terms[r] = d * d / (float)rows;
    for (int c = 0; c < cols; ++c) { dz[(int64_t)r * cols + c] = f32_to_bf16(c == 0 ? 2.f * d / (float)rows : 0.f); }
}

// dz = da * (z > 0), rounded to bf16 (da FP32 from the GEMM).
__global__ void relu_back_kernel(const float* __restrict__ da, const float* __restrict__ z, uint16_t* __restrict__ dz,
    int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) { return; }
    dz[i] = f32_to_bf16(z[i] > 0.f ? da[i] : 0.f);
}

// db[c] = sum over rows of dz[r][c] (one block per column, FP32).
__global__ void bias_grad_kernel(const uint16_t* __restrict__ dz, float* __restrict__ db, int rows, int cols) {
    __shared__ float part[EW_BLOCK];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += EW_BLOCK) { s += bf16_to_f32(dz[(int64_t)r * cols + c]); }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = EW_BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) { part[threadIdx.x] += part[threadIdx.x + off]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { db[c] = part[0]; }
}

// Adam on the FP32 master parameters, then their bf16 copy (weights only; masked entries stay 0).
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
    float* __restrict__ m2, const float* __restrict__ mask, uint16_t* __restrict__ pb, int64_t n, float lr, float b1,
    float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) { return; }
    const float gi = g[i] * (mask ? mask[i] : 1.f);
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    if (pb) { pb[i] = f32_to_bf16(v); }
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
    int64_t step = 0;
    // batch buffers
    int cap = 0;
    std::vector<float*> z;  // pre-activations per layer [cap][dims[l + 1]]
    std::vector<uint16_t*> a;  // activations: a[0] = input, a[l + 1] = output of layer l
    float* da = nullptr;  // FP32 gradient from a GEMM [cap][hidden]
    uint16_t* dz = nullptr;  // bf16 gradient of the current layer's pre-activation
    uint16_t* dz2 = nullptr;
    float* xin = nullptr;
    float* t = nullptr;
    float* loss = nullptr;
    // device address of this GPU's `scratch` array (psim_mlp_forward_noalloc), looked up at create
    unsigned char* scratch = nullptr;
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

void free_batch(psim_mlp* h) {
    for (float* q : h->z) { (void)hipFree(q); }
    for (uint16_t* q : h->a) { (void)hipFree(q); }
    h->z.clear();
    h->a.clear();
    void* rest[] = { h->da, h->dz, h->dz2, h->xin, h->t, h->loss };
    for (void* q : rest) {
        if (q) { (void)hipFree(q); }
    }
    h->da = nullptr;
    h->dz = h->dz2 = nullptr;
    h->xin = h->t = h->loss = nullptr;
    h->cap = 0;
}

int ensure_batch(psim_mlp* h, int rows) {
    if (rows <= h->cap) { return PSIM_OK; }
    free_batch(h);
    const int cap = ((rows + BATCH_ROUND - 1) / BATCH_ROUND) * BATCH_ROUND;
    const int L = (int)h->dims.size() - 1;
    for (int l = 0; l < L; ++l) { h->z.push_back(mlp_alloc<float>((size_t)cap * h->dims[l + 1])); }
    for (int l = 0; l <= L; ++l) { h->a.push_back(mlp_alloc<uint16_t>((size_t)cap * h->dims[l])); }
    int widest = 0;
    for (int d : h->dims) { widest = std::max(widest, d); }
    h->da = mlp_alloc<float>((size_t)cap * widest);
    h->dz = mlp_alloc<uint16_t>((size_t)cap * widest);
    h->dz2 = mlp_alloc<uint16_t>((size_t)cap * widest);
    h->xin = mlp_alloc<float>((size_t)cap * h->n_in);
    h->t = mlp_alloc<float>((size_t)cap);
    h->loss = mlp_alloc<float>((size_t)cap);
    bool ok = h->da && h->dz && h->dz2 && h->xin && h->t && h->loss;
    for (float* q : h->z) { ok = ok && q; }
    for (uint16_t* q : h->a) { ok = ok && q; }
    if (!ok) {
        free_batch(h);
        return mlp_fail(h, PSIM_E_DEVICE, "hipMalloc failed for the surrogate batch buffers");
    }
    h->cap = cap;
    return PSIM_OK;
}

template <bool TA, bool TB>
int gemm(psim_mlp* h, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K, bool split = false) {
    launch_gemm(TA, TB, A, B, C, M, N, K, h->stream, split ? h->ws : nullptr, split ? h->ws_elems : 0);
    MLP_CHECK(h, hipGetLastError());
    return PSIM_OK;
}

// pbT from pb: every layer's weights transposed.
int refresh_transposed(psim_mlp* h) {
    for (size_t l = 0; l + 1 < h->dims.size(); ++l) {
        const int K = h->dims[l], N = h->dims[l + 1];
        const dim3 grid((unsigned)((N + 31) / 32), (unsigned)((K + 31) / 32));
        hipLaunchKernelGGL(transpose_bf16_kernel, grid, dim3(256), 0, h->stream, h->pb + h->w_off[l],
            h->pbT + h->w_off[l], K, N);
        MLP_CHECK(h, hipGetLastError());
    }
    return PSIM_OK;
}

// Forward pass over `rows` inputs already in h->xin; leaves z and a of every layer.
int forward(psim_mlp* h, int rows) {
    int rc;
    const int L = (int)h->dims.size() - 1;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_blocks((int64_t)rows * h->n_in)), dim3(EW_BLOCK), 0, h->stream, h->xin,
        h->a[0], (int64_t)rows * h->n_in);
    MLP_CHECK(h, hipGetLastError());
    for (int l = 0; l < L; ++l) {
        const int K = h->dims[l], N = h->dims[l + 1];
        // B = W^T stored [N][K]: both operands K-major
        if ((rc = gemm<false, true>(h, h->a[l], h->pbT + h->w_off[l], h->z[l], rows, N, K)) != 0) { return rc; }
        hipLaunchKernelGGL(bias_act_kernel, dim3(ew_blocks((int64_t)rows * N)), dim3(EW_BLOCK), 0, h->stream, h->z[l],
            h->p + h->b_off[l], h->a[l + 1], (int64_t)rows, N, l + 1 < L ? 1 : 0);
        MLP_CHECK(h, hipGetLastError());
    }
    return PSIM_OK;
}

// Loss and gradients of every parameter (into h->g) for targets in h->t.
int backward(psim_mlp* h, int rows) {
    int rc;
    const int L = (int)h->dims.size() - 1;
    hipLaunchKernelGGL(loss_grad_kernel, dim3(ew_blocks(rows)), dim3(EW_BLOCK), 0, h->stream, h->z[L - 1], h->t, h->dz,
        h->loss, rows, h->dims[L]);
    MLP_CHECK(h, hipGetLastError());
    uint16_t* dz = h->dz;
    uint16_t* spare = h->dz2;
    for (int l = L - 1; l >= 0; --l) {
        const int K = h->dims[l], N = h->dims[l + 1];
        // dW[K][N] = a_l^T [K][rows] x dz [rows][N]; db[N] = column sums of dz
        if ((rc = gemm<true, false>(h, h->a[l], dz, h->g + h->w_off[l], K, N, rows, true)) != 0) { return rc; }
        hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)N), dim3(EW_BLOCK), 0, h->stream, dz, h->g + h->b_off[l],
            rows, N);
        MLP_CHECK(h, hipGetLastError());
        if (l == 0) { break; }
        // da[rows][K] = dz [rows][N] x W^T [N][K] (W stored [K][N]), then through the ReLU of layer l - 1
        if ((rc = gemm<false, true>(h, dz, h->pb + h->w_off[l], h->da, rows, K, N)) != 0) { return rc; }
        hipLaunchKernelGGL(relu_back_kernel, dim3(ew_blocks((int64_t)rows * K)), dim3(EW_BLOCK), 0, h->stream, h->da,
            h->z[l - 1], spare, (int64_t)rows * K);
        MLP_CHECK(h, hipGetLastError());
        std::swap(dz, spare);
    }
    return PSIM_OK;
}

int load_batch(psim_mlp* h, const float* x, const float* t, int rows) {
    int rc;
    if ((rc = ensure_batch(h, rows)) != 0) { return rc; }
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
// BATCH_ROUND): every piece starts on a 256-byte boundary, as hipMalloc'd buffers do, so the GEMM
// takes the same vector paths as in psim_mlp_forward.
struct NoallocLayout {
    size_t xin, a0, act0, act1, z, y, end;
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
    l.z = at;
    at = up(at + (size_t)rows * wide * sizeof(float));
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
    if (!out || n_in <= 0 || hidden <= 0 || layers <= 0 || n_in > 4096 || hidden > 8192 || layers > 16) {
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
    if (!h->p || !h->g || !h->m1 || !h->m2 || !h->mask || !h->pb || !h->pbT || !h->ws) {
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
        || hipMemset(h->pbT, 0, sizeof(uint16_t) * (size_t)off) != hipSuccess) {
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
    free_batch(h);
    void* ps[] = { h->p, h->g, h->m1, h->m2, h->mask, h->pb, h->pbT, h->ws };
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
    hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_blocks(h->n_params)), dim3(EW_BLOCK), 0, h->stream, h->p, h->pb,
        h->n_params);
    MLP_CHECK(h, hipGetLastError());
    int rc;
    if ((rc = refresh_transposed(h)) != 0) { return rc; }
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
    if ((rc = forward(h, rows)) != 0) { return rc; }
    const int L = (int)h->dims.size() - 1;
    std::vector<float> z((size_t)rows * PAD_OUT);
    MLP_CHECK(h, hipMemcpyAsync(z.data(), h->z[L - 1], sizeof(float) * z.size(), hipMemcpyDeviceToHost, h->stream));
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
    uint16_t* act[3] = { reinterpret_cast<uint16_t*>(h->scratch + lay.a0),
        reinterpret_cast<uint16_t*>(h->scratch + lay.act0), reinterpret_cast<uint16_t*>(h->scratch + lay.act1) };
    float* z = reinterpret_cast<float*>(h->scratch + lay.z);
    float* yd = reinterpret_cast<float*>(h->scratch + lay.y);
    // the kernels and launch shapes of forward(), with one pre-activation buffer and two
    // activation buffers used in turn
    MLP_CHECK(h, hipMemcpyAsync(xin, x, sizeof(float) * (size_t)rows * h->n_in, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_blocks((int64_t)rows * h->n_in)), dim3(EW_BLOCK), 0, h->stream, xin,
        act[0], (int64_t)rows * h->n_in);
    MLP_CHECK(h, hipGetLastError());
    const int L = (int)h->dims.size() - 1;
    const uint16_t* in = act[0];
    for (int l = 0; l < L; ++l) {
        const int K = h->dims[l], N = h->dims[l + 1];
        uint16_t* out = act[1 + (l % 2)];
        int rc;
        if ((rc = gemm<false, true>(h, in, h->pbT + h->w_off[l], z, rows, N, K)) != 0) { return rc; }
        hipLaunchKernelGGL(bias_act_kernel, dim3(ew_blocks((int64_t)rows * N)), dim3(EW_BLOCK), 0, h->stream, z,
            h->p + h->b_off[l], out, (int64_t)rows, N, l + 1 < L ? 1 : 0);
        MLP_CHECK(h, hipGetLastError());
        in = out;
    }
    hipLaunchKernelGGL(column0_kernel, dim3(ew_blocks(rows)), dim3(EW_BLOCK), 0, h->stream, z, yd, rows, PAD_OUT);
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
    if ((rc = forward(h, rows)) != 0) { return rc; }
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
    if ((rc = forward(h, rows)) != 0) { return rc; }
    if ((rc = backward(h, rows)) != 0) { return rc; }
    ++h->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)h->step), c2 = 1.f - std::pow(b2, (float)h->step);
    hipLaunchKernelGGL(adam_kernel, dim3(ew_blocks(h->n_params)), dim3(EW_BLOCK), 0, h->stream, h->p, h->g, h->m1, h->m2,
        h->mask, h->pb, h->n_params, lr, b1, b2, 1e-8f, c1, c2);
    MLP_CHECK(h, hipGetLastError());
    if ((rc = refresh_transposed(h)) != 0) { return rc; }
    if (loss) { return read_loss(h, rows, loss); }
    MLP_CHECK(h, hipStreamSynchronize(h->stream));
    return PSIM_OK;
}

// C = op(A) x op(B) for bf16 operands given as FP32 (rounded to bf16 here); exposed so the MFMA
// lane layout and the transposed loads can be checked against a host product.
int psim_mlp_gemm_check(int device, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float* A,
    const float* B, float* C) {
    if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0 || M > MAX_ROWS) { return PSIM_E_INVALID; }
    if (hipSetDevice(device) != hipSuccess) { return PSIM_E_DEVICE; }
    const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
    std::vector<uint16_t> ha(na), hb(nb);
    for (size_t i = 0; i < na; ++i) { ha[i] = f32_to_bf16(A[i]); }
    for (size_t i = 0; i < nb; ++i) { hb[i] = f32_to_bf16(B[i]); }
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *ws = nullptr;
    const size_t ws_elems = (size_t)MAX_SPLITS * nc;  // products split over K are checked too
    int rc = PSIM_OK;
    if (hipMalloc(&da, na * 2) != hipSuccess || hipMalloc(&db, nb * 2) != hipSuccess
        || hipMalloc(&dc, nc * 4) != hipSuccess || hipMalloc(&ws, ws_elems * 4) != hipSuccess) {
        rc = PSIM_E_DEVICE;
    } else if (hipMemcpy(da, ha.data(), na * 2, hipMemcpyHostToDevice) != hipSuccess
               || hipMemcpy(db, hb.data(), nb * 2, hipMemcpyHostToDevice) != hipSuccess) {
        rc = PSIM_E_DEVICE;
    } else {
        launch_gemm(ta != 0, tb != 0, da, db, dc, M, N, K, 0, ws, ws_elems);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess
            || hipMemcpy(C, dc, nc * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = PSIM_E_DEVICE;
        }
    }
    if (da) { (void)hipFree(da); }
    if (db) { (void)hipFree(db); }
    if (dc) { (void)hipFree(dc); }
    if (ws) { (void)hipFree(ws); }
    return rc;
}

// Average duration of one GEMM of the network for op(A)[M][K] x op(B)[K][N] (a product split over
// K includes its split reduce), from HIP events around `reps` products on a stream of its own
// (operands filled on the device).
int psim_mlp_time_gemm(int device, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, int32_t reps,
    float* ms_per_launch) {
    if (!ms_per_launch || M <= 0 || N <= 0 || K <= 0 || reps <= 0 || M > MAX_ROWS) { return PSIM_E_INVALID; }
    if (hipSetDevice(device) != hipSuccess) { return PSIM_E_DEVICE; }
    const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *src = nullptr, *ws = nullptr;
    const size_t ws_elems = (size_t)MAX_SPLITS * nc;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PSIM_OK;
    const size_t nsrc = std::max(na, nb);
    if (hipMalloc(&da, na * 2) != hipSuccess || hipMalloc(&db, nb * 2) != hipSuccess
        || hipMalloc(&dc, nc * 4) != hipSuccess || hipMalloc(&src, nsrc * 4) != hipSuccess
        || hipMalloc(&ws, ws_elems * 4) != hipSuccess || hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess
        || hipEventCreate(&e1) != hipSuccess) {
        rc = PSIM_E_DEVICE;
    } else {
        std::vector<float> h(nsrc);
        for (size_t i = 0; i < nsrc; ++i) { h[i] = (float)((i * 2654435761u) % 1000) * 1e-3f - 0.5f; }
        auto launch = [&] { launch_gemm(ta != 0, tb != 0, da, db, dc, M, N, K, st, ws, ws_elems); };
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
    void* bufs[] = { da, db, dc, src, ws };
    for (void* q : bufs) {
        if (q) { (void)hipFree(q); }
    }
    return rc;
}

}  // extern "C"
