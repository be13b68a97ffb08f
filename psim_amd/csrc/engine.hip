// MI355X phonon engine: emission, windowed transport and tally reduction kernels, and the
// psim_gpu_* C ABI (include/psim_gpu.h).
//
// Data layout in HBM:
//   model      cells, edges (3 per cell), sub-surfaces, sensors, materials, frequency/velocity
//              arrays and the CDF / LA-fraction tables: uploaded once, read-only, L2-resident.
//   pool       emission output, one slot per phonon id, structure of arrays: px py dx dy freq vel
//              age tts (f64), cell (i32, -1 = never starts), meta (u32: bin, polarization, sign,
//              started, rate sensor).  Read once, when the phonon's birth window fetches it.
//   tallies    counts [S][R] (i64) and flux [S][R][2] (f64), as Sensor::inc_energy_/inc_flux_.
//   partials   per transport block: [W][S] count (i32) and flux (2 x f64) of one launch window.
//
//   live       two dense buffers with the pool's fields plus the phonon id: each transport launch
//              reads the survivors of the previous window from one, in order, plus the phonons
//              born in its own window (ids sorted by birth step, counting sort), and appends its
//              survivors to the other.
//
// A run is: emission kernel (every phonon's birth state and birth step), birth-step scan and
// scatter, then ceil(N / W) transport launches each advancing the listed phonons across W
// measurement steps with their state in registers, each followed (when the window holds
// recorded steps) by a reduction of the block partials into the tallies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "engine_device.hpp"
#include "psim_gpu.h"

using namespace psim_dev;

namespace {

constexpr int BLOCK = 256;
constexpr int WAVE = 64;
constexpr int LDS_TALLY_BYTES = 48 * 1024;  // budget for one block's [W][S] window tally
constexpr int GLOBAL_TALLY_STEPS = 16;  // fewer LDS-tallied steps per window than this: HBM tallies
// Fewer phonons per sensor than this: global atomics too.  Small runs lose more to the many short
// LDS windows (launches, reductions, sorts) than HBM atomics lose to contention, large ones the
// reverse (measured at 1e4 and 1e5 phonons per sensor: 1.8-3x faster in HBM below, 2.7x slower
// above).
constexpr int64_t LDS_MIN_PHONONS_PER_SENSOR = 30000;
constexpr uint32_t QUEUE_CLAIM = 32;  // work-queue items a wave claims per atomic
constexpr uint32_t OUT_CLAIM = 64;  // survivor slots a wave reserves per atomic

// Tallies of a launch window.  A boundary a flight passes is recorded as part of a segment: the
// flight adds its sign and velocity over boundaries a..b, which a difference array stores as two
// entries (+ at a, - at b + 1) whatever the flight's length.  Normally the segment lies in the
// window and goes to a [W + 1][S] difference array per block in LDS, summed into per-block
// partials [W][S] at the block's end.  A segment past the window's end (a flight that crossed it),
// or every segment when even GLOBAL_TALLY_STEPS steps of S sensors do not fit the LDS budget
// (thousands of sensors, e.g. a batch of geometries), goes to [S][R + 1] difference arrays in HBM
// with atomics, summed into the tallies when the run ends (diff_prefix_kernel).
// Flux arrays hold FT elements (double, or float in the FP32 + Kahan tally precision).
struct Tally {
    int32_t* __restrict__ pcnt;  // [grid][W*S]
    void* __restrict__ pfx;  // FT
    void* __restrict__ pfy;
    float* __restrict__ pcx;  // FP32: Kahan compensation of pfx / pfy (the negative of the lost low part)
    float* __restrict__ pcy;
    int64_t* __restrict__ dcnt;  // [S][R + 1] (HBM difference arrays)
    void* __restrict__ dfx;  // FT
    void* __restrict__ dfy;
    int32_t rec;  // 1 if this window records
    int32_t global;  // 1: every segment into the HBM difference arrays
    int32_t W;
};

struct Fate {
    int64_t* __restrict__ step;
    int32_t* __restrict__ cause;
    double* __restrict__ xy;
};

// Kahan-compensated FP32 addition: s is the running sum and c its compensation, the negative of
// the low part s has lost, so s - c carries the sum to about twice FP32's precision (the library
// is compiled with -ffp-contract=off, so the four operations round as written).
__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ void fate_write(const Fate& f, int64_t id, int64_t step, int cause, double x, double y) {
    if (f.step) {
        f.step[id] = step;
        f.cause[id] = cause;
        f.xy[2 * id] = x;
        f.xy[2 * id + 1] = y;
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) { v += __shfl_xor(v, off, WAVE); }
    return v;
}

// The wave's sum of per-lane 32-bit counters, in 64 bits: 64 lanes of a long launch can together
// pass 2^32 phonon-steps.
__device__ __forceinline__ unsigned long long wave_sum64(uint32_t v) {
    unsigned long long x = v;
    for (int off = WAVE / 2; off > 0; off >>= 1) { x += __shfl_xor(x, off, WAVE); }
    return x;
}

// Phonon builders, psim/src/phononBuilder.cpp:6-49, with EmitSurface::getPhononTime
// (surface.cpp:67-69) and Line/Triangle::getRandPoint (geometry.cpp:140-143, 234-242).
__global__ void __launch_bounds__(BLOCK) emit_kernel(Model m, Pool pool, int64_t total,
    const uint64_t* __restrict__ seed_p, unsigned long long* __restrict__ counters, Fate fate,
    int32_t* __restrict__ birth_key, uint32_t* __restrict__ birth_hist) {
    const uint64_t seed = *seed_p;
    const int64_t id = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t not_started = 0;
    int32_t key = -1;
    if (id < total) {
        int lo = 0, hi = m.n_sources - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (m.sources[mid].offset <= id) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        const psim_source& src = m.sources[lo];
        const psim_cell& cl = m.cells[src.cell];
        const int mat = m.sensors[cl.sensor].material;
        const uint32_t sid = (uint32_t)(id - src.id_base);
        psim_stream st;
        psim_stream_init(&st, seed, psim_stream_tag(PSIM_TAG_EMIT, (uint32_t)src.stream), sid, 0);
        Ph p;
        p.cell = src.cell;
        p.sign = src.sign;
        double birth = 0.;
        if (src.kind == 0) {  // CellOriginBuilder::operator()
            resample(m, mat, src.table, p, st);
            double r1, r2;
            psim_urand2(&st, &r1, &r2);
            if (r1 + r2 > 1.) {
                r1 = 1 - r1;
                r2 = 1 - r2;
            }
            p.px = cl.x1 + (cl.x2 - cl.x1) * r1 + (cl.x3 - cl.x1) * r2;
            p.py = cl.y1 + (cl.y2 - cl.y1) * r1 + (cl.y3 - cl.y1) * r2;
            set_rand_direction(p, st);
        } else {  // SurfaceOriginBuilder::operator(), PhasorBuilder::operator()
            birth = src.start + src.duration * psim_urand(&st);
            resample(m, mat, src.table, p, st);
            const double r1 = psim_urand(&st);
            const double r2 = 1. - r1;
            p.px = src.p1x * r1 + src.p2x * r2;
            p.py = src.p1y * r1 + src.p2y * r2;
            redirect(src.nx, src.ny, p, st);
            if (m.phasor) {
                p.bin = 1;
                p.freq = 1;
                p.vel = 1000;
                p.pol = POL_LA;
                p.dx = src.nx;
                p.dy = src.ny;
            }
        }
        const int64_t step = (int64_t)(birth / m.step_time);
        int32_t cell = p.cell;
        if (step >= m.N) {  // the reference would read step_times_[step] out of bounds
            cell = -1;
            not_started = 1;
            fate_write(fate, id, step, FATE_NOT_STARTED, p.px, p.py);
        } else {
            key = (int32_t)step;
        }
        birth_key[id] = key;
        pool.pos[id] = make_double2(p.px, p.py);
        pool.dir[id] = make_double2(p.dx, p.dy);
        pool.fv[id] = make_double2(p.freq, p.vel);
        pool.aux[id] = PoolAux{ birth, cell, cl.sensor };
        pool.ms[id] = make_uint2((uint32_t)p.bin | (p.pol == POL_TA ? META_TA : 0u) | (p.sign < 0 ? META_NEG : 0u)
                                     | ((uint32_t)src.stream << META_STREAM_SHIFT),
            sid);
    }
    // birth-step histogram; most waves of a cell-origin source share one key (step 0)
    const int lane = threadIdx.x % WAVE;
    const int key0 = __shfl(key, 0, WAVE);
    const uint64_t same = __ballot(key == key0);
    if (same == ~0ull) {
        if (lane == 0 && key0 >= 0) { atomicAdd(&birth_hist[key0], (uint32_t)WAVE); }
    } else if (key >= 0) {
        atomicAdd(&birth_hist[key], 1u);
    }
    const uint32_t ns = wave_sum(not_started);
    if (lane == 0 && ns) { atomicAdd(&counters[PSIM_CNT_NOT_STARTED], (unsigned long long)ns); }
}

// Exclusive scan of the birth-step histogram (one block): offset[k] = phonons born before
// step k, offset[N] = all that start; cursor starts as a copy of offset.
__global__ void __launch_bounds__(1024) birth_scan_kernel(const uint32_t* __restrict__ hist, int64_t N,
    uint32_t* __restrict__ offset, uint32_t* __restrict__ cursor) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (N + 1023) / 1024;
    const int64_t lo = t * per, hi = std::min<int64_t>(N, lo + per);
    uint32_t sum = 0;
    for (int64_t k = lo; k < hi; ++k) { sum += hist[k]; }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan of the chunk sums
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (int64_t k = lo; k < hi; ++k) {
        offset[k] = run;
        cursor[k] = run;
        run += hist[k];
    }
    if (t == 1023) { offset[N] = part[1023]; }
}

// Places every phonon id in born_order by birth step (order inside one step is arbitrary).
__global__ void __launch_bounds__(BLOCK) birth_scatter_kernel(const int32_t* __restrict__ birth_key, int64_t total,
    uint32_t* __restrict__ cursor, uint32_t* __restrict__ born_order) {
    const int64_t id = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int32_t key = (id < total) ? birth_key[id] : -1;
    const int lane = threadIdx.x % WAVE;
    const int key0 = __shfl(key, 0, WAVE);
    const uint64_t same = __ballot(key == key0);
    if (same == ~0ull) {
        uint32_t base = 0;
        if (lane == 0 && key0 >= 0) { base = atomicAdd(&cursor[key0], (uint32_t)WAVE); }
        base = __shfl(base, 0, WAVE);
        if (key0 >= 0) { born_order[base + lane] = (uint32_t)id; }
    } else if (key >= 0) {
        born_order[atomicAdd(&cursor[key], 1u)] = (uint32_t)id;
    }
}

// State of the phonons that are alive between two launch windows, densely packed in the order
// the previous launch paused them, with the phonon id (the key of its Philox stream).  A launch
// reads its started phonons from one buffer in work-queue order and appends its survivors to the
// other, so the phonons that neighbouring lanes and waves fetch or park share cache lines.  Seven
// arrays of 16-byte records (112 B per phonon): a lane moves a phonon in seven vector loads or
// stores, and a launch holds seven pointers per buffer.
struct __attribute__((aligned(16))) LiveAux {
    double rate2;  // third relaxation rate of the last scattering-time draw
    int32_t cell;  // -1: a hole (a reserved slot nobody filled)
    int32_t sensor;  // sensor of the cell
};

struct __attribute__((aligned(16))) LiveTag {
    uint32_t meta;
    uint32_t id;
    uint32_t sid;  // id within the phonon's random stream
    int32_t step;  // measurement step (past the window's end after a long flight)
};

struct Live {
    double2* __restrict__ pos;  // px, py
    double2* __restrict__ dir;  // dx, dy
    double2* __restrict__ fv;  // freq, vel
    double2* __restrict__ at;  // age, time to the next scatter
    double2* __restrict__ r01;  // first two relaxation rates of the last scattering-time draw
    LiveAux* __restrict__ aux;
    LiveTag* __restrict__ tag;
};

// Work of one transport launch: the started phonons (Live buffer `in`, *n_in entries) followed by
// the phonons born in the window (born_order, read from the emission pool).  Survivors go to
// `out` in slots that waves reserve OUT_CLAIM at a time from *n_out; a wave marks the slots it
// reserved but did not fill as holes (cell -1), which the cell sort drops.
struct Lists {
    Live in;  // phonons alive at step k0 (started)
    Live out;  // survivors at step k1, with holes
    const uint32_t* __restrict__ n_in;
    uint32_t* __restrict__ n_out;
    uint32_t* __restrict__ queue;  // [windows]: work-queue cursor of launch w
    const uint32_t* __restrict__ born_order;  // ids sorted by birth step
    const uint32_t* __restrict__ born_offset;  // [N + 1]
    int32_t window;
};

// Tally path of a transport launch, fixed at compile time so that each variant carries only its
// own code: none (a window before the first recorded step), LDS-privatised window partials, or
// FP64 atomics into the global tallies (too many sensors for LDS).
constexpr int TALLY_NONE = 0, TALLY_LDS = 1, TALLY_GLOBAL = 2;

// Waves per SIMD the compiler must leave room for.  Without tally code the kernel fits 128 VGPRs
// with a few values spilled to scratch, and four waves per SIMD hide more of its load latency;
// with the tally code, 128 VGPRs spill in the tally path itself and three waves (~143 VGPRs) are
// faster.
#ifndef PSIM_UNTALLIED_WAVES_PER_EU
#define PSIM_UNTALLIED_WAVES_PER_EU 4
#endif
#ifndef PSIM_TRANSPORT_WAVES_PER_EU
#define PSIM_TRANSPORT_WAVES_PER_EU 2
#endif

// Lane states of the scheduled event loop.
enum : int { S_FETCH = 0, S_TTS = 1, S_NI = 2, S_COLL = 3, S_SCAT = 4, S_END = 5, S_EXIT = 6, S_PAUSED = 7, S_DEAD = 8 };

// Advances the listed phonons across measurement steps [k0, k1).  Each lane carries one phonon
// through ModelSimulator::simulatePhonon (psim/src/modelSimulator.cpp:139-198), pausing right
// after the measurement that brings its step to k1, then takes the next phonon of the launch's
// work queue.  The loop is split into stages -- fetch, surface handling after a hit, scatter and
// the scattering-time draw, one nextImpact -- and every iteration of the wave runs them in that
// order, each for the lanes waiting in it, so lanes that finish one stage join the next in the
// same iteration.  Each lane still performs exactly the reference's sequence of operations for
// its phonon, so trajectories are unchanged.
// TALLY: the tally path of this launch (TALLY_NONE, TALLY_LDS, TALLY_GLOBAL).
// MULTI: some source has a stream index or id base (a batch of models).  Otherwise a phonon's
// id within its stream is its id and the stream word is the bare tag, so the single-model
// variant neither loads nor stores the per-phonon stream id and keeps that word a constant;
// both variants draw the same numbers for such a model.
// NT: threads per block (128, 256 or 512), a launch shape the autotuner chooses.
// FT: the flux tallies' element type, double or float (FP32 with Kahan-compensated sums).
template <int TALLY, bool MULTI, int NT, typename FT>
__global__ void __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(TALLY == TALLY_NONE ? PSIM_UNTALLIED_WAVES_PER_EU : PSIM_TRANSPORT_WAVES_PER_EU)))
transport_kernel(Model m, Pool pool, Lists lists, int32_t k0, int32_t k1, const uint64_t* __restrict__ seed_p,
    Tally tally, unsigned long long* __restrict__ counters, Fate fate) {
    const uint64_t seed = *seed_p;
    extern __shared__ unsigned char lds_raw[];
    const int S = m.n_sensors;
    // LDS window tallies as difference arrays over the window's steps: a flight that passes
    // boundaries a..b adds its count and flux at row a and subtracts them at row b + 1, so its
    // tallies cost six atomics whatever its length; rows 0..W-1 plus one for b = k1.
    const int WS = (TALLY == TALLY_LDS && tally.rec) ? tally.W * S : 0;
    const int WS1 = WS ? WS + S : 0;
    FT* lfx = reinterpret_cast<FT*>(lds_raw);
    FT* lfy = lfx + WS1;
    int32_t* lcnt = reinterpret_cast<int32_t*>(lfy + WS1);
    for (int i = threadIdx.x; i < WS1; i += NT) {
        lfx[i] = FT(0);
        lfy[i] = FT(0);
        lcnt[i] = 0;
    }
    __syncthreads();

    LaneCounters c = { 0, 0, 0, 0, 0, 0, 0, 0 };
    const int lane = threadIdx.x % WAVE;
    const uint64_t lane_below = (1ull << lane) - 1ull;
    const int64_t n_live = *lists.n_in;
    const int64_t b0 = lists.born_offset[k0];
    const int64_t n_items = n_live + (int64_t)lists.born_offset[k1] - b0;
    uint32_t* __restrict__ queue = lists.queue + lists.window;

    int state = S_FETCH;
    uint32_t slot = 0;
    Ph p;
    p.px = p.py = p.dx = p.dy = p.freq = p.vel = 0.;
    p.cell = p.bin = p.pol = 0;
    p.sign = 1;
    p.step = 0;
    double age = 0., tts = 0., ttm = 0.;
    double relax[3] = { 0., 0., 0. };
    int events = 0;
    double drift_time = 0., drifted = 0., t_hit = 0.;
    int sensor0 = 0, collisions = 0;
    int cur_sensor = 0;  // sensor of p.cell
    psim_stream st;
    psim_stream_init(&st, seed, PSIM_TAG_TRANSPORT, 0, 0);
    // Only untallied launches hold phonons below `stop`: every window that records starts at step
    // adj - 1 = stop or later.
    const int32_t stop = TALLY == TALLY_NONE ? m.stop : 0;
    const double steps_per_time = (double)m.N / m.sim_time;
    // A flight ends at boundary `stop` (steady state, while the step lies below it), from there on at
    // every cap_len-th boundary after it (the model's flight cap), and at boundary N.  `cap` is the
    // current flight's cap, set when a phonon is fetched and when a flight reaches its cap.
    int32_t cap_len = 1, cap = 0;
    auto cap_for = [&](int32_t step) -> int32_t {
        if (step < stop) { return stop; }
        const int32_t c2 = m.stop + cap_len * ((step - m.stop) / cap_len + 1);
        return c2 < m.N ? c2 : m.N;
    };
    // Work-queue items the wave has claimed but not handed out yet (wave-uniform): a refill takes
    // them in order and claims QUEUE_CLAIM more with one atomic when they run out.
    uint32_t res_next = 0, res_end = 0;
    uint32_t out_next = 0, out_end = 0;  // survivor slots reserved by the wave, not filled yet

    // Top of simulatePhonon's loop up to the nextImpact call (modelSimulator.cpp:154-167): the
    // stall bound, S_TTS when a scattering time is due (drawn once, then drift_setup), else
    // drift_setup directly.
    auto drift_setup = [&]() {
        ttm = (double)cap * m.sim_time / (double)m.N - age;  // time to the flight's cap
        drift_time = (ttm < tts) ? ttm : tts;
        sensor0 = cur_sensor;
        drifted = 0.;
        collisions = 0;
        state = S_NI;
    };
    auto begin_event = [&]() {
        if (++events > PSIM_MAX_EVENTS_PER_STEP) {
            c.stalled++;
            state = S_DEAD + FATE_STALLED * 8;
            return;
        }
        if (tts <= 0.) {
            state = S_TTS;
            return;
        }
        drift_setup();
    };
    // Sensor::updateHeatParams (sensor.cpp:43-52) for the boundaries a..b a flight passed (a >= adj),
    // with the sensor and velocity of that moment: the recorded ones (index below R) as one segment
    // of a difference array, the part in (k0, k1] in the window's LDS, the rest (past k1, after a
    // flight that crossed the window's end, or all of it in TALLY_GLOBAL) in HBM.
    auto tally_range = [&](int32_t a, int32_t b) {
        if constexpr (TALLY != TALLY_NONE) {
            const int32_t last = b < m.adj + m.R - 1 ? b : m.adj + m.R - 1;
            if (a > last || !tally.rec) { return; }
            const double sg = (double)p.sign;
            const double vx = p.dx * p.vel * sg, vy = p.dy * p.vel * sg;
            int32_t first = a;
            if constexpr (TALLY == TALLY_LDS) {
                const int32_t in_last = last < k1 ? last : k1;
                if (a <= in_last) {
                    const int e0 = (a - k0 - 1) * S + cur_sensor, e1 = (in_last - k0) * S + cur_sensor;
                    atomicAdd(&lcnt[e0], p.sign);
                    atomicAdd(&lcnt[e1], -p.sign);
                    atomicAdd(&lfx[e0], (FT)vx);
                    atomicAdd(&lfx[e1], (FT)-vx);
                    atomicAdd(&lfy[e0], (FT)vy);
                    atomicAdd(&lfy[e1], (FT)-vy);
                    first = in_last + 1;
                }
            }
            if (first <= last) {
                const int64_t row = (int64_t)cur_sensor * (m.R + 1) - m.adj;
                const int64_t t0 = row + first, t1 = row + last + 1;
                atomicAdd(reinterpret_cast<unsigned long long*>(&tally.dcnt[t0]), (unsigned long long)(int64_t)p.sign);
                atomicAdd(reinterpret_cast<unsigned long long*>(&tally.dcnt[t1]), (unsigned long long)(int64_t)-p.sign);
                FT* dfx = static_cast<FT*>(tally.dfx);
                FT* dfy = static_cast<FT*>(tally.dfy);
                atomicAdd(&dfx[t0], (FT)vx);
                atomicAdd(&dfx[t1], (FT)-vx);
                atomicAdd(&dfy[t0], (FT)vy);
                atomicAdd(&dfy[t1], (FT)-vy);
            }
            c.tallies += (uint32_t)(last - a + 1);
        }
    };
    // The measurement events of the boundaries a flight passed up to time t
    // (modelSimulator.cpp:182-188): the step moves to the last boundary k < cap with
    // k * sim_time / N <= t (never backwards) and the stream to that step; every boundary passed
    // is a phonon-step and, from adj on, a tally with the sensor and velocity of that moment.
    // Returns whether the step moved.
    auto advance_step = [&](double t, int32_t cap) -> bool {
        int32_t k = p.step;
        const int32_t est = (int32_t)(t * steps_per_time) - 1;  // a whole step below t: at or below the answer
        if (est > k) { k = est < cap - 1 ? est : cap - 1; }
        while (k + 1 < cap && (double)(k + 1) * m.sim_time / (double)m.N <= t) { ++k; }
        if (k == p.step) { return false; }
        if constexpr (TALLY != TALLY_NONE) { tally_range(p.step + 1 > m.adj ? p.step + 1 : m.adj, k); }
        c.steps += (uint32_t)(k - p.step);
        p.step = k;
        events = 0;
        psim_stream_set_step(&st, (uint32_t)k);
        return true;
    };
    // The rest of simulatePhonon's loop body once handleImpacts returned `result`, with the
    // measurement events of the boundaries the flight passed.  A flight may carry the phonon past
    // k1: it pauses right after a boundary crossing that left its stream at the start of a step
    // with no draw made (so a resumed phonon continues exactly), at the latest at boundary
    // `stop`, where every window before the recorded steps ends.
    bool relocated = false;  // the flight ended frozen after a relocation (same sensor, same scattering time)
    double end_arg = 0.;  // S_END: handleImpacts' result, applied by end_drift
    auto end_drift = [&](double result) {
        if (cur_sensor != sensor0) { drift_time = result; }
        const double factor = p.vel * (drift_time - result);  // Phonon::drift
        p.px += p.dx * factor;
        p.py += p.dy * factor;
        age += drift_time;
        ttm -= drift_time;
        tts -= drift_time;
        const bool advanced = advance_step(age, cap);
        if (ttm == 0. && cap == m.N) {  // boundary N: the run is over (step_times_ has N entries)
            c.steps++;
            c.expired++;
            p.step = m.N;
            state = S_DEAD + FATE_EXPIRED * 8;
            return;
        }
        if (ttm == 0.) {  // the flight reached boundary `cap`
            c.steps++;
            if constexpr (TALLY != TALLY_NONE) {
                if (cap >= m.adj) { tally_range(cap, cap); }
            }
            p.step = cap;
            events = 0;
            psim_stream_set_step(&st, (uint32_t)cap);
            cap = cap_for(p.step);
            if (p.step >= k1) {
                state = S_PAUSED;
                return;
            }
        } else {
            if (!m.phasor && tts == 0.) {
                state = S_SCAT;
                return;
            }
            if (!relocated) { tts = 0.; }  // a new sensor: a new scattering time
            if (advanced && p.step >= k1) {
                state = S_PAUSED;
                return;
            }
        }
        begin_event();
    };

    // The stages as functions of the lanes waiting in them (each acts only on its own lanes).
    // Refill: the next phonons of the work queue for the lanes in S_FETCH (mask m_fetch).
    auto stage_fetch = [&](uint64_t m_fetch) {
        const uint32_t n_fetch = (uint32_t)__popcll(m_fetch);
        const uint32_t avail = res_end - res_next;
        const uint32_t rank = (uint32_t)__popcll(m_fetch & lane_below);
        int64_t item;
        if (n_fetch <= avail) {
            item = (int64_t)res_next + rank;
            res_next += n_fetch;
        } else {
            const uint32_t need = n_fetch - avail;
            const uint32_t take = need > QUEUE_CLAIM ? need : QUEUE_CLAIM;
            const int leader = __ffsll((unsigned long long)m_fetch) - 1;
            uint32_t base = 0;
            if (lane == leader) { base = atomicAdd(queue, take); }
            base = __shfl(base, leader, WAVE);
            item = rank < avail ? (int64_t)res_next + rank : (int64_t)base + (rank - avail);
            res_next = base + need;
            res_end = base + take;
        }
        if (state == S_FETCH) {
            if (item >= n_items) {
                state = S_EXIT;
            } else {
                // started phonons from the dense buffer, phonons born in the window from the pool
                const bool started = item < n_live;
                const Live& L = lists.in;
                uint32_t meta, sid;
                double2 pos, dir, fv;
                if (started) {
                    const LiveTag tg = L.tag[item];
                    slot = tg.id;
                    meta = tg.meta;
                    sid = tg.sid;
                    p.step = tg.step;
                    pos = L.pos[item];
                    dir = L.dir[item];
                    fv = L.fv[item];
                    const double2 at = L.at[item];
                    age = at.x;
                    tts = at.y;
                    const double2 r = L.r01[item];  // rates of the last scattering-time draw
                    const LiveAux ax = L.aux[item];
                    relax[0] = r.x;
                    relax[1] = r.y;
                    relax[2] = ax.rate2;
                    p.cell = ax.cell;
                    cur_sensor = ax.sensor;
                } else {  // a new phonon has no scattering time and no rates yet
                    slot = lists.born_order[b0 + (item - n_live)];
                    const uint2 ms = pool.ms[slot];
                    meta = ms.x;
                    sid = ms.y;
                    pos = pool.pos[slot];
                    dir = pool.dir[slot];
                    fv = pool.fv[slot];
                    const PoolAux ax = pool.aux[slot];
                    age = ax.age;
                    tts = 0.;
                    p.cell = ax.cell;
                    cur_sensor = ax.sensor;
                    p.step = (int32_t)(age / m.step_time);
                }
                if (!MULTI) { sid = slot; }
                cap_len = m.sensors[cur_sensor].flight_cap;
                cap = cap_for(p.step);
                p.px = pos.x;
                p.py = pos.y;
                p.dx = dir.x;
                p.dy = dir.y;
                p.freq = fv.x;
                p.vel = fv.y;
                p.bin = (int)(meta & META_BIN_MASK);
                p.pol = (meta & META_TA) ? POL_TA : POL_LA;
                p.sign = (meta & META_NEG) ? -1 : 1;
                ttm = 0.;
                events = 0;
                // the stream keeps (stream index, id within it) in its counter words c3 and c2
                psim_stream_init(&st, seed,
                    MULTI ? psim_stream_tag(PSIM_TAG_TRANSPORT, meta >> META_STREAM_SHIFT) : PSIM_TAG_TRANSPORT, sid,
                    (uint32_t)p.step);
                if (p.step >= k1) {  // a long flight already carried it past this window
                    state = S_PAUSED;
                } else {
                    begin_event();
                }
            }
        }
    };
    // One nextImpact of handleImpacts.
    auto stage_ni = [&]() {
        if (state == S_NI) {  // one nextImpact of handleImpacts
            if (find_impact(m, p, drift_time - drifted, t_hit, c)) {
                state = S_COLL;
            } else {
                end_arg = drifted;
                state = S_END;
            }
        }
    };
    // Surface handling after a hit, then handleImpacts' bookkeeping.
    auto stage_coll = [&]() {
        if (state == S_COLL) {  // surface handling, then handleImpacts' bookkeeping
            // the boundaries passed before the impact see the old direction; the surface sees
            // the step of the moment of impact
            // the reference counts collisions per drift, and a drift ends at a boundary
            if (advance_step(age + (drifted + t_hit), cap)) { collisions = 0; }
            handle_surface_collision(m, p, p.px, p.py, st);
            if (p.cell < 0) {
                c.absorbed++;
                state = S_DEAD + FATE_ABSORBED * 8;
            } else {
                cur_sensor = m.cells[p.cell].sensor;
                drifted += t_hit;
                if (++collisions > MAX_COLLISIONS) {  // Phonon::setRandPoint -> Triangle::getRandPoint
                    double r1, r2;
                    psim_urand2(&st, &r1, &r2);
                    const psim_cell& cl = m.cells[p.cell];
                    if (r1 + r2 > 1.) {
                        r1 = 1 - r1;
                        r2 = 1 - r2;
                    }
                    p.px = cl.x1 + (cl.x2 - cl.x1) * r1 + (cl.x3 - cl.x1) * r2;
                    p.py = cl.y1 + (cl.y2 - cl.y1) * r1 + (cl.y3 - cl.y1) * r2;
                    c.relocations++;
                    // the phonon stays at the new point for the rest of the reference's drift:
                    // up to the next boundary, or the flight's end if that comes first
                    const double frozen = (double)(p.step + 1) * m.sim_time / (double)m.N - age;
                    if (frozen < drift_time) { drift_time = frozen; }
                    relocated = true;
                    end_arg = drift_time;
                    state = S_END;
                } else if (cur_sensor != sensor0) {
                    end_arg = drifted;
                    state = S_END;
                } else {
                    state = S_NI;
                }
            }
        }
    };
    // The rest of simulatePhonon's loop body once handleImpacts has returned (end_drift), for the
    // lanes whose drift ended in this iteration's impact search or collision.
    auto stage_end = [&]() {
        if (state == S_END) {
            end_drift(end_arg);
            relocated = false;
        }
    };
    // Scatter (lanes in S_SCAT), then the scattering-time draw (lanes in S_SCAT or S_TTS).
    auto stage_scat = [&]() {
        // One block for every lane of the stage: a scattering lane takes the scatter's draw
        // and its next scattering-time draw from it, a lane that only draws a scattering time
        // takes the first double.
        double u_a = 0., u_b = 0.;
        if (state == S_SCAT || state == S_TTS) { psim_urand2(&st, &u_a, &u_b); }
        if (state == S_SCAT) {
            c.scatters++;
            scatter(m, p, relax, u_a, st);
            u_a = u_b;
            if (++events > PSIM_MAX_EVENTS_PER_STEP) {  // begin_event's stall bound
                c.stalled++;
                state = S_DEAD + FATE_STALLED * 8;
            } else {
                state = S_TTS;  // tts is 0 here, so begin_event would ask for a draw
            }
        }
        if (state == S_TTS) {
            relax_rates(m, cur_sensor, p, relax);
            tts = SCALING_FACTOR * -psim_log(u_a) / (0. + relax[0] + relax[1] + relax[2]);
            drift_setup();
        }
    };

    for (;;) {
        // Every iteration runs the stages in the order a phonon passes through them -- refill, one
        // impact search, the collision after a hit, the end of a drift, scatter and the
        // scattering-time draw -- each for the lanes waiting in it, so that lanes that leave one
        // stage join the next one in the same iteration and each stage's code runs once per
        // iteration for all of them: a wave alternating between impact searches and scatters
        // runs each with most of its lanes instead of one of them with the lanes that wait for
        // it.  Branches are on wave-uniform ballots; a stage with no waiting lane is skipped.
        if (!__ballot(state < S_EXIT)) { break; }  // every lane has exited
        const uint64_t m_fetch = __ballot(state == S_FETCH);
        if (m_fetch) { stage_fetch(m_fetch); }
        if (__ballot(state == S_NI)) { stage_ni(); }
        if (__ballot(state == S_COLL)) { stage_coll(); }
        if (__ballot(state == S_END)) { stage_end(); }
        if (__ballot(state == S_SCAT || state == S_TTS)) { stage_scat(); }

        // Phonons that paused this iteration go to the next window's buffer, in slots the wave
        // has reserved; the dead leave.
        const bool survive = state == S_PAUSED;
        const uint64_t m_surv = __ballot(survive);
        if (m_surv) {
            const uint32_t n_surv = (uint32_t)__popcll(m_surv);
            const uint32_t avail = out_end - out_next;
            const uint32_t rank = (uint32_t)__popcll(m_surv & lane_below);
            uint32_t k;
            if (n_surv <= avail) {
                k = out_next + rank;
                out_next += n_surv;
            } else {
                const int leader = __ffsll((unsigned long long)m_surv) - 1;
                uint32_t base = 0;
                if (lane == leader) { base = atomicAdd(lists.n_out, OUT_CLAIM); }
                base = __shfl(base, leader, WAVE);
                k = rank < avail ? out_next + rank : base + (rank - avail);
                out_next = base + (n_surv - avail);
                out_end = base + OUT_CLAIM;
            }
            if (survive) {
                const Live& O = lists.out;
                O.pos[k] = make_double2(p.px, p.py);
                O.dir[k] = make_double2(p.dx, p.dy);
                O.fv[k] = make_double2(p.freq, p.vel);
                O.at[k] = make_double2(age, tts);
                O.r01[k] = make_double2(relax[0], relax[1]);
                O.aux[k] = LiveAux{ relax[2], p.cell, cur_sensor };
                O.tag[k] = LiveTag{ (uint32_t)p.bin | (p.pol == POL_TA ? META_TA : 0u) | (p.sign < 0 ? META_NEG : 0u)
                                        | META_STARTED | (MULTI ? (st.c3 >> 8) << META_STREAM_SHIFT : 0u),
                    slot, MULTI ? st.c2 : slot, p.step };
                state = S_FETCH;
            }
        }
        if (state >= S_DEAD) {
            const int cause = (state - S_DEAD) / 8;
            fate_write(fate, slot, cause == FATE_EXPIRED ? p.step - 1 : p.step, cause, p.px, p.py);
            state = S_FETCH;
        }
    }

    // Reserved survivor slots the wave did not fill become holes.
    for (uint32_t k = out_next + (uint32_t)lane; k < out_end; k += WAVE) { lists.out.aux[k].cell = -1; }

    // Counters: wave sums, then one global atomic per wave and counter.
    const uint32_t v[8] = { c.steps, c.scatters, c.impacts, c.relocations, c.absorbed, c.expired, c.tallies,
        c.stalled };
    const int slots[8] = { PSIM_CNT_PHONON_STEPS, PSIM_CNT_SCATTERS, PSIM_CNT_IMPACTS, PSIM_CNT_RELOCATIONS,
        PSIM_CNT_ABSORBED, PSIM_CNT_EXPIRED, PSIM_CNT_TALLIES, PSIM_CNT_STALLED };
    for (int i = 0; i < 8; ++i) {
        const unsigned long long sum = wave_sum64(v[i]);
        if (lane == 0 && sum) { atomicAdd(&counters[slots[i]], sum); }
    }
    if (WS) {  // the difference arrays summed over the window's steps, per sensor, into the partials
        __syncthreads();
        const int64_t base = (int64_t)blockIdx.x * WS;
        FT* pfx = static_cast<FT*>(tally.pfx);
        FT* pfy = static_cast<FT*>(tally.pfy);
        for (int sn = threadIdx.x; sn < S; sn += NT) {
            int32_t cnt = 0;
            FT fx = FT(0), fy = FT(0);
            float cx = 0.f, cy = 0.f;  // FP32: the running sums' compensations
            for (int e = sn; e < WS; e += S) {
                cnt += lcnt[e];
                if constexpr (std::is_same<FT, float>::value) {
                    kahan_add(fx, cx, lfx[e]);
                    kahan_add(fy, cy, lfy[e]);
                    tally.pcx[base + e] = cx;
                    tally.pcy[base + e] = cy;
                } else {
                    fx += lfx[e];
                    fy += lfy[e];
                }
                tally.pcnt[base + e] = cnt;
                pfx[base + e] = fx;
                pfy[base + e] = fy;
            }
        }
    }
}

// ---- sort of the live buffer by cell ---------------------------------------------------------
// The phonons a wave works on should sit in few cells, so that the edge records and tables they
// read stay in L1: phonon ids are emitted cell by cell, but a window appends its survivors in
// the order they pause.  Between windows a counting sort moves them back into cell order (key =
// cell >> shift, at most SORT_KEYS keys): per-block LDS histograms, a scan, then a scatter in which
// every block reserves its share of each key with one atomic.  Order inside a key is arbitrary;
// no result depends on the order in which phonons are processed.
constexpr int SORT_KEYS = 4096;
constexpr int SORT_BLOCK = 1024;

__device__ __forceinline__ void sort_range(uint32_t n, uint32_t& lo, uint32_t& hi) {
    const uint64_t per = ((uint64_t)n + gridDim.x - 1) / gridDim.x;
    lo = (uint32_t)std::min<uint64_t>(n, (uint64_t)blockIdx.x * per);
    hi = (uint32_t)std::min<uint64_t>(n, (uint64_t)lo + per);
}

__global__ void __launch_bounds__(SORT_BLOCK) cell_hist_kernel(const LiveAux* __restrict__ aux,
    const uint32_t* __restrict__ n_ptr, int shift, int n_keys, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[SORT_KEYS];
    for (int k = threadIdx.x; k < n_keys; k += SORT_BLOCK) { h[k] = 0; }
    __syncthreads();
    uint32_t lo, hi;
    sort_range(*n_ptr, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_BLOCK) {
        const int32_t cl = aux[i].cell;
        if (cl >= 0) { atomicAdd(&h[(uint32_t)cl >> shift], 1u); }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n_keys; k += SORT_BLOCK) {
        if (h[k]) { atomicAdd(&hist[k], h[k]); }
    }
}

// Exclusive scan of the key counts in place (one block, n_keys <= SORT_KEYS); *total = their sum.
__global__ void __launch_bounds__(1024) cell_scan_kernel(uint32_t* __restrict__ hist, int n_keys,
    uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    constexpr int PER = SORT_KEYS / 1024;
    const int t = threadIdx.x;
    uint32_t v[PER];
    uint32_t sum = 0;
    for (int q = 0; q < PER; ++q) {
        const int k = t * PER + q;
        v[q] = (k < n_keys) ? hist[k] : 0u;
        sum += v[q];
    }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t x = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (int q = 0; q < PER; ++q) {
        const int k = t * PER + q;
        if (k < n_keys) { hist[k] = run; }
        run += v[q];
    }
    if (t == 1023) { *total = part[1023]; }
}

__global__ void __launch_bounds__(SORT_BLOCK) cell_scatter_kernel(Live src, Live dst,
    const uint32_t* __restrict__ n_ptr, int shift, int n_keys, uint32_t* __restrict__ cursor) {
    __shared__ uint32_t cnt[SORT_KEYS];
    __shared__ uint32_t base[SORT_KEYS];
    for (int k = threadIdx.x; k < n_keys; k += SORT_BLOCK) { cnt[k] = 0; }
    __syncthreads();
    uint32_t lo, hi;
    sort_range(*n_ptr, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_BLOCK) {
        const int32_t cl = src.aux[i].cell;
        if (cl >= 0) { atomicAdd(&cnt[(uint32_t)cl >> shift], 1u); }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n_keys; k += SORT_BLOCK) {
        base[k] = cnt[k] ? atomicAdd(&cursor[k], cnt[k]) : 0u;
        cnt[k] = 0;
    }
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_BLOCK) {
        const LiveAux ax = src.aux[i];
        if (ax.cell < 0) { continue; }  // hole
        const uint32_t key = (uint32_t)ax.cell >> shift;
        const uint32_t k = base[key] + atomicAdd(&cnt[key], 1u);
        dst.pos[k] = src.pos[i];
        dst.dir[k] = src.dir[i];
        dst.fv[k] = src.fv[i];
        dst.at[k] = src.at[i];
        dst.r01[k] = src.r01[i];
        dst.aux[k] = ax;
        dst.tag[k] = src.tag[i];
    }
}

// Adds the HBM difference arrays [S][R + 1] of a run, summed along the recorded steps, into the
// tallies [S][R] (one thread per sensor, steps in order; FP32 arrays are summed with Kahan
// compensation and added as sum - compensation).
template <typename FT>
__global__ void __launch_bounds__(BLOCK) diff_prefix_kernel(const int64_t* __restrict__ dcnt,
    const FT* __restrict__ dfx, const FT* __restrict__ dfy, int S, int64_t R, int64_t* __restrict__ counts,
    double* __restrict__ flux) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= S) { return; }
    int64_t cnt = 0;
    FT fx = FT(0), fy = FT(0);
    float cx = 0.f, cy = 0.f;
    for (int64_t r = 0; r < R; ++r) {
        const int64_t d = (int64_t)s * (R + 1) + r;
        cnt += dcnt[d];
        const int64_t t = (int64_t)s * R + r;
        counts[t] += cnt;
        if constexpr (std::is_same<FT, float>::value) {
            kahan_add(fx, cx, dfx[d]);
            kahan_add(fy, cy, dfy[d]);
            flux[2 * t] += (double)fx - (double)cx;
            flux[2 * t + 1] += (double)fy - (double)cy;
        } else {
            fx += dfx[d];
            fy += dfy[d];
            flux[2 * t] += fx;
            flux[2 * t + 1] += fy;
        }
    }
}

// Clears n 32-bit words (used instead of memset nodes, so that a run can be captured in a graph).
__global__ void __launch_bounds__(BLOCK) zero_words_kernel(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) { p[i] = 0u; }
}

// Sums the block partials of one window into the tallies: entry e = w * S + s of window step
// k0 + 1 + w goes to tally index (k0 + 1 + w) - adj of sensor s.
// Stores the run's seed where the kernels read it; the value travels as a kernel argument, so an
// asynchronous run needs no host buffer that outlives the call.
__global__ void seed_kernel(uint64_t* __restrict__ seed_p, uint64_t seed) {
    if (threadIdx.x == 0) { *seed_p = seed; }
}

// Sums the per-block window partials into the tallies.  Block (x, y) adds transport blocks
// [y * per, (y + 1) * per) of window entries [x * BLOCK, (x + 1) * BLOCK) (e = step-in-window * S +
// sensor) and adds its sums with atomics: enough blocks to fill the chip even for a window of a
// few sensors.  Counts stay exact (integer atomics); flux sums in another order, as the LDS
// atomics of the transport kernel already do.
// FP32 partials (FT float) are summed with Kahan compensation, each block's partial entering as its
// sum and its negated compensation, and added as sum - compensation.
template <typename FT>
__global__ void __launch_bounds__(BLOCK) reduce_kernel(const int32_t* __restrict__ pcnt,
    const FT* __restrict__ pfx, const FT* __restrict__ pfy, const float* __restrict__ pcx, const float* __restrict__ pcy,
    int n_blocks, int per, int WS, int S, int64_t k0, int64_t adj, int64_t R, int64_t* __restrict__ counts,
    double* __restrict__ flux) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= WS) { return; }
    const int w = e / S, s = e % S;
    const int64_t idx = k0 + 1 + w - adj;
    if (idx < 0 || idx >= R) { return; }
    const int g0 = blockIdx.y * per, g1 = min(n_blocks, g0 + per);
    int64_t cnt = 0;
    double fx = 0., fy = 0.;
    if constexpr (std::is_same<FT, float>::value) {
        float sx = 0.f, cx = 0.f, sy = 0.f, cy = 0.f;
        for (int g = g0; g < g1; ++g) {
            const int64_t k = (int64_t)g * WS + e;
            cnt += pcnt[k];
            kahan_add(sx, cx, pfx[k]);
            kahan_add(sx, cx, -pcx[k]);
            kahan_add(sy, cy, pfy[k]);
            kahan_add(sy, cy, -pcy[k]);
        }
        fx = (double)sx - (double)cx;
        fy = (double)sy - (double)cy;
    } else {
        for (int g = g0; g < g1; ++g) {
            const int64_t k = (int64_t)g * WS + e;
            cnt += pcnt[k];
            fx += pfx[k];
            fy += pfy[k];
        }
    }
    const int64_t t = (int64_t)s * R + idx;
    if (cnt != 0) { atomicAdd(reinterpret_cast<unsigned long long*>(counts + t), (unsigned long long)cnt); }
    if (fx != 0.) { atomicAdd(flux + 2 * t, fx); }
    if (fy != 0.) { atomicAdd(flux + 2 * t + 1, fy); }
}

template <typename T>
T* dalloc(size_t n) {
    T* p = nullptr;
    if (n == 0) { n = 1; }
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { return nullptr; }
    return p;
}

}  // namespace

struct psim_gpu_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = { nullptr, nullptr, nullptr, nullptr };
    std::string err;
    // model
    bool uploaded = false;
    bool multi_stream = false;  // some source has a stream index or id base (a batch)
    psim_run_params params{};
    Model m{};
    std::vector<void*> model_allocs;
    int32_t n_sensors = 0;
    int32_t n_cells = 0;
    // run buffers
    Pool pool{};
    int64_t pool_cap = 0;
    int64_t* d_counts = nullptr;
    double* d_flux = nullptr;
    // HBM difference arrays [S][R + 1] of the tally segments that do not go to LDS
    int64_t* d_dcnt = nullptr;
    double* d_dfx = nullptr;
    double* d_dfy = nullptr;
    unsigned long long* d_counters = nullptr;
    int32_t* d_pcnt = nullptr;
    double* d_pfx = nullptr;  // window partials; float in the FP32 tally precision
    double* d_pfy = nullptr;
    float* d_pcx = nullptr;  // FP32 tally precision: Kahan compensations of the partials
    float* d_pcy = nullptr;
    int64_t partial_cap = 0;
    bool fate_enabled = false;
    Fate fate{ nullptr, nullptr, nullptr };
    int64_t fate_cap = 0;
    // birth ordering and live lists
    int32_t* d_birth_key = nullptr;
    uint32_t* d_born_order = nullptr;
    Live live[2] = {};  // dense state of the phonons alive between launch windows (ping-pong)
    int64_t list_cap = 0;
    uint32_t* d_birth_hist = nullptr;
    uint32_t* d_born_offset = nullptr;
    uint32_t* d_cursor = nullptr;
    uint32_t* d_n_live = nullptr;
    uint32_t* d_queue = nullptr;
    uint32_t* d_sort_keys = nullptr;  // [SORT_KEYS] counts, then cursors, of the cell sort
    int64_t steps_cap = 0;
    int32_t window = 0;
    int32_t blocks_per_cu = 8;  // transport grid: n_CU x blocks_per_cu blocks
    int32_t block = 256;  // threads per transport block
    bool tally_fp32 = false;  // flux tallies in FP32 with Kahan-compensated sums (else FP64)
    int32_t tally_mode = PSIM_TALLY_AUTO;  // where recorded steps are tallied
    // seed of the current run, read by the kernels from device memory (pinned host staging copy)
    uint64_t* d_seed = nullptr;
    uint64_t* h_seed = nullptr;
    // captured run (psim_gpu_run_graph); valid while `gen` (bumped by anything that changes the
    // launch sequence or the buffers it names) equals graph_gen
    hipGraphExec_t graph_exec = nullptr;
    uint64_t gen = 1, graph_gen = 0;
    // timing of the last run
    double step_ms = 0., run_ms = 0.;
    int64_t step_launches = 0;
};

namespace {

int fail(psim_gpu_ctx* c, int code, const std::string& msg) {
    if (c) { c->err = msg; }
    return code;
}

#define HIP_CHECK(c, expr)                                                                                   \
    do {                                                                                                     \
        const hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                              \
            return fail((c), PSIM_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));                \
        }                                                                                                    \
    } while (0)

void free_model(psim_gpu_ctx* c) {
    for (void* p : c->model_allocs) { (void)hipFree(p); }
    c->model_allocs.clear();
    c->uploaded = false;
}

void free_pool(psim_gpu_ctx* c) {
    void* ps[] = { c->pool.pos, c->pool.dir, c->pool.fv, c->pool.aux, c->pool.ms };
    for (void* p : ps) {
        if (p) { (void)hipFree(p); }
    }
    c->pool = Pool{};
    c->pool_cap = 0;
}

void free_fate(psim_gpu_ctx* c) {
    if (c->fate.step) { (void)hipFree(c->fate.step); }
    if (c->fate.cause) { (void)hipFree(c->fate.cause); }
    if (c->fate.xy) { (void)hipFree(c->fate.xy); }
    c->fate = Fate{ nullptr, nullptr, nullptr };
    c->fate_cap = 0;
}

void free_lists(psim_gpu_ctx* c) {
    void* ps[] = { c->d_birth_key, c->d_born_order };
    for (void* p : ps) {
        if (p) { (void)hipFree(p); }
    }
    for (Live& L : c->live) {
        void* ls[] = { L.pos, L.dir, L.fv, L.at, L.r01, L.aux, L.tag };
        for (void* p : ls) {
            if (p) { (void)hipFree(p); }
        }
        L = Live{};
    }
    c->d_birth_key = nullptr;
    c->d_born_order = nullptr;
    c->list_cap = 0;
}

void free_steps(psim_gpu_ctx* c) {
    void* ps[] = { c->d_birth_hist, c->d_born_offset, c->d_cursor, c->d_n_live, c->d_queue, c->d_sort_keys };
    for (void* p : ps) {
        if (p) { (void)hipFree(p); }
    }
    c->d_birth_hist = c->d_born_offset = c->d_cursor = c->d_n_live = c->d_queue = c->d_sort_keys = nullptr;
    c->steps_cap = 0;
}

void free_tallies(psim_gpu_ctx* c) {
    void* ps[] = { c->d_counts, c->d_flux, c->d_dcnt, c->d_dfx, c->d_dfy };
    for (void* p : ps) {
        if (p) { (void)hipFree(p); }
    }
    c->d_counts = nullptr;
    c->d_flux = nullptr;
    c->d_dcnt = nullptr;
    c->d_dfx = nullptr;
    c->d_dfy = nullptr;
}

void free_partials(psim_gpu_ctx* c) {
    if (c->d_pcnt) { (void)hipFree(c->d_pcnt); }
    if (c->d_pfx) { (void)hipFree(c->d_pfx); }
    if (c->d_pfy) { (void)hipFree(c->d_pfy); }
    if (c->d_pcx) { (void)hipFree(c->d_pcx); }
    if (c->d_pcy) { (void)hipFree(c->d_pcy); }
    c->d_pcnt = nullptr;
    c->d_pfx = nullptr;
    c->d_pfy = nullptr;
    c->d_pcx = nullptr;
    c->d_pcy = nullptr;
    c->partial_cap = 0;
}

template <typename T>
int upload_array(psim_gpu_ctx* c, const T* host, size_t n, const T** out) {
    T* d = dalloc<T>(n);
    if (!d) { return fail(c, PSIM_E_DEVICE, "hipMalloc failed for model data"); }
    c->model_allocs.push_back(d);
    if (n) { HIP_CHECK(c, hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice)); }
    *out = d;
    return PSIM_OK;
}

// Measurement steps per transport launch.  A recording window (one that reaches a tallied step)
// keeps W x S tally entries of 20 bytes in LDS; windows before the first recorded step carry no
// tally, so by default they are merged into one launch (lanes refill from the work queue, so
// phonon lifetimes of any length balance within it; each launch ends with a tail in which the
// last phonons finish, which is why fewer launches run faster).  An explicit window length
// applies to both kinds.
constexpr int32_t DEFAULT_FREE_WINDOW = 1 << 30;

int32_t record_window_for(const psim_gpu_ctx* c) {
    // (W + 1) rows of 20 B per sensor: the difference arrays need one row past the window
    const int64_t by_lds = std::max<int64_t>(1, LDS_TALLY_BYTES / (20 * std::max(1, c->n_sensors)) - 1);
    return (int32_t)std::min<int64_t>(c->window > 0 ? c->window : 128, by_lds);
}

int32_t free_window_for(const psim_gpu_ctx* c) { return c->window > 0 ? c->window : DEFAULT_FREE_WINDOW; }

// Tally path of a run: LDS window partials for runs with enough phonons per sensor when enough
// steps of tallies fit the LDS budget, else FP64 atomics into the HBM tallies (one launch for the
// whole run).  psim_gpu_set_tally_mode forces either (LDS only where one step fits).
bool use_global_tally(const psim_gpu_ctx* c) {
    const int64_t sensors = std::max(1, c->n_sensors);
    const int64_t lds_steps = LDS_TALLY_BYTES / (20 * sensors) - 1;
    if (c->tally_mode == PSIM_TALLY_HBM || lds_steps < 1) { return true; }
    if (c->tally_mode == PSIM_TALLY_LDS) { return false; }
    return lds_steps < GLOBAL_TALLY_STEPS || c->params.total_phonons < LDS_MIN_PHONONS_PER_SENSOR * sensors;
}

// Launch windows [k0, k1) over N steps: step values k0+1..k1 are reached inside window [k0, k1);
// windows whose step values all lie below adj (the first recorded step) are tally-free.
std::vector<std::pair<int64_t, int64_t>> plan_windows(int64_t N, int64_t adj, int32_t w_free, int32_t w_rec) {
    std::vector<std::pair<int64_t, int64_t>> out;
    int64_t k0 = 0;
    while (k0 < N) {
        int64_t k1;
        if (k0 + w_free <= adj - 1) {
            k1 = k0 + w_free;
        } else if (k0 < adj - 1) {
            k1 = adj - 1;
        } else {
            k1 = std::min<int64_t>(N, k0 + w_rec);
        }
        out.emplace_back(k0, k1);
        k0 = k1;
    }
    return out;
}

// Launch plan of one run of the uploaded model.
struct RunPlan {
    int64_t n = 0, N = 0, R = 0, blocks_needed = 0, steps_need = 0, WS = 0;
    int S = 0, grid = 1;
    int32_t W = 0;
    bool global_tally = false;
    std::vector<std::pair<int64_t, int64_t>> windows;
};

// Allocates (or reuses) every buffer the run needs and plans its windows.  Any reallocation
// bumps c->gen, which invalidates a captured graph.
int prepare_run(psim_gpu_ctx* c, RunPlan& plan) {
    if (!c->uploaded) { return fail(c, PSIM_E_STATE, "psim_gpu_upload must be called before running"); }
    HIP_CHECK(c, hipSetDevice(c->device));
    const psim_run_params& P = c->params;
    const int64_t n = P.total_phonons;
    const int S = c->n_sensors;
    const int64_t R = P.num_record;
    // pool
    if (c->pool_cap < n) {
        free_pool(c);
        const size_t cap = (size_t)std::max<int64_t>(n, 1);
        c->pool.pos = dalloc<double2>(cap);
        c->pool.dir = dalloc<double2>(cap);
        c->pool.fv = dalloc<double2>(cap);
        c->pool.aux = dalloc<PoolAux>(cap);
        c->pool.ms = dalloc<uint2>(cap);
        if (!c->pool.pos || !c->pool.dir || !c->pool.fv || !c->pool.aux || !c->pool.ms) {
            free_pool(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for the phonon pool");
        }
        c->pool_cap = (int64_t)cap;
        c->gen++;
    }
    if (c->fate_enabled && c->fate_cap < n) {
        free_fate(c);
        const size_t cap = (size_t)std::max<int64_t>(n, 1);
        c->fate.step = dalloc<int64_t>(cap);
        c->fate.cause = dalloc<int32_t>(cap);
        c->fate.xy = dalloc<double>(2 * cap);
        if (!c->fate.step || !c->fate.cause || !c->fate.xy) {
            free_fate(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for the fate arrays");
        }
        c->fate_cap = (int64_t)cap;
        c->gen++;
    }
    const int64_t blocks_needed = (n + BLOCK - 1) / BLOCK;  // emission and birth sort: one thread per phonon
    const int tb = c->tally_fp32 ? 256 : c->block;  // transport threads per block (FP32 tallies: 256)
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + tb - 1) / tb, (int64_t)c->n_cu * c->blocks_per_cu));
    // Each lane counts its phonon-steps, tallies and events in 32 bits until the launch ends (the
    // wave sums are 64-bit).  The work queue spreads the phonons evenly over the lanes, so a lane
    // carries about n * N / lanes phonon-steps at most; runs where that average comes within a
    // factor of 2 of 2^32 are refused rather than risk a wrapped counter.
    const double lane_share = (double)n * (double)P.num_measurements / ((double)grid * tb);
    if (lane_share >= 2147483648.0) {
        return fail(c, PSIM_E_INVALID, "run too large: phonons x measurement steps per lane exceeds 2^31");
    }
    // A live buffer holds every phonon plus the survivor slots the waves may reserve but not fill.
    const int64_t live_need = n + (int64_t)OUT_CLAIM * grid * (tb / WAVE);
    if (c->list_cap < live_need) {
        free_lists(c);
        const size_t cap = (size_t)std::max<int64_t>(live_need, 1);
        c->d_birth_key = dalloc<int32_t>(cap);
        c->d_born_order = dalloc<uint32_t>(cap);
        bool ok = c->d_birth_key && c->d_born_order;
        for (Live& L : c->live) {
            L.pos = dalloc<double2>(cap);
            L.dir = dalloc<double2>(cap);
            L.fv = dalloc<double2>(cap);
            L.at = dalloc<double2>(cap);
            L.r01 = dalloc<double2>(cap);
            L.aux = dalloc<LiveAux>(cap);
            L.tag = dalloc<LiveTag>(cap);
            ok = ok && L.pos && L.dir && L.fv && L.at && L.r01 && L.aux && L.tag;
        }
        if (!ok) {
            free_lists(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for the phonon lists");
        }
        c->list_cap = (int64_t)cap;
        c->gen++;
    }
    const bool global_tally = use_global_tally(c);
    const int32_t W = global_tally ? free_window_for(c) : record_window_for(c);
    const int64_t N = P.num_measurements;
    const auto windows = plan_windows(N, P.step_adjustment, free_window_for(c), W);
    const int64_t n_windows = (int64_t)windows.size();
    const int64_t steps_need = std::max<int64_t>(N + 1, n_windows + 1);
    if (c->steps_cap < steps_need) {
        free_steps(c);
        c->d_birth_hist = dalloc<uint32_t>((size_t)steps_need);
        c->d_born_offset = dalloc<uint32_t>((size_t)steps_need);
        c->d_cursor = dalloc<uint32_t>((size_t)steps_need);
        c->d_n_live = dalloc<uint32_t>((size_t)steps_need);
        c->d_queue = dalloc<uint32_t>((size_t)steps_need);
        c->d_sort_keys = dalloc<uint32_t>((size_t)SORT_KEYS + 1);  // + the sorted count
        if (!c->d_birth_hist || !c->d_born_offset || !c->d_cursor || !c->d_n_live || !c->d_queue || !c->d_sort_keys) {
            free_steps(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for the step arrays");
        }
        c->steps_cap = steps_need;
        c->gen++;
    }
    const int64_t WS = global_tally ? 0 : (int64_t)W * S;
    if (c->partial_cap < grid * WS) {
        free_partials(c);
        const size_t cap = (size_t)(grid * WS);
        c->d_pcnt = dalloc<int32_t>(cap);
        c->d_pfx = dalloc<double>(cap);
        c->d_pfy = dalloc<double>(cap);
        c->d_pcx = dalloc<float>(cap);
        c->d_pcy = dalloc<float>(cap);
        if (!c->d_pcnt || !c->d_pfx || !c->d_pfy || !c->d_pcx || !c->d_pcy) {
            free_partials(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for tally partials");
        }
        c->partial_cap = (int64_t)cap;
        c->gen++;
    }
    plan.n = n;
    plan.N = N;
    plan.R = R;
    plan.S = S;
    plan.blocks_needed = blocks_needed;
    plan.grid = grid;
    plan.steps_need = steps_need;
    plan.WS = WS;
    plan.W = W;
    plan.global_tally = global_tally;
    plan.windows = windows;
    return PSIM_OK;
}

int zero_fill(psim_gpu_ctx* c, void* p, size_t bytes) {
    const int64_t words = (int64_t)(bytes / 4);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (words + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL(zero_words_kernel, dim3(blocks), dim3(BLOCK), 0, c->stream, static_cast<uint32_t*>(p), words);
    HIP_CHECK(c, hipGetLastError());
    return PSIM_OK;
}

// One transport launch of the variant for this tally path, stream layout, block size and flux type.
template <int NT, typename FT>
void launch_transport_nt(psim_gpu_ctx* c, int tally, int grid, size_t lds, const Lists& lists, int32_t k0, int32_t k1,
    const Tally& t, const Fate& fate) {
    const dim3 g((unsigned)grid), b(NT);
#define PSIM_TRANSPORT(T, M, F)                                                                              \
    hipLaunchKernelGGL((transport_kernel<T, M, NT, F>), g, b, lds, c->stream, c->m, c->pool, lists, k0, k1, c->d_seed, t, \
        c->d_counters, fate)
    // the untallied variant has no flux arrays: it is the same kernel for either flux type
    if (c->multi_stream) {
        if (tally == TALLY_NONE) {
            PSIM_TRANSPORT(TALLY_NONE, true, double);
        } else if (tally == TALLY_GLOBAL) {
            PSIM_TRANSPORT(TALLY_GLOBAL, true, FT);
        } else {
            PSIM_TRANSPORT(TALLY_LDS, true, FT);
        }
    } else if (tally == TALLY_NONE) {
        PSIM_TRANSPORT(TALLY_NONE, false, double);
    } else if (tally == TALLY_GLOBAL) {
        PSIM_TRANSPORT(TALLY_GLOBAL, false, FT);
    } else {
        PSIM_TRANSPORT(TALLY_LDS, false, FT);
    }
#undef PSIM_TRANSPORT
}

// The FP32 tally precision runs 256-thread blocks only (one more set of kernels, not three).
void launch_transport(psim_gpu_ctx* c, int tally, int grid, size_t lds, const Lists& lists, int32_t k0, int32_t k1,
    const Tally& t, const Fate& fate) {
    if (c->tally_fp32) {
        launch_transport_nt<256, float>(c, tally, grid, lds, lists, k0, k1, t, fate);
    } else if (c->block == 128) {
        launch_transport_nt<128, double>(c, tally, grid, lds, lists, k0, k1, t, fate);
    } else if (c->block == 512) {
        launch_transport_nt<512, double>(c, tally, grid, lds, lists, k0, k1, t, fate);
    } else {
        launch_transport_nt<256, double>(c, tally, grid, lds, lists, k0, k1, t, fate);
    }
}

// Every launch of one run, on c->stream.  `eager`: also times each transport window with events
// (and prints a progress line per window when PSIM_PROGRESS is set); otherwise the sequence is
// free of host synchronisation, so it can be captured into a graph.
int enqueue_run(psim_gpu_ctx* c, const RunPlan& plan, bool eager) {
    const psim_run_params& P = c->params;
    const int64_t n = plan.n, N = plan.N, R = plan.R, WS = plan.WS, blocks_needed = plan.blocks_needed;
    const int S = plan.S, grid = plan.grid;
    const int32_t W = plan.W;
    const bool global_tally = plan.global_tally;
    const auto& windows = plan.windows;
    const int64_t n_windows = (int64_t)windows.size();
    const Fate fate = c->fate_enabled ? c->fate : Fate{ nullptr, nullptr, nullptr };
    int rc;
    if ((rc = zero_fill(c, c->d_counts, sizeof(int64_t) * (size_t)std::max<int64_t>(1, S * R))) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_flux, sizeof(double) * (size_t)std::max<int64_t>(1, 2 * S * R))) != 0) { return rc; }
    const size_t SR1 = (size_t)S * (size_t)(R + 1);
    if ((rc = zero_fill(c, c->d_dcnt, sizeof(int64_t) * SR1)) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_dfx, sizeof(double) * SR1)) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_dfy, sizeof(double) * SR1)) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_counters, sizeof(unsigned long long) * PSIM_NUM_COUNTERS)) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_birth_hist, sizeof(uint32_t) * (size_t)plan.steps_need)) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_n_live, sizeof(uint32_t) * (size_t)plan.steps_need)) != 0) { return rc; }
    if ((rc = zero_fill(c, c->d_queue, sizeof(uint32_t) * (size_t)plan.steps_need)) != 0) { return rc; }
    if (n > 0) {
        hipLaunchKernelGGL(emit_kernel, dim3((unsigned)blocks_needed), dim3(BLOCK), 0, c->stream, c->m, c->pool, n,
            c->d_seed, c->d_counters, fate, c->d_birth_key, c->d_birth_hist);
        HIP_CHECK(c, hipGetLastError());
        hipLaunchKernelGGL(birth_scan_kernel, dim3(1), dim3(1024), 0, c->stream, c->d_birth_hist, N, c->d_born_offset,
            c->d_cursor);
        HIP_CHECK(c, hipGetLastError());
        hipLaunchKernelGGL(birth_scatter_kernel, dim3((unsigned)blocks_needed), dim3(BLOCK), 0, c->stream,
            c->d_birth_key, n, c->d_cursor, c->d_born_order);
        HIP_CHECK(c, hipGetLastError());
    }
    const bool progress = std::getenv("PSIM_PROGRESS") != nullptr;  // one stderr line per window
    for (int32_t w = 0; w < (int32_t)n_windows && n > 0; ++w) {
        const int64_t k0 = windows[(size_t)w].first, k1 = windows[(size_t)w].second;
        // window 0 starts from nothing (d_n_live[0] = 0); later ones from the sorted survivors
        const uint32_t* n_in = (w == 0) ? c->d_n_live : c->d_sort_keys + SORT_KEYS;
        const Lists lists{ c->live[0], c->live[1], n_in, c->d_n_live + w + 1, c->d_queue, c->d_born_order,
            c->d_born_offset, w };
        // window steps k0+1 .. k1 map to tally indices (k - adj) in [0, R)
        const bool rec = (k1 >= P.step_adjustment) && (k0 + 1 - P.step_adjustment < R) && (k1 - k0 <= W);
        Tally t{ c->d_pcnt, c->d_pfx, c->d_pfy, c->d_pcx, c->d_pcy, c->d_dcnt, c->d_dfx, c->d_dfy, rec ? 1 : 0,
            global_tally ? 1 : 0, W };
        // (W + 1) rows per sensor of a count and two flux values
        const size_t lds = (rec && !global_tally) ? (size_t)(WS + S) * (c->tally_fp32 ? 12 : 20) : 0;
        if (eager) { HIP_CHECK(c, hipEventRecord(c->ev[1], c->stream)); }
        launch_transport(c, !rec ? TALLY_NONE : global_tally ? TALLY_GLOBAL : TALLY_LDS, grid, !rec ? 0 : lds, lists,
            (int32_t)k0, (int32_t)k1, t, fate);
        HIP_CHECK(c, hipGetLastError());
        if (eager) { HIP_CHECK(c, hipEventRecord(c->ev[2], c->stream)); }
        if (rec && !global_tally) {
            const int rblocks = (int)((WS + BLOCK - 1) / BLOCK);
            // split the block partials so that the reduce has ~4 blocks per CU
            const int splits = std::max(1, std::min(grid, (4 * c->n_cu + rblocks - 1) / rblocks));
            const int per = (grid + splits - 1) / splits;
            const dim3 rg(rblocks, (unsigned)((grid + per - 1) / per));
            if (c->tally_fp32) {
                hipLaunchKernelGGL(reduce_kernel<float>, rg, dim3(BLOCK), 0, c->stream, c->d_pcnt,
                    reinterpret_cast<const float*>(c->d_pfx), reinterpret_cast<const float*>(c->d_pfy), c->d_pcx, c->d_pcy,
                    grid, per, (int)WS, S, k0, P.step_adjustment, R, c->d_counts, c->d_flux);
            } else {
                hipLaunchKernelGGL(reduce_kernel<double>, rg, dim3(BLOCK), 0, c->stream, c->d_pcnt, c->d_pfx, c->d_pfy,
                    c->d_pcx, c->d_pcy, grid, per, (int)WS, S, k0, P.step_adjustment, R, c->d_counts, c->d_flux);
            }
            HIP_CHECK(c, hipGetLastError());
        }
        if (w + 1 < (int32_t)n_windows) {  // survivors back into live[0], in cell order
            int shift = 0;
            while (((c->n_cells - 1) >> shift) >= SORT_KEYS) { ++shift; }
            const int n_keys = ((c->n_cells - 1) >> shift) + 1;
            const uint32_t* n_ptr = c->d_n_live + w + 1;
            const int sgrid = std::max(1, c->n_cu * 2);
            if ((rc = zero_fill(c, c->d_sort_keys, sizeof(uint32_t) * (size_t)n_keys)) != 0) { return rc; }
            hipLaunchKernelGGL(cell_hist_kernel, dim3(sgrid), dim3(SORT_BLOCK), 0, c->stream, c->live[1].aux, n_ptr,
                shift, n_keys, c->d_sort_keys);
            HIP_CHECK(c, hipGetLastError());
            hipLaunchKernelGGL(cell_scan_kernel, dim3(1), dim3(1024), 0, c->stream, c->d_sort_keys, n_keys,
                c->d_sort_keys + SORT_KEYS);
            HIP_CHECK(c, hipGetLastError());
            hipLaunchKernelGGL(cell_scatter_kernel, dim3(sgrid), dim3(SORT_BLOCK), 0, c->stream, c->live[1], c->live[0],
                n_ptr, shift, n_keys, c->d_sort_keys);
            HIP_CHECK(c, hipGetLastError());
        }
        if (!eager) { continue; }
        HIP_CHECK(c, hipEventSynchronize(c->ev[2]));
        float ms = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
        c->step_ms += ms;
        c->step_launches++;
        if (progress) {
            std::fprintf(stderr, "psim_gpu: window %d/%d steps [%lld, %lld) %.1f ms\n", (int)w + 1, (int)n_windows,
                (long long)k0, (long long)k1, (double)ms);
            std::fflush(stderr);
        }
    }
    if (n > 0 && R > 0) {  // the segments tallied in HBM
        const dim3 dg((unsigned)((S + BLOCK - 1) / BLOCK));
        if (c->tally_fp32) {
            hipLaunchKernelGGL(diff_prefix_kernel<float>, dg, dim3(BLOCK), 0, c->stream, c->d_dcnt,
                reinterpret_cast<const float*>(c->d_dfx), reinterpret_cast<const float*>(c->d_dfy), S, R, c->d_counts,
                c->d_flux);
        } else {
            hipLaunchKernelGGL(diff_prefix_kernel<double>, dg, dim3(BLOCK), 0, c->stream, c->d_dcnt, c->d_dfx, c->d_dfy,
                S, R, c->d_counts, c->d_flux);
        }
        HIP_CHECK(c, hipGetLastError());
    }
    return PSIM_OK;
}

int set_seed(psim_gpu_ctx* c, uint64_t seed) {
    *c->h_seed = seed;
    HIP_CHECK(c, hipMemcpyAsync(c->d_seed, c->h_seed, sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
    return PSIM_OK;
}

int finish_run(psim_gpu_ctx* c) {
    HIP_CHECK(c, hipEventRecord(c->ev[3], c->stream));
    HIP_CHECK(c, hipEventSynchronize(c->ev[3]));
    float total = 0.f;
    HIP_CHECK(c, hipEventElapsedTime(&total, c->ev[0], c->ev[3]));
    c->run_ms = total;
    return PSIM_OK;
}

int run_impl(psim_gpu_ctx* c, uint64_t seed) {
    RunPlan plan;
    int rc;
    if ((rc = prepare_run(c, plan)) != 0) { return rc; }
    if ((rc = set_seed(c, seed)) != 0) { return rc; }
    c->step_ms = 0.;
    c->step_launches = 0;
    HIP_CHECK(c, hipEventRecord(c->ev[0], c->stream));
    if ((rc = enqueue_run(c, plan, true)) != 0) { return rc; }
    return finish_run(c);
}

// Captures the uploaded run as a graph on first use, and again whenever c->gen moves.
int ensure_graph(psim_gpu_ctx* c, RunPlan& plan) {
    int rc;
    if (!c->graph_exec || c->graph_gen != c->gen) {
        if (c->graph_exec) {
            (void)hipGraphExecDestroy(c->graph_exec);
            c->graph_exec = nullptr;
        }
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_run(c, plan, false);
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(c->stream, &graph);
        if (rc != 0) {
            if (graph) { (void)hipGraphDestroy(graph); }
            return rc;
        }
        if (end != hipSuccess) { return fail(c, PSIM_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(end)); }
        const hipError_t inst = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (inst != hipSuccess) {
            c->graph_exec = nullptr;
            return fail(c, PSIM_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(inst));
        }
        c->graph_gen = c->gen;
    }
    return PSIM_OK;
}

// The same run replayed from its captured graph.
int run_graph_impl(psim_gpu_ctx* c, uint64_t seed) {
    RunPlan plan;
    int rc;
    if ((rc = prepare_run(c, plan)) != 0) { return rc; }
    if ((rc = ensure_graph(c, plan)) != 0) { return rc; }
    if ((rc = set_seed(c, seed)) != 0) { return rc; }
    HIP_CHECK(c, hipEventRecord(c->ev[0], c->stream));
    HIP_CHECK(c, hipGraphLaunch(c->graph_exec, c->stream));
    if ((rc = finish_run(c)) != 0) { return rc; }
    c->step_ms = c->run_ms;
    c->step_launches = (int64_t)plan.windows.size();
    return PSIM_OK;
}

}  // namespace

extern "C" {

int psim_gpu_create(psim_gpu_ctx** out, int device) {
    if (!out) { return PSIM_E_INVALID; }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) { return PSIM_E_DEVICE; }
    if (device < 0 || device >= count) { return PSIM_E_INVALID; }
    auto* c = new psim_gpu_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return PSIM_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return PSIM_E_DEVICE;
    }
    c->n_cu = prop.multiProcessorCount;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return PSIM_E_DEVICE;
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            psim_gpu_destroy(c);
            return PSIM_E_DEVICE;
        }
    }
    c->d_counters = dalloc<unsigned long long>(PSIM_NUM_COUNTERS);
    c->d_seed = dalloc<uint64_t>(1);
    if (!c->d_counters || !c->d_seed
        || hipHostMalloc(reinterpret_cast<void**>(&c->h_seed), sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
        c->h_seed = nullptr;
        psim_gpu_destroy(c);
        return PSIM_E_DEVICE;
    }
    *out = c;
    return PSIM_OK;
}

void psim_gpu_destroy(psim_gpu_ctx* c) {
    if (!c) { return; }
    (void)hipSetDevice(c->device);
    if (c->stream) { (void)hipStreamSynchronize(c->stream); }
    free_model(c);
    free_pool(c);
    free_fate(c);
    free_partials(c);
    free_lists(c);
    free_steps(c);
    free_tallies(c);
    if (c->d_counters) { (void)hipFree(c->d_counters); }
    if (c->d_seed) { (void)hipFree(c->d_seed); }
    if (c->h_seed) { (void)hipHostFree(c->h_seed); }
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); }
    for (auto& e : c->ev) {
        if (e) { (void)hipEventDestroy(e); }
    }
    if (c->stream) { (void)hipStreamDestroy(c->stream); }
    delete c;
}

const char* psim_gpu_last_error(const psim_gpu_ctx* c) { return c ? c->err.c_str() : "null context"; }

int psim_gpu_upload(psim_gpu_ctx* c, const psim_geom* g, const psim_tables* t, const psim_sources* s,
    const psim_run_params* p) {
    if (!c || !g || !t || !s || !p) { return fail(c, PSIM_E_INVALID, "null descriptor"); }
    if (g->n_cells <= 0 || g->n_sensors <= 0 || !g->cells || !g->edges || !g->sensors) {
        return fail(c, PSIM_E_INVALID, "geometry needs cells, edges and sensors");
    }
    if (g->n_subs < 0 || (g->n_subs > 0 && !g->subs)) { return fail(c, PSIM_E_INVALID, "bad sub-surfaces"); }
    if (t->n_materials <= 0 || t->n_tables <= 0 || !t->materials || !t->freq || !t->vel_la || !t->vel_ta || !t->cdf
        || !t->la_frac) {
        return fail(c, PSIM_E_INVALID, "tables incomplete");
    }
    if (s->n_sources < 0 || (s->n_sources > 0 && !s->sources)) { return fail(c, PSIM_E_INVALID, "bad sources"); }
    if (p->num_measurements <= 0 || p->num_measurements >= INT32_MAX || p->num_record < 0
        || p->num_record > p->num_measurements || p->step_adjustment < 0 || p->step_adjustment > p->num_measurements
        || p->step_time <= 0. || p->total_phonons < 0
        || p->total_phonons > (int64_t)UINT32_MAX) {
        return fail(c, PSIM_E_INVALID, "bad run parameters");
    }
    if ((int64_t)g->n_sensors * p->num_record >= ((int64_t)1 << 40)) {
        return fail(c, PSIM_E_INVALID, "sensors x recorded steps too large");
    }
    // Every index the kernels follow is checked here, so a bad descriptor cannot make them
    // read outside the uploaded arrays.
    for (int i = 0; i < g->n_cells; ++i) {
        if (g->cells[i].sensor < 0 || g->cells[i].sensor >= g->n_sensors) {
            return fail(c, PSIM_E_INVALID, "cell sensor out of range");
        }
        for (int b = 0; b < 3; ++b) {
            const psim_edge& e = g->edges[3 * i + b];
            if (e.sub_begin < 0 || e.n_trans < 0 || e.n_emit < 0 || e.sub_begin + e.n_trans + e.n_emit > g->n_subs) {
                return fail(c, PSIM_E_INVALID, "edge sub-surface range out of range");
            }
            for (int k = 0; k < e.n_trans + e.n_emit; ++k) {
                const psim_subsurface& ss = g->subs[e.sub_begin + k];
                if ((k < e.n_trans) != (ss.kind == 0)) {
                    return fail(c, PSIM_E_INVALID, "transition sub-surfaces must precede emitting ones");
                }
                if (ss.kind == 0 && (ss.target < 0 || ss.target >= g->n_cells)) {
                    return fail(c, PSIM_E_INVALID, "transition target out of range");
                }
            }
        }
    }
    for (int i = 0; i < g->n_sensors; ++i) {
        const psim_sensor& ss = g->sensors[i];
        if (ss.material < 0 || ss.material >= t->n_materials || ss.base_table < 0 || ss.base_table >= t->n_tables
            || ss.scatter_table < 0 || ss.scatter_table >= t->n_tables) {
            return fail(c, PSIM_E_INVALID, "sensor material or table out of range");
        }
        if (ss.flight_cap < 0) { return fail(c, PSIM_E_INVALID, "negative flight cap"); }
    }
    int64_t expect = 0;
    bool multi = false;
    for (int i = 0; i < s->n_sources; ++i) {
        const psim_source& src = s->sources[i];
        multi = multi || src.stream != 0 || src.id_base != 0;
        if (src.cell < 0 || src.cell >= g->n_cells || src.table < 0 || src.table >= t->n_tables || src.count < 0
            || src.offset != expect || (src.kind != 0 && src.kind != 1) || src.stream < 0 || src.stream > 0xffff
            || src.id_base < 0 || src.id_base > src.offset
            || src.offset + src.count - src.id_base > (int64_t)UINT32_MAX) {
            return fail(c, PSIM_E_INVALID, "bad emission source");
        }
        expect += src.count;
    }
    if (expect != p->total_phonons) { return fail(c, PSIM_E_INVALID, "source counts do not add up to total_phonons"); }
    if (p->total_phonons > 0 && s->n_sources == 0) { return fail(c, PSIM_E_INVALID, "phonons without sources"); }
    HIP_CHECK(c, hipSetDevice(c->device));
    free_model(c);
    Model m{};
    int rc;
#define PSIM_UPLOAD(field, host, count)                                                                  \
    do {                                                                                                 \
        const std::remove_const_t<std::remove_pointer_t<decltype(host)>>* d_ = nullptr;                  \
        if ((rc = upload_array(c, (host), (size_t)(count), &d_)) != 0) { return rc; }                    \
        m.field = d_;                                                                                    \
    } while (0)
    const size_t nb = (size_t)PSIM_NUM_FREQ_BINS;
    PSIM_UPLOAD(cells, g->cells, g->n_cells);
    PSIM_UPLOAD(edges, g->edges, (size_t)g->n_cells * 3);
    PSIM_UPLOAD(subs, g->subs, g->n_subs);
    std::vector<psim_sensor> sensors(g->sensors, g->sensors + g->n_sensors);
    for (psim_sensor& se : sensors) {  // 0: the cap of one model with every sensor of the upload
        if (se.flight_cap == 0) { se.flight_cap = psim_flight_cap(g->n_sensors); }
    }
    PSIM_UPLOAD(sensors, sensors.data(), g->n_sensors);
    PSIM_UPLOAD(mats, t->materials, t->n_materials);
    PSIM_UPLOAD(freq, t->freq, nb * t->n_materials);
    PSIM_UPLOAD(vel_la, t->vel_la, nb * t->n_materials);
    PSIM_UPLOAD(vel_ta, t->vel_ta, nb * t->n_materials);
    PSIM_UPLOAD(cdf, t->cdf, nb * t->n_tables);
    PSIM_UPLOAD(frac, t->la_frac, nb * t->n_tables);
    PSIM_UPLOAD(sources, s->sources, s->n_sources);
    // Guide tables of the frequency search, when every cumulative table is non-decreasing (a
    // NaN fails the test too).
    bool monotone = true;
    for (int64_t i = 0; i + 1 < (int64_t)nb * t->n_tables && monotone; ++i) {
        if ((i + 1) % (int64_t)nb != 0 && !(t->cdf[i] <= t->cdf[i + 1])) { monotone = false; }
    }
    if (monotone) {
        std::vector<uint16_t> guide((size_t)(GUIDE + 1) * t->n_tables);
        for (int k = 0; k < t->n_tables; ++k) {
            const double* cdf = t->cdf + (size_t)k * nb;
            int i = 0;
            for (int g = 0; g <= GUIDE; ++g) {  // first index whose value exceeds g / GUIDE
                const double x = (double)g / (double)GUIDE;
                while (i < (int)nb && !(cdf[i] > x)) { ++i; }
                guide[(size_t)k * (GUIDE + 1) + g] = (uint16_t)i;
            }
        }
        PSIM_UPLOAD(guide, guide.data(), guide.size());
    } else {
        m.guide = nullptr;
    }
#undef PSIM_UPLOAD
    m.n_sources = s->n_sources;
    m.n_sensors = g->n_sensors;
    m.full = p->full;
    m.phasor = p->phasor;
    m.N = (int32_t)p->num_measurements;
    m.adj = (int32_t)p->step_adjustment;
    m.stop = p->step_adjustment >= 2 ? (int32_t)(p->step_adjustment - 1) : 0;
    m.R = (int32_t)p->num_record;
    m.sim_time = p->sim_time;
    m.step_time = p->step_time;
    m.t_eq = p->t_eq;
    c->m = m;
    c->params = *p;
    const int64_t SR = (int64_t)g->n_sensors * p->num_record;
    const int64_t SR1 = (int64_t)g->n_sensors * (p->num_record + 1);
    free_tallies(c);
    c->d_counts = dalloc<int64_t>((size_t)std::max<int64_t>(1, SR));
    c->d_flux = dalloc<double>((size_t)std::max<int64_t>(1, 2 * SR));
    c->d_dcnt = dalloc<int64_t>((size_t)SR1);
    c->d_dfx = dalloc<double>((size_t)SR1);
    c->d_dfy = dalloc<double>((size_t)SR1);
    if (!c->d_counts || !c->d_flux || !c->d_dcnt || !c->d_dfx || !c->d_dfy) {
        free_tallies(c);
        return fail(c, PSIM_E_DEVICE, "hipMalloc failed for tallies");
    }
    c->n_sensors = g->n_sensors;
    c->n_cells = g->n_cells;
    c->uploaded = true;
    c->multi_stream = multi;
    c->gen++;
    return PSIM_OK;
}

int psim_gpu_run_resident(psim_gpu_ctx* c, uint64_t seed) {
    if (!c) { return PSIM_E_INVALID; }
    return run_impl(c, seed);
}

int psim_gpu_fetch(psim_gpu_ctx* c, int64_t* counts, double* flux, int64_t* counters) {
    if (!c) { return PSIM_E_INVALID; }
    if (!c->uploaded) { return fail(c, PSIM_E_STATE, "nothing uploaded"); }
    HIP_CHECK(c, hipSetDevice(c->device));
    const int64_t SR = (int64_t)c->n_sensors * c->params.num_record;
    if (counts && SR) {
        HIP_CHECK(c, hipMemcpy(counts, c->d_counts, sizeof(int64_t) * (size_t)SR, hipMemcpyDeviceToHost));
    }
    if (flux && SR) { HIP_CHECK(c, hipMemcpy(flux, c->d_flux, sizeof(double) * (size_t)(2 * SR), hipMemcpyDeviceToHost)); }
    if (counters) {
        unsigned long long tmp[PSIM_NUM_COUNTERS];
        HIP_CHECK(c, hipMemcpy(tmp, c->d_counters, sizeof tmp, hipMemcpyDeviceToHost));
        for (int i = 0; i < PSIM_NUM_COUNTERS; ++i) { counters[i] = (int64_t)tmp[i]; }
    }
    return PSIM_OK;
}

int psim_gpu_run(psim_gpu_ctx* c, uint64_t seed, int64_t* counts, double* flux, int64_t* counters) {
    if (!c) { return PSIM_E_INVALID; }
    const int rc = run_impl(c, seed);
    if (rc != PSIM_OK) { return rc; }
    return psim_gpu_fetch(c, counts, flux, counters);
}

int psim_gpu_set_fate(psim_gpu_ctx* c, int enable) {
    if (!c) { return PSIM_E_INVALID; }
    if (c->fate_enabled != (enable != 0)) { c->gen++; }
    c->fate_enabled = enable != 0;
    return PSIM_OK;
}

int psim_gpu_fetch_fate(psim_gpu_ctx* c, int64_t* step, int32_t* cause, double* xy) {
    if (!c) { return PSIM_E_INVALID; }
    if (!c->fate_enabled || !c->fate.step) { return fail(c, PSIM_E_STATE, "fate recording was not enabled"); }
    HIP_CHECK(c, hipSetDevice(c->device));
    const size_t n = (size_t)c->params.total_phonons;
    if (n == 0) { return PSIM_OK; }
    HIP_CHECK(c, hipMemcpy(step, c->fate.step, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    HIP_CHECK(c, hipMemcpy(cause, c->fate.cause, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    HIP_CHECK(c, hipMemcpy(xy, c->fate.xy, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
    return PSIM_OK;
}

int psim_gpu_last_timing(const psim_gpu_ctx* c, double* step_ms, int64_t* launches, double* run_ms) {
    if (!c) { return PSIM_E_INVALID; }
    if (step_ms) { *step_ms = c->step_ms; }
    if (launches) { *launches = c->step_launches; }
    if (run_ms) { *run_ms = c->run_ms; }
    return PSIM_OK;
}

int psim_gpu_set_window(psim_gpu_ctx* c, int32_t steps) {
    if (!c || steps < 0) { return PSIM_E_INVALID; }
    if (c->window != steps) { c->gen++; }
    c->window = steps;
    return PSIM_OK;
}

int psim_gpu_set_grid(psim_gpu_ctx* c, int32_t blocks_per_cu) {
    if (!c || blocks_per_cu < 1 || blocks_per_cu > 64) { return PSIM_E_INVALID; }
    if (c->blocks_per_cu != blocks_per_cu) { c->gen++; }
    c->blocks_per_cu = blocks_per_cu;
    return PSIM_OK;
}

int psim_gpu_set_block(psim_gpu_ctx* c, int32_t threads) {
    if (!c || (threads != 128 && threads != 256 && threads != 512)) { return PSIM_E_INVALID; }
    if (c->block != threads) { c->gen++; }
    c->block = threads;
    return PSIM_OK;
}

int psim_gpu_set_tally_precision(psim_gpu_ctx* c, int32_t precision) {
    if (!c || (precision != PSIM_TALLY_FP64 && precision != PSIM_TALLY_FP32_KAHAN)) { return PSIM_E_INVALID; }
    const bool fp32 = precision == PSIM_TALLY_FP32_KAHAN;
    if (c->tally_fp32 != fp32) { c->gen++; }
    c->tally_fp32 = fp32;
    return PSIM_OK;
}

int psim_gpu_get_launch_shape(const psim_gpu_ctx* c, int32_t* window, int32_t* blocks_per_cu, int32_t* block) {
    if (!c) { return PSIM_E_INVALID; }
    if (window) { *window = c->window; }
    if (blocks_per_cu) { *blocks_per_cu = c->blocks_per_cu; }
    if (block) { *block = c->block; }
    return PSIM_OK;
}

int psim_gpu_set_tally_mode(psim_gpu_ctx* c, int32_t mode) {
    if (!c || mode < PSIM_TALLY_AUTO || mode > PSIM_TALLY_HBM) { return PSIM_E_INVALID; }
    if (c->tally_mode != mode) { c->gen++; }
    c->tally_mode = mode;
    return PSIM_OK;
}

int psim_gpu_run_async(psim_gpu_ctx* c, uint64_t seed, void* stream) {
    if (!c) { return PSIM_E_INVALID; }
    HIP_CHECK(c, hipSetDevice(c->device));
    RunPlan plan;
    int rc;
    if ((rc = prepare_run(c, plan)) != 0) { return rc; }
    if ((rc = ensure_graph(c, plan)) != 0) { return rc; }
    // work the engine queued on its own stream (uploads, clears, a capture) comes first
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(64), 0, st, c->d_seed, seed);
    HIP_CHECK(c, hipGetLastError());
    HIP_CHECK(c, hipGraphLaunch(c->graph_exec, st));
    c->run_ms = -1.;  // not timed
    return PSIM_OK;
}

int psim_gpu_run_graph(psim_gpu_ctx* c, uint64_t seed, int64_t* counts, double* flux, int64_t* counters) {
    if (!c) { return PSIM_E_INVALID; }
    const int rc = run_graph_impl(c, seed);
    if (rc != PSIM_OK) { return rc; }
    return psim_gpu_fetch(c, counts, flux, counters);
}

int psim_gpu_autotune(psim_gpu_ctx* c, uint64_t seed, int32_t* window, int32_t* blocks_per_cu, double* best_ms) {
    if (!c) { return PSIM_E_INVALID; }
    // Launch shapes that leave every result unchanged: the length of the untallied windows
    // (0 = one launch; shorter ones re-sort the survivors by cell more often), the number of
    // transport blocks per CU and the threads per transport block.  Each candidate runs the
    // uploaded model once; the fastest is kept.  Window and grid are searched at the current block
    // size, then the other block sizes at the best window and grid.
    const int32_t windows[] = { 0, 1024, 256 };
    const int32_t grids[] = { 8, 3 };
    const int32_t blocks[] = { 128, 256, 512 };
    const int32_t w0 = c->window, g0 = c->blocks_per_cu, b0 = c->block;
    double best = -1.;
    int32_t bw = w0, bg = g0, bb = b0;
    auto trial = [&](int32_t w, int32_t g, int32_t b) -> int {
        c->window = w;
        c->blocks_per_cu = g;
        c->block = b;
        c->gen++;
        const int rc = run_impl(c, seed);
        if (rc != PSIM_OK) {
            c->window = w0;
            c->blocks_per_cu = g0;
            c->block = b0;
            c->gen++;
            return rc;
        }
        if (best < 0. || c->run_ms < best) {
            best = c->run_ms;
            bw = w;
            bg = g;
            bb = b;
        }
        return PSIM_OK;
    };
    int rc;
    for (int32_t g : grids) {
        for (int32_t w : windows) {
            if ((rc = trial(w, g, b0)) != PSIM_OK) { return rc; }
        }
    }
    const int32_t wb = bw, gb = bg;
    for (int32_t b : blocks) {
        if (b != b0 && (rc = trial(wb, gb, b)) != PSIM_OK) { return rc; }
    }
    c->window = bw;
    c->blocks_per_cu = bg;
    c->block = bb;
    c->gen++;
    if (window) { *window = bw; }
    if (blocks_per_cu) { *blocks_per_cu = bg; }
    if (best_ms) { *best_ms = best; }
    return PSIM_OK;
}

}  // extern "C"
