// MI355X phonon engine: emission, windowed transport and tally reduction kernels, and the
// psim_gpu_* C ABI (include/psim_gpu.h).
//
// Data layout in HBM:
//   model      cells, edges (3 per cell), sub-surfaces, sensors, materials, frequency/velocity
//              arrays and the CDF / LA-fraction tables: uploaded once, read-only, L2-resident.
//   pool       one slot per phonon id, structure of arrays: px py dx dy freq vel age tts (f64),
//              cell (i32, -1 = gone), meta (u32: bin, polarization, sign, started, rate sensor).
//   tallies    counts [S][R] (i64) and flux [S][R][2] (f64), as Sensor::inc_energy_/inc_flux_.
//   partials   per transport block: [W][S] count (i32) and flux (2 x f64) of one launch window.
//
// A run is: emission kernel (every phonon's birth state), then ceil(N / W) transport launches
// each advancing every live phonon across W measurement steps with its state in registers,
// each followed (when the window holds recorded steps) by a reduction of the block partials
// into the tallies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
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
constexpr int LDS_TALLY_BYTES = 32 * 1024;  // budget for one block's [W][S] window tally

struct Tally {
    int32_t* __restrict__ pcnt;  // [grid][W*S]
    double* __restrict__ pfx;
    double* __restrict__ pfy;
    int32_t rec;  // 1 if this window records
    int32_t W;
};

struct Fate {
    int64_t* __restrict__ step;
    int32_t* __restrict__ cause;
    double* __restrict__ xy;
};

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

// Phonon builders, psim/src/phononBuilder.cpp:6-49, with EmitSurface::getPhononTime
// (surface.cpp:67-69) and Line/Triangle::getRandPoint (geometry.cpp:140-143, 234-242).
__global__ void __launch_bounds__(BLOCK) emit_kernel(Model m, Pool pool, int64_t total, uint64_t seed,
    unsigned long long* __restrict__ counters, Fate fate) {
    const int64_t id = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t not_started = 0;
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
        psim_stream st;
        psim_stream_init(&st, seed, PSIM_TAG_EMIT, (uint32_t)id, 0);
        Ph p;
        p.cell = src.cell;
        p.sign = src.sign;
        double birth = 0.;
        if (src.kind == 0) {  // CellOriginBuilder::operator()
            resample(m, mat, src.table, p, st);
            double r1 = psim_urand(&st);
            double r2 = psim_urand(&st);
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
        }
        pool.px[id] = p.px;
        pool.py[id] = p.py;
        pool.dx[id] = p.dx;
        pool.dy[id] = p.dy;
        pool.freq[id] = p.freq;
        pool.vel[id] = p.vel;
        pool.age[id] = birth;
        pool.tts[id] = 0.;
        pool.cell[id] = cell;
        pool.meta[id] = (uint32_t)p.bin | (p.pol == POL_TA ? META_TA : 0u) | (p.sign < 0 ? META_NEG : 0u);
    }
    const uint32_t ns = wave_sum(not_started);
    if ((threadIdx.x % WAVE) == 0 && ns) { atomicAdd(&counters[PSIM_CNT_NOT_STARTED], (unsigned long long)ns); }
}

// Advances every live phonon of the pool across measurement steps [k0, k1): the body of
// ModelSimulator::simulatePhonon (psim/src/modelSimulator.cpp:139-198) with the lane's state
// in registers, pausing right after the measurement that brings its step to k1.
__global__ void __launch_bounds__(BLOCK) transport_kernel(Model m, Pool pool, int64_t n_slots, int64_t k0,
    int64_t k1, uint64_t seed, Tally tally, unsigned long long* __restrict__ counters, Fate fate) {
    extern __shared__ unsigned char lds_raw[];
    const int S = m.n_sensors;
    const int WS = tally.rec ? tally.W * S : 0;
    double* lfx = reinterpret_cast<double*>(lds_raw);
    double* lfy = lfx + WS;
    int32_t* lcnt = reinterpret_cast<int32_t*>(lfy + WS);
    for (int i = threadIdx.x; i < WS; i += BLOCK) {
        lfx[i] = 0.;
        lfy[i] = 0.;
        lcnt[i] = 0;
    }
    __syncthreads();

    LaneCounters c = { 0, 0, 0, 0, 0, 0, 0, 0 };
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t slot = (int64_t)blockIdx.x * BLOCK + threadIdx.x; slot < n_slots; slot += stride) {
        const int32_t cell0 = pool.cell[slot];
        if (cell0 < 0) { continue; }
        const uint32_t meta = pool.meta[slot];
        double age = pool.age[slot];
        int64_t step;
        if (meta & META_STARTED) {
            step = k0;
        } else {
            step = (int64_t)(age / m.step_time);
            if (step >= k1) { continue; }
        }
        Ph p;
        p.cell = cell0;
        p.px = pool.px[slot];
        p.py = pool.py[slot];
        p.dx = pool.dx[slot];
        p.dy = pool.dy[slot];
        p.freq = pool.freq[slot];
        p.bin = (int)(meta & META_BIN_MASK);
        p.pol = (meta & META_TA) ? POL_TA : POL_LA;
        p.sign = (meta & META_NEG) ? -1 : 1;
        p.step = step;
        p.vel = pool.vel[slot];
        double tts = pool.tts[slot];
        int rate_sensor = (int)(meta >> META_RATE_SHIFT);
        double relax[3] = { 0., 0., 0. };
        if (tts > 0.) { relax_rates(m, rate_sensor, p, relax); }
        double ttm = 0.;
        psim_stream st;
        psim_stream_init(&st, seed, PSIM_TAG_TRANSPORT, (uint32_t)slot, (uint32_t)step);
        int fate_cause = 0;
        int events = 0;
        for (;;) {
            if (++events > PSIM_MAX_EVENTS_PER_STEP) {
                c.stalled++;
                fate_cause = FATE_STALLED;
                break;
            }
            if (tts <= 0.) {
                rate_sensor = m.cells[p.cell].sensor;
                relax_rates(m, rate_sensor, p, relax);
                tts = SCALING_FACTOR * -psim_log(psim_urand(&st)) / (0. + relax[0] + relax[1] + relax[2]);
            }
            if (ttm <= 0.) { ttm = (double)(p.step + 1) * m.sim_time / (double)m.N - age; }
            double drift_time = (ttm < tts) ? ttm : tts;
            const int sensor = m.cells[p.cell].sensor;
            double drifted;
            if (handle_impacts(m, p, drift_time, sensor, drifted, st, c)) {
                if (m.cells[p.cell].sensor != sensor) { drift_time = drifted; }
                const double factor = p.vel * (drift_time - drifted);  // Phonon::drift
                p.px += p.dx * factor;
                p.py += p.dy * factor;
                age += drift_time;
                ttm -= drift_time;
                tts -= drift_time;
                if (ttm == 0.) {
                    c.steps++;
                    events = 0;
                    if (++p.step < m.N) {
                        psim_stream_set_step(&st, (uint32_t)p.step);
                        if (p.step >= m.adj) {  // Sensor::updateHeatParams, sensor.cpp:43-52
                            const int64_t idx = p.step - m.adj;
                            if (idx < m.R && tally.rec) {
                                const int s = m.cells[p.cell].sensor;
                                const int e = (int)(p.step - k0 - 1) * S + s;
                                const double sg = (double)p.sign;
                                atomicAdd(&lcnt[e], p.sign);
                                atomicAdd(&lfx[e], p.dx * p.vel * sg);
                                atomicAdd(&lfy[e], p.dy * p.vel * sg);
                                c.tallies++;
                            }
                        }
                        if (p.step == k1) { break; }
                    } else {
                        c.expired++;
                        fate_cause = FATE_EXPIRED;
                        break;
                    }
                } else if (!m.phasor && tts == 0.) {
                    c.scatters++;
                    scatter(m, p, relax, st);
                } else {
                    tts = 0.;
                }
            } else {
                c.absorbed++;
                fate_cause = FATE_ABSORBED;
                break;
            }
        }
        if (fate_cause) {
            pool.cell[slot] = -1;
            fate_write(fate, slot, fate_cause == FATE_EXPIRED ? p.step - 1 : p.step, fate_cause, p.px, p.py);
        } else {
            pool.px[slot] = p.px;
            pool.py[slot] = p.py;
            pool.dx[slot] = p.dx;
            pool.dy[slot] = p.dy;
            pool.freq[slot] = p.freq;
            pool.vel[slot] = p.vel;
            pool.age[slot] = age;
            pool.tts[slot] = tts;
            pool.cell[slot] = p.cell;
            pool.meta[slot] = (uint32_t)p.bin | (p.pol == POL_TA ? META_TA : 0u) | (p.sign < 0 ? META_NEG : 0u)
                              | META_STARTED | ((uint32_t)rate_sensor << META_RATE_SHIFT);
        }
    }

    // Counters: wave sums, then one global atomic per wave and counter.
    const uint32_t v[8] = { c.steps, c.scatters, c.impacts, c.relocations, c.absorbed, c.expired, c.tallies,
        c.stalled };
    const int slots[8] = { PSIM_CNT_PHONON_STEPS, PSIM_CNT_SCATTERS, PSIM_CNT_IMPACTS, PSIM_CNT_RELOCATIONS,
        PSIM_CNT_ABSORBED, PSIM_CNT_EXPIRED, PSIM_CNT_TALLIES, PSIM_CNT_STALLED };
    for (int i = 0; i < 8; ++i) {
        const uint32_t s = wave_sum(v[i]);
        if ((threadIdx.x % WAVE) == 0 && s) { atomicAdd(&counters[slots[i]], (unsigned long long)s); }
    }
    if (WS) {
        __syncthreads();
        const int64_t base = (int64_t)blockIdx.x * WS;
        for (int i = threadIdx.x; i < WS; i += BLOCK) {
            tally.pcnt[base + i] = lcnt[i];
            tally.pfx[base + i] = lfx[i];
            tally.pfy[base + i] = lfy[i];
        }
    }
}

// Sums the block partials of one window into the tallies: entry e = w * S + s of window step
// k0 + 1 + w goes to tally index (k0 + 1 + w) - adj of sensor s.
__global__ void __launch_bounds__(BLOCK) reduce_kernel(const int32_t* __restrict__ pcnt,
    const double* __restrict__ pfx, const double* __restrict__ pfy, int n_blocks, int WS, int S, int64_t k0,
    int64_t adj, int64_t R, int64_t* __restrict__ counts, double* __restrict__ flux) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= WS) { return; }
    const int w = e / S, s = e % S;
    const int64_t idx = k0 + 1 + w - adj;
    if (idx < 0 || idx >= R) { return; }
    int64_t cnt = 0;
    double fx = 0., fy = 0.;
    for (int g = 0; g < n_blocks; ++g) {
        const int64_t k = (int64_t)g * WS + e;
        cnt += pcnt[k];
        fx += pfx[k];
        fy += pfy[k];
    }
    const int64_t t = (int64_t)s * R + idx;
    counts[t] += cnt;
    flux[2 * t] += fx;
    flux[2 * t + 1] += fy;
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
    psim_run_params params{};
    Model m{};
    std::vector<void*> model_allocs;
    int32_t n_sensors = 0;
    // run buffers
    Pool pool{};
    int64_t pool_cap = 0;
    int64_t* d_counts = nullptr;
    double* d_flux = nullptr;
    unsigned long long* d_counters = nullptr;
    int32_t* d_pcnt = nullptr;
    double* d_pfx = nullptr;
    double* d_pfy = nullptr;
    int64_t partial_cap = 0;
    bool fate_enabled = false;
    Fate fate{ nullptr, nullptr, nullptr };
    int64_t fate_cap = 0;
    int32_t window = 0;
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
    void* ps[] = { c->pool.px, c->pool.py, c->pool.dx, c->pool.dy, c->pool.freq, c->pool.vel, c->pool.age, c->pool.tts,
        c->pool.cell, c->pool.meta };
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

void free_partials(psim_gpu_ctx* c) {
    if (c->d_pcnt) { (void)hipFree(c->d_pcnt); }
    if (c->d_pfx) { (void)hipFree(c->d_pfx); }
    if (c->d_pfy) { (void)hipFree(c->d_pfy); }
    c->d_pcnt = nullptr;
    c->d_pfx = nullptr;
    c->d_pfy = nullptr;
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

// Measurement steps per transport launch: the LDS tally of a recording window holds
// W x S entries of 20 bytes.
int32_t window_for(const psim_gpu_ctx* c) {
    if (c->window > 0) { return c->window; }
    const int64_t by_lds = std::max<int64_t>(1, LDS_TALLY_BYTES / (20 * std::max(1, c->n_sensors)));
    return (int32_t)std::min<int64_t>(64, by_lds);
}

int run_impl(psim_gpu_ctx* c, uint64_t seed) {
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
        c->pool.px = dalloc<double>(cap);
        c->pool.py = dalloc<double>(cap);
        c->pool.dx = dalloc<double>(cap);
        c->pool.dy = dalloc<double>(cap);
        c->pool.freq = dalloc<double>(cap);
        c->pool.vel = dalloc<double>(cap);
        c->pool.age = dalloc<double>(cap);
        c->pool.tts = dalloc<double>(cap);
        c->pool.cell = dalloc<int32_t>(cap);
        c->pool.meta = dalloc<uint32_t>(cap);
        if (!c->pool.px || !c->pool.py || !c->pool.dx || !c->pool.dy || !c->pool.freq || !c->pool.vel || !c->pool.age
            || !c->pool.tts || !c->pool.cell || !c->pool.meta) {
            free_pool(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for the phonon pool");
        }
        c->pool_cap = (int64_t)cap;
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
    }
    const Fate fate = c->fate_enabled ? c->fate : Fate{ nullptr, nullptr, nullptr };
    const int32_t W = window_for(c);
    const int64_t blocks_needed = (n + BLOCK - 1) / BLOCK;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(blocks_needed, (int64_t)c->n_cu * 8));
    const int64_t WS = (int64_t)W * S;
    if (c->partial_cap < grid * WS) {
        free_partials(c);
        const size_t cap = (size_t)(grid * WS);
        c->d_pcnt = dalloc<int32_t>(cap);
        c->d_pfx = dalloc<double>(cap);
        c->d_pfy = dalloc<double>(cap);
        if (!c->d_pcnt || !c->d_pfx || !c->d_pfy) {
            free_partials(c);
            return fail(c, PSIM_E_DEVICE, "hipMalloc failed for tally partials");
        }
        c->partial_cap = (int64_t)cap;
    }
    HIP_CHECK(c, hipMemsetAsync(c->d_counts, 0, sizeof(int64_t) * (size_t)std::max<int64_t>(1, S * R), c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->d_flux, 0, sizeof(double) * (size_t)std::max<int64_t>(1, 2 * S * R), c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * PSIM_NUM_COUNTERS, c->stream));
    HIP_CHECK(c, hipEventRecord(c->ev[0], c->stream));
    if (n > 0) {
        hipLaunchKernelGGL(emit_kernel, dim3((unsigned)blocks_needed), dim3(BLOCK), 0, c->stream, c->m, c->pool, n,
            seed, c->d_counters, fate);
        HIP_CHECK(c, hipGetLastError());
    }
    c->step_ms = 0.;
    c->step_launches = 0;
    const int64_t N = P.num_measurements;
    for (int64_t k0 = 0; k0 < N && n > 0; k0 += W) {
        const int64_t k1 = std::min<int64_t>(N, k0 + W);
        // window steps k0+1 .. k1 map to tally indices (k - adj) in [0, R)
        const bool rec = (k1 >= P.step_adjustment) && (k0 + 1 - P.step_adjustment < R);
        Tally t{ c->d_pcnt, c->d_pfx, c->d_pfy, rec ? 1 : 0, W };
        const size_t lds = rec ? (size_t)WS * 20 : 0;
        HIP_CHECK(c, hipEventRecord(c->ev[1], c->stream));
        hipLaunchKernelGGL(transport_kernel, dim3(grid), dim3(BLOCK), lds, c->stream, c->m, c->pool, n, k0, k1, seed,
            t, c->d_counters, fate);
        HIP_CHECK(c, hipGetLastError());
        HIP_CHECK(c, hipEventRecord(c->ev[2], c->stream));
        if (rec) {
            const int rblocks = (int)((WS + BLOCK - 1) / BLOCK);
            hipLaunchKernelGGL(reduce_kernel, dim3(rblocks), dim3(BLOCK), 0, c->stream, c->d_pcnt, c->d_pfx, c->d_pfy,
                grid, (int)WS, S, k0, P.step_adjustment, R, c->d_counts, c->d_flux);
            HIP_CHECK(c, hipGetLastError());
        }
        HIP_CHECK(c, hipEventSynchronize(c->ev[2]));
        float ms = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
        c->step_ms += ms;
        c->step_launches++;
    }
    HIP_CHECK(c, hipEventRecord(c->ev[3], c->stream));
    HIP_CHECK(c, hipEventSynchronize(c->ev[3]));
    float total = 0.f;
    HIP_CHECK(c, hipEventElapsedTime(&total, c->ev[0], c->ev[3]));
    c->run_ms = total;
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
    if (!c->d_counters) {
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
    if (c->d_counts) { (void)hipFree(c->d_counts); }
    if (c->d_flux) { (void)hipFree(c->d_flux); }
    if (c->d_counters) { (void)hipFree(c->d_counters); }
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
    if (p->num_measurements <= 0 || p->num_record < 0 || p->step_time <= 0. || p->total_phonons < 0
        || p->total_phonons > (int64_t)UINT32_MAX) {
        return fail(c, PSIM_E_INVALID, "bad run parameters");
    }
    if (g->n_sensors > 3200) {
        return fail(c, PSIM_E_INVALID, "at most 3200 sensors (one window step of tallies must fit in LDS)");
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
    }
    int64_t expect = 0;
    for (int i = 0; i < s->n_sources; ++i) {
        const psim_source& src = s->sources[i];
        if (src.cell < 0 || src.cell >= g->n_cells || src.table < 0 || src.table >= t->n_tables || src.count < 0
            || src.offset != expect || (src.kind != 0 && src.kind != 1)) {
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
    PSIM_UPLOAD(sensors, g->sensors, g->n_sensors);
    PSIM_UPLOAD(mats, t->materials, t->n_materials);
    PSIM_UPLOAD(freq, t->freq, nb * t->n_materials);
    PSIM_UPLOAD(vel_la, t->vel_la, nb * t->n_materials);
    PSIM_UPLOAD(vel_ta, t->vel_ta, nb * t->n_materials);
    PSIM_UPLOAD(cdf, t->cdf, nb * t->n_tables);
    PSIM_UPLOAD(frac, t->la_frac, nb * t->n_tables);
    PSIM_UPLOAD(sources, s->sources, s->n_sources);
#undef PSIM_UPLOAD
    m.n_sources = s->n_sources;
    m.n_sensors = g->n_sensors;
    m.full = p->full;
    m.phasor = p->phasor;
    m.N = p->num_measurements;
    m.adj = p->step_adjustment;
    m.R = p->num_record;
    m.sim_time = p->sim_time;
    m.step_time = p->step_time;
    m.t_eq = p->t_eq;
    c->m = m;
    c->params = *p;
    const int64_t SR = (int64_t)g->n_sensors * p->num_record;
    if (c->n_sensors != g->n_sensors || !c->d_counts) {
        if (c->d_counts) { (void)hipFree(c->d_counts); }
        if (c->d_flux) { (void)hipFree(c->d_flux); }
        c->d_counts = nullptr;
        c->d_flux = nullptr;
    }
    if (c->d_counts) {
        (void)hipFree(c->d_counts);
        (void)hipFree(c->d_flux);
    }
    c->d_counts = dalloc<int64_t>((size_t)std::max<int64_t>(1, SR));
    c->d_flux = dalloc<double>((size_t)std::max<int64_t>(1, 2 * SR));
    if (!c->d_counts || !c->d_flux) { return fail(c, PSIM_E_DEVICE, "hipMalloc failed for tallies"); }
    c->n_sensors = g->n_sensors;
    c->uploaded = true;
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
    c->window = steps;
    return PSIM_OK;
}

}  // extern "C"
