// C++17 host of the MI355X phonon engine (see host_model.hpp).  file:line citations refer to
// the reference, GwGibson/Psim.  Built with -ffp-contract=off so every expression rounds as the
// reference's does.
#include "host_model.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "json.hpp"
#include "psim_rng.h"

namespace psim {

namespace {

constexpr double GEOEPS = DBL_EPSILON * 1E9;  // utils.h:10
constexpr double PI = 3.1415926535897932384626433832795028841971693993751058209749445923;  // utils.h:11
constexpr double HBAR = 1.054517e-34;  // material.cpp:12
constexpr double BOLTZ = 1.38065e-23;  // material.cpp:13
constexpr double TEMP_BOUND_EPS = 10.;  // model.cpp:19
constexpr double PHASOR_TEMP_BOUND_EPS = TEMP_BOUND_EPS * 100.;  // model.cpp:20
constexpr double SS_STEPS_PERCENT = 0.1;  // model.cpp:22
constexpr float TEMP_INTERVAL = .1F;  // model.cpp:25
constexpr double INV_EPS = .0001;  // sensorInterpreter.cpp:9
constexpr std::size_t INV_MAX_ITERS = 40;  // sensorInterpreter.cpp:10
constexpr uint64_t RUN_SEED_STRIDE = 0x9E3779B97F4A7C15ull;

// ---- geometry (geometry.cpp) -----------------------------------------------------------

bool ptEq(Point a, Point b) {  // :34-39
    const double dx = a.x - b.x;
    const double dy = a.y - b.y;
    return dx * dx + dy * dy < GEOEPS * GEOEPS;
}
double slopeOf(Point a, Point b) { return (std::abs(a.x - b.x) < GEOEPS) ? 0. : (a.y - b.y) / (a.x - b.x); }
double cross(Point a, Point b) { return a.x * b.y - b.x * a.y; }
double dot(Point a, Point b) { return a.x * b.x + a.y * b.y; }
Point sub(Point a, Point b) { return { a.x - b.x, a.y - b.y }; }
bool onLine(const Line& l, Point p) { return std::fabs(cross(sub(l.p2, l.p1), sub(p, l.p1))) < GEOEPS; }
bool rightOfLine(const Line& l, Point p) { return cross(sub(l.p2, l.p1), sub(p, l.p1)) < 0.; }
bool segCrossesLine(const Line& a, const Line& b) {
    return onLine(a, b.p1) || onLine(a, b.p2) || ((rightOfLine(a, b.p1) ^ rightOfLine(a, b.p2)) != 0);
}
bool boxesIntersect(const Line& a, const Line& b) {
    return a.bl.x <= b.tr.x - GEOEPS && a.tr.x >= b.bl.x + GEOEPS && a.bl.y <= b.tr.y - GEOEPS
           && a.tr.y >= b.bl.y + GEOEPS;
}
bool relEquals(double a, double b) { return std::fabs(a - b) <= ((std::fabs(a) < std::fabs(b) ? std::fabs(b) : std::fabs(a)) * GEOEPS); }

bool intersection(const Line& s, const Line& o, Point& poi) {  // :104-138
    if (!(boxesIntersect(s, o) && segCrossesLine(s, o) && segCrossesLine(o, s))) { return false; }
    auto vert = [&](Point p1, Point p3, Point p4) {
        if (relEquals(p3.x, p4.x)) { return false; }
        const double m = slopeOf(p3, p4);
        const double b = p3.y - m * p3.x;
        poi = { p1.x, m * p1.x + b };
        return true;
    };
    if (relEquals(s.p1.x, s.p2.x)) { return vert(s.p1, o.p1, o.p2); }
    if (relEquals(o.p1.x, o.p2.x)) { return vert(o.p1, s.p1, s.p2); }
    if (relEquals(s.slope, o.slope)) { return false; }
    const double x = (o.intercept - s.intercept) / (s.slope - o.slope);
    poi = { x, s.slope * x + s.intercept };
    return true;
}

Point normalOf(const Line& l, int norm_sign) {  // :97-100
    norm_sign = (norm_sign >= 0) ? 1 : -1;
    return { norm_sign * (l.p2.y - l.p1.y) / l.length, -norm_sign * (l.p2.x - l.p1.x) / l.length };
}

bool lineContains(const Line& s, const Line& o) {  // :74-86
    if (onLine(s, o.p1) && onLine(s, o.p2) && s.length >= o.length) {
        const auto [l1_x_min, l1_x_max] = std::minmax(s.p1.x, s.p2.x);
        const auto [l2_x_min, l2_x_max] = std::minmax(o.p1.x, o.p2.x);
        const auto [l1_y_min, l1_y_max] = std::minmax(s.p1.y, s.p2.y);
        const auto [l2_y_min, l2_y_max] = std::minmax(o.p1.y, o.p2.y);
        return (l1_x_max >= l2_x_max - GEOEPS && l1_x_min <= l2_x_min + GEOEPS && l1_y_max >= l2_y_max - GEOEPS
                && l1_y_min <= l2_y_min + GEOEPS);
    }
    return false;
}

bool lineOverlaps(const Line& s, const Line& o) {  // :56-71
    if (ptEq(s.p1, o.p1) && ptEq(s.p2, o.p2)) { return true; }
    auto mm = [](Point a, Point b) {
        return (a.x + a.y - (b.x + b.y) >= 0.) ? std::pair{ b, a } : std::pair{ a, b };
    };
    const auto [pa, pb] = mm(s.p1, s.p2);
    const auto [pc, pd] = mm(o.p1, o.p2);
    if (!(ptEq(pd, pa) || ptEq(pb, pc))) {
        if ((pd.x + pd.y - (pa.x + pa.y) >= 0.) && (pb.x + pb.y - (pc.x + pc.y) >= 0.)) { return true; }
    }
    return false;
}

std::string fmtPoint(Point p) {  // geometry.cpp:29-32
    std::ostringstream os;
    os << "Point (" << p.x << ", " << p.y << ')';
    return os.str();
}

std::string fmtTriangle(Point a, Point b, Point c) {  // geometry.cpp:268-271
    return "Triangle: [" + fmtPoint(a) + ", " + fmtPoint(b) + ", " + fmtPoint(c) + "]";
}

// CompositeSurfaceError message, compositeSurface.cpp:100-108
[[noreturn]] void surfaceConflict(const Line& main, const Line& inc) {
    throw std::runtime_error(
        "An existing surface conflicts with the location of the incoming surface.\nExisting surface points are: "
        + fmtPoint(main.p1) + ' ' + fmtPoint(main.p2) + "\nIncoming surface points are: " + fmtPoint(inc.p1) + ' '
        + fmtPoint(inc.p2) + '\n');
}

bool triValid(Point p1, Point p2, Point p3) {  // :150-158
    return !((std::fabs(slopeOf(p1, p2)) == std::fabs(slopeOf(p2, p3))
                 && std::fabs(slopeOf(p2, p3)) == std::fabs(slopeOf(p3, p1)))
             || ptEq(p1, p2) || ptEq(p2, p3) || ptEq(p3, p1));
}

double triArea(Point p1, Point p2, Point p3) {  // :252-259
    const double a = Line(p1, p2).length, b = Line(p2, p3).length, c = Line(p3, p1).length;
    const double p = (a + b + c) / 2.;
    return std::sqrt(p * (p - a) * (p - b) * (p - c));
}

bool triContainsPoint(const Triangle& t, Point p) {  // :189-215
    auto area_or_zero = [](Point a, Point b, Point c) { return triValid(a, b, c) ? triArea(a, b, c) : 0.; };
    const Point bp1 = sub(t.p2, t.p1), bp2 = sub(t.p3, t.p1);
    const double d00 = dot(bp1, bp1), d01 = dot(bp1, bp2), d11 = dot(bp2, bp2);
    const double denom = d00 * d11 - d01 * d01;
    const Point bp3 = sub(p, t.p1);
    const double d20 = dot(bp3, bp1), d21 = dot(bp3, bp2);
    const double u = (d11 * d20 - d01 * d21) / denom;
    const double v = (d00 * d21 - d01 * d20) / denom;
    if (u >= GEOEPS && u <= 1. + GEOEPS && v >= GEOEPS && v <= 1. + GEOEPS && u + v <= 1. + GEOEPS) {
        if ((triArea(t.p1, t.p2, t.p3) - area_or_zero(p, t.p2, t.p3) + area_or_zero(t.p1, p, t.p3)
                + area_or_zero(t.p1, t.p2, p))
            < GEOEPS) {
            return true;
        }
    }
    return false;
}

bool triIntersects(const Triangle& a, const Triangle& b) {  // :163-185
    const Line la[3] = { Line(a.p1, a.p2), Line(a.p2, a.p3), Line(a.p3, a.p1) };
    const Line lb[3] = { Line(b.p1, b.p2), Line(b.p2, b.p3), Line(b.p3, b.p1) };
    for (const auto& l1 : la) {
        for (const auto& l2 : lb) {
            Point poi;
            if (intersection(l1, l2, poi)
                && !(ptEq(poi, l1.p1) || ptEq(poi, l1.p2) || ptEq(poi, l2.p1) || ptEq(poi, l2.p2))) {
                return true;
            }
        }
    }
    return false;
}

bool triEq(const Triangle& a, const Triangle& b) {  // :244-249
    auto has = [](const Triangle& t, Point p) { return ptEq(p, t.p1) || ptEq(p, t.p2) || ptEq(p, t.p3); };
    return has(b, a.p1) && has(b, a.p2) && has(b, a.p3);
}

bool triClockwise(const Triangle& t) {  // :217-222
    const double a = (t.p2.x - t.p1.x) * (t.p2.y + t.p1.y);
    const double b = (t.p3.x - t.p2.x) * (t.p3.y + t.p2.y);
    const double c = (t.p1.x - t.p3.x) * (t.p1.y + t.p3.y);
    return a + b + c >= 0.;
}

// libstdc++'s std::reduce / transform_reduce order for random-access ranges (bits/numeric),
// which the reference's std::execution::seq calls reach.
double reduce4(const double* a, std::size_t n) {
    double init = 0.;
    std::size_t i = 0;
    for (; n - i >= 4; i += 4) {
        const double v1 = a[i] + a[i + 1];
        const double v2 = a[i + 2] + a[i + 3];
        init = init + (v1 + v2);
    }
    for (; i < n; ++i) { init = init + a[i]; }
    return init;
}

std::pair<double, double> avgAndStdError(const std::vector<double>& data) {  // sensorInterpreter.cpp:33-46
    const double size = static_cast<double>(data.size());
    const double avg = reduce4(data.data(), data.size()) / size;
    std::vector<double> sq(data.size());
    for (std::size_t i = 0; i < data.size(); ++i) { sq[i] = (avg - data[i]) * (avg - data[i]); }
    const double std_dev = std::sqrt(reduce4(sq.data(), sq.size()) / size);
    return { avg, std_dev / std::sqrt(size) };
}

}  // namespace

Line::Line(Point a, Point b) : p1(a), p2(b) {  // geometry.cpp:45-53
    slope = slopeOf(a, b);
    intercept = a.y - slope * a.x;
    double blx = a.x, bly = a.y, trx = b.x, try_ = b.y;
    if (b.x <= a.x) { std::swap(blx, trx); }
    if (b.y <= a.y) { std::swap(bly, try_); }
    bl = { blx, bly };
    tr = { trx, try_ };
    length = std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
    if (ptEq(a, b)) {  // LineError, geometry.cpp:333-338
        throw std::runtime_error("Cannot create a line using 2 identical points -> " + fmtPoint(a) + ' ' + fmtPoint(b));
    }
}

Triangle::Triangle(Point a, Point b, Point c) : p1(a), p2(b), p3(c) {
    if (!triValid(a, b, c)) {  // TriangleError, geometry.cpp:340-345
        throw std::runtime_error("These 3 points do not allow for a valid triangle -> " + fmtTriangle(a, b, c));
    }
}

// ---- Material (material.cpp) -------------------------------------------------------------

namespace {
double getK(double freq, const std::array<double, 3>& c) {  // :86-91 (pow(x, 2) == x * x)
    const double d = c[1] * c[1] - 4. * c[0] * (c[2] - freq);
    const double a = (-c[1] - std::sqrt(d)) / (2. * c[0]);
    const double b = (-c[1] + std::sqrt(d)) / (2. * c[0]);
    return (a < b) ? a : b;
}
double getGv(double freq, const std::array<double, 3>& c) { return 2. * c[0] * getK(freq, c) + c[1]; }  // :93-95
}  // namespace

Material::Material(std::size_t id, const DispersionData& d, const RelaxationData& r)  // :20-51
    : id_(id), d_(d), r_(r), freq_width_(std::max(d.w_max_la, d.w_max_ta) / NUM_FREQ_BINS) {
    const double pi2 = PI * PI;
    for (int i = 0; i < NUM_FREQ_BINS; ++i) {
        const double freq = (2 * i + 1) * freq_width_ / 2.;
        freq_[i] = freq;
        const double la_gv = getGv(freq, d.LA_data);
        vel_la_[i] = la_gv;
        const double kl = getK(freq, d.LA_data);
        dens_la_[i] = kl * kl / 2. / pi2 / la_gv;
        const double ta_gv = getGv(freq, d.TA_data);
        if (!std::isnan(ta_gv)) {
            vel_ta_[i] = ta_gv;
            const double kt = getK(freq, d.TA_data);
            dens_ta_[i] = kt * kt / pi2 / ta_gv;
        }
    }
}

void Material::phononDist(double temp, bool ta, double* out) const {  // :184-204
    const auto& dens = ta ? dens_ta_ : dens_la_;
    const double cc = HBAR / (BOLTZ * temp);
    for (int i = 0; i < NUM_FREQ_BINS; ++i) {
        const double freq = freq_[i];
        double dist = freq * HBAR / std::expm1(cc * freq) * freq_width_ * dens[i];
        if (!full_) { dist *= cc * freq * std::exp(cc * freq) / (std::expm1(cc * freq) * temp); }
        out[i] = dist;
    }
}

void Material::initializeTables(double low, double high) {  // :101-123
    if (built_ && low == built_low_ && high == built_high_ && full_ == built_full_) { return; }
    const auto num_steps = static_cast<std::size_t>((high - low) / static_cast<double>(TEMP_INTERVAL));
    temps_.assign(num_steps, 0.);
    double n = -1.;
    for (auto& t : temps_) { t = low + static_cast<double>(TEMP_INTERVAL) * (n += 1); }
    temps_.push_back(high);
    base_sums_.resize(temps_.size());
    std::array<double, NUM_FREQ_BINS> la{}, ta{};
    for (std::size_t i = 0; i < temps_.size(); ++i) {
        phononDist(temps_[i], false, la.data());
        phononDist(temps_[i], true, ta.data());
        base_sums_[i] = std::accumulate(la.begin(), la.end(), 0.) + std::accumulate(ta.begin(), ta.end(), 0.);
    }
    cache_.clear();
    built_ = true;
    built_full_ = full_;
    built_low_ = low;
    built_high_ = high;
}

std::size_t Material::tempIndex(double temp) const noexcept {  // :241-246
    const auto idx = static_cast<std::size_t>(std::lower_bound(temps_.begin(), temps_.end(), temp) - temps_.begin());
    return std::min(idx, temps_.size() - 1);
}

const Table& Material::table(int kind, std::size_t index) const {  // :111-122, 140-179
    const std::size_t key = index * 3 + static_cast<std::size_t>(kind);
    auto it = cache_.find(key);
    if (it != cache_.end()) { return *it->second; }
    const double temp = temps_.at(index);
    std::array<double, NUM_FREQ_BINS> la{}, ta{};
    phononDist(temp, false, la.data());
    phononDist(temp, true, ta.data());
    auto rates = [&](double freq, bool is_ta) {  // tauNInv + tauUInv + tauIInv, :207-239
        double n, u;
        if (!is_ta) {
            n = r_.b_l * freq * freq * std::pow(temp, 3);
            u = r_.b_l * freq * freq * std::pow(temp, 3);
        } else {
            n = (freq < r_.w) ? r_.b_tn * freq * std::pow(temp, 4) : 0.;
            u = (freq >= r_.w) ? r_.b_tu * freq * freq / std::sinh(HBAR * freq / (temp * BOLTZ)) : 0.;
        }
        const double i = r_.b_i * std::pow(freq, 4);
        return 0. + n + u + i;
    };
    if (kind == 1) {
        for (int i = 0; i < NUM_FREQ_BINS; ++i) {
            la[i] = la[i] * vel_la_[i];
            ta[i] = ta[i] * vel_ta_[i];
        }
    } else if (kind == 2) {
        for (int i = 0; i < NUM_FREQ_BINS; ++i) { la[i] = la[i] * rates(freq_[i], false); }
        for (int i = 0; i < NUM_FREQ_BINS; ++i) { ta[i] = ta[i] * rates(freq_[i], true); }
    }
    auto t = std::make_shared<Table>();
    const double cumul_sum = std::accumulate(la.begin(), la.end(), 0.) + std::accumulate(ta.begin(), ta.end(), 0.);
    t->cdf[0] = (la[0] + ta[0]) / cumul_sum;  // buildCumulDist, :170-179
    t->la_frac[0] = la[0] / (la[0] + ta[0]);
    for (int i = 1; i < NUM_FREQ_BINS; ++i) {
        t->cdf[i] = t->cdf[i - 1] + (la[i] + ta[i]) / cumul_sum;
        t->la_frac[i] = la[i] / (la[i] + ta[i]);
    }
    t->sum = cumul_sum;
    auto& ref = *t;
    cache_.emplace(key, std::move(t));
    return ref;
}

psim_material Material::descriptor() const {
    psim_material m{};
    m.b_l = r_.b_l;
    m.b_tn = r_.b_tn;
    m.b_tu = r_.b_tu;
    m.b_i = r_.b_i;
    m.w = r_.w;
    m.w_max_la = d_.w_max_la;
    m.w_max_ta = d_.w_max_ta;
    m.freq_width = freq_width_;
    return m;
}

// ---- surfaces, sensors, cells ----------------------------------------------------------------

bool CompositeSurface::verify(const Line& l) const {  // compositeSurface.cpp:68-87
    if (!lineContains(line, l)) { return false; }
    for (const auto& t : transitions) {
        if (lineOverlaps(l, t.line)) { surfaceConflict(t.line, l); }
    }
    for (const auto& e : emits) {
        if (lineOverlaps(l, e.line)) { surfaceConflict(e.line, l); }
    }
    return true;
}

Sensor::Sensor(std::size_t id, const Material& mat, SimulationType type, std::size_t num_record, double t_init)
    : inc_energy(num_record, 0), inc_flux(num_record, std::array<double, 2>{ 0., 0. }), id_(id), material_(&mat),
      type_(type), t_init_(t_init), t_steady_(t_init) {}

double Sensor::getInitTemp() const noexcept {  // sensorController.h:63-67, 85-87, 103-105
    return (type_ == SimulationType::SteadyState) ? t_steady_ : t_init_;
}

void Sensor::updateTables() {  // sensorController.cpp:38-50
    table_index_ = material_->tempIndex(t_init_);
    heat_capacity_ = material_->baseSums()[table_index_];
}

Cell::Cell(const Triangle& t, Sensor& sensor, double spec) : tri_(t), sensor_(&sensor) {  // cell.cpp:13-19, 114-124
    spec = std::clamp(spec, 0., 1.);
    const int ns = normSign();
    const Line lines[3] = { Line(t.p1, t.p2), Line(t.p2, t.p3), Line(t.p3, t.p1) };
    for (int i = 0; i < 3; ++i) {
        b_[i].line = lines[i];
        b_[i].normal = normalOf(lines[i], ns);
        b_[i].spec = spec;
    }
    sensor_->addToArea(getArea());
}

double Cell::getArea() const noexcept { return triArea(tri_.p1, tri_.p2, tri_.p3); }
int Cell::normSign() const noexcept { return triClockwise(tri_) ? 1 : -1; }

double Cell::getInitEnergy(double t_eq) const {  // cell.cpp:38-41
    const double init_energy = getArea() * sensor_->getHeatCapacity();
    return (t_eq == 0.) ? init_energy : init_energy * std::fabs(sensor_->getInitTemp() - t_eq);
}

double Cell::getEmitEnergy(double t_eq) const {  // cell.cpp:45-63
    double emit_energy = 0.;
    for (const auto& b : b_) {
        std::vector<double> vals;
        for (const auto& es : b.emits) {
            const double energy = es.line.length * es.duration * sensor_->getMaterial().emitEnergy(es.temp) / 4.;
            vals.push_back((t_eq == 0.) ? energy : energy * std::fabs(es.temp - t_eq));
        }
        emit_energy += reduce4(vals.data(), vals.size());
    }
    return emit_energy;
}

// ---- GpuModelSimulator ---------------------------------------------------------------------

GpuModelSimulator::GpuModelSimulator(std::size_t measurement_steps, double simulation_time, bool phasor_sim)
    : measurement_steps_(measurement_steps), simulation_time_(simulation_time),
      step_time_(simulation_time / static_cast<double>(measurement_steps)), phasor_sim_(phasor_sim) {}

void GpuModelSimulator::initPhononBuilders(std::vector<Cell>& cells, std::vector<Sensor>& sensors, double t_eq,
    double eff_energy) {
    // materials and tables, deduplicated by (material, kind, grid index)
    std::vector<const Material*> mats;
    auto mat_index = [&](const Material& m) {
        for (std::size_t i = 0; i < mats.size(); ++i) {
            if (mats[i] == &m) { return static_cast<int32_t>(i); }
        }
        mats.push_back(&m);
        return static_cast<int32_t>(mats.size() - 1);
    };
    std::vector<std::tuple<const Material*, int, std::size_t>> table_keys;
    auto table_index = [&](const Material& m, int kind, std::size_t idx) {
        for (std::size_t i = 0; i < table_keys.size(); ++i) {
            if (std::get<0>(table_keys[i]) == &m && std::get<1>(table_keys[i]) == kind
                && std::get<2>(table_keys[i]) == idx) {
                return static_cast<int32_t>(i);
            }
        }
        table_keys.emplace_back(&m, kind, idx);
        return static_cast<int32_t>(table_keys.size() - 1);
    };
    auto sensor_index = [&](const Sensor& s) { return static_cast<int32_t>(&s - sensors.data()); };

    std::vector<psim_sensor> d_sensors(sensors.size());
    for (std::size_t i = 0; i < sensors.size(); ++i) {
        const Sensor& s = sensors[i];
        psim_sensor& d = d_sensors[i];
        d.t_init = s.t_init();
        d.t3 = std::pow(s.t_init(), 3);
        d.t4 = std::pow(s.t_init(), 4);
        d.area = s.getArea();
        d.material = mat_index(s.getMaterial());
        d.base_table = table_index(s.getMaterial(), 0, s.tableIndex());
        d.scatter_table = table_index(s.getMaterial(), 2, s.tableIndex());
        // the model's own cap, which a batch keeps for this model's phonons
        d.flight_cap = psim_flight_cap(static_cast<int32_t>(sensors.size()));
    }
    std::vector<psim_cell> d_cells(cells.size());
    std::vector<psim_edge> d_edges(cells.size() * 3);
    std::vector<psim_subsurface> d_subs;
    std::vector<psim_source> d_sources;
    int64_t offset = 0;
    uint32_t budget_index = 0;
    // getPhonons lambda (modelSimulator.cpp:44-49): floor(E / E_eff) plus a Bernoulli draw
    auto get_phonons = [&](double energy) {
        double ipart = 0;
        const double frac = std::modf(energy / eff_energy, &ipart);
        int64_t n = static_cast<int64_t>(ipart);
        psim_stream st;
        psim_stream_init(&st, seed_, psim_stream_tag(PSIM_TAG_BUDGET, (uint32_t)stream_), budget_index++, 0);
        if (psim_urand(&st) < frac) { ++n; }
        return n;
    };
    for (std::size_t ci = 0; ci < cells.size(); ++ci) {
        Cell& cell = cells[ci];
        const Triangle& t = cell.triangle();
        d_cells[ci] = psim_cell{ t.p1.x, t.p1.y, t.p2.x, t.p2.y, t.p3.x, t.p3.y, sensor_index(cell.sensor()), 0 };
        const Material& mat = cell.sensor().getMaterial();
        {  // cell-origin builder
            const int64_t n = get_phonons(cell.getInitEnergy(t_eq));
            psim_source src{};
            src.kind = 0;
            src.cell = static_cast<int32_t>(ci);
            src.table = table_index(mat, 0, cell.sensor().tableIndex());
            src.sign = (cell.sensor().getInitTemp() > t_eq) ? 1 : -1;
            src.offset = offset;
            src.count = n;
            src.stream = stream_;
            offset += n;
            d_sources.push_back(src);
        }
        for (int b = 0; b < 3; ++b) {
            const CompositeSurface& cs = cell.boundaries()[b];
            psim_edge& e = d_edges[3 * ci + b];
            e.p1x = cs.line.p1.x;
            e.p1y = cs.line.p1.y;
            e.p2x = cs.line.p2.x;
            e.p2y = cs.line.p2.y;
            e.slope = cs.line.slope;
            e.intercept = cs.line.intercept;
            e.blx = cs.line.bl.x;
            e.bly = cs.line.bl.y;
            e.trx = cs.line.tr.x;
            e.try_ = cs.line.tr.y;
            e.nx = cs.normal.x;
            e.ny = cs.normal.y;
            e.spec = cs.spec;
            e.sub_begin = static_cast<int32_t>(d_subs.size());
            e.n_trans = static_cast<int32_t>(cs.transitions.size());
            e.n_emit = static_cast<int32_t>(cs.emits.size());
            for (const auto& ts : cs.transitions) {
                psim_subsurface s{};
                s.p1x = ts.line.p1.x;
                s.p1y = ts.line.p1.y;
                s.p2x = ts.line.p2.x;
                s.p2y = ts.line.p2.y;
                s.nx = ts.normal.x;
                s.ny = ts.normal.y;
                s.kind = 0;
                s.target = static_cast<int32_t>(ts.cell);
                const Material& tm = cells[ts.cell].sensor().getMaterial();
                s.same_material = (tm.id() == mat.id()) ? 1 : 0;
                s.max_freq_la = tm.max_freq_la();
                s.max_freq_ta = tm.max_freq_ta();
                d_subs.push_back(s);
            }
            for (const auto& es : cs.emits) {
                psim_subsurface s{};
                s.p1x = es.line.p1.x;
                s.p1y = es.line.p1.y;
                s.p2x = es.line.p2.x;
                s.p2y = es.line.p2.y;
                s.nx = es.normal.x;
                s.ny = es.normal.y;
                s.spec = es.spec;
                s.start = es.start;
                s.duration = es.duration;
                s.kind = 1;
                d_subs.push_back(s);
                // surface-origin builder (energy as modelSimulator.cpp:67-70)
                const double factor = mat.emitEnergy(es.temp) * es.duration * es.line.length / 4.;
                const double energy = (t_eq == 0.) ? factor : factor * std::fabs(t_eq - es.temp);
                psim_source src{};
                src.kind = 1;
                src.cell = static_cast<int32_t>(ci);
                src.table = table_index(mat, 1, mat.tempIndex(es.temp));
                src.sign = (es.temp > t_eq) ? 1 : -1;
                src.offset = offset;
                src.count = get_phonons(energy);
                src.stream = stream_;
                offset += src.count;
                src.start = es.start;
                src.duration = es.duration;
                src.p1x = es.line.p1.x;
                src.p1y = es.line.p1.y;
                src.p2x = es.line.p2.x;
                src.p2y = es.line.p2.y;
                src.nx = es.normal.x;
                src.ny = es.normal.y;
                d_sources.push_back(src);
            }
        }
    }
    total_phonons_ = offset;
    // tables
    std::vector<psim_material> d_mats;
    std::vector<double> freq, vel_la, vel_ta, cdf, frac;
    for (const Material* m : mats) {
        d_mats.push_back(m->descriptor());
        freq.insert(freq.end(), m->frequencies().begin(), m->frequencies().end());
        vel_la.insert(vel_la.end(), m->velocitiesLA().begin(), m->velocitiesLA().end());
        vel_ta.insert(vel_ta.end(), m->velocitiesTA().begin(), m->velocitiesTA().end());
    }
    for (const auto& [m, kind, idx] : table_keys) {
        const Table& t = m->table(kind, idx);
        cdf.insert(cdf.end(), t.cdf.begin(), t.cdf.end());
        frac.insert(frac.end(), t.la_frac.begin(), t.la_frac.end());
    }
    const bool full = !mats.empty() && mats.front()->full();
    psim_run_params p{};
    p.num_measurements = static_cast<int64_t>(measurement_steps_);
    p.step_adjustment = static_cast<int64_t>(step_adjustment_);
    p.num_record = static_cast<int64_t>(num_record_);
    p.total_phonons = total_phonons_;
    p.sim_time = simulation_time_;
    p.step_time = step_time_;
    p.t_eq = t_eq;
    p.full = full ? 1 : 0;
    p.phasor = phasor_sim_ ? 1 : 0;
    source_counts_.clear();
    for (const auto& src : d_sources) { source_counts_.push_back(src.count); }
    flat_ = FlatModel{ std::move(d_cells), std::move(d_edges), std::move(d_subs), std::move(d_sensors), std::move(d_mats),
        std::move(freq), std::move(vel_la), std::move(vel_ta), std::move(cdf), std::move(frac), std::move(d_sources), p,
        static_cast<int32_t>(table_keys.size()) };
    if (!ctx_) { return; }  // budget only (no engine attached)
    const FlatModel& f = flat_;
    psim_geom g{ static_cast<int32_t>(f.cells.size()), static_cast<int32_t>(f.sensors.size()),
        static_cast<int32_t>(f.subs.size()), 0, f.cells.data(), f.edges.data(), f.subs.data(), f.sensors.data() };
    psim_tables t{ static_cast<int32_t>(f.mats.size()), f.n_tables, f.mats.data(), f.freq.data(), f.vel_la.data(),
        f.vel_ta.data(), f.cdf.data(), f.frac.data() };
    psim_sources s{ static_cast<int32_t>(f.sources.size()), 0, f.sources.data() };
    if (psim_gpu_upload(ctx_, &g, &t, &s, &p) != PSIM_OK) {
        throw std::runtime_error(std::string("engine upload failed: ") + psim_gpu_last_error(ctx_));
    }
}

void GpuModelSimulator::runSimulation(double /*t_eq*/, std::vector<Sensor>& sensors) {
    const std::size_t S = sensors.size(), R = num_record_;
    std::vector<int64_t> counts(S * R);
    std::vector<double> flux(2 * S * R);
    if (psim_gpu_run(ctx_, seed_, counts.data(), flux.data(), counters_.data()) != PSIM_OK) {
        throw std::runtime_error(std::string("engine run failed: ") + psim_gpu_last_error(ctx_));
    }
    for (std::size_t s = 0; s < S; ++s) {
        for (std::size_t k = 0; k < R; ++k) {
            sensors[s].inc_energy[k] = counts[s * R + k];
            sensors[s].inc_flux[k] = { flux[2 * (s * R + k)], flux[2 * (s * R + k) + 1] };
        }
    }
}

// ---- Model ----------------------------------------------------------------------------------

Model::Model(const ModelParams& p)  // model.cpp:31-44
    : num_runs_(p.num_runs), num_cells_(p.num_cells), measurement_steps_(p.measurement_steps),
      num_phonons_(p.num_phonons), simulation_time_(p.simulation_time), t_eq_(p.t_eq), phasor_sim_(p.phasor_sim),
      simulator_(p.measurement_steps, p.simulation_time, p.phasor_sim) {
    cells_.reserve(p.num_cells);
    sensors_.reserve(p.num_sensors);
}

void Model::setSimulationType(SimulationType type, std::size_t step_interval) {  // model.cpp:46-70
    sim_type_ = type;
    const double n = static_cast<double>(measurement_steps_);
    if (type == SimulationType::Transient || type == SimulationType::Periodic) {
        start_step_ = static_cast<std::size_t>(n - n * SS_STEPS_PERCENT);
        if (step_interval == 0) {
            throw std::runtime_error("Step interval of 0 is invalid for transient and periodic simulations.\n");
        }
    }
    if (type == SimulationType::Transient && t_eq_ == 0) {
        throw std::runtime_error("Transient simulations must be run using the deviational approach.\n");
    }
    if (type == SimulationType::SteadyState) {
        simulator_.setStepAdjustment(static_cast<std::size_t>(n - n * SS_STEPS_PERCENT));
        if (step_interval > 0) { throw std::runtime_error("Step interval > 0 is invalid for steady-state simulations\n"); }
    } else {
        simulator_.setStepAdjustment(0);
    }
    step_interval_ = step_interval;
    num_record_ = (type == SimulationType::SteadyState) ? static_cast<std::size_t>(n * SS_STEPS_PERCENT)
                                                         : measurement_steps_;
    simulator_.setRecordLength(num_record_);
}

void Model::addMaterial(const std::string& name, const Material& material) {  // model.cpp:72-78
    for (const auto& m : materials_) {
        if (m.first == name) { throw std::runtime_error("A duplicate material name was detected.\n"); }
    }
    materials_.emplace_back(name, std::make_unique<Material>(material));
}

void Model::addSensor(std::size_t id, const std::string& material_name, double t_init, SimulationType type) {
    for (const auto& s : sensors_) {  // model.cpp:81-95
        if (s.getID() == id) { throw std::runtime_error("Sensor with this ID already exists\n"); }
    }
    const Material* mat = nullptr;
    for (const auto& m : materials_) {
        if (m.first == material_name) { mat = m.second.get(); }
    }
    if (!mat) { throw std::runtime_error("_Map_base::at"); }  // std::unordered_map::at, model.cpp:91
    if (sensors_.size() == sensors_.capacity()) { throw std::runtime_error("Too many sensors\n"); }
    const double n = static_cast<double>(measurement_steps_);
    const std::size_t steps_to_record =
        (type == SimulationType::SteadyState) ? static_cast<std::size_t>(n * SS_STEPS_PERCENT) : measurement_steps_;
    sensors_.emplace_back(id, *mat, type, steps_to_record, t_init);
}

Sensor& Model::getSensor(std::size_t id) {  // model.cpp:188-194
    for (auto& s : sensors_) {
        if (s.getID() == id) { return s; }
    }
    throw std::runtime_error("Sensor does not exist\n");
}

void Model::addCell(const Triangle& triangle, std::size_t sensor_id, double spec) {  // model.cpp:98-113
    if (cells_.size() >= num_cells_) { throw std::runtime_error("Too many cells\n"); }
    cells_.emplace_back(triangle, getSensor(sensor_id), spec);
    const std::size_t inc = cells_.size() - 1;
    std::size_t identical = 0;
    for (std::size_t i = 0; i < cells_.size(); ++i) {
        const Triangle& a = cells_[inc].triangle();
        const Triangle& b = cells_[i].triangle();
        if (!triEq(a, b)) {
            // Cell::validate, cell.cpp:21-25, with IntersectError / OverlapError (cell.cpp:150-164);
            // IntersectError(existing = this cell, incoming = the other one) prints them that way.
            // The messages are formatted only when thrown: this loop visits every pair of cells.
            const auto ta = [&] { return fmtTriangle(a.p1, a.p2, a.p3); };
            const auto tb = [&] { return fmtTriangle(b.p1, b.p2, b.p3); };
            if (triIntersects(a, b)) {
                throw std::runtime_error("Incoming " + tb() + "\nintersects\nExisting" + ta() + '\n');
            }
            if (triContainsPoint(a, b.p1) || triContainsPoint(a, b.p2) || triContainsPoint(a, b.p3)) {
                throw std::runtime_error(tb() + "\nis contained within\n" + ta() + '\n');
            }
            if (triContainsPoint(b, a.p1) || triContainsPoint(b, a.p2) || triContainsPoint(b, a.p3)) {
                throw std::runtime_error(ta() + "\nis contained within\n" + tb() + '\n');
            }
            // Cell::findTransitionSurface, cell.cpp:81-98
            const std::array<Line, 3> l_inc = { cells_[inc].boundaries()[0].line, cells_[inc].boundaries()[1].line,
                cells_[inc].boundaries()[2].line };
            const std::array<Line, 3> l_oth = { cells_[i].boundaries()[0].line, cells_[i].boundaries()[1].line,
                cells_[i].boundaries()[2].line };
            auto set_transition = [&](std::size_t owner, const Line& l, std::size_t target) {
                const int ns = cells_[owner].normSign();
                for (auto& b : cells_[owner].boundaries()) {
                    if (b.verify(l)) {
                        b.transitions.push_back(TransitionSurface{ l, normalOf(l, ns), target });
                        return true;
                    }
                }
                return false;
            };
            for (const auto& l1 : l_inc) {
                for (const auto& l2 : l_oth) {
                    if (lineContains(l1, l2)) {
                        if (!set_transition(inc, l2, i) || !set_transition(i, l2, inc)) { surfaceConflict(l1, l2); }
                    } else if (lineContains(l2, l1)) {
                        if (!set_transition(inc, l1, i) || !set_transition(i, l1, inc)) { surfaceConflict(l2, l1); }
                    }
                }
            }
        } else if (++identical == 2) {
            throw std::runtime_error("Duplicate cell detected.\n");
        }
    }
}

bool Model::setEmitSurface(const Point& p1, const Point& p2, double temp, double duration, double start_time) {
    if ((start_time < 0. || start_time >= simulation_time_)  // model.cpp:125-138
        || (duration < 0. || duration > simulation_time_ - start_time)) {
        throw std::runtime_error("Transient Surface start_time or duration specifications are invalid.\n");
    }
    if ((start_time > 0. || duration < simulation_time_) && sim_type_ != SimulationType::Transient) {
        throw std::runtime_error("Cannot add a transient surface to a non transient simulation.\n");
    }
    const Line line(p1, p2);
    for (auto& cell : cells_) {  // Cell::setEmitSurface, cell.cpp:29-35
        for (auto& b : cell.boundaries()) {
            if (b.verify(line)) {
                b.emits.push_back(EmitSurface{ line, b.normal, b.spec, temp, duration, start_time });
                return true;
            }
        }
    }
    return false;
}

double Model::getTotalInitialEnergy() const {  // model.cpp:196-201
    std::vector<double> per_cell(cells_.size());
    for (std::size_t i = 0; i < cells_.size(); ++i) {
        per_cell[i] = cells_[i].getInitEnergy(t_eq_) + cells_[i].getEmitEnergy(t_eq_);
    }
    return reduce4(per_cell.data(), per_cell.size());
}

std::pair<double, double> Model::setTemperatureBounds() {  // model.cpp:203-218
    std::vector<double> temps;
    for (const auto& cell : cells_) {
        temps.push_back(cell.sensor().getInitTemp());
        for (const auto& b : cell.boundaries()) {
            for (const auto& es : b.emits) { temps.push_back(es.temp); }
        }
    }
    const auto [mn, mx] = std::minmax_element(temps.begin(), temps.end());
    const double bound = phasor_sim_ ? PHASOR_TEMP_BOUND_EPS : TEMP_BOUND_EPS;
    lb_ = std::max(*mn - bound, 0.);
    ub_ = *mx + bound;
    return { *mn, *mx };
}

std::vector<double> Model::findTemperature(const Sensor& s, double eff_energy, std::size_t start) const {
    const auto& energies = s.inc_energy;  // sensorInterpreter.cpp:80-112
    std::vector<double> temps(energies.size() - std::min(start, energies.size()));
    const double area = s.getArea();
    const Material& mat = s.getMaterial();
    for (std::size_t i = 0; i < temps.size(); ++i) {
        const double energy = eff_energy * static_cast<double>(energies[start + i]);
        if (t_eq_ != 0.) {
            temps[i] = energy / (area * s.getHeatCapacity()) + t_eq_;
        } else {
            double temp = 0., ub = ub_, lb = lb_;
            std::size_t iter = 0;
            while ((ub - lb >= INV_EPS) && (++iter != INV_MAX_ITERS)) {
                temp = (ub + lb) / 2.;
                const double de = (mat.baseEnergy(temp) * area) - energy;
                (de < 0.) ? lb = temp : ub = temp;
            }
            temps[i] = temp;
        }
    }
    return temps;
}

SensorMeasurements Model::scaleHeatParams(const Sensor& s, double eff_energy) const {  // sensorInterpreter.cpp:19-66
    SensorMeasurements sm;
    sm.id = s.getID();
    sm.final_temps = findTemperature(s, eff_energy, 0);
    if (!sm.final_temps.empty()) { sm.final_temps.front() = s.getInitTemp(); }
    const double f_factor = eff_energy / s.getArea();
    sm.final_fluxes.resize(s.inc_flux.size());
    std::vector<double> fx(s.inc_flux.size()), fy(s.inc_flux.size());
    for (std::size_t k = 0; k < s.inc_flux.size(); ++k) {
        sm.final_fluxes[k] = { s.inc_flux[k][0] * f_factor, s.inc_flux[k][1] * f_factor };
        fx[k] = sm.final_fluxes[k][0];
        fy[k] = sm.final_fluxes[k][1];
    }
    std::tie(sm.t_steady, sm.std_t_steady) = avgAndStdError(sm.final_temps);
    std::tie(sm.x_flux, sm.std_x_flux) = avgAndStdError(fx);
    std::tie(sm.y_flux, sm.std_y_flux) = avgAndStdError(fy);
    return sm;
}

double Model::prepare(uint64_t seed) {  // model.cpp:144-153, 160
    if (cells_.empty() || sensors_.empty()) { throw std::runtime_error("model has no cells or sensors"); }
    const auto [mn, mx] = setTemperatureBounds();
    for (auto& m : materials_) { m.second->initializeTables(mn, mx); }
    for (auto& s : sensors_) { s.updateTables(); }
    const double eff0 = getTotalInitialEnergy() / static_cast<double>(num_phonons_);
    simulator_.setSeed(seed);
    simulator_.initPhononBuilders(cells_, sensors_, t_eq_, eff0);
    return eff0;
}

double Model::prepareBudget(uint64_t seed) {
    simulator_.setEngine(nullptr);
    return prepare(seed);
}

double Model::prepareAndUpload(psim_gpu_ctx* ctx, uint64_t seed) {
    if (!ctx) { throw std::runtime_error("no engine context"); }
    simulator_.setEngine(ctx);
    return prepare(seed);
}

void Model::runSimulation(psim_gpu_ctx* ctx, uint64_t seed) {  // model.cpp:141-182
    if (!ctx) { throw std::runtime_error("no engine context"); }
    runs_.clear();
    for (std::size_t run = 0; run < num_runs_; ++run) {
        simulator_.setEngine(ctx);
        const double eff0 = prepare(seed + run * RUN_SEED_STRIDE);
        simulator_.runSimulation(t_eq_, sensors_);
        endRun(eff0, simulator_.totalPhonons(), simulator_.counters());
    }
}

void Model::endRun(double eff0, int64_t total_phonons, const std::array<int64_t, PSIM_NUM_COUNTERS>& counters) {
    RunRecord rec;
    rec.eff_energy = eff0;
    rec.total_phonons = total_phonons;
    rec.counters = counters;
    const std::size_t R = num_record_;
    for (const auto& s : sensors_) {
        rec.counts.insert(rec.counts.end(), s.inc_energy.begin(), s.inc_energy.end());
        for (const auto& f : s.inc_flux) {
            rec.flux.push_back(f[0]);
            rec.flux.push_back(f[1]);
        }
    }
    // Model::resetRequired (model.cpp:250-272): getFinalTemp with the budget's energy sets
    // each controller's t_steady_; refresh() then recomputes the energy per phonon with it.
    if (sim_type_ != SimulationType::Transient) {
        for (auto& s : sensors_) {
            double ft = 0.;
            if (s.getArea() != 0.) {
                const auto temps = findTemperature(s, eff0, start_step_);
                ft = std::accumulate(temps.begin(), temps.end(), 0.) / static_cast<double>(R - start_step_);
            }
            s.setSteadyTemp(ft);
        }
    }
    const double eff1 = getTotalInitialEnergy() / static_cast<double>(num_phonons_);
    for (const auto& s : sensors_) { rec.measurements.push_back(scaleHeatParams(s, eff1)); }
    std::sort(rec.measurements.begin(), rec.measurements.end(),
        [](const auto& a, const auto& b) { return a.id < b.id; });
    runs_.push_back(std::move(rec));
    // Model::reset(true): controllers back to t_init, tallies cleared
    for (auto& s : sensors_) {
        s.setSteadyTemp(s.t_init());
        std::fill(s.inc_energy.begin(), s.inc_energy.end(), 0);
        std::fill(s.inc_flux.begin(), s.inc_flux.end(), std::array<double, 2>{ 0., 0. });
    }
    simulator_.reset();
}

void Model::runBatch(const std::vector<Model*>& models, psim_gpu_ctx* ctx, uint64_t seed,
    std::array<int64_t, PSIM_NUM_COUNTERS>& counters) {
    if (!ctx) { throw std::runtime_error("no engine context"); }
    if (models.empty()) { throw std::runtime_error("empty batch"); }
    if (models.size() > 65536) { throw std::runtime_error("at most 65536 models per batch"); }
    const Model& m0 = *models.front();
    for (const Model* m : models) {
        if (!m) { throw std::runtime_error("null model in batch"); }
        if (m->measurement_steps_ != m0.measurement_steps_ || m->simulation_time_ != m0.simulation_time_
            || m->sim_type_ != m0.sim_type_ || m->step_interval_ != m0.step_interval_
            || m->phasor_sim_ != m0.phasor_sim_ || m->num_runs_ != 1) {
            throw std::runtime_error("batched models must share num_measurements, sim_time, sim_type, step_interval "
                                     "and phasor_sim, and have num_runs 1");
        }
    }
    // Each model's pre-run state and budget, exactly as its single run with stream m builds them.
    // The models share nothing, so their tables and budgets are built on a few host threads.
    std::vector<double> eff0(models.size());
    std::vector<int32_t> own_stream(models.size());
    std::vector<std::string> errors(models.size());
    std::atomic<std::size_t> next{ 0 };
    auto work = [&]() {
        for (std::size_t i = next++; i < models.size(); i = next++) {
            Model& m = *models[i];
            try {
                own_stream[i] = m.simulator_.stream();
                m.runs_.clear();
                m.simulator_.setEngine(nullptr);
                m.simulator_.setStream(static_cast<int32_t>(i));
                eff0[i] = m.prepare(seed);
            } catch (const std::exception& e) { errors[i] = e.what(); }
        }
    };
    const std::size_t n_threads = std::min<std::size_t>(models.size(),
        std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (std::size_t t = 1; t < n_threads; ++t) { pool.emplace_back(work); }
    work();
    for (auto& th : pool) { th.join(); }
    for (const auto& e : errors) {
        if (!e.empty()) { throw std::runtime_error(e); }
    }
    // One set of descriptors: every index shifted by the sizes of the models before it.
    FlatModel all;
    std::vector<int64_t> first_id(models.size()), sensor_base(models.size());
    for (std::size_t i = 0; i < models.size(); ++i) {
        const FlatModel& f = models[i]->simulator_.flat();
        if (f.params.full != m0.simulator_.flat().params.full) {
            throw std::runtime_error("batched models must all be full or all deviational");
        }
        const int32_t cell0 = static_cast<int32_t>(all.cells.size());
        const int32_t sensor0 = static_cast<int32_t>(all.sensors.size());
        const int32_t sub0 = static_cast<int32_t>(all.subs.size());
        const int32_t mat0 = static_cast<int32_t>(all.mats.size());
        const int32_t table0 = all.n_tables;
        const int64_t id0 = all.params.total_phonons;
        first_id[i] = id0;
        sensor_base[i] = sensor0;
        for (psim_cell c : f.cells) {
            c.sensor += sensor0;
            all.cells.push_back(c);
        }
        for (psim_edge e : f.edges) {
            e.sub_begin += sub0;
            all.edges.push_back(e);
        }
        for (psim_subsurface ss : f.subs) {
            if (ss.kind == 0) { ss.target += cell0; }
            all.subs.push_back(ss);
        }
        for (psim_sensor se : f.sensors) {
            se.material += mat0;
            se.base_table += table0;
            se.scatter_table += table0;
            all.sensors.push_back(se);
        }
        all.mats.insert(all.mats.end(), f.mats.begin(), f.mats.end());
        all.freq.insert(all.freq.end(), f.freq.begin(), f.freq.end());
        all.vel_la.insert(all.vel_la.end(), f.vel_la.begin(), f.vel_la.end());
        all.vel_ta.insert(all.vel_ta.end(), f.vel_ta.begin(), f.vel_ta.end());
        all.cdf.insert(all.cdf.end(), f.cdf.begin(), f.cdf.end());
        all.frac.insert(all.frac.end(), f.frac.begin(), f.frac.end());
        all.n_tables += f.n_tables;
        for (psim_source src : f.sources) {
            src.cell += cell0;
            src.table += table0;
            src.offset += id0;
            src.id_base = id0;
            all.sources.push_back(src);
        }
        all.params.total_phonons += f.params.total_phonons;
    }
    psim_run_params p = m0.simulator_.flat().params;
    p.total_phonons = all.params.total_phonons;
    psim_geom g{ static_cast<int32_t>(all.cells.size()), static_cast<int32_t>(all.sensors.size()),
        static_cast<int32_t>(all.subs.size()), 0, all.cells.data(), all.edges.data(), all.subs.data(),
        all.sensors.data() };
    psim_tables t{ static_cast<int32_t>(all.mats.size()), all.n_tables, all.mats.data(), all.freq.data(),
        all.vel_la.data(), all.vel_ta.data(), all.cdf.data(), all.frac.data() };
    psim_sources s{ static_cast<int32_t>(all.sources.size()), 0, all.sources.data() };
    if (psim_gpu_upload(ctx, &g, &t, &s, &p) != PSIM_OK) {
        throw std::runtime_error(std::string("engine upload failed: ") + psim_gpu_last_error(ctx));
    }
    const std::size_t S = all.sensors.size(), R = m0.num_record_;
    std::vector<int64_t> counts(S * R);
    std::vector<double> flux(2 * S * R);
    if (psim_gpu_run(ctx, seed, counts.data(), flux.data(), counters.data()) != PSIM_OK) {
        throw std::runtime_error(std::string("engine run failed: ") + psim_gpu_last_error(ctx));
    }
    // Each model's slice of the tallies into its sensors, then the rest of its run.
    const std::array<int64_t, PSIM_NUM_COUNTERS> none{};
    for (std::size_t i = 0; i < models.size(); ++i) {
        Model& m = *models[i];
        for (std::size_t si = 0; si < m.sensors_.size(); ++si) {
            const std::size_t row = static_cast<std::size_t>(sensor_base[i]) + si;
            for (std::size_t k = 0; k < R; ++k) {
                m.sensors_[si].inc_energy[k] = counts[row * R + k];
                m.sensors_[si].inc_flux[k] = { flux[2 * (row * R + k)], flux[2 * (row * R + k) + 1] };
            }
        }
        m.endRun(eff0[i], m.simulator_.flat().params.total_phonons, none);
        m.simulator_.setStream(own_stream[i]);
    }
}

void Model::exportResults(const std::filesystem::path& filepath, double time) const {  // outputManager.cpp:13-38
    if (runs_.empty()) { throw std::runtime_error("no results to export"); }
    const bool ss = sim_type_ == SimulationType::SteadyState;
    auto out_path = filepath;
    out_path.replace_extension(".txt");
    out_path.replace_filename((ss ? "ss_" : "per_") + out_path.filename().string());
    std::ofstream out(out_path, std::ios_base::trunc);
    if (!out) { throw std::runtime_error("cannot write " + out_path.string()); }
    const auto now = std::chrono::floor<std::chrono::seconds>(std::chrono::system_clock::now());
    const auto tt = std::chrono::system_clock::to_time_t(now);
    std::tm tm_time{};
    localtime_r(&tt, &tm_time);
    std::ostringstream date;
    date << std::put_time(&tm_time, "%Y-%m-%d %X");
    out << (ss ? "Steady State" : "Periodic") << " Results from " << filepath.filename() << " @ " << date.str()
        << " - Time Taken " << time << "[s] over " << num_runs_ << " runs\n";
    // OutputManager::calculateAverages (outputManager.cpp:40-70), summing runs in run order and
    // also averaging the per-step series (the reference drops them).
    const std::size_t S = runs_.front().measurements.size();
    std::vector<SensorMeasurements> avg(S);
    for (const auto& r : runs_) {
        for (std::size_t i = 0; i < S; ++i) {
            const auto& m = r.measurements[i];
            avg[i].id = m.id;
            avg[i].t_steady += m.t_steady;
            avg[i].std_t_steady += m.std_t_steady;
            avg[i].x_flux += m.x_flux;
            avg[i].std_x_flux += m.std_x_flux;
            avg[i].y_flux += m.y_flux;
            avg[i].std_y_flux += m.std_y_flux;
            if (avg[i].final_temps.empty()) {
                avg[i].final_temps.assign(m.final_temps.size(), 0.);
                avg[i].final_fluxes.assign(m.final_fluxes.size(), { 0., 0. });
            }
            for (std::size_t k = 0; k < m.final_temps.size(); ++k) {
                avg[i].final_temps[k] += m.final_temps[k];
                avg[i].final_fluxes[k][0] += m.final_fluxes[k][0];
                avg[i].final_fluxes[k][1] += m.final_fluxes[k][1];
            }
        }
    }
    const double nr = static_cast<double>(runs_.size());
    for (auto& m : avg) {
        m.t_steady /= nr;
        m.std_t_steady /= nr;
        m.x_flux /= nr;
        m.std_x_flux /= nr;
        m.y_flux /= nr;
        m.std_y_flux /= nr;
        for (auto& t : m.final_temps) { t /= nr; }
        for (auto& f : m.final_fluxes) {
            f[0] /= nr;
            f[1] /= nr;
        }
    }
    if (ss) {  // steadyStateExport, outputManager.cpp:72-78
        for (const auto& m : avg) {
            out << m.t_steady << ' ' << m.std_t_steady << ' ' << m.x_flux << ' ' << m.std_x_flux << ' ' << m.y_flux
                << ' ' << m.std_y_flux << '\n';
        }
        return;
    }
    // periodicExport, outputManager.cpp:82-112
    const std::size_t si = step_interval_;
    const std::size_t steps = avg.back().final_temps.size();
    for (std::size_t step = 0; si > 0 && si <= steps && step < steps - si + 1; step += si) {
        out << step + si / 2 << '\n' << S << '\n';
        for (const auto& m : avg) {
            const double temp = std::accumulate(m.final_temps.begin() + static_cast<long>(step),
                                    m.final_temps.begin() + static_cast<long>(step + si), 0.)
                                / static_cast<double>(si);
            std::vector<double> fx(si), fy(si);
            for (std::size_t k = 0; k < si; ++k) {
                fx[k] = m.final_fluxes[step + k][0];
                fy[k] = m.final_fluxes[step + k][1];
            }
            const double xf = reduce4(fx.data(), si) / static_cast<double>(si);
            const double yf = reduce4(fy.data(), si) / static_cast<double>(si);
            out << temp << ' ' << xf << ' ' << yf << '\n';
        }
    }
}

std::unique_ptr<Model> deserialize(const std::string& text) {  // inputManager.cpp:14-115
    const psim_json::Value j = psim_json::parse(text.c_str());
    const auto& settings = j.at("settings");
    const auto& sensors = j.at("sensors").array();
    const auto& cells = j.at("cells").array();
    const int st = static_cast<int>(settings.at("sim_type").number());
    const SimulationType type = st == 1   ? SimulationType::Periodic
                                : st == 2 ? SimulationType::Transient
                                          : SimulationType::SteadyState;
    const double t_eq = settings.at("t_eq").number();
    ModelParams params;
    params.num_cells = cells.size();
    params.num_sensors = sensors.size();
    params.measurement_steps = static_cast<std::size_t>(settings.at("num_measurements").number());
    params.num_phonons = static_cast<std::size_t>(settings.at("num_phonons").number());
    params.simulation_time = settings.at("sim_time").number();
    params.t_eq = t_eq;
    const auto& ph = settings.at("phasor_sim");
    params.phasor_sim = ph.kind == psim_json::Value::Bool && ph.b;
    if (settings.contains("num_runs")) { params.num_runs = static_cast<std::size_t>(settings.at("num_runs").number()); }
    auto model = std::make_unique<Model>(params);
    if (type == SimulationType::SteadyState) {
        model->setSimulationType(type);
    } else {
        model->setSimulationType(type, static_cast<std::size_t>(settings.at("step_interval").number()));
    }
    std::size_t mid = 0;
    for (const auto& m : j.at("materials").array()) {
        const auto& d = m.at("d_data");
        const auto& r = m.at("r_data");
        DispersionData dd;
        for (int i = 0; i < 3; ++i) {
            dd.LA_data[i] = d.at("la_data").array().at(i).number();
            dd.TA_data[i] = d.at("ta_data").array().at(i).number();
        }
        dd.w_max_la = d.at("max_freq_la").number();
        dd.w_max_ta = d.at("max_freq_ta").number();
        RelaxationData rd{ r.at("b_l").number(), r.at("b_tn").number(), r.at("b_tu").number(), r.at("b_i").number(),
            r.at("w").number() };
        Material mat(mid++, dd, rd);
        if (t_eq == 0.) { mat.setFullSimulation(); }
        model->addMaterial(m.at("name").string(), mat);
    }
    for (const auto& s : sensors) {
        model->addSensor(static_cast<std::size_t>(s.at("id").number()), s.at("material").string(),
            s.at("t_init").number(), type);
    }
    auto pt = [](const psim_json::Value& v) { return Point{ v.at("x").number(), v.at("y").number() }; };
    for (const auto& c : cells) {
        const auto& t = c.at("triangle");
        model->addCell(Triangle(pt(t.at("p1")), pt(t.at("p2")), pt(t.at("p3"))),
            static_cast<std::size_t>(c.at("sensorID").number()), c.at("specularity").number());
    }
    for (const auto& s : j.at("emit_surfaces").array()) {
        if (!model->setEmitSurface(pt(s.at("p1")), pt(s.at("p2")), s.at("temp").number(), s.at("duration").number(),
                s.at("start_time").number())) {
            throw std::runtime_error("Unable to add emitting surface.\n");
        }
    }
    return model;
}

}  // namespace psim
