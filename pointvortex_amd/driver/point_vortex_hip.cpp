// point_vortex_hip.cpp -- a PointVortex run driver on the point_vortex_hip C-ABI.
//
// `point_vortex_hip [params]` runs what the reference's executables run (src/main.cpp:112-389)
// with the state resident on the GPU from start to end:
//   * the params file (strict `key value` lines, the reference's keys, aliases, defaults and
//     validation messages: src/read.cpp:55-211, src/params.h:37-78);
//   * the initial state from an initial-condition file with its metadata checks, or the
//     built-in ring / seeded periodic populations (src/read.cpp:212-305, src/main.cpp:73-108);
//   * fixed-step RK4 or adaptive DOPRI5 on the device, each step landing on the next output
//     event or the end time (src/main.cpp:156-191, 310-377);
//   * dipole events after every accepted step through pv_dipole_process: the candidate search
//     runs on the device state and only the state of an event crosses to the host; the
//     generator's text goes into checkpoints (src/main.cpp:327-337, src/dipole.cpp:42-182);
//   * trajectory.csv, the 19-column diagnostics.csv and the console line in the reference's
//     formats (src/print.cpp:24-72, 147-173), run records (src/print.cpp:73-146), versioned
//     checkpoints written atomically and read back from versions 1 to 6
//     (src/checkpoint.cpp:23-239), run-directory ownership and overwrite protection
//     (src/main.cpp:24-53, 193-243).
// Invariants are computed on the device (pv_state_invariants).  PV_HIP_DEVICE selects the GPU
// and PV_HIP_PRECISION the arithmetic (fp64, the default, or mixed), as for the link-time
// backend adapter.
#include "pv_hip.h"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <optional>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace fs = std::filesystem;

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr std::size_t kSizeMax = std::numeric_limits<std::size_t>::max();

// ---------------------------------------------------------------- the C-ABI as exceptions

// Status codes back into the reference's exception types.
void check(pv_status status) {
    switch (status) {
    case PV_OK:
        return;
    case PV_ERR_INVALID_ARGUMENT:
        throw std::invalid_argument(pv_last_error());
    case PV_ERR_OUT_OF_RANGE:
        throw std::out_of_range(pv_last_error());
    case PV_ERR_LOGIC:
        throw std::logic_error(pv_last_error());
    default:
        throw std::runtime_error(pv_last_error());
    }
}

// ---------------------------------------------------------------- parameters

enum class Integrator { rk4, dopri5 };
enum class Reinjection { none, independent, paired };

const char *name_of(Integrator k) { return k == Integrator::rk4 ? "rk4" : "dopri5"; }
const char *name_of(Reinjection r) {
    return r == Reinjection::independent ? "independent"
           : r == Reinjection::paired    ? "paired"
                                         : "none";
}
std::optional<Integrator> integrator_named(const std::string &s) {
    if (s == "rk4")
        return Integrator::rk4;
    if (s == "dopri5")
        return Integrator::dopri5;
    return std::nullopt;
}
std::optional<Reinjection> reinjection_named(const std::string &s) {
    if (s == "none")
        return Reinjection::none;
    if (s == "independent")
        return Reinjection::independent;
    if (s == "paired")
        return Reinjection::paired;
    return std::nullopt;
}

// The run settings and their defaults (src/params.h:39-76).
struct Params {
    std::size_t n = 100;
    double time_step = 1.0e-3, end_time = 1.0, output_time = 0.1;
    std::optional<double> diagnostics_time, checkpoint_time;
    double core_radius = 0.0;
    double atol = 1.0e-10, rtol = 1.0e-8, min_dt = 1.0e-12, max_dt = 0.1;
    int threads = 0;
    Integrator integrator = Integrator::dopri5;
    std::string geometry = "infinite";
    double box_x = 2.0, box_y = 2.0, disk_radius = 1.0;
    int layers = 8;
    std::uint64_t seed = 1234567;
    bool dipoles = false;
    double dipole_distance = 0.01;
    Reinjection reinjection = Reinjection::none;
    std::string initial_file, restart_file, run_directory = "runs/default";
    bool overwrite = false;

    double diagnostics_interval() const { return diagnostics_time.value_or(output_time); }
    double checkpoint_interval() const { return checkpoint_time.value_or(output_time); }
    // The geometry fields a checkpoint records (src/checkpoint.cpp:60-69).
    double length_x() const {
        return geometry == "periodic" ? box_x : (geometry == "disk" ? disk_radius : 0.0);
    }
    double length_y() const { return geometry == "periodic" ? box_y : 0.0; }
    int image_layers() const { return geometry == "periodic" ? layers : 0; }
};

// The reference's validation, in its order and with its messages (src/read.cpp:55-105).
void validate(const Params &p) {
    const auto require = [](bool ok, const char *message) {
        if (!ok)
            throw std::invalid_argument(message);
    };
    for (double v : {p.time_step, p.end_time, p.output_time, p.core_radius, p.atol, p.rtol,
                     p.min_dt, p.max_dt, p.box_x, p.box_y, p.disk_radius, p.dipole_distance})
        require(std::isfinite(v), "simulation parameters must be finite");
    for (double r : {p.core_radius, p.disk_radius, p.dipole_distance})
        require(std::isfinite(r * r) && !(r > 0.0 && r * r == 0.0),
                "radius or distance is outside the supported numeric range");
    require(!(p.n == 0 && p.initial_file.empty() && p.restart_file.empty()), "N must be positive");
    require(p.time_step > 0.0, "timeStep must be positive");
    require(p.end_time >= 0.0, "endTime must be non-negative");
    require(p.output_time > 0.0, "outputTime must be positive");
    require(std::isfinite(p.output_time) && std::isfinite(p.diagnostics_interval()) &&
                p.diagnostics_interval() > 0.0 && std::isfinite(p.checkpoint_interval()) &&
                p.checkpoint_interval() > 0.0,
            "output intervals must be finite and positive");
    require(p.core_radius >= 0.0, "coreRadius must be non-negative");
    require(p.atol > 0.0 && p.rtol >= 0.0, "invalid integration tolerances");
    require(p.min_dt > 0.0 && p.max_dt >= p.min_dt, "invalid timestep bounds");
    require(p.threads >= 0, "numThreads must be non-negative");
    require(p.geometry == "infinite" || p.geometry == "periodic" || p.geometry == "disk",
            "invalid boundaryCondition");
    require(p.box_x > 0.0 && p.box_y > 0.0 && p.disk_radius > 0.0 && p.layers >= 0 &&
                p.layers <= 64,
            "invalid geometry parameters or periodicImageLayers > 64");
    require(p.geometry == "infinite" || p.core_radius == 0.0,
            "periodic and disk boundaries currently require coreRadius 0");
    require(p.dipole_distance > 0.0, "dipoleRemovalDistance must be positive");
    require(p.dipoles || p.reinjection == Reinjection::none,
            "dipoleReinjection requires dipoleRemoval true");
    require(!(p.geometry == "infinite" && p.reinjection != Reinjection::none),
            "dipole reinjection is available only for periodic and disk geometries");
    require(!(p.geometry == "periodic" &&
              std::fabs(p.box_x - p.box_y) > 1e-13 * std::max(p.box_x, p.box_y)),
            "Weiss-McWilliams periodic geometry requires a square box");
    require(!p.run_directory.empty(), "runDirectory must not be empty");
}

// Values are read with the standard conversions the reference uses, so that their error
// texts are the reference's too; the whole token must convert.
double to_double(const std::string &s) {
    std::size_t used = 0;
    const double v = std::stod(s, &used);
    if (used != s.size() || !std::isfinite(v))
        throw std::invalid_argument("expected a finite floating-point value");
    return v;
}
unsigned long long to_unsigned(const std::string &s) {
    if (!s.empty() && s[0] == '-')
        throw std::invalid_argument("expected a non-negative integer");
    std::size_t used = 0;
    const unsigned long long v = std::stoull(s, &used);
    if (used != s.size())
        throw std::invalid_argument("expected an integer");
    return v;
}
int to_int(const std::string &s) {
    std::size_t used = 0;
    const int v = std::stoi(s, &used);
    if (used != s.size())
        throw std::invalid_argument("expected an integer");
    return v;
}

// src/read.cpp:106-211: one `key value` per line, '#' comments, unknown keys are errors.
Params load_params(const std::string &file) {
    std::ifstream in(file);
    if (!in)
        throw std::runtime_error("cannot open parameter file: " + file);
    Params p;
    std::optional<unsigned long long> legacy_steps;
    bool end_given = false;
    using Setter = std::function<void(const std::string &)>;
    const auto dbl = [](double &field) { return Setter([&field](const std::string &v) { field = to_double(v); }); };
    const auto opt = [](std::optional<double> &field) {
        return Setter([&field](const std::string &v) { field = to_double(v); });
    };
    const auto integer = [](int &field) { return Setter([&field](const std::string &v) { field = to_int(v); }); };
    const auto text = [](std::string &field) { return Setter([&field](const std::string &v) { field = v; }); };
    const auto flag = [](const std::string &key, bool &field) {
        return Setter([key, &field](const std::string &v) {
            const bool on = v == "true" || v == "1";
            if (!on && v != "false" && v != "0")
                throw std::invalid_argument(key + " must be true or false");
            field = on;
        });
    };
    const std::map<std::string, Setter> keys = {
        {"timeStep", dbl(p.time_step)},
        {"endTime", [&](const std::string &v) { p.end_time = to_double(v); end_given = true; }},
        {"outputTime", dbl(p.output_time)},
        {"OutputTime", dbl(p.output_time)},
        {"coreRadius", dbl(p.core_radius)},
        {"absoluteTolerance", dbl(p.atol)},
        {"relativeTolerance", dbl(p.rtol)},
        {"minimumTimeStep", dbl(p.min_dt)},
        {"maximumTimeStep", dbl(p.max_dt)},
        {"boxLengthX", dbl(p.box_x)},
        {"boxLengthY", dbl(p.box_y)},
        {"diskRadius", dbl(p.disk_radius)},
        {"dipoleRemovalDistance", dbl(p.dipole_distance)},
        {"diagnosticsTime", opt(p.diagnostics_time)},
        {"checkpointTime", opt(p.checkpoint_time)},
        {"numThreads", integer(p.threads)},
        {"periodicImageLayers", integer(p.layers)},
        {"boundaryCondition", text(p.geometry)},
        {"initialConditionFile", text(p.initial_file)},
        {"restartFile", text(p.restart_file)},
        {"runDirectory", text(p.run_directory)},
        {"N", [&](const std::string &v) {
             const unsigned long long count = to_unsigned(v);
             if (count > kSizeMax)
                 throw std::out_of_range("N is too large");
             p.n = (std::size_t)count;
         }},
        {"numSteps", [&](const std::string &v) { legacy_steps = to_unsigned(v); }},
        {"coreSize", [&](const std::string &v) { p.core_radius = std::sqrt(to_double(v)); }},
        {"randomSeed", [&](const std::string &v) { p.seed = to_unsigned(v); }},
        {"dipoleRemoval", flag("dipoleRemoval", p.dipoles)},
        {"overwriteRun", flag("overwriteRun", p.overwrite)},
        {"dipoleReinjection", [&](const std::string &v) {
             const auto mode = reinjection_named(v);
             if (!mode)
                 throw std::invalid_argument("dipoleReinjection must be none, independent, or paired");
             p.reinjection = *mode;
         }},
        {"integrator", [&](const std::string &v) {
             const auto kind = integrator_named(v);
             if (!kind)
                 throw std::invalid_argument("integrator must be rk4 or dopri5");
             p.integrator = *kind;
         }},
    };
    std::string line;
    for (std::size_t number = 1; std::getline(in, line); ++number) {
        line = line.substr(0, line.find('#'));
        std::istringstream fields(line);
        std::string key, value, extra;
        if (!(fields >> key))
            continue;
        if (!(fields >> value))
            throw std::runtime_error("missing value on parameter line " + std::to_string(number));
        try {
            const auto it = keys.find(key);
            if (it == keys.end())
                throw std::invalid_argument("unknown parameter: " + key);
            it->second(value);
            if (fields >> extra)
                throw std::invalid_argument("unexpected extra value: " + extra);
        } catch (const std::exception &e) {
            throw std::runtime_error("parameter line " + std::to_string(number) + ": " + e.what());
        }
    }
    if (in.bad())
        throw std::runtime_error("failed while reading parameter file: " + file);
    if (legacy_steps && !end_given)
        p.end_time = p.time_step * (double)*legacy_steps;
    validate(p);
    return p;
}

// ---------------------------------------------------------------- vortex state

struct Invariants {
    double circulation = 0, ix = 0, iy = 0, angular = 0, hamiltonian = 0;
};

struct State {
    std::vector<double> x, y, g;
    std::size_t size() const { return x.size(); }
    void resize(std::size_t n) {
        x.resize(n);
        y.resize(n);
        g.resize(n);
    }
    void add(double a, double b, double c) {
        x.push_back(a);
        y.push_back(b);
        g.push_back(c);
    }
};

void require_finite(const State &s) {
    for (std::size_t i = 0; i < s.size(); ++i)
        if (!std::isfinite(s.x[i]) || !std::isfinite(s.y[i]) || !std::isfinite(s.g[i]))
            throw std::invalid_argument("vortex data must be finite");
}

// src/read.cpp:212-247: `x y circulation` rows (commas allowed as separators, '#' comments).
State load_vortices(const std::string &file) {
    std::ifstream in(file);
    if (!in)
        throw std::runtime_error("cannot open initial-condition file: " + file);
    State s;
    std::string line;
    for (std::size_t number = 1; std::getline(in, line); ++number) {
        line = line.substr(0, line.find('#'));
        std::replace(line.begin(), line.end(), ',', ' ');
        std::istringstream fields(line);
        fields >> std::ws;
        if (fields.eof())
            continue;
        double a, b, c;
        std::string extra;
        const bool ok = static_cast<bool>(fields >> a >> b >> c);
        if (!ok || (fields >> extra) || !std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c))
            throw std::runtime_error("invalid initial condition on line " + std::to_string(number));
        s.add(a, b, c);
    }
    if (in.bad())
        throw std::runtime_error("failed while reading initial-condition file: " + file);
    if (s.size() == 0)
        throw std::runtime_error("initial-condition file is empty");
    return s;
}

// `key=value` fields in comments of an initial-condition file (src/read.cpp:248-280).
struct Metadata {
    std::optional<std::string> geometry;
    std::optional<double> box_length, disk_radius;
};

Metadata read_metadata(const std::string &file) {
    std::ifstream in(file);
    if (!in)
        throw std::runtime_error("cannot open initial-condition file: " + file);
    Metadata m;
    std::string line;
    while (std::getline(in, line)) {
        const std::size_t hash = line.find('#');
        if (hash == std::string::npos)
            continue;
        std::istringstream fields(line.substr(hash + 1));
        std::string field;
        while (fields >> field) {
            const std::size_t eq = field.find('=');
            if (eq == std::string::npos)
                continue;
            const std::string key = field.substr(0, eq), value = field.substr(eq + 1);
            try {
                if (key == "geometry")
                    m.geometry = value;
                else if (key == "box_length")
                    m.box_length = to_double(value);
                else if (key == "disk_radius")
                    m.disk_radius = to_double(value);
            } catch (const std::exception &e) {
                throw std::runtime_error(std::string("invalid initial-condition metadata: ") +
                                         e.what());
            }
        }
    }
    return m;
}

bool close_to(double a, double b) { return std::fabs(a - b) <= 1e-13 * std::max(a, b); }

// src/main.cpp:73-108: the file (checked against the params), or a built-in population: the
// seeded +-1 periodic population (src/read.cpp:290-305), else a ring of alternating +-1
// vortices, at radius R/2 in a disk (src/read.cpp:281-289).
State initial_state(const Params &p) {
    if (!p.initial_file.empty()) {
        const Metadata m = read_metadata(p.initial_file);
        if (m.geometry && *m.geometry != p.geometry)
            throw std::invalid_argument("initial-condition geometry is " + *m.geometry +
                                        " but boundaryCondition is " + p.geometry);
        if (p.geometry == "periodic" && m.box_length &&
            (!close_to(*m.box_length, p.box_x) || !close_to(*m.box_length, p.box_y)))
            throw std::invalid_argument(
                "initial-condition box length does not match simulation parameters");
        if (p.geometry == "disk" && m.disk_radius && !close_to(*m.disk_radius, p.disk_radius))
            throw std::invalid_argument(
                "initial-condition disk radius does not match simulation parameters");
        return load_vortices(p.initial_file);
    }
    State s;
    s.resize(p.n);
    if (p.geometry == "periodic") {
        if (p.n % 2 != 0)
            throw std::invalid_argument(
                "the built-in periodic initial condition requires an even N");
        std::mt19937_64 rng(p.seed);
        std::uniform_real_distribution<double> ux(-0.5 * p.box_x, 0.5 * p.box_x);
        std::uniform_real_distribution<double> uy(-0.5 * p.box_y, 0.5 * p.box_y);
        for (std::size_t i = 0; i < p.n; ++i) {
            s.x[i] = ux(rng);
            s.y[i] = uy(rng);
            s.g[i] = i % 2 == 0 ? 1.0 : -1.0;
        }
        return s;
    }
    const double radius = p.geometry == "disk" ? 0.5 * p.disk_radius : 1.0;
    for (std::size_t i = 0; i < p.n; ++i) {
        const double angle = 2.0 * kPi * (double)i / (double)p.n;
        s.x[i] = radius * std::cos(angle);
        s.y[i] = radius * std::sin(angle);
        s.g[i] = i % 2 == 0 ? 1.0 : -1.0;
    }
    return s;
}

// ---------------------------------------------------------------- checkpoints

constexpr const char *kMagic = "POINT_VORTEX_CHECKPOINT";
constexpr unsigned kVersion = 6;

struct Schedule {
    double trajectory_interval = 0, diagnostics_interval = 0, checkpoint_interval = 0;
    double next_diagnostics = 0, next_checkpoint = 0;
};

struct Progress {
    double time, suggested_dt, next_output;
    std::size_t accepted, output_index, event_index;
};

struct DipoleCounts {
    std::size_t removed = 0, reinjected = 0;
    std::string generator;     // text of the reinjection generator's state
};

struct Checkpoint {
    State state;
    Invariants initial, segment;
    bool has_segment = false;
    double time = 0, suggested_dt = 0, next_output = 0;
    bool has_schedule = false;
    Schedule schedule;
    std::size_t accepted = 0, output_index = 0, event_index = 0;
    double core_radius = 0;
    Integrator integrator = Integrator::dopri5;
    std::string geometry = "infinite";
    double length_x = 0, length_y = 0;
    int layers = 0;
    bool dipoles = false;
    double dipole_distance = 0;
    Reinjection reinjection = Reinjection::none;
    DipoleCounts counts;
};

fs::path checkpoint_path(const fs::path &directory, std::size_t index) {
    std::ostringstream name;
    name << "checkpoint_" << std::setw(8) << std::setfill('0') << index << ".dat";
    return directory / name.str();
}

void put(std::ostream &out, const Invariants &v) {
    out << v.circulation << ' ' << v.ix << ' ' << v.iy << ' ' << v.angular << ' '
        << v.hamiltonian << '\n';
}

// Version 6 (src/checkpoint.cpp:23-102): a temporary file, synchronised, then renamed over
// the destination, whose directory is synchronised too.  An existing checkpoint is refused.
void write_checkpoint(const fs::path &directory, const State &s, const Invariants &initial,
                      const Params &p, const Invariants &segment, const DipoleCounts &counts,
                      const Schedule &sched, const Progress &at) {
    require_finite(s);
    fs::create_directories(directory);
    const fs::path destination = checkpoint_path(directory, at.output_index);
    if (fs::exists(destination))
        throw std::runtime_error("refusing to overwrite checkpoint: " + destination.string());
    fs::path temporary = destination;
    temporary += ".tmp";
    {
        std::ofstream out(temporary, std::ios::trunc);
        if (!out)
            throw std::runtime_error("cannot create checkpoint: " + temporary.string());
        out << std::setprecision(17) << kMagic << ' ' << kVersion << '\n'
            << "time " << at.time << '\n'
            << "suggested_time_step " << at.suggested_dt << '\n'
            << "next_output_time " << at.next_output << '\n'
            << "output_schedule " << sched.trajectory_interval << ' '
            << sched.diagnostics_interval << ' ' << sched.checkpoint_interval << ' '
            << sched.next_diagnostics << ' ' << sched.next_checkpoint << '\n'
            << "accepted_steps " << at.accepted << '\n'
            << "output_index " << at.output_index << '\n'
            << "event_index " << at.event_index << '\n'
            << "core_radius " << p.core_radius << '\n'
            << "integrator " << name_of(p.integrator) << '\n'
            << "geometry " << p.geometry << ' ' << p.length_x() << ' ' << p.length_y() << ' '
            << p.image_layers() << '\n'
            << "dipole_config " << p.dipoles << ' ' << p.dipole_distance << ' '
            << name_of(p.reinjection) << '\n'
            << "dipole_counts " << counts.removed << ' ' << counts.reinjected << '\n'
            << "random_engine_state " << counts.generator << '\n'
            << "initial_invariants ";
        put(out, initial);
        out << "segment_invariants ";
        put(out, segment);
        out << "vortex_count " << s.size() << '\n' << "vortices\n";
        for (std::size_t i = 0; i < s.size(); ++i)
            out << s.x[i] << ' ' << s.y[i] << ' ' << s.g[i] << '\n';
        out.flush();
        if (!out)
            throw std::runtime_error("failed while writing checkpoint: " + temporary.string());
    }
    const int fd = ::open(temporary.c_str(), O_RDONLY);
    const bool synced = fd >= 0 && ::fsync(fd) == 0;
    if (fd >= 0)
        ::close(fd);
    if (!synced)
        throw std::runtime_error("cannot synchronize checkpoint: " + temporary.string());
    fs::rename(temporary, destination);
    const int dfd = ::open(directory.c_str(), O_RDONLY | O_DIRECTORY);
    if (dfd >= 0) {
        ::fsync(dfd);
        ::close(dfd);
    }
}

// Versions 1 to 6 with the reference's field order, validation and messages
// (src/checkpoint.cpp:103-239).
Checkpoint read_checkpoint(const fs::path &file) {
    std::ifstream in(file);
    if (!in)
        throw std::runtime_error("cannot open checkpoint: " + file.string());
    std::string magic;
    unsigned version = 0;
    if (!(in >> magic >> version) || magic != kMagic || version < 1 || version > kVersion)
        throw std::runtime_error("unsupported or corrupt checkpoint header");
    Checkpoint c;
    const auto field = [&](const char *name) {
        std::string key;
        if (!(in >> key) || key != name)
            throw std::runtime_error(std::string("checkpoint is missing field: ") + name);
    };
    const auto count = [&]() {
        std::string token;
        if (!(in >> token))
            throw std::runtime_error("checkpoint has a missing integer");
        std::size_t v = 0;
        const auto r = std::from_chars(token.data(), token.data() + token.size(), v);
        if (r.ec != std::errc{} || r.ptr != token.data() + token.size())
            throw std::runtime_error("checkpoint has an invalid non-negative integer");
        return v;
    };
    const auto invariants = [&](Invariants &v) {
        in >> v.circulation >> v.ix >> v.iy >> v.angular >> v.hamiltonian;
    };
    field("time");
    in >> c.time;
    field("suggested_time_step");
    in >> c.suggested_dt;
    field("next_output_time");
    in >> c.next_output;
    if (version >= 5) {
        field("output_schedule");
        in >> c.schedule.trajectory_interval >> c.schedule.diagnostics_interval >>
            c.schedule.checkpoint_interval >> c.schedule.next_diagnostics >>
            c.schedule.next_checkpoint;
        c.has_schedule = true;
    }
    field("accepted_steps");
    c.accepted = count();
    field("output_index");
    c.output_index = count();
    if (version >= 6) {
        field("event_index");
        c.event_index = count();
    } else {
        c.event_index = c.output_index;   // one shared counter before version 6
    }
    field("core_radius");
    in >> c.core_radius;
    field("integrator");
    std::string word;
    in >> word;
    const auto kind = integrator_named(word);
    if (!kind)
        throw std::runtime_error("unsupported checkpoint integrator: " + word);
    c.integrator = *kind;
    if (version >= 2) {
        field("geometry");
        in >> c.geometry >> c.length_x >> c.length_y >> c.layers;
    }
    if (version >= 3) {
        field("dipole_config");
        in >> c.dipoles >> c.dipole_distance >> word;
        const auto mode = reinjection_named(word);
        if (!mode)
            throw std::runtime_error("unsupported checkpoint reinjection mode: " + word);
        c.reinjection = *mode;
        field("dipole_counts");
        c.counts.removed = count();
        c.counts.reinjected = count();
        field("random_engine_state");
        std::getline(in >> std::ws, c.counts.generator);
    }
    field("initial_invariants");
    invariants(c.initial);
    if (version >= 4) {
        field("segment_invariants");
        invariants(c.segment);
        c.has_segment = true;
    }
    field("vortex_count");
    const std::size_t n = count();
    field("vortices");
    for (std::size_t i = 0; i < n; ++i) {        // rows first: a corrupt count allocates nothing
        double a, b, g;
        if (!(in >> a >> b >> g))
            throw std::runtime_error("checkpoint vortex data is truncated or invalid");
        c.state.add(a, b, g);
    }
    if (in >> word)
        throw std::runtime_error("checkpoint contains unexpected trailing data");
    if (in.bad())
        throw std::runtime_error("failed while reading checkpoint");
    in.clear();
    if (!in || !std::isfinite(c.time) || c.time < 0.0 || !std::isfinite(c.suggested_dt) ||
        !(c.suggested_dt > 0.0) || !std::isfinite(c.next_output) || !(c.next_output > c.time) ||
        !std::isfinite(c.core_radius) || c.core_radius < 0.0)
        throw std::runtime_error("checkpoint is truncated or invalid");
    if (c.has_schedule) {
        const Schedule &s = c.schedule;
        for (double iv : {s.trajectory_interval, s.diagnostics_interval, s.checkpoint_interval})
            if (!std::isfinite(iv) || !(iv > 0.0))
                throw std::runtime_error("checkpoint has invalid output intervals");
        for (double next : {s.next_diagnostics, s.next_checkpoint})
            if (!std::isfinite(next) || !(next > c.time))
                throw std::runtime_error("checkpoint has invalid output schedule");
    }
    if (version >= 3 && (!std::isfinite(c.dipole_distance) || c.dipole_distance <= 0.0 ||
                         c.counts.generator.empty()))
        throw std::runtime_error("checkpoint has invalid dipole-removal state");
    for (const Invariants &v : {c.initial, c.segment})
        for (double q : {v.circulation, v.ix, v.iy, v.angular, v.hamiltonian})
            if (!std::isfinite(q))
                throw std::runtime_error("checkpoint contains non-finite invariants");
    if (c.counts.reinjected > c.counts.removed)
        throw std::runtime_error("checkpoint has invalid dipole event counts");
    for (std::size_t i = 0; i < n; ++i)
        if (!std::isfinite(c.state.x[i]) || !std::isfinite(c.state.y[i]) ||
            !std::isfinite(c.state.g[i]))
            throw std::runtime_error("checkpoint contains non-finite vortex data");
    return c;
}

// src/main.cpp:55-71.
bool checkpoint_matches(const Checkpoint &c, const Params &p) {
    return c.core_radius == p.core_radius && c.integrator == p.integrator &&
           c.geometry == p.geometry && c.length_x == p.length_x() &&
           c.length_y == p.length_y() && c.layers == p.image_layers() &&
           c.dipoles == p.dipoles &&
           (!p.dipoles ||
            (c.dipole_distance == p.dipole_distance && c.reinjection == p.reinjection));
}

// ---------------------------------------------------------------- outputs

struct RunPaths {
    fs::path directory, trajectory, diagnostics, checkpoints;
};

RunPaths run_paths(const Params &p) {
    const fs::path d(p.run_directory);
    return {d, d / "trajectory.csv", d / "diagnostics.csv", d / "checkpoints"};
}

bool same_file(const fs::path &a, const fs::path &b) {
    return fs::weakly_canonical(a) == fs::weakly_canonical(b) ||
           (fs::exists(a) && fs::exists(b) && fs::equivalent(a, b));
}

// The artefacts a run owns (src/main.cpp:31-53).
std::vector<fs::path> managed_files(const RunPaths &r) {
    return {r.trajectory, r.diagnostics, r.directory / "resolved_parameters.txt"};
}
std::vector<fs::path> managed_directories(const RunPaths &r) {
    return {r.checkpoints, r.directory / "segments"};
}

bool has_run_output(const RunPaths &r) {
    for (const auto &f : managed_files(r))
        if (fs::exists(f))
            return true;
    for (const auto &d : managed_directories(r))
        if (fs::exists(d))
            return true;
    return false;
}

void remove_run_output(const RunPaths &r) {
    std::error_code error;
    const auto fail = [&](const fs::path &path) {
        throw std::runtime_error("cannot remove previous run output: " + path.string() + ": " +
                                 error.message());
    };
    for (const auto &f : managed_files(r))
        if (fs::remove(f, error), error)
            fail(f);
    for (const auto &d : managed_directories(r))
        if (fs::remove_all(d, error), error)
            fail(d);
}

std::ofstream open_output(const fs::path &file, const char *what) {
    if (fs::exists(file))
        throw std::runtime_error("refusing to overwrite " + std::string(what) + ": " +
                                 file.string());
    if (!file.parent_path().empty())
        fs::create_directories(file.parent_path());
    std::ofstream out(file, std::ios::trunc);
    if (!out)
        throw std::runtime_error("cannot open " + std::string(what) + ": " + file.string());
    out << std::setprecision(17);
    return out;
}

// trajectory.csv and diagnostics.csv (src/print.cpp:24-72).
class Csv {
  public:
    Csv(const RunPaths &r, const Invariants &initial)
        : traj_(open_output(r.trajectory, "output file")),
          diag_(open_output(r.diagnostics, "diagnostics file")), initial_(initial) {
        traj_ << "time,frame,index,x,y,circulation,u,v\n" << std::flush;
        diag_ << "time,frame,circulation,linear_impulse_x,linear_impulse_y,angular_impulse,"
                 "hamiltonian,delta_circulation,delta_linear_impulse_x,delta_linear_impulse_y,"
                 "delta_angular_impulse,delta_hamiltonian,segment_delta_circulation,"
                 "segment_delta_linear_impulse_x,segment_delta_linear_impulse_y,"
                 "segment_delta_angular_impulse,segment_delta_hamiltonian,removed_pairs,"
                 "reinjected_pairs\n"
              << std::flush;
    }
    void trajectory(double t, std::size_t frame, const State &s, const std::vector<double> &u,
                    const std::vector<double> &v) {
        for (std::size_t i = 0; i < s.size(); ++i)
            traj_ << t << ',' << frame << ',' << i << ',' << s.x[i] << ',' << s.y[i] << ','
                  << s.g[i] << ',' << u[i] << ',' << v[i] << '\n';
        traj_.flush();
        if (!traj_)
            throw std::runtime_error("failed while writing trajectory output");
    }
    void diagnostics(double t, std::size_t frame, const Invariants &v, const Invariants &seg,
                     std::size_t removed, std::size_t reinjected) {
        diag_ << t << ',' << frame << ',' << v.circulation << ',' << v.ix << ',' << v.iy << ','
              << v.angular << ',' << v.hamiltonian;
        for (const Invariants *ref : {static_cast<const Invariants *>(&initial_), &seg})
            diag_ << ',' << v.circulation - ref->circulation << ',' << v.ix - ref->ix << ','
                  << v.iy - ref->iy << ',' << v.angular - ref->angular << ','
                  << v.hamiltonian - ref->hamiltonian;
        diag_ << ',' << removed << ',' << reinjected << '\n';
        diag_.flush();
        if (!diag_)
            throw std::runtime_error("failed while writing diagnostics output");
    }

  private:
    std::ofstream traj_, diag_;
    Invariants initial_;
};

// The console line of every diagnostics event (src/print.cpp:147-173).
void print_diagnostics(double t, std::size_t steps, const Invariants &v, const Invariants &initial,
                       const std::string &geometry, const Invariants &seg, std::size_t removed,
                       std::size_t reinjected) {
    std::cout << std::setprecision(10) << "time=" << t << " steps=" << steps
              << " circulation=" << v.circulation
              << " dCirculation=" << v.circulation - initial.circulation
              << " segmentDCirculation=" << v.circulation - seg.circulation
              << " H=" << v.hamiltonian << " dH=" << v.hamiltonian - initial.hamiltonian
              << " segmentDH=" << v.hamiltonian - seg.hamiltonian;
    if (geometry != "disk")
        std::cout << " Ix=" << v.ix << " dIx=" << v.ix - initial.ix
                  << " segmentDIx=" << v.ix - seg.ix << " Iy=" << v.iy
                  << " dIy=" << v.iy - initial.iy << " segmentDIy=" << v.iy - seg.iy;
    if (geometry != "periodic")
        std::cout << " L=" << v.angular << " dL=" << v.angular - initial.angular
                  << " segmentDL=" << v.angular - seg.angular;
    std::cout << " removedPairs=" << removed << " reinjectedPairs=" << reinjected << '\n'
              << std::flush;
}

std::string absolute(const std::string &path) {
    return path.empty() ? std::string() : fs::absolute(path).lexically_normal().string();
}

// resolved_parameters.txt in the run directory and in a new segments/segment_NNNNNNNN
// directory (src/print.cpp:73-146), each written to a temporary and renamed.
void write_run_records(const Params &p, const std::string &param_file, const std::string &details,
                       double start_time, std::size_t start_frame, bool restarting) {
    const RunPaths r = run_paths(p);
    const fs::path segments = r.directory / "segments";
    fs::create_directories(segments);
    fs::path segment;
    for (std::size_t k = 1;; ++k) {
        std::ostringstream name;
        name << "segment_" << std::setw(8) << std::setfill('0') << k;
        segment = segments / name.str();
        if (fs::create_directory(segment))
            break;
    }
    const std::string seg_abs = fs::absolute(segment).lexically_normal().string();
    for (const fs::path &file : {segment / "resolved_parameters.txt",
                                 r.directory / "resolved_parameters.txt"}) {
        const fs::path temporary(file.string() + ".tmp");
        std::ofstream out(temporary, std::ios::trunc);
        if (!out)
            throw std::runtime_error("cannot write run record: " + temporary.string());
        const auto q = [](const std::string &s) { return std::quoted(s); };
        out << std::setprecision(17) << "POINT_VORTEX_RUN_RECORD 1\n"
            << "parameter_file " << q(absolute(param_file)) << '\n'
            << "run_directory " << q(absolute(p.run_directory)) << '\n'
            << "backend HIP\n"
            << details << '\n'
            << "start_time " << start_time << '\n'
            << "start_frame " << start_frame << '\n'
            << "restarting " << restarting << '\n'
            << "restart_file " << q(absolute(p.restart_file)) << '\n'
            << "N " << p.n << '\n'
            << "timeStep " << p.time_step << '\n'
            << "endTime " << p.end_time << '\n'
            << "outputTime " << p.output_time << '\n'
            << "diagnosticsTime " << p.diagnostics_interval() << '\n'
            << "checkpointTime " << p.checkpoint_interval() << '\n'
            << "integrator " << name_of(p.integrator) << '\n'
            << "coreRadius " << p.core_radius << '\n'
            << "numThreads " << p.threads << '\n'
            << "boundaryCondition " << p.geometry << '\n'
            << "boxLengthX " << p.box_x << '\n'
            << "boxLengthY " << p.box_y << '\n'
            << "periodicImageLayers " << p.layers << '\n'
            << "diskRadius " << p.disk_radius << '\n'
            << "randomSeed " << p.seed << '\n'
            << "dipoleRemoval " << p.dipoles << '\n'
            << "dipoleRemovalDistance " << p.dipole_distance << '\n'
            << "dipoleReinjection " << name_of(p.reinjection) << '\n'
            << "initialConditionFile " << q(absolute(p.initial_file)) << '\n'
            << "trajectory_file " << q(absolute(r.trajectory.string())) << '\n'
            << "diagnostics_file " << q(absolute(r.diagnostics.string())) << '\n'
            << "checkpoint_directory " << q(absolute(r.checkpoints.string())) << '\n'
            << "segmentDirectory " << q(seg_abs) << '\n';
        out.close();
        if (!out)
            throw std::runtime_error("failed while writing run record: " + temporary.string());
        fs::rename(temporary, file);
    }
}

// ---------------------------------------------------------------- the device

int env_device() {
    const char *v = std::getenv("PV_HIP_DEVICE");
    return v ? std::atoi(v) : 0;
}

int env_precision() {
    const char *v = std::getenv("PV_HIP_PRECISION");
    if (!v || std::strcmp(v, "fp64") == 0)
        return PV_FP64;
    if (std::strcmp(v, "mixed") == 0)
        return PV_MIXED;
    throw std::invalid_argument("PV_HIP_PRECISION must be fp64 or mixed");
}

// The resident run: a pv_handle holding the state, and a host copy refreshed only when an
// output needs it.
class Device {
  public:
    explicit Device(const Params &p) {
        pv_config c{};
        c.geometry = p.geometry == "periodic" ? PV_PERIODIC
                     : p.geometry == "disk"   ? PV_DISK
                                              : PV_INFINITE;
        c.precision = env_precision();
        c.core_radius = p.core_radius;
        c.box_length_x = p.box_x;
        c.box_length_y = p.box_y;
        c.image_layers = p.layers;
        c.disk_radius = p.disk_radius;
        c.device = env_device();
        check(pv_create(&c, &h_));
    }
    ~Device() { pv_destroy(h_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;

    void upload(const State &s) {
        check(pv_upload(h_, s.x.data(), s.y.data(), s.g.data(), s.size()));
        host_ = s;
        current_ = true;
    }
    // The resident state on the host (downloaded after the last step, if any).
    const State &host() {
        if (!current_) {
            std::size_t n = 0;
            check(pv_state_size(h_, &n));
            host_.resize(n);
            check(pv_download(h_, host_.x.data(), host_.y.data(), host_.g.data(), n));
            current_ = true;
        }
        return host_;
    }
    void velocities(std::vector<double> &u, std::vector<double> &v) {
        const std::size_t n = host().size();
        u.assign(n, 0.0);
        v.assign(n, 0.0);
        check(pv_evaluate_device_state(h_, u.data(), v.data(), n));
    }
    // computeInvariants of the resident state, and of a host state (src/compute.cpp:157-163).
    Invariants invariants() {
        pv_invariants r{};
        check(pv_state_invariants(h_, &r));
        return {r.circulation, r.linear_impulse_x, r.linear_impulse_y, r.angular_impulse,
                r.hamiltonian};
    }
    Invariants invariants_of(const State &s) {
        require_finite(s);
        pv_invariants r{};
        check(pv_invariants_of(h_, s.x.data(), s.y.data(), s.g.data(), s.size(), &r));
        return {r.circulation, r.linear_impulse_x, r.linear_impulse_y, r.angular_impulse,
                r.hamiltonian};
    }
    // One step of the device state; returns {accepted, suggested} (RK4 keeps dt).
    std::pair<double, double> step(const Params &p, double dt, double suggested) {
        current_ = false;
        if (p.integrator == Integrator::rk4) {
            check(pv_rk4_step(h_, dt));
            return {dt, suggested};
        }
        pv_step_result r{};
        check(pv_dopri5_step(h_, dt, p.atol, p.rtol, p.min_dt, p.max_dt, &r));
        return {r.accepted_time_step, r.suggested_time_step};
    }
    // DipoleManager::process on the resident state; returns the pairs removed.
    std::size_t dipoles(const Params &p) {
        if (!p.dipoles || host_size() == 0)
            return 0;
        std::size_t events = 0;
        const int mode = p.reinjection == Reinjection::independent ? PV_REINJECT_INDEPENDENT
                         : p.reinjection == Reinjection::paired    ? PV_REINJECT_PAIRED
                                                                   : PV_REINJECT_NONE;
        check(pv_dipole_process(h_, p.dipole_distance, mode, &events));
        if (events)
            current_ = false;
        return events;
    }
    void seed_dipoles(std::uint64_t seed, const std::string &state) {
        check(pv_dipole_set_random(h_, seed, state.c_str()));
    }
    std::string dipole_generator() {
        std::size_t need = 0;
        check(pv_dipole_random_state(h_, nullptr, 0, &need));
        std::string text(need, '\0');
        check(pv_dipole_random_state(h_, text.data(), text.size(), &need));
        text.resize(std::strlen(text.c_str()));
        return text;
    }

  private:
    std::size_t host_size() {
        std::size_t n = 0;
        check(pv_state_size(h_, &n));
        return n;
    }
    pv_handle *h_ = nullptr;
    State host_;
    bool current_ = false;
};

std::string runtime_details() {
    char buffer[1024] = {};
    check(pv_runtime_details(env_device(), buffer, sizeof buffer));
    return std::string(buffer) + "\nhip_precision " +
           (env_precision() == PV_MIXED ? "mixed" : "fp64");
}

// ---------------------------------------------------------------- the run

void run(int argc, char **argv) {
    if (argc > 2)
        throw std::invalid_argument("expected at most one parameter file argument");
    const std::string param_file = argc > 1 ? argv[1] : "params.txt";
    const Params p = load_params(param_file);
    const RunPaths paths = run_paths(p);
    const bool restarting = !p.restart_file.empty();
    Checkpoint restart;
    if (restarting) {
        restart = read_checkpoint(p.restart_file);
        if (!checkpoint_matches(restart, p))
            throw std::runtime_error(
                "checkpoint geometry or integrator settings do not match parameters");
    }
    const State start = restarting ? restart.state : initial_state(p);
    Device dev(p);
    // The reference's order: the initial invariants of the host state, then its upload.
    const Invariants initial = restarting ? restart.initial : dev.invariants_of(start);
    dev.upload(start);
    DipoleCounts counts = restarting ? restart.counts : DipoleCounts{};
    dev.seed_dipoles(p.seed, restarting ? restart.counts.generator : std::string());
    const auto record_events = [&](std::size_t events) {
        counts.removed += events;
        if (p.reinjection != Reinjection::none)
            counts.reinjected += events;
    };
    if (!restarting)
        record_events(dev.dipoles(p));
    Invariants segment =
        restarting && restart.has_segment ? restart.segment : dev.invariants();

    double t = restarting ? restart.time : 0.0;
    if (t > p.end_time)
        throw std::runtime_error("checkpoint time is later than endTime");
    double dt = restarting ? restart.suggested_dt
                           : (p.integrator == Integrator::dopri5
                                  ? std::clamp(p.time_step, p.min_dt, p.max_dt)
                                  : p.time_step);
    // A clock keeps its saved next event when its interval is unchanged; a changed interval
    // starts a new cadence at the restart time.  Legacy checkpoints hold one shared clock.
    const auto first_event = [&](double interval, double saved_interval, double saved_next) {
        const double next = restarting && interval == saved_interval ? saved_next : t + interval;
        if (!std::isfinite(next) || !(next > t))
            throw std::runtime_error("output interval cannot advance simulation time");
        return next;
    };
    const bool saved = restarting && restart.has_schedule;
    const Schedule &ss = restart.schedule;
    double next_output = first_event(p.output_time, saved ? ss.trajectory_interval : p.output_time,
                                     restart.next_output);
    Schedule sched{p.output_time, p.diagnostics_interval(), p.checkpoint_interval(), 0.0, 0.0};
    sched.next_diagnostics =
        first_event(sched.diagnostics_interval, saved ? ss.diagnostics_interval : p.output_time,
                    saved ? ss.next_diagnostics : restart.next_output);
    sched.next_checkpoint =
        first_event(sched.checkpoint_interval, saved ? ss.checkpoint_interval : p.output_time,
                    saved ? ss.next_checkpoint : restart.next_output);
    std::size_t accepted = restarting ? restart.accepted : 0;
    std::size_t output_index = restarting ? restart.output_index : 0;   // checkpoint files
    std::size_t event_index = restarting ? restart.event_index : 0;     // output frames

    const auto protect_inputs = [&](const fs::path &destination) {
        for (const std::string &input : {param_file, p.initial_file, p.restart_file})
            if (!input.empty() && same_file(destination, input))
                throw std::runtime_error("output path would overwrite input file: " + input);
    };
    const auto check_checkpoint_destination = [&](std::size_t index) {
        const fs::path destination = checkpoint_path(paths.checkpoints, index);
        protect_inputs(destination);
        if (same_file(destination, paths.trajectory) || same_file(destination, paths.diagnostics))
            throw std::runtime_error("checkpoint and CSV output paths must be different");
    };
    if (restarting && restart.output_index == kSizeMax)
        throw std::runtime_error("checkpoint index overflow");
    // Every destination is checked before any earlier artefact is replaced.
    const fs::path traj_path = fs::weakly_canonical(paths.trajectory);
    const fs::path diag_path = fs::weakly_canonical(paths.diagnostics);
    if (same_file(traj_path, diag_path))
        throw std::runtime_error("managed trajectory and diagnostics paths must be different");
    protect_inputs(traj_path);
    protect_inputs(diag_path);
    check_checkpoint_destination(restarting ? restart.output_index + 1 : 0);
    if (has_run_output(paths)) {
        if (!p.overwrite)
            throw std::runtime_error("run directory already contains solver output: " +
                                     paths.directory.string() +
                                     " (choose another runDirectory or set overwriteRun true)");
        remove_run_output(paths);
    }
    Csv csv(paths, initial);
    std::cout << "backend=HIP\n"
              << "trajectory=" << fs::absolute(paths.trajectory).lexically_normal()
              << " interval=" << sched.trajectory_interval << '\n'
              << "diagnostics=" << fs::absolute(paths.diagnostics).lexically_normal()
              << " interval=" << sched.diagnostics_interval << '\n'
              << "checkpoints=" << fs::absolute(paths.checkpoints).lexically_normal()
              << " interval=" << sched.checkpoint_interval << '\n'
              << std::flush;

    std::vector<double> u, v;
    const auto write_trajectory = [&] {
        const State &s = dev.host();
        dev.velocities(u, v);
        csv.trajectory(t, event_index, s, u, v);
    };
    const auto write_diagnostics = [&] {
        const Invariants now = dev.invariants();
        csv.diagnostics(t, event_index, now, segment, counts.removed, counts.reinjected);
        print_diagnostics(t, accepted, now, initial, p.geometry, segment, counts.removed,
                          counts.reinjected);
    };
    const auto write_checkpoint_now = [&] {
        check_checkpoint_destination(output_index);
        DipoleCounts with_generator = counts;
        with_generator.generator = dev.dipole_generator();
        write_checkpoint(paths.checkpoints, dev.host(), initial, p, segment, with_generator,
                         sched, {t, dt, next_output, accepted, output_index, event_index});
    };
    // A restarted branch records its first frame but does not repeat its source checkpoint.
    write_trajectory();
    write_diagnostics();
    if (!restarting)
        write_checkpoint_now();
    write_run_records(p, param_file, runtime_details(), t, event_index, restarting);

    while (t < p.end_time) {
        // Land on the next event of any output stream, or on the end time.
        const double h = std::min({dt, p.end_time - t, next_output - t,
                                   sched.next_diagnostics - t, sched.next_checkpoint - t});
        if (!(t + h > t))
            throw std::runtime_error("timestep cannot advance simulation time");
        const auto [taken, suggested] = dev.step(p, h, dt);
        dt = suggested;
        if (!std::isfinite(taken) || !(t + taken > t))
            throw std::runtime_error("accepted timestep cannot advance simulation time");
        t += taken;
        if (accepted == kSizeMax)
            throw std::runtime_error("accepted-step counter overflow");
        ++accepted;
        const std::size_t events = dev.dipoles(p);
        if (events) {
            record_events(events);
            segment = dev.invariants();
        }
        const double slack = 16.0 * std::numeric_limits<double>::epsilon() * std::fabs(t);
        const bool final_frame = t + slack >= p.end_time;
        if (final_frame)
            t = p.end_time;     // one final frame, not two separated by roundoff
        const bool traj_due = t + slack >= next_output;
        const bool diag_due = t + slack >= sched.next_diagnostics;
        const bool ckpt_due = t + slack >= sched.next_checkpoint;
        const auto advance = [&](double &next, double interval) {
            do {
                const double following = next + interval;
                if (!std::isfinite(following) || !(following > next))
                    throw std::runtime_error("output interval cannot advance simulation time");
                next = following;
            } while (next <= t);
        };
        // Every due clock moves on before anything is saved, coincident events included.
        if (traj_due)
            advance(next_output, sched.trajectory_interval);
        if (diag_due)
            advance(sched.next_diagnostics, sched.diagnostics_interval);
        if (ckpt_due)
            advance(sched.next_checkpoint, sched.checkpoint_interval);
        if (traj_due || diag_due || ckpt_due || final_frame) {
            if (event_index == kSizeMax)
                throw std::runtime_error("output event-frame counter overflow");
            ++event_index;
        }
        if (traj_due || final_frame)
            write_trajectory();
        if (diag_due || final_frame)
            write_diagnostics();
        if (ckpt_due || final_frame) {
            if (output_index == kSizeMax)
                throw std::runtime_error("checkpoint index overflow");
            ++output_index;
            write_checkpoint_now();
        }
    }
}

} // namespace

int main(int argc, char **argv) {
    try {
        run(argc, argv);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << '\n';
        return 1;
    }
    return 0;
}
