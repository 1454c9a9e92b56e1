#include "pj_store.hpp"

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace mzh_host
{

namespace
{

// The descriptor record of the desc file (polyjectory.cpp:89-106), in the
// layout of the x86-64 Linux ABI.
constexpr uint32_t kPjVersion = 0;
struct PjDesc {
    uint32_t version;
    bool persist;
    uint64_t n_objs;
    uint32_t poly_op1;
    double maxT;
    double epoch;
    double epoch2;
};
static_assert(std::is_trivially_copyable_v<PjDesc>);
static_assert(sizeof(PjDesc) == 48 && offsetof(PjDesc, persist) == 4 && offsetof(PjDesc, n_objs) == 8
                  && offsetof(PjDesc, poly_op1) == 16 && offsetof(PjDesc, maxT) == 24 && offsetof(PjDesc, epoch) == 32
                  && offsetof(PjDesc, epoch2) == 40,
              "descriptor layout");
static_assert(sizeof(TrajOffset) == 16, "traj_offset layout");

void mark_read_only(const fs::path &p) { fs::permissions(p, fs::perms::owner_read, fs::perm_options::replace); }

std::mutex g_tmpdir_m;
fs::path g_tmpdir = [] {
    const char *e = std::getenv("MIZUBA_TMPDIR");
    return e ? fs::path(e) : fs::path();
}();

// Write `bytes` bytes to a new file, then mark it read-only
// (polyjectory.cpp:310-326).
void write_file(const fs::path &p, const void *data, size_t bytes)
{
    if (fs::exists(p)) throw std::runtime_error("Cannot create the file '" + p.string() + "', as it exists already");
    std::FILE *f = std::fopen(p.c_str(), "wb");
    if (!f) throw std::runtime_error("Could not open the file '" + p.string() + "'");
    const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    const int c = std::fclose(f);
    if (w != bytes || c != 0)
        throw std::runtime_error("An error was detected while trying to write into the file '" + p.string() + "'");
    mark_read_only(p);
}

template <typename T>
void write_vector(const fs::path &p, const std::vector<T> &v)
{
    write_file(p, v.data(), v.size() * sizeof(T));
}

// Read a whole file of T values (its size must be a multiple of sizeof(T)).
template <typename T>
std::vector<T> read_vector(const fs::path &p, uint64_t expect_count)
{
    const uint64_t bytes = fs::file_size(p);
    if (bytes != expect_count * sizeof(T))
        throw std::invalid_argument("The file '" + p.string() + "' has a size of " + std::to_string(bytes)
                                    + " bytes, but " + std::to_string(expect_count * sizeof(T))
                                    + " bytes were expected");
    std::vector<T> v(expect_count);
    std::FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) throw std::runtime_error("Could not open the file '" + p.string() + "'");
    const size_t r = bytes ? std::fread(v.data(), 1, bytes, f) : 0;
    std::fclose(f);
    if (r != bytes) throw std::runtime_error("Could not read the file '" + p.string() + "'");
    return v;
}

std::string hex_group(std::mt19937_64 &rng)
{
    static const char *digits = "0123456789abcdef";
    std::string s(4, '0');
    for (auto &ch : s) ch = digits[rng() & 15u];
    return s;
}

// Checks of the time and trajectory data of object i (polyjectory.cpp:541-599,
// 852-907), the messages of the reference.
void check_object(const PjArrays &a, uint64_t i)
{
    const uint64_t op1 = a.order + 1u, ns = a.nsteps[i];
    const double *tr = a.traj.data() + a.traj_off[i];
    for (uint64_t k = 0; k < ns * op1 * 7u; ++k)
        if (!std::isfinite(tr[k]))
            throw std::invalid_argument("A non-finite value was found in the trajectory at index " + std::to_string(i));
    const double *tm = a.time.data() + a.time_off[i];
    const uint64_t nt = ns ? ns + 1u : 0u;
    for (uint64_t j = 0; j < nt; ++j) {
        if (!std::isfinite(tm[j]))
            throw std::invalid_argument("A non-finite time coordinate was found for the object at index "
                                        + std::to_string(i));
        if (tm[j] < 0)
            throw std::invalid_argument("A negative time coordinate was found for the object at index "
                                        + std::to_string(i));
        if (j > 0 && !(tm[j] > tm[j - 1]))
            throw std::invalid_argument("The sequence of times for the object at index " + std::to_string(i)
                                        + " is not monotonically increasing");
    }
}

} // namespace

std::pair<double, double> normalise_epoch(double hi, double lo)
{
    // Knuth's two-sum: exact for any magnitudes of the two parts
    const double s = hi + lo;
    const double bp = s - hi;
    const double ap = s - bp;
    const double e = (hi - ap) + (lo - bp);
    if (!std::isfinite(s) || !std::isfinite(e)) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), "(%.17g, %.17g)", hi, lo);
        throw std::invalid_argument(std::string("The normalisation of the double-length number with components ")
                                    + buf + " produced a non-finite result");
    }
    return {s, e};
}

fs::path create_dir_0700(const fs::path &path)
{
    if (!fs::create_directory(path))
        throw std::runtime_error("Error while creating the directory '" + path.string()
                                 + "': the directory already exists");
    try {
        fs::permissions(path, fs::perms::owner_all, fs::perm_options::replace);
        // anything written between creation and the permission change goes
        for (const auto &e : fs::directory_iterator(path)) fs::remove_all(e.path());
        return fs::canonical(path);
    } catch (...) {
        std::error_code ec;
        fs::remove_all(path, ec);
        throw;
    }
}

fs::path get_global_tmpdir()
{
    std::lock_guard<std::mutex> lk(g_tmpdir_m);
    return g_tmpdir;
}

void set_global_tmpdir(fs::path p)
{
    std::lock_guard<std::mutex> lk(g_tmpdir_m);
    g_tmpdir = std::move(p);
}

void write_file_ro(const fs::path &p, const void *data, size_t bytes) { write_file(p, data, bytes); }

fs::path create_temp_dir(const std::optional<fs::path> &tmpdir, const char *prefix)
{
    fs::path base;
    if (tmpdir && !tmpdir->empty()) {
        base = *tmpdir;
    } else {
        base = get_global_tmpdir();
        if (base.empty()) base = fs::temp_directory_path();
    }
    try {
        base = fs::canonical(base);
    } catch (...) {
        throw std::invalid_argument("Unable to canonicalise the path to the temporary dir '" + base.string()
                                    + "' (does it exist?)");
    }
    std::mt19937_64 rng(std::random_device{}());
    for (int attempt = 0; attempt < 10; ++attempt) {
        const fs::path p = base
                           / (std::string(prefix) + "-" + hex_group(rng) + "-" + hex_group(rng) + "-" + hex_group(rng)
                              + "-" + hex_group(rng));
        if (fs::create_directory(p)) {
            try {
                fs::permissions(p, fs::perms::owner_all, fs::perm_options::replace);
                for (const auto &e : fs::directory_iterator(p)) fs::remove_all(e.path());
                return p;
            } catch (...) {
                std::error_code ec;
                fs::remove_all(p, ec);
                throw;
            }
        }
    }
    throw std::invalid_argument("Unable to create a unique temporary directory in the path '" + base.string() + "'");
}

void write_pj_files(const fs::path &dir, const PjArrays &a, bool persist)
{
    std::vector<TrajOffset> offs(a.n_objs);
    for (uint32_t i = 0; i < a.n_objs; ++i) offs[i] = TrajOffset{a.traj_off[i], a.nsteps[i]};
    write_vector(dir / "traj_offsets", offs);
    write_vector(dir / "time_offsets", a.time_off);
    write_vector(dir / "status", a.status);
    write_vector(dir / "traj", a.traj);
    write_vector(dir / "time", a.time);
    PjDesc desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.version = kPjVersion;
    desc.persist = persist;
    desc.n_objs = a.n_objs;
    desc.poly_op1 = a.order + 1u;
    desc.maxT = a.maxT;
    desc.epoch = a.epoch;
    desc.epoch2 = a.epoch2;
    write_file(dir / "desc", &desc, sizeof(desc));
}

fs::path pj_from_datafiles(const fs::path &traj_file, const fs::path &time_file, uint32_t order,
                           const std::vector<TrajOffset> &traj_offsets, const std::vector<int32_t> &status,
                           double epoch, double epoch2, const std::optional<fs::path> &data_dir, bool persist,
                           const std::optional<fs::path> &tmpdir, PjArrays &out)
{
    if (data_dir && tmpdir)
        throw std::invalid_argument(
            "The 'data_dir' and 'tmpdir' construction arguments cannot be provided at the same time");
    if (order < 2u || order == std::numeric_limits<uint32_t>::max())
        throw std::invalid_argument("Invalid polynomial order " + std::to_string(order)
                                    + " specified during the construction of a polyjectory");
    const uint64_t op1 = order + 1u;
    if (traj_offsets.empty())
        throw std::invalid_argument("Invalid trajectory offsets vector passed to the constructor of a polyjectory: the "
                                    "vector cannot be empty");
    if (status.size() != traj_offsets.size())
        throw std::invalid_argument("Invalid status vector passed to the constructor of a polyjectory: the expected "
                                    "size is "
                                    + std::to_string(traj_offsets.size()) + ", but the actual size is "
                                    + std::to_string(status.size()));
    if (!std::isfinite(epoch)) {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%g", epoch);
        throw std::invalid_argument(std::string("The initial epoch of a polyjectory must be finite, but instead a "
                                                "value of ")
                                    + buf + " was provided");
    }
    if (!std::isfinite(epoch2)) {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%g", epoch2);
        throw std::invalid_argument(std::string("The second component of the initial epoch of a polyjectory must be "
                                                "finite, but instead a value of ")
                                    + buf + " was provided");
    }
    const auto ep = normalise_epoch(epoch, epoch2);
    fs::path traj_path, time_path;
    try {
        traj_path = fs::canonical(traj_file);
        time_path = fs::canonical(time_file);
    } catch (...) {
        throw std::invalid_argument("Invalid data file(s) passed to the constructor of a polyjectory: the path(s) "
                                    "could not be canonicalised");
    }
    if (traj_path == time_path)
        throw std::invalid_argument("Invalid data file(s) passed to the constructor of a polyjectory: the trajectory "
                                    "data file and the time data file are the same file");
    // offsets (polyjectory.cpp:720-785), with overflow checks on the totals
    const uint64_t umax = std::numeric_limits<uint64_t>::max();
    uint64_t tot_traj = 0, tot_time = 0;
    for (size_t i = 0; i < traj_offsets.size(); ++i) {
        const auto [offset, n_steps] = traj_offsets[i];
        if (i == 0) {
            if (offset != 0)
                throw std::invalid_argument("Invalid trajectory offsets vector passed to the constructor of a "
                                            "polyjectory: the initial offset value must be zero but it is "
                                            + std::to_string(offset) + " instead");
        } else {
            const auto [prev_offset, prev_n_steps] = traj_offsets[i - 1];
            if (offset < prev_offset)
                throw std::invalid_argument("Invalid trajectory offsets vector passed to the constructor of a "
                                            "polyjectory: the offset of the object at index "
                                            + std::to_string(i) + " is less than the offset of the previous object");
            if (prev_n_steps > umax / op1 / 7u || offset - prev_offset != prev_n_steps * op1 * 7u)
                throw std::invalid_argument(
                    "Inconsistent data detected in the trajectory offsets vector passed to the constructor of a "
                    "polyjectory: the offset of the object at index "
                    + std::to_string(i) + " is inconsistent with the offset and number of steps of the previous object");
        }
        if (n_steps > (umax - tot_traj) / op1 / 7u || n_steps > umax - tot_time - 1u)
            throw std::overflow_error("Overflow detected in the trajectory offsets passed to the constructor of a "
                                      "polyjectory");
        tot_traj += n_steps * op1 * 7u;
        tot_time += n_steps + (n_steps != 0 ? 1u : 0u);
    }
    if (tot_traj == 0)
        throw std::invalid_argument(
            "All the trajectories in a polyjectory have a number of steps equal to zero: this is not allowed");
    if (traj_offsets.size() > 0xffffffffull) throw std::overflow_error("Too many objects in a polyjectory");
    const fs::path dir = (data_dir && !data_dir->empty()) ? create_dir_0700(*data_dir) : create_temp_dir(tmpdir);
    try {
        PjArrays a;
        a.n_objs = (uint32_t)traj_offsets.size();
        a.order = order;
        a.epoch = ep.first;
        a.epoch2 = ep.second;
        a.status = status;
        a.traj_off.resize(a.n_objs);
        a.nsteps.resize(a.n_objs);
        a.time_off.resize(a.n_objs);
        uint64_t mo = 0;
        for (uint32_t i = 0; i < a.n_objs; ++i) {
            a.traj_off[i] = traj_offsets[i].offset;
            a.nsteps[i] = traj_offsets[i].n_steps;
            a.time_off[i] = mo;
            mo += a.nsteps[i] + (a.nsteps[i] != 0 ? 1u : 0u);
        }
        // move the original files in (polyjectory.cpp:806-845)
        const fs::path tp = dir / "traj", mp = dir / "time";
        fs::rename(traj_path, tp);
        fs::rename(time_path, mp);
        if (!fs::is_regular_file(tp))
            throw std::invalid_argument("Invalid trajectory data file passed to the constructor of a polyjectory: the "
                                        "file is not a regular file");
        if (!fs::is_regular_file(mp))
            throw std::invalid_argument(
                "Invalid time data file passed to the constructor of a polyjectory: the file is not a regular file");
        mark_read_only(tp);
        mark_read_only(mp);
        if (fs::file_size(tp) != tot_traj * sizeof(double))
            throw std::invalid_argument("Invalid trajectory data file passed to the constructor of a polyjectory: the "
                                        "expected size in bytes is "
                                        + std::to_string(tot_traj * sizeof(double)) + " but the actual size is "
                                        + std::to_string(fs::file_size(tp)) + " instead");
        if (fs::file_size(mp) != tot_time * sizeof(double))
            throw std::invalid_argument("Invalid time data file passed to the constructor of a polyjectory: the "
                                        "expected size in bytes is "
                                        + std::to_string(tot_time * sizeof(double)) + " but the actual size is "
                                        + std::to_string(fs::file_size(mp)) + " instead");
        a.traj = read_vector<double>(tp, tot_traj);
        a.time = read_vector<double>(mp, tot_time);
        double maxT = -std::numeric_limits<double>::infinity();
        for (uint32_t i = 0; i < a.n_objs; ++i) {
            check_object(a, i);
            if (a.nsteps[i]) maxT = std::max(maxT, a.time[a.time_off[i] + a.nsteps[i]]);
        }
        a.maxT = maxT;
        std::vector<TrajOffset> offs(traj_offsets);
        write_vector(dir / "traj_offsets", offs);
        write_vector(dir / "time_offsets", a.time_off);
        write_vector(dir / "status", a.status);
        PjDesc desc;
        std::memset(&desc, 0, sizeof(desc));
        desc.version = kPjVersion;
        desc.persist = persist;
        desc.n_objs = a.n_objs;
        desc.poly_op1 = (uint32_t)op1;
        desc.maxT = a.maxT;
        desc.epoch = a.epoch;
        desc.epoch2 = a.epoch2;
        write_file(dir / "desc", &desc, sizeof(desc));
        out = std::move(a);
    } catch (...) {
        std::error_code ec;
        fs::remove_all(dir, ec);
        throw;
    }
    return dir;
}

fs::path pj_mount(const fs::path &data_dir, PjArrays &out)
{
    fs::path path;
    try {
        path = fs::canonical(data_dir);
    } catch (...) {
        throw std::invalid_argument("Error while trying to mount a polyjectory on path '" + data_dir.string()
                                    + "': the path could not be canonicalised (does it exist?)");
    }
    if (!fs::is_directory(path))
        throw std::invalid_argument("Cannot mount a polyjectory on path '" + path.string()
                                    + "': the path is not a directory");
    // the version first, read on its own (polyjectory.cpp:203-218)
    const fs::path dp = path / "desc";
    if (!fs::is_regular_file(dp) || fs::file_size(dp) < sizeof(uint32_t))
        throw std::invalid_argument("Cannot mount a polyjectory on path '" + path.string()
                                    + "': no valid descriptor file was found");
    uint32_t version = 0;
    {
        std::ifstream f(dp, std::ios::binary);
        f.read(reinterpret_cast<char *>(&version), sizeof(version));
    }
    if (version != kPjVersion)
        throw std::invalid_argument("Version mismatch when trying to mount() a polyjectory: the current polyjectory "
                                    "version is "
                                    + std::to_string(kPjVersion) + ", while the polyjectory in the data dir '"
                                    + path.string() + "' has a version number of " + std::to_string(version));
    const auto dv = read_vector<PjDesc>(dp, 1);
    const PjDesc desc = dv[0];
    if (!desc.persist)
        throw std::invalid_argument("Cannot mount() a polyjectory on the the data dir '" + path.string()
                                    + "': the data is not persistent");
    if (desc.n_objs == 0 || desc.n_objs > 0xffffffffull || desc.poly_op1 < 3u)
        throw std::invalid_argument("Cannot mount a polyjectory on path '" + path.string()
                                    + "': the descriptor is invalid");
    PjArrays a;
    a.n_objs = (uint32_t)desc.n_objs;
    a.order = desc.poly_op1 - 1u;
    a.maxT = desc.maxT;
    a.epoch = desc.epoch;
    a.epoch2 = desc.epoch2;
    const auto offs = read_vector<TrajOffset>(path / "traj_offsets", a.n_objs);
    a.time_off = read_vector<uint64_t>(path / "time_offsets", a.n_objs);
    a.status = read_vector<int32_t>(path / "status", a.n_objs);
    a.traj_off.resize(a.n_objs);
    a.nsteps.resize(a.n_objs);
    uint64_t tot_traj = 0, tot_time = 0;
    for (uint32_t i = 0; i < a.n_objs; ++i) {
        a.traj_off[i] = offs[i].offset;
        a.nsteps[i] = offs[i].n_steps;
        tot_traj = std::max<uint64_t>(tot_traj, offs[i].offset + offs[i].n_steps * desc.poly_op1 * 7u);
        if (offs[i].n_steps) tot_time = std::max<uint64_t>(tot_time, a.time_off[i] + offs[i].n_steps + 1u);
    }
    a.traj = read_vector<double>(path / "traj", fs::file_size(path / "traj") / sizeof(double));
    a.time = read_vector<double>(path / "time", fs::file_size(path / "time") / sizeof(double));
    if (a.traj.size() < tot_traj || a.time.size() < tot_time)
        throw std::invalid_argument("Cannot mount a polyjectory on path '" + path.string()
                                    + "': the data files are smaller than the offsets require");
    out = std::move(a);
    return path;
}

} // namespace mzh_host
