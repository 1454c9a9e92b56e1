// On-disk storage of polyjectories in mizuba's data-dir format
// (mizuba/polyjectory.cpp:76-236): six files in a 0700 directory,
//   desc          48-byte descriptor {u32 version, bool persist, u64 n_objs,
//                 u32 order + 1, f64 maxT, f64 epoch, f64 epoch2}
//   traj_offsets  {u64 offset, u64 n_steps} per object (offset in doubles)
//   time_offsets  u64 per object (offset in doubles)
//   traj          f64 [steps][order + 1][7], objects concatenated
//   time          f64, n_steps + 1 values per object with data
//   status        i32 per object
// all marked read-only once written. Host code only.
#pragma once

#include <cstdint>
#include <filesystem>
#include <optional>
#include <utility>
#include <vector>

namespace mzh_host
{

namespace fs = std::filesystem;

// The arrays of one polyjectory, in the layout of the files above.
struct PjArrays {
    uint32_t n_objs = 0, order = 0;
    double maxT = 0, epoch = 0, epoch2 = 0;
    std::vector<double> traj, time;
    std::vector<uint64_t> traj_off, nsteps, time_off;
    std::vector<int32_t> status;
};

// polyjectory::traj_offset (mizuba/polyjectory.hpp:90-93).
struct TrajOffset {
    uint64_t offset;
    uint64_t n_steps;
};

// Normalised double-length epoch (Knuth's two-sum, detail/dfloat_utils.hpp:34-51).
std::pair<double, double> normalise_epoch(double hi, double lo);

// A new directory with owner-only permissions; it must not exist
// (detail/file_utils.cpp:165-198). Returns its canonical path.
fs::path create_dir_0700(const fs::path &path);
// A new uniquely named directory "<prefix>-XXXX-XXXX-XXXX-XXXX" in tmpdir, else
// in the global temporary dir (set_tmpdir, initially $MIZUBA_TMPDIR), else in
// the system's temporary directory (detail/file_utils.cpp:90-163).
fs::path create_temp_dir(const std::optional<fs::path> &tmpdir, const char *prefix = "mizuba_polyjectory");

// The global temporary dir (mizuba/tmpdir.cpp:40-78); empty: none.
fs::path get_global_tmpdir();
void set_global_tmpdir(fs::path p);

// Write `bytes` bytes into a new file and mark it read-only.
void write_file_ro(const fs::path &p, const void *data, size_t bytes);

// Write the six files of `a` into the existing directory `dir`.
void write_pj_files(const fs::path &dir, const PjArrays &a, bool persist);

// polyjectory::polyjectory from data files (polyjectory.cpp:652-948): checks
// the arguments, moves the two files into a new data dir (data_dir, or a
// temporary one), checks and loads their contents and writes the other files.
// On error the new data dir is removed. Returns the data dir.
fs::path pj_from_datafiles(const fs::path &traj_file, const fs::path &time_file, uint32_t order,
                           const std::vector<TrajOffset> &traj_offsets, const std::vector<int32_t> &status,
                           double epoch, double epoch2, const std::optional<fs::path> &data_dir, bool persist,
                           const std::optional<fs::path> &tmpdir, PjArrays &out);

// polyjectory::mount (polyjectory.cpp:951-980, 189-236): load a persistent
// data dir. Returns its canonical path.
fs::path pj_mount(const fs::path &data_dir, PjArrays &out);

} // namespace mzh_host
