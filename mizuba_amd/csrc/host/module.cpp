// pybind11 front end of mizuba_amd: the reference's Python API for the
// conjunction-screening path (mizuba/expose_polyjectory.cpp:127-215,
// mizuba/expose_conjunctions.cpp:62-203) in C++ over the mizuba_amd C ABI.
//
// Host code only: every device operation goes through include/mizuba_amd.h.
#include <pybind11/eval.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl/filesystem.h>

#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "../../../include/mizuba_amd.h"
#include "pj_store.hpp"
#include "synth.hpp"

namespace py = pybind11;
using namespace py::literals;

namespace
{

std::string fmt_d(double v)
{
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    return buf;
}

// Turn a C-ABI status into the exception type the reference throws.
void check(int st, const mzh_ctx *ctx)
{
    if (st == MZH_OK) return;
    char buf[1024] = {0};
    if (ctx) {
        mzh_last_error(ctx, buf, sizeof(buf));
    } else {
        mzh_global_error(buf, sizeof(buf));
    }
    const std::string msg(buf);
    if (st == MZH_EINVAL || st == MZH_ESIZE) throw std::invalid_argument(msg);
    if (st == MZH_ERANGE) throw std::out_of_range(msg);
    if (st == MZH_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

py::object numpy() { return py::module_::import("numpy"); }

py::dtype bvh_node_dtype()
{
    return py::dtype::from_args(py::eval(
        "[('begin', '<u4'), ('end', '<u4'), ('left', '<i4'), ('right', '<i4'), ('lb', '<f4', (4,)), ('ub', '<f4', (4,))]"));
}

py::dtype traj_offset_dtype() { return py::dtype::from_args(py::eval("[('offset', '<u8'), ('n_steps', '<u8')]")); }

py::dtype gpe_dtype()
{
    return py::dtype::from_args(py::eval(
        "[('norad_id', '<u8'), ('epoch_jd1', '<f8'), ('epoch_jd2', '<f8'), ('n0', '<f8'), ('e0', '<f8'), "
        "('i0', '<f8'), ('node0', '<f8'), ('omega0', '<f8'), ('m0', '<f8'), ('bstar', '<f8')]"));
}

py::dtype aabb_collision_dtype() { return py::dtype::from_args(py::eval("[('i', '<u4'), ('j', '<u4')]")); }

py::dtype conj_dtype()
{
    return py::dtype::from_args(
        py::eval("[('i', '<u4'), ('j', '<u4'), ('tca', '<f8'), ('dca', '<f8'), ('ri', '<f8', (3,)), "
                 "('vi', '<f8', (3,)), ('rj', '<f8', (3,)), ('vj', '<f8', (3,))]"));
}

// Read-only numpy view of host memory owned by `base`.
template <typename T>
py::array ro_view(const py::object &base, const T *ptr, std::vector<py::ssize_t> shape)
{
    py::array a(py::dtype::of<T>(), shape, {}, ptr, base);
    a.attr("setflags")("write"_a = false);
    return a;
}

py::array ro_view_dtype(const py::object &base, const py::dtype &dt, const void *ptr, py::ssize_t n)
{
    py::array a(dt, {n}, {}, ptr, base);
    a.attr("setflags")("write"_a = false);
    return a;
}

// ---------------------------------------------------------------------------
// polyjectory: minimal in-memory equivalent of mizuba::polyjectory
// ---------------------------------------------------------------------------

// Device copy of a polyjectory, made on its first use on a device
// (state_eval / state_meval, conjunctions) and kept until the polyjectory is
// released; screenings share its arrays (mzh_share_polyjectory).
struct PjDevice {
    mzh_ctx *ctx = nullptr;
    PjDevice() = default;
    PjDevice(const PjDevice &) = delete;
    PjDevice &operator=(const PjDevice &) = delete;
    ~PjDevice()
    {
        if (ctx) mzh_destroy(ctx);
    }
};

struct PjData {
    uint32_t n_objs = 0, order = 0;
    double maxT = 0, epoch = 0, epoch2 = 0;
    std::vector<double> traj, time;
    std::vector<uint64_t> traj_off, nsteps, time_off, time_len;
    std::vector<int32_t> status;
    std::mutex eval_m;  // evaluations on the device copies
    std::mutex dev_m;   // the map of device copies
    std::map<int, std::unique_ptr<PjDevice>> devs;
    // data dir holding the polyjectory's files (empty: in memory only); a
    // non-persistent one is removed with the data (polyjectory.cpp:245-290)
    std::filesystem::path data_dir;
    bool persist = false;
    PjData() = default;
    PjData(const PjData &) = delete;
    PjData &operator=(const PjData &) = delete;
    ~PjData()
    {
        if (!data_dir.empty() && !persist) {
            std::error_code ec;
            std::filesystem::remove_all(data_dir, ec);
        }
    }
};

struct Polyjectory {
    std::shared_ptr<PjData> d;
};

// The context holding the device copy of d on `device`, uploaded on first use.
mzh_ctx *pj_device(PjData &d, int device)
{
    std::lock_guard<std::mutex> lk(d.dev_m);
    auto &slot = d.devs[device];
    if (!slot) {
        auto dev = std::make_unique<PjDevice>();
        check(mzh_create(device, &dev->ctx), nullptr);
        check(mzh_upload_polyjectory(dev->ctx, d.n_objs, d.order, d.traj.data(), d.traj.size(), d.time.data(),
                                     d.time.size(), d.traj_off.data(), d.nsteps.data(), d.time_off.data(), d.maxT),
              dev->ctx);
        slot = std::move(dev);
    }
    return slot->ctx;
}

// polyjectory::check_attached (polyjectory.cpp:333-338).
PjData &attached(const Polyjectory &p)
{
    if (!p.d) throw std::invalid_argument("Cannot operate on a detached polyjectory");
    return *p.d;
}

std::shared_ptr<PjData> pj_from_arrays(mzh_host::PjArrays &&a)
{
    auto d = std::make_shared<PjData>();
    d->n_objs = a.n_objs;
    d->order = a.order;
    d->maxT = a.maxT;
    d->epoch = a.epoch;
    d->epoch2 = a.epoch2;
    d->traj = std::move(a.traj);
    d->time = std::move(a.time);
    d->traj_off = std::move(a.traj_off);
    d->nsteps = std::move(a.nsteps);
    d->time_off = std::move(a.time_off);
    d->status = std::move(a.status);
    d->time_len.resize(d->n_objs);
    for (uint32_t i = 0; i < d->n_objs; ++i) d->time_len[i] = d->nsteps[i] ? d->nsteps[i] + 1 : 0;
    return d;
}

// Write the files of an in-memory polyjectory into a new data dir (the user's
// data_dir, or a temporary dir in tmpdir), as the reference does for every
// polyjectory (polyjectory.cpp:389-392, 460-519). The data dir is removed
// again if writing fails.
void store_pj(PjData &d, const std::optional<std::filesystem::path> &data_dir,
              const std::optional<std::filesystem::path> &tmpdir, bool persist)
{
    const auto dir = (data_dir && !data_dir->empty()) ? mzh_host::create_dir_0700(*data_dir)
                                                      : mzh_host::create_temp_dir(tmpdir);
    try {
        mzh_host::PjArrays a;
        a.n_objs = d.n_objs;
        a.order = d.order;
        a.maxT = d.maxT;
        a.epoch = d.epoch;
        a.epoch2 = d.epoch2;
        // the file writer reads the arrays: lend them and take them back
        a.traj.swap(d.traj);
        a.time.swap(d.time);
        a.traj_off.swap(d.traj_off);
        a.nsteps.swap(d.nsteps);
        a.time_off.swap(d.time_off);
        a.status.swap(d.status);
        struct Back {
            mzh_host::PjArrays &a;
            PjData &d;
            ~Back()
            {
                a.traj.swap(d.traj);
                a.time.swap(d.time);
                a.traj_off.swap(d.traj_off);
                a.nsteps.swap(d.nsteps);
                a.time_off.swap(d.time_off);
                a.status.swap(d.status);
            }
        } back{a, d};
        mzh_host::write_pj_files(dir, a, persist);
    } catch (...) {
        std::error_code ec;
        std::filesystem::remove_all(dir, ec);
        throw;
    }
    d.data_dir = dir;
    d.persist = persist;
}

// Validation of polyjectory.cpp:375-447 and :559-599 with its messages.
std::shared_ptr<PjData> build_pj(const std::vector<py::array_t<double>> &trajs,
                                 const std::vector<py::array_t<double>> &times, const std::vector<int32_t> &status,
                                 double epoch, double epoch2)
{
    const size_t n = trajs.size();
    for (const auto &a : trajs) {
        if (a.ndim() != 3)
            throw std::invalid_argument("A trajectory array must have 3 dimensions, but instead "
                                        + std::to_string(a.ndim()) + " dimension(s) were detected");
        if (a.shape(2) != 7)
            throw std::invalid_argument("A trajectory array must have a size of 7 in the third dimension, but instead "
                                        "a size of "
                                        + std::to_string(a.shape(2)) + " was detected");
    }
    for (const auto &a : trajs)
        if (!(a.flags() & py::array::c_style))
            throw std::invalid_argument("All trajectory arrays must be C contiguous and properly aligned");
    for (const auto &a : times) {
        if (a.ndim() != 1)
            throw std::invalid_argument("A time array must have 1 dimension, but instead " + std::to_string(a.ndim())
                                        + " dimension(s) were detected");
        if (!(a.flags() & py::array::c_style))
            throw std::invalid_argument("All time arrays must be C contiguous and properly aligned");
    }
    if (n == 0) throw std::invalid_argument("Cannot initialise a polyjectory object from an empty list of trajectories");
    if (n != times.size())
        throw std::invalid_argument("In the construction of a polyjectory, the number of objects deduced from the list "
                                    "of trajectories ("
                                    + std::to_string(n)
                                    + ") is inconsistent with the number of objects deduced from the list of times ("
                                    + std::to_string(times.size()) + ")");
    if (n != status.size())
        throw std::invalid_argument("In the construction of a polyjectory, the number of objects deduced from the list "
                                    "of trajectories ("
                                    + std::to_string(n)
                                    + ") is inconsistent with the number of objects deduced from the status list ("
                                    + std::to_string(status.size()) + ")");
    if (!std::isfinite(epoch))
        throw std::invalid_argument("The initial epoch of a polyjectory must be finite, but instead a value of "
                                    + fmt_d(epoch) + " was provided");
    if (!std::isfinite(epoch2))
        throw std::invalid_argument("The second component of the initial epoch of a polyjectory must be finite, but "
                                    "instead a value of "
                                    + fmt_d(epoch2) + " was provided");
    if (n > 0xffffffffull) throw std::overflow_error("Too many objects in a polyjectory");
    auto d = std::make_shared<PjData>();
    d->n_objs = (uint32_t)n;
    const auto ep = mzh_host::normalise_epoch(epoch, epoch2);
    d->epoch = ep.first;
    d->epoch2 = ep.second;
    d->status = status;
    uint64_t op1 = 0, tot_traj = 0, tot_time = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t cur = (uint64_t)trajs[i].shape(1);
        if (i == 0) {
            if (cur < 3)
                throw std::invalid_argument(
                    "The trajectory polynomial order for the first object is less than 2 - this is not allowed");
            op1 = cur;
        } else if (cur != op1) {
            throw std::invalid_argument("The trajectory polynomial order for the object at index " + std::to_string(i)
                                        + " is inconsistent with the polynomial order deduced from the first object ("
                                        + std::to_string(op1 - 1) + ")");
        }
        tot_traj += (uint64_t)trajs[i].shape(0) * op1 * 7;
    }
    if (tot_traj == 0)
        throw std::invalid_argument(
            "All the trajectories in a polyjectory have a number of steps equal to zero: this is not allowed");
    double maxT = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t ns = (uint64_t)trajs[i].shape(0), nt = (uint64_t)times[i].shape(0);
        if (nt == 0) {
            if (ns != 0)
                throw std::invalid_argument("The trajectory of the object at index " + std::to_string(i)
                                            + " has a nonzero number of steps but no associated time data");
        } else if (nt - 1 != ns) {
            throw std::invalid_argument("The number of steps for the trajectory of the object at index "
                                        + std::to_string(i) + " is " + std::to_string(ns) + ", but the number of times is "
                                        + std::to_string(nt)
                                        + " (the number of times must be equal to the number of steps + 1)");
        }
        if (nt > 0) {
            const double cT = times[i].at(nt - 1);
            maxT = (cT > maxT) ? cT : maxT;
        }
        tot_time += nt;
    }
    d->order = (uint32_t)(op1 - 1);
    d->traj.resize(tot_traj);
    d->time.resize(tot_time);
    d->traj_off.resize(n);
    d->nsteps.resize(n);
    d->time_off.resize(n);
    d->time_len.resize(n);
    uint64_t to = 0, mo = 0;
    for (size_t i = 0; i < n; ++i) {
        const auto tr = trajs[i].unchecked<3>();
        const uint64_t ns = (uint64_t)trajs[i].shape(0);
        d->traj_off[i] = to;
        d->nsteps[i] = ns;
        for (uint64_t s = 0; s < ns; ++s)
            for (uint64_t k = 0; k < op1; ++k)
                for (uint64_t ch = 0; ch < 7; ++ch) {
                    const double v = tr(s, k, ch);
                    if (!std::isfinite(v))
                        throw std::invalid_argument("A non-finite value was found in the trajectory at index "
                                                    + std::to_string(i));
                    d->traj[to++] = v;
                }
        const auto tm = times[i].unchecked<1>();
        const uint64_t nt = (uint64_t)times[i].shape(0);
        d->time_off[i] = mo;
        d->time_len[i] = nt;
        for (uint64_t j = 0; j < nt; ++j) {
            const double v = tm(j);
            if (!std::isfinite(v))
                throw std::invalid_argument("A non-finite time coordinate was found for the object at index "
                                            + std::to_string(i));
            if (v < 0)
                throw std::invalid_argument("A negative time coordinate was found for the object at index "
                                            + std::to_string(i));
            if (j > 0 && !(v > tm(j - 1)))
                throw std::invalid_argument("The sequence of times for the object at index " + std::to_string(i)
                                            + " is not monotonically increasing");
            d->time[mo++] = v;
        }
    }
    d->maxT = maxT;
    return d;
}

std::shared_ptr<PjData> pj_from_synth(mzh_host::SynthOut &&s)
{
    auto d = std::make_shared<PjData>();
    d->n_objs = (uint32_t)s.nsteps.size();
    d->order = s.order;
    d->maxT = s.maxT;
    d->traj = std::move(s.traj);
    d->time = std::move(s.time);
    d->traj_off = std::move(s.traj_off);
    d->nsteps = std::move(s.nsteps);
    d->time_off = std::move(s.time_off);
    d->time_len.resize(d->n_objs);
    for (uint32_t i = 0; i < d->n_objs; ++i) d->time_len[i] = d->nsteps[i] ? d->nsteps[i] + 1 : 0;
    d->status.assign(d->n_objs, 0);
    return d;
}

// ---------------------------------------------------------------------------
// device context wrapper
// ---------------------------------------------------------------------------

struct Ctx {
    mzh_ctx *ctx = nullptr;
    int device = 0;
    std::shared_ptr<PjData> pj;
    explicit Ctx(int device_) : device(device_)
    {
        const int st = mzh_create(device, &ctx);
        check(st, nullptr);
    }
    ~Ctx()
    {
        if (ctx) mzh_destroy(ctx);
    }
    Ctx(const Ctx &) = delete;
    Ctx &operator=(const Ctx &) = delete;

    // a new host-to-device copy of the polyjectory in this context
    void upload(const std::shared_ptr<PjData> &d)
    {
        int st;
        {
            py::gil_scoped_release rel;
            st = mzh_upload_polyjectory(ctx, d->n_objs, d->order, d->traj.data(), d->traj.size(), d->time.data(),
                                        d->time.size(), d->traj_off.data(), d->nsteps.data(), d->time_off.data(),
                                        d->maxT);
        }
        check(st, ctx);
        pj = d;
    }

    // use the polyjectory's cached device copy instead of uploading it
    void share(const std::shared_ptr<PjData> &d, int device)
    {
        int st;
        {
            py::gil_scoped_release rel;
            mzh_ctx *src = pj_device(*d, device);
            st = mzh_share_polyjectory(ctx, src);
        }
        check(st, ctx);
        pj = d;
    }

    void screen(double thresh, double interval, const std::optional<std::vector<int32_t>> &otypes, int mode)
    {
        if (otypes && otypes->size() != pj->n_objs)
            throw std::invalid_argument("Invalid array of object types passed to the constructor of a conjunctions "
                                        "objects: the expected size is "
                                        + std::to_string(pj->n_objs) + ", but the actual size is "
                                        + std::to_string(otypes->size()) + " instead");
        int st;
        {
            py::gil_scoped_release rel;
            st = mzh_screen(ctx, thresh, interval, otypes ? otypes->data() : nullptr, mode);
        }
        check(st, ctx);
    }
};

// ---------------------------------------------------------------------------
// polyjectory evaluation (polyjectory.cpp:1094-1366, expose_polyjectory.cpp:60-95, 308-520)
// ---------------------------------------------------------------------------

// Evaluations run on the GPU through mzh_state_eval, on a device copy of the
// polyjectory made on first use (device 0).
void pj_state_eval(PjData &d, const double *tm, int mode, const uint64_t *sel, uint64_t n_rows, uint64_t n_times,
                   double *out)
{
    if (n_rows == 0 || n_times == 0) return;
    py::gil_scoped_release rel;
    mzh_ctx *ctx = pj_device(d, 0);
    std::lock_guard<std::mutex> lk(d.eval_m);
    check(mzh_state_eval(ctx, tm, mode, sel, n_rows, n_times, out), ctx);
}

void check_cc_aligned(const py::array &a, const char *msg)
{
    const py::object fl = a.attr("flags");
    if (!fl.attr("c_contiguous").cast<bool>() || !fl.attr("aligned").cast<bool>()) throw std::invalid_argument(msg);
}

bool may_share_memory(const py::array &a, const py::array &b)
{
    return numpy().attr("may_share_memory")(a, b).cast<bool>();
}

using EvalSelector = std::optional<std::variant<std::size_t, py::array_t<std::size_t>>>;

// The selector argument (expose_polyjectory.cpp:66-95): one index or a 1-D array.
struct SelView {
    bool present = false;
    uint64_t single = 0;
    const uint64_t *ptr = nullptr;
    uint64_t n = 0;
    const py::array *arr = nullptr;
};

SelView setup_selector(const EvalSelector &selector)
{
    SelView s;
    if (!selector) return s;
    s.present = true;
    if (const auto *idx = std::get_if<std::size_t>(&*selector)) {
        s.single = *idx;
        s.ptr = &s.single;
        s.n = 1;
        return s;
    }
    const auto &v = std::get<py::array_t<std::size_t>>(*selector);
    check_cc_aligned(v, "The selector array passed to state_eval() must be C contiguous and properly aligned");
    if (v.ndim() != 1)
        throw std::invalid_argument("The selector array passed to state_eval() must have 1 dimension, but the number "
                                    "of dimensions is "
                                    + std::to_string(v.ndim()) + " instead");
    s.ptr = reinterpret_cast<const uint64_t *>(v.data());
    s.n = (uint64_t)v.shape(0);
    s.arr = &v;
    return s;
}

py::array_t<double> state_eval(const py::object &self, const std::variant<double, py::array_t<double>> &tm,
                               const std::optional<py::array_t<double>> &out_, const EvalSelector &selector)
{
    auto &d = attached(self.cast<const Polyjectory &>());
    SelView sel = setup_selector(selector);
    if (sel.present && !sel.arr) sel.ptr = &sel.single;
    const auto *tm_arr = std::get_if<py::array_t<double>>(&tm);
    py::array_t<double> out;
    if (out_) {
        out = *out_;
        check_cc_aligned(out, "The output array passed to state_eval() must be C contiguous and properly aligned");
        if (out.ndim() != 2)
            throw std::invalid_argument("The output array passed to state_eval() must have 2 dimensions, but the "
                                        "number of dimensions is "
                                        + std::to_string(out.ndim()) + " instead");
        if (out.shape(1) != 7)
            throw std::invalid_argument("The output array passed to state_eval() must have a size of 7 in the second "
                                        "dimension, but the size in the second dimension is "
                                        + std::to_string(out.shape(1)) + " instead");
        if (tm_arr && may_share_memory(out, *tm_arr))
            throw std::invalid_argument(
                "Potential memory overlap detected between the output array passed to state_eval() and the time array");
        if (sel.arr && may_share_memory(out, *sel.arr))
            throw std::invalid_argument("Potential memory overlap detected between the output array passed to "
                                        "state_eval() and the array of object indices");
    } else {
        out = py::array_t<double>(std::vector<py::ssize_t>{(py::ssize_t)(sel.present ? sel.n : d.n_objs), 7});
    }
    int mode = MZH_EVAL_TIME_SCALAR;
    const double *tm_ptr = std::get_if<double>(&tm);
    uint64_t n_tm = 1;
    if (tm_arr) {
        check_cc_aligned(*tm_arr, "The time array passed to state_eval() must be C contiguous and properly aligned");
        if (tm_arr->ndim() != 1)
            throw std::invalid_argument("The time array passed to state_eval() must have 1 dimension, but the number "
                                        "of dimensions is "
                                        + std::to_string(tm_arr->ndim()) + " instead");
        mode = MZH_EVAL_TIME_PER_OBJECT;
        tm_ptr = tm_arr->data();
        n_tm = (uint64_t)tm_arr->shape(0);
    }
    // polyjectory.cpp:1158-1225
    const uint64_t n_out = (uint64_t)out.shape(0);
    if (sel.present) {
        if (n_out != sel.n)
            throw std::invalid_argument(
                "Invalid output array passed to state_eval(): the number of objects selected for evaluation is "
                + std::to_string(sel.n) + " but the size of the first dimension of the array is "
                + std::to_string(n_out) + " (the two numbers must be equal)");
        if (tm_arr && n_tm != sel.n)
            throw std::invalid_argument("Invalid time array passed to state_eval(): the number of selected objects is "
                                        + std::to_string(sel.n) + " but the size of the array is "
                                        + std::to_string(n_tm) + " (the two numbers must be equal)");
    } else {
        if (n_out != d.n_objs)
            throw std::invalid_argument("Invalid output array passed to state_eval(): the number of objects is "
                                        + std::to_string(d.n_objs)
                                        + " but the size of the first dimension of the array is "
                                        + std::to_string(n_out) + " (the two numbers must be equal)");
        if (tm_arr && n_tm != d.n_objs)
            throw std::invalid_argument("Invalid time array passed to state_eval(): the number of objects is "
                                        + std::to_string(d.n_objs) + " but the size of the array is "
                                        + std::to_string(n_tm) + " (the two numbers must be equal)");
    }
    pj_state_eval(d, tm_ptr, mode, sel.present ? sel.ptr : nullptr, n_out, 1, out.mutable_data());
    return out;
}

py::array_t<double> state_meval(const py::object &self, const py::array_t<double> &tm,
                                const std::optional<py::array_t<double>> &out_, const EvalSelector &selector)
{
    auto &d = attached(self.cast<const Polyjectory &>());
    SelView sel = setup_selector(selector);
    if (sel.present && !sel.arr) sel.ptr = &sel.single;
    check_cc_aligned(tm, "The time array passed to state_meval() must be C contiguous and properly aligned");
    if (tm.ndim() != 1 && tm.ndim() != 2)
        throw std::invalid_argument("The time array passed to state_meval() must have either 1 or 2 dimensions, but "
                                    "the number of dimensions is "
                                    + std::to_string(tm.ndim()) + " instead");
    const uint64_t n_time_evals = (uint64_t)(tm.ndim() == 1 ? tm.shape(0) : tm.shape(1));
    py::array_t<double> out;
    if (out_) {
        out = *out_;
        check_cc_aligned(out, "The output array passed to state_meval() must be C contiguous and properly aligned");
        if (out.ndim() != 3)
            throw std::invalid_argument("The output array passed to state_meval() must have 3 dimensions, but the "
                                        "number of dimensions is "
                                        + std::to_string(out.ndim()) + " instead");
        if (out.shape(2) != 7)
            throw std::invalid_argument("The output array passed to state_meval() must have a size of 7 in the third "
                                        "dimension, but the size in the third dimension is "
                                        + std::to_string(out.shape(2)) + " instead");
        if (may_share_memory(out, tm))
            throw std::invalid_argument("Potential memory overlap detected between the output array passed to "
                                        "state_meval() and the time array");
        if (sel.arr && may_share_memory(out, *sel.arr))
            throw std::invalid_argument("Potential memory overlap detected between the output array passed to "
                                        "state_meval() and the array of object indices");
    } else {
        out = py::array_t<double>(std::vector<py::ssize_t>{(py::ssize_t)(sel.present ? sel.n : d.n_objs),
                                                           (py::ssize_t)n_time_evals, 7});
    }
    // polyjectory.cpp:1255-1327
    const uint64_t n_out = (uint64_t)out.shape(0);
    if ((uint64_t)out.shape(1) != n_time_evals)
        throw std::invalid_argument("Invalid output array passed to state_meval(): the number of time evaluations per "
                                    "object is "
                                    + std::to_string(n_time_evals)
                                    + " but the size of the second dimension of the array is "
                                    + std::to_string(out.shape(1)) + " (the two numbers must be equal)");
    if (sel.present) {
        if (n_out != sel.n)
            throw std::invalid_argument("Invalid output array passed to state_meval(): the number of selected objects "
                                        "is "
                                        + std::to_string(sel.n) + " but the size of the first dimension of the array is "
                                        + std::to_string(n_out) + " (the two numbers must be equal)");
        if (tm.ndim() == 2 && (uint64_t)tm.shape(0) != sel.n)
            throw std::invalid_argument("Invalid time array passed to state_meval(): the number of selected objects "
                                        "is "
                                        + std::to_string(sel.n) + " but the size of the first dimension of the array is "
                                        + std::to_string(tm.shape(0)) + " (the two numbers must be equal)");
    } else {
        if (n_out != d.n_objs)
            throw std::invalid_argument("Invalid output array passed to state_meval(): the number of objects is "
                                        + std::to_string(d.n_objs)
                                        + " but the size of the first dimension of the array is "
                                        + std::to_string(n_out) + " (the two numbers must be equal)");
        if (tm.ndim() == 2 && (uint64_t)tm.shape(0) != d.n_objs)
            throw std::invalid_argument("Invalid time array passed to state_meval(): the number of objects is "
                                        + std::to_string(d.n_objs)
                                        + " but the size of the first dimension of the array is "
                                        + std::to_string(tm.shape(0)) + " (the two numbers must be equal)");
    }
    pj_state_eval(d, tm.data(), tm.ndim() == 1 ? MZH_EVAL_TIME_SHARED : MZH_EVAL_TIME_PER_EVAL,
                  sel.present ? sel.ptr : nullptr, n_out, n_time_evals, out.mutable_data());
    return out;
}

// Device memory of a Screener's context (mzh_device_alloc), for the
// device-pointer entry points; freed with the object.
struct DeviceBuffer {
    std::shared_ptr<Ctx> owner;
    void *p = nullptr;
    uint64_t nbytes = 0;
    DeviceBuffer(std::shared_ptr<Ctx> c, uint64_t n) : owner(std::move(c)), nbytes(n)
    {
        check(mzh_device_alloc(owner->ctx, n, &p), owner->ctx);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer()
    {
        if (p) mzh_device_free(owner->ctx, p);
    }
    void write(const py::array &a)
    {
        if (!(a.flags() & py::array::c_style) || (uint64_t)a.nbytes() > nbytes)
            throw std::invalid_argument("write() needs a C-contiguous array of at most the buffer's size");
        check(mzh_copy_to_device(owner->ctx, p, a.data(), (uint64_t)a.nbytes()), owner->ctx);
    }
    void read(py::array &a)
    {
        if (!(a.flags() & py::array::c_style) || !a.writeable() || (uint64_t)a.nbytes() > nbytes)
            throw std::invalid_argument("read() needs a writeable C-contiguous array of at most the buffer's size");
        check(mzh_copy_from_device(owner->ctx, a.mutable_data(), p, (uint64_t)a.nbytes()), owner->ctx);
    }
};

// GPEs from a NumPy array of gpe_dtype (expose_make_sgp4_polyjectory.cpp:55-70).
std::vector<mzh_gpe> gpes_from_array(const py::array &a)
{
    if (!a.dtype().equal(gpe_dtype()))
        throw py::type_error("When passing the 'gpes' argument to make_sgp4_polyjectory() as a NumPy array, the dtype "
                             "must be 'gpe_dtype', but it is '"
                             + py::str(a.dtype()).cast<std::string>() + "' instead");
    if (a.ndim() != 1)
        throw std::invalid_argument("The array of gpes passed to make_sgp4_polyjectory() must have 1 dimension, but "
                                    "the number of dimensions is "
                                    + std::to_string(a.ndim()) + " instead");
    check_cc_aligned(a, "The array of gpes passed to make_sgp4_polyjectory() must be C contiguous and properly aligned");
    static_assert(sizeof(mzh_gpe) == 80, "gpe layout");
    std::vector<mzh_gpe> v((size_t)a.shape(0));
    if (!v.empty()) std::memcpy(v.data(), a.data(), v.size() * sizeof(mzh_gpe));
    return v;
}

// The context's polyjectory as a host polyjectory (statuses and epoch included).
std::shared_ptr<PjData> pj_from_ctx(mzh_ctx *ctx)
{
    uint64_t n = 0, tl = 0, ml = 0;
    uint32_t order = 0;
    double maxT = 0, ep[2] = {0, 0};
    check(mzh_get_polyjectory_info(ctx, &n, &order, &tl, &ml, &maxT, ep), ctx);
    mzh_host::PjArrays a;
    a.n_objs = (uint32_t)n;
    a.order = order;
    a.maxT = maxT;
    a.epoch = ep[0];
    a.epoch2 = ep[1];
    a.traj.resize(tl);
    a.time.resize(ml);
    a.traj_off.resize(n);
    a.nsteps.resize(n);
    a.time_off.resize(n);
    a.status.resize(n);
    check(mzh_copy_polyjectory(ctx, a.traj.data(), a.time.data(), a.traj_off.data(), a.nsteps.data(),
                               a.time_off.data(), a.status.data()),
          ctx);
    return pj_from_arrays(std::move(a));
}

int parse_mode(const std::string &m)
{
    if (m == "fp64") return MZH_MODE_FP64;
    if (m == "mixed") return MZH_MODE_MIXED;
    throw std::invalid_argument("Invalid screening mode '" + m + "': it must be 'fp64' or 'mixed'");
}

// ---------------------------------------------------------------------------
// conjunctions
// ---------------------------------------------------------------------------

struct ConjImpl {
    std::unique_ptr<Ctx> ctx;
    std::shared_ptr<PjData> pj;
    double conj_thresh = 0, conj_det_interval = 0;
    int mode = MZH_MODE_FP64;
    uint64_t C = 0, N = 0;
    std::vector<int32_t> otypes;
    std::vector<double> cd_end_times;
    std::mutex m;
    std::unique_ptr<std::vector<float>> aabbs, srt_aabbs;
    std::unique_ptr<std::vector<uint64_t>> mcodes, srt_mcodes;
    std::unique_ptr<std::vector<uint32_t>> srt_idx;
    std::map<uint64_t, std::vector<mzh_bvh_node>> trees;
    std::map<uint64_t, std::vector<mzh_aabb_collision>> bps;
    std::unique_ptr<std::vector<mzh_conj>> conjs;
    // data dir holding the result files, removed with the object
    // (conjunctions.cpp:200-233); empty: results in HBM and host caches only
    std::filesystem::path data_dir;
    // screened in ranges of chunk_steps conjunction steps: every result is in
    // the host caches, and the counters and stage times are the ranges' sums
    uint64_t chunk_steps = 0;
    uint64_t np[MZH_NP_COUNT] = {};
    double stage[MZH_STAGE_COUNT] = {};
    ConjImpl() = default;
    ConjImpl(const ConjImpl &) = delete;
    ConjImpl &operator=(const ConjImpl &) = delete;
    ~ConjImpl()
    {
        if (!data_dir.empty()) {
            std::error_code ec;
            std::filesystem::remove_all(data_dir, ec);
        }
    }
};

struct Conjunctions {
    std::shared_ptr<ConjImpl> impl;
};

template <typename T, typename F>
const std::vector<T> &lazy(ConjImpl &im, std::unique_ptr<std::vector<T>> &slot, size_t n, F &&fetch)
{
    std::lock_guard<std::mutex> lk(im.m);
    if (!slot) {
        auto v = std::make_unique<std::vector<T>>(n ? n : 1);
        int st;
        {
            py::gil_scoped_release rel;
            st = fetch(v->data());
        }
        check(st, im.ctx->ctx);
        slot = std::move(v);
    }
    return *slot;
}

// Fill the accessor state of a screened problem.
void finish_impl(ConjImpl &im, const Polyjectory &pj, double conj_thresh, double conj_det_interval,
                 const std::optional<std::vector<int32_t>> &otypes, int m)
{
    im.pj = pj.d;
    im.conj_thresh = conj_thresh;
    im.conj_det_interval = conj_det_interval;
    im.mode = m;
    im.N = pj.d->n_objs;
    check(mzh_get_n_cd_steps(im.ctx->ctx, &im.C), im.ctx->ctx);
    im.otypes = otypes ? *otypes : std::vector<int32_t>(pj.d->n_objs, 1);
    im.cd_end_times.resize(im.C);
    check(mzh_copy_cd_end_times(im.ctx->ctx, im.cd_end_times.data()), im.ctx->ctx);
    // the object keeps only what its accessors read (the polyjectory's device
    // copy stays with the polyjectory)
    check(mzh_release(im.ctx->ctx, MZH_RELEASE_SCRATCH | MZH_RELEASE_POLYJECTORY), im.ctx->ctx);
}

// The result files of detect_conjunctions (detect_conjunctions.cpp:91-126,
// 455-534) in `dir`: aabbs, srt_aabbs, mcodes, srt_mcodes, vidx, bvh (the
// steps' trees concatenated), bp (the steps' pairs concatenated; one zeroed
// record when there are none) and conjunctions (a single byte when there are
// none), all read-only.
void store_conjunctions(ConjImpl &im, const std::filesystem::path &dir)
{
    mzh_ctx *ctx = im.ctx->ctx;
    const uint64_t C = im.C, N = im.N;
    if (im.chunk_steps) {
        // a screening in ranges of steps keeps every result in the host caches
        mzh_host::write_file_ro(dir / "aabbs", im.aabbs->data(), C * (N + 1) * 8 * sizeof(float));
        mzh_host::write_file_ro(dir / "srt_aabbs", im.srt_aabbs->data(), C * (N + 1) * 8 * sizeof(float));
        mzh_host::write_file_ro(dir / "mcodes", im.mcodes->data(), C * N * sizeof(uint64_t));
        mzh_host::write_file_ro(dir / "srt_mcodes", im.srt_mcodes->data(), C * N * sizeof(uint64_t));
        mzh_host::write_file_ro(dir / "vidx", im.srt_idx->data(), C * N * sizeof(uint32_t));
        std::vector<mzh_bvh_node> all_t;
        std::vector<mzh_aabb_collision> all_p;
        for (uint64_t k = 0; k < C; ++k) {
            const auto &t = im.trees.at(k);
            all_t.insert(all_t.end(), t.begin(), t.end());
            const auto &b = im.bps.at(k);
            all_p.insert(all_p.end(), b.begin(), b.end());
        }
        mzh_host::write_file_ro(dir / "bvh", all_t.data(), all_t.size() * sizeof(mzh_bvh_node));
        if (all_p.empty()) all_p.push_back(mzh_aabb_collision{});
        mzh_host::write_file_ro(dir / "bp", all_p.data(), all_p.size() * sizeof(mzh_aabb_collision));
        if (!im.conjs->empty()) {
            mzh_host::write_file_ro(dir / "conjunctions", im.conjs->data(), im.conjs->size() * sizeof(mzh_conj));
        } else {
            const char empty = 0;
            mzh_host::write_file_ro(dir / "conjunctions", &empty, 1);
        }
        return;
    }
    {
        std::vector<float> v(C * (N + 1) * 8);
        check(mzh_copy_aabbs(ctx, v.data()), ctx);
        mzh_host::write_file_ro(dir / "aabbs", v.data(), v.size() * sizeof(float));
        check(mzh_copy_srt_aabbs(ctx, v.data()), ctx);
        mzh_host::write_file_ro(dir / "srt_aabbs", v.data(), v.size() * sizeof(float));
    }
    {
        std::vector<uint64_t> v(C * N);
        check(mzh_copy_mcodes(ctx, v.data()), ctx);
        mzh_host::write_file_ro(dir / "mcodes", v.data(), v.size() * sizeof(uint64_t));
        check(mzh_copy_srt_mcodes(ctx, v.data()), ctx);
        mzh_host::write_file_ro(dir / "srt_mcodes", v.data(), v.size() * sizeof(uint64_t));
    }
    {
        std::vector<uint32_t> v(C * N);
        check(mzh_copy_srt_idx(ctx, v.data()), ctx);
        mzh_host::write_file_ro(dir / "vidx", v.data(), v.size() * sizeof(uint32_t));
    }
    {
        std::vector<mzh_bvh_node> all;
        for (uint64_t k = 0; k < C; ++k) {
            uint64_t n = 0;
            check(mzh_get_bvh_tree_size(ctx, k, &n), ctx);
            const size_t at = all.size();
            all.resize(at + n);
            if (n) check(mzh_copy_bvh_tree(ctx, k, all.data() + at), ctx);
        }
        mzh_host::write_file_ro(dir / "bvh", all.data(), all.size() * sizeof(mzh_bvh_node));
    }
    {
        std::vector<mzh_aabb_collision> all;
        for (uint64_t k = 0; k < C; ++k) {
            uint64_t n = 0;
            check(mzh_get_aabb_collisions_size(ctx, k, &n), ctx);
            const size_t at = all.size();
            all.resize(at + n);
            if (n) check(mzh_copy_aabb_collisions(ctx, k, all.data() + at), ctx);
        }
        if (all.empty()) all.push_back(mzh_aabb_collision{});
        mzh_host::write_file_ro(dir / "bp", all.data(), all.size() * sizeof(mzh_aabb_collision));
    }
    {
        uint64_t n = 0;
        check(mzh_get_n_conjunctions(ctx, &n), ctx);
        if (n) {
            std::vector<mzh_conj> v(n);
            check(mzh_copy_conjunctions(ctx, v.data()), ctx);
            mzh_host::write_file_ro(dir / "conjunctions", v.data(), v.size() * sizeof(mzh_conj));
        } else {
            const char empty = 0;
            mzh_host::write_file_ro(dir / "conjunctions", &empty, 1);
        }
    }
}

// hint_release (conjunctions.cpp:466-477 madvises the memory-mapped results
// away): every result is copied into the object's host caches, which the
// accessors read from then on, and the device copies are freed.
void conj_hint_release(ConjImpl &im)
{
    if (im.chunk_steps) return;  // already in host memory
    mzh_ctx *ctx = im.ctx->ctx;
    lazy<float>(im, im.aabbs, im.C * (im.N + 1) * 8, [&](float *p) { return mzh_copy_aabbs(ctx, p); });
    lazy<float>(im, im.srt_aabbs, im.C * (im.N + 1) * 8, [&](float *p) { return mzh_copy_srt_aabbs(ctx, p); });
    lazy<uint64_t>(im, im.mcodes, im.C * im.N, [&](uint64_t *p) { return mzh_copy_mcodes(ctx, p); });
    lazy<uint64_t>(im, im.srt_mcodes, im.C * im.N, [&](uint64_t *p) { return mzh_copy_srt_mcodes(ctx, p); });
    lazy<uint32_t>(im, im.srt_idx, im.C * im.N, [&](uint32_t *p) { return mzh_copy_srt_idx(ctx, p); });
    uint64_t nc = 0;
    check(mzh_get_n_conjunctions(ctx, &nc), ctx);
    lazy<mzh_conj>(im, im.conjs, nc, [&](mzh_conj *p) { return mzh_copy_conjunctions(ctx, p); });
    std::lock_guard<std::mutex> lk(im.m);
    py::gil_scoped_release rel;
    for (uint64_t k = 0; k < im.C; ++k) {
        if (!im.trees.count(k)) {
            uint64_t n = 0;
            check(mzh_get_bvh_tree_size(ctx, k, &n), ctx);
            std::vector<mzh_bvh_node> v(n ? n : 1);
            check(mzh_copy_bvh_tree(ctx, k, v.data()), ctx);
            v.resize(n);
            im.trees.emplace(k, std::move(v));
        }
        if (!im.bps.count(k)) {
            uint64_t n = 0;
            check(mzh_get_aabb_collisions_size(ctx, k, &n), ctx);
            std::vector<mzh_aabb_collision> v(n ? n : 1);
            check(mzh_copy_aabb_collisions(ctx, k, v.data()), ctx);
            v.resize(n);
            im.bps.emplace(k, std::move(v));
        }
    }
    check(mzh_release(ctx, MZH_RELEASE_RESULTS), ctx);
}

void check_otypes_size(const std::optional<std::vector<int32_t>> &otypes, uint64_t n_objs)
{
    if (otypes && otypes->size() != n_objs)
        throw std::invalid_argument("Invalid array of object types passed to the constructor of a conjunctions "
                                    "objects: the expected size is "
                                    + std::to_string(n_objs) + ", but the actual size is "
                                    + std::to_string(otypes->size()) + " instead");
}

// The constructor's checks in the reference's order (conjunctions.cpp:267-332:
// threshold, interval, then the size and values of otypes), before any device
// context is made; then this build's size limit of one batch
// (mzh_check_screen), an error unless `ranges` (then the screening runs in
// ranges of steps). Returns the number of conjunction steps.
uint64_t validate_screen(const PjData &pd, double conj_thresh, double conj_det_interval,
                         const std::optional<std::vector<int32_t>> &otypes, int mode, bool ranges = false)
{
    uint64_t C = 0;
    auto run = [&](const int32_t *ot) {
        const int st = mzh_check_screen(pd.n_objs, pd.maxT, conj_thresh, conj_det_interval, ot, mode, &C);
        if (!(st == MZH_ESIZE && ranges)) check(st, nullptr);
    };
    run(nullptr);
    check_otypes_size(otypes, pd.n_objs);
    if (otypes) run(otypes->data());
    return C;
}

// Screen the problem in ranges of K conjunction steps (mzh_screen_steps),
// each range's results copied into the host caches at their place before the
// next range is screened: the device holds one range at a time, the
// reference's chunked loop (detect_conjunctions.cpp:61-79, 184-242), with the
// results in host memory instead of files. The results are the one-batch
// screening's, bit for bit.
void screen_in_ranges(ConjImpl &im, const PjData &pd, double conj_thresh, double conj_det_interval,
                      const std::optional<std::vector<int32_t>> &otypes, int m, uint64_t C, uint64_t K)
{
    mzh_ctx *ctx = im.ctx->ctx;
    const uint64_t N = pd.n_objs;
    im.C = C;
    im.N = N;
    im.chunk_steps = K;
    im.aabbs = std::make_unique<std::vector<float>>(C * (N + 1) * 8);
    im.srt_aabbs = std::make_unique<std::vector<float>>(C * (N + 1) * 8);
    im.mcodes = std::make_unique<std::vector<uint64_t>>(C * N ? C * N : 1);
    im.srt_mcodes = std::make_unique<std::vector<uint64_t>>(C * N ? C * N : 1);
    im.srt_idx = std::make_unique<std::vector<uint32_t>>(C * N ? C * N : 1);
    im.conjs = std::make_unique<std::vector<mzh_conj>>();
    im.cd_end_times.resize(C);
    const int32_t *ot = otypes ? otypes->data() : nullptr;
    py::gil_scoped_release rel;
    for (uint64_t k0 = 0; k0 < C; k0 += K) {
        const uint64_t k1 = std::min(C, k0 + K), nk = k1 - k0;
        check(mzh_screen_steps(ctx, conj_thresh, conj_det_interval, ot, m, k0, k1), ctx);
        check(mzh_copy_cd_end_times(ctx, im.cd_end_times.data() + k0), ctx);
        check(mzh_copy_aabbs(ctx, im.aabbs->data() + k0 * (N + 1) * 8), ctx);
        check(mzh_copy_srt_aabbs(ctx, im.srt_aabbs->data() + k0 * (N + 1) * 8), ctx);
        check(mzh_copy_mcodes(ctx, im.mcodes->data() + k0 * N), ctx);
        check(mzh_copy_srt_mcodes(ctx, im.srt_mcodes->data() + k0 * N), ctx);
        check(mzh_copy_srt_idx(ctx, im.srt_idx->data() + k0 * N), ctx);
        for (uint64_t k = 0; k < nk; ++k) {
            uint64_t n = 0;
            check(mzh_get_bvh_tree_size(ctx, k, &n), ctx);
            std::vector<mzh_bvh_node> t(n ? n : 1);
            if (n) check(mzh_copy_bvh_tree(ctx, k, t.data()), ctx);
            t.resize(n);
            im.trees[k0 + k] = std::move(t);
            check(mzh_get_aabb_collisions_size(ctx, k, &n), ctx);
            std::vector<mzh_aabb_collision> b(n ? n : 1);
            if (n) check(mzh_copy_aabb_collisions(ctx, k, b.data()), ctx);
            b.resize(n);
            im.bps[k0 + k] = std::move(b);
        }
        uint64_t nc = 0;
        check(mzh_get_n_conjunctions(ctx, &nc), ctx);
        const size_t at = im.conjs->size();
        im.conjs->resize(at + nc);
        if (nc) check(mzh_copy_conjunctions(ctx, im.conjs->data() + at), ctx);
        uint64_t r[MZH_NP_COUNT];
        check(mzh_get_np_report(ctx, r), ctx);
        for (int j = 0; j < MZH_NP_COUNT; ++j) im.np[j] += r[j];
        double t[MZH_STAGE_COUNT];
        check(mzh_get_stage_times(ctx, t), ctx);
        for (int j = 0; j < MZH_STAGE_COUNT; ++j) im.stage[j] += t[j];
    }
    check(mzh_release(ctx, MZH_RELEASE_SCRATCH | MZH_RELEASE_RESULTS | MZH_RELEASE_POLYJECTORY), ctx);
}

// Host bytes that a screening's results take in host memory: the per-step
// arrays, about two 48-byte tree nodes per object and step, and the pairs
// (estimated at one per object and step).
uint64_t result_host_bytes(uint64_t N, uint64_t C)
{
    return C * ((N + 1) * 64 + N * (8 + 8 + 4 + 96 + 8));
}

Conjunctions make_conjunctions(const Polyjectory &pj, double conj_thresh, double conj_det_interval,
                               std::optional<std::vector<int32_t>> otypes,
                               std::optional<std::filesystem::path> data_dir,
                               std::optional<std::filesystem::path> tmpdir, const std::string &mode, int device)
{
    if (data_dir && tmpdir)
        throw std::invalid_argument(
            "The 'data_dir' and 'tmpdir' construction arguments cannot be provided at the same time");
    const int m = parse_mode(mode);
    const PjData &pd = attached(pj);
    const uint64_t C = validate_screen(pd, conj_thresh, conj_det_interval, otypes, m, true);
    // a screening beyond one batch's limit or the device's free memory runs
    // in ranges of steps, its results in host memory, which must hold them
    auto check_host = [&]() {
        const uint64_t need = result_host_bytes(pd.n_objs, C);
        const uint64_t avail = (uint64_t)sysconf(_SC_AVPHYS_PAGES) * (uint64_t)sysconf(_SC_PAGESIZE);
        if (need > avail)
            throw std::invalid_argument("A screening of " + std::to_string(pd.n_objs) + " objects over "
                                        + std::to_string(C) + " conjunction steps is too large: its results need about "
                                        + std::to_string(need) + " bytes of host memory, more than the "
                                        + std::to_string(avail) + " bytes available");
    };
    if (C > 0xffffffffull / (pd.n_objs ? pd.n_objs : 1)) check_host();
    auto im = std::make_shared<ConjImpl>();
    // the result files are written when a data dir or a temporary dir is
    // given; the dir is made first, as the reference does (conjunctions.cpp:336-341)
    if (data_dir && !data_dir->empty()) {
        im->data_dir = mzh_host::create_dir_0700(*data_dir);
    } else if (tmpdir) {
        im->data_dir = mzh_host::create_temp_dir(tmpdir, "mizuba_conjunctions");
    }
    im->ctx = std::make_unique<Ctx>(device);
    im->ctx->share(pj.d, device);
    // conjunction steps per range: what one batch's limit and half the
    // device's free memory allow (MIZUBA_AMD_CHUNK_STEPS caps it)
    uint64_t K = C;
    {
        uint64_t free_b = 0, total_b = 0;
        check(mzh_device_memory(im->ctx->ctx, &free_b, &total_b), im->ctx->ctx);
        check(mzh_chunk_steps(pd.n_objs, C, free_b / 2, &K), nullptr);
        if (const char *e = std::getenv("MIZUBA_AMD_CHUNK_STEPS"); e && *e) {
            const uint64_t cap = std::strtoull(e, nullptr, 10);
            if (cap) K = std::min(K, cap);
        }
    }
    if (K < C) {
        check_host();
        screen_in_ranges(*im, pd, conj_thresh, conj_det_interval, otypes, m, C, K);
        im->pj = pj.d;
        im->conj_thresh = conj_thresh;
        im->conj_det_interval = conj_det_interval;
        im->mode = m;
        im->otypes = otypes ? *otypes : std::vector<int32_t>(pd.n_objs, 1);
    } else {
        im->ctx->screen(conj_thresh, conj_det_interval, otypes, m);
        finish_impl(*im, pj, conj_thresh, conj_det_interval, otypes, m);
    }
    if (!im->data_dir.empty()) {
        py::gil_scoped_release rel;
        store_conjunctions(*im, im->data_dir);
    }
    return Conjunctions{im};
}

// Several independent problems screened together (mzh_screen_batch); the
// reference would construct one conjunctions object per problem.
std::vector<Conjunctions> make_conjunctions_batch(const std::vector<Polyjectory> &pjs,
                                                  const std::vector<double> &conj_threshs,
                                                  const std::vector<double> &conj_det_intervals,
                                                  std::optional<std::vector<std::optional<std::vector<int32_t>>>> otypes,
                                                  const std::string &mode, int device)
{
    const size_t n = pjs.size();
    if (conj_threshs.size() != n || conj_det_intervals.size() != n || (otypes && otypes->size() != n))
        throw std::invalid_argument("conjunctions_batch: the polyjectories, thresholds, intervals and object types "
                                    "must come in lists of the same length");
    const int m = parse_mode(mode);
    std::vector<std::shared_ptr<ConjImpl>> ims(n);
    std::vector<mzh_ctx *> ctxs(n);
    std::vector<const int32_t *> ots(n, nullptr);
    for (size_t k = 0; k < n; ++k)
        validate_screen(attached(pjs[k]), conj_threshs[k], conj_det_intervals[k],
                        otypes ? (*otypes)[k] : std::optional<std::vector<int32_t>>{}, m);
    for (size_t k = 0; k < n; ++k) {
        ims[k] = std::make_shared<ConjImpl>();
        ims[k]->ctx = std::make_unique<Ctx>(device);
        ims[k]->ctx->share(pjs[k].d, device);
        ctxs[k] = ims[k]->ctx->ctx;
        if (otypes && (*otypes)[k]) ots[k] = (*otypes)[k]->data();
    }
    std::vector<int> status(n, MZH_OK);
    {
        py::gil_scoped_release rel;
        (void)mzh_screen_batch(ctxs.data(), n, conj_threshs.data(), conj_det_intervals.data(), ots.data(), m,
                               status.data());
    }
    std::vector<Conjunctions> out;
    out.reserve(n);
    for (size_t k = 0; k < n; ++k) {
        check(status[k], ctxs[k]);
        finish_impl(*ims[k], pjs[k], conj_threshs[k], conj_det_intervals[k],
                    otypes ? (*otypes)[k] : std::optional<std::vector<int32_t>>{}, m);
        out.push_back(Conjunctions{ims[k]});
    }
    return out;
}

// A batch handle over Screener contexts, which it keeps alive.
struct BatchHandle {
    mzh_batch *b = nullptr;
    std::vector<std::shared_ptr<Ctx>> scrs;
    BatchHandle() = default;
    BatchHandle(const BatchHandle &) = delete;
    BatchHandle &operator=(const BatchHandle &) = delete;
    ~BatchHandle()
    {
        if (b) mzh_batch_destroy(b);
    }
};

std::shared_ptr<BatchHandle> make_batch_handle(const std::vector<std::shared_ptr<Ctx>> &scrs,
                                               const std::vector<double> &threshs,
                                               const std::vector<double> &intervals,
                                               std::optional<std::vector<std::optional<std::vector<int32_t>>>> otypes,
                                               const std::string &mode, bool graph, bool tune)
{
    const size_t n = scrs.size();
    if (threshs.size() != n || intervals.size() != n || (otypes && otypes->size() != n))
        throw std::invalid_argument("Batch: the screeners, thresholds, intervals and object types must come in lists "
                                    "of the same length");
    std::vector<mzh_ctx *> ctxs(n);
    std::vector<const int32_t *> ots(n, nullptr);
    for (size_t k = 0; k < n; ++k) {
        if (!scrs[k] || !scrs[k]->pj) throw std::invalid_argument("No polyjectory has been uploaded");
        ctxs[k] = scrs[k]->ctx;
        if (otypes && (*otypes)[k]) {
            check_otypes_size((*otypes)[k], scrs[k]->pj->n_objs);
            ots[k] = (*otypes)[k]->data();
        }
    }
    auto h = std::make_shared<BatchHandle>();
    h->scrs = scrs;
    const int flags = (graph ? MZH_BATCH_GRAPH : 0) | (tune ? MZH_BATCH_TUNE : 0);
    check(mzh_batch_create(ctxs.data(), n, threshs.data(), intervals.data(), ots.data(), parse_mode(mode), flags,
                           &h->b),
          nullptr);
    return h;
}

} // namespace

PYBIND11_MODULE(_core, m)
{
    m.doc() = "mizuba_amd: MI355X-native conjunction screening over polyjectories";

    // ---- polyjectory ---------------------------------------------------
    py::class_<Polyjectory> pt(m, "polyjectory", py::dynamic_attr());
    using OptPath = std::optional<std::filesystem::path>;
    pt.def(py::init([](const py::iterable &trajs_, const py::iterable &times_, const py::iterable &status_, double epoch,
                       double epoch2, OptPath data_dir, bool persist, OptPath tmpdir) {
               if (data_dir && tmpdir)
                   throw std::invalid_argument(
                       "The 'data_dir' and 'tmpdir' construction arguments cannot be provided at the same time");
               // expose_polyjectory.cpp:129-212: arrays are converted only when
               // their type differs; a float64 or int32 array is checked as it is
               py::array_t<int32_t> st_arr = status_.cast<py::array_t<int32_t>>();
               if (st_arr.ndim() != 1)
                   throw std::invalid_argument("A status array must have 1 dimension, but instead "
                                               + std::to_string(st_arr.ndim()) + " dimension(s) were detected");
               if (!(st_arr.flags() & py::array::c_style))
                   throw std::invalid_argument("The status array must be C contiguous and properly aligned");
               std::vector<py::array_t<double>> trajs, times;
               for (auto o : trajs_) trajs.push_back(o.cast<py::array_t<double>>());
               for (auto o : times_) times.push_back(o.cast<py::array_t<double>>());
               std::vector<int32_t> status(st_arr.data(), st_arr.data() + st_arr.shape(0));
               auto d = build_pj(trajs, times, status, epoch, epoch2);
               // files are written when a data dir, a temporary dir or persistence
               // is asked for; otherwise the polyjectory lives in memory only
               if ((data_dir && !data_dir->empty()) || tmpdir || persist) store_pj(*d, data_dir, tmpdir, persist);
               return Polyjectory{d};
           }),
           "trajs"_a, "times"_a, "status"_a, "epoch"_a = 0., "epoch2"_a = 0., "data_dir"_a = py::none(),
           "persist"_a = false, "tmpdir"_a = py::none());
    pt.def_static(
        "from_datafiles",
        [](const std::filesystem::path &traj_file, const std::filesystem::path &time_file, uint32_t order,
           const py::object &traj_offsets_, const std::vector<int32_t> &status, double epoch, double epoch2,
           OptPath data_dir, bool persist, OptPath tmpdir) {
            // expose_polyjectory.cpp:216-249
            py::array offs = numpy().attr("asarray")(traj_offsets_, "dtype"_a = traj_offset_dtype());
            if (offs.ndim() != 1)
                throw std::invalid_argument("The array of trajectory offsets passed to 'from_datafiles()' must have 1 "
                                            "dimension, but the number of dimensions is "
                                            + std::to_string(offs.ndim()) + " instead");
            offs = numpy().attr("ascontiguousarray")(offs);
            std::vector<mzh_host::TrajOffset> tv((size_t)offs.shape(0));
            if (!tv.empty()) std::memcpy(tv.data(), offs.data(), tv.size() * sizeof(mzh_host::TrajOffset));
            mzh_host::PjArrays a;
            std::filesystem::path dir;
            {
                py::gil_scoped_release rel;
                dir = mzh_host::pj_from_datafiles(traj_file, time_file, order, tv, status, epoch, epoch2, data_dir,
                                                  persist, tmpdir, a);
            }
            auto d = pj_from_arrays(std::move(a));
            d->data_dir = dir;
            d->persist = persist;
            return Polyjectory{d};
        },
        "traj_file"_a, "time_file"_a, "order"_a, "traj_offsets"_a, "status"_a, "epoch"_a = 0., "epoch2"_a = 0.,
        "data_dir"_a = py::none(), "persist"_a = false, "tmpdir"_a = py::none());
    pt.def_static(
        "mount",
        [](const std::filesystem::path &path) {
            mzh_host::PjArrays a;
            std::filesystem::path dir;
            {
                py::gil_scoped_release rel;
                dir = mzh_host::pj_mount(path, a);
            }
            auto d = pj_from_arrays(std::move(a));
            d->data_dir = dir;
            d->persist = true;  // only persistent data can be mounted
            return Polyjectory{d};
        },
        "data_dir"_a);
    pt.def_property_readonly_static("traj_offset", [](const py::object &) { return traj_offset_dtype(); });
    pt.def_property_readonly("n_objs", [](const Polyjectory &p) { return attached(p).n_objs; });
    pt.def_property_readonly("maxT", [](const Polyjectory &p) { return attached(p).maxT; });
    pt.def_property_readonly("epoch", [](const Polyjectory &p) {
        const auto &d = attached(p);
        return py::make_tuple(d.epoch, d.epoch2);
    });
    pt.def_property_readonly("poly_order", [](const Polyjectory &p) { return attached(p).order; });
    pt.def_property_readonly("data_dir", [](const Polyjectory &p) -> py::object {
        const auto &d = attached(p);
        if (d.data_dir.empty()) return py::none();  // in memory only
        return py::cast(d.data_dir);
    });
    pt.def_property_readonly("persist", [](const Polyjectory &p) { return attached(p).persist; });
    pt.def("detach", [](Polyjectory &p) { p.d.reset(); });
    pt.def_property_readonly("is_detached", [](const Polyjectory &p) { return !p.d; });
    pt.def("hint_release", [](const Polyjectory &p) {
        // the data is held in host memory, not in mapped files: the device
        // copies are what can be released (screenings made from them keep
        // what they need)
        auto &d = attached(p);
        std::lock_guard<std::mutex> le(d.eval_m);
        std::lock_guard<std::mutex> lk(d.dev_m);
        d.devs.clear();
    });
    pt.def_property_readonly("status", [](const py::object &self) {
        const auto &d = attached(self.cast<const Polyjectory &>());
        return ro_view(self, d.status.data(), {(py::ssize_t)d.n_objs});
    });
    pt.def("__getitem__", [](const py::object &self, py::ssize_t i) {
        const auto &d = attached(self.cast<const Polyjectory &>());
        if (i < 0 || (uint64_t)i >= d.n_objs)
            throw std::out_of_range("Invalid object index " + std::to_string(i)
                                    + " specified - the total number of objects in the polyjectory is only "
                                    + std::to_string(d.n_objs));
        const uint64_t ns = d.nsteps[i];
        const py::ssize_t op1 = d.order + 1;
        auto tr = ro_view(self, d.traj.data() + d.traj_off[i], {(py::ssize_t)ns, op1, 7});
        auto tm = ro_view(self, d.time.data() + d.time_off[i], {(py::ssize_t)(ns + (ns != 0 ? 1 : 0))});
        return py::make_tuple(tr, tm, d.status[i]);
    });
    pt.def("_flat", [](const py::object &self) {
        // flattened arrays (traj, time, traj_off, nsteps, time_off), for tests and tools
        const auto &d = attached(self.cast<const Polyjectory &>());
        return py::make_tuple(ro_view(self, d.traj.data(), {(py::ssize_t)d.traj.size()}),
                              ro_view(self, d.time.data(), {(py::ssize_t)d.time.size()}),
                              ro_view(self, d.traj_off.data(), {(py::ssize_t)d.n_objs}),
                              ro_view(self, d.nsteps.data(), {(py::ssize_t)d.n_objs}),
                              ro_view(self, d.time_off.data(), {(py::ssize_t)d.n_objs}));
    });
    pt.def("state_eval", &state_eval, "time"_a, "out"_a.noconvert() = py::none(), "obj_idx"_a = py::none());
    pt.def("state_meval", &state_meval, "time"_a, "out"_a.noconvert() = py::none(), "obj_idx"_a = py::none());

    // ---- polyjectories from SGP4 element sets ---------------------------
    m.attr("gpe_dtype") = gpe_dtype();
    m.def(
        "_make_sgp4_polyjectory",
        [](const py::array &gpes_, double jd_begin, double jd_end, double reentry_radius, double exit_radius,
           std::optional<std::filesystem::path> data_dir, bool persist, std::optional<std::filesystem::path> tmpdir,
           uint32_t order, int device) {
            if (data_dir && tmpdir)
                throw std::invalid_argument(
                    "The 'data_dir' and 'tmpdir' construction arguments cannot be provided at the same time");
            const auto gpes = gpes_from_array(gpes_);
            // the context that generates the polyjectory keeps it in HBM and
            // becomes the polyjectory's device copy on this device
            auto dev = std::make_unique<PjDevice>();
            check(mzh_create(device, &dev->ctx), nullptr);
            {
                py::gil_scoped_release rel;
                check(mzh_make_sgp4_polyjectory(dev->ctx, gpes.data(), gpes.size(), jd_begin, jd_end, reentry_radius,
                                                exit_radius, order),
                      dev->ctx);
            }
            auto d = pj_from_ctx(dev->ctx);
            if ((data_dir && !data_dir->empty()) || tmpdir || persist) store_pj(*d, data_dir, tmpdir, persist);
            d->devs[device] = std::move(dev);
            return Polyjectory{d};
        },
        "gpes"_a, "jd_begin"_a, "jd_end"_a, "reentry_radius"_a, "exit_radius"_a, "data_dir"_a, "persist"_a,
        "tmpdir"_a, "order"_a, "device"_a);
    m.def(
        "sgp4_propagate",
        [](const py::array &gpes_, const py::array_t<double> &tsince_, int device) {
            // states of every GPE at the minutes since its epoch tsince[g, k]
            // (or tsince[k] for all): (out [n_gpes][n_times][6], err [n_gpes][n_times])
            const auto gpes = gpes_from_array(gpes_);
            const uint64_t ng = gpes.size();
            py::array_t<double> tsince = numpy().attr("ascontiguousarray")(tsince_);
            uint64_t nt = 0;
            std::vector<double> t;
            if (tsince.ndim() == 1) {
                nt = (uint64_t)tsince.shape(0);
                t.resize(ng * nt);
                for (uint64_t g = 0; g < ng; ++g) std::memcpy(t.data() + g * nt, tsince.data(), nt * sizeof(double));
            } else if (tsince.ndim() == 2 && (uint64_t)tsince.shape(0) == ng) {
                nt = (uint64_t)tsince.shape(1);
                t.assign(tsince.data(), tsince.data() + ng * nt);
            } else {
                throw std::invalid_argument("sgp4_propagate(): the times must be a 1-D array or a 2-D array with one "
                                            "row per GPE");
            }
            py::array_t<double> out(std::vector<py::ssize_t>{(py::ssize_t)ng, (py::ssize_t)nt, 6});
            py::array_t<int32_t> err(std::vector<py::ssize_t>{(py::ssize_t)ng, (py::ssize_t)nt});
            Ctx c(device);
            {
                py::gil_scoped_release rel;
                check(mzh_sgp4_propagate(c.ctx, gpes.data(), ng, t.data(), nt, out.mutable_data(),
                                         err.mutable_data()),
                      c.ctx);
            }
            return py::make_tuple(out, err);
        },
        "gpes"_a, "tsince"_a, "device"_a = 0);

    // tmpdir getter/setter (core.cpp:63-86): None stands for no global dir
    m.def("get_tmpdir", []() -> std::optional<std::filesystem::path> {
        auto p = mzh_host::get_global_tmpdir();
        if (p.empty()) return std::nullopt;
        return p;
    });
    m.def(
        "set_tmpdir",
        [](std::optional<std::filesystem::path> p) { mzh_host::set_global_tmpdir(p ? *p : std::filesystem::path()); },
        "path"_a);

    m.def(
        "make_synthetic_polyjectory",
        [](uint32_t n_objs, double days, uint64_t seed, double frac_high, uint32_t order, uint32_t n_shell) {
            if (n_objs == 0) throw std::invalid_argument("n_objs must be positive");
            if (!(days > 0) || !std::isfinite(days)) throw std::invalid_argument("days must be finite and positive");
            if (order < 2 || order > 20) throw std::invalid_argument("order must be in [2, 20]");
            mzh_host::SynthParams p;
            p.n_objs = n_objs;
            p.days = days;
            p.seed = seed;
            p.frac_high = frac_high;
            p.order = order;
            p.n_shell = n_shell;
            mzh_host::SynthOut out;
            {
                py::gil_scoped_release rel;
                out = mzh_host::make_synthetic(p);
            }
            return Polyjectory{pj_from_synth(std::move(out))};
        },
        "n_objs"_a, "days"_a, "seed"_a = 42, "frac_high"_a = 0.2, "order"_a = 8, "n_shell"_a = 0);

    // ---- device context for repeated screening of one polyjectory ------
    py::class_<DeviceBuffer, std::shared_ptr<DeviceBuffer>>(m, "DeviceBuffer")
        .def_property_readonly("ptr", [](const DeviceBuffer &b) { return reinterpret_cast<uintptr_t>(b.p); })
        .def_property_readonly("nbytes", [](const DeviceBuffer &b) { return b.nbytes; })
        .def("write", &DeviceBuffer::write, "array"_a)
        .def("read", &DeviceBuffer::read, "array"_a);

    py::class_<Ctx, std::shared_ptr<Ctx>>(m, "Screener")
        .def(py::init<int>(), "device"_a = 0)
        .def(
            "upload",
            [](Ctx &c, const Polyjectory &p, bool copy) {
                // the polyjectory's device copy on this context's device is
                // shared (made on first use); copy=True uploads a new one
                (void)attached(p);
                if (copy) {
                    c.upload(p.d);
                } else {
                    c.share(p.d, c.device);
                }
            },
            "pj"_a, "copy"_a = false)
        .def(
            "screen",
            [](Ctx &c, double thresh, double interval, std::optional<std::vector<int32_t>> otypes,
               const std::string &mode, std::optional<std::pair<uint64_t, uint64_t>> steps) {
                if (!c.pj) throw std::invalid_argument("No polyjectory has been uploaded");
                if (!steps) {
                    c.screen(thresh, interval, otypes, parse_mode(mode));
                    return;
                }
                // the conjunction steps [steps[0], steps[1]) only (mzh_screen_steps)
                if (otypes && otypes->size() != c.pj->n_objs)
                    throw std::invalid_argument("Invalid array of object types passed to the constructor of a "
                                                "conjunctions objects: the expected size is "
                                                + std::to_string(c.pj->n_objs) + ", but the actual size is "
                                                + std::to_string(otypes->size()) + " instead");
                const int m = parse_mode(mode);
                int st;
                {
                    py::gil_scoped_release rel;
                    st = mzh_screen_steps(c.ctx, thresh, interval, otypes ? otypes->data() : nullptr, m, steps->first,
                                          steps->second);
                }
                check(st, c.ctx);
            },
            "conj_thresh"_a, "conj_det_interval"_a, "otypes"_a = py::none(), "mode"_a = "fp64",
            "steps"_a = py::none())
        .def("stage_times",
             [](Ctx &c) {
                 double t[MZH_STAGE_COUNT];
                 check(mzh_get_stage_times(c.ctx, t), c.ctx);
                 py::dict d;
                 const char *names[] = {"aabbs", "morton", "bvh", "broad", "narrow", "total", "aabb_kernel", "narrow_kernel", "broad_kernel"};
                 for (int k = 0; k < MZH_STAGE_COUNT; ++k) d[names[k]] = t[k];
                 return d;
             })
        .def("np_report",
             [](Ctx &c) {
                 uint64_t r[MZH_NP_COUNT];
                 check(mzh_get_np_report(c.ctx, r), c.ctx);
                 py::dict d;
                 const char *names[] = {"n_tot_conj_candidates", "n_dist2_check",      "n_poly_roots",
                                        "n_fex_check",           "n_poly_no_roots",    "n_tot_dist_minima",
                                        "n_tot_discarded_dist_minima"};
                 for (int k = 0; k < MZH_NP_COUNT; ++k) d[names[k]] = r[k];
                 return d;
             })
        .def("n_cd_steps",
             [](Ctx &c) {
                 uint64_t n;
                 check(mzh_get_n_cd_steps(c.ctx, &n), c.ctx);
                 return n;
             })
        .def("n_pairs",
             [](Ctx &c) {
                 uint64_t C;
                 check(mzh_get_n_cd_steps(c.ctx, &C), c.ctx);
                 std::vector<uint64_t> off(2 * C);
                 check(mzh_copy_bp_offsets(c.ctx, off.data()), c.ctx);
                 return C ? off[2 * (C - 1)] + off[2 * (C - 1) + 1] : 0;
             })
        .def(
            "state_eval_device",
            [](Ctx &c, uintptr_t tm, int time_mode, uintptr_t sel, uint64_t n_rows, uint64_t n_times, uintptr_t out) {
                // device pointers (e.g. torch tensors' data_ptr()); sel may be 0
                if (!c.pj) throw std::invalid_argument("No polyjectory has been uploaded");
                int st;
                {
                    py::gil_scoped_release rel;
                    st = mzh_state_eval_device(c.ctx, reinterpret_cast<const double *>(tm), time_mode,
                                               reinterpret_cast<const uint64_t *>(sel), n_rows, n_times,
                                               reinterpret_cast<double *>(out));
                }
                check(st, c.ctx);
            },
            "tm"_a, "time_mode"_a, "sel"_a, "n_rows"_a, "n_times"_a, "out"_a)
        .def(
            "make_sgp4",
            [](Ctx &c, const py::array &gpes_, double jd_begin, double jd_end, double reentry_radius,
               double exit_radius, uint32_t order) {
                // generate the polyjectory of the GPEs into this context, where it
                // stays for screening; returns the per-satellite statuses
                const auto gpes = gpes_from_array(gpes_);
                {
                    py::gil_scoped_release rel;
                    check(mzh_make_sgp4_polyjectory(c.ctx, gpes.data(), gpes.size(), jd_begin, jd_end,
                                                    reentry_radius, exit_radius, order),
                          c.ctx);
                }
                // the data stay in HBM: only the sizes and statuses come back
                auto d = std::make_shared<PjData>();
                uint64_t n = 0, tl = 0, ml = 0;
                double ep[2] = {0, 0};
                check(mzh_get_polyjectory_info(c.ctx, &n, &d->order, &tl, &ml, &d->maxT, ep), c.ctx);
                d->n_objs = (uint32_t)n;
                d->epoch = ep[0];
                d->epoch2 = ep[1];
                d->status.resize(n);
                check(mzh_copy_polyjectory(c.ctx, nullptr, nullptr, nullptr, nullptr, nullptr, d->status.data()),
                      c.ctx);
                c.pj = d;
                py::array_t<int32_t> status((py::ssize_t)n);
                std::copy(d->status.begin(), d->status.end(), status.mutable_data());
                return status;
            },
            "gpes"_a, "jd_begin"_a, "jd_end"_a, "reentry_radius"_a = 0.0,
            "exit_radius"_a = std::numeric_limits<double>::infinity(), "order"_a = 8)
        .def("device_buffer",
             [](const std::shared_ptr<Ctx> &c, uint64_t nbytes) { return std::make_shared<DeviceBuffer>(c, nbytes); },
             "nbytes"_a)
        .def("pj_info",
             [](Ctx &c) {
                 // sizes of the context's polyjectory
                 uint64_t n = 0, tl = 0, ml = 0;
                 uint32_t order = 0;
                 double maxT = 0, ep[2] = {0, 0};
                 check(mzh_get_polyjectory_info(c.ctx, &n, &order, &tl, &ml, &maxT, ep), c.ctx);
                 py::dict d;
                 d["n_objs"] = n;
                 d["order"] = order;
                 d["n_steps"] = tl / (7u * (order + 1u));
                 d["maxT"] = maxT;
                 d["epoch"] = py::make_tuple(ep[0], ep[1]);
                 return d;
             })
        .def("sgp4_times",
             [](Ctx &c) {
                 double t[5];
                 check(mzh_get_sgp4_times(c.ctx, t), c.ctx);
                 py::dict d;
                 const char *names[5] = {"seg_init", "count", "scan", "write", "transpose"};
                 for (int k = 0; k < 5; ++k) d[names[k]] = t[k];
                 return d;
             })
        .def("eval_time",
             [](Ctx &c) {
                 double s = 0;
                 check(mzh_get_eval_time(c.ctx, &s), c.ctx);
                 return s;
             })
        .def(
            "root_finding",
            [](Ctx &c, const py::array_t<double, py::array::c_style | py::array::forcecast> &polys,
               const py::array_t<double, py::array::c_style | py::array::forcecast> &rf_int) {
                // the narrow phase's root finder on free-standing polynomials
                // (mzh_root_finding): returns (roots [n][2 order + 2], counts, flags)
                if (polys.ndim() != 2 || rf_int.ndim() != 1 || rf_int.shape(0) != polys.shape(0))
                    throw std::invalid_argument("root_finding: polys must be [n, order + 1] and rf_int [n]");
                const uint64_t n = polys.shape(0);
                const uint32_t order = (uint32_t)polys.shape(1) - 1u;
                py::array_t<double> roots({(py::ssize_t)n, (py::ssize_t)(2 * order + 2)});
                py::array_t<uint32_t> cnt((py::ssize_t)n), flags((py::ssize_t)n);
                std::fill(roots.mutable_data(), roots.mutable_data() + roots.size(), 0.0);
                int st;
                {
                    py::gil_scoped_release rel;
                    st = mzh_root_finding(c.ctx, order, n, polys.data(), rf_int.data(), roots.mutable_data(),
                                          cnt.mutable_data(), flags.mutable_data());
                }
                check(st, c.ctx);
                return py::make_tuple(roots, cnt, flags);
            },
            "polys"_a, "rf_int"_a)
        .def("conjunctions", [](Ctx &c) {
            uint64_t n;
            check(mzh_get_n_conjunctions(c.ctx, &n), c.ctx);
            py::array out(conj_dtype(), std::vector<py::ssize_t>{(py::ssize_t)n});
            check(mzh_copy_conjunctions(c.ctx, static_cast<mzh_conj *>(out.mutable_data())), c.ctx);
            return out;
        });

    m.def(
        "screen_batch",
        [](const std::vector<Ctx *> &scrs, const std::vector<double> &threshs, const std::vector<double> &intervals,
           const std::string &mode) {
            const size_t n = scrs.size();
            if (threshs.size() != n || intervals.size() != n)
                throw std::invalid_argument("screen_batch: the screeners, thresholds and intervals must come in lists "
                                            "of the same length");
            std::vector<mzh_ctx *> ctxs(n);
            for (size_t k = 0; k < n; ++k) {
                if (!scrs[k] || !scrs[k]->pj) throw std::invalid_argument("No polyjectory has been uploaded");
                ctxs[k] = scrs[k]->ctx;
            }
            std::vector<int> status(n, MZH_OK);
            {
                py::gil_scoped_release rel;
                (void)mzh_screen_batch(ctxs.data(), n, threshs.data(), intervals.data(), nullptr, parse_mode(mode),
                                       status.data());
            }
            for (size_t k = 0; k < n; ++k) check(status[k], ctxs[k]);
        },
        "screeners"_a, "conj_threshs"_a, "conj_det_intervals"_a, "mode"_a = "fp64",
        "Screen the polyjectories resident in several Screener objects together (mzh_screen_batch).");

    // ---- a batch handle over Screener contexts ----------------------------
    py::class_<BatchHandle, std::shared_ptr<BatchHandle>>(m, "Batch")
        .def(py::init(&make_batch_handle), "screeners"_a, "conj_threshs"_a, "conj_det_intervals"_a,
             "otypes"_a = py::none(), "mode"_a = "fp64", "graph"_a = true, "tune"_a = false,
             "The same problems screened again and again (mzh_batch_create): one launch per stage for all of "
             "them; graph: replay the captured launch sequence after the first run; tune: time the launch shapes "
             "on the first run.")
        .def("run",
             [](BatchHandle &b) {
                 std::vector<int> st(b.scrs.size(), MZH_OK);
                 {
                     py::gil_scoped_release rel;
                     (void)mzh_batch_run(b.b, st.data());
                 }
                 for (size_t k = 0; k < st.size(); ++k) check(st[k], b.scrs[k]->ctx);
             })
        .def_property_readonly("replays",
                               [](const BatchHandle &b) {
                                   uint64_t r = 0;
                                   mzh_batch_info(b.b, &r, nullptr, nullptr);
                                   return r;
                               })
        .def_property_readonly("seconds",
                               [](const BatchHandle &b) {
                                   double t = 0;
                                   mzh_batch_info(b.b, nullptr, &t, nullptr);
                                   return t;
                               })
        .def_property_readonly("shape", [](const BatchHandle &b) {
            uint32_t sh[5] = {0, 0, 0, 0, 0};
            mzh_batch_info(b.b, nullptr, nullptr, sh);
            py::dict d;
            d["bucket_bits"] = sh[0];
            d["sort_wgs"] = sh[1];
            d["bvh_wgs"] = sh[2];
            d["bp_block"] = sh[3];
            d["roots_wgs"] = sh[4];
            return d;
        });

    m.def("conjunctions_batch", &make_conjunctions_batch, "pjs"_a, "conj_threshs"_a, "conj_det_intervals"_a,
          "otypes"_a = py::none(), "mode"_a = "fp64", "device"_a = 0,
          "Screen several independent polyjectories together on one GPU; returns one conjunctions object per "
          "problem, identical to constructing them one by one.");

    m.def(
        "check_screen",
        [](uint64_t n_objs, double maxT, double conj_thresh, double conj_det_interval,
           std::optional<std::vector<int32_t>> otypes, const std::string &mode) {
            // the constructor's validation without a device (mzh_check_screen)
            const int m = parse_mode(mode);
            uint64_t C = 0;
            check(mzh_check_screen(n_objs, maxT, conj_thresh, conj_det_interval, nullptr, m, &C), nullptr);
            check_otypes_size(otypes, n_objs);
            if (otypes)
                check(mzh_check_screen(n_objs, maxT, conj_thresh, conj_det_interval, otypes->data(), m, &C), nullptr);
            return C;
        },
        "n_objs"_a, "maxT"_a, "conj_thresh"_a, "conj_det_interval"_a, "otypes"_a = py::none(), "mode"_a = "fp64",
        "Validate screening arguments as conjunctions() does and return the number of conjunction steps.");

    m.def("device_count", []() {
        int n = 0;
        mzh_device_count(&n);
        return n;
    });

    // ---- conjunctions -------------------------------------------------
    py::class_<Conjunctions> cj(m, "conjunctions", py::dynamic_attr());
    cj.def(py::init(&make_conjunctions), "pj"_a, "conj_thresh"_a, "conj_det_interval"_a, "otypes"_a = py::none(),
           "data_dir"_a = py::none(), "tmpdir"_a = py::none(), "mode"_a = "fp64", "device"_a = 0);
    cj.def_property_readonly("n_cd_steps", [](const Conjunctions &c) { return c.impl->C; });
    cj.def_property_readonly("mode", [](const Conjunctions &c) {
        return std::string(c.impl->mode == MZH_MODE_FP64 ? "fp64" : "mixed");
    });
    cj.def_property_readonly("aabbs", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        const auto &v = lazy<float>(im, im.aabbs, im.C * (im.N + 1) * 8,
                                    [&](float *p) { return mzh_copy_aabbs(im.ctx->ctx, p); });
        return ro_view(self, v.data(), {(py::ssize_t)im.C, (py::ssize_t)im.N + 1, 2, 4});
    });
    cj.def_property_readonly("srt_aabbs", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        const auto &v = lazy<float>(im, im.srt_aabbs, im.C * (im.N + 1) * 8,
                                    [&](float *p) { return mzh_copy_srt_aabbs(im.ctx->ctx, p); });
        return ro_view(self, v.data(), {(py::ssize_t)im.C, (py::ssize_t)im.N + 1, 2, 4});
    });
    cj.def_property_readonly("mcodes", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        const auto &v = lazy<uint64_t>(im, im.mcodes, im.C * im.N,
                                       [&](uint64_t *p) { return mzh_copy_mcodes(im.ctx->ctx, p); });
        return ro_view(self, v.data(), {(py::ssize_t)im.C, (py::ssize_t)im.N});
    });
    cj.def_property_readonly("srt_mcodes", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        const auto &v = lazy<uint64_t>(im, im.srt_mcodes, im.C * im.N,
                                       [&](uint64_t *p) { return mzh_copy_srt_mcodes(im.ctx->ctx, p); });
        return ro_view(self, v.data(), {(py::ssize_t)im.C, (py::ssize_t)im.N});
    });
    cj.def_property_readonly("srt_idx", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        const auto &v = lazy<uint32_t>(im, im.srt_idx, im.C * im.N,
                                       [&](uint32_t *p) { return mzh_copy_srt_idx(im.ctx->ctx, p); });
        return ro_view(self, v.data(), {(py::ssize_t)im.C, (py::ssize_t)im.N});
    });
    cj.def_property_readonly("cd_end_times", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        return ro_view(self, im.cd_end_times.data(), {(py::ssize_t)im.C});
    });
    cj.def("get_bvh_tree", [](const py::object &self, uint64_t i) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        std::lock_guard<std::mutex> lk(im.m);
        auto it = im.trees.find(i);
        if (it == im.trees.end() && im.chunk_steps)
            throw std::out_of_range("Cannot fetch the BVH tree for the conjunction timestep at index "
                                    + std::to_string(i) + ": the total number of conjunction steps is only "
                                    + std::to_string(im.C));
        if (it == im.trees.end()) {
            uint64_t n = 0;
            check(mzh_get_bvh_tree_size(im.ctx->ctx, i, &n), im.ctx->ctx);
            std::vector<mzh_bvh_node> v(n ? n : 1);
            check(mzh_copy_bvh_tree(im.ctx->ctx, i, v.data()), im.ctx->ctx);
            v.resize(n);
            it = im.trees.emplace(i, std::move(v)).first;
        }
        return ro_view_dtype(self, bvh_node_dtype(), it->second.data(), (py::ssize_t)it->second.size());
    });
    cj.def("get_aabb_collisions", [](const py::object &self, uint64_t i) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        std::lock_guard<std::mutex> lk(im.m);
        auto it = im.bps.find(i);
        if (it == im.bps.end() && im.chunk_steps)
            throw std::out_of_range("Cannot fetch the list of AABB collisions for the conjunction timestep at index "
                                    + std::to_string(i) + ": the total number of conjunction steps is only "
                                    + std::to_string(im.C));
        if (it == im.bps.end()) {
            uint64_t n = 0;
            check(mzh_get_aabb_collisions_size(im.ctx->ctx, i, &n), im.ctx->ctx);
            std::vector<mzh_aabb_collision> v(n ? n : 1);
            check(mzh_copy_aabb_collisions(im.ctx->ctx, i, v.data()), im.ctx->ctx);
            v.resize(n);
            it = im.bps.emplace(i, std::move(v)).first;
        }
        return ro_view_dtype(self, aabb_collision_dtype(), it->second.data(), (py::ssize_t)it->second.size());
    });
    cj.def_property_readonly("conjunctions", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        uint64_t n = 0;
        if (im.chunk_steps) {
            n = im.conjs->size();
        } else {
            check(mzh_get_n_conjunctions(im.ctx->ctx, &n), im.ctx->ctx);
        }
        const auto &v = lazy<mzh_conj>(im, im.conjs, n, [&](mzh_conj *p) { return mzh_copy_conjunctions(im.ctx->ctx, p); });
        return ro_view_dtype(self, conj_dtype(), v.data(), (py::ssize_t)n);
    });
    cj.def_property_readonly("otypes", [](const py::object &self) {
        auto &im = *self.cast<const Conjunctions &>().impl;
        return ro_view(self, im.otypes.data(), {(py::ssize_t)im.N});
    });
    cj.def_property_readonly("conj_thresh", [](const Conjunctions &c) { return c.impl->conj_thresh; });
    cj.def_property_readonly("conj_det_interval", [](const Conjunctions &c) { return c.impl->conj_det_interval; });
    cj.def("hint_release", [](const Conjunctions &c) { conj_hint_release(*c.impl); });
    cj.def_property_readonly("data_dir", [](const Conjunctions &c) -> py::object {
        if (c.impl->data_dir.empty()) return py::none();
        return py::cast(c.impl->data_dir);
    });
    cj.def_property_readonly("np_report", [](const Conjunctions &c) {
        uint64_t r[MZH_NP_COUNT];
        if (c.impl->chunk_steps) {
            std::copy(c.impl->np, c.impl->np + MZH_NP_COUNT, r);
        } else {
            check(mzh_get_np_report(c.impl->ctx->ctx, r), c.impl->ctx->ctx);
        }
        py::dict d;
        const char *names[] = {"n_tot_conj_candidates", "n_dist2_check",     "n_poly_roots",
                               "n_fex_check",           "n_poly_no_roots",   "n_tot_dist_minima",
                               "n_tot_discarded_dist_minima"};
        for (int k = 0; k < MZH_NP_COUNT; ++k) d[names[k]] = r[k];
        return d;
    });
    cj.def_property_readonly("chunk_steps", [](const Conjunctions &c) { return c.impl->chunk_steps; },
                             "Conjunction steps per range of a screening run in ranges (0: one batch).");
    cj.def_property_readonly("stage_times", [](const Conjunctions &c) {
        double t[MZH_STAGE_COUNT];
        if (c.impl->chunk_steps) {
            std::copy(c.impl->stage, c.impl->stage + MZH_STAGE_COUNT, t);
        } else {
            check(mzh_get_stage_times(c.impl->ctx->ctx, t), c.impl->ctx->ctx);
        }
        py::dict d;
        const char *names[] = {"aabbs", "morton", "bvh", "broad", "narrow", "total", "aabb_kernel", "narrow_kernel", "broad_kernel"};
        for (int k = 0; k < MZH_STAGE_COUNT; ++k) d[names[k]] = t[k];
        return d;
    });
    cj.def_property_readonly_static("bvh_node", [](const py::object &) { return bvh_node_dtype(); });
    cj.def_property_readonly_static("aabb_collision", [](const py::object &) { return aabb_collision_dtype(); });
    cj.def_property_readonly_static("conj", [](const py::object &) { return conj_dtype(); });
}
