// host_util.h -- helpers of the drop-in C++ classes (device buffers, BC strings, errors).
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "asd2_hip.h"
#include "common.h"

namespace asd2host {

// The reference terminates on fatal solver errors (print + exit(1)); the drop-in classes
// keep that convention for GPU failures.
[[noreturn]] inline void fatal(const char* where) {
    std::cerr << where << ": " << asd2_last_error() << std::endl;
    std::exit(1);
}

inline void check(int status, const char* where) {
    if (status != ASD2_OK) fatal(where);
}

inline asd2_ctx* make_ctx(int nx, int ny, int g, double bx, double by, double dx, double dy) {
    asd2_grid gr{nx, ny, g, bx, by, dx > 0 ? dx : 1.0, dy > 0 ? dy : 1.0};
    asd2_ctx* c = nullptr;
    check(asd2_create(&gr, &c), "asd2_create");
    return c;
}

// host <-> device traffic of the classes (asd2::Transfers)
void count_transfer(bool upload, size_t bytes);

// A device copy of a host vector (or of a plain host array).
class DevVec {
public:
    explicit DevVec(size_t n) : n_(n) { check(asd2_malloc(reinterpret_cast<void**>(&p_), std::max<size_t>(n, 1) * 8), "asd2_malloc"); }
    DevVec(const std::vector<double>& h) : DevVec(h.size()) { put(h); }
    ~DevVec() { asd2_free(p_); }
    DevVec(const DevVec&) = delete;
    DevVec& operator=(const DevVec&) = delete;
    void put(const std::vector<double>& h) { put(h.data(), h.size()); }
    void put(const double* h, size_t n) {
        check(asd2_memcpy_h2d(p_, h, std::min(n, n_) * 8), "upload");
        count_transfer(true, std::min(n, n_) * 8);
    }
    void get(std::vector<double>& h) const {
        h.resize(n_);
        check(asd2_memcpy_d2h(h.data(), p_, n_ * 8), "download");
        count_transfer(false, n_ * 8);
    }
    std::vector<double> get() const {
        std::vector<double> h;
        get(h);
        return h;
    }
    double* p() const { return p_; }
    size_t size() const { return n_; }

private:
    double* p_ = nullptr;
    size_t n_ = 0;
};

// ------------------------------------------------------------------ device residency
// The process-wide table of device mirrors of callers' vectors (resident mode, asd2_resident.h),
// keyed by the vector's buffer.  Each mirror records which side is newer and the context whose
// stream wrote it last: a context that reads or writes a mirror another context wrote first
// waits for that context (the contexts' streams are not ordered with each other).
namespace mirror {
bool resident();
// the mirror of h, or nullptr; ctx's access is ordered after the mirror's last writer
double* find(asd2_ctx* ctx, const std::vector<double>& h);
// the mirror of h (created when absent), holding h's contents when `current` is asked for
double* acquire(asd2_ctx* ctx, const std::vector<double>& h, bool current);
// ctx has written the mirror of h: the device copy is the newer one
void written(asd2_ctx* ctx, const std::vector<double>& h);
void sync_host(const std::vector<double>& h);
void forget(const std::vector<double>& h);
// ctx is about to be destroyed: wait for it and clear it as a writer
void release(asd2_ctx* ctx);
}  // namespace mirror

// The device side of a drop-in solver: its context, named scratch buffers, and the access of its
// methods to the callers' vectors.  Default mode: in() uploads, wrote() downloads (the host vectors
// stay current).  Resident mode: vectors with a mirror are used in place; in-place arguments and
// outputs get one.
struct SolverDevice {
    asd2_ctx* ctx = nullptr;
    size_t n = 0;
    const char* who = "";
    std::map<std::string, std::unique_ptr<DevVec>> bufs;
    SolverDevice() = default;
    SolverDevice(const SolverDevice&) = delete;
    SolverDevice& operator=(const SolverDevice&) = delete;
    ~SolverDevice() {
        if (ctx) mirror::release(ctx);
        bufs.clear();
        if (ctx) asd2_destroy(ctx);
    }
    void check_size(const std::vector<double>& h, const std::string& name) const {
        if (h.size() != n) {
            std::cerr << who << ": array '" << name << "' has " << h.size() << " values, expected " << n << std::endl;
            std::exit(1);
        }
    }
    double* scratch(const std::string& name) {
        auto& b = bufs[name];
        if (!b) b.reset(new DevVec(n));
        return b->p();
    }
    // h for reading
    double* in(const std::string& name, const std::vector<double>& h) {
        check_size(h, name);
        if (mirror::resident())
            if (double* p = mirror::find(ctx, h)) return p;
        double* p = scratch(name);
        bufs[name]->put(h);
        return p;
    }
    // h for reading and writing
    double* inout(const std::string& name, std::vector<double>& h) {
        check_size(h, name);
        if (mirror::resident()) return mirror::acquire(ctx, h, true);
        double* p = scratch(name);
        bufs[name]->put(h);
        return p;
    }
    // h for writing every element (its contents are not read); sized n if it is not
    double* out(const std::string& name, std::vector<double>& h) {
        if (h.size() != n) h.assign(n, 0.0);
        if (mirror::resident()) return mirror::acquire(ctx, h, false);
        return scratch(name);
    }
    // after the launches that wrote h through inout() / out()
    void wrote(const std::string& name, std::vector<double>& h) {
        if (mirror::resident()) mirror::written(ctx, h);
        else bufs[name]->get(h);
    }
    // a by-value result held in a scratch buffer
    std::vector<double> result(const std::string& name) { return bufs[name]->get(); }
    // h's host copy current before a host loop reads it
    static void host_read(const std::vector<double>& h) {
        if (mirror::resident()) mirror::sync_host(h);
    }
};

inline std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
    return s;
}

// the string vocabulary of BoundaryCondition2D::SetBC_*_2D (reference src/bc2d.cpp:16-344)
inline BC2D parse_bc(const std::string& s, bool pressure_grid) {
    const std::string k = lower(s);
    if (k == "periodic") return BC2D::PERIODIC;
    if (k == "neumann") return BC2D::NEUMANN;
    if (k == "axisym") return BC2D::AXISYM;
    if (k == "wall") return BC2D::WALL;
    if (k == "dirichlet") return BC2D::DIRICHLET;
    if (k == "inlet") return BC2D::INLET;
    if (k == "outlet") return BC2D::OUTLET;
    if (pressure_grid && k == "pressure") return BC2D::PRESSURE;
    if (pressure_grid && k == "lsfree") return BC2D::LSFREE;
    return BC2D::CUSTOM;
}

inline int code(BC2D k) { return static_cast<int>(k); }

}  // namespace asd2host
