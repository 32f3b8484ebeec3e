// host_util.h -- helpers of the drop-in C++ classes (device buffers, BC strings, errors).
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <iostream>
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

// A device copy of a host vector (or of a plain host array).
class DevVec {
public:
    explicit DevVec(size_t n) : n_(n) { check(asd2_malloc(reinterpret_cast<void**>(&p_), std::max<size_t>(n, 1) * 8), "asd2_malloc"); }
    DevVec(const std::vector<double>& h) : DevVec(h.size()) { put(h); }
    ~DevVec() { asd2_free(p_); }
    DevVec(const DevVec&) = delete;
    DevVec& operator=(const DevVec&) = delete;
    void put(const std::vector<double>& h) { put(h.data(), h.size()); }
    void put(const double* h, size_t n) { check(asd2_memcpy_h2d(p_, h, std::min(n, n_) * 8), "upload"); }
    void get(std::vector<double>& h) const {
        h.resize(n_);
        check(asd2_memcpy_d2h(h.data(), p_, n_ * 8), "download");
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
