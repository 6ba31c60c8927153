// Host build of haissinskijulia_amd/csrc/hm_math.hpp for tests/test_math.py:
// the kernels' elementary functions, compiled for the CPU with the same
// source, exported over arrays.
#include <cmath>
#include <cstdint>

#define HM_QUAL inline
#include "../haissinskijulia_amd/csrc/hm_math.hpp"

struct HostOps {
    static double rcp(double a) { return 1.0 / a; }
    static double frexp_mant(double u) { int e; return std::frexp(u, &e); }
    static int frexp_exp(double u) { int e; std::frexp(u, &e); return e; }
};

extern "C" {
void t_sin(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = hm::fm::sin_reduced(x[i]);
}
void t_sincos(const double *x, double *s, double *c, int64_t n) {
    for (int64_t i = 0; i < n; ++i) hm::fm::sincos_reduced(x[i], s[i], c[i]);
}
void t_log(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = hm::fm::log_pos<HostOps>(x[i]);
}
void t_rcp(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = hm::fm::rcp<HostOps>(x[i]);
}
}
