// Shared definitions of the HIP backend (gfx950 / MI355X only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace bdh {

// Last error message of the calling thread (returned by bdh_last_error()).
void set_error(const std::string& msg);
const char* last_error();

#define BDH_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::bdh::set_error(std::string(#call) + ": " + hipGetErrorString(e_) +        \
                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");   \
            return -1;                                                                  \
        }                                                                               \
    } while (0)

#define BDH_CHECK(cond, msg)                                                            \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::bdh::set_error(std::string(msg));                                         \
            return -2;                                                                  \
        }                                                                               \
    } while (0)

#define BDH_LAUNCH_CHECK()                                                              \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::bdh::set_error(std::string("kernel launch: ") + hipGetErrorString(e_) +   \
                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");   \
            return -1;                                                                  \
        }                                                                               \
    } while (0)

// Physical constants of the reference (src/data.h:224-228, 254).
constexpr double kB = 0.001986248532617;
constexpr double kB_1_8PI = 0.00007903031804378703;
constexpr double kB_1_6PI = 0.00010537375739171604;
constexpr double SQRTPI = 1.77245385090551602729;
constexpr double PI_F = (double)3.14159265358979323846f;  // the reference's float M_PIF

// Pair separation shared by every pair kernel, evaluated without FMA contraction so the
// brute-force and cell-list paths (and a plain FP64 host loop) see identical distances.
// Minimum image of trans_dist_vec (trans.c:111-121) when pbc is set.
__device__ __forceinline__ double pair_r2(double& dx, double& dy, double& dz, int pbc, double box)
{
#pragma clang fp contract(off)
    if (pbc) {
        const double inv = 1.0 / box;
        dx = dx - box * floor(dx * inv + 0.5);
        dy = dy - box * floor(dy * inv + 0.5);
        dz = dz - box * floor(dz * inv + 0.5);
    }
    return dx * dx + dy * dy + dz * dz;
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
inline long long ceil_div_ll(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace bdh
