// xdem_common.hpp — shared host/device definitions of libxdem (gfx950).
//
// Error convention: every C entry point returns an int32 status and records a
// message retrievable with xdem_last_error() (thread-local), replacing the
// reference's print-and-exit checkCudaErrors.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/xdem.h"

namespace xdem {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
hipStream_t stream();

#define XDEM_HIP(call)                                                                 \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::xdem::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,             \
                              hipGetErrorString(e_));                                  \
            return XDEM_ERR_HIP;                                                       \
        }                                                                              \
    } while (0)

#define XDEM_REQUIRE(cond, ...)                                                        \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::xdem::set_error(__VA_ARGS__);                                            \
            return XDEM_ERR_INVALID;                                                   \
        }                                                                              \
    } while (0)

#define XDEM_TRY(call)                                                                 \
    do {                                                                               \
        int32_t rc_ = (call);                                                          \
        if (rc_ != XDEM_OK) return rc_;                                                \
    } while (0)

// Launch-error check after a kernel launch.
#define XDEM_LAUNCHED() XDEM_HIP(hipGetLastError())

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------- kernel params
// The reference keeps these in a __constant__ block (mphysics_cuda_impl.cuh:6);
// here they travel by value as a kernel argument, so every launch is
// self-contained (re-entrant, graph-capturable).
struct KParams {
    uint32_t np, nsphere, ncell;
    uint32_t gx, gy, gz;
    int32_t rolling;
    int32_t pad;
    double dt, half2dt, cs, coh;
    double gxv, gyv, gzv;   // gravity
    double wox, woy, woz;   // world origin
};

KParams params_from(const xdem_device_parameters& p);
int32_t validate_params(const xdem_device_parameters& p);
const xdem_device_parameters& host_params();
bool params_set();

// --------------------------------------------------- device vector algebra
// Same expression order as the reference helpers (mphysics_cuda_impl.cuh:19-63).
struct d3 {
    double x, y, z;
};
struct alignas(16) d4 {
    double x, y, z, w;
};

__host__ __device__ inline d3 mk3(double x, double y, double z) { return d3{x, y, z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator-(d3 a) { return d3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline d3 operator*(double s, d3 a) { return d3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline d3 operator/(d3 a, double s) { return d3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline void operator+=(d3& a, d3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double length(d3 a) { return sqrt(dot(a, a)); }
__host__ __device__ inline d3 cross(d3 a, d3 b) {
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline d3 of(const xdem_double3& v) { return d3{v.x, v.y, v.z}; }

// calcGridPos — floor((p - wo) / cs), IEEE division (no reciprocal multiply),
// mphysics_cuda_impl.cuh:75-83.
__device__ inline void grid_pos(const KParams& k, double x, double y, double z, int& cx, int& cy,
                                int& cz) {
    cx = (int)floor((x - k.wox) / k.cs);
    cy = (int)floor((y - k.woy) / k.cs);
    cz = (int)floor((z - k.woz) / k.cs);
}
// calcGridHash — wrap by mask, mphysics_cuda_impl.cuh:65-72.
__device__ inline uint32_t grid_hash(const KParams& k, int cx, int cy, int cz) {
    uint32_t x = (uint32_t)cx & (k.gx - 1);
    uint32_t y = (uint32_t)cy & (k.gy - 1);
    uint32_t z = (uint32_t)cz & (k.gz - 1);
    return __umul24(__umul24(z, k.gy), k.gx) + __umul24(y, k.gx) + x;
}

// ------------------------------------------------------------- workspace
// Grow-only device scratch owned by the library (sort ping-pong buffers,
// histograms, reduction partials).
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    int32_t reserve(size_t need);
};
Workspace& workspace(int slot);

// ------------------------------------------------------------ phase timer
// HIP-event timing of the launches of one phase, on the stream they are
// launched on.  Events are recorded around each phase and read at flush(),
// so timing adds no host synchronisation inside a timed region.
struct PhaseTimer {
    static constexpr int NPHASE = 8;
    bool on = false;
    std::vector<hipEvent_t> ev;  // pairs
    std::vector<int> phase;      // phase of each pair
    size_t used = 0;
    int open = -1;
    double ms[NPHASE] = {0};
    uint32_t count[NPHASE] = {0};
    int32_t begin(int ph, hipStream_t s);
    int32_t end(hipStream_t s);
    int32_t flush();
    void reset();
    ~PhaseTimer();
};

// ------------------------------------------------------------ sort API
// Stable sort of (keys, vals) by the low `key_bits` bits of keys (in place:
// results end in keys/vals).  Uses workspace slot 0.
int32_t sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, hipStream_t s);
// cstart/cend over sorted keys (+ sorted_index = vals); cells not present keep
// whatever the arrays held (the caller clears them).
int32_t cell_ranges(const uint32_t* keys, const uint32_t* vals, uint32_t* cstart, uint32_t* cend,
                    uint32_t* sorted_index, uint32_t n, hipStream_t s);
int32_t fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
uint32_t bits_for(uint32_t ncell);

}  // namespace xdem
