// xdem_io.hip — the reference's result and resume files (host code):
// part%04d.bin (dem_simulation::saveResult(_f), dem_simulation.cpp:491-547)
// and <name>_final.bfr (xDynamicsSolver::saveFinalResult,
// xDynamicsSolver.cpp:111-130, + dem_simulation::saveFinalResult(_f),
// dem_simulation.cpp:551-583; read by startingModel::setDEMData,
// startingModel.cpp:32-67), and the checkpoint container check.
#include <cmath>
#include <vector>

#include "../../include/xdem_io.h"
#include "xdem_io.hpp"

using namespace xdem;

namespace {

template <class T>
int32_t write_part(const char* path, double ct, uint32_t np, const T* pos4) {
    XDEM_REQUIRE(np == 0 || pos4, "null positions");
    io::Writer w;
    w.put(ct);
    w.put(np);
    w.arr(pos4, 4 * (size_t)np);
    return io::write_file(path, w.buf);
}

template <class T>
int32_t write_bfr(const char* path, double ct, uint32_t np, const T* pos4, const T* vel3,
                  const T* avel3, int32_t precision) {
    XDEM_REQUIRE(np == 0 || (pos4 && vel3 && avel3), "null array");
    io::Writer w;
    w.put(ct);
    w.put(precision);
    w.put((int32_t)1);  // flag: DEM data follows
    w.arr(pos4, 4 * (size_t)np);
    w.arr(vel3, 3 * (size_t)np);
    w.arr(avel3, 3 * (size_t)np);
    return io::write_file(path, w.buf);
}

template <class T>
bool read_widen(io::Reader& r, double* dst, size_t n) {
    if (!dst) {
        r.off += sizeof(T) * n;
        return r.off <= r.n;
    }
    for (size_t i = 0; i < n; ++i) {
        T v;
        if (!r.get(v)) return false;
        dst[i] = (double)v;
    }
    return true;
}

}  // namespace

extern "C" {

int32_t xdem_write_part_f64(const char* path, double ct, uint32_t np, const double* pos4) {
    return write_part(path, ct, np, pos4);
}

int32_t xdem_write_part_f32(const char* path, double ct, uint32_t np, const float* pos4) {
    return write_part(path, ct, np, pos4);
}

int32_t xdem_read_part(const char* path, double* ct, uint32_t* np, int32_t* precision,
                       double* pos4, uint32_t cap) {
    std::vector<uint8_t> f;
    XDEM_TRY(io::read_file(path, f));
    io::Reader r{f.data(), f.size(), 0};
    double t;
    uint32_t n;
    XDEM_REQUIRE(r.get(t) && r.get(n), "%s: too short for a part file", path);
    const size_t body = f.size() - 12;
    int32_t prec;
    if (body == 32ull * n)
        prec = 2;
    else if (body == 16ull * n)
        prec = 1;
    else {
        xdem::set_error("%s: %zu bytes after the header do not hold %u positions", path, body, n);
        return XDEM_ERR_INVALID;
    }
    if (ct) *ct = t;
    if (np) *np = n;
    if (precision) *precision = prec;
    if (pos4) {
        XDEM_REQUIRE(cap >= n, "%s: %u positions, room for %u", path, n, cap);
        bool ok = prec == 2 ? read_widen<double>(r, pos4, 4 * (size_t)n)
                            : read_widen<float>(r, pos4, 4 * (size_t)n);
        XDEM_REQUIRE(ok, "%s: truncated", path);
    }
    return XDEM_OK;
}

int32_t xdem_write_bfr_f64(const char* path, double ct, uint32_t np, const double* pos4,
                           const double* vel3, const double* avel3) {
    return write_bfr(path, ct, np, pos4, vel3, avel3, 2);
}

int32_t xdem_write_bfr_f32(const char* path, double ct, uint32_t np, const float* pos4,
                           const float* vel3, const float* avel3) {
    return write_bfr(path, ct, np, pos4, vel3, avel3, 1);
}

int32_t xdem_read_bfr(const char* path, uint32_t np, double* ct, int32_t* precision, double* pos4,
                      double* vel3, double* avel3) {
    std::vector<uint8_t> f;
    XDEM_TRY(io::read_file(path, f));
    io::Reader r{f.data(), f.size(), 0};
    double t;
    int32_t prec, flag;
    XDEM_REQUIRE(r.get(t) && r.get(prec) && r.get(flag), "%s: too short for a .bfr file", path);
    XDEM_REQUIRE(prec == 1 || prec == 2, "%s: precision %d (expected 1 or 2)", path, prec);
    XDEM_REQUIRE(flag == 1, "%s: no DEM data (flag %d)", path, flag);
    const size_t need = (prec == 2 ? 8 : 4) * 10ull * np;
    XDEM_REQUIRE(f.size() - 16 >= need, "%s: holds fewer than %u particles", path, np);
    bool ok = prec == 2 ? (read_widen<double>(r, pos4, 4ull * np) &&
                           read_widen<double>(r, vel3, 3ull * np) &&
                           read_widen<double>(r, avel3, 3ull * np))
                        : (read_widen<float>(r, pos4, 4ull * np) &&
                           read_widen<float>(r, vel3, 3ull * np) &&
                           read_widen<float>(r, avel3, 3ull * np));
    XDEM_REQUIRE(ok, "%s: truncated", path);
    if (ct) *ct = t;
    if (precision) *precision = prec;
    return XDEM_OK;
}

int32_t xdem_checkpoint_verify(const char* path, uint32_t* kind) {
    std::vector<uint8_t> f;
    io::Reader r;
    return io::read_checkpoint(path, 0, f, r, kind);
}

}  // extern "C"
