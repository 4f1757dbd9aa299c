// xdem_geom.hpp — FP64 wall geometry shared by the plain and fused paths.
#pragma once
#include "xdem_common.hpp"

namespace xdem {

// particle_plane_contact_detection — mphysics_cuda_impl.cuh:487-570 (branch
// order kept, including the unreachable seventh branch of :549).
// EXACT_POW: the reference's pow(x, 2.0) (plain path, for expression-level
// parity); otherwise x * x (fused path: identical up to the last bit of an
// edge/corner branch test).
template <bool EXACT_POW = true>
__device__ inline double plane_detect(const xdem_device_plane_info& pe, d3 xp, d3 wp, d3& u, double r) {
    // Every branch below needs |wp.z| < r (the face test) or a sum of
    // non-negative squares that includes wp.z^2 to stay below r^2; when both
    // fail no branch can be taken, so leave before the pow calls.  Exact: the
    // rounded sums are >= their last term.
    if (!(fabs(wp.z) < r) && !(wp.z * wp.z < r * r)) return -1.0;
    const double ex_l1 = wp.x - pe.l1, ey_l2 = wp.y - pe.l2;
    double a_l1 = EXACT_POW ? pow(ex_l1, 2.0) : ex_l1 * ex_l1;
    double b_l2 = EXACT_POW ? pow(ey_l2, 2.0) : ey_l2 * ey_l2;
    double sqa = wp.x * wp.x;
    double sqb = wp.y * wp.y;
    double sqc = wp.z * wp.z;
    double sqr = r * r;
    d3 xw = of(pe.xw), w2 = of(pe.w2), w3 = of(pe.w3), w4 = of(pe.w4), uw = of(pe.uw);
    if (fabs(wp.z) < r && (wp.x > 0 && wp.x < pe.l1) && (wp.y > 0 && wp.y < pe.l2)) {
        d3 dp = xp - xw;
        d3 uu = uw / length(uw);
        int pp = -(dot(dp, uw) < 0 ? -1 : 1);
        u = (double)pp * uu;
        return r - fabs(dot(dp, u));
    }
    if (wp.x < 0 && wp.y < 0 && (sqa + sqb + sqc) < sqr) {
        d3 Xsw = xp - xw;
        double h = length(Xsw);
        u = Xsw / h;
        return r - h;
    } else if (wp.x > pe.l1 && wp.y < 0 && (a_l1 + sqb + sqc) < sqr) {
        d3 Xsw = xp - w2;
        double h = length(Xsw);
        u = Xsw / h;
        return r - h;
    } else if (wp.x > pe.l1 && wp.y > pe.l2 && (a_l1 + b_l2 + sqc) < sqr) {
        d3 Xsw = xp - w3;
        double h = length(Xsw);
        u = Xsw / h;
        return r - h;
    } else if (wp.x < 0 && wp.y > pe.l2 && (sqa + b_l2 + sqc) < sqr) {
        d3 Xsw = xp - w4;
        double h = length(Xsw);
        u = Xsw / h;
        return r - h;
    }
    if ((wp.x > 0 && wp.x < pe.l1) && wp.y < 0 && (sqb + sqc) < sqr) {
        d3 Xsw = xp - xw;
        d3 wj_wi = w2 - xw;
        d3 us = wj_wi / length(wj_wi);
        d3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        return r - h;
    } else if ((wp.x > 0 && wp.x < pe.l1) && wp.y > pe.l2 && (b_l2 + sqc) < sqr) {
        d3 Xsw = xp - w4;
        d3 wj_wi = w3 - w4;
        d3 us = wj_wi / length(wj_wi);
        d3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        return r - h;
    } else if ((wp.x > 0 && wp.y < pe.l2) && wp.x < 0 && (sqr + sqc) < sqr) {
        d3 Xsw = xp - xw;
        d3 wj_wi = w4 - xw;
        d3 us = wj_wi / length(wj_wi);
        d3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        return r - h;
    } else if ((wp.x > 0 && wp.y < pe.l2) && wp.x > pe.l1 && (a_l1 + sqc) < sqr) {
        d3 Xsw = xp - w2;
        d3 wj_wi = w3 - w2;
        d3 us = wj_wi / length(wj_wi);
        d3 h_star = Xsw - (dot(Xsw, us)) * us;
        double h = length(h_star);
        u = -h_star / h;
        return r - h;
    }
    return -1.0;
}

// Closest point of triangle (a, b, c) to p — Ericson, Real-Time Collision
// Detection, 5.1.5 (the routine of closestPtPointTriangle,
// mphysics_cuda_impl.cuh:837-899), on plain vertices.
__host__ __device__ inline d3 closest_pt_tri(d3 p, d3 a, d3 b, d3 c) {
    d3 ab = b - a, ac = c - a, ap = p - a;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    d3 bp = p - b;
    double d3_ = dot(ab, bp), d4_ = dot(ac, bp);
    if (d3_ >= 0.0 && d4_ <= d3_) return b;
    double vc = d1 * d4_ - d3_ * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) return a + (d1 / (d1 - d3_)) * ab;
    d3 cp = p - c;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + (d2 / (d2 - d6)) * ac;
    double va = d3_ * d6 - d5 * d4_;
    if (va <= 0.0 && (d4_ - d3_) >= 0.0 && (d5 - d6) >= 0.0)
        return b + ((d4_ - d3_) / ((d4_ - d3_) + (d5 - d6))) * (c - b);
    double denom = 1.0 / (va + vb + vc);
    return a + (vb * denom) * ab + (vc * denom) * ac;
}

}  // namespace xdem
