// TEST INFRASTRUCTURE — the CPU oracle of the temporalCone microstep.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load or run anything under oracle/; the product never does.
//
// FP32 vector/quaternion helpers with a fixed evaluation order. The oracle is
// compiled with -ffp-contract=off so every expression below rounds exactly
// as written (no fused multiply-add, no reassociation).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

struct V3 {
    float x, y, z;
};
struct Q4 {
    float x, y, z, w;
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 mul(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float fmaxd(float a, float b) { return a > b ? a : b; }
inline float fmind(float a, float b) { return a < b ? a : b; }
inline float clampd(float x, float lo, float hi) { return fmind(fmaxd(x, lo), hi); }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    const float inv = 1.0f / std::sqrt(dot(a, a));
    return mul(inv, a);
}
inline V3 vabs(V3 a) { return V3{std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline V3 vmin(V3 a, V3 b) { return V3{fmind(a.x, b.x), fmind(a.y, b.y), fmind(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return V3{fmaxd(a.x, b.x), fmaxd(a.y, b.y), fmaxd(a.z, b.z)}; }
inline float maxabs3(V3 a) {
    return fmaxd(fmaxd(std::fabs(a.x), std::fabs(a.y)), std::fabs(a.z));
}
inline bool finite3(V3 a) {
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
}
inline bool finite4(Q4 a) {
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z) &&
        std::isfinite(a.w);
}
inline float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline void setcomp(V3& a, int i, float v) {
    if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}

inline Q4 qmul(Q4 l, Q4 r) {
    return Q4{
        l.w * r.x + l.x * r.w + l.y * r.z - l.z * r.y,
        l.w * r.y - l.x * r.z + l.y * r.w + l.z * r.x,
        l.w * r.z + l.x * r.y - l.y * r.x + l.z * r.w,
        l.w * r.w - (l.x * r.x + l.y * r.y + l.z * r.z)};
}
inline Q4 qconj(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
inline V3 qrot(Q4 q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = mul(2.0f, cross(u, v));
    return add(add(v, mul(q.w, t)), cross(u, t));
}
inline V3 qrot_inv(Q4 q, V3 v) { return qrot(qconj(q), v); }
inline float qdot(Q4 a, Q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
inline Q4 qscale(float s, Q4 q) { return Q4{s * q.x, s * q.y, s * q.z, s * q.w}; }

// Normalizes a finite quaternion whose largest component lies in
// [0.25, 1.001] (the component-domain admission of engine_types.h:96-97).
inline bool checked_quat(Q4 in, Q4& out) {
    if (!finite4(in)) return false;
    const float m = fmaxd(fmaxd(std::fabs(in.x), std::fabs(in.y)),
                          fmaxd(std::fabs(in.z), std::fabs(in.w)));
    if (m < 0.25f || m > 1.001f) return false;
    const float inv = 1.0f / std::sqrt(qdot(in, in));
    out = qscale(inv, in);
    return finite4(out);
}

// Normalizes any finite quaternion with squared norm above 1e-12.
inline bool normalized_quat(Q4 in, Q4& out) {
    if (!finite4(in)) return false;
    const float n2 = qdot(in, in);
    if (!(n2 > 1.0e-12f) || !std::isfinite(n2)) return false;
    const float inv = 1.0f / std::sqrt(n2);
    out = qscale(inv, in);
    return finite4(out);
}

// Deterministic single-precision sine/cosine: Cody-Waite reduction by pi/4
// and the minimax polynomials of the Cephes library (S. L. Moshier,
// sinf.c / cosf.c). Written out so that CPU and GPU round identically.
inline void det_sincos(float x, float& s_out, float& c_out) {
    float ax = x;
    float sign_s = 1.0f;
    if (ax < 0.0f) {
        ax = -ax;
        sign_s = -1.0f;
    }
    if (!(ax <= 1.0e5f)) {
        s_out = NAN;
        c_out = NAN;
        return;
    }
    uint32_t j = (uint32_t)(1.27323954473516f * ax);
    float y = (float)j;
    if ((j & 1u) != 0u) {
        j += 1u;
        y += 1.0f;
    }
    j &= 7u;
    float sign_c = 1.0f;
    if (j > 3u) {
        sign_s = -sign_s;
        sign_c = -sign_c;
        j -= 4u;
    }
    if (j > 1u) sign_c = -sign_c;
    const float r = ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) -
        y * 3.77489497744594108e-8f;
    const float z = r * r;
    const float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z +
                      4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
    const float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z -
                      1.6666654611e-1f) * z * r + r;
    float s, c;
    if (j == 1u || j == 2u) {
        s = pc;
        c = ps;
    } else {
        s = ps;
        c = pc;
    }
    s_out = sign_s * s;
    c_out = sign_c * c;
}

inline Q4 axis_angle(V3 axis_n, float angle) {
    float s, c;
    det_sincos(0.5f * angle, s, c);
    return Q4{axis_n.x * s, axis_n.y * s, axis_n.z * s, c};
}

// Quaternion increment exp(0.5 * h * w) used by the SO(3)-exponential update
// (ArticulatedABA.metal:1466-1490, MetalWorldContact.metal:8936-8950).
inline Q4 rotation_increment(V3 rotation_vector) {
    const float angle = length(rotation_vector);
    const float half = 0.5f * angle;
    float s, c;
    det_sincos(half, s, c);
    const float scale = angle > 1.0e-6f ? s / angle : 0.5f - angle * angle / 48.0f;
    return Q4{rotation_vector.x * scale, rotation_vector.y * scale,
              rotation_vector.z * scale, c};
}

inline uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

} // namespace oracle
