// FP32 vector / quaternion helpers for the gfx950 kernels. Every expression
// is evaluated exactly as written: the library is compiled with
// -ffp-contract=off and correctly rounded division and square root, so the
// kernels round identically to a scalar IEEE-754 host implementation of the
// same expressions (which is how parity is tested).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrd {

struct V3 {
    float x, y, z;
};
struct Q4 {
    float x, y, z, w;
};

#define MRD_INL __device__ __forceinline__

MRD_INL V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MRD_INL V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MRD_INL V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
MRD_INL V3 mul(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
MRD_INL float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MRD_INL V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
MRD_INL float fmaxd(float a, float b) { return a > b ? a : b; }
MRD_INL float fmind(float a, float b) { return a < b ? a : b; }
MRD_INL float clampd(float x, float lo, float hi) { return fmind(fmaxd(x, lo), hi); }
MRD_INL float sqrt_rn(float x) { return __builtin_sqrtf(x); }
MRD_INL float fabs_(float x) { return __builtin_fabsf(x); }
MRD_INL bool isfin(float x) { return __builtin_isfinite(x); }
MRD_INL float length(V3 a) { return sqrt_rn(dot(a, a)); }
MRD_INL V3 normalize(V3 a) {
    const float inv = 1.0f / sqrt_rn(dot(a, a));
    return mul(inv, a);
}
MRD_INL V3 vabs(V3 a) { return V3{fabs_(a.x), fabs_(a.y), fabs_(a.z)}; }
MRD_INL V3 vmin(V3 a, V3 b) { return V3{fmind(a.x, b.x), fmind(a.y, b.y), fmind(a.z, b.z)}; }
MRD_INL V3 vmax(V3 a, V3 b) { return V3{fmaxd(a.x, b.x), fmaxd(a.y, b.y), fmaxd(a.z, b.z)}; }
MRD_INL float maxabs3(V3 a) { return fmaxd(fmaxd(fabs_(a.x), fabs_(a.y)), fabs_(a.z)); }
MRD_INL bool finite3(V3 a) { return isfin(a.x) && isfin(a.y) && isfin(a.z); }
MRD_INL bool finite4(Q4 a) { return isfin(a.x) && isfin(a.y) && isfin(a.z) && isfin(a.w); }
MRD_INL float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
MRD_INL void setcomp(V3& a, int i, float v) {
    if (i == 0) a.x = v;
    else if (i == 1) a.y = v;
    else a.z = v;
}

MRD_INL Q4 qmul(Q4 l, Q4 r) {
    return Q4{l.w * r.x + l.x * r.w + l.y * r.z - l.z * r.y,
              l.w * r.y - l.x * r.z + l.y * r.w + l.z * r.x,
              l.w * r.z + l.x * r.y - l.y * r.x + l.z * r.w,
              l.w * r.w - (l.x * r.x + l.y * r.y + l.z * r.z)};
}
MRD_INL Q4 qconj(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
MRD_INL V3 qrot(Q4 q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = mul(2.0f, cross(u, v));
    return add(add(v, mul(q.w, t)), cross(u, t));
}
MRD_INL V3 qrot_inv(Q4 q, V3 v) { return qrot(qconj(q), v); }
MRD_INL float qdot(Q4 a, Q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
MRD_INL Q4 qscale(float s, Q4 q) { return Q4{s * q.x, s * q.y, s * q.z, s * q.w}; }

// Component-domain quaternion admission: finite, largest |component| in
// [0.25, 1.001], then normalized.
MRD_INL bool checked_quat(Q4 in, Q4& out) {
    if (!finite4(in)) return false;
    const float m = fmaxd(fmaxd(fabs_(in.x), fabs_(in.y)), fmaxd(fabs_(in.z), fabs_(in.w)));
    if (m < 0.25f || m > 1.001f) return false;
    const float inv = 1.0f / sqrt_rn(qdot(in, in));
    out = qscale(inv, in);
    return finite4(out);
}

MRD_INL bool normalized_quat(Q4 in, Q4& out) {
    if (!finite4(in)) return false;
    const float n2 = qdot(in, in);
    if (!(n2 > 1.0e-12f) || !isfin(n2)) return false;
    const float inv = 1.0f / sqrt_rn(n2);
    out = qscale(inv, in);
    return finite4(out);
}

// Deterministic sine/cosine: Cody-Waite reduction by pi/4 with the Cephes
// single-precision minimax polynomials (S. L. Moshier). The device libm is
// not used so that results do not depend on its implementation.
MRD_INL void det_sincos(float x, float& s_out, float& c_out) {
    float ax = x;
    float sign_s = 1.0f;
    if (ax < 0.0f) {
        ax = -ax;
        sign_s = -1.0f;
    }
    if (!(ax <= 1.0e5f)) {
        s_out = __builtin_nanf("");
        c_out = __builtin_nanf("");
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
    const float r = ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    const float z = r * r;
    const float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z -
        0.5f * z + 1.0f;
    const float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
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

MRD_INL Q4 axis_angle(V3 axis_n, float angle) {
    float s, c;
    det_sincos(0.5f * angle, s, c);
    return Q4{axis_n.x * s, axis_n.y * s, axis_n.z * s, c};
}

// exp map of a rotation vector as a quaternion increment.
MRD_INL Q4 rotation_increment(V3 rv) {
    const float angle = length(rv);
    const float half = 0.5f * angle;
    float s, c;
    det_sincos(half, s, c);
    const float scale = angle > 1.0e-6f ? s / angle : 0.5f - angle * angle / 48.0f;
    return Q4{rv.x * scale, rv.y * scale, rv.z * scale, c};
}

} // namespace mrd
