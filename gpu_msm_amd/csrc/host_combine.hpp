// Host side of the MSM finish: device field words -> host field elements,
// the Horner window combine (most significant window first, as reference
// src/curve.cu:2060-2179 combines windows) and the ABI record writers
// (reference include/curve.h:13-43 layouts). Shared by every host entry point.
#pragma once
#include <string.h>

#include "engine.hpp"
#include "host_field.hpp"

namespace gmc {

// 16 x 28-bit device limbs (value < 2^448) -> canonical 7 x 64 host element.
inline void dev_elem_to_host(gmh::Fp& r, const uint32_t* v) {
    gmh::Fp t;
    gmh::from_limbs28(t, v);
    gmh::reduce(r, t);
}
inline void dev_elem_to_host(gmh::Fp2& r, const uint32_t* v) {
    dev_elem_to_host(r.c0, v);
    dev_elem_to_host(r.c1, v + 16);
}

// 32-bit words of one device element of HF's field.
template <class HF>
constexpr int words_of() {
    return sizeof(HF) == sizeof(gmh::Fp) ? 16 : 32;
}

// Horner over nw XYZZ window sums (raw device words), most significant first,
// in Jacobian coordinates (the doublings dominate); false if the sum is the
// point at infinity.
template <class HF>
bool host_combine(HF& x, HF& y, const uint32_t* raw, int nw, int c) {
    constexpr int W = words_of<HF>();
    auto load = [&](gmh::Jac<HF>& j, int w) {
        const uint32_t* b = raw + (size_t)w * 4 * W;
        gmh::Xyzz<HF> p;
        dev_elem_to_host(p.X, b);
        dev_elem_to_host(p.Y, b + W);
        dev_elem_to_host(p.ZZ, b + 2 * W);
        dev_elem_to_host(p.ZZZ, b + 3 * W);
        gmh::from_xyzz(j, p);
    };
    gmh::Jac<HF> r, v;
    load(r, nw - 1);
    for (int w = nw - 2; w >= 0; --w) {
        for (int k = 0; k < c; ++k) gmh::jdbl(r, r);
        load(v, w);
        gmh::jadd(r, r, v);
    }
    return gmh::to_affine(x, y, r);
}

inline void put_elem(uint64_t* w, const gmh::Fp& a) { memcpy(w, a.l, 56); }
inline void put_elem(uint64_t* w, const gmh::Fp2& a) {
    memcpy(w, a.c0.l, 56);
    memcpy(w + 7, a.c1.l, 56);
}
inline void get_elem(gmh::Fp& a, const uint64_t* w) { memcpy(a.l, w, 56); }
inline void get_elem(gmh::Fp2& a, const uint64_t* w) {
    memcpy(a.c0.l, w, 56);
    memcpy(a.c1.l, w + 7, 56);
}

// ABI affine record: (x, y, infinity = false) or (0, 0, infinity = true).
template <class HF>
void write_affine_abi(void* out, bool finite, const HF& x, const HF& y) {
    constexpr int E = sizeof(HF) / 8;
    uint64_t buf[2 * E + 1];
    memset(buf, 0, sizeof buf);
    if (finite) {
        put_elem(buf, x);
        put_elem(buf + E, y);
    } else {
        buf[2 * E] = 1;
    }
    memcpy(out, buf, sizeof buf);
}

// ABI projective record: (x, y, R mod p) or all-zero (Z = 0 is infinity).
template <class HF>
void write_proj_abi(void* out, bool finite, const HF& x, const HF& y) {
    constexpr int E = sizeof(HF) / 8;
    uint64_t buf[3 * E];
    memset(buf, 0, sizeof buf);
    if (finite) {
        put_elem(buf, x);
        put_elem(buf + E, y);
        memcpy(buf + 2 * E, gmh::ONE.l, 56);
    }
    memcpy(out, buf, sizeof buf);
}

template <class O>
struct HostOf;
template <>
struct HostOf<gm::FpOps> {
    using HF = gmh::Fp;
    static constexpr int CURVE = gm::CURVE_G1;
};
template <>
struct HostOf<gm::Fp2Ops> {
    using HF = gmh::Fp2;
    static constexpr int CURVE = gm::CURVE_G2;
};

template <class HF>
struct AffResult {
    bool finite;
    HF x, y;
};

}  // namespace gmc
