// fastsurf.h — filtered evaluation of a hit's binary64 surface quantities.
//
// The reference reconstructs a hit in binary64: t, u, v from the watertight
// test (reference src/geometry.cpp:147-182), then position = a*w + b*u + c*v
// and the interpolated shading normal normalized in binary64 (surfaceAt,
// src/geometry.cpp:401-459), and finally rounds them to float.  Only the
// float results are used by a path whose origin is not split (no precision
// replay), so instead of emulating every binary64 step this filter
// evaluates the same expressions in double-float with a forward error bound
// that covers both its own rounding and the reference's binary64 rounding
// (per operation 2^-42 + 2^-52, both far above the true bounds).  When the
// interval around each result contains no float rounding boundary, the float
// is the one the reference produces; otherwise (or for zeros, extreme
// magnitudes, near-degenerate determinants) the caller uses the exact
// emulation.  The bound follows the same derivation as the FP32 triangle
// filter in geom.h with u replaced by the per-operation bound.
#pragma once
#include "df.h"
#include "geom.h"

namespace cr {

constexpr float kFsEta = 2.0f * (df::kDfRel + 2.220446049250313e-16f);  // chain-vs-chain per-op bound

// RN32 of every real within `err` of x is x.hi: true when no float midpoint
// lies within err of x (zeros and extreme magnitudes are left to the exact path).
CR_HDI bool round_certain(df::DF x, float err, float& out) {
    const float f = x.hi;
    const float a = fabsf(f);
    if (!(a >= 0x1p-90f) || !(a <= 0x1p90f) || !(err >= 0.0f) || !finite1(err)) return false;
    const float up = next_up(a) - a, dn = a - next_down(a);
    const float r = f > 0 ? x.lo : -x.lo;  // offset of x from f toward larger magnitude
    if (!(r + err < 0.499f * up) || !(r - err > -0.499f * dn)) return false;
    out = f;
    return true;
}

// position, interpolated shading normal (before the flip toward Ng) and
// float(t) of the closest hit on triangle (va, vb, vc) with vertex normals
// (n0, n1, n2), for a query without split origin.  Returns false when any of
// them is not certain; outputs are then unspecified.
CR_HD inline bool fast_hit_surface(const RayQuery& q, V3 va, V3 vb, V3 vc, V3 n0, V3 n1, V3 n2, V3& position,
                                   V3& normal, float& tf) {
    using df::DF;
    const float eta = kFsEta;
    if (q.precise || !q.valid) return false;
    const int kx = q.kx, ky = q.ky, kz = q.kz;
    const float dz = at(q.d, kz);
    const DF sx = df::div(df::from(at(q.d, kx)), df::from(dz));
    const DF sy = df::div(df::from(at(q.d, ky)), df::from(dz));
    const DF sz = df::div(df::from(1.0f), df::from(dz));
    // A = V - O (exact as a float pair), sheared coordinates
    const V3 verts[3] = {va, vb, vc};
    DF X[3], Y[3], Z[3];
    float E[3];
    float maxA = 0;
    for (int k = 0; k < 3; ++k) {
        const DF Ax = df::two_sum(at(verts[k], kx), -at(q.o, kx));
        const DF Ay = df::two_sum(at(verts[k], ky), -at(q.o, ky));
        const DF Az = df::two_sum(at(verts[k], kz), -at(q.o, kz));
        X[k] = df::sub(Ax, df::mul(sx, Az));
        Y[k] = df::sub(Ay, df::mul(sy, Az));
        Z[k] = Az;
        const float m = df::mag(Ax) + df::mag(Ay) + 3.0f * df::mag(Az);
        maxA = fmaxf(maxA, m);
        E[k] = 1.1f * eta * (m + df::mag(X[k]) + df::mag(Y[k]));
    }
    if (!(maxA <= 0x1p40f) || !(maxA >= 0x1p-40f)) return false;
    // edge functions w0 = cx*by - cy*bx, w1 = ax*cy - ay*cx, w2 = bx*ay - by*ax
    const int i0[3] = {2, 0, 1}, i1[3] = {1, 2, 0};  // (c,b), (a,c), (b,a)
    DF W[3];
    float EW[3];
    for (int k = 0; k < 3; ++k) {
        const int p = i0[k], r = i1[k];
        const DF t1 = df::mul(X[p], Y[r]), t2 = df::mul(Y[p], X[r]);
        W[k] = df::sub(t1, t2);
        const float mp = df::mag(X[p]) + df::mag(Y[p]), mr = df::mag(X[r]) + df::mag(Y[r]);
        EW[k] = 1.1f * (mp * E[r] + mr * E[p] + 2.0f * E[p] * E[r] +
                        eta * (df::mag(t1) + df::mag(t2) + df::mag(W[k])));
    }
    const DF det = df::add(df::add(W[0], W[1]), W[2]);
    const float adet = df::mag(det);
    const float Edet = 1.1f * ((EW[0] + EW[1]) + EW[2] + 2.0f * eta * (df::mag(W[0]) + df::mag(W[1]) + df::mag(W[2])));
    if (!(adet > 4.0f * Edet) || !(adet >= 0x1p-100f)) return false;
    const float denom = adet - Edet;
    const DF u = df::div(W[1], det), v = df::div(W[2], det);
    const float Eu = 1.1f * ((EW[1] + df::mag(u) * Edet) / denom + eta * df::mag(u));
    const float Ev = 1.1f * ((EW[2] + df::mag(v) * Edet) / denom + eta * df::mag(v));
    // t = ((w0*(sz*az) + w1*(sz*bz)) + w2*(sz*cz)) / det
    DF num = df::from(0.0f);
    float Enum = 0, msum = 0;
    for (int k = 0; k < 3; ++k) {
        const DF zk = df::mul(sz, Z[k]);
        const float Ez = 3.3f * eta * df::mag(zk);
        const DF term = df::mul(W[k], zk);
        num = k == 0 ? term : df::add(num, term);
        Enum += df::mag(W[k]) * Ez + df::mag(zk) * EW[k] + EW[k] * Ez;
        msum += df::mag(term);
    }
    Enum = 1.1f * (Enum + 3.0f * eta * msum);
    const DF t = df::div(num, det);
    const float Et = 1.1f * ((Enum + df::mag(t) * Edet) / denom + eta * df::mag(t));
    if (!round_certain(t, Et, tf)) return false;
    // w = (1 - u) - v, then position and normal interpolation
    const DF w = df::sub(df::sub(df::from(1.0f), u), v);
    const float Ew = 1.1f * (Eu + Ev + 2.0f * eta * (1.0f + df::mag(u) + df::mag(v)));
    float pos[3], nrm[3];
    DF N[3];
    float EN[3];
    // Interpolating three +0.0 values gives +0.0 in binary64 for an accepted hit:
    // u and v are >= 0 or -0, and when both are -0, w = 1, so at least one
    // product is +0 and the sum is +0.
    bool nzero[3];
    for (int i = 0; i < 3; ++i) {
        const float a_i = at(va, i), b_i = at(vb, i), c_i = at(vc, i);
        if ((sf::fbits(a_i) | sf::fbits(b_i) | sf::fbits(c_i)) == 0) {
            pos[i] = 0.0f;
        } else {
            const DF pa = df::mulf(w, a_i), pb = df::mulf(u, b_i), pc = df::mulf(v, c_i);
            const DF P = df::add(df::add(pa, pb), pc);
            const float EP = 1.1f * (fabsf(a_i) * Ew + fabsf(b_i) * Eu + fabsf(c_i) * Ev +
                                     3.0f * eta * (df::mag(pa) + df::mag(pb) + df::mag(pc)));
            if (!round_certain(P, EP, pos[i])) return false;
        }
        const float m_i = at(n0, i), o_i = at(n1, i), r_i = at(n2, i);
        nzero[i] = (sf::fbits(m_i) | sf::fbits(o_i) | sf::fbits(r_i)) == 0;
        const DF na = df::mulf(w, m_i), nb = df::mulf(u, o_i), nc = df::mulf(v, r_i);
        N[i] = nzero[i] ? df::from(0.0f) : df::add(df::add(na, nb), nc);
        EN[i] = nzero[i] ? 0.0f
                         : 1.1f * (fabsf(m_i) * Ew + fabsf(o_i) * Eu + fabsf(r_i) * Ev +
                                   3.0f * eta * (df::mag(na) + df::mag(nb) + df::mag(nc)));
    }
    // unit(N): sq = (Nx*Nx + Ny*Ny) + Nz*Nz, inv = 1/sqrt(sq), n_i = N_i*inv
    const DF s2 = df::add(df::add(df::mul(N[0], N[0]), df::mul(N[1], N[1])), df::mul(N[2], N[2]));
    const float S = df::mag(s2);
    float ES = 0;
    for (int i = 0; i < 3; ++i) ES += 2.0f * df::mag(N[i]) * EN[i] + EN[i] * EN[i];
    ES = 1.1f * (ES + 3.0f * eta * S);
    if (!(S >= 0x1p-60f) || !(S <= 0x1p60f) || !(ES < 0.25f * S)) return false;
    const DF root = df::sqrt_(s2);
    const DF inv = df::div(df::from(1.0f), root);
    const float rho = 1.2f * (0.5f * ES / (S - ES) + 2.0f * eta);  // relative error of inv
    for (int i = 0; i < 3; ++i) {
        if (nzero[i]) {  // +0 * inv = +0
            nrm[i] = 0.0f;
            continue;
        }
        const DF c = df::mul(N[i], inv);
        const float Ec = 1.1f * (df::mag(inv) * EN[i] + df::mag(c) * rho + eta * df::mag(c));
        if (!round_certain(c, Ec, nrm[i])) return false;
    }
    position = V3{pos[0], pos[1], pos[2]};
    normal = V3{nrm[0], nrm[1], nrm[2]};
    return true;
}

}  // namespace cr
