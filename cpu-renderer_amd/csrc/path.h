// path.h — one camera sample's transport as pure per-path steps (host+device).
//
// begin_sample() is the reference's per-sample setup (reference
// src/render.cpp:396-401) and bounce() one iteration of trace()'s loop
// (reference src/render.cpp:240-344) given the closest hit of the current ray.
// The wavefront kernels load a PathState from HBM, call these, and store it
// back; the same functions compiled for the CPU drive the emulation used by
// the CPU test-suite, so the device logic is checked without a GPU.
#pragma once
#include "geom.h"
#include "shade.h"

namespace cr {

struct PathState {
    V3 o, d, lo;      // current ray: origin, direction, origin_low
    float tmin, tmax;
    V3 T, L;          // throughput, radiance
    float eta_scale, last_pdf;
    V3 last_pos, sigma;
    int32_t medium;
    uint32_t flags;  // depth | kLastDelta | kPrecise
    uint32_t prev;   // excluded triangle id (precise replay)
    uint64_t rng;
};

// Shadow ray emitted by NEE; `c` is added to L when the segment is unoccluded.
struct ShadowOut {
    bool emit;
    V3 o, d, lo, c;
    float tmax;
    uint32_t ignore;
};

struct ErrorOut {
    uint32_t code;  // 0 = none
    uint32_t depth, triangle, material;
    sf::f64 t;
    V3 point, origin, direction;
};

enum Action : uint32_t { kContinue = 0, kRetire = 1, kReplay = 2 };

CR_HDI void set_error(ErrorOut& e, uint32_t code, uint32_t depth, uint32_t tri, uint32_t mat, sf::f64 t, V3 point,
                      V3 origin, V3 dir) {
    e.code = code;
    e.depth = depth;
    e.triangle = tri;
    e.material = mat;
    e.t = t;
    e.point = point;
    e.origin = origin;
    e.direction = dir;
}

// Camera sample setup; returns false (with st zeroed) if the camera cannot form the ray.
CR_HDI bool begin_sample(const DevJob& J, uint32_t pixel, uint32_t sample, bool precise, PathState& st) {
    const int x = int(pixel % uint32_t(J.width)), y = int(pixel / uint32_t(J.width));
    Rng rng{sample_seed(J.seed, pixel, sample)};
    const float dx = rng.uniform(), dy = rng.uniform();
    const bool ok = camera_ray(J.cam, float(x) + dx, float(y) + dy, st.o, st.d, st.tmin, st.tmax);
    if (!ok) {
        st.o = st.d = V3{0, 0, 0};
        st.tmin = st.tmax = 0;
    }
    st.lo = V3{0, 0, 0};
    st.T = v3(1);
    st.L = v3(0);
    st.eta_scale = 1;
    st.last_pdf = 0;
    st.last_pos = st.sigma = V3{0, 0, 0};
    st.medium = -1;
    st.flags = kLastDelta | (precise ? kPrecise : 0u);
    st.prev = kNoTriangle;
    st.rng = rng.state;
    return ok;
}

// One bounce.  `prim` is the closest-hit primitive of st's ray (kNoTriangle
// on a miss).  Returns kContinue (st holds the next ray), kRetire (st.L is
// final once `sh`'s contribution, if any, is applied) or kReplay (restart the
// sample with precise origins).  Hard failures set `err` and retire.
CR_HDI uint32_t bounce(const DevJob& J, PathState& st, uint32_t prim, ShadowOut& sh, ErrorOut& err) {
    sh.emit = false;
    err.code = 0;
    const DevScene& sc = J.scene;
    const DevEnv& env = J.env;
    const uint32_t depth = st.flags & kDepthMask;
    const bool last_delta = (st.flags & kLastDelta) != 0, precise = (st.flags & kPrecise) != 0;
    const bool nee = J.nee != 0;
    V3 T = st.T, L = st.L;
    const V3 ray_o = st.o, ray_d = st.d;

    if (prim == kNoTriangle) {
        if (st.medium >= 0) T = T * beer_lambert(st.sigma, kInf);
        const float w = last_delta || !nee ? 1 : power_heuristic_ff(st.last_pdf, environment_pdf(J.lights, env, ray_d));
        st.L = L + T * env_radiance(env, ray_d) * w;
        return kRetire;
    }
    RayQuery q;
    setup_query(q, ray_o, ray_d, st.tmin, st.tmax, st.lo, st.prev);
    const uint32_t tri_id = sc.prims[prim].id;
    // t only feeds the medium's attenuation as float(t) (and error reports,
    // which recompute it); u and v always feed the surface reconstruction.
    // Both come from float pairs where their bounds decide the results (no
    // maps, no split origin), else from the binary64 chain.
    const bool in_medium = st.medium >= 0;
    float tf = 0.0f;
    Surf s;
    bool have_surface = false;
    if (!precise && !sc.materials[sc.prims[prim].material].maps) {
        const PrimRec& p = sc.prims[prim];
        HitUvFF h;
        if (hit_uv_ff(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz}, h) &&
            (!in_medium || hit_t_ff(q, h, tf)))
            have_surface = surface_at_ff(sc, prim, ray_d, h, s);
    }
    if (!have_surface) {
        const ExactHit hit = exact_uv_of(sc, q, prim, in_medium);
        if (in_medium) tf = sf::to_float(hit.t);
        s = surface_at(sc, prim, ray_d, hit.u, hit.v, precise);
    }
    if (s.texture_error) {
        set_error(err, kErrTexture, depth, tri_id, s.material, sf::zero(0), s.position, ray_o, ray_d);
        st.L = v3(0);
        return kRetire;
    }
    if (in_medium) T = T * beer_lambert(st.sigma, tf);
    if ((s.front || s.double_sided) && max_component(s.emission) > 0) {
        const float w = last_delta || !nee
                            ? 1
                            : power_heuristic_ff(st.last_pdf, triangle_pdf(sc, J.lights, prim, st.last_pos, s.position));
        L = L + T * s.emission * w;
    }
    if (int(depth) == J.bounces) {
        st.L = L;
        return kRetire;
    }
    const V3 wo = -ray_d;
    Rng rng{st.rng};
    if (nee) {
        const LightSample light = sample_light(sc, J.lights, env, s.position, rng);
        if (light.tex_error) {
            set_error(err, kErrTexture, depth, tri_id, s.material, sf::zero(0), s.position, ray_o, ray_d);
            st.L = v3(0);
            return kRetire;
        }
        const bool enclosed_infinite = st.medium >= 0 && s.thickness <= 0 && !finite1(light.distance);
        if (!enclosed_infinite && light.pdf > 0 && max_component(light.radiance) > 0) {
            const BsdfEval bsdf = evaluate_bsdf(s, wo, light.direction);
            if (max_component(bsdf.value) > 0) {
                V3 so, sd, slo{0, 0, 0};
                float stmax = kInf;
                if (precise) {
                    so = s.position;
                    sd = light.direction;
                    slo = s.position_low;
                } else {
                    const PrimAux& ax = sc.aux[prim];
                    const uint32_t e = offset_ray(s.position, s.ng, V3{ax.ung[0], ax.ung[1], ax.ung[2]},
                                                  light.direction, kInf, so, sd, stmax);
                    if (e) {
                        set_error(err, e, depth, tri_id, s.material, sf::zero(0), s.position, ray_o, ray_d);
                        st.L = v3(0);
                        return kRetire;
                    }
                }
                if (finite1(light.distance)) {
                    V3 end = light.target;
                    if (length_sq(light.target_normal) > 0) {
                        V3 eo, ed;
                        float et;
                        const uint32_t e = offset_ray(end, light.target_normal, light.target_unormal, -light.direction,
                                                      kInf, eo, ed, et);
                        if (e) {
                            set_error(err, e, depth, tri_id, s.material, sf::zero(0), end, ray_o, ray_d);
                            st.L = v3(0);
                            return kRetire;
                        }
                        end = eo;
                    }
                    const V3 delta = end - so;
                    stmax = length(delta) * (1 - 2e-6f);
                    sd = normalize(delta);
                }
                const float w = light.delta ? 1 : power_heuristic_ff(light.pdf, bsdf.pdf);
                V3 shadow_sigma = st.sigma;
                bool shadow_in_medium = st.medium >= 0;
                if (s.thickness > 0 && dot(wo, s.ng) * dot(light.direction, s.ng) < 0) {
                    shadow_in_medium = s.front;
                    const DevMaterial& m = sc.materials[s.material];
                    shadow_sigma = s.front ? V3{m.sigma[0], m.sigma[1], m.sigma[2]} : V3{0, 0, 0};
                }
                const V3 tr = shadow_in_medium ? beer_lambert(shadow_sigma, light.distance) : v3(1);
                sh.emit = true;
                sh.o = so;
                sh.d = sd;
                sh.lo = slo;
                sh.tmax = stmax;
                sh.ignore = precise ? tri_id : kNoTriangle;
                sh.c = T * bsdf.value * light.radiance * tr * (fabsf(dot(s.normal, light.direction)) * w / light.pdf);
            }
        }
    }
    const BsdfSample smp = sample_bsdf(s, wo, rng);
    if (!sample_ok(smp)) {
        st.L = L;
        return kRetire;
    }
    T = T * smp.weight;
    if (!finite(T)) {
        set_error(err, kErrNonFiniteThroughput, depth, tri_id, s.material, sf::zero(0), s.position, ray_o, ray_d);
        sh.emit = false;
        st.L = v3(0);
        return kRetire;
    }
    int32_t medium = st.medium;
    V3 sigma = st.sigma;
    float eta_scale = st.eta_scale;
    if (smp.transmitted && s.thickness > 0) {
        uint32_t boundary = 0;
        if (s.front) {
            if (medium >= 0) boundary = kErrNested;
            else {
                medium = int32_t(s.material);
                const DevMaterial& m = sc.materials[s.material];
                sigma = V3{m.sigma[0], m.sigma[1], m.sigma[2]};
            }
        } else {
            if (medium < 0) boundary = kErrExitWithoutEntry;
            else {
                medium = -1;
                sigma = V3{0, 0, 0};
            }
        }
        if (boundary) {
            sh.emit = false;
            if (!precise) return kReplay;
            set_error(err, boundary, depth, tri_id, s.material, exact_of(sc, q, prim).t, s.position, ray_o, ray_d);
            st.L = v3(0);
            return kRetire;
        }
        eta_scale *= smp.eta * smp.eta;
    }
    V3 no, nd, nlo{0, 0, 0};
    float ntmax = kInf;
    if (precise) {
        no = s.position;
        nd = smp.direction;
        nlo = s.position_low;
    } else {
        const PrimAux& ax = sc.aux[prim];
        const uint32_t e =
            offset_ray(s.position, s.ng, V3{ax.ung[0], ax.ung[1], ax.ung[2]}, smp.direction, kInf, no, nd, ntmax);
        if (e) {
            set_error(err, e, depth, tri_id, s.material, sf::zero(0), s.position, ray_o, ray_d);
            sh.emit = false;
            st.L = v3(0);
            return kRetire;
        }
    }
    uint32_t action = kContinue;
    if (J.rr && depth >= 3) {
        const float survive = clampf(max_component(T) * eta_scale, 0.05f, 0.95f);
        if (rng.uniform() >= survive) action = kRetire;
        else T = T * (1 / survive);
    }
    st.rng = rng.state;
    st.T = T;
    st.L = L;
    st.eta_scale = eta_scale;
    st.last_pdf = smp.pdf;
    st.last_pos = s.position;
    st.sigma = sigma;
    st.medium = medium;
    st.o = no;
    st.d = nd;
    st.lo = nlo;
    st.tmin = 0;
    st.tmax = ntmax;
    st.prev = precise ? tri_id : kNoTriangle;
    st.flags = (depth + 1) | (smp.delta ? kLastDelta : 0u) | (precise ? kPrecise : 0u);
    return action;
}

// Final radiance check (reference src/render.cpp:346).
CR_HDI bool radiance_ok(V3 l) { return finite(l) && min_component(l) >= 0; }

}  // namespace cr
