// integrator.h — Langevin dynamics with NO_SQUISH rigid-body rotation, as device
// functions shared by the one-system step kernels and the batched replica kernel, so
// both advance a particle with the same arithmetic and the same random stream.
//
// Semantics follow the host engine's Langevin method (velocity Verlet on the total
// force): the Langevin force and torque drawn in one step's second half are added to
// the conservative force again in the next step's first half.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oxd {

// Counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}
__device__ __forceinline__ float u01f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }

// q1 (x) q2 for quaternions stored (w, x, y, z) in double4 (x=w, y=x, z=y, w=z).
__device__ __forceinline__ double4 qmul(double4 a, double4 b) {
    return make_double4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w,
                        a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
                        a.x * b.z - a.y * b.w + a.z * b.x + a.w * b.y,
                        a.x * b.w + a.y * b.z - a.z * b.y + a.w * b.x);
}
__device__ __forceinline__ double4 qconj(double4 a) { return make_double4(a.x, -a.y, -a.z, -a.w); }
// (0, w) (x) q and q (x) (0, w) for a pure vector w. For a unit q, q (x) (0, R^T w) =
// (0, w) (x) q: the body-frame torque kick of NO_SQUISH needs no rotation matrix.
__device__ __forceinline__ double4 qmul_vq(double wx, double wy, double wz, const double4& q) {
    return make_double4(-wx * q.y - wy * q.z - wz * q.w,
                        wx * q.x + wy * q.w - wz * q.z,
                        -wx * q.w + wy * q.x + wz * q.y,
                        wx * q.z - wy * q.y + wz * q.x);
}
__device__ __forceinline__ double4 qmul_qv(const double4& q, double wx, double wy, double wz) {
    return make_double4(-q.y * wx - q.z * wy - q.w * wz,
                        q.x * wx + q.z * wz - q.w * wy,
                        q.x * wy - q.y * wz + q.w * wx,
                        q.x * wz + q.y * wy - q.z * wx);
}
// cos and sin of x: Taylor series through x^7 / x^6 for |x| <= 2^-5 (the first dropped
// terms, x^9/9! and x^8/8!, are below 2.5e-18 and 2.3e-17 relative, under half an FP64
// ulp), through x^13 for |x| <= 1/4 (truncation below 2^-60), the library functions
// otherwise. A Langevin rotor turns by |omega| dt / 2 ~ 1e-3 rad: the short series is
// the rule, and a wavefront rarely runs a second branch.
__device__ __forceinline__ void sincos_small(double x, double& s, double& c) {
    if (fabs(x) <= 0.03125) {
        double x2 = x * x;
        s = x * (1.0 + x2 * (-1.0 / 6 + x2 * (1.0 / 120 + x2 * (-1.0 / 5040))));
        c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24 + x2 * (-1.0 / 720)));
    } else if (fabs(x) <= 0.25) {
        double x2 = x * x;
        s = x * (1.0 + x2 * (-1.0 / 6 + x2 * (1.0 / 120 + x2 * (-1.0 / 5040 + x2 * (1.0 / 362880 +
                 x2 * (-1.0 / 39916800 + x2 * (1.0 / 6227020800.0)))))));
        c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24 + x2 * (-1.0 / 720 + x2 * (1.0 / 40320 + x2 * (-1.0 / 3628800 +
                 x2 * (1.0 / 479001600.0))))));
    } else {
        sincos(x, &s, &c);
    }
}
// Free-rotor sub-step about body axis k for time tau (Miller et al., J. Chem. Phys.
// 116, 8649 (2002), NO_SQUISH): P_k permutes quaternion components.
__device__ __forceinline__ double4 permute(int k, double4 q) {
    if (k == 1) return make_double4(-q.y, q.x, q.w, -q.z);
    if (k == 2) return make_double4(-q.z, -q.w, q.x, q.y);
    return make_double4(-q.w, q.z, -q.y, q.x);
}
// inv4I = 1 / (4 I_k).
__device__ __forceinline__ void rotor(int k, double tau, double inv4I, double4& q, double4& p) {
    double4 Pq = permute(k, q), Pp = permute(k, p);
    double zeta = (p.x * Pq.x + p.y * Pq.y + p.z * Pq.z + p.w * Pq.w) * inv4I;
    double s, c;
    sincos_small(zeta * tau, s, c);
    q = make_double4(c * q.x + s * Pq.x, c * q.y + s * Pq.y, c * q.z + s * Pq.z, c * q.w + s * Pq.w);
    p = make_double4(c * p.x + s * Pp.x, c * p.y + s * Pp.y, c * p.z + s * Pp.z, c * p.w + s * Pp.w);
}

struct Langevin {
    double dt, kT, gamma, gamma_r, mass, inertia;
    // derived constants (make_langevin), so the per-particle chain holds no division
    // or square root of its own
    double hm;         // dt / (2 m)
    double ct;         // sqrt(6 gamma kT / dt): half-width of the uniform force noise
    double cr;         // sqrt(2 gamma_r kT / dt): deviation of the Gaussian torque noise
    double gr_half_I;  // gamma_r / (2 I): drag on a component of q* (x) p = (0, 2 L_body)
    double inv4I;      // 1 / (4 I)
};

__host__ __device__ inline Langevin make_langevin(double dt, double kT, double gamma, double gamma_r, double mass,
                                                  double inertia) {
    Langevin lv;
    lv.dt = dt; lv.kT = kT; lv.gamma = gamma; lv.gamma_r = gamma_r; lv.mass = mass; lv.inertia = inertia;
    lv.hm = 0.5 * dt / mass;
    lv.ct = sqrt(6.0 * gamma * kT / dt);
    lv.cr = sqrt(2.0 * gamma_r * kT / dt);
    lv.gr_half_I = 0.5 * gamma_r / inertia;
    lv.inv4I = 0.25 / inertia;
    return lv;
}

// 1 / sqrt(s) for the squared norm of a quaternion that left the rotor chain: within
// 2^-20 of 1 (always, but for a caller's unnormalised input) by 1 - d/2 + 3 d^2/8,
// d = s - 1, whose truncation (5/16 |d|^3 < 2^-61) is below half an FP64 ulp.
__device__ __forceinline__ double rsqrt_near1(double s) {
    const double d = s - 1.0;
    if (fabs(d) <= 0x1p-20) return 1.0 + d * (-0.5 + 0.375 * d);
    return 1.0 / sqrt(s);
}

// x += dt v, wrapped into the box centred on the origin.
__device__ __forceinline__ void lv_drift(const Langevin& lv, const double L[3], const double Linv[3], const double4& v,
                                         double4& x) {
    const double dt = lv.dt;
    x.x += dt * v.x; x.y += dt * v.y; x.z += dt * v.z;
    x.x -= L[0] * floor(x.x * Linv[0] + 0.5);
    x.y -= L[1] * floor(x.y * Linv[1] + 0.5);
    x.z -= L[2] * floor(x.z * Linv[2] + 0.5);
}

// First half of a step for one particle: v += dt/2 (F + F_L)/m, x += dt v (wrapped
// into the box centred on the origin), p += dt q (x) (tau + tau_L)_body, NO_SQUISH
// rotation (axis 3: dt/2, 2: dt/2, 1: dt, 2: dt/2, 3: dt/2), renormalised q.
// (f, t): conservative force and torque (lab frame); (bd, bdt): the Langevin force and
// torque drawn in the previous second half. x.w (type bits) is left untouched.
// The translational part of lv_first_half (kick and drift).
__device__ __forceinline__ void lv_first_half_translate(const Langevin& lv, const double L[3], const double Linv[3],
                                                        float fx, float fy, float fz, const double4& bd,
                                                        double4& x, double4& v) {
    const double hm = lv.hm;
    v.x += hm * (fx + bd.x); v.y += hm * (fy + bd.y); v.z += hm * (fz + bd.z);
    lv_drift(lv, L, Linv, v, x);
}

// NO_SQUISH free rotation (axis 3: dt/2, 2: dt/2, 1: dt, 2: dt/2, 3: dt/2) and renormalisation.
__device__ __forceinline__ void lv_rotate(const Langevin& lv, double4& q, double4& p) {
    const double dt = lv.dt;
    const double inv4I = lv.inv4I;
    rotor(3, 0.5 * dt, inv4I, q, p);
    rotor(2, 0.5 * dt, inv4I, q, p);
    rotor(1, dt, inv4I, q, p);
    rotor(2, 0.5 * dt, inv4I, q, p);
    rotor(3, 0.5 * dt, inv4I, q, p);
    const double nq = rsqrt_near1(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    q.x *= nq; q.y *= nq; q.z *= nq; q.w *= nq;
}

// The rotational part of lv_first_half (kick, NO_SQUISH rotation, renormalisation).
__device__ __forceinline__ void lv_first_half_rotate(const Langevin& lv, float tx, float ty, float tz,
                                                     const double4& bdt, double4& q, double4& p) {
    const double dt = lv.dt;
    // p += dt q (x) (0, tau_body) = dt (0, tau_lab) (x) q
    const double4 kick = qmul_vq(tx + bdt.x, ty + bdt.y, tz + bdt.z, q);
    p.x += dt * kick.x; p.y += dt * kick.y; p.z += dt * kick.z; p.w += dt * kick.w;
    lv_rotate(lv, q, p);
}

__device__ __forceinline__ void lv_first_half(const Langevin& lv, const double L[3], const double Linv[3],
                                              float fx, float fy, float fz, float tx, float ty, float tz,
                                              const double4& bd, const double4& bdt,
                                              double4& x, double4& v, double4& q, double4& p) {
    lv_first_half_translate(lv, L, Linv, fx, fy, fz, bd, x, v);
    lv_first_half_rotate(lv, tx, ty, tz, bdt, q, p);
}

// Langevin force and torque of the second half-kick of step `step` for particle `i`
// with random key `seed`, from the state (v, q, p) before that kick: b = -gamma v plus
// uniform noise of variance 2 gamma kT / dt per component (lab frame); rb = -gamma_r
// L_body / I plus Gaussian noise of variance 2 gamma_r kT / dt (body frame).
__device__ __forceinline__ void lv_noise(const Langevin& lv, uint32_t i, unsigned long long step,
                                         unsigned long long seed, const double4& v, const double4& q,
                                         const double4& p, double b[3], double rb[3]) {
    uint32_t c[4] = {i, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t d[4] = {i, (uint32_t)step, (uint32_t)(step >> 32), 1u};
    philox(d, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double ct = lv.ct;
    b[0] = -lv.gamma * v.x + ct * (2.0 * u01f(c[0]) - 1.0);
    b[1] = -lv.gamma * v.y + ct * (2.0 * u01f(c[1]) - 1.0);
    b[2] = -lv.gamma * v.z + ct * (2.0 * u01f(c[2]) - 1.0);
    const double4 Lb = qmul(qconj(q), p);  // = (0, 2 L_body)
    const double cr = lv.cr, gr = lv.gr_half_I;
    // Box-Muller in FP32: the noise needs no FP64 accuracy
    float r1 = sqrtf(-2.f * __logf(u01f(c[3]))), r2 = sqrtf(-2.f * __logf(u01f(d[1])));
    float s0, c0, s2, c2;
    __sincosf(6.2831853f * u01f(d[0]), &s0, &c0);
    __sincosf(6.2831853f * u01f(d[2]), &s2, &c2);
    rb[0] = cr * (double)(r1 * c0) - gr * Lb.y;
    rb[1] = cr * (double)(r1 * s0) - gr * Lb.z;
    rb[2] = cr * (double)(r2 * c2) - gr * Lb.w;
}

// Second half of step `step` for particle `i`: v += dt/2 (F + b)/m, p += dt q (x) (tau_body
// + rb), with (b, rb) from lv_noise. Returns the Langevin force and torque (lab frame) in
// bd / bdt for the next step's first half.
__device__ __forceinline__ void lv_second_half(const Langevin& lv, uint32_t i, unsigned long long step,
                                               unsigned long long seed, float fx, float fy, float fz,
                                               float tx, float ty, float tz, const double4& q,
                                               double4& v, double4& p, double4& bd, double4& bdt) {
    const double dt = lv.dt;
    double b[3], rb[3];
    lv_noise(lv, i, step, seed, v, q, p, b, rb);
    const double hm = lv.hm;
    v.x += hm * (fx + b[0]); v.y += hm * (fy + b[1]); v.z += hm * (fz + b[2]);
    bd = make_double4(b[0], b[1], b[2], 0.0);
    const double4 K = qmul_qv(q, rb[0], rb[1], rb[2]);  // q (x) (0, rb)
    const double4 T = qmul_vq(tx, ty, tz, q);            // q (x) (0, R^T tau)
    p.x += dt * (T.x + K.x); p.y += dt * (T.y + K.y); p.z += dt * (T.z + K.z); p.w += dt * (T.w + K.w);
    const double4 l = qmul(K, qconj(q));  // (0, R rb): the Langevin torque in the lab frame
    bdt = make_double4(l.y, l.z, l.w, 0.0);
}

// The second half of step k merged with the first half of step k + 1 (velocity Verlet
// as leapfrog: the two half-kicks with the same force add up), in two parts. lv_merge_pre
// needs only the state before the force evaluation and may run ahead of it: with (b, rb)
// of step k, v += dt b / m and p += 2 dt q (x) (0, rb). lv_merge_kick then adds the force
// and torque of step k: v += dt F / m, p += 2 dt (0, tau_lab) (x) q.
__device__ __forceinline__ void lv_merge_pre(const Langevin& lv, uint32_t i, unsigned long long step,
                                             unsigned long long seed, const double4& q, double4& v, double4& p) {
    double b[3], rb[3];
    lv_noise(lv, i, step, seed, v, q, p, b, rb);
    const double h = 2.0 * lv.hm, h2 = 2.0 * lv.dt;
    const double4 K = qmul_qv(q, rb[0], rb[1], rb[2]);
    v.x += h * b[0]; v.y += h * b[1]; v.z += h * b[2];
    p.x += h2 * K.x; p.y += h2 * K.y; p.z += h2 * K.z; p.w += h2 * K.w;
}
// lv_merge_pre in two independent parts, for two lanes: the translational one (b from the
// first Philox block) and the rotational one (rb from the first block's last word and the
// second block). Each performs the same operations as lv_merge_pre on its quantities.
__device__ __forceinline__ void lv_merge_pre_translate(const Langevin& lv, uint32_t i, unsigned long long step,
                                                       unsigned long long seed, double4& v) {
    uint32_t c[4] = {i, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double ct = lv.ct;
    const double b0 = -lv.gamma * v.x + ct * (2.0 * u01f(c[0]) - 1.0);
    const double b1 = -lv.gamma * v.y + ct * (2.0 * u01f(c[1]) - 1.0);
    const double b2 = -lv.gamma * v.z + ct * (2.0 * u01f(c[2]) - 1.0);
    const double h = 2.0 * lv.hm;
    v.x += h * b0; v.y += h * b1; v.z += h * b2;
}
__device__ __forceinline__ void lv_merge_pre_rotate(const Langevin& lv, uint32_t i, unsigned long long step,
                                                    unsigned long long seed, const double4& q, double4& p) {
    uint32_t c[4] = {i, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t d[4] = {i, (uint32_t)step, (uint32_t)(step >> 32), 1u};
    philox(d, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double4 Lb = qmul(qconj(q), p);  // = (0, 2 L_body)
    const double cr = lv.cr, gr = lv.gr_half_I;
    float r1 = sqrtf(-2.f * __logf(u01f(c[3]))), r2 = sqrtf(-2.f * __logf(u01f(d[1])));
    float s0, c0, s2, c2;
    __sincosf(6.2831853f * u01f(d[0]), &s0, &c0);
    __sincosf(6.2831853f * u01f(d[2]), &s2, &c2);
    const double rb0 = cr * (double)(r1 * c0) - gr * Lb.y;
    const double rb1 = cr * (double)(r1 * s0) - gr * Lb.z;
    const double rb2 = cr * (double)(r2 * c2) - gr * Lb.w;
    const double h2 = 2.0 * lv.dt;
    const double4 K = qmul_qv(q, rb0, rb1, rb2);
    p.x += h2 * K.x; p.y += h2 * K.y; p.z += h2 * K.z; p.w += h2 * K.w;
}
__device__ __forceinline__ void lv_merge_kick(const Langevin& lv, float fx, float fy, float fz, float tx, float ty,
                                              float tz, const double4& q, double4& v, double4& p) {
    const double h = 2.0 * lv.hm, h2 = 2.0 * lv.dt;
    v.x += h * fx; v.y += h * fy; v.z += h * fz;
    const double4 T = qmul_vq(tx, ty, tz, q);
    p.x += h2 * T.x; p.y += h2 * T.y; p.z += h2 * T.z; p.w += h2 * T.w;
}

}  // namespace oxd
