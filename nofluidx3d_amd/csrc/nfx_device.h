// Device-side constitutive and contact math shared by the plain and the batched kernels.
//
// Every per-tetra law works on the displacement gradient H = F - I of the tetra, never on F itself.
// The fused and batched steps keep node DISPLACEMENTS u = x - X0 as their FP32 position state, so
//     H = [u1 - u0 | u2 - u0 | u3 - u0] * Dm^-1
// is formed from small differences of small numbers, and the laws below are written so that the
// strain measures they need (B - I, det F - 1, the elastic part of the viscoplastic split) come out
// of H without subtracting the identity from an O(1) quantity.  In FP32 an absolute position of
// magnitude ~50 carries ~4e-6 of rounding, which F = r Dm^-1 passes on to the strain directly; the
// displacement form keeps that error relative to the (small) strain instead.  The per-plugin
// kernels, which hold absolute positions as the reference's buffers do, form H = r Dm^-1 - I.
//
// Corner forces.  The reference computes rows f1..f3 = -V0 J r^-1 sigma with J = det(r)/(6 V0)
// (tetra.cl:159-168), re-inverting r every step.  Here f_k = -(1/6) sigma c_k with c_k the k-th
// column of cof(r) = r_{k+1} x r_{k+2}.  Since r = F Dm, cof(r) = cof(F) cof(Dm) and
// cof(Dm) = det(Dm) Dm^-T, so c_k = det(Dm) cof(F) d_k with d_k the k-th row of Dm^-1 and
// det(Dm) = 6 V0: no inversion and no current edges needed.  f0 = -(f1 + f2 + f3) (tetra.cl:168).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace nfx {

// Hardware transcendental/division approximations (v_rcp/v_sqrt/v_exp/v_log, about 1 ulp, no
// denormal rescaling).  Every operand they see here is a normal number or a case whose IEEE
// result they also give (log of a negative -> NaN, log 0 -> -inf, exp2 of +inf -> +inf).
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }

// Streaming accesses: data that a pass touches once (per-tetra state, tile tables) may take the
// nontemporal cache policy, so that they do not evict the node positions that neighbouring tiles
// gather again.  That pays when the state exceeds the 256 MB memory-side cache; a state that fits
// is better served by the default policy, which keeps it there from step to step.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
// Element i of an array whose base address every lane of the wave holds (a workgroup's tile or
// state plane).  The base is made explicitly wave-uniform and the lane offset is a 32-bit byte
// offset, so the access takes the scalar-base + vector-offset form with no per-lane 64-bit address
// arithmetic (the compiler otherwise derives each of a tetra's nine state planes from one 64-bit
// per-lane address).  Only for bases that are the same on every lane.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> uat(T* base, uint32_t i) {
    const uint64_t a = reinterpret_cast<uint64_t>(base);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    const gptr<B> b = (gptr<B>)((hi << 32) | lo);
    return (gptr<T>)(b + (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <bool NT, class T>
__device__ __forceinline__ T ld_g(gptr<const T> p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st_g(gptr<T> p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}
// a store through a plain pointer, with the policy a compile-time choice of the caller
template <bool NT, class T>
__device__ __forceinline__ void st_s(T* p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// symmetric 3x3: xx yy zz xy yz xz
struct S6 {
    float xx, yy, zz, xy, yz, xz;
};
__device__ __forceinline__ V3 smul(const S6& s, V3 v) {
    return V3{s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
              s.xz * v.x + s.yz * v.y + s.zz * v.z};
}

// general 3x3, row-major m[i][j]
struct M9 {
    float m[3][3];
};

// F = r * D where r has columns r1, r2, r3 (r = edge vectors: F; r = displacement differences: H)
__device__ __forceinline__ M9 deformation_gradient(V3 r1, V3 r2, V3 r3, const float* D) {
    M9 F;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        F.m[0][j] = r1.x * D[0 * 3 + j] + r2.x * D[1 * 3 + j] + r3.x * D[2 * 3 + j];
        F.m[1][j] = r1.y * D[0 * 3 + j] + r2.y * D[1 * 3 + j] + r3.y * D[2 * 3 + j];
        F.m[2][j] = r1.z * D[0 * 3 + j] + r2.z * D[1 * 3 + j] + r3.z * D[2 * 3 + j];
    }
    return F;
}

// Dm^-1 of a tetra from its reference edges R_k = X0_k - X0_0 (the columns of Dm), in FP32: rows
// (R2 x R3, R3 x R1, R1 x R2) / det with det = R1 . (R2 x R3) = 6 V0, returned.  The edges of FP32
// reference positions are correctly rounded differences (relative error 2^-24), so D differs from the
// FP64-computed, FP32-rounded Dm^-1 by a few units in the last place relative to |D|.
__device__ __forceinline__ float dminv_from_ref(V3 r1, V3 r2, V3 r3, float* D) {
    const V3 g1 = cross(r2, r3), g2 = cross(r3, r1), g3 = cross(r1, r2);
    const float det = dot(r1, g1);
    const float id = frcp(det);
    D[0] = g1.x * id; D[1] = g1.y * id; D[2] = g1.z * id;
    D[3] = g2.x * id; D[4] = g2.y * id; D[5] = g2.z * id;
    D[6] = g3.x * id; D[7] = g3.y * id; D[8] = g3.z * id;
    return det;
}
// The same rows unscaled, G = det(Dm) Dm^-1 = rows (R2 x R3, R3 x R1, R1 x R2); returns det.  The laws
// need Dm^-1 only as H = dU Dm^-1 = (dU G) / det and as the edge cofactors c_k = cof(F) g_k, so a
// caller holding G scales H once and passes G with scale 1 (no division of the nine entries).
__device__ __forceinline__ float shape_rows(V3 r1, V3 r2, V3 r3, float* G) {
    const V3 g1 = cross(r2, r3), g2 = cross(r3, r1), g3 = cross(r1, r2);
    G[0] = g1.x; G[1] = g1.y; G[2] = g1.z;
    G[3] = g2.x; G[4] = g2.y; G[5] = g2.z;
    G[6] = g3.x; G[7] = g3.y; G[8] = g3.z;
    return dot(r1, g1);
}

// H = F - I from the four node displacements (the fused and batched steps)
__device__ __forceinline__ M9 grad_disp(V3 u0, V3 u1, V3 u2, V3 u3, const float* D) {
    return deformation_gradient(u1 - u0, u2 - u0, u3 - u0, D);
}
// H = r D - I from absolute node positions (the per-plugin kernels)
__device__ __forceinline__ M9 grad_abs(V3 p0, V3 p1, V3 p2, V3 p3, const float* D) {
    M9 H = deformation_gradient(p1 - p0, p2 - p0, p3 - p0, D);
    H.m[0][0] -= 1.0f;
    H.m[1][1] -= 1.0f;
    H.m[2][2] -= 1.0f;
    return H;
}

// A * A^T (symmetric)
__device__ __forceinline__ S6 aat(const M9& A) {
    S6 s;
    s.xx = A.m[0][0] * A.m[0][0] + A.m[0][1] * A.m[0][1] + A.m[0][2] * A.m[0][2];
    s.yy = A.m[1][0] * A.m[1][0] + A.m[1][1] * A.m[1][1] + A.m[1][2] * A.m[1][2];
    s.zz = A.m[2][0] * A.m[2][0] + A.m[2][1] * A.m[2][1] + A.m[2][2] * A.m[2][2];
    s.xy = A.m[0][0] * A.m[1][0] + A.m[0][1] * A.m[1][1] + A.m[0][2] * A.m[1][2];
    s.yz = A.m[1][0] * A.m[2][0] + A.m[1][1] * A.m[2][1] + A.m[1][2] * A.m[2][2];
    s.xz = A.m[0][0] * A.m[2][0] + A.m[0][1] * A.m[2][1] + A.m[0][2] * A.m[2][2];
    return s;
}

// (I + H)(I + H)^T - I = H + H^T + H H^T: the left Cauchy-Green tensor less the identity
__device__ __forceinline__ S6 cauchy_green_m1(const M9& H) {
    S6 s = aat(H);
    s.xx += 2.0f * H.m[0][0];
    s.yy += 2.0f * H.m[1][1];
    s.zz += 2.0f * H.m[2][2];
    s.xy += H.m[0][1] + H.m[1][0];
    s.yz += H.m[1][2] + H.m[2][1];
    s.xz += H.m[0][2] + H.m[2][0];
    return s;
}

// det(I + H) - 1 as the sum of the principal invariants of H (no cancellation against the 1)
__device__ __forceinline__ float det_ipH_m1(const M9& H) {
    const float i1 = H.m[0][0] + H.m[1][1] + H.m[2][2];
    const float i2 = (H.m[0][0] * H.m[1][1] - H.m[0][1] * H.m[1][0]) + (H.m[0][0] * H.m[2][2] - H.m[0][2] * H.m[2][0]) +
                     (H.m[1][1] * H.m[2][2] - H.m[1][2] * H.m[2][1]);
    const float i3 = H.m[0][0] * (H.m[1][1] * H.m[2][2] - H.m[1][2] * H.m[2][1]) -
                     H.m[0][1] * (H.m[1][0] * H.m[2][2] - H.m[1][2] * H.m[2][0]) +
                     H.m[0][2] * (H.m[1][0] * H.m[2][1] - H.m[1][1] * H.m[2][0]);
    return i1 + (i2 + i3);
}
// the same for a symmetric S (xx yy zz xy yz xz)
__device__ __forceinline__ float det_ipS_m1(const S6& S) {
    const float i1 = S.xx + S.yy + S.zz;
    const float i2 = (S.xx * S.yy - S.xy * S.xy) + (S.xx * S.zz - S.xz * S.xz) + (S.yy * S.zz - S.yz * S.yz);
    const float i3 = S.xx * (S.yy * S.zz - S.yz * S.yz) - S.xy * (S.xy * S.zz - S.yz * S.xz) +
                     S.xz * (S.xy * S.yz - S.yy * S.xz);
    return i1 + (i2 + i3);
}

// cofactor matrix C = det(F) F^-T of F = I + H: row i is (row i+1 of F) x (row i+2 of F)
__device__ __forceinline__ M9 cof_ipH(const M9& H) {
    const V3 f0 = mk(1.0f + H.m[0][0], H.m[0][1], H.m[0][2]);
    const V3 f1 = mk(H.m[1][0], 1.0f + H.m[1][1], H.m[1][2]);
    const V3 f2 = mk(H.m[2][0], H.m[2][1], 1.0f + H.m[2][2]);
    const V3 c0 = cross(f1, f2), c1 = cross(f2, f0), c2 = cross(f0, f1);
    M9 C;
    C.m[0][0] = c0.x; C.m[0][1] = c0.y; C.m[0][2] = c0.z;
    C.m[1][0] = c1.x; C.m[1][1] = c1.y; C.m[1][2] = c1.z;
    C.m[2][0] = c2.x; C.m[2][1] = c2.y; C.m[2][2] = c2.z;
    return C;
}

// 1 / det(Dm^-1) = det(Dm) = 6 V0 (signed)
__device__ __forceinline__ float det_dm(const float* D) {
    const float d = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                    D[2] * (D[3] * D[7] - D[4] * D[6]);
    return frcp(d);
}

// the current edge cofactors c_k = r_{k+1} x r_{k+2} = s C d_k (s = det Dm = 6 V0, d_k = row k of Dm^-1)
struct Cof3 {
    V3 c1, c2, c3;
};
// c_k = C (s d_k) for rows d_k of D (D = Dm^-1 with s = det Dm, or D = G with s = 1)
__device__ __forceinline__ Cof3 edge_cofactors(const M9& C, const float* D, float s) {
    Cof3 o;
    V3* c[3] = {&o.c1, &o.c2, &o.c3};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const V3 d = mk(D[3 * k + 0] * s, D[3 * k + 1] * s, D[3 * k + 2] * s);
        *c[k] = mk(C.m[0][0] * d.x + C.m[0][1] * d.y + C.m[0][2] * d.z, C.m[1][0] * d.x + C.m[1][1] * d.y + C.m[1][2] * d.z,
                   C.m[2][0] * d.x + C.m[2][1] * d.y + C.m[2][2] * d.z);
    }
    return o;
}

// inverse by cofactors / det
__device__ __forceinline__ M9 inv9(const M9& A) {
    M9 R;
    const float c00 = A.m[1][1] * A.m[2][2] - A.m[1][2] * A.m[2][1];
    const float c01 = A.m[1][2] * A.m[2][0] - A.m[1][0] * A.m[2][2];
    const float c02 = A.m[1][0] * A.m[2][1] - A.m[1][1] * A.m[2][0];
    const float d = A.m[0][0] * c00 + A.m[0][1] * c01 + A.m[0][2] * c02;
    const float id = frcp(d);
    R.m[0][0] = c00 * id;
    R.m[1][0] = c01 * id;
    R.m[2][0] = c02 * id;
    R.m[0][1] = (A.m[0][2] * A.m[2][1] - A.m[0][1] * A.m[2][2]) * id;
    R.m[1][1] = (A.m[0][0] * A.m[2][2] - A.m[0][2] * A.m[2][0]) * id;
    R.m[2][1] = (A.m[0][1] * A.m[2][0] - A.m[0][0] * A.m[2][1]) * id;
    R.m[0][2] = (A.m[0][1] * A.m[1][2] - A.m[0][2] * A.m[1][1]) * id;
    R.m[1][2] = (A.m[0][2] * A.m[1][0] - A.m[0][0] * A.m[1][2]) * id;
    R.m[2][2] = (A.m[0][0] * A.m[1][1] - A.m[0][1] * A.m[1][0]) * id;
    return R;
}

__device__ __forceinline__ M9 mul9(const M9& A, const M9& B) {
    M9 C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
    return C;
}

__device__ __forceinline__ M9 ipH(const M9& H) {
    M9 F = H;
    F.m[0][0] += 1.0f;
    F.m[1][1] += 1.0f;
    F.m[2][2] += 1.0f;
    return F;
}

// compressible Mooney-Rivlin Cauchy stress (tetra.cl:118-156; Neo-Hookean when mu2 = 0) from
// Bd = B - I and Jm1 = J - 1.  The reference's
//     sigma = mu1 J^-5/3 dev B + mu2 J^-7/3 (I1 B - B^2 - 2/3 I2 I) + kappa (J - 1) I
// is rewritten with B = I + Bd, t = tr Bd: dev B = dev Bd and
//     I1 B - B^2 - 2/3 I2 I = dev Bd + dev(t Bd - Bd^2),
// so every term is of the order of the strain and no O(1) parts cancel.  J^-5/3 and J^-7/3 come
// from one cube root J^-1/3 = exp2(-log2(J)/3), so that J < 0 yields NaN and J = 0 yields Inf, as
// pow() does in the reference.
__device__ __forceinline__ S6 mr_stress(const S6& Bd, float Jm1, float mu1, float mu2, float kappa) {
    const float J = 1.0f + Jm1;
    const float t = fexp2(flog2(J) * (-1.0f / 3.0f));
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    const float tr = Bd.xx + Bd.yy + Bd.zz;
    const float d3 = tr * (1.0f / 3.0f);
    const float vol = kappa * Jm1;
    if (mu2 == 0.0f) {  // Neo-Hookean (uniform branch): no B^2 terms
        const float a = mu1 * t5;
        return S6{a * (Bd.xx - d3) + vol, a * (Bd.yy - d3) + vol, a * (Bd.zz - d3) + vol, a * Bd.xy, a * Bd.yz,
                  a * Bd.xz};
    }
    // W = t Bd - Bd^2 (symmetric)
    S6 W;
    W.xx = tr * Bd.xx - (Bd.xx * Bd.xx + Bd.xy * Bd.xy + Bd.xz * Bd.xz);
    W.yy = tr * Bd.yy - (Bd.xy * Bd.xy + Bd.yy * Bd.yy + Bd.yz * Bd.yz);
    W.zz = tr * Bd.zz - (Bd.xz * Bd.xz + Bd.yz * Bd.yz + Bd.zz * Bd.zz);
    W.xy = tr * Bd.xy - (Bd.xx * Bd.xy + Bd.xy * Bd.yy + Bd.xz * Bd.yz);
    W.yz = tr * Bd.yz - (Bd.xy * Bd.xz + Bd.yy * Bd.yz + Bd.yz * Bd.zz);
    W.xz = tr * Bd.xz - (Bd.xx * Bd.xz + Bd.xy * Bd.yz + Bd.xz * Bd.zz);
    const float w3 = (W.xx + W.yy + W.zz) * (1.0f / 3.0f);
    const float a = mu1 * t5;       // dU/dI1 J^-5/3
    const float k = mu2 * t5 * t2;  // dU/dI2 J^-7/3
    const float ak = a + k;
    S6 s;
    s.xx = ak * (Bd.xx - d3) + k * (W.xx - w3) + vol;
    s.yy = ak * (Bd.yy - d3) + k * (W.yy - w3) + vol;
    s.zz = ak * (Bd.zz - d3) + k * (W.zz - w3) + vol;
    s.xy = ak * Bd.xy + k * W.xy;
    s.yz = ak * Bd.yz + k * W.yz;
    s.xz = ak * Bd.xz + k * W.xz;
    return s;
}

struct Corners {
    V3 f0, f1, f2, f3;
};

// f_k = scale * sigma c_k, f0 = -(f1 + f2 + f3)
__device__ __forceinline__ Corners cofactor_forces(const S6& sig, const Cof3& c, float scale) {
    const S6 s{sig.xx * scale, sig.yy * scale, sig.zz * scale, sig.xy * scale, sig.yz * scale, sig.xz * scale};
    Corners o;
    o.f1 = smul(s, c.c1);
    o.f2 = smul(s, c.c2);
    o.f3 = smul(s, c.c3);
    o.f0 = (o.f1 + o.f2 + o.f3) * -1.0f;
    return o;
}
// the same from the current edges r1..r3 (c_k = r_{k+1} x r_{k+2})
__device__ __forceinline__ Corners cofactor_forces_r(const S6& sig, V3 r1, V3 r2, V3 r3, float scale) {
    return cofactor_forces(sig, Cof3{cross(r2, r3), cross(r3, r1), cross(r1, r2)}, scale);
}

// Mooney-Rivlin / Neo-Hookean branch of tetra.cl:115-168: J = det F (tetra.cl:132 computes
// det(r) / (6 V0), the same number), rows -V0 J r^-1 sigma = -(1/6) sigma c_k.
__device__ __forceinline__ Corners mr_tetra(const M9& H, const float* D, float v0, float mu1, float mu2,
                                            float kappa) {
    const S6 sig = mr_stress(cauchy_green_m1(H), det_ipH_m1(H), mu1, mu2, kappa);
    return cofactor_forces(sig, edge_cofactors(cof_ipH(H), D, 6.0f * v0), -1.0f / 6.0f);
}

// rows f1..f3 of  -V0 mu * D * M * F^T  (tetra.cl:93 and :114 share this shape), M symmetric or not
__device__ __forceinline__ Corners dmf_rows(const float* D, const M9& M, const M9& F, float scale) {
    // G = D * M
    M9 Dm;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Dm.m[i][j] = D[i * 3 + j];
    const M9 G = mul9(Dm, M);
    V3 rows[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // row i of G * F^T: (G F^T)_{il} = sum_j G_ij F_lj
        rows[i] = mk(G.m[i][0] * F.m[0][0] + G.m[i][1] * F.m[0][1] + G.m[i][2] * F.m[0][2],
                     G.m[i][0] * F.m[1][0] + G.m[i][1] * F.m[1][1] + G.m[i][2] * F.m[1][2],
                     G.m[i][0] * F.m[2][0] + G.m[i][1] * F.m[2][1] + G.m[i][2] * F.m[2][2]) *
                  scale;
    }
    Corners c;
    c.f1 = rows[0];
    c.f2 = rows[1];
    c.f3 = rows[2];
    c.f0 = (c.f1 + c.f2 + c.f3) * -1.0f;
    return c;
}

// St-Venant-Kirchhoff branch (tetra.cl:94-114): rows = -V0 mu * D F^T (kappa tr(B-I) I + (B-I)),
// kappa = nu/(1-2nu) (the ratio lambda/(2 mu)).  B - I comes from H directly.
__device__ __forceinline__ Corners svk_tetra(const M9& H, const float* D, float v0, float mu, float kappa) {
    const M9 F = ipH(H);
    const S6 Bd = cauchy_green_m1(H);
    const float trBP = Bd.xx + Bd.yy + Bd.zz;
    M9 FT;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FT.m[i][j] = F.m[j][i];
    M9 M;
    M.m[0][0] = kappa * trBP + Bd.xx;
    M.m[1][1] = kappa * trBP + Bd.yy;
    M.m[2][2] = kappa * trBP + Bd.zz;
    M.m[0][1] = M.m[1][0] = Bd.xy;
    M.m[1][2] = M.m[2][1] = Bd.yz;
    M.m[0][2] = M.m[2][0] = Bd.xz;
    const M9 FM = mul9(FT, M);
    // rows of D * FM
    Corners c;
    V3 rows[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        rows[i] = mk(D[i * 3 + 0] * FM.m[0][0] + D[i * 3 + 1] * FM.m[1][0] + D[i * 3 + 2] * FM.m[2][0],
                     D[i * 3 + 0] * FM.m[0][1] + D[i * 3 + 1] * FM.m[1][1] + D[i * 3 + 2] * FM.m[2][1],
                     D[i * 3 + 0] * FM.m[0][2] + D[i * 3 + 1] * FM.m[1][2] + D[i * 3 + 2] * FM.m[2][2]) *
                  (-v0 * mu);
    c.f1 = rows[0];
    c.f2 = rows[1];
    c.f3 = rows[2];
    c.f0 = (c.f1 + c.f2 + c.f3) * -1.0f;
    return c;
}

// "Kellnberger" St-Venant-Kirchhoff variant (tetra.cl:76-93): S = polar stretch of B by ten Newton
// steps X <- (X + X^-T)/2 (tetra.cl:5-12), inner = I + S^-1 (kappa (tr S - 3) - 1),
// rows = -V0 mu * D * inner * F^T.
__device__ __forceinline__ Corners svkk_tetra(const M9& H, const float* D, float v0, float mu, float kappa) {
    const M9 F = ipH(H);
    const S6 Bd = cauchy_green_m1(H);
    M9 B;
    B.m[0][0] = 1.0f + Bd.xx; B.m[1][1] = 1.0f + Bd.yy; B.m[2][2] = 1.0f + Bd.zz;
    B.m[0][1] = B.m[1][0] = Bd.xy;
    B.m[1][2] = B.m[2][1] = Bd.yz;
    B.m[0][2] = B.m[2][0] = Bd.xz;
    M9 X = B;
    for (int it = 0; it < 10; ++it) {
        const M9 Xi = inv9(X);  // (X^T)^-1 = (X^-1)^T
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) X.m[i][j] = 0.5f * (X.m[i][j] + Xi.m[j][i]);
    }
    M9 XT;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) XT.m[i][j] = X.m[j][i];
    const M9 S = mul9(XT, B);
    const M9 Si = inv9(S);
    const float sf = kappa * (S.m[0][0] + S.m[1][1] + S.m[2][2] - 3.0f) - 1.0f;
    M9 inner;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) inner.m[i][j] = (i == j ? 1.0f : 0.0f) + Si.m[i][j] * sf;
    return dmf_rows(D, inner, F, -v0 * mu);
}

// the face-area nodal forces of the legacy and viscoplastic kernels, -(sgn(det r)/6) sigma c_k
// (orientation free, as their outward normals are), with det r = det(Dm) det F
__device__ __forceinline__ Corners face_forces(const S6& sig, const M9& H, const float* D) {
    const float s = det_dm(D);
    const float detr = s * (1.0f + det_ipH_m1(H));
    return cofactor_forces(sig, edge_cofactors(cof_ipH(H), D, s), detr < 0.0f ? (1.0f / 6.0f) : (-1.0f / 6.0f));
}

__device__ __forceinline__ void stress_outputs(const S6& sig, float& von_mises, float& pressure) {
    const float tr3 = (sig.xx + sig.yy + sig.zz) * (1.0f / 3.0f);
    const float dx = sig.xx - tr3, dy = sig.yy - tr3, dz = sig.zz - tr3;
    von_mises = fsqrt(1.5f * (dx * dx + dy * dy + dz * dz + 2.0f * (sig.xy * sig.xy + sig.yz * sig.yz + sig.xz * sig.xz)));
    pressure = tr3;
}

// ---- legacy stress laws (kernels/InteractionLinearElastic.cl, InteractionLinearElasticDeviatoric.cl,
// InteractionSecondOrderNeoHookean.cl, InteractionPoroNeoHookean.cl) with node 0 as the edge base.
// Each gives a Cauchy stress from F = I + H; the nodal forces are the kernels' face-area forces.
struct LegacyConst {
    float mu, lam;    // linear elastic: 2 mu eps + lam tr(eps) I
    float K;          // deviatoric and second-order laws: bulk modulus
    float mu2;        // second-order coefficient
    float phi;        // porous volume fraction
};

template <int LAW>  // 4 linear elastic, 5 linear deviatoric, 6 second-order NH, 7 poro-NH
__device__ __forceinline__ Corners legacy_tetra(const M9& H, const float* D, const LegacyConst& L, float& von_mises,
                                                float& pressure) {
    S6 sig;
    if constexpr (LAW == 4 || LAW == 5) {
        // small strain eps = sym(F) - I = sym(H)  (InteractionLinearElastic.cl:60-62)
        const S6 e{H.m[0][0], H.m[1][1], H.m[2][2], 0.5f * (H.m[0][1] + H.m[1][0]), 0.5f * (H.m[1][2] + H.m[2][1]),
                   0.5f * (H.m[0][2] + H.m[2][0])};
        const float tr = e.xx + e.yy + e.zz;
        if constexpr (LAW == 4) {
            const float a = 2.0f * L.mu, b = L.lam * tr;
            sig = S6{a * e.xx + b, a * e.yy + b, a * e.zz + b, a * e.xy, a * e.yz, a * e.xz};
        } else {
            const float a = 2.0f * L.mu, t3 = tr * (1.0f / 3.0f), b = L.K * tr;
            sig = S6{a * (e.xx - t3) + b, a * (e.yy - t3) + b, a * (e.zz - t3) + b, a * e.xy, a * e.yz, a * e.xz};
        }
    } else {
        const S6 Bd = cauchy_green_m1(H);
        // J = sqrt(det B) = |det F| as the kernels take it
        const float Jm1s = det_ipH_m1(H);
        const float Jm1 = (Jm1s < -1.0f) ? (-2.0f - Jm1s) : Jm1s;
        const float J = 1.0f + Jm1;
        const float trBd = Bd.xx + Bd.yy + Bd.zz;
        const float t = fexp2(flog2(J) * (-1.0f / 3.0f));  // J^-1/3
        if constexpr (LAW == 6) {
            // a = J^-5/3 (mu + mu2 (trB J^-2/3 - 3)) on dev B, b = K (J - 1)
            const float t2 = t * t;
            const float a = t2 * t2 * t * (L.mu + L.mu2 * ((3.0f + trBd) * t2 - 3.0f));
            const float b = L.K * Jm1;
            const float d3 = trBd * (1.0f / 3.0f);
            sig = S6{a * (Bd.xx - d3) + b, a * (Bd.yy - d3) + b, a * (Bd.zz - d3) + b, a * Bd.xy, a * Bd.yz, a * Bd.xz};
        } else {
            // Danielsson-Parks-Boyce porous Neo-Hookean (InteractionPoroNeoHookean.cl:72-78)
            const float phi = L.phi;
            const float trB = 3.0f + trBd;
            const float eta = 1.0f + Jm1 * frcp(phi);
            const float ie3 = fexp2(flog2(eta) * (-1.0f / 3.0f));  // eta^-1/3
            const float iJ = frcp(J);
            const float t2 = t * t;
            const float a = 0.5f * L.mu * (4.0f * iJ - 2.0f * iJ * iJ - 2.0f * (phi + 2.0f * Jm1) * t2 * t2 * t * ie3);
            const float q = ((4.0f - phi) * eta + (1.0f - phi)) * frcp(phi * eta * eta + (1.0f - phi) * eta);
            const float b = 0.5f * L.mu * trB * (iJ * iJ - (1.0f / 3.0f) * t2 * ie3 * q);
            sig = S6{a * (1.0f + Bd.xx) + b, a * (1.0f + Bd.yy) + b, a * (1.0f + Bd.zz) + b, a * Bd.xy, a * Bd.yz,
                     a * Bd.xz};
        }
    }
    stress_outputs(sig, von_mises, pressure);
    return face_forces(sig, H, D);
}

struct VPOut {
    Corners f;
    float Fp[9];  // the updated state: Fp (vp_tetra), Q = Fp^-1 - I (vp_tetra_q) or Gd (vp_tetra_gd)
    float epsbar;
    float von_mises, pressure;
    bool flows;  // Dp != 0: the plastic state changes this step
};

struct VPConst {
    float mu1, mu2, kappa, m, Y, rate, eps0, n, sy;
    int linear;  // m == 1 and n == 0
};

// Finite-strain viscoplastic law (InteractionFiniteStrainViscoplastic.cl:57-206) with node 0 as
// the edge base.  The face-area nodal forces of :110-141 equal -(sgn(det r)/6) sigma c_k for a
// closed tetra surface (both are 1/3 of sigma times the outward area vector of the opposite face).
//
// vp_stress_flow: from Bed = Be - I of the elastic part, the Mooney-Rivlin stress with
// J = sqrt(det Be) (:82-107), von Mises stress and pressure (:144-151), the flow rate (:153,
// optionally with the yield offset) and epsbar += rate (:156).  Returns the Cauchy stress and the
// plastic stretching Dp = 1.5 rate dev(sigma) / svm (:160-169) as a symmetric tensor.
__device__ __forceinline__ S6 vp_stress_flow(const S6& Bed, float epsbar, const VPConst& P, VPOut& o, S6& sig) {
    // det Be - 1 = x, J = sqrt(1 + x): J - 1 = x / (1 + J)
    const float x = det_ipS_m1(Bed);
    const float Jr = fsqrt(1.0f + x);
    const float Jm1 = x * frcp(1.0f + Jr);
    sig = mr_stress(Bed, Jm1, P.mu1, P.mu2, P.kappa);
    const float tr3 = (sig.xx + sig.yy + sig.zz) * (1.0f / 3.0f);
    S6 dv = sig;
    dv.xx -= tr3;
    dv.yy -= tr3;
    dv.zz -= tr3;
    const float dd = dv.xx * dv.xx + dv.yy * dv.yy + dv.zz * dv.zz + 2.0f * (dv.xy * dv.xy + dv.yz * dv.yz + dv.xz * dv.xz);
    const float svm = fsqrt(1.5f * dd);
    o.von_mises = svm;
    o.pressure = tr3;
    float rate, g;  // Dp = g dev(sigma)
    if (P.linear && !(P.sy > 0.0f)) {
        // m = 1, n = 0, no yield offset (uniform branch): rate = rate0 svm / Y, so
        // g = 1.5 rate / svm = 1.5 rate0 / Y needs no von Mises stress (svm = 0 means dev = 0)
        const float k = P.rate * frcp(P.Y);
        rate = k * svm;
        g = 1.5f * k;
        o.epsbar = epsbar + rate;
    } else {
        const float over = P.sy > 0.0f ? fmaxf(svm - P.sy, 0.0f) : svm;
        if (P.linear) {  // m = 1, n = 0: sigma0 = Y, rate linear in the overstress
            rate = P.rate * over * frcp(P.Y);
        } else {
            // sigma0 = Y (1 + epsbar/eps0)^n and (over/sigma0)^m as 2^(p log2 x) on the hardware
            // exp2/log2 (about 1e-6 relative; a libm powf here is hundreds of instructions, part of
            // which the compiler hoists into the linear path as well); 0^m = 0 for m > 0, 1 for m = 0
            const float sigma0 = P.Y * fexp2(P.n * flog2(1.0f + epsbar * frcp(P.eps0)));
            const float ratio = over * frcp(sigma0);
            rate = P.rate * (ratio > 0.0f ? fexp2(P.m * flog2(ratio)) : (P.m == 0.0f ? 1.0f : 0.0f));
        }
        o.epsbar = epsbar + rate;
        g = (svm != 0.0f) ? rate * 1.5f * frcp(svm) : 0.0f;
    }
    // Dp = 0 (no overstress, or no deviatoric stress) gives N = 0: the state update adds exact zeros
    o.flows = g != 0.0f && svm != 0.0f;
    return S6{g * dv.xx, g * dv.yy, g * dv.zz, g * dv.xy, g * dv.yz, g * dv.xz};
}

// N = (I + Dp)^-1 Dp (symmetric: the two commute), by cofactors
__device__ __forceinline__ S6 flow_n(const S6& Dp) {
    const float axx = 1.0f + Dp.xx, ayy = 1.0f + Dp.yy, azz = 1.0f + Dp.zz;
    const float cxx = ayy * azz - Dp.yz * Dp.yz, cyy = axx * azz - Dp.xz * Dp.xz, czz = axx * ayy - Dp.xy * Dp.xy;
    const float cxy = Dp.xz * Dp.yz - Dp.xy * azz, cyz = Dp.xy * Dp.xz - axx * Dp.yz, cxz = Dp.xy * Dp.yz - ayy * Dp.xz;
    const float ia = frcp(axx * cxx + Dp.xy * cxy + Dp.xz * cxz);
    const S6 Ai{cxx * ia, cyy * ia, czz * ia, cxy * ia, cyz * ia, cxz * ia};
    return S6{Ai.xx * Dp.xx + Ai.xy * Dp.xy + Ai.xz * Dp.xz, Ai.xy * Dp.xy + Ai.yy * Dp.yy + Ai.yz * Dp.yz,
              Ai.xz * Dp.xz + Ai.yz * Dp.yz + Ai.zz * Dp.zz, Ai.xx * Dp.xy + Ai.xy * Dp.yy + Ai.xz * Dp.yz,
              Ai.xy * Dp.xz + Ai.yy * Dp.yz + Ai.yz * Dp.zz, Ai.xx * Dp.xz + Ai.xy * Dp.yz + Ai.xz * Dp.zz};
}

// State Fp (the reference's plasticDeformationTensor; the per-plugin kernel): Fe = F Fp^-1,
// Fp += Fe^-1 Dp F (:171-206).
__device__ __forceinline__ VPOut vp_tetra(const M9& H, const float* D, const float* Fp, float epsbar,
                                          const VPConst& P) {
    VPOut o;
    const M9 F = ipH(H);
    M9 Fpm;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fpm.m[i][j] = Fp[i * 3 + j];
    const M9 Fe = mul9(F, inv9(Fpm));
    const S6 Be = aat(Fe);
    S6 sig;
    const S6 Dp = vp_stress_flow(S6{Be.xx - 1.0f, Be.yy - 1.0f, Be.zz - 1.0f, Be.xy, Be.yz, Be.xz}, epsbar, P, o, sig);
    o.f = face_forces(sig, H, D);
    M9 Dpm;
    Dpm.m[0][0] = Dp.xx; Dpm.m[1][1] = Dp.yy; Dpm.m[2][2] = Dp.zz;
    Dpm.m[0][1] = Dpm.m[1][0] = Dp.xy;
    Dpm.m[1][2] = Dpm.m[2][1] = Dp.yz;
    Dpm.m[0][2] = Dpm.m[2][0] = Dp.xz;
    const M9 dF = mul9(mul9(inv9(Fe), Dpm), F);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) o.Fp[i * 3 + j] = Fp[i * 3 + j] + dF.m[i][j];
    return o;
}

// The same law on the plastic deviation Q = Fp^-1 - I (9 floats per tetra; Q = 0 at the start).
// Fe = F Fp^-1 = (I + H)(I + Q), so Fe - I = He = H + Q + H Q with no O(1) part.  The reference's
// update Fp' = Fp + Fe^-1 Dp F equals Fp F^-1 (I + Dp) F, so exactly
//     I + Q' = F^-1 (I + Dp)^-1 Fe,   i.e.   Q' = Q - F^-1 N Fe,   N = (I + Dp)^-1 Dp,
// an increment that is exactly zero when nothing flows.  F^-1 = C^T / det F with the cofactor
// matrix C the forces already use.  Stored as Fp itself in FP32, a small plastic increment would
// be rounded against the O(1) entries of Fp every step (a drift of ~1e-7 of Fp per step); stored
// as Q it keeps its own relative precision.
// s = det Dm (= 6 V0).  A tetra that does not flow keeps Q bit for bit; a wave whose tetra all do not
// flow (the far field under a yield offset) skips the update.
// D with scale gs gives the edge cofactors (D = Dm^-1, gs = s; or D = G, gs = 1); s = det Dm.
// zero_state (uniform over the caller's workgroup): Q is known to be zero, so Fe = F and He = H
__device__ __forceinline__ VPOut vp_tetra_q(const M9& H, const float* D, float gs, float s, const float* Q,
                                            float epsbar, const VPConst& P, bool zero_state = false) {
    VPOut o;
    M9 He = H;
    if (!zero_state) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                He.m[i][j] = H.m[i][j] + Q[3 * i + j] +
                             (H.m[i][0] * Q[0 * 3 + j] + H.m[i][1] * Q[1 * 3 + j] + H.m[i][2] * Q[2 * 3 + j]);
    }
    S6 sig;
    const S6 Dp = vp_stress_flow(cauchy_green_m1(He), epsbar, P, o, sig);
    const M9 C = cof_ipH(H);
    // det F from the cofactor row the forces need anyway: only its sign and 1/det F enter below (the
    // elastic volume change is taken from det Be), so the cancellation in det F - 1 does not matter
    const float J = (1.0f + H.m[0][0]) * C.m[0][0] + H.m[0][1] * C.m[0][1] + H.m[0][2] * C.m[0][2];
    o.f = cofactor_forces(sig, edge_cofactors(C, D, gs), s * J < 0.0f ? (1.0f / 6.0f) : (-1.0f / 6.0f));
    if (!o.flows) {
#pragma unroll
        for (int k = 0; k < 9; ++k) o.Fp[k] = Q[k];
        return o;
    }
    // W = N Fe = N (I + He)
    const S6 N = flow_n(Dp);
    const V3 n0 = mk(N.xx, N.xy, N.xz), n1 = mk(N.xy, N.yy, N.yz), n2 = mk(N.xz, N.yz, N.zz);
    M9 W;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const V3 h = mk(He.m[0][j], He.m[1][j], He.m[2][j]);
        const V3 nh = mk(dot(n0, h), dot(n1, h), dot(n2, h));
        W.m[0][j] = (j == 0 ? N.xx : j == 1 ? N.xy : N.xz) + nh.x;
        W.m[1][j] = (j == 0 ? N.xy : j == 1 ? N.yy : N.yz) + nh.y;
        W.m[2][j] = (j == 0 ? N.xz : j == 1 ? N.yz : N.zz) + nh.z;
    }
    // Q' = Q - C^T W / det F
    const float idet = frcp(J);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const V3 ci = mk(C.m[0][i], C.m[1][i], C.m[2][i]);  // column i of C = row i of C^T
#pragma unroll
        for (int j = 0; j < 3; ++j)
            o.Fp[3 * i + j] = Q[3 * i + j] - dot(ci, mk(W.m[0][j], W.m[1][j], W.m[2][j])) * idet;
    }
    return o;
}

// The same law on the symmetric plastic deviation Gd = Fp^-1 Fp^-T - I (6 floats: xx, yy, zz, xy,
// yz, xz; zero at the start).  The elastic left Cauchy-Green tensor is
//     Be = Fe Fe^T = F (I + Gd) F^T,   Be - I = (B - I) + F Gd F^T,
// so the isotropic stress, the flow and the forces depend on Gd alone (the plastic rotation it drops
// enters none of them), and I + Q' = F^-1 (I - N) Fe gives exactly
//     Gd' = Gd + F^-1 X F^-T,   X = -N Be - Be N + N Be N = (N^2 - 2N) Be,
// (N is a function of the deviatoric stress of an isotropic law, a polynomial in Be, so it commutes
// with Be), an increment that is zero when nothing flows.  Fp is not recoverable from Gd.
__device__ __forceinline__ VPOut vp_tetra_gd(const M9& H, const float* D, float gs, float s, const float* G6,
                                             float epsbar, const VPConst& P) {
    VPOut o;
    const S6 G{G6[0], G6[1], G6[2], G6[3], G6[4], G6[5]};
    const M9 F = ipH(H);
    const V3 fx = mk(F.m[0][0], F.m[0][1], F.m[0][2]), fy = mk(F.m[1][0], F.m[1][1], F.m[1][2]),
             fz = mk(F.m[2][0], F.m[2][1], F.m[2][2]);  // rows of F
    const V3 tx = smul(G, fx), ty = smul(G, fy), tz = smul(G, fz);
    S6 Bed = cauchy_green_m1(H);
    Bed.xx += dot(tx, fx);
    Bed.yy += dot(ty, fy);
    Bed.zz += dot(tz, fz);
    Bed.xy += dot(tx, fy);
    Bed.yz += dot(ty, fz);
    Bed.xz += dot(tx, fz);
    S6 sig;
    const S6 Dp = vp_stress_flow(Bed, epsbar, P, o, sig);
    const M9 C = cof_ipH(H);
    // det F from the cofactor row the forces need anyway: only its sign and 1/det F enter below (the
    // elastic volume change is taken from det Be), so the cancellation in det F - 1 does not matter
    const float J = (1.0f + H.m[0][0]) * C.m[0][0] + H.m[0][1] * C.m[0][1] + H.m[0][2] * C.m[0][2];
    o.f = cofactor_forces(sig, edge_cofactors(C, D, gs), s * J < 0.0f ? (1.0f / 6.0f) : (-1.0f / 6.0f));
    if (!o.flows) {
#pragma unroll
        for (int k = 0; k < 6; ++k) o.Fp[k] = G6[k];
        return o;
    }
    const S6 N = flow_n(Dp);
    // X = (N^2 - 2N)(I + Bed)
    const V3 n0 = mk(N.xx, N.xy, N.xz), n1 = mk(N.xy, N.yy, N.yz), n2 = mk(N.xz, N.yz, N.zz);
    const V3 m0 = smul(N, n0) - n0 * 2.0f, m1 = smul(N, n1) - n1 * 2.0f, m2 = smul(N, n2) - n2 * 2.0f;
    const V3 b0 = mk(1.0f + Bed.xx, Bed.xy, Bed.xz), b1 = mk(Bed.xy, 1.0f + Bed.yy, Bed.yz),
             b2 = mk(Bed.xz, Bed.yz, 1.0f + Bed.zz);
    const S6 X{dot(m0, b0), dot(m1, b1), dot(m2, b2), dot(m0, b1), dot(m1, b2), dot(m0, b2)};
    // F^-1 X F^-T = C^T X C / det F^2, with the columns k_i of C
    const V3 k0 = mk(C.m[0][0], C.m[1][0], C.m[2][0]), k1 = mk(C.m[0][1], C.m[1][1], C.m[2][1]),
             k2 = mk(C.m[0][2], C.m[1][2], C.m[2][2]);
    const float idet = frcp(J);
    const float id2 = idet * idet;
    const V3 u0 = smul(X, k0), u1 = smul(X, k1), u2 = smul(X, k2);
    o.Fp[0] = G.xx + dot(k0, u0) * id2;
    o.Fp[1] = G.yy + dot(k1, u1) * id2;
    o.Fp[2] = G.zz + dot(k2, u2) * id2;
    o.Fp[3] = G.xy + dot(k0, u1) * id2;
    o.Fp[4] = G.yz + dot(k1, u2) * id2;
    o.Fp[5] = G.xz + dot(k0, u2) * id2;
    return o;
}

// ---- node terms -------------------------------------------------------------------------------
// The per-plugin kernels take absolute FP32 positions (the reference's buffers).  The fused and
// batched node passes hold displacements u and the FP32 reference positions X0; their contact
// terms evaluate the penetration in FP64 on x = X0 + u (exact: both are FP32) and the FP64
// indenter position, because the penetration of a node near the tip (a few 1e-2 at first contact)
// is a difference of coordinates near 75, where FP32 keeps only ~4e-6.

// Sphere.cl:9-16; returns the force added to the node
__device__ __forceinline__ V3 sphere_term(V3 x, float k, float radius, float ypos) {
    const V3 r = mk(x.x, x.y - ypos, x.z);
    const float len = sqrtf(dot(r, r));
    const float dis = fmaxf(radius - len, 0.0f);
    if (dis <= 0.0f || len <= 0.0f) return mk(0.f, 0.f, 0.f);
    return r * (k * dis / len);
}

// Substrate.cl:4-5 (y component)
__device__ __forceinline__ float substrate_term(float y, float k, float wall) { return k * fmaxf(wall - y, 0.0f); }

// Plane.cl:4-5 (y component)
__device__ __forceinline__ float plane_term(float y, float k, float wall) { return -k * fmaxf(y - wall, 0.0f); }

// FP64 forms for x = X0 + u.  sphere_term_d: the force (fx, fy, fz) and the penetration; y is the
// FP64 vertical coordinate, x0/z0 and ux/uz the horizontal parts (read by the caller only for
// nodes with y > ypos - radius, the only ones the sphere can touch)
struct Contact {
    double fx, fy, fz, dis;
};
__device__ __forceinline__ Contact sphere_term_d(double x, double y, double z, double k, double radius, double ypos) {
    const double ry = y - ypos;
    const double len = sqrt(x * x + ry * ry + z * z);
    const double dis = fmax(radius - len, 0.0);
    if (dis <= 0.0 || len <= 0.0) return Contact{0.0, 0.0, 0.0, 0.0};
    const double s = k * dis / len;
    return Contact{x * s, ry * s, z * s, dis};
}

// VelocityVerlet.cl:11-15 with xi = 0.005, unit mass, dt = 1 (x may be a position or a displacement)
struct VV {
    static constexpr float a = 1.0f - 0.5f * 0.005f;
    static constexpr float inv_b = 1.0f / (1.0f + 0.5f * 0.005f);
    static constexpr float a_over_b = a * inv_b;
};
__device__ __forceinline__ void vv_update(V3& x, V3& v, V3 f, V3 fo) {
    v = (v * VV::a + (f + fo) * 0.5f) * VV::inv_b;
    x = x + v * VV::a + f * 0.5f;
}
// The same recurrence on (x, w), w = a v + f/2 of the previous step: v' = (w + f'/2)/b, so
// x' = x + a v' + f'/2 = x + w' with w' = (a/b)(w + f'/2) + f'/2.  No previous force is kept.
__device__ __forceinline__ void vv_update_w(V3& x, V3& w, V3 f) {
    const V3 h = f * 0.5f;
    w = (w + h) * VV::a_over_b + h;
    x = x + w;
}
// Entering the (x, w) form: the step's velocity from (v, f_old) as vv_update has it, then
// w' = a v' + f'/2 (the caller adds it to x).
__device__ __forceinline__ V3 vv_w_from_v(V3 v, V3 f, V3 fo) {
    const V3 vn = (v * VV::a + (f + fo) * 0.5f) * VV::inv_b;
    return vn * VV::a + f * 0.5f;
}
// Leaving it: v' = (w + f'/2)/b, x' = x + a v' + f'/2; w is replaced by v' (f' becomes f_old).
__device__ __forceinline__ void vv_v_from_w(V3& x, V3& w, V3 f) {
    const V3 h = f * 0.5f;
    const V3 vn = (w + h) * VV::inv_b;
    x = x + (vn * VV::a + h);
    w = vn;
}

// Simulation.py:173-183 / 117-132, in double
__device__ __forceinline__ double indenter_pos(double start, double speed, double move_steps, int64_t time) {
    const double t = (double)time;
    return (t <= move_steps) ? start - speed * t : start - speed * move_steps;
}

}  // namespace nfx
