// Device-side constitutive and contact math shared by the plain and the batched kernels.
//
// All per-tetra functions take the four gathered node positions and the tetra's inverse
// reference shape matrix D = Dm^-1 (row-major, Dm = [R1 | R2 | R3] as columns, TetraCell.py:33-38)
// and return the four corner forces f0..f3.  They restate the reference's laws in an algebraically
// equivalent form that avoids the two per-step 3x3 inversions of tetra.cl:53,159:
//
//   reference (tetra.cl:159-168):  rows f1..f3 = -V0 * J * r^-1 * sigma,  J = det(r)/(6 V0)
//   here:                          f_k = -(1/6) * sigma * c_k,  c_k = r_{k+1} x r_{k+2}
//
// because V0 * J * r^-1 = det(r)/6 * adj(r)/det(r) and the rows of adj(r) are c_1, c_2, c_3;
// sigma is symmetric.  f0 = -(f1 + f2 + f3) (tetra.cl:168).
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

// F = r * D where r has columns r1, r2, r3
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

// compressible Mooney-Rivlin Cauchy stress (tetra.cl:118-156; NH when mu2 = 0) from the left
// Cauchy-Green tensor B and J.  J^-5/3 and J^-7/3 come from one cube root t = J^-1/3, taken as
// exp2(-log2(J)/3) so that J < 0 yields NaN and J = 0 yields Inf, as pow() does in the reference.
__device__ __forceinline__ S6 mr_stress(const S6& B, float J, float mu1, float mu2, float kappa) {
    const float trB = B.xx + B.yy + B.zz;
    if (mu2 == 0.0f) {  // Neo-Hookean (uniform branch): no B^2 terms
        const float t = fexp2(flog2(J) * (-1.0f / 3.0f));
        const float t2 = t * t;
        const float a = mu1 * t2 * t2 * t;
        const float iso = kappa * (J - 1.0f) - trB * (1.0f / 3.0f) * a;
        return S6{a * B.xx + iso, a * B.yy + iso, a * B.zz + iso, a * B.xy, a * B.yz, a * B.xz};
    }
    // B^2 (symmetric)
    S6 B2;
    B2.xx = B.xx * B.xx + B.xy * B.xy + B.xz * B.xz;
    B2.yy = B.xy * B.xy + B.yy * B.yy + B.yz * B.yz;
    B2.zz = B.xz * B.xz + B.yz * B.yz + B.zz * B.zz;
    B2.xy = B.xx * B.xy + B.xy * B.yy + B.xz * B.yz;
    B2.yz = B.xy * B.xz + B.yy * B.yz + B.yz * B.zz;
    B2.xz = B.xx * B.xz + B.xy * B.yz + B.xz * B.zz;
    const float trB2 = B2.xx + B2.yy + B2.zz;
    const float t = fexp2(flog2(J) * (-1.0f / 3.0f));
    const float t2 = t * t;
    const float t5 = t2 * t2 * t;
    const float a = mu1 * t5;        // dU/dI1
    const float k = mu2 * t5 * t2;   // dU/dK
    const float iso = kappa * (J - 1.0f) - trB * (1.0f / 3.0f) * a + k * (trB2 - trB * trB) * (1.0f / 3.0f);
    const float bc = a + k * trB;    // coefficient of B
    S6 s;
    s.xx = bc * B.xx - k * B2.xx + iso;
    s.yy = bc * B.yy - k * B2.yy + iso;
    s.zz = bc * B.zz - k * B2.zz + iso;
    s.xy = bc * B.xy - k * B2.xy;
    s.yz = bc * B.yz - k * B2.yz;
    s.xz = bc * B.xz - k * B2.xz;
    return s;
}

struct Corners {
    V3 f0, f1, f2, f3;
};

// f_k = -(sgn/6) sigma c_k
__device__ __forceinline__ Corners cofactor_forces(const S6& sig, V3 r1, V3 r2, V3 r3, float scale) {
    const S6 s{sig.xx * scale, sig.yy * scale, sig.zz * scale, sig.xy * scale, sig.yz * scale, sig.xz * scale};
    Corners c;
    c.f1 = smul(s, cross(r2, r3));
    c.f2 = smul(s, cross(r3, r1));
    c.f3 = smul(s, cross(r1, r2));
    c.f0 = (c.f1 + c.f2 + c.f3) * -1.0f;
    return c;
}

// Mooney-Rivlin / Neo-Hookean branch of tetra.cl:115-168.
__device__ __forceinline__ Corners mr_tetra(V3 p0, V3 p1, V3 p2, V3 p3, const float* D, float v0, float mu1,
                                            float mu2, float kappa) {
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const M9 F = deformation_gradient(r1, r2, r3, D);
    const S6 B = aat(F);
    const float J = dot(r1, cross(r2, r3)) * frcp(6.0f * v0);  // tetra.cl:132
    const S6 sig = mr_stress(B, J, mu1, mu2, kappa);
    return cofactor_forces(sig, r1, r2, r3, -1.0f / 6.0f);
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
// kappa = nu/(1-2nu) (the ratio lambda/(2 mu)).
__device__ __forceinline__ Corners svk_tetra(V3 p0, V3 p1, V3 p2, V3 p3, const float* D, float v0, float mu,
                                             float kappa) {
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const M9 F = deformation_gradient(r1, r2, r3, D);
    const S6 B = aat(F);
    const float trBP = B.xx + B.yy + B.zz - 3.0f;
    // H = D F^T; Force = H * (kappa trBP I + BP) * (-V0 mu)  ==  D * F^T * M
    M9 FT;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FT.m[i][j] = F.m[j][i];
    M9 M;
    M.m[0][0] = kappa * trBP + B.xx - 1.0f;
    M.m[1][1] = kappa * trBP + B.yy - 1.0f;
    M.m[2][2] = kappa * trBP + B.zz - 1.0f;
    M.m[0][1] = M.m[1][0] = B.xy;
    M.m[1][2] = M.m[2][1] = B.yz;
    M.m[0][2] = M.m[2][0] = B.xz;
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
__device__ __forceinline__ Corners svkk_tetra(V3 p0, V3 p1, V3 p2, V3 p3, const float* D, float v0, float mu,
                                              float kappa) {
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const M9 F = deformation_gradient(r1, r2, r3, D);
    const S6 Bs = aat(F);
    M9 B;
    B.m[0][0] = Bs.xx; B.m[1][1] = Bs.yy; B.m[2][2] = Bs.zz;
    B.m[0][1] = B.m[1][0] = Bs.xy;
    B.m[1][2] = B.m[2][1] = Bs.yz;
    B.m[0][2] = B.m[2][0] = Bs.xz;
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

// ---- legacy stress laws (kernels/InteractionLinearElastic.cl, InteractionLinearElasticDeviatoric.cl,
// InteractionSecondOrderNeoHookean.cl, InteractionPoroNeoHookean.cl) with node 0 as the edge base.
// Each gives a Cauchy stress from F = r D; the nodal forces are the kernels' face-area forces,
// -(sgn(det r)/6) sigma c_k (orientation free, as the outward normals there are).
struct LegacyConst {
    float mu, lam;    // linear elastic: 2 mu eps + lam tr(eps) I
    float K;          // deviatoric and second-order laws: bulk modulus
    float mu2;        // second-order coefficient
    float phi;        // porous volume fraction
};

template <int LAW>  // 4 linear elastic, 5 linear deviatoric, 6 second-order NH, 7 poro-NH
__device__ __forceinline__ Corners legacy_tetra(V3 p0, V3 p1, V3 p2, V3 p3, const float* D, const LegacyConst& L,
                                                float& von_mises, float& pressure) {
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const M9 F = deformation_gradient(r1, r2, r3, D);
    S6 sig;
    if constexpr (LAW == 4 || LAW == 5) {
        // small strain eps = sym(F) - I  (InteractionLinearElastic.cl:60-62)
        const S6 e{F.m[0][0] - 1.0f, F.m[1][1] - 1.0f, F.m[2][2] - 1.0f, 0.5f * (F.m[0][1] + F.m[1][0]),
                   0.5f * (F.m[1][2] + F.m[2][1]), 0.5f * (F.m[0][2] + F.m[2][0])};
        const float tr = e.xx + e.yy + e.zz;
        if constexpr (LAW == 4) {
            const float a = 2.0f * L.mu, b = L.lam * tr;
            sig = S6{a * e.xx + b, a * e.yy + b, a * e.zz + b, a * e.xy, a * e.yz, a * e.xz};
        } else {
            const float a = 2.0f * L.mu, t3 = tr * (1.0f / 3.0f), b = L.K * tr;
            sig = S6{a * (e.xx - t3) + b, a * (e.yy - t3) + b, a * (e.zz - t3) + b, a * e.xy, a * e.yz, a * e.xz};
        }
    } else {
        const S6 B = aat(F);
        const float detB = B.xx * (B.yy * B.zz - B.yz * B.yz) - B.xy * (B.xy * B.zz - B.yz * B.xz) +
                           B.xz * (B.xy * B.yz - B.yy * B.xz);
        const float J = fsqrt(detB);
        const float trB = B.xx + B.yy + B.zz;
        const float t = fexp2(flog2(J) * (-1.0f / 3.0f));  // J^-1/3
        if constexpr (LAW == 6) {
            // a = J^-5/3 (mu + mu2 (trB J^-2/3 - 3)) on dev B, b = K (J - 1)
            const float t2 = t * t;
            const float a = t2 * t2 * t * (L.mu + L.mu2 * (trB * t2 - 3.0f));
            const float b = L.K * (J - 1.0f);
            const float d3 = trB * (1.0f / 3.0f);
            sig = S6{a * (B.xx - d3) + b, a * (B.yy - d3) + b, a * (B.zz - d3) + b, a * B.xy, a * B.yz, a * B.xz};
        } else {
            // Danielsson-Parks-Boyce porous Neo-Hookean (InteractionPoroNeoHookean.cl:72-78)
            const float phi = L.phi;
            const float eta = 1.0f + (J - 1.0f) * frcp(phi);
            const float ie3 = fexp2(flog2(eta) * (-1.0f / 3.0f));  // eta^-1/3
            const float iJ = frcp(J);
            const float t2 = t * t;
            const float a = 0.5f * L.mu * (4.0f * iJ - 2.0f * iJ * iJ - 2.0f * (phi + 2.0f * (J - 1.0f)) * t2 * t2 * t * ie3);
            const float q = ((4.0f - phi) * eta + (1.0f - phi)) * frcp(phi * eta * eta + (1.0f - phi) * eta);
            const float b = 0.5f * L.mu * trB * (iJ * iJ - (1.0f / 3.0f) * t2 * ie3 * q);
            sig = S6{a * B.xx + b, a * B.yy + b, a * B.zz + b, a * B.xy, a * B.yz, a * B.xz};
        }
    }
    const float tr3 = (sig.xx + sig.yy + sig.zz) * (1.0f / 3.0f);
    const float dx = sig.xx - tr3, dy = sig.yy - tr3, dz = sig.zz - tr3;
    von_mises = fsqrt(1.5f * (dx * dx + dy * dy + dz * dz + 2.0f * (sig.xy * sig.xy + sig.yz * sig.yz + sig.xz * sig.xz)));
    pressure = tr3;
    const float detr = dot(r1, cross(r2, r3));
    return cofactor_forces(sig, r1, r2, r3, detr < 0.0f ? (1.0f / 6.0f) : (-1.0f / 6.0f));
}

struct VPOut {
    Corners f;
    float Fp[9];  // the updated state: Fp (vp_tetra) or the shape state S (vp_tetra_shape)
    float epsbar;
    float von_mises, pressure;
};

struct VPConst {
    float mu1, mu2, kappa, m, Y, rate, eps0, n, sy;
    int linear;  // m == 1 and n == 0
};

// Finite-strain viscoplastic law (InteractionFiniteStrainViscoplastic.cl:57-206) with node 0 as
// the edge base.  The face-area nodal forces of :110-141 equal -(sgn(det r)/6) sigma c_k for a
// closed tetra surface (both are 1/3 of sigma times the outward area vector of the opposite face).
//
// vp_stress_flow: from the elastic part Fe, the Mooney-Rivlin stress on Be = Fe Fe^T with
// J = sqrt(det Be) (:82-107), the corner forces, von Mises stress and pressure (:144-151), the
// flow rate (:153, optionally with the yield offset) and epsbar += rate (:156).  Returns the
// plastic stretching Dp = 1.5 rate dev(sigma) / svm (:160-169) as a symmetric tensor.
__device__ __forceinline__ S6 vp_stress_flow_be(const S6& B, V3 r1, V3 r2, V3 r3, float epsbar, const VPConst& P,
                                                VPOut& o) {
    const float detB = B.xx * (B.yy * B.zz - B.yz * B.yz) - B.xy * (B.xy * B.zz - B.yz * B.xz) +
                       B.xz * (B.xy * B.yz - B.yy * B.xz);
    const float J = fsqrt(detB);
    const S6 sig = mr_stress(B, J, P.mu1, P.mu2, P.kappa);
    const float detr = dot(r1, cross(r2, r3));
    o.f = cofactor_forces(sig, r1, r2, r3, detr < 0.0f ? (1.0f / 6.0f) : (-1.0f / 6.0f));
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
    S6 Dp;
    Dp.xx = g * dv.xx;
    Dp.yy = g * dv.yy;
    Dp.zz = g * dv.zz;
    Dp.xy = g * dv.xy;
    Dp.yz = g * dv.yz;
    Dp.xz = g * dv.xz;
    return Dp;
}
__device__ __forceinline__ S6 vp_stress_flow(const M9& Fe, V3 r1, V3 r2, V3 r3, float epsbar, const VPConst& P,
                                             VPOut& o) {
    return vp_stress_flow_be(aat(Fe), r1, r2, r3, epsbar, P, o);
}

// State Fp (the reference's plasticDeformationTensor): Fe = F Fp^-1, Fp += Fe^-1 Dp F (:171-206).
__device__ __forceinline__ VPOut vp_tetra(V3 p0, V3 p1, V3 p2, V3 p3, const float* D, const float* Fp, float epsbar,
                                          const VPConst& P) {
    VPOut o;
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const M9 F = deformation_gradient(r1, r2, r3, D);
    M9 Fpm;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fpm.m[i][j] = Fp[i * 3 + j];
    const M9 Fe = mul9(F, inv9(Fpm));
    const S6 Dp = vp_stress_flow(Fe, r1, r2, r3, epsbar, P, o);
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

// The same law on the shape state S = Dm^-1 Fp^-1 (9 floats per tetra, replacing both Dm^-1 and
// Fp): Fe = r S, one 3x3 product.  The reference's update Fp' = Fp + Fe^-1 Dp F equals
// Fp (I + F^-1 Dp F), so exactly
//     S' = Dm^-1 Fp'^-1 = r^-1 (I + Dp)^-1 Fe,   i.e.   S' = S - r^-1 (I + Dp)^-1 Dp Fe,
// with r^-1 = [c1; c2; c3] / det r (the cofactor rows the forces use).  The increment form is
// exactly zero when nothing flows, so S carries no rounding drift between flow events.
__device__ __forceinline__ VPOut vp_tetra_shape(V3 p0, V3 p1, V3 p2, V3 p3, const float* S, float epsbar,
                                                const VPConst& P) {
    VPOut o;
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const M9 Fe = deformation_gradient(r1, r2, r3, S);
    const S6 Dp = vp_stress_flow(Fe, r1, r2, r3, epsbar, P, o);
    // Ai = (I + Dp)^-1 (symmetric, by cofactors)
    const float axx = 1.0f + Dp.xx, ayy = 1.0f + Dp.yy, azz = 1.0f + Dp.zz;
    const float cxx = ayy * azz - Dp.yz * Dp.yz, cyy = axx * azz - Dp.xz * Dp.xz, czz = axx * ayy - Dp.xy * Dp.xy;
    const float cxy = Dp.xz * Dp.yz - Dp.xy * azz, cyz = Dp.xy * Dp.xz - axx * Dp.yz, cxz = Dp.xy * Dp.yz - ayy * Dp.xz;
    const float ia = frcp(axx * cxx + Dp.xy * cxy + Dp.xz * cxz);
    const S6 Ai{cxx * ia, cyy * ia, czz * ia, cxy * ia, cyz * ia, cxz * ia};
    // Nm = Ai Dp (symmetric in exact arithmetic: Ai and Dp commute), full 3x3 here
    M9 Nm;
    const V3 d0 = mk(Dp.xx, Dp.xy, Dp.xz), d1 = mk(Dp.xy, Dp.yy, Dp.yz), d2 = mk(Dp.xz, Dp.yz, Dp.zz);
    const V3 n0 = smul(Ai, d0), n1 = smul(Ai, d1), n2 = smul(Ai, d2);  // columns of Ai Dp
    Nm.m[0][0] = n0.x; Nm.m[1][0] = n0.y; Nm.m[2][0] = n0.z;
    Nm.m[0][1] = n1.x; Nm.m[1][1] = n1.y; Nm.m[2][1] = n1.z;
    Nm.m[0][2] = n2.x; Nm.m[1][2] = n2.y; Nm.m[2][2] = n2.z;
    const M9 W = mul9(Nm, Fe);
    const V3 c1 = cross(r2, r3), c2 = cross(r3, r1), c3 = cross(r1, r2);
    const float idet = -frcp(dot(r1, c1));
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const V3 w = mk(W.m[0][j], W.m[1][j], W.m[2][j]);
        o.Fp[0 * 3 + j] = S[0 * 3 + j] + dot(c1, w) * idet;
        o.Fp[1 * 3 + j] = S[1 * 3 + j] + dot(c2, w) * idet;
        o.Fp[2 * 3 + j] = S[2 * 3 + j] + dot(c3, w) * idet;
    }
    return o;
}

// The same law on the symmetric state G = S S^T (6 floats: xx, yy, zz, xy, yz, xz).  The elastic
// left Cauchy-Green tensor is Be = Fe Fe^T = r G r^T, so the isotropic stress, the flow and the
// forces depend on G alone (the plastic rotation that G drops enters none of them), and
// S' = r^-1 (I + Dp)^-1 r S gives exactly
//     G' = G + r^-1 (-N Be - Be N + N Be N) r^-T,   N = (I + Dp)^-1 Dp,
// an increment that is zero when nothing flows.  Dp is proportional to the deviatoric stress of an
// isotropic law, a polynomial in Be, so N commutes with Be and the bracket is (N^2 - 2N) Be.
// Fp is not recoverable from G.
__device__ __forceinline__ VPOut vp_tetra_sym(V3 p0, V3 p1, V3 p2, V3 p3, const float* G6, float epsbar,
                                              const VPConst& P) {
    VPOut o;
    const V3 r1 = p1 - p0, r2 = p2 - p0, r3 = p3 - p0;
    const S6 G{G6[0], G6[1], G6[2], G6[3], G6[4], G6[5]};
    // Be = r G r^T with the rows of r (component i of the three edges)
    const V3 qx = mk(r1.x, r2.x, r3.x), qy = mk(r1.y, r2.y, r3.y), qz = mk(r1.z, r2.z, r3.z);
    const V3 tx = smul(G, qx), ty = smul(G, qy), tz = smul(G, qz);
    const S6 Be{dot(tx, qx), dot(ty, qy), dot(tz, qz), dot(tx, qy), dot(ty, qz), dot(tx, qz)};
    const S6 Dp = vp_stress_flow_be(Be, r1, r2, r3, epsbar, P, o);
    // Ai = (I + Dp)^-1 (symmetric, by cofactors) and N = Ai Dp (symmetric: they commute)
    const float axx = 1.0f + Dp.xx, ayy = 1.0f + Dp.yy, azz = 1.0f + Dp.zz;
    const float cxx = ayy * azz - Dp.yz * Dp.yz, cyy = axx * azz - Dp.xz * Dp.xz, czz = axx * ayy - Dp.xy * Dp.xy;
    const float cxy = Dp.xz * Dp.yz - Dp.xy * azz, cyz = Dp.xy * Dp.xz - axx * Dp.yz, cxz = Dp.xy * Dp.yz - ayy * Dp.xz;
    const float ia = frcp(axx * cxx + Dp.xy * cxy + Dp.xz * cxz);
    const S6 Ai{cxx * ia, cyy * ia, czz * ia, cxy * ia, cyz * ia, cxz * ia};
    const S6 N{Ai.xx * Dp.xx + Ai.xy * Dp.xy + Ai.xz * Dp.xz, Ai.xy * Dp.xy + Ai.yy * Dp.yy + Ai.yz * Dp.yz,
               Ai.xz * Dp.xz + Ai.yz * Dp.yz + Ai.zz * Dp.zz, Ai.xx * Dp.xy + Ai.xy * Dp.yy + Ai.xz * Dp.yz,
               Ai.xy * Dp.xz + Ai.yy * Dp.yz + Ai.yz * Dp.zz, Ai.xx * Dp.xz + Ai.xy * Dp.yz + Ai.xz * Dp.zz};
    // X = (N^2 - 2N) Be (symmetric: the factors commute)
    const V3 n0 = mk(N.xx, N.xy, N.xz), n1 = mk(N.xy, N.yy, N.yz), n2 = mk(N.xz, N.yz, N.zz);
    const V3 m0 = smul(N, n0) - n0 * 2.0f, m1 = smul(N, n1) - n1 * 2.0f, m2 = smul(N, n2) - n2 * 2.0f;
    const V3 b0 = mk(Be.xx, Be.xy, Be.xz), b1 = mk(Be.xy, Be.yy, Be.yz), b2 = mk(Be.xz, Be.yz, Be.zz);
    const S6 X{dot(m0, b0), dot(m1, b1), dot(m2, b2), dot(m0, b1), dot(m1, b2), dot(m0, b2)};
    // r^-1 = [c1; c2; c3] / det r:  dG = C X C^T / det^2
    const V3 c1 = cross(r2, r3), c2 = cross(r3, r1), c3 = cross(r1, r2);
    const float idet = frcp(dot(r1, c1));
    const float id2 = idet * idet;
    const V3 u1 = smul(X, c1), u2 = smul(X, c2), u3 = smul(X, c3);
    o.Fp[0] = G.xx + dot(c1, u1) * id2;
    o.Fp[1] = G.yy + dot(c2, u2) * id2;
    o.Fp[2] = G.zz + dot(c3, u3) * id2;
    o.Fp[3] = G.xy + dot(c1, u2) * id2;
    o.Fp[4] = G.yz + dot(c2, u3) * id2;
    o.Fp[5] = G.xz + dot(c1, u3) * id2;
    return o;
}

// ---- node terms (per node, FP32 positions) -------------------------------------------------

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

// VelocityVerlet.cl:11-15 with xi = 0.005, unit mass, dt = 1
struct VV {
    static constexpr float a = 1.0f - 0.5f * 0.005f;
    static constexpr float inv_b = 1.0f / (1.0f + 0.5f * 0.005f);
    static constexpr float a_over_b = a * inv_b;
};
__device__ __forceinline__ void vv_update(V3& x, V3& v, V3 f, V3 fo) {
    v = (v * VV::a + (f + fo) * 0.5f) * VV::inv_b;
    x = x + v * VV::a + f * 0.5f;
}
// The same recurrence on (x, u), u = a v + f/2 of the previous step: v' = (u + f'/2)/b, so
// x' = x + a v' + f'/2 = x + u' with u' = (a/b)(u + f'/2) + f'/2.  No previous force is kept.
__device__ __forceinline__ void vv_update_u(V3& x, V3& u, V3 f) {
    const V3 h = f * 0.5f;
    u = (u + h) * VV::a_over_b + h;
    x = x + u;
}

// Simulation.py:173-183 / 117-132, in double
__device__ __forceinline__ double indenter_pos(double start, double speed, double move_steps, int64_t time) {
    const double t = (double)time;
    return (t <= move_steps) ? start - speed * t : start - speed * move_steps;
}

}  // namespace nfx
