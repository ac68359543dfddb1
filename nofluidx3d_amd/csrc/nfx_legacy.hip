// libnfx: the legacy contacts and integrators of noFluidX3D (the kernels that
// SimulationLegacy.py:139-330, 550-617 and 884-940 wire up), FP32 state on the device:
//   k_round_tip        Interactions/RoundTip.cl          sphere below its centre, cylinder above
//   k_adhesive_plane   InteractionAdhesivePlane.cl       Lennard-Jones wall on surface points
//   k_point_zero       PointZeroForce.cl                 pull by node 0's x and z
//   k_com_*            CalcCOM.cl, COMForce.cl           centre of mass (x, z) and its spring
//   k_over_damped      OverDamped.cl                     x += f
//   k_vv2              velocityVerlet2.cl                mass-weighted damped Verlet
//   k_volume_quad      InteractionSphereIntegral.cl, InteractionTipIntegral.cl
//                      exp(-k dis) n body force, Bower's 4-point rule per tetra
//   k_lve              InteractionLinearViscoelastic.cl   small-strain Prony-series law (node-0 base)
//   k_surface_adhesion InteractionAdhesivePlaneSurfaceIntegral.cl (Lennard-Jones area density) and
//                      InteractionHardAdhesivePlaneSurfaceIntegral.cl (exp(-k d) - A exp(-d)),
//                      3-point edge-midpoint rule per surface triangle
// Scatters into the force array use float atomics, as the reference's atomicAdd; the volume
// integrals can also write per-corner records for the deterministic CSR gather.
#include <hip/hip_runtime.h>

#include <cmath>

#include "nfx_common.h"
#include "nfx_device.h"

namespace nfx {

constexpr int kB = 256;
static inline unsigned int nb_of(int64_t n) { return (unsigned int)((n + kB - 1) / kB); }

__device__ __forceinline__ V3 load3(const float* __restrict__ x, int64_t N, int64_t n) {
    return mk(x[n], x[N + n], x[2 * N + n]);
}

__global__ __launch_bounds__(kB) void k_round_tip(float* f, const float* __restrict__ x, int64_t N, float R,
                                                  float ypos) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= N) return;
    const V3 pos = load3(x, N, p);
    V3 r = mk(pos.x, pos.y - ypos, pos.z);
    if (pos.y > ypos) r.y = 0.0f;  // cylinder above the sphere centre (RoundTip.cl:10-12)
    const float len = sqrtf(dot(r, r));
    if (len <= 0.0f) return;
    const float dis = fmaxf(len - R, 0.0f);  // as written in RoundTip.cl:14
    const V3 fc = r * (0.1f * dis / len);     // prefactor 0.1 (RoundTip.cl:15)
    f[p] += fc.x;
    f[N + p] += fc.y;
    f[2 * N + p] += fc.z;
}

__global__ __launch_bounds__(kB) void k_adhesive_plane(float* f, const float* __restrict__ x, int64_t N, float wall,
                                                       float eps, float sigma, const int32_t* __restrict__ on_surface) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= N || on_surface[p] != 1) return;
    const float d = x[N + p] - wall;
    const float s = sigma / d;
    const float s6 = s * s * s * s * s * s;
    f[N + p] += 24.0f * eps / d * (2.0f * s6 * s6 - s6);
}

__global__ __launch_bounds__(kB) void k_point_zero(float* f, const float* __restrict__ x, int64_t N, float k) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= N) return;
    f[p] += -k * x[0];
    f[2 * N + p] += -k * x[2 * N];
}

// centre of mass of x (component 0) and z (component 2), FP64 per-block partial sums, then one
// block sums them in block order: deterministic
__global__ __launch_bounds__(kB) void k_com_partial(const float* __restrict__ x, int64_t N, double* partial) {
    double sx = 0.0, sz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < N; i += (int64_t)gridDim.x * kB) {
        sx += (double)x[i];
        sz += (double)x[2 * N + i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sz += __shfl_xor(sz, off);
    }
    __shared__ double wx[kB / 64], wz[kB / 64];
    if ((threadIdx.x & 63) == 0) {
        wx[threadIdx.x >> 6] = sx;
        wz[threadIdx.x >> 6] = sz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kB / 64; ++w) {
            a += wx[w];
            b += wz[w];
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

__global__ __launch_bounds__(64) void k_com_finish(const double* __restrict__ partial, int nblocks, int64_t N,
                                                   double* com) {
    if (threadIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nblocks; ++i) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    com[0] = a / (double)N;
    com[1] = b / (double)N;
}

__global__ __launch_bounds__(kB) void k_com_force(float* f, int64_t N, float k, const double* __restrict__ com) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= N) return;
    f[p] += (float)(-k * com[0]);
    f[2 * N + p] += (float)(-k * com[1]);
}

__global__ __launch_bounds__(kB) void k_over_damped(float* x, float* f, int64_t N, int32_t point3fix) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= N) return;
    V3 fv = load3(f, N, p);
    if (point3fix && p == 3) fv.x = fv.z = 0.0f;
    x[p] += fv.x;
    x[N + p] += fv.y;
    x[2 * N + p] += fv.z;
    f[p] = f[N + p] = f[2 * N + p] = 0.0f;
}

__global__ __launch_bounds__(kB) void k_vv2(float* x, float* v, float* f, float* fo, const float* __restrict__ m,
                                            int64_t N, float damping) {
    const int64_t p = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (p >= N) return;
    const float im = 1.0f / m[p];
    const V3 a = load3(f, N, p) * im;
    V3 vv = (load3(v, N, p) + (load3(fo, N, p) + a) * 0.5f) * (1.0f / (1.0f + 0.5f * damping));
    const V3 acc = a - vv * damping;
    const V3 xx = load3(x, N, p) + vv + acc * 0.5f;
    x[p] = xx.x; x[N + p] = xx.y; x[2 * N + p] = xx.z;
    v[p] = vv.x; v[N + p] = vv.y; v[2 * N + p] = vv.z;
    fo[p] = acc.x; fo[N + p] = acc.y; fo[2 * N + p] = acc.z;
    f[p] = f[N + p] = f[2 * N + p] = 0.0f;
}

struct QuadTip {
    float radius, half_angle, sin_a, cos_a, k;
    float tx, ty, tz;
};

template <bool TIP>
__device__ __forceinline__ V3 body_force(V3 pos, const QuadTip& q) {
    if constexpr (TIP) {  // InteractionTipIntegral.cl:1-20
        const V3 rho = mk(pos.x - q.tx, 0.0f, pos.z - q.tz);
        const float lr = sqrtf(dot(rho, rho));
        if (pos.y < q.ty - q.sin_a * lr) {
            const V3 r = mk(pos.x - q.tx, pos.y - q.ty, pos.z - q.tz);
            const float len = sqrtf(dot(r, r));
            return r * (expf(-q.k * (len - q.radius)) / len);
        }
        const V3 e = rho * (1.0f / lr);
        const V3 n = mk(q.cos_a * e.x, -q.sin_a, q.cos_a * e.z);
        const float dis = dot(pos, n) - dot(mk(q.tx, q.ty, q.tz), n) - q.radius;
        return n * expf(-q.k * dis);
    } else {  // InteractionSphereIntegral.cl:1-8
        const V3 r = mk(pos.x, pos.y - q.ty, pos.z);
        const float len = sqrtf(dot(r, r));
        return r * (expf(-q.k * (len - q.radius)) / len);
    }
}

template <bool TIP, bool CORNERS>
__global__ __launch_bounds__(kB) void k_volume_quad(float* f, float* corners, const float* __restrict__ x,
                                                    const uint32_t* __restrict__ tets, int64_t N, int64_t T,
                                                    QuadTip q) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t >= T) return;
    uint32_t id[4];
    V3 p[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        id[a] = tets[a * T + t];
        p[a] = load3(x, N, id[a]);
    }
    const V3 r1 = p[0] - p[3], r2 = p[1] - p[3], r3 = p[2] - p[3];
    const float vol6 = fabsf(dot(r1, cross(r2, r3)));  // |det| as the kernels compute it
    constexpr float al = 0.58541020f, be = 0.13819669f;
    V3 F[4] = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
#pragma unroll
    for (int qi = 0; qi < 4; ++qi) {
        float Nq[4] = {be, be, be, be};
        if (qi < 3) Nq[qi] = al;
        Nq[3] = 1.0f - Nq[0] - Nq[1] - Nq[2];
        const V3 xq = p[0] * Nq[0] + p[1] * Nq[1] + p[2] * Nq[2] + p[3] * Nq[3];
        const V3 fq = body_force<TIP>(xq, q);
#pragma unroll
        for (int a = 0; a < 4; ++a) F[a] = F[a] + fq * Nq[a];
    }
    const float s = 0.25f * vol6;
    if constexpr (CORNERS) {
        float* c = corners + 12 * t;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            c[3 * a] = F[a].x * s;
            c[3 * a + 1] = F[a].y * s;
            c[3 * a + 2] = F[a].z * s;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            unsafeAtomicAdd(&f[id[a]], F[a].x * s);
            unsafeAtomicAdd(&f[N + id[a]], F[a].y * s);
            unsafeAtomicAdd(&f[2 * N + id[a]], F[a].z * s);
        }
    }
}

// HARD = false: Lennard-Jones density 24 eps/d (2 (sigma/d)^12 - (sigma/d)^6) with (eps, sigma) =
// (a, b) (InteractionAdhesivePlaneSurfaceIntegral.cl:1-4); HARD = true: exp(-k d) - A exp(-d) with
// (A, k) = (a, b) (InteractionHardAdhesivePlaneSurfaceIntegral.cl:1-4).  d = y - wall at each point.
template <bool HARD>
__global__ __launch_bounds__(kB) void k_surface_adhesion(float* f, const float* __restrict__ x,
                                                         const uint32_t* __restrict__ tri, int64_t N, int64_t NT,
                                                         float wall, float a, float b) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t >= NT) return;
    uint32_t id[3];
    V3 p[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        id[c] = tri[c * NT + t];
        p[c] = load3(x, N, id[c]);
    }
    const V3 cr = cross(p[0] - p[2], p[1] - p[2]);
    const float area = 0.5f * sqrtf(dot(cr, cr));
    const float xi[3][2] = {{0.5f, 0.0f}, {0.0f, 0.5f}, {0.5f, 0.5f}};
    float F[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int qi = 0; qi < 3; ++qi) {
        const float Nq[3] = {xi[qi][0], xi[qi][1], 1.0f - xi[qi][0] - xi[qi][1]};
        const float d = Nq[0] * p[0].y + Nq[1] * p[1].y + Nq[2] * p[2].y - wall;
        float fq;
        if constexpr (HARD) {
            fq = expf(-b * d) - a * expf(-d);
        } else {
            const float s = b / d;
            const float s6 = s * s * s * s * s * s;
            fq = 24.0f * a / d * (2.0f * s6 * s6 - s6);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) F[c] += fq * Nq[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) unsafeAtomicAdd(&f[N + id[c]], area * (1.0f / 3.0f) * F[c]);
}

// InteractionLinearViscoelastic.cl: small-strain generalized Maxwell (Prony) law, node-0 base
struct LveConst {
    float two_mu_inf, kappa, two_mu;
    int nterms;
    float w[8], decay[8];
};

template <bool CORNERS>
__global__ __launch_bounds__(kB) void k_lve(float* f, float* corners, const float* __restrict__ x,
                                            const uint32_t* __restrict__ tets, const float* __restrict__ dminv,
                                            int64_t N, int64_t T, LveConst c, float* dev_old, float* transient,
                                            float* vm, float* pr) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t >= T) return;
    const uint32_t n0 = tets[t], n1 = tets[T + t], n2 = tets[2 * T + t], n3 = tets[3 * T + t];
    float D[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) D[k] = dminv[k * T + t];
    const V3 p0 = load3(x, N, n0), r1 = load3(x, N, n1) - p0, r2 = load3(x, N, n2) - p0, r3 = load3(x, N, n3) - p0;
    const M9 F = deformation_gradient(r1, r2, r3, D);
    float e[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) e[i][k] = 0.5f * (F.m[i][k] + F.m[k][i]) - (i == k ? 1.0f : 0.0f);
    const float tr = e[0][0] + e[1][1] + e[2][2];
    float sig[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int q = 3 * i + k;
            const float dv = e[i][k] - (i == k ? tr * (1.0f / 3.0f) : 0.0f);
            const float rate = dv - dev_old[q * T + t];
            dev_old[q * T + t] = dv;
            float s = c.two_mu_inf * dv + (i == k ? c.kappa * tr : 0.0f);
            for (int j = 0; j < c.nterms; ++j) {
                float* h = &transient[(9 * (int64_t)j + q) * T + t];
                const float hn = c.decay[j] * *h + c.two_mu * c.w[j] * rate;
                *h = hn;
                s += hn;
            }
            sig[i][k] = s;
        }
    // the legacy kernel keeps sigma as computed (symmetric up to rounding); its face forces use the rows
    const S6 S{sig[0][0], sig[1][1], sig[2][2], 0.5f * (sig[0][1] + sig[1][0]), 0.5f * (sig[1][2] + sig[2][1]),
               0.5f * (sig[0][2] + sig[2][0])};
    if (vm) {
        const float t3 = (S.xx + S.yy + S.zz) * (1.0f / 3.0f);
        const float dx = S.xx - t3, dy = S.yy - t3, dz = S.zz - t3;
        vm[t] = sqrtf(1.5f * (dx * dx + dy * dy + dz * dz + 2.0f * (S.xy * S.xy + S.yz * S.yz + S.xz * S.xz)));
        pr[t] = t3;
    }
    const float detr = dot(r1, cross(r2, r3));
    const Corners fc = cofactor_forces_r(S, r1, r2, r3, detr < 0.0f ? (1.0f / 6.0f) : (-1.0f / 6.0f));
    const uint32_t id[4] = {n0, n1, n2, n3};
    const V3 fa[4] = {fc.f0, fc.f1, fc.f2, fc.f3};
    if constexpr (CORNERS) {
        float* cc = corners + 12 * t;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            cc[3 * a] = fa[a].x;
            cc[3 * a + 1] = fa[a].y;
            cc[3 * a + 2] = fa[a].z;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            unsafeAtomicAdd(&f[id[a]], fa[a].x);
            unsafeAtomicAdd(&f[N + id[a]], fa[a].y);
            unsafeAtomicAdd(&f[2 * N + id[a]], fa[a].z);
        }
    }
}

static int check_n(int64_t n, const char* what) {
    NFX_REQUIRE(n > 0 && n < (int64_t(1) << 31), "%s count %lld out of range", what, (long long)n);
    return NFX_OK;
}

}  // namespace nfx

using namespace nfx;

extern "C" {

int nfx_round_tip_force(nfx_ctx* c, float* force, const float* points, int64_t N, double radius, double sphere_pos) {
    NFX_REQUIRE(c && force && points, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    k_round_tip<<<nb_of(N), kB, 0, c->stream>>>(force, points, N, (float)radius, (float)sphere_pos);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_adhesive_plane_force(nfx_ctx* c, float* force, const float* points, int64_t N, double wall_pos,
                             double adhesion, double distance, const int32_t* on_surface) {
    NFX_REQUIRE(c && force && points && on_surface, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    const double sigma = std::pow(0.5, 1.0 / 6.0) * distance;  // InteractionAdhesivePlane.cl:6
    k_adhesive_plane<<<nb_of(N), kB, 0, c->stream>>>(force, points, N, (float)wall_pos, (float)adhesion, (float)sigma,
                                                     on_surface);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_point_zero_force(nfx_ctx* c, float* force, const float* points, int64_t N, double k) {
    NFX_REQUIRE(c && force && points, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    k_point_zero<<<nb_of(N), kB, 0, c->stream>>>(force, points, N, (float)k);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_center_of_mass(nfx_ctx* c, const float* points, int64_t N, double* com) {
    NFX_REQUIRE(c && points && com, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    const int nb = (int)std::min<int64_t>(nb_of(N), kMaxReduceBlocks);
    k_com_partial<<<nb, kB, 0, c->stream>>>(points, N, c->partials);
    NFX_CHECK_LAUNCH();
    k_com_finish<<<1, 64, 0, c->stream>>>(c->partials, nb, N, com);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_com_force(nfx_ctx* c, float* force, int64_t N, double k, const double* com) {
    NFX_REQUIRE(c && force && com, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    k_com_force<<<nb_of(N), kB, 0, c->stream>>>(force, N, (float)k, com);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_over_damped(nfx_ctx* c, float* points, float* force, int64_t N, int32_t point3fix) {
    NFX_REQUIRE(c && points && force, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    k_over_damped<<<nb_of(N), kB, 0, c->stream>>>(points, force, N, point3fix);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_velocity_verlet2(nfx_ctx* c, float* points, float* vel, float* force, float* fold, const float* masses,
                         int64_t N, double damping) {
    NFX_REQUIRE(c && points && vel && force && fold && masses, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    k_vv2<<<nb_of(N), kB, 0, c->stream>>>(points, vel, force, fold, masses, N, (float)damping);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_volume_integral_force(nfx_ctx* c, int32_t tip_shape, const float* points, const uint32_t* tetras, int64_t N,
                              int64_t T, double radius, double half_angle, double tip_x, double tip_y, double tip_z,
                              double k, float* force, float* corners) {
    NFX_REQUIRE(c && points && tetras && (force || corners), "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    if (int rc = check_n(T, "tetra")) return rc;
    QuadTip q{(float)radius, (float)half_angle, (float)std::sin(half_angle), (float)std::cos(half_angle), (float)k,
              (float)tip_x, (float)tip_y, (float)tip_z};
    const unsigned int nb = nb_of(T);
    if (tip_shape && corners) k_volume_quad<true, true><<<nb, kB, 0, c->stream>>>(force, corners, points, tetras, N, T, q);
    else if (tip_shape) k_volume_quad<true, false><<<nb, kB, 0, c->stream>>>(force, corners, points, tetras, N, T, q);
    else if (corners) k_volume_quad<false, true><<<nb, kB, 0, c->stream>>>(force, corners, points, tetras, N, T, q);
    else k_volume_quad<false, false><<<nb, kB, 0, c->stream>>>(force, corners, points, tetras, N, T, q);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_linear_viscoelastic_force(nfx_ctx* c, const nfx_viscoelastic* p, const float* points, const uint32_t* tetras,
                                  const float* dminv, int64_t N, int64_t T, float* dev_old, float* transient,
                                  float* von_mises, float* pressure, float* force, float* corners) {
    NFX_REQUIRE(c && p && points && tetras && dminv && dev_old && (force || corners), "NULL argument");
    NFX_REQUIRE(p->nterms >= 0 && p->nterms <= 8, "viscoelastic terms %d outside 0..8", p->nterms);
    NFX_REQUIRE(p->nterms == 0 || transient, "NULL transient stresses");
    NFX_REQUIRE(!von_mises == !pressure, "von_mises and pressure go together");
    if (int rc = check_n(N, "node")) return rc;
    if (int rc = check_n(T, "tetra")) return rc;
    LveConst k{};
    const double mu = p->E / (2 * (1 + p->nu));
    double winf = 1.0;
    for (int j = 0; j < p->nterms; ++j) {
        NFX_REQUIRE(p->taus[j] > 0, "relaxation time %d must be positive", j);
        winf -= p->weights[j];
        k.w[j] = (float)p->weights[j];
        k.decay[j] = (float)std::exp(-1.0 / p->taus[j]);
    }
    k.two_mu_inf = (float)(2 * mu * winf);
    k.two_mu = (float)(2 * mu);
    k.kappa = (float)(p->E / (3 * (1 - 2 * p->nu)));
    k.nterms = p->nterms;
    if (corners)
        k_lve<true><<<nb_of(T), kB, 0, c->stream>>>(force, corners, points, tetras, dminv, N, T, k, dev_old, transient,
                                                   von_mises, pressure);
    else
        k_lve<false><<<nb_of(T), kB, 0, c->stream>>>(force, corners, points, tetras, dminv, N, T, k, dev_old, transient,
                                                    von_mises, pressure);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_adhesive_plane_surface_force(nfx_ctx* c, float* force, const float* points, const uint32_t* triangles,
                                     int64_t N, int64_t NT, double wall_pos, double adhesion, double distance) {
    NFX_REQUIRE(c && force && points && triangles, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    if (int rc = check_n(NT, "triangle")) return rc;
    const double sigma = std::pow(0.5, 1.0 / 6.0) * distance;
    k_surface_adhesion<false><<<nb_of(NT), kB, 0, c->stream>>>(force, points, triangles, N, NT, (float)wall_pos,
                                                              (float)adhesion, (float)sigma);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_hard_adhesive_plane_surface_force(nfx_ctx* c, float* force, const float* points, const uint32_t* triangles,
                                          int64_t N, int64_t NT, double wall_pos, double adhesion,
                                          double force_const) {
    NFX_REQUIRE(c && force && points && triangles, "NULL argument");
    if (int rc = check_n(N, "node")) return rc;
    if (int rc = check_n(NT, "triangle")) return rc;
    k_surface_adhesion<true><<<nb_of(NT), kB, 0, c->stream>>>(force, points, triangles, N, NT, (float)wall_pos,
                                                             (float)adhesion, (float)force_const);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

}  // extern "C"
