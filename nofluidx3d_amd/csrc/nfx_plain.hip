// libnfx: plain (single-cell) path of the explicit tetrahedral-FEM step on gfx950.
//
// Passes per step (Simulation.timeStep, Simulation.py:232-237):
//   force pass  gather 4 nodes, F = I + [u1-u0|u2-u0|u3-u0] Dm^-1 from node displacements u,
//               stress, 4 corner forces, reduced to nodes by one of three scatters: float atomics
//               into the node force (the reference's atomicAdd.cl:1-11), a 12-float corner record
//               per tetra gathered through the node CSR, or per-tile LDS sums (the default).
//   node pass   the node's force (atomic sum, CSR gather with 4 lanes per node and a fixed xor
//               butterfly, or its tile partials in order); substrate + indenter terms in FP64 at
//               x = X0 + u and the device-computed indenter position; the damped velocity-Verlet
//               update of u; the FP64 indenter-force reduction.
// The per-plugin kernels (one per reference kernel) take absolute positions, as the reference's
// buffers hold them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "nfx_common.h"
#include "nfx_device.h"

using namespace nfx;

namespace nfx {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

constexpr int kBlock = 256;

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// logical blocks so that tetras sharing nodes share an L2.
__device__ __forceinline__ int64_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
    return (int64_t)x * q + (x < r ? x : r) + i;
}

struct TetArgs {
    int64_t N, T;
    const uint32_t* __restrict__ tetras;
    const float* __restrict__ dminv;
    const float* __restrict__ v0;
    const float* __restrict__ x;
    float* force;
    float* corners;
    float* vm;  // legacy laws: per-tetra von Mises stress and pressure (nullable)
    float* pr;
};

struct ElasticDev {
    float mu1, mu2, kappa;  // MR
    float mu, kappa_svk;    // SVK / SVKK (kappa = nu / (1 - 2 nu))
    LegacyConst L;          // legacy laws
};

static bool is_legacy(int model) { return model >= NFX_LINEAR_ELASTIC && model <= NFX_PORO_NEO_HOOKEAN; }

static ElasticDev elastic_consts(const nfx_plan& p) {
    const double E = p.E, nu = p.nu, ratio = p.ratio;
    const double mu = E / (2 * (1 + nu));  // tetra.cl:118-121 (FP64 on the host)
    ElasticDev e;
    e.mu1 = (float)(ratio * mu);
    e.mu2 = (float)((1 - ratio) * mu);
    e.kappa = (float)(E / (3 * (1 - 2 * nu)));
    e.mu = (float)mu;
    e.kappa_svk = (float)(nu / (1 - 2 * nu));
    // InteractionLinearElastic.cl:66-67: a = 2E/(2+2nu) = 2 mu, b = E nu/((1+nu)(1-2nu)) = lambda
    e.L.mu = (float)(p.model == NFX_LINEAR_ELASTIC ? mu : p.shear);
    e.L.lam = (float)(E * nu / ((1 + nu) * (1 - 2 * nu)));
    e.L.K = (float)p.bulk;
    e.L.mu2 = (float)p.shear2;
    e.L.phi = (float)p.volume_fraction;
    return e;
}

static VPConst vp_consts(const nfx_viscoplastic& v) {
    VPConst c;
    c.mu1 = (float)v.mu1;
    c.mu2 = (float)v.mu2;
    c.kappa = (float)v.kappa;
    c.m = (float)v.flow_exponent;
    c.Y = (float)v.yield_stress;
    c.rate = (float)v.flow_rate;
    c.eps0 = (float)v.hardening_threshold;
    c.n = (float)v.hardening_exponent;
    c.sy = (float)v.yield_offset;
    c.linear = (v.flow_exponent == 1.0 && v.hardening_exponent == 0.0) ? 1 : 0;
    return c;
}

__device__ __forceinline__ V3 gather3(const float* __restrict__ x, int64_t N, uint32_t n) {
    return mk(x[n], x[N + n], x[2 * N + n]);
}

__device__ __forceinline__ void scatter_atomic(float* f, int64_t N, uint32_t n, V3 v) {
    unsafeAtomicAdd(&f[n], v.x);
    unsafeAtomicAdd(&f[N + n], v.y);
    unsafeAtomicAdd(&f[2 * N + n], v.z);
}

__device__ __forceinline__ void store_corners(float* __restrict__ c, int64_t t, const Corners& f) {
    float* p = c + 12 * t;
    p[0] = f.f0.x; p[1] = f.f0.y; p[2] = f.f0.z;
    p[3] = f.f1.x; p[4] = f.f1.y; p[5] = f.f1.z;
    p[6] = f.f2.x; p[7] = f.f2.y; p[8] = f.f2.z;
    p[9] = f.f3.x; p[10] = f.f3.y; p[11] = f.f3.z;
}

// ---- one-time FP64 precompute -----------------------------------------------------------------
// Inputs are the reference's Tetra buffers (Tetra.py:53-58): edge vectors R1..R3 as 9 SoA FP64
// arrays and the signed volumes.  Dm = [R1 | R2 | R3] (tetra.cl:40-47) is inverted once in FP64
// (the reference re-inverts it every step, tetra.cl:53); rows of Dm^-1 are (R2xR3, R3xR1, R1xR2)/det.
// Dm^-1 of tetra t from its edge vectors R[k] (column k+1 of Dm), in FP64, stored as FP32
__device__ __forceinline__ void dminv_from_edges(const double (&R)[3][3], int64_t T, int64_t t, float* __restrict__ dminv) {
    double cr[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double* a = R[(k + 1) % 3];
        const double* b = R[(k + 2) % 3];
        cr[k][0] = a[1] * b[2] - a[2] * b[1];
        cr[k][1] = a[2] * b[0] - a[0] * b[2];
        cr[k][2] = a[0] * b[1] - a[1] * b[0];
    }
    const double det = R[0][0] * cr[0][0] + R[0][1] * cr[0][1] + R[0][2] * cr[0][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dminv[(3 * i + j) * T + t] = (float)(cr[i][j] / det);
}

__global__ __launch_bounds__(kBlock) void k_precompute_f64(const double* __restrict__ ev, const double* __restrict__ vol,
                                                           int64_t T, float* __restrict__ dminv, float* __restrict__ v0) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    double R[3][3];  // R[k] = column k+1 of Dm
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[k][c] = ev[(3 * k + c) * T + t];
    dminv_from_edges(R, T, t, dminv);
    v0[t] = (float)vol[t];
}

// A batch's geometry in simulation units from its cells' recentred positions p (FP64, micrometres)
// and per-cell scales f (RadiusSIM / |p[1].x|, TetraCell.py:15-20): the FP32 reference positions
// X0 = fl32(p f), and per tetra, from X0 in FP64, the edges R_k = X0_k - X0_0, V0 = R1 . (R2 x R3) / 6
// and Dm^-1, in the operation order of the host construction (no contraction), so that a reset
// rebuilds exactly what construction built.  Padding tetra (past tile_ntet in their tile) get the
// identity shape and V0 = 1/6.
__global__ __launch_bounds__(kBlock) void k_batch_points(const double* __restrict__ p, const double* __restrict__ f,
                                                         const uint32_t* __restrict__ node_cell, int64_t N,
                                                         float* __restrict__ x) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    const double s = f[node_cell[n]];
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c * N + n] = (float)(p[c * N + n] * s);
}

__global__ __launch_bounds__(kBlock) void k_batch_edges(const double* __restrict__ p, const double* __restrict__ f,
                                                        const uint32_t* __restrict__ node_cell,
                                                        const uint32_t* __restrict__ tets,
                                                        const uint16_t* __restrict__ tile_ntet, int32_t K, int64_t N,
                                                        int64_t T, float* __restrict__ dminv, float* __restrict__ v0) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    double R[3][3];
    double vol;
    if (t % K >= (int64_t)tile_ntet[t / K]) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) R[k][c] = (k == c) ? 1.0 : 0.0;
        vol = 1.0 / 6.0;
    } else {
        const uint32_t n0 = tets[t];
        const double s = f[node_cell[n0]];
        double x0[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) x0[c] = (double)(float)(p[c * N + n0] * s);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t nk = tets[(int64_t)(k + 1) * T + t];
#pragma unroll
            for (int c = 0; c < 3; ++c) R[k][c] = (double)(float)(p[c * N + nk] * s) - x0[c];
        }
        const double cx = R[1][1] * R[2][2] - R[1][2] * R[2][1];
        const double cy = R[1][2] * R[2][0] - R[1][0] * R[2][2];
        const double cz = R[1][0] * R[2][1] - R[1][1] * R[2][0];
        vol = (R[0][0] * cx + R[0][1] * cy + R[0][2] * cz) / 6.0;
    }
    dminv_from_edges(R, T, t, dminv);
    v0[t] = (float)vol;
}

// ---- force pass, one thread per tetra (atomic or corner-record scatter) --------------------------
// DISP: the node array holds displacements from the reference positions (fused steps); otherwise
// absolute positions (the per-plugin kernels)
template <bool DISP>
__device__ __forceinline__ M9 tetra_grad(const TetArgs& a, uint32_t n0, uint32_t n1, uint32_t n2, uint32_t n3,
                                         const float* D) {
    const V3 p0 = gather3(a.x, a.N, n0), p1 = gather3(a.x, a.N, n1), p2 = gather3(a.x, a.N, n2),
             p3 = gather3(a.x, a.N, n3);
    if constexpr (DISP) return grad_disp(p0, p1, p2, p3, D);
    else return grad_abs(p0, p1, p2, p3, D);
}

__device__ __forceinline__ void emit_corners(const TetArgs& a, int64_t t, uint32_t n0, uint32_t n1, uint32_t n2,
                                             uint32_t n3, const Corners& c, bool corners) {
    if (corners) {
        store_corners(a.corners, t, c);
    } else {
        scatter_atomic(a.force, a.N, n0, c.f0);
        scatter_atomic(a.force, a.N, n1, c.f1);
        scatter_atomic(a.force, a.N, n2, c.f2);
        scatter_atomic(a.force, a.N, n3, c.f3);
    }
}

template <int MODEL, bool CORNERS, bool DISP>
__global__ __launch_bounds__(kBlock) void k_tetra(TetArgs a, ElasticDev e) {
    const int64_t t = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (t >= a.T) return;
    const int64_t T = a.T;
    const uint32_t n0 = a.tetras[t], n1 = a.tetras[T + t], n2 = a.tetras[2 * T + t], n3 = a.tetras[3 * T + t];
    float D[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) D[k] = a.dminv[k * T + t];
    const M9 H = tetra_grad<DISP>(a, n0, n1, n2, n3, D);
    Corners c;
    if constexpr (MODEL >= NFX_LINEAR_ELASTIC) {
        float vm, pr;
        c = legacy_tetra<MODEL>(H, D, e.L, vm, pr);
        if (a.vm) {
            a.vm[t] = vm;
            a.pr[t] = pr;
        }
    } else {
        const float v0 = a.v0[t];
        if constexpr (MODEL == NFX_MR)
            c = mr_tetra(H, D, v0, e.mu1, e.mu2, e.kappa);
        else if constexpr (MODEL == NFX_SVK)
            c = svk_tetra(H, D, v0, e.mu, e.kappa_svk);
        else
            c = svkk_tetra(H, D, v0, e.mu, e.kappa_svk);
    }
    emit_corners(a, t, n0, n1, n2, n3, c, CORNERS);
}

// ---- force pass, viscoplastic -----------------------------------------------------------------
// FUSED: displacements and the plastic deviation Q = Fp^-1 - I (state); otherwise absolute
// positions and the reference's Fp (the per-plugin kernel)
template <bool CORNERS, bool STRESS, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_vp(TetArgs a, VPConst P, float* __restrict__ state, float* __restrict__ eps,
                                               float* __restrict__ vm, float* __restrict__ pr) {
    const int64_t t = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (t >= a.T) return;
    const int64_t T = a.T;
    const uint32_t n0 = a.tetras[t], n1 = a.tetras[T + t], n2 = a.tetras[2 * T + t], n3 = a.tetras[3 * T + t];
    float D[9], st[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) D[k] = a.dminv[k * T + t];
#pragma unroll
    for (int k = 0; k < 9; ++k) st[k] = state[k * T + t];
    const float e0 = eps ? eps[t] : 0.0f;
    const M9 H = tetra_grad<FUSED>(a, n0, n1, n2, n3, D);
    float sdm = 0.f;
    if constexpr (FUSED) sdm = det_dm(D);
    const VPOut o = FUSED ? vp_tetra_q(H, D, sdm, sdm, st, e0, P) : vp_tetra(H, D, st, e0, P);
#pragma unroll
    for (int k = 0; k < 9; ++k) state[k * T + t] = o.Fp[k];
    if (eps) eps[t] = o.epsbar;
    if constexpr (STRESS) {
        vm[t] = o.von_mises;
        pr[t] = o.pressure;
    }
    emit_corners(a, t, n0, n1, n2, n3, o.f, CORNERS);
}

// ---- state conversions ---------------------------------------------------------------------------
// Q = Fp^-1 - I  <->  Fp = (I + Q)^-1, per tetra (9T SoA each)
template <bool TO_Q>
__global__ __launch_bounds__(kBlock) void k_vp_convert(const float* __restrict__ in, int64_t T, float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    M9 X;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) X.m[i][j] = in[(3 * i + j) * T + t] + ((!TO_Q && i == j) ? 1.0f : 0.0f);
    const M9 Y = inv9(X);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[(3 * i + j) * T + t] = Y.m[i][j] - ((TO_Q && i == j) ? 1.0f : 0.0f);
}

// The plastic-activity flag of each tile from the state itself: 1 when any of its tetra has a
// nonzero plastic deviation or plastic strain.  One workgroup per tile of K tetra.
__global__ __launch_bounds__(kBlock) void k_vp_tile_flags(const float* __restrict__ state, int planes,
                                                          const float* __restrict__ eps, int64_t T, int K,
                                                          uint32_t* __restrict__ flags) {
    const int64_t t0 = (int64_t)blockIdx.x * K;
    int nz = 0;
    for (int i = threadIdx.x; i < K; i += kBlock) {
        const int64_t t = t0 + i;
        if (t >= T) break;
        for (int k = 0; k < planes; ++k) nz |= state[k * T + t] != 0.0f;
        if (eps) nz |= eps[t] != 0.0f;
    }
    nz = __syncthreads_or(nz);
    if (threadIdx.x == 0) flags[blockIdx.x] = nz ? 1u : 0u;
}

// x = X0 + u (TO_POINTS) or u = x - X0, component-wise over 3N
template <bool TO_POINTS>
__global__ __launch_bounds__(kBlock) void k_disp_convert(const float* __restrict__ ref, const float* __restrict__ in,
                                                         int64_t n3, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    out[i] = TO_POINTS ? ref[i] + in[i] : in[i] - ref[i];
}

// ---- node helpers shared by the node pass and the node-fused force pass -----------------------
// Sum of a node's tile partials [b, e) in order, four records' loads in flight at a time (zeros past
// the end, which leave the sum unchanged): a node has 1-8 partials, and a loop of dependent single
// loads would wait a full memory latency per partial.
__device__ __forceinline__ V3 sum_partials(const float* __restrict__ partials, uint32_t b, uint32_t e) {
    V3 f = mk(0.f, 0.f, 0.f);
    for (uint32_t j = b; j < e; j += 4) {
        V3 q[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float* p = partials + 3 * (int64_t)(j + i);
            q[i] = (j + i < e) ? mk(p[0], p[1], p[2]) : mk(0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) f = f + q[i];
    }
    return f;
}

// The same sum with the first eight records' loads issued together (one memory round for nodes of
// up to eight partials, most of them), then four at a time; the additions run in the same order
__device__ __forceinline__ V3 sum_partials8(const float* __restrict__ partials, uint32_t b, uint32_t e) {
    V3 q[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float* p = partials + 3 * (int64_t)(b + i);
        q[i] = (b + i < e) ? mk(p[0], p[1], p[2]) : mk(0.f, 0.f, 0.f);
    }
    V3 f = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 8; ++i) f = f + q[i];
    for (uint32_t j = b + 8; j < e; j += 4) {
        V3 r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float* p = partials + 3 * (int64_t)(j + i);
            r[i] = (j + i < e) ? mk(p[0], p[1], p[2]) : mk(0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) f = f + r[i];
    }
    return f;
}

// The contact terms of a node at x = X0 + u, in FP64: the substrate (Substrate.cl:4-5), the
// indenter (Sphere.cl:9-16 or Plane.cl:4-5) at the FP64 position pos, added to f; the indenter's
// own force and the penetration for the tip reduction.  X0.x/z and u.x/z are read only for nodes
// the sphere can touch (y > pos - R): a few per cent of a cell.
template <bool SUB, int IND>
__device__ __forceinline__ void node_contacts(const float* __restrict__ ref, int64_t N, int64_t n, V3 u, double sub_k,
                                              double wall, double ind_k, double ind_radius, double pos, V3& f,
                                              double& tip_f, double& tip_dis) {
    const double y = (double)ref[N + n] + (double)u.y;
    if constexpr (SUB) f.y += (float)(sub_k * fmax(wall - y, 0.0));
    if constexpr (IND == NFX_INDENTER_SPHERE) {
        if (y > pos - ind_radius) {
            const Contact c = sphere_term_d((double)ref[n] + (double)u.x, y, (double)ref[2 * N + n] + (double)u.z,
                                            ind_k, ind_radius, pos);
            f = f + mk((float)c.fx, (float)c.fy, (float)c.fz);
            tip_f = -c.fy;
            tip_dis = c.dis;
        }
    } else if constexpr (IND == NFX_INDENTER_PLANE) {
        const double dis = fmax(y - pos, 0.0);
        f.y -= (float)(ind_k * dis);
        tip_f = ind_k * dis;
        tip_dis = dis;
    }
}

// The previous step's node update riding in a node-fused force pass: every tile that touches node n
// updates it from the same inputs with the same operations (so the copies agree bitwise) and the
// node's first tile (its partial is the node's first, dst == node_off[n]) stores the result.
struct NodeStep {
    const float* __restrict__ prev_partials;  // the previous force pass's partials, node order
    const uint2* __restrict__ range;          // per tile-node, its node's partials [b, e)
    const float* __restrict__ w;     // w = a v + f_old/2, or v when vv_first
    const float* __restrict__ fold;  // vv_first: f_old
    float* __restrict__ x_next;      // the displacements after the step
    float* __restrict__ w_next;
    double* __restrict__ tip_row;  // nullable: per tile, the indenter force on the nodes it stores
    int32_t substrate, indenter, fix, vv_first;
    double sub_k, wall, ind_k, ind_radius, ind_start, ind_speed, ind_move;
    int64_t step;  // the step being completed (step_dev + step_off when step_dev is set)
    const int64_t* step_dev;
    int32_t step_off;
};

// node n's update (k_node_update's arithmetic, the substrate/indenter/fix choices uniform at run
// time): x holds u before the step on entry and after it on return; tip receives the node's indenter
// force when this tile stores the node
__device__ __forceinline__ void node_step(const NodeStep& s, const float* __restrict__ ref, int64_t N, uint32_t n,
                                          uint32_t dst, uint2 rg, double pos, V3& x, double& tip) {
    const uint32_t b = rg.x, e = rg.y;
    V3 f = sum_partials8(s.prev_partials, b, e);
    double tf = 0.0, td = 0.0;
    if (s.substrate) {
        if (s.indenter == NFX_INDENTER_SPHERE)
            node_contacts<true, NFX_INDENTER_SPHERE>(ref, N, n, x, s.sub_k, s.wall, s.ind_k, s.ind_radius, pos, f, tf, td);
        else if (s.indenter == NFX_INDENTER_PLANE)
            node_contacts<true, NFX_INDENTER_PLANE>(ref, N, n, x, s.sub_k, s.wall, s.ind_k, s.ind_radius, pos, f, tf, td);
        else
            node_contacts<true, NFX_INDENTER_NONE>(ref, N, n, x, s.sub_k, s.wall, s.ind_k, s.ind_radius, pos, f, tf, td);
    } else {
        if (s.indenter == NFX_INDENTER_SPHERE)
            node_contacts<false, NFX_INDENTER_SPHERE>(ref, N, n, x, s.sub_k, s.wall, s.ind_k, s.ind_radius, pos, f, tf, td);
        else if (s.indenter == NFX_INDENTER_PLANE)
            node_contacts<false, NFX_INDENTER_PLANE>(ref, N, n, x, s.sub_k, s.wall, s.ind_k, s.ind_radius, pos, f, tf, td);
    }
    V3 w = mk(s.w[n], s.w[N + n], s.w[2 * N + n]);
    if (s.vv_first) {
        w = vv_w_from_v(w, f, mk(s.fold[n], s.fold[N + n], s.fold[2 * N + n]));
        x = x + w;
    } else {
        vv_update_w(x, w, f);
    }
    if (s.fix && (n == 3 || n == 4)) {  // x = z = 0 (VelocityVerlet.cl:17-22): u = -X0
        x.x = -ref[n];
        x.z = -ref[2 * N + n];
    }
    if (dst == b) {
        s.x_next[n] = x.x; s.x_next[N + n] = x.y; s.x_next[2 * N + n] = x.z;
        s.w_next[n] = w.x; s.w_next[N + n] = w.y; s.w_next[2 * N + n] = w.z;
        tip = tf;
    }
}

// ---- force pass, tiled deterministic scatter ---------------------------------------------------
// One workgroup per tile of K consecutive (Morton-ordered) tetra, one thread per tetra.  The
// tile's distinct nodes are listed on the host (tile_nodes, ascending, MX <= 256 entries per tile)
// with, for each, the end of its run in the tile's corner list sorted by (node, slot) (tile_lend)
// and the position of its partial sum in node order (tile_dst).  Each corner slot 4*(t-t0)+a
// carries its tile-local node and its rank in that node's run (tile_corners), so its place in the
// sorted list is run start + rank.  Phases:
//   0  issue the tetra's streaming loads (Dm^-1 and V0 unless GEO, the viscoplastic state and epsbar),
//      its corner records and the tile metadata
//   1  node displacements (GEO: and reference positions) and run starts -> LDS
//   2  (GEO: Dm^-1 and V0 from the reference edges) material law on LDS displacements; each corner
//      force -> LDS at its place in the sorted list
//   3  one thread per tile-node: sum its contiguous run -> partials[tile_dst]
// Viscoplastic state: a tetra writes its state back only when it flows this step (otherwise the
// update is the identity and memory already holds it), and a tile whose vp_active flag is 0 reads
// none (its tetra have never flowed: Q = 0, epsbar = 0, which memory also holds).  Far from the
// contact, below the yield offset, a tile costs its corner records and node data only.
// GEO trades the 36 + 4 B per tetra of streamed Dm^-1 and V0 for 12 B per tile-node of reference
// positions (~4 B per tetra) and ~60 FLOPs per tetra.
// The node pass sums each node's contiguous partials in tile order: no atomics, fixed order.
struct TileArgs {
    int64_t N, T;
    const float* __restrict__ dminv;
    const float* __restrict__ v0;
    const float* __restrict__ x;    // node displacements u
    const float* __restrict__ ref;  // reference positions X0
    const uint32_t* __restrict__ tile_count;
    const uint32_t* __restrict__ tile_nodes;
    const uint16_t* __restrict__ tile_lend;
    const uint16_t* __restrict__ tile_corners;
    const uint32_t* __restrict__ tile_dst;
    float* __restrict__ partials;
    int32_t max_nodes;
    // viscoplastic, nullable: per tile, 0 = every tetra of the tile has a zero plastic state (its
    // state is neither read nor written until one of them flows; the tile then sets its flag)
    uint32_t* __restrict__ vp_active;
    // batches: cell-aligned tiles with per-cell material constants (NULL for a single cell)
    const uint16_t* __restrict__ tile_ntet;
    const uint32_t* __restrict__ tile_cell;
    const nfx_cell* __restrict__ cells;
    // batches (NARROW): tile-node ids and partial positions as 16-bit offsets from per-tile bases
    const uint16_t* __restrict__ tile_nodes16;
    const uint16_t* __restrict__ tile_dst16;
    const uint32_t* __restrict__ tile_node_base;
    const uint32_t* __restrict__ tile_dst_base;
    // batches: the step's curve samples ride in the same launch, one workgroup per cell appended
    // after the tiles (curve_blocks = 0: none); they read the positions the tiles read
    int32_t curve_blocks;
    int32_t curve_ind;
    const uint32_t* __restrict__ cell_node_off;
    double* curves;
    int32_t ncurve;
    int64_t curve_steps, curve_step;
    const int64_t* curve_step_dev;
    int32_t curve_off;
    // NODE: the previous step's node update (x holds the displacements before it)
    NodeStep ns;
};

// One cell's sample of the batched curves, by one workgroup of any size (<= 512): on the first
// step of each of the ncurve sample intervals, the indenter force on the cell's nodes at the
// current positions, summed in FP64 in a fixed order (strided per-thread sums, then lanes and
// waves in order).  Every exit before the barrier is uniform over the workgroup.
__device__ void batch_curve_sample(const float* __restrict__ x, const float* __restrict__ ref, int64_t N,
                                   const uint32_t* __restrict__ cell_node_off, const nfx_cell* __restrict__ cells,
                                   double* curves, int32_t ncurve, int64_t curve_steps, int64_t time, int cell,
                                   int ind) {
    if (time < 0 || time >= curve_steps) return;
    const int64_t k = time * ncurve / curve_steps;
    if (time > 0 && (time - 1) * ncurve / curve_steps == k) return;  // not the first step of sample k
    const nfx_cell cp = cells[cell];
    const double pos = indenter_pos(cp.tip_start, cp.tip_speed, cp.tip_move_steps, time);
    double s = 0.0;
    for (int64_t n = cell_node_off[cell] + threadIdx.x; n < cell_node_off[cell + 1]; n += blockDim.x) {
        const double y = (double)ref[N + n] + (double)x[N + n];
        if (ind == NFX_INDENTER_SPHERE) {
            if (y > pos - (double)cp.tip_radius)
                s -= sphere_term_d((double)ref[n] + (double)x[n], y, (double)ref[2 * N + n] + (double)x[2 * N + n],
                                   cp.tip_k, cp.tip_radius, pos).fy;
        } else {
            s += (double)cp.tip_k * fmax(y - pos, 0.0);
        }
    }
    __shared__ double ws[8];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += ws[w];
        curves[(int64_t)cell * ncurve + k] = t;
    }
}

// LDS of a tile: corner forces 12K floats, then per tile-node its displacement and X0.x (16 B), with
// GEO X0.y and X0.z (8 B), and its run start (u16)
template <int K>
__host__ __device__ constexpr size_t tile_lds_bytes_fixed() {
    return 4 * 12 * K;
}
__host__ __device__ constexpr size_t tile_lds_bytes_per_node(bool geo) { return 16 + (geo ? 8 : 0) + 2; }

// GEO: Dm^-1 and V0 from the staged reference positions; NT: cache policy of the streams; EPS: the
// viscoplastic law keeps its accumulated plastic strain (the linear law does not read it); SYM: its
// state is the symmetric deviation Gd (6 planes) instead of Q (9 planes); NARROW: 16-bit tile-node
// metadata (batches); NODE: the previous step's node update rides in this pass.  Waves per SIMD: the
// viscoplastic forms fit 72 VGPRs (7 waves; measured 5 % faster than 6, while the 64-register bound
// of 8 waves spills and is slower) except with stress outputs or NODE (6), and the symmetric one
// without epsbar fits 64 (8 waves)
template <int MODEL, int K, bool GEO, bool STRESS, bool NT = false, bool EPS = true, bool SYM = false,
          bool NARROW = false, bool NODE = false>
__global__ __launch_bounds__(K, MODEL != NFX_VISCOPLASTIC ? 8 : (STRESS || NODE) ? 6 : (!SYM || EPS) ? 7 : 8) void k_tile_force(TileArgs a, ElasticDev e, VPConst P, float* __restrict__ Sst,
                                                  float* __restrict__ eps, float* __restrict__ vm,
                                                  float* __restrict__ pr) {
    // one thread per tetra of the tile (blockDim.x == K)
    extern __shared__ __align__(16) unsigned char smem[];
    float* s_cf = reinterpret_cast<float*>(smem);  // [component][4K], sorted-corner order
    const int MX = a.max_nodes;
    // per tile-node, node-major: (ux, uy, uz, X0x) as one 16-byte LDS word and, with GEO, (X0y, X0z)
    // as one 8-byte word, so a corner's node data is one or two LDS reads at one address
    float4* s_n = reinterpret_cast<float4*>(s_cf + 12 * K);
    __shared__ int s_flow;                         // viscoplastic: a tetra of the tile flowed this step
    __shared__ double s_tip[NODE ? K / 64 : 1];    // NODE: per wave, the indenter force of the previous step
    float2* s_r = reinterpret_cast<float2*>(s_n + MX);
    uint16_t* s_lb = reinterpret_cast<uint16_t*>(s_r + (GEO ? MX : 0));

    const uint32_t ntile_blocks = gridDim.x - (uint32_t)a.curve_blocks;
    if (blockIdx.x >= ntile_blocks) {  // the step's curve sample of cell blockIdx.x - ntile_blocks
        const int64_t time = a.curve_step_dev ? (*a.curve_step_dev + a.curve_off) : a.curve_step;
        batch_curve_sample(a.x, a.ref, a.N, a.cell_node_off, a.cells, a.curves, a.ncurve, a.curve_steps, time,
                           (int)(blockIdx.x - ntile_blocks), a.curve_ind);
        return;
    }
    const int64_t c = xcd_block(blockIdx.x, ntile_blocks);
    const int64_t T = a.T, N = a.N;
    const int64_t t0 = c * K;
    int nt = (int)((T - t0) < K ? (T - t0) : K);
    if (a.cells) {  // batch: this tile's cell and its constants (uniform over the workgroup)
        nt = a.tile_ntet[c];
        const nfx_cell cp = a.cells[a.tile_cell[c]];
        e.mu1 = cp.mu1;
        e.mu2 = cp.mu2;
        e.kappa = cp.kappa;
        P.mu1 = cp.mu1;
        P.mu2 = cp.mu2;
        P.kappa = cp.kappa;
        P.rate = cp.vp_rate;
        P.Y = cp.vp_yield;
        P.sy = cp.vp_offset;
        P.linear = 1;
    }
    const int tl = threadIdx.x;
    const bool active = tl < nt;
    const uint32_t u = (uint32_t)tl;
    // the tile's plastic-activity flag (uniform): without it every tile reads its state
    bool plastic = true;
    if constexpr (MODEL == NFX_VISCOPLASTIC) {
        if (a.vp_active) plastic = a.vp_active[c] != 0u;
    }

    // issue this tetra's streaming loads first: they overlap the staging phases below
    float D[9], Sv[MODEL == NFX_VISCOPLASTIC ? 9 : 1], e0 = 0.f, v0 = 0.f, sdm = 0.f;
    if (active) {
        if constexpr (!GEO) {
#pragma unroll
            for (int k = 0; k < 9; ++k) D[k] = ld_g<NT, float>(uat(a.dminv + t0 + k * T, u));
        }
        if constexpr (MODEL == NFX_VISCOPLASTIC) {
            if (plastic) {
#pragma unroll
                for (int k = 0; k < (SYM ? 6 : 9); ++k)
                    Sv[k] = ld_g<NT, float>(uat((const float*)(Sst + t0 + k * T), u));
                if constexpr (EPS) e0 = ld_g<NT, float>(uat((const float*)(eps + t0), u));
            } else {
#pragma unroll
                for (int k = 0; k < (SYM ? 6 : 9); ++k) Sv[k] = 0.0f;
            }
        } else if constexpr (MODEL < NFX_VISCOPLASTIC && !GEO) {
            v0 = ld_g<NT, float>(uat(a.v0 + t0, u));
        }
    }

    // tile metadata: thread tl owns tile-node tl (MX <= K, checked on the host)
    const int64_t j0 = c * (int64_t)MX;
    const int M = (int)a.tile_count[c];
    const bool owner = tl < MX;
    uint32_t nid = 0, dst = 0;
    int lb = 0, le = 0;
    uint2 rg = make_uint2(0u, 0u);  // NODE: the node's partials of the previous force pass
    if (owner) {
        if constexpr (NODE) rg = a.ns.range[j0 + tl];
        if constexpr (NARROW) {
            nid = a.tile_node_base[c] + (uint32_t)ld_g<NT, uint16_t>(uat(a.tile_nodes16 + j0, u));
            dst = a.tile_dst_base[c] + (uint32_t)ld_g<NT, uint16_t>(uat(a.tile_dst16 + j0, u));
        } else {
            nid = ld_g<NT, uint32_t>(uat(a.tile_nodes + j0, u));
            dst = ld_g<NT, uint32_t>(uat(a.tile_dst + j0, u));
        }
        lb = tl ? ld_g<NT, uint16_t>(uat(a.tile_lend + j0, u - 1)) : 0;
        le = ld_g<NT, uint16_t>(uat(a.tile_lend + j0, u));
    }
    // this tetra's 4 corner records (tile-node | rank << 8), as one 8-byte load
    uint64_t cr = 0;
    if (active) cr = ld_g<NT, uint64_t>(uat(reinterpret_cast<const uint64_t*>(a.tile_corners + 4 * K * c), u));
    double tip = 0.0;  // NODE: the indenter force on the node this thread stores
    if (owner) {
        float ux = a.x[nid], uy = a.x[N + nid], uz = a.x[2 * N + nid];
        if (NODE && tl < M) {  // complete the previous step at this node (padding entries past M are
                               // not nodes of the tile); the forces below see the result
            double pos = 0.0;
            if (a.ns.indenter) {
                const int64_t time = a.ns.step_dev ? (*a.ns.step_dev + a.ns.step_off) : a.ns.step;
                pos = indenter_pos(a.ns.ind_start, a.ns.ind_speed, a.ns.ind_move, time);
            }
            V3 xn = mk(ux, uy, uz);
            node_step(a.ns, a.ref, N, nid, dst, rg, pos, xn, tip);
            ux = xn.x;
            uy = xn.y;
            uz = xn.z;
        }
        if constexpr (GEO) {
            s_n[tl] = make_float4(ux, uy, uz, a.ref[nid]);
            s_r[tl] = make_float2(a.ref[N + nid], a.ref[2 * N + nid]);
        } else {
            s_n[tl] = make_float4(ux, uy, uz, 0.0f);
        }
        s_lb[tl] = (uint16_t)lb;
    }
    if (MODEL == NFX_VISCOPLASTIC && tl == 0) s_flow = 0;
    if constexpr (NODE) {
        if (a.ns.tip_row) {  // the tile's share of the indenter force: lanes by a fixed butterfly, then waves
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) tip += __shfl_xor(tip, off);
            if ((tl & 63) == 0) s_tip[tl >> 6] = tip;
        }
    }
    __syncthreads();
    if constexpr (NODE) {
        if (a.ns.tip_row && tl == 0) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < K / 64; ++w) t += s_tip[w];
            a.ns.tip_row[c] = t;
        }
    }
    int flows = 0;
    if (active) {
        const int l0 = (int)(cr & 0xFF), l1 = (int)((cr >> 16) & 0xFF), l2 = (int)((cr >> 32) & 0xFF),
                  l3 = (int)((cr >> 48) & 0xFF);
        // ROWS: D holds the unscaled rows G = det(Dm) Dm^-1 (the laws whose forces are linear in
        // V0 Dm^-1 and whose cofactors are cof(F) g_k); the legacy laws take Dm^-1 itself
        constexpr bool ROWS = GEO && MODEL <= NFX_VISCOPLASTIC;
        if constexpr (GEO) {
            const float4 n0 = s_n[l0], n1 = s_n[l1], n2 = s_n[l2], n3 = s_n[l3];
            const float2 m0 = s_r[l0], m1 = s_r[l1], m2 = s_r[l2], m3 = s_r[l3];
            const V3 r0 = mk(n0.w, m0.x, m0.y), r1 = mk(n1.w, m1.x, m1.y), r2 = mk(n2.w, m2.x, m2.y),
                     r3 = mk(n3.w, m3.x, m3.y);
            if constexpr (ROWS) {
                sdm = shape_rows(r1 - r0, r2 - r0, r3 - r0, D);
            } else {
                sdm = dminv_from_ref(r1 - r0, r2 - r0, r3 - r0, D);
                v0 = sdm * (1.0f / 6.0f);
            }
        } else if constexpr (MODEL == NFX_VISCOPLASTIC) {
            sdm = det_dm(D);
        }
        const float4 w0 = s_n[l0], w1 = s_n[l1], w2 = s_n[l2], w3 = s_n[l3];
        const V3 u0 = mk(w0.x, w0.y, w0.z), u1 = mk(w1.x, w1.y, w1.z), u2 = mk(w2.x, w2.y, w2.z),
                 u3 = mk(w3.x, w3.y, w3.z);
        M9 H = grad_disp(u0, u1, u2, u3, D);
        if constexpr (ROWS) {
            const float id = frcp(sdm);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) H.m[i][j] *= id;
        }
        // with ROWS, V0 Dm^-1 = G / 6 and the edge cofactors take G with scale 1
        const float vs = ROWS ? (1.0f / 6.0f) : v0;
        const float gs = ROWS ? 1.0f : sdm;
        Corners f;
        if constexpr (MODEL == NFX_VISCOPLASTIC) {
            VPOut o;
            if constexpr (SYM) o = vp_tetra_gd(H, D, gs, sdm, Sv, e0, P);
            else o = vp_tetra_q(H, D, gs, sdm, Sv, e0, P, !plastic);
            if (o.flows) {
                flows = 1;
#pragma unroll
                for (int k = 0; k < (SYM ? 6 : 9); ++k) st_g<NT, float>(uat(Sst + t0 + k * T, u), o.Fp[k]);
                if constexpr (EPS) st_g<NT, float>(uat(eps + t0, u), o.epsbar);
            }
            if constexpr (STRESS) {
                st_s<NT>(vm + t0 + u, o.von_mises);
                st_s<NT>(pr + t0 + u, o.pressure);
            }
            f = o.f;
        } else if constexpr (MODEL >= NFX_LINEAR_ELASTIC) {
            float svm, spr;
            f = legacy_tetra<MODEL>(H, D, e.L, svm, spr);
            if constexpr (STRESS) {
                st_s<NT>(vm + t0 + u, svm);
                st_s<NT>(pr + t0 + u, spr);
            }
        } else if constexpr (MODEL == NFX_MR) {
            f = mr_tetra(H, D, vs, e.mu1, e.mu2, e.kappa);
        } else if constexpr (MODEL == NFX_SVK) {
            f = svk_tetra(H, D, vs, e.mu, e.kappa_svk);
        } else {
            f = svkk_tetra(H, D, vs, e.mu, e.kappa_svk);
        }
        // each corner's place in the tile's (node, slot)-sorted corner list: run start + rank
        const int q0 = s_lb[l0] + (int)((cr >> 8) & 0xFF), q1 = s_lb[l1] + (int)((cr >> 24) & 0xFF);
        const int q2 = s_lb[l2] + (int)((cr >> 40) & 0xFF), q3 = s_lb[l3] + (int)((cr >> 56) & 0xFF);
        s_cf[q0] = f.f0.x; s_cf[4 * K + q0] = f.f0.y; s_cf[8 * K + q0] = f.f0.z;
        s_cf[q1] = f.f1.x; s_cf[4 * K + q1] = f.f1.y; s_cf[8 * K + q1] = f.f1.z;
        s_cf[q2] = f.f2.x; s_cf[4 * K + q2] = f.f2.y; s_cf[8 * K + q2] = f.f2.z;
        s_cf[q3] = f.f3.x; s_cf[4 * K + q3] = f.f3.y; s_cf[8 * K + q3] = f.f3.z;
    }
    if (MODEL == NFX_VISCOPLASTIC && flows) s_flow = 1;  // any flowing lane; all write the same value
    __syncthreads();
    if (MODEL == NFX_VISCOPLASTIC && a.vp_active && !plastic && tl == 0 && s_flow) a.vp_active[c] = 1u;  // sticky
    if (tl < M) {
        float fx = 0.f, fy = 0.f, fz = 0.f;
        int q = lb;
        for (; q + 4 <= le; q += 4) {  // the run in slot order, four corners per trip
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                fx += s_cf[q + i];
                fy += s_cf[4 * K + q + i];
                fz += s_cf[8 * K + q + i];
            }
        }
        for (; q < le; ++q) {
            fx += s_cf[q];
            fy += s_cf[4 * K + q];
            fz += s_cf[8 * K + q];
        }
        float* o = a.partials + 3 * (int64_t)dst;
        o[0] = fx;
        o[1] = fy;
        o[2] = fz;
    }
}

// ---- deterministic block reduction (sum, max) finished by the last block --------------------
__device__ __forceinline__ void wave_sum_max(double& s, double& m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        m = fmax(m, __shfl_xor(m, off));
    }
}

__device__ void reduce_finish(double s, double m, double* partials, unsigned int* counter, double* out_sum,
                              double* out_max) {
    __shared__ double ss[kBlock / 64], sm[kBlock / 64];
    __shared__ bool last;
    wave_sum_max(s, m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        ss[wid] = s;
        sm[wid] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
            a += ss[w];
            b = fmax(b, sm[w]);
        }
        partials[2 * blockIdx.x] = a;
        partials[2 * blockIdx.x + 1] = b;
        __threadfence();
        const unsigned int prev = atomicAdd(counter, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a = 0.0, b = 0.0;
    for (unsigned int i = threadIdx.x; i < gridDim.x; i += kBlock) {
        a += __hip_atomic_load(&partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b = fmax(b, __hip_atomic_load(&partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    wave_sum_max(a, b);
    __syncthreads();
    if (lane == 0) {
        ss[wid] = a;
        sm[wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
            x += ss[w];
            y = fmax(y, sm[w]);
        }
        if (out_sum) *out_sum = x;
        if (out_max) *out_max = y;
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- node pass ------------------------------------------------------------------------------
struct NodeArgs {
    int64_t N;
    float* x;  // displacements u (fused node pass) or absolute positions (per-plugin kernels)
    const float* __restrict__ ref;
    float* v;
    float* xo;  // where the node pass stores x and v (x and v themselves, or the other buffers of a
    float* vo;  // node-fused call's last step)
    float* force;
    float* fold;
    const uint32_t* __restrict__ off;
    const uint32_t* __restrict__ slots;
    const float* __restrict__ corners;
    const uint32_t* __restrict__ node_off;
    const float* __restrict__ partials;
    double sub_k, wall;
    int32_t indenter;
    double ind_k, ind_radius;
    double ind_start, ind_speed, ind_move;
};

// Sum of node n's corner records over the G lanes of its group, combined by a fixed butterfly.
template <int G>
__device__ __forceinline__ V3 gather_node(const NodeArgs& a, int64_t n, int sub) {
    V3 f = mk(0.f, 0.f, 0.f);
    const uint32_t b = a.off[n], e = a.off[n + 1];
    for (uint32_t j = b + sub; j < e; j += G) {
        const uint32_t s = a.slots[j];
        const float* p = a.corners + 12 * (int64_t)(s >> 2) + 3 * (s & 3);
        f = f + mk(p[0], p[1], p[2]);
    }
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        f.x += __shfl_xor(f.x, o);
        f.y += __shfl_xor(f.y, o);
        f.z += __shfl_xor(f.z, o);
    }
    return f;
}

// GATHER: 0 = read the atomically accumulated force (and clear it), 1 = corner records via the
// node CSR (4 lanes per node), 2 = tile partials via node_parts (one lane per node).
// VVM, the velocity-Verlet state this step reads and writes (the reference's (v, f_old) between
// calls, the (u, w) form of the batch inside a call, w = a v + f_old/2, which needs no previous force
// and so moves 24 B per node less):
//   VV_PLAIN   reads v, f_old and writes v, f_old (a one-step call)
//   VV_FIRST   reads v, f_old and writes w (into v's buffer)
//   VV_MIDDLE  reads w and writes w
//   VV_LAST    reads w and writes v, f_old
enum { VV_PLAIN = 0, VV_FIRST = 1, VV_MIDDLE = 2, VV_LAST = 3 };
template <int GATHER, bool SUB, int IND, bool FIX, bool TIP, int VVM = VV_PLAIN>
__global__ __launch_bounds__(kBlock) void k_node_update(NodeArgs a, int64_t step, const int64_t* step_dev, int32_t step_off,
                                                        double* partials, unsigned int* counter, double* tip_sum,
                                                        double* tip_max) {
    constexpr bool CSR = GATHER == 1;
    constexpr int G = CSR ? 4 : 1;
    const int64_t time = step_dev ? (*step_dev + step_off) : step;
    const double pos = IND ? indenter_pos(a.ind_start, a.ind_speed, a.ind_move, time) : 0.0;
    const int64_t N = a.N;
    const int sub = threadIdx.x % G;
    double tsum = 0.0, tmax = 0.0;
    const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
    // XCD-aware block order, as the tile force pass has it: an XCD updates the contiguous node range
    // whose partials its own tiles wrote (nodes and tetra follow one Morton order); -2.5 % per launch
    for (int64_t n = (xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) / G; n < N; n += stride) {
        V3 f;
        if constexpr (GATHER == 1) {
            f = gather_node<G>(a, n, sub);
        } else if constexpr (GATHER == 2) {
            f = sum_partials8(a.partials, a.node_off[n], a.node_off[n + 1]);
        } else {
            f = mk(a.force[n], a.force[N + n], a.force[2 * N + n]);
        }
        if (sub != 0) continue;
        V3 x = mk(a.x[n], a.x[N + n], a.x[2 * N + n]);
        double tf = 0.0, td = 0.0;
        node_contacts<SUB, IND>(a.ref, N, n, x, a.sub_k, a.wall, a.ind_k, a.ind_radius, pos, f, tf, td);
        if constexpr (TIP) {
            tsum += tf;
            tmax = fmax(tmax, td);
        }
        V3 v = mk(a.v[n], a.v[N + n], a.v[2 * N + n]);  // v, or w inside a call
        if constexpr (VVM == VV_PLAIN) {
            vv_update(x, v, f, mk(a.fold[n], a.fold[N + n], a.fold[2 * N + n]));
        } else if constexpr (VVM == VV_FIRST) {
            v = vv_w_from_v(v, f, mk(a.fold[n], a.fold[N + n], a.fold[2 * N + n]));
            x = x + v;
        } else if constexpr (VVM == VV_MIDDLE) {
            vv_update_w(x, v, f);
        } else {
            vv_v_from_w(x, v, f);
        }
        if constexpr (FIX) {
            if (n == 3 || n == 4) {  // x = z = 0 (VelocityVerlet.cl:17-22): u = -X0
                x.x = -a.ref[n];
                x.z = -a.ref[2 * N + n];
            }
        }
        a.xo[n] = x.x; a.xo[N + n] = x.y; a.xo[2 * N + n] = x.z;
        a.vo[n] = v.x; a.vo[N + n] = v.y; a.vo[2 * N + n] = v.z;
        if constexpr (VVM == VV_PLAIN || VVM == VV_LAST) {
            a.fold[n] = f.x; a.fold[N + n] = f.y; a.fold[2 * N + n] = f.z;
        }
        if constexpr (GATHER == 0) {
            a.force[n] = 0.0f; a.force[N + n] = 0.0f; a.force[2 * N + n] = 0.0f;
        }
    }
    if constexpr (TIP) reduce_finish(tsum, tmax, partials, counter, tip_sum, tip_max);
}

// ---- standalone per-plugin node kernels --------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_gather(NodeArgs a, float* __restrict__ force) {
    constexpr int G = 4;
    const int sub = threadIdx.x % G;
    const int64_t N = a.N;
    const int64_t n = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G;
    if (n >= N) return;
    const V3 f = gather_node<G>(a, n, sub);
    if (sub == 0) {
        force[n] += f.x;
        force[N + n] += f.y;
        force[2 * N + n] += f.z;
    }
}

__global__ __launch_bounds__(kBlock) void k_sphere(float* f, const float* __restrict__ x, int64_t N, float k, float R,
                                                   float pos) {
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    const V3 fs = sphere_term(mk(x[n], x[N + n], x[2 * N + n]), k, R, pos);
    f[n] += fs.x;
    f[N + n] += fs.y;
    f[2 * N + n] += fs.z;
}

__global__ __launch_bounds__(kBlock) void k_substrate(float* f, const float* __restrict__ x, int64_t N, float k, float w) {
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    f[N + n] += substrate_term(x[N + n], k, w);
}

__global__ __launch_bounds__(kBlock) void k_plane(float* f, const float* __restrict__ x, int64_t N, float k, float w) {
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    f[N + n] += plane_term(x[N + n], k, w);
}

__global__ __launch_bounds__(kBlock) void k_vv(float* x, float* v, float* f, float* fo, int64_t N, int32_t fix) {
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    V3 xx = mk(x[n], x[N + n], x[2 * N + n]);
    V3 vv = mk(v[n], v[N + n], v[2 * N + n]);
    const V3 ff = mk(f[n], f[N + n], f[2 * N + n]);
    const V3 oo = mk(fo[n], fo[N + n], fo[2 * N + n]);
    vv_update(xx, vv, ff, oo);
    if (fix && (n == 3 || n == 4)) {
        xx.x = 0.0f;
        xx.z = 0.0f;
    }
    x[n] = xx.x; x[N + n] = xx.y; x[2 * N + n] = xx.z;
    v[n] = vv.x; v[N + n] = vv.y; v[2 * N + n] = vv.z;
    fo[n] = ff.x; fo[N + n] = ff.y; fo[2 * N + n] = ff.z;
    f[n] = 0.0f; f[N + n] = 0.0f; f[2 * N + n] = 0.0f;
}

// The indenter force and maximum penetration at the current positions, in FP64 (Simulation.py:
// 198-212 evaluates them on the host in FP64).  REF: x holds displacements from ref; otherwise
// absolute positions.
template <bool REF>
__global__ __launch_bounds__(kBlock) void k_tip(const float* __restrict__ x, const float* __restrict__ ref, int64_t N,
                                                int32_t ind, double k, double R, double pos, double* partials,
                                                unsigned int* counter, double* out_sum, double* out_max) {
    double s = 0.0, m = 0.0;
    for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < N; n += (int64_t)gridDim.x * kBlock) {
        double px = (double)x[n], py = (double)x[N + n], pz = (double)x[2 * N + n];
        if constexpr (REF) {
            px += (double)ref[n];
            py += (double)ref[N + n];
            pz += (double)ref[2 * N + n];
        }
        if (ind == NFX_INDENTER_SPHERE) {
            const Contact c = sphere_term_d(px, py, pz, k, R, pos);
            s -= c.fy;
            m = fmax(m, c.dis);
        } else {
            const double dis = fmax(py - pos, 0.0);
            s += k * dis;
            m = fmax(m, dis);
        }
    }
    reduce_finish(s, m, partials, counter, out_sum, out_max);
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

// A node-fused step's indenter force: row blockIdx.x of rows (one FP64 value per tile) summed in a
// fixed order (strided per-thread sums, lanes by a fixed butterfly, waves in order) into out.
__global__ __launch_bounds__(kBlock) void k_tip_rows(const double* __restrict__ rows, int64_t ntiles,
                                                     double* __restrict__ out) {
    const double* r = rows + (int64_t)blockIdx.x * ntiles;
    double t = 0.0;
    for (int64_t i = threadIdx.x; i < ntiles; i += kBlock) t += r[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    __shared__ double ws[kBlock / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) x += ws[w];
        out[blockIdx.x] = x;
    }
}

// ---- scratch of the allocation-free tip-force reduction (nfx_tip_force_noalloc) ----------------
// The reduction's per-block FP64 (sum, max) partials, kMaxReduceBlocks pairs, then its completion
// counter in the first word of the 32-double tail: 32.25 KiB.  Each GPU the library runs on holds
// its own copy of this array until the process ends.  In host code the name is not a device
// address: host code uses only the per-device address hipGetSymbolAddress returns (noalloc_scratch).
constexpr int kNoAllocPartials = 2 * kMaxReduceBlocks;
static __device__ __align__(256) double scratch[kNoAllocPartials + 32];

__global__ void k_clear_u32(unsigned int* p) { *p = 0u; }

// STREAM-style copy, the measured HBM ceiling: one 16-byte element per thread, nontemporal loads
// and stores (the fastest of the copy forms measured on MI355X: grid-stride loops and unrolled
// forms were slower)
typedef float nfx_f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void k_stream_copy(const nfx_f4* __restrict__ a, nfx_f4* __restrict__ b,
                                                        int64_t n4) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n4) st_nt(b + i, ld_nt(a + i));
}

// ---- host helpers -----------------------------------------------------------------------------
static inline unsigned int nblocks(int64_t n, int per = kBlock) { return (unsigned int)((n + per - 1) / per); }

static int check_sizes(int64_t N, int64_t T) {
    NFX_REQUIRE(N > 0 && N < (int64_t(1) << 31), "node count %lld out of range", (long long)N);
    NFX_REQUIRE(T > 0 && T < (int64_t(1) << 30), "tetra count %lld out of range", (long long)T);
    return NFX_OK;
}

static TetArgs tet_args(const nfx_plan& p) {
    return TetArgs{p.N, p.T, p.tetras, p.dminv, p.v0, p.disp, p.force, p.corners, nullptr, nullptr};
}

static NodeArgs node_args(const nfx_plan& p) {
    NodeArgs a;
    a.N = p.N;
    a.x = p.disp;
    a.ref = p.ref;
    a.v = p.vel;
    a.xo = p.disp;
    a.vo = p.vel;
    a.force = p.force;
    a.fold = p.fold;
    a.off = p.csr_offsets;
    a.slots = p.csr_slots;
    a.corners = p.corners;
    a.node_off = p.node_off;
    a.partials = p.partials;
    a.sub_k = p.substrate_k;
    a.wall = p.wall;
    a.indenter = p.indenter;
    a.ind_k = p.ind_k;
    a.ind_radius = p.ind_radius;
    a.ind_start = p.ind_start;
    a.ind_speed = p.ind_speed;
    a.ind_move = p.ind_move_steps;
    return a;
}

static int validate_plan(const nfx_plan* p) {
    NFX_REQUIRE(p != nullptr, "plan is NULL");
    int rc = check_sizes(p->N, p->T);
    if (rc) return rc;
    NFX_REQUIRE(p->tetras && p->dminv && p->ref && p->disp && p->vel && p->fold, "plan: NULL mesh/state pointer");
    NFX_REQUIRE(p->model >= NFX_SVKK && p->model <= NFX_PORO_NEO_HOOKEAN, "plan: unknown model %d", p->model);
    NFX_REQUIRE(p->model > NFX_MR || p->v0, "plan: NULL v0");
    NFX_REQUIRE(p->model != NFX_VISCOPLASTIC || p->epsbar ||
                    (p->scatter == NFX_SCATTER_TILED && p->vp.flow_exponent == 1.0 && p->vp.hardening_exponent == 0.0),
                "plan: viscoplastic state epsbar missing (it may be NULL only for the linear law on the tiled path)");
    NFX_REQUIRE(p->model != NFX_VISCOPLASTIC || p->vp_state, "plan: viscoplastic state vp_state missing");
    NFX_REQUIRE(!p->vp_sym || (p->model == NFX_VISCOPLASTIC && p->scatter == NFX_SCATTER_TILED),
                "plan: the symmetric state (vp_sym) is for the tiled viscoplastic force only");
    NFX_REQUIRE(!p->vp_active || (p->model == NFX_VISCOPLASTIC && p->scatter == NFX_SCATTER_TILED),
                "plan: the plastic-activity flags (vp_active) are for the tiled viscoplastic force only");
    NFX_REQUIRE(p->scatter >= NFX_SCATTER_ATOMIC && p->scatter <= NFX_SCATTER_TILED, "plan: unknown scatter %d",
                p->scatter);
    if (p->scatter == NFX_SCATTER_TILED) {
        NFX_REQUIRE(p->tile_k == 128 || p->tile_k == 256 || p->tile_k == 512, "plan: tile_k must be 128, 256 or 512");
        NFX_REQUIRE(p->tile_count && p->tile_nodes && p->tile_lend && p->tile_corners && p->tile_dst && p->node_off &&
                        p->partials,
                    "plan: tiled scatter needs the tile arrays");
        NFX_REQUIRE(p->tile_max_nodes > 0 && p->tile_max_nodes <= p->tile_k && p->tile_max_nodes <= 256,
                    "plan: tile_max_nodes %d exceeds tile_k %d or 256 (use another tile size or scatter mode)",
                    p->tile_max_nodes, p->tile_k);
        const size_t lds = (size_t)p->tile_k * 48 + tile_lds_bytes_per_node(p->shape_from_ref != 0) * (size_t)p->tile_max_nodes;
        NFX_REQUIRE(lds <= 65536, "plan: tile needs %zu B of LDS (max 65536); use a smaller tile_k", lds);
    }
    NFX_REQUIRE(p->scatter != NFX_SCATTER_ATOMIC || p->force, "plan: atomic scatter needs force");
    NFX_REQUIRE(p->scatter != NFX_SCATTER_CSR || (p->csr_offsets && p->csr_slots && p->corners),
                "plan: CSR scatter needs offsets, slots and corners");
    NFX_REQUIRE(p->indenter >= 0 && p->indenter <= 2, "plan: unknown indenter %d", p->indenter);
    const int nb = (p->disp_b != nullptr) + (p->vel_b != nullptr) + (p->partials_b != nullptr) +
                   (p->tile_range != nullptr);
    NFX_REQUIRE(nb == 0 || nb == 4,
                "plan: the node-fused arrays disp_b, vel_b, partials_b and tile_range are set together");
    NFX_REQUIRE(nb == 0 || (p->scatter == NFX_SCATTER_TILED && p->model >= NFX_SVKK && p->model <= NFX_VISCOPLASTIC),
                "plan: node-fused steps need the tiled scatter and model NFX_SVKK, NFX_SVK, NFX_MR or NFX_VISCOPLASTIC");
    NFX_REQUIRE(nb == 0 || (p->disp_b != p->disp && p->vel_b != p->vel && p->partials_b != p->partials),
                "plan: the node-fused buffers must differ from disp, vel and partials");
    return NFX_OK;
}

template <int MODEL, int K, bool GEO>
static void launch_tile_legacy(unsigned int nb, size_t lds, hipStream_t s, const TileArgs& a, const ElasticDev& e,
                               const VPConst& c, float* vm, float* pr) {
    if (vm && pr)
        k_tile_force<MODEL, K, GEO, true><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr, vm, pr);
    else
        k_tile_force<MODEL, K, GEO, false><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr, nullptr, nullptr);
}

template <int K, bool GEO, bool STRESS, bool NT, bool NODE>
static void launch_tile_vp(unsigned int nb, size_t lds, hipStream_t s, const TileArgs& a, const ElasticDev& e,
                           const VPConst& c, const nfx_plan& p) {
    float* vm = STRESS ? p.von_mises : nullptr;
    float* pr = STRESS ? p.pressure : nullptr;
    if constexpr (!STRESS && !NODE) {  // the batch's form: symmetric state with 16-bit tile-node metadata
        if (p.vp_sym && a.tile_nodes16) {
            if (p.epsbar)
                k_tile_force<NFX_VISCOPLASTIC, K, GEO, false, NT, true, true, true><<<nb, K, lds, s>>>(a, e, c, p.vp_state,
                                                                                                p.epsbar, vm, pr);
            else
                k_tile_force<NFX_VISCOPLASTIC, K, GEO, false, NT, false, true, true><<<nb, K, lds, s>>>(a, e, c, p.vp_state,
                                                                                                 nullptr, vm, pr);
            return;
        }
    }
    constexpr int VP = NFX_VISCOPLASTIC;
    if (p.vp_sym) {
        if (p.epsbar)
            k_tile_force<VP, K, GEO, STRESS, NT, true, true, false, NODE><<<nb, K, lds, s>>>(a, e, c, p.vp_state,
                                                                                       p.epsbar, vm, pr);
        else
            k_tile_force<VP, K, GEO, STRESS, NT, false, true, false, NODE><<<nb, K, lds, s>>>(a, e, c, p.vp_state,
                                                                                        nullptr, vm, pr);
    } else if (p.epsbar) {
        k_tile_force<VP, K, GEO, STRESS, NT, true, false, false, NODE><<<nb, K, lds, s>>>(a, e, c, p.vp_state, p.epsbar,
                                                                                    vm, pr);
    } else {
        k_tile_force<VP, K, GEO, STRESS, NT, false, false, false, NODE><<<nb, K, lds, s>>>(a, e, c, p.vp_state, nullptr,
                                                                                     vm, pr);
    }
}

// the buffers of one node-fused launch (launch k >= 1 of a call completes step k-1 at the nodes
// and runs the force pass of step k)
struct FusedIO {
    float* x;       // displacements before the node step (read)
    float* x_next;  // ... after it
    float* w;       // w, or v on the call's first launch
    float* w_next;
    float* prev_partials;
    float* partials;
    double* tip_row;  // nullable
    int32_t vv_first;
    int64_t step;
    const int64_t* step_dev;
    int32_t step_off;
};

template <int K, bool GEO, bool NODE = false>
static void launch_tiled_k(nfx_ctx* ctx, const nfx_plan& p, bool st, const FusedIO* io = nullptr) {
    // ctx->batch is set only while a batched step is being enqueued
    TileArgs a{p.N,          p.T,         p.dminv,        p.v0,       p.disp,     p.ref,
               p.tile_count, p.tile_nodes, p.tile_lend,    p.tile_corners, p.tile_dst, p.partials,
               p.tile_max_nodes};
    if (p.model == NFX_VISCOPLASTIC) a.vp_active = p.vp_active;
    if constexpr (NODE) {
        a.x = io->x;
        a.partials = io->partials;
        NodeStep& n = a.ns;
        n.prev_partials = io->prev_partials;
        n.range = reinterpret_cast<const uint2*>(p.tile_range);
        n.w = io->w;
        n.fold = p.fold;
        n.x_next = io->x_next;
        n.w_next = io->w_next;
        n.tip_row = io->tip_row;
        n.substrate = p.substrate;
        n.indenter = p.indenter;
        n.fix = p.fix_top_bottom;
        n.vv_first = io->vv_first;
        n.sub_k = p.substrate_k;
        n.wall = p.wall;
        n.ind_k = p.ind_k;
        n.ind_radius = p.ind_radius;
        n.ind_start = p.ind_start;
        n.ind_speed = p.ind_speed;
        n.ind_move = p.ind_move_steps;
        n.step = io->step;
        n.step_dev = io->step_dev;
        n.step_off = io->step_off;
    }
    if (ctx->batch) {
        a.tile_ntet = ctx->batch->tile_ntet;
        a.tile_cell = ctx->batch->tile_cell;
        a.cells = ctx->batch->cells;
        if (ctx->batch->tile_nodes16) {
            a.tile_nodes16 = ctx->batch->tile_nodes16;
            a.tile_dst16 = ctx->batch->tile_dst16;
            a.tile_node_base = ctx->batch->tile_node_base;
            a.tile_dst_base = ctx->batch->tile_dst_base;
        }
        if (ctx->curve_fused) {
            a.curve_blocks = ctx->batch->ncells;
            a.curve_ind = p.indenter;
            a.cell_node_off = ctx->batch->cell_node_off;
            a.curves = ctx->batch->curves;
            a.ncurve = ctx->batch->ncurve;
            a.curve_steps = ctx->batch->curve_steps;
            a.curve_step = ctx->curve_step;
            a.curve_step_dev = ctx->curve_step_dev;
            a.curve_off = ctx->curve_off;
        }
    }
    const unsigned int nb = (unsigned int)((p.T + K - 1) / K) + (unsigned int)a.curve_blocks;
    const size_t lds = tile_lds_bytes_fixed<K>() + tile_lds_bytes_per_node(GEO) * (size_t)p.tile_max_nodes;
    const ElasticDev e = elastic_consts(p);
    const VPConst c = vp_consts(p.vp);
    hipStream_t s = ctx->stream;
    float* vm = st ? p.von_mises : nullptr;
    float* pr = st ? p.pressure : nullptr;
    const bool nt = p.stream_nt != 0;
    if constexpr (NODE) {  // the single-cell laws of a node-fused step (validate_plan admits no other)
        constexpr bool N_ = true;
        switch (p.model) {
            case NFX_MR:
                if (nt)
                    k_tile_force<NFX_MR, K, GEO, false, true, true, false, false, N_><<<nb, K, lds, s>>>(a, e, c, nullptr,
                                                                                             nullptr, nullptr, nullptr);
                else
                    k_tile_force<NFX_MR, K, GEO, false, false, true, false, false, N_><<<nb, K, lds, s>>>(a, e, c, nullptr,
                                                                                              nullptr, nullptr, nullptr);
                break;
            case NFX_SVK:
                k_tile_force<NFX_SVK, K, GEO, false, false, true, false, false, N_><<<nb, K, lds, s>>>(a, e, c, nullptr,
                                                                                           nullptr, nullptr, nullptr);
                break;
            case NFX_SVKK:
                k_tile_force<NFX_SVKK, K, GEO, false, false, true, false, false, N_><<<nb, K, lds, s>>>(a, e, c, nullptr,
                                                                                            nullptr, nullptr, nullptr);
                break;
            default:
                if (st && nt) launch_tile_vp<K, GEO, true, true, true>(nb, lds, s, a, e, c, p);
                else if (st) launch_tile_vp<K, GEO, true, false, true>(nb, lds, s, a, e, c, p);
                else if (nt) launch_tile_vp<K, GEO, false, true, true>(nb, lds, s, a, e, c, p);
                else launch_tile_vp<K, GEO, false, false, true>(nb, lds, s, a, e, c, p);
        }
        return;
    }
    switch (p.model) {
        case NFX_MR:
            if (a.tile_nodes16 && nt)
                k_tile_force<NFX_MR, K, GEO, false, true, true, false, true><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr,
                                                                                         nullptr, nullptr);
            else if (a.tile_nodes16)
                k_tile_force<NFX_MR, K, GEO, false, false, true, false, true><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr,
                                                                                          nullptr, nullptr);
            else if (nt)
                k_tile_force<NFX_MR, K, GEO, false, true><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr, nullptr, nullptr);
            else k_tile_force<NFX_MR, K, GEO, false><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr, nullptr, nullptr);
            break;
        case NFX_SVK: k_tile_force<NFX_SVK, K, GEO, false><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr, nullptr, nullptr); break;
        case NFX_SVKK: k_tile_force<NFX_SVKK, K, GEO, false><<<nb, K, lds, s>>>(a, e, c, nullptr, nullptr, nullptr, nullptr); break;
        case NFX_LINEAR_ELASTIC: launch_tile_legacy<NFX_LINEAR_ELASTIC, K, GEO>(nb, lds, s, a, e, c, vm, pr); break;
        case NFX_LINEAR_DEVIATORIC: launch_tile_legacy<NFX_LINEAR_DEVIATORIC, K, GEO>(nb, lds, s, a, e, c, vm, pr); break;
        case NFX_NEO_HOOKEAN_2: launch_tile_legacy<NFX_NEO_HOOKEAN_2, K, GEO>(nb, lds, s, a, e, c, vm, pr); break;
        case NFX_PORO_NEO_HOOKEAN: launch_tile_legacy<NFX_PORO_NEO_HOOKEAN, K, GEO>(nb, lds, s, a, e, c, vm, pr); break;
        default:
            if (st && nt) launch_tile_vp<K, GEO, true, true, false>(nb, lds, s, a, e, c, p);
            else if (st) launch_tile_vp<K, GEO, true, false, false>(nb, lds, s, a, e, c, p);
            else if (nt) launch_tile_vp<K, GEO, false, true, false>(nb, lds, s, a, e, c, p);
            else launch_tile_vp<K, GEO, false, false, false>(nb, lds, s, a, e, c, p);
    }
}

// One thread per tetra with the atomic (csr false) or corner-record scatter.  DISP: a.x holds
// displacements and a viscoplastic state is Q (the fused steps); otherwise absolute positions and Fp
// (the per-plugin kernels).  a.vm/a.pr select the legacy laws' stress outputs.
template <bool DISP>
static void launch_tetra_pass(hipStream_t s, const TetArgs& a, int model, bool csr, const ElasticDev& e,
                              const VPConst& c, float* state, float* eps, float* vm, float* pr) {
    const unsigned int nb = nblocks(a.T);
    if (model == NFX_VISCOPLASTIC) {
        const bool st = vm && pr;
        if (csr && st) k_vp<true, true, DISP><<<nb, kBlock, 0, s>>>(a, c, state, eps, vm, pr);
        else if (csr) k_vp<true, false, DISP><<<nb, kBlock, 0, s>>>(a, c, state, eps, nullptr, nullptr);
        else if (st) k_vp<false, true, DISP><<<nb, kBlock, 0, s>>>(a, c, state, eps, vm, pr);
        else k_vp<false, false, DISP><<<nb, kBlock, 0, s>>>(a, c, state, eps, nullptr, nullptr);
        return;
    }
    switch (model * 2 + (csr ? 1 : 0)) {
        case NFX_MR * 2: k_tetra<NFX_MR, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_MR * 2 + 1: k_tetra<NFX_MR, true, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_SVK * 2: k_tetra<NFX_SVK, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_SVK * 2 + 1: k_tetra<NFX_SVK, true, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_SVKK * 2: k_tetra<NFX_SVKK, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_SVKK * 2 + 1: k_tetra<NFX_SVKK, true, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_LINEAR_ELASTIC * 2: k_tetra<NFX_LINEAR_ELASTIC, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_LINEAR_ELASTIC * 2 + 1: k_tetra<NFX_LINEAR_ELASTIC, true, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_LINEAR_DEVIATORIC * 2: k_tetra<NFX_LINEAR_DEVIATORIC, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_LINEAR_DEVIATORIC * 2 + 1:
            k_tetra<NFX_LINEAR_DEVIATORIC, true, DISP><<<nb, kBlock, 0, s>>>(a, e);
            break;
        case NFX_NEO_HOOKEAN_2 * 2: k_tetra<NFX_NEO_HOOKEAN_2, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_NEO_HOOKEAN_2 * 2 + 1: k_tetra<NFX_NEO_HOOKEAN_2, true, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        case NFX_PORO_NEO_HOOKEAN * 2: k_tetra<NFX_PORO_NEO_HOOKEAN, false, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
        default: k_tetra<NFX_PORO_NEO_HOOKEAN, true, DISP><<<nb, kBlock, 0, s>>>(a, e); break;
    }
}

// the fused step's force pass (displacement state)
static int launch_force_pass(nfx_ctx* ctx, const nfx_plan& p, bool stress) {
    const bool st = stress && p.von_mises && p.pressure;
    if (p.scatter == NFX_SCATTER_TILED) {
        const bool geo = p.shape_from_ref != 0;
        if (p.tile_k == 128) geo ? launch_tiled_k<128, true>(ctx, p, st) : launch_tiled_k<128, false>(ctx, p, st);
        else if (p.tile_k == 256) geo ? launch_tiled_k<256, true>(ctx, p, st) : launch_tiled_k<256, false>(ctx, p, st);
        else geo ? launch_tiled_k<512, true>(ctx, p, st) : launch_tiled_k<512, false>(ctx, p, st);
        NFX_CHECK_LAUNCH();
        return NFX_OK;
    }
    TetArgs a = tet_args(p);
    const bool legacy_st = is_legacy(p.model) && st;
    if (legacy_st) {
        a.vm = p.von_mises;
        a.pr = p.pressure;
    }
    launch_tetra_pass<true>(ctx->stream, a, p.model, p.scatter == NFX_SCATTER_CSR, elastic_consts(p),
                            vp_consts(p.vp), p.vp_state, p.epsbar, st ? p.von_mises : nullptr,
                            st ? p.pressure : nullptr);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

static bool node_fused(const nfx_plan& p) { return p.disp_b != nullptr; }  // validate_plan checked the rest

// launch k (k >= 1) of a node-fused call starting at step0: it completes step step0 + k - 1 at the
// nodes and runs the force pass of step step0 + k.  Launch k reads the buffers of parity (k - 1) & 1 and
// writes those of parity k & 1, where parity 0 is the plan's (disp, vel, partials) and 1 the second
// buffers; the call's first force pass (k = 0) writes the plan's partials.
static FusedIO fused_io(const nfx_plan& p, int64_t k, double* tip_row, int64_t step, const int64_t* step_dev,
                        int32_t step_off) {
    float* X[2] = {p.disp, p.disp_b};
    float* W[2] = {p.vel, p.vel_b};
    float* P[2] = {p.partials, p.partials_b};
    const int r = (int)((k - 1) & 1), w = (int)(k & 1);
    FusedIO io;
    io.x = X[r];
    io.x_next = X[w];
    io.w = W[r];
    io.w_next = W[w];
    io.prev_partials = P[r];
    io.partials = P[w];
    io.tip_row = tip_row;
    io.vv_first = k == 1;
    io.step = step;
    io.step_dev = step_dev;
    io.step_off = step_off;
    return io;
}

static int launch_fused(nfx_ctx* ctx, const nfx_plan& p, bool stress, const FusedIO& io) {
    const bool st = stress && p.von_mises && p.pressure;
    const bool geo = p.shape_from_ref != 0;
    if (p.tile_k == 128) geo ? launch_tiled_k<128, true, true>(ctx, p, st, &io) : launch_tiled_k<128, false, true>(ctx, p, st, &io);
    else if (p.tile_k == 256) geo ? launch_tiled_k<256, true, true>(ctx, p, st, &io) : launch_tiled_k<256, false, true>(ctx, p, st, &io);
    else geo ? launch_tiled_k<512, true, true>(ctx, p, st, &io) : launch_tiled_k<512, false, true>(ctx, p, st, &io);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

template <int CSR, bool SUB, int IND, bool FIX, int VVM>
static void launch_node_m(nfx_ctx* ctx, const NodeArgs& a, int64_t step, const int64_t* step_dev, int32_t off,
                          double* tip_sum, double* tip_max) {
    constexpr int G = CSR == 1 ? 4 : 1;
    const bool tip = (tip_sum || tip_max) && IND != NFX_INDENTER_NONE;
    unsigned int nb = nblocks(a.N * G);
    if (tip) {
        nb = std::min<unsigned int>(nb, kMaxReduceBlocks);
        k_node_update<CSR, SUB, IND, FIX, (IND != NFX_INDENTER_NONE), VVM><<<nb, kBlock, 0, ctx->stream>>>(
            a, step, step_dev, off, ctx->partials, ctx->counter, tip_sum, tip_max);
    } else {
        k_node_update<CSR, SUB, IND, FIX, false, VVM><<<nb, kBlock, 0, ctx->stream>>>(a, step, step_dev, off, nullptr,
                                                                                     nullptr, nullptr, nullptr);
    }
}

template <int CSR, bool SUB, int IND, bool FIX>
static void launch_node_t(nfx_ctx* ctx, const NodeArgs& a, int64_t step, const int64_t* step_dev, int32_t off,
                          double* tip_sum, double* tip_max, int vvm) {
    switch (vvm) {
        case VV_FIRST: launch_node_m<CSR, SUB, IND, FIX, VV_FIRST>(ctx, a, step, step_dev, off, tip_sum, tip_max); break;
        case VV_MIDDLE: launch_node_m<CSR, SUB, IND, FIX, VV_MIDDLE>(ctx, a, step, step_dev, off, tip_sum, tip_max); break;
        case VV_LAST: launch_node_m<CSR, SUB, IND, FIX, VV_LAST>(ctx, a, step, step_dev, off, tip_sum, tip_max); break;
        default: launch_node_m<CSR, SUB, IND, FIX, VV_PLAIN>(ctx, a, step, step_dev, off, tip_sum, tip_max);
    }
}

template <int CSR, bool SUB>
static void launch_node_ind(nfx_ctx* ctx, const nfx_plan& p, const NodeArgs& a, int64_t step, const int64_t* sd,
                            int32_t off, double* ts, double* tm, int vvm) {
    const bool fix = p.fix_top_bottom != 0;
    switch (p.indenter * 2 + (fix ? 1 : 0)) {
        case 0: launch_node_t<CSR, SUB, 0, false>(ctx, a, step, sd, off, ts, tm, vvm); break;
        case 1: launch_node_t<CSR, SUB, 0, true>(ctx, a, step, sd, off, ts, tm, vvm); break;
        case 2: launch_node_t<CSR, SUB, 1, false>(ctx, a, step, sd, off, ts, tm, vvm); break;
        case 3: launch_node_t<CSR, SUB, 1, true>(ctx, a, step, sd, off, ts, tm, vvm); break;
        case 4: launch_node_t<CSR, SUB, 2, false>(ctx, a, step, sd, off, ts, tm, vvm); break;
        default: launch_node_t<CSR, SUB, 2, true>(ctx, a, step, sd, off, ts, tm, vvm); break;
    }
}

// the buffers of a node pass that reads and stores different ones (a node-fused call's last step)
struct NodeIO {
    float* x;
    float* xo;
    float* v;
    float* vo;
    const float* partials;
};

// the node pass ending a node-fused call of n steps: step n-1 from the buffers of parity (n - 1) & 1,
// stored into the plan's disp and vel (and fold)
static NodeIO fused_last_io(const nfx_plan& p, int64_t n) {
    const int r = (int)((n - 1) & 1);
    return NodeIO{r ? p.disp_b : p.disp, p.disp, r ? p.vel_b : p.vel, p.vel, r ? p.partials_b : p.partials};
}

// vvm: VV_PLAIN for a one-step call; a call of n >= 2 steps runs VV_FIRST, VV_MIDDLE x (n - 2), VV_LAST
static int launch_node_pass(nfx_ctx* ctx, const nfx_plan& p, int64_t step, const int64_t* step_dev, int32_t off,
                            double* tip_sum, double* tip_max, int vvm = VV_PLAIN, const NodeIO* io = nullptr) {
    NodeArgs a = node_args(p);
    if (io) {
        a.x = io->x;
        a.xo = io->xo;
        a.v = io->v;
        a.vo = io->vo;
        a.partials = io->partials;
    }
    const bool sub = p.substrate != 0;
    switch (p.scatter * 2 + (sub ? 1 : 0)) {
        case 0: launch_node_ind<0, false>(ctx, p, a, step, step_dev, off, tip_sum, tip_max, vvm); break;
        case 1: launch_node_ind<0, true>(ctx, p, a, step, step_dev, off, tip_sum, tip_max, vvm); break;
        case 2: launch_node_ind<1, false>(ctx, p, a, step, step_dev, off, tip_sum, tip_max, vvm); break;
        case 3: launch_node_ind<1, true>(ctx, p, a, step, step_dev, off, tip_sum, tip_max, vvm); break;
        case 4: launch_node_ind<2, false>(ctx, p, a, step, step_dev, off, tip_sum, tip_max, vvm); break;
        default: launch_node_ind<2, true>(ctx, p, a, step, step_dev, off, tip_sum, tip_max, vvm); break;
    }
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

// ---- batched node pass and curve sampling ----------------------------------------------------
struct BatchNodeArgs {
    int64_t N;
    float* x;  // displacements u
    const float* __restrict__ ref;
    float* w;  // velocity-Verlet state w = a v + f/2
    const uint32_t* __restrict__ node_off;
    const float* __restrict__ partials;
    const uint32_t* __restrict__ node_cell;
    const nfx_cell* __restrict__ cells;
};

// Node pass of a batch: each node applies its own cell's substrate and indenter (position from
// the step index with the cell's law, contacts in FP64 at X0 + u), then the damped velocity-Verlet
// update in the (u, w) form.
template <bool SUB, int IND>
__global__ __launch_bounds__(kBlock) void k_batch_node(BatchNodeArgs a, int64_t step, const int64_t* step_dev,
                                                       int32_t step_off) {
    const int64_t time = step_dev ? (*step_dev + step_off) : step;
    const int64_t N = a.N;
    const int64_t n = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;  // XCD-aware, as k_node_update
    if (n >= N) return;
    V3 f = sum_partials8(a.partials, a.node_off[n], a.node_off[n + 1]);
    const nfx_cell* cp = a.cells + a.node_cell[n];
    V3 x = mk(a.x[n], a.x[N + n], a.x[2 * N + n]);
    const double pos = IND ? indenter_pos(cp->tip_start, cp->tip_speed, cp->tip_move_steps, time) : 0.0;
    double tf = 0.0, td = 0.0;
    node_contacts<SUB, IND>(a.ref, N, n, x, cp->substrate_k, cp->wall, cp->tip_k, cp->tip_radius, pos, f, tf, td);
    V3 w = mk(a.w[n], a.w[N + n], a.w[2 * N + n]);
    vv_update_w(x, w, f);
    a.x[n] = x.x; a.x[N + n] = x.y; a.x[2 * N + n] = x.z;
    a.w[n] = w.x; a.w[N + n] = w.y; a.w[2 * N + n] = w.z;
}

// One workgroup per cell: on a sampling step, the cell's indenter force (Simulation.py:198-212 in
// simulation units) from the positions at the start of the step, summed in FP64 in a fixed order.
template <int IND>
__global__ __launch_bounds__(kBlock) void k_batch_curves(const float* __restrict__ x, const float* __restrict__ ref,
                                                         int64_t N, const uint32_t* __restrict__ cell_node_off,
                                                         const nfx_cell* __restrict__ cells, double* curves,
                                                         int32_t ncurve, int64_t curve_steps, int64_t step,
                                                         const int64_t* step_dev, int32_t step_off) {
    const int64_t time = step_dev ? (*step_dev + step_off) : step;
    batch_curve_sample(x, ref, N, cell_node_off, cells, curves, ncurve, curve_steps, time, blockIdx.x, IND);
}

static int validate_batch(const nfx_plan* p, const nfx_batch* b) {
    NFX_REQUIRE(b != nullptr, "batch is NULL");
    NFX_REQUIRE(p->scatter == NFX_SCATTER_TILED, "batched stepping needs the tiled scatter");
    NFX_REQUIRE(p->model == NFX_MR || p->model == NFX_VISCOPLASTIC, "batched model must be MR or viscoplastic");
    NFX_REQUIRE(b->ncells > 0 && b->tile_ntet && b->tile_cell && b->node_cell && b->cell_node_off && b->cells,
                "batch: NULL table");
    NFX_REQUIRE(!b->curves || (b->ncurve > 0 && b->curve_steps > 0 && p->indenter != NFX_INDENTER_NONE),
                "batch: curves need ncurve, curve_steps and an indenter");
    const int narrow = (b->tile_nodes16 != nullptr) + (b->tile_dst16 != nullptr) + (b->tile_node_base != nullptr) +
                       (b->tile_dst_base != nullptr);
    NFX_REQUIRE(narrow == 0 || narrow == 4, "batch: the 16-bit tile arrays come all four or none");
    NFX_REQUIRE(p->T % p->tile_k == 0, "batch: padded tetra count must be a multiple of tile_k");
    return NFX_OK;
}

template <bool SUB>
static void launch_batch_node_sub(nfx_ctx* ctx, const nfx_plan& p, const BatchNodeArgs& a, int64_t step,
                                  const int64_t* sd, int32_t off) {
    const unsigned int nb = nblocks(p.N);
    hipStream_t s = ctx->stream;
    if (p.indenter == NFX_INDENTER_SPHERE) k_batch_node<SUB, 1><<<nb, kBlock, 0, s>>>(a, step, sd, off);
    else if (p.indenter == NFX_INDENTER_PLANE) k_batch_node<SUB, 2><<<nb, kBlock, 0, s>>>(a, step, sd, off);
    else k_batch_node<SUB, 0><<<nb, kBlock, 0, s>>>(a, step, sd, off);
}

static int launch_batch_step(nfx_ctx* ctx, const nfx_plan& p, const nfx_batch& b, int64_t step, const int64_t* sd,
                             int32_t off, int which = 7) {
    // a whole step enqueues the curve sample inside the force pass's launch (it reads the same
    // positions, before the node pass moves them); the curve pass alone keeps its own kernel
    const bool fuse = (which & 3) == 3 && b.curves;
    if ((which & 1) && b.curves && !fuse) {
        if (p.indenter == NFX_INDENTER_SPHERE)
            k_batch_curves<1><<<b.ncells, kBlock, 0, ctx->stream>>>(p.disp, p.ref, p.N, b.cell_node_off, b.cells,
                                                                    b.curves, b.ncurve, b.curve_steps, step, sd, off);
        else
            k_batch_curves<2><<<b.ncells, kBlock, 0, ctx->stream>>>(p.disp, p.ref, p.N, b.cell_node_off, b.cells,
                                                                    b.curves, b.ncurve, b.curve_steps, step, sd, off);
        NFX_CHECK_LAUNCH();
    }
    if (which & 2) {
        ctx->batch = &b;
        ctx->curve_fused = fuse;
        ctx->curve_step = step;
        ctx->curve_step_dev = sd;
        ctx->curve_off = off;
        const int rc = launch_force_pass(ctx, p, false);
        ctx->batch = nullptr;
        ctx->curve_fused = false;
        if (rc) return rc;
    }
    if (which & 4) {
        BatchNodeArgs a{p.N, p.disp, p.ref, p.vel, p.node_off, p.partials, b.node_cell, b.cells};
        if (p.substrate) launch_batch_node_sub<true>(ctx, p, a, step, sd, off);
        else launch_batch_node_sub<false>(ctx, p, a, step, sd, off);
        NFX_CHECK_LAUNCH();
    }
    return NFX_OK;
}

// ---- captured step blocks --------------------------------------------------------------------
// A context keeps the graphs of the last kMaxGraphs (plan, batch, history mode) combinations it
// stepped, so that callers alternating between several batches (cell groups) or plans replay their
// own graphs instead of capturing again.
static void drop_graphs(nfx_ctx* ctx) {
    for (auto& g : ctx->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    ctx->graphs.clear();
}

static nfx_ctx::GraphEntry* find_graph(nfx_ctx* ctx, const nfx_plan& p, const nfx_batch* b, bool hist) {
    for (auto& g : ctx->graphs)
        if (g.is_batch == (b != nullptr) && g.hist == hist && std::memcmp(&g.plan, &p, sizeof(nfx_plan)) == 0 &&
            (!b || std::memcmp(&g.batch, b, sizeof(nfx_batch)) == 0)) {
            g.last_use = ++ctx->graph_clock;
            return &g;
        }
    return nullptr;
}

// capture kGraphBlock steps of the plan (b == NULL) or of the batch; steps read the device step
// counter ctx->step_dev, so a replay is valid at any step offset
static int capture_graph(nfx_ctx* ctx, const nfx_plan& p, const nfx_batch* b, bool hist, nfx_ctx::GraphEntry** out) {
    if ((int)ctx->graphs.size() >= kMaxGraphs) {  // evict the least recently used
        size_t old = 0;
        for (size_t i = 1; i < ctx->graphs.size(); ++i)
            if (ctx->graphs[i].last_use < ctx->graphs[old].last_use) old = i;
        if (ctx->graphs[old].exec) (void)hipGraphExecDestroy(ctx->graphs[old].exec);
        if (ctx->graphs[old].graph) (void)hipGraphDestroy(ctx->graphs[old].graph);
        ctx->graphs.erase(ctx->graphs.begin() + (std::ptrdiff_t)old);
    }
    NFX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeGlobal));
    int rc = NFX_OK;
    for (int64_t i = 0; i < kGraphBlock && rc == NFX_OK; ++i) {
        if (b) {
            rc = launch_batch_step(ctx, p, *b, 0, ctx->step_dev, (int32_t)i);
        } else if (node_fused(p)) {
            // launch i of a block is launch k0 + i of its call with k0 even (run_fused starts blocks at
            // k = 2, 102, ...), so its buffers are those of launch 2 + i
            const int64_t ntiles = (p.T + p.tile_k - 1) / p.tile_k;
            rc = launch_fused(ctx, p, false,
                              fused_io(p, 2 + i, hist ? ctx->tip_rows + i * ntiles : nullptr, 0, ctx->step_dev, (int32_t)i));
        } else {
            rc = launch_force_pass(ctx, p, false);
            if (rc == NFX_OK)
                rc = launch_node_pass(ctx, p, 0, ctx->step_dev, (int32_t)i, hist ? ctx->hist_dev + i : nullptr,
                                      nullptr, VV_MIDDLE);
        }
    }
    if (rc == NFX_OK && !b && hist && node_fused(p)) {
        k_tip_rows<<<kGraphBlock, kBlock, 0, ctx->stream>>>(ctx->tip_rows, (p.T + p.tile_k - 1) / p.tile_k, ctx->hist_dev);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            set_error("kernel launch failed: %s", hipGetErrorString(le));
            rc = NFX_ERR_HIP;
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != NFX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return NFX_ERR_HIP;
    }
    nfx_ctx::GraphEntry ge{};
    ge.graph = g;
    const hipError_t ei = hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        return NFX_ERR_HIP;
    }
    ge.plan = p;
    if (b) ge.batch = *b;
    ge.is_batch = b != nullptr;
    ge.hist = hist;
    ge.last_use = ++ctx->graph_clock;
    ctx->graphs.push_back(ge);
    ctx->captures++;
    *out = &ctx->graphs.back();
    return NFX_OK;
}

// the context's per-tile indenter-force rows for kGraphBlock steps of a plan with ntiles tiles;
// growing them drops the captured graphs (they hold the old address)
static int ensure_tip_rows(nfx_ctx* ctx, int64_t ntiles) {
    const int64_t need = ntiles * kGraphBlock;
    if (ctx->tip_rows_cap >= need) return NFX_OK;
    NFX_HIP(hipStreamSynchronize(ctx->stream));
    drop_graphs(ctx);
    if (ctx->tip_rows) NFX_HIP(hipFree(ctx->tip_rows));
    ctx->tip_rows = nullptr;
    ctx->tip_rows_cap = 0;
    NFX_HIP(hipMalloc(&ctx->tip_rows, sizeof(double) * (size_t)need));
    ctx->tip_rows_cap = need;
    return NFX_OK;
}

// A node-fused call of n >= 2 steps: the force pass of step 0, launches k = 1 .. n-1 (each completes
// step k-1 at the nodes and runs the force pass of step k; launch 1 turns (v, f_old) into w), then
// the node pass of step n-1, which turns w back into (v, f_old) in the plan's buffers.  Launches
// 2 .. n-2 replay captured blocks of kGraphBlock where whole blocks fit (every block starts at an even
// k, so one captured graph serves them all); the last launch stays eager so that it can write the
// stress outputs.  With a tip history each launch writes one value per tile, reduced in a fixed
// order per step (k_tip_rows).
static int run_fused(nfx_ctx* c, const nfx_plan& p, int64_t step0, int64_t n, double* tip_hist, bool use_graph,
                     bool stress) {
    const int64_t ntiles = (p.T + p.tile_k - 1) / p.tile_k;
    int rc = NFX_OK;
    if (tip_hist) {
        rc = ensure_tip_rows(c, ntiles);
        if (rc) return rc;
    }
    rc = launch_force_pass(c, p, false);
    if (rc) return rc;
    auto eager = [&](int64_t k) -> int {
        int r = launch_fused(c, p, stress && k == n - 1,
                             fused_io(p, k, tip_hist ? c->tip_rows : nullptr, step0 + k - 1, nullptr, 0));
        if (r == NFX_OK && tip_hist) {
            k_tip_rows<<<1, kBlock, 0, c->stream>>>(c->tip_rows, ntiles, tip_hist + k - 1);
            NFX_CHECK_LAUNCH();
        }
        return r;
    };
    rc = eager(1);
    if (rc) return rc;
    int64_t k = 2;
    if (use_graph && n - 1 - k >= kGraphBlock) {
        const bool hist = tip_hist != nullptr;
        nfx_ctx::GraphEntry* ge = find_graph(c, p, nullptr, hist);
        if (!ge) {
            rc = capture_graph(c, p, nullptr, hist, &ge);
            if (rc) return rc;
        }
        while (n - 1 - k >= kGraphBlock) {
            k_set_i64<<<1, 1, 0, c->stream>>>(c->step_dev, step0 + k - 1);
            NFX_CHECK_LAUNCH();
            NFX_HIP(hipGraphLaunch(ge->exec, c->stream));
            if (hist)
                NFX_HIP(hipMemcpyAsync(tip_hist + k - 1, c->hist_dev, sizeof(double) * kGraphBlock,
                                       hipMemcpyDeviceToDevice, c->stream));
            k += kGraphBlock;
        }
    }
    for (; k < n; ++k) {
        rc = eager(k);
        if (rc) return rc;
    }
    const NodeIO io = fused_last_io(p, n);
    return launch_node_pass(c, p, step0 + n - 1, nullptr, 0, tip_hist ? tip_hist + n - 1 : nullptr, nullptr, VV_LAST,
                            &io);
}

}  // namespace nfx

// =================================================================================================
// C ABI
// =================================================================================================
extern "C" {

const char* nfx_last_error(void) { return nfx::g_err.c_str(); }

int nfx_device_count(int* count) {
    NFX_REQUIRE(count, "count is NULL");
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return NFX_OK;
}

namespace nfx {
// device resources of a new context (on the current device); on failure the caller destroys it
static int init_ctx(nfx_ctx* c, void* stream) {
    hipDeviceProp_t prop;
    NFX_HIP(hipGetDeviceProperties(&prop, c->device));
    c->num_cu = prop.multiProcessorCount;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        NFX_HIP(hipStreamCreate(&c->stream));
        c->own_stream = true;
    }
    NFX_HIP(hipMalloc(&c->partials, sizeof(double) * 2 * kMaxReduceBlocks));
    NFX_HIP(hipMalloc(&c->counter, sizeof(unsigned int) * 4));
    NFX_HIP(hipMalloc(&c->step_dev, sizeof(int64_t) * 2));
    NFX_HIP(hipMalloc(&c->hist_dev, sizeof(double) * (kGraphBlock + 2)));
    NFX_HIP(hipMemset(c->counter, 0, sizeof(unsigned int) * 4));
    NFX_HIP(hipMemset(c->step_dev, 0, sizeof(int64_t) * 2));
    NFX_HIP(hipEventCreate(&c->ev0));
    NFX_HIP(hipEventCreate(&c->ev1));
    NFX_HIP(hipDeviceSynchronize());
    return NFX_OK;
}
}  // namespace nfx

int nfx_init(int device_id, void* stream, nfx_ctx** out) {
    NFX_REQUIRE(out, "ctx out-pointer is NULL");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device available");
        return NFX_ERR_DEVICE;
    }
    if (device_id < 0 || device_id >= count) {
        set_error("device %d not found (%d HIP devices)", device_id, count);
        return NFX_ERR_DEVICE;
    }
    NFX_HIP(hipSetDevice(device_id));
    nfx_ctx* c = new nfx_ctx();
    c->device = device_id;
    const int rc = nfx::init_ctx(c, stream);
    if (rc != NFX_OK) {
        nfx_destroy(c);  // frees what was set up (the error message stays)
        return rc;
    }
    *out = c;
    return NFX_OK;
}

int nfx_destroy(nfx_ctx* c) {
    if (!c) return NFX_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    nfx::drop_graphs(c);
    (void)hipFree(c->partials);
    (void)hipFree(c->counter);
    (void)hipFree(c->step_dev);
    (void)hipFree(c->hist_dev);
    if (c->tip_rows) (void)hipFree(c->tip_rows);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NFX_OK;
}

int nfx_graph_captures(nfx_ctx* c, int64_t* count) {
    NFX_REQUIRE(c && count, "NULL argument");
    *count = c->captures;
    return NFX_OK;
}

int nfx_stream_sync(nfx_ctx* c) {
    NFX_REQUIRE(c, "ctx is NULL");
    NFX_HIP(hipStreamSynchronize(c->stream));
    return NFX_OK;
}

int nfx_device_name(nfx_ctx* c, char* buf, int32_t len) {
    NFX_REQUIRE(c && buf && len > 0, "bad arguments");
    hipDeviceProp_t prop;
    NFX_HIP(hipGetDeviceProperties(&prop, c->device));
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return NFX_OK;
}

int nfx_batch_geometry(nfx_ctx* c, const double* p_rel, const double* cell_scale, const uint32_t* node_cell,
                       const uint32_t* tetras, const uint16_t* tile_ntet, int32_t tile_k, int64_t N, int64_t T,
                       float* ref, float* dminv, float* v0) {
    NFX_REQUIRE(c && p_rel && cell_scale && node_cell && tetras && tile_ntet && ref && dminv && v0, "NULL argument");
    NFX_REQUIRE(tile_k == 128 || tile_k == 256 || tile_k == 512, "tile_k must be 128, 256 or 512");
    int rc = check_sizes(N, T);
    if (rc) return rc;
    NFX_REQUIRE(T % tile_k == 0, "padded tetra count must be a multiple of tile_k");
    k_batch_points<<<nblocks(N), kBlock, 0, c->stream>>>(p_rel, cell_scale, node_cell, N, ref);
    NFX_CHECK_LAUNCH();
    k_batch_edges<<<nblocks(T), kBlock, 0, c->stream>>>(p_rel, cell_scale, node_cell, tetras, tile_ntet, tile_k, N, T,
                                                        dminv, v0);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_precompute_f64(nfx_ctx* c, const double* edge_vectors, const double* volumes, int64_t T, float* dminv,
                       float* v0) {
    NFX_REQUIRE(c && edge_vectors && volumes && dminv && v0, "NULL argument");
    NFX_REQUIRE(T > 0 && T < (int64_t(1) << 30), "tetra count %lld out of range", (long long)T);
    k_precompute_f64<<<nblocks(T), kBlock, 0, c->stream>>>(edge_vectors, volumes, T, dminv, v0);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

static ElasticDev elastic_of(const nfx_elastic& m) {
    nfx_plan p{};
    p.model = m.model;
    p.E = m.E;
    p.nu = m.nu;
    p.ratio = m.ratio;
    p.shear = m.shear;
    p.bulk = m.bulk;
    p.shear2 = m.shear2;
    p.volume_fraction = m.volume_fraction;
    return elastic_consts(p);
}

static int tetra_common(nfx_ctx* c, const nfx_elastic* m, const float* points, const uint32_t* tetras,
                        const float* dminv, const float* v0, int64_t N, int64_t T, float* force, float* corners) {
    NFX_REQUIRE(c && m && points && tetras && dminv && v0 && (force || corners), "NULL argument");
    NFX_REQUIRE(m->model >= NFX_SVKK && m->model <= NFX_MR, "unknown elastic model %d", m->model);
    int rc = check_sizes(N, T);
    if (rc) return rc;
    const TetArgs a{N, T, tetras, dminv, v0, points, force, corners, nullptr, nullptr};
    launch_tetra_pass<false>(c->stream, a, m->model, corners != nullptr, elastic_of(*m), VPConst{}, nullptr, nullptr,
                             nullptr, nullptr);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_legacy_force(nfx_ctx* c, const nfx_elastic* m, const float* points, const uint32_t* tetras,
                     const float* dminv, int64_t N, int64_t T, float* von_mises, float* pressure, float* force,
                     float* corners) {
    NFX_REQUIRE(c && m && points && tetras && dminv && (force || corners), "NULL argument");
    NFX_REQUIRE(is_legacy(m->model), "not a legacy law: model %d", m->model);
    int rc = check_sizes(N, T);
    if (rc) return rc;
    const bool st = von_mises && pressure;
    const TetArgs a{N, T, tetras, dminv, nullptr, points, force, corners, st ? von_mises : nullptr,
                    st ? pressure : nullptr};
    launch_tetra_pass<false>(c->stream, a, m->model, corners != nullptr, elastic_of(*m), VPConst{}, nullptr, nullptr,
                             nullptr, nullptr);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_tetra_force(nfx_ctx* c, const nfx_elastic* m, const float* points, const uint32_t* tetras, const float* dminv,
                    const float* v0, int64_t N, int64_t T, float* force) {
    NFX_REQUIRE(force, "force is NULL");
    return tetra_common(c, m, points, tetras, dminv, v0, N, T, force, nullptr);
}

int nfx_tetra_corner_forces(nfx_ctx* c, const nfx_elastic* m, const float* points, const uint32_t* tetras,
                            const float* dminv, const float* v0, int64_t N, int64_t T, float* corners) {
    NFX_REQUIRE(corners, "corners is NULL");
    return tetra_common(c, m, points, tetras, dminv, v0, N, T, nullptr, corners);
}

int nfx_viscoplastic_force(nfx_ctx* c, const nfx_viscoplastic* vp, const float* points, const uint32_t* tetras,
                           const float* dminv, int64_t N, int64_t T, float* Fp, float* epsbar, float* von_mises,
                           float* pressure, float* force, float* corners) {
    NFX_REQUIRE(c && vp && points && tetras && dminv && Fp && epsbar && (force || corners), "NULL argument");
    int rc = check_sizes(N, T);
    if (rc) return rc;
    const bool st = von_mises && pressure;
    const TetArgs a{N, T, tetras, dminv, nullptr, points, force, corners, nullptr, nullptr};
    launch_tetra_pass<false>(c->stream, a, NFX_VISCOPLASTIC, corners != nullptr, ElasticDev{}, vp_consts(*vp), Fp,
                             epsbar, st ? von_mises : nullptr, st ? pressure : nullptr);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_vp_q_from_fp(nfx_ctx* c, const float* Fp, int64_t T, float* Q) {
    NFX_REQUIRE(c && Fp && Q, "NULL argument");
    NFX_REQUIRE(T > 0 && T < (int64_t(1) << 30), "tetra count %lld out of range", (long long)T);
    k_vp_convert<true><<<nblocks(T), kBlock, 0, c->stream>>>(Fp, T, Q);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_vp_fp_from_q(nfx_ctx* c, const float* Q, int64_t T, float* Fp) {
    NFX_REQUIRE(c && Q && Fp, "NULL argument");
    NFX_REQUIRE(T > 0 && T < (int64_t(1) << 30), "tetra count %lld out of range", (long long)T);
    k_vp_convert<false><<<nblocks(T), kBlock, 0, c->stream>>>(Q, T, Fp);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_vp_tile_flags(nfx_ctx* c, const nfx_plan* p) {
    NFX_REQUIRE(c && p, "NULL argument");
    const int rc = validate_plan(p);
    if (rc) return rc;
    NFX_REQUIRE(p->model == NFX_VISCOPLASTIC && p->scatter == NFX_SCATTER_TILED && p->vp_active,
                "plan: plastic-activity flags need the tiled viscoplastic force and vp_active");
    const int64_t ntiles = (p->T + p->tile_k - 1) / p->tile_k;
    k_vp_tile_flags<<<(unsigned int)ntiles, kBlock, 0, c->stream>>>(p->vp_state, p->vp_sym ? 6 : 9, p->epsbar, p->T,
                                                                     p->tile_k, p->vp_active);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_points_from_disp(nfx_ctx* c, const float* ref, const float* disp, int64_t N, float* points) {
    NFX_REQUIRE(c && ref && disp && points, "NULL argument");
    NFX_REQUIRE(N > 0 && N < (int64_t(1) << 31), "node count out of range");
    k_disp_convert<true><<<nblocks(3 * N), kBlock, 0, c->stream>>>(ref, disp, 3 * N, points);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_disp_from_points(nfx_ctx* c, const float* ref, const float* points, int64_t N, float* disp) {
    NFX_REQUIRE(c && ref && disp && points, "NULL argument");
    NFX_REQUIRE(N > 0 && N < (int64_t(1) << 31), "node count out of range");
    k_disp_convert<false><<<nblocks(3 * N), kBlock, 0, c->stream>>>(ref, points, 3 * N, disp);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_gather_corners(nfx_ctx* c, const uint32_t* off, const uint32_t* slots, const float* corners, int64_t N,
                       float* force) {
    NFX_REQUIRE(c && off && slots && corners && force, "NULL argument");
    NFX_REQUIRE(N > 0 && N < (int64_t(1) << 31), "node count out of range");
    NodeArgs a{};
    a.N = N;
    a.off = off;
    a.slots = slots;
    a.corners = corners;
    k_gather<<<nblocks(N * 4), kBlock, 0, c->stream>>>(a, force);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_build_node_csr_host(const uint32_t* tet, int64_t N, int64_t T, uint32_t* off, uint32_t* slots) {
    NFX_REQUIRE(tet && off && slots, "NULL argument");
    int rc = check_sizes(N, T);
    if (rc) return rc;
    std::vector<uint32_t> cnt((size_t)N + 1, 0);
    for (int64_t a = 0; a < 4; ++a)
        for (int64_t t = 0; t < T; ++t) {
            const uint32_t n = tet[a * T + t];
            NFX_REQUIRE(n < (uint64_t)N, "tetra %lld corner %lld references node %u >= N", (long long)t, (long long)a, n);
            cnt[n + 1]++;
        }
    for (int64_t n = 0; n < N; ++n) cnt[n + 1] += cnt[n];
    std::memcpy(off, cnt.data(), sizeof(uint32_t) * (size_t)(N + 1));
    // ascending slot order per node: iterate slots in increasing order (t major, corner minor)
    for (int64_t t = 0; t < T; ++t)
        for (int64_t a = 0; a < 4; ++a) {
            const uint32_t n = tet[a * T + t];
            slots[cnt[n]++] = (uint32_t)(4 * t + a);
        }
    return NFX_OK;
}

int nfx_build_tiles_host(const uint32_t* tet, int64_t N, int64_t T, int32_t K, int32_t* max_nodes, int64_t* m_total,
                         uint32_t* tile_count, uint32_t* tile_nodes, uint16_t* tile_lend, uint16_t* tile_corners,
                         uint32_t* tile_dst, uint32_t* node_off) {
    NFX_REQUIRE(tet && max_nodes && m_total && tile_count, "NULL argument");
    NFX_REQUIRE(K == 128 || K == 256 || K == 512, "tile_k must be 128, 256 or 512");
    int rc = check_sizes(N, T);
    if (rc) return rc;
    const bool fill = tile_nodes != nullptr;
    NFX_REQUIRE(!fill || (tile_lend && tile_corners && tile_dst && node_off), "NULL output array");
    NFX_REQUIRE(!fill || *max_nodes <= 256, "max_nodes %d > 256: a tile-node index must fit 8 bits (use tile_k 256)",
                *max_nodes);
    const int64_t ntiles = (T + K - 1) / K;
    std::vector<uint64_t> key((size_t)4 * K);
    const int32_t MX = fill ? *max_nodes : 0;
    uint64_t total = 0;
    int32_t mx = 0;
    std::vector<uint32_t> cnt;
    if (fill) cnt.assign((size_t)N + 1, 0);
    for (int64_t c = 0; c < ntiles; ++c) {
        const int64_t t0 = c * K;
        const int nt = (int)std::min<int64_t>(K, T - t0);
        for (int tl = 0; tl < nt; ++tl)
            for (int a = 0; a < 4; ++a) {
                const uint32_t n = tet[(int64_t)a * T + t0 + tl];
                NFX_REQUIRE(n < (uint64_t)N, "tetra %lld references node %u >= N", (long long)(t0 + tl), n);
                key[(size_t)(4 * tl + a)] = ((uint64_t)n << 12) | (uint64_t)(4 * tl + a);
            }
        std::sort(key.begin(), key.begin() + 4 * nt);
        int M = 0, run0 = 0;
        if (fill)
            for (int q = 4 * nt; q < 4 * K; ++q) tile_corners[(size_t)4 * K * c + q] = 0;  // padding slots
        for (int q = 0; q < 4 * nt; ++q) {
            const uint32_t n = (uint32_t)(key[(size_t)q] >> 12);
            const bool first = q == 0 || (uint32_t)(key[(size_t)q - 1] >> 12) != n;
            if (first) {
                if (fill) {
                    NFX_REQUIRE(M < MX, "tile %lld has more than max_nodes = %d nodes", (long long)c, MX);
                    tile_nodes[c * MX + M] = n;
                    cnt[n + 1]++;
                }
                ++M;
                run0 = q;
            }
            if (fill) {
                const int rank = q - run0;
                NFX_REQUIRE(rank < 256, "tile %lld: node %u has more than 256 corners (use tile_k 256)", (long long)c,
                            n);
                tile_corners[(size_t)4 * K * c + (key[(size_t)q] & 0xFFF)] = (uint16_t)((rank << 8) | (M - 1));
                tile_lend[c * MX + M - 1] = (uint16_t)(q + 1);
            }
        }
        if (fill) {
            for (int i = M; i < MX; ++i) {  // padding: a valid node id, an empty run, never read
                tile_nodes[c * MX + i] = tile_nodes[c * MX];
                tile_lend[c * MX + i] = (uint16_t)(4 * nt);
                tile_dst[c * MX + i] = 0;
            }
        }
        tile_count[c] = (uint32_t)M;
        total += (uint64_t)M;
        NFX_REQUIRE(total < (uint64_t(1) << 32), "too many tile-nodes");
        mx = std::max(mx, M);
    }
    *m_total = (int64_t)total;
    if (!fill) {
        *max_nodes = mx;
        return NFX_OK;
    }
    // node-ordered partial positions: node n's tile-nodes in ascending tile order
    for (int64_t n = 0; n < N; ++n) cnt[n + 1] += cnt[n];
    std::memcpy(node_off, cnt.data(), sizeof(uint32_t) * (size_t)(N + 1));
    for (int64_t c = 0; c < ntiles; ++c)
        for (uint32_t i = 0; i < tile_count[c]; ++i) tile_dst[c * MX + i] = cnt[tile_nodes[c * MX + i]]++;
    return NFX_OK;
}

int nfx_sphere_force(nfx_ctx* c, float* force, const float* points, int64_t N, double k, double radius, double pos) {
    NFX_REQUIRE(c && force && points && N > 0, "bad argument");
    k_sphere<<<nblocks(N), kBlock, 0, c->stream>>>(force, points, N, (float)k, (float)radius, (float)pos);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_substrate_force(nfx_ctx* c, float* force, const float* points, int64_t N, double k, double wall) {
    NFX_REQUIRE(c && force && points && N > 0, "bad argument");
    k_substrate<<<nblocks(N), kBlock, 0, c->stream>>>(force, points, N, (float)k, (float)wall);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_plane_force(nfx_ctx* c, float* force, const float* points, int64_t N, double k, double wall) {
    NFX_REQUIRE(c && force && points && N > 0, "bad argument");
    k_plane<<<nblocks(N), kBlock, 0, c->stream>>>(force, points, N, (float)k, (float)wall);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_velocity_verlet(nfx_ctx* c, float* points, float* vel, float* force, float* fold, int64_t N, int32_t fix) {
    NFX_REQUIRE(c && points && vel && force && fold && N > 0, "bad argument");
    k_vv<<<nblocks(N), kBlock, 0, c->stream>>>(points, vel, force, fold, N, fix);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_tip_force(nfx_ctx* c, const float* points, const float* ref, int64_t N, int32_t ind, double k, double radius,
                  double pos, double* out2) {
    NFX_REQUIRE(c && points && out2 && N > 0, "bad argument");
    NFX_REQUIRE(ind == NFX_INDENTER_SPHERE || ind == NFX_INDENTER_PLANE, "unknown indenter %d", ind);
    const unsigned int nb = std::min<unsigned int>(nblocks(N), kMaxReduceBlocks);
    if (ref)
        k_tip<true><<<nb, kBlock, 0, c->stream>>>(points, ref, N, ind, k, radius, pos, c->partials, c->counter, out2,
                                                  out2 + 1);
    else
        k_tip<false><<<nb, kBlock, 0, c->stream>>>(points, nullptr, N, ind, k, radius, pos, c->partials, c->counter,
                                                   out2, out2 + 1);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

namespace nfx {
constexpr int kMaxNoAllocDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_addr[kMaxNoAllocDevices] = {};

// the current device's copy of ``scratch``, looked up once per device
static int noalloc_scratch(double** out) {
    int dev = -1;
    NFX_HIP(hipGetDevice(&dev));
    NFX_REQUIRE(dev >= 0 && dev < kMaxNoAllocDevices, "device %d: the allocation-free entry points serve devices "
                "0..%d", dev, kMaxNoAllocDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* a = nullptr;
        const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !a) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return NFX_ERR_HIP;
        }
        g_scratch_addr[dev] = static_cast<double*>(a);
    }
    *out = g_scratch_addr[dev];
    return NFX_OK;
}
}  // namespace nfx

int nfx_tip_force_noalloc(const float* points, const float* ref, int64_t N, int32_t ind, double k, double radius,
                          double pos, double* out2, void* stream) {
    NFX_REQUIRE(points && out2, "NULL argument");
    NFX_REQUIRE(N > 0 && N < (int64_t(1) << 31), "node count %lld out of range (1 .. 2^31 - 1)", (long long)N);
    NFX_REQUIRE(ind == NFX_INDENTER_SPHERE || ind == NFX_INDENTER_PLANE, "unknown indenter %d", ind);
    double* s = nullptr;
    const int rc = nfx::noalloc_scratch(&s);
    if (rc) return rc;
    // the grid is capped at kMaxReduceBlocks, so its partials always fit the kNoAllocPartials slots
    const unsigned int nb = std::min<unsigned int>(nblocks(N), kMaxReduceBlocks);
    unsigned int* counter = reinterpret_cast<unsigned int*>(s + kNoAllocPartials);
    hipStream_t st = static_cast<hipStream_t>(stream);
    k_clear_u32<<<1, 1, 0, st>>>(counter);  // a kernel, not a memset: callers may capture this call
    if (ref)
        k_tip<true><<<nb, kBlock, 0, st>>>(points, ref, N, ind, k, radius, pos, s, counter, out2, out2 + 1);
    else
        k_tip<false><<<nb, kBlock, 0, st>>>(points, nullptr, N, ind, k, radius, pos, s, counter, out2, out2 + 1);
    NFX_CHECK_LAUNCH();
    return NFX_OK;
}

int nfx_node_update(nfx_ctx* c, const nfx_plan* p, int64_t step, double* tip_out) {
    NFX_REQUIRE(c, "ctx is NULL");
    int rc = validate_plan(p);
    if (rc) return rc;
    NFX_REQUIRE(!tip_out || p->indenter != NFX_INDENTER_NONE, "tip force output needs an indenter");
    return launch_node_pass(c, *p, step, nullptr, 0, tip_out, tip_out ? tip_out + 1 : nullptr);
}

int nfx_run_steps(nfx_ctx* c, const nfx_plan* p, int64_t step0, int64_t nsteps, double* tip_hist, int32_t use_graph) {
    NFX_REQUIRE(c, "ctx is NULL");
    int rc = validate_plan(p);
    if (rc) return rc;
    NFX_REQUIRE(nsteps >= 0, "nsteps < 0");
    NFX_REQUIRE(!tip_hist || p->indenter != NFX_INDENTER_NONE, "tip force history needs an indenter");
    if (nsteps == 0) return NFX_OK;
    // with stress outputs in the plan, the last step of the call writes the per-tetra von Mises
    // stress and pressure (the state a frame written afterwards shows)
    const bool stress = p->von_mises && p->pressure;
    if (nsteps == 1) {
        rc = launch_force_pass(c, *p, stress);
        if (rc) return rc;
        return launch_node_pass(c, *p, step0, nullptr, 0, tip_hist, nullptr, VV_PLAIN);
    }
    if (node_fused(*p)) return run_fused(c, *p, step0, nsteps, tip_hist, use_graph != 0, stress);
    // n >= 2: the first step turns (v, f_old) into w, the middle steps advance (u, w), the last turns
    // w back into (v, f_old); the first and last steps run eagerly, the middle ones in whole blocks of
    // kGraphBlock steps that replay one captured graph (kept for as long as the plan and the history
    // mode stay the same) and an eager remainder, so no call re-captures for a new step count
    rc = launch_force_pass(c, *p, false);
    if (rc) return rc;
    rc = launch_node_pass(c, *p, step0, nullptr, 0, tip_hist, nullptr, VV_FIRST);
    if (rc) return rc;
    int64_t done = 1;
    const int64_t ngraph = nsteps - 1;  // middle steps run while done < ngraph
    if (use_graph && ngraph - done >= kGraphBlock) {
        const int64_t B = kGraphBlock;
        const bool hist = tip_hist != nullptr;
        nfx_ctx::GraphEntry* ge = find_graph(c, *p, nullptr, hist);
        if (!ge) {
            rc = capture_graph(c, *p, nullptr, hist, &ge);
            if (rc) return rc;
        }
        while (ngraph - done >= B) {
            k_set_i64<<<1, 1, 0, c->stream>>>(c->step_dev, step0 + done);
            NFX_CHECK_LAUNCH();
            NFX_HIP(hipGraphLaunch(ge->exec, c->stream));
            if (hist)
                NFX_HIP(hipMemcpyAsync(tip_hist + done, c->hist_dev, sizeof(double) * B, hipMemcpyDeviceToDevice,
                                       c->stream));
            done += B;
        }
    }
    for (; done < nsteps; ++done) {
        const bool last = done == nsteps - 1;
        rc = launch_force_pass(c, *p, stress && last);
        if (rc) return rc;
        rc = launch_node_pass(c, *p, step0 + done, nullptr, 0, tip_hist ? tip_hist + done : nullptr, nullptr,
                              last ? VV_LAST : VV_MIDDLE);
        if (rc) return rc;
    }
    return NFX_OK;
}

int nfx_time_passes(nfx_ctx* c, const nfx_plan* p, int64_t step0, int32_t nrep, double* ms_force, double* ms_node) {
    NFX_REQUIRE(c && ms_force && ms_node && nrep > 0, "bad argument");
    int rc = validate_plan(p);
    if (rc) return rc;
    std::vector<hipEvent_t> ev(3 * (size_t)nrep);
    for (auto& e : ev) NFX_HIP(hipEventCreate(&e));
    if (node_fused(*p) && nrep >= 3) {
        // the force pass of step 0 (ev 0-1), node-fused launches 1 .. nrep-1 (ev[i] - ev[i+1]), the
        // last node pass (ev nrep - ev nrep+1)
        NFX_HIP(hipEventRecord(ev[0], c->stream));
        rc = launch_force_pass(c, *p, false);
        for (int k = 1; k < nrep && rc == NFX_OK; ++k) {
            NFX_HIP(hipEventRecord(ev[k], c->stream));
            rc = launch_fused(c, *p, false, fused_io(*p, k, nullptr, step0 + k - 1, nullptr, 0));
        }
        if (rc == NFX_OK) {
            NFX_HIP(hipEventRecord(ev[nrep], c->stream));
            const NodeIO io = fused_last_io(*p, nrep);
            rc = launch_node_pass(c, *p, step0 + nrep - 1, nullptr, 0, nullptr, nullptr, VV_LAST, &io);
            NFX_HIP(hipEventRecord(ev[nrep + 1], c->stream));
        }
        if (rc == NFX_OK) {
            NFX_HIP(hipEventSynchronize(ev[nrep + 1]));
            double sf = 0;
            for (int k = 2; k < nrep; ++k) {
                float a = 0;
                NFX_HIP(hipEventElapsedTime(&a, ev[k], ev[k + 1]));
                sf += a;
            }
            float b = 0;
            NFX_HIP(hipEventElapsedTime(&b, ev[nrep], ev[nrep + 1]));
            *ms_force = sf / (nrep - 2);
            *ms_node = b;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        return rc;
    }
    for (int i = 0; i < nrep && rc == NFX_OK; ++i) {
        NFX_HIP(hipEventRecord(ev[3 * i], c->stream));
        rc = launch_force_pass(c, *p, false);
        if (rc) break;
        NFX_HIP(hipEventRecord(ev[3 * i + 1], c->stream));
        const int vvm = nrep == 1 ? VV_PLAIN : i == 0 ? VV_FIRST : i == nrep - 1 ? VV_LAST : VV_MIDDLE;
        rc = launch_node_pass(c, *p, step0 + i, nullptr, 0, nullptr, nullptr, vvm);
        NFX_HIP(hipEventRecord(ev[3 * i + 2], c->stream));
    }
    if (rc == NFX_OK) {
        NFX_HIP(hipEventSynchronize(ev[3 * (size_t)nrep - 1]));
        // the node pass is averaged over the middle steps (the (u, w) form a long call runs) when
        // there are any
        double sf = 0, sn = 0;
        int nn = 0;
        for (int i = 0; i < nrep; ++i) {
            float a = 0, b = 0;
            NFX_HIP(hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]));
            NFX_HIP(hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]));
            sf += a;
            if (nrep < 3 || (i > 0 && i < nrep - 1)) {
                sn += b;
                ++nn;
            }
        }
        *ms_force = sf / nrep;
        *ms_node = sn / nn;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}


int nfx_batch_run_steps(nfx_ctx* c, const nfx_plan* p, const nfx_batch* b, int64_t step0, int64_t nsteps,
                        int32_t use_graph) {
    NFX_REQUIRE(c, "ctx is NULL");
    int rc = validate_plan(p);
    if (rc) return rc;
    rc = validate_batch(p, b);
    if (rc) return rc;
    NFX_REQUIRE(nsteps >= 0, "nsteps < 0");
    int64_t done = 0;
    if (use_graph && nsteps >= kGraphBlock) {  // as in nfx_run_steps: whole blocks replay, the rest is eager
        const int64_t B = kGraphBlock;
        nfx_ctx::GraphEntry* ge = find_graph(c, *p, b, false);
        if (!ge) {
            rc = capture_graph(c, *p, b, false, &ge);
            if (rc) return rc;
        }
        while (nsteps - done >= B) {
            k_set_i64<<<1, 1, 0, c->stream>>>(c->step_dev, step0 + done);
            NFX_CHECK_LAUNCH();
            NFX_HIP(hipGraphLaunch(ge->exec, c->stream));
            done += B;
        }
    }
    for (; done < nsteps; ++done) {
        rc = launch_batch_step(c, *p, *b, step0 + done, nullptr, 0);
        if (rc) return rc;
    }
    return NFX_OK;
}

int nfx_batch_time_passes(nfx_ctx* c, const nfx_plan* p, const nfx_batch* b, int64_t step0, int32_t nrep,
                          double* ms_force, double* ms_node, double* ms_curve) {
    NFX_REQUIRE(c && ms_force && ms_node && ms_curve && nrep > 0, "bad argument");
    int rc = validate_plan(p);
    if (rc) return rc;
    rc = validate_batch(p, b);
    if (rc) return rc;
    std::vector<hipEvent_t> ev(4 * (size_t)nrep);
    for (auto& e : ev) NFX_HIP(hipEventCreate(&e));
    for (int i = 0; i < nrep && rc == NFX_OK; ++i) {
        NFX_HIP(hipEventRecord(ev[4 * i], c->stream));
        rc = launch_batch_step(c, *p, *b, step0 + i, nullptr, 0, 1);
        NFX_HIP(hipEventRecord(ev[4 * i + 1], c->stream));
        if (rc == NFX_OK) rc = launch_batch_step(c, *p, *b, step0 + i, nullptr, 0, 2);
        NFX_HIP(hipEventRecord(ev[4 * i + 2], c->stream));
        if (rc == NFX_OK) rc = launch_batch_step(c, *p, *b, step0 + i, nullptr, 0, 4);
        NFX_HIP(hipEventRecord(ev[4 * i + 3], c->stream));
    }
    if (rc == NFX_OK) {
        NFX_HIP(hipEventSynchronize(ev[4 * (size_t)nrep - 1]));
        double sc = 0, sf = 0, sn = 0;
        for (int i = 0; i < nrep; ++i) {
            float a = 0, bb = 0, cc = 0;
            NFX_HIP(hipEventElapsedTime(&a, ev[4 * i], ev[4 * i + 1]));
            NFX_HIP(hipEventElapsedTime(&bb, ev[4 * i + 1], ev[4 * i + 2]));
            NFX_HIP(hipEventElapsedTime(&cc, ev[4 * i + 2], ev[4 * i + 3]));
            sc += a;
            sf += bb;
            sn += cc;
        }
        *ms_curve = sc / nrep;
        *ms_force = sf / nrep;
        *ms_node = sn / nrep;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}


int nfx_stream_copy(nfx_ctx* c, const void* src, void* dst, int64_t bytes, int32_t nrep, double* ms) {
    NFX_REQUIRE(c && src && dst && ms && nrep > 0 && bytes > 0 && bytes % 16 == 0, "bad argument");
    const int64_t n4 = bytes / 16;
    NFX_REQUIRE((n4 + kBlock - 1) / kBlock < (int64_t(1) << 31), "copy too large");
    const unsigned int nb = (unsigned int)((n4 + kBlock - 1) / kBlock);
    k_stream_copy<<<nb, kBlock, 0, c->stream>>>((const nfx_f4*)src, (nfx_f4*)dst, n4);
    NFX_CHECK_LAUNCH();
    NFX_HIP(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < nrep; ++i) k_stream_copy<<<nb, kBlock, 0, c->stream>>>((const nfx_f4*)src, (nfx_f4*)dst, n4);
    NFX_CHECK_LAUNCH();
    NFX_HIP(hipEventRecord(c->ev1, c->stream));
    NFX_HIP(hipEventSynchronize(c->ev1));
    float t = 0;
    NFX_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t / nrep;
    return NFX_OK;
}

}  // extern "C"
