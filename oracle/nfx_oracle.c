/*
 * nfx_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of noFluidX3D's explicit tetrahedral-FEM time step.  It is the checker
 * for the HIP kernels (tests, smoke check) and the timed CPU baseline of bench.py; nothing in
 * the product links or calls it.  Each function follows the reference's arithmetic in the
 * order the OpenCL source writes it (same buffers, same SoA layouts, FP64 throughout):
 *
 *   orc_tetra_force      kernels/Interactions/tetra.cl:14-182 (SVKK, SVK, MR/NH branches)
 *                        with kernels/matrix.cl:190-290 (Double3x3 algebra, 1-based rows)
 *   orc_viscoplastic     kernels/InteractionFiniteStrainViscoplastic.cl:1-220, restated with
 *                        node 0 as the edge base (the legacy kernel used node 3; TetraCell.py:33-38
 *                        uses node 0), plus an optional yield offset (0 = reference law)
 *   orc_linear_viscoelastic  kernels/InteractionLinearViscoelastic.cl (node-0 base)
 *   orc_legacy_force     kernels/InteractionLinearElastic.cl, InteractionLinearElasticDeviatoric.cl,
 *                        InteractionSecondOrderNeoHookean.cl, InteractionPoroNeoHookean.cl (node-0 base)
 *   orc_round_tip, orc_adhesive_plane, orc_point_zero_force, orc_center_of_mass, orc_com_force,
 *   orc_over_damped, orc_velocity_verlet2, orc_volume_quadrature, orc_adhesive_plane_surface,
 *   orc_hard_adhesive_plane_surface
 *                        the legacy contacts and integrators (RoundTip.cl, InteractionAdhesivePlane.cl,
 *                        PointZeroForce.cl, CalcCOM.cl, COMForce.cl, OverDamped.cl, velocityVerlet2.cl,
 *                        InteractionSphereIntegral.cl, InteractionTipIntegral.cl,
 *                        InteractionAdhesivePlaneSurfaceIntegral.cl,
 *                        InteractionHardAdhesivePlaneSurfaceIntegral.cl)
 *   orc_sphere           kernels/Interactions/Sphere.cl:1-17
 *   orc_substrate        kernels/Interactions/Substrate.cl:1-7
 *   orc_plane            kernels/Interactions/Plane.cl:1-7
 *   orc_velocity_verlet  kernels/Interactions/VelocityVerlet.cl:1-36
 *   orc_tip_force        Simulation.py:198-212 (forceSI, without the p0*L0^2 unit factor)
 *   orc_run              Simulation.py:232-237 timeStep in registration order
 *                        (Tetra, Substrate, Sphere|Plane, VelocityVerlet)
 *
 * The node scatter (kernels/atomicAdd.cl:1-11 CAS loop) is restated as a plain sequential add
 * when nthreads == 1 (deterministic order t = 0..T-1) and as an OpenMP atomic add otherwise.
 * Parity of each function is pinned by tests/test_oracle.py (closed forms, energy consistency by
 * central differences, objectivity, momentum balance, exact integrator recurrences).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } d3;
typedef struct { double m11, m12, m13, m21, m22, m23, m31, m32, m33; } M3;

static inline d3 v3(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 vadd(d3 a, d3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 vsub(d3 a, d3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 vscale(d3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline double vdot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 vcross(d3 a, d3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vlen(d3 a) { return sqrt(vdot(a, a)); }
static inline d3 vdivs(d3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }

/* matrix.cl:190-204 */
static inline M3 from_rows(d3 r1, d3 r2, d3 r3) {
    M3 m = {r1.x, r1.y, r1.z, r2.x, r2.y, r2.z, r3.x, r3.y, r3.z};
    return m;
}
static inline M3 from_cols(d3 c1, d3 c2, d3 c3) {
    M3 m = {c1.x, c2.x, c3.x, c1.y, c2.y, c3.y, c1.z, c2.z, c3.z};
    return m;
}
/* matrix.cl:214-232 */
static inline d3 row(M3 m, int i) {
    return i == 1 ? v3(m.m11, m.m12, m.m13) : i == 2 ? v3(m.m21, m.m22, m.m23) : v3(m.m31, m.m32, m.m33);
}
static inline d3 col(M3 m, int j) {
    return j == 1 ? v3(m.m11, m.m21, m.m31) : j == 2 ? v3(m.m12, m.m22, m.m32) : v3(m.m13, m.m23, m.m33);
}
/* matrix.cl:234-248 */
static inline double mtr(M3 m) { return m.m11 + m.m22 + m.m33; }
static inline double mdet(M3 m) { return vdot(col(m, 1), vcross(col(m, 2), col(m, 3))); }
static inline M3 mT(M3 m) {
    M3 r = {m.m11, m.m21, m.m31, m.m12, m.m22, m.m32, m.m13, m.m23, m.m33};
    return r;
}
/* matrix.cl:280-290 */
static inline M3 mscale(M3 m, double s) {
    M3 r = {m.m11 * s, m.m12 * s, m.m13 * s, m.m21 * s, m.m22 * s, m.m23 * s, m.m31 * s, m.m32 * s, m.m33 * s};
    return r;
}
static inline M3 mdiv(M3 m, double s) { return mscale(m, 1.0 / s); }
/* matrix.cl:250-253: cofactor rows divided by det, no singularity check */
static inline M3 minv(M3 m) {
    return mdiv(from_rows(vcross(col(m, 2), col(m, 3)), vcross(col(m, 3), col(m, 1)), vcross(col(m, 1), col(m, 2))),
                mdet(m));
}
/* matrix.cl:255-265 */
static inline M3 madd(M3 a, M3 b) {
    return from_rows(vadd(row(a, 1), row(b, 1)), vadd(row(a, 2), row(b, 2)), vadd(row(a, 3), row(b, 3)));
}
static inline M3 msub(M3 a, M3 b) {
    return from_rows(vsub(row(a, 1), row(b, 1)), vsub(row(a, 2), row(b, 2)), vsub(row(a, 3), row(b, 3)));
}
/* matrix.cl:267-273 */
static inline M3 mmul(M3 a, M3 b) {
    M3 r = {vdot(row(a, 1), col(b, 1)), vdot(row(a, 1), col(b, 2)), vdot(row(a, 1), col(b, 3)),
            vdot(row(a, 2), col(b, 1)), vdot(row(a, 2), col(b, 2)), vdot(row(a, 2), col(b, 3)),
            vdot(row(a, 3), col(b, 1)), vdot(row(a, 3), col(b, 2)), vdot(row(a, 3), col(b, 3))};
    return r;
}
static inline M3 munit(void) {
    M3 r = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    return r;
}

/* tetra.cl:5-12: ten Newton steps X <- (X + X^-T)/2, then X^T * initial */
static M3 approximate_root(M3 initial) {
    M3 it = initial;
    for (int i = 0; i < 10; i++) it = mdiv(madd(it, minv(mT(it))), 2.0);
    return mmul(mT(it), initial);
}

/* exported for the SVKK pin: approximateRoot of a row-major 3x3 */
void orc_polar_stretch(const double* m9, double* out9) {
    M3 m = {m9[0], m9[1], m9[2], m9[3], m9[4], m9[5], m9[6], m9[7], m9[8]};
    M3 r = approximate_root(m);
    out9[0] = r.m11; out9[1] = r.m12; out9[2] = r.m13;
    out9[3] = r.m21; out9[4] = r.m22; out9[5] = r.m23;
    out9[6] = r.m31; out9[7] = r.m32; out9[8] = r.m33;
}

enum { ORC_SVKK = 0, ORC_SVK = 1, ORC_MR = 2 };

static inline void scatter(double* f, int64_t idx, double v, int atomic) {
    if (atomic) {
#pragma omp atomic
        f[idx] += v;
    } else {
        f[idx] += v;
    }
}

/* kernels/Interactions/tetra.cl:14-182 for tetra t. */
static void tetra_one(int64_t t, int model, double E, double nu, double ratio, int64_t N, int64_t T,
                      const double* points, const uint32_t* tetras, const double* ev, const double* vol,
                      double* force, int atomic) {
    const uint32_t i0 = tetras[t], i1 = tetras[T + t], i2 = tetras[2 * T + t], i3 = tetras[3 * T + t];
    const d3 p0 = v3(points[i0], points[N + i0], points[2 * N + i0]);
    const d3 p1 = v3(points[i1], points[N + i1], points[2 * N + i1]);
    const d3 p2 = v3(points[i2], points[N + i2], points[2 * N + i2]);
    const d3 p3 = v3(points[i3], points[N + i3], points[2 * N + i3]);
    const d3 r1 = vsub(p1, p0), r2 = vsub(p2, p0), r3 = vsub(p3, p0); /* :31-35 */
    const M3 rT = from_cols(r1, r2, r3);                                /* :37 */
    const d3 R1 = v3(ev[t], ev[T + t], ev[2 * T + t]);                  /* :40-43 */
    const d3 R2 = v3(ev[3 * T + t], ev[4 * T + t], ev[5 * T + t]);
    const d3 R3 = v3(ev[6 * T + t], ev[7 * T + t], ev[8 * T + t]);
    const double V0 = vol[t];
    const M3 J = from_cols(R1, R2, R3); /* :47 */
    const M3 invJ = minv(J);            /* :53 */
    const M3 F = mmul(rT, invJ);        /* :63 */
    const M3 FT = mT(F);
    const M3 B = mmul(F, FT); /* :68 */
    const M3 uT = munit();
    M3 Force;
    if (model == ORC_SVKK) { /* :76-93 */
        const double kappa = nu / (1 - 2 * nu);
        const double mu = E / (2 * (1 + nu));
        const M3 S = approximate_root(B);
        const M3 invS = minv(S);
        const double sf = kappa * (mtr(S) - 3) - 1;
        const M3 inner = madd(uT, mscale(invS, sf));
        Force = mscale(mmul(mmul(invJ, inner), FT), -V0 * mu);
    } else if (model == ORC_SVK) { /* :94-114 */
        const double kappa = nu / (1 - 2 * nu);
        const double mu = E / (2 * (1 + nu));
        const M3 BP = msub(B, uT);
        const M3 H = mmul(invJ, FT);
        const M3 term1 = mscale(H, kappa * mtr(BP));
        const M3 term2 = mmul(H, BP);
        Force = mscale(madd(term1, term2), -V0 * mu);
    } else { /* :115-163, MR and NH */
        const double mu = E / (2 * (1 + nu));
        const double mu1 = ratio * mu;
        const double mu2 = (1 - ratio) * mu;
        const double kappa = E / (3 * (1 - 2 * nu));
        const M3 Bsq = mmul(B, B);
        const double trB = mtr(B);
        const double trBsq = mtr(Bsq);
        const double Jr = mdet(rT) / (6 * V0);
        const double dUdI1 = mu1 / pow(Jr, 5 / 3.0);
        const double dUdK = mu2 / pow(Jr, 7 / 3.0);
        const double dUdJ = kappa * (Jr - 1);
        const double dI1dJ = -trB / 3;
        const M3 term1 = mscale(B, dUdI1);
        const M3 term2 = mscale(uT, dUdJ + dI1dJ * dUdI1);
        const M3 inner1 = msub(mscale(B, trB), Bsq);
        const M3 inner2 = mscale(uT, (trBsq - trB * trB) / 3);
        const M3 term3 = mscale(madd(inner1, inner2), dUdK);
        const M3 sigma = madd(madd(term1, term2), term3);
        const M3 helper = mmul(minv(rT), sigma); /* :159 */
        Force = mscale(helper, -V0 * Jr);        /* :162 */
    }
    const d3 f1 = row(Force, 1), f2 = row(Force, 2), f3 = row(Force, 3);
    const d3 f0 = vscale(vadd(vadd(f1, f2), f3), -1.0); /* :168 */
    scatter(force, i0, f0.x, atomic);                    /* :170-181 */
    scatter(force, N + i0, f0.y, atomic);
    scatter(force, 2 * N + i0, f0.z, atomic);
    scatter(force, i1, f1.x, atomic);
    scatter(force, N + i1, f1.y, atomic);
    scatter(force, 2 * N + i1, f1.z, atomic);
    scatter(force, i2, f2.x, atomic);
    scatter(force, N + i2, f2.y, atomic);
    scatter(force, 2 * N + i2, f2.z, atomic);
    scatter(force, i3, f3.x, atomic);
    scatter(force, N + i3, f3.y, atomic);
    scatter(force, 2 * N + i3, f3.z, atomic);
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    return nthreads;
#else
    (void)nthreads;
    return 1;
#endif
}

void orc_tetra_force(int model, double E, double nu, double ratio, int64_t N, int64_t T, const double* points,
                     const uint32_t* tetras, const double* ev, const double* vol, double* force, int nthreads) {
    nthreads = set_threads(nthreads);
    if (nthreads == 1) {
        for (int64_t t = 0; t < T; t++) tetra_one(t, model, E, nu, ratio, N, T, points, tetras, ev, vol, force, 0);
        return;
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t t = 0; t < T; t++) tetra_one(t, model, E, nu, ratio, N, T, points, tetras, ev, vol, force, 1);
}

/* Viscoplastic material parameters; uniform over tetra as the reference wrapper builds them
 * (Interactions.py:270-277, values from SimulationLegacy.py:842-882). */
typedef struct {
    double mu1, mu2, kappa;      /* shear moduli and bulk modulus */
    double flow_exponent;        /* m   */
    double yield_stress;         /* Y   (sigma0 at zero accumulated strain) */
    double flow_rate;            /* eps0dot, per step */
    double hardening_threshold;  /* eps0 */
    double hardening_exponent;   /* n   */
    double yield_offset;         /* sigma_y: rate uses max(svm - sigma_y, 0); 0 = reference law */
} orc_vp_params;

/* InteractionFiniteStrainViscoplastic.cl:1-220 with the edge base moved to node 0: the legacy
 * nodes (p1, p2, p3, p4) are (tetras[1], tetras[2], tetras[3], tetras[0]). */
static void vp_one(int64_t t, const orc_vp_params* P, int64_t N, int64_t T, const double* points,
                   const uint32_t* tetras, const double* ev, const double* vol, double* force, double* acc,
                   double* Fp, double* vonMises, double* pressure, int atomic) {
    const uint32_t id1 = tetras[T + t], id2 = tetras[2 * T + t], id3 = tetras[3 * T + t], id4 = tetras[t];
    const d3 p1 = v3(points[id1], points[N + id1], points[2 * N + id1]);
    const d3 p2 = v3(points[id2], points[N + id2], points[2 * N + id2]);
    const d3 p3 = v3(points[id3], points[N + id3], points[2 * N + id3]);
    const d3 p4 = v3(points[id4], points[N + id4], points[2 * N + id4]);
    const d3 r1 = vsub(p1, p4), r2 = vsub(p2, p4), r3 = vsub(p3, p4); /* :18-22 */
    const d3 R1 = v3(ev[t], ev[T + t], ev[2 * T + t]);                  /* :25-28 */
    const d3 R2 = v3(ev[3 * T + t], ev[4 * T + t], ev[5 * T + t]);
    const d3 R3 = v3(ev[6 * T + t], ev[7 * T + t], ev[8 * T + t]);
    (void)vol;
    const d3 u1 = vsub(r1, R1), u2 = vsub(r2, R2), u3 = vsub(r3, R3); /* :32-34 */
    const double divisor = vdot(vcross(R1, R2), R3);                  /* :38-42 */
    const d3 invJ1 = vdivs(vcross(R2, R3), divisor);
    const d3 invJ2 = vdivs(vcross(R3, R1), divisor);
    const d3 invJ3 = vdivs(vcross(R1, R2), divisor);
    /* :47-55 */
    const d3 du1 = vadd(vadd(vscale(invJ1, u1.x), vscale(invJ2, u2.x)), vscale(invJ3, u3.x));
    const d3 du2 = vadd(vadd(vscale(invJ1, u1.y), vscale(invJ2, u2.y)), vscale(invJ3, u3.y));
    const d3 du3 = vadd(vadd(vscale(invJ1, u1.z), vscale(invJ2, u2.z)), vscale(invJ3, u3.z));
    const d3 F1 = vadd(du1, v3(1, 0, 0)), F2 = vadd(du2, v3(0, 1, 0)), F3 = vadd(du3, v3(0, 0, 1));
    /* :59-75 */
    double a = Fp[0 * T + t], b = Fp[1 * T + t], c = Fp[2 * T + t];
    double d = Fp[3 * T + t], e = Fp[4 * T + t], f = Fp[5 * T + t];
    double g = Fp[6 * T + t], h = Fp[7 * T + t], i = Fp[8 * T + t];
    const double detFp = vdot(vcross(v3(a, b, c), v3(d, e, f)), v3(g, h, i));
    const d3 iF1 = vscale(v3(e * i - f * h, c * h - b * i, b * f - c * e), 1.0 / detFp);
    const d3 iF2 = vscale(v3(f * g - d * i, a * i - c * g, c * d - a * f), 1.0 / detFp);
    const d3 iF3 = vscale(v3(d * h - e * g, b * g - a * h, a * e - b * d), 1.0 / detFp);
    /* :77-79  Fe = F * Fp^-1 */
    const d3 Fe1 = v3(F1.x * iF1.x + F1.y * iF2.x + F1.z * iF3.x, F1.x * iF1.y + F1.y * iF2.y + F1.z * iF3.y,
                      F1.x * iF1.z + F1.y * iF2.z + F1.z * iF3.z);
    const d3 Fe2 = v3(F2.x * iF1.x + F2.y * iF2.x + F2.z * iF3.x, F2.x * iF1.y + F2.y * iF2.y + F2.z * iF3.y,
                      F2.x * iF1.z + F2.y * iF2.z + F2.z * iF3.z);
    const d3 Fe3 = v3(F3.x * iF1.x + F3.y * iF2.x + F3.z * iF3.x, F3.x * iF1.y + F3.y * iF2.y + F3.z * iF3.y,
                      F3.x * iF1.z + F3.y * iF2.z + F3.z * iF3.z);
    /* :82-96 */
    const d3 B1 = v3(vdot(Fe1, Fe1), vdot(Fe1, Fe2), vdot(Fe1, Fe3));
    const d3 B2 = v3(vdot(Fe2, Fe1), vdot(Fe2, Fe2), vdot(Fe2, Fe3));
    const d3 B3 = v3(vdot(Fe3, Fe1), vdot(Fe3, Fe2), vdot(Fe3, Fe3));
    const double J = sqrt(vdot(vcross(B1, B2), B3));
    const d3 Bsq1 = v3(B1.x * B1.x + B1.y * B2.x + B1.z * B3.x, B1.x * B1.y + B1.y * B2.y + B1.z * B3.y,
                       B1.x * B1.z + B1.y * B2.z + B1.z * B3.z);
    const d3 Bsq2 = v3(B2.x * B1.x + B2.y * B2.x + B2.z * B3.x, B2.x * B1.y + B2.y * B2.y + B2.z * B3.y,
                       B2.x * B1.z + B2.y * B2.z + B2.z * B3.z);
    const d3 Bsq3 = v3(B3.x * B1.x + B3.y * B2.x + B3.z * B3.x, B3.x * B1.y + B3.y * B2.y + B3.z * B3.y,
                       B3.x * B1.z + B3.y * B2.z + B3.z * B3.z);
    const double trB = B1.x + B2.y + B3.z;
    const double trBsq = Bsq1.x + Bsq2.y + Bsq3.z;
    const d3 E1 = v3(1, 0, 0), E2 = v3(0, 1, 0), E3 = v3(0, 0, 1);
    /* :105-107 */
    const double j53 = pow(J, -5.0 / 3.0), j73 = pow(J, -7.0 / 3.0);
    const d3 s1 = vadd(vadd(vscale(vsub(B1, vscale(E1, 1. / 3. * trB)), P->mu1 * j53), vscale(E1, P->kappa * (J - 1))),
                       vscale(vadd(vsub(vsub(vscale(B1, trB), vscale(E1, 1.0 / 3.0 * trB * trB)), Bsq1),
                                   vscale(E1, 1.0 / 3.0 * trBsq)),
                              P->mu2 * j73));
    const d3 s2 = vadd(vadd(vscale(vsub(B2, vscale(E2, 1. / 3. * trB)), P->mu1 * j53), vscale(E2, P->kappa * (J - 1))),
                       vscale(vadd(vsub(vsub(vscale(B2, trB), vscale(E2, 1.0 / 3.0 * trB * trB)), Bsq2),
                                   vscale(E2, 1.0 / 3.0 * trBsq)),
                              P->mu2 * j73));
    const d3 s3 = vadd(vadd(vscale(vsub(B3, vscale(E3, 1. / 3. * trB)), P->mu1 * j53), vscale(E3, P->kappa * (J - 1))),
                       vscale(vadd(vsub(vsub(vscale(B3, trB), vscale(E3, 1.0 / 3.0 * trB * trB)), Bsq3),
                                   vscale(E3, 1.0 / 3.0 * trBsq)),
                              P->mu2 * j73));
    /* :110-123 outward unit normals */
    d3 n1 = vcross(r2, r3), n2 = vcross(r3, r1), n3 = vcross(r1, r2), n4 = vcross(vsub(r3, r1), vsub(r3, r2));
    n1 = vdivs(n1, vlen(n1));
    n2 = vdivs(n2, vlen(n2));
    n3 = vdivs(n3, vlen(n3));
    n4 = vdivs(n4, vlen(n4));
    if (vdot(n1, vscale(r1, -1.0)) < 0) n1 = vscale(n1, -1.0);
    if (vdot(n2, vscale(r2, -1.0)) < 0) n2 = vscale(n2, -1.0);
    if (vdot(n3, vscale(r3, -1.0)) < 0) n3 = vscale(n3, -1.0);
    if (vdot(n4, r1) < 0) n4 = vscale(n4, -1.0);
    /* :126-141 */
    const double A1 = 0.5 * vlen(vcross(r2, r3)), A2 = 0.5 * vlen(vcross(r3, r1)), A3 = 0.5 * vlen(vcross(r1, r2));
    const double A4 = 0.5 * vlen(vcross(vsub(r3, r1), vsub(r3, r2)));
    const d3 fA1 = vscale(v3(vdot(s1, n1), vdot(s2, n1), vdot(s3, n1)), A1);
    const d3 fA2 = vscale(v3(vdot(s1, n2), vdot(s2, n2), vdot(s3, n2)), A2);
    const d3 fA3 = vscale(v3(vdot(s1, n3), vdot(s2, n3), vdot(s3, n3)), A3);
    const d3 fA4 = vscale(v3(vdot(s1, n4), vdot(s2, n4), vdot(s3, n4)), A4);
    const d3 f1 = vscale(vadd(vadd(fA2, fA3), fA4), -1.0 / 3.0);
    const d3 f2 = vscale(vadd(vadd(fA3, fA4), fA1), -1.0 / 3.0);
    const d3 f3 = vscale(vadd(vadd(fA4, fA1), fA2), -1.0 / 3.0);
    const d3 f4 = vscale(vadd(vadd(fA1, fA2), fA3), -1.0 / 3.0);
    /* :144-156 */
    const double trS = s1.x + s2.y + s3.z;
    const d3 ds1 = vsub(s1, vscale(E1, trS / 3.0)), ds2 = vsub(s2, vscale(E2, trS / 3.0)),
             ds3 = vsub(s3, vscale(E3, trS / 3.0));
    const double svm = sqrt(3. / 2. * (vdot(ds1, ds1) + vdot(ds2, ds2) + vdot(ds3, ds3)));
    const double sigma0 = P->yield_stress * pow(1 + acc[t] / P->hardening_threshold, P->hardening_exponent);
    if (vonMises) vonMises[t] = svm;
    if (pressure) pressure[t] = trS / 3.0;
    const double over = P->yield_offset > 0 ? fmax(svm - P->yield_offset, 0.0) : svm;
    const double rate = P->flow_rate * pow(over / sigma0, P->flow_exponent);
    acc[t] += rate;
    /* :160-169 */
    d3 Dp1, Dp2, Dp3;
    if (svm != 0) {
        Dp1 = vdivs(vscale(ds1, rate * 3. / 2.), svm);
        Dp2 = vdivs(vscale(ds2, rate * 3. / 2.), svm);
        Dp3 = vdivs(vscale(ds3, rate * 3. / 2.), svm);
    } else {
        Dp1 = Dp2 = Dp3 = v3(0, 0, 0);
    }
    /* :173-206  Fp += Fe^-1 * Dp * F */
    a = Fe1.x; b = Fe1.y; c = Fe1.z; d = Fe2.x; e = Fe2.y; f = Fe2.z; g = Fe3.x; h = Fe3.y; i = Fe3.z;
    const double detFe = vdot(vcross(Fe1, Fe2), Fe3);
    const d3 iE1 = vscale(v3(e * i - f * h, c * h - b * i, b * f - c * e), 1. / detFe);
    const d3 iE2 = vscale(v3(f * g - d * i, a * i - c * g, c * d - a * f), 1. / detFe);
    const d3 iE3 = vscale(v3(d * h - e * g, b * g - a * h, a * e - b * d), 1. / detFe);
    const d3 tm1 = v3(iE1.x * Dp1.x + iE1.y * Dp2.x + iE1.z * Dp3.x, iE1.x * Dp1.y + iE1.y * Dp2.y + iE1.z * Dp3.y,
                      iE1.x * Dp1.z + iE1.y * Dp2.z + iE1.z * Dp3.z);
    const d3 tm2 = v3(iE2.x * Dp1.x + iE2.y * Dp2.x + iE2.z * Dp3.x, iE2.x * Dp1.y + iE2.y * Dp2.y + iE2.z * Dp3.y,
                      iE2.x * Dp1.z + iE2.y * Dp2.z + iE2.z * Dp3.z);
    const d3 tm3 = v3(iE3.x * Dp1.x + iE3.y * Dp2.x + iE3.z * Dp3.x, iE3.x * Dp1.y + iE3.y * Dp2.y + iE3.z * Dp3.y,
                      iE3.x * Dp1.z + iE3.y * Dp2.z + iE3.z * Dp3.z);
    const d3 dF1 = v3(tm1.x * F1.x + tm1.y * F2.x + tm1.z * F3.x, tm1.x * F1.y + tm1.y * F2.y + tm1.z * F3.y,
                      tm1.x * F1.z + tm1.y * F2.z + tm1.z * F3.z);
    const d3 dF2 = v3(tm2.x * F1.x + tm2.y * F2.x + tm2.z * F3.x, tm2.x * F1.y + tm2.y * F2.y + tm2.z * F3.y,
                      tm2.x * F1.z + tm2.y * F2.z + tm2.z * F3.z);
    const d3 dF3 = v3(tm3.x * F1.x + tm3.y * F2.x + tm3.z * F3.x, tm3.x * F1.y + tm3.y * F2.y + tm3.z * F3.y,
                      tm3.x * F1.z + tm3.y * F2.z + tm3.z * F3.z);
    Fp[0 * T + t] += dF1.x; Fp[1 * T + t] += dF1.y; Fp[2 * T + t] += dF1.z;
    Fp[3 * T + t] += dF2.x; Fp[4 * T + t] += dF2.y; Fp[5 * T + t] += dF2.z;
    Fp[6 * T + t] += dF3.x; Fp[7 * T + t] += dF3.y; Fp[8 * T + t] += dF3.z;
    /* :209-220 (legacy p1..p4 -> node ids id1..id4) */
    scatter(force, id1, f1.x, atomic); scatter(force, N + id1, f1.y, atomic); scatter(force, 2 * N + id1, f1.z, atomic);
    scatter(force, id2, f2.x, atomic); scatter(force, N + id2, f2.y, atomic); scatter(force, 2 * N + id2, f2.z, atomic);
    scatter(force, id3, f3.x, atomic); scatter(force, N + id3, f3.y, atomic); scatter(force, 2 * N + id3, f3.z, atomic);
    scatter(force, id4, f4.x, atomic); scatter(force, N + id4, f4.y, atomic); scatter(force, 2 * N + id4, f4.z, atomic);
}

void orc_viscoplastic(const orc_vp_params* P, int64_t N, int64_t T, const double* points, const uint32_t* tetras,
                      const double* ev, const double* vol, double* force, double* acc, double* Fp, double* vonMises,
                      double* pressure, int nthreads) {
    nthreads = set_threads(nthreads);
    if (nthreads == 1) {
        for (int64_t t = 0; t < T; t++)
            vp_one(t, P, N, T, points, tetras, ev, vol, force, acc, Fp, vonMises, pressure, 0);
        return;
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t t = 0; t < T; t++) vp_one(t, P, N, T, points, tetras, ev, vol, force, acc, Fp, vonMises, pressure, 1);
}

/* Legacy stress laws (SimulationLegacy.py:696-840, Interactions.py:101-247 and 352-408), each a
 * Cauchy stress from the displacement gradient of the node-0-based edge vectors, applied with the
 * face-area nodal forces of the same kernels, restated with node 0 as the edge base:
 *   law 0  InteractionLinearElastic.cl:30-83           sigma = 2 mu eps + lambda tr(eps) I
 *   law 1  InteractionLinearElasticDeviatoric.cl:55-80 sigma = 2 mu dev(eps) + K tr(eps) I
 *   law 2  InteractionSecondOrderNeoHookean.cl:57-96   sigma = J^-5/3 (mu + mu2 (trB J^-2/3 - 3)) dev(B) + K (J - 1) I
 *   law 3  InteractionPoroNeoHookean.cl:57-82          sigma = a B + b I (Danielsson, Parks, Boyce 2002)
 * with eps = (grad u + grad u^T)/2, B = F F^T, J = sqrt(det B). */
typedef struct {
    double E, nu;     /* law 0 */
    double shear;     /* laws 1-3 */
    double bulk;      /* laws 1-2 */
    double shear2;    /* law 2: second-order coefficient */
    double phi;       /* law 3: porous volume fraction */
} orc_legacy_params;

static void legacy_one(int64_t t, int law, const orc_legacy_params* P, int64_t N, int64_t T, const double* points,
                       const uint32_t* tetras, const double* ev, double* force, double* vonMises, double* pressure,
                       int atomic) {
    const uint32_t id1 = tetras[T + t], id2 = tetras[2 * T + t], id3 = tetras[3 * T + t], id4 = tetras[t];
    const d3 p1 = v3(points[id1], points[N + id1], points[2 * N + id1]);
    const d3 p2 = v3(points[id2], points[N + id2], points[2 * N + id2]);
    const d3 p3 = v3(points[id3], points[N + id3], points[2 * N + id3]);
    const d3 p4 = v3(points[id4], points[N + id4], points[2 * N + id4]);
    const d3 r1 = vsub(p1, p4), r2 = vsub(p2, p4), r3 = vsub(p3, p4);
    const d3 R1 = v3(ev[t], ev[T + t], ev[2 * T + t]);
    const d3 R2 = v3(ev[3 * T + t], ev[4 * T + t], ev[5 * T + t]);
    const d3 R3 = v3(ev[6 * T + t], ev[7 * T + t], ev[8 * T + t]);
    const d3 u1 = vsub(r1, R1), u2 = vsub(r2, R2), u3 = vsub(r3, R3);
    const double divisor = vdot(vcross(R1, R2), R3);
    const d3 invJ1 = vdivs(vcross(R2, R3), divisor);
    const d3 invJ2 = vdivs(vcross(R3, R1), divisor);
    const d3 invJ3 = vdivs(vcross(R1, R2), divisor);
    /* rows of grad u: du_i/dR */
    const d3 du1 = vadd(vadd(vscale(invJ1, u1.x), vscale(invJ2, u2.x)), vscale(invJ3, u3.x));
    const d3 du2 = vadd(vadd(vscale(invJ1, u1.y), vscale(invJ2, u2.y)), vscale(invJ3, u3.y));
    const d3 du3 = vadd(vadd(vscale(invJ1, u1.z), vscale(invJ2, u2.z)), vscale(invJ3, u3.z));
    const d3 E1 = v3(1, 0, 0), E2 = v3(0, 1, 0), E3 = v3(0, 0, 1);
    d3 s1, s2, s3;
    if (law == 0 || law == 1) {
        const d3 e1 = vscale(vadd(du1, v3(du1.x, du2.x, du3.x)), 0.5);
        const d3 e2 = vscale(vadd(du2, v3(du1.y, du2.y, du3.y)), 0.5);
        const d3 e3 = vscale(vadd(du3, v3(du1.z, du2.z, du3.z)), 0.5);
        const double trE = e1.x + e2.y + e3.z;
        if (law == 0) {
            const double a = 2 * P->E / (2 + 2 * P->nu);
            const double b = P->E * P->nu / ((1 + P->nu) * (1 - 2 * P->nu));
            s1 = vadd(vscale(e1, a), vscale(E1, b * trE));
            s2 = vadd(vscale(e2, a), vscale(E2, b * trE));
            s3 = vadd(vscale(e3, a), vscale(E3, b * trE));
        } else {
            const d3 d1 = vsub(e1, vscale(E1, trE / 3.)), d2 = vsub(e2, vscale(E2, trE / 3.)),
                     d3_ = vsub(e3, vscale(E3, trE / 3.));
            s1 = vadd(vscale(d1, 2 * P->shear), vscale(E1, P->bulk * trE));
            s2 = vadd(vscale(d2, 2 * P->shear), vscale(E2, P->bulk * trE));
            s3 = vadd(vscale(d3_, 2 * P->shear), vscale(E3, P->bulk * trE));
        }
    } else {
        const d3 F1 = vadd(du1, E1), F2 = vadd(du2, E2), F3 = vadd(du3, E3);
        const d3 B1 = v3(vdot(F1, F1), vdot(F1, F2), vdot(F1, F3));
        const d3 B2 = v3(vdot(F2, F1), vdot(F2, F2), vdot(F2, F3));
        const d3 B3 = v3(vdot(F3, F1), vdot(F3, F2), vdot(F3, F3));
        const double J = sqrt(vdot(vcross(B1, B2), B3));
        const double trB = B1.x + B2.y + B3.z;
        if (law == 2) {
            const double a = 1.0 / pow(J, 5.0 / 3.0) * (P->shear + P->shear2 * (trB / pow(J, 2.0 / 3.0) - 3.0));
            const double b = P->bulk * (J - 1.0);
            s1 = vadd(vscale(vsub(B1, vscale(E1, trB / 3.0)), a), vscale(E1, b));
            s2 = vadd(vscale(vsub(B2, vscale(E2, trB / 3.0)), a), vscale(E2, b));
            s3 = vadd(vscale(vsub(B3, vscale(E3, trB / 3.0)), a), vscale(E3, b));
        } else {
            const double phi = P->phi, mu = P->shear;
            const double eta = 1 + (J - 1) / phi;
            const double a = mu / 2.0 * (4.0 / J - 2.0 / (J * J) -
                                         2.0 * (phi + 2.0 * (J - 1.0)) / (pow(J, 5.0 / 3.0) * pow(eta, 1.0 / 3.0)));
            const double b = mu / 2.0 * trB *
                             (1.0 / (J * J) - 1.0 / (3.0 * pow(J, 2.0 / 3.0) * pow(eta, 1.0 / 3.0)) *
                                                  ((4.0 - phi) * eta + (1.0 - phi)) /
                                                  (phi * eta * eta + (1.0 - phi) * eta));
            s1 = vadd(vscale(B1, a), vscale(E1, b));
            s2 = vadd(vscale(B2, a), vscale(E2, b));
            s3 = vadd(vscale(B3, a), vscale(E3, b));
        }
    }
    const double trS = s1.x + s2.y + s3.z;
    const d3 ds1 = vsub(s1, vscale(E1, trS / 3.0)), ds2 = vsub(s2, vscale(E2, trS / 3.0)),
             ds3 = vsub(s3, vscale(E3, trS / 3.0));
    if (vonMises) vonMises[t] = sqrt(3. / 2. * (vdot(ds1, ds1) + vdot(ds2, ds2) + vdot(ds3, ds3)));
    if (pressure) pressure[t] = trS / 3.0;
    /* outward unit normals, face areas and the 1/3 face-force split of the same kernels */
    d3 n1 = vcross(r2, r3), n2 = vcross(r3, r1), n3 = vcross(r1, r2), n4 = vcross(vsub(r3, r1), vsub(r3, r2));
    n1 = vdivs(n1, vlen(n1));
    n2 = vdivs(n2, vlen(n2));
    n3 = vdivs(n3, vlen(n3));
    n4 = vdivs(n4, vlen(n4));
    if (vdot(n1, vscale(r1, -1.0)) < 0) n1 = vscale(n1, -1.0);
    if (vdot(n2, vscale(r2, -1.0)) < 0) n2 = vscale(n2, -1.0);
    if (vdot(n3, vscale(r3, -1.0)) < 0) n3 = vscale(n3, -1.0);
    if (vdot(n4, r1) < 0) n4 = vscale(n4, -1.0);
    const double A1 = 0.5 * vlen(vcross(r2, r3)), A2 = 0.5 * vlen(vcross(r3, r1)), A3 = 0.5 * vlen(vcross(r1, r2));
    const double A4 = 0.5 * vlen(vcross(vsub(r3, r1), vsub(r3, r2)));
    const d3 fA1 = vscale(v3(vdot(s1, n1), vdot(s2, n1), vdot(s3, n1)), A1);
    const d3 fA2 = vscale(v3(vdot(s1, n2), vdot(s2, n2), vdot(s3, n2)), A2);
    const d3 fA3 = vscale(v3(vdot(s1, n3), vdot(s2, n3), vdot(s3, n3)), A3);
    const d3 fA4 = vscale(v3(vdot(s1, n4), vdot(s2, n4), vdot(s3, n4)), A4);
    const d3 f1 = vscale(vadd(vadd(fA2, fA3), fA4), -1.0 / 3.0);
    const d3 f2 = vscale(vadd(vadd(fA3, fA4), fA1), -1.0 / 3.0);
    const d3 f3 = vscale(vadd(vadd(fA4, fA1), fA2), -1.0 / 3.0);
    const d3 f4 = vscale(vadd(vadd(fA1, fA2), fA3), -1.0 / 3.0);
    scatter(force, id1, f1.x, atomic); scatter(force, N + id1, f1.y, atomic); scatter(force, 2 * N + id1, f1.z, atomic);
    scatter(force, id2, f2.x, atomic); scatter(force, N + id2, f2.y, atomic); scatter(force, 2 * N + id2, f2.z, atomic);
    scatter(force, id3, f3.x, atomic); scatter(force, N + id3, f3.y, atomic); scatter(force, 2 * N + id3, f3.z, atomic);
    scatter(force, id4, f4.x, atomic); scatter(force, N + id4, f4.y, atomic); scatter(force, 2 * N + id4, f4.z, atomic);
}

/* InteractionLinearViscoelastic.cl:1-224 (wrapper Interactions.py:518-600): small-strain
 * generalized Maxwell (Prony) law, node-0 base.  sigma = 2 mu w_inf dev(eps) + kappa tr(eps) I
 * + sum_j h_j with h_j <- exp(-1/tau_j) h_j + 2 mu w_j (dev(eps) - dev(eps)_old), w_inf = 1 - sum w_j;
 * state per tetra: dev(eps)_old (9 SoA) and h_j (9 SoA per term, term j at 9 j T). */
typedef struct {
    double E, nu;
    int32_t nterms;
    double weights[8], taus[8];
} orc_lve_params;

static void lve_one(int64_t t, const orc_lve_params* P, int64_t N, int64_t T, const double* points,
                    const uint32_t* tetras, const double* ev, double* force, double* devold, double* transient,
                    double* vonMises, double* pressure, int atomic) {
    const uint32_t id1 = tetras[T + t], id2 = tetras[2 * T + t], id3 = tetras[3 * T + t], id4 = tetras[t];
    const d3 p1 = v3(points[id1], points[N + id1], points[2 * N + id1]);
    const d3 p2 = v3(points[id2], points[N + id2], points[2 * N + id2]);
    const d3 p3 = v3(points[id3], points[N + id3], points[2 * N + id3]);
    const d3 p4 = v3(points[id4], points[N + id4], points[2 * N + id4]);
    const d3 r1 = vsub(p1, p4), r2 = vsub(p2, p4), r3 = vsub(p3, p4);
    const d3 R1 = v3(ev[t], ev[T + t], ev[2 * T + t]);
    const d3 R2 = v3(ev[3 * T + t], ev[4 * T + t], ev[5 * T + t]);
    const d3 R3 = v3(ev[6 * T + t], ev[7 * T + t], ev[8 * T + t]);
    const d3 u1 = vsub(r1, R1), u2 = vsub(r2, R2), u3 = vsub(r3, R3);
    const double divisor = vdot(vcross(R1, R2), R3);
    const d3 invJ1 = vdivs(vcross(R2, R3), divisor);
    const d3 invJ2 = vdivs(vcross(R3, R1), divisor);
    const d3 invJ3 = vdivs(vcross(R1, R2), divisor);
    const d3 du1 = vadd(vadd(vscale(invJ1, u1.x), vscale(invJ2, u2.x)), vscale(invJ3, u3.x));
    const d3 du2 = vadd(vadd(vscale(invJ1, u1.y), vscale(invJ2, u2.y)), vscale(invJ3, u3.y));
    const d3 du3 = vadd(vadd(vscale(invJ1, u1.z), vscale(invJ2, u2.z)), vscale(invJ3, u3.z));
    const d3 E1 = v3(1, 0, 0), E2 = v3(0, 1, 0), E3 = v3(0, 0, 1);
    const d3 e1 = vscale(vadd(du1, v3(du1.x, du2.x, du3.x)), 0.5);
    const d3 e2 = vscale(vadd(du2, v3(du1.y, du2.y, du3.y)), 0.5);
    const d3 e3 = vscale(vadd(du3, v3(du1.z, du2.z, du3.z)), 0.5);
    const double trE = e1.x + e2.y + e3.z;
    const d3 de[3] = {vsub(e1, vscale(E1, trE / 3.)), vsub(e2, vscale(E2, trE / 3.)), vsub(e3, vscale(E3, trE / 3.))};
    const double mu = P->E / (2 * (1 + P->nu)), kappa = P->E / (3 * (1 - 2 * P->nu));
    double wel = 1.0;
    for (int j = 0; j < P->nterms; j++) wel -= P->weights[j];
    double sig[3][3], dot_[3][3];
    for (int i = 0; i < 3; i++) {
        const double dv[3] = {de[i].x, de[i].y, de[i].z};
        for (int k = 0; k < 3; k++) {
            const int c = 3 * i + k;
            dot_[i][k] = dv[k] - devold[c * T + t];
            devold[c * T + t] = dv[k];
            sig[i][k] = 2 * mu * wel * dv[k] + (i == k ? kappa * trE : 0.0);
        }
    }
    for (int j = 0; j < P->nterms; j++) {
        const double decay = exp(-1. / P->taus[j]);
        for (int c = 0; c < 9; c++) {
            double* h = &transient[(9 * (int64_t)j + c) * T + t];
            *h = decay * *h + 2 * mu * P->weights[j] * dot_[c / 3][c % 3];
            sig[c / 3][c % 3] += *h;
        }
    }
    const d3 s1 = v3(sig[0][0], sig[0][1], sig[0][2]), s2 = v3(sig[1][0], sig[1][1], sig[1][2]),
             s3 = v3(sig[2][0], sig[2][1], sig[2][2]);
    const double trS = s1.x + s2.y + s3.z;
    const d3 ds1 = vsub(s1, vscale(E1, trS / 3.0)), ds2 = vsub(s2, vscale(E2, trS / 3.0)),
             ds3 = vsub(s3, vscale(E3, trS / 3.0));
    if (vonMises) vonMises[t] = sqrt(3. / 2. * (vdot(ds1, ds1) + vdot(ds2, ds2) + vdot(ds3, ds3)));
    if (pressure) pressure[t] = trS / 3.0;
    d3 n1 = vcross(r2, r3), n2 = vcross(r3, r1), n3 = vcross(r1, r2), n4 = vcross(vsub(r3, r1), vsub(r3, r2));
    n1 = vdivs(n1, vlen(n1));
    n2 = vdivs(n2, vlen(n2));
    n3 = vdivs(n3, vlen(n3));
    n4 = vdivs(n4, vlen(n4));
    if (vdot(n1, vscale(r1, -1.0)) < 0) n1 = vscale(n1, -1.0);
    if (vdot(n2, vscale(r2, -1.0)) < 0) n2 = vscale(n2, -1.0);
    if (vdot(n3, vscale(r3, -1.0)) < 0) n3 = vscale(n3, -1.0);
    if (vdot(n4, r1) < 0) n4 = vscale(n4, -1.0);
    const double A1 = 0.5 * vlen(vcross(r2, r3)), A2 = 0.5 * vlen(vcross(r3, r1)), A3 = 0.5 * vlen(vcross(r1, r2));
    const double A4 = 0.5 * vlen(vcross(vsub(r3, r1), vsub(r3, r2)));
    const d3 fA1 = vscale(v3(vdot(s1, n1), vdot(s2, n1), vdot(s3, n1)), A1);
    const d3 fA2 = vscale(v3(vdot(s1, n2), vdot(s2, n2), vdot(s3, n2)), A2);
    const d3 fA3 = vscale(v3(vdot(s1, n3), vdot(s2, n3), vdot(s3, n3)), A3);
    const d3 fA4 = vscale(v3(vdot(s1, n4), vdot(s2, n4), vdot(s3, n4)), A4);
    const d3 f1 = vscale(vadd(vadd(fA2, fA3), fA4), -1.0 / 3.0);
    const d3 f2 = vscale(vadd(vadd(fA3, fA4), fA1), -1.0 / 3.0);
    const d3 f3 = vscale(vadd(vadd(fA4, fA1), fA2), -1.0 / 3.0);
    const d3 f4 = vscale(vadd(vadd(fA1, fA2), fA3), -1.0 / 3.0);
    scatter(force, id1, f1.x, atomic); scatter(force, N + id1, f1.y, atomic); scatter(force, 2 * N + id1, f1.z, atomic);
    scatter(force, id2, f2.x, atomic); scatter(force, N + id2, f2.y, atomic); scatter(force, 2 * N + id2, f2.z, atomic);
    scatter(force, id3, f3.x, atomic); scatter(force, N + id3, f3.y, atomic); scatter(force, 2 * N + id3, f3.z, atomic);
    scatter(force, id4, f4.x, atomic); scatter(force, N + id4, f4.y, atomic); scatter(force, 2 * N + id4, f4.z, atomic);
}

void orc_linear_viscoelastic(const orc_lve_params* P, int64_t N, int64_t T, const double* points,
                             const uint32_t* tetras, const double* ev, double* force, double* devold,
                             double* transient, double* vonMises, double* pressure) {
    for (int64_t t = 0; t < T; t++)
        lve_one(t, P, N, T, points, tetras, ev, force, devold, transient, vonMises, pressure, 0);
}

void orc_legacy_force(int law, const orc_legacy_params* P, int64_t N, int64_t T, const double* points,
                      const uint32_t* tetras, const double* ev, double* force, double* vonMises, double* pressure,
                      int nthreads) {
    nthreads = set_threads(nthreads);
    if (nthreads == 1) {
        for (int64_t t = 0; t < T; t++) legacy_one(t, law, P, N, T, points, tetras, ev, force, vonMises, pressure, 0);
        return;
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t t = 0; t < T; t++) legacy_one(t, law, P, N, T, points, tetras, ev, force, vonMises, pressure, 1);
}

/* ---- legacy contacts and integrators (the kernels under kernels/ used by SimulationLegacy.py:139-330, 550-617,
 *      884-940), restated as written ------------------------------------------------------------ */

/* Interactions/RoundTip.cl:1-21: sphere below its centre, a cylinder above it; prefactor 0.1 (the
 * kernel ignores its forceConst argument); force k max(|r| - R, 0) r/|r| as written */
void orc_round_tip(int64_t N, double* force, const double* points, double radius, double spherePos) {
    for (int64_t p = 0; p < N; p++) {
        const d3 pos = v3(points[p], points[N + p], points[2 * N + p]);
        d3 r = vsub(pos, v3(0, spherePos, 0));
        if (pos.y > spherePos) r.y = 0;
        const double len = vlen(r);
        const d3 n = len > 0 ? vscale(r, 1.0 / len) : v3(0, 0, 0);
        const double dis = fmax(len - radius, 0.);
        const d3 fc = vscale(n, 0.1 * dis);
        force[p] += fc.x;
        force[N + p] += fc.y;
        force[2 * N + p] += fc.z;
    }
}

/* InteractionAdhesivePlane.cl:1-14: Lennard-Jones wall on surface points, sigma = 2^(-1/6) rm */
void orc_adhesive_plane(int64_t N, double* force, const double* points, double wallPos, double adhesion,
                        double distance, const int32_t* onSurface) {
    const double sigma = pow(0.5, 1.0 / 6.0) * distance;
    for (int64_t p = 0; p < N; p++) {
        if (onSurface[p] != 1) continue;
        const double d = points[N + p] - wallPos;
        force[N + p] += 24.0 * adhesion / d * (2 * pow(sigma / d, 12) - pow(sigma / d, 6));
    }
}

/* PointZeroForce.cl:1-10: every point is pulled by -k times node 0's x and z */
void orc_point_zero_force(int64_t N, double* force, const double* points, double k) {
    for (int64_t p = 0; p < N; p++) {
        force[p] += -k * points[0];
        force[2 * N + p] += -k * points[2 * N];
    }
}

/* CalcCOM.cl:1-8 (com[0] = mean x, com[1] = mean z) and COMForce.cl:1-7 */
void orc_center_of_mass(int64_t N, const double* points, double* com) {
    for (int id = 0; id < 2; id++) {
        double sum = 0.0;
        for (int64_t i = 0; i < N; i++) sum += points[2 * id * N + i];
        com[id] = sum / (double)N;
    }
}
void orc_com_force(int64_t N, double* force, double k, const double* com) {
    for (int64_t p = 0; p < N; p++) {
        force[p] += -k * com[0];
        force[2 * N + p] += -k * com[1];
    }
}

/* OverDamped.cl:1-21: x += f (point 3 keeps x and z when point3fix), f = 0 */
void orc_over_damped(int64_t N, int point3fix, double* points, double* force) {
    for (int64_t p = 0; p < N; p++) {
        d3 f = v3(force[p], force[N + p], force[2 * N + p]);
        if (point3fix && p == 3) f.x = f.z = 0.0;
        points[p] += f.x;
        points[N + p] += f.y;
        points[2 * N + p] += f.z;
        force[p] = force[N + p] = force[2 * N + p] = 0.0;
    }
}

/* velocityVerlet2.cl:1-26: per-point masses, damping xi; fOld holds the previous acceleration */
void orc_velocity_verlet2(int64_t N, double* points, double* vel, double* force, double* fold, const double* masses,
                          double damping) {
    for (int64_t p = 0; p < N; p++) {
        d3 x = v3(points[p], points[N + p], points[2 * N + p]);
        d3 v = v3(vel[p], vel[N + p], vel[2 * N + p]);
        const d3 f = v3(force[p], force[N + p], force[2 * N + p]);
        const d3 fo = v3(fold[p], fold[N + p], fold[2 * N + p]);
        const double m = masses[p];
        v = vscale(vadd(v, vscale(vadd(fo, vdivs(f, m)), 0.5)), 1.0 / (1 + damping / 2.0));
        const d3 acc = vsub(vdivs(f, m), vscale(v, damping));
        x = vadd(vadd(x, v), vscale(acc, 0.5));
        points[p] = x.x; points[N + p] = x.y; points[2 * N + p] = x.z;
        vel[p] = v.x; vel[N + p] = v.y; vel[2 * N + p] = v.z;
        fold[p] = acc.x; fold[N + p] = acc.y; fold[2 * N + p] = acc.z;
        force[p] = force[N + p] = force[2 * N + p] = 0.0;
    }
}

/* InteractionSphereIntegral.cl:1-85 and InteractionTipIntegral.cl:1-114: the body-force density
 * exp(-k dis) n integrated over each tetra with Bower's 4-point rule; node order of the legacy
 * kernels (p1..p3 = corners 0..2 of the tetra array, p4 = corner 3 as there). */
static d3 force_sphere(d3 pos, double radius, double spherePos, double k) {
    const d3 r = vsub(pos, v3(0, spherePos, 0));
    const d3 n = vdivs(r, vlen(r));
    return vscale(n, exp(-k * (vlen(r) - radius)));
}
static d3 force_tip(d3 pos, double radius, double halfAngle, d3 tip, double k) {
    const d3 rho = v3(pos.x - tip.x, 0, pos.z - tip.z);
    if (pos.y < (tip.y - sin(halfAngle) * vlen(rho))) {
        const d3 r = vsub(pos, tip);
        const d3 n = vdivs(r, vlen(r));
        return vscale(n, exp(-k * (vlen(r) - radius)));
    }
    const d3 eRho = vdivs(rho, vlen(rho));
    const d3 n = vadd(vscale(v3(0, 1, 0), -sin(halfAngle)), vscale(eRho, cos(halfAngle)));
    const double dis = vdot(pos, n) - vdot(tip, n) - radius;
    return vscale(n, exp(-k * dis));
}
void orc_volume_quadrature(int tip_shape, int64_t N, int64_t T, double* force, const double* points,
                           const uint32_t* tetras, double radius, double halfAngle, double tipX, double tipY,
                           double tipZ, double k) {
    const double alpha = 0.58541020, beta = 0.13819669;
    const double xi[4][3] = {{alpha, beta, beta}, {beta, alpha, beta}, {beta, beta, alpha}, {beta, beta, beta}};
    for (int64_t t = 0; t < T; t++) {
        uint32_t id[4];
        d3 p[4];
        for (int a = 0; a < 4; a++) {
            id[a] = tetras[a * T + t];
            p[a] = v3(points[id[a]], points[N + id[a]], points[2 * N + id[a]]);
        }
        const d3 r1 = vsub(p[0], p[3]), r2 = vsub(p[1], p[3]), r3 = vsub(p[2], p[3]);
        const double volume = fabs(vdot(r1, vcross(r2, r3)));
        d3 F[4] = {v3(0, 0, 0), v3(0, 0, 0), v3(0, 0, 0), v3(0, 0, 0)};
        for (int q = 0; q < 4; q++) {
            const double Nq[4] = {xi[q][0], xi[q][1], xi[q][2], 1.0 - xi[q][0] - xi[q][1] - xi[q][2]};
            const d3 x = vadd(vadd(vadd(vscale(p[0], Nq[0]), vscale(p[1], Nq[1])), vscale(p[2], Nq[2])),
                              vscale(p[3], Nq[3]));
            const d3 fq = tip_shape ? force_tip(x, radius, halfAngle, v3(tipX, tipY, tipZ), k)
                                    : force_sphere(x, radius, tipY, k);
            for (int a = 0; a < 4; a++) F[a] = vadd(F[a], vscale(fq, Nq[a]));
        }
        for (int a = 0; a < 4; a++) {
            force[id[a]] += 0.25 * volume * F[a].x;
            force[N + id[a]] += 0.25 * volume * F[a].y;
            force[2 * N + id[a]] += 0.25 * volume * F[a].z;
        }
    }
}

/* Bower's 3-point (edge mid-point) rule of an area density over each surface triangle, y component:
 * node a of triangle t gets area/3 * sum_q f(y_q - wall) N_a(xi_q).  law 0: the Lennard-Jones density
 * 24 A/d (2 (s/d)^12 - (s/d)^6), s = 2^-1/6 p (InteractionAdhesivePlaneSurfaceIntegral.cl:1-56);
 * law 1: exp(-p d) - A exp(-d) (InteractionHardAdhesivePlaneSurfaceIntegral.cl:1-55). */
static void surface_density_rule(int law, int64_t N, int64_t NT, double* force, const double* points,
                                 const uint32_t* tri, double wallPos, double A, double p_) {
    const double sigma = pow(0.5, 1.0 / 6.0) * p_;
    const double xi[3][2] = {{0.5, 0.0}, {0.0, 0.5}, {0.5, 0.5}};
    for (int64_t t = 0; t < NT; t++) {
        uint32_t id[3];
        d3 p[3];
        for (int a = 0; a < 3; a++) {
            id[a] = tri[a * NT + t];
            p[a] = v3(points[id[a]], points[N + id[a]], points[2 * N + id[a]]);
        }
        const double area = 0.5 * vlen(vcross(vsub(p[0], p[2]), vsub(p[1], p[2])));
        double F[3] = {0, 0, 0};
        for (int q = 0; q < 3; q++) {
            const double Nq[3] = {xi[q][0], xi[q][1], 1.0 - xi[q][0] - xi[q][1]};
            const double y = Nq[0] * p[0].y + Nq[1] * p[1].y + Nq[2] * p[2].y;
            const double d = y - wallPos;
            const double fq = law == 0 ? 24.0 * A / d * (2 * pow(sigma / d, 12) - pow(sigma / d, 6))
                                       : exp(-p_ * d) - A * exp(-d);
            for (int a = 0; a < 3; a++) F[a] += fq * Nq[a];
        }
        for (int a = 0; a < 3; a++) force[N + id[a]] += area / 3.0 * F[a];
    }
}

void orc_adhesive_plane_surface(int64_t N, int64_t NT, double* force, const double* points, const uint32_t* tri,
                                double wallPos, double adhesion, double distance) {
    surface_density_rule(0, N, NT, force, points, tri, wallPos, adhesion, distance);
}

void orc_hard_adhesive_plane_surface(int64_t N, int64_t NT, double* force, const double* points, const uint32_t* tri,
                                     double wallPos, double adhesion, double forceConst) {
    surface_density_rule(1, N, NT, force, points, tri, wallPos, adhesion, forceConst);
}

/* Sphere.cl:1-17.  OpenCL normalize() of a zero vector returns the zero vector. */
void orc_sphere(int64_t N, double k, double* force, const double* points, double radius, double spherePos) {
    for (int64_t p = 0; p < N; p++) {
        const d3 pos = v3(points[p], points[N + p], points[2 * N + p]);
        const d3 r = vsub(pos, v3(0, spherePos, 0));
        const double len = vlen(r);
        const d3 n = len > 0 ? vscale(r, 1.0 / len) : v3(0, 0, 0);
        const double dis = fmax(radius - len, 0.);
        const d3 fc = vscale(n, k * dis);
        force[p] += fc.x;
        force[N + p] += fc.y;
        force[2 * N + p] += fc.z;
    }
}

/* Substrate.cl:1-7 */
void orc_substrate(int64_t N, double k, double* force, const double* points, double wallPos) {
    for (int64_t p = 0; p < N; p++) {
        const double dis = fmax(wallPos - points[N + p], 0.);
        force[N + p] += k * dis;
    }
}

/* Plane.cl:1-7 */
void orc_plane(int64_t N, double k, double* force, const double* points, double wallPos) {
    for (int64_t p = 0; p < N; p++) {
        const double dis = fmax(points[N + p] - wallPos, 0.);
        force[N + p] += -k * dis;
    }
}

/* VelocityVerlet.cl:1-36 (xi = 0.005, unit mass, dt = 1) */
void orc_velocity_verlet(int64_t N, int fixTopBottom, double* points, double* vel, double* force, double* fold) {
    const double xi = 0.005;
    for (int64_t p = 0; p < N; p++) {
        d3 x = v3(points[p], points[N + p], points[2 * N + p]);
        d3 v = v3(vel[p], vel[N + p], vel[2 * N + p]);
        const d3 f = v3(force[p], force[N + p], force[2 * N + p]);
        const d3 fo = v3(fold[p], fold[N + p], fold[2 * N + p]);
        v = vscale(vadd(vscale(v, 1 - 0.5 * xi), vscale(vadd(f, fo), 0.5)), 1.0 / (1 + 0.5 * xi));
        x = vadd(vadd(x, vscale(v, 1 - 0.5 * xi)), vscale(f, 0.5));
        if (fixTopBottom && (p == 3 || p == 4)) {
            x.x = 0.0;
            x.z = 0.0;
        }
        points[p] = x.x; points[N + p] = x.y; points[2 * N + p] = x.z;
        vel[p] = v.x; vel[N + p] = v.y; vel[2 * N + p] = v.z;
        fold[p] = f.x; fold[N + p] = f.y; fold[2 * N + p] = f.z;
        force[p] = 0.0; force[N + p] = 0.0; force[2 * N + p] = 0.0;
    }
}

/* Simulation.py:198-212 in simulation units: -sum_p (k * max(R - |x-c|, 0) * n)_y and the
 * maximum penetration.  indenter 1 = sphere at (0, pos, 0); 2 = plane at y = pos
 * (Simulation.py:146-153, which returns +sum(k * max(y - pos, 0))). */
void orc_tip_force(int64_t N, const double* points, int indenter, double k, double radius, double pos,
                   double* out_force, double* out_maxdis) {
    double s = 0, mx = 0;
    for (int64_t p = 0; p < N; p++) {
        if (indenter == 1) {
            const d3 r = vsub(v3(points[p], points[N + p], points[2 * N + p]), v3(0, pos, 0));
            const double len = vlen(r);
            const double dis = fmax(radius - len, 0.);
            const double ny = len > 0 ? r.y / len : 0.0;
            s += -(k * dis * ny);
            if (dis > mx) mx = dis;
        } else {
            const double dis = fmax(points[N + p] - pos, 0.);
            s += k * dis;
            if (dis > mx) mx = dis;
        }
    }
    *out_force = s;
    *out_maxdis = mx;
}

/* One simulation run with the plugin sequence Tetra|Viscoplastic, Substrate, Sphere|Plane,
 * VelocityVerlet (Simulation.py:232-237), with the indenter law of Simulation.py:117-132 /
 * 173-183.  tip_hist[s] receives the indenter force of step s (computed from the positions at
 * the start of step s, i.e. forceSI at time s). */
typedef struct {
    int model;              /* ORC_SVKK / ORC_SVK / ORC_MR, or 3 = viscoplastic */
    double E, nu, ratio;    /* elastic constants in simulation units */
    orc_vp_params vp;       /* model 3 */
    int substrate;          /* 1 = on */
    double substrate_k, wall;
    int indenter;           /* 0 none, 1 sphere, 2 plane */
    double ind_k, ind_radius, ind_start, ind_speed, ind_move_steps;
    int fix_top_bottom;
    int nthreads;
    orc_legacy_params legacy; /* models 4-7: legacy laws 0-3 */
} orc_run_params;

static double indenter_pos(const orc_run_params* R, int64_t time) {
    /* Simulation.py:173-183 (sphere) and :117-132 (plane) */
    if ((double)time <= R->ind_move_steps) return R->ind_start - R->ind_speed * (double)time;
    return R->ind_start - R->ind_speed * R->ind_move_steps;
}

void orc_run(const orc_run_params* R, int64_t N, int64_t T, double* points, const uint32_t* tetras, const double* ev,
             const double* vol, double* vel, double* force, double* fold, double* acc, double* Fp, int64_t step0,
             int64_t nsteps, double* tip_hist) {
    for (int64_t s = 0; s < nsteps; s++) {
        const int64_t time = step0 + s;
        if (R->model == 3)
            orc_viscoplastic(&R->vp, N, T, points, tetras, ev, vol, force, acc, Fp, NULL, NULL, R->nthreads);
        else if (R->model >= 4)
            orc_legacy_force(R->model - 4, &R->legacy, N, T, points, tetras, ev, force, NULL, NULL, R->nthreads);
        else
            orc_tetra_force(R->model, R->E, R->nu, R->ratio, N, T, points, tetras, ev, vol, force, R->nthreads);
        if (R->substrate) orc_substrate(N, R->substrate_k, force, points, R->wall);
        if (R->indenter) {
            const double pos = indenter_pos(R, time);
            if (tip_hist) {
                double fz, mx;
                orc_tip_force(N, points, R->indenter, R->ind_k, R->ind_radius, pos, &fz, &mx);
                tip_hist[s] = fz;
            }
            if (R->indenter == 1)
                orc_sphere(N, R->ind_k, force, points, R->ind_radius, pos);
            else
                orc_plane(N, R->ind_k, force, points, pos);
        }
        orc_velocity_verlet(N, R->fix_top_bottom, points, vel, force, fold);
    }
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
