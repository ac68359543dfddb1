/*
 * nfx.h -- C ABI of libnfx.so, the MI355X (gfx950) implementation of noFluidX3D's explicit
 * tetrahedral-FEM time step.
 *
 * This boundary replaces PyOpenCL, nofluidx3d/openCL.py, nofluidx3d/KernelBuilder.py and the
 * OpenCL kernels under nofluidx3d/kernels/ (reference paths).  The Python classes Simulation,
 * TetraCell, Interaction, Tetra, Sphere, Substrate, Plane, VelocityVerlet and Bridge keep their
 * API and bind these entry points with ctypes.
 *
 * Conventions
 * - Every array argument is a DEVICE pointer owned by the caller (torch tensors in the Python
 *   layer) unless the name ends in _host.  The library never frees caller memory.
 * - Layouts are the reference's structure-of-arrays: component c of node n at c*N + n
 *   (Simulation.py:82-84), corner a of tetra t at a*T + t (Simulation.py:85-87), matrix entry
 *   (i,j) of tetra t at (3*i+j)*T + t (Tetra.py:53-58).
 * - State is FP32 on the device; the one-time inverse-shape-matrix/volume precompute and the
 *   tip-force reduction are FP64.
 * - All calls are asynchronous on the context's stream; nfx_stream_sync waits.  One host thread
 *   drives one context (the reference drives one in-order queue, openCL.py:34).
 * - Return value 0 on success, a negative NFX_ERR_* code otherwise; nfx_last_error() then holds
 *   a thread-local message.  The reference's hard exits (openCL.py:30-32, util/IO.py:17-28) are
 *   reproduced by the Python layer only.
 */
#ifndef NFX_H
#define NFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NFX_OK 0
#define NFX_ERR_ARG (-1)
#define NFX_ERR_DEVICE (-2)
#define NFX_ERR_HIP (-3)
#define NFX_ERR_RANGE (-4)

/* material models: Tetra.materialModel (interactions/Tetra.py:25-33); NEO_HOOKEAN is MOONEY_RIVLIN
 * with ratio 1 (Tetra.py:77-78, tetra.cl:115) */
#define NFX_SVKK 0
#define NFX_SVK 1
#define NFX_MR 2
#define NFX_VISCOPLASTIC 3
/* legacy stress laws (SimulationLegacy.py:696-840; Interactions.py:101-247, 352-408), node-0 base */
#define NFX_LINEAR_ELASTIC 4     /* InteractionLinearElastic.cl: E, nu */
#define NFX_LINEAR_DEVIATORIC 5  /* InteractionLinearElasticDeviatoric.cl: shear, bulk */
#define NFX_NEO_HOOKEAN_2 6      /* InteractionSecondOrderNeoHookean.cl: shear, shear2, bulk */
#define NFX_PORO_NEO_HOOKEAN 7   /* InteractionPoroNeoHookean.cl: shear, volume_fraction */

#define NFX_SCATTER_ATOMIC 0 /* float atomics, as the reference's atomicAdd (atomicAdd.cl:1-11) */
#define NFX_SCATTER_CSR 1    /* deterministic node<-corner gather, fixed summation order */
#define NFX_SCATTER_TILED 2  /* deterministic two-level gather: per-tile LDS sums, then per-node sums */

#define NFX_INDENTER_NONE 0
#define NFX_INDENTER_SPHERE 1 /* Sphere.cl */
#define NFX_INDENTER_PLANE 2  /* Plane.cl */

typedef struct nfx_ctx nfx_ctx;

/* Elastic constants in simulation units: the #defines of interactions/Tetra.py:66-79, and the
 * per-tetra parameter arrays of the legacy laws (uniform here). */
typedef struct {
    int32_t model; /* NFX_SVKK, NFX_SVK, NFX_MR or a legacy law NFX_LINEAR_ELASTIC..NFX_PORO_NEO_HOOKEAN */
    double E, nu, ratio;
    double shear, bulk, shear2, volume_fraction; /* legacy laws */
} nfx_elastic;

/* Finite-strain viscoplastic parameters (uniform per tetra; Interactions.py:270-277 builds eight
 * uniform per-tetra arrays, SimulationLegacy.py:842-882 sets their values). */
typedef struct {
    double mu1, mu2, kappa;
    double flow_exponent;       /* m */
    double yield_stress;        /* Y: sigma0 = Y (1 + epsbar/eps0)^n */
    double flow_rate;           /* eps0dot per step */
    double hardening_threshold; /* eps0 */
    double hardening_exponent;  /* n */
    double yield_offset;        /* sigma_y >= 0; rate uses max(svm - sigma_y, 0); 0 = reference law */
} nfx_viscoplastic;

/* Everything one fused step of one cell touches.
 *
 * Position state.  The fused and batched steps keep each node's DISPLACEMENT u = x - X0 from its
 * reference position X0 (ref, FP32; the positions the FP64 precompute of dminv saw) in disp, and form
 * the deformation gradient as F = I + [u1-u0 | u2-u0 | u3-u0] Dm^-1.  An FP32 absolute position near
 * |x| = 50 carries ~4e-6 of rounding, which F = r Dm^-1 would pass on to the strain; a displacement
 * keeps its rounding relative to its own (small) size.  nfx_points_from_disp / nfx_disp_from_points
 * convert to and from the absolute positions the per-plugin kernels take (the reference's buffers).
 *
 * Viscoplastic state.  vp_state holds the plastic deviation Q = Fp^-1 - I (9T SoA, zero at the
 * start), or with vp_sym the symmetric deviation Gd = Fp^-1 Fp^-T - I (6T: xx, yy, zz, xy, yz, xz).
 * Stored as Fp itself in FP32, each step's small plastic increment would be rounded against Fp's
 * O(1) entries; as a deviation it keeps its own relative precision.  nfx_vp_fp_from_q /
 * nfx_vp_q_from_fp convert to and from the reference's plasticDeformationTensor. */
typedef struct {
    int64_t N, T;
    const uint32_t* tetras; /* 4T  SoA */
    const float* dminv;     /* 9T  SoA, inverse reference shape matrix */
    const float* v0;        /* T   reference volume (models NFX_SVKK, NFX_SVK, NFX_MR) */
    const float* ref;       /* 3N  SoA reference positions X0 */
    float* disp;            /* 3N  SoA displacements u = x - X0 */
    float* vel;             /* 3N */
    float* force;           /* 3N  (atomic scatter target; cleared by the node update) */
    float* fold;            /* 3N */
    int32_t model;          /* NFX_SVKK / NFX_SVK / NFX_MR / NFX_VISCOPLASTIC / legacy laws */
    double E, nu, ratio;
    nfx_viscoplastic vp;
    float* vp_state;  /* model NFX_VISCOPLASTIC: Q (9T) or, with vp_sym, Gd (6T) */
    float* epsbar;    /* T; may be NULL on the tiled path when the law is linear (flow exponent 1,
                         hardening exponent 0: the flow does not read it), and then is not kept */
    float* von_mises; /* T, nullable: when set (with pressure), nfx_run_steps writes the per-tetra von
                         Mises stress and pressure of the last step of the call (viscoplastic and
                         legacy laws) */
    float* pressure;  /* T, nullable */
    int32_t scatter;  /* NFX_SCATTER_ATOMIC, NFX_SCATTER_CSR or NFX_SCATTER_TILED */
    const uint32_t* csr_offsets; /* N+1, incidences of node n at [off[n], off[n+1]) */
    const uint32_t* csr_slots;   /* 4T, slot = 4*t + corner, ascending per node */
    float* corners;              /* 12T scratch: force of corner a of tetra t at 12*t + 3*a + c */
    int32_t tile_k;              /* NFX_SCATTER_TILED: tetra per tile (128, 256 or 512; 128 fills the chip
                                    for small meshes) */
    int32_t tile_max_nodes;      /* MX: tile-node capacity per tile (largest distinct-node count,
                                    at most 256) */
    const uint32_t* tile_count;  /* ntiles: distinct nodes of tile c (<= MX) */
    const uint32_t* tile_nodes;  /* ntiles*MX: node ids, ascending within a tile, padded */
    const uint16_t* tile_lend;   /* ntiles*MX: end of each tile-node's run in the tile's slot list */
    const uint16_t* tile_corners; /* 4*tile_k*ntiles: for corner a of tetra t (entry 4*(t-t0)+a),
                                    (rank << 8) | tile-node: the tile-local node (< 256) and the
                                    corner's place among that node's corners in slot order */
    const uint32_t* tile_dst;    /* ntiles*MX: position of each tile-node's partial in node order */
    const uint32_t* node_off;    /* N+1: partials of node n at [node_off[n], node_off[n+1]) */
    float* partials;             /* 3*M scratch, node order (tile order within a node) */
    int32_t substrate;
    double substrate_k, wall;
    int32_t indenter; /* NFX_INDENTER_* */
    double ind_k, ind_radius, ind_start, ind_speed, ind_move_steps;
    int32_t fix_top_bottom;
    double shear, bulk, shear2, volume_fraction; /* legacy laws (as in nfx_elastic) */
    int32_t vp_sym;    /* NFX_VISCOPLASTIC with NFX_SCATTER_TILED: 1 = vp_state holds Gd.  The stress,
                          flow and forces depend on Gd alone, so the run is the same with a third less
                          state; Fp is not recoverable from it */
    int32_t stream_nt; /* NFX_SCATTER_TILED with NFX_MR or NFX_VISCOPLASTIC: 1 = the force pass reads
                          and writes the per-tetra state and tile tables with the nontemporal cache
                          policy (faster when they exceed the memory-side cache), 0 = default policy;
                          an autotuner choice, results are identical */
    int32_t shape_from_ref; /* NFX_SCATTER_TILED: 1 = the force pass forms each tetra's Dm^-1 and V0 from the
                               reference positions ref of its nodes (FP32 cross products) instead of reading
                               dminv and v0: 40 B per tetra less traffic for ~60 FLOPs.  Both describe the
                               same reference shape; they differ by FP32 rounding (~1e-7 relative).
                               dminv must still be set (the other scatter modes read it) */
    uint32_t* vp_active;    /* NFX_VISCOPLASTIC with NFX_SCATTER_TILED, nullable: per tile (ceil(T/tile_k)), 0 =
                               every tetra of the tile has a zero plastic state (vp_state and epsbar).  The
                               force pass then reads no state for the tile and sets the flag once one of its
                               tetra flows.  A caller that writes vp_state or epsbar itself recomputes the
                               flags with nfx_vp_tile_flags.  Independently of the flags, a tetra writes its
                               state back only in a step in which it flows (the state is unchanged otherwise).
                               Results are bitwise those without flags */
    /* Node-fused steps (NFX_SCATTER_TILED with NFX_MR, NFX_SVK, NFX_SVKK or NFX_VISCOPLASTIC; the four
       below set or all NULL; every node must be a corner of some tetra).  With them, a call of n >= 2 steps
       runs n - 1 of its node passes inside the next step's force pass: each tile first updates the
       nodes it touches from the previous pass's partials (every tile that touches a node computes
       the same values; the node's first tile stores them) and then computes its forces from the new
       displacements.  A step is then one launch instead of two.  The launches alternate between
       (disp, vel, partials) and these second buffers; between calls the state is in disp, vel and
       fold as without them.  Scratch: their contents between calls mean nothing. */
    float* disp_b;     /* 3N */
    float* vel_b;      /* 3N */
    float* partials_b; /* 3*M, as partials */
    const uint32_t* tile_range; /* node-fused steps: ntiles*MX*2, for tile-node j of tile c the range
                                   [node_off[n], node_off[n+1]) of its node n's partials at 2*(c*MX + j) */
} nfx_plan;

/* Per-cell parameters of a batch, in that cell's simulation units.  The batched material is the
 * Mooney-Rivlin/Neo-Hookean law, optionally with the Maxwell-type viscoplastic flow (m = 1, n = 0:
 * rate = vp_rate * max(svm - vp_offset, 0) / vp_yield per step). */
typedef struct {
    float mu1, mu2, kappa;
    float vp_rate, vp_yield, vp_offset;
    float wall, substrate_k;
    float tip_radius, tip_k;
    double tip_start, tip_speed, tip_move_steps;
} nfx_cell;

/* A batch is the concatenation of many cells' meshes, with tiles aligned to cells (a tile never
 * spans two cells, so each cell sums its forces in the same order as when it runs alone).
 * Batched steps keep the damped velocity-Verlet state as (u, w) with w = a v + f/2 of the last
 * step (a = 1 - xi/2) in plan.vel: the same recurrence as VelocityVerlet.cl (u += w,
 * w' = (a/b)(w + f'/2) + f'/2, b = 1 + xi/2) without the previous force; plan.fold is unused.
 * Zero v and f give w = 0. */
typedef struct {
    int32_t ncells;
    const uint16_t* tile_ntet;     /* ntiles: tetra in each tile */
    const uint32_t* tile_cell;     /* ntiles: owning cell of each tile */
    const uint32_t* node_cell;     /* N: owning cell of each node */
    const uint32_t* cell_node_off; /* ncells+1: node range of each cell */
    const nfx_cell* cells;         /* ncells (device) */
    double* curves;                /* ncells*ncurve FP64 indenter-force samples */
    int32_t ncurve;
    int64_t curve_steps; /* sample k is taken at the first step s with floor(s*ncurve/curve_steps) == k */
    /* nullable, all four or none: the tile-node ids and partial positions as 16-bit offsets from
     * per-tile bases (a tile never leaves its cell; every offset below 65536), which the force pass
     * then reads in place of plan.tile_nodes and plan.tile_dst */
    const uint16_t* tile_nodes16;   /* ntiles*MX: tile_nodes - tile_node_base[tile] */
    const uint16_t* tile_dst16;     /* ntiles*MX: tile_dst - tile_dst_base[tile] (padding 0) */
    const uint32_t* tile_node_base; /* ntiles */
    const uint32_t* tile_dst_base;  /* ntiles */
} nfx_batch;

/* ---- device and queue: replaces openCL.py:17-50 -------------------------------------------- */
int nfx_device_count(int* count);                        /* openCL.py:20-28 device walk */
int nfx_init(int device_id, void* stream, nfx_ctx** ctx); /* openCL.py:17-34 initializeOpenCLObjects;
                                                             stream: hipStream_t to use, or NULL for a
                                                             library-owned default-priority stream */
int nfx_destroy(nfx_ctx* ctx);
int nfx_stream_sync(nfx_ctx* ctx); /* getCommandQueue().finish(), Simulation.py:266 */
const char* nfx_last_error(void);
int nfx_device_name(nfx_ctx* ctx, char* buf, int32_t len);

/* ---- one-time FP64 precompute: the inversion of Dm = [R1|R2|R3] that tetra.cl:47-53 repeats
 *      every step.  Inputs are Tetra's buffers (interactions/Tetra.py:53-58): edge_vectors 9T FP64
 *      SoA (TetraCell.py:26-30) and volumes T FP64 (TetraCell.py:40-41).  Outputs FP32. ---------- */
int nfx_precompute_f64(nfx_ctx* ctx, const double* edge_vectors, const double* volumes, int64_t T, float* dminv,
                       float* v0);

/* ---- per-plugin kernels, one per reference kernel ------------------------------------------ */
/* tetra.cl:14-182 calcTetraForce with the atomic node scatter (force += ...) */
int nfx_tetra_force(nfx_ctx* ctx, const nfx_elastic* mat, const float* points, const uint32_t* tetras,
                    const float* dminv, const float* v0, int64_t N, int64_t T, float* force);
/* same forces written per corner (no scatter): corners[12*t + 3*a + c] */
int nfx_tetra_corner_forces(nfx_ctx* ctx, const nfx_elastic* mat, const float* points, const uint32_t* tetras,
                            const float* dminv, const float* v0, int64_t N, int64_t T, float* corners);
/* the legacy laws' kernels (InteractionLinearElastic.cl, InteractionLinearElasticDeviatoric.cl,
 * InteractionSecondOrderNeoHookean.cl, InteractionPoroNeoHookean.cl: Cauchy stress, face-area forces,
 * vonMises and pressure), node-0 edge base; atomic scatter when corners is NULL, per-corner output
 * otherwise; von_mises/pressure nullable */
int nfx_legacy_force(nfx_ctx* ctx, const nfx_elastic* mat, const float* points, const uint32_t* tetras,
                     const float* dminv, int64_t N, int64_t T, float* von_mises, float* pressure, float* force,
                     float* corners);
/* InteractionLinearViscoelastic.cl (wrapper Interactions.py:518-600), node-0 edge base: small-strain
 * generalized Maxwell law sigma = 2 mu w_inf dev(eps) + kappa tr(eps) I + sum_j h_j,
 * h_j <- exp(-1/tau_j) h_j + 2 mu w_j (dev(eps) - dev(eps)_old), w_inf = 1 - sum_j w_j; tau_j in steps.
 * State: dev_old 9T SoA, transient 9*nterms*T (term j, entry c of tetra t at (9j + c)T + t), zero at
 * the start.  Atomic scatter when corners is NULL; von_mises/pressure nullable. */
typedef struct {
    double E, nu;
    int32_t nterms; /* 0..8 */
    double weights[8];
    double taus[8];
} nfx_viscoelastic;
int nfx_linear_viscoelastic_force(nfx_ctx* ctx, const nfx_viscoelastic* p, const float* points, const uint32_t* tetras,
                                  const float* dminv, int64_t N, int64_t T, float* dev_old, float* transient,
                                  float* von_mises, float* pressure, float* force, float* corners);
/* InteractionFiniteStrainViscoplastic.cl:1-220 (node-0 edge base); atomic scatter when corners
 * is NULL, per-corner output otherwise */
int nfx_viscoplastic_force(nfx_ctx* ctx, const nfx_viscoplastic* vp, const float* points, const uint32_t* tetras,
                           const float* dminv, int64_t N, int64_t T, float* Fp, float* epsbar, float* von_mises,
                           float* pressure, float* force, float* corners);
/* The fused viscoplastic state Q = Fp^-1 - I and the reference's Fp (plasticDeformationTensor,
 * Interactions.py:279-281), 9T SoA FP32 each */
int nfx_vp_q_from_fp(nfx_ctx* ctx, const float* Fp, int64_t T, float* Q);
int nfx_vp_fp_from_q(nfx_ctx* ctx, const float* Q, int64_t T, float* Fp);
/* plan.vp_active from plan.vp_state and plan.epsbar: 1 for a tile with any nonzero entry (no
 * reference counterpart: the reference updates every tetra's state every step) */
int nfx_vp_tile_flags(nfx_ctx* ctx, const nfx_plan* plan);
/* absolute positions x = X0 + u (3N SoA) from displacements, and back (no reference counterpart:
 * the reference keeps absolute FP64 positions) */
int nfx_points_from_disp(nfx_ctx* ctx, const float* ref, const float* disp, int64_t N, float* points);
int nfx_disp_from_points(nfx_ctx* ctx, const float* ref, const float* points, int64_t N, float* disp);
/* A batch's geometry from its cells' recentred FP64 positions p_rel (3N SoA, micrometres) and per-cell
 * scales (RadiusSIM / |p[1].x|): the FP32 reference positions X0 = fl32(p_rel * scale) and, from X0 in
 * FP64, Dm^-1 and V0, exactly as the host construction computes them (TetraCell.py:15-41); padding
 * tetra get the identity shape (a reset on the device) */
int nfx_batch_geometry(nfx_ctx* ctx, const double* p_rel, const double* cell_scale, const uint32_t* node_cell,
                       const uint32_t* tetras, const uint16_t* tile_ntet, int32_t tile_k, int64_t N, int64_t T,
                       float* ref, float* dminv, float* v0);
/* deterministic gather: force[c*N+n] += sum over slots of node n, in slot order */
int nfx_gather_corners(nfx_ctx* ctx, const uint32_t* csr_offsets, const uint32_t* csr_slots, const float* corners,
                       int64_t N, float* force);
/* host-side tiling for NFX_SCATTER_TILED (host pointers).  Call once with tile_nodes == NULL to get
 * *max_nodes (MX), *m_total and tile_count, then again with arrays of ntiles*MX entries
 * (tile_nodes, tile_lend, tile_dst), 4*tile_k*ntiles (tile_corners) and N+1 (node_off).  The second
 * call fails when MX > 256 or a node has more than 256 corners in one tile (use tile_k 256). */
int nfx_build_tiles_host(const uint32_t* tetras_host, int64_t N, int64_t T, int32_t tile_k, int32_t* max_nodes,
                         int64_t* m_total, uint32_t* tile_count, uint32_t* tile_nodes, uint16_t* tile_lend,
                         uint16_t* tile_corners, uint32_t* tile_dst, uint32_t* node_off);
/* host-side CSR build (host pointers): offsets N+1, slots 4T ascending per node */
int nfx_build_node_csr_host(const uint32_t* tetras_host, int64_t N, int64_t T, uint32_t* offsets_host,
                            uint32_t* slots_host);
int nfx_sphere_force(nfx_ctx* ctx, float* force, const float* points, int64_t N, double k, double radius,
                     double sphere_pos);                                                       /* Sphere.cl */
int nfx_substrate_force(nfx_ctx* ctx, float* force, const float* points, int64_t N, double k, double wall_pos); /* Substrate.cl */
int nfx_plane_force(nfx_ctx* ctx, float* force, const float* points, int64_t N, double k, double wall_pos);     /* Plane.cl */
int nfx_velocity_verlet(nfx_ctx* ctx, float* points, float* vel, float* force, float* fold, int64_t N,
                        int32_t fix_top_bottom); /* VelocityVerlet.cl */
/* Simulation.py:146-153 / 198-212 forceSI and "Max dis" in simulation units, reduced in FP64 on the
 * device: out2[0] = indenter force, out2[1] = max penetration (device pointer to 2 doubles).  With
 * ref non-NULL, points holds displacements from ref and the penetrations are taken in FP64 on
 * x = ref + points. */
int nfx_tip_force(nfx_ctx* ctx, const float* points, const float* ref, int64_t N, int32_t indenter, double k,
                  double radius, double pos, double* out2);

/* nfx_tip_force without a context and without any allocation: the same FP64 reduction (bitwise the
 * same result) with its scratch -- 2048 per-block (sum, max) pairs and a completion counter, 32.25
 * KiB -- in a fixed array of the library that each GPU holds for the life of the process.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Enqueued on `stream`
 * (a hipStream_t, NULL for the null stream) on the current device.  N up to 2^31 - 1 (the grid is
 * capped at 2048 blocks, so every N in range fits the scratch).  All calls on one GPU share that
 * scratch: calls that overlap in time on one GPU race. */
int nfx_tip_force_noalloc(const float* points, const float* ref, int64_t N, int32_t indenter, double k, double radius,
                          double pos, double* out2, void* stream);

/* ---- legacy contacts and integrators (the kernels SimulationLegacy.py:139-330, 550-617 and
 *      884-940 use), restated as written; force/points/vel/fold are 3N SoA FP32 ------------------ */
/* Interactions/RoundTip.cl: sphere below its centre, cylinder above; force 0.1 max(|r| - R, 0) r/|r| */
int nfx_round_tip_force(nfx_ctx* ctx, float* force, const float* points, int64_t N, double radius, double sphere_pos);
/* InteractionAdhesivePlane.cl: Lennard-Jones wall, sigma = 2^-1/6 distance, on points with on_surface == 1 */
int nfx_adhesive_plane_force(nfx_ctx* ctx, float* force, const float* points, int64_t N, double wall_pos,
                             double adhesion, double distance, const int32_t* on_surface);
/* PointZeroForce.cl: force -= k (x of node 0, 0, z of node 0) on every node */
int nfx_point_zero_force(nfx_ctx* ctx, float* force, const float* points, int64_t N, double k);
/* CalcCOM.cl: com[0] = mean x, com[1] = mean z (FP64, device, deterministic order) */
int nfx_center_of_mass(nfx_ctx* ctx, const float* points, int64_t N, double* com);
/* COMForce.cl: force -= k (com[0], 0, com[1]) on every node */
int nfx_com_force(nfx_ctx* ctx, float* force, int64_t N, double k, const double* com);
/* OverDamped.cl: x += f (node 3 keeps x and z when point3fix), f = 0 */
int nfx_over_damped(nfx_ctx* ctx, float* points, float* force, int64_t N, int32_t point3fix);
/* velocityVerlet2.cl: per-node masses (N), damping; fold holds the previous acceleration */
int nfx_velocity_verlet2(nfx_ctx* ctx, float* points, float* vel, float* force, float* fold, const float* masses,
                         int64_t N, double damping);
/* InteractionSphereIntegral.cl (tip_shape 0: sphere at (0, tip_y, 0)) and InteractionTipIntegral.cl
 * (tip_shape 1: sphere of radius R with a cone of half angle half_angle, apex region at tip_x/y/z):
 * body force exp(-k dis) n by Bower's 4-point rule per tetra (scaled by |det r| / 4 as written);
 * float-atomic scatter into force, or per-corner records (corners[12*t + 3*a + c]) when corners != NULL */
int nfx_volume_integral_force(nfx_ctx* ctx, int32_t tip_shape, const float* points, const uint32_t* tetras, int64_t N,
                              int64_t T, double radius, double half_angle, double tip_x, double tip_y, double tip_z,
                              double k, float* force, float* corners);
/* InteractionAdhesivePlaneSurfaceIntegral.cl: Lennard-Jones area density over the surface triangles
 * (3 x NT SoA) by the 3-point edge-midpoint rule, y component, float-atomic scatter */
int nfx_adhesive_plane_surface_force(nfx_ctx* ctx, float* force, const float* points, const uint32_t* triangles,
                                     int64_t N, int64_t NT, double wall_pos, double adhesion, double distance);
/* InteractionHardAdhesivePlaneSurfaceIntegral.cl (wrapper Interactions.py:778-817, set up by
 * SimulationLegacy.py:657-692): area density exp(-force_const d) - adhesion exp(-d), d = y - wall_pos,
 * by the same 3-point rule over the surface triangles, y component, float-atomic scatter */
int nfx_hard_adhesive_plane_surface_force(nfx_ctx* ctx, float* force, const float* points, const uint32_t* triangles,
                                          int64_t N, int64_t NT, double wall_pos, double adhesion,
                                          double force_const);

/* ---- fused stepping: Simulation.timeStep (Simulation.py:232-237) for steps
 *      [step0, step0 + nsteps): force pass, then one node pass that applies the substrate and
 *      indenter terms at the device-computed indenter position (Simulation.py:117-132, 173-183)
 *      and the damped velocity-Verlet update, on the displacement state (plan.ref, plan.disp).
 *      tip_hist (device, nullable, nsteps doubles) receives the FP64 indenter force of each step.
 *      use_graph != 0 replays whole blocks of 100 steps from a captured hipGraph (captured once per
 *      plan and history mode, and reused whatever nsteps is); the remainder runs eagerly.
 *      Between calls the velocity-Verlet state is the reference's (vel, fold).  Inside a call of
 *      n >= 2 steps the first step turns it into w = a v + f_old/2 (kept in vel), the middle steps
 *      advance (disp, w) without a previous force (24 B per node less), and the last step turns w
 *      back into (vel, fold): the same recurrence, with rounding that depends on how a run is cut
 *      into calls (the same cuts give the same bits).  With the node-fused buffers in the plan
 *      (disp_b, vel_b, partials_b), steps 1 .. n-1 of the call are one launch each (the node update
 *      of the step before rides in the force pass) and a last node pass ends the call; the first
 *      call with a tip history allocates the context's per-tile indenter-force rows (ceil(T/tile_k)
 *      x 100 doubles), so a caller that captures this call makes one such call beforehand. ------ */
int nfx_run_steps(nfx_ctx* ctx, const nfx_plan* plan, int64_t step0, int64_t nsteps, double* tip_hist,
                  int32_t use_graph);
/* number of step blocks the context has captured into graphs so far (a replay of a cached block
 * does not count) */
int nfx_graph_captures(nfx_ctx* ctx, int64_t* count);
/* the node pass alone (for callers that enqueue the force pass themselves), on (vel, fold) */
int nfx_node_update(nfx_ctx* ctx, const nfx_plan* plan, int64_t step, double* tip_out);

/* ---- batched stepping (no reference counterpart; the reference simulates one cell per process):
 *      all cells of the batch advance together, one force pass + one node pass per step over the
 *      concatenated meshes, plus a per-cell curve-sampling pass; plan holds the concatenated mesh,
 *      state and tile arrays (scatter must be NFX_SCATTER_TILED, model NFX_MR or NFX_VISCOPLASTIC,
 *      the substrate/indenter flags select the terms; their values come from the cell table). ---- */
int nfx_batch_run_steps(nfx_ctx* ctx, const nfx_plan* plan, const nfx_batch* batch, int64_t step0, int64_t nsteps,
                        int32_t use_graph);
/* average device time of the three batched passes over nrep steps (ms), HIP events */
int nfx_batch_time_passes(nfx_ctx* ctx, const nfx_plan* plan, const nfx_batch* batch, int64_t step0, int32_t nrep,
                          double* ms_force, double* ms_node, double* ms_curve);

/* ---- kernel timing for the benchmark: average device time of the force pass over nrep steps and
 *      of the node pass over the middle steps of those nrep (the (disp, w) form; all of them when
 *      nrep < 3), measured with HIP events on the context's stream (ms).  For a plan with the
 *      node-fused buffers and nrep >= 3, ms_force is the average node-fused launch over the middle
 *      steps (node update of the step before + force pass) and ms_node the call's last node pass. */
int nfx_time_passes(nfx_ctx* ctx, const nfx_plan* plan, int64_t step0, int32_t nrep, double* ms_force,
                    double* ms_node);

/* ---- MLP surrogate (no reference counterpart): BF16-on-MFMA GEMMs with fused epilogues --------
 * C[M x N] = A[M x K] * B[N x K]^T, A and B row-major BF16 (K contiguous), FP32 accumulation.
 * M a multiple of 128, N and K of 32, lda/ldb multiples of 8. */
#define NFX_MLP_STORE 0    /* C = acc (+ bias if non-NULL) */
#define NFX_MLP_FWD_GELU 1 /* C (nullable) = Z = acc + bias; H = gelu(Z) BF16 [M x N]; HT (nullable) = H^T */
#define NFX_MLP_BWD_GELU 2 /* H = acc * gelu'(Z) BF16 [M x N]; HT = H^T */
#define NFX_MLP_ADAM 3     /* acc = weight gradient: Adam on FP32 master C, BF16 copies H [M x N], HT [N x M] */

typedef struct {
    int32_t epilogue;
    int32_t M, N, K;
    const void* A; /* BF16 */
    int32_t lda;
    const void* B; /* BF16 */
    int32_t ldb;
    const float* bias; /* N */
    float* C;          /* FP32 [M x N] */
    const float* Z;    /* FP32 [M x N], BWD_GELU */
    void* H;           /* BF16 [M x N] */
    void* HT;          /* BF16 [N x M] */
    float* adam_m;     /* ADAM moments, FP32 [M x N] */
    float* adam_v;
    float lr_t;        /* lr / (1 - beta1^t) */
    float bc2;         /* 1 / sqrt(1 - beta2^t): w -= lr_t m / (sqrt(v) bc2 + eps) */
    float beta1, beta2, eps;
    float grad_scale;  /* multiplies the gradient before the update */
    float* workspace;  /* FP32 [split_k x M x N], split_k > 1 only */
    int32_t split_k;   /* 0 or 1: one pass over K; S > 1 (S <= 64, K % (32 S) == 0): K in S slices
                          accumulated into the workspace, then summed in slice order and the
                          epilogue applied (for GEMMs with a small M x N and a long K) */
} nfx_mlp_gemm_args;

int nfx_mlp_gemm(nfx_ctx* ctx, const nfx_mlp_gemm_args* args);
/* row-weighted mean squared error sum_ij w_i (Y - T)^2 / (M N) over M x N FP32 (accumulated into
 * *loss_sum, FP64, device) and its gradient 2 w_i (Y - T) / (M N) as BF16 in [M x N] (dZ) and
 * [N x M] (dZT).  row_weight (M FP32) is nullable: w_i = 1.  M and N are multiples of 32. */
int nfx_mlp_mse(nfx_ctx* ctx, const float* Y, const float* target, const float* row_weight, int32_t M, int32_t N,
                void* dZ, void* dZT, double* loss_sum);
/* bias gradient = row sums of dZT [N x M] (BF16), then Adam on the FP32 bias b with moments m, v */
int nfx_mlp_bias_adam(nfx_ctx* ctx, const void* dZT, int32_t M, int32_t N, float* b, float* m, float* v, float lr_t,
                      float bc2, float beta1, float beta2, float eps, float grad_scale);

/* ---- measured HBM ceiling: average time of nrep device-to-device copies of `bytes` (a multiple of
 *      16) with 16-byte vector accesses, HIP events (ms).  No reference counterpart. ----------- */
int nfx_stream_copy(nfx_ctx* ctx, const void* src, void* dst, int64_t bytes, int32_t nrep, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* NFX_H */
