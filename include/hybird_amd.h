/*
 * hybird_amd: MI355X-native coupled lattice-Boltzmann / discrete-element step.
 *
 * C ABI of libhybird_amd.so. Plain pointers, sizes and POD structs only; every
 * function returns 0 on success or a negative HYB_E* code, with a message
 * available from hyb_last_error(). All work for one handle is serialised on
 * that handle's HIP stream; host synchronisation happens only in the *_get_*
 * calls, the timing queries and hyb_sync().
 *
 * The entry points replace the methods of the reference's solver classes
 * (reference file:line of the interface each one stands in for):
 *   hyb_lb_create        LB::LB + LB::latticeBoltzmannGet + LB::latticeDefinition  (LB.h:104-155, LB.cpp:42-188)
 *   hyb_dem_create       DEM::DEM + DEM::discreteElementGet                         (DEM.h:98-138, DEM.cpp:13-183)
 *   hyb_dem_set_objects  the object file part of DEM::discreteElementGet               (DEM.cpp:131-173)
 *   hyb_dem_init         DEM::discreteElementInit                                   (DEM.h:139, DEM.cpp:185-297)
 *   hyb_dem_init_lattice DEM::discreteElementInit with its own arguments
 *                        (boundary, lbSize, unit, lbF) instead of an LB handle   (DEM.h:139, DEM.cpp:185-297)
 *   hyb_lb_init          LB::latticeBolzmannInit                                    (LB.h:158, LB.cpp:190-219)
 *   hyb_dem_step         DEM::discreteElementStep                                   (DEM.h:137, DEM.cpp:331-376)
 *   hyb_lb_fs_step       LB::latticeBoltzmannFreeSurfaceStep                        (LB.h:161, LB.cpp:235-245)
 *   hyb_lb_couple        LB::latticeBoltzmannCouplingStep                           (LB.h:160, LB.cpp:247-280)
 *   hyb_lb_step          LB::latticeBolzmannStep                                    (LB.h:159, LB.cpp:221-233)
 *   hyb_cycle            goCycle without the IO call                                (hybird.cpp:35-66)
 *   hyb_lb_set_capture   keeps the node fields IO reads (IO.cpp:711-831) while stepping
 *   hyb_lb_set_force_field, hyb_lb_set_populations, hyb_adj_forward, hyb_adj_backward
 *                        discrete adjoint of the LB step (not in the reference)
 *   hyb_state_save / hyb_state_load   full-precision checkpoint and restart (not in the reference)
 *   hyb_lb_get / hyb_dem_get   host views of LB::nodes/types and DEM::elmts/particles read by IO
 *                        (IO.cpp:711-831, 835-1106)
 * The reference has no foreign-function interface; a C++ caller reaches these
 * through the facade classes in hybird_amd/host (same class and method names as
 * the reference), and Python reaches them through ctypes (hybird_amd/_abi.py).
 */
#ifndef HYBIRD_AMD_H
#define HYBIRD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYB_OK 0
#define HYB_E_ARG (-1)     /* invalid argument or configuration */
#define HYB_E_HIP (-2)     /* HIP runtime error */
#define HYB_E_STATE (-3)   /* call out of order (e.g. step before init) */
#define HYB_E_CAPACITY (-4) /* a device list overflowed its capacity */
#define HYB_E_UNSUPPORTED (-5)
#define HYB_E_NUMERIC (-6) /* non-finite value detected */
/* Device-side failures (a DEM list overflowing its capacity, a particle leaving
 * the cell grid, a non-finite density or element velocity) are detected by the
 * kernels while the stream runs ahead of the host. The kernel that finds one
 * records it in the handle's error word (pinned host memory) and every later
 * DEM kernel of that handle does nothing, so the state stays that of the
 * failing sub-step. The error is returned by the first call on the handle (or
 * on hyb_cycle/hyb_sync with it) made after the device has run that kernel; a
 * call that synchronises (hyb_sync, the views) always returns it. */

/* problemName values (vector.h:21), plus an extension for a generic liquid box */
enum hyb_problem {
    HYB_PROBLEM_NONE = 0,
    HYB_PROBLEM_SPLASH = 2,     /* gas above z = 20 (LB.cpp:628-640) */
    HYB_PROBLEM_DRUM = 4,       /* rotating drum: cylinder through (0.95,0,1.26)-(0.95,1,1.26), R = 1.243,
                                   angular velocity (0, drumSpeed, 0) (DEM.cpp:876-887); liquid inside the
                                   cylinder through (0.45,0,-0.96) with the radius of LB.cpp:734-741; the
                                   global mass fix towards fluidMass after every free-surface pass
                                   (LB.cpp:1806-1822); y walls only static (boundary 5 or 7) */
    HYB_PROBLEM_NET = 6,        /* static cylinder through (0,4,4)-(1,4,4), R = 4 (physical units), fluid
                                   inside with curved walls (DEM.cpp:916-927, LB.cpp:552-602, 1278-1319);
                                   without free surface */
    HYB_PROBLEM_LIQUID_BOX = 100 /* liquid inside liquidBox[], gas elsewhere (not in the reference) */
};

enum hyb_contact { HYB_CONTACT_LINEAR = 0, HYB_CONTACT_HERTZIAN = 1 }; /* elmt.h:141 */
enum hyb_collision { HYB_COLLISION_BGK = 0, HYB_COLLISION_TRT = 1 };
enum hyb_precision { HYB_FP64 = 0, HYB_FP32_SHIFTED = 1 };
/* Apparent-viscosity law of non-Newtonian nodes (node.cpp:125-140 is the Bingham law). FRICTIONAL (not in
 * the reference) replaces the Bingham yield stress by a Coulomb-like one proportional to the gauge pressure:
 * nu = plasticVisc + frictionCoefFluid * max(n - rho0, 0) / 3 / (2 n shearRate), clamped like the Bingham
 * viscosity; it needs no nonNewtonianSolve flag. */
enum hyb_rheology { HYB_RHEOLOGY_BINGHAM = 0, HYB_RHEOLOGY_FRICTIONAL = 1 };
/* Tangential contact law. VISCOUS is the reference's history-free law (DEM.cpp:2176-2222):
 * F_t = min(2 viscTang sqrt(m_eff k_s) |v_t|, mu F_n) along v_t. SPRING (not in the reference) keeps a
 * tangential spring xi per particle pair and per particle's wall contact: xi is projected on the current tangent
 * plane, advanced by v_t dt, F_t = k_s xi + 2 viscTang sqrt(m_eff k_s) v_t with the same k_s, limited to
 * mu F_n (slip resets xi to the limit); xi is carried across neighbour rebuilds and restarts. Objects and
 * cylinders keep the viscous law. */
enum hyb_tangential { HYB_TANGENTIAL_VISCOUS = 0, HYB_TANGENTIAL_SPRING = 1 };
/* Walls a particle can touch. FIRST is the reference's near-wall list (DEM.cpp:1496-1513): only the first
 * wall (in wall order) within nebrRange of a particle is listed, so a particle in the corner of a side wall
 * and the floor has no floor contact and sinks through it under gravity. ALL (not in the reference) lists
 * every wall within nebrRange and applies their contacts in wall order (with SPRING, one spring per particle
 * and wall). */
enum hyb_wall_contacts { HYB_WALL_CONTACTS_FIRST = 0, HYB_WALL_CONTACTS_ALL = 1 };

/* Configuration: the .cfg keys of the reference, in the reference's (physical)
 * units; the library converts to lattice units exactly as LB.cpp:85-188 and
 * DEM.cpp:13-183 do. Fields after `collision` are extensions whose defaults
 * reproduce the reference. */
typedef struct hyb_config {
    /* run control, hybird.cpp:136-190 */
    int32_t problem;
    int32_t demSolve, lbSolve, freeSurface, forceField, nonNewtonian, turbulence;
    double demInitialRepeat;
    /* lattice, LB.cpp:93-185 */
    double unitLength, unitTime, unitDensity;
    double lbSize[3];
    double plasticVisc, yieldStress, initVisc;
    double initVelocity[3];
    double lbF[3];
    int32_t boundary[6];
    double slipCoefficient, turbConst;
    /* DEM, DEM.cpp:15-61, 177-181 */
    double density;
    int32_t contactModel;
    double youngMod, poisson, linearStiff, restitution, viscTang, frictionCoefPart, frictionCoefWall;
    double numVisc;
    int32_t multiStep;
    double criticalRatio;
    double translate[3];
    double scale;
    /* extensions */
    int32_t collision;      /* HYB_COLLISION_BGK (reference) or HYB_COLLISION_TRT */
    double trtMagic;        /* TRT magic parameter Lambda (default 1/4) */
    int32_t precision;      /* HYB_FP64 or HYB_FP32_SHIFTED */
    double liquidBox[6];    /* HYB_PROBLEM_LIQUID_BOX: xmin,xmax,ymin,ymax,zmin,zmax (node coordinates, inclusive) */
    double avalanchePosit;  /* HYB_PROBLEM_NET: front of the liquid (physical units, LB.h avalanchePosit) */
    double drumSpeed;       /* HYB_PROBLEM_DRUM: drum angular velocity (rad/s, cfg key drumSpeed) */
    double fluidMass;       /* HYB_PROBLEM_DRUM: target liquid mass (physical units, cfg key fluidMass) */
    int32_t pairCapacity;   /* DEM neighbour-pair capacity (0: sized from the particle count) */
    int32_t rheology;       /* HYB_RHEOLOGY_BINGHAM (reference) or HYB_RHEOLOGY_FRICTIONAL */
    double frictionCoefFluid; /* HYB_RHEOLOGY_FRICTIONAL: yield stress per unit gauge pressure */
    int32_t tangentialModel;  /* HYB_TANGENTIAL_VISCOUS (reference) or HYB_TANGENTIAL_SPRING */
    int32_t wallContacts;     /* HYB_WALL_CONTACTS_FIRST (reference) or HYB_WALL_CONTACTS_ALL */
    int32_t reserved[2];
} hyb_config;

/* Fills *cfg with the reference's parse defaults (LB.cpp:93-185, DEM.cpp:15-181)
 * overridden by the lattice-unit conventions (unit* = 1). */
void hyb_config_defaults(hyb_config* cfg);

/* One particle-file record, DEM.cpp:85-120:
 * index size radius x y z vx vy vz wx wy wz q0 q1 q2 q3 dq0 dq1 dq2 dq3 */
#define HYB_ELEMENT_RECORD 20
/* One object (static sphere) of the object file (DEM.cpp:136-173): index, size,
 * r, x, y, z, vx, vy, vz (the line's remaining rotation values are unused). */
#define HYB_OBJECT_RECORD 9

typedef struct hyb_lb hyb_lb;
typedef struct hyb_dem hyb_dem;

const char* hyb_last_error(void);
int hyb_device_count(int* count);
int hyb_sync(hyb_lb* lb);

/* ---- DEM ---- */
int hyb_dem_create(const hyb_config* cfg, int32_t n_elements, const double* records, hyb_dem** out);
/* The object file's static spheres (DEM::discreteElementGet, DEM.cpp:131-173),
 * n records of HYB_OBJECT_RECORD, before hyb_dem_init: lattice nodes inside them
 * become static walls (LB.cpp:535-550) and particles touching them get the
 * object contact (DEM.cpp:1747-1770, 2060-2136). Indices must be 0..n-1. */
int hyb_dem_set_objects(hyb_dem* dem, int32_t n_objects, const double* records);
int hyb_dem_init(hyb_dem* dem, const hyb_lb* lb);
/* boundary: the six box boundary codes; lbSize: node counts; unit*: base units
 * (after the automatic unitTime of LB.cpp:124-133); lbF: body force in lattice units. */
int hyb_dem_init_lattice(hyb_dem* dem, const int32_t boundary[6], const int32_t lbSize[3], double unitLength,
                         double unitTime, double unitDensity, const double lbF[3]);
int hyb_dem_step(hyb_dem* dem);
int hyb_dem_destroy(hyb_dem* dem);

/* ---- LB ---- */
int hyb_lb_create(const hyb_config* cfg, hyb_lb** out);
int hyb_lb_init(hyb_lb* lb, hyb_dem* dem);
int hyb_lb_fs_step(hyb_lb* lb);
int hyb_lb_couple(hyb_lb* lb, hyb_dem* dem, int32_t* newNeighborList);
int hyb_lb_step(hyb_lb* lb, hyb_dem* dem);
int hyb_lb_destroy(hyb_lb* lb);

/* goCycle (hybird.cpp:35-66) repeated n times, IO excluded. */
int hyb_cycle(hyb_lb* lb, hyb_dem* dem, int32_t n);

/* ---- host views ----
 * Field names and layouts:
 *   LB  "f"          f64 [N][19]  populations in the reference's step-end (post-stream) state
 *       "n"          f64 [N]      density from the last reconstruction
 *       "u"          f64 [N][3]   velocity kept by the node (Guo-shifted, node.cpp:85-88)
 *       "hydroForce" f64 [N][3]
 *       "mass"       f64 [N]      free-surface mass function
 *       "fluidMass"  f64 [1]      sum of "mass" over active nodes outside particles (IO.cpp:987-999)
 *       "totalMass"  f64 [1]      the same sum at initialisation (LB.cpp:203-216)
 *       "visc"       f64 [N]      kinematic viscosity (lattice units)
 *       "types"      u8  [N]      node type code (node.h:73-83)
 *       "pflag"      u8  [N]      inside-particle flag
 *       "solidIndex" u32 [N]      particle index for pflag nodes
 *   DEM "e_x" f64[E][6][3], "e_q" f64[E][6][4], "e_w" f64[E][6][3], "e_forces" f64[E][7][3]
 *       (FHydro, MHydro, FParticle, MParticle, FWall, MWall, FGrav), "e_misc" f64[E][19],
 *       "p_data" f64[P][10] (x0, x1, radiusVec, r), "p_idx" u32[P][4],
 *       "neighborTable" u32[2*pairs], "nearWallTable" u32[2*k], "ghostIndex" u32[G], "ghostVec" f64[G][3],
 *       "nearObjectTable" u32[2*k] (particle, object), "objects" f64[O][8] (index, r, x0, x1),
 *       "dem_scalars" f64[17], "energy" f64[6] (DEM::particleEnergy after the last sub-step: kin, rotKin,
 *       trKin, grav, elastic, total; DEM.cpp:2226-2276), "walls" f64[W][12] (p, n, FParticle, FHydro = 0:
 *       walls[] after the last sub-step, DEM.cpp:1936-1979), "objectForces" f64[O][3] (objects[].FParticle,
 *       DEM.cpp:2100-2134; rows as "objects")
 *   LB  "wallFHydro" f64[W][3]: walls[].FHydro of the last LB step (moving no-slip walls, LB.cpp:1321-1341,
 *       summed exactly); "status" f64[3]: max |u|^2 over active nodes, active nodes, active nodes with
 *       viscosity > 0.95 (maxTau - 0.5) / 3 (IO.cpp:835-850, 969-985; free surface: needs capture)
 * *bytes_needed receives the size; nothing is copied when capacity is too small.
 * N = lbSize product; E elements; P particles including ghosts. */
int hyb_lb_get(hyb_lb* lb, const char* name, void* host, size_t capacity, size_t* bytes_needed);
int hyb_dem_get(hyb_dem* dem, const char* name, void* host, size_t capacity, size_t* bytes_needed);

/* MercuryDPM-format saving, the data of DEM::exportDataFile, exportRestartFile
 * and the .fstat statistics (DEM.cpp:355-360, 1131-1146, 2280-2565), which the
 * reference writes inside a sub-step when `demTimeStep % saveCount == 0`.
 * With save_count > 0, each such sub-step of hyb_dem_step (or hyb_cycle)
 * captures, after the predictor, the particle state and the contact statistics
 * of its force evaluation; a call keeps only its own captures (call hyb_cycle
 * one cycle at a time to see every one). 0 turns saving off. */
int hyb_dem_set_save_count(hyb_dem* dem, int32_t save_count);
int hyb_dem_saved(hyb_dem* dem, int32_t* count);
/* Part of capture `event` (0 <= event < count), copied like hyb_dem_get:
 *   "info"     f64[10]: demTime, demTimeStep, pairs, particles, objects in the
 *              object file, dampCoeff, deltat, demSize[3]
 *   "elements" f64[E][9]: predicted x0, predicted x1, wGlobal
 *   "pairs"    f64[pairs][12] (neighbour-table order), "walls" f64[E][12] (the
 *              element's near wall), "objects" f64[E][16][12] (its near objects
 *              in list order; empty without objects): a, b, overlap, |Fn|, |Ft|,
 *              en[3], et[3], flag, where a, b are the pair's particle indices
 *              (ghosts >= E), or the wall index (-1: none) / object row of the
 *              "objects" view (periodic copies included) and the element, and
 *              flag is 1 for a contact with tangential motion. */
int hyb_dem_saved_get(hyb_dem* dem, int32_t event, const char* part, void* host, size_t capacity,
                      size_t* bytes_needed);

/* Profiling of the fused stream-collide kernel: when enabled, every LB step
 * records HIP events around that kernel on the handle's stream. */
int hyb_lb_set_profiling(hyb_lb* lb, int32_t enabled);
int hyb_lb_kernel_time(hyb_lb* lb, double* total_ms, int64_t* launches);
/* Keep the density, velocity and node force of every collision for the host
 * views "n", "u", "hydroForce" (what IO reads from the nodes, IO.cpp:711-831).
 * Without free surface they can also be recomputed on demand; with free
 * surface they must be captured while stepping. Costs 56 bytes per node and
 * step of extra HBM writes while enabled. */
int hyb_lb_set_capture(hyb_lb* lb, int32_t enabled);
/* HBM reference point (not in the reference): bytes read + written per second
 * (GB/s) of a plain non-temporal device copy of `bytes` with `width` (4, 8 or
 * 16) bytes per lane, averaged over `reps` launches timed with HIP events.
 * Allocates and frees its two buffers; synchronises. */
int hyb_copy_bandwidth(size_t bytes, int32_t width, int32_t reps, double* gbps);
/* Active (fluid + interface) node count, the MLUPS denominator. */
int hyb_lb_active_nodes(hyb_lb* lb, int64_t* count);

/* walls[].FHydro of the last LB step (the "wallFHydro" view: moving no-slip box
 * walls, LB.cpp:1321-1341, exact sums) written to DEVICE memory d_out[3 * walls]
 * (physical units) on the handle's stream, without synchronising and without
 * allocating. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * The sums use one fixed device array per GPU, sized for at most
 * HYB_WALL_FHYDRO_MAX_WALLS walls (more: HYB_E_CAPACITY); calls that overlap on
 * one GPU race on it. n_out: capacity of d_out in doubles (at least 3 * walls). */
#define HYB_WALL_FHYDRO_MAX_WALLS 6
int hyb_lb_wall_fhydro_into(hyb_lb* lb, double* d_out, int32_t n_out);

/* ---- extensions for the adjoint (no counterpart in the reference) ----
 * Per-node body force F[N][3] in lattice units, added to lbF in every
 * collision (NULL removes it). */
int hyb_lb_set_force_field(hyb_lb* lb, const float* F);
/* Replaces the initial populations f[N][19] (post-stream convention of the
 * "f" view); only before the first step. */
int hyb_lb_set_populations(hyb_lb* lb, const double* f);
/* Runs `steps` LB steps from the initial state (particles frozen: no DEM or
 * coupling steps) and keeps the state every `every` steps for the reverse pass
 * (FP64, no free surface, constant viscosity). It also allocates the buffers
 * of the reverse pass (the states of one checkpoint interval, two adjoint
 * population buffers, the gradient accumulators), kept until the next
 * hyb_adj_forward or hyb_lb_destroy. */
int hyb_adj_forward(hyb_lb* lb, hyb_dem* dem, int32_t steps, int32_t every);
/* Gradients of a loss L(f_T) of the final post-stream populations, given
 * seed = dL/df_T [N][19]: grad_force [N][3] (per-node body force), *grad_tau
 * (relaxation time tau = 1/2 + 3 nu), grad_f0 [N][19] (initial populations).
 * Any output may be NULL. Recomputes the states between checkpoints. */
int hyb_adj_backward(hyb_lb* lb, hyb_dem* dem, const double* seed, double* grad_force, double* grad_tau,
                     double* grad_f0);
/* Device time (HIP events on the handle's stream) of the step loop of the last
 * hyb_adj_forward (LB steps and checkpoint copies) and of the reverse loop of the
 * last hyb_adj_backward (segment replays and adjoint steps), in milliseconds;
 * neither includes allocation or host transfers. */
int hyb_adj_device_time(hyb_lb* lb, double* forward_ms, double* backward_ms);

/* ---- checkpoint and restart (no counterpart in the reference, which cannot
 * resume a run) ----
 * hyb_state_save writes the complete device state of initialised handles
 * (either may be NULL) to a binary file: populations, node types and link
 * words, free-surface masses, densities, lists and counters, the element
 * state, neighbour tables and every step counter. hyb_state_load restores it
 * into handles created and initialised from the same configuration and element
 * records (HYB_E_ARG otherwise); stepping on continues bit-identically to the
 * run that was saved. Both synchronise the handles' stream. A failed load
 * may leave the handles partly overwritten. */
int hyb_state_save(hyb_lb* lb, hyb_dem* dem, const char* path);
int hyb_state_load(hyb_lb* lb, hyb_dem* dem, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* HYBIRD_AMD_H */
