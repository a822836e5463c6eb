/*
 * agrav.h -- C ABI of the MI355X-native TreePM gravity step.
 *
 * A drop-in for AREPO's gravity hot path: tree build, tree walk, periodic PM
 * and the FORCETEST direct-summation check.  Every entry point names the
 * reference function it replaces (paths relative to AREPO's src/).
 *
 * Conventions follow the reference: tree accelerations and potentials are in
 * G=1 units until agrav_gravity_force_finalize (accel.c:258-265); PM results
 * include G (pm_periodic.c:1590); OldAcc is |a_tree + a_PM/G| in G=1 units
 * (accel.c:237-255).  Particle arrays are in the caller's order; positions
 * are interleaved x,y,z per particle like P[i].Pos (allvars.h:1216-1274).
 *
 * A context is not thread-safe; distinct contexts may be used from distinct
 * host threads.  All device work of a context runs on its own HIP stream.
 */
#ifndef AGRAV_H
#define AGRAV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Errors replace terminate_program()/MPI_Abort (allvars.h:173) and the
 * negative returns of force_treebuild (forcetree.c:186-228). */
typedef enum
{
  AGRAV_OK            = 0,
  AGRAV_ERR_BAD_PARAM = -1, /* invalid parameter or argument */
  AGRAV_ERR_CAPACITY  = -2, /* a capacity limit was exceeded (forcetree.c:930-933) */
  AGRAV_ERR_OUT_OF_BOX = -3, /* particle outside the periodic box (forcetree.c:432-438) */
  AGRAV_ERR_HIP       = -4, /* HIP runtime error */
  AGRAV_ERR_OOM       = -5, /* device (or page-locked host) allocation failed */
  AGRAV_ERR_STATE     = -6, /* call order violated (e.g. walk before build) */
  AGRAV_ERR_FFT       = -7, /* rocFFT error */
  AGRAV_ERR_NO_DEVICE = -8  /* no usable GPU */
} agrav_status;

typedef enum
{
  AGRAV_FP64 = 0, /* DOUBLEPRECISION=1 semantics throughout */
  AGRAV_FP32 = 1  /* FP32 interactions on node-relative offsets, compensated sums;
                     FP64 kept for positions, COM accumulation, Green's function */
} agrav_precision;

typedef enum
{
  AGRAV_WALK_PLAIN = 0, /* one lane per target, AREPO's per-particle opening test */
  AGRAV_WALK_GROUP = 1  /* one wavefront per Hilbert-ordered target group, node tiles in LDS */
} agrav_walk_kind;

#define AGRAV_MAX_SOFT_TYPES 8

/* Runtime replacement of the gravity Config.sh flags and param.txt entries
 * (Template-Config.sh:66-136, parameters.c:214-363). */
typedef struct
{
  int periodic;                   /* 1 unless GRAVITY_NOT_PERIODIC */
  double box_size;                /* BoxSize; isolated runs normally pass 0 */
  int pmgrid;                     /* PMGRID, 0 = tree only (power of two, periodic only) */
  double asmth;                   /* ASMTH in mesh cells (allvars.h:461), default 1.25 */
  double rcut;                    /* RCUT in units of Asmth (allvars.h:469), default 4.5 */
  double G;                       /* GravityConstantInternal */
  double err_tol_theta;           /* ErrTolTheta (BH, and the first-step seeding pass) */
  double err_tol_force_acc;       /* ErrTolForceAcc */
  int type_of_opening_criterion;  /* 0 = Barnes-Hut, 1 = relative */
  int n_soft_types;               /* number of softening types in use (>=1) */
  double softening[AGRAV_MAX_SOFT_TYPES]; /* Plummer-equivalent softening per type (SofteningComovingType*) */
  int precision;                  /* agrav_precision */
  int walk;                       /* agrav_walk_kind */
  int group_size;                 /* targets per group for AGRAV_WALK_GROUP (1..64; 0 = by workload, the default:
                                     64, or 16 for isolated sets below 2^19 particles) */
  double tree_alloc_factor;       /* TreeAllocFactor (init.c:110); unused: node arrays are sized exactly */
  int leaf_bucket;                /* group walk: an opened node with at most this many particles is expanded
                                     straight into particles (0..64, default 8) */
  int eval_potential;             /* EVALPOTENTIAL (Template-Config.sh): 1 = potentials too (default), 0 = forces
                                     only; the FP32 group walk then skips the potential arithmetic */
  int group_waves;                /* FP32 group walk of one set: wavefronts sharing each group's traversal (1, 2
                                     or 4; 0 = by workload: isolated sets 4 below 2^17 particles, periodic
                                     4 below 2^18 and 2 below 2^20, else 1).  More waves cut the latency of
                                     small sets; results are deterministic for a given value but differ in
                                     the last bits between values */
  int placehighresregion;         /* PLACEHIGHRESREGION (Template-Config.sh): bit t set for the particle types t that
                                     define a second, high-resolution PM grid of 2 x PMGRID cells (pm_nonperiodic.c);
                                     0 = none.  Periodic TreePM only; the types come from agrav_set_particle_types */
  double enlargeregion;           /* ENLARGEREGION: the high-resolution region's extent factor (<= 0: not set) */
} agrav_params;

/* Per-stage device times (ms) of the last call, named like AREPO's cpu.txt
 * timers (utils/timer.h:174-187). */
typedef struct
{
  double treebuild;   /* keys, sort, octree, moments */
  double treewalk;    /* the walk kernel(s) of the last agrav_gravity_tree */
  double pm_grav;     /* long-range PM force */
  double finalize;    /* gravity_force_finalize */
  double walk_kernel; /* average duration of one walk kernel launch */
  long long interactions; /* sum of per-target interaction counts of the last walk */
  int numnodes;       /* octree nodes of the last build */
  int walk_retries;   /* groups re-run by the fallback walk (stack overflow) */
  long long walk_popped;  /* group walk: stack entries evaluated (traversal work) */
  long long tile_entries; /* group walk: interaction-tile entries streamed by all 64 lanes */
  int unbind_rounds;      /* agrav_subfind_unbind: passes of the batched loop (the longest group's) */
  int unbind_solves;      /* agrav_subfind_unbind: batched potential solves (tree build + walk) */
  double unbind_solve_ms; /* agrav_subfind_unbind: device time of those solves */
  double unbind_pass_ms;  /* agrav_subfind_unbind: device time of the per-group pass kernels */
  long long device_bytes; /* HBM held by the context's device buffers now (they are kept between calls) */
  int walk_ordered;       /* group walk: 1 when its groups were launched costliest first, by the tile entries each
                             group streamed in the last walk of the same grouping (every target of as many particles,
                             same group size; like GravCost); the order changes no result */
} agrav_timings;

typedef struct agrav_ctx agrav_ctx;

const char *agrav_status_string(agrav_status s);
/* Human-readable detail of the last error on this context. */
const char *agrav_last_error(const agrav_ctx *ctx);

/* Fill p with the defaults of the reference (allvars.h:461-469, init.c:110). */
void agrav_default_params(agrav_params *p);

/* begrun1(): set_softenings (grav_softening.c:55-96), force_short_range_init
 * (forcetree_walk.c:72-85), ewald_init (forcetree_ewald.c:89-252) and
 * pm_init_periodic (pm_periodic.c:173-255).  Selects the current HIP device. */
agrav_status agrav_create(const agrav_params *p, agrav_ctx **out);
void agrav_destroy(agrav_ctx *ctx);

/* Change the softening table between calls: set_softenings (grav_softening.c:55-96). */
agrav_status agrav_set_softenings(agrav_ctx *ctx, int n_soft_types, const double *softening);

/* Particle load: replaces the reads of P[i].Pos/Mass/SofteningType
 * (forcetree.c:394-417).  Host pointers; data are copied to HBM.
 * softtype may be NULL (all type 0). */
agrav_status agrav_set_particles(agrav_ctx *ctx, int64_t n, const double *pos, const double *mass, const uint8_t *softtype);
/* Page-locked host memory for the arrays a caller stages particle data and results in
 * (no counterpart in the reference, whose particle data never leave the host): copies
 * from and to it run at the host link's DMA rate instead of through the runtime's
 * pageable staging.  Any host pointer argument of this interface may point into it.
 * agrav_host_free takes only pointers from agrav_host_alloc (NULL is ignored). */
agrav_status agrav_host_alloc(size_t bytes, void **ptr);
agrav_status agrav_host_free(void *ptr);
/* Same with device pointers (already resident in HBM); copied device-to-device on the ctx stream. */
agrav_status agrav_set_particles_device(agrav_ctx *ctx, int64_t n, const double *d_pos, const double *d_mass,
                                        const uint8_t *d_softtype);
/* P[i].Type (allvars.h:1216-1274), host pointer, caller order: with placehighresregion, the types whose
 * particles size the high-resolution region (pm_init_regionsize, pm_nonperiodic.c:135-245) and count as
 * high-resolution particles wherever they are (pmforce_is_particle_high_res, :344-371).  Needed after
 * every set_particles of a context with placehighresregion. */
agrav_status agrav_set_particle_types(agrav_ctx *ctx, const uint8_t *type);
/* The high-resolution region of the last PM step (All.Xmintot[1], All.Xmaxtot[1], All.Corner[1],
 * All.UpperCorner[1], All.TotalMeshSize[1], All.Asmth[1], All.Rcut[1]): region[0..2] xmin, [3..5] xmax,
 * [6..8] corner, [9..11] upper corner, [12] mesh size, [13] Asmth[1], [14] Rcut[1] (region may be NULL).
 * AGRAV_ERR_STATE while it is not sized.  reset = 1 drops it: the next PM sizes it from the particles,
 * as pm_init_regionsize does at start-up (long_range_init_regionsize, longrange.c:72-86); without a reset
 * it is re-sized only when a high-resolution particle has left it (longrange.c:129-140). */
agrav_status agrav_highres_region(agrav_ctx *ctx, int reset, double *region);
/* P[i].OldAcc (accel.c:237-255), host pointer, caller order. */
agrav_status agrav_set_old_acc(agrav_ctx *ctx, const float *old_acc);

/* The active particles, gravity_tree's TargetList (gravtree.c:494-520, built from
 * TimeBinsGravity.ActiveParticleList, allvars.h:664): caller indices, each at most
 * once.  The next agrav_gravity_tree, agrav_gravity_force_finalize and
 * agrav_compute_grav_accelerations evaluate and write only these; the tree is built
 * from every particle and the long-range PM force covers every particle, as in the
 * reference.  Outputs of inactive particles keep their last values (zero after
 * set_particles).  active = NULL or n_active < 0 makes every particle active again
 * (the default); set_particles resets the list.  Not for batches. */
agrav_status agrav_set_active(agrav_ctx *ctx, const int32_t *active, int64_t n_active);

/* HIERARCHICAL_GRAVITY's partial tree (forcetree.c:91-95, :166-171; the driver
 * do_gravity_hydro.c:126-266 calls compute_grav_accelerations(timebin,
 * FLAG_PARTIAL_TREE) per timebin).  partial = 1: while an active list is set, the
 * next builds insert only the active particles, in the domain cube of every
 * particle, so the walk gives the forces among the active particles; the
 * long-range PM force still covers every particle, and compute_grav_accelerations
 * reuses the last one while the particles are unchanged (GravPM between PM steps,
 * do_gravity_hydro.c:105-123).  A list changed after such a
 * build makes the walk fail (AGRAV_ERR_STATE) until the next build.
 * set_oldacc = 0: gravity_force_finalize leaves OldAcc as it is, as on every
 * timebin below the highest occupied one (accel.c:231-232).  Defaults: partial
 * = 0, set_oldacc = 1.  The setting stays across set_particles. */
agrav_status agrav_set_partial_tree(agrav_ctx *ctx, int partial, int set_oldacc);

/* construct_forcetree (forcetree.c:110-155): domain cube, Peano-Hilbert keys,
 * sort, octree, bottom-up moments.  numnodes may be NULL. */
agrav_status agrav_construct_forcetree(agrav_ctx *ctx, int *numnodes);

/* gravity_tree (gravtree.c:357-707) over all particles with force_treeevaluate
 * (forcetree_walk.c:115-514).  use_bh=1: Barnes-Hut with ErrTolTheta;
 * use_bh=0: the configured criterion (relative needs OldAcc). */
agrav_status agrav_gravity_tree(agrav_ctx *ctx, int use_bh);

/* long_range_force / pmforce_periodic (longrange.c:98-198, pm_periodic.c:1572-1936). */
agrav_status agrav_long_range_force(agrav_ctx *ctx);

/* gravity_force_finalize (accel.c:204-347): OldAcc, x G, PM potential constant, + PM. */
agrav_status agrav_gravity_force_finalize(agrav_ctx *ctx);

/* compute_grav_accelerations (accel.c:58-80) + gravity (accel.c:94-192):
 * first_step=1 with the relative criterion runs PM and a Barnes-Hut pass to
 * seed OldAcc first (accel.c:62-72). */
agrav_status agrav_compute_grav_accelerations(agrav_ctx *ctx, int first_step);

/* Results in the caller's order (host pointers, any may be NULL):
 * acc[3n] = GravAccel (+GravPM), pot[n] = Potential, nint[n], old_acc[n]. */
agrav_status agrav_get_accelerations(agrav_ctx *ctx, double *acc, double *pot, int32_t *nint, float *old_acc);
/* Device pointers to the same results (valid until the next call on ctx). */
agrav_status agrav_get_accelerations_device(agrav_ctx *ctx, const double **acc, const double **pot);
/* Timestep criteria from the accelerations of the last step, before AREPO's integer
 * timeline: find_long_range_step_constraint (timestep.c:525-565; PMGRID only, every
 * particle counted as collisionless) gives dt_displacement = 2 x the minimum over
 * particles of sqrt(2 ErrTolIntAccuracy a ForceSoftening / 2.8 / |a_PM,phys|) H(a),
 * clamped to [MinSizeTimestep, MaxSizeTimestep]; get_timestep_gravity
 * (timestep.c:284-344) gives dt[i] (caller order, may be NULL) from the total
 * acceleration (GravAccel + GravPM), capped at MaxSizeTimestep and, with PMGRID, at
 * dt_displacement.  Comoving: a2inv = 1/a^2 on accelerations, a and H(a) as in
 * set_cosmo_factors_for_current_time (timestep.c:77-90).  A massless particle gets
 * MaxSizeTimestep.  A dt below MinSizeTimestep is raised to it and counted in
 * n_below_min (the reference stops there unless NOSTOP_WHEN_BELOW_MINTIMESTEP). */
typedef struct
{
  double err_tol_int_accuracy; /* ErrTolIntAccuracy */
  double max_size_timestep;    /* MaxSizeTimestep */
  double min_size_timestep;    /* MinSizeTimestep */
  double time_a;               /* All.Time, comoving only */
  double hubble_a;             /* hubble_function(All.Time), comoving only */
  int comoving;                /* ComovingIntegrationOn */
} agrav_timestep_params;

agrav_status agrav_gravity_timesteps(agrav_ctx *ctx, const agrav_timestep_params *tp, double *dt, double *dt_displacement,
                                     int64_t *n_below_min);

/* Raw parts: tree force (G=1 units) and PM force, caller order, host pointers (may be NULL).
 * After a partial tree the tree parts are written for the tree's particles only. */
agrav_status agrav_get_components(agrav_ctx *ctx, double *acc_tree, double *pot_tree, double *acc_pm, double *pot_pm);

/* gravity_forcetest / gravity_forcetest_evaluate (gravtree_forcetest.c:257-648):
 * FP64 direct summation for ns selected particles (caller indices) over all
 * particles, with the 129^3 Ewald table (gravtree_forcetest.c:933-1085) when
 * periodic.  Outputs in G=1 units (any may be NULL): true force and
 * potential, and the exact-erfc short-range parts (PMGRID only). */
agrav_status agrav_gravity_forcetest(agrav_ctx *ctx, int ns, const int32_t *sel, double *acc, double *acc_short, double *pot,
                                     double *pot_short);

/* gravity_direct (gravdirect.c:61-257) with force_evaluate_direct
 * (forcetree_walk.c:557-708): all-pairs tree-equivalent force, G=1 units,
 * caller order (host pointers). */
agrav_status agrav_gravity_direct(agrav_ctx *ctx, double *acc, double *pot);

/* agrav_gravity_direct for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The same kernels and summation order as
 * agrav_gravity_direct (the results are bitwise equal), but nothing is allocated:
 * the partial sums live in a fixed 4 MiB device array of the library, set up by
 * agrav_create on each device.  Inputs and outputs are caller-owned device
 * pointers in caller order: d_pos[3n], d_mass[n], d_softtype[n] (or NULL: every
 * particle type 0); d_acc[3n] and d_pot[n] (d_pot may be NULL) in G=1 units.  The
 * context supplies the parameters and tables (its particles are not used).
 * Asynchronous on `stream` (a hipStream_t; NULL = the context's stream): no
 * allocation, no host synchronisation, no memset.  n <= AGRAV_DIRECT_NOALLOC_MAX_N,
 * else AGRAV_ERR_CAPACITY.  Softening types are not checked on the host: a type
 * >= n_soft_types makes the results NaN.  All calls on one GPU share its scratch
 * array, so calls that overlap in time on one GPU race; serialise them (one
 * stream, or events). */
#define AGRAV_DIRECT_NOALLOC_MAX_N 16384
agrav_status agrav_gravity_direct_noalloc(agrav_ctx *ctx, int64_t n, const double *d_pos, const double *d_mass,
                                          const uint8_t *d_softtype, double *d_acc, double *d_pot, void *stream);

agrav_status agrav_get_timings(agrav_ctx *ctx, agrav_timings *t);

/* Tree inspection for tests: node count and per-node data (host pointers of
 * numnodes entries; any may be NULL).  Node 0 is the root; order is pre-order. */
agrav_status agrav_get_tree(agrav_ctx *ctx, int *numnodes, double *center, double *len, double *com, double *mass, int32_t *level,
                            int32_t *npart);
/* Particle permutation: sorted position k holds caller index perm[k]; keys[k] its 63-bit Hilbert key.
 * One entry per particle of the tree (the active ones after a partial tree). */
agrav_status agrav_get_order(agrav_ctx *ctx, int32_t *perm, uint64_t *keys);

/* The tables the kernels use: short-range (forcetree_walk.c:72-85), 2 x 128 floats;
 * Ewald tree table (forcetree_ewald.c:89-252), 4 x 65^3 doubles (periodic tree-only);
 * FORCETEST table (gravtree_forcetest.c:933-1008), 4 x 129^3 doubles (periodic). */
agrav_status agrav_get_tables(agrav_ctx *ctx, float *tab_force, float *tab_pot, double *ewald_tree, double *ewald_forcetest);

/* ---- batch: many independent isolated sets in one build and one walk -------
 * Replaces the per-group loop of subfind_unbind (subfind_serial.c:429-560):
 * subfind_loctree_treebuild (subfind_loctree.c:192) and
 * subfind_loctree_treeevaluate_potential (subfind_loctree.c:564-730) for every
 * member of every group.  Set s holds the particles [offsets[s], offsets[s+1])
 * (nsets+1 offsets from 0); positions are FP32 offsets from the set's centre
 * (centres: 3 doubles per set, NULL = origin); softening[s] is the set's
 * Plummer-equivalent softening.  Sets do not interact.  The parameters must be
 * isolated (periodic = 0, pmgrid = 0).  After this call
 * agrav_compute_grav_accelerations solves the whole batch (Barnes-Hut with
 * err_tol_theta when type_of_opening_criterion = 0, as the unbinding walk at
 * subfind_loctree.c:639, else the relative criterion seeded by a Barnes-Hut
 * pass); results come back through agrav_get_accelerations(_device) in the
 * caller's order.  The potential includes each particle's own softened term, as
 * in the reference walk. */
#define AGRAV_BATCH_CLASSES 4
typedef struct
{
  int group_size[AGRAV_BATCH_CLASSES];         /* targets per wavefront group, 1..64 */
  int leaf_bucket[AGRAV_BATCH_CLASSES];        /* opened nodes of at most this many particles go to the tile as particles */
  int64_t class_max_n[AGRAV_BATCH_CLASSES];    /* class k: sets with class_max_n[k-1] < n <= class_max_n[k]; -1 = no limit */
  int64_t sample_particles[AGRAV_BATCH_CLASSES]; /* autotune: particles in the timed sample of the class */
  double best_ms[AGRAV_BATCH_CLASSES];         /* autotune: walk time of the sample with the chosen setting */
  int tuned;
} agrav_batch_tuning;

agrav_status agrav_batch_set_particles(agrav_ctx *ctx, int nsets, const int64_t *offsets, const float *pos_rel,
                                       const double *centres, const double *mass, const double *softening);
/* Times every candidate group size (16, 32, 64) and leaf bucket (0, 4, 8, 16) on a
 * deterministic sample of about sample_particles particles per set-size class
 * (<= 0: 4M) and keeps the fastest per class for the following solves.  Builds
 * the tree first if needed.  out may be NULL. */
agrav_status agrav_batch_autotune(agrav_ctx *ctx, int64_t sample_particles, agrav_batch_tuning *out);
agrav_status agrav_batch_get_tuning(agrav_ctx *ctx, agrav_batch_tuning *out);
agrav_status agrav_batch_set_tuning(agrav_ctx *ctx, const agrav_batch_tuning *t);

/* ---- SUBFIND unbinding of many candidate groups ---------------------------
 * Replaces subfind_unbind (subfind_serial.c:429-637) called once per candidate
 * (subfind_serial.c:291, subfind_collective.c:1442).  Group g holds the
 * particles [offsets[g], offsets[g+1]).  Every pass of the reference's loop is
 * restated: the potential of all members (G/a x the Barnes-Hut walk of a local
 * tree, err_tol_theta of the context = ErrTolThetaSubfind, subfind_loctree.c:639)
 * in a full pass; the centre of mass about the member of minimum potential and
 * the bulk velocity; binding energies with the Hubble flow, the self-energy
 * G/a m/eps and, for gas (type 0), utherm; removal of the members with E > 0
 * above the 0.25 len-th largest energy; the phase switch between full passes and
 * energy-only passes; the stop when a full pass removes nobody or fewer than
 * des_link_ngb members remain.  In an energy-only pass the reference evaluates
 * the weakly bound members' potentials again but keeps the old values scaled by
 * G/a once more (subfind_serial.c:491-497); that scaling is kept, the discarded
 * walk is not run.  All groups advance together: the potentials of every group
 * in a full pass come from one batched tree build and walk (the context's
 * precision and walk kind), the rest of a pass runs one workgroup per group.
 *
 * Positions are FP64 in an isolated frame: a group in a periodic box is passed
 * relative to a point inside it, wrapped to the nearest image (the reference's
 * GRAVITY_NEAREST distances for groups smaller than half the box).
 * softening[g] is the group's Plummer-equivalent softening (ForceSoftening/2.8
 * of its members).  cubes (4 doubles per group {centre[3], len}, or NULL) is the
 * local tree's root cube kept for every pass (subfind_loctree_findExtent,
 * subfind_loctree.c:101-160); NULL computes it from each group's members.
 * utherm and type may be NULL (no gas).  The context must be isolated
 * (periodic = 0, pmgrid = 0).  Outputs: bound[n] (1 = member left), len[g]
 * (members left; below des_link_ngb the group is dropped by the caller, as the
 * reference does), and optionally len_non_gas[g], iterations[g] (passes counted
 * as the reference's iter), potential[n] and bnd_energy[n] (PS.Potential and
 * PS.BindingEnergy as last computed).  Members of equal minimum potential: the
 * first in caller order is the reference point. */
typedef struct
{
  double time_a;    /* All.Time: the scale factor when comoving */
  double hubble_a;  /* hubble_function(All.Time), comoving only */
  int comoving;     /* All.ComovingIntegrationOn */
  int des_link_ngb; /* All.DesLinkNgb */
  int max_iter;     /* MAXITER (allvars.h:482); <= 0: 300000 */
} agrav_unbind_params;

agrav_status agrav_subfind_unbind(agrav_ctx *ctx, int ngroups, const int64_t *offsets, const double *pos, const double *vel,
                                  const double *mass, const double *softening, const double *utherm, const uint8_t *type,
                                  const double *cubes, const agrav_unbind_params *up, uint8_t *bound, int32_t *len,
                                  int32_t *len_non_gas, int32_t *iterations, double *potential, double *bnd_energy);

/* ---- serving: continuous batching of independent solves ------------------
 * No counterpart in the reference, which solves independent sets one after
 * another (subfind_unbind, subfind_serial.c:429-560).  Requests are queued in
 * arrival order; `slots` resident contexts (own stream, own device buffers kept
 * between requests) admit the oldest request that fits the HBM budget: each slot
 * is charged the bytes its context holds, or its running request's footprint;
 * idle contexts are replaced by fresh ones when their resident buffers alone
 * keep a request out.  Each request runs compute_grav_accelerations(first_step=1)
 * on its set with the server's parameters and softening table (type 0 from the
 * request when it gives one).  Results are identical to a solve of the same set
 * alone. */
typedef struct agrav_server agrav_server;

typedef struct
{
  int64_t n;
  const double *pos;        /* 3n, host; kept alive by the caller until the request completes */
  const double *mass;       /* n */
  const uint8_t *softtype;  /* n or NULL */
  double softening;         /* softening of type 0 for this set; 0 = the server's parameter */
  double *acc;              /* 3n output (host) */
  double *pot;              /* n output (host) or NULL */
} agrav_request;

typedef struct
{
  agrav_status status;
  double queue_ms, run_ms, total_ms; /* submit->start, start->end, submit->end */
  int slot;
} agrav_request_result;

typedef struct
{
  long long completed, failed;
  int queued, running, peak_running, slots;
  size_t hbm_in_use;    /* what admission charges now: each slot's resident buffers, or its running request's footprint */
  size_t hbm_peak;      /* the largest hbm_in_use so far (never above hbm_budget) */
  size_t hbm_budget;
  size_t hbm_resident;  /* device bytes the slots' contexts hold now (measured) */
  long long evictions;  /* idle contexts replaced by fresh ones to admit a request */
  long long model_exceeded; /* solves that left more resident than their footprint model */
} agrav_server_info;

/* The server's footprint model: device bytes one solve of n particles adds to a fresh context (inputs,
 * sorted copies, keys, octree, results, scratch).  A request is charged a fresh context's bytes plus this. */
size_t agrav_solve_footprint(const agrav_params *p, int64_t n);
/* AGRAV_ERR_CAPACITY when hbm_budget_bytes (0 = 64 GiB) cannot hold `slots` fresh contexts. */
agrav_status agrav_server_create(const agrav_params *p, int slots, size_t hbm_budget_bytes, agrav_server **out);
/* AGRAV_ERR_CAPACITY when the request's footprint, with every other slot at a fresh context, exceeds the budget. */
agrav_status agrav_server_submit(agrav_server *s, const agrav_request *r, uint64_t *ticket);
/* Blocks until the request is done; releases its ticket. */
agrav_status agrav_server_wait(agrav_server *s, uint64_t ticket, agrav_request_result *res);
agrav_status agrav_server_poll(agrav_server *s, uint64_t ticket, int *done);
agrav_status agrav_server_stats(agrav_server *s, agrav_server_info *st);
/* Finishes the queued requests (they stay waitable until then), then stops the workers and frees the slots. */
void agrav_server_destroy(agrav_server *s);

/* Standalone 63-bit Peano-Hilbert key of 21-bit integer coordinates (host side, for checks). */
uint64_t agrav_hilbert_key21(uint32_t x, uint32_t y, uint32_t z);

#ifdef __cplusplus
}
#endif
#endif
