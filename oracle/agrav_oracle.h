/*
 * agrav_oracle.h -- CPU FP64 restatement of AREPO's gravity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path (libagrav): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function follows the reference file:line given in its comment
 * (paths relative to the AREPO source tree).  Semantics are AREPO with
 * DOUBLEPRECISION=1, DOUBLEPRECISION_FFTW, EVALPOTENTIAL, one MPI rank,
 * DO_NOT_RANDOMIZE_DOMAINCENTER and zero random box displacement.
 */
#ifndef AGRAV_ORACLE_H
#define AGRAV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NTAB 127          /* forcetree.h:86 */
#define ORC_MAX_TREE_LEVEL 30 /* forcetree.h:106 */
#define ORC_MAXSOFT 8

typedef struct
{
  int periodic;          /* 1: periodic box (GRAVITY_NOT_PERIODIC unset) */
  double box_size;       /* All.BoxSize */
  int pmgrid;            /* 0: tree only, else PMGRID */
  double asmth;          /* ASMTH in mesh cells (allvars.h:461) */
  double rcut;           /* RCUT in units of Asmth (allvars.h:469) */
  double G;              /* All.G */
  double err_tol_theta;  /* All.ErrTolTheta */
  double err_tol_force_acc; /* All.ErrTolForceAcc */
  int opening;           /* TypeOfOpeningCriterion: 0 BH, 1 relative */
  int nsoft;             /* number of softening types */
  double soft[ORC_MAXSOFT]; /* Plummer-equivalent softening per type */
} orc_params;

typedef struct orc_tree orc_tree;

/* The high-resolution PM region of PLACEHIGHRESREGION (pm_nonperiodic.c:135-245): mask and enlarge are
 * inputs; orc_highres_regionsize fills the rest (All.Xmintot[1], All.Xmaxtot[1], All.Corner[1],
 * All.UpperCorner[1], All.TotalMeshSize[1], All.Asmth[1], All.Rcut[1]). */
#define ORC_GRIDBOOST 2 /* GRIDBOOST default, pm_nonperiodic.c:75-77 */
typedef struct
{
  int mask;       /* PLACEHIGHRESREGION: bit t set for the high-resolution particle types t */
  double enlarge; /* ENLARGEREGION (<= 0: not set) */
  int valid;
  double xmin[3], xmax[3], corner[3], upper[3];
  double meshsize, asmth, rcut;
} orc_highres;

/* grav_softening.c:55-96: ForceSoftening[i] = 2.8 eps_i, sentinel 0 at [nsoft] */
void orc_set_softenings(const orc_params *p, double *force_softening);

/* forcetree_walk.c:72-85 */
void orc_short_range_tables(double rcut, float *tab_force, float *tab_pot);

/* pm_periodic.c:190-191 */
void orc_asmth_rcut(const orc_params *p, double *asmth_len, double *rcut_len);

/* domain_box.c:107-197 */
void orc_domain_extent(const orc_params *p, int n, const double *pos, double corner[3], double center[3], double *len);

/* forcetree.c:110-823 (+ insert :835-938, moments :1217-1429).
 * Returns number of nodes, or -1 when alloc_factor*n nodes are too few. */
int orc_tree_build(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype, double alloc_factor,
                   orc_tree **out);
/* Same with a given root cube {centre[3], len} (NULL: the domain cube of domain_box.c) */
int orc_tree_build_cube(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype,
                        double alloc_factor, const double *cube, orc_tree **out);
void orc_tree_free(orc_tree *t);
int orc_tree_numnodes(const orc_tree *t);
/* Copy node data: center[3M], len[M], com[3M], mass[M], level[M], npart[M] (any may be NULL) */
void orc_tree_export(const orc_tree *t, double *center, double *len, double *com, double *mass, int *level, int *npart);

/* forcetree_walk.c:115-514 for each target (all particles, in input order).
 * use_bh: 1 -> Barnes-Hut with err_tol_theta, 0 -> relative with old_acc.
 * ewald_tab: NULL or the 4x65^3 table from orc_ewald_table_tree (periodic tree-only).
 * Returns the total interaction count.  nthreads<=0 -> all OpenMP threads. */
long long orc_tree_walk(const orc_params *p, const orc_tree *t, int use_bh, const float *old_acc, const double *ewald_tab,
                        int ntarget, const int *targets, double *acc, double *pot, int *nint, int nthreads);

/* pm_periodic.c:1572-1936 with CIC deposit :1049-1170 and readout :1322-1550.
 * acc_pm includes G (GravPM), pot_pm includes 4 pi G / V (PM_Potential). */
int orc_pm_periodic(const orc_params *p, int n, const double *pos, const double *mass, double *acc_pm, double *pot_pm);

/* forcetree_ewald.c:393-450 (psi) and :463-527 (force), unscaled (unit box) */
double orc_ewald_psi(double x, double y, double z);
void orc_ewald_force(double x, double y, double z, double force[3]);
/* forcetree_ewald.c:89-252: 4 x (EN+1)^3 table [fx,fy,fz,pot][i][j][k], scaled to box */
void orc_ewald_table_tree(int en, double box, double *tab, int nthreads);
/* forcetree_ewald.c:268-382: trilinear lookups on that table */
void orc_ewald_corr(const double *tab, int en, double box, double dx, double dy, double dz, double fcorr[3]);
double orc_ewald_pot_corr(const double *tab, int en, double box, double dx, double dy, double dz);

/* gravtree_forcetest.c:820-916: exact series, force[4] = {fx,fy,fz,pot} with sign convention of the table */
void orc_forcetest_ewald_exact(double box, double x, double y, double z, double force[4]);
/* gravtree_forcetest.c:933-1008: 4 x (tew_n+1)^3 table */
void orc_forcetest_ewald_table(int tew_n, double box, double *tab, int nthreads);
/* gravtree_forcetest.c:1019-1085 */
void orc_forcetest_ewald_lookup(const double *tab, int tew_n, double box, double dx, double dy, double dz, double force[4]);

/* gravtree_forcetest.c:421-648 for selected targets.  ewald: 0 none, 1 table (tab/tew_n), 2 exact series.
 * Outputs (each may be NULL): acc[3*ns], acc_short[3*ns], pot[ns], pot_short[ns] -- all in G=1 units. */
void orc_forcetest(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype, int ns,
                   const int *sel, int ewald, const double *tab, int tew_n, double *acc, double *acc_short, double *pot,
                   double *pot_short, int nthreads);

/* accel.c:58-80 + gravity() :94-192 + gravity_force_finalize() :204-347 + long_range_force() longrange.c:98-198.
 * first_step=1 runs the Barnes-Hut seeding pass (accel.c:62-72) before the relative pass.
 * old_acc (in/out, may be NULL when first_step=1 and opening=1 or opening=0).
 * Outputs: acc[3n] (G included, tree+PM), pot[n], nint[n] of the final pass.
 * Timings (seconds) of the final pass written to t_out[3] = {treebuild, treewalk, pm} when non-NULL.
 * Returns total interactions of the final pass, or a negative error. */
long long orc_compute_grav_accelerations(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype,
                                         int first_step, float *old_acc, double *acc, double *pot, int *nint, int nthreads,
                                         double *t_out);

/* Direct summation pair kernel of forcetree_walk.c:557-708 (ALLOW_DIRECT_SUMMATION), all pairs, G=1 units */
void orc_direct_summation(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype, double *acc,
                          double *pot, int nthreads);

/* Unbinding potential of one isolated group with a local BH tree: subfind_loctree.c:564-730 (potential only,
 * ErrTolThetaSubfind criterion, spline softening, G=1).  Returns interactions. */
long long orc_group_potential_bh(int n, const double *pos, const double *mass, double h, double theta, double *pot);
/* Same with the local tree's root cube {centre[3], len} given (NULL: isolated domain cube) */
long long orc_group_potential_bh_cube(int n, const double *pos, const double *mass, double h, double theta, const double *cube,
                                      double *pot);

/* subfind_loctree_findExtent, subfind_loctree.c:101-160: cube[4] = {centre[3], 1.001 x largest extent} */
void orc_loctree_extent(int n, const double *pos, double cube[4]);

typedef struct
{
  double G;          /* All.G */
  double time_a;     /* All.Time (scale factor when comoving) */
  double hubble_a;   /* hubble_function(All.Time) */
  int comoving;      /* All.ComovingIntegrationOn */
  double theta;      /* All.ErrTolThetaSubfind */
  int des_link_ngb;  /* All.DesLinkNgb */
  int max_iter;      /* MAXITER, allvars.h:482 */
} orc_unbind_params;

/* subfind_unbind, subfind_serial.c:429-637, one isolated group; eps = Plummer-equivalent softening
 * (ForceSoftening/2.8).  cube: local-tree root {centre[3], len} or NULL for the group's own extent.
 * utherm/type may be NULL (no gas).  bound[n] = 1 for the members left.  potential / bnd_energy (n, may
 * be NULL): PS.Potential and PS.BindingEnergy as last computed (removed members keep their last values).
 * Returns the number left, or -1 past max_iter iterations. */
int orc_subfind_unbind(const orc_unbind_params *up, int n, const double *pos, const double *vel, const double *mass,
                       const double *utherm, const uint8_t *type, double eps, const double *cube, uint8_t *bound,
                       double *potential, double *bnd_energy, int *len_non_gas, int *iterations);

typedef struct
{
  double err_tol_int_accuracy; /* All.ErrTolIntAccuracy */
  double max_size_timestep;    /* All.MaxSizeTimestep */
  double min_size_timestep;    /* All.MinSizeTimestep */
  double time_a;               /* All.Time when comoving */
  double hubble_a;             /* hubble_function(All.Time) when comoving */
  int comoving;                /* All.ComovingIntegrationOn */
} orc_timestep_params;

/* find_long_range_step_constraint (timestep.c:525-565) over acc_pm (NULL: no PM) and get_timestep_gravity's
 * dt (timestep.c:284-344) before the integer timeline, for n particles with total acceleration acc
 * (GravAccel + GravPM).  A dt below MinSizeTimestep is raised to it and counted (the return value). */
int orc_gravity_timesteps(const orc_timestep_params *tp, const orc_params *p, int n, const double *acc, const double *acc_pm,
                          const double *mass, const uint8_t *stype, double *dt, double *dt_displacement);

/* Textbook in-place complex FFT of a cube (n^3, n power of two); sign -1 forward (FFTW_FORWARD), +1 backward; unnormalised */
void orc_fft3d(int n, double *re, double *im, int sign);

/* pm_init_regionsize for the high-resolution grid (pm_nonperiodic.c:135-245) from the particles of the
 * mask's types; -1 when there are none */
int orc_highres_regionsize(const orc_params *p, orc_highres *r, int n, const double *pos, const uint8_t *type);
/* pmforce_is_particle_high_res (pm_nonperiodic.c:344-371) */
int orc_highres_flag(const orc_params *p, const orc_highres *r, int type, const double *pos);
void orc_highres_flags(const orc_params *p, const orc_highres *r, int n, const double *pos, const uint8_t *type, uint8_t *flags);
/* pmforce_nonperiodic(1) with pm_setup_nonperiodic_kernel (pm_nonperiodic.c:1508-1987): adds the
 * high-resolution PM force (G included) and potential of the high-resolution particles into acc_pm and
 * pot_pm; returns 1 (nothing added) when one of them lies outside the region */
int orc_pm_highres(const orc_params *p, const orc_highres *r, int n, const double *pos, const double *mass, const uint8_t *type,
                   double *acc_pm, double *pot_pm);
/* the walk, FORCETEST, direct summation and the step with a high-resolution region: a target with hrflag
 * set uses Asmth[1] and Rcut[1] (forcetree_walk.c:143-150, gravtree_forcetest.c:480-483,
 * forcetree_walk.c:576-581); the step sizes the region when hr->valid is 0 */
long long orc_tree_walk_hr(const orc_params *p, const orc_tree *t, int use_bh, const float *old_acc, const double *ewald_tab,
                           const orc_highres *hr, const uint8_t *hrflag, int ntarget, const int *targets, double *acc,
                           double *pot, int *nint, int nthreads);
void orc_forcetest_hr(const orc_params *p, const orc_highres *hr, const uint8_t *hrflag, int n, const double *pos,
                      const double *mass, const uint8_t *stype, int ns, const int *sel, int ewald, const double *tab, int tew_n,
                      double *acc, double *acc_short, double *pot, double *pot_short, int nthreads);
void orc_direct_summation_hr(const orc_params *p, const orc_highres *hr, const uint8_t *hrflag, int n, const double *pos,
                             const double *mass, const uint8_t *stype, double *acc, double *pot, int nthreads);
long long orc_compute_grav_accelerations_hr(const orc_params *p, orc_highres *hr, const uint8_t *type, int n, const double *pos,
                                            const double *mass, const uint8_t *stype, int first_step, float *old_acc, double *acc,
                                            double *pot, int *nint, int nthreads, double *t_out);

/* peano_hilbert_key (peano.c:374-390): the `bits`-level key (bits <= 42) of integer
 * coordinates, as high and low 64-bit words; and its inverse (peano.c:446-496). */
void orc_peano_hilbert_key(uint64_t x, uint64_t y, uint64_t z, int bits, uint64_t *hi, uint64_t *lo);
void orc_peano_hilbert_key_inverse(uint64_t hi, uint64_t lo, int bits, uint64_t *x, uint64_t *y, uint64_t *z);

#ifdef __cplusplus
}
#endif
#endif
