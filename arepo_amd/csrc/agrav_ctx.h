// agrav_ctx.h -- the context behind the C ABI: parameters, the HIP stream,
// and every device buffer of one gravity solver instance.  Buffers grow on
// demand and are kept between calls, so a context that is reused (the serving
// loop) keeps its tree and meshes resident in HBM.
#pragma once

#include <rocfft/rocfft.h>

#include <utility>
#include <vector>

#include "agrav_internal.h"

namespace agrav
{

// Every device buffer of a context registers itself in the context's list while
// the context is constructed (agrav_create sets dbuf_registry for that), so the
// HBM a context holds is known exactly: the sum of `bytes` over the list.
struct DBufBase;
extern thread_local DBufBase **dbuf_registry;

struct DBufBase
{
  DBufBase *reg_next = nullptr;
  size_t bytes       = 0;  // device bytes currently allocated
  DBufBase()
  {
    if(dbuf_registry)
      {
        reg_next       = *dbuf_registry;
        *dbuf_registry = this;
      }
  }
  DBufBase(const DBufBase &)            = delete;
  DBufBase &operator=(const DBufBase &) = delete;
};

template <class T> struct DBuf : DBufBase
{
  T *p       = nullptr;
  size_t cap = 0;  // elements
  agrav_status ensure(size_t count)
  {
    if(count <= cap && p)
      return AGRAV_OK;
    release();
    size_t nb = (count ? count : 1) * sizeof(T);
    if(hipMalloc((void **)&p, nb) != hipSuccess)
      {
        (void)hipGetLastError();
        p = nullptr;
        set_error_msg("hipMalloc failed");
        return AGRAV_ERR_OOM;
      }
    cap   = count ? count : 1;
    bytes = nb;
    return AGRAV_OK;
  }
  void release()
  {
    if(p)
      (void)hipFree(p);
    p     = nullptr;
    cap   = 0;
    bytes = 0;
  }
  // exchange the allocations of two buffers (each stays in its own registry)
  void swap_storage(DBuf &o)
  {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
    std::swap(bytes, o.bytes);
  }
};

struct PmState
{
  int grid = 0;
  int prec = -1;  // agrav_precision of the plans
  rocfft_plan fwd = nullptr, bwd = nullptr;
  rocfft_execution_info info_f = nullptr, info_b = nullptr;
  DBuf<char> work;
  DBuf<unsigned long long> rho_fix;  // fixed-point CIC mass grid (deterministic deposit)
  DBuf<char> real;                   // G^3 real grid (float or double)
  DBuf<char> cplx;                   // G*G*(G/2+1) complex grid
  DBuf<char> force;                  // 3 x G^3 real force grids
  DBuf<double> green;                // G*G*(G/2+1) Green's function x deconvolution (FP64)
  double mass_scale = 0;             // fixed-point units per mass unit
};

// The high-resolution PM region of PLACEHIGHRESREGION (agrav_pmhr.hip): its geometry
// (pm_init_regionsize, pm_nonperiodic.c:135-245), the zero-padded grid of 2 PMGRID
// cells with its rocFFT plans and the transformed two-scale kernel, kept between calls
// while the region holds its particles.
struct HrState
{
  bool valid = false;
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0}, corner[3] = {0, 0, 0}, upper[3] = {0, 0, 0};
  double meshsize = 0, asmth = 0, rcut = 0;
  int grid = 0, prec = -1;
  bool kernel_ready = false;
  rocfft_plan fwd = nullptr, bwd = nullptr;
  rocfft_execution_info info_f = nullptr, info_b = nullptr;
  DBuf<char> work;
  DBuf<unsigned long long> rho_fix;  // fixed-point mass on the padded grid
  DBuf<char> real;                   // padded real grid (density, then potential)
  DBuf<char> cplx;                   // its transform
  DBuf<char> kern;                   // transformed kernel with the CIC deconvolution
  DBuf<char> mesh;                   // {F_x, F_y, F_z, phi} over the inner half (GRID/2)^3
  DBuf<unsigned long long> ext;      // extent of the region's particles: min[3], max[3] (ordered bits), count, outside
  DBuf<double> mpart;                // per-block masses of the high-resolution particles
  double m_high = 0;                 // mass of the high-resolution particles (forcetree.c:440-456)
};

// device buffers of the unbinding loop (agrav_unbind.hip), kept between calls
struct UnbindBufs
{
  DBuf<double> pos, vel, mass, uth, pot, E, cube, epsr, wbl, lim, ppos, sv, cpart, cmin;
  DBuf<long long> cimin;
  DBuf<uint8_t> type, alive;
  DBuf<int> gid, flag, dst, cidx, glist, gphase, res, cnt, glen, cbeg, ccnt;
  DBuf<int2> chunk;
  DBuf<int64_t> off, soff;
  DBuf<char> scan_tmp;
  std::vector<int2> h_chunk;  // host copy of chunk (source of its asynchronous upload)
  void release()
  {
    pos.release(), vel.release(), mass.release(), uth.release(), pot.release(), E.release();
    cube.release(), epsr.release(), wbl.release(), lim.release(), ppos.release(), sv.release();
    cpart.release(), cmin.release(), cimin.release(), type.release(), alive.release();
    gid.release(), flag.release(), dst.release(), cidx.release(), glist.release(), gphase.release();
    res.release(), cnt.release(), glen.release(), cbeg.release(), ccnt.release(), chunk.release();
    off.release(), soff.release(), scan_tmp.release();
  }
};

}  // namespace agrav

struct agrav_ctx
{
  agrav_params par{};
  hipStream_t stream = nullptr;
  // page-locked host block for the small read-backs of a step (counters, node count,
  // domain): copies to it are asynchronous, so each read-back costs one stream sync
  static constexpr size_t kStageBytes = 4096;
  void *hstage = nullptr;
  template <typename T> T *stage(size_t off) { return reinterpret_cast<T *>(static_cast<char *>(hstage) + off); }
  int device         = 0;
  std::string err;

  agrav::SoftTable soft{};
  agrav::DBuf<double> d_soft;   // ForceSoftening on device (nsoft+1)
  float tab_force[AG_NTAB + 1];
  float tab_pot[AG_NTAB + 1];
  agrav::DBuf<float> d_tab;     // [force 128 | pot 128]
  agrav::DBuf<double> d_ewald;  // tree-walk Ewald table 4 x 65^3 (FP64)
  agrav::DBuf<float> d_ewald_f; // same in FP32 for the FP32 walk
  agrav::DBuf<double> d_ewald_ft; // FORCETEST table 4 x 129^3
  bool ewald_ready = false, ewald_ft_ready = false;

  // particles in caller order
  int64_t n = 0;
  agrav::DBuf<double> d_pos_in, d_mass_in;
  agrav::DBuf<uint8_t> d_stype_in;
  agrav::DBuf<float> d_oldacc;
  bool have_stype = false, have_oldacc = false;
  double mass_total = 0;

  // batch of independent isolated sets (batch_nsets = 0: one particle set)
  int batch_nsets = 0;
  int geo_sbits = 0, geo_levels = AG_MAX_LEVELS;
  std::vector<int64_t> h_soff;       // set offsets (nsets + 1), caller and sorted order alike
  agrav::DBuf<int64_t> d_soff;
  agrav::DBuf<int> d_pset_in, d_pset; // set of each particle in caller / sorted order
  agrav::DBuf<int> d_sroot, d_send;   // top entry of each set and the entry after it
  agrav::DBuf<agrav::Domain> d_sdom;  // domain cube per set
  agrav::DBuf<double> d_shd;          // softening length h = 2.8 eps per set
  const double *batch_cubes = nullptr; // device {centre[3], len} per set: given root cubes, or null
  agrav::DBuf<int4> d_groups;         // walk groups {first target, count, set, leaf bucket}
  agrav::DBuf<int2> d_gsets;          // {set, first walk group} of each non-empty set, largest first
  std::vector<int2> h_gsets;          // host copy of d_gsets (source of its asynchronous upload)
  int batch_ngroups = 0;
  bool batch_groups_dirty = true;
  int batch_gs[4] = {64, 64, 64, 64}; // group size per set-size class
  int batch_lb[4] = {4, 4, 4, 4};     // leaf bucket per set-size class
  int batch_tuned = 0;

  // keys and sorted particles
  agrav::DBuf<uint64_t> d_key, d_key2;
  agrav::DBuf<uint64_t> d_key_lo, d_key_lo_s; // second key words: caller order, sorted order (tie runs only)
  agrav::DBuf<int> d_ties;                    // first sorted position of every run of equal first words, then
                                              // of the runs too long for one thread
  agrav::DBuf<int> d_idx, d_idx2;
  agrav::DBuf<double> d_px, d_py, d_pz, d_pm;
  agrav::DBuf<uint8_t> d_ps;
  agrav::DBuf<int> d_pnext;
  agrav::DBuf<uint4> d_prec;
  agrav::DBuf<float> d_oldacc_s;

  // octree construction: per particle, the levels of the cubes starting there
  agrav::DBuf<unsigned> d_lvlmask;
  agrav::DBuf<int> d_pcnt, d_pbase;
  agrav::DBuf<int> d_bcnt;            // per-block partial counts

  // octree (pre-order)
  int m = 0;
  agrav::DBuf<int> d_nfirst, d_nlast, d_nlevel;
  agrav::DBuf<double> d_sx, d_sy, d_sz, d_nm, d_cx, d_cy, d_cz, d_len;
  agrav::DBuf<int> d_nextnode, d_sibling;
  agrav::DBuf<uint8_t> d_maxsoft;
  agrav::DBuf<uint4> d_recA, d_recB;
  agrav::DBuf<int2> d_nrange;
  agrav::DBuf<int> d_cbeg, d_ccnt, d_cnn, d_child;
  agrav::DBuf<int4> d_meta;
  agrav::DBuf<uint4> d_cprec;
  agrav::DBuf<int4> d_nmeta;           // node walk records at their child-list positions
  agrav::DBuf<uint4> d_nrecA, d_nrecB;
  agrav::DBuf<double> d_mass_part;      // block sums of the total mass
  agrav::DBuf<int> d_npos;               // position of each node in its parent's child list
  // moments from prefix sums: per-tile sums, prefix high (FP64) and low (FP32)
  // parts of (m, m x, m y, m z), per-type particle counts (several softening types)
  agrav::DBuf<char> d_mom_tiles;
  agrav::DBuf<int4> d_mom_tab;  // batch: tiles aligned to the sets' first particles
  agrav::DBuf<int> d_mom_tbeg;  // batch: first tile of every set
  agrav::DBuf<double4> d_mom_hi;
  agrav::DBuf<float4> d_mom_lo;
  agrav::DBuf<int> d_tcnt;
  int lvlcnt_h[AG_MAX_LEVELS + 2];

  // walk outputs in sorted order (G=1 units) and per-target counts
  agrav::DBuf<double> d_at, d_pt;
  agrav::DBuf<int> d_nint;
  // Walk groups are launched in descending order of their estimated cost, the tile
  // entries each group streamed in the last walk of the same grouping (as AREPO balances
  // its domains by the last step's GravCost): the costliest groups start first, so that
  // no long group is left running alone at the end of the launch.  Group g of the next
  // walk holds about the same particles while n and the group size are unchanged.
  agrav::DBuf<int> d_gcost;               // per group: tile entries of the last walk
  agrav::DBuf<uint64_t> d_gkey, d_gkey2;  // group sort keys
  agrav::DBuf<int> d_gord, d_gord2;       // group indices; sorted: the launch order
  agrav::DBuf<char> d_gtemp;              // the group sort's look-back states and digit counts
  int gcost_groups = -1;     // d_gcost holds the costs of a walk of this many groups (-1: none) ...
  int gcost_gs     = 0;      // ... of this group size ...
  int64_t gcost_n  = -1;     // ... over every one of this many particles
  const int *walk_gorder = nullptr;  // launch order of the current group walk, or null
  int *walk_gcost        = nullptr;  // where the current group walk records its groups' costs, or null
  agrav::DBuf<int> d_stack;       // group-walk spill/overflow bookkeeping
  agrav::DBuf<int> d_failed;      // groups that overflowed their LDS stack
  agrav::DBuf<int> d_counters;    // [0] group work counter, [1] failed count, [2..] flags
  agrav::DBuf<unsigned long long> d_isum; // interaction total

  // PM
  agrav::PmState pm;
  // high-resolution region: particle types (caller order), the high-resolution flag of
  // every particle (caller and sorted order), the walk's target lists split by it
  agrav::HrState hr;
  agrav::DBuf<uint8_t> d_ptype_in, d_hr, d_hr_s;
  agrav::DBuf<int> d_hr_flag, d_hr_pos, d_hr_list;
  bool have_ptype = false;
  const int *walk_tmap = nullptr;  // the target list of the current walk pass (walk_nt >= 0)
  int64_t walk_nt      = -1;
  agrav::DBuf<double> d_acc_pm, d_pot_pm;  // caller order, or sorted order when pm_sorted
  agrav::DBuf<double> d_acc_pm2, d_pot_pm2;  // caller-order copies made before a partial tree replaces the order
  bool pm_sorted = false;
  bool pm_time_pending = false;  // the last PM's events are recorded, its time not yet read
  // a full step's periodic PM runs on its own stream beside the octree kernels: forked
  // once the gather has written the Hilbert-ordered particles, joined before the walk
  hipStream_t pm_stream = nullptr;  // created with a periodic PMGRID (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_count = nullptr;  // the build's node-count read-back (main stream)
  // a one-set build runs the prefix sums of the moments on a stream of their own beside
  // the octree kernels (they need only the sorted particles); joined before the moments
  hipStream_t mom_stream = nullptr;  // default priority
  hipEvent_t ev_mfork = nullptr, ev_mjoin = nullptr;
  bool mom_forked = false;  // ev_mjoin marks the end of the last work forked to mom_stream
  bool bcopy_side = false;  // the build's read-back copies went to mom_stream (ev_mjoin)
  // a full step's walk order, sorted on mom_stream beside the build (event ev_order)
  hipEvent_t ev_order = nullptr;
  const int *pre_gorder = nullptr;
  int pre_ngroups = -1, pre_gs = 0;  // its grouping (-1: none pending)
  bool isum_clear = false;           // the walk counters were cleared before that sort
  bool pm_fork   = false;  // the next build forks the PM (set by compute_grav_accelerations)
  bool pm_forked = false;  // the build forked it; the main stream has not joined it yet
  // a full step reads the build's error flags back with the walk's counters (one
  // synchronisation); the walks test the flags on the device first
  bool defer_build_check   = false;  // set by compute_grav_accelerations around its build
  bool build_check_pending = false;  // the build's read-back is queued, not yet checked
  bool build_time_pending  = false;  // its duration (events ev[7], ev[1]) is read with the check
  // finalize fused into the group walk's epilogue (set by compute_grav_accelerations)
  bool fuse_final   = false;  // the next group walk may write the final outputs
  bool final_fused  = false;  // the last walk wrote them for every target
  agrav::FinalOut fin{};      // what the walk kernel writes (perm null: nothing)

  // active targets of the walk and the finalize (TimeBinsGravity.ActiveParticleList);
  // n_active < 0: every particle
  int64_t n_active = -1;
  agrav::DBuf<int> d_active;      // caller indices as given
  agrav::DBuf<int> d_tmap;        // their sorted positions, in Hilbert order
  agrav::DBuf<int> d_aflag, d_apos, d_ainv;
  bool tmap_dirty   = true;       // the list or the sorted order changed
  bool outputs_zero = false;      // the caller-order outputs hold values for every particle
  int64_t active_version = 0;     // bumped by every new list

  // HIERARCHICAL_GRAVITY (agrav_set_partial_tree): builds insert only the active
  // particles; the finalize may leave OldAcc as it is
  bool partial_tree  = false;     // the setting
  bool keep_oldacc   = false;
  bool tree_partial  = false;     // the current tree holds the active particles only
  int64_t n_tree     = 0;         // particles in the current tree
  int64_t tree_active_version = -1;

  // final results in caller order
  agrav::DBuf<double> d_acc, d_pot;
  agrav::DBuf<int> d_nint_out;

  // scratch
  agrav::DBuf<char> d_temp;
  agrav::DBuf<agrav::Domain> d_dom;
  agrav::Domain h_dom{};

  // state
  bool have_particles = false, have_tree = false, have_tree_force = false, have_pm = false, have_final = false;

  // unbinding loop
  agrav::UnbindBufs unb;

  // every DBuf member above, linked at construction (device_bytes sums them)
  agrav::DBufBase *bufs = nullptr;

  // timing
  hipEvent_t ev[8] = {};
  agrav_timings tim{};
  int tim_group_waves = 1;  // wavefronts per group of the last group walk
};

namespace agrav
{
agrav_status fail(agrav_ctx *c, agrav_status s, const std::string &m);
size_t device_bytes(const agrav_ctx *c);
agrav_status build_tree(agrav_ctx *c);
agrav_status build_check(agrav_ctx *c);
agrav_status walk_tree(agrav_ctx *c, int use_bh);
agrav_status active_targets(agrav_ctx *c);
agrav_status pm_init(agrav_ctx *c);
void pm_free(agrav_ctx *c);
agrav_status pm_force(agrav_ctx *c);
agrav_status pm_fork(agrav_ctx *c);
agrav_status pm_time_settle(agrav_ctx *c);
bool hr_on(const agrav_ctx *c);
void hr_free(agrav_ctx *c);
agrav_status hr_pm(agrav_ctx *c);
agrav_status hr_flags(agrav_ctx *c);
agrav_status hr_split_targets(agrav_ctx *c, const int *tmap, int64_t nt, const int **lo, int64_t *n_lo, const int **hi,
                              int64_t *n_hi);
agrav_status ewald_tree_table(agrav_ctx *c);
agrav_status ewald_forcetest_table(agrav_ctx *c);
agrav_status forcetest(agrav_ctx *c, int ns, const int *sel, double *acc, double *acc_short, double *pot, double *pot_short);
agrav_status direct_all(agrav_ctx *c, double *acc, double *pot);
agrav_status noalloc_scratch_init(agrav_ctx *c);
agrav_status direct_noalloc(agrav_ctx *c, int64_t n, const double *pos, const double *mass, const uint8_t *stype, double *acc,
                            double *pot, hipStream_t s);
agrav_status finalize(agrav_ctx *c);
agrav_status final_out(agrav_ctx *c, FinalOut &f);
WalkConst walk_const(const agrav_ctx *c, int use_bh);
BatchView batch_view(const agrav_ctx *c);
agrav_status batch_groups(agrav_ctx *c);
// fin: the final outputs the walk's epilogue writes (perm null: none); *fused is set
// when every target's outputs were written so (no group re-walked)
agrav_status walk_preorder(agrav_ctx *c);
agrav_status walk_batch_groups(agrav_ctx *c, int use_bh, const int4 *groups, int ngroups, float *kernel_ms,
                               const FinalOut &fin = FinalOut{}, bool *fused = nullptr);
TreeView tree_view(const agrav_ctx *c);
}  // namespace agrav
