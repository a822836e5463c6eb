// agrav_api.hip -- the C ABI (include/agrav.h): context life cycle, particle
// upload, the compute_grav_accelerations sequence (accel.c:58-347) and
// result download.  Device work of a context runs on its own stream.
#include "agrav_ctx.h"

#include <math.h>
#include <string.h>

#include <vector>

namespace agrav
{

static thread_local std::string g_last_msg;
thread_local DBufBase **dbuf_registry = nullptr;

size_t device_bytes(const agrav_ctx *c)
{
  size_t b = 0;
  for(const DBufBase *d = c->bufs; d; d = d->reg_next)
    b += d->bytes;
  return b;
}

void set_error_msg(const std::string &m) { g_last_msg = m; }

agrav_status hip_fail(hipError_t e, const char *what, const char *file, int line)
{
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  g_last_msg = buf;
  return AGRAV_ERR_HIP;
}

agrav_status fail(agrav_ctx *c, agrav_status s, const std::string &m)
{
  if(c)
    c->err = m;
  g_last_msg = m;
  return s;
}

static constexpr int B = 256;

// gravity_force_finalize, accel.c:204-347 (no comoving terms, OmegaLambda = 0), for
// every particle or for the active targets (sorted positions in tmap)
__global__ void k_finalize(int64_t n, const int *__restrict__ tmap, FinalOut f, const double *__restrict__ at,
                           const double *__restrict__ pt, const int *__restrict__ nint)
{
  int64_t j = (int64_t)blockIdx.x * B + threadIdx.x;
  if(j >= n)
    return;
  const int64_t k = tmap ? tmap[j] : j;
  finalize_one(f, k, at[3 * k], at[3 * k + 1], at[3 * k + 2], pt[k], nint[k]);
}

// what finalize writes, with its output buffers allocated
agrav_status final_out(agrav_ctx *c, FinalOut &f)
{
  const int64_t n = c->n;
  agrav_status st;
  if((st = c->d_acc.ensure(3 * (n ? n : 1))) || (st = c->d_pot.ensure(n ? n : 1)) || (st = c->d_nint_out.ensure(n ? n : 1)) ||
     (st = c->d_oldacc.ensure(n ? n : 1)))
    return st;
  if(c->n_active >= 0 && !c->outputs_zero && n > 0)
    {
      // inactive particles keep their outputs; after set_particles they have none yet
      AG_HIP(hipMemsetAsync(c->d_acc.p, 0, 3 * n * sizeof(double), c->stream));
      AG_HIP(hipMemsetAsync(c->d_pot.p, 0, n * sizeof(double), c->stream));
      AG_HIP(hipMemsetAsync(c->d_nint_out.p, 0, n * sizeof(int), c->stream));
    }
  c->outputs_zero = true;
  const bool pm = c->par.pmgrid > 0;
  double pot_const = 0;
  if(pm && c->par.periodic)
    {
      // accel.c:269-274, with the mass of each PM region (forcetree.c:440-456)
      const double asmth = c->par.asmth * c->par.box_size / c->par.pmgrid;
      const double L     = c->par.box_size, vol = L * L * L;
      const double m_hr  = hr_on(c) ? c->hr.m_high : 0.0;
      pot_const          = (c->mass_total - m_hr) * M_PI / (asmth * asmth * vol);
      if(hr_on(c))
        pot_const += m_hr * M_PI / (c->hr.asmth * c->hr.asmth * vol);
    }
  // OldAcc is left as it is on request (accel.c:231-232: set on the highest occupied timebin only)
  f = FinalOut{c->d_idx2.p, pm ? c->d_acc_pm.p : nullptr, pm ? c->d_pot_pm.p : nullptr, c->pm_sorted ? 1 : 0, c->par.G,
               pot_const, c->d_acc.p, c->d_pot.p, c->d_nint_out.p, c->keep_oldacc ? nullptr : c->d_oldacc.p, c->d_oldacc_s.p};
  return AGRAV_OK;
}

agrav_status finalize(agrav_ctx *c)
{
  if(!c->have_tree_force)
    return fail(c, AGRAV_ERR_STATE, "finalize before gravity_tree");
  const int64_t n = c->n;
  agrav_status st;
  const bool pm = c->par.pmgrid > 0;
  if(pm && !c->have_pm)
    return fail(c, AGRAV_ERR_STATE, "finalize with PMGRID before long_range_force");
  if(c->final_fused)
    {
      // the walk's epilogue has written the outputs of every target
      c->final_fused  = false;
      c->tim.finalize = 0;
      c->have_oldacc  = c->have_oldacc || !c->keep_oldacc;
      c->have_final   = true;
      return AGRAV_OK;
    }
  FinalOut f;
  if((st = final_out(c, f)) != AGRAV_OK)
    return st;
  AG_HIP(hipEventRecord(c->ev[0], c->stream));
  const bool tm    = c->n_active >= 0 && !c->tree_partial;
  const int64_t nt = c->tree_partial ? c->n_tree : c->n_active >= 0 ? c->n_active : n;
  if(nt > 0)
    k_finalize<<<(unsigned)((nt + B - 1) / B), B, 0, c->stream>>>(nt, tm ? c->d_tmap.p : nullptr, f, c->d_at.p, c->d_pt.p,
                                                                  c->d_nint.p);
  AG_HIP(hipGetLastError());
  AG_HIP(hipEventRecord(c->ev[1], c->stream));
  AG_HIP(hipEventSynchronize(c->ev[1]));
  float ms = 0;
  AG_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->tim.finalize = ms;
  c->have_oldacc  = c->have_oldacc || !c->keep_oldacc;
  c->have_final   = true;
  return AGRAV_OK;
}

// timestep criteria after a step (timestep.c:284-344, :525-565): the physical
// acceleration a2inv |a|, dt = sqrt(2 ErrTolIntAccuracy a ForceSoftening / 2.8 / ac) H(a)
struct TsArgs
{
  double eta2a;  // 2 ErrTolIntAccuracy cf_atime
  double a2inv, hub, maxdt, mindt;
};

// find_long_range_step_constraint: the minimum over every particle of the dt of its PM
// acceleration (positive doubles order like their bit patterns: one atomic minimum)
__global__ void k_pm_dtmin(int64_t n, const double *__restrict__ acc_pm, const uint8_t *__restrict__ st, const double *__restrict__ fsoft,
                           TsArgs a, unsigned long long *dtmin)
{
  __shared__ unsigned long long s_min[B / 64];
  const int64_t i        = (int64_t)blockIdx.x * B + threadIdx.x;
  unsigned long long v   = ~0ull;
  if(i < n)
    {
      const double ax = a.a2inv * acc_pm[3 * i], ay = a.a2inv * acc_pm[3 * i + 1], az = a.a2inv * acc_pm[3 * i + 2];
      double ac       = sqrt(ax * ax + ay * ay + az * az);
      if(ac < 1e-37)  // MIN_FLOAT_NUMBER
        ac = 1e-37;
      const double d = sqrt(a.eta2a * fsoft[st ? st[i] : 0] / (2.8 * ac)) * a.hub;
      v              = (unsigned long long)__double_as_longlong(d);
    }
  for(int o = 32; o > 0; o >>= 1)
    {
      const unsigned long long w = __shfl_xor(v, o, 64);
      v                          = w < v ? w : v;
    }
  if((threadIdx.x & 63) == 0)
    s_min[threadIdx.x >> 6] = v;
  __syncthreads();
  if(threadIdx.x == 0)
    {
      for(int k = 1; k < B / 64; k++)
        v = s_min[k] < v ? s_min[k] : v;
      atomicMin(dtmin, v);
    }
}

// get_timestep_gravity before the integer timeline, caller order
__global__ void k_grav_dt(int64_t n, const double *__restrict__ acc, const double *__restrict__ mass, const uint8_t *__restrict__ st,
                          const double *__restrict__ fsoft, TsArgs a, int have_pm, double dtdisp, double *dt, int *below)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  const double ax = a.a2inv * acc[3 * i], ay = a.a2inv * acc[3 * i + 1], az = a.a2inv * acc[3 * i + 2];
  double ac       = sqrt(ax * ax + ay * ay + az * az);
  if(ac == 0)
    ac = 1.0e-30;
  double d = sqrt(a.eta2a * fsoft[st ? st[i] : 0] / 2.8 / ac) * a.hub;
  if(mass[i] == 0)
    d = a.maxdt;  // a swallowed particle (timestep.c:328-329)
  if(d >= a.maxdt)
    d = a.maxdt;
  if(d < a.mindt)
    {
      atomicAdd(below, 1);
      d = a.mindt;
    }
  if(have_pm && d >= dtdisp)
    d = dtdisp;
  dt[i] = d;
}

// softening types of a device-resident input must index the softening table
__global__ void k_check_stype(const uint8_t *__restrict__ st, int64_t n, int nsoft, int *bad)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i < n && st[i] >= nsoft)
    atomicOr(bad, 1);
}

// Total mass in an order fixed by n alone: blocks of kMassChunk particles, each summed
// by 256 threads over a fixed stride and then pairwise in LDS; the block sums likewise
// by one block.  Host and device inputs give the same bits.
static constexpr int kMassChunk = 4096;
__global__ __launch_bounds__(B) void k_mass_partial(const double *__restrict__ m, int64_t n, double *part)
{
  __shared__ double sh[B];
  const int64_t b0 = (int64_t)blockIdx.x * kMassChunk;
  double v         = 0;
  for(int k = threadIdx.x; k < kMassChunk; k += B)
    if(b0 + k < n)
      v += m[b0 + k];
  sh[threadIdx.x] = v;
  __syncthreads();
  for(int w = B / 2; w > 0; w >>= 1)
    {
      if((int)threadIdx.x < w)
        sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
  if(threadIdx.x == 0)
    part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(B) void k_mass_final(const double *__restrict__ part, int64_t nb, double *out)
{
  __shared__ double sh[B];
  double v = 0;
  for(int64_t k = threadIdx.x; k < nb; k += B)
    v += part[k];
  sh[threadIdx.x] = v;
  __syncthreads();
  for(int w = B / 2; w > 0; w >>= 1)
    {
      if((int)threadIdx.x < w)
        sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
  if(threadIdx.x == 0)
    out[0] = sh[0];
}

static void soft_from(agrav_ctx *c, int nsoft, const double *eps)
{
  // set_softenings, grav_softening.c:90-95
  c->soft.nsoft = nsoft;
  for(int i = 0; i < nsoft; i++)
    c->soft.h[i] = 2.8 * eps[i];
  c->soft.h[nsoft] = 0;
}

}  // namespace agrav

using namespace agrav;

extern "C" {

const char *agrav_status_string(agrav_status s)
{
  switch(s)
    {
      case AGRAV_OK:
        return "ok";
      case AGRAV_ERR_BAD_PARAM:
        return "bad parameter";
      case AGRAV_ERR_CAPACITY:
        return "capacity exceeded";
      case AGRAV_ERR_OUT_OF_BOX:
        return "particle out of box";
      case AGRAV_ERR_HIP:
        return "HIP error";
      case AGRAV_ERR_OOM:
        return "out of device memory";
      case AGRAV_ERR_STATE:
        return "call order violated";
      case AGRAV_ERR_FFT:
        return "rocFFT error";
      case AGRAV_ERR_NO_DEVICE:
        return "no GPU device";
    }
  return "unknown";
}

const char *agrav_last_error(const agrav_ctx *ctx)
{
  if(ctx && !ctx->err.empty())
    return ctx->err.c_str();
  return g_last_msg.c_str();
}

void agrav_default_params(agrav_params *p)
{
  memset(p, 0, sizeof(*p));
  p->periodic                  = 1;
  p->box_size                  = 1.0;
  p->pmgrid                    = 0;
  p->asmth                     = 1.25;  // allvars.h:461
  p->rcut                      = 4.5;   // allvars.h:469
  p->G                         = 1.0;
  p->err_tol_theta             = 0.7;
  p->err_tol_force_acc         = 0.0025;
  p->type_of_opening_criterion = 1;
  p->n_soft_types              = 1;
  p->softening[0]              = 0.01;
  p->precision                 = AGRAV_FP64;
  p->walk                      = AGRAV_WALK_GROUP;
  p->group_size                = 0;  // by workload (walk_tree)
  p->tree_alloc_factor         = 0;
  p->leaf_bucket               = 8;
  p->eval_potential            = 1;
  p->group_waves               = 0;
}

static agrav_status check_params(const agrav_params *p)
{
  if(!p)
    return AGRAV_ERR_BAD_PARAM;
  if(p->n_soft_types < 1 || p->n_soft_types > AGRAV_MAX_SOFT_TYPES)
    return AGRAV_ERR_BAD_PARAM;
  for(int i = 0; i < p->n_soft_types; i++)
    if(!(p->softening[i] > 0))
      return AGRAV_ERR_BAD_PARAM;
  if(p->periodic && !(p->box_size > 0))
    return AGRAV_ERR_BAD_PARAM;
  if(p->pmgrid < 0 || (p->pmgrid > 0 && (!p->periodic || (p->pmgrid & (p->pmgrid - 1)) || p->pmgrid < 8 || p->pmgrid > 1024)))
    return AGRAV_ERR_BAD_PARAM;
  if(p->pmgrid > 0 && (!(p->asmth > 0) || !(p->rcut > 0)))
    return AGRAV_ERR_BAD_PARAM;
  if(!(p->G > 0) || p->err_tol_theta < 0 || p->err_tol_force_acc < 0)
    return AGRAV_ERR_BAD_PARAM;
  if(p->type_of_opening_criterion != 0 && p->type_of_opening_criterion != 1)
    return AGRAV_ERR_BAD_PARAM;
  if(p->type_of_opening_criterion == 0 && !(p->err_tol_theta > 0))
    return AGRAV_ERR_BAD_PARAM;
  if(p->precision != AGRAV_FP64 && p->precision != AGRAV_FP32)
    return AGRAV_ERR_BAD_PARAM;
  if(p->walk != AGRAV_WALK_PLAIN && p->walk != AGRAV_WALK_GROUP)
    return AGRAV_ERR_BAD_PARAM;
  if(p->precision == AGRAV_FP32 && p->walk == AGRAV_WALK_PLAIN)
    return AGRAV_ERR_BAD_PARAM;  // the per-target walk is the FP64 restatement only
  if(p->group_size < 0 || p->group_size > 64)
    return AGRAV_ERR_BAD_PARAM;
  if(p->leaf_bucket < 0 || p->leaf_bucket > 64)
    return AGRAV_ERR_BAD_PARAM;
  if(p->group_waves != 0 && p->group_waves != 1 && p->group_waves != 2 && p->group_waves != 4)
    return AGRAV_ERR_BAD_PARAM;
  // PLACEHIGHRESREGION: a second grid inside a periodic TreePM box (longrange.c:57-59)
  if(p->placehighresregion < 0 || p->placehighresregion > 255 ||
     (p->placehighresregion != 0 && (!p->periodic || p->pmgrid <= 0)))
    return AGRAV_ERR_BAD_PARAM;
  return AGRAV_OK;
}

agrav_status agrav_create(const agrav_params *p, agrav_ctx **out)
{
  if(!out)
    return AGRAV_ERR_BAD_PARAM;
  *out            = nullptr;
  agrav_status st = check_params(p);
  if(st != AGRAV_OK)
    {
      set_error_msg("invalid agrav_params");
      return st;
    }
  int ndev = 0;
  if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    {
      (void)hipGetLastError();
      set_error_msg("no HIP device available");
      return AGRAV_ERR_NO_DEVICE;
    }
  DBufBase *reg = nullptr;
  dbuf_registry = &reg;
  agrav_ctx *c  = new agrav_ctx();
  dbuf_registry = nullptr;
  c->bufs       = reg;
  c->par        = *p;
  if(hipGetDevice(&c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
     hipHostMalloc(&c->hstage, agrav_ctx::kStageBytes, hipHostMallocDefault) != hipSuccess)
    {
      delete c;
      set_error_msg("cannot create a HIP stream");
      return AGRAV_ERR_HIP;
    }
  for(auto &e : c->ev)
    if(hipEventCreate(&e) != hipSuccess)
      {
        agrav_destroy(c);
        return AGRAV_ERR_HIP;
      }
  if(hipEventCreateWithFlags(&c->ev_count, hipEventDisableTiming) != hipSuccess ||
     hipStreamCreateWithFlags(&c->mom_stream, hipStreamNonBlocking) != hipSuccess ||
     hipEventCreateWithFlags(&c->ev_mfork, hipEventDisableTiming) != hipSuccess ||
     hipEventCreateWithFlags(&c->ev_mjoin, hipEventDisableTiming) != hipSuccess ||
     hipEventCreateWithFlags(&c->ev_order, hipEventDisableTiming) != hipSuccess)
    {
      agrav_destroy(c);
      return AGRAV_ERR_HIP;
    }
  // the PM stream of a periodic PMGRID (same default priority as the main stream)
  if(p->pmgrid > 0 && p->periodic &&
     (hipStreamCreateWithFlags(&c->pm_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess))
    {
      agrav_destroy(c);
      set_error_msg("cannot create the PM stream");
      return AGRAV_ERR_HIP;
    }
  soft_from(c, p->n_soft_types, p->softening);
  // force_short_range_init, forcetree_walk.c:72-85 (FP64 compute, FP32 store)
  for(int i = 0; i <= AG_NTAB; i++)
    {
      double u        = ((p->rcut / 2.0) / AG_NTAB) * i;
      c->tab_pot[i]   = (float)(-erfc(u));
      c->tab_force[i] = u > 0 ? (float)((erfc(u) + 2.0 * u / sqrt(M_PI) * exp(-u * u) - 1.0) / (u * u)) : 0.f;
    }
  // [force 128 | potential 128] as above, then the FP32 TreePM group walk's tables on the
  // same u grid, computed in FP64: the force table times du^2 (that walk works in units
  // of the grid step du), T_i = tab_force[i] du^2, as {T_i, T_(i+1) - T_i} x 128, and
  // {P_i, P_(i+1) - P_i} x 128 with P = -erfc(u); entry NTAB (beyond the cut) is zero
  float tabs[256 + 512];
  memcpy(tabs, c->tab_force, 128 * sizeof(float));
  memcpy(tabs + 128, c->tab_pot, 128 * sizeof(float));
  {
    const double du = (p->rcut / 2.0) / AG_NTAB;
    auto tf = [&](int i) {
      const double u = du * i;
      return u > 0 ? (erfc(u) + 2.0 * u / sqrt(M_PI) * exp(-u * u) - 1.0) / (u * u) * du * du : 0.0;
    };
    for(int i = 0; i <= AG_NTAB; i++)
      {
        float *a = tabs + 256 + 2 * i, *b = tabs + 512 + 2 * i;
        a[0] = a[1] = b[0] = b[1] = 0;
        if(i == AG_NTAB)
          continue;
        a[0] = (float)tf(i);
        a[1] = (float)(tf(i + 1) - tf(i));
        b[0] = (float)(-erfc(du * i));
        b[1] = (float)(-erfc(du * (i + 1)) + erfc(du * i));
      }
  }
  if((st = c->d_tab.ensure(256 + 512)) || (st = c->d_soft.ensure(AG_MAXSOFT + 1)))
    {
      agrav_destroy(c);
      return st;
    }
  if(hipMemcpy(c->d_tab.p, tabs, sizeof(tabs), hipMemcpyHostToDevice) != hipSuccess ||
     hipMemcpy(c->d_soft.p, c->soft.h, sizeof(c->soft.h), hipMemcpyHostToDevice) != hipSuccess)
    {
      agrav_destroy(c);
      return AGRAV_ERR_HIP;
    }
  if(c->par.pmgrid > 0 && (st = pm_init(c)) != AGRAV_OK)
    {
      set_error_msg(c->err);
      agrav_destroy(c);
      return st;
    }
  // ewald_init (begrun.c:157, forcetree_ewald.c:89-252): periodic tree-only
  if(c->par.periodic && c->par.pmgrid == 0 &&
     ((st = ewald_tree_table(c)) != AGRAV_OK || hipStreamSynchronize(c->stream) != hipSuccess))
    {
      set_error_msg(c->err);
      agrav_destroy(c);
      return st != AGRAV_OK ? st : AGRAV_ERR_HIP;
    }
  // the device address of the fixed scratch array of agrav_gravity_direct_noalloc
  if((st = noalloc_scratch_init(c)) != AGRAV_OK)
    {
      set_error_msg(c->err);
      agrav_destroy(c);
      return st;
    }
  *out = c;
  return AGRAV_OK;
}

void agrav_destroy(agrav_ctx *c)
{
  if(!c)
    return;
  if(c->stream)
    (void)hipStreamSynchronize(c->stream);
  if(c->pm_stream)
    (void)hipStreamSynchronize(c->pm_stream);
  if(c->mom_stream)
    (void)hipStreamSynchronize(c->mom_stream);
  if(c->hstage)
    (void)hipHostFree(c->hstage);
  pm_free(c);
  hr_free(c);
  for(auto &e : c->ev)
    if(e)
      (void)hipEventDestroy(e);
  for(hipEvent_t e : {c->ev_fork, c->ev_join, c->ev_count, c->ev_mfork, c->ev_mjoin, c->ev_order})
    if(e)
      (void)hipEventDestroy(e);
  // DBuf members release their memory explicitly
  c->d_soft.release();
  c->d_tab.release();
  c->d_ewald.release();
  c->d_ewald_f.release();
  c->d_ewald_ft.release();
  c->d_pos_in.release();
  c->d_mass_in.release();
  c->d_stype_in.release();
  c->d_oldacc.release();
  c->d_key.release();
  c->d_key2.release();
  c->d_key_lo.release();
  c->d_key_lo_s.release();
  c->d_ties.release();
  c->d_idx.release();
  c->d_idx2.release();
  c->d_px.release();
  c->d_py.release();
  c->d_pz.release();
  c->d_pm.release();
  c->d_ps.release();
  c->d_pnext.release();
  c->d_prec.release();
  c->d_oldacc_s.release();
  c->d_lvlmask.release();
  c->d_pcnt.release();
  c->d_pbase.release();
  c->d_bcnt.release();
  c->d_nfirst.release();
  c->d_nlast.release();
  c->d_nlevel.release();
  c->d_sx.release();
  c->d_sy.release();
  c->d_sz.release();
  c->d_nm.release();
  c->d_cx.release();
  c->d_cy.release();
  c->d_cz.release();
  c->d_len.release();
  c->d_nextnode.release();
  c->d_sibling.release();
  c->d_maxsoft.release();
  c->d_recA.release();
  c->d_recB.release();
  c->d_nrange.release();
  c->d_cbeg.release();
  c->d_ccnt.release();
  c->d_cnn.release();
  c->d_child.release();
  c->d_meta.release();
  c->d_cprec.release();
  c->d_nmeta.release();
  c->d_nrecA.release();
  c->d_nrecB.release();
  c->d_npos.release();
  c->d_mass_part.release();
  c->d_mom_tiles.release();
  c->d_mom_tab.release();
  c->d_mom_tbeg.release();
  c->d_mom_hi.release();
  c->d_mom_lo.release();
  c->d_tcnt.release();
  c->d_soff.release();
  c->d_pset_in.release();
  c->d_pset.release();
  c->d_sroot.release();
  c->d_send.release();
  c->d_sdom.release();
  c->d_shd.release();
  c->d_groups.release();
  c->d_gsets.release();
  c->d_at.release();
  c->d_pt.release();
  c->d_nint.release();
  c->d_stack.release();
  c->d_failed.release();
  c->d_counters.release();
  c->d_isum.release();
  c->d_acc_pm.release();
  c->d_pot_pm.release();
  c->d_acc_pm2.release();
  c->d_pot_pm2.release();
  c->d_acc.release();
  c->d_pot.release();
  c->d_nint_out.release();
  c->d_temp.release();
  c->d_dom.release();
  c->d_active.release();
  c->d_tmap.release();
  c->d_aflag.release();
  c->d_apos.release();
  c->d_ainv.release();
  c->unb.release();
  if(c->pm_stream)
    (void)hipStreamDestroy(c->pm_stream);
  if(c->mom_stream)
    (void)hipStreamDestroy(c->mom_stream);
  if(c->stream)
    (void)hipStreamDestroy(c->stream);
  delete c;
}

agrav_status agrav_set_softenings(agrav_ctx *c, int nsoft, const double *eps)
{
  if(!c || !eps || nsoft < 1 || nsoft > AGRAV_MAX_SOFT_TYPES)
    return AGRAV_ERR_BAD_PARAM;
  for(int i = 0; i < nsoft; i++)
    if(!(eps[i] > 0))
      return fail(c, AGRAV_ERR_BAD_PARAM, "softening must be positive");
  if(c->have_stype && c->n > 0 && nsoft < c->soft.nsoft)
    {
      // a shorter table must still cover the types of the particles held
      agrav_status st;
      if((st = c->d_counters.ensure(128)))
        return st;
      int bad = 0, *d_bad = c->d_counters.p + 100;
      AG_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), c->stream));
      k_check_stype<<<(unsigned)((c->n + 255) / 256), 256, 0, c->stream>>>(c->d_stype_in.p, c->n, nsoft, d_bad);
      AG_HIP(hipGetLastError());
      AG_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      AG_HIP(hipStreamSynchronize(c->stream));
      if(bad)
        return fail(c, AGRAV_ERR_BAD_PARAM, "softening type out of range of the new table");
    }
  c->par.n_soft_types = nsoft;
  for(int i = 0; i < nsoft; i++)
    c->par.softening[i] = eps[i];
  soft_from(c, nsoft, eps);
  AG_HIP(hipMemcpyAsync(c->d_soft.p, c->soft.h, sizeof(c->soft.h), hipMemcpyHostToDevice, c->stream));
  AG_HIP(hipStreamSynchronize(c->stream));
  c->have_tree = c->have_tree_force = c->have_final = false;
  return AGRAV_OK;
}

static agrav_status set_particles_common(agrav_ctx *c, int64_t n, const double *pos, const double *mass, const uint8_t *stype,
                                         hipMemcpyKind kind)
{
  if(!c || n < 0 || (n > 0 && (!pos || !mass)))
    return AGRAV_ERR_BAD_PARAM;
  if(n >= (1ll << 30))
    return fail(c, AGRAV_ERR_CAPACITY, "particle count exceeds 2^30 per context");
  // host input is checked before any state changes
  if(stype && kind == hipMemcpyHostToDevice)
    for(int64_t i = 0; i < n; i++)
      if(stype[i] >= c->soft.nsoft)
        return fail(c, AGRAV_ERR_BAD_PARAM, "softening type out of range");
  // from here on the old particle set is gone: a failure below leaves the context
  // without particles (every later call reports AGRAV_ERR_STATE until a
  // set_particles succeeds), never with a mix of old and new buffers
  c->n              = 0;
  c->n_tree         = 0;
  c->batch_nsets    = 0;
  c->n_active       = -1;
  c->active_version++;
  c->tree_partial   = false;
  c->tmap_dirty     = true;
  c->outputs_zero   = false;
  c->have_particles = false;
  c->have_stype     = false;
  c->have_ptype     = false;
  c->have_oldacc    = false;
  c->have_tree = c->have_tree_force = c->have_pm = c->have_final = false;
  agrav_status st;
  const int64_t nn = n ? n : 1;
  if((st = c->d_pos_in.ensure(3 * nn)) || (st = c->d_mass_in.ensure(nn)) || (st = c->d_oldacc.ensure(nn)))
    return st;
  hipStream_t s = c->stream;
  if(n > 0)
    {
      AG_HIP(hipMemcpyAsync(c->d_pos_in.p, pos, 3 * n * sizeof(double), kind, s));
      AG_HIP(hipMemcpyAsync(c->d_mass_in.p, mass, n * sizeof(double), kind, s));
    }
  if(stype)
    {
      if((st = c->d_stype_in.ensure(nn)))
        return st;
      AG_HIP(hipMemcpyAsync(c->d_stype_in.p, stype, n, kind, s));
      if(kind != hipMemcpyHostToDevice && n > 0)
        {
          // the same range check on the device copy
          if((st = c->d_counters.ensure(128)))
            return st;
          int bad = 0, *d_bad = c->d_counters.p + 100;
          AG_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
          k_check_stype<<<(unsigned)((n + B - 1) / B), B, 0, s>>>(c->d_stype_in.p, n, c->soft.nsoft, d_bad);
          AG_HIP(hipGetLastError());
          AG_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
          AG_HIP(hipStreamSynchronize(s));
          if(bad)
            return fail(c, AGRAV_ERR_BAD_PARAM, "softening type out of range");
        }
    }
  // total mass on the device (the PM mesh scale and the potential constant use it)
  double mt = 0;
  if(n > 0)
    {
      const int64_t nb = (n + kMassChunk - 1) / kMassChunk;
      if((st = c->d_mass_part.ensure(nb + 1)))
        return st;
      k_mass_partial<<<(unsigned)nb, B, 0, s>>>(c->d_mass_in.p, n, c->d_mass_part.p);
      k_mass_final<<<1, B, 0, s>>>(c->d_mass_part.p, nb, c->d_mass_part.p + nb);
      AG_HIP(hipGetLastError());
      AG_HIP(hipMemcpyAsync(c->stage<double>(0), c->d_mass_part.p + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    }
  AG_HIP(hipMemsetAsync(c->d_oldacc.p, 0, nn * sizeof(float), s));
  AG_HIP(hipStreamSynchronize(s));
  if(n > 0)
    mt = *c->stage<double>(0);
  c->mass_total     = mt;
  c->n              = n;
  c->have_stype     = stype != nullptr;
  c->have_particles = true;
  return AGRAV_OK;
}

agrav_status agrav_set_particles(agrav_ctx *c, int64_t n, const double *pos, const double *mass, const uint8_t *stype)
{
  return set_particles_common(c, n, pos, mass, stype, hipMemcpyHostToDevice);
}

agrav_status agrav_set_particles_device(agrav_ctx *c, int64_t n, const double *pos, const double *mass, const uint8_t *stype)
{
  return set_particles_common(c, n, pos, mass, stype, hipMemcpyDeviceToDevice);
}

agrav_status agrav_set_particle_types(agrav_ctx *c, const uint8_t *type)
{
  if(!c || !type)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "set_particle_types before set_particles");
  agrav_status st;
  if((st = c->d_ptype_in.ensure(c->n ? c->n : 1)))
    return st;
  if(c->n > 0)
    AG_HIP(hipMemcpyAsync(c->d_ptype_in.p, type, c->n, hipMemcpyHostToDevice, c->stream));
  AG_HIP(hipStreamSynchronize(c->stream));
  c->have_ptype = true;
  c->have_pm = c->have_tree_force = c->have_final = false;  // the high-resolution flags change
  return AGRAV_OK;
}

agrav_status agrav_highres_region(agrav_ctx *c, int reset, double *region)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!hr_on(c))
    return fail(c, AGRAV_ERR_STATE, "no high-resolution region: placehighresregion is 0");
  if(reset)
    {
      c->hr.valid        = false;
      c->hr.kernel_ready = false;
      return AGRAV_OK;
    }
  if(!c->hr.valid)
    return fail(c, AGRAV_ERR_STATE, "the high-resolution region is not sized yet");
  if(region)
    {
      for(int j = 0; j < 3; j++)
        {
          region[j]     = c->hr.xmin[j];
          region[3 + j] = c->hr.xmax[j];
          region[6 + j] = c->hr.corner[j];
          region[9 + j] = c->hr.upper[j];
        }
      region[12] = c->hr.meshsize;
      region[13] = c->hr.asmth;
      region[14] = c->hr.rcut;
    }
  return AGRAV_OK;
}

agrav_status agrav_host_alloc(size_t bytes, void **ptr)
{
  if(!ptr)
    return AGRAV_ERR_BAD_PARAM;
  *ptr = nullptr;
  if(bytes == 0)
    return AGRAV_OK;
  if(hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess)
    {
      *ptr = nullptr;
      set_error_msg("cannot allocate page-locked host memory");
      return AGRAV_ERR_OOM;
    }
  return AGRAV_OK;
}

agrav_status agrav_host_free(void *ptr)
{
  if(ptr && hipHostFree(ptr) != hipSuccess)
    return AGRAV_ERR_BAD_PARAM;
  return AGRAV_OK;
}

agrav_status agrav_set_old_acc(agrav_ctx *c, const float *old_acc)
{
  if(!c || !old_acc)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "set_old_acc before set_particles");
  if(c->n == 0)
    return AGRAV_OK;
  AG_HIP(hipMemcpyAsync(c->d_oldacc.p, old_acc, c->n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  AG_HIP(hipStreamSynchronize(c->stream));
  c->have_oldacc = true;
  c->have_tree   = false;  // the sorted copy is refreshed by the next build
  return AGRAV_OK;
}

agrav_status agrav_construct_forcetree(agrav_ctx *c, int *numnodes)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "construct_forcetree before set_particles");
  AG_HIP(hipEventRecord(c->ev[7], c->stream));
  agrav_status st = build_tree(c);
  if(st != AGRAV_OK)
    {
      c->have_tree = false;
      return st;
    }
  if(c->build_check_pending)
    c->build_time_pending = true;  // read with the deferred check
  else
    {
      AG_HIP(hipEventSynchronize(c->ev[1]));  // recorded by the build (already passed when it read back)
      float ms = 0;
      AG_HIP(hipEventElapsedTime(&ms, c->ev[7], c->ev[1]));
      c->tim.treebuild = ms;
    }
  c->tim.numnodes  = c->m;
  c->tmap_dirty    = true;  // the sorted positions of the active targets changed
  if(numnodes)
    *numnodes = c->m;
  c->have_tree_force = false;
  return AGRAV_OK;
}

agrav_status agrav_set_active(agrav_ctx *c, const int32_t *active, int64_t n_active)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "set_active before set_particles");
  c->active_version++;
  if(!active || n_active < 0)
    {
      c->n_active   = -1;
      c->tmap_dirty = true;
      return AGRAV_OK;
    }
  if(c->batch_nsets > 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "active lists select targets of one particle set, not of a batch");
  if(n_active > c->n)
    return fail(c, AGRAV_ERR_BAD_PARAM, "more active particles than particles");
  agrav_status st;
  if((st = c->d_active.ensure(n_active ? n_active : 1)) != AGRAV_OK)
    return st;
  if(n_active > 0)
    {
      AG_HIP(hipMemcpyAsync(c->d_active.p, active, n_active * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
      AG_HIP(hipStreamSynchronize(c->stream));
    }
  c->n_active        = n_active;
  c->tmap_dirty      = true;
  c->have_tree_force = c->have_final = false;
  return AGRAV_OK;
}

agrav_status agrav_set_partial_tree(agrav_ctx *c, int partial, int set_oldacc)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  c->partial_tree = partial != 0;
  c->keep_oldacc  = set_oldacc == 0;
  return AGRAV_OK;
}

agrav_status agrav_gravity_timesteps(agrav_ctx *c, const agrav_timestep_params *tp, double *dt, double *dt_displacement,
                                     int64_t *n_below_min)
{
  if(!c || !tp || !(tp->err_tol_int_accuracy > 0) || !(tp->max_size_timestep > 0) || tp->min_size_timestep < 0)
    return AGRAV_ERR_BAD_PARAM;
  if(tp->comoving && !(tp->time_a > 0))
    return fail(c, AGRAV_ERR_BAD_PARAM, "comoving timesteps need a positive scale factor");
  if(!c->have_final)
    return fail(c, AGRAV_ERR_STATE, "gravity_timesteps before the accelerations of a step");
  if(c->batch_nsets > 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "timesteps are for one particle set, not for a batch");
  const int64_t n = c->n;
  hipStream_t s   = c->stream;
  TsArgs a;
  a.eta2a = 2.0 * tp->err_tol_int_accuracy * (tp->comoving ? tp->time_a : 1.0);
  a.a2inv = tp->comoving ? 1.0 / (tp->time_a * tp->time_a) : 1.0;
  a.hub   = tp->comoving ? tp->hubble_a : 1.0;
  a.maxdt = tp->max_size_timestep;
  a.mindt = tp->min_size_timestep;
  agrav_status st;
  if((st = c->d_counters.ensure(128)) || (st = c->d_temp.ensure(n ? 8 * n : 8)))
    return st;
  unsigned long long *d_min = reinterpret_cast<unsigned long long *>(c->d_counters.p + 104);  // 8-byte aligned
  int *d_below              = c->d_counters.p + 106;
  const bool pm             = c->par.pmgrid > 0 && c->have_pm;
  double dtdisp             = tp->max_size_timestep;
  unsigned long long init   = ~0ull;
  int zero                  = 0;
  AG_HIP(hipMemcpyAsync(d_min, &init, sizeof(init), hipMemcpyHostToDevice, s));
  AG_HIP(hipMemcpyAsync(d_below, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
  const unsigned g = (unsigned)((n + B - 1) / B);
  if(pm && n > 0)
    {
      // the PM accelerations are in sorted order when the tree of the step was current
      const uint8_t *stp = c->pm_sorted ? (c->soft.nsoft > 1 ? c->d_ps.p : nullptr)
                                        : (c->have_stype ? c->d_stype_in.p : nullptr);
      k_pm_dtmin<<<g, B, 0, s>>>(n, c->d_acc_pm.p, stp, c->d_soft.p, a, d_min);
      AG_HIP(hipGetLastError());
      unsigned long long bits = 0;
      AG_HIP(hipMemcpyAsync(&bits, d_min, sizeof(bits), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      double m = 0;
      memcpy(&m, &bits, sizeof(m));
      dtdisp = 2.0 * m;  // one timebin higher (timestep.c:558)
      if(dtdisp > tp->max_size_timestep)
        dtdisp = tp->max_size_timestep;
      if(dtdisp < tp->min_size_timestep)
        dtdisp = tp->min_size_timestep;
    }
  double *d_dt = reinterpret_cast<double *>(c->d_temp.p);
  if(n > 0)
    {
      k_grav_dt<<<g, B, 0, s>>>(n, c->d_acc.p, c->d_mass_in.p, c->have_stype ? c->d_stype_in.p : nullptr, c->d_soft.p, a, pm ? 1 : 0,
                                dtdisp, d_dt, d_below);
      AG_HIP(hipGetLastError());
      if(dt)
        AG_HIP(hipMemcpyAsync(dt, d_dt, n * sizeof(double), hipMemcpyDeviceToHost, s));
    }
  int below = 0;
  AG_HIP(hipMemcpyAsync(&below, d_below, sizeof(below), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(dt_displacement)
    *dt_displacement = dtdisp;
  if(n_below_min)
    *n_below_min = below;
  return AGRAV_OK;
}

agrav_status agrav_gravity_tree(agrav_ctx *c, int use_bh)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  return walk_tree(c, use_bh);
}

agrav_status agrav_long_range_force(agrav_ctx *c)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "long_range_force before set_particles");
  if(c->par.pmgrid <= 0)
    return fail(c, AGRAV_ERR_STATE, "long_range_force without PMGRID");
  return pm_force(c);
}

agrav_status agrav_gravity_force_finalize(agrav_ctx *c)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  return finalize(c);
}

agrav_status agrav_compute_grav_accelerations(agrav_ctx *c, int first_step)
{
  // accel.c:58-80 and gravity() accel.c:94-192
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "compute_grav_accelerations before set_particles");
  // the PM does not depend on the tree; running it after the build's gather lets it use
  // the Hilbert-ordered particles (the reference calls long_range_force first).  A full
  // tree's periodic PM runs on the PM stream beside the octree kernels; the main stream
  // joins it before the walk, whose epilogue adds the PM force.
  const bool partial_build = c->partial_tree && c->n_active >= 0 && c->batch_nsets == 0;
  const bool need_pm       = c->par.pmgrid > 0 && !(partial_build && c->have_pm);
  const bool plain_step    = c->n_active < 0 && c->batch_nsets == 0 && !hr_on(c) && c->n > 0;
  c->pm_fork               = need_pm && c->pm_stream && plain_step;
  // a full step of every particle checks its build after the walk's read-back (the walk
  // tests the build's flags on the device before it starts)
  c->defer_build_check     = plain_step;
  agrav_status st          = agrav_construct_forcetree(c, nullptr);
  c->pm_fork               = false;
  c->defer_build_check     = false;
  if(c->pm_forked)
    {
      AG_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
      c->pm_forked = false;
    }
  else if(st == AGRAV_OK && need_pm)
    st = pm_force(c);
  // a partial-tree step keeps the long-range force of the last PM while the particles
  // are unchanged, as GravPM between PM steps (do_gravity_hydro.c:105-123)
  else if(st == AGRAV_OK && c->par.pmgrid > 0 && (st = pm_time_settle(c)) == AGRAV_OK)
    c->tim.pm_grav = 0;  // no PM in this call
  // the walks of a full step write the final outputs from their epilogue
  c->fuse_final = true;
  const bool seed = c->par.type_of_opening_criterion == 1 && first_step && c->par.err_tol_theta > 0;
  if(st == AGRAV_OK && seed)
    {
      // Barnes-Hut pass that seeds OldAcc (accel.c:62-72); the tree of the
      // second gravity() call would be identical, so it is reused.
      if((st = walk_tree(c, 1)) == AGRAV_OK)
        st = finalize(c);
    }
  if(st == AGRAV_OK && (st = walk_tree(c, c->par.type_of_opening_criterion == 0)) == AGRAV_OK)
    st = finalize(c);
  c->fuse_final  = false;
  c->final_fused = false;
  if(c->pre_ngroups >= 0)
    {
      // an order sorted for a walk that did not run (a failed build): later work on the
      // main stream waits for its sort
      (void)hipStreamWaitEvent(c->stream, c->ev_order, 0);
      c->pre_ngroups = -1;
    }
  c->isum_clear = false;
  if(c->build_check_pending)
    {
      // normally read with the walk's counters; otherwise here
      AG_HIP(hipStreamSynchronize(c->stream));
      const agrav_status sb = build_check(c);
      if(st == AGRAV_OK)
        st = sb;
    }
  return st;
}

agrav_status agrav_get_accelerations(agrav_ctx *c, double *acc, double *pot, int32_t *nint, float *old_acc)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_final)
    return fail(c, AGRAV_ERR_STATE, "no results: run gravity_force_finalize first");
  const int64_t n = c->n;
  hipStream_t s   = c->stream;
  if(n > 0)
    {
      if(acc)
        AG_HIP(hipMemcpyAsync(acc, c->d_acc.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
      if(pot)
        AG_HIP(hipMemcpyAsync(pot, c->d_pot.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
      if(nint)
        AG_HIP(hipMemcpyAsync(nint, c->d_nint_out.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
      if(old_acc)
        AG_HIP(hipMemcpyAsync(old_acc, c->d_oldacc.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    }
  AG_HIP(hipStreamSynchronize(s));
  return AGRAV_OK;
}

agrav_status agrav_get_accelerations_device(agrav_ctx *c, const double **acc, const double **pot)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_final)
    return fail(c, AGRAV_ERR_STATE, "no results");
  if(acc)
    *acc = c->d_acc.p;
  if(pot)
    *pot = c->d_pot.p;
  return AGRAV_OK;
}

agrav_status agrav_get_components(agrav_ctx *c, double *acc_tree, double *pot_tree, double *acc_pm, double *pot_pm)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  const int64_t n = c->n;
  hipStream_t s   = c->stream;
  if(n == 0)
    return AGRAV_OK;
  if((acc_tree || pot_tree) && !c->have_tree_force)
    return fail(c, AGRAV_ERR_STATE, "no tree force");
  if((acc_pm || pot_pm) && !c->have_pm)
    return fail(c, AGRAV_ERR_STATE, "no PM force");
  if((acc_tree || pot_tree) && c->n_tree > 0)
    {
      // a partial tree holds, and writes, only its own particles
      const int64_t nt = c->n_tree;
      std::vector<double> at(3 * nt), pt(nt);
      std::vector<int> perm(nt);
      AG_HIP(hipMemcpyAsync(at.data(), c->d_at.p, 3 * nt * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(pt.data(), c->d_pt.p, nt * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(perm.data(), c->d_idx2.p, nt * sizeof(int), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      for(int64_t k = 0; k < nt; k++)
        {
          int i = perm[k];
          if(acc_tree)
            for(int d = 0; d < 3; d++)
              acc_tree[3 * (int64_t)i + d] = at[3 * k + d];
          if(pot_tree)
            pot_tree[i] = pt[k];
        }
    }
  if(acc_pm || pot_pm)
    {
      std::vector<double> ap(3 * n), pp(n);
      std::vector<int> perm;
      AG_HIP(hipMemcpyAsync(ap.data(), c->d_acc_pm.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(pp.data(), c->d_pot_pm.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
      if(c->pm_sorted)
        {
          perm.resize(n);
          AG_HIP(hipMemcpyAsync(perm.data(), c->d_idx2.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
        }
      AG_HIP(hipStreamSynchronize(s));
      for(int64_t k = 0; k < n; k++)
        {
          int64_t i = c->pm_sorted ? perm[k] : k;
          if(acc_pm)
            for(int d = 0; d < 3; d++)
              acc_pm[3 * i + d] = ap[3 * k + d];
          if(pot_pm)
            pot_pm[i] = pp[k];
        }
    }
  AG_HIP(hipStreamSynchronize(s));
  return AGRAV_OK;
}

agrav_status agrav_gravity_forcetest(agrav_ctx *c, int ns, const int32_t *sel, double *acc, double *acc_short, double *pot,
                                     double *pot_short)
{
  if(!c || ns < 0 || (ns > 0 && !sel))
    return AGRAV_ERR_BAD_PARAM;
  return forcetest(c, ns, sel, acc, acc_short, pot, pot_short);
}

agrav_status agrav_gravity_direct(agrav_ctx *c, double *acc, double *pot)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  return direct_all(c, acc, pot);
}

agrav_status agrav_gravity_direct_noalloc(agrav_ctx *c, int64_t n, const double *d_pos, const double *d_mass,
                                          const uint8_t *d_softtype, double *d_acc, double *d_pot, void *stream)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  return direct_noalloc(c, n, d_pos, d_mass, d_softtype, d_acc, d_pot, stream ? (hipStream_t)stream : c->stream);
}

agrav_status agrav_get_timings(agrav_ctx *c, agrav_timings *t)
{
  if(!c || !t)
    return AGRAV_ERR_BAD_PARAM;
  agrav_status st = pm_time_settle(c);
  if(st != AGRAV_OK)
    return st;
  *t              = c->tim;
  t->device_bytes = (long long)device_bytes(c);
  return AGRAV_OK;
}

agrav_status agrav_get_tree(agrav_ctx *c, int *numnodes, double *center, double *len, double *com, double *mass, int32_t *level,
                            int32_t *npart)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_tree)
    return fail(c, AGRAV_ERR_STATE, "no tree");
  const int M   = c->m;
  hipStream_t s = c->stream;
  if(numnodes)
    *numnodes = M;
  if(M == 0)
    return AGRAV_OK;
  std::vector<double> a(M), b(M), d(M);
  if(center)
    {
      AG_HIP(hipMemcpyAsync(a.data(), c->d_cx.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(b.data(), c->d_cy.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(d.data(), c->d_cz.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      for(int j = 0; j < M; j++)
        {
          center[3 * j]     = a[j];
          center[3 * j + 1] = b[j];
          center[3 * j + 2] = d[j];
        }
    }
  if(com)
    {
      AG_HIP(hipMemcpyAsync(a.data(), c->d_sx.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(b.data(), c->d_sy.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(d.data(), c->d_sz.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      for(int j = 0; j < M; j++)
        {
          com[3 * j]     = a[j];
          com[3 * j + 1] = b[j];
          com[3 * j + 2] = d[j];
        }
    }
  if(len)
    AG_HIP(hipMemcpyAsync(len, c->d_len.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
  if(mass)
    AG_HIP(hipMemcpyAsync(mass, c->d_nm.p, M * sizeof(double), hipMemcpyDeviceToHost, s));
  if(level)
    AG_HIP(hipMemcpyAsync(level, c->d_nlevel.p, M * sizeof(int), hipMemcpyDeviceToHost, s));
  if(npart)
    {
      std::vector<int> f(M), l(M);
      AG_HIP(hipMemcpyAsync(f.data(), c->d_nfirst.p, M * sizeof(int), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(l.data(), c->d_nlast.p, M * sizeof(int), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      for(int j = 0; j < M; j++)
        npart[j] = l[j] - f[j] + 1;
    }
  AG_HIP(hipStreamSynchronize(s));
  return AGRAV_OK;
}

agrav_status agrav_get_order(agrav_ctx *c, int32_t *perm, uint64_t *keys)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(!c->have_tree)
    return fail(c, AGRAV_ERR_STATE, "no tree");
  const int64_t n = c->n_tree;
  if(n == 0)
    return AGRAV_OK;
  if(perm)
    AG_HIP(hipMemcpyAsync(perm, c->d_idx2.p, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if(keys)
    AG_HIP(hipMemcpyAsync(keys, c->d_key2.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  AG_HIP(hipStreamSynchronize(c->stream));
  return AGRAV_OK;
}

agrav_status agrav_get_tables(agrav_ctx *c, float *tab_force, float *tab_pot, double *ewald_tree, double *ewald_forcetest)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(tab_force)
    memcpy(tab_force, c->tab_force, sizeof(c->tab_force));
  if(tab_pot)
    memcpy(tab_pot, c->tab_pot, sizeof(c->tab_pot));
  agrav_status st;
  if(ewald_tree)
    {
      if(!c->ewald_ready && (st = ewald_tree_table(c)) != AGRAV_OK)
        return st;
      AG_HIP(hipMemcpyAsync(ewald_tree, c->d_ewald.p, 4 * 65 * 65 * 65 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
  if(ewald_forcetest)
    {
      if(!c->ewald_ft_ready && (st = ewald_forcetest_table(c)) != AGRAV_OK)
        return st;
      AG_HIP(hipMemcpyAsync(ewald_forcetest, c->d_ewald_ft.p, (size_t)4 * 129 * 129 * 129 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    }
  AG_HIP(hipStreamSynchronize(c->stream));
  return AGRAV_OK;
}

uint64_t agrav_hilbert_key21(uint32_t x, uint32_t y, uint32_t z)
{
  static const uint8_t tab[192] = AG_HILBERT_TABLE;
  uint64_t top = 0, low = 0;
  hilbert_digits(x, y, z, AG_KEY_LEVELS, tab, top, low);
  return low;
}

}  // extern "C"
