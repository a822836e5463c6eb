// agrav_batch.hip -- many independent isolated particle sets solved together.
//
// The reference computes the potential of each candidate group on its own, one
// group after another on one core: a local tree build (subfind_loctree_treebuild,
// subfind_loctree.c:192) and a Barnes-Hut walk per member
// (subfind_loctree_treeevaluate_potential, subfind_loctree.c:564-730), driven by
// subfind_unbind (subfind_serial.c:429-560).  Here every set of a batch goes
// through ONE tree build -- keys carry the set index above the Hilbert digits, so
// each set becomes its own subtree with its own domain cube -- and ONE persistent
// walk launch, whose wavefronts take target groups from a queue that holds the
// groups of the largest sets first.  The group size and the leaf bucket are
// chosen per set-size class by a timing autotuner on a sample of the batch.
#include "agrav_ctx.h"

#include <algorithm>
#include <numeric>

namespace agrav
{

static constexpr int B = 256;

// set-size classes of the tuner: n <= 128, <= 2048, <= 32768, larger
static const int64_t kClassMax[AGRAV_BATCH_CLASSES] = {128, 2048, 32768, -1};

static int size_class(int64_t n)
{
  for(int k = 0; k < AGRAV_BATCH_CLASSES - 1; k++)
    if(n <= kClassMax[k])
      return k;
  return AGRAV_BATCH_CLASSES - 1;
}

// caller-order positions: FP32 offsets from each set's centre -> FP64 absolute
// positions, and the set index of every particle
__global__ void k_batch_ingest(int64_t n, int nsets, const int64_t *__restrict__ off, const float *__restrict__ rel,
                               const double *__restrict__ centre, double *pos, int *pset)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  int lo = 0, hi = nsets - 1;  // last set with off[s] <= i (empty sets share offsets)
  while(lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if(off[mid] <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
  pset[i] = lo;
  for(int d = 0; d < 3; d++)
    pos[3 * i + d] = (double)rel[3 * i + d] + (centre ? centre[3 * lo + d] : 0.0);
}

// walk groups of the given sets: each set is cut into groups of its class's size,
// largest sets first
static void make_groups(const agrav_ctx *c, const std::vector<int> &sets, const int *gs, const int *lb, std::vector<int4> &out)
{
  std::vector<int> order(sets);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return c->h_soff[a + 1] - c->h_soff[a] > c->h_soff[b + 1] - c->h_soff[b];
  });
  out.clear();
  for(int s : order)
    {
      const int64_t f = c->h_soff[s], n = c->h_soff[s + 1] - f;
      const int k = size_class(n), g = gs[k];
      for(int64_t p = 0; p < n; p += g)
        out.push_back(make_int4((int)(f + p), (int)std::min<int64_t>(g, n - p), s, lb[k]));
    }
}

struct ClassTuning
{
  int gs[AGRAV_BATCH_CLASSES], lb[AGRAV_BATCH_CLASSES];
};

// the walk groups of one set per workgroup: set q of the size order starts at group
// sg[q].y and is cut into groups of its class's size
__global__ void k_expand_groups(const int2 *__restrict__ sg, const int64_t *__restrict__ off, ClassTuning ct, int4 *groups)
{
  const int s = sg[blockIdx.x].x, base = sg[blockIdx.x].y;
  const int64_t f = off[s], n = off[s + 1] - f;
  int k = AGRAV_BATCH_CLASSES - 1;
  if(n <= 128)
    k = 0;
  else if(n <= 2048)
    k = 1;
  else if(n <= 32768)
    k = 2;
  const int g = ct.gs[k], ng = (int)((n + g - 1) / g);
  for(int j = threadIdx.x; j < ng; j += blockDim.x)
    {
      const int64_t p = (int64_t)j * g;
      groups[base + j] = make_int4((int)(f + p), (int)(n - p < g ? n - p : g), s, ct.lb[k]);
    }
}

agrav_status batch_groups(agrav_ctx *c)
{
  if(!c->batch_groups_dirty)
    return AGRAV_OK;
  // the same table as make_groups, written on the device: the host only orders the sets
  // (largest first) and places each set's first group
  static_assert(sizeof(kClassMax) / sizeof(kClassMax[0]) == 4, "the expansion kernel knows four classes");
  const int nsets = c->batch_nsets;
  std::vector<int> order(nsets);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return c->h_soff[a + 1] - c->h_soff[a] > c->h_soff[b + 1] - c->h_soff[b];
  });
  // kept in the context: the upload below is asynchronous
  std::vector<int2> &sg = c->h_gsets;
  sg.clear();
  sg.reserve(nsets);
  int64_t total = 0;
  for(int s : order)
    {
      const int64_t n = c->h_soff[s + 1] - c->h_soff[s];
      if(n == 0)
        continue;
      sg.push_back(make_int2(s, (int)total));
      const int g = c->batch_gs[size_class(n)];
      total += (n + g - 1) / g;
    }
  agrav_status st;
  if((st = c->d_groups.ensure(total)) != AGRAV_OK || (st = c->d_gsets.ensure(sg.size())) != AGRAV_OK)
    return st;
  if(!sg.empty())
    {
      ClassTuning ct;
      for(int k = 0; k < AGRAV_BATCH_CLASSES; k++)
        {
          ct.gs[k] = c->batch_gs[k];
          ct.lb[k] = c->batch_lb[k];
        }
      AG_HIP(hipMemcpyAsync(c->d_gsets.p, sg.data(), sg.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
      k_expand_groups<<<(unsigned)sg.size(), 256, 0, c->stream>>>(c->d_gsets.p, c->d_soff.p, ct, c->d_groups.p);
      AG_HIP(hipGetLastError());
    }
  c->batch_ngroups      = (int)total;
  c->batch_groups_dirty = false;
  return AGRAV_OK;
}

static void fill_tuning(const agrav_ctx *c, agrav_batch_tuning *t)
{
  for(int k = 0; k < AGRAV_BATCH_CLASSES; k++)
    {
      t->group_size[k]  = c->batch_gs[k];
      t->leaf_bucket[k] = c->batch_lb[k];
      t->class_max_n[k] = kClassMax[k];
    }
  t->tuned = c->batch_tuned;
}

}  // namespace agrav

using namespace agrav;

extern "C" {

agrav_status agrav_batch_set_particles(agrav_ctx *c, int nsets, const int64_t *offsets, const float *pos_rel, const double *centres,
                                       const double *mass, const double *softening)
{
  if(!c || nsets < 1 || !offsets || !softening)
    return AGRAV_ERR_BAD_PARAM;
  if(c->par.periodic || c->par.pmgrid > 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "a batch holds isolated sets: periodic = 0 and pmgrid = 0 required");
  if(nsets > (1 << 21))
    return fail(c, AGRAV_ERR_CAPACITY, "more than 2^21 sets in one batch");
  if(offsets[0] != 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "batch offsets must start at 0");
  for(int s = 0; s < nsets; s++)
    {
      if(offsets[s + 1] < offsets[s])
        return fail(c, AGRAV_ERR_BAD_PARAM, "batch offsets must not decrease");
      if(!(softening[s] > 0))
        return fail(c, AGRAV_ERR_BAD_PARAM, "batch softening must be positive");
    }
  const int64_t n = offsets[nsets];
  if(n > 0 && (!pos_rel || !mass))
    return AGRAV_ERR_BAD_PARAM;
  if(n >= (1ll << 30))
    return fail(c, AGRAV_ERR_CAPACITY, "particle count exceeds 2^30 per context");
  const int64_t nn = n ? n : 1;
  // the old particle set is gone from here on: a failure below leaves the context
  // without particles, never with a mix of old and new buffers
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
  c->have_oldacc    = false;
  c->gcost_groups   = -1;
  c->have_tree = c->have_tree_force = c->have_pm = c->have_final = false;
  agrav_status st;
  if((st = c->d_pos_in.ensure(3 * nn)) || (st = c->d_mass_in.ensure(nn)) || (st = c->d_oldacc.ensure(nn)) ||
     (st = c->d_pset_in.ensure(nn)) || (st = c->d_soff.ensure(nsets + 1)) || (st = c->d_shd.ensure(nsets)) ||
     (st = c->d_at.ensure(3 * nn)) || (st = c->d_sdom.ensure(nsets)))
    return st;
  hipStream_t s = c->stream;
  std::vector<double> h(nsets);
  for(int k = 0; k < nsets; k++)
    h[k] = 2.8 * softening[k];  // set_softenings, grav_softening.c:90-95
  c->h_soff.assign(offsets, offsets + nsets + 1);
  AG_HIP(hipMemcpyAsync(c->d_soff.p, offsets, (nsets + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  AG_HIP(hipMemcpyAsync(c->d_shd.p, h.data(), nsets * sizeof(double), hipMemcpyHostToDevice, s));
  double *d_cen = nullptr;
  if(n > 0)
    {
      // the FP32 offsets are staged in the (not yet used) walk output buffer
      float *d_rel = reinterpret_cast<float *>(c->d_at.p);
      AG_HIP(hipMemcpyAsync(d_rel, pos_rel, 3 * n * sizeof(float), hipMemcpyHostToDevice, s));
      AG_HIP(hipMemcpyAsync(c->d_mass_in.p, mass, n * sizeof(double), hipMemcpyHostToDevice, s));
      if(centres)
        {
          // centres go through the per-set domain buffer, which the build overwrites
          static_assert(sizeof(Domain) >= 3 * sizeof(double), "domain record holds a centre");
          d_cen = reinterpret_cast<double *>(c->d_sdom.p);
          AG_HIP(hipMemcpyAsync(d_cen, centres, 3 * nsets * sizeof(double), hipMemcpyHostToDevice, s));
        }
      k_batch_ingest<<<(unsigned)((n + B - 1) / B), B, 0, s>>>(n, nsets, c->d_soff.p, d_rel, d_cen, c->d_pos_in.p, c->d_pset_in.p);
      AG_HIP(hipGetLastError());
    }
  double mt = 0;
  for(int64_t i = 0; i < n; i++)
    mt += mass[i];
  AG_HIP(hipMemsetAsync(c->d_oldacc.p, 0, nn * sizeof(float), s));
  AG_HIP(hipStreamSynchronize(s));
  // one softening per set; the context's own softening table is not used by a batch
  c->have_stype         = false;
  c->batch_nsets        = nsets;
  c->batch_groups_dirty = true;
  c->n_active           = -1;
  c->tmap_dirty         = true;
  c->mass_total         = mt;
  c->n                  = n;
  c->have_particles     = true;
  c->have_oldacc        = false;
  c->gcost_groups      = -1;
  c->have_tree = c->have_tree_force = c->have_pm = c->have_final = false;
  return AGRAV_OK;
}

agrav_status agrav_batch_get_tuning(agrav_ctx *c, agrav_batch_tuning *t)
{
  if(!c || !t)
    return AGRAV_ERR_BAD_PARAM;
  *t = agrav_batch_tuning{};
  fill_tuning(c, t);
  return AGRAV_OK;
}

agrav_status agrav_batch_set_tuning(agrav_ctx *c, const agrav_batch_tuning *t)
{
  if(!c || !t)
    return AGRAV_ERR_BAD_PARAM;
  for(int k = 0; k < AGRAV_BATCH_CLASSES; k++)
    if(t->group_size[k] < 1 || t->group_size[k] > 64 || t->leaf_bucket[k] < 0 || t->leaf_bucket[k] > 64)
      return fail(c, AGRAV_ERR_BAD_PARAM, "group size must be in [1, 64] and leaf bucket in [0, 64]");
  for(int k = 0; k < AGRAV_BATCH_CLASSES; k++)
    {
      c->batch_gs[k] = t->group_size[k];
      c->batch_lb[k] = t->leaf_bucket[k];
    }
  c->batch_tuned        = t->tuned;
  c->batch_groups_dirty = true;
  return AGRAV_OK;
}

agrav_status agrav_batch_autotune(agrav_ctx *c, int64_t sample_particles, agrav_batch_tuning *out)
{
  if(!c)
    return AGRAV_ERR_BAD_PARAM;
  if(c->batch_nsets < 1 || !c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "batch_autotune before batch_set_particles");
  if(c->par.walk == AGRAV_WALK_PLAIN)
    return fail(c, AGRAV_ERR_STATE, "batch_autotune tunes the group walk");
  if(sample_particles <= 0)
    sample_particles = 4 << 20;
  agrav_status st;
  if(!c->have_tree && (st = agrav_construct_forcetree(c, nullptr)) != AGRAV_OK)
    return st;
  const bool use_bh = c->par.type_of_opening_criterion == 0;
  if(!use_bh && !c->have_final)
    {
      // the relative criterion needs OldAcc: one Barnes-Hut pass with the current tuning
      if((st = walk_tree(c, 1)) != AGRAV_OK || (st = finalize(c)) != AGRAV_OK)
        return st;
    }
  static const int cand_gs[] = {16, 32, 64};
  static const int cand_lb[] = {0, 4, 8, 16};
  agrav_batch_tuning res{};
  fill_tuning(c, &res);
  std::vector<int4> g;
  for(int k = 0; k < AGRAV_BATCH_CLASSES; k++)
    {
      // deterministic sample of the class: every stride-th set, about sample_particles in all
      std::vector<int> cls;
      int64_t tot = 0;
      for(int s = 0; s < c->batch_nsets; s++)
        {
          const int64_t n = c->h_soff[s + 1] - c->h_soff[s];
          if(n > 0 && size_class(n) == k)
            {
              cls.push_back(s);
              tot += n;
            }
        }
      if(cls.empty())
        continue;
      const int64_t stride = std::max<int64_t>(1, (tot + sample_particles - 1) / sample_particles);
      std::vector<int> sample;
      int64_t ns = 0;
      for(size_t q = 0; q < cls.size(); q += (size_t)stride)
        {
          sample.push_back(cls[q]);
          ns += c->h_soff[cls[q] + 1] - c->h_soff[cls[q]];
        }
      double best = 1e300;
      for(int gsv : cand_gs)
        for(int lbv : cand_lb)
          {
            int gs[AGRAV_BATCH_CLASSES], lb[AGRAV_BATCH_CLASSES];
            for(int q = 0; q < AGRAV_BATCH_CLASSES; q++)
              {
                gs[q] = gsv;
                lb[q] = lbv;
              }
            make_groups(c, sample, gs, lb, g);
            if((st = c->d_groups.ensure(g.size())) != AGRAV_OK)
              return st;
            AG_HIP(hipMemcpyAsync(c->d_groups.p, g.data(), g.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
            double t_best = 1e300;
            for(int rep = 0; rep < 2; rep++)
              {
                float ms = 0;
                if((st = walk_batch_groups(c, use_bh, c->d_groups.p, (int)g.size(), &ms)) != AGRAV_OK)
                  return st;
                t_best = std::min(t_best, (double)ms);
              }
            if(t_best < best)
              {
                best               = t_best;
                res.group_size[k]  = gsv;
                res.leaf_bucket[k] = lbv;
              }
          }
      res.best_ms[k]          = best;
      res.sample_particles[k] = ns;
    }
  for(int k = 0; k < AGRAV_BATCH_CLASSES; k++)
    {
      c->batch_gs[k] = res.group_size[k];
      c->batch_lb[k] = res.leaf_bucket[k];
    }
  c->batch_tuned        = 1;
  res.tuned             = 1;
  c->batch_groups_dirty = true;
  // the tuning walks overwrote the walk outputs of the sampled targets
  c->have_tree_force = false;
  if(out)
    *out = res;
  return AGRAV_OK;
}

}  // extern "C"
