// agrav_unbind.hip -- SUBFIND's unbinding loop for many candidate groups at once.
//
// The reference unbinds one candidate after another on one core
// (subfind_unbind, subfind_serial.c:429-637): a local tree and a Barnes-Hut
// potential per member (subfind_loctree.c:192-730), the bulk velocity and the
// centre of mass, binding energies, a sort of the energies, and removal of at
// most a quarter of the members per pass, until a full pass removes nobody.
// Here every group of the call advances one pass per iteration together:
//   * groups whose next pass recomputes the potential are compacted into one
//     batch and solved by ONE tree build and ONE persistent walk (the batch
//     machinery of agrav_batch.hip), with each group's root cube fixed for all
//     passes as the reference keeps its local-tree cube;
//   * one workgroup per group then does the rest of the pass in place: the
//     minimum-potential member, FP64 sums for the centre of mass and the bulk
//     velocity, the binding energies, the two order statistics the reference
//     takes from its sorted energy list (an exact 8-bit radix select on the
//     energies' order-preserving keys instead of a sort), and the removal.
// The host keeps each group's phase and reads back three counts per group per
// iteration.
#include "agrav_ctx.h"

#include <algorithm>
#include <numeric>

namespace agrav
{

static constexpr int B  = 256;
static constexpr int UB = 1024;  // threads of the per-group pass kernel
static constexpr int UW = UB / 64;

// group of every particle (offsets are sorted; empty groups share offsets)
__global__ void k_unb_gid(int64_t n, int ng, const int64_t *__restrict__ off, int *gid, uint8_t *alive)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  int lo = 0, hi = ng - 1;
  while(lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if(off[mid] <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
  gid[i]   = lo;
  alive[i] = 1;
}

// subfind_loctree_findExtent (subfind_loctree.c:101-160): 1.001 x the largest
// extent of the bounding box, centred on it; one workgroup per group
__global__ __launch_bounds__(B) void k_unb_cubes(const double *__restrict__ pos, const int64_t *__restrict__ off, double *cube)
{
  __shared__ double s_lo[3][B / 64], s_hi[3][B / 64];
  const int g = blockIdx.x;
  double lo[3] = {1.0e306, 1.0e306, 1.0e306}, hi[3] = {-1.0e306, -1.0e306, -1.0e306};
  for(int64_t i = off[g] + threadIdx.x; i < off[g + 1]; i += B)
    for(int d = 0; d < 3; d++)
      {
        const double x = pos[3 * i + d];
        lo[d]          = x < lo[d] ? x : lo[d];
        hi[d]          = x > hi[d] ? x : hi[d];
      }
  for(int d = 0; d < 3; d++)
    {
      for(int o = 32; o > 0; o >>= 1)
        {
          const double a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
          lo[d]          = a < lo[d] ? a : lo[d];
          hi[d]          = b > hi[d] ? b : hi[d];
        }
      if((threadIdx.x & 63) == 0)
        {
          s_lo[d][threadIdx.x >> 6] = lo[d];
          s_hi[d][threadIdx.x >> 6] = hi[d];
        }
    }
  __syncthreads();
  if(threadIdx.x == 0)
    {
      double a[3], b[3], len = 0;
      for(int d = 0; d < 3; d++)
        {
          a[d] = s_lo[d][0];
          b[d] = s_hi[d][0];
          for(int w = 1; w < B / 64; w++)
            {
              a[d] = s_lo[d][w] < a[d] ? s_lo[d][w] : a[d];
              b[d] = s_hi[d][w] > b[d] ? s_hi[d][w] : b[d];
            }
          if(b[d] - a[d] > len)
            len = b[d] - a[d];
        }
      len *= 1.001;
      if(!(len > 0))
        len = 1.0;  // coincident members: any cube serves
      const bool empty = off[g + 1] == off[g];
      for(int d = 0; d < 3; d++)
        cube[4 * g + d] = empty ? 0.0 : 0.5 * (a[d] + b[d]);
      cube[4 * g + 3] = len;
    }
}

// members that take part in this iteration's potential solve
__global__ void k_unb_flag(int64_t n, const uint8_t *__restrict__ alive, const int *__restrict__ gid,
                           const int *__restrict__ gphase, int *flag)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i < n)
    flag[i] = alive[i] && gphase[gid[i]] == 0;
}

// offsets of the compacted batch: a group's members start where its first particle lands
__global__ void k_unb_soff(int ng, int64_t n, const int64_t *__restrict__ off, const int *__restrict__ dst,
                           const int *__restrict__ total, int64_t *soff)
{
  const int g = blockIdx.x * B + threadIdx.x;
  if(g > ng)
    return;
  soff[g] = (g < ng && off[g] < n) ? (int64_t)dst[off[g]] : (int64_t)*total;
}

__global__ void k_unb_compact(int64_t n, const int *__restrict__ flag, const int *__restrict__ dst, const int *__restrict__ gid,
                              const double *__restrict__ pos, const double *__restrict__ mass, double *pos_in, double *mass_in,
                              int *pset_in, int *cidx)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n || !flag[i])
    return;
  const int64_t d = dst[i];
  for(int j = 0; j < 3; j++)
    pos_in[3 * d + j] = pos[3 * i + j];
  mass_in[d] = mass[i];
  pset_in[d] = gid[i];
  cidx[d]    = (int)i;
}

// PS[p].Potential = G / cf_atime * pot (subfind_serial.c:466-468), from the walk's
// sorted-order potential back to the member
__global__ void k_unb_scatter(int64_t m, const int *__restrict__ perm, const int *__restrict__ cidx, const double *__restrict__ pt,
                              double g_a, double *pot)
{
  const int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k < m)
    pot[cidx[perm[k]]] = g_a * pt[k];
}

// order-preserving 64-bit key of a double (larger value, larger key)
__device__ inline uint64_t okey(double x)
{
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

__device__ inline double okey_inv(uint64_t k)
{
  const uint64_t b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __longlong_as_double((long long)b);
}

struct UnbArgs
{
  const int64_t *off;
  const int *glist, *gphase;
  const double *pos, *vel, *mass, *uth, *epsr;
  const uint8_t *type;
  double *pot, *E, *wbl, *ppos;
  uint8_t *alive;
  int *res;                        // per group {unbound, len left, len_non_gas}
  double g_a, vel_to_phys, atime, hubble;
};

// fixed-order block sum of NV doubles (wave butterflies, then the waves in order)
template <int NV> __device__ inline void block_sum(double (&v)[NV], double (*s_red)[8])
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for(int q = 0; q < NV; q++)
    for(int o = 32; o > 0; o >>= 1)
      v[q] += __shfl_xor(v[q], o, 64);
  if(lane == 0)
    for(int q = 0; q < NV; q++)
      s_red[w][q] = v[q];
  __syncthreads();
  for(int q = 0; q < NV; q++)
    {
      double t = s_red[0][q];
      for(int k = 1; k < UW; k++)
        t += s_red[k][q];
      v[q] = t;
    }
  __syncthreads();
}

// One pass of subfind_unbind (subfind_serial.c:455-607) for one group per workgroup.
__global__ __launch_bounds__(UB) void k_unb_pass(UnbArgs a)
{
  __shared__ double s_red[UW][8];
  __shared__ double s_pmin[UW];
  __shared__ long long s_imin[UW];
  __shared__ int s_hist[2][256];
  __shared__ unsigned long long s_pref[2];
  __shared__ int s_rank[2];
  const int g = a.glist[blockIdx.x];
  const int64_t f = a.off[g], e = a.off[g + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int phase = a.gphase[g];

  double px, py, pz;
  if(phase == 1)
    {
      // :491-497: the weakly bound members' potentials are scaled by G/a again (the
      // reference evaluates them anew but keeps the old value)
      const double lim = a.wbl[g];
      for(int64_t i = f + t; i < e; i += UB)
        if(a.alive[i] && a.E[i] >= lim)
          a.pot[i] *= a.g_a;
      px = a.ppos[3 * g];
      py = a.ppos[3 * g + 1];
      pz = a.ppos[3 * g + 2];
    }
  else
    {
      // member of minimum potential (:460-471): the first one in member order on ties
      double best  = 1.0e30;
      long long bi = -1;
      for(int64_t i = f + t; i < e; i += UB)
        if(a.alive[i])
          {
            const double p = a.pot[i];
            if(bi < 0 || p < best)
              {
                best = p;
                bi   = i;
              }
          }
      for(int o = 32; o > 0; o >>= 1)
        {
          const double ob     = __shfl_xor(best, o, 64);
          const long long oi  = __shfl_xor(bi, o, 64);
          const bool take     = oi >= 0 && (bi < 0 || ob < best || (ob == best && oi < bi));
          best                = take ? ob : best;
          bi                  = take ? oi : bi;
        }
      if(lane == 0)
        {
          s_pmin[w] = best;
          s_imin[w] = bi;
        }
      __syncthreads();
      best = s_pmin[0];
      bi   = s_imin[0];
      for(int k = 1; k < UW; k++)
        if(s_imin[k] >= 0 && (bi < 0 || s_pmin[k] < best || (s_pmin[k] == best && s_imin[k] < bi)))
          {
            best = s_pmin[k];
            bi   = s_imin[k];
          }
      // a group in the loop always has a member left (the removal keeps the rank-0.25 len one)
      px = bi >= 0 ? a.pos[3 * bi] : 0.0;
      py = bi >= 0 ? a.pos[3 * bi + 1] : 0.0;
      pz = bi >= 0 ? a.pos[3 * bi + 2] : 0.0;
      __syncthreads();
      if(t == 0)
        {
          a.ppos[3 * g]     = px;
          a.ppos[3 * g + 1] = py;
          a.ppos[3 * g + 2] = pz;
        }
    }

  // bulk velocity and centre of mass (:501-532)
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for(int64_t i = f + t; i < e; i += UB)
    if(a.alive[i])
      {
        const double m = a.mass[i];
        v[0] += m * (a.pos[3 * i] - px);
        v[1] += m * (a.pos[3 * i + 1] - py);
        v[2] += m * (a.pos[3 * i + 2] - pz);
        v[3] += m * a.vel[3 * i];
        v[4] += m * a.vel[3 * i + 1];
        v[5] += m * a.vel[3 * i + 2];
        v[6] += m;
      }
  block_sum<7>(v, s_red);
  const double tot = v[6];
  const double vx = v[3] / tot, vy = v[4] / tot, vz = v[5] / tot;
  const double sx = v[0] / tot + px, sy = v[1] / tot + py, sz = v[2] / tot + pz;

  // binding energies (:534-556), with the self-energy G/a m / eps added back
  int nl = 0, npos = 0;
  for(int64_t i = f + t; i < e; i += UB)
    if(a.alive[i])
      {
        double dvx = a.vel_to_phys * (a.vel[3 * i] - vx);
        double dvy = a.vel_to_phys * (a.vel[3 * i + 1] - vy);
        double dvz = a.vel_to_phys * (a.vel[3 * i + 2] - vz);
        dvx += a.hubble * (a.atime * (a.pos[3 * i] - sx));
        dvy += a.hubble * (a.atime * (a.pos[3 * i + 1] - sy));
        dvz += a.hubble * (a.atime * (a.pos[3 * i + 2] - sz));
        double en = a.pot[i] + 0.5 * (dvx * dvx + dvy * dvy + dvz * dvz);
        en += a.g_a * a.mass[i] / a.epsr[g];
        if(a.type && a.uth && a.type[i] == 0)
          en += a.uth[i];
        a.E[i] = en;
        nl++;
        npos += en > 0;
      }
  double cnt[2] = {(double)nl, (double)npos};
  block_sum<2>(cnt, s_red);
  const int len = (int)cnt[0], P = (int)cnt[1];

  // energy_limit = 0.25 len-th largest energy (:560); weakly_bound_limit = the
  // energy where the running count of (E > 0) minus (E <= 0) over the sorted list
  // first drops to zero (:562-572), i.e. the min(max(2P - 1, 0), len - 1)-th largest
  if(t < 2)
    {
      s_pref[t] = 0;
      s_rank[t] = t == 0 ? (int)(0.25 * len) : min(max(2 * P - 1, 0), len - 1);
    }
  for(int shift = 56; shift >= 0; shift -= 8)
    {
      for(int k = t; k < 512; k += UB)
        (&s_hist[0][0])[k] = 0;
      __syncthreads();
      const uint64_t mask = shift == 56 ? 0ull : (~0ull << (shift + 8));
      const uint64_t p0 = s_pref[0], p1 = s_pref[1];
      for(int64_t i = f + t; i < e; i += UB)
        if(a.alive[i])
          {
            const uint64_t k = okey(a.E[i]);
            const int d      = (int)((k >> shift) & 255);
            if((k & mask) == p0)
              atomicAdd(&s_hist[0][d], 1);
            if((k & mask) == p1)
              atomicAdd(&s_hist[1][d], 1);
          }
      __syncthreads();
      if(lane == 0 && w < 2)
        {
          // the digit holding the rank-th largest key, counting from the top digit
          int r = s_rank[w], d = 255;
          for(; d > 0; d--)
            {
              const int c = s_hist[w][d];
              if(r < c)
                break;
              r -= c;
            }
          s_rank[w] = r;
          s_pref[w] |= (unsigned long long)d << shift;
        }
      __syncthreads();
    }
  const double energy_limit = okey_inv(s_pref[0]);

  // omit unbound members, at most a quarter of them (:574-590)
  int unb = 0, lng = 0;
  for(int64_t i = f + t; i < e; i += UB)
    if(a.alive[i])
      {
        const double en = a.E[i];
        if(en > 0 && en > energy_limit)
          {
            a.alive[i] = 0;
            unb++;
          }
        else if(!a.type || a.type[i] != 0)
          lng++;
      }
  double c2[2] = {(double)unb, (double)lng};
  block_sum<2>(c2, s_red);
  if(t == 0)
    {
      a.res[3 * g]     = (int)c2[0];
      a.res[3 * g + 1] = len - (int)c2[0];
      a.res[3 * g + 2] = (int)c2[1];
      a.wbl[g]         = okey_inv(s_pref[1]);
    }
}

// the potential solve of the members of every phase-0 group, through the batch path
static agrav_status unbind_potential(agrav_ctx *c, int ng, int64_t n, double g_a)
{
  UnbindBufs &u = c->unb;
  hipStream_t s = c->stream;
  agrav_status st;
  const unsigned nb = (unsigned)((n + B - 1) / B);
  int *d_total = u.cnt.p, *d_err = u.cnt.p + 1;
  k_unb_flag<<<nb, B, 0, s>>>(n, u.alive.p, u.gid.p, u.gphase.p, u.flag.p);
  AG_HIP(hipMemsetAsync(u.cnt.p, 0, 2 * sizeof(int), s));
  if((st = exclusive_scan_i32(u.flag.p, u.dst.p, n, d_total, u.scan_tmp.p, d_err, s)) != AGRAV_OK)
    return st;
  k_unb_soff<<<(ng + 1 + B - 1) / B, B, 0, s>>>(ng, n, u.off.p, u.dst.p, d_total, u.soff.p);
  AG_HIP(hipGetLastError());
  std::vector<int64_t> soff(ng + 1);
  int herr[2];
  AG_HIP(hipMemcpyAsync(soff.data(), u.soff.p, (ng + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  AG_HIP(hipMemcpyAsync(herr, u.cnt.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(herr[1])
    return fail(c, AGRAV_ERR_HIP, "unbinding: scan look-back timed out");
  const int64_t m  = soff[ng];
  const int64_t mm = m ? m : 1;
  if((st = c->d_pos_in.ensure(3 * mm)) || (st = c->d_mass_in.ensure(mm)) || (st = c->d_oldacc.ensure(mm)) ||
     (st = c->d_pset_in.ensure(mm)) || (st = c->d_soff.ensure(ng + 1)) || (st = c->d_at.ensure(3 * mm)))
    return st;
  k_unb_compact<<<nb, B, 0, s>>>(n, u.flag.p, u.dst.p, u.gid.p, u.pos.p, u.mass.p, c->d_pos_in.p, c->d_mass_in.p, c->d_pset_in.p,
                                 u.cidx.p);
  AG_HIP(hipGetLastError());
  AG_HIP(hipMemcpyAsync(c->d_soff.p, u.soff.p, (ng + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
  c->h_soff             = soff;
  c->batch_nsets        = ng;
  c->have_stype         = false;  // one softening per group
  c->batch_groups_dirty = true;
  c->n                  = m;
  c->have_particles     = true;
  c->have_oldacc        = false;
  c->have_tree = c->have_tree_force = c->have_pm = c->have_final = false;
  if(m == 0)
    return AGRAV_OK;
  if((st = agrav_construct_forcetree(c, nullptr)) != AGRAV_OK)
    return st;
  if((st = walk_tree(c, 1)) != AGRAV_OK)
    return st;
  k_unb_scatter<<<(unsigned)((m + B - 1) / B), B, 0, s>>>(m, c->d_idx2.p, u.cidx.p, c->d_pt.p, g_a, u.pot.p);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
}

}  // namespace agrav

using namespace agrav;

extern "C" {

agrav_status agrav_subfind_unbind(agrav_ctx *c, int ngroups, const int64_t *offsets, const double *pos, const double *vel,
                                  const double *mass, const double *softening, const double *utherm, const uint8_t *type,
                                  const double *cubes, const agrav_unbind_params *up, uint8_t *bound, int32_t *len_out,
                                  int32_t *len_non_gas, int32_t *iterations, double *potential, double *bnd_energy)
{
  if(!c || ngroups < 1 || !offsets || !softening || !up || !bound || !len_out)
    return AGRAV_ERR_BAD_PARAM;
  if(c->par.periodic || c->par.pmgrid > 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "unbinding works on isolated groups: periodic = 0 and pmgrid = 0 required");
  if(ngroups > (1 << 21))
    return fail(c, AGRAV_ERR_CAPACITY, "more than 2^21 groups in one call");
  if(offsets[0] != 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "group offsets must start at 0");
  for(int g = 0; g < ngroups; g++)
    {
      if(offsets[g + 1] < offsets[g])
        return fail(c, AGRAV_ERR_BAD_PARAM, "group offsets must not decrease");
      if(!(softening[g] > 0))
        return fail(c, AGRAV_ERR_BAD_PARAM, "softening must be positive");
      if(cubes && !(cubes[4 * g + 3] > 0))
        return fail(c, AGRAV_ERR_BAD_PARAM, "root cube lengths must be positive");
    }
  if(up->comoving && !(up->time_a > 0))
    return fail(c, AGRAV_ERR_BAD_PARAM, "comoving unbinding needs a positive scale factor");
  const int64_t n = offsets[ngroups];
  if(n > 0 && (!pos || !vel || !mass))
    return AGRAV_ERR_BAD_PARAM;
  if(n >= (1ll << 30))
    return fail(c, AGRAV_ERR_CAPACITY, "particle count exceeds 2^30 per context");
  if(utherm && !type)
    return fail(c, AGRAV_ERR_BAD_PARAM, "utherm needs the particle types (gas is type 0)");
  const int max_iter = up->max_iter > 0 ? up->max_iter : 300000;  // MAXITER, allvars.h:482

  UnbindBufs &u   = c->unb;
  hipStream_t s   = c->stream;
  const int64_t nn = n ? n : 1;
  agrav_status st;
  if((st = u.pos.ensure(3 * nn)) || (st = u.vel.ensure(3 * nn)) || (st = u.mass.ensure(nn)) || (st = u.pot.ensure(nn)) ||
     (st = u.E.ensure(nn)) || (st = u.alive.ensure(nn)) || (st = u.gid.ensure(nn)) || (st = u.flag.ensure(nn)) ||
     (st = u.dst.ensure(nn)) || (st = u.cidx.ensure(nn)) || (st = u.cube.ensure(4 * ngroups)) ||
     (st = u.epsr.ensure(ngroups)) || (st = u.wbl.ensure(ngroups)) || (st = u.ppos.ensure(3 * ngroups)) ||
     (st = u.glist.ensure(ngroups)) || (st = u.gphase.ensure(ngroups)) || (st = u.res.ensure(3 * ngroups)) ||
     (st = u.cnt.ensure(4)) || (st = u.off.ensure(ngroups + 1)) || (st = u.soff.ensure(ngroups + 1)) ||
     (st = u.scan_tmp.ensure(scan_temp_bytes(nn) + 4096)) || (st = c->d_shd.ensure(ngroups)) ||
     (st = c->d_sdom.ensure(ngroups)))
    return st;
  if(utherm && (st = u.uth.ensure(nn)))
    return st;
  if(type && (st = u.type.ensure(nn)))
    return st;

  // inputs; the softening length h = 2.8 eps per group (grav_softening.c:90-95) and
  // the eps the self-energy divides by, ForceSoftening / 2.8 (subfind_serial.c:553)
  std::vector<double> h(ngroups), epsr(ngroups);
  for(int g = 0; g < ngroups; g++)
    {
      h[g]    = 2.8 * softening[g];
      epsr[g] = h[g] / 2.8;
    }
  AG_HIP(hipMemcpyAsync(u.off.p, offsets, (ngroups + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  AG_HIP(hipMemcpyAsync(c->d_shd.p, h.data(), ngroups * sizeof(double), hipMemcpyHostToDevice, s));
  AG_HIP(hipMemcpyAsync(u.epsr.p, epsr.data(), ngroups * sizeof(double), hipMemcpyHostToDevice, s));
  if(n > 0)
    {
      AG_HIP(hipMemcpyAsync(u.pos.p, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
      AG_HIP(hipMemcpyAsync(u.vel.p, vel, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
      AG_HIP(hipMemcpyAsync(u.mass.p, mass, n * sizeof(double), hipMemcpyHostToDevice, s));
      if(utherm)
        AG_HIP(hipMemcpyAsync(u.uth.p, utherm, n * sizeof(double), hipMemcpyHostToDevice, s));
      if(type)
        AG_HIP(hipMemcpyAsync(u.type.p, type, n * sizeof(uint8_t), hipMemcpyHostToDevice, s));
      k_unb_gid<<<(unsigned)((n + B - 1) / B), B, 0, s>>>(n, ngroups, u.off.p, u.gid.p, u.alive.p);
      AG_HIP(hipGetLastError());
      AG_HIP(hipMemsetAsync(u.pot.p, 0, n * sizeof(double), s));
      AG_HIP(hipMemsetAsync(u.E.p, 0, n * sizeof(double), s));
    }
  if(cubes)
    AG_HIP(hipMemcpyAsync(u.cube.p, cubes, 4 * ngroups * sizeof(double), hipMemcpyHostToDevice, s));
  else
    {
      k_unb_cubes<<<ngroups, B, 0, s>>>(u.pos.p, u.off.p, u.cube.p);
      AG_HIP(hipGetLastError());
    }

  // the walk: Barnes-Hut with ErrTolThetaSubfind (subfind_loctree.c:639) and the
  // potential; the context's group-walk tuning applies
  const int keep_pot   = c->par.eval_potential;
  c->par.eval_potential = 1;
  c->batch_cubes        = u.cube.p;
  const double cf_atime = up->comoving ? up->time_a : 1.0;
  const double g_a      = c->par.G / cf_atime;

  // host state per group: phase 0 (full potential next), 1 (energies only), 2 (done)
  std::vector<int> phase(ngroups, 0), len(ngroups), lng(ngroups, 0), iter(ngroups, 0), res(3 * ngroups);
  for(int g = 0; g < ngroups; g++)
    {
      len[g] = (int)(offsets[g + 1] - offsets[g]);
      if(len[g] == 0)
        phase[g] = 2;
    }
  std::vector<int> glist;
  glist.reserve(ngroups);
  int rounds = 0, solves = 0;
  double solve_ms = 0, pass_ms = 0;
  st = AGRAV_OK;
  while(st == AGRAV_OK)
    {
      glist.clear();
      bool need_pot = false;
      for(int g = 0; g < ngroups; g++)
        if(phase[g] != 2)
          {
            glist.push_back(g);
            need_pot |= phase[g] == 0;
          }
      if(glist.empty())
        break;
      // largest groups first, so the longest workgroups start early
      std::stable_sort(glist.begin(), glist.end(), [&](int a, int b) { return len[a] > len[b]; });
      if(hipMemcpyAsync(u.gphase.p, phase.data(), ngroups * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(u.glist.p, glist.data(), glist.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
        {
          st = hip_fail(hipGetLastError(), "unbinding upload", __FILE__, __LINE__);
          break;
        }
      if(need_pot)
        {
          (void)hipEventRecord(c->ev[5], s);
          if((st = unbind_potential(c, ngroups, n, g_a)) != AGRAV_OK)
            break;
          solves++;
        }
      (void)hipEventRecord(c->ev[6], s);
      UnbArgs a{u.off.p,  u.glist.p, u.gphase.p, u.pos.p, u.vel.p,   u.mass.p, utherm ? u.uth.p : nullptr, u.epsr.p,
                type ? u.type.p : nullptr, u.pot.p, u.E.p, u.wbl.p, u.ppos.p, u.alive.p, u.res.p, g_a,
                up->comoving ? 1.0 / up->time_a : 1.0, cf_atime, up->comoving ? up->hubble_a : 0.0};
      k_unb_pass<<<(unsigned)glist.size(), UB, 0, s>>>(a);
      if(hipGetLastError() != hipSuccess || hipEventRecord(c->ev[4], s) != hipSuccess ||
         hipMemcpyAsync(res.data(), u.res.p, 3 * ngroups * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess)
        {
          st = hip_fail(hipGetLastError(), "unbinding pass", __FILE__, __LINE__);
          break;
        }
      float ms = 0;
      if(need_pot && hipEventElapsedTime(&ms, c->ev[5], c->ev[6]) == hipSuccess)
        solve_ms += ms;
      if(hipEventElapsedTime(&ms, c->ev[6], c->ev[4]) == hipSuccess)
        pass_ms += ms;
      rounds++;
      // the loop control of subfind_serial.c:592-607
      for(int g : glist)
        {
          int unbound = res[3 * g];
          len[g]      = res[3 * g + 1];
          lng[g]      = res[3 * g + 2];
          if(len[g] < up->des_link_ngb)
            {
              phase[g] = 2;
              continue;
            }
          if(phase[g] == 0)
            {
              if(unbound > 0)
                phase[g] = 1;
            }
          else if(unbound == 0)
            {
              phase[g] = 0;
              unbound  = 1;
            }
          if(iter[g]++ > max_iter)
            {
              st = fail(c, AGRAV_ERR_CAPACITY, "unbinding passed MAXITER iterations");
              break;
            }
          if(unbound == 0)
            phase[g] = 2;
        }
    }
  c->tim.unbind_rounds   = rounds;
  c->tim.unbind_solves   = solves;
  c->tim.unbind_solve_ms = solve_ms;
  c->tim.unbind_pass_ms  = pass_ms;
  // the context's particle state was the last compacted batch: callers set particles anew
  c->batch_cubes        = nullptr;
  c->par.eval_potential = keep_pot;
  c->batch_nsets        = 0;
  c->have_particles = c->have_tree = c->have_tree_force = c->have_pm = c->have_final = false;
  if(st != AGRAV_OK)
    return st;

  if(n > 0)
    {
      AG_HIP(hipMemcpyAsync(bound, u.alive.p, n * sizeof(uint8_t), hipMemcpyDeviceToHost, s));
      if(potential)
        AG_HIP(hipMemcpyAsync(potential, u.pot.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
      if(bnd_energy)
        AG_HIP(hipMemcpyAsync(bnd_energy, u.E.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
    }
  for(int g = 0; g < ngroups; g++)
    {
      len_out[g] = len[g];
      if(len_non_gas)
        len_non_gas[g] = lng[g];
      if(iterations)
        iterations[g] = iter[g];
    }
  return AGRAV_OK;
}

}  // extern "C"
