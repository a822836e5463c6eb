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
  const int *gphase;               // per group: 0 full pass, 1 energies only, 2 done
  const int2 *chunk;               // {group, first member} per chunk of at most CH members
  const int *cbeg;                 // first chunk of every group (ngroups + 1)
  const double *pos, *vel, *mass, *uth, *epsr;
  const uint8_t *type;
  double *pot, *E, *wbl, *lim, *ppos, *sv;
  uint8_t *alive;
  double *cpart;                   // per chunk: 7 partial sums (centre of mass, bulk velocity, mass)
  double *cmin;                    // per chunk: minimum potential ...
  long long *cimin;                //            ... and its member (-1: none)
  int *ccnt;                       // per chunk: two counts
  int *glen;                       // per group: {len, P = #(E > 0)}
  int *res;                        // per group {unbound, len left, len_non_gas}
  double g_a, vel_to_phys, atime, hubble;
};

static constexpr int CB = 256;          // threads of the per-chunk kernels
static constexpr int CH = 16 * CB;      // members per chunk

// fixed-order sum over the block of NV doubles (wave butterflies, then the waves in order)
template <int NT, int NV> __device__ inline void block_sum(double (&v)[NV], double (*s_red)[8])
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
      for(int k = 1; k < NT / 64; k++)
        t += s_red[k][q];
      v[q] = t;
    }
  __syncthreads();
}

// (minimum, member) pairs: the smaller potential, the first member on ties, -1 = none
__device__ inline void min_merge(double &best, long long &bi, double ob, long long oi)
{
  if(oi >= 0 && (bi < 0 || ob < best || (ob == best && oi < bi)))
    {
      best = ob;
      bi   = oi;
    }
}

__device__ inline void wave_min(double &best, long long &bi)
{
  for(int o = 32; o > 0; o >>= 1)
    min_merge(best, bi, __shfl_xor(best, o, 64), __shfl_xor(bi, o, 64));
}

// the members of a chunk of an active group: [f, e), or an empty range
__device__ inline int chunk_range(const UnbArgs &a, int c, int64_t &f, int64_t &e)
{
  const int2 ch = a.chunk[c];
  f             = ch.y;
  e             = f + CH < a.off[ch.x + 1] ? f + CH : a.off[ch.x + 1];
  return ch.x;
}

// 1. energy-only pass (:491-497): the weakly bound members' potentials are scaled by
// G/a again (the reference evaluates them anew but keeps the old value); full pass:
// each chunk's member of minimum potential (:460-471)
__global__ __launch_bounds__(CB) void k_uc_min(UnbArgs a)
{
  __shared__ double s_min[CB / 64];
  __shared__ long long s_imin[CB / 64];
  int64_t f, e;
  const int g = chunk_range(a, blockIdx.x, f, e), ph = a.gphase[g];
  if(ph == 2)
    return;
  if(ph == 1)
    {
      if(a.g_a != 1.0)
        {
          const double lim = a.wbl[g];
          for(int64_t i = f + threadIdx.x; i < e; i += CB)
            if(a.alive[i] && a.E[i] >= lim)
              a.pot[i] *= a.g_a;
        }
      return;
    }
  double best  = 0;
  long long bi = -1;
  for(int64_t i = f + threadIdx.x; i < e; i += CB)
    if(a.alive[i])
      min_merge(best, bi, a.pot[i], i);
  wave_min(best, bi);
  if((threadIdx.x & 63) == 0)
    {
      s_min[threadIdx.x >> 6]  = best;
      s_imin[threadIdx.x >> 6] = bi;
    }
  __syncthreads();
  if(threadIdx.x == 0)
    {
      for(int w = 1; w < CB / 64; w++)
        min_merge(best, bi, s_min[w], s_imin[w]);
      a.cmin[blockIdx.x]  = best;
      a.cimin[blockIdx.x] = bi;
    }
}

// 2. per group (one wavefront): the member of minimum potential over its chunks in order
__global__ void k_ug_min(UnbArgs a, int ng)
{
  const int g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if(g >= ng || a.gphase[g] != 0)
    return;
  double best  = 0;
  long long bi = -1;
  for(int c = a.cbeg[g] + lane; c < a.cbeg[g + 1]; c += 64)
    min_merge(best, bi, a.cmin[c], a.cimin[c]);
  wave_min(best, bi);
  // a group in the loop always has a member left (the removal keeps the rank-0.25 len one)
  if(lane == 0)
    for(int d = 0; d < 3; d++)
      a.ppos[3 * g + d] = bi >= 0 ? a.pos[3 * bi + d] : 0.0;
}

// 3. bulk velocity and centre of mass (:501-532): chunk partial sums
__global__ __launch_bounds__(CB) void k_uc_sums(UnbArgs a)
{
  __shared__ double s_red[CB / 64][8];
  int64_t f, e;
  const int g = chunk_range(a, blockIdx.x, f, e);
  if(a.gphase[g] == 2)
    return;
  const double px = a.ppos[3 * g], py = a.ppos[3 * g + 1], pz = a.ppos[3 * g + 2];
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for(int64_t i = f + threadIdx.x; i < e; i += CB)
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
  block_sum<CB, 7>(v, s_red);
  if(threadIdx.x < 7)
    a.cpart[8 * (int64_t)blockIdx.x + threadIdx.x] = v[threadIdx.x];
}

// 4. per group (one wavefront): s = ppos + sum m (x - ppos) / M, v = sum m vel / M
__global__ void k_ug_sums(UnbArgs a, int ng)
{
  const int g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if(g >= ng || a.gphase[g] == 2)
    return;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for(int c = a.cbeg[g] + lane; c < a.cbeg[g + 1]; c += 64)
    for(int q = 0; q < 7; q++)
      v[q] += a.cpart[8 * (int64_t)c + q];
  for(int q = 0; q < 7; q++)
    for(int o = 32; o > 0; o >>= 1)
      v[q] += __shfl_xor(v[q], o, 64);
  if(lane == 0)
    {
      const double tot = v[6];
      for(int d = 0; d < 3; d++)
        {
          a.sv[6 * g + d]     = v[d] / tot + a.ppos[3 * g + d];
          a.sv[6 * g + 3 + d] = v[3 + d] / tot;
        }
    }
}

// 5. binding energies (:534-556), with the self-energy G/a m / eps added back; chunk
// counts of members and of positive energies
__global__ __launch_bounds__(CB) void k_uc_energy(UnbArgs a)
{
  __shared__ double s_red[CB / 64][8];
  int64_t f, e;
  const int g = chunk_range(a, blockIdx.x, f, e);
  if(a.gphase[g] == 2)
    return;
  const double sx = a.sv[6 * g], sy = a.sv[6 * g + 1], sz = a.sv[6 * g + 2];
  const double vx = a.sv[6 * g + 3], vy = a.sv[6 * g + 4], vz = a.sv[6 * g + 5];
  const double eps = a.epsr[g];
  double cnt[2] = {0, 0};
  for(int64_t i = f + threadIdx.x; i < e; i += CB)
    if(a.alive[i])
      {
        double dvx = a.vel_to_phys * (a.vel[3 * i] - vx);
        double dvy = a.vel_to_phys * (a.vel[3 * i + 1] - vy);
        double dvz = a.vel_to_phys * (a.vel[3 * i + 2] - vz);
        dvx += a.hubble * (a.atime * (a.pos[3 * i] - sx));
        dvy += a.hubble * (a.atime * (a.pos[3 * i + 1] - sy));
        dvz += a.hubble * (a.atime * (a.pos[3 * i + 2] - sz));
        double en = a.pot[i] + 0.5 * (dvx * dvx + dvy * dvy + dvz * dvz);
        en += a.g_a * a.mass[i] / eps;
        if(a.type && a.uth && a.type[i] == 0)
          en += a.uth[i];
        a.E[i] = en;
        cnt[0] += 1.0;
        cnt[1] += en > 0 ? 1.0 : 0.0;
      }
  block_sum<CB, 2>(cnt, s_red);
  if(threadIdx.x == 0)
    {
      a.ccnt[2 * blockIdx.x]     = (int)cnt[0];
      a.ccnt[2 * blockIdx.x + 1] = (int)cnt[1];
    }
}

// 6. per group (one wavefront): members and positive energies
__global__ void k_ug_count(UnbArgs a, int ng)
{
  const int g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if(g >= ng || a.gphase[g] == 2)
    return;
  int n = 0, p = 0;
  for(int c = a.cbeg[g] + lane; c < a.cbeg[g + 1]; c += 64)
    {
      n += a.ccnt[2 * c];
      p += a.ccnt[2 * c + 1];
    }
  for(int o = 32; o > 0; o >>= 1)
    {
      n += __shfl_xor(n, o, 64);
      p += __shfl_xor(p, o, 64);
    }
  if(lane == 0)
    {
      a.glen[2 * g]     = n;
      a.glen[2 * g + 1] = p;
    }
}

// 7. energy_limit = 0.25 len-th largest energy (:560); weakly_bound_limit = the energy
// where the running count of (E > 0) minus (E <= 0) over the sorted list first drops
// to zero (:562-572), i.e. the min(max(2P - 1, 0), len - 1)-th largest.  With at most
// 0.25 len positive energies the energy limit is <= 0 and the removal is exactly
// {E > 0}; the weakly bound limit only scales potentials by G/a, which is 1 in
// physical runs with G = 1.  Either order statistic is selected (exact 8-bit radix
// select on order-preserving keys, one workgroup per group) only when it changes
// the outcome.
__global__ __launch_bounds__(UB) void k_ug_select(UnbArgs a, const int *__restrict__ glist)
{
  __shared__ int s_hist[2][256];
  __shared__ unsigned long long s_pref[2];
  __shared__ int s_rank[2];
  const int g = glist[blockIdx.x];
  const int len = a.glen[2 * g], P = a.glen[2 * g + 1];
  const int k1        = (int)(0.25 * len);
  const bool need_lim = P > k1;
  const bool need_wbl = a.g_a != 1.0;
  if(!need_lim && !need_wbl)
    return;
  const int64_t f = a.off[g], e = a.off[g + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if(t < 2)
    {
      s_pref[t] = 0;
      s_rank[t] = t == 0 ? k1 : min(max(2 * P - 1, 0), len - 1);
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
            if(need_lim && (k & mask) == p0)
              atomicAdd(&s_hist[0][d], 1);
            if(need_wbl && (k & mask) == p1)
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
  if(t == 0)
    {
      a.lim[g] = need_lim ? okey_inv(s_pref[0]) : 0.0;
      a.wbl[g] = need_wbl ? okey_inv(s_pref[1]) : 0.0;
    }
}

// 8. omit unbound members, at most a quarter of them (:574-590); chunk counts of the
// removed members and of the non-gas members left
__global__ __launch_bounds__(CB) void k_uc_remove(UnbArgs a)
{
  __shared__ double s_red[CB / 64][8];
  int64_t f, e;
  const int g = chunk_range(a, blockIdx.x, f, e);
  if(a.gphase[g] == 2)
    return;
  const int len = a.glen[2 * g], P = a.glen[2 * g + 1];
  const double lim = P > (int)(0.25 * len) ? a.lim[g] : 0.0;
  double cnt[2] = {0, 0};
  for(int64_t i = f + threadIdx.x; i < e; i += CB)
    if(a.alive[i])
      {
        const double en = a.E[i];
        if(en > 0 && en > lim)
          {
            a.alive[i] = 0;
            cnt[0] += 1.0;
          }
        else if(!a.type || a.type[i] != 0)
          cnt[1] += 1.0;
      }
  block_sum<CB, 2>(cnt, s_red);
  if(threadIdx.x == 0)
    {
      a.ccnt[2 * blockIdx.x]     = (int)cnt[0];
      a.ccnt[2 * blockIdx.x + 1] = (int)cnt[1];
    }
}

// 9. per group (one wavefront): {unbound, len left, len_non_gas}
__global__ void k_ug_result(UnbArgs a, int ng)
{
  const int g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if(g >= ng || a.gphase[g] == 2)
    return;
  int u = 0, l = 0;
  for(int c = a.cbeg[g] + lane; c < a.cbeg[g + 1]; c += 64)
    {
      u += a.ccnt[2 * c];
      l += a.ccnt[2 * c + 1];
    }
  for(int o = 32; o > 0; o >>= 1)
    {
      u += __shfl_xor(u, o, 64);
      l += __shfl_xor(l, o, 64);
    }
  if(lane == 0)
    {
      a.res[3 * g]     = u;
      a.res[3 * g + 1] = a.glen[2 * g] - u;
      a.res[3 * g + 2] = l;
    }
}

// one pass of subfind_unbind (subfind_serial.c:455-590) for every active group
static agrav_status unbind_pass(const UnbArgs &a, int ng, int nchunks, int nactive, const int *glist, hipStream_t s)
{
  const unsigned gb = (unsigned)((ng + 3) / 4);
  k_uc_min<<<nchunks, CB, 0, s>>>(a);
  k_ug_min<<<gb, 256, 0, s>>>(a, ng);
  k_uc_sums<<<nchunks, CB, 0, s>>>(a);
  k_ug_sums<<<gb, 256, 0, s>>>(a, ng);
  k_uc_energy<<<nchunks, CB, 0, s>>>(a);
  k_ug_count<<<gb, 256, 0, s>>>(a, ng);
  k_ug_select<<<nactive, UB, 0, s>>>(a, glist);
  k_uc_remove<<<nchunks, CB, 0, s>>>(a);
  k_ug_result<<<gb, 256, 0, s>>>(a, ng);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
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
  c->n_active           = -1;
  c->have_stype         = false;  // one softening per group
  c->batch_groups_dirty = true;
  c->n                  = m;
  c->have_particles     = true;
  c->have_oldacc        = false;
  c->gcost_groups      = -1;
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
     (st = c->d_sdom.ensure(ngroups)) || (st = u.lim.ensure(ngroups)) || (st = u.sv.ensure(6 * ngroups)) ||
     (st = u.glen.ensure(2 * ngroups)) || (st = u.cbeg.ensure(ngroups + 1)))
    return st;
  // chunks of at most CH members, each inside one group
  std::vector<int> cbeg(ngroups + 1);
  std::vector<int2> &chunks = u.h_chunk;
  chunks.clear();
  for(int g = 0; g < ngroups; g++)
    {
      cbeg[g] = (int)chunks.size();
      for(int64_t q = offsets[g]; q < offsets[g + 1]; q += CH)
        chunks.push_back(make_int2(g, (int)q));
    }
  cbeg[ngroups]     = (int)chunks.size();
  const int nchunks = (int)chunks.size();
  if((st = u.chunk.ensure(nchunks)) || (st = u.cpart.ensure(8 * (size_t)nchunks)) || (st = u.cmin.ensure(nchunks)) ||
     (st = u.cimin.ensure(nchunks)) || (st = u.ccnt.ensure(2 * (size_t)nchunks)))
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
  AG_HIP(hipMemcpyAsync(u.cbeg.p, cbeg.data(), (ngroups + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  if(nchunks > 0)
    AG_HIP(hipMemcpyAsync(u.chunk.p, chunks.data(), nchunks * sizeof(int2), hipMemcpyHostToDevice, s));
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
      UnbArgs a{u.off.p,   u.gphase.p, u.chunk.p, u.cbeg.p, u.pos.p,  u.vel.p,  u.mass.p,
                utherm ? u.uth.p : nullptr, u.epsr.p, type ? u.type.p : nullptr, u.pot.p, u.E.p, u.wbl.p,
                u.lim.p,   u.ppos.p,   u.sv.p,    u.alive.p, u.cpart.p, u.cmin.p, u.cimin.p, u.ccnt.p,
                u.glen.p,  u.res.p,    g_a,       up->comoving ? 1.0 / up->time_a : 1.0, cf_atime,
                up->comoving ? up->hubble_a : 0.0};
      if((st = unbind_pass(a, ngroups, nchunks, (int)glist.size(), u.glist.p, s)) != AGRAV_OK)
        break;
      if(hipEventRecord(c->ev[4], s) != hipSuccess ||
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
