// agrav_build.hip -- octree construction (construct_forcetree, forcetree.c:110-1429).
//
// AREPO inserts particles one by one into a pointer octree (forcetree.c:835-938)
// and then computes moments with a recursive depth-first pass (:1217-1429).
// Here the same octree -- every cube that holds two or more particles, single
// particles hanging in the child slots -- is derived in parallel from sorted
// 63-bit Peano-Hilbert keys:
//   1. domain cube (domain_box.c:107-197) and keys from AREPO's 52-bit integer
//      coordinates (forcetree.c:493-505), top 21 bits per axis;
//   2. LSD radix sort of the keys (Hilbert order, like domain.c:140);
//   3. binary radix tree over the sorted keys (Karras, HPG 2012);
//   4. each binary node whose prefix crosses 3-bit digit boundaries yields the
//      octree cubes in between; cubes are placed in pre-order (first particle,
//      then level) so AREPO's nextnode/sibling threading (forcetree.c:1288-1330)
//      follows from array positions;
//   5. moments level by level, deepest first, in FP64.
// Cube centres and lengths use AREPO's integer formula (forcetree.c:903-906),
// so node geometry is bit-identical to the reference's.  Cells of the
// 21st level that still hold several particles become buckets whose
// particles are direct children (AREPO would split them further, up to level 30).
#include "agrav_ctx.h"

// The domain cube, integer coordinates and cube centres must round exactly like
// the reference's C code (no fused multiply-add), so that node geometry is
// bit-identical.
#pragma clang fp contract(off)

namespace agrav
{

static constexpr int B = 256;

static inline unsigned nblk(int64_t n, int b = B) { return (unsigned)((n + b - 1) / b); }

__device__ inline unsigned long long ordered_bits(double v)
{
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double from_ordered(unsigned long long u)
{
  u = (u & 0x8000000000000000ull) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
  return __longlong_as_double((long long)u);
}

__device__ inline double wrap_box(double x, double L)
{
  if(x < 0)
    x += L;
  if(x >= L)
    x -= L;
  return x;
}

// ---------------------------------------------------------------------------
// 1. domain extent (domain_box.c:107-197, DO_NOT_RANDOMIZE_DOMAINCENTER)
// ---------------------------------------------------------------------------

__global__ void k_extent(const double *__restrict__ pos, int64_t n, unsigned long long *mm)
{
  // mm[0..2] = ordered min, mm[3..5] = ordered max (initialised by the host)
  __shared__ unsigned long long smin[3][B / 64], smax[3][B / 64];
  unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0, 0, 0};
  for(int64_t i = (int64_t)blockIdx.x * B + threadIdx.x; i < n; i += (int64_t)gridDim.x * B)
    for(int d = 0; d < 3; d++)
      {
        unsigned long long o = ordered_bits(pos[3 * i + d]);
        lo[d]                = o < lo[d] ? o : lo[d];
        hi[d]                = o > hi[d] ? o : hi[d];
      }
  for(int d = 0; d < 3; d++)
    {
      for(int off = 32; off > 0; off >>= 1)
        {
          unsigned long long a = __shfl_xor(lo[d], off, 64), b = __shfl_xor(hi[d], off, 64);
          lo[d]                = a < lo[d] ? a : lo[d];
          hi[d]                = b > hi[d] ? b : hi[d];
        }
      if((threadIdx.x & 63) == 0)
        {
          smin[d][threadIdx.x >> 6] = lo[d];
          smax[d][threadIdx.x >> 6] = hi[d];
        }
    }
  __syncthreads();
  if(threadIdx.x < 3)
    {
      int d                = threadIdx.x;
      unsigned long long a = smin[d][0], b = smax[d][0];
      for(int w = 1; w < B / 64; w++)
        {
          a = smin[d][w] < a ? smin[d][w] : a;
          b = smax[d][w] > b ? smax[d][w] : b;
        }
      atomicMin(&mm[d], a);
      atomicMax(&mm[3 + d], b);
    }
}

__global__ void k_domain_finish(const unsigned long long *mm, Domain *dom, int periodic, double box, int64_t n)
{
  double xmin[3], xmax[3];
  for(int d = 0; d < 3; d++)
    {
      xmin[d] = 0;  // preset to the simulation box (domain_box.c:112-116)
      xmax[d] = box;
      if(!periodic && n > 0)
        {
          double a = from_ordered(mm[d]), b = from_ordered(mm[3 + d]);
          xmin[d]  = a < xmin[d] ? a : xmin[d];
          xmax[d]  = b > xmax[d] ? b : xmax[d];
        }
      dom->xmin[d] = xmin[d];
      dom->xmax[d] = xmax[d];
    }
  double len = 0;
  for(int d = 0; d < 3; d++)
    if(xmax[d] - xmin[d] > len)
      len = xmax[d] - xmin[d];
  if(len <= 0)
    len = 1.0;
  len *= periodic ? 1.00001 : 1.2;  // domain_box.c:161-168
  for(int d = 0; d < 3; d++)
    {
      dom->center[d] = 0.5 * (xmin[d] + xmax[d]);
      dom->corner[d] = 0.5 * (xmin[d] + xmax[d]) - 0.5 * len;
    }
  dom->len     = len;
  dom->inv_len = 1.0 / len;
  dom->bigfac  = len / (double)(1ll << 52);
  if(periodic)
    {
      dom->fix_unit = box / 4294967296.0;  // 2^32 steps per box: wrap-around is the nearest image
      for(int d = 0; d < 3; d++)
        dom->fix_org[d] = 0;
    }
  else
    {
      dom->fix_unit = len / 2147483648.0;  // 2^31 steps per domain: differences fit in int32
      for(int d = 0; d < 3; d++)
        dom->fix_org[d] = dom->corner[d];
    }
  dom->fix_inv    = 1.0 / dom->fix_unit;
  dom->out_of_box = 0;
}

__device__ inline uint32_t to_fix(double x, const Domain &dom, int d)
{
  return (uint32_t)(long long)llrint((x - dom.fix_org[d]) * dom.fix_inv);
}

// Geometry of one build: a single particle set (one domain, 21 Hilbert levels) or
// a batch of independent sets, each with its own domain cube; batch keys carry the
// set index in their top `sbits` bits above `levels` Hilbert levels per axis.
struct BuildGeo
{
  const Domain *dom;      // 1 domain, or one per set
  const int *pset_in;     // set of each caller-order particle (batch only)
  const int *pset;        // set of each sorted particle (batch only)
  int sbits;              // key bits above the Hilbert digits
  int levels;             // Hilbert levels per axis
};

// batch: each set's domain cube (domain_box.c:107-197, isolated rule with
// BoxSize = 0: the cube holds the set's particles and the origin)
__global__ __launch_bounds__(256) void k_set_domains(const double *__restrict__ pos, const int64_t *__restrict__ off, Domain *sdom)
{
  __shared__ double s_lo[3][4], s_hi[3][4];
  const int s = blockIdx.x;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // preset [0, BoxSize] with BoxSize = 0
  for(int64_t i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x)
    for(int d = 0; d < 3; d++)
      {
        double x = pos[3 * i + d];
        lo[d]    = x < lo[d] ? x : lo[d];
        hi[d]    = x > hi[d] ? x : hi[d];
      }
  for(int d = 0; d < 3; d++)
    {
      for(int o = 32; o > 0; o >>= 1)
        {
          double a = __shfl_xor(lo[d], o, 64), bb = __shfl_xor(hi[d], o, 64);
          lo[d]    = a < lo[d] ? a : lo[d];
          hi[d]    = bb > hi[d] ? bb : hi[d];
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
      Domain dm{};
      double len = 0;
      for(int d = 0; d < 3; d++)
        {
          double a = s_lo[d][0], bb = s_hi[d][0];
          for(int w = 1; w < 4; w++)
            {
              a  = s_lo[d][w] < a ? s_lo[d][w] : a;
              bb = s_hi[d][w] > bb ? s_hi[d][w] : bb;
            }
          dm.xmin[d] = a;
          dm.xmax[d] = bb;
          if(bb - a > len)
            len = bb - a;
        }
      if(len <= 0)
        len = 1.0;
      len *= 1.2;  // domain_box.c:165
      for(int d = 0; d < 3; d++)
        {
          dm.center[d]  = 0.5 * (dm.xmin[d] + dm.xmax[d]);
          dm.corner[d]  = dm.center[d] - 0.5 * len;
          dm.fix_org[d] = dm.corner[d];
        }
      dm.len      = len;
      dm.inv_len  = 1.0 / len;
      dm.bigfac   = len / (double)(1ll << 52);
      dm.fix_unit = len / 2147483648.0;
      dm.fix_inv  = 1.0 / dm.fix_unit;
      sdom[s]     = dm;
    }
}

// ---------------------------------------------------------------------------
// 2. keys (forcetree.c:394-417 wrap, :493-505 integer coordinates)
// ---------------------------------------------------------------------------

__device__ inline uint64_t int52(double x, const Domain &dom, int d)
{
  double v = (x - dom.corner[d]) * dom.inv_len + 1.0;
  return (uint64_t)__double_as_longlong(v) & 0xFFFFFFFFFFFFFull;
}

__global__ void k_keys(const double *__restrict__ pos, int64_t n, BuildGeo geo, int periodic, double box,
                       uint64_t *__restrict__ keys, int *__restrict__ idx, Domain *dom_w)
{
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  const int set    = geo.pset_in ? geo.pset_in[i] : 0;
  const Domain dom = geo.dom[set];
  uint32_t c[3];
  bool bad = false;
  for(int d = 0; d < 3; d++)
    {
      double x = pos[3 * i + d];
      if(periodic)
        x = wrap_box(x, box);
      if(!(x >= dom.corner[d] && x < dom.corner[d] + dom.len))
        bad = true;
      c[d] = (uint32_t)(int52(x, dom, d) >> (52 - AG_KEY_LEVELS));
    }
  if(bad)
    {
      atomicExch(&dom_w->out_of_box, 1);
      c[0] = c[1] = c[2] = 0;
    }
  const int hb = 3 * geo.levels;  // Hilbert bits kept; their top digits are the coarser cells
  keys[i]      = ((uint64_t)set << hb) | (hilbert_key21(c[0], c[1], c[2]) >> (3 * AG_KEY_LEVELS - hb));
  idx[i]       = (int)i;
}

__global__ void k_gather(const int *__restrict__ idx, const double *__restrict__ pos, const double *__restrict__ mass,
                         const uint8_t *__restrict__ stype, const float *__restrict__ oldacc, int64_t n, int periodic, double box,
                         BuildGeo geo, int *pset_sorted, double *px, double *py, double *pz, double *pm, uint8_t *ps, uint4 *prec,
                         float *oldacc_s)
{
  int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k >= n)
    return;
  int i            = idx[k];
  const int set    = geo.pset_in ? geo.pset_in[i] : 0;
  const Domain dom = geo.dom[set];
  if(pset_sorted)
    pset_sorted[k] = set;
  double x = pos[3 * (int64_t)i], y = pos[3 * (int64_t)i + 1], z = pos[3 * (int64_t)i + 2];
  if(periodic)
    {
      x = wrap_box(x, box);
      y = wrap_box(y, box);
      z = wrap_box(z, box);
    }
  px[k]     = x;
  py[k]     = y;
  pz[k]     = z;
  double m  = mass[i];
  pm[k]     = m;
  if(ps)
    ps[k] = stype ? stype[i] : 0;
  prec[k] = make_uint4(to_fix(x, dom, 0), to_fix(y, dom, 1), to_fix(z, dom, 2), __float_as_uint((float)m));
  if(oldacc_s)
    oldacc_s[k] = oldacc ? oldacc[i] : 0.f;
}

// ---------------------------------------------------------------------------
// 3. binary radix tree (Karras 2012)
// ---------------------------------------------------------------------------

__device__ inline int kdelta(const uint64_t *__restrict__ k, int n, int i, int j)
{
  if(j < 0 || j >= n)
    return -1;
  uint64_t a = k[i], b = k[j];
  if(a == b)
    return 64 + __clz((unsigned)(i ^ j));
  return __clzll((long long)(a ^ b)) - 1;  // keys use bits 62..0
}

// octree level of a binary node with the given common-prefix length; -1 while the
// prefix has not left the set-index bits of a batch key
__device__ inline int oct_level(int prefix, int sbits, int levels)
{
  return prefix < sbits ? -1 : min((prefix - sbits) / 3, levels);
}

__global__ void k_binary(const uint64_t *__restrict__ keys, int n, int *bfirst, int *blast, int *bprefix, int *bparent)
{
  int i = blockIdx.x * B + threadIdx.x;
  if(i >= n - 1)
    return;
  int dl = kdelta(keys, n, i, i - 1), dr = kdelta(keys, n, i, i + 1);
  int d    = dr > dl ? 1 : -1;
  int dmin = dr > dl ? dl : dr;
  int lmax = 2;
  while(kdelta(keys, n, i, i + lmax * d) > dmin)
    lmax <<= 1;
  int l = 0;
  for(int t = lmax >> 1; t >= 1; t >>= 1)
    if(kdelta(keys, n, i, i + (l + t) * d) > dmin)
      l += t;
  int j     = i + l * d;
  int dnode = kdelta(keys, n, i, j);
  int s = 0, t = l;
  do
    {
      t = (t + 1) >> 1;
      if(kdelta(keys, n, i, i + (s + t) * d) > dnode)
        s += t;
    }
  while(t > 1);
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int first = i < j ? i : j, last = i < j ? j : i;
  bfirst[i]  = first;
  blast[i]   = last;
  bprefix[i] = dnode;
  if(first != gamma)
    bparent[gamma] = i;
  if(last != gamma + 1)
    bparent[gamma + 1] = i;
  if(i == 0)
    bparent[0] = -1;
}

// number of octree cubes each binary node stands for, and the cubes per level
// (aggregated in LDS: one global atomic per block and level)
__global__ __launch_bounds__(B) void k_oct_count(int n, const int *__restrict__ bprefix, const int *__restrict__ bparent, int sbits,
                                                 int levels, int *bcount, int *lvlcnt)
{
  __shared__ int s_cnt[AG_KEY_LEVELS + 1];
  if(threadIdx.x <= AG_KEY_LEVELS)
    s_cnt[threadIdx.x] = 0;
  __syncthreads();
  int i = blockIdx.x * B + threadIdx.x;
  if(i < n - 1)
    {
      const int lp = (i == 0) ? -1 : oct_level(bprefix[bparent[i]], sbits, levels);
      const int lv = oct_level(bprefix[i], sbits, levels);
      bcount[i]    = lv - lp;
      for(int L = lp + 1; L <= lv; L++)
        atomicAdd(&s_cnt[L], 1);
    }
  __syncthreads();
  if(threadIdx.x <= AG_KEY_LEVELS && s_cnt[threadIdx.x])
    atomicAdd(&lvlcnt[threadIdx.x], s_cnt[threadIdx.x]);
}

// the cubes of each binary node: levels lp+1 .. its own, all over its particle range
__global__ __launch_bounds__(B) void k_oct_emit(int n, const int *__restrict__ bfirst, const int *__restrict__ blast,
                                                const int *__restrict__ bprefix, const int *__restrict__ bparent,
                                                const int *__restrict__ bcount, const int *__restrict__ bbase, int sbits,
                                                int levels, int *tfirst, int *tlast, int *tlevel)
{
  int i = blockIdx.x * B + threadIdx.x;
  int c = (i < n - 1) ? bcount[i] : 0;
  if(c > 0)
    {
      int lp = (i == 0) ? -1 : oct_level(bprefix[bparent[i]], sbits, levels);
      int f = bfirst[i], la = blast[i], b = bbase[i];
      for(int k = 0; k < c; k++)
        {
          tfirst[b + k] = f;
          tlast[b + k]  = la;
          tlevel[b + k] = lp + 1 + k;
        }
    }
}

// Levels at which a cube starts at particle f, from its neighbours' keys alone:
// the level-L cell of f (prefix of sbits + 3L bits) starts at f when f-1 differs
// within that prefix, and is a cube (>= 2 particles) when f+1 shares it.
__global__ void k_lvlmask(const uint64_t *__restrict__ keys, int n, int sbits, int levels, unsigned *lvlmask)
{
  int f = blockIdx.x * B + threadIdx.x;
  if(f >= n)
    return;
  const int a = kdelta(keys, n, f, f - 1), b = kdelta(keys, n, f, f + 1);  // -1 beyond the ends
  const int lo = a < sbits ? 0 : (a - sbits) / 3 + 1;
  const int hi = b < sbits ? -1 : min((b - sbits) / 3, levels);
  unsigned m   = 0;
  for(int L = lo; L <= hi; L++)
    m |= 1u << L;
  lvlmask[f] = m;
}

__global__ void k_single_root(int *tfirst, int *tlast, int *tlevel, unsigned *lvlmask, int *lvlcnt)
{
  tfirst[0]  = 0;
  tlast[0]   = 0;
  tlevel[0]  = 0;
  lvlmask[0] = 1u;
  lvlcnt[0]  = 1;
}

__global__ void k_first_count(const unsigned *__restrict__ lvlmask, int n, int *pcnt)
{
  int f = blockIdx.x * B + threadIdx.x;
  if(f < n)
    pcnt[f] = __popc(lvlmask[f]);
}

__device__ inline int pnext_of(int p, int n, const unsigned *__restrict__ lvlmask, const int *__restrict__ pbase)
{
  int q = p + 1;
  if(q >= n)
    return -1;
  return lvlmask[q] ? n + pbase[q] : q;
}

// place cubes in pre-order and derive geometry and threading
__global__ void k_place(int mt, int n, const int *__restrict__ tfirst, const int *__restrict__ tlast, const int *__restrict__ tlevel,
                        const unsigned *__restrict__ lvlmask, const int *__restrict__ pbase, const double *__restrict__ px,
                        const double *__restrict__ py, const double *__restrict__ pz, BuildGeo geo, int *nfirst,
                        int *nlast, int *nlevel, double *cx, double *cy, double *cz, double *len, int *nextnode, int *sibling,
                        int2 *nrange)
{
  int t = blockIdx.x * B + threadIdx.x;
  if(t >= mt)
    return;
  int f = tfirst[t], la = tlast[t], L = tlevel[t];
  const Domain dom = geo.dom[geo.pset ? geo.pset[f] : 0];
  unsigned mask = lvlmask[f];
  int j         = pbase[f] + __popc(mask & ((1u << L) - 1u));
  nfirst[j]     = f;
  nlast[j]      = la;
  nrange[j]     = make_int2(f, la - f + 1);
  nlevel[j]     = L;
  len[j]        = ldexp(dom.len, -L);
  if(L == 0)
    {
      cx[j] = dom.center[0];
      cy[j] = dom.center[1];
      cz[j] = dom.center[2];
    }
  else
    {
      // forcetree.c:903-906: ((ipos & centermask) | mask) * BigFac + corner
      const int sh       = 52 - L;
      const uint64_t half = 1ull << (51 - L);
      double p[3]        = {px[f], py[f], pz[f]};
      double c[3];
      for(int d = 0; d < 3; d++)
        {
          uint64_t ip = int52(p[d], dom, d);
          c[d]        = (double)(((ip >> sh) << sh) | half) * dom.bigfac + dom.corner[d];
        }
      cx[j] = c[0];
      cy[j] = c[1];
      cz[j] = c[2];
    }
  bool deeper = (mask >> (L + 1)) != 0u;
  nextnode[j] = deeper ? n + j + 1 : f;
  sibling[j]  = pnext_of(la, n, lvlmask, pbase);
}

__global__ void k_pnext(int n, const unsigned *__restrict__ lvlmask, const int *__restrict__ pbase, int *pnext)
{
  int p = blockIdx.x * B + threadIdx.x;
  if(p < n)
    pnext[p] = pnext_of(p, n, lvlmask, pbase);
}

__global__ __launch_bounds__(B) void k_level_scatter(int m, const int *__restrict__ nlevel, const int *__restrict__ lvloff,
                                                     int *lvlfill, int *lvlnodes)
{
  // block-local ranks per level in LDS, one global reservation per block and level
  __shared__ int s_cnt[AG_KEY_LEVELS + 1], s_base[AG_KEY_LEVELS + 1];
  if(threadIdx.x <= AG_KEY_LEVELS)
    s_cnt[threadIdx.x] = 0;
  __syncthreads();
  int j    = blockIdx.x * B + threadIdx.x;
  int L    = (j < m) ? nlevel[j] : -1;
  int rank = (L >= 0) ? atomicAdd(&s_cnt[L], 1) : 0;
  __syncthreads();
  if(threadIdx.x <= AG_KEY_LEVELS)
    s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(&lvlfill[threadIdx.x], s_cnt[threadIdx.x]) : 0;
  __syncthreads();
  if(L >= 0)
    lvlnodes[lvloff[L] + s_base[L] + rank] = j;
}

// ---------------------------------------------------------------------------
// 5. child lists and moments (forcetree.c:1217-1429)
// ---------------------------------------------------------------------------

// CSR child lists, node children first (the group walk pushes them) and then
// particle children (the group walk streams them straight into its tile)
__global__ void k_fill_children(int m, int n, const int *__restrict__ nextnode, const int *__restrict__ sibling,
                                const int *__restrict__ pnext, const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                int *child, int *cnn)
{
  int j = blockIdx.x * B + threadIdx.x;
  if(j >= m)
    return;
  const int end = sibling[j], stop = cbeg[j] + ccnt[j];
  int o = cbeg[j], steps = 0;
  for(int c = nextnode[j]; c != end && c >= 0 && steps < ccnt[j]; steps++)
    {
      if(c >= n && o < stop)
        child[o++] = c;
      c = (c < n) ? pnext[c] : sibling[c - n];
    }
  cnn[j] = o - cbeg[j];
  steps  = 0;
  for(int c = nextnode[j]; c != end && c >= 0 && steps < ccnt[j]; steps++)
    {
      if(c < n && o < stop)
        child[o++] = c;
      c = (c < n) ? pnext[c] : sibling[c - n];
    }
}

// children of each node, by walking its threading from nextnode to sibling
__global__ void k_count_children(int m, int n, const int *__restrict__ nextnode, const int *__restrict__ sibling,
                                 const int *__restrict__ pnext, int *ccnt, int limit, int *err)
{
  int j = blockIdx.x * B + threadIdx.x;
  if(j >= m)
    return;
  int k = 0;
  for(int c = nextnode[j], end = sibling[j]; c != end; k++)
    {
      if(k >= limit || c < 0)
        {
          atomicExch(err, 1);  // broken threading: stop instead of looping
          break;
        }
      c = (c < n) ? pnext[c] : sibling[c - n];
    }
  ccnt[j] = k;
}

// One node's moments from its CSR children (forcetree.c:1217-1429): mass, FP64
// centre of mass (geometric centre when empty) and maxsofttype, plus the two FP32
// walk records.  Children are summed in their fixed list order.
template <bool VOLATILE>
__device__ inline void node_moments(int j, int n, const int *__restrict__ cbeg, const int *__restrict__ ccnt, const int *__restrict__ child,
                                    const double *__restrict__ px, const double *__restrict__ py, const double *__restrict__ pz,
                                    const double *__restrict__ pm, const uint8_t *__restrict__ ps, const double *__restrict__ hsoft,
                                    int nsoft, const double *__restrict__ cx, const double *__restrict__ cy,
                                    const double *__restrict__ cz, const double *__restrict__ len, double *sx, double *sy,
                                    double *sz, double *nm, uint8_t *maxsoft, uint4 *recA, uint4 *recB, const Domain *__restrict__ doms,
                                    const int *__restrict__ nfirst, const int *__restrict__ pset)
{
  const Domain &dom = doms[pset ? pset[nfirst[j]] : 0];
  double m = 0, s0 = 0, s1 = 0, s2 = 0;
  int ms = nsoft;  // sentinel type with zero softening (forcetree.c:1254)
  for(int q = cbeg[j], e = cbeg[j] + ccnt[j]; q < e; q++)
    {
      const int c = child[q];
      double mm, x, y, z;
      int st;
      if(c < n)
        {
          mm = pm[c];
          x  = px[c];
          y  = py[c];
          z  = pz[c];
          st = ps ? ps[c] : 0;
        }
      else
        {
          const int k = c - n;
          if(VOLATILE)
            {
              // written earlier in this launch by another lane: agent-coherent loads
              mm = __hip_atomic_load(&nm[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              x  = __hip_atomic_load(&sx[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              y  = __hip_atomic_load(&sy[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              z  = __hip_atomic_load(&sz[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              st = __hip_atomic_load(&maxsoft[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          else
            {
              mm = nm[k];
              x  = sx[k];
              y  = sy[k];
              z  = sz[k];
              st = maxsoft[k];
            }
        }
      m += mm;
      s0 += mm * x;
      s1 += mm * y;
      s2 += mm * z;
      if(hsoft[ms] < hsoft[st])
        ms = st;
    }
  if(m != 0)
    {
      s0 /= m;
      s1 /= m;
      s2 /= m;
    }
  else
    {
      s0 = cx[j];
      s1 = cy[j];
      s2 = cz[j];
    }
  sx[j]      = s0;
  sy[j]      = s1;
  sz[j]      = s2;
  nm[j]      = m;
  maxsoft[j] = (uint8_t)ms;
  recA[j]    = make_uint4(to_fix(s0, dom, 0), to_fix(s1, dom, 1), to_fix(s2, dom, 2), __float_as_uint((float)m));
  recB[j]    = make_uint4(to_fix(cx[j], dom, 0), to_fix(cy[j], dom, 1), to_fix(cz[j], dom, 2), __float_as_uint((float)len[j]));
}

#define MOMENT_ARGS                                                                                                            \
  n, cbeg, ccnt, child, px, py, pz, pm, ps, hsoft, nsoft, cx, cy, cz, len, sx, sy, sz, nm, maxsoft, recA, recB, domp, nfirst, pset
#define MOMENT_PARAMS                                                                                                          \
  int n, const int *__restrict__ cbeg, const int *__restrict__ ccnt, const int *__restrict__ child, const double *__restrict__ px, \
      const double *__restrict__ py, const double *__restrict__ pz, const double *__restrict__ pm, const uint8_t *__restrict__ ps,  \
      const double *__restrict__ hsoft, int nsoft, const double *__restrict__ cx, const double *__restrict__ cy,                   \
      const double *__restrict__ cz, const double *__restrict__ len, double *sx, double *sy, double *sz, double *nm,               \
      uint8_t *maxsoft, uint4 *recA, uint4 *recB, const Domain *__restrict__ domp, const int *__restrict__ nfirst,       \
      const int *__restrict__ pset

// one level with many nodes: a lane per node
__global__ __launch_bounds__(B) void k_moments_level(const int *__restrict__ list, int cnt, MOMENT_PARAMS)
{
  int t = blockIdx.x * B + threadIdx.x;
  if(t >= cnt)
    return;
  node_moments<false>(list[t], MOMENT_ARGS);
}

// a run of consecutive small levels (deepest first) in one workgroup: levels are
// separated by workgroup barriers, child data are read past L1
__global__ __launch_bounds__(1024) void k_moments_levels_small(const int *__restrict__ list, const int *__restrict__ lvloff, int lhi,
                                                                int llo, MOMENT_PARAMS)
{
  for(int L = lhi; L >= llo; L--)
    {
      for(int t = lvloff[L] + threadIdx.x; t < lvloff[L + 1]; t += blockDim.x)
        node_moments<true>(list[t], MOMENT_ARGS);
      __threadfence_block();
      __syncthreads();
    }
}

// batch: the entry where each set's walk starts (its top cube, or its particle
// when alone) and the entry that follows the set in the threaded order
__global__ void k_set_roots(int nsets, int n, const int64_t *__restrict__ off, const unsigned *__restrict__ lvlmask,
                            const int *__restrict__ pbase, int *root, int *end)
{
  int s = blockIdx.x * B + threadIdx.x;
  if(s >= nsets)
    return;
  const int f = (int)off[s], l = (int)off[s + 1];
  if(l <= f)
    {
      root[s] = -1;
      end[s]  = -1;
      return;
    }
  root[s] = (l - f > 1 && lvlmask[f]) ? n + pbase[f] : f;
  end[s]  = pnext_of(l - 1, n, lvlmask, pbase);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

agrav_status build_tree(agrav_ctx *c)
{
  hipStream_t s   = c->stream;
  const int64_t n = c->n;
  const int N     = (int)n;
  agrav_status st;
  if(n <= 0)
    {
      c->m         = 0;
      c->have_tree = true;
      return AGRAV_OK;
    }
  if(n >= (1ll << 30))
    return fail(c, AGRAV_ERR_CAPACITY, "particle count exceeds 2^30 per context");

  // scratch
  size_t temp = sort_temp_bytes(n);
  size_t t2   = scan_temp_bytes(n) + 4096;
  if(t2 > temp)
    temp = t2;
  if((st = c->d_temp.ensure(temp + 4096)) != AGRAV_OK)
    return st;
  if((st = c->d_dom.ensure(1)) != AGRAV_OK)
    return st;
  if((st = c->d_counters.ensure(64)) != AGRAV_OK)
    return st;

  // 1. domain (one cube, or one per set of a batch)
  const int nsets = c->batch_nsets;
  BuildGeo geo{c->d_dom.p, nullptr, nullptr, 0, AG_KEY_LEVELS};
  unsigned long long mm_init[6] = {~0ull, ~0ull, ~0ull, 0, 0, 0};
  unsigned long long *d_mm      = (unsigned long long *)c->d_counters.p;  // 6 x u64 = 12 ints
  AG_HIP(hipMemcpyAsync(d_mm, mm_init, sizeof(mm_init), hipMemcpyHostToDevice, s));
  if(!c->par.periodic && nsets == 0)
    {
      unsigned g = nblk(n);
      if(g > 1024)
        g = 1024;
      k_extent<<<g, B, 0, s>>>(c->d_pos_in.p, n, d_mm);
    }
  k_domain_finish<<<1, 1, 0, s>>>(d_mm, c->d_dom.p, c->par.periodic, c->par.box_size, n);
  AG_HIP(hipGetLastError());
  if(nsets > 0)
    {
      if((st = c->d_sdom.ensure(nsets)) || (st = c->d_pset.ensure(n)))
        return st;
      int bits = 0;
      while((1ll << bits) < nsets)
        bits++;
      geo.levels  = (63 - bits) / 3;
      geo.sbits   = 63 - 3 * geo.levels;
      geo.dom     = c->d_sdom.p;
      geo.pset_in = c->d_pset_in.p;
      geo.pset    = c->d_pset.p;
      k_set_domains<<<nsets, 256, 0, s>>>(c->d_pos_in.p, c->d_soff.p, c->d_sdom.p);
      AG_HIP(hipGetLastError());
    }
  c->geo_sbits  = geo.sbits;
  c->geo_levels = geo.levels;

  // 2. keys and sort
  if((st = c->d_key.ensure(n)) || (st = c->d_key2.ensure(n)) || (st = c->d_idx.ensure(n)) || (st = c->d_idx2.ensure(n)))
    return st;
  k_keys<<<nblk(n), B, 0, s>>>(c->d_pos_in.p, n, geo, c->par.periodic, c->par.box_size, c->d_key.p, c->d_idx.p, c->d_dom.p);
  AG_HIP(hipGetLastError());
  if((st = radix_sort_u64(c->d_key.p, c->d_idx.p, c->d_key2.p, c->d_idx2.p, n, 63, c->d_temp.p, s)) != AGRAV_OK)
    return st;
  // sorted keys/perm are now in d_key2/d_idx2

  if((st = c->d_px.ensure(n)) || (st = c->d_py.ensure(n)) || (st = c->d_pz.ensure(n)) || (st = c->d_pm.ensure(n)) ||
     (st = c->d_prec.ensure(n)) || (st = c->d_pnext.ensure(n)) || (st = c->d_oldacc_s.ensure(n)))
    return st;
  const bool multisoft = c->soft.nsoft > 1 || c->have_stype;
  if(multisoft && (st = c->d_ps.ensure(n)))
    return st;
  k_gather<<<nblk(n), B, 0, s>>>(c->d_idx2.p, c->d_pos_in.p, c->d_mass_in.p, c->have_stype ? c->d_stype_in.p : nullptr,
                                 c->have_oldacc ? c->d_oldacc.p : nullptr, n, c->par.periodic, c->par.box_size, geo,
                                 nsets > 0 ? c->d_pset.p : nullptr, c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, multisoft ? c->d_ps.p : nullptr, c->d_prec.p,
                                 c->d_oldacc_s.p);
  AG_HIP(hipGetLastError());

  // 3./4. binary tree -> octree cubes
  const int nb = N > 1 ? N - 1 : 1;
  if((st = c->d_bfirst.ensure(nb)) || (st = c->d_blast.ensure(nb)) || (st = c->d_bprefix.ensure(nb)) ||
     (st = c->d_bparent.ensure(nb)) || (st = c->d_bcount.ensure(nb)) || (st = c->d_bbase.ensure(nb)) ||
     (st = c->d_lvlmask.ensure(n)) || (st = c->d_pcnt.ensure(n)) || (st = c->d_pbase.ensure(n)) ||
     (st = c->d_lvlcnt.ensure(AG_KEY_LEVELS + 2)) || (st = c->d_lvloff.ensure(AG_KEY_LEVELS + 2)))
    return st;
  AG_HIP(hipMemsetAsync(c->d_lvlcnt.p, 0, (AG_KEY_LEVELS + 2) * sizeof(int), s));
  int *d_total = c->d_counters.p + 16;
  int mt       = 1;
  for(int L = 0; L <= AG_KEY_LEVELS; L++)
    c->lvlcnt_h[L] = L == 0 ? 1 : 0;  // a single particle: the root alone
  if(N > 1)
    {
      k_binary<<<nblk(N - 1), B, 0, s>>>(c->d_key2.p, N, c->d_bfirst.p, c->d_blast.p, c->d_bprefix.p, c->d_bparent.p);
      k_oct_count<<<nblk(N - 1), B, 0, s>>>(N, c->d_bprefix.p, c->d_bparent.p, geo.sbits, geo.levels, c->d_bcount.p,
                                            c->d_lvlcnt.p);
      AG_HIP(hipGetLastError());
      if((st = exclusive_scan_i32(c->d_bcount.p, c->d_bbase.p, N - 1, d_total, c->d_temp.p, s)) != AGRAV_OK)
        return st;
      // the one host synchronisation of the octree: node count and cubes per level
      AG_HIP(hipMemcpyAsync(&mt, d_total, sizeof(int), hipMemcpyDeviceToHost, s));
      AG_HIP(hipMemcpyAsync(c->lvlcnt_h, c->d_lvlcnt.p, (AG_KEY_LEVELS + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
    }
  // capacity check in the spirit of TreeAllocFactor (forcetree.c:186-228): the
  // exact count is known here, so the node arrays are sized to it.
  if(mt <= 0 || mt > 24 * N + 1)
    return fail(c, AGRAV_ERR_CAPACITY, "octree node count out of range");
  const int M = mt;
  if((st = c->d_tfirst.ensure(M)) || (st = c->d_tlast.ensure(M)) || (st = c->d_tlevel.ensure(M)) ||
     (st = c->d_nfirst.ensure(M)) || (st = c->d_nlast.ensure(M)) || (st = c->d_nlevel.ensure(M)) || (st = c->d_sx.ensure(M)) ||
     (st = c->d_sy.ensure(M)) || (st = c->d_sz.ensure(M)) || (st = c->d_nm.ensure(M)) || (st = c->d_cx.ensure(M)) ||
     (st = c->d_cy.ensure(M)) || (st = c->d_cz.ensure(M)) || (st = c->d_len.ensure(M)) || (st = c->d_nextnode.ensure(M)) ||
     (st = c->d_sibling.ensure(M)) || (st = c->d_maxsoft.ensure(M)) || (st = c->d_recA.ensure(M)) ||
     (st = c->d_recB.ensure(M)) || (st = c->d_nrange.ensure(M)) || (st = c->d_cbeg.ensure(M)) || (st = c->d_ccnt.ensure(M)) || (st = c->d_cnn.ensure(M)) ||
     (st = c->d_child.ensure((size_t)M + n)) || (st = c->d_lvlnodes.ensure(M)))
    return st;
  if(N > 1)
    {
      k_oct_emit<<<nblk(N - 1), B, 0, s>>>(N, c->d_bfirst.p, c->d_blast.p, c->d_bprefix.p, c->d_bparent.p, c->d_bcount.p,
                                           c->d_bbase.p, geo.sbits, geo.levels, c->d_tfirst.p, c->d_tlast.p, c->d_tlevel.p);
      k_lvlmask<<<nblk(N), B, 0, s>>>(c->d_key2.p, N, geo.sbits, geo.levels, c->d_lvlmask.p);
    }
  else
    k_single_root<<<1, 1, 0, s>>>(c->d_tfirst.p, c->d_tlast.p, c->d_tlevel.p, c->d_lvlmask.p, c->d_lvlcnt.p);
  k_first_count<<<nblk(N), B, 0, s>>>(c->d_lvlmask.p, N, c->d_pcnt.p);
  AG_HIP(hipGetLastError());
  if((st = exclusive_scan_i32(c->d_pcnt.p, c->d_pbase.p, N, nullptr, c->d_temp.p, s)) != AGRAV_OK)
    return st;
  k_place<<<nblk(M), B, 0, s>>>(M, N, c->d_tfirst.p, c->d_tlast.p, c->d_tlevel.p, c->d_lvlmask.p, c->d_pbase.p, c->d_px.p,
                                c->d_py.p, c->d_pz.p, geo, c->d_nfirst.p, c->d_nlast.p, c->d_nlevel.p, c->d_cx.p, c->d_cy.p,
                                c->d_cz.p, c->d_len.p, c->d_nextnode.p, c->d_sibling.p, c->d_nrange.p);
  k_pnext<<<nblk(N), B, 0, s>>>(N, c->d_lvlmask.p, c->d_pbase.p, c->d_pnext.p);
  if(nsets > 0)
    {
      if((st = c->d_sroot.ensure(nsets)) || (st = c->d_send.ensure(nsets)))
        return st;
      k_set_roots<<<nblk(nsets), B, 0, s>>>(nsets, N, c->d_soff.p, c->d_lvlmask.p, c->d_pbase.p, c->d_sroot.p, c->d_send.p);
    }
  AG_HIP(hipGetLastError());

  // 5. child lists, then moments level by level (deepest first); runs of small
  //    levels share one single-workgroup launch
  int *d_berr = c->d_counters.p + 50;
  AG_HIP(hipMemsetAsync(d_berr, 0, sizeof(int), s));
  k_count_children<<<nblk(M), B, 0, s>>>(M, N, c->d_nextnode.p, c->d_sibling.p, c->d_pnext.p, c->d_ccnt.p, N + M, d_berr);
  AG_HIP(hipGetLastError());
  if((st = exclusive_scan_i32(c->d_ccnt.p, c->d_cbeg.p, M, nullptr, c->d_temp.p, s)) != AGRAV_OK)
    return st;
  k_fill_children<<<nblk(M), B, 0, s>>>(M, N, c->d_nextnode.p, c->d_sibling.p, c->d_pnext.p, c->d_cbeg.p, c->d_ccnt.p,
                                         c->d_child.p, c->d_cnn.p);
  int off[AG_KEY_LEVELS + 2];
  off[0] = 0;
  for(int L = 0; L <= AG_KEY_LEVELS; L++)
    off[L + 1] = off[L] + c->lvlcnt_h[L];
  if(off[AG_KEY_LEVELS + 1] != M)
    return fail(c, AGRAV_ERR_CAPACITY, "inconsistent octree level counts");
  AG_HIP(hipMemcpyAsync(c->d_lvloff.p, off, (AG_KEY_LEVELS + 2) * sizeof(int), hipMemcpyHostToDevice, s));
  int *d_fill = c->d_counters.p + 24;  // 22 ints
  AG_HIP(hipMemsetAsync(d_fill, 0, (AG_KEY_LEVELS + 1) * sizeof(int), s));
  k_level_scatter<<<nblk(M), B, 0, s>>>(M, c->d_nlevel.p, c->d_lvloff.p, d_fill, c->d_lvlnodes.p);
  AG_HIP(hipGetLastError());
  const uint8_t *d_ps = multisoft ? c->d_ps.p : nullptr;
  const int small = 4096;
  for(int L = AG_KEY_LEVELS; L >= 0;)
    {
      if(c->lvlcnt_h[L] >= small)
        {
          k_moments_level<<<nblk(c->lvlcnt_h[L]), B, 0, s>>>(c->d_lvlnodes.p + off[L], c->lvlcnt_h[L], N, c->d_cbeg.p, c->d_ccnt.p,
                                                              c->d_child.p, c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, d_ps,
                                                              c->d_soft.p, c->soft.nsoft, c->d_cx.p, c->d_cy.p, c->d_cz.p,
                                                              c->d_len.p, c->d_sx.p, c->d_sy.p, c->d_sz.p, c->d_nm.p,
                                                              c->d_maxsoft.p, c->d_recA.p, c->d_recB.p, geo.dom, c->d_nfirst.p,
                                                              geo.pset);
          L--;
        }
      else
        {
          int lo = L;
          while(lo - 1 >= 0 && c->lvlcnt_h[lo - 1] < small)
            lo--;
          k_moments_levels_small<<<1, 1024, 0, s>>>(c->d_lvlnodes.p, c->d_lvloff.p, L, lo, N, c->d_cbeg.p, c->d_ccnt.p,
                                                    c->d_child.p, c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, d_ps, c->d_soft.p,
                                                    c->soft.nsoft, c->d_cx.p, c->d_cy.p, c->d_cz.p, c->d_len.p, c->d_sx.p,
                                                    c->d_sy.p, c->d_sz.p, c->d_nm.p, c->d_maxsoft.p, c->d_recA.p, c->d_recB.p,
                                                    geo.dom, c->d_nfirst.p, geo.pset);
          L = lo - 1;
        }
    }
  AG_HIP(hipGetLastError());

  // out-of-box check (forcetree.c:432-438 aborts; here it is an error code)
  int berr = 0;
  AG_HIP(hipMemcpyAsync(&c->h_dom, c->d_dom.p, sizeof(Domain), hipMemcpyDeviceToHost, s));
  AG_HIP(hipMemcpyAsync(&berr, d_berr, sizeof(int), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(berr)
    return fail(c, AGRAV_ERR_CAPACITY, "octree threading inconsistent (moment pass stopped)");
  if(c->h_dom.out_of_box)
    return fail(c, AGRAV_ERR_OUT_OF_BOX, "particle outside the periodic box after one wrap");
  c->m         = M;
  c->have_tree = true;
  return AGRAV_OK;
}

TreeView tree_view(const agrav_ctx *c)
{
  TreeView v;
  v.n        = (int)c->n;
  v.m        = c->m;
  v.px       = c->d_px.p;
  v.py       = c->d_py.p;
  v.pz       = c->d_pz.p;
  v.pm       = c->d_pm.p;
  v.ps       = (c->soft.nsoft > 1 || c->have_stype) ? c->d_ps.p : nullptr;
  v.pnext    = c->d_pnext.p;
  v.prec     = c->d_prec.p;
  v.sx       = c->d_sx.p;
  v.sy       = c->d_sy.p;
  v.sz       = c->d_sz.p;
  v.nm       = c->d_nm.p;
  v.cx       = c->d_cx.p;
  v.cy       = c->d_cy.p;
  v.cz       = c->d_cz.p;
  v.len      = c->d_len.p;
  v.nextnode = c->d_nextnode.p;
  v.sibling  = c->d_sibling.p;
  v.maxsoft  = c->d_maxsoft.p;
  v.recA     = c->d_recA.p;
  v.recB     = c->d_recB.p;
  v.nrange   = c->d_nrange.p;
  v.cbeg     = c->d_cbeg.p;
  v.ccnt     = c->d_ccnt.p;
  v.cnn      = c->d_cnn.p;
  v.child    = c->d_child.p;
  return v;
}

}  // namespace agrav
