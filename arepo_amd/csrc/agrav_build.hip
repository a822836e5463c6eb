// agrav_build.hip -- octree construction (construct_forcetree, forcetree.c:110-1429).
//
// AREPO inserts particles one by one into a pointer octree (forcetree.c:835-938)
// and then computes moments with a recursive depth-first pass (:1217-1429).
// Here the same octree -- every cube that holds two or more particles, single
// particles hanging in the child slots -- is derived in parallel from sorted
// Peano-Hilbert keys:
//   1. domain cube (domain_box.c:107-197) and keys from AREPO's 52-bit integer
//      coordinates (forcetree.c:493-505): 30 levels (MAX_TREE_LEVEL,
//      forcetree.h:106) in two 63-bit words;
//   2. LSD radix sort of the first words (Hilbert order, like domain.c:140), then
//      runs of equal first words ordered by the second words;
//   3. the levels at which cubes start at each particle, from its neighbours'
//      keys; cubes are placed in pre-order (first particle, then level) so
//      AREPO's nextnode/sibling threading (forcetree.c:1288-1330) follows from
//      array positions;
//   4. moments of every cube at once from double-double prefix sums over the
//      Hilbert-ordered particles.
// Cube centres and lengths use AREPO's integer formula (forcetree.c:903-906),
// so node geometry is bit-identical to the reference's.  Cells of level 30 that
// still hold several particles become buckets whose particles are direct
// children (AREPO goes on putting them into free slots, forcetree.c:861-877).
#include "agrav_ctx.h"

#include <algorithm>

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

// batch: root cubes given per set {centre[3], len}, as SUBFIND keeps one local-tree
// cube per group for every unbinding pass (subfind_loctree.c:101-206)
__global__ void k_set_domains_fixed(const double *__restrict__ cube, int nsets, Domain *sdom)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if(s >= nsets)
    return;
  Domain dm{};
  const double len = cube[4 * s + 3];
  for(int d = 0; d < 3; d++)
    {
      dm.center[d]  = cube[4 * s + d];
      dm.corner[d]  = cube[4 * s + d] - 0.5 * len;
      dm.fix_org[d] = dm.corner[d];
      dm.xmin[d]    = dm.corner[d];
      dm.xmax[d]    = dm.corner[d] + len;
    }
  dm.len      = len;
  dm.inv_len  = 1.0 / len;
  dm.bigfac   = len / (double)(1ll << 52);
  dm.fix_unit = len / 2147483648.0;
  dm.fix_inv  = 1.0 / dm.fix_unit;
  sdom[s]     = dm;
}

// ---------------------------------------------------------------------------
// 2. keys (forcetree.c:394-417 wrap, :493-505 integer coordinates)
// ---------------------------------------------------------------------------

__device__ inline uint64_t int52(double x, const Domain &dom, int d)
{
  double v = (x - dom.corner[d]) * dom.inv_len + 1.0;
  return (uint64_t)__double_as_longlong(v) & 0xFFFFFFFFFFFFFull;
}

// Keys of 126 bits in two 63-bit words {keys, keys_lo}: the set index in the top
// sbits bits, then `levels` Peano-Hilbert digits (3 bits each, coarsest first),
// then zeros.  The octree splits cells down to `levels` (AREPO's insertion tree
// splits down to MAX_TREE_LEVEL, forcetree.c:861-877); only particles sharing a
// cell below the first word need the second one.
__constant__ uint8_t c_hilbert_tab[192] = AG_HILBERT_TABLE;

__global__ void k_keys(const double *__restrict__ pos, const int *__restrict__ src, int64_t n, BuildGeo geo, int periodic,
                       double box, uint64_t *__restrict__ keys, uint64_t *__restrict__ keys_lo, int *__restrict__ idx, Domain *dom_w)
{
  __shared__ uint8_t s_tab[192];  // the curve's state table, read once per level per particle
  if(threadIdx.x < 192)
    s_tab[threadIdx.x] = c_hilbert_tab[threadIdx.x];
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k >= n)
    return;
  const int64_t i = src ? (int64_t)src[k] : k;  // caller index (src: a partial tree's checked list)
  const int set    = geo.pset_in ? geo.pset_in[i] : 0;
  const Domain dom = geo.dom[set];
  const int L      = geo.levels;
  uint32_t c[3];
  bool bad = false;
  for(int d = 0; d < 3; d++)
    {
      double x = pos[3 * i + d];
      if(periodic)
        x = wrap_box(x, box);
      if(!(x >= dom.corner[d] && x < dom.corner[d] + dom.len))
        bad = true;
      c[d] = (uint32_t)(int52(x, dom, d) >> (52 - L));
    }
  if(bad)
    {
      atomicExch(&dom_w->out_of_box, 1);
      c[0] = c[1] = c[2] = 0;
    }
  uint64_t top, low;
  hilbert_digits(c[0], c[1], c[2], L, s_tab, top, low);
  const int nlow = L < AG_KEY_LEVELS ? L : AG_KEY_LEVELS;
  unsigned __int128 key = ((unsigned __int128)(unsigned)set << (3 * L)) | ((unsigned __int128)top << (3 * nlow)) | low;
  key <<= (126 - geo.sbits - 3 * L);
  keys[k]    = (uint64_t)(key >> 63);
  keys_lo[i] = (uint64_t)key & ((1ull << 63) - 1);  // by caller index: the tie-run sorts read keys_lo[idx[.]]
  idx[k]     = (int)i;
}

// Tie runs.  The radix sort orders the first key words from bit tshift up (tshift = 0:
// all of them; tshift = AG_TSHIFT for one set, whose second words have their low 36
// bits zero).  A run of keys equal in those bits is then ordered by a composite secondary
// key: the first word's low tshift bits on top of the second word shifted down by
// tshift, which loses nothing; after the run is sorted both words are written back.
__device__ inline uint64_t tie_sec(uint64_t k1, uint64_t klo, int tshift)
{
  return tshift ? (((k1 & ((1ull << tshift) - 1)) << (63 - tshift)) | (klo >> tshift)) : klo;
}

// the first word (from the run's shared high bits) and the second word of a composite key
__device__ inline void tie_unpack(uint64_t sec, uint64_t k1_run, int tshift, uint64_t &k1, uint64_t &klo)
{
  if(tshift)
    {
      k1  = (k1_run & ~((1ull << tshift) - 1)) | (sec >> (63 - tshift));
      klo = (sec << tshift) & ((1ull << 63) - 1);
    }
  else
    {
      k1  = k1_run;
      klo = sec;
    }
}

// runs of first key words equal from bit tshift up in the sorted order (particles
// sharing a cell at that depth): the start of each run.  Each wavefront takes `span`
// consecutive keys (a multiple of 64) and reserves the slots of all its run starts with
// one atomic: 64 keys fill the device at a few million keys, while at 1e8 keys one
// atomic per 64 serialised on the counter for 5.7 ms (1024: 0.2 ms).
static inline int tie_span(int64_t n) { return n < (1 << 23) ? 64 : 1024; }
__global__ void k_tie_runs(const uint64_t *__restrict__ keys, int n, int tshift, int span, int *runs, int *nruns)
{
  const int lane   = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)blockIdx.x * (B / 64) + (threadIdx.x >> 6)) * span;
  const uint64_t lt = (1ull << lane) - 1;
  auto head = [&](int64_t k) {
    return k + 1 < n && (keys[k] >> tshift) == (keys[k + 1] >> tshift) && (k == 0 || (keys[k - 1] >> tshift) != (keys[k] >> tshift));
  };
  int cnt = 0;
  for(int j = 0; j < span; j += 64)
    cnt += __popcll(__ballot(head(w0 + j + lane)));
  if(cnt == 0)
    return;
  int base = 0;
  if(lane == 0)
    base = atomicAdd(nruns, cnt);
  base = __shfl(base, 0, 64);
  for(int j = 0; j < span; j += 64)
    {
      const bool h     = head(w0 + j + lane);
      const uint64_t m = __ballot(h);
      if(h)
        runs[base + __popcll(m & lt)] = (int)(w0 + j + lane);
      base += __popcll(m);
    }
}

__device__ inline bool tie_less(uint64_t ka, int va, uint64_t kb, int vb);

// a sorted value read as a caller index: a sort whose look-back timed out (the build
// reports it) may leave entries that are not indices; they read index 0, so that no
// read leaves the arrays
__device__ inline int caller_index(int i, int nall) { return (unsigned)i < (unsigned)nall ? i : 0; }

// Short runs (the usual case: a few particles sharing a deep cell), one thread
// each: the run's second key words, then an insertion sort by (second word,
// caller index) in place.  Runs longer than kShortRun go to the workgroup sort.
static constexpr int kShortRun = 32;
__global__ void k_tie_sort_short(uint64_t *keys, const uint64_t *__restrict__ keys_lo, int n, int nall, int tshift,
                                 const int *__restrict__ runs, const int *__restrict__ nruns, int *idx, uint64_t *lo_s,
                                 int *long_runs, int *nlong)
{
  const int nr = *nruns;
  for(int r = blockIdx.x * B + threadIdx.x; r < nr; r += gridDim.x * B)
    {
      const int k0      = runs[r];
      const uint64_t kr = keys[k0];
      int e             = k0 + 1;
      while(e < n && e - k0 <= kShortRun && (keys[e] >> tshift) == (kr >> tshift))
        e++;
      if(e - k0 > kShortRun)
        {
          long_runs[atomicAdd(nlong, 1)] = k0;
          continue;
        }
      for(int t = k0; t < e; t++)
        lo_s[t] = tie_sec(keys[t], keys_lo[caller_index(idx[t], nall)], tshift);
      for(int t = k0 + 1; t < e; t++)
        {
          const uint64_t kk = lo_s[t];
          const int vv      = idx[t];
          int j             = t - 1;
          while(j >= k0 && tie_less(kk, vv, lo_s[j], idx[j]))
            {
              lo_s[j + 1] = lo_s[j];
              idx[j + 1]  = idx[j];
              j--;
            }
          lo_s[j + 1] = kk;
          idx[j + 1]  = vv;
        }
      if(tshift)
        for(int t = k0; t < e; t++)
          {
            uint64_t k1, klo;
            tie_unpack(lo_s[t], kr, tshift, k1, klo);
            keys[t] = k1;
            lo_s[t] = klo;
          }
    }
}

__device__ inline bool tie_less(uint64_t ka, int va, uint64_t kb, int vb) { return ka < kb || (ka == kb && va < vb); }

// Sort every run by its second key words, ties by caller index: the order a stable
// sort of the full 126-bit keys gives.  One workgroup per run; a bitonic network
// whose every compare-exchange puts the smaller pair first, so that the padding
// to a power of two (virtual +inf entries) never moves; in LDS up to 1024
// entries, in place in global memory beyond.  Also the sorted second words.
__global__ __launch_bounds__(256) void k_tie_sort(uint64_t *keys, const uint64_t *__restrict__ keys_lo, int n, int nall, int tshift,
                                                  const int *__restrict__ runs, const int *__restrict__ nruns, int *idx,
                                                  uint64_t *lo_s)
{
  __shared__ uint64_t s_k[1024];
  __shared__ int s_v[1024];
  __shared__ int s_end;
  const int nr = *nruns;
  for(int r = blockIdx.x; r < nr; r += gridDim.x)
    {
      const int k0      = runs[r];
      const uint64_t kr = keys[k0];
      if(threadIdx.x == 0)
        {
          int e = k0 + 1;
          while(e < n && (keys[e] >> tshift) == (kr >> tshift))
            e++;
          s_end = e;
        }
      __syncthreads();
      const int len = s_end - k0;
      int P         = 1;
      while(P < len)
        P <<= 1;
      const bool lds = P <= 1024;
      uint64_t *gk   = lo_s + k0;
      int *gv        = idx + k0;
      for(int t = threadIdx.x; t < P; t += 256)
        {
          const uint64_t kk = t < len ? tie_sec(keys[k0 + t], keys_lo[caller_index(idx[k0 + t], nall)], tshift) : ~0ull;
          if(lds)
            {
              s_k[t] = kk;
              s_v[t] = t < len ? idx[k0 + t] : 0x7fffffff;
            }
          else if(t < len)
            gk[t] = kk;
        }
      __syncthreads();
      uint64_t *K = lds ? s_k : gk;
      int *V      = lds ? s_v : gv;
      for(int size = 2; size <= P; size <<= 1)
        for(int j = size >> 1; j > 0; j >>= 1)
          {
            for(int t = threadIdx.x; t < (P >> 1); t += 256)
              {
                // first step of each merge: mirror pairs inside blocks of `size`; then pairs at distance j
                int a, b;
                if(j == (size >> 1))
                  {
                    const int half = size >> 1, blk = t / half, o = t % half;
                    a = blk * size + o;
                    b = blk * size + size - 1 - o;
                  }
                else
                  {
                    a = 2 * j * (t / j) + (t % j);
                    b = a + j;
                  }
                if(b < (lds ? P : len))
                  {
                    const uint64_t ka = K[a], kb = K[b];
                    const int va = V[a], vb = V[b];
                    if(tie_less(kb, vb, ka, va))
                      {
                        K[a] = kb;
                        K[b] = ka;
                        V[a] = vb;
                        V[b] = va;
                      }
                  }
              }
            __syncthreads();
          }
      if(lds)
        for(int t = threadIdx.x; t < len; t += 256)
          {
            gk[t] = s_k[t];
            gv[t] = s_v[t];
          }
      __syncthreads();
      if(tshift)
        for(int t = threadIdx.x; t < len; t += 256)
          {
            uint64_t k1, klo;
            tie_unpack(gk[t], kr, tshift, k1, klo);
            keys[k0 + t] = k1;
            gk[t]        = klo;
          }
      __syncthreads();
    }
}

__global__ void k_gather(const int *__restrict__ idx, const double *__restrict__ pos, const double *__restrict__ mass,
                         const uint8_t *__restrict__ stype, const float *__restrict__ oldacc, int64_t n, int64_t nall, int periodic,
                         double box, BuildGeo geo, int *pset_sorted, double *px, double *py, double *pz, double *pm, uint8_t *ps,
                         uint4 *prec, float *oldacc_s)
{
  int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k >= n)
    return;
  const int i = caller_index(idx[k], (int)nall);
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
// 3./4. octree cubes from the sorted keys
// ---------------------------------------------------------------------------

// number of equal leading bits of the 126-bit keys of sorted particles i and j
// (each word uses bits 62..0; second words are read only where the first ones
// are equal, i.e. inside a tie run); identical keys count as longer than any
// key prefix; -1 beyond the ends
__device__ inline int key_lcp2(const uint64_t *__restrict__ k, const uint64_t *__restrict__ klo, int i, int j)
{
  uint64_t a = k[i], b = k[j];
  if(a != b)
    return __clzll((long long)(a ^ b)) - 1;
  a = klo[i];
  b = klo[j];
  return a == b ? 126 : 63 + __clzll((long long)(a ^ b)) - 1;
}

__device__ inline int kdelta(const uint64_t *__restrict__ k, const uint64_t *__restrict__ klo, int n, int i, int j)
{
  if(j < 0 || j >= n)
    return -1;
  const int l = key_lcp2(k, klo, i, j);
  return l < 126 ? l : 128 + __clz((unsigned)(i ^ j));
}


// Levels at which a cube starts at particle f, from its neighbours' keys alone:
// the level-L cell of f (prefix of sbits + 3L bits) starts at f when f-1 differs
// within that prefix, and is a cube (>= 2 particles, as AREPO splits every cell
// holding two or more) when f+1 shares it.  Also the cubes per particle, and per
// block the cubes per level (ballots, then one row of partial counts per block:
// global atomics on the same few words would serialise).
__global__ __launch_bounds__(B) void k_lvlmask(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ klo, int n, int sbits,
                                               int levels, unsigned *lvlmask, int *pcnt, int *lvlpart)
{
  __shared__ int s_cnt[AG_MAX_LEVELS + 1];
  if(threadIdx.x <= AG_MAX_LEVELS)
    s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int f = blockIdx.x * B + threadIdx.x;
  unsigned m  = 0;
  if(f < n)
    {
      const int a = kdelta(keys, klo, n, f, f - 1), b = kdelta(keys, klo, n, f, f + 1);  // -1 beyond the ends
      const int lo = a < sbits ? 0 : (a - sbits) / 3 + 1;
      const int hi = b < sbits ? -1 : min((b - sbits) / 3, levels);
      // levels lo..hi (0 <= lo <= hi <= 30 when any)
      m          = lo <= hi ? (((2u << hi) - 1u) & ~((1u << lo) - 1u)) : 0u;
      lvlmask[f] = m;
      pcnt[f]    = __popc(m);
    }
  // per level, the wavefront's count: only the levels present in the wavefront
  unsigned u = m;
#pragma unroll
  for(int o = 32; o > 0; o >>= 1)
    u |= (unsigned)__shfl_xor((int)u, o, 64);
  u = (unsigned)__builtin_amdgcn_readfirstlane((int)u);
  while(u)
    {
      const int L = __ffs(u) - 1;
      u &= u - 1;
      const int c = __popcll(__ballot((m >> L) & 1u));
      if((threadIdx.x & 63) == 0)
        atomicAdd(&s_cnt[L], c);
    }
  __syncthreads();
  // level-major rows: k_level_totals reads each level's counts contiguously
  if(threadIdx.x <= AG_MAX_LEVELS)
    lvlpart[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s_cnt[threadIdx.x];
}

// cubes per level: block L sums row L of the partial counts (one per block of k_lvlmask)
__global__ __launch_bounds__(B) void k_level_totals(const int *__restrict__ lvlpart, int nrows, int *lvlcnt)
{
  __shared__ int s_w[B / 64];
  const int L    = blockIdx.x;
  const int *row = lvlpart + (int64_t)L * nrows;
  int c          = 0;
  for(int r0 = threadIdx.x; r0 < nrows; r0 += B * 8)
    {
      int v[8];
#pragma unroll
      for(int u = 0; u < 8; u++)
        v[u] = r0 + u * B < nrows ? row[r0 + u * B] : 0;
#pragma unroll
      for(int u = 0; u < 8; u++)
        c += v[u];
    }
  for(int o = 32; o > 0; o >>= 1)
    c += __shfl_xor(c, o, 64);
  if((threadIdx.x & 63) == 0)
    s_w[threadIdx.x >> 6] = c;
  __syncthreads();
  if(threadIdx.x == 0)
    {
      int t = 0;
      for(int w = 0; w < B / 64; w++)
        t += s_w[w];
      lvlcnt[L] = t;
    }
}

__global__ void k_single_root(unsigned *lvlmask, int *pcnt, int *lvlcnt)
{
  lvlmask[0] = 1u;
  pcnt[0]    = 1;
  lvlcnt[0]  = 1;
}

__device__ inline int pnext_of(int p, int n, const unsigned *__restrict__ lvlmask, const int *__restrict__ pbase)
{
  int q = p + 1;
  if(q >= n)
    return -1;
  return lvlmask[q] ? n + pbase[q] : q;
}


// Place the cubes starting at each particle in pre-order (first particle, then
// level) and derive their geometry and first link.
__global__ void k_place(int n, const unsigned *__restrict__ lvlmask, const int *__restrict__ pbase, const double *__restrict__ px,
                        const double *__restrict__ py, const double *__restrict__ pz, BuildGeo geo, int *nfirst, int *nlevel,
                        double *cx, double *cy, double *cz, double *len, int *nextnode)
{
  const int f = blockIdx.x * B + threadIdx.x;
  if(f >= n)
    return;
  const unsigned mask = lvlmask[f];
  if(!mask)
    return;
  const Domain dom = geo.dom[geo.pset ? geo.pset[f] : 0];
  double ipf[3]    = {px[f], py[f], pz[f]};
  uint64_t ip[3];
  for(int d = 0; d < 3; d++)
    ip[d] = int52(ipf[d], dom, d);
  int j = pbase[f];
  for(unsigned rest = mask; rest; j++)
    {
      const int L = __ffs(rest) - 1;  // shallowest remaining cube first: pre-order
      rest &= rest - 1;
      nfirst[j] = f;
      nlevel[j] = L;
      len[j]    = ldexp(dom.len, -L);
      if(L == 0)
        {
          cx[j] = dom.center[0];
          cy[j] = dom.center[1];
          cz[j] = dom.center[2];
        }
      else
        {
          // forcetree.c:903-906: ((ipos & centermask) | mask) * BigFac + corner
          const int sh        = 52 - L;
          const uint64_t half = 1ull << (51 - L);
          cx[j] = (double)(((ip[0] >> sh) << sh) | half) * dom.bigfac + dom.corner[0];
          cy[j] = (double)(((ip[1] >> sh) << sh) | half) * dom.bigfac + dom.corner[1];
          cz[j] = (double)(((ip[2] >> sh) << sh) | half) * dom.bigfac + dom.corner[2];
        }
      nextnode[j] = rest ? n + j + 1 : f;  // a deeper cube starts at the same particle, or the particle itself
    }
}

// The last particle of every cube: the end of the run of sorted keys that share
// the cube's prefix, by a galloping search from its first particle; then the
// particle range and the sibling link (the entry after the cube).
__global__ void k_cube_last(int m, int n, const uint64_t *__restrict__ keys, const uint64_t *__restrict__ klo,
                            const int *__restrict__ nfirst, const int *__restrict__ nlevel, const unsigned *__restrict__ lvlmask,
                            const int *__restrict__ pbase, int sbits, int *nlast, int2 *nrange, int *sibling)
{
  const int j = blockIdx.x * B + threadIdx.x;
  if(j >= m)
    return;
  const int f = nfirst[j], q = sbits + 3 * nlevel[j];
  int end     = f;
  if(q == 0)
    end = n - 1;  // the root of a single set holds every particle
  else
    {
      int step = 1;
      while(end + step < n && key_lcp2(keys, klo, f, end + step) >= q)
        {
          end += step;
          step <<= 1;
        }
      int hi = end + step < n ? end + step : n;  // first index known outside the cube
      while(hi - end > 1)
        {
          const int mid = (end + hi) >> 1;
          if(key_lcp2(keys, klo, f, mid) >= q)
            end = mid;
          else
            hi = mid;
        }
    }
  nlast[j]   = end;
  nrange[j]  = make_int2(f, end - f + 1);
  sibling[j] = pnext_of(end, n, lvlmask, pbase);
}

__global__ void k_pnext(int n, const unsigned *__restrict__ lvlmask, const int *__restrict__ pbase, int *pnext)
{
  int p = blockIdx.x * B + threadIdx.x;
  if(p < n)
    pnext[p] = pnext_of(p, n, lvlmask, pbase);
}

// ---------------------------------------------------------------------------
// 5. child lists and moments (forcetree.c:1217-1429)
// ---------------------------------------------------------------------------

// CSR child lists, node children first (the group walk pushes them) and then
// particle children (the group walk streams them straight into its tile), and the
// position of every node in its parent's list (npos; -1 for a cube without parent)
// Also the walk's per-node meta record and a copy of the FP32 records of the
// particle children next to their list entries, so that the group walk reads a
// node and its particle children in one memory round trip each.
__global__ void k_fill_children(int m, int n, const int *__restrict__ nextnode, const int *__restrict__ sibling,
                                const int *__restrict__ pnext, const int *__restrict__ cbeg, const int *__restrict__ ccnt,
                                const int2 *__restrict__ nrange, const uint4 *__restrict__ prec, int *child,
                                const int *__restrict__ cnn, int4 *meta, uint4 *cprec, int *npos)
{
  int j = blockIdx.x * B + threadIdx.x;
  if(j >= m)
    return;
  // one walk of the threading: node children from the list's start, particle children
  // after the nn node children (counted by k_count_children), each in threading order
  const int end = sibling[j], beg = cbeg[j], stop = beg + ccnt[j], nn = cnn[j];
  int on = beg, op = beg + nn, steps = 0;
  for(int c = nextnode[j]; c != end && c >= 0 && steps < ccnt[j]; steps++)
    {
      if(c >= n)
        {
          if(on < beg + nn)
            {
              npos[c - n] = on;
              child[on++] = c;
            }
        }
      else if(op < stop)
        {
          cprec[op]   = prec[c];
          child[op++] = c;
        }
      c = (c < n) ? pnext[c] : sibling[c - n];
    }
  const int2 rg = nrange[j];
  meta[j]       = make_int4(beg + nn, op - beg - nn, rg.x, min(nn, 15) | (min(rg.y, 255) << 4));
}

// children of each node, by walking its threading from nextnode to sibling
__global__ void k_count_children(int m, int n, const int *__restrict__ nextnode, const int *__restrict__ sibling,
                                 const int *__restrict__ pnext, int *ccnt, int *cnn, int *npos, int limit, int *err)
{
  int j = blockIdx.x * B + threadIdx.x;
  if(j >= m)
    return;
  npos[j] = -1;  // set for every node with a parent by k_fill_children
  int k = 0, kn = 0;
  for(int c = nextnode[j], end = sibling[j]; c != end; k++)
    {
      if(k >= limit || c < 0)
        {
          atomicExch(err, 1);  // broken threading: stop instead of looping
          break;
        }
      kn += c >= n;
      c = (c < n) ? pnext[c] : sibling[c - n];
    }
  ccnt[j] = k;
  cnn[j]  = kn;  // node children, listed first
}

// Moments (forcetree.c:1217-1429) from prefix sums.  Every cube holds a
// contiguous range of the Hilbert-ordered particles, so its mass and its
// mass-weighted position are differences of two prefix sums over the particles.
// The prefix sums of (m, m x, m y, m z) run in double-double arithmetic (an FP64
// value and its rounding error; products split exactly with an FMA), so they are
// exact to about 2^-104 of the running total; stored with the low part in FP32
// they keep about 2^-77.  A cube's sums taken as differences then match a direct
// FP64 sum over its particles (or its children, as the reference sums them) to
// FP64 rounding, and every cube is done at once, with no pass per level.
struct DD
{
  double hi, lo;
};

__device__ inline DD dd_add(DD a, DD b)
{
  const double s  = a.hi + b.hi;
  const double v  = s - a.hi;
  const double e  = ((a.hi - (s - v)) + (b.hi - v)) + (a.lo + b.lo);  // two-sum error plus the low parts
  const double hi = s + e;
  return DD{hi, e - (hi - s)};
}

struct DD4
{
  DD v[4];  // m, m x, m y, m z
};

__device__ inline DD4 dd4_add(const DD4 &a, const DD4 &b)
{
  DD4 r;
#pragma unroll
  for(int k = 0; k < 4; k++)
    r.v[k] = dd_add(a.v[k], b.v[k]);
  return r;
}

__device__ inline DD4 dd4_zero()
{
  DD4 r;
#pragma unroll
  for(int k = 0; k < 4; k++)
    r.v[k] = DD{0, 0};
  return r;
}

__device__ inline DD4 particle_dd4(double m, double x, double y, double z)
{
  DD4 r;
  r.v[0]         = DD{m, 0};
  const double q[3] = {x, y, z};
#pragma unroll
  for(int k = 0; k < 3; k++)
    {
      const double p = m * q[k];
      r.v[1 + k]     = DD{p, fma(m, q[k], -p)};  // exact product
    }
  return r;
}

static constexpr int kMomItems = 4;
static constexpr int kMomTile  = B * kMomItems;

// The prefix sums restart at every set of a batch, over tiles aligned to the set's
// first particle, and every step below depends only on the set's own particles:
// a set's moments come out bitwise the same in a batch as in a solve of that set
// alone.  Entries: set s holds its n_s + 1 prefix values at [off_s + s, off_(s+1) + s].
struct MomTiles
{
  const int4 *tab;   // {first particle, count, set, last tile of its set}; null: one set of n particles
  const int *tbeg;   // first tile of every set (nsets + 1), or null
  int n;
};

__device__ inline int4 mom_tile(const MomTiles &mt, int t, int ntiles)
{
  if(mt.tab)
    return mt.tab[t];
  const int st = t * kMomTile;
  return make_int4(st, min(kMomTile, mt.n - st), 0, t == ntiles - 1);
}

// exclusive scan of one value per thread over the block (Hillis-Steele in LDS)
__device__ inline DD4 block_excl_scan_dd4(DD4 v, DD4 *s_buf)
{
  const int t = threadIdx.x;
  s_buf[t]    = v;
  __syncthreads();
  for(int d = 1; d < B; d <<= 1)
    {
      DD4 o = t >= d ? s_buf[t - d] : dd4_zero();
      __syncthreads();
      if(t >= d)
        s_buf[t] = dd4_add(o, s_buf[t]);
      __syncthreads();
    }
  const DD4 r = t > 0 ? s_buf[t - 1] : dd4_zero();
  __syncthreads();
  return r;
}

__device__ inline DD4 dd4_shfl_up(const DD4 &a, int d)
{
  DD4 r;
#pragma unroll
  for(int k = 0; k < 4; k++)
    r.v[k] = DD{__shfl_up(a.v[k].hi, d, 64), __shfl_up(a.v[k].lo, d, 64)};
  return r;
}

__device__ inline DD4 dd4_shfl_xor(const DD4 &a, int d)
{
  DD4 r;
#pragma unroll
  for(int k = 0; k < 4; k++)
    r.v[k] = DD{__shfl_xor(a.v[k].hi, d, 64), __shfl_xor(a.v[k].lo, d, 64)};
  return r;
}

// exclusive scan of one value per thread over the block: inclusive scans inside each
// wavefront by shuffles, then the wavefront totals in order through LDS (one barrier)
__device__ inline DD4 block_excl_scan_dd4_waves(DD4 v, DD4 *s_wave)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  DD4 inc = v;
#pragma unroll
  for(int d = 1; d < 64; d <<= 1)
    {
      const DD4 o = dd4_shfl_up(inc, d);
      if(lane >= d)
        inc = dd4_add(o, inc);
    }
  if(lane == 63)
    s_wave[w] = inc;
  __syncthreads();
  DD4 base = dd4_zero();
  for(int k = 0; k < w; k++)
    base = dd4_add(base, s_wave[k]);
  DD4 ex = dd4_shfl_up(inc, 1);
  if(lane == 0)
    ex = dd4_zero();
  return dd4_add(base, ex);
}

// 1. per tile: the sums of its (m, m x, m y, m z): coalesced reads, then a fixed-order
//    reduction (wave butterflies, then the wavefronts in order)
__global__ __launch_bounds__(B) void k_mom_tile_sums(MomTiles mt, const double *__restrict__ px, const double *__restrict__ py,
                                                     const double *__restrict__ pz, const double *__restrict__ pm, DD4 *tsum)
{
  __shared__ DD4 s_wave[B / 64];
  const int4 tl = mom_tile(mt, blockIdx.x, gridDim.x);
  DD4 acc       = dd4_zero();
  for(int u = 0; u < kMomItems; u++)
    {
      const int l = threadIdx.x + u * B;
      if(l < tl.y)
        {
          const int i = tl.x + l;
          acc         = dd4_add(acc, particle_dd4(pm[i], px[i], py[i], pz[i]));
        }
    }
#pragma unroll
  for(int d = 32; d > 0; d >>= 1)
    acc = dd4_add(acc, dd4_shfl_xor(acc, d));
  if((threadIdx.x & 63) == 0)
    s_wave[threadIdx.x >> 6] = acc;
  __syncthreads();
  if(threadIdx.x == 0)
    {
      DD4 t = s_wave[0];
      for(int k = 1; k < B / 64; k++)
        t = dd4_add(t, s_wave[k]);
      tsum[blockIdx.x] = t;
    }
}

// 2. per set (one block each): exclusive prefix of its tile sums, in place, and the
//    set's total; an empty set writes its single (zero) prefix entry
__global__ __launch_bounds__(B) void k_mom_set_scan(MomTiles mt, int ntiles, const int64_t *__restrict__ soff, DD4 *tsum,
                                                    DD4 *settot, double4 *phi, float4 *plo)
{
  __shared__ DD4 s_wave[B / 64];
  const int set = blockIdx.x;
  const int t0 = mt.tbeg ? mt.tbeg[set] : 0, t1 = mt.tbeg ? mt.tbeg[set + 1] : ntiles;
  const int nt = t1 - t0, per = (nt + B - 1) / B, b0 = t0 + threadIdx.x * per;
  const int e0 = b0 + per < t1 ? b0 + per : t1;  // this thread's tiles [b0, e0)
  // runs of up to 8 tile sums, all read before the first is added (one memory round
  // trip per run instead of one per tile)
  constexpr int kRun = 8;
  DD4 acc = dd4_zero();
  for(int k0 = b0; k0 < e0; k0 += kRun)
    {
      DD4 v[kRun];
#pragma unroll
      for(int u = 0; u < kRun; u++)
        v[u] = k0 + u < e0 ? tsum[k0 + u] : dd4_zero();
#pragma unroll
      for(int u = 0; u < kRun; u++)
        if(k0 + u < e0)
          acc = dd4_add(acc, v[u]);
    }
  DD4 run = block_excl_scan_dd4_waves(acc, s_wave);
  for(int k0 = b0; k0 < e0; k0 += kRun)
    {
      DD4 v[kRun];
#pragma unroll
      for(int u = 0; u < kRun; u++)
        v[u] = k0 + u < e0 ? tsum[k0 + u] : dd4_zero();
#pragma unroll
      for(int u = 0; u < kRun; u++)
        if(k0 + u < e0)
          {
            tsum[k0 + u] = run;
            run          = dd4_add(run, v[u]);
          }
    }
  if(threadIdx.x == B - 1 && settot)
    settot[set] = run;  // the last thread's running value has passed every tile
  if(nt == 0 && threadIdx.x == 0 && soff)
    {
      const int64_t e = soff[set] + set;
      phi[e]          = make_double4(0, 0, 0, 0);
      plo[e]          = make_float4(0, 0, 0, 0);
    }
}

// 2b. batch: exclusive prefix of the set totals (nsets + 1 entries), for the cubes
//     above the set level, which span whole sets
__global__ __launch_bounds__(B) void k_mom_settot_scan(int nsets, const DD4 *__restrict__ settot, DD4 *stp)
{
  __shared__ DD4 s_buf[B];
  const int per = (nsets + B - 1) / B, b0 = threadIdx.x * per;
  DD4 acc       = dd4_zero();
  for(int k = b0; k < b0 + per && k < nsets; k++)
    acc = dd4_add(acc, settot[k]);
  DD4 run = block_excl_scan_dd4(acc, s_buf);
  for(int k = b0; k < b0 + per && k < nsets; k++)
    {
      stp[k] = run;
      run    = dd4_add(run, settot[k]);
    }
  if(threadIdx.x == B - 1)
    stp[nsets] = run;
}

// 3. the prefix before every particle, and each set's total after its last particle:
//    high parts in FP64, low parts in FP32.  The tile passes through LDS: coalesced
//    loads of {m, x, y, z}, a serial run of kMomItems entries per thread with the
//    block scan between the runs, the prefix written over each input slot once it is
//    read, then coalesced stores.  Entry l sits at slot l + l / kMomItems: the pad
//    slot per run spreads the threads' strided slots over the LDS banks.
__device__ inline int mom_slot(int l) { return l + l / kMomItems; }

__global__ __launch_bounds__(B) void k_mom_prefix(MomTiles mt, const double *__restrict__ px, const double *__restrict__ py,
                                                  const double *__restrict__ pz, const double *__restrict__ pm,
                                                  const DD4 *__restrict__ toff, double4 *phi, float4 *plo)
{
  static constexpr int kSlots = kMomTile + kMomTile / kMomItems;
  __shared__ double4 s_q[kSlots];
  __shared__ float4 s_lo[kSlots];
  __shared__ DD4 s_wave[B / 64];
  const int4 tl = mom_tile(mt, blockIdx.x, gridDim.x);
  for(int l = threadIdx.x; l < tl.y; l += B)
    {
      const int i       = tl.x + l;
      s_q[mom_slot(l)]  = make_double4(pm[i], px[i], py[i], pz[i]);
    }
  __syncthreads();
  const int l0 = threadIdx.x * kMomItems;
  DD4 acc      = dd4_zero();
#pragma unroll
  for(int u = 0; u < kMomItems; u++)
    if(l0 + u < tl.y)
      {
        const double4 q = s_q[mom_slot(l0 + u)];
        acc             = dd4_add(acc, particle_dd4(q.x, q.y, q.z, q.w));
      }
  DD4 run = dd4_add(toff[blockIdx.x], block_excl_scan_dd4_waves(acc, s_wave));
#pragma unroll
  for(int u = 0; u < kMomItems; u++)
    {
      const int l = l0 + u;
      if(l < tl.y)
        {
          const int k      = mom_slot(l);
          const double4 q  = s_q[k];
          s_q[k]           = make_double4(run.v[0].hi, run.v[1].hi, run.v[2].hi, run.v[3].hi);
          s_lo[k]          = make_float4((float)run.v[0].lo, (float)run.v[1].lo, (float)run.v[2].lo, (float)run.v[3].lo);
          run              = dd4_add(run, particle_dd4(q.x, q.y, q.z, q.w));
          if(l == tl.y - 1 && tl.w)
            {
              // the set's total, one entry past its last particle
              const int64_t e = (int64_t)tl.x + tl.y + tl.z;
              phi[e]          = make_double4(run.v[0].hi, run.v[1].hi, run.v[2].hi, run.v[3].hi);
              plo[e]          = make_float4((float)run.v[0].lo, (float)run.v[1].lo, (float)run.v[2].lo, (float)run.v[3].lo);
            }
        }
    }
  __syncthreads();
  for(int l = threadIdx.x; l < tl.y; l += B)
    {
      const int64_t e = (int64_t)tl.x + l + tl.z;
      const int k     = mom_slot(l);
      phi[e]          = s_q[k];
      plo[e]          = s_lo[k];
    }
}

// count of particles of one softening type before every particle (input of a scan)
__global__ void k_type_flags(int n, const uint8_t *__restrict__ ps, int t, int *flag)
{
  const int i = blockIdx.x * B + threadIdx.x;
  if(i < n)
    flag[i] = ps[i] == t ? 1 : 0;
}

// 4. every cube: sums as differences of the prefix at its last particle + 1 and at
//    its first; FP64 centre of mass (geometric centre when massless), the softening
//    type with the largest softening among its particles (sentinel nsoft when it has
//    none, forcetree.c:1254), and the two FP32 walk records
__global__ __launch_bounds__(B) void k_moments_prefix(int m, const int2 *__restrict__ nrange, const double4 *__restrict__ phi,
                                                      const float4 *__restrict__ plo, const DD4 *__restrict__ stp,
                                                      const int *__restrict__ tcnt, int np1, int nsoft, const double *__restrict__ hsoft,
                                                      const double *__restrict__ cx, const double *__restrict__ cy,
                                                      const double *__restrict__ cz, const double *__restrict__ len, double *sx,
                                                      double *sy, double *sz, double *nm, uint8_t *maxsoft, uint4 *recA,
                                                      uint4 *recB, const Domain *__restrict__ doms, const int *__restrict__ pset,
                                                      const int *__restrict__ npos, const int4 *__restrict__ meta, int4 *nmeta,
                                                      uint4 *nrecA, uint4 *nrecB)
{
  const int j = blockIdx.x * B + threadIdx.x;
  if(j >= m)
    return;
  const int2 rg = nrange[j];
  const int a = rg.x, b = rg.x + rg.y;
  const int sa = pset ? pset[a] : 0, sb = (pset && rg.y > 0) ? pset[b - 1] : sa;
  DD lo4[4], hi4[4];
  if(sa == sb)
    {
      // a cube inside one set: its entries are shifted by the set index
      const double4 ha = phi[(int64_t)a + sa], hb = phi[(int64_t)b + sa];
      const float4 la = plo[(int64_t)a + sa], lb = plo[(int64_t)b + sa];
      lo4[0] = DD{ha.x, (double)la.x};
      lo4[1] = DD{ha.y, (double)la.y};
      lo4[2] = DD{ha.z, (double)la.z};
      lo4[3] = DD{ha.w, (double)la.w};
      hi4[0] = DD{hb.x, (double)lb.x};
      hi4[1] = DD{hb.y, (double)lb.y};
      hi4[2] = DD{hb.z, (double)lb.z};
      hi4[3] = DD{hb.w, (double)lb.w};
    }
  else
    {
      // a cube above the set level holds whole sets
      const DD4 p0 = stp[sa], p1 = stp[sb + 1];
#pragma unroll
      for(int k = 0; k < 4; k++)
        {
          lo4[k] = p0.v[k];
          hi4[k] = p1.v[k];
        }
    }
  double q[4];
#pragma unroll
  for(int k = 0; k < 4; k++)
    {
      const DD d = dd_add(hi4[k], DD{-lo4[k].hi, -lo4[k].lo});
      q[k]       = d.hi + d.lo;
    }
  const double mass = q[0];
  double s0, s1, s2;
  if(mass != 0)
    {
      s0 = q[1] / mass;
      s1 = q[2] / mass;
      s2 = q[3] / mass;
    }
  else
    {
      s0 = cx[j];
      s1 = cy[j];
      s2 = cz[j];
    }
  int ms = nsoft;
  if(tcnt)
    {
      for(int t = 0; t < nsoft; t++)
        if(tcnt[(int64_t)t * np1 + b] - tcnt[(int64_t)t * np1 + a] > 0 && hsoft[ms] < hsoft[t])
          ms = t;
    }
  else if(rg.y > 0)
    ms = 0;
  const Domain &dom = doms[pset ? pset[a] : 0];
  sx[j]      = s0;
  sy[j]      = s1;
  sz[j]      = s2;
  nm[j]      = mass;
  maxsoft[j] = (uint8_t)ms;
  const uint4 ra = make_uint4(to_fix(s0, dom, 0), to_fix(s1, dom, 1), to_fix(s2, dom, 2), __float_as_uint((float)mass));
  const uint4 rb = make_uint4(to_fix(cx[j], dom, 0), to_fix(cy[j], dom, 1), to_fix(cz[j], dom, 2), __float_as_uint((float)len[j]));
  recA[j]        = ra;
  recB[j]        = rb;
  // the same records at the node's position in its parent's child list (the FP32 group
  // walk pushes child-list positions and reads a popped entry's records there)
  const int p = npos[j];
  if(p >= 0)
    {
      nmeta[p] = meta[j];
      nrecA[p] = ra;
      nrecB[p] = rb;
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

// A partial tree reads positions through the active list, so the list is checked
// before anything indexes with it: every entry in range and listed once.
__global__ void k_check_active(const int *__restrict__ active, int64_t na, int64_t n, int *flag, int *bad)
{
  const int64_t j = (int64_t)blockIdx.x * B + threadIdx.x;
  if(j >= na)
    return;
  const int i = active[j];
  if(i < 0 || i >= n)
    {
      atomicOr(bad, 1);
      return;
    }
  if(atomicExch(&flag[i], 1) != 0)
    atomicOr(bad, 2);
}

// build counters at the start of a build: the three extent minima (u64, ints 0..5)
// start at all ones, the maxima and every other counter (ints 6..63) at zero, and the
// key sort's digit counts (ints kSortHist .. + 2048) at zero
static constexpr int kSortHist = 128;
// offset of the build's read-back in the context's staging block (the walk reads back
// its counters at offset 0)
static constexpr size_t kBuildStage = 1024;
__global__ void k_counters_init(int *cnt)
{
  const int t = threadIdx.x;
  if(t < 64)
    cnt[t] = t < 6 ? -1 : 0;
  for(int q = t; q < 2048; q += blockDim.x)
    cnt[kSortHist + q] = 0;
}

// PM results kept in the sorted order of the tree about to be replaced, back to caller order
__global__ void k_unsort_pm(const int *__restrict__ perm, int64_t n, const double *__restrict__ acc_s,
                            const double *__restrict__ pot_s, double *acc, double *pot)
{
  const int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k >= n)
    return;
  const int64_t i = perm[k];
  acc[3 * i]      = acc_s[3 * k];
  acc[3 * i + 1]  = acc_s[3 * k + 1];
  acc[3 * i + 2]  = acc_s[3 * k + 2];
  pot[i]          = pot_s[k];
}

static agrav_status partial_tree_prepare(agrav_ctx *c)
{
  hipStream_t s    = c->stream;
  const int64_t n  = c->n, na = c->n_active;
  agrav_status st;
  if((st = c->d_counters.ensure(128)) || (st = c->d_aflag.ensure(n ? n : 1)))
    return st;
  int *d_bad = c->d_counters.p + 102;
  AG_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
  if(na > 0)
    {
      AG_HIP(hipMemsetAsync(c->d_aflag.p, 0, n * sizeof(int), s));
      k_check_active<<<nblk(na), B, 0, s>>>(c->d_active.p, na, n, c->d_aflag.p, d_bad);
      AG_HIP(hipGetLastError());
    }
  if(c->have_pm && c->pm_sorted && c->n_tree == n && n > 0)
    {
      if((st = c->d_acc_pm2.ensure(3 * n)) || (st = c->d_pot_pm2.ensure(n)))
        return st;
      k_unsort_pm<<<nblk(n), B, 0, s>>>(c->d_idx2.p, n, c->d_acc_pm.p, c->d_pot_pm.p, c->d_acc_pm2.p, c->d_pot_pm2.p);
      AG_HIP(hipGetLastError());
      c->d_acc_pm.swap_storage(c->d_acc_pm2);
      c->d_pot_pm.swap_storage(c->d_pot_pm2);
      c->pm_sorted = false;
    }
  int bad = 0;
  AG_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(bad & 1)
    return fail(c, AGRAV_ERR_BAD_PARAM, "active index out of range");
  if(bad & 2)
    return fail(c, AGRAV_ERR_BAD_PARAM, "active list repeats a particle");
  return AGRAV_OK;
}

agrav_status build_tree(agrav_ctx *c)
{
  hipStream_t s = c->stream;
  // HIERARCHICAL_GRAVITY's partial tree (forcetree.c:91-95, :166-171): only the active
  // particles are inserted, in a domain cube that still spans every particle
  const bool part    = c->partial_tree && c->n_active >= 0 && c->batch_nsets == 0;
  const int64_t nall = c->n;
  const int64_t n    = part ? c->n_active : nall;
  const int N        = (int)n;
  agrav_status st;
  c->have_tree = false;
  if(part && (st = partial_tree_prepare(c)) != AGRAV_OK)
    return st;
  c->tree_partial        = part;
  c->tree_active_version = c->active_version;
  c->n_tree              = n;
  if(n <= 0)
    {
      c->m         = 0;
      c->have_tree = true;
      AG_HIP(hipEventRecord(c->ev[1], s));  // end of the build (timed by the caller)
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
  // device counters: [0..11] domain extent (6 u64), [12] octree node count, [13] a scan or
  // sort look-back that timed out, [14..35] cubes per level (read back together),
  // [48..55] walk queue and failures, [50] threading error
  if((st = c->d_counters.ensure(kSortHist + 2048)) != AGRAV_OK)
    return st;
  int *d_total = c->d_counters.p + 12, *d_perr = c->d_counters.p + 13, *d_lvlcnt = c->d_counters.p + 14;
  // prefix sums of a build that failed after forking them still read its particles
  if(c->mom_forked)
    {
      AG_HIP(hipStreamWaitEvent(s, c->ev_mjoin, 0));
      c->mom_forked = false;
    }
  // one small kernel sets the extent's start values and clears every other build counter
  unsigned long long *d_mm = (unsigned long long *)c->d_counters.p;  // 6 x u64 = 12 ints
  k_counters_init<<<1, 256, 0, s>>>(c->d_counters.p);

  // 1. domain (one cube, or one per set of a batch)
  const int nsets = c->batch_nsets;
  BuildGeo geo{c->d_dom.p, nullptr, nullptr, 0, AG_MAX_LEVELS};
  if(!c->par.periodic && nsets == 0)
    {
      unsigned g = nblk(nall);
      if(g > 1024)
        g = 1024;
      k_extent<<<g, B, 0, s>>>(c->d_pos_in.p, nall, d_mm);
    }
  k_domain_finish<<<1, 1, 0, s>>>(d_mm, c->d_dom.p, c->par.periodic, c->par.box_size, nall);
  AG_HIP(hipGetLastError());
  if(nsets > 0)
    {
      if((st = c->d_sdom.ensure(nsets)) || (st = c->d_pset.ensure(n)))
        return st;
      int bits = 0;
      while((1ll << bits) < nsets)
        bits++;
      geo.levels  = std::min(AG_MAX_LEVELS, (126 - bits) / 3);
      geo.sbits   = bits;
      geo.dom     = c->d_sdom.p;
      geo.pset_in = c->d_pset_in.p;
      geo.pset    = c->d_pset.p;
      if(c->batch_cubes)
        k_set_domains_fixed<<<(nsets + 255) / 256, 256, 0, s>>>(c->batch_cubes, nsets, c->d_sdom.p);
      else
        k_set_domains<<<nsets, 256, 0, s>>>(c->d_pos_in.p, c->d_soff.p, c->d_sdom.p);
      AG_HIP(hipGetLastError());
    }
  c->geo_sbits  = geo.sbits;
  c->geo_levels = geo.levels;

  // 2. keys and sort
  if((st = c->d_key.ensure(n)) || (st = c->d_key2.ensure(n)) || (st = c->d_idx.ensure(n)) || (st = c->d_idx2.ensure(n)) ||
     (st = c->d_key_lo.ensure(nall)) || (st = c->d_key_lo_s.ensure(n)) || (st = c->d_ties.ensure(2 * (n / 2 + 1))))
    return st;
  k_keys<<<nblk(n), B, 0, s>>>(c->d_pos_in.p, part ? c->d_active.p : nullptr, n, geo, c->par.periodic, c->par.box_size,
                               c->d_key.p, c->d_key_lo.p, c->d_idx.p, c->d_dom.p);
  AG_HIP(hipGetLastError());
  // one set: the first words are sorted from bit 23 up (5 passes instead of 8); runs
  // sharing a level-13 cell are finished by the tie-run sorts (from bit 15: 6 passes,
  // 30 us slower at 2M clustered particles; from bit 31 the runs in the densest cores
  // grow long enough to cost more than the pass saved)
#ifndef AG_TSHIFT
#define AG_TSHIFT 23
#endif
  static_assert(AG_TSHIFT <= 36, "the composite tie key holds the first word's low tshift bits in the second "
                                 "word's zero low bits");
  const int tshift = nsets == 0 ? AG_TSHIFT : 0;
  bool sorted_in_keys = false;
  if((st = radix_sort_u64(c->d_key.p, c->d_idx.p, c->d_key2.p, c->d_idx2.p, n, 63, c->d_temp.p, d_perr, s, tshift,
                          &sorted_in_keys, c->d_counters.p + kSortHist)) != AGRAV_OK)
    return st;
  if(sorted_in_keys)
    {
      // an even pass count ends in the input pair: the buffers trade places instead of a copy
      c->d_key.swap_storage(c->d_key2);
      c->d_idx.swap_storage(c->d_idx2);
    }
  // a full step's walk order is sorted beside the rest of the build (queued here, when the
  // main stream holds enough work that the launches cost the device no time)
  if(c->defer_build_check && !part && nsets == 0 && (st = walk_preorder(c)) != AGRAV_OK)
    return st;
  // sorted first key words and perm are now in d_key2/d_idx2; runs of equal first
  // words are put in the order of their second words
  int *d_nties = c->d_counters.p + 60;  // cleared at the start of the build
  if(n > 1)
    {
      int *d_nlong = c->d_counters.p + 61, *d_long = c->d_ties.p + (n / 2 + 1);
      const int64_t per_block = (int64_t)(B / 64) * tie_span(n);
      k_tie_runs<<<(unsigned)((n + per_block - 1) / per_block), B, 0, s>>>(c->d_key2.p, (int)n, tshift, tie_span(n), c->d_ties.p,
                                                                          d_nties);
      k_tie_sort_short<<<512, B, 0, s>>>(c->d_key2.p, c->d_key_lo.p, (int)n, (int)nall, tshift, c->d_ties.p, d_nties, c->d_idx2.p,
                                         c->d_key_lo_s.p, d_long, d_nlong);
      k_tie_sort<<<256, 256, 0, s>>>(c->d_key2.p, c->d_key_lo.p, (int)n, (int)nall, tshift, d_long, d_nlong, c->d_idx2.p,
                                     c->d_key_lo_s.p);
      AG_HIP(hipGetLastError());
    }

  if((st = c->d_px.ensure(n)) || (st = c->d_py.ensure(n)) || (st = c->d_pz.ensure(n)) || (st = c->d_pm.ensure(n)) ||
     (st = c->d_prec.ensure(n)) || (st = c->d_pnext.ensure(n)) || (st = c->d_oldacc_s.ensure(n)))
    return st;
  // per-particle softening types matter only with several table entries: types are
  // checked against the table, so with one entry every particle has type 0
  const bool multisoft = c->soft.nsoft > 1;
  if(multisoft && (st = c->d_ps.ensure(n)))
    return st;
  k_gather<<<nblk(n), B, 0, s>>>(c->d_idx2.p, c->d_pos_in.p, c->d_mass_in.p, c->have_stype ? c->d_stype_in.p : nullptr,
                                 c->have_oldacc ? c->d_oldacc.p : nullptr, n, nall, c->par.periodic, c->par.box_size, geo,
                                 nsets > 0 ? c->d_pset.p : nullptr, c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, multisoft ? c->d_ps.p : nullptr, c->d_prec.p,
                                 c->d_oldacc_s.p);
  AG_HIP(hipGetLastError());
  // a full step's periodic PM needs only the sorted particles: it runs on the PM stream
  // beside the octree kernels below
  if(c->pm_fork && !part && nsets == 0 && (st = pm_fork(c)) != AGRAV_OK)
    return st;

  // moments of all cubes from prefix sums over the particles, restarted at every set
  MomTiles mtl{nullptr, nullptr, N};
  int ntiles = (N + kMomTile - 1) / kMomTile;
  const int nseg = nsets > 0 ? nsets : 1;
  DD4 *stp        = nullptr;  // prefix sums at the set starts (batches)
  auto moment_prefix = [&](hipStream_t ms) -> agrav_status {
    if(nsets > 0)
      {
        std::vector<int4> tab;
        std::vector<int> tbeg(nsets + 1);
        for(int q = 0; q < nsets; q++)
          {
            tbeg[q]         = (int)tab.size();
            const int64_t f = c->h_soff[q], e = c->h_soff[q + 1];
            for(int64_t st0 = f; st0 < e; st0 += kMomTile)
              tab.push_back(make_int4((int)st0, (int)std::min<int64_t>(kMomTile, e - st0), q, st0 + kMomTile >= e));
          }
        tbeg[nsets] = (int)tab.size();
        ntiles      = (int)tab.size();
        if((st = c->d_mom_tab.ensure(tab.size() ? tab.size() : 1)) || (st = c->d_mom_tbeg.ensure(nsets + 1)))
          return st;
        if(!tab.empty())
          AG_HIP(hipMemcpyAsync(c->d_mom_tab.p, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice, ms));
        AG_HIP(hipMemcpyAsync(c->d_mom_tbeg.p, tbeg.data(), (nsets + 1) * sizeof(int), hipMemcpyHostToDevice, ms));
        AG_HIP(hipStreamSynchronize(ms));  // the host vectors go out of scope
        mtl.tab  = c->d_mom_tab.p;
        mtl.tbeg = c->d_mom_tbeg.p;
      }
    const size_t ntile_bytes = (size_t)(ntiles ? ntiles : 1) * sizeof(DD4), nset_bytes = (size_t)(nseg + 1) * sizeof(DD4);
    if((st = c->d_mom_tiles.ensure(ntile_bytes + 2 * nset_bytes)) || (st = c->d_mom_hi.ensure((size_t)N + nseg)) ||
       (st = c->d_mom_lo.ensure((size_t)N + nseg)))
      return st;
    DD4 *tsum   = reinterpret_cast<DD4 *>(c->d_mom_tiles.p);
    DD4 *settot = reinterpret_cast<DD4 *>(c->d_mom_tiles.p + ntile_bytes);
    stp         = reinterpret_cast<DD4 *>(c->d_mom_tiles.p + ntile_bytes + nset_bytes);
    if(ntiles > 0)
      k_mom_tile_sums<<<ntiles, B, 0, ms>>>(mtl, c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, tsum);
    k_mom_set_scan<<<nseg, B, 0, ms>>>(mtl, ntiles, nsets > 0 ? c->d_soff.p : nullptr, tsum, settot, c->d_mom_hi.p, c->d_mom_lo.p);
    if(nsets > 0)
      k_mom_settot_scan<<<1, B, 0, ms>>>(nsets, settot, stp);
    if(ntiles > 0)
      k_mom_prefix<<<ntiles, B, 0, ms>>>(mtl, c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, tsum, c->d_mom_hi.p, c->d_mom_lo.p);
    AG_HIP(hipGetLastError());
    return AGRAV_OK;
  };
  const bool fork_moments = nsets == 0 && c->mom_stream != nullptr;
  if(fork_moments)
    {
      AG_HIP(hipEventRecord(c->ev_mfork, s));
      AG_HIP(hipStreamWaitEvent(c->mom_stream, c->ev_mfork, 0));
      if((st = moment_prefix(c->mom_stream)) != AGRAV_OK)
        return st;
      AG_HIP(hipEventRecord(c->ev_mjoin, c->mom_stream));
      c->mom_forked = true;
    }

  // 3./4. octree cubes straight from the sorted keys
  if((st = c->d_lvlmask.ensure(n)) || (st = c->d_pcnt.ensure(n)) || (st = c->d_pbase.ensure(n)))
    return st;
  if(N > 1)
    {
      int *lvlpart = c->d_bcnt.p;  // AG_MAX_LEVELS + 1 rows of nblk(N) partial counts
      if((st = c->d_bcnt.ensure((size_t)nblk(N) * (AG_MAX_LEVELS + 1))) != AGRAV_OK)
        return st;
      lvlpart = c->d_bcnt.p;
      k_lvlmask<<<nblk(N), B, 0, s>>>(c->d_key2.p, c->d_key_lo_s.p, N, geo.sbits, geo.levels, c->d_lvlmask.p, c->d_pcnt.p,
                                      lvlpart);
      k_level_totals<<<AG_MAX_LEVELS + 1, B, 0, s>>>(lvlpart, (int)nblk(N), d_lvlcnt);
    }
  else
    k_single_root<<<1, 1, 0, s>>>(c->d_lvlmask.p, c->d_pcnt.p, d_lvlcnt);
  AG_HIP(hipGetLastError());
  if((st = exclusive_scan_i32(c->d_pcnt.p, c->d_pbase.p, N, d_total, c->d_temp.p, d_perr, s)) != AGRAV_OK)
    return st;
  // the one host synchronisation of the octree: node count and cubes per level
  static_assert(14 + AG_MAX_LEVELS + 1 <= 48, "cubes per level fit below the walk counters");
  int *hc = c->stage<int>(0);
  AG_HIP(hipMemcpyAsync(hc, d_total, (2 + AG_MAX_LEVELS + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
  AG_HIP(hipEventRecord(c->ev_count, s));
  // the kernels that need only the sorted particles run while the host waits for the
  // count: the particle threading (and the prefix sums of the moments, unless forked)
  k_pnext<<<nblk(N), B, 0, s>>>(N, c->d_lvlmask.p, c->d_pbase.p, c->d_pnext.p);
  if(!fork_moments && (st = moment_prefix(s)) != AGRAV_OK)
    return st;
  AG_HIP(hipEventSynchronize(c->ev_count));
  for(int L = 0; L <= AG_MAX_LEVELS; L++)
    c->lvlcnt_h[L] = hc[2 + L];
  if(hc[1])
    return fail(c, AGRAV_ERR_HIP, "sort/scan look-back did not complete");
  const int mt = hc[0];
  // capacity check in the spirit of TreeAllocFactor (forcetree.c:186-228): the
  // exact count is known here, so the node arrays are sized to it.  The cubes per
  // level are counted without overflow (each at most N); their 64-bit sum must
  // equal the 32-bit scan total, and nodes and particles share one int32 entry
  // space (entries >= N are nodes).
  int64_t mt64 = 0;
  for(int L = 0; L <= AG_MAX_LEVELS; L++)
    mt64 += c->lvlcnt_h[L];
  if(mt <= 0 || (int64_t)mt != mt64 || mt64 > 24 * (int64_t)N + 1 || mt64 + (int64_t)N >= ((int64_t)1 << 31))
    return fail(c, AGRAV_ERR_CAPACITY, "octree node count out of range");
  const int M = mt;
  if((st = c->d_nfirst.ensure(M)) || (st = c->d_nlast.ensure(M)) || (st = c->d_nlevel.ensure(M)) || (st = c->d_sx.ensure(M)) ||
     (st = c->d_sy.ensure(M)) || (st = c->d_sz.ensure(M)) || (st = c->d_nm.ensure(M)) || (st = c->d_cx.ensure(M)) ||
     (st = c->d_cy.ensure(M)) || (st = c->d_cz.ensure(M)) || (st = c->d_len.ensure(M)) || (st = c->d_nextnode.ensure(M)) ||
     (st = c->d_sibling.ensure(M)) || (st = c->d_maxsoft.ensure(M)) || (st = c->d_recA.ensure(M)) ||
     (st = c->d_recB.ensure(M)) || (st = c->d_nrange.ensure(M)) || (st = c->d_cbeg.ensure(M)) || (st = c->d_ccnt.ensure(M)) ||
     (st = c->d_cnn.ensure(M)) || (st = c->d_child.ensure((size_t)M + n)))
    return st;
  k_place<<<nblk(N), B, 0, s>>>(N, c->d_lvlmask.p, c->d_pbase.p, c->d_px.p, c->d_py.p, c->d_pz.p, geo, c->d_nfirst.p,
                                c->d_nlevel.p, c->d_cx.p, c->d_cy.p, c->d_cz.p, c->d_len.p, c->d_nextnode.p);
  k_cube_last<<<nblk(M), B, 0, s>>>(M, N, c->d_key2.p, c->d_key_lo_s.p, c->d_nfirst.p, c->d_nlevel.p, c->d_lvlmask.p,
                                    c->d_pbase.p, geo.sbits, c->d_nlast.p, c->d_nrange.p, c->d_sibling.p);
  if(nsets > 0)
    {
      if((st = c->d_sroot.ensure(nsets)) || (st = c->d_send.ensure(nsets)))
        return st;
      k_set_roots<<<nblk(nsets), B, 0, s>>>(nsets, N, c->d_soff.p, c->d_lvlmask.p, c->d_pbase.p, c->d_sroot.p, c->d_send.p);
    }
  AG_HIP(hipGetLastError());

  // 5. child lists, then moments level by level (deepest first); runs of small
  //    levels share one single-workgroup launch
  if(scan_temp_bytes(M) + 4096 > c->d_temp.cap && (st = c->d_temp.ensure(scan_temp_bytes(M) + 4096)) != AGRAV_OK)
    return st;
  int *d_berr = c->d_counters.p + 50;  // cleared at the start of the build
  if((st = c->d_npos.ensure(M)))
    return st;
  k_count_children<<<nblk(M), B, 0, s>>>(M, N, c->d_nextnode.p, c->d_sibling.p, c->d_pnext.p, c->d_ccnt.p, c->d_cnn.p,
                                         c->d_npos.p, N + M, d_berr);
  AG_HIP(hipGetLastError());
  if((st = exclusive_scan_i32(c->d_ccnt.p, c->d_cbeg.p, M, nullptr, c->d_temp.p, d_perr, s)) != AGRAV_OK)
    return st;
  if((st = c->d_meta.ensure(M)) || (st = c->d_cprec.ensure((size_t)M + n)))
    return st;
  k_fill_children<<<nblk(M), B, 0, s>>>(M, N, c->d_nextnode.p, c->d_sibling.p, c->d_pnext.p, c->d_cbeg.p, c->d_ccnt.p,
                                         c->d_nrange.p, c->d_prec.p, c->d_child.p, c->d_cnn.p, c->d_meta.p, c->d_cprec.p,
                                         c->d_npos.p);
  int mtot = 0;
  for(int L = 0; L <= AG_MAX_LEVELS; L++)
    mtot += c->lvlcnt_h[L];
  if(mtot != M)
    return fail(c, AGRAV_ERR_CAPACITY, "inconsistent octree level counts");
  int *tcnt = nullptr;
  if(multisoft)
    {
      // per softening type, the count of its particles before every particle
      const int ns = c->soft.nsoft;
      if((st = c->d_tcnt.ensure((size_t)ns * (N + 1))) || (st = c->d_pcnt.ensure(n)))
        return st;
      tcnt = c->d_tcnt.p;
      for(int t = 0; t < ns; t++)
        {
          k_type_flags<<<nblk(N), B, 0, s>>>(N, c->d_ps.p, t, c->d_pcnt.p);
          if((st = exclusive_scan_i32(c->d_pcnt.p, tcnt + (size_t)t * (N + 1), N, tcnt + (size_t)t * (N + 1) + N, c->d_temp.p,
                                      d_perr, s)) != AGRAV_OK)
            return st;
        }
    }
  if((st = c->d_nmeta.ensure((size_t)M + n)) || (st = c->d_nrecA.ensure((size_t)M + n)) ||
     (st = c->d_nrecB.ensure((size_t)M + n)))
    return st;
  if(c->mom_forked)
    {
      AG_HIP(hipStreamWaitEvent(s, c->ev_mjoin, 0));
      c->mom_forked = false;
    }
  k_moments_prefix<<<nblk(M), B, 0, s>>>(M, c->d_nrange.p, c->d_mom_hi.p, c->d_mom_lo.p, stp, tcnt, N + 1, c->soft.nsoft, c->d_soft.p,
                                         c->d_cx.p, c->d_cy.p, c->d_cz.p, c->d_len.p, c->d_sx.p, c->d_sy.p, c->d_sz.p, c->d_nm.p,
                                         c->d_maxsoft.p, c->d_recA.p, c->d_recB.p, geo.dom, geo.pset, c->d_npos.p, c->d_meta.p,
                                         c->d_nmeta.p, c->d_nrecA.p, c->d_nrecB.p);
  AG_HIP(hipGetLastError());

  // end of the build, recorded before the read-back so that timing it needs no second
  // round trip to the device
  AG_HIP(hipEventRecord(c->ev[1], s));
  // error flags and domain (out-of-box: forcetree.c:432-438 aborts; here it is an error
  // code), read back into their own part of the staging block; a deferred check's copies
  // go to the side stream, off the walk's path (the next build and the check wait for them)
  static_assert(kBuildStage + 64 * sizeof(int) + sizeof(Domain) <= agrav_ctx::kStageBytes, "read-backs fit the staging block");
  const bool side_copy = c->defer_build_check && c->mom_stream != nullptr;
  hipStream_t rs       = side_copy ? c->mom_stream : s;
  if(side_copy)
    AG_HIP(hipStreamWaitEvent(rs, c->ev[1], 0));
  AG_HIP(hipMemcpyAsync(c->stage<int>(kBuildStage), c->d_counters.p, 64 * sizeof(int), hipMemcpyDeviceToHost, rs));
  AG_HIP(hipMemcpyAsync(c->stage<Domain>(kBuildStage + 256), c->d_dom.p, sizeof(Domain), hipMemcpyDeviceToHost, rs));
  if(side_copy)
    {
      AG_HIP(hipEventRecord(c->ev_mjoin, rs));
      c->mom_forked = true;
    }
  c->bcopy_side = side_copy;
  c->m                   = M;
  c->have_tree           = true;
  c->build_check_pending = true;
  // a deferred check is read after the next synchronisation (the walk's read-back):
  // the walks test the same flags on the device before they start
  if(c->defer_build_check)
    return AGRAV_OK;
  AG_HIP(hipStreamSynchronize(s));
  return build_check(c);
}

agrav_status build_check(agrav_ctx *c)
{
  if(!c->build_check_pending)
    return AGRAV_OK;
  c->build_check_pending = false;
  if(c->bcopy_side)
    {
      AG_HIP(hipEventSynchronize(c->ev_mjoin));  // the read-back copies on the side stream
      c->bcopy_side = false;
      c->mom_forked = false;  // the side stream's work so far is done: the next build need not wait
    }
  if(c->build_time_pending)
    {
      float ms = 0;
      AG_HIP(hipEventElapsedTime(&ms, c->ev[7], c->ev[1]));
      c->tim.treebuild      = ms;
      c->build_time_pending = false;
    }
  const int *cnow        = c->stage<int>(kBuildStage);
  c->h_dom               = *c->stage<Domain>(kBuildStage + 256);
  const int perr = cnow[13], berr = cnow[50];
  agrav_status st = AGRAV_OK;
  if(berr)
    st = fail(c, AGRAV_ERR_CAPACITY, "octree threading inconsistent (moment pass stopped)");
  else if(perr)
    st = fail(c, AGRAV_ERR_HIP, "scan look-back did not complete");
  else if(c->h_dom.out_of_box)
    st = fail(c, AGRAV_ERR_OUT_OF_BOX, "particle outside the periodic box after one wrap");
  if(st != AGRAV_OK)
    c->have_tree = c->have_tree_force = c->have_final = false;
  return st;
}

TreeView tree_view(const agrav_ctx *c)
{
  TreeView v;
  v.n        = (int)c->n_tree;
  v.m        = c->m;
  v.px       = c->d_px.p;
  v.py       = c->d_py.p;
  v.pz       = c->d_pz.p;
  v.pm       = c->d_pm.p;
  v.ps       = (c->soft.nsoft > 1) ? c->d_ps.p : nullptr;
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
  v.meta     = c->d_meta.p;
  v.cprec    = c->d_cprec.p;
  v.nmeta    = c->d_nmeta.p;
  v.nrecA    = c->d_nrecA.p;
  v.nrecB    = c->d_nrecB.p;
  v.dom      = c->d_dom.p;
  v.cnt      = c->d_counters.p;
  return v;
}

}  // namespace agrav
