// agrav_prims.hip -- device scan and LSD radix sort used by the tree build.
//
// Both are written for 64-wide CDNA wavefronts: block scans combine per-lane
// serial scans with cross-lane shuffles; the radix sort ranks keys stably
// inside a wavefront with 8 ballots (one per digit bit) and per-wave digit
// counters in LDS, so no key is ever compared against another lane's key.
#include "agrav_internal.h"

namespace agrav
{

static constexpr int kScanBlock = 256;
static constexpr int kScanItems = 8;
static constexpr int kScanTile  = kScanBlock * kScanItems;

__device__ inline int wave_incl_scan(int v)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for(int d = 1; d < 64; d <<= 1)
    {
      int t = __shfl_up(v, d, 64);
      if(lane >= d)
        v += t;
    }
  return v;
}

// exclusive scan of one block's 2048-item tile; returns the tile total
__device__ inline int block_excl_scan_tile(int vals[kScanItems], int *s_wave)
{
  int sum = 0;
#pragma unroll
  for(int i = 0; i < kScanItems; i++)
    {
      int v   = vals[i];
      vals[i] = sum;
      sum += v;
    }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl       = wave_incl_scan(sum);
  if(lane == 63)
    s_wave[w] = incl;
  __syncthreads();
  int wbase = 0, total = 0;
#pragma unroll
  for(int k = 0; k < kScanBlock / 64; k++)
    {
      int t = s_wave[k];
      if(k < w)
        wbase += t;
      total += t;
    }
  int excl = wbase + incl - sum;
#pragma unroll
  for(int i = 0; i < kScanItems; i++)
    vals[i] += excl;
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kScanBlock) void k_tile_sums(const int *__restrict__ in, int64_t n, int *__restrict__ sums)
{
  __shared__ int s_wave[kScanBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int vals[kScanItems];
#pragma unroll
  for(int i = 0; i < kScanItems; i++)
    vals[i] = (base + i < n) ? in[base + i] : 0;
  int total = block_excl_scan_tile(vals, s_wave);
  if(threadIdx.x == 0)
    sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const int *__restrict__ in, int *__restrict__ out, int64_t n,
                                                          const int *__restrict__ offsets, int *__restrict__ total_out)
{
  __shared__ int s_wave[kScanBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int vals[kScanItems];
#pragma unroll
  for(int i = 0; i < kScanItems; i++)
    vals[i] = (base + i < n) ? in[base + i] : 0;
  int last_in = (n > 0) ? 0 : 0;
  (void)last_in;
  int total = block_excl_scan_tile(vals, s_wave);
  int off   = offsets ? offsets[blockIdx.x] : 0;
#pragma unroll
  for(int i = 0; i < kScanItems; i++)
    if(base + i < n)
      out[base + i] = vals[i] + off;
  if(total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *total_out = off + total;
}

static int64_t ntiles_of(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

size_t scan_temp_bytes(int64_t n)
{
  size_t bytes = 0;
  int64_t t    = ntiles_of(n);
  while(t > 1)
    {
      bytes += 2 * (size_t)t * sizeof(int) + 256;
      t = ntiles_of(t);
    }
  return bytes + 256;
}

agrav_status exclusive_scan_i32(const int *d_in, int *d_out, int64_t n, int *d_total, void *d_temp, hipStream_t s)
{
  if(n <= 0)
    {
      if(d_total)
        AG_HIP(hipMemsetAsync(d_total, 0, sizeof(int), s));
      return AGRAV_OK;
    }
  int64_t nt = ntiles_of(n);
  if(nt == 1)
    {
      k_tile_scan<<<1, kScanBlock, 0, s>>>(d_in, d_out, n, nullptr, d_total);
      AG_HIP(hipGetLastError());
      return AGRAV_OK;
    }
  int *sums    = (int *)d_temp;
  int *offsets = sums + nt;
  char *rest   = (char *)(offsets + nt);
  rest         = (char *)(((uintptr_t)rest + 255) & ~(uintptr_t)255);
  k_tile_sums<<<(unsigned)nt, kScanBlock, 0, s>>>(d_in, n, sums);
  AG_HIP(hipGetLastError());
  agrav_status st = exclusive_scan_i32(sums, offsets, nt, nullptr, rest, s);
  if(st != AGRAV_OK)
    return st;
  k_tile_scan<<<(unsigned)nt, kScanBlock, 0, s>>>(d_in, d_out, n, offsets, d_total);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------

static constexpr int kSortBlock = 256;
static constexpr int kSortWaves = kSortBlock / 64;
static constexpr int kSortRounds = 8;  // keys per lane
static constexpr int kSortTile  = kSortBlock * kSortRounds;
static_assert(kSortBlock == 256, "the scatter gives each thread one of the 256 digits");

__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                           int ntiles, int *__restrict__ hist)
{
  __shared__ int s_h[256];
  s_h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSortTile;
  for(int i = threadIdx.x; i < kSortTile; i += kSortBlock)
    {
      int64_t k = base + i;
      if(k < n)
        atomicAdd(&s_h[(keys[k] >> shift) & 255], 1);
    }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = s_h[threadIdx.x];
}

// One tile of 2048 keys: rank stably by the current digit, stage the tile in LDS
// in digit order, then write each digit's run to its global position with
// consecutive lanes on consecutive addresses (coalesced stores).
__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                                                              uint64_t *__restrict__ kout, int *__restrict__ vout, int64_t n,
                                                              int shift, int ntiles, const int *__restrict__ offs)
{
  __shared__ uint64_t s_key[kSortTile];
  __shared__ int s_val[kSortTile];
  __shared__ int s_cnt[kSortWaves][256];
  __shared__ int s_dbase[256], s_gbase[256], s_wsum[kSortWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for(int i = lane; i < 256; i += 64)
    s_cnt[w][i] = 0;
  __builtin_amdgcn_wave_barrier();

  const uint64_t lt    = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t tbase  = (int64_t)blockIdx.x * kSortTile;
  const int tcount     = (int)(n - tbase < kSortTile ? n - tbase : kSortTile);
  const int64_t wbase  = tbase + (int64_t)w * (64 * kSortRounds);
  uint64_t key[kSortRounds];
  int val[kSortRounds];
  int pos[kSortRounds];
  int dig[kSortRounds];

  // 1. rank inside the wave: 8 ballots give the lanes with the same digit
#pragma unroll
  for(int r = 0; r < kSortRounds; r++)
    {
      int64_t k   = wbase + r * 64 + lane;
      bool valid  = k < n;
      key[r]      = valid ? kin[k] : 0;
      val[r]      = valid ? vin[k] : 0;
      int d       = (int)((key[r] >> shift) & 255);
      dig[r]      = valid ? d : -1;
      uint64_t m  = __ballot(valid);
#pragma unroll
      for(int b = 0; b < 8; b++)
        {
          uint64_t bal = __ballot(valid && ((d >> b) & 1));
          m &= ((d >> b) & 1) ? bal : ~bal;
        }
      int before = 0;
      if(valid)
        before = s_cnt[w][d];
      __builtin_amdgcn_wave_barrier();
      pos[r] = before + __popcll(m & lt);
      if(valid && (m & lt) == 0)
        s_cnt[w][d] = before + __popcll(m);
      __builtin_amdgcn_wave_barrier();
    }
  __syncthreads();

  // 2. per digit: wave offsets, tile-local start (scan over digits), global start
  {
    const int d = threadIdx.x;  // kSortBlock == 256 digits
    int run     = 0;
#pragma unroll
    for(int ww = 0; ww < kSortWaves; ww++)
      {
        int c        = s_cnt[ww][d];
        s_cnt[ww][d] = run;
        run += c;
      }
    const int incl = wave_incl_scan(run);
    if(lane == 63)
      s_wsum[w] = incl;
    __syncthreads();
    int wpre = 0;
#pragma unroll
    for(int ww = 0; ww < kSortWaves; ww++)
      wpre += ww < w ? s_wsum[ww] : 0;
    s_dbase[d] = wpre + incl - run;
    s_gbase[d] = offs[(int64_t)d * ntiles + blockIdx.x];
  }
  __syncthreads();

  // 3. stage the tile in digit order
#pragma unroll
  for(int r = 0; r < kSortRounds; r++)
    if(dig[r] >= 0)
      {
        const int d  = dig[r];
        const int lr = s_dbase[d] + s_cnt[w][d] + pos[r];
        s_key[lr]    = key[r];
        s_val[lr]    = val[r];
      }
  __syncthreads();

  // 4. runs of equal digits go out to consecutive global addresses
  for(int i = threadIdx.x; i < tcount; i += kSortBlock)
    {
      const uint64_t k = s_key[i];
      const int d      = (int)((k >> shift) & 255);
      const int64_t dst = (int64_t)s_gbase[d] + (i - s_dbase[d]);
      kout[dst]        = k;
      vout[dst]        = s_val[i];
    }
}

size_t sort_temp_bytes(int64_t n)
{
  int64_t nt = (n + kSortTile - 1) / kSortTile;
  size_t h   = (size_t)256 * nt * sizeof(int);
  return 2 * h + scan_temp_bytes(256 * nt) + 1024;
}

agrav_status radix_sort_u64(uint64_t *keys_in, int *vals_in, uint64_t *keys_out, int *vals_out, int64_t n, int key_bits,
                            void *d_temp, hipStream_t s)
{
  if(n <= 0)
    return AGRAV_OK;
  int ntiles   = (int)((n + kSortTile - 1) / kSortTile);
  int *hist    = (int *)d_temp;
  int *offs    = hist + (size_t)256 * ntiles;
  char *stemp  = (char *)(offs + (size_t)256 * ntiles);
  stemp        = (char *)(((uintptr_t)stemp + 255) & ~(uintptr_t)255);
  int passes   = (key_bits + 7) / 8;
  uint64_t *ka = keys_in, *kb = keys_out;
  int *va = vals_in, *vb = vals_out;
  for(int p = 0; p < passes; p++)
    {
      int shift = 8 * p;
      k_radix_hist<<<ntiles, kSortBlock, 0, s>>>(ka, n, shift, ntiles, hist);
      AG_HIP(hipGetLastError());
      agrav_status st = exclusive_scan_i32(hist, offs, (int64_t)256 * ntiles, nullptr, stemp, s);
      if(st != AGRAV_OK)
        return st;
      k_radix_scatter<<<ntiles, kSortBlock, 0, s>>>(ka, va, kb, vb, n, shift, ntiles, offs);
      AG_HIP(hipGetLastError());
      uint64_t *tk = ka;
      ka           = kb;
      kb           = tk;
      int *tv      = va;
      va           = vb;
      vb           = tv;
    }
  if(ka != keys_out)
    {
      AG_HIP(hipMemcpyAsync(keys_out, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
      AG_HIP(hipMemcpyAsync(vals_out, va, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
  return AGRAV_OK;
}

}  // namespace agrav
