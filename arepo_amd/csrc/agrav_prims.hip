// agrav_prims.hip -- device scan and LSD radix sort used by the tree build.
//
// Both are written for 64-wide CDNA wavefronts: block scans combine per-lane
// serial scans with cross-lane shuffles; the radix sort ranks keys stably
// inside a wavefront with 8 ballots (one per digit bit) and per-wave digit
// counters in LDS, so no key is ever compared against another lane's key.
//
// Both are single-pass over their data: tiles take tickets in order and find
// the sum of all earlier tiles by decoupled look-back (Merrill and Garland,
// "Single-pass parallel prefix scan with decoupled look-back", NVR-2016-002):
// each tile publishes its own total, then walks back over its predecessors'
// published states until it meets an inclusive prefix.  A state is one 64-bit
// word {2-bit flag, 62-bit value} read and written with agent-scope atomics, so
// it is coherent across the XCDs.  A tile only waits on tiles with earlier
// tickets, which are already running; every wait is bounded and reports an
// error instead of spinning on.
#include "agrav_internal.h"

namespace agrav
{

static constexpr int kScanBlock = 256;
static constexpr int kScanItems = 16;
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

static constexpr uint64_t kFlagAgg  = 1ull << 62;
static constexpr uint64_t kFlagIncl = 2ull << 62;
static constexpr uint64_t kValMask  = (1ull << 62) - 1;
static constexpr int kSpinLimit     = 1 << 22;

// 32-bit states {2-bit flag, 30-bit value} for the sort, whose counts stay below 2^30
static constexpr uint32_t kFlagAgg32  = 1u << 30;
static constexpr uint32_t kFlagIncl32 = 2u << 30;
static constexpr uint32_t kValMask32  = (1u << 30) - 1;

template <typename W> __device__ inline W lb_load(const W *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename W> __device__ inline void lb_store(W *p, W v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Sum of the values of tiles 0 .. t-1 from their states (state of tile q at
// state[q * stride]); the caller has published its own aggregate.  WIN states
// are read at once (independent loads, one memory round trip), then summed in
// order up to the nearest inclusive prefix; a window with a state not yet
// published is read again.
template <typename W, int WIN>
__device__ inline int64_t look_back(const W *state, int64_t stride, int t, int *err)
{
  constexpr int sh  = 8 * sizeof(W) - 2;
  constexpr W vmask = (W(1) << sh) - 1;
  int64_t excl      = 0;
  int spins         = 0;
  for(int q = t - 1; q >= 0;)
    {
      W v[WIN];
#pragma unroll
      for(int k = 0; k < WIN; k++)
        v[k] = q - k >= 0 ? lb_load(&state[(int64_t)(q - k) * stride]) : (W(2) << sh);  // before tile 0: prefix 0
      bool ready = true, incl = false;
      int64_t sum = 0;
#pragma unroll
      for(int k = 0; k < WIN; k++)
        if(ready && !incl)
          {
            const W f = v[k] >> sh;
            if(f == 0)
              ready = false;
            else
              {
                sum += (int64_t)(v[k] & vmask);
                incl = f == 2;
              }
          }
      if(!ready)
        {
          if(++spins > kSpinLimit)
            {
              atomicExch(err, 1);  // a predecessor never published: give up (error, not a hang)
              break;
            }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
      excl += sum;
      if(incl)
        break;
      q -= WIN;
    }
  return excl;
}

// The same look-back by a whole wavefront: lane k reads the state of tile q-k, so
// one round trip covers 64 predecessors.  All lanes return the sum.
__device__ inline int64_t look_back_wave(const uint64_t *state, int t, int *err)
{
  const int lane = threadIdx.x & 63;
  int64_t excl   = 0;
  int spins      = 0;
  for(int q = t - 1; q >= 0;)
    {
      const int qq     = q - lane;
      const uint64_t v = qq >= 0 ? lb_load(&state[qq]) : kFlagIncl;  // before tile 0: prefix 0
      const uint64_t f = v >> 62;
      const uint64_t incl_m  = __ballot(f == 2);
      const uint64_t unready = __ballot(f == 0);
      const int k_incl       = incl_m ? __ffsll((long long)incl_m) - 1 : 63;  // nearest inclusive prefix
      const uint64_t need    = k_incl == 63 ? ~0ull : ((2ull << k_incl) - 1);
      if(unready & need)
        {
          if(++spins > kSpinLimit)
            {
              if(lane == 0)
                atomicExch(err, 1);  // a predecessor never published: give up (error, not a hang)
              break;
            }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
      int64_t part = lane <= k_incl ? (int64_t)(v & kValMask) : 0;
      for(int o = 32; o > 0; o >>= 1)
        part += __shfl_xor(part, o, 64);
      excl += part;
      if(incl_m)
        break;
      q -= 64;
    }
  return excl;
}

// single-pass exclusive scan of int32, 2048 items per tile
__global__ __launch_bounds__(kScanBlock) void k_scan_lb(const int *__restrict__ in, int *__restrict__ out, int64_t n, int ntiles,
                                                        int *__restrict__ total_out, uint64_t *state, int *ticket, int *err)
{
  __shared__ int s_wave[kScanBlock / 64];
  __shared__ int s_t;
  __shared__ int64_t s_excl;
  if(threadIdx.x == 0)
    s_t = atomicAdd(ticket, 1);
  __syncthreads();
  const int t        = s_t;
  const int64_t base = (int64_t)t * kScanTile + (int64_t)threadIdx.x * kScanItems;
  // each thread's 16 consecutive items as four 16-byte loads when the buffers allow it
  // (one instruction then covers 1 KB of the tile instead of 64 scattered words)
  const bool vec = ((((uintptr_t)in) | ((uintptr_t)out)) & 15) == 0 && base + kScanItems <= n;
  int vals[kScanItems];
  if(vec)
    {
      const int4 *p = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
      for(int q = 0; q < kScanItems / 4; q++)
        {
          const int4 v    = p[q];
          vals[4 * q]     = v.x;
          vals[4 * q + 1] = v.y;
          vals[4 * q + 2] = v.z;
          vals[4 * q + 3] = v.w;
        }
    }
  else
    {
#pragma unroll
      for(int i = 0; i < kScanItems; i++)
        vals[i] = (base + i < n) ? in[base + i] : 0;
    }
  const int total = block_excl_scan_tile(vals, s_wave);
  if(threadIdx.x < 64)
    {
      int64_t excl = 0;
      if(t > 0)
        {
          if(threadIdx.x == 0)
            lb_store(&state[t], kFlagAgg | (uint64_t)(uint32_t)total);
          excl = look_back_wave(state, t, err);
        }
      if(threadIdx.x == 0)
        {
          lb_store(&state[t], kFlagIncl | ((uint64_t)(excl + total) & kValMask));
          s_excl = excl;
        }
    }
  __syncthreads();
  const int off = (int)s_excl;
  if(vec)
    {
      int4 *p = reinterpret_cast<int4 *>(out + base);
#pragma unroll
      for(int q = 0; q < kScanItems / 4; q++)
        p[q] = make_int4(vals[4 * q] + off, vals[4 * q + 1] + off, vals[4 * q + 2] + off, vals[4 * q + 3] + off);
    }
  else
    {
#pragma unroll
      for(int i = 0; i < kScanItems; i++)
        if(base + i < n)
          out[base + i] = vals[i] + off;
    }
  if(total_out && t == ntiles - 1 && threadIdx.x == 0)
    *total_out = off + total;
}

static int64_t ntiles_of(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

size_t scan_temp_bytes(int64_t n) { return (size_t)(ntiles_of(n) + 1) * sizeof(uint64_t) + 256; }

agrav_status exclusive_scan_i32(const int *d_in, int *d_out, int64_t n, int *d_total, void *d_temp, int *d_err, hipStream_t s)
{
  if(n <= 0)
    {
      if(d_total)
        AG_HIP(hipMemsetAsync(d_total, 0, sizeof(int), s));
      return AGRAV_OK;
    }
  const int64_t nt = ntiles_of(n);
  uint64_t *state  = (uint64_t *)d_temp;
  int *ticket      = (int *)(state + nt);
  AG_HIP(hipMemsetAsync(d_temp, 0, (size_t)nt * sizeof(uint64_t) + sizeof(uint64_t), s));
  k_scan_lb<<<(unsigned)nt, kScanBlock, 0, s>>>(d_in, d_out, n, (int)nt, d_total, state, ticket, d_err);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------

static constexpr int kSortBlock = 256;
static constexpr int kSortWaves = kSortBlock / 64;
#ifndef AG_SORT_ROUNDS
#define AG_SORT_ROUNDS 16
#endif
static constexpr int kSortRounds = AG_SORT_ROUNDS;  // keys per lane
static constexpr int kSortTile  = kSortBlock * kSortRounds;
static_assert(kSortBlock == 256, "the scatter gives each thread one of the 256 digits");

// digit histograms of all passes at once: per-block partials, one row per block
static constexpr int kHistBlocks = 256;

// Digit counts of every pass from one read of the keys: per-block counts in LDS, then
// one global atomic per non-zero bin into ghist (passes x 256, zero on entry).  The
// blocks also clear the passes' look-back states (nstates words) for the onesweep
// kernels that follow.
__global__ __launch_bounds__(kSortBlock) void k_radix_hist_all(const uint64_t *__restrict__ keys, int64_t n, int passes, int lo_bit,
                                                               int *__restrict__ ghist, uint32_t *__restrict__ states,
                                                               int64_t nstates)
{
  __shared__ int s_h[8][256];
  for(int q = threadIdx.x; q < 8 * 256; q += kSortBlock)
    s_h[q >> 8][q & 255] = 0;
  for(int64_t q = (int64_t)blockIdx.x * kSortBlock + threadIdx.x; q < nstates; q += (int64_t)gridDim.x * kSortBlock)
    states[q] = 0;
  __syncthreads();
  // kHistBatch keys per thread in flight at once (loads from an index clamped to the
  // last key, whose duplicates are not counted), then their digits counted
  constexpr int kHistBatch = 8;
  const int64_t stride     = (int64_t)gridDim.x * kSortBlock;
  for(int64_t k0 = (int64_t)blockIdx.x * kSortBlock + threadIdx.x; k0 < n; k0 += kHistBatch * stride)
    {
      uint64_t kv[kHistBatch];
#pragma unroll
      for(int j = 0; j < kHistBatch; j++)
        {
          const int64_t k = k0 + j * stride;
          kv[j]           = keys[k < n ? k : n - 1];
        }
#pragma unroll
      for(int j = 0; j < kHistBatch; j++)
        if(k0 + j * stride < n)
          for(int p = 0; p < passes; p++)
            atomicAdd(&s_h[p][(kv[j] >> (lo_bit + 8 * p)) & 255], 1);
    }
  __syncthreads();
  for(int q = threadIdx.x; q < passes * 256; q += kSortBlock)
    {
      const int v = s_h[q >> 8][q & 255];
      if(v)
        atomicAdd(&ghist[q], v);
    }
}

// One pass of the sort over one tile of 2048 keys: rank stably by the digit,
// publish the tile's digit counts and look back for the counts of all earlier
// tiles, stage the tile in LDS in digit order, then write each digit's run to
// its global position with consecutive lanes on consecutive addresses.
__global__ __launch_bounds__(kSortBlock) void k_radix_onesweep(const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                                                               uint64_t *__restrict__ kout, int *__restrict__ vout, int64_t n,
                                                               int shift, const int *__restrict__ ghist, uint32_t *state,
                                                               int *ticket, int *err)
{
  __shared__ uint64_t s_key[kSortTile];
  __shared__ int s_val[kSortTile];
  __shared__ int s_cnt[kSortWaves][256];
  __shared__ int s_dbase[256], s_gbase[256], s_wsum[kSortWaves], s_gsum[kSortWaves];
  __shared__ int s_t;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if(threadIdx.x == 0)
    s_t = atomicAdd(ticket, 1);
  for(int i = lane; i < 256; i += 64)
    s_cnt[w][i] = 0;
  __syncthreads();
  const int t = s_t;

  const uint64_t lt   = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t tbase = (int64_t)t * kSortTile;
  const int tcount    = (int)(n - tbase < kSortTile ? n - tbase : kSortTile);
  const int64_t wbase = tbase + (int64_t)w * (64 * kSortRounds);
  uint64_t key[kSortRounds];
  int val[kSortRounds];
  int pos[kSortRounds];
  int dig[kSortRounds];

  // 1. rank inside the wave: 8 ballots give the lanes with the same digit.  All the
  //    wave's keys are loaded first, unconditionally from an index clamped to the last
  //    key (n >= 1), so that every load is in flight before the first ballot (a load
  //    under the bounds test was issued and waited for round by round)
#pragma unroll
  for(int r = 0; r < kSortRounds; r++)
    {
      const int64_t k  = wbase + r * 64 + lane;
      const int64_t kc = k < n ? k : n - 1;
      key[r]           = kin[kc];
      val[r]           = vin[kc];
    }
#pragma unroll
  for(int r = 0; r < kSortRounds; r++)
    {
      int64_t k   = wbase + r * 64 + lane;
      bool valid  = k < n;
      key[r]      = valid ? key[r] : 0;
      val[r]      = valid ? val[r] : 0;
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

  // 2. per digit (one per thread): wave offsets, the tile's count published, the
  //    earlier tiles' counts by look-back, the tile-local start and the digit's start in
  //    the output (exclusive scan of the pass's digit counts) by scans over digits
  {
    const int d    = threadIdx.x;
    const int gh   = ghist[d];
    const int ginc = wave_incl_scan(gh);
    if(lane == 63)
      s_gsum[w] = ginc;
    int run     = 0;
#pragma unroll
    for(int ww = 0; ww < kSortWaves; ww++)
      {
        int c        = s_cnt[ww][d];
        s_cnt[ww][d] = run;
        run += c;
      }
    int64_t excl = 0;
    if(t > 0)
      {
        lb_store(&state[(int64_t)t * 256 + d], kFlagAgg32 | (uint32_t)run);
        excl = look_back<uint32_t, 8>(state + d, 256, t, err);
      }
    lb_store(&state[(int64_t)t * 256 + d], kFlagIncl32 | ((uint32_t)(excl + run) & kValMask32));
    const int incl = wave_incl_scan(run);
    if(lane == 63)
      s_wsum[w] = incl;
    __syncthreads();
    int wpre = 0, gpre = 0;
#pragma unroll
    for(int ww = 0; ww < kSortWaves; ww++)
      {
        wpre += ww < w ? s_wsum[ww] : 0;
        gpre += ww < w ? s_gsum[ww] : 0;
      }
    s_dbase[d] = wpre + incl - run;
    s_gbase[d] = gpre + ginc - gh + (int)excl;
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
      const uint64_t k  = s_key[i];
      const int d       = (int)((k >> shift) & 255);
      const int64_t dst = (int64_t)s_gbase[d] + (i - s_dbase[d]);
      kout[dst]         = k;
      vout[dst]         = s_val[i];
    }
}

size_t sort_temp_bytes(int64_t n)
{
  const int64_t nt = (n + kSortTile - 1) / kSortTile;
  // per pass: look-back states (nt x 256 words) + ticket; digit counts of every pass
  return (size_t)8 * (nt * 256 + 1) * sizeof(uint32_t) + 8 * 256 * sizeof(int) + 1024;
}

agrav_status radix_sort_u64(uint64_t *keys_in, int *vals_in, uint64_t *keys_out, int *vals_out, int64_t n, int key_bits,
                            void *d_temp, int *d_err, hipStream_t s, int lo_bit, bool *in_input, int *ghist_zero)
{
  if(in_input)
    *in_input = false;
  if(n <= 0)
    return AGRAV_OK;
  const int64_t nt  = (n + kSortTile - 1) / kSortTile;
  const int passes  = (key_bits - lo_bit + 7) / 8;
  uint32_t *states  = (uint32_t *)d_temp;                     // passes x (nt x 256 + 1)
  const int64_t per = nt * 256 + 1;
  int *ghist        = ghist_zero;                              // passes x 256 digit counts
  if(!ghist)
    {
      ghist = (int *)(states + (size_t)passes * per);
      AG_HIP(hipMemsetAsync(ghist, 0, (size_t)passes * 256 * sizeof(int), s));
    }
  const int hb = (int)(nt < kHistBlocks ? nt : kHistBlocks);
  k_radix_hist_all<<<hb, kSortBlock, 0, s>>>(keys_in, n, passes, lo_bit, ghist, states, (int64_t)passes * per);
  AG_HIP(hipGetLastError());
  uint64_t *ka = keys_in, *kb = keys_out;
  int *va = vals_in, *vb = vals_out;
  for(int p = 0; p < passes; p++)
    {
      uint32_t *st = states + (size_t)p * per;
      int *ticket  = (int *)(st + nt * 256);
      k_radix_onesweep<<<(unsigned)nt, kSortBlock, 0, s>>>(ka, va, kb, vb, n, lo_bit + 8 * p, ghist + 256 * p, st, ticket, d_err);
      AG_HIP(hipGetLastError());
      uint64_t *tk = ka;
      ka           = kb;
      kb           = tk;
      int *tv      = va;
      va           = vb;
      vb           = tv;
    }
  if(ka != keys_out && in_input)
    *in_input = true;
  else if(ka != keys_out)
    {
      AG_HIP(hipMemcpyAsync(keys_out, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
      AG_HIP(hipMemcpyAsync(vals_out, va, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
  return AGRAV_OK;
}

}  // namespace agrav
