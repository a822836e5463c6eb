// agrav_pmhr.hip -- the high-resolution PM region of PLACEHIGHRESREGION
// (pm_nonperiodic.c with grnr = 1, longrange.c:129-140).
//
// A second, non-periodic PM grid of GRID = 2 PMGRID cells covers the particles of
// the high-resolution types.  Its kernel is the force band between the coarse split
// scale Asmth[0] and the fine one Asmth[1]; targets in the region take the short-range
// force with Asmth[1] and Rcut[1] (the walks split their targets by the flag computed
// here).  Steps and their reference counterparts:
//   region             pm_init_regionsize, pm_nonperiodic.c:135-245 -> k_hr_extent (one
//                      pass over the particles of the mask's types) and the host formula;
//                      re-sized only when such a particle has left it (:1527-1561);
//   kernel             pm_setup_nonperiodic_kernel, :1778-1987 -> k_hr_kernel_real, rocFFT
//                      r2c, k_hr_deconv (CIC deconvolution twice), once per region;
//   deposit            :810-1170 -> k_hr_deposit (64-bit fixed point, order independent);
//   convolution        :1597-1627 -> rocFFT r2c, complex product, c2r;
//   differences        :1652-1700 -> k_hr_fd on the cells 2 .. GRID/2 - 3 of the inner half;
//   readout            :1283-1500 -> k_hr_readout, added to the periodic PM result, for the
//                      high-resolution particles only (pmforce_is_particle_high_res, :344-371).
#include "agrav_ctx.h"

#include <math.h>
#include <string.h>

#include <vector>

namespace agrav
{

static constexpr int B         = 256;
static constexpr int kGridBoost = 2;  // GRIDBOOST default, pm_nonperiodic.c:75-77
static inline unsigned nb(int64_t n) { return (unsigned)((n + B - 1) / B); }

bool hr_on(const agrav_ctx *c) { return c->par.placehighresregion != 0 && c->par.pmgrid > 0 && c->par.periodic; }

__device__ inline double hr_wrap(double x, double L)
{
  if(x < 0)
    x += L;
  if(x >= L)
    x -= L;
  return x;
}

// order-preserving 64-bit key of a double (min/max by unsigned integer atomics)
__device__ inline unsigned long long ord_key(double x)
{
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
static inline double ord_val(unsigned long long k)
{
  unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  double x;
  memcpy(&x, &b, sizeof(x));
  return x;
}

__device__ inline unsigned long long wave_min_u64(unsigned long long v)
{
  for(int o = 32; o > 0; o >>= 1)
    {
      unsigned long long w = __shfl_xor(v, o, 64);
      v                    = w < v ? w : v;
    }
  return v;
}
__device__ inline unsigned long long wave_max_u64(unsigned long long v)
{
  for(int o = 32; o > 0; o >>= 1)
    {
      unsigned long long w = __shfl_xor(v, o, 64);
      v                    = w > v ? w : v;
    }
  return v;
}

struct HrBox
{
  double lo[3], hi[3];
};

// extent of the particles of the mask's types (wrapped positions), their count, and
// whether one of them lies outside the current region [xmin, xmax] (pm_nonperiodic.c:1527-1561)
__global__ __launch_bounds__(B) void k_hr_extent(int64_t n, const double *__restrict__ pos, const uint8_t *__restrict__ type,
                                                 int mask, double L, int valid, HrBox cur, unsigned long long *ext)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  const bool sel  = i < n && ((1 << type[i]) & mask);
  unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
  unsigned long long out = 0;
  if(sel)
    for(int j = 0; j < 3; j++)
      {
        const double x = hr_wrap(pos[3 * i + j], L);
        mn[j] = mx[j] = ord_key(x);
        if(valid && (x < cur.lo[j] || x > cur.hi[j]))
          out = 1;
      }
  const unsigned long long cnt = __popcll(__ballot(sel));
  for(int j = 0; j < 3; j++)
    {
      mn[j] = wave_min_u64(mn[j]);
      mx[j] = wave_max_u64(mx[j]);
    }
  out = wave_max_u64(out);
  if((threadIdx.x & 63) == 0 && cnt > 0)
    {
      for(int j = 0; j < 3; j++)
        {
          atomicMin(&ext[j], mn[j]);
          atomicMax(&ext[3 + j], mx[j]);
        }
      atomicAdd(&ext[6], cnt);
      if(out)
        atomicOr(&ext[7], 1ull);
    }
}

__global__ void k_hr_ext_init(unsigned long long *ext)
{
  const int j = threadIdx.x;
  if(j < 8)
    ext[j] = j < 3 ? ~0ull : 0ull;
}

// pmforce_is_particle_high_res (pm_nonperiodic.c:344-371) for every particle, and per
// block the mass of the flagged particles summed in a fixed order
__global__ __launch_bounds__(B) void k_hr_flags(int64_t n, const double *__restrict__ pos, const uint8_t *__restrict__ type,
                                                const double *__restrict__ mass, int mask, double L, HrBox reg, uint8_t *flag,
                                                double *mpart)
{
  __shared__ double s_m[B];
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  double m        = 0;
  if(i < n)
    {
      bool hr = ((1 << type[i]) & mask) != 0;
      if(!hr)
        {
          hr = true;
          for(int j = 0; j < 3; j++)
            {
              const double x = hr_wrap(pos[3 * i + j], L);
              if(x < reg.lo[j] || x > reg.hi[j])
                hr = false;
            }
        }
      flag[i] = hr ? 1 : 0;
      m       = hr ? mass[i] : 0.0;
    }
  s_m[threadIdx.x] = m;
  __syncthreads();
  for(int w = B / 2; w > 0; w >>= 1)
    {
      if(threadIdx.x < w)
        s_m[threadIdx.x] += s_m[threadIdx.x + w];
      __syncthreads();
    }
  if(threadIdx.x == 0)
    mpart[blockIdx.x] = s_m[0];
}

// the kernel in real space (pm_nonperiodic.c:1850-1905): coordinates in units of the
// mesh size wrapped to [-1/2, 1/2), u = r / (2 ASMTH / GRID);
// -(erfc(u Asmth[1] / Asmth[0]) - erfc(u)) / r, and -(1 - Asmth[1] / Asmth[0]) / (sqrt(pi) ASMTH / GRID) at r = 0
template <typename R> __global__ void k_hr_kernel_real(int G, double asmth_cells, double ratio, R *out)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= (int64_t)G * G * G)
    return;
  const int z = (int)(i % G), y = (int)((i / G) % G), x = (int)(i / ((int64_t)G * G));
  double xx = (double)x / G, yy = (double)y / G, zz = (double)z / G;
  if(xx >= 0.5)
    xx -= 1.0;
  if(yy >= 0.5)
    yy -= 1.0;
  if(zz >= 0.5)
    zz -= 1.0;
  const double r = sqrt(xx * xx + yy * yy + zz * zz);
  const double u = 0.5 * r / (asmth_cells / G);
  double v;
  if(r > 0)
    v = -(erfc(u * ratio) - erfc(u)) / r;
  else
    v = -(1 - ratio) / (sqrt(M_PI) * (asmth_cells / G));
  out[i] = (R)v;
}

// CIC deconvolution of the transformed kernel, twice (pm_nonperiodic.c:1936-1981):
// [sinc(pi k_x / GRID) sinc(pi k_y / GRID) sinc(pi k_z / GRID)]^-4, frequencies above GRID / 2 negative
template <typename R> __global__ void k_hr_deconv(int G, R *kern)
{
  const int gz    = G / 2 + 1;
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= (int64_t)G * G * gz)
    return;
  const int z = (int)(i % gz), y = (int)((i / gz) % G), x = (int)(i / ((int64_t)gz * G));
  const double kx = x > G / 2 ? x - G : x, ky = y > G / 2 ? y - G : y, kz = z > G / 2 ? z - G : z;
  if(kx * kx + ky * ky + kz * kz <= 0)
    return;
  double fx = 1, fy = 1, fz = 1;
  if(kx != 0)
    fx = sin(M_PI * kx / G) / (M_PI * kx / G);
  if(ky != 0)
    fy = sin(M_PI * ky / G) / (M_PI * ky / G);
  if(kz != 0)
    fz = sin(M_PI * kz / G) / (M_PI * kz / G);
  double ff = 1 / (fx * fy * fz);
  ff        = ff * ff * ff * ff;
  kern[2 * i]     = (R)((double)kern[2 * i] * ff);
  kern[2 * i + 1] = (R)((double)kern[2 * i + 1] * ff);
}

struct HrGrid
{
  double corner[3], upper[3];
  double fs;  // GRID / mesh size (to_slab_fac)
  int G, H;   // padded grid, its inner half
};

// CIC deposit of every particle inside [Corner, UpperCorner) (pm_nonperiodic.c:810-1170), no wrapping
__global__ void k_hr_deposit(int64_t n, const double *__restrict__ pos, const double *__restrict__ mass, double L, HrGrid g,
                             double scale, unsigned long long *rho)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  double x[3];
  for(int k = 0; k < 3; k++)
    {
      x[k] = hr_wrap(pos[3 * i + k], L);
      if(!(x[k] >= g.corner[k] && x[k] < g.upper[k]))
        return;
    }
  int s[3];
  double d[3];
  for(int k = 0; k < 3; k++)
    {
      s[k] = (int)(g.fs * (x[k] - g.corner[k]));  // at most GRID/2 - 1 below UpperCorner: s + 1 < GRID
      d[k] = g.fs * (x[k] - g.corner[k]) - s[k];
    }
  const double m = mass[i] * scale;
  for(int c = 0; c < 8; c++)
    {
      const int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
      const double w = (bx ? d[0] : 1.0 - d[0]) * (by ? d[1] : 1.0 - d[1]) * (bz ? d[2] : 1.0 - d[2]);
      atomicAdd(&rho[((size_t)(s[0] + bx) * g.G + (s[1] + by)) * g.G + (s[2] + bz)], (unsigned long long)llrint(m * w));
    }
}

template <typename R>
__global__ void k_hr_fix_to_real(const unsigned long long *__restrict__ rho, int64_t g3, double scale, R *out)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i < g3)
    out[i] = (R)((double)rho[i] / scale);
}

template <typename R> __global__ void k_hr_cmul(R *a, const R *__restrict__ k, int64_t nc)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= nc)
    return;
  const double ar = a[2 * i], ai = a[2 * i + 1], kr = k[2 * i], ki = k[2 * i + 1];
  a[2 * i]     = (R)(ar * kr - ai * ki);
  a[2 * i + 1] = (R)(ar * ki + ai * kr);
}

template <typename R> struct HrMesh4
{
  R v[4];
};

// 4-point differences (pm_nonperiodic.c:1652-1700) on the cells 2 .. GRID/2 - 3 of each
// axis, zero elsewhere, and the potential, over the inner half of the padded grid
template <typename R> __global__ void k_hr_fd(const R *__restrict__ phi, int G, int H, double fac, HrMesh4<R> *mesh)
{
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= (int64_t)H * H * H)
    return;
  const int z = (int)(i % H), y = (int)((i / H) % H), x = (int)(i / ((int64_t)H * H));
  auto at = [&](int a, int b, int c) -> double { return (double)phi[((int64_t)a * G + b) * G + c]; };
  HrMesh4<R> m;
  m.v[0] = m.v[1] = m.v[2] = 0;
  if(x >= 2 && x < H - 2 && y >= 2 && y < H - 2 && z >= 2 && z < H - 2)
    {
      m.v[0] = (R)(fac * ((4.0 / 3) * (at(x - 1, y, z) - at(x + 1, y, z)) - (1.0 / 6) * (at(x - 2, y, z) - at(x + 2, y, z))));
      m.v[1] = (R)(fac * ((4.0 / 3) * (at(x, y - 1, z) - at(x, y + 1, z)) - (1.0 / 6) * (at(x, y - 2, z) - at(x, y + 2, z))));
      m.v[2] = (R)(fac * ((4.0 / 3) * (at(x, y, z - 1) - at(x, y, z + 1)) - (1.0 / 6) * (at(x, y, z - 2) - at(x, y, z + 2))));
    }
  m.v[3]  = (R)at(x, y, z);
  mesh[i] = m;
}

// CIC readout for the high-resolution particles (pm_nonperiodic.c:1283-1500), added to
// the periodic result: entry q of the PM output is particle perm[q] (sorted order) or q
template <typename R>
__global__ void k_hr_readout(int64_t n, const int *__restrict__ perm, const double *__restrict__ pos,
                             const uint8_t *__restrict__ flag, double L, HrGrid g, const HrMesh4<R> *__restrict__ mesh,
                             double fac_pot, double *acc_pm, double *pot_pm)
{
  const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
  if(q >= n)
    return;
  const int64_t i = perm ? perm[q] : q;
  if(!flag[i])
    return;
  int s[3];
  double d[3];
  for(int k = 0; k < 3; k++)
    {
      const double x = hr_wrap(pos[3 * i + k], L);
      s[k]           = (int)(g.fs * (x - g.corner[k]));
      d[k]           = g.fs * (x - g.corner[k]) - s[k];
      s[k]           = s[k] < 0 ? 0 : (s[k] < g.H - 1 ? s[k] : g.H - 2);  // the region lies inside; kept in bounds
    }
  double f0 = 0, f1 = 0, f2 = 0, pp = 0;
  for(int c = 0; c < 8; c++)
    {
      const int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
      const double w = (bx ? d[0] : 1.0 - d[0]) * (by ? d[1] : 1.0 - d[1]) * (bz ? d[2] : 1.0 - d[2]);
      const HrMesh4<R> m = mesh[((int64_t)(s[0] + bx) * g.H + (s[1] + by)) * g.H + (s[2] + bz)];
      f0 += (double)m.v[0] * w;
      f1 += (double)m.v[1] * w;
      f2 += (double)m.v[2] * w;
      pp += (double)m.v[3] * w;
    }
  acc_pm[3 * q + 0] += f0;
  acc_pm[3 * q + 1] += f1;
  acc_pm[3 * q + 2] += f2;
  pot_pm[q] += pp * fac_pot;
}

static agrav_status hr_fft_check(agrav_ctx *c, rocfft_status s, const char *what)
{
  if(s != rocfft_status_success)
    return fail(c, AGRAV_ERR_FFT, std::string("rocFFT (high-resolution grid): ") + what);
  return AGRAV_OK;
}

void hr_free(agrav_ctx *c)
{
  HrState &h = c->hr;
  if(h.fwd)
    rocfft_plan_destroy(h.fwd);
  if(h.bwd)
    rocfft_plan_destroy(h.bwd);
  if(h.info_f)
    rocfft_execution_info_destroy(h.info_f);
  if(h.info_b)
    rocfft_execution_info_destroy(h.info_b);
  h.fwd = h.bwd = nullptr;
  h.info_f = h.info_b = nullptr;
  h.work.release();
  h.rho_fix.release();
  h.real.release();
  h.cplx.release();
  h.kern.release();
  h.mesh.release();
  h.ext.release();
  h.mpart.release();
  h.grid         = 0;
  h.kernel_ready = false;
}

// plans and grids of the padded GRID^3 mesh (pm_init_nonperiodic, pm_nonperiodic.c:255-330)
static agrav_status hr_init(agrav_ctx *c)
{
  HrState &h  = c->hr;
  const int G = kGridBoost * c->par.pmgrid;
  if(h.grid == G && h.prec == c->par.precision)
    return AGRAV_OK;
  const bool keep_region = h.valid;
  hr_free(c);
  h.valid             = keep_region;
  const bool sp       = c->par.precision == AGRAV_FP32;
  const size_t rsz    = sp ? sizeof(float) : sizeof(double);
  const int H         = G / 2;
  const int64_t g3    = (int64_t)G * G * G;
  const int64_t nc    = (int64_t)G * G * (G / 2 + 1);
  const size_t len[3] = {(size_t)G, (size_t)G, (size_t)G};
  rocfft_precision pr = sp ? rocfft_precision_single : rocfft_precision_double;
  agrav_status st;
  if((st = pm_init(c)) != AGRAV_OK)  // rocfft_setup
    return st;
  if((st = hr_fft_check(c, rocfft_plan_create(&h.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, pr, 3,
                                              len, 1, nullptr),
                        "plan forward")) ||
     (st = hr_fft_check(c, rocfft_plan_create(&h.bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, pr, 3,
                                              len, 1, nullptr),
                        "plan inverse")))
    return st;
  size_t w1 = 0, w2 = 0;
  rocfft_plan_get_work_buffer_size(h.fwd, &w1);
  rocfft_plan_get_work_buffer_size(h.bwd, &w2);
  const size_t w = w1 > w2 ? w1 : w2;
  if((st = h.work.ensure(w ? w : 1)) || (st = h.rho_fix.ensure(g3)) || (st = h.real.ensure(g3 * rsz)) ||
     (st = h.cplx.ensure(nc * 2 * rsz)) || (st = h.kern.ensure(nc * 2 * rsz)) ||
     (st = h.mesh.ensure((size_t)H * H * H * 4 * rsz)) || (st = h.ext.ensure(8)))
    return st;
  if((st = hr_fft_check(c, rocfft_execution_info_create(&h.info_f), "info")) ||
     (st = hr_fft_check(c, rocfft_execution_info_create(&h.info_b), "info")))
    return st;
  if(w)
    {
      rocfft_execution_info_set_work_buffer(h.info_f, h.work.p, w);
      rocfft_execution_info_set_work_buffer(h.info_b, h.work.p, w);
    }
  rocfft_execution_info_set_stream(h.info_f, c->stream);
  rocfft_execution_info_set_stream(h.info_b, c->stream);
  h.grid = G;
  h.prec = c->par.precision;
  return AGRAV_OK;
}

// pm_init_regionsize (pm_nonperiodic.c:135-245) for the high-resolution grid, from the
// extent [lo, hi] of the particles of the mask's types
static void hr_size_region(agrav_ctx *c, const double lo[3], const double hi[3])
{
  HrState &h     = c->hr;
  const int GRID = kGridBoost * c->par.pmgrid;
  double size    = 0;
  for(int j = 0; j < 3; j++)
    size = hi[j] - lo[j] > size ? hi[j] - lo[j] : size;
  if(c->par.enlargeregion > 0)
    size *= c->par.enlargeregion;  // ENLARGEREGION
  for(int j = 0; j < 3; j++)
    {
      h.xmin[j] = (lo[j] + hi[j]) / 2 - size / 2;  // symmetrised onto the centre
      h.xmax[j] = h.xmin[j] + size;
    }
  const double meshinner = size;
  double tms             = size * 2.001 * GRID / ((double)(GRID - 2 - 8));
  for(int j = 0; j < 3; j++)
    {
      h.corner[j] = h.xmin[j] - 2.0005 * tms / GRID;
      h.upper[j]  = h.corner[j] + (GRID / 2 - 1) * (tms / GRID);
    }
  const double asmth0 = c->par.asmth * c->par.box_size / c->par.pmgrid, rcut0 = c->par.rcut * asmth0;
  if(2 * tms / GRID < rcut0)
    {
      // a high-resolution cell below half the coarse cut: the region is widened by the coarse cut
      tms = 2 * (meshinner + 2 * rcut0) * GRID / ((double)(GRID - 2));
      for(int j = 0; j < 3; j++)
        {
          h.corner[j] = h.xmin[j] - 1.0001 * rcut0;
          h.upper[j]  = h.corner[j] + (GRID / 2 - 1) * (tms / GRID);
        }
      if(2 * tms / GRID > rcut0)
        {
          tms = 2 * (meshinner + 2 * rcut0) * GRID / ((double)(GRID - 10));
          for(int j = 0; j < 3; j++)
            {
              h.corner[j] = h.xmin[j] - 1.0001 * (rcut0 + 2 * tms / GRID);
              h.upper[j]  = h.corner[j] + (GRID / 2 - 1) * (tms / GRID);
            }
        }
    }
  h.meshsize     = tms;
  h.asmth        = c->par.asmth * tms / GRID;  // All.Asmth[1]
  h.rcut         = c->par.rcut * h.asmth;      // All.Rcut[1]
  h.valid        = true;
  h.kernel_ready = false;
}

template <typename R> static agrav_status hr_kernel(agrav_ctx *c)
{
  HrState &h        = c->hr;
  const int G       = h.grid;
  const int64_t g3  = (int64_t)G * G * G;
  const int64_t nc  = (int64_t)G * G * (G / 2 + 1);
  hipStream_t s     = c->stream;
  const double a0   = c->par.asmth * c->par.box_size / c->par.pmgrid;
  k_hr_kernel_real<R><<<nb(g3), B, 0, s>>>(G, c->par.asmth, h.asmth / a0, (R *)h.real.p);
  AG_HIP(hipGetLastError());
  void *in[1]  = {h.real.p};
  void *out[1] = {h.kern.p};
  agrav_status st;
  if((st = hr_fft_check(c, rocfft_execute(h.fwd, in, out, h.info_f), "kernel forward")))
    return st;
  k_hr_deconv<R><<<nb(nc), B, 0, s>>>(G, (R *)h.kern.p);
  AG_HIP(hipGetLastError());
  h.kernel_ready = true;
  return AGRAV_OK;
}

// the region for the particles now loaded: sized when unset or when a particle of the
// mask's types has left it (longrange.c:129-140), then the kernel for it
static agrav_status hr_region(agrav_ctx *c)
{
  HrState &h      = c->hr;
  const int64_t n = c->n;
  hipStream_t s   = c->stream;
  if(!c->have_ptype)
    return fail(c, AGRAV_ERR_STATE, "placehighresregion needs the particle types (agrav_set_particle_types)");
  HrBox cur{};
  for(int j = 0; j < 3; j++)
    {
      cur.lo[j] = h.xmin[j];
      cur.hi[j] = h.xmax[j];
    }
  k_hr_ext_init<<<1, 64, 0, s>>>(h.ext.p);
  if(n > 0)
    k_hr_extent<<<nb(n), B, 0, s>>>(n, c->d_pos_in.p, c->d_ptype_in.p, c->par.placehighresregion, c->par.box_size,
                                    h.valid ? 1 : 0, cur, h.ext.p);
  AG_HIP(hipGetLastError());
  unsigned long long e[8];
  AG_HIP(hipMemcpyAsync(e, h.ext.p, sizeof(e), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(e[6] == 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "placehighresregion: no particle of the high-resolution types");
  if(!h.valid || e[7])
    {
      double lo[3], hi[3];
      for(int j = 0; j < 3; j++)
        {
          lo[j] = ord_val(e[j]);
          hi[j] = ord_val(e[3 + j]);
        }
      hr_size_region(c, lo, hi);
      // the reference stops when even the new region misses a particle (longrange.c:139)
      for(int j = 0; j < 3; j++)
        if(lo[j] < h.xmin[j] || hi[j] > h.xmax[j])
          {
            h.valid = false;
            return fail(c, AGRAV_ERR_OUT_OF_BOX, "high-resolution particles outside the high-resolution region after re-sizing");
          }
    }
  if(!h.kernel_ready)
    return c->par.precision == AGRAV_FP32 ? hr_kernel<float>(c) : hr_kernel<double>(c);
  return AGRAV_OK;
}

agrav_status hr_flags(agrav_ctx *c)
{
  HrState &h      = c->hr;
  const int64_t n = c->n;
  hipStream_t s   = c->stream;
  if(!h.valid)
    return fail(c, AGRAV_ERR_STATE, "the high-resolution region is not sized: run the PM (long_range_force) first");
  if(!c->have_ptype)
    return fail(c, AGRAV_ERR_STATE, "placehighresregion needs the particle types (agrav_set_particle_types)");
  agrav_status st;
  const unsigned nblk = nb(n > 0 ? n : 1);
  if((st = c->d_hr.ensure(n ? n : 1)) || (st = h.mpart.ensure(nblk)))
    return st;
  HrBox reg{};
  for(int j = 0; j < 3; j++)
    {
      reg.lo[j] = h.xmin[j];
      reg.hi[j] = h.xmax[j];
    }
  h.m_high = 0;
  if(n == 0)
    return AGRAV_OK;
  k_hr_flags<<<nblk, B, 0, s>>>(n, c->d_pos_in.p, c->d_ptype_in.p, c->d_mass_in.p, c->par.placehighresregion, c->par.box_size,
                                reg, c->d_hr.p, h.mpart.p);
  AG_HIP(hipGetLastError());
  std::vector<double> mp(nblk);
  AG_HIP(hipMemcpyAsync(mp.data(), h.mpart.p, nblk * sizeof(double), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  double m = 0;
  for(unsigned b = 0; b < nblk; b++)  // fixed order
    m += mp[b];
  h.m_high = m;
  return AGRAV_OK;
}

template <typename R> static agrav_status hr_run(agrav_ctx *c)
{
  HrState &h       = c->hr;
  const int G      = h.grid, H = G / 2;
  const int64_t n  = c->n;
  const int64_t g3 = (int64_t)G * G * G;
  const int64_t nc = (int64_t)G * G * (G / 2 + 1);
  const double L   = c->par.box_size;
  hipStream_t s    = c->stream;
  agrav_status st;
  HrGrid g{};
  for(int j = 0; j < 3; j++)
    {
      g.corner[j] = h.corner[j];
      g.upper[j]  = h.upper[j];
    }
  g.fs = G / h.meshsize;
  g.G  = G;
  g.H  = H;
  int e = 0;
  frexp(c->mass_total > 0 ? c->mass_total : 1.0, &e);
  const double scale = ldexp(1.0, 61 - e);
  AG_HIP(hipMemsetAsync(h.rho_fix.p, 0, g3 * sizeof(unsigned long long), s));
  if(n > 0)
    k_hr_deposit<<<nb(n), B, 0, s>>>(n, c->d_pos_in.p, c->d_mass_in.p, L, g, scale, h.rho_fix.p);
  k_hr_fix_to_real<R><<<nb(g3), B, 0, s>>>(h.rho_fix.p, g3, scale, (R *)h.real.p);
  AG_HIP(hipGetLastError());
  void *in[1]  = {h.real.p};
  void *out[1] = {h.cplx.p};
  if((st = hr_fft_check(c, rocfft_execute(h.fwd, in, out, h.info_f), "execute forward")))
    return st;
  k_hr_cmul<R><<<nb(nc), B, 0, s>>>((R *)h.cplx.p, (const R *)h.kern.p, nc);
  AG_HIP(hipGetLastError());
  void *in2[1]  = {h.cplx.p};
  void *out2[1] = {h.real.p};
  if((st = hr_fft_check(c, rocfft_execute(h.bwd, in2, out2, h.info_b), "execute inverse")))
    return st;
  // G / L^4 (L / GRID)^3 for the potential, / (2 L / GRID) for the differences (pm_nonperiodic.c:1522-1523)
  const double cell    = h.meshsize / G;
  const double fac_pot = c->par.G / pow(h.meshsize, 4) * pow(cell, 3);
  const double fac_fd  = fac_pot / (2 * cell);
  HrMesh4<R> *mesh     = reinterpret_cast<HrMesh4<R> *>(h.mesh.p);
  k_hr_fd<R><<<nb((int64_t)H * H * H), B, 0, s>>>((const R *)h.real.p, G, H, fac_fd, mesh);
  if(n > 0)
    k_hr_readout<R><<<nb(n), B, 0, s>>>(n, c->pm_sorted ? c->d_idx2.p : nullptr, c->d_pos_in.p, c->d_hr.p, L, g, mesh, fac_pot,
                                        c->d_acc_pm.p, c->d_pot_pm.p);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
}

// the high-resolution grid's part of long_range_force (longrange.c:129-140), after the
// periodic grid: its force and potential are added to the periodic ones
agrav_status hr_pm(agrav_ctx *c)
{
  if(!hr_on(c))
    return AGRAV_OK;
  agrav_status st;
  if((st = hr_init(c)) != AGRAV_OK || (st = hr_region(c)) != AGRAV_OK || (st = hr_flags(c)) != AGRAV_OK)
    return st;
  return c->par.precision == AGRAV_FP32 ? hr_run<float>(c) : hr_run<double>(c);
}

// The walk's targets split by the high-resolution flag, each list in the order given
// (Hilbert order): targets tmap[0 .. nt) (sorted positions; null: 0 .. nt), flags
// gathered to sorted order through the tree's permutation.
__global__ void k_hr_gather_flags(int64_t n, const int *__restrict__ perm, const uint8_t *__restrict__ flag, uint8_t *flag_s)
{
  const int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k < n)
    flag_s[k] = flag[perm[k]];
}

__global__ void k_hr_target_flags(int64_t nt, const int *__restrict__ tmap, const uint8_t *__restrict__ flag_s, int *f)
{
  const int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k < nt)
    f[k] = flag_s[tmap ? tmap[k] : k];
}

__global__ void k_hr_split(int64_t nt, const int *__restrict__ tmap, const int *__restrict__ f, const int *__restrict__ pos,
                           int *lo, int *hi)
{
  const int64_t k = (int64_t)blockIdx.x * B + threadIdx.x;
  if(k >= nt)
    return;
  const int t = tmap ? tmap[k] : (int)k;
  if(f[k])
    hi[pos[k]] = t;
  else
    lo[k - pos[k]] = t;
}

agrav_status hr_split_targets(agrav_ctx *c, const int *tmap, int64_t nt, const int **lo, int64_t *n_lo, const int **hi,
                              int64_t *n_hi)
{
  hipStream_t s       = c->stream;
  const int64_t ntree = c->n_tree;
  agrav_status st;
  if((st = hr_flags(c)) != AGRAV_OK)
    return st;
  const int64_t nn = nt > 0 ? nt : 1;
  if((st = c->d_hr_s.ensure(ntree ? ntree : 1)) || (st = c->d_hr_flag.ensure(nn)) || (st = c->d_hr_pos.ensure(nn)) ||
     (st = c->d_hr_list.ensure(2 * nn)) || (st = c->d_temp.ensure(scan_temp_bytes(nn) + 4096)) ||
     (st = c->d_counters.ensure(128)))
    return st;
  int *d_tot = c->d_counters.p + 104, *d_err = c->d_counters.p + 105;
  AG_HIP(hipMemsetAsync(d_tot, 0, 2 * sizeof(int), s));
  if(ntree > 0)
    k_hr_gather_flags<<<nb(ntree), B, 0, s>>>(ntree, c->d_idx2.p, c->d_hr.p, c->d_hr_s.p);
  if(nt > 0)
    {
      k_hr_target_flags<<<nb(nt), B, 0, s>>>(nt, tmap, c->d_hr_s.p, c->d_hr_flag.p);
      AG_HIP(hipGetLastError());
      if((st = exclusive_scan_i32(c->d_hr_flag.p, c->d_hr_pos.p, nt, d_tot, c->d_temp.p, d_err, s)) != AGRAV_OK)
        return st;
      k_hr_split<<<nb(nt), B, 0, s>>>(nt, tmap, c->d_hr_flag.p, c->d_hr_pos.p, c->d_hr_list.p, c->d_hr_list.p + nt);
    }
  AG_HIP(hipGetLastError());
  int h[2] = {0, 0};
  AG_HIP(hipMemcpyAsync(h, d_tot, sizeof(h), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(h[1])
    return fail(c, AGRAV_ERR_HIP, "high-resolution target scan look-back did not complete");
  *n_hi = h[0];
  *n_lo = nt - h[0];
  *lo   = c->d_hr_list.p;
  *hi   = c->d_hr_list.p + nt;
  return AGRAV_OK;
}

}  // namespace agrav
