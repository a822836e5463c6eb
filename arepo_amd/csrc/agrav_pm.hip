// agrav_pm.hip -- periodic particle-mesh long-range force
// (long_range_force, longrange.c:98-198; pmforce_periodic, pm_periodic.c:1572-1936).
//
// Steps and their reference counterparts:
//   CIC deposit        pm_periodic.c:1049-1170  -> k_cic_deposit: 64-bit fixed-point
//                      atomics (order-independent, so the mesh is bit-reproducible
//                      whatever the atomic order), converted to FP32/FP64 afterwards;
//   r2c / c2r 3-D FFT  pm_mpi_fft.c:400-496     -> rocFFT out-of-place plans;
//   Green's function   pm_periodic.c:1626-1707  -> FP64 table built once, applied in FP64;
//   4-point FD         pm_periodic.c:1731-1800  -> k_fd4 (all three components at once);
//   CIC readout        pm_periodic.c:1322-1550  -> k_cic_readout from one {F_x, F_y, F_z, phi} mesh.
// FFTs are unnormalised like FFTW; 4 pi G / V and 1/(2d) are applied as in the reference.
#include "agrav_ctx.h"

#include <math.h>

namespace agrav
{

static constexpr int B = 256;
static inline unsigned nb(int64_t n) { return (unsigned)((n + B - 1) / B); }

__device__ inline double wrap_pm(double x, double L)
{
  if(x < 0)
    x += L;
  if(x >= L)
    x -= L;
  return x;
}

struct Cic
{
  int s[3], ss[3];
  double d[3];
};

// A position still outside [0, L) after one wrap (or not finite) is reported
// through `ok` and mapped to cell 0, so that no mesh access leaves the mesh.
__device__ inline Cic cic_of(const double *p, int G, double L, bool &ok)
{
  // slab = (int)(x*PMGRID/L), dx = x*PMGRID/L - slab, periodic neighbour (pm_periodic.c:1051-1107)
  Cic c;
  const double to_slab = G / L;
  ok                   = true;
  for(int k = 0; k < 3; k++)
    {
      double x = wrap_pm(p[k], L);
      if(!(x >= 0 && x < L))
        {
          ok = false;
          x  = 0;
        }
      c.s[k] = (int)(to_slab * x);
      c.d[k] = to_slab * x - c.s[k];
      if(c.s[k] >= G)
        c.s[k] -= G;
      c.ss[k] = c.s[k] + 1;
      if(c.ss[k] >= G)
        c.ss[k] -= G;
    }
  return c;
}
__device__ inline Cic cic_of(const double *p, int G, double L)
{
  bool ok;
  return cic_of(p, G, L, ok);
}

__global__ void k_cic_deposit(const double *__restrict__ pos, const double *__restrict__ mass, int64_t n, int G, double L,
                              double scale, unsigned long long *rho, int *bad)
{
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  bool ok;
  Cic c              = cic_of(&pos[3 * i], G, L, ok);
  if(!ok)
    {
      atomicOr(bad, 1);
      return;
    }
  double m           = mass[i] * scale;
  const double *d    = c.d;
#define FI(x, y, z) (((size_t)(x)*G + (y)) * G + (z))
  atomicAdd(&rho[FI(c.s[0], c.s[1], c.s[2])], (unsigned long long)llrint(m * (1.0 - d[0]) * (1.0 - d[1]) * (1.0 - d[2])));
  atomicAdd(&rho[FI(c.s[0], c.s[1], c.ss[2])], (unsigned long long)llrint(m * (1.0 - d[0]) * (1.0 - d[1]) * (d[2])));
  atomicAdd(&rho[FI(c.s[0], c.ss[1], c.s[2])], (unsigned long long)llrint(m * (1.0 - d[0]) * (d[1]) * (1.0 - d[2])));
  atomicAdd(&rho[FI(c.s[0], c.ss[1], c.ss[2])], (unsigned long long)llrint(m * (1.0 - d[0]) * (d[1]) * (d[2])));
  atomicAdd(&rho[FI(c.ss[0], c.s[1], c.s[2])], (unsigned long long)llrint(m * (d[0]) * (1.0 - d[1]) * (1.0 - d[2])));
  atomicAdd(&rho[FI(c.ss[0], c.s[1], c.ss[2])], (unsigned long long)llrint(m * (d[0]) * (1.0 - d[1]) * (d[2])));
  atomicAdd(&rho[FI(c.ss[0], c.ss[1], c.s[2])], (unsigned long long)llrint(m * (d[0]) * (d[1]) * (1.0 - d[2])));
  atomicAdd(&rho[FI(c.ss[0], c.ss[1], c.ss[2])], (unsigned long long)llrint(m * (d[0]) * (d[1]) * (d[2])));
#undef FI
}

// Deposit of Hilbert-ordered particles: the 256 particles of a block occupy a
// compact region, so their CIC footprint is first accumulated in an LDS tile with
// 64-bit LDS atomics and then added to the mesh with one global atomic per
// touched cell.  Blocks whose footprint does not fit the tile use global atomics.
static constexpr int kTile = 4096;

__device__ inline int wrap_cell(int v, int G) { return v < 0 ? v + G : (v >= G ? v - G : v); }

__global__ __launch_bounds__(B) void k_cic_deposit_sorted(const double *__restrict__ px, const double *__restrict__ py,
                                                          const double *__restrict__ pz, const double *__restrict__ pm,
                                                          int64_t n, int G, double L, double scale, unsigned long long *rho)
{
  __shared__ unsigned long long s_tile[kTile];
  __shared__ int s_lo[3], s_hi[3];
  const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  const bool has  = i < n;
  const int64_t i0 = (int64_t)blockIdx.x * B;
  double p[3]     = {0, 0, 0};
  if(has)
    {
      p[0] = px[i];
      p[1] = py[i];
      p[2] = pz[i];
    }
  Cic c = cic_of(p, G, L);
  // cell offsets relative to the block's first particle, unwrapped to (-G/2, G/2]
  double q[3] = {px[i0], py[i0], pz[i0]};
  Cic c0      = cic_of(q, G, L);
  int off[3];
  for(int k = 0; k < 3; k++)
    {
      int d = c.s[k] - c0.s[k];
      if(d > G / 2)
        d -= G;
      if(d <= -G / 2)
        d += G;
      off[k] = d;
    }
  if(threadIdx.x < 3)
    {
      s_lo[threadIdx.x] = 1 << 30;
      s_hi[threadIdx.x] = -(1 << 30);
    }
  __syncthreads();
  // the block's footprint: wave reductions, then one LDS atomic per wave (atomics of
  // all lanes on the same word serialise)
  const int lane = threadIdx.x & 63;
  for(int k = 0; k < 3; k++)
    {
      int lo = has ? off[k] : (1 << 30), hi = has ? off[k] : -(1 << 30);
#pragma unroll
      for(int o = 32; o > 0; o >>= 1)
        {
          lo = min(lo, __shfl_xor(lo, o, 64));
          hi = max(hi, __shfl_xor(hi, o, 64));
        }
      if(lane == 0)
        {
          atomicMin(&s_lo[k], lo);
          atomicMax(&s_hi[k], hi);
        }
    }
  __syncthreads();
  const int ex = s_hi[0] - s_lo[0] + 2, ey = s_hi[1] - s_lo[1] + 2, ez = s_hi[2] - s_lo[2] + 2;
  const bool tiled = (int64_t)ex * ey * ez <= kTile && ex <= G && ey <= G && ez <= G;
  const double m     = has ? pm[i] * scale : 0.0;
  const double *d    = c.d;
  const double w[8]  = {(1.0 - d[0]) * (1.0 - d[1]) * (1.0 - d[2]), (1.0 - d[0]) * (1.0 - d[1]) * (d[2]),
                        (1.0 - d[0]) * (d[1]) * (1.0 - d[2]),       (1.0 - d[0]) * (d[1]) * (d[2]),
                        (d[0]) * (1.0 - d[1]) * (1.0 - d[2]),       (d[0]) * (1.0 - d[1]) * (d[2]),
                        (d[0]) * (d[1]) * (1.0 - d[2]),             (d[0]) * (d[1]) * (d[2])};
  if(tiled)
    {
      for(int t = threadIdx.x; t < ex * ey * ez; t += B)
        s_tile[t] = 0;
      __syncthreads();
      if(has)
        {
          const int bx = off[0] - s_lo[0], by = off[1] - s_lo[1], bz = off[2] - s_lo[2];
#pragma unroll
          for(int q8 = 0; q8 < 8; q8++)
            {
              int ox = (q8 >> 2) & 1, oy = (q8 >> 1) & 1, oz = q8 & 1;
              int t  = ((bx + ox) * ey + (by + oy)) * ez + (bz + oz);
              atomicAdd(&s_tile[t], (unsigned long long)llrint(m * w[q8]));
            }
        }
      __syncthreads();
      for(int t = threadIdx.x; t < ex * ey * ez; t += B)
        {
          unsigned long long v = s_tile[t];
          if(v)
            {
              int tz = t % ez, ty = (t / ez) % ey, tx = t / (ez * ey);
              int gx = wrap_cell(c0.s[0] + s_lo[0] + tx, G), gy = wrap_cell(c0.s[1] + s_lo[1] + ty, G),
                  gz = wrap_cell(c0.s[2] + s_lo[2] + tz, G);
              atomicAdd(&rho[((size_t)gx * G + gy) * G + gz], v);
            }
        }
    }
  else if(has)
    {
#pragma unroll
      for(int q8 = 0; q8 < 8; q8++)
        {
          int gx = ((q8 >> 2) & 1) ? c.ss[0] : c.s[0], gy = ((q8 >> 1) & 1) ? c.ss[1] : c.s[1],
              gz = (q8 & 1) ? c.ss[2] : c.s[2];
          atomicAdd(&rho[((size_t)gx * G + gy) * G + gz], (unsigned long long)llrint(m * w[q8]));
        }
    }
}

// one 4-component mesh {F_x, F_y, F_z, phi} per cell: a particle's readout is 8 vector loads
template <typename R> struct Mesh4
{
  R v[4];
};

template <typename R>
__device__ inline void cic_gather4(const Cic &c, int G, const Mesh4<R> *__restrict__ mesh, double &f0, double &f1, double &f2,
                                   double &pp)
{
  const double *d = c.d;
  const int64_t idx[8] = {((int64_t)c.s[0] * G + c.s[1]) * G + c.s[2],   ((int64_t)c.s[0] * G + c.s[1]) * G + c.ss[2],
                          ((int64_t)c.s[0] * G + c.ss[1]) * G + c.s[2],  ((int64_t)c.s[0] * G + c.ss[1]) * G + c.ss[2],
                          ((int64_t)c.ss[0] * G + c.s[1]) * G + c.s[2],  ((int64_t)c.ss[0] * G + c.s[1]) * G + c.ss[2],
                          ((int64_t)c.ss[0] * G + c.ss[1]) * G + c.s[2], ((int64_t)c.ss[0] * G + c.ss[1]) * G + c.ss[2]};
  const double w[8] = {(1.0 - d[0]) * (1.0 - d[1]) * (1.0 - d[2]), (1.0 - d[0]) * (1.0 - d[1]) * (d[2]),
                       (1.0 - d[0]) * (d[1]) * (1.0 - d[2]),       (1.0 - d[0]) * (d[1]) * (d[2]),
                       (d[0]) * (1.0 - d[1]) * (1.0 - d[2]),       (d[0]) * (1.0 - d[1]) * (d[2]),
                       (d[0]) * (d[1]) * (1.0 - d[2]),             (d[0]) * (d[1]) * (d[2])};
  pp = f0 = f1 = f2 = 0;
#pragma unroll
  for(int q = 0; q < 8; q++)
    {
      const Mesh4<R> m = mesh[idx[q]];
      f0 += (double)m.v[0] * w[q];
      f1 += (double)m.v[1] * w[q];
      f2 += (double)m.v[2] * w[q];
      pp += (double)m.v[3] * w[q];
    }
}

template <typename R>
__global__ void k_cic_readout_sorted(const double *__restrict__ px, const double *__restrict__ py, const double *__restrict__ pz,
                                     int64_t n, int G, double L, const Mesh4<R> *__restrict__ mesh, double fac_pot, double *acc_pm,
                                     double *pot_pm)
{
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  double p[3] = {px[i], py[i], pz[i]};
  double f0, f1, f2, pp;
  cic_gather4<R>(cic_of(p, G, L), G, mesh, f0, f1, f2, pp);
  pot_pm[i]         = pp * fac_pot;
  acc_pm[3 * i + 0] = f0;
  acc_pm[3 * i + 1] = f1;
  acc_pm[3 * i + 2] = f2;
}

template <typename R>
__global__ void k_fix_to_real(const unsigned long long *__restrict__ rho, int64_t g3, double scale, R *out)
{
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i < g3)
    out[i] = (R)((double)rho[i] / scale);
}

__global__ void k_green_table(int G, double L, double asmth, double *green)
{
  // pm_periodic.c:1626-1707; index = kz + (G/2+1)*(y + G*x)
  int64_t i   = (int64_t)blockIdx.x * B + threadIdx.x;
  const int gz = G / 2 + 1;
  if(i >= (int64_t)G * G * gz)
    return;
  int z = (int)(i % gz), y = (int)((i / gz) % G), x = (int)(i / ((int64_t)gz * G));
  double d     = L / G;
  double dhalf = 0.5 * d;
  double kfac  = 2.0 * M_PI / (G * d);
  int xx = x >= G / 2 ? x - G : x, yy = y >= G / 2 ? y - G : y, zz = z >= G / 2 ? z - G : z;
  double kx = kfac * xx, ky = kfac * yy, kz = kfac * zz;
  double k2 = kx * kx + ky * ky + kz * kz;
  double v  = 0;
  if(k2 > 0)
    {
      double smth = -exp(-k2 * asmth * asmth) / k2;
      double fx = 1, fy = 1, fz = 1;
      if(xx != 0)
        {
          fx = kx * dhalf;
          fx = sin(fx) / fx;
        }
      if(yy != 0)
        {
          fy = ky * dhalf;
          fy = sin(fy) / fy;
        }
      if(zz != 0)
        {
          fz = kz * dhalf;
          fz = sin(fz) / fz;
        }
      double ff = 1 / (fx * fy * fz);
      v         = smth * ff * ff * ff * ff;
    }
  green[i] = v;  // k = 0 -> 0 (pm_periodic.c:1706-1707)
}

template <typename R> __global__ void k_green_apply(R *cplx, const double *__restrict__ green, int64_t nc)
{
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= nc)
    return;
  double g        = green[i];
  cplx[2 * i]     = (R)((double)cplx[2 * i] * g);
  cplx[2 * i + 1] = (R)((double)cplx[2 * i + 1] * g);
}

template <typename R> __global__ void k_fd4(const R *__restrict__ phi, int G, double fac, Mesh4<R> *mesh)
{
  // F = fac/(2d) [ 4/3 (phi_{-1} - phi_{+1}) - 1/6 (phi_{-2} - phi_{+2}) ]   (fac includes 1/(2d))
  int64_t i  = (int64_t)blockIdx.x * B + threadIdx.x;
  int64_t g3 = (int64_t)G * G * G;
  if(i >= g3)
    return;
  int z = (int)(i % G), y = (int)((i / G) % G), x = (int)(i / ((int64_t)G * G));
  auto at = [&](int a, int b, int c) -> double { return (double)phi[((int64_t)a * G + b) * G + c]; };
  int xl = (x - 1 + G) % G, xr = (x + 1) % G, xll = (x - 2 + G) % G, xrr = (x + 2) % G;
  int yl = (y - 1 + G) % G, yr = (y + 1) % G, yll = (y - 2 + G) % G, yrr = (y + 2) % G;
  int zl = (z - 1 + G) % G, zr = (z + 1) % G, zll = (z - 2 + G) % G, zrr = (z + 2) % G;
  Mesh4<R> m;
  m.v[0]  = (R)(fac * ((4.0 / 3) * (at(xl, y, z) - at(xr, y, z)) - (1.0 / 6) * (at(xll, y, z) - at(xrr, y, z))));
  m.v[1]  = (R)(fac * ((4.0 / 3) * (at(x, yl, z) - at(x, yr, z)) - (1.0 / 6) * (at(x, yll, z) - at(x, yrr, z))));
  m.v[2]  = (R)(fac * ((4.0 / 3) * (at(x, y, zl) - at(x, y, zr)) - (1.0 / 6) * (at(x, y, zll) - at(x, y, zrr))));
  m.v[3]  = phi[i];
  mesh[i] = m;
}

template <typename R>
__global__ void k_cic_readout(const double *__restrict__ pos, int64_t n, int G, double L, const Mesh4<R> *__restrict__ mesh,
                              double fac_pot, double *acc_pm, double *pot_pm)
{
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if(i >= n)
    return;
  double f0, f1, f2, pp;
  bool ok;  // positions outside the box were reported by the deposit; read cell 0 instead
  cic_gather4<R>(cic_of(&pos[3 * i], G, L, ok), G, mesh, f0, f1, f2, pp);
  pot_pm[i]         = pp * fac_pot;
  acc_pm[3 * i + 0] = f0;
  acc_pm[3 * i + 1] = f1;
  acc_pm[3 * i + 2] = f2;
}

static agrav_status fft_check(agrav_ctx *c, rocfft_status s, const char *what)
{
  if(s != rocfft_status_success)
    return fail(c, AGRAV_ERR_FFT, std::string("rocFFT: ") + what);
  return AGRAV_OK;
}

void pm_free(agrav_ctx *c)
{
  PmState &p = c->pm;
  if(p.fwd)
    rocfft_plan_destroy(p.fwd);
  if(p.bwd)
    rocfft_plan_destroy(p.bwd);
  if(p.info_f)
    rocfft_execution_info_destroy(p.info_f);
  if(p.info_b)
    rocfft_execution_info_destroy(p.info_b);
  p.fwd = p.bwd = nullptr;
  p.info_f = p.info_b = nullptr;
  p.work.release();
  p.rho_fix.release();
  p.real.release();
  p.cplx.release();
  p.force.release();
  p.green.release();
  p.grid = 0;
}

agrav_status pm_init(agrav_ctx *c)
{
  // pm_init_periodic, pm_periodic.c:173-255
  const int G = c->par.pmgrid;
  if(G <= 0)
    return AGRAV_OK;
  PmState &p = c->pm;
  if(p.grid == G && p.prec == c->par.precision)
    return AGRAV_OK;
  pm_free(c);
  static bool setup_done = false;
  if(!setup_done)
    {
      if(rocfft_setup() != rocfft_status_success)
        return fail(c, AGRAV_ERR_FFT, "rocfft_setup failed");
      setup_done = true;
    }
  const bool sp       = c->par.precision == AGRAV_FP32;
  const size_t rsz    = sp ? sizeof(float) : sizeof(double);
  const int64_t g3    = (int64_t)G * G * G;
  const int64_t nc    = (int64_t)G * G * (G / 2 + 1);
  const size_t len[3] = {(size_t)G, (size_t)G, (size_t)G};
  rocfft_precision pr = sp ? rocfft_precision_single : rocfft_precision_double;
  agrav_status st;
  if((st = fft_check(c, rocfft_plan_create(&p.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, pr, 3, len, 1,
                                           nullptr),
                     "plan forward")))
    return st;
  if((st = fft_check(c, rocfft_plan_create(&p.bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, pr, 3, len, 1,
                                           nullptr),
                     "plan inverse")))
    return st;
  size_t w1 = 0, w2 = 0;
  rocfft_plan_get_work_buffer_size(p.fwd, &w1);
  rocfft_plan_get_work_buffer_size(p.bwd, &w2);
  size_t w = w1 > w2 ? w1 : w2;
  if((st = p.work.ensure(w ? w : 1)) || (st = p.rho_fix.ensure(g3)) || (st = p.real.ensure(g3 * rsz)) ||
     (st = p.cplx.ensure(nc * 2 * rsz)) || (st = p.force.ensure(4 * g3 * rsz)) || (st = p.green.ensure(nc)))
    return st;
  if((st = fft_check(c, rocfft_execution_info_create(&p.info_f), "info")) ||
     (st = fft_check(c, rocfft_execution_info_create(&p.info_b), "info")))
    return st;
  if(w)
    {
      rocfft_execution_info_set_work_buffer(p.info_f, p.work.p, w);
      rocfft_execution_info_set_work_buffer(p.info_b, p.work.p, w);
    }
  rocfft_execution_info_set_stream(p.info_f, c->stream);
  rocfft_execution_info_set_stream(p.info_b, c->stream);
  double asmth = c->par.asmth * c->par.box_size / G;  // pm_periodic.c:190
  k_green_table<<<nb(nc), B, 0, c->stream>>>(G, c->par.box_size, asmth, p.green.p);
  AG_HIP(hipGetLastError());
  p.grid = G;
  p.prec = c->par.precision;
  return AGRAV_OK;
}

// sorted: the particles are in the Hilbert order of a current full tree (px, py, pz, pm)
template <typename R> static agrav_status pm_run(agrav_ctx *c, hipStream_t s, bool sorted)
{
  PmState &p      = c->pm;
  const int G     = p.grid;
  const double L  = c->par.box_size;
  const int64_t n = c->n;
  const int64_t g3 = (int64_t)G * G * G;
  const int64_t nc = (int64_t)G * G * (G / 2 + 1);
  agrav_status st;
  if((st = c->d_acc_pm.ensure(3 * (n ? n : 1))) || (st = c->d_pot_pm.ensure(n ? n : 1)))
    return st;
  // fixed-point scale of the mass mesh: a power of two leaving two bits of headroom
  // below 2^63 for the total mass (kept on the host since the particle load)
  int e = 0;
  frexp(c->mass_total > 0 ? c->mass_total : 1.0, &e);
  const double scale = ldexp(1.0, 61 - e);
  AG_HIP(hipMemsetAsync(p.rho_fix.p, 0, g3 * sizeof(unsigned long long), s));
  // with a current tree the particles are available in Hilbert order: tiled deposit and
  // coherent readout, results kept in sorted order (the finalize kernel maps them back)
  if(n > 0 && sorted)
    k_cic_deposit_sorted<<<nb(n), B, 0, s>>>(c->d_px.p, c->d_py.p, c->d_pz.p, c->d_pm.p, n, G, L, scale, p.rho_fix.p);
  int *d_bad = nullptr;
  if(n > 0 && !sorted)
    {
      // unsorted input has not passed the tree build's box check: the deposit checks it
      if((st = c->d_counters.ensure(128)))
        return st;
      d_bad = c->d_counters.p + 101;
      AG_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
      k_cic_deposit<<<nb(n), B, 0, s>>>(c->d_pos_in.p, c->d_mass_in.p, n, G, L, scale, p.rho_fix.p, d_bad);
    }
  k_fix_to_real<R><<<nb(g3), B, 0, s>>>(p.rho_fix.p, g3, scale, (R *)p.real.p);
  AG_HIP(hipGetLastError());
  void *in[1]  = {p.real.p};
  void *out[1] = {p.cplx.p};
  rocfft_execution_info_set_stream(p.info_f, s);
  rocfft_execution_info_set_stream(p.info_b, s);
  if((st = fft_check(c, rocfft_execute(p.fwd, in, out, p.info_f), "execute forward")))
    return st;
  k_green_apply<R><<<nb(nc), B, 0, s>>>((R *)p.cplx.p, p.green.p, nc);
  AG_HIP(hipGetLastError());
  void *in2[1]  = {p.cplx.p};
  void *out2[1] = {p.real.p};
  if((st = fft_check(c, rocfft_execute(p.bwd, in2, out2, p.info_b), "execute inverse")))
    return st;
  const double fac_pot = 4 * M_PI * c->par.G / (L * L * L);  // pm_periodic.c:1590
  const double fac_fd  = fac_pot / (2 * (L / G));            // :1592
  Mesh4<R> *mesh = reinterpret_cast<Mesh4<R> *>(p.force.p);
  k_fd4<R><<<nb(g3), B, 0, s>>>((const R *)p.real.p, G, fac_fd, mesh);
  if(n > 0 && sorted)
    k_cic_readout_sorted<R><<<nb(n), B, 0, s>>>(c->d_px.p, c->d_py.p, c->d_pz.p, n, G, L, mesh, fac_pot, c->d_acc_pm.p,
                                                c->d_pot_pm.p);
  else if(n > 0)
    k_cic_readout<R><<<nb(n), B, 0, s>>>(c->d_pos_in.p, n, G, L, mesh, fac_pot, c->d_acc_pm.p, c->d_pot_pm.p);
  AG_HIP(hipGetLastError());
  if(d_bad)
    {
      int bad = 0;
      AG_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      if(bad)
        return fail(c, AGRAV_ERR_OUT_OF_BOX, "particle outside the periodic box after one wrap");
    }
  c->pm_sorted = sorted;
  return AGRAV_OK;
}

// the PM's duration is read from its events once they have passed, so that the
// host does not wait for the PM before it queues the walk
agrav_status pm_time_settle(agrav_ctx *c)
{
  if(!c->pm_time_pending)
    return AGRAV_OK;
  AG_HIP(hipEventSynchronize(c->ev[6]));
  float ms = 0;
  AG_HIP(hipEventElapsedTime(&ms, c->ev[5], c->ev[6]));
  c->tim.pm_grav      = ms;
  c->pm_time_pending = false;
  return AGRAV_OK;
}

// the periodic grid on stream s; the high-resolution grid follows on the main stream
static agrav_status pm_force_on(agrav_ctx *c, hipStream_t s, bool sorted)
{
  if(c->par.pmgrid <= 0)
    return AGRAV_OK;
  agrav_status st = pm_init(c);
  if(st != AGRAV_OK || (st = pm_time_settle(c)) != AGRAV_OK)
    return st;
  AG_HIP(hipEventRecord(c->ev[5], s));
  st = (c->par.precision == AGRAV_FP32) ? pm_run<float>(c, s, sorted) : pm_run<double>(c, s, sorted);
  // the high-resolution grid after the periodic one (longrange.c:129-140)
  if(st == AGRAV_OK && hr_on(c))
    st = hr_pm(c);
  if(st != AGRAV_OK)
    return st;
  AG_HIP(hipEventRecord(c->ev[6], s));
  c->pm_time_pending = true;
  c->have_pm         = true;
  return AGRAV_OK;
}

agrav_status pm_force(agrav_ctx *c)
{
  // a partial tree lacks the inactive particles: the PM then reads caller order
  return pm_force_on(c, c->stream, c->have_tree && !c->tree_partial);
}

// a full step's PM on the PM stream, once the build's gather has written the sorted
// particles (the build calls this right after the gather; the caller joins ev_join)
agrav_status pm_fork(agrav_ctx *c)
{
  if(!c->pm_stream || hr_on(c))
    return fail(c, AGRAV_ERR_STATE, "no PM stream for this context");
  AG_HIP(hipEventRecord(c->ev_fork, c->stream));
  AG_HIP(hipStreamWaitEvent(c->pm_stream, c->ev_fork, 0));
  c->pm_forked    = true;  // from here on the main stream must join, whatever happens
  agrav_status st = pm_force_on(c, c->pm_stream, true);
  AG_HIP(hipEventRecord(c->ev_join, c->pm_stream));
  return st;
}

}  // namespace agrav
