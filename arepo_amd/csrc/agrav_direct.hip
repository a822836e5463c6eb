// agrav_direct.hip -- FP64 direct summation and the Ewald tables.
//
//  * k_ewald_tree_table: the 65^3 x 4 periodic correction table of
//    ewald_init (forcetree_ewald.c:89-252) with ewald_psi (:393-450) and
//    ewald_force (:463-527), one lane per table point, FP64.
//  * k_ewald_ft_table: the 129^3 x 4 FORCETEST table
//    (gravtree_forcetest.c:933-1008) from the exact series (:820-916).
//  * k_forcetest: gravity_forcetest_evaluate (gravtree_forcetest.c:421-648):
//    every selected target against every particle, sources staged in LDS,
//    split over source chunks with per-chunk partial sums reduced in a fixed
//    order (deterministic).
//  * k_direct: force_evaluate_direct (forcetree_walk.c:557-708) for all pairs.
#include "agrav_ctx.h"

#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

namespace agrav
{

#define SF1 10.666666666667
#define SF2 32.0
#define SF3 (-38.4)
#define SF4 (-2.8)
#define SF5 5.333333333333
#define SF6 6.4
#define SF7 (-9.6)
#define SF8 21.333333333333
#define SF9 (-48.0)
#define SF10 38.4
#define SF11 (-10.666666666667)
#define SF12 (-0.066666666667)
#define SF13 (-3.2)
#define SF14 0.066666666667
#define SF15 (-16.0)
#define SF16 9.6
#define SF17 (-2.133333333333)

static constexpr int B = 256;

__device__ double ewald_psi_d(double x, double y, double z)
{
  double r = sqrt(x * x + y * y + z * z);
  if(r == 0)
    return 0;
  const double alpha = 3.0;
  double sum1        = 0;
  for(int nx = -4; nx <= 4; nx++)
    for(int ny = -4; ny <= 4; ny++)
      for(int nz = -4; nz <= 4; nz++)
        {
          double dx = x - nx, dy = y - ny, dz = z - nz;
          double rr = sqrt(dx * dx + dy * dy + dz * dz);
          sum1 += erfc(alpha * rr) / rr;
        }
  const int kmax = (int)(2 * alpha + 0.5);
  double sum2    = 0;
  for(int nx = -kmax; nx <= kmax; nx++)
    for(int ny = -kmax; ny <= kmax; ny++)
      for(int nz = -kmax; nz <= kmax; nz++)
        {
          double kx = 2.0 * M_PI * nx, ky = 2.0 * M_PI * ny, kz = 2.0 * M_PI * nz;
          double k2 = kx * kx + ky * ky + kz * kz;
          if(k2 > 0)
            sum2 += 4.0 * M_PI / k2 * exp(-k2 / (4.0 * alpha * alpha)) * cos(x * kx + y * ky + z * kz);
        }
  return M_PI / (alpha * alpha) - sum1 - sum2 + 1.0 / r;
}

__device__ void ewald_force_d(double x, double y, double z, double f[3])
{
  f[0] = f[1] = f[2] = 0;
  double r2          = x * x + y * y + z * z;
  if(r2 == 0)
    return;
  const double alpha = 2.0, alpha2 = alpha * alpha;
  double r3inv = 1.0 / (r2 * sqrt(r2));
  f[0] += r3inv * x;
  f[1] += r3inv * y;
  f[2] += r3inv * z;
  for(int nx = -4; nx <= 4; nx++)
    for(int ny = -4; ny <= 4; ny++)
      for(int nz = -4; nz <= 4; nz++)
        {
          double dx = x - nx, dy = y - ny, dz = z - nz;
          double rr2 = dx * dx + dy * dy + dz * dz, rr = sqrt(rr2);
          double val = erfc(alpha * rr) + 2.0 * alpha * rr / sqrt(M_PI) * exp(-alpha2 * rr2);
          double v2  = val / (rr2 * rr);
          f[0] -= dx * v2;
          f[1] -= dy * v2;
          f[2] -= dz * v2;
        }
  const int hmax = (int)(2 * alpha + 0.5);
  for(int hx = -hmax; hx <= hmax; hx++)
    for(int hy = -hmax; hy <= hmax; hy++)
      for(int hz = -hmax; hz <= hmax; hz++)
        {
          double h2 = hx * hx + hy * hy + hz * hz;
          if(h2 > 0)
            {
              double val = 2.0 / h2 * exp(-M_PI * M_PI * h2 / alpha2) * sin(2.0 * M_PI * (x * hx + y * hy + z * hz));
              f[0] -= hx * val;
              f[1] -= hy * val;
              f[2] -= hz * val;
            }
        }
}

__global__ void k_ewald_tree_table(double box, double *tab, float *tabf)
{
  const int en = 64, n1 = 65, size = n1 * n1 * n1;
  int n = blockIdx.x * B + threadIdx.x;
  if(n >= size)
    return;
  int i = n / (n1 * n1), j = (n / n1) % n1, k = n % n1;
  double xx = 0.5 * i / en, yy = 0.5 * j / en, zz = 0.5 * k / en;
  double f[3];
  ewald_force_d(xx, yy, zz, f);
  double v[4] = {f[0] / (box * box), f[1] / (box * box), f[2] / (box * box), ewald_psi_d(xx, yy, zz) / box};
  for(int c = 0; c < 4; c++)
    {
      tab[c * size + n]  = v[c];
      tabf[c * size + n] = (float)v[c];
    }
}

__device__ void forcetest_exact_d(double box, double x, double y, double z, double f[4])
{
  for(int i = 0; i < 4; i++)
    f[i] = 0;
  double r2 = x * x + y * y + z * z;
  if(r2 == 0)
    return;
  double sx = -1, sy = -1, sz = -1;
  if(x < 0)
    {
      x  = -x;
      sx = 1;
    }
  if(y < 0)
    {
      y  = -y;
      sy = 1;
    }
  if(z < 0)
    {
      z  = -z;
      sz = 1;
    }
  const double alpha = 2.0 / box, alpha2 = alpha * alpha;
  double r     = sqrt(r2);
  double r3inv = 1.0 / (r2 * r);
  f[0] += r3inv * x;
  f[1] += r3inv * y;
  f[2] += r3inv * z;
  for(int nx = -6; nx <= 6; nx++)
    for(int ny = -6; ny <= 6; ny++)
      for(int nz = -6; nz <= 6; nz++)
        {
          double dx = x - nx * box, dy = y - ny * box, dz = z - nz * box;
          double rr2 = dx * dx + dy * dy + dz * dz, rr = sqrt(rr2);
          double ea  = erfc(alpha * rr);
          double val = ea + 2.0 * alpha * rr / sqrt(M_PI) * exp(-alpha2 * rr2);
          double v2  = val / (rr2 * rr);
          f[0] -= dx * v2;
          f[1] -= dy * v2;
          f[2] -= dz * v2;
          f[3] -= ea / rr;
        }
  const int kmax    = (int)(4 * alpha * box + 0.5);
  const double vol  = box * box * box;
  const double kfac = 2.0 * M_PI / box;
  for(int nx = -kmax; nx <= kmax; nx++)
    for(int ny = -kmax; ny <= kmax; ny++)
      for(int nz = -kmax; nz <= kmax; nz++)
        {
          double kx = kfac * nx, ky = kfac * ny, kz = kfac * nz;
          double k2 = kx * kx + ky * ky + kz * kz;
          if(k2 > 0)
            {
              double kdotx = x * kx + y * ky + z * kz;
              double vv    = 4.0 * M_PI / (k2 * vol) * exp(-k2 / (4.0 * alpha2));
              double sn, cs;
              sincos(kdotx, &sn, &cs);
              f[0] -= kx * vv * sn;
              f[1] -= ky * vv * sn;
              f[2] -= kz * vv * sn;
              f[3] -= vv * cs;
            }
        }
  f[3] += M_PI / (alpha2 * vol) + 1.0 / r;
  f[0] *= sx;
  f[1] *= sy;
  f[2] *= sz;
}

__global__ void k_ewald_ft_table(double box, double *tab)
{
  const int tn = 128, n1 = 129;
  const int size = n1 * n1 * n1;
  int n          = blockIdx.x * B + threadIdx.x;
  if(n >= size)
    return;
  int i = n / (n1 * n1), j = (n / n1) % n1, k = n % n1;
  double f[4];
  forcetest_exact_d(box, 0.5 * i / tn * box, 0.5 * j / tn * box, 0.5 * k / tn * box, f);
  for(int c = 0; c < 4; c++)
    tab[(size_t)c * size + n] = f[c];
}

__device__ inline void ft_lookup(const double *__restrict__ tab, double box, double dx, double dy, double dz, double f[4])
{
  // gravtree_forcetest.c:1019-1085
  const int tn = 128, n1 = 129, size = n1 * n1 * n1;
  double sx = 1, sy = 1, sz = 1;
  if(dx < 0)
    {
      dx = -dx;
      sx = -1;
    }
  if(dy < 0)
    {
      dy = -dy;
      sy = -1;
    }
  if(dz < 0)
    {
      dz = -dz;
      sz = -1;
    }
  double fi = 2.0 * tn / box;
  double u = dx * fi, v = dy * fi, w = dz * fi;
  int i = (int)u, j = (int)v, k = (int)w;
  i = i >= tn ? tn - 1 : i;
  j = j >= tn ? tn - 1 : j;
  k = k >= tn ? tn - 1 : k;
  u -= i;
  v -= j;
  w -= k;
  double ww[8] = {(1 - u) * (1 - v) * (1 - w), (1 - u) * (1 - v) * w, (1 - u) * v * (1 - w), (1 - u) * v * w,
                  u * (1 - v) * (1 - w),       u * (1 - v) * w,       u * v * (1 - w),       u * v * w};
  int o[8]     = {(i * n1 + j) * n1 + k,           (i * n1 + j) * n1 + k + 1,           (i * n1 + j + 1) * n1 + k,
                  (i * n1 + j + 1) * n1 + k + 1,   ((i + 1) * n1 + j) * n1 + k,         ((i + 1) * n1 + j) * n1 + k + 1,
                  ((i + 1) * n1 + j + 1) * n1 + k, ((i + 1) * n1 + j + 1) * n1 + k + 1};
  for(int c = 0; c < 4; c++)
    {
      double a = 0;
#pragma unroll
      for(int q = 0; q < 8; q++)
        a += tab[(size_t)c * size + o[q]] * ww[q];
      f[c] = a;
    }
  f[0] *= sx;
  f[1] *= sy;
  f[2] *= sz;
}

struct DirectArgs
{
  const double *pos;  // caller order, interleaved
  const double *mass;
  const uint8_t *stype;
  const double *hsoft;
  int n;
  int ns;
  const int *sel;     // nullptr -> all
  int periodic;
  double box, half;
  int pm;
  double asmth, vol;
  const uint8_t *hr;    // high-resolution flags (caller order) or null: Asmth[1] = hr_asmth for them
  double hr_asmth;
  const double *ewtab;  // FORCETEST table (periodic)
  int chunk;            // sources per chunk
  double *part;         // [nchunk][ns][8]
};

__device__ inline double near_d(double x, int per, double box, double half)
{
  if(!per)
    return x;
  return x > half ? x - box : (x < -half ? x + box : x);
}

__device__ inline double wrapb(double x, int per, double L)
{
  if(!per)
    return x;
  if(x < 0)
    x += L;
  if(x >= L)
    x -= L;
  return x;
}

// gravity_forcetest_evaluate, gravtree_forcetest.c:500-639
__global__ __launch_bounds__(B) void k_forcetest(DirectArgs a)
{
  __shared__ double s_x[B], s_y[B], s_z[B], s_m[B], s_h[B];
  const int q     = blockIdx.x * B + threadIdx.x;
  const int chunk = blockIdx.y;
  const bool has  = q < a.ns;
  const int i     = has ? (a.sel ? a.sel[q] : q) : 0;
  const double px = wrapb(a.pos[3 * (int64_t)i], a.periodic, a.box);
  const double py = wrapb(a.pos[3 * (int64_t)i + 1], a.periodic, a.box);
  const double pz = wrapb(a.pos[3 * (int64_t)i + 2], a.periodic, a.box);
  const double h_i = a.hsoft[a.stype ? a.stype[i] : 0];
  // gravtree_forcetest.c:480-483: Asmth[1] for a target in the high-resolution region
  const double asmth = (a.hr && a.hr[i]) ? a.hr_asmth : a.asmth;
  double acc[3] = {0, 0, 0}, accs[3] = {0, 0, 0}, pot = 0, pots = 0;
  const int j0 = chunk * a.chunk;
  const int j1 = min(a.n, j0 + a.chunk);
  for(int tb = j0; tb < j1; tb += B)
    {
      int j = tb + threadIdx.x;
      __syncthreads();
      if(j < j1)
        {
          s_x[threadIdx.x] = wrapb(a.pos[3 * (int64_t)j], a.periodic, a.box);
          s_y[threadIdx.x] = wrapb(a.pos[3 * (int64_t)j + 1], a.periodic, a.box);
          s_z[threadIdx.x] = wrapb(a.pos[3 * (int64_t)j + 2], a.periodic, a.box);
          s_m[threadIdx.x] = a.mass[j];
          s_h[threadIdx.x] = a.hsoft[a.stype ? a.stype[j] : 0];
        }
      __syncthreads();
      int cnt = min(B, j1 - tb);
      for(int k = 0; k < cnt; k++)
        {
          double h_j  = s_h[k];
          double hmax = h_j > h_i ? h_j : h_i;
          double dx   = near_d(s_x[k] - px, a.periodic, a.box, a.half);
          double dy   = near_d(s_y[k] - py, a.periodic, a.box, a.half);
          double dz   = near_d(s_z[k] - pz, a.periodic, a.box, a.half);
          double r2   = dx * dx + dy * dy + dz * dz;
          double m    = s_m[k];
          double r    = sqrt(r2);
          double fac_newton = 0, wp_newton = 0, fac, wp;
          if(r > 0)
            {
              fac_newton = m / (r2 * r);
              wp_newton  = -m / r;
            }
          if(r >= hmax)
            {
              fac = fac_newton;
              wp  = wp_newton;
            }
          else
            {
              double h_inv = 1.0 / hmax, h3_inv = h_inv * h_inv * h_inv, u = r * h_inv;
              if(u < 0.5)
                {
                  double u2 = u * u;
                  fac       = m * h3_inv * (SF1 + u2 * (SF2 * u + SF3));
                  wp        = m * h_inv * (SF4 + u2 * (SF5 + u2 * (SF6 * u + SF7)));
                }
              else
                {
                  double u2 = u * u, u3 = u2 * u;
                  fac = m * h3_inv * (SF8 + SF9 * u + SF10 * u2 + SF11 * u3 + SF12 / u3);
                  wp  = m * h_inv * (SF13 + SF14 / u + u2 * (SF1 + u * (SF15 + u * (SF16 + SF17 * u))));
                }
            }
          if(a.pm)
            {
              double u  = 0.5 / asmth * r;
              double ec = erfc(u);
              double ff = ec + 2.0 * u / sqrt(M_PI) * exp(-u * u) - 1.0;
              double fs = fac + fac_newton * ff;
              accs[0] += dx * fs;
              accs[1] += dy * fs;
              accs[2] += dz * fs;
              pots += wp + wp_newton * (ec - 1.0) + m * M_PI / (asmth * asmth * a.vol);
            }
          double fc[4] = {0, 0, 0, 0};
          if(a.periodic)
            ft_lookup(a.ewtab, a.box, dx, dy, dz, fc);
          acc[0] += dx * fac + m * fc[0];
          acc[1] += dy * fac + m * fc[1];
          acc[2] += dz * fac + m * fc[2];
          pot += wp + m * fc[3];
        }
    }
  if(has)
    {
      double *o = a.part + ((size_t)chunk * a.ns + q) * 8;
      o[0]      = acc[0];
      o[1]      = acc[1];
      o[2]      = acc[2];
      o[3]      = pot;
      o[4]      = accs[0];
      o[5]      = accs[1];
      o[6]      = accs[2];
      o[7]      = pots;
    }
}

__global__ void k_chunk_reduce(const double *__restrict__ part, int ns, int nchunk, double *out)
{
  int q = blockIdx.x * B + threadIdx.x;
  if(q >= ns)
    return;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for(int c = 0; c < nchunk; c++)
    for(int k = 0; k < 8; k++)
      v[k] += part[((size_t)c * ns + q) * 8 + k];
  for(int k = 0; k < 8; k++)
    out[(size_t)q * 8 + k] = v[k];
}

// force_evaluate_direct, forcetree_walk.c:557-708 (table erfc, tree Ewald table)
struct DirectAllArgs
{
  const double *pos, *mass, *hsoft;
  const uint8_t *stype;
  int nsoft;  // softening types in the table; a larger type gives NaN results
  int n, periodic, pm, ewald;
  double box, half, asmthfac, asmthinv2, ewald_fac;
  const uint8_t *hr;  // high-resolution flags or null: those targets take the Asmth[1] constants below
  double hr_asmthfac, hr_asmthinv2;
  const float *tab;
  const double *ewtab;
  int chunk;
  double *part;  // [nchunk][n][4]
};

__global__ __launch_bounds__(B) void k_direct(DirectAllArgs a)
{
  __shared__ double s_x[B], s_y[B], s_z[B], s_m[B], s_h[B];
  __shared__ float s_tab[256];
  s_tab[threadIdx.x] = a.tab[threadIdx.x];
  const int i      = blockIdx.x * B + threadIdx.x;
  const int chunk  = blockIdx.y;
  const bool has   = i < a.n;
  const int ii     = has ? i : 0;
  const double px = wrapb(a.pos[3 * (int64_t)ii], a.periodic, a.box), py = wrapb(a.pos[3 * (int64_t)ii + 1], a.periodic, a.box),
               pz = wrapb(a.pos[3 * (int64_t)ii + 2], a.periodic, a.box);
  // softening types index the table only in range; a type out of range makes
  // the results NaN (through the target's outputs or the source's mass)
  const int t_i    = a.stype ? a.stype[ii] : 0;
  const double h_i = a.hsoft[t_i < a.nsoft ? t_i : 0];
  // forcetree_walk.c:576-581: Asmth[1] for a target in the high-resolution region
  const bool hr_i         = a.hr && a.hr[ii];
  const double asmthfac  = hr_i ? a.hr_asmthfac : a.asmthfac;
  const double asmthinv2 = hr_i ? a.hr_asmthinv2 : a.asmthinv2;
  double ax = 0, ay = 0, az = 0, pot = 0;
  const int j0 = chunk * a.chunk, j1 = min(a.n, j0 + a.chunk);
  for(int tb = j0; tb < j1; tb += B)
    {
      int j = tb + threadIdx.x;
      __syncthreads();
      if(j < j1)
        {
          const int t_j    = a.stype ? a.stype[j] : 0;
          s_x[threadIdx.x] = wrapb(a.pos[3 * (int64_t)j], a.periodic, a.box);
          s_y[threadIdx.x] = wrapb(a.pos[3 * (int64_t)j + 1], a.periodic, a.box);
          s_z[threadIdx.x] = wrapb(a.pos[3 * (int64_t)j + 2], a.periodic, a.box);
          s_m[threadIdx.x] = t_j < a.nsoft ? a.mass[j] : __builtin_nan("");
          s_h[threadIdx.x] = a.hsoft[t_j < a.nsoft ? t_j : 0];
        }
      __syncthreads();
      int cnt = min(B, j1 - tb);
      for(int k = 0; k < cnt; k++)
        {
          double hmax = s_h[k] > h_i ? s_h[k] : h_i;
          double dx = near_d(s_x[k] - px, a.periodic, a.box, a.half), dy = near_d(s_y[k] - py, a.periodic, a.box, a.half),
                 dz = near_d(s_z[k] - pz, a.periodic, a.box, a.half);
          double r  = sqrt(dx * dx + dy * dy + dz * dz);
          double m  = s_m[k];
          double ff = 0, fp = 0, fac, wp;
          if(a.pm)
            {
              double te = asmthfac * r;
              int ti    = (int)te;
              if(ti >= AG_NTAB)
                continue;
              double w = te - ti;
              ff       = (1.0 - w) * s_tab[ti] + w * s_tab[ti + 1];
              fp       = (1.0 - w) * s_tab[128 + ti] + w * s_tab[128 + ti + 1];
            }
          if(r >= hmax)
            {
              double rinv = 1.0 / r;
              fac         = a.pm ? rinv * rinv * rinv + rinv * ff * asmthinv2 : rinv * rinv * rinv;
              wp          = a.pm ? rinv * fp : -rinv;
            }
          else
            {
              double h_inv = 1.0 / hmax, h3_inv = h_inv * h_inv * h_inv, u = r * h_inv;
              if(u < 0.5)
                {
                  double u2 = u * u;
                  fac       = h3_inv * (SF1 + u2 * (SF2 * u + SF3));
                  wp        = h_inv * (SF4 + u2 * (SF5 + u2 * (SF6 * u + SF7)));
                }
              else
                {
                  double u2 = u * u, u3 = u2 * u;
                  fac = h3_inv * (SF8 + SF9 * u + SF10 * u2 + SF11 * u3 + SF12 / u3);
                  wp  = h_inv * (SF13 + SF14 / u + u2 * (SF1 + u * (SF15 + u * (SF16 + SF17 * u))));
                }
              if(a.pm && r > 0)
                {
                  double rinv = 1.0 / r;
                  fac += rinv * ff * asmthinv2;
                  wp += rinv * (fp + 1.0);
                }
            }
          pot += m * wp;
          fac *= m;
          ax += dx * fac;
          ay += dy * fac;
          az += dz * fac;
          if(a.ewald)
            {
              // forcetree_ewald.c:268-382
              const int EN = 64, n1 = 65, size = n1 * n1 * n1;
              double sx = dx < 0 ? 1 : -1, sy = dy < 0 ? 1 : -1, sz = dz < 0 ? 1 : -1;
              double u = fabs(dx) * a.ewald_fac, v = fabs(dy) * a.ewald_fac, w = fabs(dz) * a.ewald_fac;
              int ei = (int)u, ej = (int)v, ek = (int)w;
              ei = ei >= EN ? EN - 1 : ei;
              ej = ej >= EN ? EN - 1 : ej;
              ek = ek >= EN ? EN - 1 : ek;
              u -= ei;
              v -= ej;
              w -= ek;
              double f8[8] = {(1 - u) * (1 - v) * (1 - w), (1 - u) * (1 - v) * w, (1 - u) * v * (1 - w), (1 - u) * v * w,
                              u * (1 - v) * (1 - w),       u * (1 - v) * w,       u * v * (1 - w),       u * v * w};
              int o[8] = {(ei * n1 + ej) * n1 + ek,           (ei * n1 + ej) * n1 + ek + 1,
                          (ei * n1 + ej + 1) * n1 + ek,       (ei * n1 + ej + 1) * n1 + ek + 1,
                          ((ei + 1) * n1 + ej) * n1 + ek,     ((ei + 1) * n1 + ej) * n1 + ek + 1,
                          ((ei + 1) * n1 + ej + 1) * n1 + ek, ((ei + 1) * n1 + ej + 1) * n1 + ek + 1};
              double c0 = 0, c1 = 0, c2 = 0, c3 = 0;
              for(int qq = 0; qq < 8; qq++)
                {
                  c0 += a.ewtab[o[qq]] * f8[qq];
                  c1 += a.ewtab[size + o[qq]] * f8[qq];
                  c2 += a.ewtab[2 * size + o[qq]] * f8[qq];
                  c3 += a.ewtab[3 * size + o[qq]] * f8[qq];
                }
              ax += m * sx * c0;
              ay += m * sy * c1;
              az += m * sz * c2;
              pot += m * c3;
            }
        }
    }
  if(has)
    {
      if(t_i >= a.nsoft)
        ax = ay = az = pot = __builtin_nan("");
      double *o = a.part + ((size_t)chunk * a.n + i) * 4;
      o[0]      = ax;
      o[1]      = ay;
      o[2]      = az;
      o[3]      = pot;
    }
}

__global__ void k_chunk_reduce4(const double *__restrict__ part, int n, int nchunk, double *out)
{
  int q = blockIdx.x * B + threadIdx.x;
  if(q >= n)
    return;
  double v[4] = {0, 0, 0, 0};
  for(int c = 0; c < nchunk; c++)
    for(int k = 0; k < 4; k++)
      v[k] += part[((size_t)c * n + q) * 4 + k];
  for(int k = 0; k < 4; k++)
    out[(size_t)q * 4 + k] = v[k];
}

agrav_status ewald_tree_table(agrav_ctx *c)
{
  const int size = 65 * 65 * 65;
  agrav_status st;
  if((st = c->d_ewald.ensure(4 * size)) || (st = c->d_ewald_f.ensure(4 * size)))
    return st;
  k_ewald_tree_table<<<(size + B - 1) / B, B, 0, c->stream>>>(c->par.box_size, c->d_ewald.p, c->d_ewald_f.p);
  AG_HIP(hipGetLastError());
  c->ewald_ready = true;
  return AGRAV_OK;
}

agrav_status ewald_forcetest_table(agrav_ctx *c)
{
  const int size = 129 * 129 * 129;
  agrav_status st;
  if((st = c->d_ewald_ft.ensure((size_t)4 * size)))
    return st;
  k_ewald_ft_table<<<(size + B - 1) / B, B, 0, c->stream>>>(c->par.box_size, c->d_ewald_ft.p);
  AG_HIP(hipGetLastError());
  c->ewald_ft_ready = true;
  return AGRAV_OK;
}

static int pick_chunks(int64_t ntarget_blocks, int64_t n)
{
  // enough blocks to fill 256 CUs several times over, chunks of at least 2048 sources
  int64_t want = (2048 + ntarget_blocks - 1) / ntarget_blocks;
  int64_t maxc = (n + 2047) / 2048;
  if(want > maxc)
    want = maxc;
  if(want < 1)
    want = 1;
  return (int)want;
}

// Batch of isolated sets: FP64 direct summation of each target over its own set
// (Newton + spline softening with the set's h, gravtree_forcetest.c:542-581; no
// periodic or short-range parts), one block per target, sources strided over the
// threads and reduced in a fixed order.
__global__ __launch_bounds__(B) void k_direct_batch(int ns, const int *__restrict__ sel, const double *__restrict__ pos,
                                                    const double *__restrict__ mass, const int *__restrict__ pset,
                                                    const int64_t *__restrict__ off, const double *__restrict__ sh, double *out)
{
  __shared__ double s_v[4][B];
  const int q = blockIdx.x;
  const int i = sel ? sel[q] : q;
  const int set = pset[i];
  const double px = pos[3 * (int64_t)i], py = pos[3 * (int64_t)i + 1], pz = pos[3 * (int64_t)i + 2];
  const double h = sh[set], h_inv = 1.0 / h, h3_inv = h_inv * h_inv * h_inv;
  double v[4] = {0, 0, 0, 0};
  for(int64_t j = off[set] + threadIdx.x; j < off[set + 1]; j += B)
    {
      const double dx = pos[3 * j] - px, dy = pos[3 * j + 1] - py, dz = pos[3 * j + 2] - pz;
      const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2), m = mass[j];
      double fac, wp;
      if(r >= h)
        {
          fac = m / (r2 * r);
          wp  = -m / r;
        }
      else
        {
          const double u = r * h_inv, u2 = u * u;
          if(u < 0.5)
            {
              fac = m * h3_inv * (SF1 + u2 * (SF2 * u + SF3));
              wp  = m * h_inv * (SF4 + u2 * (SF5 + u2 * (SF6 * u + SF7)));
            }
          else
            {
              const double u3 = u2 * u;
              fac = m * h3_inv * (SF8 + SF9 * u + SF10 * u2 + SF11 * u3 + SF12 / u3);
              wp  = m * h_inv * (SF13 + SF14 / u + u2 * (SF1 + u * (SF15 + u * (SF16 + SF17 * u))));
            }
        }
      v[0] += dx * fac;
      v[1] += dy * fac;
      v[2] += dz * fac;
      v[3] += wp;
    }
  for(int k = 0; k < 4; k++)
    s_v[k][threadIdx.x] = v[k];
  __syncthreads();
  for(int w = B / 2; w > 0; w >>= 1)
    {
      if(threadIdx.x < w)
        for(int k = 0; k < 4; k++)
          s_v[k][threadIdx.x] += s_v[k][threadIdx.x + w];
      __syncthreads();
    }
  if(threadIdx.x < 4)
    out[(size_t)q * 4 + threadIdx.x] = s_v[threadIdx.x][0];
}

static agrav_status direct_batch(agrav_ctx *c, int ns, const int *sel_h, double *acc, double *pot)
{
  hipStream_t s = c->stream;
  agrav_status st;
  DBuf<int> dsel;
  DBuf<double> dout;
  if((sel_h && (st = dsel.ensure(ns))) || (st = dout.ensure((size_t)ns * 4)))
    {
      dsel.release();
      dout.release();
      return st;
    }
  if(sel_h)
    AG_HIP(hipMemcpyAsync(dsel.p, sel_h, ns * sizeof(int), hipMemcpyHostToDevice, s));
  k_direct_batch<<<ns, B, 0, s>>>(ns, sel_h ? dsel.p : nullptr, c->d_pos_in.p, c->d_mass_in.p, c->d_pset_in.p, c->d_soff.p,
                                  c->d_shd.p, dout.p);
  AG_HIP(hipGetLastError());
  std::vector<double> h((size_t)ns * 4);
  AG_HIP(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  dsel.release();
  dout.release();
  for(int q = 0; q < ns; q++)
    {
      if(acc)
        for(int d = 0; d < 3; d++)
          acc[3 * (size_t)q + d] = h[(size_t)q * 4 + d];
      if(pot)
        pot[q] = h[(size_t)q * 4 + 3];
    }
  return AGRAV_OK;
}

agrav_status forcetest(agrav_ctx *c, int ns, const int *sel_h, double *acc, double *acc_short, double *pot, double *pot_short)
{
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "forcetest before set_particles");
  if(ns <= 0)
    return AGRAV_OK;
  for(int q = 0; q < ns; q++)
    if(sel_h[q] < 0 || sel_h[q] >= c->n)
      return fail(c, AGRAV_ERR_BAD_PARAM, "forcetest index out of range");
  if(c->batch_nsets > 0)
    {
      // isolated sets: no short-range split
      if(acc_short)
        std::fill(acc_short, acc_short + 3 * (size_t)ns, 0.0);
      if(pot_short)
        std::fill(pot_short, pot_short + ns, 0.0);
      return direct_batch(c, ns, sel_h, acc, pot);
    }
  agrav_status st;
  if(c->par.periodic && !c->ewald_ft_ready)
    if((st = ewald_forcetest_table(c)) != AGRAV_OK)
      return st;
  if(!c->par.periodic && (st = c->d_ewald_ft.ensure(1)))
    return st;
  hipStream_t s   = c->stream;
  const int64_t n = c->n;
  int tblocks     = (ns + B - 1) / B;
  int nchunk      = pick_chunks(tblocks, n);
  int chunk       = (int)((n + nchunk - 1) / nchunk);
  DBuf<int> dsel;
  DBuf<double> dpart, dout;
  if((st = dsel.ensure(ns)) || (st = dpart.ensure((size_t)nchunk * ns * 8)) || (st = dout.ensure((size_t)ns * 8)))
    {
      dsel.release();
      dpart.release();
      dout.release();
      return st;
    }
  AG_HIP(hipMemcpyAsync(dsel.p, sel_h, ns * sizeof(int), hipMemcpyHostToDevice, s));
  DirectArgs a{};
  a.pos      = c->d_pos_in.p;
  a.mass     = c->d_mass_in.p;
  a.stype    = c->have_stype ? c->d_stype_in.p : nullptr;
  a.hsoft    = c->d_soft.p;
  a.n        = (int)n;
  a.ns       = ns;
  a.sel      = dsel.p;
  a.periodic = c->par.periodic;
  a.box      = c->par.box_size;
  a.half     = 0.5 * c->par.box_size;
  a.pm       = c->par.pmgrid > 0;
  a.asmth    = a.pm ? c->par.asmth * c->par.box_size / c->par.pmgrid : 0;
  a.vol      = a.box * a.box * a.box;
  if(hr_on(c))
    {
      if((st = hr_flags(c)) != AGRAV_OK)
        return st;
      a.hr       = c->d_hr.p;
      a.hr_asmth = c->hr.asmth;
    }
  a.ewtab    = c->d_ewald_ft.p;
  a.chunk    = chunk;
  a.part     = dpart.p;
  dim3 grid(tblocks, nchunk);
  k_forcetest<<<grid, B, 0, s>>>(a);
  k_chunk_reduce<<<tblocks, B, 0, s>>>(dpart.p, ns, nchunk, dout.p);
  AG_HIP(hipGetLastError());
  std::vector<double> h((size_t)ns * 8);
  AG_HIP(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  dsel.release();
  dpart.release();
  dout.release();
  for(int q = 0; q < ns; q++)
    {
      const double *v = &h[(size_t)q * 8];
      for(int d = 0; d < 3; d++)
        {
          if(acc)
            acc[3 * q + d] = v[d];
          if(acc_short)
            acc_short[3 * q + d] = v[4 + d];
        }
      if(pot)
        pot[q] = v[3];
      if(pot_short)
        pot_short[q] = v[7];
    }
  return AGRAV_OK;
}

// k_direct over n particles given by device pointers, partial sums into part
// ([nchunk][n][4]); the context supplies the softening and short-range tables
// and, for a periodic tree-only context, the Ewald table (ready beforehand)
static agrav_status direct_launch(agrav_ctx *c, int64_t n, const double *pos, const double *mass, const uint8_t *stype,
                                  double *part, int nchunk, int chunk, hipStream_t s)
{
  const bool ew = c->par.periodic && c->par.pmgrid == 0;
  WalkConst wc  = walk_const(c, 0);
  DirectAllArgs a{};
  a.pos       = pos;
  a.mass      = mass;
  a.hsoft     = c->d_soft.p;
  a.stype     = stype;
  a.nsoft     = c->soft.nsoft;
  a.n         = (int)n;
  a.periodic  = c->par.periodic;
  a.pm        = c->par.pmgrid > 0;
  a.ewald     = ew;
  a.box       = c->par.box_size;
  a.half      = 0.5 * c->par.box_size;
  a.asmthfac  = wc.asmthfac;
  a.asmthinv2 = wc.asmthinv2;
  a.ewald_fac = wc.ewald_fac;
  a.tab       = c->d_tab.p;
  a.ewtab     = ew ? c->d_ewald.p : c->d_soft.p;  // read only when ewald
  if(hr_on(c) && pos == c->d_pos_in.p)
    {
      // the context's own particles: their high-resolution flags
      agrav_status st;
      if((st = hr_flags(c)) != AGRAV_OK)
        return st;
      const double ai = 0.5 / c->hr.asmth;
      a.hr            = c->d_hr.p;
      a.hr_asmthinv2  = ai * ai;
      a.hr_asmthfac   = ai * (AG_NTAB / (c->par.rcut / 2.0));
    }
  a.chunk     = chunk;
  a.part      = part;
  dim3 grid((unsigned)((n + B - 1) / B), nchunk);
  k_direct<<<grid, B, 0, s>>>(a);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
}

agrav_status direct_all(agrav_ctx *c, double *acc, double *pot)
{
  if(!c->have_particles)
    return fail(c, AGRAV_ERR_STATE, "gravity_direct before set_particles");
  const int64_t n = c->n;
  if(n == 0)
    return AGRAV_OK;
  if(c->batch_nsets > 0)
    return direct_batch(c, (int)n, nullptr, acc, pot);
  agrav_status st;
  const bool ew = c->par.periodic && c->par.pmgrid == 0;
  if(ew && !c->ewald_ready)
    if((st = ewald_tree_table(c)) != AGRAV_OK)
      return st;
  hipStream_t s = c->stream;
  int tblocks   = (int)((n + B - 1) / B);
  int nchunk    = pick_chunks(tblocks, n);
  int chunk     = (int)((n + nchunk - 1) / nchunk);
  DBuf<double> dpart, dout;
  if((st = dpart.ensure((size_t)nchunk * n * 4)) || (st = dout.ensure((size_t)n * 4)))
    {
      dpart.release();
      dout.release();
      return st;
    }
  if((st = direct_launch(c, n, c->d_pos_in.p, c->d_mass_in.p, c->have_stype ? c->d_stype_in.p : nullptr, dpart.p, nchunk,
                         chunk, s)) != AGRAV_OK)
    {
      dpart.release();
      dout.release();
      return st;
    }
  k_chunk_reduce4<<<tblocks, B, 0, s>>>(dpart.p, (int)n, nchunk, dout.p);
  AG_HIP(hipGetLastError());
  std::vector<double> h((size_t)n * 4);
  AG_HIP(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  dpart.release();
  dout.release();
  for(int64_t q = 0; q < n; q++)
    {
      if(acc)
        for(int d = 0; d < 3; d++)
          acc[3 * q + d] = h[q * 4 + d];
      if(pot)
        pot[q] = h[q * 4 + 3];
    }
  return AGRAV_OK;
}

// ---- gravity_direct without allocation -------------------------------------
// For callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops): the partial sums of k_direct go to `scratch`, a
// file-scope device array of fixed size, instead of a buffer allocated per call.
// Every call on a GPU shares that GPU's copy, so calls that overlap on one GPU
// race.  Nothing in it needs clearing: k_direct writes every partial sum that
// the reduction reads.
//
// Largest partial-sum array: pick_chunks gives at most ceil(n / 2048) chunks, so
// n <= AGRAV_DIRECT_NOALLOC_MAX_N needs at most 8 x 16384 x 4 doubles (4 MiB).
static constexpr size_t kNoallocDoubles = (size_t)((AGRAV_DIRECT_NOALLOC_MAX_N + 2047) / 2048) * AGRAV_DIRECT_NOALLOC_MAX_N * 4;
static __device__ __align__(256) double scratch[kNoallocDoubles];

// the device address of each GPU's copy of `scratch` (host code never uses the
// array's name itself: on the host it is only a stand-in for the symbol)
static constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

agrav_status noalloc_scratch_init(agrav_ctx *c)
{
  if(c->device < 0 || c->device >= kMaxDevices)
    return fail(c, AGRAV_ERR_NO_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if(g_scratch_ptr[c->device])
    return AGRAV_OK;
  void *ptr = nullptr;
  AG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if(!ptr)
    return fail(c, AGRAV_ERR_HIP, "no device address for the scratch array");
  g_scratch_ptr[c->device] = (double *)ptr;
  return AGRAV_OK;
}

// sum the partials of each particle in chunk order (as k_chunk_reduce4) and
// write GravAccel and Potential straight to the caller's arrays
__global__ void k_chunk_reduce_out(const double *__restrict__ part, int n, int nchunk, double *acc, double *pot)
{
  int q = blockIdx.x * B + threadIdx.x;
  if(q >= n)
    return;
  double v[4] = {0, 0, 0, 0};
  for(int c = 0; c < nchunk; c++)
    for(int k = 0; k < 4; k++)
      v[k] += part[((size_t)c * n + q) * 4 + k];
  acc[3 * (size_t)q]     = v[0];
  acc[3 * (size_t)q + 1] = v[1];
  acc[3 * (size_t)q + 2] = v[2];
  if(pot)
    pot[q] = v[3];
}

agrav_status direct_noalloc(agrav_ctx *c, int64_t n, const double *pos, const double *mass, const uint8_t *stype, double *acc,
                            double *pot, hipStream_t s)
{
  if(n < 0 || (n > 0 && (!pos || !mass || !acc)))
    return fail(c, AGRAV_ERR_BAD_PARAM, "gravity_direct_noalloc: null device pointer");
  if(n > AGRAV_DIRECT_NOALLOC_MAX_N)
    return fail(c, AGRAV_ERR_CAPACITY, "gravity_direct_noalloc: more than AGRAV_DIRECT_NOALLOC_MAX_N particles");
  if(c->batch_nsets > 0)
    return fail(c, AGRAV_ERR_BAD_PARAM, "gravity_direct_noalloc: not for a batch context");
  if(hr_on(c))
    return fail(c, AGRAV_ERR_BAD_PARAM, "gravity_direct_noalloc: not for a context with a high-resolution region");
  if(c->par.periodic && c->par.pmgrid == 0 && !c->ewald_ready)
    return fail(c, AGRAV_ERR_STATE, "gravity_direct_noalloc: the Ewald table is not ready");
  double *part = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if(!part)
    return fail(c, AGRAV_ERR_STATE, "gravity_direct_noalloc: scratch not set up on this device");
  if(n == 0)
    return AGRAV_OK;
  const int tblocks = (int)((n + B - 1) / B);
  const int nchunk  = pick_chunks(tblocks, n);
  const int chunk   = (int)((n + nchunk - 1) / nchunk);
  if((size_t)nchunk * (size_t)n * 4 > kNoallocDoubles)
    return fail(c, AGRAV_ERR_CAPACITY, "gravity_direct_noalloc: partial sums exceed the scratch array");
  agrav_status st;
  if((st = direct_launch(c, n, pos, mass, stype, part, nchunk, chunk, s)) != AGRAV_OK)
    return st;
  k_chunk_reduce_out<<<tblocks, B, 0, s>>>(part, (int)n, nchunk, acc, pot);
  AG_HIP(hipGetLastError());
  return AGRAV_OK;
}

}  // namespace agrav
