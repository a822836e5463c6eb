// agrav_walk.hip -- the tree walk: gravity_tree (gravtree.c:357-707) with
// force_treeevaluate (forcetree_walk.c:115-514).
//
// Two kernels:
//  * k_walk_plain: one lane per target, AREPO's stackless nextnode/sibling
//    walk with the per-particle opening test, FP64 throughout.  This is the
//    plain GPU restatement of the reference loop (same interaction set).
//  * k_walk_group: one 64-lane wavefront per group of <= 64 consecutive
//    targets in Peano-Hilbert order.  The wavefront evaluates up to 64 nodes
//    at a time against the group's bounding box with conservative versions of
//    AREPO's tests (a node is accepted only if every target would accept it),
//    keeps the pending nodes in an LDS stack, and stages accepted nodes and
//    particles as float4 tiles in LDS that every lane (= target) consumes by
//    broadcast reads.  FP32 variant: offsets are formed in fixed point
//    relative to the group centre and converted to float, forces are summed
//    per tile and the tile sums are added with compensated (Kahan) sums.
#include "agrav_ctx.h"

#include <string.h>

namespace agrav
{

// spline-softening coefficients (allvars.h:557-573)
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

__device__ inline double nearest_d(double x, const WalkConst &wc)
{
  if(!wc.periodic)
    return x;
  return x > wc.half_box ? x - wc.box : (x < -wc.half_box ? x + wc.box : x);
}

// pair kernel of forcetree_walk.c:364-487 in type T; returns false beyond the table
template <typename T, bool PM>
__device__ inline bool pair_kernel(T r, T hmax, const float *__restrict__ s_tab, T asmthfac, T asmthinv2, T &fac, T &wp)
{
  T ff = 0, fp = 0;
  if(PM)
    {
      T tabentry = asmthfac * r;
      int ti     = (int)tabentry;
      if(ti >= AG_NTAB)
        return false;
      T w = tabentry - (T)ti;
      ff  = ((T)1 - w) * (T)s_tab[ti] + w * (T)s_tab[ti + 1];
      fp  = ((T)1 - w) * (T)s_tab[128 + ti] + w * (T)s_tab[128 + ti + 1];
    }
  if(r >= hmax)
    {
      T rinv  = (T)1 / r;
      T rinv3 = rinv * rinv * rinv;
      if(PM)
        {
          fac = rinv3 + rinv * ff * asmthinv2;
          wp  = rinv * fp;
        }
      else
        {
          fac = rinv3;
          wp  = -rinv;
        }
    }
  else
    {
      T h_inv  = (T)1 / hmax;
      T h3_inv = h_inv * h_inv * h_inv;
      T u      = r * h_inv;
      if(u < (T)0.5)
        {
          T u2 = u * u;
          fac  = h3_inv * ((T)SF1 + u2 * ((T)SF2 * u + (T)SF3));
          wp   = h_inv * ((T)SF4 + u2 * ((T)SF5 + u2 * ((T)SF6 * u + (T)SF7)));
        }
      else
        {
          T u2 = u * u, u3 = u2 * u;
          fac  = h3_inv * ((T)SF8 + (T)SF9 * u + (T)SF10 * u2 + (T)SF11 * u3 + (T)SF12 / u3);
          wp   = h_inv * ((T)SF13 + (T)SF14 / u + u2 * ((T)SF1 + u * ((T)SF15 + u * ((T)SF16 + (T)SF17 * u))));
        }
      if(PM && r > (T)0)
        {
          T rinv = (T)1 / r;
          fac += rinv * ff * asmthinv2;
          wp += rinv * (fp + (T)1);
        }
    }
  return true;
}

// Ewald correction (forcetree_ewald.c:268-382), table [4][65][65][65]
template <typename T, typename TT>
__device__ inline void ewald_add(const TT *__restrict__ tab, T fac_intp, T dx, T dy, T dz, T m, T &ax, T &ay, T &az, T &pot)
{
  const int EN = 64, n1 = 65, size = n1 * n1 * n1;
  T sx = dx < 0 ? (T)1 : (T)-1, sy = dy < 0 ? (T)1 : (T)-1, sz = dz < 0 ? (T)1 : (T)-1;
  T u = fabs(dx) * fac_intp, v = fabs(dy) * fac_intp, w = fabs(dz) * fac_intp;
  int i = (int)u, j = (int)v, k = (int)w;
  i = i >= EN ? EN - 1 : i;
  j = j >= EN ? EN - 1 : j;
  k = k >= EN ? EN - 1 : k;
  u -= i;
  v -= j;
  w -= k;
  T f[8] = {(1 - u) * (1 - v) * (1 - w), (1 - u) * (1 - v) * w, (1 - u) * v * (1 - w), (1 - u) * v * w,
            u * (1 - v) * (1 - w),       u * (1 - v) * w,       u * v * (1 - w),       u * v * w};
  int o[8] = {(i * n1 + j) * n1 + k,           (i * n1 + j) * n1 + k + 1,           (i * n1 + j + 1) * n1 + k,
              (i * n1 + j + 1) * n1 + k + 1,   ((i + 1) * n1 + j) * n1 + k,         ((i + 1) * n1 + j) * n1 + k + 1,
              ((i + 1) * n1 + j + 1) * n1 + k, ((i + 1) * n1 + j + 1) * n1 + k + 1};
  T c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
  for(int q = 0; q < 8; q++)
    {
      c0 += (T)tab[o[q]] * f[q];
      c1 += (T)tab[size + o[q]] * f[q];
      c2 += (T)tab[2 * size + o[q]] * f[q];
      c3 += (T)tab[3 * size + o[q]] * f[q];
    }
  ax += m * sx * c0;
  ay += m * sy * c1;
  az += m * sz * c2;
  pot += m * c3;
}

// ---------------------------------------------------------------------------
// plain walk: one lane per target (forcetree_walk.c:115-514)
// ---------------------------------------------------------------------------

template <bool PM, bool EWALD>
__global__ __launch_bounds__(256) void k_walk_plain(TreeView tv, WalkConst wc, const double *__restrict__ hsoft,
                                                    const float *__restrict__ oldacc_s, const float *__restrict__ tabs,
                                                    const double *__restrict__ ewald, int ntarget, const int *__restrict__ tlist,
                                                    int tl_group, const int *__restrict__ tmap, int ntarg, BatchView bv, double *at,
                                                    double *pt, int *nint, unsigned long long *isum)
{
  __shared__ float s_tab[256];
  if(tree_invalid(tv))
    return;
  s_tab[threadIdx.x] = tabs[threadIdx.x];
  __syncthreads();
  int t = blockIdx.x * 256 + threadIdx.x;
  int i = -1;
  if(t < ntarget)
    {
      if(tlist && bv.pset)
        {
          // targets of failed groups of a batch: up to tl_group targets per group
          const int4 d = bv.groups[tlist[t / tl_group]];
          i            = (t % tl_group) < d.y ? d.x + (t % tl_group) : -1;
        }
      else if(tlist)
        {
          // targets of failed groups: tl_group consecutive targets per group id
          const int j = tlist[t / tl_group] * tl_group + (t % tl_group);
          i           = j < ntarg ? (tmap ? tmap[j] : j) : -1;
        }
      else
        i = tmap ? tmap[t] : t;  // every particle, or the active-target list
    }
  int nin = 0;
  if(i >= 0)
    {
      const int N       = tv.n;
      const double posx = tv.px[i], posy = tv.py[i], posz = tv.pz[i];
      // a set of a batch is walked from its own top entry to the entry after it, with its softening
      const int set     = bv.pset ? bv.pset[i] : 0;
      const double *hs  = bv.pset ? bv.sh + set : hsoft;
      const double h_i  = hs[tv.ps ? tv.ps[i] : 0];
      const double aold = wc.errtol * (double)oldacc_s[i];
      double ax = 0, ay = 0, az = 0, pot = 0;
      int no         = bv.pset ? bv.sroot[set] : N;
      const int stop = bv.pset ? bv.send[set] : -1;
      int steps = 0;
      const int max_steps = N + tv.m + 1;  // every entry is visited at most once per target
      while(no >= 0 && no != stop)
        {
          if(++steps > max_steps)
            {
              nin = -1;  // broken threading: give up on this target
              break;
            }
          double dx, dy, dz, r2, mass, hmax;
          if(no < N)
            {
              dx         = nearest_d(tv.px[no] - posx, wc);
              dy         = nearest_d(tv.py[no] - posy, wc);
              dz         = nearest_d(tv.pz[no] - posz, wc);
              r2         = dx * dx + dy * dy + dz * dz;
              mass       = tv.pm[no];
              double h_j = hs[tv.ps ? tv.ps[no] : 0];
              hmax       = h_j > h_i ? h_j : h_i;
              no         = tv.pnext[no];
            }
          else
            {
              const int k = no - N;
              mass        = tv.nm[k];
              dx          = nearest_d(tv.sx[k] - posx, wc);
              dy          = nearest_d(tv.sy[k] - posy, wc);
              dz          = nearest_d(tv.sz[k] - posz, wc);
              r2          = dx * dx + dy * dy + dz * dz;
              double ln   = tv.len[k];
              if(PM && r2 > wc.rcut2)
                {
                  double eff = wc.rcut + 0.5 * ln;
                  double d0  = nearest_d(tv.cx[k] - posx, wc);
                  double d1  = nearest_d(tv.cy[k] - posy, wc);
                  double d2  = nearest_d(tv.cz[k] - posz, wc);
                  if(d0 < -eff || d0 > eff || d1 < -eff || d1 > eff || d2 < -eff || d2 > eff)
                    {
                      no = tv.sibling[k];
                      continue;
                    }
                }
              if(wc.use_bh)
                {
                  if(ln * ln > r2 * wc.theta2)
                    {
                      no = tv.nextnode[k];
                      continue;
                    }
                }
              else
                {
                  double len2 = ln * ln;
                  if(len2 > r2 * (1.2 * 1.2))
                    {
                      no = tv.nextnode[k];
                      continue;
                    }
                  if(mass * len2 > r2 * r2 * aold)
                    {
                      no = tv.nextnode[k];
                      continue;
                    }
                  if(fabs(nearest_d(tv.cx[k] - posx, wc)) < 0.60 * ln && fabs(nearest_d(tv.cy[k] - posy, wc)) < 0.60 * ln &&
                     fabs(nearest_d(tv.cz[k] - posz, wc)) < 0.60 * ln)
                    {
                      no = tv.nextnode[k];
                      continue;
                    }
                }
              double h_j = bv.pset ? h_i : hsoft[tv.maxsoft[k]];
              if(h_j > h_i)
                {
                  if(r2 < h_j * h_j)
                    {
                      no = tv.nextnode[k];
                      continue;
                    }
                  hmax = h_j;
                }
              else
                hmax = h_i;
              no = tv.sibling[k];
            }
          if(mass != 0)
            {
              double r = sqrt(r2), fac, wp;
              if(!pair_kernel<double, PM>(r, hmax, s_tab, wc.asmthfac, wc.asmthinv2, fac, wp))
                continue;
              pot += mass * wp;
              fac *= mass;
              ax += dx * fac;
              ay += dy * fac;
              az += dz * fac;
              if(EWALD)
                ewald_add<double, double>(ewald, wc.ewald_fac, dx, dy, dz, mass, ax, ay, az, pot);
              nin++;
            }
        }
      at[3 * (int64_t)i + 0] = ax;
      at[3 * (int64_t)i + 1] = ay;
      at[3 * (int64_t)i + 2] = az;
      pt[i]                  = pot;
      nint[i]                = nin;
    }
  // interaction total
  unsigned long long v = (unsigned long long)(nin > 0 ? nin : 0);
  for(int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  if((threadIdx.x & 63) == 0 && v)
    atomicAdd(isum, v);
}

// ---------------------------------------------------------------------------
// group walk
// ---------------------------------------------------------------------------

static constexpr int kStackCap = 1024;

// build-time switches of the group walk (kept for A/B timing)
#ifndef AG_NODE_RCUT_PRUNE
#define AG_NODE_RCUT_PRUNE 1
#endif
#ifndef AG_XCD_REMAP
#define AG_XCD_REMAP 1
#endif
#ifndef AG_WALK_POP
#define AG_WALK_POP 1
#endif
#ifndef AG_SPLIT_MIN
#define AG_SPLIT_MIN 64
#endif
static constexpr int kListCap  = 128;
// FP32 TreePM walks of one softening type split the tile: a main list and, after it, a
// list of the entries within the softening length of the group's box (only its loop has
// the spline branch).  One staging round adds at most 64 entries to either.
#ifndef AG_FAST_MAIN
#define AG_FAST_MAIN 64
#endif
static constexpr int kSoftCap  = 64;
static constexpr int kListAlloc = AG_FAST_MAIN + kSoftCap > kListCap ? AG_FAST_MAIN + kSoftCap : kListCap;

// Inclusive prefix sum over the 64 lanes of a wavefront (all lanes active) with DPP
// moves: shifts by 1, 2, 4, 8 within each row of 16 lanes, then the last lane of row
// 0 (2) added to row 1 (3) and the last lane of row 1 to rows 2 and 3.  A lane whose
// source lies outside its row, or a row outside the row mask, adds the `old` operand 0.
__device__ inline int wave_incl_scan_i(int v)
{
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return v;
}
template <typename T> __device__ inline T wave_min(T v)
{
#pragma unroll
  for(int off = 32; off > 0; off >>= 1)
    {
      T o = __shfl_xor(v, off, 64);
      v   = o < v ? o : v;
    }
  return v;
}
template <typename T> __device__ inline T wave_max(T v)
{
#pragma unroll
  for(int off = 32; off > 0; off >>= 1)
    {
      T o = __shfl_xor(v, off, 64);
      v   = o > v ? o : v;
    }
  return v;
}

// Kahan-compensated accumulation of a tile sum into a running total
template <typename T> __device__ inline void kahan_add(T &sum, T &comp, T x)
{
  T y = x - comp;
  T t = sum + y;
  comp = (t - sum) - y;
  sum  = t;
}

template <typename T> struct Vec4
{
  T x, y, z, w;
};

// Nearest periodic image of a pair offset (GRAVITY_NEAREST, allvars.h:603-612).
// The group walk holds offsets from the group centre, each within half a box,
// so a target-source offset can reach a full box; tree-only periodic runs add
// the Ewald correction of exactly this image, whose table covers half a box.
template <typename T> __device__ inline T nearest_image(T d, T box, T half)
{
  return d > half ? d - box : (d < -half ? d + box : d);
}

// FP64 tile loop (FP64 group walks): 1/r from the hardware reciprocal square root
// (about 2^-23 relative) refined by two Newton steps, which reach FP64 rounding,
// instead of a square root and a division per pair; the pair kernel is otherwise
// pair_kernel's (forcetree_walk.c:364-487).  The tiny addend keeps 1/r finite for
// the target itself (r2 = 0), which then takes the softened branch with r = 0.
template <bool PM, bool EWALD, bool MSOFT>
__device__ inline void tile_compute_f64(const Vec4<double> *__restrict__ s_list, const double *__restrict__ s_h, int nl_in,
                                        double tx, double ty, double tz, double hi, const float *__restrict__ s_tab,
                                        const WalkConst &wc, double asmthfac, double asmthinv2, const double *__restrict__ ewald_d,
                                        double &ax, double &ay, double &az, double &pp, double &cax, double &cay, double &caz,
                                        double &cpp, int &nin)
{
  double fx = 0, fy = 0, fz = 0, fp = 0;
  const int nl       = __builtin_amdgcn_readfirstlane(nl_in);
  const double box = wc.box, half = wc.half_box;
  for(int k = 0; k < nl; k++)
    {
      const Vec4<double> e = s_list[k];
      double dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
      if(EWALD)
        {
          dx = nearest_image(dx, box, half);
          dy = nearest_image(dy, box, half);
          dz = nearest_image(dz, box, half);
        }
      const double r2  = dx * dx + dy * dy + dz * dz;
      const double r2s = r2 + 1e-200;
      double y         = __builtin_amdgcn_rsq(r2s);
#pragma unroll
      for(int it = 0; it < 2; it++)
        y = fma(0.5 * y, fma(-r2s * y, y, 1.0), y);
      const double r = r2 * y;
      double hmax    = hi;
      if(MSOFT)
        hmax = s_h[k] > hi ? s_h[k] : hi;
      if(e.w == 0.0)
        continue;
      double ff = 0, fpt = 0;
      if(PM)
        {
          const double te = asmthfac * r;
          const int ti    = (int)te;
          if(ti >= AG_NTAB)
            continue;
          const double w = te - (double)ti;
          ff             = (1.0 - w) * (double)s_tab[ti] + w * (double)s_tab[ti + 1];
          fpt            = (1.0 - w) * (double)s_tab[128 + ti] + w * (double)s_tab[128 + ti + 1];
        }
      double fac, wp;
      if(r >= hmax)
        {
          const double rinv3 = y * y * y;
          fac                = PM ? rinv3 + y * ff * asmthinv2 : rinv3;
          wp                 = PM ? y * fpt : -y;
        }
      else
        {
          // spline softening (allvars.h:557-573), rare: exact divisions
          const double h_inv = 1.0 / hmax, h3_inv = h_inv * h_inv * h_inv;
          const double u = r * h_inv, u2 = u * u;
          if(u < 0.5)
            {
              fac = h3_inv * (SF1 + u2 * (SF2 * u + SF3));
              wp  = h_inv * (SF4 + u2 * (SF5 + u2 * (SF6 * u + SF7)));
            }
          else
            {
              const double u3 = u2 * u;
              fac = h3_inv * (SF8 + SF9 * u + SF10 * u2 + SF11 * u3 + SF12 / u3);
              wp  = h_inv * (SF13 + SF14 / u + u2 * (SF1 + u * (SF15 + u * (SF16 + SF17 * u))));
            }
          if(PM && r > 0.0)
            {
              fac += y * ff * asmthinv2;
              wp += y * (fpt + 1.0);
            }
        }
      fp += e.w * wp;
      fac *= e.w;
      fx += dx * fac;
      fy += dy * fac;
      fz += dz * fac;
      if(EWALD)
        ewald_add<double, double>(ewald_d, wc.ewald_fac, dx, dy, dz, e.w, fx, fy, fz, fp);
      nin++;
    }
  // tile sums join the running totals through compensated additions
  kahan_add(ax, cax, fx);
  kahan_add(ay, cay, fy);
  kahan_add(az, caz, fz);
  kahan_add(pp, cpp, fp);
}

// FP32 tile loop: one reciprocal square root per interaction (v_rsq_f32, about
// 1 ulp), the short-range factors from the LDS tables without branches, and the
// softened kernel only for the rare entries inside the softening length.  The
// arithmetic is the reference's pair kernel (forcetree_walk.c:364-487).
// Table layout (FP32 walks): float2 {f_i, f_(i+1) - f_i} for i = 0..127 with
// entry 127 = {f_127, 0}, so an index clamped to 127 reads in bounds, then the
// potential table as float2 {p_i, p_(i+1) - p_i} in the same way (one read at a
// fixed offset from the force entry).  Zero-mass entries never enter the tile, so
// validity is only the table test.
static constexpr float kTinyR2 = 1e-30f;

// SPLINE = false: the caller has shown that no entry lies within the softening length
// of any target (the entries near the group go to a list of their own).
template <bool PM, bool EWALD, bool MSOFT, bool POT, bool SPLINE = true>
__device__ inline void tile_compute_f32(const Vec4<float> *__restrict__ s_list, const float *__restrict__ s_h, int nl_in, float tx,
                                        float ty, float tz, float hi, const float *__restrict__ s_tab, const WalkConst &wc,
                                        float asmthfac, float asmthinv2, const float *__restrict__ ewald_f, float &ax, float &ay,
                                        float &az, float &pp, float &cax, float &cay, float &caz, float &cpp, int &nin)
{
  float fx = 0, fy = 0, fz = 0, fp = 0;
  const int nl         = __builtin_amdgcn_readfirstlane(nl_in);  // wave-uniform: scalar loop control
  const float2 *ftab   = reinterpret_cast<const float2 *>(s_tab);
  const float2 *ptab   = reinterpret_cast<const float2 *>(s_tab) + 128;
  const float hi_inv   = __builtin_amdgcn_rcpf(hi);
  const float box = (float)wc.box, half = (float)wc.half_box;
  int cnt = 0;
  // one pair of the target with tile entry e at position k
  auto pair = [&](const Vec4<float> e, int k) {
      float dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
      if(EWALD)
        {
          dx = nearest_image(dx, box, half);
          dy = nearest_image(dy, box, half);
          dz = nearest_image(dz, box, half);
        }
      // the 1e-30 term keeps the reciprocal square root finite for a pair at distance
      // zero (the target itself); any other pair has r2 >= (2^-32 box)^2, far above it
      const float r2   = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kTinyR2)));
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float r    = r2 * rinv;
      float hmax = hi, hinv = hi_inv;
      if(MSOFT)
        {
          hmax = s_h[k] > hi ? s_h[k] : hi;
          hinv = __builtin_amdgcn_rcpf(hmax);
        }
      float ff = 0, fpt = 0;
      bool valid = true;
      if(PM)
        {
          const float te = asmthfac * r;
          const int ti   = (int)te;
          valid          = ti < AG_NTAB;
          const int tc   = min(ti, AG_NTAB);
          // te - ti exactly (te >= 0); for te past the table the pair is invalid anyway
          const float w   = __builtin_amdgcn_fractf(te);
          const float2 fd = ftab[tc];
          ff = fd.x + w * fd.y;  // force table is pre-scaled by asmthinv2
          if(POT)
            {
              const float2 pd = ptab[tc];
              fpt             = pd.x + w * pd.y;
            }
        }
      const float rinv2 = rinv * rinv;
      float fac, wp;
      if(PM)
        {
          fac = rinv * (rinv2 + ff);
          wp  = rinv * fpt;
        }
      else
        {
          fac = rinv2 * rinv;
          wp  = -rinv;
        }
      if(SPLINE && r < hmax)
        {
          // spline softening (allvars.h:557-573); u < 1 here
          const float h3 = hinv * hinv * hinv;
          const float u  = r * hinv;
          const float u2 = u * u;
          if(u < 0.5f)
            {
              fac = h3 * ((float)SF1 + u2 * ((float)SF2 * u + (float)SF3));
              wp  = hinv * ((float)SF4 + u2 * ((float)SF5 + u2 * ((float)SF6 * u + (float)SF7)));
            }
          else
            {
              const float ui = __builtin_amdgcn_rcpf(u);
              const float u3 = u2 * u;
              fac = h3 * ((float)SF8 + (float)SF9 * u + (float)SF10 * u2 + (float)SF11 * u3 + (float)SF12 * ui * ui * ui);
              wp  = hinv * ((float)SF13 + (float)SF14 * ui + u2 * ((float)SF1 + u * ((float)SF15 + u * ((float)SF16 + (float)SF17 * u))));
            }
          if(PM && r2 > kTinyR2)
            {
              fac += rinv * ff;
              wp += rinv * (fpt + 1.0f);
            }
        }
      const float m = valid ? e.w : 0.f;
      if(POT)
        fp += m * wp;
      const float mf = m * fac;
      fx += dx * mf;
      fy += dy * mf;
      fz += dz * mf;
      if(EWALD && valid)
        ewald_add<float, float>(ewald_f, (float)wc.ewald_fac, dx, dy, dz, e.w, fx, fy, fz, fp);
      cnt += valid ? 1 : 0;
  };
  // two entries per round, both read before either pair is computed (as in
  // tile_loop_f32_pm)
  int k = 0;
  for(; k + 1 < nl; k += 2)
    {
      const Vec4<float> e0 = s_list[k], e1 = s_list[k + 1];
      pair(e0, k);
      pair(e1, k + 1);
    }
  if(k < nl)
    pair(s_list[k], k);
  nin += cnt;
  kahan_add(ax, cax, fx);
  kahan_add(ay, cay, fy);
  kahan_add(az, caz, fz);
  kahan_add(pp, cpp, fp);
}

// FP32 TreePM tile loop of one softening type, in table units: coordinates are
// scaled by asmthfac, so r' = r2' rsq(r2') is the table position directly.  The pair
// factor is fac = rinv (rinv^2 + T), T the interpolated short-range correction
// (S(u) - 1) / u^2 du^2 (forcetree_walk.c:367-380 with u = r' du), so that fac r'^3 = S(u) =
// erfc(u) + 2u/sqrt(pi) exp(-u^2).  The LDS table holds {T_i, T_(i+1) - T_i} and
// {P_i, P_(i+1) - P_i} for the potential P = -erfc(u), zero at entry NTAB; a pair beyond
// the cut (tabindex >= NTAB, forcetree_walk.c:372) adds nothing and counts nothing.  In these units a tile sum is the physical one times
// asmthfac^-2 (accelerations) and asmthfac^-1 (potentials); the caller rescales.  The
// spline branch (allvars.h:557-573, forcetree_walk.c:429-456) is entered only by tiles
// that hold an entry within the softening length of the group (`soft`).  Without
// CLAMP the table index is not clamped to NTAB: the caller has shown that every pair of
// the group lies below 2 NTAB table units, and the LDS table is zero up to there.
template <bool POT, bool SOFT, bool CLAMP>
__device__ inline void tile_loop_f32_pm(const Vec4<float> *__restrict__ s_list, int nl, float tx, float ty, float tz, float hs,
                                        float hs_inv, const float *__restrict__ tab, float &fx, float &fy, float &fz,
                                        float &fp, float &cnt)
{
  const float h3 = hs_inv * hs_inv * hs_inv;
  // one pair of the target with tile entry e
  auto pair = [&](const Vec4<float> e) {
      const float dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
      const float r2   = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kTinyR2)));
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float te   = r2 * rinv;
      const int tc     = CLAMP ? min((int)te, AG_NTAB) : (int)te;
      const float w    = __builtin_amdgcn_fractf(te);
      // {T_i, dT_i} (b64 reads: fewer LDS bank conflicts than wider entries); rc = rinv
      // within the cut and 0 beyond it, where the Newtonian term of fac must vanish too
      const float2 q = reinterpret_cast<const float2 *>(tab)[tc];
      const float ff = fmaf(w, q.y, q.x);
      float fpt      = 0.f;
      if(POT)
        {
          const float2 pq = reinterpret_cast<const float2 *>(tab)[128 + tc];
          fpt             = fmaf(w, pq.y, pq.x);
        }
      const float rc = tc < AG_NTAB ? rinv : 0.f;
      float fac = rc * fmaf(rinv, rinv, ff);
      float wp  = POT ? rinv * fpt : 0.f;
      if(SOFT && te < hs && tc < AG_NTAB)
        {
          // spline softening; u < 1 here
          const float u  = te * hs_inv;
          const float u2 = u * u;
          float fs, ws;
          if(u < 0.5f)
            {
              fs = h3 * ((float)SF1 + u2 * ((float)SF2 * u + (float)SF3));
              ws = hs_inv * ((float)SF4 + u2 * ((float)SF5 + u2 * ((float)SF6 * u + (float)SF7)));
            }
          else
            {
              const float ui = __builtin_amdgcn_rcpf(u);
              const float u3 = u2 * u;
              fs = h3 * ((float)SF8 + (float)SF9 * u + (float)SF10 * u2 + (float)SF11 * u3 + (float)SF12 * ui * ui * ui);
              ws = hs_inv * ((float)SF13 + (float)SF14 * ui + u2 * ((float)SF1 + u * ((float)SF15 + u * ((float)SF16 + (float)SF17 * u))));
            }
          if(r2 > kTinyR2)
            {
              fs += rinv * ff;
              if(POT)
                ws += rinv * (fpt + 1.0f);
            }
          fac = fs;
          wp  = ws;
        }
      const float mf = e.w * fac;
      fx = fmaf(dx, mf, fx);
      fy = fmaf(dy, mf, fy);
      fz = fmaf(dz, mf, fz);
      if(POT)
        fp = fmaf(e.w, wp, fp);
      cnt = fmaf(rc, te, cnt);  // rinv r = 1 within the cut (to a few ulp), 0 beyond
  };
  // two entries per round, unrolled by hand: both entry reads go out before either
  // pair is computed, at fixed offsets from one address (the loop-pragma unroll with
  // its remainder handling measured 5 % slower); deeper unrolling costs occupancy
  int k = 0;
  for(; k + 1 < nl; k += 2)
    {
      const Vec4<float> e0 = s_list[k], e1 = s_list[k + 1];
      pair(e0);
      pair(e1);
    }
  if(k < nl)
    pair(s_list[k]);
}

template <bool POT>
__device__ inline void tile_compute_f32_pm(const Vec4<float> *__restrict__ s_list, int nl_in, float tx, float ty, float tz, float hs,
                                           bool soft, bool compact, const float *__restrict__ s_tab, float &ax, float &ay, float &az, float &pp,
                                           float &cax, float &cay, float &caz, float &cpp, int &nin)
{
  float fx = 0, fy = 0, fz = 0, fp = 0, cnt = 0;
  const int nl          = __builtin_amdgcn_readfirstlane(nl_in);  // wave-uniform: scalar loop control
  const float *tab      = s_tab;
  const float hs_inv    = __builtin_amdgcn_rcpf(hs);
  if(!POT && compact)
    {
      if(soft)
        tile_loop_f32_pm<false, true, false>(s_list, nl, tx, ty, tz, hs, hs_inv, tab, fx, fy, fz, fp, cnt);
      else
        tile_loop_f32_pm<false, false, false>(s_list, nl, tx, ty, tz, hs, hs_inv, tab, fx, fy, fz, fp, cnt);
    }
  else if(soft)
    tile_loop_f32_pm<POT, true, true>(s_list, nl, tx, ty, tz, hs, hs_inv, tab, fx, fy, fz, fp, cnt);
  else
    tile_loop_f32_pm<POT, false, true>(s_list, nl, tx, ty, tz, hs, hs_inv, tab, fx, fy, fz, fp, cnt);
  nin += (int)(cnt + 0.5f);  // a sum of tile-many terms 1 +- 2^-21: round to the count
  kahan_add(ax, cax, fx);
  kahan_add(ay, cay, fy);
  kahan_add(az, caz, fz);
  kahan_add(pp, cpp, fp);
}

template <typename T, bool PM, bool EWALD, bool MSOFT>
__device__ inline void tile_compute(const Vec4<T> *__restrict__ s_list, const T *__restrict__ s_h, int nl, T tx, T ty, T tz, T hi,
                                    const float *__restrict__ s_tab, const WalkConst &wc, T asmthfac, T asmthinv2,
                                    const float *__restrict__ ewald_f, const double *__restrict__ ewald_d, T &ax, T &ay, T &az, T &pp,
                                    T &cax, T &cay, T &caz, T &cpp, int &nin, bool spline = true)
{
  if constexpr(sizeof(T) == 4)
    {
      if(wc.eval_pot && spline)
        tile_compute_f32<PM, EWALD, MSOFT, true>(s_list, s_h, nl, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_f, ax, ay,
                                                 az, pp, cax, cay, caz, cpp, nin);
      else if(wc.eval_pot)
        tile_compute_f32<PM, EWALD, MSOFT, true, false>(s_list, s_h, nl, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_f, ax,
                                                        ay, az, pp, cax, cay, caz, cpp, nin);
      else if(spline)
        tile_compute_f32<PM, EWALD, MSOFT, false>(s_list, s_h, nl, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_f, ax,
                                                  ay, az, pp, cax, cay, caz, cpp, nin);
      else
        tile_compute_f32<PM, EWALD, MSOFT, false, false>(s_list, s_h, nl, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_f,
                                                         ax, ay, az, pp, cax, cay, caz, cpp, nin);
    }
  else
    tile_compute_f64<PM, EWALD, MSOFT>(s_list, s_h, nl, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_d, ax, ay, az, pp, cax,
                                       cay, caz, cpp, nin);
}

// LDS of one group walk (one wavefront per block)
template <typename T, bool MSOFT> struct WalkLds
{
  alignas(16) float tab[512];  // short-range tables: FP64 walks [force 128 | potential 128], FP32 walks see tile_compute_f32
  int stack[kStackCap];
  int map[64];  // list and index offset of each slot of a particle-append chunk
  Vec4<T> list[kListAlloc];
  T h[MSOFT ? kListCap : 1];
};

// load the short-range tables (tabs = [force 128 | potential 128 | the FP32 TreePM
// loop's {T_i, dT_i} x 128 | {P_i, dP_i} x 128])
template <typename T, bool PM, bool FAST>
__device__ inline void load_tables(float *s_tab, const float *__restrict__ tabs, const WalkConst &wc)
{
  if constexpr(FAST)
    {
      // {T_i, dT_i} x 128, then {P_i, dP_i} x 128 when potentials are evaluated, else zeros
      // (the force table read past NTAB by loops without an index clamp)
      for(int q = threadIdx.x & 63; q < 512; q += 64)
        s_tab[q] = (q < 256 || wc.eval_pot) ? tabs[256 + q] : 0.f;
    }
  else if constexpr(sizeof(T) == 4)
    {
      // the FP32 tile loop uses the force table pre-multiplied by (0.5/Asmth)^2
      const float tscale = PM ? (float)wc.asmthinv2 : 1.0f;
      float2 *ftab       = reinterpret_cast<float2 *>(s_tab);
      for(int q = threadIdx.x & 63; q < 128; q += 64)
        {
          const float f0 = tabs[q] * tscale;
          const float f1 = q < 127 ? tabs[q + 1] * tscale : f0;
          ftab[q]        = make_float2(f0, f1 - f0);
          const float p0 = tabs[128 + q];
          const float p1 = q < 127 ? tabs[128 + q + 1] : p0;
          ftab[128 + q]  = make_float2(p0, p1 - p0);
        }
    }
  else
    for(int q = threadIdx.x & 63; q < 256; q += 64)
      s_tab[q] = tabs[q];
}

// Walk of one group: the np (<= 64) Hilbert-consecutive targets from p0, one per
// lane, against the tree below the entry `root` (the tree root N, or the top entry
// of one set of a batch).  `hsoft` is indexed by softening type; `unit` is the
// fixed-point step of the targets' domain.  With nsplit > 1 the group's traversal
// is shared by nsplit wavefronts, this one being number wsplit: all of them expand
// the first levels alike (only wave 0 keeps their interactions), then the stack is
// dealt out round-robin and each wave walks its share; every entry is evaluated by
// exactly one wave, and the caller adds the waves' partial sums in wave order.
// Returns false
// when the LDS stack overflowed (the caller re-walks the group one lane per
// target).
// per-lane partial sums of one walk (Kahan pairs) and its interaction count
template <typename T> struct GroupAcc
{
  T ax, ay, az, pp, cax, cay, caz, cpp;
  int nin;
};

template <typename T, bool PM, bool EWALD, bool MSOFT>
__device__ inline bool walk_group(WalkLds<T, MSOFT> &lds, const TreeView &tv, const WalkConst &wc, const double *__restrict__ hsoft,
                                  const float *__restrict__ oldacc_s, const float *__restrict__ ewald_f,
                                  const double *__restrict__ ewald_d, const int *__restrict__ tmap, int p0, int np, int root,
                                  T unit, int leaf_bucket, int nsplit, int wsplit, GroupAcc<T> &out,
                                  unsigned long long (&cnt)[3])
{
  int *s_stack     = lds.stack;
  int *s_map       = lds.map;
  Vec4<T> *s_list  = lds.list;
  T *s_h           = lds.h;
  const float *s_tab = lds.tab;
  const int lane   = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  const int N  = tv.n;
  const bool has = lane < np;
  // the lane's target: Hilbert-consecutive particles, or entries of the active-target list
  const int pi   = tmap ? tmap[p0 + (has ? lane : 0)] : p0 + (has ? lane : 0);

  // ---- targets relative to the group centre
  T tx, ty, tz, hx, hy, hz;
  uint32_t cfx = 0, cfy = 0, cfz = 0;  // fixed-point group centre (FP32 path)
  double c0 = 0, c1 = 0, c2 = 0;       // FP64 group centre
  if constexpr(sizeof(T) == 4)
    {
      uint4 tr  = tv.prec[pi];
      uint32_t rx = __shfl(tr.x, 0, 64), ry = __shfl(tr.y, 0, 64), rz = __shfl(tr.z, 0, 64);
      int ox = (int)(tr.x - rx), oy = (int)(tr.y - ry), oz = (int)(tr.z - rz);
      int64_t mnx = wave_min((int64_t)ox), mxx = wave_max((int64_t)ox);
      int64_t mny = wave_min((int64_t)oy), mxy = wave_max((int64_t)oy);
      int64_t mnz = wave_min((int64_t)oz), mxz = wave_max((int64_t)oz);
      cfx = rx + (uint32_t)(int32_t)((mnx + mxx) >> 1);
      cfy = ry + (uint32_t)(int32_t)((mny + mxy) >> 1);
      cfz = rz + (uint32_t)(int32_t)((mnz + mxz) >> 1);
      tx  = (T)(int)(tr.x - cfx) * unit;
      ty  = (T)(int)(tr.y - cfy) * unit;
      tz  = (T)(int)(tr.z - cfz) * unit;
    }
  else
    {
      double rx = __shfl(tv.px[pi], 0, 64), ry = __shfl(tv.py[pi], 0, 64), rz = __shfl(tv.pz[pi], 0, 64);
      double ox = nearest_d(tv.px[pi] - rx, wc), oy = nearest_d(tv.py[pi] - ry, wc), oz = nearest_d(tv.pz[pi] - rz, wc);
      c0 = rx + 0.5 * (wave_min(ox) + wave_max(ox));
      c1 = ry + 0.5 * (wave_min(oy) + wave_max(oy));
      c2 = rz + 0.5 * (wave_min(oz) + wave_max(oz));
      tx = (T)nearest_d(tv.px[pi] - c0, wc);
      ty = (T)nearest_d(tv.py[pi] - c1, wc);
      tz = (T)nearest_d(tv.pz[pi] - c2, wc);
    }
  // half-extent of the target box, padded by one fixed-point step and a relative epsilon
  const T pad = sizeof(T) == 4 ? (T)1e-6 : (T)1e-13;
  hx = wave_max((T)fabs(tx)) * ((T)1 + pad) + unit;
  hy = wave_max((T)fabs(ty)) * ((T)1 + pad) + unit;
  hz = wave_max((T)fabs(tz)) * ((T)1 + pad) + unit;
  const T hi     = (T)hsoft[tv.ps ? tv.ps[pi] : 0];
  const T hi_min = wave_min(hi);
  T aold_min     = 0;
  if(!wc.use_bh)
    aold_min = (T)wc.errtol * wave_min((T)oldacc_s[pi]);

  const T rcut = (T)wc.rcut, rcut2 = (T)wc.rcut2, theta2 = (T)wc.theta2;
  const T asmthfac = (T)wc.asmthfac, asmthinv2 = (T)wc.asmthinv2;
  // FP32 TreePM of one softening type: tile entries and targets in table units (see
  // tile_compute_f32_pm), a tile marked `soft` when an entry lies within the
  // softening length of the group's box
  constexpr bool kFast = sizeof(T) == 4 && PM && !EWALD && !MSOFT;
  // FP32 walks of one softening type keep node entries on the stack as positions in
  // their parent's child list (e >= N: position e - N, records in nmeta/nrecA/nrecB),
  // so that pushing the children of an opened node reads nothing; the top entry is a
  // node index (e <= -2: node -e - 2).  Other walks push node indices (e >= N: node e - N).
  constexpr bool kPos  = sizeof(T) == 4 && !MSOFT;
  const T cs           = kFast ? asmthfac : (T)1;
  const T hs2          = hi_min * hi_min * (T)1.01;
  // every tile entry lies within Rcut (x 1.00005) of the group's box, so no pair is
  // farther than that plus the box diagonal: below 2 NTAB - 1 table units the loop needs
  // no index clamp (the table is zero from NTAB to 2 NTAB)
  const bool compact = kFast && (rcut * (T)1.0001 + (T)2 * sqrt(hx * hx + hy * hy + hz * hz)) * asmthfac < (T)(2 * AG_NTAB - 1);
  // TreePM pairs take their periodic image relative to the group centre, which is every
  // target's nearest image of every source within Rcut only while the group's
  // half-extent plus Rcut stays below half a box.  Groups of Hilbert-consecutive
  // particles are far smaller; a wider group (sparse active targets) is reported as
  // failed before its walk starts (every wavefront of the group decides alike) and is
  // walked per target by the caller.
  if(PM && wc.periodic)
    {
      const T hmax = hx > hy ? (hx > hz ? hx : hz) : (hy > hz ? hy : hz);
      if(hmax + rcut >= (T)(0.49 * wc.box))
        return false;
    }

  // ---- walk
  int sp = 1;
  __builtin_amdgcn_wave_barrier();
  if(lane == 0)
    s_stack[0] = (kPos && root >= N) ? -(root - N) - 2 : root;
  int nl = 0;
  T ax = 0, ay = 0, az = 0, pp = 0, cax = 0, cay = 0, caz = 0, cpp = 0;
  int nin        = 0;
  unsigned long long n_popped = 0, n_tiled = 0;  // diagnostics: stack entries evaluated, tile entries computed
  // the main list and (FP32 TreePM of one softening type) the list of entries within
  // the softening length of the group's box, whose loop alone has the spline branch
  constexpr int kMainCap = kFast ? AG_FAST_MAIN : kListCap;
  // isolated FP32 walks of one softening type also keep the entries near the group in a
  // list of their own (in the LDS of the short-range table, which they do not read), so
  // that the main list's loop has no softening test
  constexpr bool kNearList = sizeof(T) == 4 && !PM && !EWALD && !MSOFT;
  constexpr bool kSplit    = kFast || kNearList;
  static_assert(!kNearList || sizeof(lds.tab) >= kSoftCap * sizeof(Vec4<float>), "near list fits the table's LDS");
  Vec4<T> *s_soft        = kNearList ? reinterpret_cast<Vec4<T> *>(lds.tab) : s_list + kMainCap;
  int ns                 = 0;
  auto fast_tile         = [&](const Vec4<T> *lst, int cnt, bool soft) {
    if constexpr(kFast)
      {
        if(wc.eval_pot)
          tile_compute_f32_pm<true>(lst, cnt, tx * cs, ty * cs, tz * cs, hi * cs, soft, false, s_tab, ax, ay, az, pp, cax, cay, caz,
                                    cpp, nin);
        else
          tile_compute_f32_pm<false>(lst, cnt, tx * cs, ty * cs, tz * cs, hi * cs, soft, compact, s_tab, ax, ay, az, pp, cax, cay,
                                     caz, cpp, nin);
      }
  };
  auto flush_tile = [&]() {
    if constexpr(kFast)
      fast_tile(s_list, nl, false);
    else
      tile_compute<T, PM, EWALD, MSOFT>(s_list, s_h, nl, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_f, ewald_d, ax, ay,
                                        az, pp, cax, cay, caz, cpp, nin, !kNearList);
  };
  auto soft_tile = [&]() {
    if constexpr(kFast)
      fast_tile(s_soft, ns, true);
    else if constexpr(kNearList)
      tile_compute<T, PM, EWALD, MSOFT>(s_soft, s_h, ns, tx, ty, tz, hi, s_tab, wc, asmthfac, asmthinv2, ewald_f, ewald_d, ax, ay,
                                        az, pp, cax, cay, caz, cpp, nin, true);
  };
  // stage one entry per lane (acc: the lane has one; nr: within the softening length of
  // the group's box), flushing a list first when the entries would not fit
  auto stage = [&](bool acc, bool nr, const Vec4<T> &it, T hv) {
    const bool to_soft = kSplit && nr;
    const uint64_t bm  = __ballot(acc && !to_soft);
    const int cnt      = __popcll(bm);
    if(nl + cnt > kMainCap)
      {
        __builtin_amdgcn_wave_barrier();
        n_tiled += nl;
        flush_tile();
        nl = 0;
        __builtin_amdgcn_wave_barrier();
      }
    if constexpr(kSplit)
      {
        const uint64_t bs = __ballot(acc && to_soft);
        const int sc      = __popcll(bs);
        if(ns + sc > kSoftCap)
          {
            __builtin_amdgcn_wave_barrier();
            n_tiled += ns;
            soft_tile();
            ns = 0;
            __builtin_amdgcn_wave_barrier();
          }
        if(acc && to_soft)
          s_soft[ns + __popcll(bs & lt)] = it;
        ns += sc;
      }
    if(acc && !to_soft)
      {
        const int slot = nl + __popcll(bm & lt);
        s_list[slot]   = it;
        if(MSOFT)
          s_h[slot] = hv;
      }
    nl += cnt;
  };

  bool overflow  = false;
  int pushed     = 1;
  const int max_push = N + tv.m;  // each entry is pushed at most once per group
  __syncthreads();

  // entries popped per lane and round (FP32: AG_WALK_POP); the records of all of them
  // are read before any is used.  More than one per lane deepens the stack enough to
  // overflow it and measured slower, so one is the default.
  constexpr int kPop = sizeof(T) == 4 ? AG_WALK_POP : 1;
  // An opened node pushes at most 8 node children.  Near the top of the stack fewer
  // entries are popped, so that the pushes can never pass the last kReserve slots; a
  // single entry (a depth-first step) adds at most 7 per tree level, which the
  // reserve holds down to the deepest level.  Deep walks (large isolated sets) then
  // narrow instead of overflowing.
  constexpr int kReserve = 8 * (AG_MAX_LEVELS + 2);
  // split group: all waves expand the same entries (only wave 0 keeps their
  // interactions) until the stack holds kSplitMin entries; then each wave keeps every
  // nsplit-th of them and walks on alone
  const int kSplitMin = AG_SPLIT_MIN * nsplit;
  bool split_done     = nsplit <= 1;
  while(sp > 0)
    {
      if(!split_done && sp >= kSplitMin)
        {
          int kept = 0;
          for(int c0 = 0; c0 < sp; c0 += 64)
            {
              const int i       = c0 + lane;
              const int e       = i < sp ? s_stack[i] : -1;
              const bool keep   = i < sp && i % nsplit == wsplit;
              const uint64_t km = __ballot(keep);
              __builtin_amdgcn_wave_barrier();  // the chunk is read before it is overwritten
              if(keep)
                s_stack[kept + __popcll(km & lt)] = e;
              kept += __popcll(km);
              __builtin_amdgcn_wave_barrier();
            }
          sp         = kept;
          split_done = true;
          if(sp == 0)
            break;
        }
      int take       = sp < 64 * kPop ? sp : 64 * kPop;
      const int room = (kStackCap - kReserve - sp) >> 3;
      take           = take < room ? take : (room > 1 ? room : 1);
      const int base = sp - take;
      n_popped += take;
      int ent[kPop];
#pragma unroll
      for(int h = 0; h < kPop; h++)
        ent[h] = lane + 64 * h < take ? s_stack[base + 64 * h + lane] : -1;
      sp = base;
      __builtin_amdgcn_wave_barrier();
      const bool keep_ia = split_done || wsplit == 0;

      uint4 rA[kPop], rB[kPop];
      int4 rM[kPop];
      if constexpr(sizeof(T) == 4)
        {
#pragma unroll
          for(int h = 0; h < kPop; h++)
            {
              const int e = ent[h];
              rA[h]       = make_uint4(0, 0, 0, 0);
              rB[h]       = make_uint4(0, 0, 0, 0);
              rM[h]       = make_int4(0, 0, 0, 0);
              if(kPos && e >= N)
                {
                  rM[h] = tv.nmeta[e - N];
                  rA[h] = tv.nrecA[e - N];
                  rB[h] = tv.nrecB[e - N];
                }
              else if(e >= N || e < -1)
                {
                  const int jn = e >= N ? e - N : -e - 2;
                  rM[h]        = tv.meta[jn];
                  rA[h]        = tv.recA[jn];
                  rB[h]        = tv.recB[jn];
                }
              else if(e >= 0)
                rA[h] = tv.prec[e];
            }
        }

      int nchs[kPop], cbs[kPop], nch_all = 0;
#pragma unroll
      for(int h = 0; h < kPop; h++)
        {
          const int e = ent[h];
          int act = 0;  // 1 accept node, 2 open node, 3 particle, 4 expand small node into particles
          int j   = 0;
          int pbase = 0, pcnt = 0;
          int4 mt   = make_int4(0, 0, 0, 0);  // node meta {particle children start, count, first particle, cnn | count8 << 4}
          Vec4<T> item{0, 0, 0, 0};
          T hj      = 0;
          bool near = false;  // within the softening length of the group's box
          if(e >= N || (kPos && e < -1))
            {
              j = e >= N ? e - N : -e - 2;  // node index (walks without kPos) or child-list position
              T m, ln, dcx, dcy, dcz, sxr, syr, szr;
              if constexpr(sizeof(T) == 4)
                {
                  const uint4 A = rA[h], Bq = rB[h];
                  mt      = rM[h];
                  m       = __uint_as_float(A.w);
                  ln      = __uint_as_float(Bq.w);
                  dcx     = (T)(int)(Bq.x - cfx) * unit;
                  dcy     = (T)(int)(Bq.y - cfy) * unit;
                  dcz     = (T)(int)(Bq.z - cfz) * unit;
                  sxr     = (T)(int)(A.x - cfx) * unit;
                  syr     = (T)(int)(A.y - cfy) * unit;
                  szr     = (T)(int)(A.z - cfz) * unit;
                }
              else
                {
                  mt  = tv.meta[j];
                  m   = tv.nm[j];
                  ln  = tv.len[j];
                  dcx = nearest_d(tv.cx[j] - c0, wc);
                  dcy = nearest_d(tv.cy[j] - c1, wc);
                  dcz = nearest_d(tv.cz[j] - c2, wc);
                  sxr = nearest_d(tv.sx[j] - c0, wc);
                  syr = nearest_d(tv.sy[j] - c1, wc);
                  szr = nearest_d(tv.sz[j] - c2, wc);
                }
              const T gx = (T)fabs(dcx) - hx, gy = (T)fabs(dcy) - hy, gz = (T)fabs(dcz) - hz;
              bool discard = false;
              if(PM)
                {
                  // forcetree_walk.c:217-244 for every target of the group
                  T eff   = rcut + (T)0.5 * ln;
                  discard = gx > eff || gy > eff || gz > eff;
                }
              if(!discard)
                {
                  T ddx = (T)fabs(sxr) - hx, ddy = (T)fabs(syr) - hy, ddz = (T)fabs(szr) - hz;
                  ddx   = ddx > 0 ? ddx : 0;
                  ddy   = ddy > 0 ? ddy : 0;
                  ddz   = ddz > 0 ? ddz : 0;
                  T r2  = ddx * ddx + ddy * ddy + ddz * ddz;  // nearest target distance bound
                  T len2 = ln * ln;
                  bool open;
                  if(wc.use_bh)
                    open = len2 > r2 * theta2;  // forcetree_walk.c:246-254
                  else
                    open = (len2 > r2 * (T)1.44) || (m * len2 > r2 * r2 * aold_min) ||
                           (gx < (T)0.6 * ln && gy < (T)0.6 * ln && gz < (T)0.6 * ln);  // :259-289
                  if(MSOFT)
                    {
                      hj = (T)hsoft[tv.maxsoft[j]];
                      if(!open && hj > hi_min && r2 < hj * hj)  // :294-320
                        open = true;
                    }
                  act  = open ? 2 : 1;
                  // an accepted node whose COM lies beyond Rcut of the whole group fails the
                  // table test (tabindex < NTAB, forcetree_walk.c:372) for every target
                  if(PM && AG_NODE_RCUT_PRUNE && !open && r2 > rcut2 * (T)1.0001)
                    act = 0;
                  item = Vec4<T>{sxr * cs, syr * cs, szr * cs, m};
                  near = r2 < hs2;
                  if(open)
                    {
                      // a small opened node is expanded straight into its particles, which are
                      // contiguous in Hilbert order
                      const int cnt8 = mt.w >> 4;  // particle count, saturated at 255
                      if(cnt8 <= leaf_bucket)
                        {
                          act   = 4;
                          pbase = mt.z;
                          pcnt  = cnt8;
                        }
                    }
                }
            }
          else if(e >= 0)
            {
              T x, y, z, m;
              if constexpr(sizeof(T) == 4)
                {
                  const uint4 P = rA[h];
                  x             = (T)(int)(P.x - cfx) * unit;
                  y             = (T)(int)(P.y - cfy) * unit;
                  z             = (T)(int)(P.z - cfz) * unit;
                  m             = __uint_as_float(P.w);
                }
              else
                {
                  x = nearest_d(tv.px[e] - c0, wc);
                  y = nearest_d(tv.py[e] - c1, wc);
                  z = nearest_d(tv.pz[e] - c2, wc);
                  m = tv.pm[e];
                }
              act  = 3;
              item = Vec4<T>{x * cs, y * cs, z * cs, m};
              if(MSOFT)
                hj = (T)hsoft[tv.ps[e]];
              if(PM || kNearList)
                {
                  // a particle farther than Rcut from the whole group fails the table test for every target
                  T ddx = (T)fabs(x) - hx, ddy = (T)fabs(y) - hy, ddz = (T)fabs(z) - hz;
                  ddx   = ddx > 0 ? ddx : 0;
                  ddy   = ddy > 0 ? ddy : 0;
                  ddz   = ddz > 0 ? ddz : 0;
                  const T dd2 = ddx * ddx + ddy * ddy + ddz * ddz;
                  if(PM && dd2 > rcut2 * (T)1.0001)
                    act = 0;
                  near = dd2 < hs2;
                }
            }

          // stage accepted entries in the LDS tile
          const bool accp = (act == 1 || act == 3) && item.w != (T)0 && keep_ia;
          stage(accp, near, item, hj);

          // particles of small opened nodes (contiguous ranges) and the particle children of
          // opened nodes go straight into the tile: the ranges of all lanes are laid end to
          // end and read 64 particles per chunk, one per lane, so that a chunk never adds
          // more than 64 entries and every read of a chunk is in flight at once; particles
          // beyond Rcut of the whole group are dropped here
          {
            int pc = 0;
            bool from_list = false;
            if(act == 4)
              pc = pcnt;
            else if(act == 2)
              {
                pc        = mt.y;
                pbase     = mt.x;
                from_list = true;
              }
            if(!keep_ia)
              pc = 0;
            const int pincl = wave_incl_scan_i(pc);
            const int ptot  = __builtin_amdgcn_readlane(pincl, 63);
            const int pst   = pincl - pc;  // first slot of this lane's range
            for(int k0 = 0; k0 < ptot; k0 += 64)
              {
                // slot t of the chunk holds particle k0 + t, at index pbase - pst + k0 + t of
                // the list (child list or particles) of the lane whose range covers it: that
                // lane marks the slot with pbase - pst + 512 (>= 1) and the list in bit 31
                const int lo = pst - k0 > 0 ? pst - k0 : 0, hi = pst + pc - k0 < 64 ? pst + pc - k0 : 64;
                const int mark = (pbase - pst + 512) | (from_list ? (int)0x80000000 : 0);
                for(int t = lo; t < hi; t++)
                  s_map[t] = mark;
                __builtin_amdgcn_wave_barrier();
                bool keep    = k0 + lane < ptot;
                const int mk = keep ? s_map[lane] : 0;
                __builtin_amdgcn_wave_barrier();  // the map is read before the next chunk writes it
                const bool fl = mk < 0;
                const int pi2 = (mk & 0x7fffffff) - 512 + k0 + lane;
                T x = 0, y = 0, z = 0, m = 0, hp = 0;
                bool pnear = false;
                if(keep)
                  {
                    if constexpr(sizeof(T) == 4)
                      {
                        // particle children come as records copied next to the child list
                        const uint4 P = fl ? tv.cprec[pi2] : tv.prec[pi2];
                        x             = (T)(int)(P.x - cfx) * unit;
                        y             = (T)(int)(P.y - cfy) * unit;
                        z             = (T)(int)(P.z - cfz) * unit;
                        m             = __uint_as_float(P.w);
                      }
                    else
                      {
                        const int e2 = fl ? tv.child[pi2] : pi2;
                        x            = nearest_d(tv.px[e2] - c0, wc);
                        y            = nearest_d(tv.py[e2] - c1, wc);
                        z            = nearest_d(tv.pz[e2] - c2, wc);
                        m            = tv.pm[e2];
                      }
                    if(MSOFT)
                      hp = (T)hsoft[tv.ps[fl ? tv.child[pi2] : pi2]];
                    keep = m != (T)0;
                    if((PM || kNearList) && keep)
                      {
                        T ddx = (T)fabs(x) - hx, ddy = (T)fabs(y) - hy, ddz = (T)fabs(z) - hz;
                        ddx   = ddx > 0 ? ddx : 0;
                        ddy   = ddy > 0 ? ddy : 0;
                        ddz   = ddz > 0 ? ddz : 0;
                        const T dd2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if(PM)
                          keep = dd2 <= rcut2 * (T)1.0001;
                        pnear = dd2 < hs2;
                      }
                  }
                stage(keep, pnear, Vec4<T>{x * cs, y * cs, z * cs, m}, hp);
              }
          }
          nchs[h] = (act == 2) ? (mt.w & 15) : 0;  // node children only (at most 8)
          cbs[h]  = mt.x - nchs[h];                  // node children precede the particle children
          nch_all += nchs[h];
        }

      // push the children of opened nodes
      const int incl = wave_incl_scan_i(nch_all);
      const int tot  = __builtin_amdgcn_readlane(incl, 63);
      pushed += tot;
      if(sp + tot > kStackCap || pushed > max_push)
        {
          overflow = true;
          break;
        }
      int o = sp + incl - nch_all;
#pragma unroll
      for(int h = 0; h < kPop; h++)
        {
          const int nc = nchs[h];
          if(kPos)
            {
              // node children are consecutive in the child list: push their positions
              for(int k = 0; k < nc; k++)
                s_stack[o + k] = N + cbs[h] + k;
              o += nc;
            }
          else if(nc > 0)
            {
              // all child indices are read before the first is stored (one memory round trip)
              int cv[8];
#pragma unroll
              for(int k = 0; k < 8; k++)
                cv[k] = tv.child[cbs[h] + (k < nc ? k : nc - 1)];
#pragma unroll
              for(int k = 0; k < 8; k++)
                if(k < nc)
                  s_stack[o + k] = cv[k];
              o += nc;
            }
        }
      sp += tot;
      __builtin_amdgcn_wave_barrier();
    }

  if(overflow)
    return false;
  __builtin_amdgcn_wave_barrier();
  n_tiled += nl;
  flush_tile();
  if constexpr(kSplit)
    {
      n_tiled += ns;
      soft_tile();
    }

  if constexpr(kFast)
    {
      // back from table units: accelerations x asmthfac^2, potentials x asmthfac
      const T c2 = cs * cs;
      ax *= c2, ay *= c2, az *= c2, cax *= c2, cay *= c2, caz *= c2;
      pp *= cs, cpp *= cs;
    }
  out = GroupAcc<T>{ax, ay, az, pp, cax, cay, caz, cpp, nin};
  unsigned long long v = has ? (unsigned long long)nin : 0ull;
  for(int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  cnt[0] += v;  // wave-uniform: interactions, stack entries evaluated, tile entries
  cnt[1] += n_popped;
  cnt[2] += n_tiled;
  return true;
}

__device__ inline void flush_counts(const unsigned long long (&cnt)[3], unsigned long long *isum)
{
  if((threadIdx.x & 63) == 0)
    for(int k = 0; k < 3; k++)
      if(cnt[k])
        atomicAdd(isum + k, cnt[k]);
}


// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumbering gives every XCD one contiguous run of groups: Hilbert-neighbouring
// groups, which read largely the same nodes, then share an L2.
__device__ inline int xcd_block(int b, int nb)
{
  if(!AG_XCD_REMAP)
    return b;
  const int per = nb >> 3, rem = nb & 7, x = b & 7, k = b >> 3;
  return x * per + (x < rem ? x : rem) + k;
}

// One group of gs consecutive targets per block of W wavefronts (W > 1: the
// group's traversal is shared, see walk_group); wave 0 adds the waves' partial
// sums in wave order, so the result is deterministic for a given W.
template <typename T, bool PM, bool EWALD, bool MSOFT, int W>
// the FP32 TreePM walk of one softening type is held to 96 VGPRs (5 waves per SIMD): measured faster
__global__ __launch_bounds__(64 * W, (sizeof(T) == 4 && PM && !EWALD && !MSOFT) ? 5 : 1) void k_walk_group(TreeView tv, WalkConst wc, const double *__restrict__ hsoft,
                                                       const float *__restrict__ oldacc_s, const float *__restrict__ tabs,
                                                       const float *__restrict__ ewald_f, const double *__restrict__ ewald_d, int gs,
                                                       const int *__restrict__ tmap, int ntarg, double *at, double *pt, int *nint,
                                                       int *failed, int *nfailed, unsigned long long *isum, FinalOut fin,
                                                       const int *__restrict__ gorder, int *gcost)
{
  __shared__ WalkLds<T, MSOFT> lds[W];
  __shared__ int s_fail;
  if(tree_invalid(tv))
    return;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  load_tables<T, PM, sizeof(T) == 4 && PM && !EWALD && !MSOFT>(lds[wv].tab, tabs, wc);
  if(threadIdx.x == 0)
    s_fail = 0;
  // the group: by the launch order (costliest first) when given, else in Hilbert order
  // with consecutive groups on one XCD.  Blocks go to the 8 XCDs in turn, so within every
  // full run of 64 blocks the order positions are transposed: each XCD then takes 8
  // consecutive positions (groups of similar cost, in Hilbert order) and shares their
  // nodes in its L2.  An order entry out of range (its sort did not complete, which the
  // host reports) walks nothing.
  int gpos = blockIdx.x;
  if(gorder && (int)(gpos | 63) < (int)gridDim.x)
    gpos = (gpos & ~63) | ((gpos & 7) << 3) | ((gpos >> 3) & 7);
  const int g = gorder ? gorder[gpos] : xcd_block(blockIdx.x, gridDim.x);
  if(g < 0 || g >= (int)gridDim.x)
    return;
  const int p0 = g * gs;
  const int np = min(gs, ntarg - p0);
  unsigned long long cnt[3] = {0, 0, 0};
  GroupAcc<T> acc;
  const bool ok = walk_group<T, PM, EWALD, MSOFT>(lds[wv], tv, wc, hsoft, oldacc_s, ewald_f, ewald_d, tmap, p0, np, tv.n,
                                                  (T)(float)tv.dom->fix_unit, wc.leaf_bucket, W, wv, acc, cnt);
  flush_counts(cnt, isum);
  // the group's cost for the launch order of the next walk: its tile entries (summed
  // over the wavefronts of a split group below)
  if(W == 1 && gcost && threadIdx.x == 0)
    gcost[g] = (int)min(cnt[2], (unsigned long long)INT_MAX);
  const int64_t pi = lane < np ? (tmap ? tmap[p0 + lane] : p0 + lane) : 0;
  if(W == 1)
    {
      if(ok && lane < np)
        {
          const double a0 = (double)acc.ax - (double)acc.cax, a1 = (double)acc.ay - (double)acc.cay,
                       a2 = (double)acc.az - (double)acc.caz, p = (double)acc.pp - (double)acc.cpp;
          at[3 * pi + 0] = a0;
          at[3 * pi + 1] = a1;
          at[3 * pi + 2] = a2;
          pt[pi]         = p;
          nint[pi]       = acc.nin;
          if(fin.perm)
            finalize_one(fin, pi, a0, a1, a2, p, acc.nin);
        }
      else if(!ok && lane == 0)
        {
          int slot     = atomicAdd(nfailed, 1);
          failed[slot] = g;
        }
      return;
    }
  // partial sums through LDS (the waves' stacks are free now)
  if(!ok && lane == 0)
    atomicOr(&s_fail, 1);
  double *red = reinterpret_cast<double *>(lds[wv].stack);  // 4 x 64 doubles + 64 ints < 4 KB
  red[lane]        = (double)acc.ax - (double)acc.cax;
  red[64 + lane]   = (double)acc.ay - (double)acc.cay;
  red[128 + lane]  = (double)acc.az - (double)acc.caz;
  red[192 + lane]  = (double)acc.pp - (double)acc.cpp;
  reinterpret_cast<int *>(red + 256)[lane] = acc.nin;
  if(lane == 0)
    reinterpret_cast<unsigned long long *>(red + 288)[0] = cnt[2];  // after the 64 counts; < 4 KB
  __syncthreads();
  if(wv == 0 && gcost && lane == 0)
    {
      unsigned long long te = 0;
      for(int v = 0; v < W; v++)
        te += reinterpret_cast<const unsigned long long *>(reinterpret_cast<const double *>(lds[v].stack) + 288)[0];
      gcost[g] = (int)min(te, (unsigned long long)INT_MAX);
    }
  if(wv == 0)
    {
      if(s_fail)
        {
          if(lane == 0)
            {
              int slot     = atomicAdd(nfailed, 1);
              failed[slot] = g;
            }
        }
      else if(lane < np)
        {
          double s[4] = {0, 0, 0, 0};
          int ni      = 0;
          for(int v = 0; v < W; v++)
            {
              const double *r = reinterpret_cast<const double *>(lds[v].stack);
              for(int k = 0; k < 4; k++)
                s[k] += r[64 * k + lane];
              ni += reinterpret_cast<const int *>(r + 256)[lane];
            }
          at[3 * pi + 0] = s[0];
          at[3 * pi + 1] = s[1];
          at[3 * pi + 2] = s[2];
          pt[pi]         = s[3];
          nint[pi]       = ni;
          if(fin.perm)
            finalize_one(fin, pi, s[0], s[1], s[2], s[3], ni);
        }
    }
}

// Batch of independent isolated sets: a persistent grid whose waves take groups
// {first target, count, set, leaf bucket} from one queue (largest sets first)
// until it is empty; every wave leaves when the counter passes the group count.
// (Per-XCD queues over contiguous parts of the list measured 3 % slower: the
// sets are small enough to stay in L2 either way, and the split adds a tail.)
// Within a full solve the epilogue also writes the final outputs (fin.perm set).
// The FP32 kernel is held to 96 VGPRs (5 wavefronts per SIMD; its LDS allows 19 per
// CU): 4096-set batch walk 63.7 against 66.1 ms at the compiler's 98.
#ifndef AG_BATCH_MINW
#define AG_BATCH_MINW 5
#endif
template <typename T>
__global__ __launch_bounds__(64, sizeof(T) == 4 ? AG_BATCH_MINW : 1) void k_walk_batch(TreeView tv, WalkConst wc, BatchView bv, const float *__restrict__ oldacc_s,
                                                   const float *__restrict__ tabs, int ngroups, int *queue, double *at, double *pt,
                                                   int *nint, int *failed, int *nfailed, unsigned long long *isum, FinalOut fin)
{
  __shared__ WalkLds<T, false> lds;
  load_tables<T, false, false>(lds.tab, tabs, wc);
  unsigned long long cnt[3] = {0, 0, 0};
  for(;;)
    {
      int g = 0;
      if(threadIdx.x == 0)
        g = atomicAdd(queue, 1);
      g = __shfl(g, 0, 64);
      if(g >= ngroups)
        break;
      const int4 d = bv.groups[g];
      const T unit = (T)bv.dom[d.z].fix_unit;
      GroupAcc<T> acc;
      if(walk_group<T, false, false, false>(lds, tv, wc, bv.sh + d.z, oldacc_s, nullptr, nullptr, nullptr, d.x, d.y, bv.sroot[d.z],
                                            unit, d.w, 1, 0, acc, cnt))
        {
          if((int)threadIdx.x < d.y)
            {
              const int64_t pi = (int64_t)d.x + threadIdx.x;
              const double a0 = (double)acc.ax - (double)acc.cax, a1 = (double)acc.ay - (double)acc.cay,
                           a2 = (double)acc.az - (double)acc.caz, p = (double)acc.pp - (double)acc.cpp;
              at[3 * pi + 0] = a0;
              at[3 * pi + 1] = a1;
              at[3 * pi + 2] = a2;
              pt[pi]         = p;
              nint[pi]       = acc.nin;
              if(fin.perm)
                finalize_one(fin, pi, a0, a1, a2, p, acc.nin);
            }
        }
      else if(threadIdx.x == 0)
        {
          int slot     = atomicAdd(nfailed, 1);
          failed[slot] = g;
        }
    }
  flush_counts(cnt, isum);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

WalkConst walk_const(const agrav_ctx *c, int use_bh)
{
  WalkConst wc{};
  wc.periodic = c->par.periodic;
  wc.box      = c->par.box_size;
  wc.half_box = 0.5 * c->par.box_size;
  wc.pm       = c->par.pmgrid > 0;
  if(wc.pm)
    {
      double asmth  = c->par.asmth * c->par.box_size / c->par.pmgrid;  // pm_periodic.c:190-191
      double rcut   = c->par.rcut * asmth;
      wc.rcut       = rcut;
      wc.rcut2      = rcut * rcut;
      double ai     = 0.5 / asmth;
      wc.asmthinv2  = ai * ai;
      wc.asmthfac   = ai * (AG_NTAB / (c->par.rcut / 2.0));
    }
  wc.theta2    = c->par.err_tol_theta * c->par.err_tol_theta;
  wc.errtol    = c->par.err_tol_force_acc;
  wc.use_bh    = use_bh;
  wc.ewald     = c->par.periodic && c->par.pmgrid == 0;
  wc.ewald_fac = wc.ewald ? 2.0 * 64 / c->par.box_size : 0;
  wc.multisoft = c->soft.nsoft > 1;
  wc.leaf_bucket = c->par.leaf_bucket > 0 ? c->par.leaf_bucket : 0;
  wc.eval_pot    = c->par.eval_potential ? 1 : 0;
  return wc;
}

// the walk's targets: every particle, or the active-target list (sorted positions in
// Hilbert order); a partial tree's targets are its own particles
// (during one pass of a walk split by the high-resolution flag: that pass's list)
static const int *target_map(const agrav_ctx *c)
{
  if(c->walk_nt >= 0)
    return c->walk_tmap;
  return c->n_active >= 0 && !c->tree_partial ? c->d_tmap.p : nullptr;
}
static int target_count(const agrav_ctx *c)
{
  if(c->walk_nt >= 0)
    return (int)c->walk_nt;
  return c->tree_partial ? (int)c->n_tree : c->n_active >= 0 ? (int)c->n_active : (int)c->n;
}

template <typename T, bool PM, bool EW, bool MS, int W>
static void launch_group_w(int ngroups, hipStream_t s, const TreeView &tv, const WalkConst &wc, agrav_ctx *c, int gs)
{
  k_walk_group<T, PM, EW, MS, W><<<ngroups, 64 * W, 0, s>>>(tv, wc, c->d_soft.p, c->d_oldacc_s.p, c->d_tab.p, c->d_ewald_f.p,
                                                            c->d_ewald.p, gs, target_map(c), target_count(c), c->d_at.p,
                                                            c->d_pt.p, c->d_nint.p, c->d_failed.p,
                                                            reinterpret_cast<int *>(c->d_isum.p + 3), c->d_isum.p, c->fin,
                                                            c->walk_gorder, c->walk_gcost);
}

template <typename T, bool PM, bool EW, bool MS>
static void launch_group(int ngroups, hipStream_t s, const TreeView &tv, const WalkConst &wc, agrav_ctx *c, int gs, int waves)
{
  if constexpr(sizeof(T) == 4)
    {
      if(waves == 4)
        return launch_group_w<T, PM, EW, MS, 4>(ngroups, s, tv, wc, c, gs);
      if(waves == 2)
        return launch_group_w<T, PM, EW, MS, 2>(ngroups, s, tv, wc, c, gs);
    }
  launch_group_w<T, PM, EW, MS, 1>(ngroups, s, tv, wc, c, gs);
}

template <typename T>
static void dispatch_group(bool pm, bool ew, bool ms, int ngroups, hipStream_t s, const TreeView &tv, const WalkConst &wc,
                           agrav_ctx *c, int gs, int waves)
{
  if(pm)
    {
      if(ms)
        launch_group<T, true, false, true>(ngroups, s, tv, wc, c, gs, waves);
      else
        launch_group<T, true, false, false>(ngroups, s, tv, wc, c, gs, waves);
    }
  else if(ew)
    {
      if(ms)
        launch_group<T, false, true, true>(ngroups, s, tv, wc, c, gs, waves);
      else
        launch_group<T, false, true, false>(ngroups, s, tv, wc, c, gs, waves);
    }
  else
    {
      if(ms)
        launch_group<T, false, false, true>(ngroups, s, tv, wc, c, gs, waves);
      else
        launch_group<T, false, false, false>(ngroups, s, tv, wc, c, gs, waves);
    }
}

static void launch_plain(bool pm, bool ew, int nt, const int *tlist, int tl_group, hipStream_t s, const TreeView &tv,
                         const WalkConst &wc, agrav_ctx *c, const BatchView &bv)
{
  unsigned g = (unsigned)((nt + 255) / 256);
  if(g == 0)
    return;
  const int *tm  = bv.pset ? nullptr : target_map(c);
  const int ntar = bv.pset ? (int)c->n : target_count(c);
  if(pm)
    k_walk_plain<true, false><<<g, 256, 0, s>>>(tv, wc, c->d_soft.p, c->d_oldacc_s.p, c->d_tab.p, c->d_ewald.p, nt, tlist,
                                                 tl_group, tm, ntar, bv, c->d_at.p, c->d_pt.p, c->d_nint.p, c->d_isum.p);
  else if(ew)
    k_walk_plain<false, true><<<g, 256, 0, s>>>(tv, wc, c->d_soft.p, c->d_oldacc_s.p, c->d_tab.p, c->d_ewald.p, nt, tlist,
                                                 tl_group, tm, ntar, bv, c->d_at.p, c->d_pt.p, c->d_nint.p, c->d_isum.p);
  else
    k_walk_plain<false, false><<<g, 256, 0, s>>>(tv, wc, c->d_soft.p, c->d_oldacc_s.p, c->d_tab.p, c->d_ewald.p, nt, tlist,
                                                  tl_group, tm, ntar, bv, c->d_at.p, c->d_pt.p, c->d_nint.p, c->d_isum.p);
}

BatchView batch_view(const agrav_ctx *c)
{
  BatchView bv{};
  if(c->batch_nsets > 0)
    bv = BatchView{c->d_groups.p, c->d_pset.p, c->d_sroot.p, c->d_send.p, c->d_shd.p, c->d_sdom.p};
  return bv;
}

// persistent grid of the batched walk: as many one-wave blocks as fit on the device at once
template <typename T> static int batch_grid(agrav_ctx *c)
{
  int per_cu = 0, cus = 0;
  if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_walk_batch<T>), 64, 0) != hipSuccess ||
     per_cu < 1)
    {
      (void)hipGetLastError();
      per_cu = 8;
    }
  if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
    {
      (void)hipGetLastError();
      cus = 256;
    }
  return per_cu * cus;
}

// One batched group walk over the groups [0, ngroups) of `groups` (all sets of the
// batch, or a sample of them while tuning); failed groups are re-walked one lane
// per target.  Interaction totals accumulate in d_isum.
agrav_status walk_batch_groups(agrav_ctx *c, int use_bh, const int4 *groups, int ngroups, float *kernel_ms, const FinalOut &fin,
                               bool *fused)
{
  hipStream_t s      = c->stream;
  agrav_status st;
  WalkConst wc       = walk_const(c, use_bh);
  wc.multisoft       = 0;
  const TreeView tv  = tree_view(c);
  BatchView bv       = batch_view(c);
  bv.groups          = groups;
  if(!c->have_tree || c->batch_nsets < 1)
    return fail(c, AGRAV_ERR_STATE, "batched walk without a batch tree");
  // every buffer the kernels write (the tuner may walk before any solve has)
  const int64_t nn = c->n ? c->n : 1;
  if((st = c->d_at.ensure(3 * nn)) || (st = c->d_pt.ensure(nn)) || (st = c->d_nint.ensure(nn)) || (st = c->d_isum.ensure(4)) ||
     (st = c->d_failed.ensure(ngroups > 0 ? ngroups : 1)))
    return st;
  int *queue = c->d_counters.p + 52, *nfailed = c->d_counters.p + 48;
  AG_HIP(hipMemsetAsync(c->d_counters.p + 48, 0, 8 * sizeof(int), s));
  AG_HIP(hipEventRecord(c->ev[2], s));
  if(ngroups > 0)
    {
      if(c->par.precision == AGRAV_FP32)
        k_walk_batch<float><<<batch_grid<float>(c), 64, 0, s>>>(tv, wc, bv, c->d_oldacc_s.p, c->d_tab.p, ngroups, queue, c->d_at.p,
                                                                 c->d_pt.p, c->d_nint.p, c->d_failed.p, nfailed, c->d_isum.p, fin);
      else
        k_walk_batch<double><<<batch_grid<double>(c), 64, 0, s>>>(tv, wc, bv, c->d_oldacc_s.p, c->d_tab.p, ngroups, queue,
                                                                   c->d_at.p, c->d_pt.p, c->d_nint.p, c->d_failed.p, nfailed,
                                                                   c->d_isum.p, fin);
      AG_HIP(hipGetLastError());
    }
  AG_HIP(hipEventRecord(c->ev[3], s));
  int nf = 0;
  AG_HIP(hipMemcpyAsync(&nf, nfailed, sizeof(int), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(nf > 0)
    {
      c->tim.walk_retries += nf;
      launch_plain(false, false, nf * 64, c->d_failed.p, 64, s, tv, wc, c, bv);
      AG_HIP(hipGetLastError());
    }
  // re-walked targets still need the finalize kernel
  if(fused)
    *fused = fin.perm != nullptr && nf == 0;
  if(kernel_ms)
    AG_HIP(hipEventElapsedTime(kernel_ms, c->ev[2], c->ev[3]));
  return AGRAV_OK;
}

// active targets (caller indices) -> their sorted positions in Hilbert order: a flag
// per sorted position, a scan, a scatter; a repeated or out-of-range index is an error
__global__ void k_flag_active(const int *__restrict__ active, int64_t na, int64_t n, const int *__restrict__ inv, int *flag,
                              int *bad)
{
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(j >= na)
    return;
  const int i = active[j];
  if(i < 0 || i >= n)
    {
      atomicOr(bad, 1);
      return;
    }
  if(atomicExch(&flag[inv[i]], 1) != 0)
    atomicOr(bad, 2);
}

__global__ void k_inverse_perm(const int *__restrict__ perm, int64_t n, int *inv)
{
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(k < n)
    inv[perm[k]] = (int)k;
}

__global__ void k_fill_tmap(const int *__restrict__ flag, const int *__restrict__ pos, int64_t n, int *tmap)
{
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(k < n && flag[k])
    tmap[pos[k]] = (int)k;
}

agrav_status active_targets(agrav_ctx *c)
{
  if(!c->tmap_dirty)
    return AGRAV_OK;
  const int64_t n = c->n, na = c->n_active;
  hipStream_t s   = c->stream;
  agrav_status st;
  if((st = c->d_tmap.ensure(na ? na : 1)) || (st = c->d_aflag.ensure(n ? n : 1)) || (st = c->d_apos.ensure(n ? n : 1)) ||
     (st = c->d_ainv.ensure(n ? n : 1)) || (st = c->d_temp.ensure(scan_temp_bytes(n) + 4096)))
    return st;
  int *d_bad = c->d_counters.p + 102, *d_err = c->d_counters.p + 103;
  AG_HIP(hipMemsetAsync(d_bad, 0, 2 * sizeof(int), s));
  if(na > 0)
    {
      const unsigned gn = (unsigned)((n + 255) / 256), ga = (unsigned)((na + 255) / 256);
      AG_HIP(hipMemsetAsync(c->d_aflag.p, 0, n * sizeof(int), s));
      k_inverse_perm<<<gn, 256, 0, s>>>(c->d_idx2.p, n, c->d_ainv.p);
      k_flag_active<<<ga, 256, 0, s>>>(c->d_active.p, na, n, c->d_ainv.p, c->d_aflag.p, d_bad);
      AG_HIP(hipGetLastError());
      if((st = exclusive_scan_i32(c->d_aflag.p, c->d_apos.p, n, nullptr, c->d_temp.p, d_err, s)) != AGRAV_OK)
        return st;
      k_fill_tmap<<<gn, 256, 0, s>>>(c->d_aflag.p, c->d_apos.p, n, c->d_tmap.p);
      AG_HIP(hipGetLastError());
    }
  int bad[2];
  AG_HIP(hipMemcpyAsync(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, s));
  AG_HIP(hipStreamSynchronize(s));
  if(bad[0] & 1)
    return fail(c, AGRAV_ERR_BAD_PARAM, "active index out of range");
  if(bad[0] & 2)
    return fail(c, AGRAV_ERR_BAD_PARAM, "active list repeats a particle");
  if(bad[1])
    return fail(c, AGRAV_ERR_HIP, "active-target scan look-back did not complete");
  c->tmap_dirty = false;
  return AGRAV_OK;
}

// Sort key of each walk group: its cost (tile entries of the last walk of the same
// grouping) as an 8-bit code falling with log2 of the cost (1/8 octave per step), so
// that one stable radix pass orders the groups costliest first and keeps groups of
// similar cost in Hilbert order.  Block 0 also clears the pass's digit counts and the
// walk's counters (where the sort reports a look-back that did not complete).
__global__ void k_group_keys(int ngroups, const int *__restrict__ gcost, uint64_t *keys, int *ids, int *ghist,
                             unsigned long long *isum)
{
  if(blockIdx.x == 0)
    {
      ghist[threadIdx.x] = 0;  // 256 threads, one digit each
      if(threadIdx.x < 4)
        isum[threadIdx.x] = 0;
    }
  const int g = blockIdx.x * 256 + threadIdx.x;
  if(g >= ngroups)
    return;
  const int v    = gcost[g];
  const int code = min(255, (int)(8.0f * __log2f(1.0f + (float)(v > 0 ? v : 0))));
  keys[g]        = (uint64_t)(255 - code);
  ids[g]         = g;
}

// The group sort's look-back error flag is the upper half of the walk counters' last
// word (the lower half counts the failed groups; the walk clears both), so the walk's
// one read-back of its counters carries it.
static int *group_sort_err(agrav_ctx *c) { return reinterpret_cast<int *>(c->d_isum.p + 3) + 1; }
// fewer wavefronts than this (groups x wavefronts per group) fill the device in about
// one round and gain nothing from an order
static constexpr int kOrderMinWaves = 8192;

// launch order of the group walk from the costs in d_gcost, or null
static agrav_status group_order(agrav_ctx *c, int ngroups, const int **order, hipStream_t s)
{
  *order = nullptr;
  agrav_status st;
  const size_t tb = sort_temp_bytes(ngroups);  // look-back states, then 256 digit counts
  if((st = c->d_gkey.ensure(ngroups)) || (st = c->d_gkey2.ensure(ngroups)) || (st = c->d_gord.ensure(ngroups)) ||
     (st = c->d_gord2.ensure(ngroups)) || (st = c->d_gtemp.ensure(tb + 256 * sizeof(int))))
    return st;
  int *ghist = reinterpret_cast<int *>(c->d_gtemp.p + tb);
  k_group_keys<<<(ngroups + 255) / 256, 256, 0, s>>>(ngroups, c->d_gcost.p, c->d_gkey.p, c->d_gord.p, ghist, c->d_isum.p);
  AG_HIP(hipGetLastError());
  bool in_input = false;
  if((st = radix_sort_u64(c->d_gkey.p, c->d_gord.p, c->d_gkey2.p, c->d_gord2.p, ngroups, 8, c->d_gtemp.p,
                          group_sort_err(c), s, 0, &in_input, ghist)) != AGRAV_OK)
    return st;
  *order = in_input ? c->d_gord.p : c->d_gord2.p;
  return AGRAV_OK;
}

static agrav_status walk_pass(agrav_ctx *c, int use_bh, const WalkConst &wc, const TreeView &tv);

// Group size and wavefronts per group of a group walk over nt of a tree's ntree
// particles, by workload when not set (MI355X timings of FP32 walks): periodic sets
// and large isolated sets fill the device with groups of 64; small isolated sets are
// bound by the latency of each group's walk and gain from groups of 16 (200k Hernquist
// particles: 1.3 ms against 8.3 ms) and from sharing a group over 4 wavefronts
static void group_shape(const agrav_ctx *c, int64_t nt, int64_t ntree, bool isolated, int *gs_out, int *waves_out)
{
  int gs = c->par.group_size;
  if(gs <= 0)
    {
      gs = (isolated && nt < (1 << 19)) ? 16 : 64;
      // a sparse active list spreads a group over a larger volume (more nodes opened)
      // and gives fewer groups: about 256 N_active / N targets per group, 8 at least
      // (2M-particle clustered TreePM box, 3 % active: walk 1.7 ms against 5.0 ms)
      if(nt < ntree / 2)
        {
          int sg = 64;
          while(sg > 8 && (int64_t)sg * ntree > 256 * nt)
            sg >>= 1;
          gs = std::min(gs, sg);
        }
    }
  // wavefronts per group: small sets have too few groups to fill the device
  int waves = c->par.group_waves;
  if(waves <= 0)
    waves = isolated ? (nt < (1 << 17) ? 4 : 1) : (nt < (1 << 18) ? 4 : (nt < (1 << 20) ? 2 : 1));
  *gs_out    = gs;
  *waves_out = waves;
}

// A full step's launch order, sorted on the side stream while the tree is built (the
// build calls this once its sort is queued): it needs only the costs of the last walk of
// the same grouping (every particle, groups of the size this step's walk will take).  The
// key kernel clears the walk's counters, where the sort reports a look-back that did not
// complete; the walk waits for the sort's event and leaves the counters as they are.
agrav_status walk_preorder(agrav_ctx *c)
{
  c->pre_ngroups = -1;
  const int64_t n = c->n;
  if(c->par.walk == AGRAV_WALK_PLAIN || !c->mom_stream || n <= 0 || c->batch_nsets > 0)
    return AGRAV_OK;
  const bool pm = c->par.pmgrid > 0, ew = c->par.periodic && !pm;
  int gs = 0, waves = 0;
  group_shape(c, n, n, !pm && !ew, &gs, &waves);
  const int ngroups     = (int)((n + gs - 1) / gs);
  const int order_waves = c->par.precision == AGRAV_FP32 ? waves : 1;
  if((int64_t)ngroups * order_waves < kOrderMinWaves || c->gcost_groups != ngroups || c->gcost_gs != gs || c->gcost_n != n)
    return AGRAV_OK;
  agrav_status st;
  if((st = c->d_isum.ensure(4)) != AGRAV_OK)
    return st;
  c->isum_clear = true;
  AG_HIP(hipEventRecord(c->ev_mfork, c->stream));
  AG_HIP(hipStreamWaitEvent(c->mom_stream, c->ev_mfork, 0));
  const int *order = nullptr;
  if((st = group_order(c, ngroups, &order, c->mom_stream)) != AGRAV_OK)
    return st;
  AG_HIP(hipEventRecord(c->ev_order, c->mom_stream));
  c->pre_gorder  = order;
  c->pre_ngroups = ngroups;
  c->pre_gs      = gs;
  return AGRAV_OK;
}

agrav_status walk_tree(agrav_ctx *c, int use_bh)
{
  if(!c->have_tree)
    return fail(c, AGRAV_ERR_STATE, "gravity_tree called before construct_forcetree");
  const int64_t n = c->n;
  hipStream_t s   = c->stream;
  agrav_status st;
  if((st = c->d_at.ensure(3 * (n ? n : 1))) || (st = c->d_pt.ensure(n ? n : 1)) || (st = c->d_nint.ensure(n ? n : 1)) ||
     (st = c->d_isum.ensure(4)))
    return st;
  if(!c->isum_clear)
    AG_HIP(hipMemsetAsync(c->d_isum.p, 0, 4 * sizeof(unsigned long long), s));
  c->isum_clear = false;  // cleared by the pre-order of this step's first walk
  if(n == 0)
    {
      c->have_tree_force = true;
      return AGRAV_OK;
    }
  if(!use_bh && c->par.type_of_opening_criterion == 0)
    use_bh = 1;
  const WalkConst wc = walk_const(c, use_bh);
  const TreeView tv  = tree_view(c);
  const bool ew      = wc.ewald;
  if(ew && !c->ewald_ready)
    {
      if((st = ewald_tree_table(c)) != AGRAV_OK)
        return st;
    }
  if(!ew)
    {
      // kernels take the pointers unconditionally
      if((st = c->d_ewald.ensure(1)) || (st = c->d_ewald_f.ensure(1)))
        return st;
    }
  if(c->tree_partial && c->tree_active_version != c->active_version)
    return fail(c, AGRAV_ERR_STATE, "the active list changed after the partial tree was built");
  // the active-target list in Hilbert order, after each build or change of the list
  if(c->n_active >= 0 && !c->tree_partial && (st = active_targets(c)) != AGRAV_OK)
    return st;
  const int64_t nt = target_count(c);
  if(nt == 0)
    {
      c->tim.treewalk = c->tim.walk_kernel = 0;
      c->tim.interactions                  = 0;
      c->have_tree_force                   = true;
      return AGRAV_OK;
    }
  c->tim.walk_retries = 0;
  if(!hr_on(c) || c->batch_nsets > 0)
    return walk_pass(c, use_bh, wc, tv);
  // a high-resolution region: its targets take the short-range force with Asmth[1] and
  // Rcut[1] (forcetree_walk.c:143-150), the others with Asmth[0] and Rcut[0]; one pass
  // over each list, each in Hilbert order
  const int *lo = nullptr, *hi = nullptr;
  int64_t n_lo = 0, n_hi = 0;
  if((st = hr_split_targets(c, target_map(c), nt, &lo, &n_lo, &hi, &n_hi)) != AGRAV_OK)
    return st;
  WalkConst wc1   = wc;
  const double ai = 0.5 / c->hr.asmth;
  wc1.rcut        = c->hr.rcut;
  wc1.rcut2       = c->hr.rcut * c->hr.rcut;
  wc1.asmthinv2   = ai * ai;
  wc1.asmthfac    = ai * (AG_NTAB / (c->par.rcut / 2.0));
  c->walk_tmap = lo;
  c->walk_nt   = n_lo;
  st           = n_lo > 0 ? walk_pass(c, use_bh, wc, tv) : AGRAV_OK;
  const bool fused0   = n_lo == 0 || c->final_fused;
  const double t_walk = c->tim.treewalk, t_kernel = c->tim.walk_kernel;
  const int retries   = c->tim.walk_retries;
  if(st == AGRAV_OK && n_hi > 0)
    {
      c->walk_tmap = hi;
      c->walk_nt   = n_hi;
      AG_HIP(hipMemsetAsync(c->d_isum.p + 3, 0, sizeof(unsigned long long), s));  // the failed-group counter
      st = walk_pass(c, use_bh, wc1, tv);
      c->tim.treewalk += n_lo > 0 ? t_walk : 0;
      c->tim.walk_kernel += n_lo > 0 ? t_kernel : 0;
      c->tim.walk_retries += retries;
      c->final_fused = c->final_fused && fused0;
    }
  c->walk_tmap = nullptr;
  c->walk_nt   = -1;
  return st;
}

// one walk over the targets of target_map / target_count with the constants wc; the
// interaction counters accumulate over the passes of a step (each pass reads their totals)
static agrav_status walk_pass(agrav_ctx *c, int use_bh, const WalkConst &wc, const TreeView &tv)
{
  hipStream_t s   = c->stream;
  agrav_status st;
  const bool pm = wc.pm, ew = wc.ewald, ms = wc.multisoft;
  const int64_t nt = target_count(c);
  c->tim.walk_retries = 0;
  c->tim.walk_ordered = 0;
  unsigned long long cnt[4] = {0, 0, 0, 0};
  bool have_cnt       = false;  // the counters are read already (group walk, no re-walk)
  c->final_fused      = false;
  c->fin.perm         = nullptr;
  // a launch order sorted beside the build (walk_preorder) is taken by the group walk of
  // its grouping below; any other walk clears the counters its sort may have flagged
  const bool pre = c->pre_ngroups >= 0;
  if(pre)
    AG_HIP(hipStreamWaitEvent(s, c->ev_order, 0));
  if(pre && (c->batch_nsets > 0 || c->par.walk == AGRAV_WALK_PLAIN))
    {
      AG_HIP(hipMemsetAsync(c->d_isum.p, 0, 4 * sizeof(unsigned long long), s));
      c->pre_ngroups = -1;
    }
  if(c->batch_nsets > 0 && c->par.walk != AGRAV_WALK_PLAIN)
    {
      // batch of sets: one persistent launch over the group queue; within a full solve
      // its epilogue writes the final outputs too
      FinalOut fin{};
      bool fused = false;
      if((st = batch_groups(c)) != AGRAV_OK || (c->fuse_final && (st = final_out(c, fin)) != AGRAV_OK) ||
         (st = walk_batch_groups(c, use_bh, c->d_groups.p, c->batch_ngroups, nullptr, fin, &fused)) != AGRAV_OK)
        return st;
      c->final_fused = fused;
    }
  else if(c->par.walk == AGRAV_WALK_PLAIN)
    {
      AG_HIP(hipEventRecord(c->ev[2], s));
      launch_plain(pm, ew, (int)nt, nullptr, 1, s, tv, wc, c, batch_view(c));
      AG_HIP(hipGetLastError());
      AG_HIP(hipEventRecord(c->ev[3], s));
    }
  else
    {
      // by workload when not set (MI355X timings of FP32 walks): periodic sets and large
      // isolated sets fill the device with groups of 64; small isolated sets are bound
      // by the latency of each group's walk and gain from groups of 16 (200k Hernquist
      // particles: 1.3 ms against 8.3 ms) and from sharing a group over 4 wavefronts
      int gs = 0, waves = 0;
      group_shape(c, nt, c->n_tree, !pm && !ew, &gs, &waves);
      int ngroups = (int)((nt + gs - 1) / gs);
      if((st = c->d_failed.ensure(ngroups)) != AGRAV_OK)
        return st;
      c->tim_group_waves = c->par.precision == AGRAV_FP32 ? waves : 1;
      // within a full step the walk's epilogue also writes the final outputs (the PM
      // has run, in sorted order); the finalize kernel then has nothing left to do
      if(c->fuse_final && (!pm || (c->have_pm && c->pm_sorted)) && (st = final_out(c, c->fin)) != AGRAV_OK)
        return st;
      // costliest groups first, by the costs of the last walk of the same grouping: every
      // target of as many particles, in groups of the same size; this walk's costs order
      // the next one
      const bool whole  = c->batch_nsets == 0 && target_map(c) == nullptr && nt == c->n_tree;
      const int *gorder = nullptr;
      const int order_waves = c->par.precision == AGRAV_FP32 ? waves : 1;
      if(pre && whole && c->pre_ngroups == ngroups && c->pre_gs == gs)
        gorder = c->pre_gorder;  // sorted beside the build
      else if(pre)
        AG_HIP(hipMemsetAsync(c->d_isum.p, 0, 4 * sizeof(unsigned long long), s));
      c->pre_ngroups = -1;
      if(!gorder && whole && (int64_t)ngroups * order_waves >= kOrderMinWaves && c->gcost_groups == ngroups &&
         c->gcost_gs == gs && c->gcost_n == c->n && (st = group_order(c, ngroups, &gorder, s)) != AGRAV_OK)
        return st;
      if(whole && (st = c->d_gcost.ensure(ngroups)) != AGRAV_OK)
        return st;
      c->gcost_groups     = -1;  // valid again once this walk has run
      // the walk's time starts here: the order kernels above are not part of it
      AG_HIP(hipEventRecord(c->ev[2], s));
      c->walk_gorder      = gorder;
      c->walk_gcost       = whole ? c->d_gcost.p : nullptr;
      c->tim.walk_ordered = gorder != nullptr;
      if(c->par.precision == AGRAV_FP32)
        dispatch_group<float>(pm, ew, ms, ngroups, s, tv, wc, c, gs, waves);
      else
        dispatch_group<double>(pm, ew, ms, ngroups, s, tv, wc, c, gs, 1);
      c->walk_gorder = nullptr;
      c->walk_gcost  = nullptr;
      AG_HIP(hipGetLastError());
      if(whole)
        {
          c->gcost_groups = ngroups;
          c->gcost_gs     = gs;
          c->gcost_n      = c->n;
        }
      AG_HIP(hipEventRecord(c->ev[3], s));
      AG_HIP(hipEventRecord(c->ev[4], s));  // the walk's end, unless groups are re-walked below
      // groups whose LDS stack overflowed (their count is isum[3]) are re-walked one
      // lane per target
      AG_HIP(hipMemcpyAsync(c->stage<unsigned long long>(0), c->d_isum.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      memcpy(cnt, c->stage<unsigned long long>(0), sizeof(cnt));
      if(gorder && (cnt[3] >> 32) != 0)
        return fail(c, AGRAV_ERR_HIP, "walk group order: sort look-back did not complete");
      const int nf = (int)(unsigned)cnt[3];
      have_cnt     = nf == 0;
      if(nf > 0)
        {
          c->tim.walk_retries += nf;
          launch_plain(pm, ew, nf * gs, c->d_failed.p, gs, s, tv, wc, c, BatchView{});
          AG_HIP(hipGetLastError());
        }
      // re-walked targets still need the finalize kernel
      c->final_fused = c->fin.perm != nullptr && nf == 0;
      c->fin.perm    = nullptr;
    }
  if(!have_cnt)
    {
      // other walks, or a group walk with re-walked groups: one more read-back
      AG_HIP(hipEventRecord(c->ev[4], s));
      AG_HIP(hipMemcpyAsync(c->stage<unsigned long long>(0), c->d_isum.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
      AG_HIP(hipStreamSynchronize(s));
      memcpy(cnt, c->stage<unsigned long long>(0), sizeof(cnt));
    }
  // a step's build is checked here, with the walk's read-back (its flags were tested by
  // the walk kernels on the device)
  if((st = build_check(c)) != AGRAV_OK)
    return st;
  const unsigned long long tot = cnt[0];
  c->tim.walk_popped           = (long long)cnt[1];
  c->tim.tile_entries          = (long long)cnt[2];
  float ms_walk = 0, ms_kernel = 0;
  AG_HIP(hipEventElapsedTime(&ms_kernel, c->ev[2], c->ev[3]));
  AG_HIP(hipEventElapsedTime(&ms_walk, c->ev[2], c->ev[4]));
  c->tim.treewalk     = ms_walk;
  c->tim.walk_kernel  = ms_kernel;
  c->tim.interactions = (long long)tot;
  c->have_tree_force  = true;
  return AGRAV_OK;
}

}  // namespace agrav
