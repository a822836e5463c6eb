/*
 * agrav_oracle.c -- CPU FP64 restatement of AREPO's gravity step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * The code restates the algorithm of the AREPO sources cited per function
 * (paths relative to AREPO's src/).  It is written from those semantics, not
 * copied: data live in plain arrays owned by an orc_tree, there is no MPI,
 * no top-level domain tree (one rank: its empty top nodes carry no mass and
 * do not change forces), and domain randomisation is off.
 */
#include "agrav_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* spline-softening coefficients, allvars.h:557-573 */
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

static int omp_threads(int nthreads)
{
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* parameters and tables                                               */
/* ------------------------------------------------------------------ */

void orc_set_softenings(const orc_params *p, double *fs)
{
  /* grav_softening.c:90-95 */
  for(int i = 0; i < p->nsoft; i++)
    fs[i] = 2.8 * p->soft[i];
  fs[p->nsoft] = 0;
}

void orc_short_range_tables(double rcut, float *tf, float *tp)
{
  /* forcetree_walk.c:72-85 */
  for(int i = 0; i <= ORC_NTAB; i++)
    {
      double u = ((rcut / 2.0) / ORC_NTAB) * i;
      tp[i]    = (float)(-erfc(u));
      if(u > 0)
        tf[i] = (float)((erfc(u) + 2.0 * u / sqrt(M_PI) * exp(-u * u) - 1.0) / (u * u));
      else
        tf[i] = 0;
    }
}

void orc_asmth_rcut(const orc_params *p, double *asmth_len, double *rcut_len)
{
  /* pm_periodic.c:190-191 */
  double a   = p->pmgrid > 0 ? p->asmth * p->box_size / p->pmgrid : 0;
  *asmth_len = a;
  *rcut_len  = p->rcut * a;
}

static inline double nearest(double x, double box, int periodic)
{
  /* GRAVITY_NEAREST_X, allvars.h:603-612 */
  if(!periodic)
    return x;
  double half = 0.5 * box;
  return x > half ? x - box : (x < -half ? x + box : x);
}

void orc_domain_extent(const orc_params *p, int n, const double *pos, double corner[3], double center[3], double *len_out)
{
  /* domain_box.c:107-197 with DO_NOT_RANDOMIZE_DOMAINCENTER */
  /* preset to the simulation box [0, BoxSize] (domain_box.c:112-116); isolated
   * runs pass BoxSize = 0 so that the cube is the particles' extent plus the origin */
  double xmin[3], xmax[3];
  for(int j = 0; j < 3; j++)
    {
      xmin[j] = 0;
      xmax[j] = p->box_size;
    }
  for(int i = 0; i < n; i++)
    for(int j = 0; j < 3; j++)
      {
        double x = pos[3 * i + j];
        if(p->periodic)
          {
            if(x < 0)
              x += p->box_size;
            if(x >= p->box_size)
              x -= p->box_size;
          }
        if(xmin[j] > x)
          xmin[j] = x;
        if(xmax[j] < x)
          xmax[j] = x;
      }
  double len = 0;
  for(int j = 0; j < 3; j++)
    if(xmax[j] - xmin[j] > len)
      len = xmax[j] - xmin[j];
  if(len <= 0)
    len = 1.0;
  len *= p->periodic ? 1.00001 : 1.2;
  for(int j = 0; j < 3; j++)
    {
      center[j] = 0.5 * (xmin[j] + xmax[j]);
      corner[j] = 0.5 * (xmin[j] + xmax[j]) - 0.5 * len;
    }
  *len_out = len;
}

/* ------------------------------------------------------------------ */
/* tree                                                                */
/* ------------------------------------------------------------------ */

struct orc_tree
{
  int n, maxnodes, numnodes;
  int *suns;          /* 8 per node */
  double *center;     /* 3 per node */
  double *len;
  double *s;          /* centre of mass, 3 per node */
  double *mass;
  int *sibling, *nextnode, *father;
  int *level;
  int *npart;
  uint8_t *maxsoft;
  int *pnext;         /* Nextnode[] of particles */
  uint64_t *intpos;   /* 3 per particle */
  double *pos;        /* wrapped positions, 3 per particle */
  const double *pmass;
  const uint8_t *stype;
  double force_softening[ORC_MAXSOFT + 1];
  int nsoft;
  double corner[3], dcenter[3], dlen, bigfac;
  int periodic;
  double box;
};

static uint64_t dbl_mantissa(double d)
{
  /* force_double_to_int, forcetree.h:122-131 */
  uint64_t u;
  memcpy(&u, &d, 8);
  return u & 0xFFFFFFFFFFFFFull;
}

static int tree_new_node(orc_tree *t)
{
  if(t->numnodes >= t->maxnodes)
    return -1;
  int k = t->numnodes++;
  for(int j = 0; j < 8; j++)
    t->suns[8 * k + j] = -1;
  return k;
}

/* forcetree.c:835-938; node k is referenced as n + k like Nodes[Tree_MaxPart + k] */
static int tree_insert(orc_tree *t, int i)
{
  const int N      = t->n;
  const uint64_t *ip = &t->intpos[3 * i];
  int th = N + 0, parent = -1, levels = 0;
  unsigned subnode = 0;
  uint64_t mask     = 1ull << 51;
  int shx = 51, shy = 50, shz = 49;
  int64_t centermask = (int64_t)0xFFF0000000000000ull;

  for(;;)
    {
      if(th >= N)
        {
          int k   = th - N;
          subnode = (unsigned)(((ip[0] & mask) >> shx) | ((ip[1] & mask) >> shy) | ((ip[2] & mask) >> shz));
          shx--;
          shy--;
          shz--;
          centermask >>= 1;
          mask >>= 1;
          levels++;
          if(levels > ORC_MAX_TREE_LEVEL)
            {
              /* coincident points: take the next free slot (forcetree.c:861-877) */
              for(int j = 0; j < 8; j++)
                {
                  if(t->suns[8 * k + subnode] < 0)
                    break;
                  subnode++;
                  if(subnode >= 8)
                    subnode = 7;
                }
            }
          int nn = t->suns[8 * k + subnode];
          if(nn >= 0)
            {
              parent = th;
              th     = nn;
            }
          else
            {
              t->suns[8 * k + subnode] = i;
              return 0;
            }
        }
      else
        {
          /* the slot holds one particle: replace it by a new node (forcetree.c:893-929) */
          int k = tree_new_node(t);
          if(k < 0)
            return -1;
          t->suns[8 * (parent - N) + subnode] = N + k;
          t->len[k]                            = (double)(mask << 1) * t->bigfac;
          for(int d = 0; d < 3; d++)
            t->center[3 * k + d] = (double)((ip[d] & (uint64_t)centermask) | mask) * t->bigfac + t->corner[d];
          t->level[k]       = levels;
          const uint64_t *jp = &t->intpos[3 * th];
          unsigned sub2 = (unsigned)(((jp[0] & mask) >> shx) | ((jp[1] & mask) >> shy) | ((jp[2] & mask) >> shz));
          t->suns[8 * k + sub2] = th;
          th                     = N + k;
        }
    }
}

/* forcetree.c:1217-1429 */
static void tree_update_recursive(orc_tree *t, int no, int sib, int father, int *last)
{
  const int N = t->n;
  if(no >= N)
    {
      int k = no - N;
      int suns[8];
      memcpy(suns, &t->suns[8 * k], sizeof(suns));
      if(*last >= 0)
        {
          if(*last >= N)
            t->nextnode[*last - N] = no;
          else
            t->pnext[*last] = no;
        }
      *last = no;

      double mass = 0, s[3] = {0, 0, 0};
      int np = 0;
      int maxsoft = t->nsoft; /* sentinel with zero softening (forcetree.c:1254) */
      for(int j = 0; j < 8; j++)
        {
          int p = suns[j];
          if(p < 0)
            continue;
          int jj, nextsib = sib;
          for(jj = j + 1; jj < 8; jj++)
            if(suns[jj] >= 0)
              {
                nextsib = suns[jj];
                break;
              }
          tree_update_recursive(t, p, nextsib, no, last);
          if(p < N)
            {
              double m = t->pmass[p];
              mass += m;
              for(int d = 0; d < 3; d++)
                s[d] += m * t->pos[3 * p + d];
              int st = t->stype ? t->stype[p] : 0;
              if(t->force_softening[maxsoft] < t->force_softening[st])
                maxsoft = st;
              np++;
            }
          else
            {
              int c = p - N;
              mass += t->mass[c];
              for(int d = 0; d < 3; d++)
                s[d] += t->mass[c] * t->s[3 * c + d];
              if(t->force_softening[maxsoft] < t->force_softening[t->maxsoft[c]])
                maxsoft = t->maxsoft[c];
              np += t->npart[c];
            }
        }
      if(mass)
        for(int d = 0; d < 3; d++)
          s[d] /= mass;
      else
        for(int d = 0; d < 3; d++)
          s[d] = t->center[3 * k + d];
      t->mass[k] = mass;
      for(int d = 0; d < 3; d++)
        t->s[3 * k + d] = s[d];
      t->maxsoft[k] = (uint8_t)maxsoft;
      t->sibling[k] = sib;
      t->father[k]  = father;
      t->npart[k]   = np;
    }
  else
    {
      if(*last >= 0)
        {
          if(*last >= N)
            t->nextnode[*last - N] = no;
          else
            t->pnext[*last] = no;
        }
      *last = no;
    }
}

void orc_tree_free(orc_tree *t)
{
  if(!t)
    return;
  free(t->suns);
  free(t->center);
  free(t->len);
  free(t->s);
  free(t->mass);
  free(t->sibling);
  free(t->nextnode);
  free(t->father);
  free(t->level);
  free(t->npart);
  free(t->maxsoft);
  free(t->pnext);
  free(t->intpos);
  free(t->pos);
  free(t);
}

int orc_tree_build(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype, double alloc_factor,
                   orc_tree **out)
{
  return orc_tree_build_cube(p, n, pos, mass, stype, alloc_factor, NULL, out);
}

int orc_tree_build_cube(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype,
                        double alloc_factor, const double *cube, orc_tree **out)
{
  *out = NULL;
  if(n <= 0)
    return 0;
  orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
  t->n        = n;
  t->maxnodes = (int)(alloc_factor * n) + 1;
  if(t->maxnodes < 2)
    t->maxnodes = 2;
  int M       = t->maxnodes;
  t->suns     = (int *)malloc(sizeof(int) * 8 * M);
  t->center   = (double *)malloc(sizeof(double) * 3 * M);
  t->len      = (double *)malloc(sizeof(double) * M);
  t->s        = (double *)malloc(sizeof(double) * 3 * M);
  t->mass     = (double *)malloc(sizeof(double) * M);
  t->sibling  = (int *)malloc(sizeof(int) * M);
  t->nextnode = (int *)malloc(sizeof(int) * M);
  t->father   = (int *)malloc(sizeof(int) * M);
  t->level    = (int *)malloc(sizeof(int) * M);
  t->npart    = (int *)malloc(sizeof(int) * M);
  t->maxsoft  = (uint8_t *)malloc(M);
  t->pnext    = (int *)malloc(sizeof(int) * n);
  t->intpos   = (uint64_t *)malloc(sizeof(uint64_t) * 3 * n);
  t->pos      = (double *)malloc(sizeof(double) * 3 * n);
  t->pmass    = mass;
  t->stype    = stype;
  t->nsoft    = p->nsoft;
  t->periodic = p->periodic;
  t->box      = p->box_size;
  orc_set_softenings(p, t->force_softening);

  if(cube)
    {
      /* a given root cube {centre[3], len}, as the subfind local tree keeps one per
       * group (subfind_loctree.c:169-179, :203-206) */
      t->dlen = cube[3];
      for(int j = 0; j < 3; j++)
        {
          t->dcenter[j] = cube[j];
          t->corner[j]  = cube[j] - 0.5 * cube[3];
        }
    }
  else
    orc_domain_extent(p, n, pos, t->corner, t->dcenter, &t->dlen);
  t->bigfac   = t->dlen / (double)(1ll << 52);
  double ilen = 1.0 / t->dlen;

  /* wrap into the box and form integer coordinates (forcetree.c:394-417, :493-505) */
  for(int i = 0; i < n; i++)
    for(int d = 0; d < 3; d++)
      {
        double x = pos[3 * i + d];
        if(p->periodic)
          {
            if(x < 0)
              x += p->box_size;
            if(x >= p->box_size)
              x -= p->box_size;
          }
        t->pos[3 * i + d]    = x;
        t->intpos[3 * i + d] = dbl_mantissa((x - t->corner[d]) * ilen + 1.0);
      }

  /* root node (forcetree.c:339-351) */
  int root = tree_new_node(t);
  t->len[root] = t->dlen;
  for(int d = 0; d < 3; d++)
    t->center[3 * root + d] = t->dcenter[d];
  t->level[root] = 0;

  for(int i = 0; i < n; i++)
    if(tree_insert(t, i) < 0)
      {
        orc_tree_free(t);
        return -1;
      }

  int last = -1;
  tree_update_recursive(t, n + 0, -1, -1, &last);
  if(last >= n)
    t->nextnode[last - n] = -1;
  else
    t->pnext[last] = -1;

  *out = t;
  return t->numnodes;
}

int orc_tree_numnodes(const orc_tree *t) { return t ? t->numnodes : 0; }

void orc_tree_export(const orc_tree *t, double *center, double *len, double *com, double *mass, int *level, int *npart)
{
  int M = t->numnodes;
  if(center)
    memcpy(center, t->center, sizeof(double) * 3 * M);
  if(len)
    memcpy(len, t->len, sizeof(double) * M);
  if(com)
    memcpy(com, t->s, sizeof(double) * 3 * M);
  if(mass)
    memcpy(mass, t->mass, sizeof(double) * M);
  if(level)
    memcpy(level, t->level, sizeof(int) * M);
  if(npart)
    memcpy(npart, t->npart, sizeof(int) * M);
}

/* ------------------------------------------------------------------ */
/* Ewald tables                                                        */
/* ------------------------------------------------------------------ */

double orc_ewald_psi(double x, double y, double z)
{
  /* forcetree_ewald.c:393-450 (cubic unit box) */
  double r = sqrt(x * x + y * y + z * z);
  if(r == 0)
    return 0;
  const double alpha = 3.0;
  const int nmax     = 4;
  double sum1        = 0;
  for(int nx = -nmax; nx <= nmax; nx++)
    for(int ny = -nmax; ny <= nmax; ny++)
      for(int nz = -nmax; nz <= nmax; nz++)
        {
          double dx = x - nx, dy = y - ny, dz = z - nz;
          double rr = sqrt(dx * dx + dy * dy + dz * dz);
          sum1 += erfc(alpha * rr) / rr;
        }
  double alpha2 = alpha * alpha;
  int kmax      = (int)(2 * alpha + 0.5);
  double sum2   = 0;
  for(int nx = -kmax; nx <= kmax; nx++)
    for(int ny = -kmax; ny <= kmax; ny++)
      for(int nz = -kmax; nz <= kmax; nz++)
        {
          double kx = 2.0 * M_PI * nx, ky = 2.0 * M_PI * ny, kz = 2.0 * M_PI * nz;
          double k2 = kx * kx + ky * ky + kz * kz;
          if(k2 > 0)
            sum2 += 4.0 * M_PI / k2 * exp(-k2 / (4.0 * alpha2)) * cos(x * kx + y * ky + z * kz);
        }
  return M_PI / (alpha * alpha) - sum1 - sum2 + 1.0 / r;
}

void orc_ewald_force(double x, double y, double z, double force[3])
{
  /* forcetree_ewald.c:463-527 (cubic unit box) */
  force[0] = force[1] = force[2] = 0;
  double r2 = x * x + y * y + z * z;
  if(r2 == 0)
    return;
  const double alpha  = 2.0;
  const double alpha2 = alpha * alpha;
  double r3inv        = 1.0 / (r2 * sqrt(r2));
  force[0] += r3inv * x;
  force[1] += r3inv * y;
  force[2] += r3inv * z;
  const int nmax = 4;
  for(int nx = -nmax; nx <= nmax; nx++)
    for(int ny = -nmax; ny <= nmax; ny++)
      for(int nz = -nmax; nz <= nmax; nz++)
        {
          double dx = x - nx, dy = y - ny, dz = z - nz;
          double rr2 = dx * dx + dy * dy + dz * dz;
          double rr  = sqrt(rr2);
          double val = erfc(alpha * rr) + 2.0 * alpha * rr / sqrt(M_PI) * exp(-alpha2 * rr2);
          double v2  = val / (rr2 * rr);
          force[0] -= dx * v2;
          force[1] -= dy * v2;
          force[2] -= dz * v2;
        }
  int hmax = (int)(2 * alpha + 0.5);
  for(int hx = -hmax; hx <= hmax; hx++)
    for(int hy = -hmax; hy <= hmax; hy++)
      for(int hz = -hmax; hz <= hmax; hz++)
        {
          double h2 = hx * hx + hy * hy + hz * hz;
          if(h2 > 0)
            {
              double hdotx = x * hx + y * hy + z * hz;
              double val   = 2.0 / h2 * exp(-M_PI * M_PI * h2 / alpha2) * sin(2.0 * M_PI * hdotx);
              force[0] -= hx * val;
              force[1] -= hy * val;
              force[2] -= hz * val;
            }
        }
}

void orc_ewald_table_tree(int en, double box, double *tab, int nthreads)
{
  /* forcetree_ewald.c:158-175 (table on one octant) and :239-249 (scaling) */
  int n1   = en + 1;
  int size = n1 * n1 * n1;
  (void)omp_threads;
#pragma omp parallel for schedule(dynamic, 64) num_threads(omp_threads(nthreads))
  for(int n = 0; n < size; n++)
    {
      int i = n / (n1 * n1), j = (n / n1) % n1, k = n % n1;
      double xx = 0.5 * i / en, yy = 0.5 * j / en, zz = 0.5 * k / en;
      double f[3];
      orc_ewald_force(xx, yy, zz, f);
      tab[0 * size + n] = f[0] / (box * box);
      tab[1 * size + n] = f[1] / (box * box);
      tab[2 * size + n] = f[2] / (box * box);
      tab[3 * size + n] = orc_ewald_psi(xx, yy, zz) / box;
    }
}

static void trilinear_setup(double fac_intp, int en, double dx, double dy, double dz, int *i, int *j, int *k, double w[8])
{
  double u = dx * fac_intp;
  *i       = (int)u;
  if(*i >= en)
    *i = en - 1;
  u -= *i;
  double v = dy * fac_intp;
  *j       = (int)v;
  if(*j >= en)
    *j = en - 1;
  v -= *j;
  double ww = dz * fac_intp;
  *k        = (int)ww;
  if(*k >= en)
    *k = en - 1;
  ww -= *k;
  w[0] = (1 - u) * (1 - v) * (1 - ww);
  w[1] = (1 - u) * (1 - v) * (ww);
  w[2] = (1 - u) * (v) * (1 - ww);
  w[3] = (1 - u) * (v) * (ww);
  w[4] = (u) * (1 - v) * (1 - ww);
  w[5] = (u) * (1 - v) * (ww);
  w[6] = (u) * (v) * (1 - ww);
  w[7] = (u) * (v) * (ww);
}

static double trilinear(const double *t, int n1, int i, int j, int k, const double w[8])
{
#define T(a, b, c) t[((a)*n1 + (b)) * n1 + (c)]
  return T(i, j, k) * w[0] + T(i, j, k + 1) * w[1] + T(i, j + 1, k) * w[2] + T(i, j + 1, k + 1) * w[3] + T(i + 1, j, k) * w[4] +
         T(i + 1, j, k + 1) * w[5] + T(i + 1, j + 1, k) * w[6] + T(i + 1, j + 1, k + 1) * w[7];
#undef T
}

void orc_ewald_corr(const double *tab, int en, double box, double dx, double dy, double dz, double fcorr[3])
{
  /* forcetree_ewald.c:268-328 */
  double sx = -1, sy = -1, sz = -1;
  if(dx < 0)
    {
      dx = -dx;
      sx = +1;
    }
  if(dy < 0)
    {
      dy = -dy;
      sy = +1;
    }
  if(dz < 0)
    {
      dz = -dz;
      sz = +1;
    }
  int i, j, k, n1 = en + 1, size = n1 * n1 * n1;
  double w[8];
  trilinear_setup(2.0 * en / box, en, dx, dy, dz, &i, &j, &k, w);
  fcorr[0] = sx * trilinear(tab + 0 * size, n1, i, j, k, w);
  fcorr[1] = sy * trilinear(tab + 1 * size, n1, i, j, k, w);
  fcorr[2] = sz * trilinear(tab + 2 * size, n1, i, j, k, w);
}

double orc_ewald_pot_corr(const double *tab, int en, double box, double dx, double dy, double dz)
{
  /* forcetree_ewald.c:344-382 */
  int i, j, k, n1 = en + 1, size = n1 * n1 * n1;
  double w[8];
  trilinear_setup(2.0 * en / box, en, fabs(dx), fabs(dy), fabs(dz), &i, &j, &k, w);
  return trilinear(tab + 3 * size, n1, i, j, k, w);
}

void orc_forcetest_ewald_exact(double box, double x, double y, double z, double force[4])
{
  /* gravtree_forcetest.c:820-916 */
  for(int i = 0; i < 4; i++)
    force[i] = 0;
  double r2 = x * x + y * y + z * z;
  if(r2 == 0)
    return;
  double sx = -1, sy = -1, sz = -1;
  if(x < 0)
    {
      x  = -x;
      sx = +1;
    }
  if(y < 0)
    {
      y  = -y;
      sy = +1;
    }
  if(z < 0)
    {
      z  = -z;
      sz = +1;
    }
  double alpha  = 2.0 / box;
  double alpha2 = alpha * alpha;
  double r      = sqrt(r2);
  double r3inv  = 1.0 / (r2 * r);
  force[0] += r3inv * x;
  force[1] += r3inv * y;
  force[2] += r3inv * z;
  const int nmax = 6;
  for(int nx = -nmax; nx <= nmax; nx++)
    for(int ny = -nmax; ny <= nmax; ny++)
      for(int nz = -nmax; nz <= nmax; nz++)
        {
          double dx = x - nx * box, dy = y - ny * box, dz = z - nz * box;
          double rr2 = dx * dx + dy * dy + dz * dz;
          double rr  = sqrt(rr2);
          double ea  = erfc(alpha * rr);
          double val = ea + 2.0 * alpha * rr / sqrt(M_PI) * exp(-alpha2 * rr2);
          double v2  = val / (rr2 * rr);
          force[0] -= dx * v2;
          force[1] -= dy * v2;
          force[2] -= dz * v2;
          force[3] -= ea / rr;
        }
  int kmax = (int)(4 * alpha * box + 0.5);
  double vol = box * box * box;
  for(int nx = -kmax; nx <= kmax; nx++)
    for(int ny = -kmax; ny <= kmax; ny++)
      for(int nz = -kmax; nz <= kmax; nz++)
        {
          double kx = (2.0 * M_PI / box) * nx, ky = (2.0 * M_PI / box) * ny, kz = (2.0 * M_PI / box) * nz;
          double k2 = kx * kx + ky * ky + kz * kz;
          if(k2 > 0)
            {
              double kdotx = x * kx + y * ky + z * kz;
              double vv    = 4.0 * M_PI / (k2 * vol) * exp(-k2 / (4.0 * alpha2));
              double val   = vv * sin(kdotx);
              force[0] -= kx * val;
              force[1] -= ky * val;
              force[2] -= kz * val;
              force[3] -= vv * cos(kdotx);
            }
        }
  force[3] += M_PI / (alpha2 * vol) + 1.0 / r;
  force[0] *= sx;
  force[1] *= sy;
  force[2] *= sz;
}

void orc_forcetest_ewald_table(int tew_n, double box, double *tab, int nthreads)
{
  /* gravtree_forcetest.c:933-1008 */
  int n1   = tew_n + 1;
  int size = n1 * n1 * n1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(omp_threads(nthreads))
  for(int n = 0; n < size; n++)
    {
      int i = n / (n1 * n1), j = (n / n1) % n1, k = n % n1;
      double f[4];
      orc_forcetest_ewald_exact(box, 0.5 * i / tew_n * box, 0.5 * j / tew_n * box, 0.5 * k / tew_n * box, f);
      for(int c = 0; c < 4; c++)
        tab[c * size + n] = f[c];
    }
}

void orc_forcetest_ewald_lookup(const double *tab, int tew_n, double box, double dx, double dy, double dz, double force[4])
{
  /* gravtree_forcetest.c:1019-1085 */
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
  int i, j, k, n1 = tew_n + 1, size = n1 * n1 * n1;
  double w[8];
  trilinear_setup(2.0 * tew_n / box, tew_n, dx, dy, dz, &i, &j, &k, w);
  for(int c = 0; c < 4; c++)
    force[c] = trilinear(tab + c * size, n1, i, j, k, w);
  force[0] *= sx;
  force[1] *= sy;
  force[2] *= sz;
}

/* ------------------------------------------------------------------ */
/* tree walk                                                           */
/* ------------------------------------------------------------------ */

typedef struct
{
  int pm;
  double rcut2, rcut, asmthinv2, asmthfac;
  float tf[ORC_NTAB + 1], tp[ORC_NTAB + 1];
} shortrange_ctx;

/* the short-range constants of forcetree_walk.c:142-156 for split scale asmth and cut rcut (lengths) */
static void shortrange_setup_len(const orc_params *p, double asmth, double rcut, shortrange_ctx *c)
{
  c->pm        = 1;
  c->rcut      = rcut;
  c->rcut2     = rcut * rcut;
  double ai    = 0.5 / asmth;
  c->asmthinv2 = ai * ai;
  c->asmthfac  = ai * (ORC_NTAB / (p->rcut / 2.0));
  orc_short_range_tables(p->rcut, c->tf, c->tp);
}

static void shortrange_setup(const orc_params *p, shortrange_ctx *c)
{
  c->pm = p->pmgrid > 0;
  if(!c->pm)
    return;
  double asmth, rcut;
  orc_asmth_rcut(p, &asmth, &rcut);
  c->rcut      = rcut;
  c->rcut2     = rcut * rcut;
  double ai    = 0.5 / asmth;
  c->asmthinv2 = ai * ai;
  c->asmthfac  = ai * (ORC_NTAB / (p->rcut / 2.0));
  orc_short_range_tables(p->rcut, c->tf, c->tp);
}

/* pair kernel of forcetree_walk.c:364-487; returns 0 when beyond the table (no interaction) */
static inline int pair_kernel(const shortrange_ctx *c, double r, double hmax, double *fac_out, double *wp_out)
{
  double ff = 0, fp = 0;
  if(c->pm)
    {
      double tabentry = c->asmthfac * r;
      int ti          = (int)tabentry;
      if(ti >= ORC_NTAB)
        return 0;
      double w = tabentry - ti;
      ff       = (1.0 - w) * c->tf[ti] + w * c->tf[ti + 1];
      fp       = (1.0 - w) * c->tp[ti] + w * c->tp[ti + 1];
    }
  double fac, wp;
  if(r >= hmax)
    {
      double rinv  = 1.0 / r;
      double rinv3 = rinv * rinv * rinv;
      if(c->pm)
        {
          fac = rinv3 + rinv * ff * c->asmthinv2;
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
      double h_inv  = 1.0 / hmax;
      double h3_inv = h_inv * h_inv * h_inv;
      double u      = r * h_inv;
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
      if(c->pm && r > 0)
        {
          double rinv = 1.0 / r;
          fac += rinv * ff * c->asmthinv2;
          wp += rinv * (fp + 1.0);
        }
    }
  *fac_out = fac;
  *wp_out  = wp;
  return 1;
}

/* force_treeevaluate, forcetree_walk.c:115-514, mode 0, one rank */
static int walk_one(const orc_params *p, const orc_tree *t, const shortrange_ctx *c, int use_bh, double aold_abs,
                    const double *ewald_tab, const double pos[3], int stype, double out_acc[3], double *out_pot)
{
  const int N    = t->n;
  const int per  = p->periodic;
  const double L = p->box_size;
  double ax = 0, ay = 0, az = 0, pot = 0;
  int nint      = 0;
  double aold   = p->err_tol_force_acc * aold_abs;
  double h_i    = t->force_softening[stype];
  double theta2 = p->err_tol_theta * p->err_tol_theta;
  int no        = N + 0;

  while(no >= 0)
    {
      double dx, dy, dz, r2, mass, hmax;
      if(no < N)
        {
          dx   = nearest(t->pos[3 * no + 0] - pos[0], L, per);
          dy   = nearest(t->pos[3 * no + 1] - pos[1], L, per);
          dz   = nearest(t->pos[3 * no + 2] - pos[2], L, per);
          r2   = dx * dx + dy * dy + dz * dz;
          mass = t->pmass[no];
          double h_j = t->force_softening[t->stype ? t->stype[no] : 0];
          hmax       = h_j > h_i ? h_j : h_i;
          no         = t->pnext[no];
        }
      else
        {
          int k     = no - N;
          mass      = t->mass[k];
          dx        = nearest(t->s[3 * k + 0] - pos[0], L, per);
          dy        = nearest(t->s[3 * k + 1] - pos[1], L, per);
          dz        = nearest(t->s[3 * k + 2] - pos[2], L, per);
          r2        = dx * dx + dy * dy + dz * dz;
          double ln = t->len[k];
          if(c->pm && r2 > c->rcut2)
            {
              double eff = c->rcut + 0.5 * ln;
              int skip   = 0;
              for(int d = 0; d < 3 && !skip; d++)
                {
                  double dist = nearest(t->center[3 * k + d] - pos[d], L, per);
                  if(dist < -eff || dist > eff)
                    skip = 1;
                }
              if(skip)
                {
                  no = t->sibling[k];
                  continue;
                }
            }
          if(use_bh)
            {
              if(ln * ln > r2 * theta2)
                {
                  no = t->nextnode[k];
                  continue;
                }
            }
          else
            {
              double len2 = ln * ln;
              if(len2 > r2 * (1.2 * 1.2))
                {
                  no = t->nextnode[k];
                  continue;
                }
              if(mass * len2 > r2 * r2 * aold)
                {
                  no = t->nextnode[k];
                  continue;
                }
              if(fabs(nearest(t->center[3 * k + 0] - pos[0], L, per)) < 0.60 * ln &&
                 fabs(nearest(t->center[3 * k + 1] - pos[1], L, per)) < 0.60 * ln &&
                 fabs(nearest(t->center[3 * k + 2] - pos[2], L, per)) < 0.60 * ln)
                {
                  no = t->nextnode[k];
                  continue;
                }
            }
          double h_j = t->force_softening[t->maxsoft[k]];
          if(h_j > h_i)
            {
              if(r2 < h_j * h_j)
                {
                  no = t->nextnode[k];
                  continue;
                }
              hmax = h_j;
            }
          else
            hmax = h_i;
          no = t->sibling[k];
        }

      if(mass)
        {
          double r = sqrt(r2), fac, wp;
          if(!pair_kernel(c, r, hmax, &fac, &wp))
            continue;
          pot += mass * wp;
          fac *= mass;
          ax += dx * fac;
          ay += dy * fac;
          az += dz * fac;
          if(ewald_tab)
            {
              double fc[3];
              orc_ewald_corr(ewald_tab, 64, L, dx, dy, dz, fc);
              ax += mass * fc[0];
              ay += mass * fc[1];
              az += mass * fc[2];
              pot += mass * orc_ewald_pot_corr(ewald_tab, 64, L, dx, dy, dz);
            }
          nint++;
        }
    }
  out_acc[0] = ax;
  out_acc[1] = ay;
  out_acc[2] = az;
  *out_pot   = pot;
  return nint;
}

long long orc_tree_walk(const orc_params *p, const orc_tree *t, int use_bh, const float *old_acc, const double *ewald_tab,
                        int ntarget, const int *targets, double *acc, double *pot, int *nint, int nthreads)
{
  return orc_tree_walk_hr(p, t, use_bh, old_acc, ewald_tab, NULL, NULL, ntarget, targets, acc, pot, nint, nthreads);
}

long long orc_tree_walk_hr(const orc_params *p, const orc_tree *t, int use_bh, const float *old_acc, const double *ewald_tab,
                           const orc_highres *hr, const uint8_t *hrflag, int ntarget, const int *targets, double *acc,
                           double *pot, int *nint, int nthreads)
{
  shortrange_ctx c, c1;
  shortrange_setup(p, &c);
  /* forcetree_walk.c:143-150: a target in the high-resolution region uses Asmth[1] and Rcut[1] */
  if(hr && hrflag)
    shortrange_setup_len(p, hr->asmth, hr->rcut, &c1);
  long long total = 0;
  int nt          = targets ? ntarget : t->n;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) num_threads(omp_threads(nthreads))
  for(int q = 0; q < nt; q++)
    {
      int i = targets ? targets[q] : q;
      double a[3], pt;
      double aold = old_acc ? old_acc[i] : 0;
      const shortrange_ctx *ci = (hr && hrflag && hrflag[i]) ? &c1 : &c;
      int ni = walk_one(p, t, ci, use_bh, aold, ewald_tab, &t->pos[3 * i], t->stype ? t->stype[i] : 0, a, &pt);
      acc[3 * q + 0] = a[0];
      acc[3 * q + 1] = a[1];
      acc[3 * q + 2] = a[2];
      pot[q]         = pt;
      if(nint)
        nint[q] = ni;
      total += ni;
    }
  return total;
}

/* ------------------------------------------------------------------ */
/* FFT and periodic PM                                                 */
/* ------------------------------------------------------------------ */

static void fft1d(int n, double *re, double *im, int stride, int sign)
{
  /* iterative radix-2 Cooley-Tukey, unnormalised, exp(sign*2*pi*i*j*k/n) */
  for(int i = 1, j = 0; i < n; i++)
    {
      int bit = n >> 1;
      for(; j & bit; bit >>= 1)
        j ^= bit;
      j ^= bit;
      if(i < j)
        {
          double tr = re[i * stride], ti = im[i * stride];
          re[i * stride] = re[j * stride];
          im[i * stride] = im[j * stride];
          re[j * stride] = tr;
          im[j * stride] = ti;
        }
    }
  for(int len = 2; len <= n; len <<= 1)
    {
      double ang = sign * 2.0 * M_PI / len;
      for(int i = 0; i < n; i += len)
        for(int k = 0; k < len / 2; k++)
          {
            double wr = cos(ang * k), wi = sin(ang * k);
            int a = (i + k) * stride, b = (i + k + len / 2) * stride;
            double xr = re[b] * wr - im[b] * wi;
            double xi = re[b] * wi + im[b] * wr;
            re[b]     = re[a] - xr;
            im[b]     = im[a] - xi;
            re[a] += xr;
            im[a] += xi;
          }
    }
}

void orc_fft3d(int n, double *re, double *im, int sign)
{
  size_t n2 = (size_t)n * n;
  for(size_t a = 0; a < n2; a++) /* along z (contiguous) */
    fft1d(n, re + a * n, im + a * n, 1, sign);
  for(int x = 0; x < n; x++) /* along y */
    for(int z = 0; z < n; z++)
      fft1d(n, re + (size_t)x * n2 + z, im + (size_t)x * n2 + z, n, sign);
  for(size_t yz = 0; yz < n2; yz++) /* along x */
    fft1d(n, re + yz, im + yz, (int)n2, sign);
}

int orc_pm_periodic(const orc_params *p, int n, const double *pos, const double *mass, double *acc_pm, double *pot_pm)
{
  const int G    = p->pmgrid;
  const double L = p->box_size;
  if(G <= 0 || (G & (G - 1)))
    return -1;
  size_t G3  = (size_t)G * G * G;
  double *re = (double *)calloc(G3, sizeof(double));
  double *im = (double *)calloc(G3, sizeof(double));
  double *fg = (double *)malloc(G3 * sizeof(double));
  if(!re || !im || !fg)
    return -2;
#define IDX(x, y, z) (((size_t)(x)*G + (y)) * G + (z))
  double to_slab = G / L;

  /* CIC deposit, pm_periodic.c:1049-1170 */
  for(int i = 0; i < n; i++)
    {
      double x[3];
      int s[3], ss[3];
      double d[3];
      for(int k = 0; k < 3; k++)
        {
          x[k] = pos[3 * i + k];
          if(x[k] < 0)
            x[k] += L;
          if(x[k] >= L)
            x[k] -= L;
          s[k] = (int)(to_slab * x[k]);
          d[k] = to_slab * x[k] - s[k];
          if(s[k] >= G)
            s[k] -= G;
          ss[k] = s[k] + 1;
          if(ss[k] >= G)
            ss[k] -= G;
        }
      double m = mass[i];
      re[IDX(s[0], s[1], s[2])] += m * (1.0 - d[0]) * (1.0 - d[1]) * (1.0 - d[2]);
      re[IDX(s[0], s[1], ss[2])] += m * (1.0 - d[0]) * (1.0 - d[1]) * (d[2]);
      re[IDX(s[0], ss[1], s[2])] += m * (1.0 - d[0]) * (d[1]) * (1.0 - d[2]);
      re[IDX(s[0], ss[1], ss[2])] += m * (1.0 - d[0]) * (d[1]) * (d[2]);
      re[IDX(ss[0], s[1], s[2])] += m * (d[0]) * (1.0 - d[1]) * (1.0 - d[2]);
      re[IDX(ss[0], s[1], ss[2])] += m * (d[0]) * (1.0 - d[1]) * (d[2]);
      re[IDX(ss[0], ss[1], s[2])] += m * (d[0]) * (d[1]) * (1.0 - d[2]);
      re[IDX(ss[0], ss[1], ss[2])] += m * (d[0]) * (d[1]) * (d[2]);
    }

  orc_fft3d(G, re, im, -1);

  /* Green's function and CIC deconvolution, pm_periodic.c:1626-1707 */
  double asmth, rcut;
  orc_asmth_rcut(p, &asmth, &rcut);
  double asmth2 = asmth * asmth;
  double dcell  = L / G;
  double dhalf  = 0.5 * dcell;
  double kfac   = 2.0 * M_PI / (G * dcell);
  for(int x = 0; x < G; x++)
    for(int y = 0; y < G; y++)
      for(int z = 0; z < G; z++)
        {
          int xx = x >= G / 2 ? x - G : x, yy = y >= G / 2 ? y - G : y, zz = z >= G / 2 ? z - G : z;
          double kx = kfac * xx, ky = kfac * yy, kz = kfac * zz;
          double k2 = kx * kx + ky * ky + kz * kz;
          size_t ip = IDX(x, y, z);
          if(k2 > 0)
            {
              double smth = -exp(-k2 * asmth2) / k2;
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
              smth *= ff * ff * ff * ff;
              re[ip] *= smth;
              im[ip] *= smth;
            }
          else
            re[ip] = im[ip] = 0;
        }

  orc_fft3d(G, re, im, +1); /* re now holds the (unnormalised) potential */

  double fac_pot = 4 * M_PI * p->G / (L * L * L);
  double fac_fd  = fac_pot / (2 * dcell);

  /* readout helper, pm_periodic.c:1322-1550 */
  for(int i = 0; i < n; i++)
    {
      acc_pm[3 * i + 0] = acc_pm[3 * i + 1] = acc_pm[3 * i + 2] = 0;
      pot_pm[i]                                                = 0;
    }
  for(int dim = -1; dim < 3; dim++)
    {
      const double *grid = re;
      if(dim >= 0)
        {
          /* 4-point finite difference, pm_periodic.c:1731-1800 */
          for(int x = 0; x < G; x++)
            for(int y = 0; y < G; y++)
              for(int z = 0; z < G; z++)
                {
                  int c[3] = {x, y, z}, l[3], r[3], ll[3], rr[3];
                  for(int k = 0; k < 3; k++)
                    l[k] = r[k] = ll[k] = rr[k] = c[k];
                  l[dim]  = (c[dim] - 1 + G) % G;
                  r[dim]  = (c[dim] + 1) % G;
                  ll[dim] = (c[dim] - 2 + G) % G;
                  rr[dim] = (c[dim] + 2) % G;
                  fg[IDX(x, y, z)] = fac_fd * ((4.0 / 3) * (re[IDX(l[0], l[1], l[2])] - re[IDX(r[0], r[1], r[2])]) -
                                               (1.0 / 6) * (re[IDX(ll[0], ll[1], ll[2])] - re[IDX(rr[0], rr[1], rr[2])]));
                }
          grid = fg;
        }
      for(int i = 0; i < n; i++)
        {
          double x[3];
          int s[3], ss[3];
          double d[3];
          for(int k = 0; k < 3; k++)
            {
              x[k] = pos[3 * i + k];
              if(x[k] < 0)
                x[k] += L;
              if(x[k] >= L)
                x[k] -= L;
              s[k] = (int)(to_slab * x[k]);
              d[k] = to_slab * x[k] - s[k];
              if(s[k] >= G)
                s[k] -= G;
              ss[k] = s[k] + 1;
              if(ss[k] >= G)
                ss[k] -= G;
            }
          double v = grid[IDX(s[0], s[1], s[2])] * (1.0 - d[0]) * (1.0 - d[1]) * (1.0 - d[2]) +
                     grid[IDX(s[0], s[1], ss[2])] * (1.0 - d[0]) * (1.0 - d[1]) * (d[2]) +
                     grid[IDX(s[0], ss[1], s[2])] * (1.0 - d[0]) * (d[1]) * (1.0 - d[2]) +
                     grid[IDX(s[0], ss[1], ss[2])] * (1.0 - d[0]) * (d[1]) * (d[2]) +
                     grid[IDX(ss[0], s[1], s[2])] * (d[0]) * (1.0 - d[1]) * (1.0 - d[2]) +
                     grid[IDX(ss[0], s[1], ss[2])] * (d[0]) * (1.0 - d[1]) * (d[2]) +
                     grid[IDX(ss[0], ss[1], s[2])] * (d[0]) * (d[1]) * (1.0 - d[2]) +
                     grid[IDX(ss[0], ss[1], ss[2])] * (d[0]) * (d[1]) * (d[2]);
          if(dim < 0)
            pot_pm[i] += v * fac_pot;
          else
            acc_pm[3 * i + dim] += v;
        }
    }
#undef IDX
  free(re);
  free(im);
  free(fg);
  return 0;
}

/* ------------------------------------------------------------------ */
/* high-resolution PM region (PLACEHIGHRESREGION), pm_nonperiodic.c    */
/* ------------------------------------------------------------------ */

static double wrap_box(const orc_params *p, double x)
{
  if(!p->periodic)
    return x;
  if(x < 0)
    x += p->box_size;
  if(x >= p->box_size)
    x -= p->box_size;
  return x;
}

int orc_highres_regionsize(const orc_params *p, orc_highres *r, int n, const double *pos, const uint8_t *type)
{
  /* pm_init_regionsize, pm_nonperiodic.c:135-245, for the high-resolution grid */
  const int GRID = ORC_GRIDBOOST * p->pmgrid;
  double xmin[3], xmax[3];
  for(int j = 0; j < 3; j++)
    {
      xmin[j] = 1.0e36;
      xmax[j] = -1.0e36;
    }
  int found = 0;
  for(int i = 0; i < n; i++)
    if((1 << type[i]) & r->mask)
      {
        found = 1;
        for(int j = 0; j < 3; j++)
          {
            double x = wrap_box(p, pos[3 * i + j]);
            if(x > xmax[j])
              xmax[j] = x;
            if(x < xmin[j])
              xmin[j] = x;
          }
      }
  if(!found)
    return -1;
  double size = xmax[0] - xmin[0];
  for(int j = 1; j < 3; j++)
    if(xmax[j] - xmin[j] > size)
      size = xmax[j] - xmin[j];
  if(r->enlarge > 0)
    size *= r->enlarge; /* ENLARGEREGION */
  for(int j = 0; j < 3; j++)
    {
      /* symmetrised onto the centre */
      r->xmin[j] = (xmin[j] + xmax[j]) / 2 - size / 2;
      r->xmax[j] = r->xmin[j] + size;
    }
  const double meshinner = size;
  double tms             = size * 2.001 * GRID / ((double)(GRID - 2 - 8));
  for(int j = 0; j < 3; j++)
    {
      r->corner[j] = r->xmin[j] - 2.0005 * tms / GRID;
      r->upper[j]  = r->corner[j] + (GRID / 2 - 1) * (tms / GRID);
    }
  double asmth0, rcut0;
  orc_asmth_rcut(p, &asmth0, &rcut0);
  if(2 * tms / GRID < rcut0)
    {
      /* a high-resolution cell below half the coarse cut: the region is widened by the coarse cut */
      tms = 2 * (meshinner + 2 * rcut0) * GRID / ((double)(GRID - 2));
      for(int j = 0; j < 3; j++)
        {
          r->corner[j] = r->xmin[j] - 1.0001 * rcut0;
          r->upper[j]  = r->corner[j] + (GRID / 2 - 1) * (tms / GRID);
        }
      if(2 * tms / GRID > rcut0)
        {
          tms = 2 * (meshinner + 2 * rcut0) * GRID / ((double)(GRID - 10));
          for(int j = 0; j < 3; j++)
            {
              r->corner[j] = r->xmin[j] - 1.0001 * (rcut0 + 2 * tms / GRID);
              r->upper[j]  = r->corner[j] + (GRID / 2 - 1) * (tms / GRID);
            }
        }
    }
  r->meshsize = tms;
  r->asmth    = p->asmth * tms / GRID;
  r->rcut     = p->rcut * r->asmth;
  r->valid    = 1;
  return 0;
}

int orc_highres_flag(const orc_params *p, const orc_highres *r, int type, const double *pos)
{
  /* pmforce_is_particle_high_res, pm_nonperiodic.c:344-371 */
  if((1 << type) & r->mask)
    return 1;
  for(int j = 0; j < 3; j++)
    {
      double x = wrap_box(p, pos[j]);
      if(x < r->xmin[j] || x > r->xmax[j])
        return 0;
    }
  return 1;
}

void orc_highres_flags(const orc_params *p, const orc_highres *r, int n, const double *pos, const uint8_t *type, uint8_t *flags)
{
  for(int i = 0; i < n; i++)
    flags[i] = (uint8_t)orc_highres_flag(p, r, type[i], &pos[3 * i]);
}

int orc_pm_highres(const orc_params *p, const orc_highres *r, int n, const double *pos, const double *mass, const uint8_t *type,
                   double *acc_pm, double *pot_pm)
{
  /* pmforce_nonperiodic(1), pm_nonperiodic.c:1508-1775, with the kernel of pm_setup_nonperiodic_kernel
   * (:1778-1987); accumulates into acc_pm (G included) and pot_pm for the high-resolution particles.
   * Returns 1 when a high-resolution particle lies outside the region (the caller re-sizes it). */
  const int GRID  = ORC_GRIDBOOST * p->pmgrid;
  const int H     = GRID / 2;
  const double fs = GRID / r->meshsize;  /* to_slab_fac */
  for(int i = 0; i < n; i++)
    if(orc_highres_flag(p, r, type[i], &pos[3 * i]))
      for(int j = 0; j < 3; j++)
        {
          double x = wrap_box(p, pos[3 * i + j]);
          if(x < r->xmin[j] || x > r->xmax[j])
            return 1;
        }
  size_t G3   = (size_t)GRID * GRID * GRID;
  double *re  = (double *)calloc(G3, sizeof(double));
  double *im  = (double *)calloc(G3, sizeof(double));
  double *kre = (double *)calloc(G3, sizeof(double));
  double *kim = (double *)calloc(G3, sizeof(double));
  double *fg  = (double *)calloc(G3, sizeof(double));
  if(!re || !im || !kre || !kim || !fg)
    return -2;
#define IDX(x, y, z) (((size_t)(x)*GRID + (y)) * GRID + (z))
  /* CIC deposit of every particle inside [Corner, UpperCorner) (:810-1170), no wrapping */
  for(int i = 0; i < n; i++)
    {
      double x[3], d[3];
      int s[3], in = 1;
      for(int k = 0; k < 3 && in; k++)
        {
          x[k] = wrap_box(p, pos[3 * i + k]);
          if(x[k] < r->corner[k] || x[k] >= r->upper[k])
            in = 0;
        }
      if(!in)
        continue;
      for(int k = 0; k < 3; k++)
        {
          s[k] = (int)(fs * (x[k] - r->corner[k]));
          d[k] = fs * (x[k] - r->corner[k]) - s[k];
        }
      double m = mass[i];
      for(int c = 0; c < 8; c++)
        {
          int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
          re[IDX(s[0] + bx, s[1] + by, s[2] + bz)] += m * (bx ? d[0] : 1.0 - d[0]) * (by ? d[1] : 1.0 - d[1]) * (bz ? d[2] : 1.0 - d[2]);
        }
    }
  orc_fft3d(GRID, re, im, -1);
  /* the kernel in real space on the whole grid (coordinates in units of the mesh size, wrapped to
   * [-1/2, 1/2)): -(erfc(u Asmth[1] / Asmth[0]) - erfc(u)) / r with u = r / (2 ASMTH / GRID) */
  double asmth0, rcut0;
  orc_asmth_rcut(p, &asmth0, &rcut0);
  const double ratio = r->asmth / asmth0;
  for(int x = 0; x < GRID; x++)
    for(int y = 0; y < GRID; y++)
      for(int z = 0; z < GRID; z++)
        {
          double xx = (double)x / GRID, yy = (double)y / GRID, zz = (double)z / GRID;
          if(xx >= 0.5)
            xx -= 1.0;
          if(yy >= 0.5)
            yy -= 1.0;
          if(zz >= 0.5)
            zz -= 1.0;
          double rr = sqrt(xx * xx + yy * yy + zz * zz);
          double u  = 0.5 * rr / (p->asmth / GRID);
          if(rr > 0)
            kre[IDX(x, y, z)] = -(erfc(u * ratio) - erfc(u)) / rr;
          else
            kre[IDX(x, y, z)] = -(1 - ratio) / (sqrt(M_PI) * (p->asmth / GRID));
        }
  orc_fft3d(GRID, kre, kim, -1);
  for(int x = 0; x < GRID; x++)
    for(int y = 0; y < GRID; y++)
      for(int z = 0; z < GRID; z++)
        {
          /* CIC deconvolution, twice (:1936-1981); frequencies above GRID / 2 are negative */
          double kx = x > GRID / 2 ? x - GRID : x, ky = y > GRID / 2 ? y - GRID : y, kz = z > GRID / 2 ? z - GRID : z;
          size_t ip = IDX(x, y, z);
          if(kx * kx + ky * ky + kz * kz > 0)
            {
              double fx = 1, fy = 1, fz = 1;
              if(kx != 0)
                fx = sin(M_PI * kx / GRID) / (M_PI * kx / GRID);
              if(ky != 0)
                fy = sin(M_PI * ky / GRID) / (M_PI * ky / GRID);
              if(kz != 0)
                fz = sin(M_PI * kz / GRID) / (M_PI * kz / GRID);
              double ff = 1 / (fx * fy * fz);
              ff        = ff * ff * ff * ff;
              kre[ip] *= ff;
              kim[ip] *= ff;
            }
          double a = re[ip], b = im[ip];
          re[ip] = a * kre[ip] - b * kim[ip];
          im[ip] = a * kim[ip] + b * kre[ip];
        }
  orc_fft3d(GRID, re, im, +1); /* re: the potential (unnormalised) */
  const double cell    = r->meshsize / GRID;
  const double fac_pot = p->G / pow(r->meshsize, 4) * pow(cell, 3);
  const double fac_fd  = fac_pot / (2 * cell);
  for(int dim = -1; dim < 3; dim++)
    {
      const double *grid = re;
      if(dim >= 0)
        {
          /* 4-point differences on the cells 2 .. GRID/2 - 3 of each axis (:1652-1700) */
          memset(fg, 0, G3 * sizeof(double));
          for(int x = 2; x < H - 2; x++)
            for(int y = 2; y < H - 2; y++)
              for(int z = 2; z < H - 2; z++)
                {
                  int l[3] = {x, y, z}, rt[3] = {x, y, z}, ll[3] = {x, y, z}, rr[3] = {x, y, z};
                  l[dim] -= 1;
                  rt[dim] += 1;
                  ll[dim] -= 2;
                  rr[dim] += 2;
                  fg[IDX(x, y, z)] = fac_fd * ((4.0 / 3) * (re[IDX(l[0], l[1], l[2])] - re[IDX(rt[0], rt[1], rt[2])]) -
                                               (1.0 / 6) * (re[IDX(ll[0], ll[1], ll[2])] - re[IDX(rr[0], rr[1], rr[2])]));
                }
          grid = fg;
        }
      /* CIC readout for the high-resolution particles (:1283-1500) */
      for(int i = 0; i < n; i++)
        {
          if(!orc_highres_flag(p, r, type[i], &pos[3 * i]))
            continue;
          double x[3], d[3];
          int s[3];
          for(int k = 0; k < 3; k++)
            {
              x[k] = wrap_box(p, pos[3 * i + k]);
              s[k] = (int)(fs * (x[k] - r->corner[k]));
              d[k] = fs * (x[k] - r->corner[k]) - s[k];
            }
          double v = 0;
          for(int c = 0; c < 8; c++)
            {
              int bx = (c >> 2) & 1, by = (c >> 1) & 1, bz = c & 1;
              v += grid[IDX(s[0] + bx, s[1] + by, s[2] + bz)] * (bx ? d[0] : 1.0 - d[0]) * (by ? d[1] : 1.0 - d[1]) *
                   (bz ? d[2] : 1.0 - d[2]);
            }
          if(dim < 0)
            pot_pm[i] += v * fac_pot;
          else
            acc_pm[3 * i + dim] += v;
        }
    }
#undef IDX
  free(re);
  free(im);
  free(kre);
  free(kim);
  free(fg);
  return 0;
}

/* ------------------------------------------------------------------ */
/* FORCETEST direct summation                                          */
/* ------------------------------------------------------------------ */

void orc_forcetest(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype, int ns,
                   const int *sel, int ewald, const double *tab, int tew_n, double *acc, double *acc_short, double *pot,
                   double *pot_short, int nthreads)
{
  orc_forcetest_hr(p, NULL, NULL, n, pos, mass, stype, ns, sel, ewald, tab, tew_n, acc, acc_short, pot, pot_short, nthreads);
}

void orc_forcetest_hr(const orc_params *p, const orc_highres *hr, const uint8_t *hrflag, int n, const double *pos,
                      const double *mass, const uint8_t *stype, int ns, const int *sel, int ewald, const double *tab, int tew_n,
                      double *acc, double *acc_short, double *pot, double *pot_short, int nthreads)
{
  double fs[ORC_MAXSOFT + 1];
  orc_set_softenings(p, fs);
  double asmth0 = 0, rcut = 0;
  if(p->pmgrid > 0)
    orc_asmth_rcut(p, &asmth0, &rcut);
  const double L   = p->box_size;
  const int per    = p->periodic;
  const double vol = L * L * L;
#pragma omp parallel for schedule(dynamic, 4) num_threads(omp_threads(nthreads))
  for(int q = 0; q < ns; q++)
    {
      int i      = sel[q];
      double h_i = fs[stype ? stype[i] : 0];
      /* gravtree_forcetest.c:480-483: Asmth[1] for a target in the high-resolution region */
      const double asmth = (hr && hrflag && hrflag[i]) ? hr->asmth : asmth0;
      double a[3] = {0, 0, 0}, as[3] = {0, 0, 0}, pt = 0, pts = 0;
      for(int j = 0; j < n; j++)
        {
          double h_j  = fs[stype ? stype[j] : 0];
          double hmax = h_j > h_i ? h_j : h_i;
          double dx   = nearest(pos[3 * j + 0] - pos[3 * i + 0], L, per);
          double dy   = nearest(pos[3 * j + 1] - pos[3 * i + 1], L, per);
          double dz   = nearest(pos[3 * j + 2] - pos[3 * i + 2], L, per);
          double r2   = dx * dx + dy * dy + dz * dz;
          double m    = mass[j];
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
          double anx = dx * fac, any = dy * fac, anz = dz * fac, pn = wp;
          if(p->pmgrid > 0)
            {
              double u  = 0.5 / asmth * r;
              double ff = erfc(u) + 2.0 * u / sqrt(M_PI) * exp(-u * u) - 1.0;
              double fp = erfc(u);
              double fs_ = fac + fac_newton * ff;
              double ws  = wp + wp_newton * (fp - 1.0);
              as[0] += dx * fs_;
              as[1] += dy * fs_;
              as[2] += dz * fs_;
              pts += ws + m * M_PI / (asmth * asmth * vol);
            }
          double fc[4] = {0, 0, 0, 0};
          if(per && ewald == 1)
            orc_forcetest_ewald_lookup(tab, tew_n, L, dx, dy, dz, fc);
          else if(per && ewald == 2)
            {
              /* the exact series in the table's sign convention (gravtree_forcetest.c:1040-1084 flips
               * the sign relative to the raw series) */
              orc_forcetest_ewald_exact(L, fabs(dx), fabs(dy), fabs(dz), fc);
              fc[0] *= dx < 0 ? -1 : 1;
              fc[1] *= dy < 0 ? -1 : 1;
              fc[2] *= dz < 0 ? -1 : 1;
            }
          a[0] += anx + m * fc[0];
          a[1] += any + m * fc[1];
          a[2] += anz + m * fc[2];
          pt += pn + m * fc[3];
        }
      if(acc)
        for(int d = 0; d < 3; d++)
          acc[3 * q + d] = a[d];
      if(acc_short)
        for(int d = 0; d < 3; d++)
          acc_short[3 * q + d] = as[d];
      if(pot)
        pot[q] = pt;
      if(pot_short)
        pot_short[q] = pts;
    }
}

void orc_direct_summation(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype, double *acc,
                          double *pot, int nthreads)
{
  orc_direct_summation_hr(p, NULL, NULL, n, pos, mass, stype, acc, pot, nthreads);
}

void orc_direct_summation_hr(const orc_params *p, const orc_highres *hr, const uint8_t *hrflag, int n, const double *pos,
                             const double *mass, const uint8_t *stype, double *acc, double *pot, int nthreads)
{
  /* force_evaluate_direct, forcetree_walk.c:557-708; Asmth[1] for a target in the high-resolution
   * region (:576-581) */
  shortrange_ctx c0, c1;
  shortrange_setup(p, &c0);
  if(hr && hrflag)
    shortrange_setup_len(p, hr->asmth, hr->rcut, &c1);
  double fs[ORC_MAXSOFT + 1];
  orc_set_softenings(p, fs);
  const double L = p->box_size;
  const int per  = p->periodic;
#pragma omp parallel for schedule(dynamic, 16) num_threads(omp_threads(nthreads))
  for(int i = 0; i < n; i++)
    {
      double h_i = fs[stype ? stype[i] : 0];
      const shortrange_ctx *cp = (hr && hrflag && hrflag[i]) ? &c1 : &c0;
      double ax = 0, ay = 0, az = 0, pt = 0;
      for(int j = 0; j < n; j++)
        {
          double h_j  = fs[stype ? stype[j] : 0];
          double hmax = h_j > h_i ? h_j : h_i;
          double dx   = nearest(pos[3 * j + 0] - pos[3 * i + 0], L, per);
          double dy   = nearest(pos[3 * j + 1] - pos[3 * i + 1], L, per);
          double dz   = nearest(pos[3 * j + 2] - pos[3 * i + 2], L, per);
          double r    = sqrt(dx * dx + dy * dy + dz * dz), fac, wp;
          if(!pair_kernel(cp, r, hmax, &fac, &wp))
            continue;
          pt += mass[j] * wp;
          fac *= mass[j];
          ax += dx * fac;
          ay += dy * fac;
          az += dz * fac;
        }
      acc[3 * i + 0] = ax;
      acc[3 * i + 1] = ay;
      acc[3 * i + 2] = az;
      pot[i]         = pt;
    }
}

/* ------------------------------------------------------------------ */
/* driver: accel.c + longrange.c                                       */
/* ------------------------------------------------------------------ */

static double wall(void)
{
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0;
#endif
}

long long orc_compute_grav_accelerations(const orc_params *p, int n, const double *pos, const double *mass, const uint8_t *stype,
                                         int first_step, float *old_acc, double *acc, double *pot, int *nint, int nthreads,
                                         double *t_out)
{
  return orc_compute_grav_accelerations_hr(p, NULL, NULL, n, pos, mass, stype, first_step, old_acc, acc, pot, nint, nthreads,
                                           t_out);
}

long long orc_compute_grav_accelerations_hr(const orc_params *p, orc_highres *hr, const uint8_t *type, int n, const double *pos,
                                            const double *mass, const uint8_t *stype, int first_step, float *old_acc, double *acc,
                                            double *pot, int *nint, int nthreads, double *t_out)
{
  if(n <= 0)
    return 0;
  uint8_t *hrflag = NULL;
  if(hr && (!type || p->pmgrid <= 0 || !p->periodic))
    return -3;
  double *acc_pm = (double *)calloc(3 * (size_t)n, sizeof(double));
  double *pot_pm = (double *)calloc((size_t)n, sizeof(double));
  double *at     = (double *)malloc(3 * sizeof(double) * n);
  double *pt     = (double *)malloc(sizeof(double) * n);
  float *oa      = old_acc ? old_acc : (float *)calloc((size_t)n, sizeof(float));
  double *ewald  = NULL;
  if(p->periodic && p->pmgrid == 0)
    {
      ewald = (double *)malloc(sizeof(double) * 4 * 65 * 65 * 65);
      orc_ewald_table_tree(64, p->box_size, ewald, nthreads);
    }
  double m_total = 0;
  for(int i = 0; i < n; i++)
    m_total += mass[i];
  double asmth = 0, rcut = 0;
  if(p->pmgrid > 0)
    orc_asmth_rcut(p, &asmth, &rcut);

  int passes      = (p->opening == 1 && first_step && p->err_tol_theta > 0) ? 2 : 1;
  long long total = 0;
  double tb = 0, tw = 0, tpm = 0;
  for(int pass = 0; pass < passes; pass++)
    {
      int use_bh = (p->opening == 0) || (pass == 0 && passes == 2);
      double t0  = wall();
      if(p->pmgrid > 0 && pass == 0)
        {
          orc_pm_periodic(p, n, pos, mass, acc_pm, pot_pm);
          if(hr)
            {
              /* longrange.c:129-140: the high-resolution grid after the periodic one; its region is sized
               * when unset and again when a high-resolution particle has left it */
              if(!hr->valid && orc_highres_regionsize(p, hr, n, pos, type) != 0)
                return -4;
              int rc = orc_pm_highres(p, hr, n, pos, mass, type, acc_pm, pot_pm);
              if(rc == 1)
                {
                  if(orc_highres_regionsize(p, hr, n, pos, type) != 0)
                    return -4;
                  rc = orc_pm_highres(p, hr, n, pos, mass, type, acc_pm, pot_pm);
                }
              if(rc != 0)
                return -5;
              hrflag = (uint8_t *)malloc((size_t)n);
              orc_highres_flags(p, hr, n, pos, type, hrflag);
            }
        }
      double t1 = wall();
      if(pass == 0)
        tpm = t1 - t0;
      orc_tree *t = NULL;
      double af   = n < 1000 ? 10.0 : 0.7; /* init.c:110-114 */
      int rc;
      while((rc = orc_tree_build(p, n, pos, mass, stype, af, &t)) < 0)
        {
          af *= 1.15; /* forcetree.c:210 */
          if(af > 30.0)
            return -1;
        }
      double t2 = wall();
      total     = orc_tree_walk_hr(p, t, use_bh, oa, ewald, hr, hrflag, n, NULL, at, pt, nint, nthreads);
      double t3 = wall();
      orc_tree_free(t);
      /* the potential constant of accel.c:269-274 with the mass of each PM region (forcetree.c:440-456) */
      double pot_const = 0;
      if(p->pmgrid > 0 && p->periodic)
        {
          double m_high = 0;
          for(int i = 0; hrflag && i < n; i++)
            if(hrflag[i])
              m_high += mass[i];
          const double vol = p->box_size * p->box_size * p->box_size;
          pot_const        = (m_total - m_high) * M_PI / (asmth * asmth * vol);
          if(hrflag)
            pot_const += m_high * M_PI / (hr->asmth * hr->asmth * vol);
        }
      /* gravity_force_finalize, accel.c:204-347 */
      for(int i = 0; i < n; i++)
        {
          double ax = at[3 * i + 0] + acc_pm[3 * i + 0] / p->G;
          double ay = at[3 * i + 1] + acc_pm[3 * i + 1] / p->G;
          double az = at[3 * i + 2] + acc_pm[3 * i + 2] / p->G;
          oa[i]     = (float)sqrt(ax * ax + ay * ay + az * az);
          for(int d = 0; d < 3; d++)
            acc[3 * i + d] = at[3 * i + d] * p->G + acc_pm[3 * i + d];
          pot[i] = (pt[i] + pot_const) * p->G + pot_pm[i];
        }
      tb = t2 - t1;
      tw = t3 - t2;
    }
  if(t_out)
    {
      t_out[0] = tb;
      t_out[1] = tw;
      t_out[2] = tpm;
    }
  free(acc_pm);
  free(pot_pm);
  free(at);
  free(pt);
  free(ewald);
  free(hrflag);
  if(!old_acc)
    free(oa);
  return total;
}

/* ------------------------------------------------------------------ */
/* unbinding potential: subfind_loctree.c:564-730                      */
/* ------------------------------------------------------------------ */

long long orc_group_potential_bh(int n, const double *pos, const double *mass, double h, double theta, double *pot)
{
  return orc_group_potential_bh_cube(n, pos, mass, h, theta, NULL, pot);
}

long long orc_group_potential_bh_cube(int n, const double *pos, const double *mass, double h, double theta, const double *cube,
                                      double *pot)
{
  orc_params p;
  memset(&p, 0, sizeof(p));
  p.periodic = 0;
  p.box_size = 1;
  p.nsoft    = 1;
  p.soft[0]  = h / 2.8;
  orc_tree *t = NULL;
  double af   = n < 1000 ? 10.0 : 0.7;
  while(orc_tree_build_cube(&p, n, pos, mass, NULL, af, cube, &t) < 0)
    af *= 1.15;
  if(!t)
    return 0;
  long long total = 0;
  const int N     = n;
  double theta2   = theta * theta;
  for(int i = 0; i < n; i++)
    {
      const double *x = &t->pos[3 * i];
      double pt       = 0;
      int no          = N;
      while(no >= 0)
        {
          double dx, dy, dz, r2, m;
          if(no < N)
            {
              dx = t->pos[3 * no] - x[0];
              dy = t->pos[3 * no + 1] - x[1];
              dz = t->pos[3 * no + 2] - x[2];
              m  = mass[no];
              no = t->pnext[no];
            }
          else
            {
              int k = no - N;
              m     = t->mass[k];
              dx    = t->s[3 * k] - x[0];
              dy    = t->s[3 * k + 1] - x[1];
              dz    = t->s[3 * k + 2] - x[2];
              r2    = dx * dx + dy * dy + dz * dz;
              if(t->len[k] * t->len[k] > r2 * theta2 && m)
                {
                  no = t->nextnode[k];
                  continue;
                }
              no = t->sibling[k];
            }
          r2       = dx * dx + dy * dy + dz * dz;
          double r = sqrt(r2);
          if(r >= h)
            pt -= m / r;
          else
            {
              double hi = 1.0 / h, u = r * hi, wp;
              if(u < 0.5)
                wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
              else
                wp = -3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
              pt += m * hi * wp;
            }
          total++;
        }
      pot[i] = pt;
    }
  orc_tree_free(t);
  return total;
}

/* ------------------------------------------------------------------ */
/* SUBFIND unbinding                                                   */
/* ------------------------------------------------------------------ */

void orc_loctree_extent(int n, const double *pos, double cube[4])
{
  /* subfind_loctree_findExtent, subfind_loctree.c:101-160: the largest extent of the
   * bounding box times 1.001, centred on the box */
  double xmin[3], xmax[3];
  for(int j = 0; j < 3; j++)
    {
      xmin[j] = 1.0e306; /* MAX_REAL_NUMBER */
      xmax[j] = -1.0e306;
    }
  for(int i = 0; i < n; i++)
    for(int j = 0; j < 3; j++)
      {
        if(xmin[j] > pos[3 * i + j])
          xmin[j] = pos[3 * i + j];
        if(xmax[j] < pos[3 * i + j])
          xmax[j] = pos[3 * i + j];
      }
  double len = 0;
  for(int j = 0; j < 3; j++)
    if(xmax[j] - xmin[j] > len)
      len = xmax[j] - xmin[j];
  len *= 1.001;
  /* a group whose particles all coincide has no extent; any cube then serves */
  if(!(len > 0))
    len = 1.0;
  for(int j = 0; j < 3; j++)
    cube[j] = n > 0 ? 0.5 * (xmin[j] + xmax[j]) : 0.0;
  cube[3] = len;
}

static int cmp_desc(const void *a, const void *b)
{
  /* subfind_compare_binding_energy: largest first */
  double x = *(const double *)a, y = *(const double *)b;
  return x > y ? -1 : (x < y ? 1 : 0);
}

int orc_subfind_unbind(const orc_unbind_params *up, int n, const double *pos, const double *vel, const double *mass,
                       const double *utherm, const uint8_t *type, double eps, const double *cube_in, uint8_t *bound,
                       double *potential, double *bnd_energy_out, int *len_non_gas_out, int *iterations_out)
{
  /* subfind_unbind, subfind_serial.c:429-637, for one group of an isolated set
   * (positions need no nearest-image wrap).  The local tree's root cube is kept
   * for every iteration (subfind_serial.c:107, subfind_loctree.c:192-206). */
  const double G        = up->G;
  const double cf_atime = up->comoving ? up->time_a : 1.0;
  const double vel_to_phys = up->comoving ? 1.0 / up->time_a : 1.0; /* :441-452 */
  const double atime       = up->comoving ? up->time_a : 1.0;
  const double H_of_a      = up->comoving ? up->hubble_a : 0.0;
  const double h           = 2.8 * eps; /* ForceSoftening, grav_softening.c:90-95 */
  double cube[4];
  if(cube_in)
    memcpy(cube, cube_in, sizeof(cube));
  else
    orc_loctree_extent(n, pos, cube);

  int *ud        = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  double *pot    = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  double *E      = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  double *sorted = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  double *spos   = (double *)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
  double *smass  = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  double *spot   = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for(int i = 0; i < n; i++)
    ud[i] = i;
  int len = n, phaseflag = 0, iter = 0, unbound = 0, len_non_gas = 0, rc = 0;
  double weakly_bound_limit = 0, minpot = 0, ppos[3] = {0, 0, 0};
  if(n > 0)
    do
      {
        if(phaseflag == 0)
          {
            /* potential of every member with a tree over the current members (:458-477) */
            for(int i = 0; i < len; i++)
              {
                int p = ud[i];
                for(int j = 0; j < 3; j++)
                  spos[3 * i + j] = pos[3 * p + j];
                smass[i] = mass[p];
              }
            orc_group_potential_bh_cube(len, spos, smass, h, up->theta, cube, spot);
            int minindex = -1;
            minpot       = 1.0e30;
            for(int i = 0; i < len; i++)
              {
                int p  = ud[i];
                pot[p] = G / cf_atime * spot[i];
                if(pot[p] < minpot || minindex == -1)
                  {
                    minpot   = pot[p];
                    minindex = p;
                  }
              }
            for(int j = 0; j < 3; j++)
              ppos[j] = pos[3 * minindex + j];
          }
        else
          {
            /* :491-497: the reference evaluates the potential of the weakly bound
             * members again but keeps the old value and scales it by G/a once more
             * (the new value is discarded), so only that scaling is restated */
            for(int i = 0; i < len; i++)
              {
                int p = ud[i];
                if(E[p] >= weakly_bound_limit)
                  pot[p] *= G / cf_atime;
              }
          }

        /* bulk velocity and centre of mass (:501-532) */
        double v[3] = {0, 0, 0}, s[3] = {0, 0, 0}, TotMass = 0;
        for(int i = 0; i < len; i++)
          {
            int p = ud[i];
            for(int j = 0; j < 3; j++)
              {
                double ddxx = pos[3 * p + j] - ppos[j];
                s[j] += mass[p] * ddxx;
                v[j] += mass[p] * vel[3 * p + j];
              }
            TotMass += mass[p];
          }
        for(int j = 0; j < 3; j++)
          {
            v[j] /= TotMass;
            s[j] /= TotMass;
            s[j] += ppos[j];
          }

        /* binding energies (:534-556) */
        for(int i = 0; i < len; i++)
          {
            int p = ud[i];
            double dv[3];
            for(int j = 0; j < 3; j++)
              {
                dv[j]     = vel_to_phys * (vel[3 * p + j] - v[j]);
                double dx = atime * (pos[3 * p + j] - s[j]);
                dv[j] += H_of_a * dx;
              }
            E[p] = pot[p] + 0.5 * (dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            E[p] += G / cf_atime * mass[p] / (h / 2.8); /* self-energy */
            if(type && type[p] == 0 && utherm)
              E[p] += utherm[p];
            sorted[i] = E[p];
          }

        qsort(sorted, len, sizeof(double), cmp_desc); /* :558 */
        double energy_limit = sorted[(int)(0.25 * len)];
        int i;
        for(i = 0, unbound = 0; i < len - 1; i++)
          {
            if(sorted[i] > 0)
              unbound++;
            else
              unbound--;
            if(unbound <= 0)
              break;
          }
        weakly_bound_limit = sorted[i];

        /* omit unbound members, at most a quarter of them (:574-590) */
        for(i = 0, unbound = 0, len_non_gas = 0; i < len; i++)
          {
            int p = ud[i];
            if(E[p] > 0 && E[p] > energy_limit)
              {
                unbound++;
                ud[i] = ud[len - 1];
                i--;
                len--;
              }
            else if(!type || type[p] != 0)
              len_non_gas++;
          }

        if(len < up->des_link_ngb)
          break;

        if(phaseflag == 0)
          {
            if(unbound > 0)
              phaseflag = 1;
          }
        else if(unbound == 0)
          {
            phaseflag = 0; /* one more pass over all members */
            unbound   = 1;
          }

        if(iter++ > up->max_iter)
          {
            rc = -1;
            break;
          }
      }
    while(unbound > 0);

  for(int q = 0; q < n; q++)
    bound[q] = 0;
  for(int q = 0; q < len; q++)
    bound[ud[q]] = 1;
  if(potential)
    memcpy(potential, pot, sizeof(double) * n);
  if(bnd_energy_out)
    memcpy(bnd_energy_out, E, sizeof(double) * n);
  if(len_non_gas_out)
    *len_non_gas_out = len_non_gas;
  if(iterations_out)
    *iterations_out = iter;
  free(ud);
  free(pot);
  free(E);
  free(sorted);
  free(spos);
  free(smass);
  free(spot);
  return rc < 0 ? rc : len;
}

/* ------------------------------------------------------------------ */
/* Gravity timesteps                                                   */
/* ------------------------------------------------------------------ */

int orc_gravity_timesteps(const orc_timestep_params *tp, const orc_params *p, int n, const double *acc, const double *acc_pm,
                          const double *mass, const uint8_t *stype, double *dt, double *dt_displacement)
{
  double fs[ORC_MAXSOFT + 1];
  orc_set_softenings(p, fs);
  const double a2inv = tp->comoving ? 1.0 / (tp->time_a * tp->time_a) : 1.0;
  const double atime = tp->comoving ? tp->time_a : 1.0;
  const double hub   = tp->comoving ? tp->hubble_a : 1.0;
  /* find_long_range_step_constraint, timestep.c:525-565 (every particle collisionless) */
  double dtdisp = tp->max_size_timestep;
  if(acc_pm)
    {
      double dtmin = 1e306; /* MAX_DOUBLE_NUMBER */
      for(int i = 0; i < n; i++)
        {
          double ax = a2inv * acc_pm[3 * i], ay = a2inv * acc_pm[3 * i + 1], az = a2inv * acc_pm[3 * i + 2];
          double ac = sqrt(ax * ax + ay * ay + az * az);
          if(ac < 1e-37) /* MIN_FLOAT_NUMBER */
            ac = 1e-37;
          double d = sqrt(2.0 * tp->err_tol_int_accuracy * atime * fs[stype ? stype[i] : 0] / (2.8 * ac));
          d *= hub;
          if(d < dtmin)
            dtmin = d;
        }
      dtmin *= 2.0;
      dtdisp = dtmin;
      if(dtdisp > tp->max_size_timestep)
        dtdisp = tp->max_size_timestep;
      if(dtdisp < tp->min_size_timestep)
        dtdisp = tp->min_size_timestep;
    }
  if(dt_displacement)
    *dt_displacement = dtdisp;
  /* get_timestep_gravity, timestep.c:284-344, before the integer timeline */
  int below = 0;
  for(int i = 0; i < n; i++)
    {
      double ax = a2inv * acc[3 * i], ay = a2inv * acc[3 * i + 1], az = a2inv * acc[3 * i + 2];
      double ac = sqrt(ax * ax + ay * ay + az * az);
      if(ac == 0)
        ac = 1.0e-30;
      double d = sqrt(2 * tp->err_tol_int_accuracy * atime * fs[stype ? stype[i] : 0] / 2.8 / ac);
      d *= hub;
      if(mass[i] == 0)
        d = tp->max_size_timestep;
      if(d >= tp->max_size_timestep)
        d = tp->max_size_timestep;
      if(d < tp->min_size_timestep)
        {
          below++;
          d = tp->min_size_timestep;
        }
      if(acc_pm && d >= dtdisp)
        d = dtdisp;
      dt[i] = d;
    }
  return below;
}

/* ------------------------------------------------------------------ */
/* Peano-Hilbert keys, peano.c:334-496                                  */
/* ------------------------------------------------------------------ */

/* The curve's tables, restated from peano.c: for the key, the next orientation
 * (rottable3, :334-350) and the key digit (subpix3, :352-362) of each of 48
 * orientations and each octant pix = x bit << 2 | y bit << 1 | z bit; for the
 * inverse, the octant order of 24 rotations (quadrants, :393-424) with the
 * rotation maps and the sense flips (:426-433). */
static const unsigned char ph_next_rotation[384] = {
  36, 28, 25, 27, 10, 10, 25, 27, 29, 11, 24, 24, 37, 11, 26, 26,
  8, 8, 25, 27, 30, 38, 25, 27, 9, 39, 24, 24, 9, 31, 26, 26,
  40, 24, 44, 32, 40, 6, 44, 6, 25, 7, 33, 7, 41, 41, 45, 45,
  4, 42, 4, 46, 26, 42, 34, 46, 43, 43, 47, 47, 5, 27, 5, 35,
  33, 35, 36, 28, 33, 35, 2, 2, 32, 32, 29, 3, 34, 34, 37, 3,
  33, 35, 0, 0, 33, 35, 30, 38, 32, 32, 1, 39, 34, 34, 1, 31,
  24, 42, 32, 46, 14, 42, 14, 46, 43, 43, 47, 47, 25, 15, 33, 15,
  40, 12, 44, 12, 40, 26, 44, 34, 13, 27, 13, 35, 41, 41, 45, 45,
  28, 41, 28, 22, 38, 43, 38, 22, 42, 40, 23, 23, 29, 39, 29, 39,
  41, 36, 20, 36, 43, 30, 20, 30, 37, 31, 37, 31, 42, 40, 21, 21,
  28, 18, 28, 45, 38, 18, 38, 47, 19, 19, 46, 44, 29, 39, 29, 39,
  16, 36, 45, 36, 16, 30, 47, 30, 37, 31, 37, 31, 17, 17, 46, 44,
  12, 4, 1, 3, 34, 34, 1, 3, 5, 35, 0, 0, 13, 35, 2, 2,
  32, 32, 1, 3, 6, 14, 1, 3, 33, 15, 0, 0, 33, 7, 2, 2,
  16, 0, 20, 8, 16, 30, 20, 30, 1, 31, 9, 31, 17, 17, 21, 21,
  28, 18, 28, 22, 2, 18, 10, 22, 19, 19, 23, 23, 29, 3, 29, 11,
  9, 11, 12, 4, 9, 11, 26, 26, 8, 8, 5, 27, 10, 10, 13, 27,
  9, 11, 24, 24, 9, 11, 6, 14, 8, 8, 25, 15, 10, 10, 25, 7,
  0, 18, 8, 22, 38, 18, 38, 22, 19, 19, 23, 23, 1, 39, 9, 39,
  16, 36, 20, 36, 16, 2, 20, 10, 37, 3, 37, 11, 17, 17, 21, 21,
  4, 17, 4, 46, 14, 19, 14, 46, 18, 16, 47, 47, 5, 15, 5, 15,
  17, 12, 44, 12, 19, 6, 44, 6, 13, 7, 13, 7, 18, 16, 45, 45,
  4, 42, 4, 21, 14, 42, 14, 23, 43, 43, 22, 20, 5, 15, 5, 15,
  40, 12, 21, 12, 40, 6, 23, 6, 13, 7, 13, 7, 41, 41, 22, 20};
static const unsigned char ph_digit[384] = {
  0, 7, 1, 6, 3, 4, 2, 5, 7, 4, 6, 5, 0, 3, 1, 2,
  4, 3, 5, 2, 7, 0, 6, 1, 3, 0, 2, 1, 4, 7, 5, 6,
  1, 0, 6, 7, 2, 3, 5, 4, 0, 3, 7, 4, 1, 2, 6, 5,
  3, 2, 4, 5, 0, 1, 7, 6, 2, 1, 5, 6, 3, 0, 4, 7,
  6, 1, 7, 0, 5, 2, 4, 3, 1, 2, 0, 3, 6, 5, 7, 4,
  2, 5, 3, 4, 1, 6, 0, 7, 5, 6, 4, 7, 2, 1, 3, 0,
  7, 6, 0, 1, 4, 5, 3, 2, 6, 5, 1, 2, 7, 4, 0, 3,
  5, 4, 2, 3, 6, 7, 1, 0, 4, 7, 3, 0, 5, 6, 2, 1,
  6, 7, 5, 4, 1, 0, 2, 3, 7, 0, 4, 3, 6, 1, 5, 2,
  0, 1, 3, 2, 7, 6, 4, 5, 1, 6, 2, 5, 0, 7, 3, 4,
  2, 3, 1, 0, 5, 4, 6, 7, 3, 4, 0, 7, 2, 5, 1, 6,
  4, 5, 7, 6, 3, 2, 0, 1, 5, 2, 6, 1, 4, 3, 7, 0,
  7, 0, 6, 1, 4, 3, 5, 2, 0, 3, 1, 2, 7, 4, 6, 5,
  3, 4, 2, 5, 0, 7, 1, 6, 4, 7, 5, 6, 3, 0, 2, 1,
  6, 7, 1, 0, 5, 4, 2, 3, 7, 4, 0, 3, 6, 5, 1, 2,
  4, 5, 3, 2, 7, 6, 0, 1, 5, 6, 2, 1, 4, 7, 3, 0,
  1, 6, 0, 7, 2, 5, 3, 4, 6, 5, 7, 4, 1, 2, 0, 3,
  5, 2, 4, 3, 6, 1, 7, 0, 2, 1, 3, 0, 5, 6, 4, 7,
  0, 1, 7, 6, 3, 2, 4, 5, 1, 2, 6, 5, 0, 3, 7, 4,
  2, 3, 5, 4, 1, 0, 6, 7, 3, 0, 4, 7, 2, 1, 5, 6,
  1, 0, 2, 3, 6, 7, 5, 4, 0, 7, 3, 4, 1, 6, 2, 5,
  7, 6, 4, 5, 0, 1, 3, 2, 6, 1, 5, 2, 7, 0, 4, 3,
  5, 4, 6, 7, 2, 3, 1, 0, 4, 3, 7, 0, 5, 2, 6, 1,
  3, 2, 0, 1, 4, 5, 7, 6, 2, 5, 1, 6, 3, 4, 0, 7};
static const int ph_quadrant[192] = {
  0, 7, 1, 6, 3, 4, 2, 5, 7, 4, 6, 5, 0, 3, 1, 2,
  4, 3, 5, 2, 7, 0, 6, 1, 3, 0, 2, 1, 4, 7, 5, 6,
  1, 0, 6, 7, 2, 3, 5, 4, 0, 3, 7, 4, 1, 2, 6, 5,
  3, 2, 4, 5, 0, 1, 7, 6, 2, 1, 5, 6, 3, 0, 4, 7,
  6, 1, 7, 0, 5, 2, 4, 3, 1, 2, 0, 3, 6, 5, 7, 4,
  2, 5, 3, 4, 1, 6, 0, 7, 5, 6, 4, 7, 2, 1, 3, 0,
  7, 6, 0, 1, 4, 5, 3, 2, 6, 5, 1, 2, 7, 4, 0, 3,
  5, 4, 2, 3, 6, 7, 1, 0, 4, 7, 3, 0, 5, 6, 2, 1,
  6, 7, 5, 4, 1, 0, 2, 3, 7, 0, 4, 3, 6, 1, 5, 2,
  0, 1, 3, 2, 7, 6, 4, 5, 1, 6, 2, 5, 0, 7, 3, 4,
  2, 3, 1, 0, 5, 4, 6, 7, 3, 4, 0, 7, 2, 5, 1, 6,
  4, 5, 7, 6, 3, 2, 0, 1, 5, 2, 6, 1, 4, 3, 7, 0};
static const int ph_rotx_map[24] = {
  4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15,
  0, 1, 2, 3, 17, 18, 19, 16, 23, 20, 21, 22};
static const int ph_roty_map[24] = {
  1, 2, 3, 0, 16, 17, 18, 19, 11, 8, 9, 10,
  22, 23, 20, 21, 14, 15, 12, 13, 4, 5, 6, 7};
static const int ph_rotx[8] = {
  3, 0, 0, 2, 2, 0, 0, 1};
static const int ph_roty[8] = {
  0, 1, 1, 2, 2, 3, 3, 0};
static const int ph_sense[8] = {
  -1, -1, -1, 1, 1, -1, -1, -1};

/* peano_hilbert_key, peano.c:374-390: `bits` levels, most significant first;
 * the key (up to 3 x 42 bits) is returned as its high and low 64-bit words. */
void orc_peano_hilbert_key(uint64_t x, uint64_t y, uint64_t z, int bits, uint64_t *hi, uint64_t *lo)
{
  unsigned __int128 key = 0;
  int rotation          = 0;
  for(int l = bits - 1; l >= 0; l--)
    {
      const int pix = (int)((((x >> l) & 1) << 2) | (((y >> l) & 1) << 1) | ((z >> l) & 1));
      key           = (key << 3) | ph_digit[8 * rotation + pix];
      rotation      = ph_next_rotation[8 * rotation + pix];
    }
  *hi = (uint64_t)(key >> 64);
  *lo = (uint64_t)key;
}

/* peano_hilbert_key_inverse, peano.c:446-496: the coordinates of a key, from
 * the quadrant tables (a second description of the same curve). */
void orc_peano_hilbert_key_inverse(uint64_t hi, uint64_t lo, int bits, uint64_t *x, uint64_t *y, uint64_t *z)
{
  int inv[24][8][3];
  for(int r = 0; r < 24; r++)
    for(int b = 0; b < 8; b++)
      {
        const int q = ph_quadrant[8 * r + b];
        inv[r][q][0] = (b >> 2) & 1;
        inv[r][q][1] = (b >> 1) & 1;
        inv[r][q][2] = b & 1;
      }
  const unsigned __int128 key = ((unsigned __int128)hi << 64) | lo;
  int rotation = 0, sense = 1;
  uint64_t ox = 0, oy = 0, oz = 0;
  for(int l = bits - 1; l >= 0; l--)
    {
      const int part = (int)((key >> (3 * l)) & 7);
      const int quad = sense == 1 ? part : 7 - part;
      ox             = (ox << 1) | (uint64_t)inv[rotation][quad][0];
      oy             = (oy << 1) | (uint64_t)inv[rotation][quad][1];
      oz             = (oz << 1) | (uint64_t)inv[rotation][quad][2];
      sense *= ph_sense[quad];
      for(int k = 0; k < ph_rotx[quad]; k++)
        rotation = ph_rotx_map[rotation];
      for(int k = 0; k < ph_roty[quad]; k++)
        rotation = ph_roty_map[rotation];
    }
  *x = ox;
  *y = oy;
  *z = oz;
}
