// agrav_internal.h -- shared declarations of the HIP gravity library.
//
// Layout of the device state (all struct-of-arrays, particles in Hilbert order
// after the build):
//   particles  : sorted positions (double x/y/z), mass, softening type,
//                caller index, and a 16-byte record {x,y,z fixed-point, mass f32}
//                for the FP32 walk;
//   nodes      : pre-order octree, each node covering the contiguous particle
//                range [first, last]; double COM/center/len/mass, AREPO's
//                nextnode/sibling threading (entries < N are particles, >= N
//                are nodes), and two 16-byte records for the FP32 walk;
//   children   : CSR lists of each node's child entries (group walk).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "agrav.h"

#define AG_NTAB 127
#define AG_KEY_LEVELS 21  // Hilbert levels of one set in the first 63-bit key word
#define AG_MAX_LEVELS 30  // deepest octree level, AREPO's MAX_TREE_LEVEL (forcetree.h:106)
#define AG_MAXSOFT AGRAV_MAX_SOFT_TYPES

namespace agrav
{

struct Status
{
  agrav_status code = AGRAV_OK;
  std::string msg;
};

// Device-side domain description, written by the extent kernel.
struct Domain
{
  double corner[3];
  double center[3];
  double len;
  double inv_len;
  double bigfac;    // len / 2^52 (domain_box.c:196)
  double fix_unit;  // length of one fixed-point step
  double fix_inv;   // 1 / fix_unit
  double fix_org[3]; // fixed-point origin (0 for periodic boxes, the corner otherwise)
  double xmin[3], xmax[3];
  int out_of_box;
  int pad;
};

struct SoftTable
{
  double h[AG_MAXSOFT + 1]; // ForceSoftening, sentinel 0 at [nsoft]
  int nsoft;
};

// Kernel-facing tree view.
struct TreeView
{
  int n;        // particles
  int m;        // nodes
  // particles (sorted)
  const double *px, *py, *pz, *pm;
  const uint8_t *ps;      // softening types (nullptr when single type)
  const int *pnext;       // AREPO Nextnode[] of particles
  const uint4 *prec;      // FP32 record {fx, fy, fz, mass bits}
  // nodes
  const double *sx, *sy, *sz, *nm; // COM and mass
  const double *cx, *cy, *cz, *len;
  const int *nextnode, *sibling;
  const uint8_t *maxsoft;
  const uint4 *recA;      // {comfx, comfy, comfz, mass bits}
  const uint4 *recB;      // {cenfx, cenfy, cenfz, len bits}
  const int2 *nrange;     // {first particle, particle count} per node
  const int *cbeg;        // children CSR
  const int *ccnt;         // children per node
  const int *cnn;          // of which nodes (listed first)
  const int *child;
  const int4 *meta;        // per node {start and count of its particle children in child[], first particle,
                           //           node children | min(particles, 255) << 4}
  const uint4 *cprec;      // FP32 records of particle children, at their child[] positions
  const int4 *nmeta;       // meta, recA, recB of node children, at their child[] positions
  const uint4 *nrecA, *nrecB;
  const Domain *dom;       // the domain cube of one set (its fixed-point step; out_of_box)
  const int *cnt;          // build counters: [13] a sort or scan look-back timed out, [50] threading error
};

// a tree whose build reported an error on the device (read by the walks before they
// start, so that a step whose build check is read back after the walk never walks it)
__device__ inline bool tree_invalid(const TreeView &tv)
{
  return tv.cnt[13] != 0 || tv.cnt[50] != 0 || tv.dom->out_of_box != 0;
}

struct WalkConst
{
  int periodic;
  double box, half_box;
  int pm;             // TreePM short-range walk
  double rcut, rcut2, asmthinv2, asmthfac;
  double theta2;      // BH
  double errtol;      // ErrTolForceAcc
  int use_bh;
  int ewald;          // periodic tree-only: Ewald correction per interaction
  double ewald_fac;   // 2 EN / box
  int multisoft;
  int leaf_bucket;    // opened nodes with at most this many particles are expanded into particles
  int eval_pot;       // EVALPOTENTIAL: also sum the potential
};

// Batch of independent isolated sets sharing one tree build (group walk queue
// and per-set geometry).
struct BatchView
{
  const int4 *groups;  // {first target, count, set, leaf bucket} per walk group, largest sets first
  const int *pset;     // set of each sorted particle
  const int *sroot;    // top entry of each set: its cube, or its particle when alone
  const int *send;     // entry after the set in the threaded (nextnode) order, -1 for the last
  const double *sh;    // softening length h = 2.8 eps per set
  const Domain *dom;   // domain cube per set
};

// gravity_force_finalize for one target (accel.c:204-347, no comoving terms,
// OmegaLambda = 0): OldAcc = |a_tree + a_PM / G| (accel.c:237-255), a = a_tree G
// + a_PM (accel.c:258-265), the periodic TreePM potential constant and G
// (accel.c:269-286); outputs in caller order.  Used by the finalize kernel and by
// the group walk's epilogue when the step fuses the two.
struct FinalOut
{
  const int *perm;                // sorted position -> caller index; null: nothing to write
  const double *acc_pm, *pot_pm;  // PM results, or null without PM
  int pm_sorted;                  // PM results in sorted (1) or caller (0) order
  double G, pot_const;
  double *acc, *pot;
  int *nint;
  float *oldacc, *oldacc_s;  // caller order (null: OldAcc is kept), sorted order
};

__device__ inline void finalize_one(const FinalOut &f, int64_t k, double ax, double ay, double az, double pt, int ni)
{
  const int i = f.perm[k];
  double p0 = 0, p1 = 0, p2 = 0, pp = 0;
  if(f.acc_pm)
    {
      const int64_t q = f.pm_sorted ? k : (int64_t)i;
      p0              = f.acc_pm[3 * q];
      p1              = f.acc_pm[3 * q + 1];
      p2              = f.acc_pm[3 * q + 2];
      pp              = f.pot_pm[q];
    }
  const double ox = ax + p0 / f.G, oy = ay + p1 / f.G, oz = az + p2 / f.G;
  if(f.oldacc)
    {
      const float oa = (float)sqrt(ox * ox + oy * oy + oz * oz);
      f.oldacc[i]    = oa;
      f.oldacc_s[k]  = oa;
    }
  f.acc[3 * (int64_t)i + 0] = ax * f.G + p0;
  f.acc[3 * (int64_t)i + 1] = ay * f.G + p1;
  f.acc[3 * (int64_t)i + 2] = az * f.G + p2;
  f.pot[i]                  = (pt + f.pot_const) * f.G + pp;
  f.nint[i]                 = ni;
}

// hip error helper
#define AG_HIP(call)                                                                                \
  do                                                                                                \
    {                                                                                               \
      hipError_t e_ = (call);                                                                       \
      if(e_ != hipSuccess)                                                                          \
        return ::agrav::hip_fail(e_, #call, __FILE__, __LINE__);                                    \
    }                                                                                               \
  while(0)

agrav_status hip_fail(hipError_t e, const char *what, const char *file, int line);
void set_error_msg(const std::string &m);

// scan / sort primitives (agrav_prims.hip).  Both report a look-back that timed
// out by setting *d_err (device int, cleared by the caller).
size_t scan_temp_bytes(int64_t n);
agrav_status exclusive_scan_i32(const int *d_in, int *d_out, int64_t n, int *d_total, void *d_temp, int *d_err, hipStream_t s);
size_t sort_temp_bytes(int64_t n);
// LSD radix sort of 64-bit keys with 32-bit values, stable; bits [lo_bit, key_bits)
// are sorted (keys equal in those bits keep their input order).  The passes alternate
// between the two buffer pairs.  Results end up in keys_out/vals_out; with in_input
// given, an even pass count instead leaves them in keys_in/vals_in (no copy) and sets
// *in_input.  ghist_zero: a device array of 8 x 256 ints that is zero on entry (the
// digit counts go there), or null (the sort clears its own in d_temp).
agrav_status radix_sort_u64(uint64_t *keys_in, int *vals_in, uint64_t *keys_out, int *vals_out, int64_t n, int key_bits,
                            void *d_temp, int *d_err, hipStream_t s, int lo_bit = 0, bool *in_input = nullptr,
                            int *ghist_zero = nullptr);

// AREPO's 3-D Peano-Hilbert key (peano_hilbert_key, peano.c:374-390) on b-bit
// coordinates, evaluated as a state machine: each level's key digit depends on
// that level's octant (x bit << 2 | y bit << 1 | z bit) and on an orientation
// state carried down from the levels above.  Entry [8 s + o] = (next state << 3)
// | digit, start state 0.  The 24 states are those of the reference's curve
// reachable from its start orientation; the table is derived from the curve's
// quadrant description (peano.c:393-433) by arepo_amd/gen_hilbert_table.py, and
// the keys are checked against the oracle's restatement of peano_hilbert_key.
// A key's top 3L bits are the key of the coordinates' top L bits.
#define AG_HILBERT_TABLE                                                                                          \
  {                                                                                                              \
      8,  23,  25,  38,  43,  44,  26,  37,   0,  49,  63,  70,  75,  50,  76,  69, \
     86,   7,  89,  56,  85, 100,  90,  99, 104, 115,   1,   2, 127, 116, 134, 133, \
    140, 151,   5,   6, 139, 152, 130, 129, 138, 117,   3,   4, 137, 118,  96,  79, \
    160,   9,  91,  10, 175, 102,  92, 101, 142, 113,  15,  16, 141, 114, 132, 131, \
     84,  13, 183,  14,  83,  98, 184,  97,  82,  11,  93,  12,  81, 128,  94,  47, \
     22, 167,  21,  68,  73, 168,  74,  67,  18,  51,  17, 176,  77,  52,  78, 191, \
     20,  53,  19,  66, 135,  54,  40,  65,  24, 155, 143, 156, 161, 162, 182, 181, \
     58,  57,  27, 144,  45,  46,  28, 159, 174, 173, 185, 186,  31, 148, 136, 147, \
     60,  59,  29,  34, 103,  72,  30,  33,  62,  61, 111,  36,  41,  42, 120,  35, \
    124,  39, 123, 112, 165, 166, 178, 177, 170, 169, 189, 190, 107,  32, 108, 119, \
     48,  87, 187, 188, 105, 150, 106, 149, 126, 153, 125, 154,  55,  80, 180, 179, \
    172, 171,  71,  88, 109, 146, 110, 145, 122, 157, 121, 158, 163, 164,  64,  95 \
  }

// digits of the L-level key of L-bit coordinates: the last min(L, 21) levels in
// `low`, the levels above them in `top` (key = top << 3 min(L, 21) | low)
__host__ __device__ inline void hilbert_digits(uint32_t x, uint32_t y, uint32_t z, int L, const uint8_t *tab, uint64_t &top,
                                               uint64_t &low)
{
  unsigned s = 0;
  uint64_t a = 0, c = 0;
  const int lo_levels = L < AG_KEY_LEVELS ? L : AG_KEY_LEVELS;
  int l = L - 1;
  for(; l >= lo_levels; l--)
    {
      const unsigned e = tab[8 * s + ((((x >> l) & 1u) << 2) | (((y >> l) & 1u) << 1) | ((z >> l) & 1u))];
      a = (a << 3) | (e & 7u);
      s = e >> 3;
    }
  for(; l >= 0; l--)
    {
      const unsigned e = tab[8 * s + ((((x >> l) & 1u) << 2) | (((y >> l) & 1u) << 1) | ((z >> l) & 1u))];
      c = (c << 3) | (e & 7u);
      s = e >> 3;
    }
  top = a;
  low = c;
}

}  // namespace agrav
