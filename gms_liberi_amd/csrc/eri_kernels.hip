// Per-class ERI + Fock digestion kernels for gfx950 (CDNA4), FP64.
//
// Method: Rys quadrature (Rys, Dupuis & King, J. Comput. Chem. 4, 154 (1983))
// on the primitive "(e0|f0)" integrals, e spanning angular momenta
// la..la+lb and f lc..lc+ld, contracted over primitives, then the horizontal
// recurrence (Head-Gordon & Pople, J. Chem. Phys. 89, 5777 (1988)) once per
// contracted quartet.  Roots/weights come from piecewise polynomials in the
// Rys argument X generated offline (gen/rys_tables.py).  This replaces the
// reference's three kernel families (rotated-axis, ERIC, Rys; src/rhf/) with
// one engine specialised per class at compile time.
//
// Two execution shapes:
//  * lane-per-quartet (small classes): one lane owns one contracted quartet,
//    all intermediates in registers, loops fully unrolled per class;
//  * workgroup-per-quartet (large classes): the (e0|f0) accumulators are
//    spread over the workgroup's lanes, the 2-D Rys tables of each primitive
//    quartet are staged in LDS, the HRR runs on a per-workgroup scratch area.
//
// Digestion (the reference's DIRFCK_RHF, e.g. src/rhf/rot_axis/int0000.F90:241-311)
// contracts each shell block with the six density blocks and adds the result
// to the packed Fock matrix with one hardware FP64 atomic per Fock element of
// the block (not one per integral).
#include <hip/hip_runtime.h>

#include "liberi_internal.hpp"

#define RYS_CONST static __device__ const
#include "rys_table.inc"
// host copy of the end of the tabulated range per root count
static const double rys_xmax_host[RYS_NMAX + 1] = {0.0, 39.0, 46.5, 53.5, 60.0, 66.0, 72.0, 77.5};

namespace liberi {

// polynomial degree, inverse interval width and intervals of the table of
// the N-root rule (the one- and two-root rules have their own, finer grid)
__host__ __device__ constexpr int rys_deg(int n) { return n <= RYS_NSMALL ? RYS_DEG_SMALL : RYS_DEG; }
__host__ __device__ constexpr double rys_inv_h(int n) { return n <= RYS_NSMALL ? RYS_INV_H_SMALL : RYS_INV_H; }
__host__ __device__ constexpr int rys_ncoef(int n) { return 2 * n * (rys_deg(n) + 1); }

// ------------------------------------------------------------ cartesian maps
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nrange(int l0, int l1) {
  int n = 0;
  for (int l = l0; l <= l1; ++l) n += ncart(l);
  return n;
}
// internal order inside a shell of angular momentum L: i = L - lx, index i(i+1)/2 + lz
__host__ __device__ constexpr int cidx(int lx, int ly, int lz) {
  return ((ly + lz) * (ly + lz + 1)) / 2 + lz;
}

// component tables of the angular-momentum range [L0, L1] (internal order)
template <int L0, int L1>
struct Range {
  static constexpr int n = nrange(L0, L1);
  int lx[n], ly[n], lz[n], l[n];
  constexpr Range() : lx(), ly(), lz(), l() {
    int k = 0;
    for (int L = L0; L <= L1; ++L)
      for (int i = 0; i <= L; ++i)
        for (int z = 0; z <= i; ++z) {
          lx[k] = L - i;
          ly[k] = i - z;
          lz[k] = z;
          l[k] = L;
          ++k;
        }
  }
};

// GAMESS component order (p x,y,z; d xx,yy,zz,xy,xz,yz; f xxx,yyy,zzz,xxy,xxz,
// xyy,yyz,xzz,yzz,xyz -- src/rhf/rys/int3300_rysgen.F90 ix/iy/iz tables) and the
// kernels' scale factors (sqrt3 mixed d, sqrt5 / sqrt15 mixed f,
// src/rhf/eric/int0030_ericgen.F90:49-57).
__host__ __device__ constexpr int gamess_pos(int lx, int ly, int lz) {
  const int L = lx + ly + lz;
  if (L <= 1) return lx ? 0 : (ly ? 1 : (lz ? 2 : 0));
  if (L == 2) {
    if (lx == 2) return 0;
    if (ly == 2) return 1;
    if (lz == 2) return 2;
    if (lx && ly) return 3;
    if (lx && lz) return 4;
    return 5;
  }
  if (lx == 3) return 0;
  if (ly == 3) return 1;
  if (lz == 3) return 2;
  if (lx == 2) return ly ? 3 : 4;
  if (ly == 2) return lx ? 5 : 6;
  if (lz == 2) return lx ? 7 : 8;
  return 9;
}
__host__ __device__ constexpr double cart_scale(int lx, int ly, int lz) {
  const int L = lx + ly + lz;
  const int mx = lx > ly ? (lx > lz ? lx : lz) : (ly > lz ? ly : lz);
  if (L == 2) return mx == 2 ? 1.0 : 1.7320508075688772935;
  if (L == 3) return mx == 3 ? 1.0 : (mx == 2 ? 2.2360679774997896964 : 3.8729833462074168852);
  return 1.0;
}

// packed lower-triangle index of (i, j); i, j < kMaxBasis keeps i(i+1) below
// 2^32, and the 24-bit multiply (full rate, where a 32-bit one is a quarter)
// holds i + 1 < 2^24
__device__ __forceinline__ uint32_t tri(uint32_t i, uint32_t j) {
  const uint32_t m = i >= j ? i : j, n = i >= j ? j : i;
  return (__umul24(m, m + 1) >> 1) + n;
}

// Reciprocal and reciprocal square root for the primitive loops: the
// hardware estimate (v_rcp_f64 / v_rsq_f64) refined by two Newton-Raphson
// steps, r <- r + r (1 - x r) and r <- r + r/2 (1 - x r^2), which leave the
// result within a few ulp of the correctly rounded value (the IEEE divide
// and square root sequences cost about twice as many instructions).
__device__ __forceinline__ double nr_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ double nr_rsqrt(double x) {
  double r = __builtin_amdgcn_rsq(x);
  r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
  return fma(0.5 * r, fma(-x * r, r, 1.0), r);
}
__device__ __forceinline__ float nr_rcp(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, r, 1.0f), r, r);
}

// --------------------------------------------------------------- Rys rule
// Tabulated rule for X < xswitch, asymptotic (Hermite) rule above.  xswitch
// never exceeds rys_xmax[N]; the asymptotic rule reproduces exactly the
// large-argument Boys values F_m = (2m-1)!!/(2X)^m sqrt(pi/4X) that the
// reference uses past its switch points.
// tab: the N-root part of the table (rys_table + rys_offset[N], or a copy of
// its first intervals in LDS), istride doubles per interval
template <int N>
__device__ __forceinline__ void rys_rule(double X, double xswitch, double (&u)[N], double (&w)[N],
                                         const double *tab = rys_table + rys_offset[N],
                                         int istride = rys_ncoef(N)) {
  if (X < xswitch) {
    constexpr int DG = rys_deg(N);
    constexpr double IH = rys_inv_h(N);
    const int j = (int)(X * IH);
    const double s = X * (2.0 * IH) - (double)(2 * j + 1);
    const double *c = tab + __mul24(j, istride);
#pragma unroll
    for (int r = 0; r < 2 * N; ++r) {
      const double *cr = c + r * (DG + 1);
      double acc = cr[DG];
#pragma unroll
      for (int k = DG - 1; k >= 0; --k) acc = fma(acc, s, cr[k]);
      if (r < N)
        u[r] = acc;
      else
        w[r - N] = acc;
    }
  } else {
    // 1/sqrt(X) and 1/X from the hardware estimate with Newton steps (a few
    // ulp; the IEEE divide and square root cost twice the instructions)
    const double isx = nr_rsqrt(X), ix = isx * isx;
#pragma unroll
    for (int r = 0; r < N; ++r) {
      u[r] = rys_herm_u[N][r] * ix;
      w[r] = rys_herm_w[N][r] * isx;
    }
  }
}

// one root of the rule (used by the cooperative kernels: one lane per root)
template <int N>
__device__ __forceinline__ void rys_root(double X, double xswitch, int r, double &u, double &w) {
  if (X < xswitch) {
    constexpr int DG = rys_deg(N);
    constexpr double IH = rys_inv_h(N);
    const int j = (int)(X * IH);
    const double s = X * (2.0 * IH) - (double)(2 * j + 1);
    const double *c = rys_table + rys_offset[N] + j * rys_ncoef(N);
    const double *cu = c + r * (DG + 1), *cw = c + (N + r) * (DG + 1);
    double au = cu[DG], aw = cw[DG];
#pragma unroll
    for (int k = DG - 1; k >= 0; --k) {
      au = fma(au, s, cu[k]);
      aw = fma(aw, s, cw[k]);
    }
    u = au;
    w = aw;
  } else {
    const double ix = 1.0 / X;
    u = rys_herm_u[N][r] * ix;
    w = rys_herm_w[N][r] * sqrt(ix);
  }
}

// 2-D Rys integrals G(i,k), i <= NI-1 (bra), k <= NK-1 (ket), for one root and
// one cartesian direction:  G(i+1,0) = C00 G(i,0) + i B10 G(i-1,0),
// G(i,k+1) = C00' G(i,k) + k B01 G(i,k-1) + i B00 G(i-1,k).
template <int NI, int NK, class T>
__device__ __forceinline__ void rys_2d(T (&g)[NI][NK], T c00, T c00p, T b10, T b01, T b00, T g00) {
  g[0][0] = g00;
  if (NI > 1) g[1][0] = c00 * g00;
#pragma unroll
  for (int i = 1; i + 1 < NI; ++i) g[i + 1][0] = c00 * g[i][0] + (T)i * b10 * g[i - 1][0];
  if (NK > 1) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      T v = c00p * g[i][0];
      if (i > 0) v += (T)i * b00 * g[i - 1][0];
      g[i][1] = v;
    }
#pragma unroll
    for (int k = 1; k + 1 < NK; ++k)
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        T v = c00p * g[i][k] + (T)k * b01 * g[i][k - 1];
        if (i > 0) v += (T)i * b00 * g[i - 1][k];
        g[i][k + 1] = v;
      }
  }
}

// segment holding quartet q, searched in [lo, hi): seg_start[lo] <= q < seg_start[hi]
__device__ __forceinline__ int find_segment(const ClassDev &cd, int64_t q, int lo, int hi) {
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cd.seg_start[mid] <= q)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void quartet_of_segment(const ClassDev &cd, int s, int64_t q, int32_t &bra,
                                                   int32_t &ket) {
  const int32_t row = cd.seg_bra[s];
  const int32_t col = cd.ket_list[cd.seg_ket0[s] + (int32_t)(q - cd.seg_start[s])];
  bra = cd.transposed ? col : row;
  ket = cd.transposed ? row : col;
}

// quartet index -> (bra pair, ket pair) through the class's segment table
__device__ __forceinline__ void decode_quartet(const ClassDev &cd, int64_t q, int32_t &bra, int32_t &ket) {
  int lo = 0, hi = cd.nseg;  // seg_start[lo] <= q < seg_start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cd.seg_start[mid] <= q)
      lo = mid;
    else
      hi = mid;
  }
  quartet_of_segment(cd, lo, q, bra, ket);
}

__device__ __forceinline__ void fock_add(double *F, uint32_t i, uint32_t j, double v) {
  unsafeAtomicAdd(F + tri(i, j), v);
}

// Read-only views of the pair and primitive records in the constant address
// space: a load at a wave-uniform address (the fixed pair of a row tile and
// its primitives) then goes through the scalar data cache into SGPRs,
// instead of one vector load per lane.  The kernels never write the records.
typedef __attribute__((address_space(4))) const PrimRec ConstPrimRec;
typedef __attribute__((address_space(4))) const PairRec ConstPairRec;
__device__ __forceinline__ ConstPrimRec *const_view(const PrimRec *p) { return (ConstPrimRec *)p; }
__device__ __forceinline__ ConstPairRec *const_view(const PairRec *p) { return (ConstPairRec *)p; }
// record i of a constant view, by value
__device__ __forceinline__ PrimRec load_rec(ConstPrimRec *v, int64_t i) {
  ConstPrimRec *q = v + i;
  PrimRec r;
  r.p = q->p;
  r.Px = q->Px;
  r.Py = q->Py;
  r.Pz = q->Pz;
  r.PAx = q->PAx;
  r.PAy = q->PAy;
  r.PAz = q->PAz;
  r.K = q->K;
  return r;
}
__device__ __forceinline__ PairRec load_rec(ConstPairRec *v, int64_t i) {
  ConstPairRec *q = v + i;
  PairRec r;
  r.prim_off = q->prim_off;
  r.nprim = q->nprim;
  r.bfA = q->bfA;
  r.bfB = q->bfB;
  r.ABx = q->ABx;
  r.ABy = q->ABy;
  r.ABz = q->ABz;
  r.sameAB = q->sameAB;
  r.shA = q->shA;
  r.cA = q->cA;
  r.cB = q->cB;
  r.x = q->x;
  return r;
}

// ===================================================== lane-per-quartet path
// size of the largest intermediate of the contracted HRR (both directions)
__host__ __device__ constexpr int hrr_max_buffer(int LA, int LB, int LC, int LD) {
  const int NF = nrange(LC, LC + LD);
  int m = nrange(LA, LA + LB) * NF;
  for (int j = 1; j <= LB; ++j) {
    const int s = nrange(LA, LA + LB - j) * ncart(j) * NF;
    m = s > m ? s : m;
  }
  for (int j = 1; j <= LD; ++j) {
    const int s = ncart(LA) * ncart(LB) * nrange(LC, LC + LD - j) * ncart(j);
    m = s > m ? s : m;
  }
  return m;
}

template <int LA, int LB, int LC, int LD>
struct Small {
  static constexpr int LAB = LA + LB, LCD = LC + LD;
  static constexpr int NR = (LAB + LCD) / 2 + 1;
  static constexpr int NE = nrange(LA, LAB), NF = nrange(LC, LCD);
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NABCD = NA * NB * NC * ND;
  static constexpr int NBUF = hrr_max_buffer(LA, LB, LC, LD);
};

// bra HRR level J: out[a in range(LA, LAB-J)][b in cart(J)][f] from level J-1
template <int LA, int LAB, int J, int NF, class T>
__device__ __forceinline__ void hrr_bra_level(const T *in, T *out, T abx, T aby, T abz) {
  constexpr Range<LA, LAB - J> ra{};
  constexpr Range<J, J> rb{};
  constexpr Range<J - 1, J - 1> rbm{};
  constexpr int nbm = ncart(J - 1), nb = ncart(J);
  (void)rbm;
#pragma unroll
  for (int a = 0; a < ra.n; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b) {
      const int bx = rb.lx[b], by = rb.ly[b], bz = rb.lz[b];
      const int dir = bx ? 0 : (by ? 1 : 2);
      const int bmi = cidx(bx - (dir == 0), by - (dir == 1), bz - (dir == 2));
      const int ax = ra.lx[a] + (dir == 0), ay = ra.ly[a] + (dir == 1), az = ra.lz[a] + (dir == 2);
      const int api = nrange(LA, ra.l[a]) + cidx(ax, ay, az);  // index in range(LA, LAB-J+1)
      const T ab = dir == 0 ? abx : (dir == 1 ? aby : abz);
#pragma unroll
      for (int f = 0; f < NF; ++f)
        out[(a * nb + b) * NF + f] = in[(api * nbm + bmi) * NF + f] + ab * in[(a * nbm + bmi) * NF + f];
    }
}

template <int LA, int LB, int NF, class T>
__device__ __forceinline__ void hrr_bra(T *buf0, T *buf1, T abx, T aby, T abz) {
  // level 0 lives in buf0 as [e in range(LA, LA+LB)][1][f]; result in buf0 when LB is even, buf1 when odd
  if constexpr (LB >= 1) hrr_bra_level<LA, LA + LB, 1, NF>(buf0, buf1, abx, aby, abz);
  if constexpr (LB >= 2) hrr_bra_level<LA, LA + LB, 2, NF>(buf1, buf0, abx, aby, abz);
  if constexpr (LB >= 3) hrr_bra_level<LA, LA + LB, 3, NF>(buf0, buf1, abx, aby, abz);
}

// ket HRR level J on the inner index: out[ab][c in range(LC, LCD-J)][d in cart(J)]
template <int LC, int LCD, int J, int NAB, class T>
__device__ __forceinline__ void hrr_ket_level(const T *in, T *out, T cdx, T cdy, T cdz) {
  constexpr Range<LC, LCD - J> rc{};
  constexpr Range<J, J> rd{};
  constexpr int nin = nrange(LC, LCD - J + 1) * ncart(J - 1);
  constexpr int nout = nrange(LC, LCD - J) * ncart(J);
  constexpr int ndm = ncart(J - 1), nd = ncart(J);
#pragma unroll
  for (int c = 0; c < rc.n; ++c)
#pragma unroll
    for (int d = 0; d < nd; ++d) {
      const int dx = rd.lx[d], dy = rd.ly[d], dz = rd.lz[d];
      const int dir = dx ? 0 : (dy ? 1 : 2);
      const int dmi = cidx(dx - (dir == 0), dy - (dir == 1), dz - (dir == 2));
      const int cx = rc.lx[c] + (dir == 0), cy = rc.ly[c] + (dir == 1), cz = rc.lz[c] + (dir == 2);
      const int cpi = nrange(LC, rc.l[c]) + cidx(cx, cy, cz);
      const T cdv = dir == 0 ? cdx : (dir == 1 ? cdy : cdz);
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab)
        out[ab * nout + c * nd + d] = in[ab * nin + cpi * ndm + dmi] + cdv * in[ab * nin + c * ndm + dmi];
    }
}

template <int LC, int LD, int NAB, class T>
__device__ __forceinline__ void hrr_ket(T *buf0, T *buf1, T cdx, T cdy, T cdz) {
  if constexpr (LD >= 1) hrr_ket_level<LC, LC + LD, 1, NAB>(buf0, buf1, cdx, cdy, cdz);
  if constexpr (LD >= 2) hrr_ket_level<LC, LC + LD, 2, NAB>(buf1, buf0, cdx, cdy, cdz);
  if constexpr (LD >= 3) hrr_ket_level<LC, LC + LD, 3, NAB>(buf0, buf1, cdx, cdy, cdz);
}

// Mixed precision: quartets whose Schwarz bound times max|D| is below tau
// may be evaluated in FP32 (tau = 0: never), in the classes the build enables
// it for.  The form is chosen per wave -- FP32 only when the quartet of every
// active lane qualifies -- so that a wave never runs both forms of the
// primitive loop.
__device__ __forceinline__ bool mixed_wave(const ClassDev &cd, const BuildArgs &ba, const PairRec &bra,
                                           const PairRec &ket) {
  if (!cd.fp32 || !(ba.fp32_tau > 0.0)) return false;
  const bool lo = bra.x * ket.x * ba.dmax[0] < ba.fp32_tau;
  return __ballot(!lo) == 0;
}
// adds the wave's FP32 quartets (n per lane) to the build's count: one atomic
// per wave; every lane of the wave must call it
__device__ __forceinline__ void count_fp32(const BuildArgs &ba, unsigned n) {
  if (!(ba.fp32_tau > 0.0)) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
  if (__lane_id() == 0 && n != 0) atomicAdd((unsigned long long *)ba.fp32_count, (unsigned long long)n);
}

// One contracted quartet held by one lane: primitive loop with Rys quadrature,
// contracted HRR, cartesian scale factors and the value cutoff.  Result
// v[((a*NB + b)*NC + c)*ND + d], components in internal order.  T is the
// arithmetic type of the 2-D integrals, the primitive accumulation and the
// HRR; the Rys argument, roots and weights are always FP64 (they stand where
// the reference evaluates its Boys function), and so is the result.
template <class T, int LA, int LB, int LC, int LD, bool KET_OUTER = false, bool FIXED_KET = false>
__device__ __forceinline__ void small_integrals(const PairRec &bra, const PairRec &ket, const PrimRec *prims,
                                                double xswitch, double cut,
                                                double (&v)[Small<LA, LB, LC, LD>::NABCD],
                                                const double *tab = rys_table + rys_offset[Small<LA, LB, LC, LD>::NR],
                                                int istride = rys_ncoef(Small<LA, LB, LC, LD>::NR)) {
  using S = Small<LA, LB, LC, LD>;
  constexpr int NE = S::NE, NF = S::NF, NR = S::NR, NI = S::LAB + 1, NK = S::LCD + 1;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND;
  constexpr int NBUF = S::NBUF;
  constexpr Range<LA, LA + LB> re{};
  constexpr Range<LC, LC + LD> rf{};

  T acc[NBUF];
#pragma unroll
  for (int i = 0; i < NE * NF; ++i) acc[i] = (T)0;

  // one primitive quartet into acc
  auto prim = [&](const PrimRec &P, const PrimRec &Q) {
    const double rs = nr_rsqrt(P.p + Q.p), ipq = rs * rs;  // 1/sqrt(p+q), 1/(p+q)
    const double pqx = P.Px - Q.Px, pqy = P.Py - Q.Py, pqz = P.Pz - Q.Pz;
    const double rho = P.p * Q.p * ipq;
    const double X = rho * (pqx * pqx + pqy * pqy + pqz * pqz);
    const double pref = kTwoPi52 * P.K * Q.K * rs;
    double u[NR], w[NR];
    rys_rule<NR>(X, xswitch, u, w, tab, istride);
    const T tp = (T)P.p, tq = (T)Q.p, tipq = (T)ipq;
    const T hp = (T)0.5 * nr_rcp(tp), hq = (T)0.5 * nr_rcp(tq);
    const T tpqx = (T)pqx, tpqy = (T)pqy, tpqz = (T)pqz;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const T t2 = (T)u[r];
      const T b00 = (T)0.5 * t2 * tipq;
      const T b10 = hp - hp * tq * t2 * tipq;
      const T b01 = hq - hq * tp * t2 * tipq;
      const T fa = tq * tipq * t2, fc = tp * tipq * t2;
      T gx[NI][NK], gy[NI][NK], gz[NI][NK];
      rys_2d<NI, NK>(gx, (T)P.PAx - fa * tpqx, (T)Q.PAx + fc * tpqx, b10, b01, b00, (T)1);
      rys_2d<NI, NK>(gy, (T)P.PAy - fa * tpqy, (T)Q.PAy + fc * tpqy, b10, b01, b00, (T)1);
      rys_2d<NI, NK>(gz, (T)P.PAz - fa * tpqz, (T)Q.PAz + fc * tpqz, b10, b01, b00, (T)(w[r] * pref));
#pragma unroll
      for (int e = 0; e < NE; ++e)
#pragma unroll
        for (int f = 0; f < NF; ++f)
          acc[e * NF + f] += gx[re.lx[e]][rf.lx[f]] * gy[re.ly[e]][rf.ly[f]] * gz[re.lz[e]][rf.lz[f]];
    }
  };
  if constexpr (KET_OUTER) {
    // the inner loop walks the bra pair's primitives (the pair a row tile
    // keeps fixed: every lane then reads the same records, through the
    // scalar cache)
    ConstPrimRec *cprims = const_view(prims);
    for (int iq = 0; iq < ket.nprim; ++iq) {
      const PrimRec Q = prims[ket.prim_off + iq];
      for (int ip = 0; ip < bra.nprim; ++ip) prim(load_rec(cprims, bra.prim_off + ip), Q);
    }
  } else {
    for (int ip = 0; ip < bra.nprim; ++ip) {
      const PrimRec P = prims[bra.prim_off + ip];
      // (ss|ss): the next ket primitive is loaded one iteration ahead, its
      // latency hidden behind this iteration's arithmetic (in the other
      // classes the extra registers cost more occupancy than the latency it
      // hides)
      constexpr bool ahead = LA + LB + LC + LD == 0;
      // a row tile's fixed pair is the ket here (uniform: the scalar cache)
      auto ket_rec = [&](int iq) -> PrimRec {
        if constexpr (FIXED_KET)
          return load_rec(const_view(prims), ket.prim_off + iq);
        else
          return prims[ket.prim_off + iq];
      };
      PrimRec Qn = ket_rec(0);
      for (int iq = 0; iq < ket.nprim; ++iq) {
        const PrimRec Q = ahead ? Qn : ket_rec(iq);
        if (ahead && iq + 1 < ket.nprim) Qn = ket_rec(iq + 1);
        prim(P, Q);
      }
    }
  }

  // contracted HRR: bra then ket, ping-pong between two register arrays; the
  // result ends in b1 when exactly one of LB, LD is odd
  T b1[NBUF];
  if constexpr (LB > 0) hrr_bra<LA, LB, NF>(acc, b1, (T)bra.ABx, (T)bra.ABy, (T)bra.ABz);
  if constexpr (LD > 0) {
    if constexpr (LB % 2)
      hrr_ket<LC, LD, NA * NB>(b1, acc, (T)ket.ABx, (T)ket.ABy, (T)ket.ABz);
    else
      hrr_ket<LC, LD, NA * NB>(acc, b1, (T)ket.ABx, (T)ket.ABy, (T)ket.ABz);
  }
  constexpr bool in_b1 = ((LB % 2) != 0) != ((LD % 2) != 0);

  // normalised block v[a][b][c][d] (internal order), value cutoff
  constexpr Range<LA, LA> ra{};
  constexpr Range<LB, LB> rb{};
  constexpr Range<LC, LC> rc{};
  constexpr Range<LD, LD> rd{};
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const double sc = cart_scale(ra.lx[a], ra.ly[a], ra.lz[a]) * cart_scale(rb.lx[b], rb.ly[b], rb.lz[b]) *
                            cart_scale(rc.lx[c], rc.ly[c], rc.lz[c]) * cart_scale(rd.lx[d], rd.ly[d], rd.lz[d]);
          const int idx = ((a * NB + b) * NC + c) * ND + d;
          const double x = (double)(in_b1 ? b1[idx] : acc[idx]) * sc;
          v[idx] = fabs(x) < cut ? 0.0 : x;
        }
}

// Lane-per-quartet kernel with direct digestion: every Fock element of the
// six blocks of a quartet is one atomic add to F.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void eri_small_kernel(ClassDev cd, BuildArgs ba) {
  using S = Small<LA, LB, LC, LD>;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND;
  constexpr Range<LA, LA> ra{};
  constexpr Range<LB, LB> rb{};
  constexpr Range<LC, LC> rc{};
  constexpr Range<LD, LD> rd{};

  // the block's quartets span segments [s_lo, s_hi]: two full searches, then
  // every lane searches only that short range
  __shared__ int s_lo, s_hi;
  const int64_t q0 = cd.q_begin + (int64_t)blockIdx.x * blockDim.x;
  const int64_t qlast = (q0 + blockDim.x < cd.q_end ? q0 + blockDim.x : cd.q_end) - 1;
  if (threadIdx.x == 0) s_lo = find_segment(cd, q0, 0, cd.nseg);
  if (threadIdx.x == 1) s_hi = find_segment(cd, qlast, 0, cd.nseg);
  __syncthreads();
  const int64_t q = q0 + threadIdx.x;
  if (q >= cd.q_end) return;
  int32_t ib, ik;
  quartet_of_segment(cd, find_segment(cd, q, s_lo, s_hi + 1), q, ib, ik);
  const PairRec bra = ba.pairs[ib];
  const PairRec ket = ba.pairs[ik];
  double res[S::NABCD];
  if (mixed_wave(cd, ba, bra, ket)) {
    const uint64_t m = __ballot(true);
    if (__lane_id() == __ffsll((unsigned long long)m) - 1)
      atomicAdd((unsigned long long *)ba.fp32_count, (unsigned long long)__popcll(m));
    small_integrals<float, LA, LB, LC, LD>(bra, ket, ba.prims, cd.xswitch, ba.value_cut, res);
  } else {
    small_integrals<double, LA, LB, LC, LD>(bra, ket, ba.prims, cd.xswitch, ba.value_cut, res);
  }

  // degeneracy weight of the shell quartet (A==B, C==D, AB==CD)
  double s = 1.0;
  if (bra.sameAB) s *= 0.5;
  if (ket.sameAB) s *= 0.5;
  if (cd.diag && ib == ik) s *= 0.5;

  int ga[NA], gb[NB], gc[NC], gd[ND];
#pragma unroll
  for (int a = 0; a < NA; ++a) ga[a] = bra.bfA + gamess_pos(ra.lx[a], ra.ly[a], ra.lz[a]);
#pragma unroll
  for (int b = 0; b < NB; ++b) gb[b] = bra.bfB + gamess_pos(rb.lx[b], rb.ly[b], rb.lz[b]);
#pragma unroll
  for (int c = 0; c < NC; ++c) gc[c] = ket.bfA + gamess_pos(rc.lx[c], rc.ly[c], rc.lz[c]);
#pragma unroll
  for (int d = 0; d < ND; ++d) gd[d] = ket.bfB + gamess_pos(rd.lx[d], rd.ly[d], rd.lz[d]);

  const double *Dm = ba.D, *Dk = ba.DK;  // Coulomb and exchange densities
  double *F = ba.F;
  // J: F[ab] += 4s sum_cd v D[cd], F[cd] += 4s sum_ab v D[ab]
  {
    double dcd[NC * ND];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int d = 0; d < ND; ++d) dcd[c * ND + d] = Dm[tri(gc[c], gd[d])];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NC * ND; ++k) t += res[(a * NB + b) * NC * ND + k] * dcd[k];
        if (t != 0.0) fock_add(F, ga[a], gb[b], 4.0 * s * t);
      }
  }
  {
    double dab[NA * NB];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) dab[a * NB + b] = Dm[tri(ga[a], gb[b])];
#pragma unroll
    for (int k = 0; k < NC * ND; ++k) {
      double t = 0.0;
#pragma unroll
      for (int ab = 0; ab < NA * NB; ++ab) t += res[ab * NC * ND + k] * dab[ab];
      if (t != 0.0) fock_add(F, gc[k / ND], gd[k % ND], 4.0 * s * t);
    }
  }
  const double sk = s * ba.k_scale;
  // K: F[ac] -= s sum_bd v D[bd];  F[bd] -= s sum_ac v D[ac]
  {
    double dbd[NB * ND], dac[NA * NC];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int d = 0; d < ND; ++d) dbd[b * ND + d] = Dk[tri(gb[b], gd[d])];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) dac[a * NC + c] = Dk[tri(ga[a], gc[c])];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int d = 0; d < ND; ++d) t += res[((a * NB + b) * NC + c) * ND + d] * dbd[b * ND + d];
        if (t != 0.0) fock_add(F, ga[a], gc[c], -sk * t);
      }
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int c = 0; c < NC; ++c) t += res[((a * NB + b) * NC + c) * ND + d] * dac[a * NC + c];
        if (t != 0.0) fock_add(F, gb[b], gd[d], -sk * t);
      }
  }
  // K: F[ad] -= s sum_bc v D[bc];  F[bc] -= s sum_ad v D[ad]
  {
    double dbc[NB * NC], dad[NA * ND];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c) dbc[b * NC + c] = Dk[tri(gb[b], gc[c])];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int d = 0; d < ND; ++d) dad[a * ND + d] = Dk[tri(ga[a], gd[d])];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int c = 0; c < NC; ++c) t += res[((a * NB + b) * NC + c) * ND + d] * dbc[b * NC + c];
        if (t != 0.0) fock_add(F, ga[a], gd[d], -sk * t);
      }
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int d = 0; d < ND; ++d) t += res[((a * NB + b) * NC + c) * ND + d] * dad[a * ND + d];
        if (t != 0.0) fock_add(F, gb[b], gc[c], -sk * t);
      }
  }
}

__device__ __forceinline__ int gpos_of(int L, int k);

__device__ __forceinline__ void lds_add(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Intervals of the N-root table that the rule reads below xswitch (the
// first ceil(xswitch / h)), and the doubles their LDS copy takes: the copy
// pads each interval to an odd stride, so lanes reading different intervals
// spread over the LDS banks.
__host__ __device__ inline int64_t rys_table_intervals(int n, double xswitch) {
  int64_t nint = (int64_t)ceil(xswitch * rys_inv_h(n));
  const double xmax = n == 1 ? 39.0 : n == 2 ? 46.5 : n == 3 ? 53.5 : n == 4 ? 60.0 : n == 5 ? 66.0 : n == 6 ? 72.0 : 77.5;
  const int64_t cap = (int64_t)(xmax * rys_inv_h(n) + 0.5);
  if (nint > cap) nint = cap;
  if (nint < 0) nint = 0;
  return nint;
}
__host__ __device__ constexpr int rys_lds_stride(int n) { return rys_ncoef(n) + 1; }
__host__ __device__ inline int64_t rys_table_used(int n, double xswitch) {
  return rys_table_intervals(n, xswitch) * rys_lds_stride(n);
}

// Shape of a row tile of class (LA LB | LC LD).  The tile fixes one shell
// pair (the bra pair, or with TR the ket pair): its shells give the rows of
// the on-chip Fock accumulators.  The other ("moving") pair varies per
// quartet; its shells' angular momenta M1, M2 give the column spaces, each the
// compact list of all functions of that angular momentum (one space when
// M1 == M2).
template <int LA, int LB, int LC, int LD, bool TR>
struct RowShape {
  static constexpr int L1 = TR ? LC : LA, L2 = TR ? LD : LB;
  static constexpr int M1 = TR ? LA : LC, M2 = TR ? LB : LD;
  static constexpr int N1 = ncart(L1), N2 = ncart(L2), NROW = N1 + N2, NFIX = N1 * N2;
  static constexpr int NM1 = ncart(M1), NM2 = ncart(M2);
  // a class whose bra and ket have the same shape is diagonal: no TR form
  static constexpr bool exists = !(TR && LA == LC && LB == LD);
};

template <int LA, int LB, int LC, int LD, bool TR>
__host__ __device__ inline int64_t row_columns(const int32_t *ncol) {
  using R = RowShape<LA, LB, LC, LD, TR>;
  return (int64_t)ncol[R::M1] + (R::M1 == R::M2 ? 0 : ncol[R::M2]);
}

// x plus the value of lane CTRL-shifted x (DPP), lanes outside the row or
// masked off adding 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, 0xf, false);
  return x + __hiloint2double(hi, lo);
}
// Sums over the 64 lanes of a wave of K values per lane, added to LDS:
// dst[idx(k)] += sum over lanes of x[k].  Two halvings first exchange whole
// half-waves and rows between the registers of two values
// (v_permlane32_swap: upper 32 lanes of the first with the lower 32 of the
// second; v_permlane16_swap: odd rows of the first with even rows of the
// second), so one add leaves value j in one half (row) and value j + h in
// the other: the number of values a lane carries halves at each step.  Row
// r then holds values (r & 1 ? K2 + j : j) + (r >> 1 ? K1 : 0), j < K2,
// each summed over its four lanes of the row position; prefix doubling
// inside the row (row_shr 1, 2, 4, 8) completes the sums in lane 15 of the
// row, which adds its K2 totals to LDS (four rows, four distinct addresses).
__device__ __forceinline__ double halves_add(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                   false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                   false);
  return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ double rows_add(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                   false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                   false);
  return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
template <int K, class Idx>
__device__ __forceinline__ void wave_sum_to_lds(const double (&x)[K], double *dst, Idx idx) {
  constexpr int K1 = (K + 1) / 2, K2 = (K1 + 1) / 2;
  double a[K1], b[K2];
#pragma unroll
  for (int j = 0; j < K1; ++j) a[j] = halves_add(x[j], j + K1 < K ? x[j + K1] : 0.0);
#pragma unroll
  for (int j = 0; j < K2; ++j) b[j] = rows_add(a[j], j + K2 < K1 ? a[j + K2] : 0.0);
  const int lane = (int)__lane_id(), r = lane >> 4;
#pragma unroll
  for (int j = 0; j < K2; ++j) {
    double t = b[j];
    t = dpp_add<0x111, 0xf>(t);
    t = dpp_add<0x112, 0xf>(t);
    t = dpp_add<0x114, 0xf>(t);
    t = dpp_add<0x118, 0xf>(t);
    const int s = (r & 1) ? K2 + j : j;
    const int k = s + ((r >> 1) ? K1 : 0);
    if ((lane & 15) == 15 && s < K1 && k < K && t != 0.0) lds_add(&dst[idx(k)], t);
  }
}

// Row-tiled lane-per-quartet kernel.  One workgroup owns one tile: a run of
// quartets that share the fixed pair.  Four of the six digestion blocks --
// the exchange blocks K(AC), K(AD), K(BC), K(BD) -- land in Fock rows of the
// fixed pair's shells; they are accumulated in LDS, rows[r][column], and
// added to F once per tile.  The Coulomb block of the fixed pair is summed in
// registers over the tile and reduced across the workgroup at the end.  Only
// the Coulomb block of the moving pair goes to F per quartet.
// Threads of a row-tiled lane-per-quartet workgroup: one copy of the
// tile's LDS rows and Rys table serves all of them, so larger workgroups
// hold more waves per CU where the LDS limits them.  Measured on C60 (256 ->
// 512 / 1024 threads): (pp|ds) 4.61 -> 4.18 ms, (pp|pp) 2.18 -> 1.97, (ps|ps)
// 6.09 -> 5.83 (1024); the other lane classes are as fast or faster at 256.
template <int LA, int LB, int LC, int LD>
constexpr int row_threads() {
  if constexpr (LA == 1 && LB == 1 && LC == 2 && LD == 0) return 512;
  if constexpr (LA == 1 && LB == 1 && LC == 1 && LD == 1) return 512;
  if constexpr (LA == 1 && LB == 0 && LC == 1 && LD == 0) return 1024;
  return 256;
}

template <int LA, int LB, int LC, int LD, bool TR>
__global__ __launch_bounds__((row_threads<LA, LB, LC, LD>())) void eri_row_kernel(ClassDev cd, BuildArgs ba) {
  using S = Small<LA, LB, LC, LD>;
  using R = RowShape<LA, LB, LC, LD, TR>;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND;
  constexpr int N1 = R::N1, N2 = R::N2, NFIX = R::NFIX, NROW = R::NROW;
  constexpr Range<LA, LA> ra{};
  constexpr Range<LB, LB> rb{};
  constexpr Range<LC, LC> rc{};
  constexpr Range<LD, LD> rd{};
  constexpr Range<R::L1, R::L1> r1{};
  constexpr Range<R::L2, R::L2> r2{};
  constexpr Range<R::M1, R::M1> rm1{};
  constexpr Range<R::M2, R::M2> rm2{};
  extern __shared__ double lds[];  // jfix[NFIX], rows[NROW][ncols], then the Rys table
  double *jfix_s = lds, *rows = lds + NFIX;
  const int ncols = (int)row_columns<LA, LB, LC, LD, TR>(ba.ncol);
  const int off2 = R::M1 == R::M2 ? 0 : ba.ncol[R::M1];
  // the class's Rys table below the switch, staged once per workgroup with
  // intervals padded to an odd stride
  double *tab = rows + NROW * ncols;
  {
    constexpr int IN = rys_ncoef(S::NR), LS = rys_lds_stride(S::NR);
    const int ntab = (int)rys_table_intervals(S::NR, cd.xswitch) * IN;
    const double *src = rys_table + rys_offset[S::NR];
    for (int i = threadIdx.x; i < ntab; i += blockDim.x) tab[(i / IN) * LS + i % IN] = src[i];
  }

  const int tile = cd.t_begin + (int)blockIdx.x;
  const int64_t q0 = cd.tile_q[tile] > cd.q_begin ? cd.tile_q[tile] : cd.q_begin;
  const int64_t q1 = cd.tile_q[tile + 1] < cd.q_end ? cd.tile_q[tile + 1] : cd.q_end;
  const int seg0 = cd.tile_seg[tile];
  const int32_t ifix = cd.seg_bra[seg0];
  const PairRec fix = load_rec(const_view(ba.pairs), ifix);
  for (int i = threadIdx.x; i < NFIX + NROW * ncols; i += blockDim.x) lds[i] = 0.0;

  const double *Dm = ba.D, *Dk = ba.DK;  // Coulomb and exchange densities
  double *F = ba.F;
  // the fixed pair's density block and Coulomb partial sums
  double dfix[NFIX], jfix[NFIX];
#pragma unroll
  for (int i = 0; i < N1; ++i)
#pragma unroll
    for (int j = 0; j < N2; ++j) {
      dfix[i * N2 + j] = Dm[tri(fix.bfA + gamess_pos(r1.lx[i], r1.ly[i], r1.lz[i]),
                                fix.bfB + gamess_pos(r2.lx[j], r2.ly[j], r2.lz[j]))];
      jfix[i * N2 + j] = 0.0;
    }
  const double sfix = fix.sameAB ? 0.5 : 1.0;
  __syncthreads();

  unsigned nlow = 0;  // quartets this lane evaluated in FP32
  int sg = seg0;      // segment of the lane's quartet (its quartets ascend)
  for (int64_t qb = q0; qb < q1; qb += blockDim.x) {
    const int64_t q = qb + threadIdx.x;
    if (q >= q1) continue;
    while (sg + 1 < cd.nseg && cd.seg_start[sg + 1] <= q) ++sg;
    const int32_t imov = cd.ket_list[cd.seg_ket0[sg] + (int32_t)(q - cd.seg_start[sg])];
    const PairRec mov = ba.pairs[imov];
    const PairRec &bra = TR ? mov : fix;
    const PairRec &ket = TR ? fix : mov;
    double v[S::NABCD];
    if (mixed_wave(cd, ba, bra, ket)) {
      ++nlow;
      small_integrals<float, LA, LB, LC, LD, !TR, TR>(bra, ket, ba.prims, cd.xswitch, ba.value_cut, v, tab,
                                             rys_lds_stride(S::NR));
    } else {
      small_integrals<double, LA, LB, LC, LD, !TR, TR>(bra, ket, ba.prims, cd.xswitch, ba.value_cut, v, tab,
                                              rys_lds_stride(S::NR));
    }

    double s = sfix;
    if (mov.sameAB) s *= 0.5;
    if (cd.diag && ifix == imov) s *= 0.5;
    int ga[NA], gb[NB], gc[NC], gd[ND];
#pragma unroll
    for (int a = 0; a < NA; ++a) ga[a] = bra.bfA + gamess_pos(ra.lx[a], ra.ly[a], ra.lz[a]);
#pragma unroll
    for (int b = 0; b < NB; ++b) gb[b] = bra.bfB + gamess_pos(rb.lx[b], rb.ly[b], rb.lz[b]);
#pragma unroll
    for (int c = 0; c < NC; ++c) gc[c] = ket.bfA + gamess_pos(rc.lx[c], rc.ly[c], rc.lz[c]);
#pragma unroll
    for (int d = 0; d < ND; ++d) gd[d] = ket.bfB + gamess_pos(rd.lx[d], rd.ly[d], rd.lz[d]);
    // compact columns of the moving pair's functions
    int cm1[R::NM1], cm2[R::NM2];
#pragma unroll
    for (int i = 0; i < R::NM1; ++i) cm1[i] = mov.cA + gamess_pos(rm1.lx[i], rm1.ly[i], rm1.lz[i]);
#pragma unroll
    for (int i = 0; i < R::NM2; ++i) cm2[i] = off2 + mov.cB + gamess_pos(rm2.lx[i], rm2.ly[i], rm2.lz[i]);

    // Coulomb: the fixed pair's block in registers, the moving pair's to F
    if constexpr (!TR) {
      double dcd[NC * ND];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) dcd[c * ND + d] = Dm[tri(gc[c], gd[d])];
#pragma unroll
      for (int ab = 0; ab < NA * NB; ++ab) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NC * ND; ++k) t += v[ab * NC * ND + k] * dcd[k];
        jfix[ab] += 4.0 * s * t;
      }
#pragma unroll
      for (int k = 0; k < NC * ND; ++k) {
        double t = 0.0;
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) t += v[ab * NC * ND + k] * dfix[ab];
        if (t != 0.0) fock_add(F, gc[k / ND], gd[k % ND], 4.0 * s * t);
      }
    } else {
      double dab[NA * NB];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) dab[a * NB + b] = Dm[tri(ga[a], gb[b])];
#pragma unroll
      for (int k = 0; k < NC * ND; ++k) {
        double t = 0.0;
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) t += v[ab * NC * ND + k] * dab[ab];
        jfix[k] += 4.0 * s * t;
      }
#pragma unroll
      for (int ab = 0; ab < NA * NB; ++ab) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NC * ND; ++k) t += v[ab * NC * ND + k] * dfix[k];
        if (t != 0.0) fock_add(F, ga[ab / NB], gb[ab % NB], 4.0 * s * t);
      }
    }
    const double sk = s * ba.k_scale;
    // exchange: K(AC) and K(BD)
    {
      double dbd[NB * ND], dac[NA * NC];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int d = 0; d < ND; ++d) dbd[b * ND + d] = Dk[tri(gb[b], gd[d])];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c) dac[a * NC + c] = Dk[tri(ga[a], gc[c])];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double t = 0.0;
#pragma unroll
          for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int d = 0; d < ND; ++d) t += v[((a * NB + b) * NC + c) * ND + d] * dbd[b * ND + d];
          if (t != 0.0) lds_add(TR ? &rows[c * ncols + cm1[a]] : &rows[a * ncols + cm1[c]], -sk * t);
        }
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          double t = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int c = 0; c < NC; ++c) t += v[((a * NB + b) * NC + c) * ND + d] * dac[a * NC + c];
          if (t != 0.0) lds_add(TR ? &rows[(NC + d) * ncols + cm2[b]] : &rows[(NA + b) * ncols + cm2[d]], -sk * t);
        }
    }
    // exchange: K(AD) and K(BC)
    {
      double dbc[NB * NC], dad[NA * ND];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c) dbc[b * NC + c] = Dk[tri(gb[b], gc[c])];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int d = 0; d < ND; ++d) dad[a * ND + d] = Dk[tri(ga[a], gd[d])];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          double t = 0.0;
#pragma unroll
          for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) t += v[((a * NB + b) * NC + c) * ND + d] * dbc[b * NC + c];
          if (t != 0.0) lds_add(TR ? &rows[(NC + d) * ncols + cm1[a]] : &rows[a * ncols + cm2[d]], -sk * t);
        }
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double t = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int d = 0; d < ND; ++d) t += v[((a * NB + b) * NC + c) * ND + d] * dad[a * ND + d];
          if (t != 0.0) lds_add(TR ? &rows[c * ncols + cm2[b]] : &rows[(NA + b) * ncols + cm1[c]], -sk * t);
        }
    }
  }

  count_fp32(ba, nlow);
  // the fixed pair's Coulomb block: reduce over each wave, then over the waves
  wave_sum_to_lds<NFIX>(jfix, jfix_s, [](int k) { return k; });
  __syncthreads();
  for (int i = threadIdx.x; i < NFIX; i += blockDim.x) {
    const double x = jfix_s[i];
    if (x != 0.0) fock_add(F, fix.bfA + gpos_of(R::L1, i / N2), fix.bfB + gpos_of(R::L2, i % N2), x);
  }
  for (int i = threadIdx.x; i < NROW * ncols; i += blockDim.x) {
    const double x = rows[i];
    if (x == 0.0) continue;
    const int r = i / ncols, k = i - r * ncols;
    const int rbf = r < N1 ? fix.bfA + gpos_of(R::L1, r) : fix.bfB + gpos_of(R::L2, r - N1);
    const int cbf = (R::M1 == R::M2 || k < off2) ? ba.col_bf[R::M1][k] : ba.col_bf[R::M2][k - off2];
    fock_add(F, rbf, cbf, x);
  }
}

// ====================================================== wave-split row path
// The row-tiled form for classes whose (e0|f0) block is too large for the
// registers of one lane: S parts share each quartet, part p accumulating one
// slice of the block.  The slices are chosen so that each part's HRR stays
// inside its slice:
//  * columns (LD = 0, no ket HRR): a run of the columns f; the bra HRR runs
//    column by column;
//  * rows (LB = 0, no bra HRR): a run of the rows e; the ket HRR runs row by
//    row;
//  * ket cone (LD = 1): one component c of shell C and the four columns
//    c, c + 1x, c + 1y, c + 1z its ket HRR reads; the ket step
//    (e|c,1i) = (e|c+1i) + CD_i (e|c) runs first, then the bra HRR on the
//    three columns (c, d).
// Every part repeats the Rys rule and the 2-D recursion of each primitive
// quartet (the compiler drops the table elements its slice never reads) and
// keeps its slice, its HRR and its part of the digestion in registers: no
// LDS traffic in the primitive loop.  The exchange blocks land in the
// workgroup's LDS rows as in eri_row_kernel; a block summed over the split
// index gets one LDS add per part.  A wave runs S / W of the parts one after
// the other over its quartets, so that W waves serve one run of 64 quartets.

enum SplitSide { SPLIT_NONE = 0, SPLIT_COLS = 1, SPLIT_ROWS = 2, SPLIT_CONE = 3 };


// largest HRR intermediate of a slice: NEP rows e and NFP columns f of
// (e0|f0), bra levels on the NFP columns, ket levels on the slice's rows ab
__host__ __device__ constexpr int split_buffer(int LA, int LB, int LC, int LD, int NEP, int NFP) {
  int m = NEP * NFP;
  for (int j = 1; j <= LB; ++j) {
    const int s = nrange(LA, LA + LB - j) * ncart(j) * NFP;
    m = s > m ? s : m;
  }
  const int nab = LB == 0 ? NEP : ncart(LA) * ncart(LB);
  for (int j = 1; j <= LD; ++j) {
    const int s = nab * nrange(LC, LC + LD - j) * ncart(j);
    m = s > m ? s : m;
  }
  return m;
}

// Parts S and waves per run W of the split form of a class (S = 0: none),
// and for the ket cone the directions of D per part (3: all; 1: one, the
// part is a (c, direction) pair).  Chosen so that the slice, the tables and
// the HRR levels of a part fit the registers of two waves per SIMD (8-wave
// workgroups) or three (12 waves).
template <int LA, int LB, int LC, int LD>
constexpr int split_cone_dirs() {
  // (dp|pp), (dp|dp): 16 rows e, four columns per part; the others (25 or
  // 31 rows) two columns per part
  return LA == 2 && LB == 1 ? 3 : 1;
}
template <int LA, int LB, int LC, int LD>
constexpr int split_parts() {
  if constexpr (LA == 2 && LB == 1 && LC == 2 && LD == 0) return 2;  // (dp|ds): 3 of the 6 d columns
  if constexpr (LA == 2 && LB == 2 && LC == 2 && LD == 0) return 6;  // (dd|ds): one d column
  if constexpr (LA == 1 && LB == 0 && LC == 2 && LD == 2) return 3;  // (ps|dd): one p row
  if constexpr (LA == 2 && LB == 1 && LC == 1 && LD == 1) return 3;  // (dp|pp): cone of one p of C
  if constexpr (LA == 2 && LB == 1 && LC == 2 && LD == 1) return 6;  // (dp|dp): cone of one d of C
  if constexpr (LA == 2 && LB == 2 && LC == 2 && LD == 1) return 18;  // (dd|dp): one d of C, one direction
  if constexpr (LA == 2 && LB == 2 && LC == 1 && LD == 1) return 9;   // (dd|pp): one p of C, one direction
  // classes with an f shell (the C-N chain)
  if constexpr (LA == 2 && LB == 1 && LC == 3 && LD == 0) return 10;  // (dp|fs): one f column
  if constexpr (LA == 2 && LB == 2 && LC == 3 && LD == 0) return 10;  // (dd|fs): one f column
  if constexpr (LA == 3 && LB == 1 && LC == 2 && LD == 0) return 6;   // (fp|ds): one d column
  if constexpr (LA == 3 && LB == 1 && LC == 3 && LD == 0) return 10;  // (fp|fs): one f column
  if constexpr (LA == 3 && LB == 1 && LC == 1 && LD == 1) return 9;   // (fp|pp): one p of C, one direction
  if constexpr (LA == 3 && LB == 1 && LC == 2 && LD == 1) return 18;  // (fp|dp): one d of C, one direction
  if constexpr (LA == 2 && LB == 2 && LC == 3 && LD == 1) return 30;  // (dd|fp): one f of C, one direction
  if constexpr (LA == 3 && LB == 1 && LC == 3 && LD == 1) return 30;  // (fp|fp): one f of C, one direction
  return 0;
}
template <int LA, int LB, int LC, int LD>
constexpr int split_waves() {
  constexpr int S = split_parts<LA, LB, LC, LD>();
  if constexpr (LA == 2 && LB == 1 && LC == 1 && LD == 1) return 1;  // (dp|pp): the 3 parts in turn
  if constexpr (LA == 2 && LB == 1 && LC == 2 && LD == 1) return 2;  // (dp|dp): 3 parts per wave
  if constexpr (LA == 2 && LB == 2 && LC == 2 && LD == 1) return 2;  // (dd|dp): 9 parts per wave
  if constexpr (LA == 2 && LB == 2 && LC == 1 && LD == 1) return 1;  // (dd|pp): 9 parts per wave
  if constexpr (LA == 3 && LB == 1 && LC == 1 && LD == 1) return 1;  // (fp|pp): 9 parts per wave
  if constexpr ((LA == 3 || LC == 3) && S % 2 == 0) return 2;  // f classes: S/2 parts per wave
  return S;
}

template <int LA, int LB, int LC, int LD>
struct SplitShape {
  static constexpr int S = split_parts<LA, LB, LC, LD>();
  static constexpr int W = split_waves<LA, LB, LC, LD>();
  static constexpr int NE = nrange(LA, LA + LB), NF = nrange(LC, LC + LD);
  static constexpr int SIDE = S == 0 ? SPLIT_NONE
                              : LD == 0 ? SPLIT_COLS
                              : LB == 0 ? SPLIT_ROWS
                              : LD == 1 ? SPLIT_CONE
                                        : SPLIT_NONE;
  // ket cone: directions of D per part, parts per component c
  static constexpr int NDI = SIDE == SPLIT_CONE ? split_cone_dirs<LA, LB, LC, LD>() : ncart(LD);
  static constexpr int PPC = 3 / (NDI > 0 ? NDI : 1);
  static constexpr bool exists = SIDE != SPLIT_NONE && W > 0 && S % W == 0 &&
                                 (SIDE == SPLIT_COLS   ? NF % S == 0
                                  : SIDE == SPLIT_ROWS ? NE % S == 0
                                                       : (NDI == 1 || NDI == 3) && S == ncart(LC) * PPC);
  // rows e and columns f of (e0|f0) one part accumulates
  static constexpr int NEP = SIDE == SPLIT_ROWS ? NE / S : NE;
  static constexpr int NFP = SIDE == SPLIT_COLS ? NF / S : (SIDE == SPLIT_CONE ? 1 + NDI : NF);
  // waves per workgroup: W per run of 64 quartets, NBAT runs; 8 waves when W
  // divides it, else 12 (two or three waves per SIMD)
  static constexpr int NBAT = !exists ? 1 : (8 % W == 0 ? 8 / W : 12 / W);
  static constexpr int NT = 64 * W * NBAT;
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  // slice of the block after the HRR: rows a (all but in the rows split),
  // columns c (one per part in the cone, all in the rows split), columns d
  // (the part's directions in the cone, else all)
  static constexpr int NAP = SIDE == SPLIT_ROWS ? NEP : ncart(LA);
  static constexpr int NCP = SIDE == SPLIT_COLS ? NFP : (SIDE == SPLIT_CONE ? 1 : ncart(LC));
  static constexpr int NB = ncart(LB), ND = ncart(LD), NDP = NDI;
  static constexpr int NV = NAP * NB * NCP * NDP;
  static constexpr int NBUF =
      SIDE == SPLIT_CONE ? split_buffer(LA, LB, LC, 0, NEP, NDP) : split_buffer(LA, LB, LC, LD, NEP, NFP);
  // first component of the slice in shell A (rows split), C (columns, cone), D (cone)
  template <int PART>
  static constexpr int a0() { return SIDE == SPLIT_ROWS ? PART * NEP : 0; }
  template <int PART>
  static constexpr int c0() { return SIDE == SPLIT_COLS ? PART * NFP : (SIDE == SPLIT_CONE ? PART / PPC : 0); }
  template <int PART>
  static constexpr int d0() { return SIDE == SPLIT_CONE ? (PART % PPC) * NDI : 0; }
};

// column j of part PART: its index in range(LC, LC + LD) (in the cone: c,
// then c + 1i for the part's directions i)
template <int LA, int LB, int LC, int LD, int PART>
constexpr int split_col(int j) {
  using P = SplitShape<LA, LB, LC, LD>;
  if (P::SIDE == SPLIT_COLS) return PART * P::NFP + j;
  if (P::SIDE != SPLIT_CONE) return j;
  constexpr Range<LC, LC> rc{};
  constexpr int c = P::template c0<PART>();
  if (j == 0) return c;
  const int i = P::template d0<PART>() + j - 1;
  return ncart(LC) + cidx(rc.lx[c] + (i == 0), rc.ly[c] + (i == 1), rc.lz[c] + (i == 2));
}

// One contracted quartet, slice PART: the part's rows and columns of (e0|f0)
// over the primitive loop, then the HRR of the slice, cartesian scales and
// the value cutoff.  Result v[((a*NB + b)*NCP + c)*NDP + d] with a, c, d
// relative to the slice's first component.
template <int LA, int LB, int LC, int LD, int PART, bool KET_OUTER>
__device__ __forceinline__ void split_integrals(const PairRec &bra, const PairRec &ket, const PrimRec *prims,
                                                double xswitch, double cut,
                                                double (&v)[SplitShape<LA, LB, LC, LD>::NV], const double *tab,
                                                int istride) {
  using P = SplitShape<LA, LB, LC, LD>;
  constexpr int NR = P::NR, NI = LA + LB + 1, NK = LC + LD + 1;
  constexpr int NEP = P::NEP, NFP = P::NFP, NBUF = P::NBUF;
  constexpr int E0 = P::template a0<PART>();
  constexpr Range<LA, LA + LB> re{};
  constexpr Range<LC, LC + LD> rf{};

  double acc[NBUF > NEP * NFP ? NBUF : NEP * NFP];
#pragma unroll
  for (int i = 0; i < NEP * NFP; ++i) acc[i] = 0.0;

  auto prim = [&](const PrimRec &Pp, const PrimRec &Q) {
    const double rs = nr_rsqrt(Pp.p + Q.p), ipq = rs * rs;
    const double pqx = Pp.Px - Q.Px, pqy = Pp.Py - Q.Py, pqz = Pp.Pz - Q.Pz;
    const double X = Pp.p * Q.p * ipq * (pqx * pqx + pqy * pqy + pqz * pqz);
    const double pref = kTwoPi52 * Pp.K * Q.K * rs;
    double u[NR], w[NR];
    rys_rule<NR>(X, xswitch, u, w, tab, istride);
    const double hp = 0.5 * nr_rcp(Pp.p), hq = 0.5 * nr_rcp(Q.p);
    const double qi = Q.p * ipq, pi = Pp.p * ipq;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double t2 = u[r];
      const double b00 = 0.5 * t2 * ipq;
      const double fa = qi * t2, fc = pi * t2;
      const double b10 = hp - hp * fa, b01 = hq - hq * fc;
      double gx[NI][NK], gy[NI][NK], gz[NI][NK];
      rys_2d<NI, NK>(gx, Pp.PAx - fa * pqx, Q.PAx + fc * pqx, b10, b01, b00, 1.0);
      rys_2d<NI, NK>(gy, Pp.PAy - fa * pqy, Q.PAy + fc * pqy, b10, b01, b00, 1.0);
      rys_2d<NI, NK>(gz, Pp.PAz - fa * pqz, Q.PAz + fc * pqz, b10, b01, b00, w[r] * pref);
#pragma unroll
      for (int e = 0; e < NEP; ++e)
#pragma unroll
        for (int f = 0; f < NFP; ++f) {
          const int ee = E0 + e, ff = split_col<LA, LB, LC, LD, PART>(f);
          acc[e * NFP + f] += gx[re.lx[ee]][rf.lx[ff]] * gy[re.ly[ee]][rf.ly[ff]] * gz[re.lz[ee]][rf.lz[ff]];
        }
    }
  };
  // the fixed pair's primitives (the inner loop) through the scalar cache
  ConstPrimRec *cprims = const_view(prims);
  if constexpr (KET_OUTER) {
    for (int iq = 0; iq < ket.nprim; ++iq) {
      const PrimRec Q = prims[ket.prim_off + iq];
      for (int ip = 0; ip < bra.nprim; ++ip) prim(load_rec(cprims, bra.prim_off + ip), Q);
    }
  } else {
    for (int ip = 0; ip < bra.nprim; ++ip) {
      const PrimRec Pp = prims[bra.prim_off + ip];
      for (int iq = 0; iq < ket.nprim; ++iq) prim(Pp, load_rec(cprims, ket.prim_off + iq));
    }
  }

  // HRR of the slice
  constexpr int NABP = LB == 0 ? NEP : ncart(LA) * ncart(LB);
  double b1[NBUF];
  // where the last HRR level left the slice
  constexpr bool in_b1 =
      P::SIDE == SPLIT_CONE ? LB % 2 == 0 : ((LB % 2) != 0) != ((LD % 2) != 0);
  if constexpr (P::SIDE == SPLIT_CONE) {
    // ket step into b1 as [e][d], then the bra levels on those columns
    constexpr int NDP = P::NDP, D0 = P::template d0<PART>();
#pragma unroll
    for (int e = 0; e < NEP; ++e)
#pragma unroll
      for (int j = 0; j < NDP; ++j) {
        const double cd = D0 + j == 0 ? ket.ABx : (D0 + j == 1 ? ket.ABy : ket.ABz);
        b1[e * NDP + j] = acc[e * NFP + 1 + j] + cd * acc[e * NFP];
      }
    hrr_bra<LA, LB, NDP>(b1, acc, bra.ABx, bra.ABy, bra.ABz);
  } else {
    if constexpr (LB > 0) hrr_bra<LA, LB, NFP>(acc, b1, bra.ABx, bra.ABy, bra.ABz);
    if constexpr (LD > 0) {
      if constexpr (LB % 2)
        hrr_ket<LC, LD, NABP>(b1, acc, ket.ABx, ket.ABy, ket.ABz);
      else
        hrr_ket<LC, LD, NABP>(acc, b1, ket.ABx, ket.ABy, ket.ABz);
    }
  }

  constexpr int A0 = P::template a0<PART>(), C0 = P::template c0<PART>(), D0 = P::template d0<PART>();
  constexpr int NAP = P::NAP, NB = P::NB, NCP = P::NCP, NDP = P::NDP;
  constexpr Range<LA, LA> ra{};
  constexpr Range<LB, LB> rb{};
  constexpr Range<LC, LC> rc{};
  constexpr Range<LD, LD> rd{};
#pragma unroll
  for (int a = 0; a < NAP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int c = 0; c < NCP; ++c)
#pragma unroll
        for (int d = 0; d < NDP; ++d) {
          const int aa = A0 + a, cc = C0 + c, dd = D0 + d;
          const double sc = cart_scale(ra.lx[aa], ra.ly[aa], ra.lz[aa]) * cart_scale(rb.lx[b], rb.ly[b], rb.lz[b]) *
                            cart_scale(rc.lx[cc], rc.ly[cc], rc.lz[cc]) * cart_scale(rd.lx[dd], rd.ly[dd], rd.lz[dd]);
          const int idx = ((a * NB + b) * NCP + c) * NDP + d;
          const double x = (in_b1 ? b1[idx] : acc[idx]) * sc;
          v[idx] = fabs(x) < cut ? 0.0 : x;
        }
}

constexpr int64_t kSplitLdsBytes = 160 * 1024;

// Quartets of one part (slice PART) of a wave-split row tile: the run of 64
// quartets qb + bat*64 + lane of every step of NBAT runs.
template <int LA, int LB, int LC, int LD, bool TR, int PART>
__device__ __forceinline__ void split_part(const ClassDev &cd, const BuildArgs &ba, const PairRec &fix, int32_t ifix,
                                           int64_t q0, int64_t q1, int seg0, int bat, double *jfix_s,
                                           const double *dfix_s, double *rows, int ncols, int off2,
                                           const double *tab) {
  using P = SplitShape<LA, LB, LC, LD>;
  using R = RowShape<LA, LB, LC, LD, TR>;
  constexpr int NA = ncart(LA), NB = P::NB, NC = ncart(LC), ND = P::ND, NDP = P::NDP;
  constexpr int NAP = P::NAP, NCP = P::NCP;
  constexpr int A0 = P::template a0<PART>(), C0 = P::template c0<PART>(), D0 = P::template d0<PART>();
  constexpr Range<LA, LA> ra{};
  constexpr Range<LB, LB> rb{};
  constexpr Range<LC, LC> rc{};
  constexpr Range<LD, LD> rd{};
  // the moving pair's shells: M1 (its slice when it is the split one), M2
  constexpr int NM1 = TR ? NAP : NCP, M10 = TR ? A0 : C0, NM2 = TR ? NB : NDP, M20 = TR ? 0 : D0;
  constexpr Range<R::M1, R::M1> rm1{};
  constexpr Range<R::M2, R::M2> rm2{};
  // this part's share of the fixed pair's Coulomb block
  constexpr int NJ = TR ? NCP * NDP : NAP * NB;
  const double *Dm = ba.D, *Dk = ba.DK;
  double *F = ba.F;
  const double sfix = fix.sameAB ? 0.5 : 1.0;
  const int lane = threadIdx.x & 63;
  int sg = seg0;
  for (int64_t qb = q0; qb < q1; qb += P::NBAT * 64) {
    const int64_t q = qb + bat * 64 + lane;
    double jf[NJ];
#pragma unroll
    for (int k = 0; k < NJ; ++k) jf[k] = 0.0;
    if (q < q1) {
      while (sg + 1 < cd.nseg && cd.seg_start[sg + 1] <= q) ++sg;
      const int32_t imov = cd.ket_list[cd.seg_ket0[sg] + (int32_t)(q - cd.seg_start[sg])];
      const PairRec mov = ba.pairs[imov];
      const PairRec &bra = TR ? mov : fix;
      const PairRec &ket = TR ? fix : mov;
      double v[P::NV];
      split_integrals<LA, LB, LC, LD, PART, !TR>(bra, ket, ba.prims, cd.xswitch, ba.value_cut, v, tab,
                                                rys_lds_stride(P::NR));
      double s = sfix;
      if (mov.sameAB) s *= 0.5;
      if (cd.diag && ifix == imov) s *= 0.5;
      int ga[NAP], gb[NB], gc[NCP], gd[NDP];
#pragma unroll
      for (int a = 0; a < NAP; ++a) ga[a] = bra.bfA + gamess_pos(ra.lx[A0 + a], ra.ly[A0 + a], ra.lz[A0 + a]);
#pragma unroll
      for (int b = 0; b < NB; ++b) gb[b] = bra.bfB + gamess_pos(rb.lx[b], rb.ly[b], rb.lz[b]);
#pragma unroll
      for (int c = 0; c < NCP; ++c) gc[c] = ket.bfA + gamess_pos(rc.lx[C0 + c], rc.ly[C0 + c], rc.lz[C0 + c]);
#pragma unroll
      for (int d = 0; d < NDP; ++d) gd[d] = ket.bfB + gamess_pos(rd.lx[D0 + d], rd.ly[D0 + d], rd.lz[D0 + d]);
      int cm1[NM1], cm2[NM2];
#pragma unroll
      for (int i = 0; i < NM1; ++i) cm1[i] = mov.cA + gamess_pos(rm1.lx[M10 + i], rm1.ly[M10 + i], rm1.lz[M10 + i]);
#pragma unroll
      for (int i = 0; i < NM2; ++i) cm2[i] = off2 + mov.cB + gamess_pos(rm2.lx[M20 + i], rm2.ly[M20 + i], rm2.lz[M20 + i]);
      auto V = [&](int a, int b, int c, int d) -> double { return v[((a * NB + b) * NCP + c) * NDP + d]; };

      // Coulomb: the fixed pair's block summed over the wave, the moving pair's to F
      if constexpr (!TR) {
        double dcd[NCP * NDP];
#pragma unroll
        for (int c = 0; c < NCP; ++c)
#pragma unroll
          for (int d = 0; d < NDP; ++d) dcd[c * NDP + d] = Dm[tri(gc[c], gd[d])];
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            double t = 0.0;
#pragma unroll
            for (int c = 0; c < NCP; ++c)
#pragma unroll
              for (int d = 0; d < NDP; ++d) t += V(a, b, c, d) * dcd[c * NDP + d];
            jf[a * NB + b] = 4.0 * s * t;
          }
#pragma unroll
        for (int c = 0; c < NCP; ++c)
#pragma unroll
          for (int d = 0; d < NDP; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < NAP; ++a)
#pragma unroll
              for (int b = 0; b < NB; ++b) t += V(a, b, c, d) * dfix_s[(A0 + a) * NB + b];
            if (t != 0.0) fock_add(F, gc[c], gd[d], 4.0 * s * t);
          }
      } else {
        double dab[NAP * NB];
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) dab[a * NB + b] = Dm[tri(ga[a], gb[b])];
#pragma unroll
        for (int c = 0; c < NCP; ++c)
#pragma unroll
          for (int d = 0; d < NDP; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < NAP; ++a)
#pragma unroll
              for (int b = 0; b < NB; ++b) t += V(a, b, c, d) * dab[a * NB + b];
            jf[c * NDP + d] = 4.0 * s * t;
          }
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            double t = 0.0;
#pragma unroll
            for (int c = 0; c < NCP; ++c)
#pragma unroll
              for (int d = 0; d < NDP; ++d) t += V(a, b, c, d) * dfix_s[(C0 + c) * ND + D0 + d];
            if (t != 0.0) fock_add(F, ga[a], gb[b], 4.0 * s * t);
          }
      }
      const double sk = s * ba.k_scale;
      // exchange: K(AC) and K(BD)
      {
        double dbd[NB * NDP], dac[NAP * NCP];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int d = 0; d < NDP; ++d) dbd[b * NDP + d] = Dk[tri(gb[b], gd[d])];
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int c = 0; c < NCP; ++c) dac[a * NCP + c] = Dk[tri(ga[a], gc[c])];
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int c = 0; c < NCP; ++c) {
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
              for (int d = 0; d < NDP; ++d) t += V(a, b, c, d) * dbd[b * NDP + d];
            if (t != 0.0)
              lds_add(TR ? &rows[(C0 + c) * ncols + cm1[a]] : &rows[(A0 + a) * ncols + cm1[c]], -sk * t);
          }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int d = 0; d < NDP; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < NAP; ++a)
#pragma unroll
              for (int c = 0; c < NCP; ++c) t += V(a, b, c, d) * dac[a * NCP + c];
            if (t != 0.0) lds_add(TR ? &rows[(NC + D0 + d) * ncols + cm2[b]] : &rows[(NA + b) * ncols + cm2[d]], -sk * t);
          }
      }
      // exchange: K(AD) and K(BC)
      {
        double dbc[NB * NCP], dad[NAP * NDP];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int c = 0; c < NCP; ++c) dbc[b * NCP + c] = Dk[tri(gb[b], gc[c])];
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int d = 0; d < NDP; ++d) dad[a * NDP + d] = Dk[tri(ga[a], gd[d])];
#pragma unroll
        for (int a = 0; a < NAP; ++a)
#pragma unroll
          for (int d = 0; d < NDP; ++d) {
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
              for (int c = 0; c < NCP; ++c) t += V(a, b, c, d) * dbc[b * NCP + c];
            if (t != 0.0) lds_add(TR ? &rows[(NC + D0 + d) * ncols + cm1[a]] : &rows[(A0 + a) * ncols + cm2[d]], -sk * t);
          }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int c = 0; c < NCP; ++c) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < NAP; ++a)
#pragma unroll
              for (int d = 0; d < NDP; ++d) t += V(a, b, c, d) * dad[a * NDP + d];
            if (t != 0.0) lds_add(TR ? &rows[(C0 + c) * ncols + cm2[b]] : &rows[(NA + b) * ncols + cm1[c]], -sk * t);
          }
      }
    }
    // this part's share of the fixed pair's Coulomb block, summed over the
    // wave (every lane takes part; lanes past the tile hold zeros)
    wave_sum_to_lds<NJ>(jf, jfix_s, [](int k) {
      return TR ? (C0 + k / NDP) * ND + D0 + k % NDP : (A0 + k / NB) * NB + k % NB;
    });
  }
}

// the parts of wave slice wpart (S / W consecutive parts), one after the other
template <int LA, int LB, int LC, int LD, bool TR, int PART>
__device__ __forceinline__ void split_dispatch(int wpart, const ClassDev &cd, const BuildArgs &ba,
                                               const PairRec &fix, int32_t ifix, int64_t q0, int64_t q1, int seg0,
                                               int bat, double *jfix_s, const double *dfix_s, double *rows,
                                               int ncols, int off2, const double *tab) {
  using P = SplitShape<LA, LB, LC, LD>;
  if constexpr (PART < P::S) {
    if (PART / (P::S / P::W) == wpart)
      split_part<LA, LB, LC, LD, TR, PART>(cd, ba, fix, ifix, q0, q1, seg0, bat, jfix_s, dfix_s, rows, ncols, off2,
                                           tab);
    split_dispatch<LA, LB, LC, LD, TR, PART + 1>(wpart, cd, ba, fix, ifix, q0, q1, seg0, bat, jfix_s, dfix_s, rows,
                                                 ncols, off2, tab);
  }
}

// LDS of the wave-split kernel: jfix and dfix (the fixed pair's Coulomb sums
// and density block), the exchange rows, the Rys table below the switch
template <int LA, int LB, int LC, int LD, bool TR>
int64_t split_lds(const int32_t *ncol, double xswitch) {
  using R = RowShape<LA, LB, LC, LD, TR>;
  return (int64_t)sizeof(double) * (2 * R::NFIX + R::NROW * row_columns<LA, LB, LC, LD, TR>(ncol) +
                                    rys_table_used(SplitShape<LA, LB, LC, LD>::NR, xswitch));
}

// Wave-split row-tiled kernel: one workgroup owns one row tile (a run of
// quartets sharing the fixed pair); wave w takes the parts of wave slice
// w % W of the quartets of run w / W of each step.
template <int LA, int LB, int LC, int LD, bool TR>
__global__ __launch_bounds__((SplitShape<LA, LB, LC, LD>::NT)) void eri_split_kernel(ClassDev cd, BuildArgs ba) {
  using P = SplitShape<LA, LB, LC, LD>;
  using R = RowShape<LA, LB, LC, LD, TR>;
  constexpr int N1 = R::N1, N2 = R::N2, NFIX = R::NFIX, NROW = R::NROW;
  extern __shared__ double lds[];
  double *jfix_s = lds, *dfix_s = lds + NFIX, *rows = lds + 2 * NFIX;
  const int ncols = (int)row_columns<LA, LB, LC, LD, TR>(ba.ncol);
  const int off2 = R::M1 == R::M2 ? 0 : ba.ncol[R::M1];
  double *tab = rows + NROW * ncols;
  {
    constexpr int IN = rys_ncoef(P::NR), LS = rys_lds_stride(P::NR);
    const int ntab = (int)rys_table_intervals(P::NR, cd.xswitch) * IN;
    const double *src = rys_table + rys_offset[P::NR];
    for (int i = threadIdx.x; i < ntab; i += blockDim.x) tab[(i / IN) * LS + i % IN] = src[i];
  }
  const int tile = cd.t_begin + (int)blockIdx.x;
  const int64_t q0 = cd.tile_q[tile] > cd.q_begin ? cd.tile_q[tile] : cd.q_begin;
  const int64_t q1 = cd.tile_q[tile + 1] < cd.q_end ? cd.tile_q[tile + 1] : cd.q_end;
  const int seg0 = cd.tile_seg[tile];
  const int32_t ifix = cd.seg_bra[seg0];
  const PairRec fix = load_rec(const_view(ba.pairs), ifix);
  for (int i = threadIdx.x; i < NFIX; i += blockDim.x) {
    jfix_s[i] = 0.0;
    dfix_s[i] = ba.D[tri(fix.bfA + gpos_of(R::L1, i / N2), fix.bfB + gpos_of(R::L2, i % N2))];
  }
  for (int i = threadIdx.x; i < NROW * ncols; i += blockDim.x) rows[i] = 0.0;
  __syncthreads();

  // slice and run of this wave (uniform over the wave)
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  split_dispatch<LA, LB, LC, LD, TR, 0>(wave % P::W, cd, ba, fix, ifix, q0, q1, seg0, wave / P::W, jfix_s, dfix_s,
                                        rows, ncols, off2, tab);
  __syncthreads();

  double *F = ba.F;
  for (int i = threadIdx.x; i < NFIX; i += blockDim.x) {
    const double x = jfix_s[i];
    if (x != 0.0) fock_add(F, fix.bfA + gpos_of(R::L1, i / N2), fix.bfB + gpos_of(R::L2, i % N2), x);
  }
  for (int i = threadIdx.x; i < NROW * ncols; i += blockDim.x) {
    const double x = rows[i];
    if (x == 0.0) continue;
    const int r = i / ncols, k = i - r * ncols;
    const int rbf = r < N1 ? fix.bfA + gpos_of(R::L1, r) : fix.bfB + gpos_of(R::L2, r - N1);
    const int cbf = (R::M1 == R::M2 || k < off2) ? ba.col_bf[R::M1][k] : ba.col_bf[R::M2][k - off2];
    fock_add(F, rbf, cbf, x);
  }
}

// ================================================ workgroup-per-quartet path
__device__ __forceinline__ void cart_decode(int L, int m, int &lx, int &ly, int &lz) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= m) ++i;
  lz = m - i * (i + 1) / 2;
  lx = L - i;
  ly = i - lz;
}
// component k of range [L0, ...] -> its angular momentum and (lx,ly,lz)
__device__ __forceinline__ void range_decode(int L0, int k, int &L, int &lx, int &ly, int &lz) {
  L = L0;
  while (k >= ncart(L)) {
    k -= ncart(L);
    ++L;
  }
  cart_decode(L, k, lx, ly, lz);
}
__device__ __forceinline__ int range_index(int L0, int lx, int ly, int lz) {
  return nrange(L0, lx + ly + lz - 1) + cidx(lx, ly, lz);
}

// component k (internal order) of a shell of angular momentum L -> GAMESS
// position and cartesian scale factor, without loops
__device__ __forceinline__ int gpos_rt(int L, int k) {
  if (L <= 1) return k;
  if (L == 2) return (int)((0x251430u >> (4 * k)) & 15u);
  return (int)((0x2861795430ull >> (4 * k)) & 15ull);
}
__device__ __forceinline__ double scale_rt(int L, int k) {
  if (L == 2) return (k == 1 || k == 2 || k == 4) ? 1.7320508075688772935 : 1.0;
  if (L == 3) return (k == 0 || k == 6 || k == 9) ? 1.0 : (k == 4 ? 3.8729833462074168852 : 2.2360679774997896964);
  return 1.0;
}

// Workgroup-per-quartet path for the classes whose group slot exceeds the
// LDS (the f classes).  One workgroup of NT lanes evaluates one quartet at a
// time: 3 NR lanes make the 2-D tables of a primitive quartet, every lane
// accumulates its components of (e0|f0); the HRR levels, the cartesian
// scales and the digestion then run over all lanes.  The HRR buffers live in
// LDS wherever the class's two buffers, tables, density blocks and step
// tables fit kCoopLds (gfx950: 160 KB per workgroup), else in a per-workgroup
// HBM scratch.  Each HRR level reads a per-row step table (source rows and
// direction, made once per workgroup), so an item costs two reads, one FMA
// and one write; the digestion reads the six density blocks from LDS with
// compile-time strides.
constexpr int kCoopLds = 160 * 1024;

template <int LA, int LB, int LC, int LD>
struct Coop {
  static constexpr int LAB = LA + LB, LCD = LC + LD;
  static constexpr int NR = (LAB + LCD) / 2 + 1;
  static constexpr int NI = LAB + 1, NK = LCD + 1;
  static constexpr int NE = nrange(LA, LAB), NF = nrange(LC, LCD);
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NABCD = NA * NB * NC * ND;
  static constexpr int NAB = NA * NB, NCD = NC * ND, NAC = NA * NC, NBD = NB * ND, NAD = NA * ND, NBC = NB * NC;
  static constexpr int DB = NAB + NCD + NAC + NBD + NAD + NBC;
  static constexpr int GT = 3 * NR * NI * NK;  // 2-D tables of one primitive quartet
  static constexpr int64_t buf_doubles() { return hrr_max_buffer(LA, LB, LC, LD); }
  // step-table entries: bra levels 1..LB (rows a, b), then ket levels 1..LD (c, d)
  static constexpr int bra_rows(int j) { return nrange(LA, LAB - j) * ncart(j); }
  static constexpr int ket_rows(int j) { return nrange(LC, LCD - j) * ncart(j); }
  static constexpr int tab_off(int lev) {  // lev 0..LB-1 bra, LB..LB+LD-1 ket
    int o = 0;
    for (int i = 0; i < lev; ++i) o += i < LB ? bra_rows(i + 1) : ket_rows(i - LB + 1);
    return o;
  }
  static constexpr int NTAB = tab_off(LB + LD);
  static constexpr int fixed_doubles() { return GT + DB + (NTAB + 1) / 2; }
  static constexpr bool IN_LDS = (int64_t)8 * (fixed_doubles() + 2 * buf_doubles()) <= kCoopLds;
  static constexpr int LDS_DOUBLES = fixed_doubles() + (IN_LDS ? 2 * (int)buf_doubles() : 0);
  static constexpr int NT = IN_LDS || NE * NF > 2048 ? 512 : (NE * NF > 512 ? 128 : 64);
  static constexpr int NKC = (NE * NF + NT - 1) / NT;  // accumulators per lane
  static constexpr int64_t scratch() { return IN_LDS ? 0 : 2 * buf_doubles(); }
};

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__((Coop<LA, LB, LC, LD>::NT)) void eri_coop_kernel(ClassDev cd, BuildArgs ba) {
  using S = Coop<LA, LB, LC, LD>;
  constexpr int NT = S::NT, NR = S::NR, NI = S::NI, NK = S::NK, NE = S::NE, NF = S::NF, NKC = S::NKC;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND;
  constexpr int NAB = S::NAB, NCD = S::NCD, NAC = S::NAC, NBD = S::NBD, NAD = S::NAD, NBC = S::NBC;
  constexpr int GSZ = NI * NK;
  __shared__ double lds[S::LDS_DOUBLES];
  double *Gt = lds;                   // [3][NR][GSZ]
  double *Dab = Gt + S::GT, *Dcd = Dab + NAB, *Dbd = Dcd + NCD, *Dac = Dbd + NBD, *Dbc = Dac + NAC, *Dad = Dbc + NBC;
  int *tab = (int *)(Dad + NAD);      // HRR step tables
  double *buf0 = S::IN_LDS ? lds + S::fixed_doubles() : ba.scratch + (int64_t)blockIdx.x * ba.scratch_per_wg;
  double *buf1 = buf0 + S::buf_doubles();

  const int t = threadIdx.x;
  // step tables: output row -> source rows s1 (shifted) and s2, direction
  for (int i = t; i < S::NTAB; i += NT) {
    int lev = 0;
    while (lev + 1 < LB + LD && S::tab_off(lev + 1) <= i) ++lev;
    const int r = i - S::tab_off(lev);
    int s1, s2, dir;
    if (lev < LB) {  // bra level j: row (a in range(LA, LAB-j), b in cart(j))
      const int j = lev + 1, nb = ncart(j), nbm = ncart(j - 1), a = r / nb, b = r % nb;
      int bx, by, bz, La, ax, ay, az;
      cart_decode(j, b, bx, by, bz);
      range_decode(LA, a, La, ax, ay, az);
      dir = bx ? 0 : (by ? 1 : 2);
      const int bmi = cidx(bx - (dir == 0), by - (dir == 1), bz - (dir == 2));
      s1 = range_index(LA, ax + (dir == 0), ay + (dir == 1), az + (dir == 2)) * nbm + bmi;
      s2 = a * nbm + bmi;
    } else {  // ket level j: column (c in range(LC, LCD-j), d in cart(j))
      const int j = lev - LB + 1, nd = ncart(j), ndm = ncart(j - 1), c = r / nd, d = r % nd;
      int dx, dy, dz, Lc, cx, cy, cz;
      cart_decode(j, d, dx, dy, dz);
      range_decode(LC, c, Lc, cx, cy, cz);
      dir = dx ? 0 : (dy ? 1 : 2);
      const int dmi = cidx(dx - (dir == 0), dy - (dir == 1), dz - (dir == 2));
      s1 = range_index(LC, cx + (dir == 0), cy + (dir == 1), cz + (dir == 2)) * ndm + dmi;
      s2 = c * ndm + dmi;
    }
    tab[i] = s1 | (s2 << 14) | (dir << 28);
  }

  // per-lane component offsets into the G tables (fixed for the kernel)
  int offx[NKC], offy[NKC], offz[NKC];
#pragma unroll
  for (int k = 0; k < NKC; ++k) {
    const int c = t + NT * k;
    if (c < NE * NF) {
      int Le, ex, ey, ez, Lf, fx, fy, fz;
      range_decode(LA, c / NF, Le, ex, ey, ez);
      range_decode(LC, c % NF, Lf, fx, fy, fz);
      offx[k] = ex * NK + fx;
      offy[k] = ey * NK + fy;
      offz[k] = ez * NK + fz;
    } else {
      offx[k] = offy[k] = offz[k] = 0;
    }
  }
  const double *Dm = ba.D, *Dk = ba.DK;
  double *F = ba.F;
  const double ks = ba.k_scale;

  // a contiguous chunk of the class's quartets per workgroup: the segment is
  // searched once and then followed, and consecutive quartets share the bra
  // pair (the cached bra records)
  const int64_t nq = cd.q_end - cd.q_begin;
  const int64_t per = (nq + gridDim.x - 1) / gridDim.x;
  const int64_t q0 = cd.q_begin + (int64_t)blockIdx.x * per;
  const int64_t q1 = q0 + per < cd.q_end ? q0 + per : cd.q_end;
  int sg = q0 < q1 ? find_segment(cd, q0, 0, cd.nseg) : 0;
  for (int64_t q = q0; q < q1; ++q) {
    while (sg + 1 < cd.nseg && cd.seg_start[sg + 1] <= q) ++sg;
    int32_t ib, ik;
    quartet_of_segment(cd, sg, q, ib, ik);
    const PairRec bra = ba.pairs[ib];
    const PairRec ket = ba.pairs[ik];
    // the six density blocks of the digestion (Coulomb: D(ab), D(cd); exchange: the rest)
    for (int i = t; i < S::DB; i += NT) {
      int k = i;
      const double *src = Dm;
      uint32_t r0, r1;
      if (k < NAB) {
        r0 = bra.bfA + gpos_rt(LA, k / NB), r1 = bra.bfB + gpos_rt(LB, k % NB);
      } else if ((k -= NAB) < NCD) {
        r0 = ket.bfA + gpos_rt(LC, k / ND), r1 = ket.bfB + gpos_rt(LD, k % ND);
      } else if (src = Dk, (k -= NCD) < NBD) {
        r0 = bra.bfB + gpos_rt(LB, k / ND), r1 = ket.bfB + gpos_rt(LD, k % ND);
      } else if ((k -= NBD) < NAC) {
        r0 = bra.bfA + gpos_rt(LA, k / NC), r1 = ket.bfA + gpos_rt(LC, k % NC);
      } else if ((k -= NAC) < NBC) {
        r0 = bra.bfB + gpos_rt(LB, k / NC), r1 = ket.bfA + gpos_rt(LC, k % NC);
      } else {
        k -= NBC;
        r0 = bra.bfA + gpos_rt(LA, k / ND), r1 = ket.bfB + gpos_rt(LD, k % ND);
      }
      Dab[i] = src[tri(r0, r1)];
    }
    double acc[NKC];
#pragma unroll
    for (int k = 0; k < NKC; ++k) acc[k] = 0.0;

    for (int ip = 0; ip < bra.nprim; ++ip) {
      for (int iq = 0; iq < ket.nprim; ++iq) {
        if (t < 3 * NR) {
          const PrimRec P = ba.prims[bra.prim_off + ip];
          const PrimRec Q = ba.prims[ket.prim_off + iq];
          const int r = t / 3, dim = t - 3 * (t / 3);
          const double pq = P.p + Q.p, ipq = 1.0 / pq;
          const double pqx = P.Px - Q.Px, pqy = P.Py - Q.Py, pqz = P.Pz - Q.Pz;
          const double rho = P.p * Q.p * ipq;
          const double X = rho * (pqx * pqx + pqy * pqy + pqz * pqz);
          double t2, wr;
          rys_root<NR>(X, cd.xswitch, r, t2, wr);
          const double hp = 0.5 / P.p, hq = 0.5 / Q.p;
          const double b00 = 0.5 * t2 * ipq;
          const double b10 = hp - hp * Q.p * t2 * ipq;
          const double b01 = hq - hq * P.p * t2 * ipq;
          const double fa = Q.p * ipq * t2, fc = P.p * ipq * t2;
          const double pqd = dim == 0 ? pqx : (dim == 1 ? pqy : pqz);
          const double pa = dim == 0 ? P.PAx : (dim == 1 ? P.PAy : P.PAz);
          const double qc = dim == 0 ? Q.PAx : (dim == 1 ? Q.PAy : Q.PAz);
          const double g00 = dim == 2 ? wr * kTwoPi52 * P.K * Q.K * sqrt(ipq) : 1.0;
          double g[NI][NK];
          rys_2d<NI, NK>(g, pa - fa * pqd, qc + fc * pqd, b10, b01, b00, g00);
#pragma unroll
          for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int k = 0; k < NK; ++k) Gt[(dim * NR + r) * GSZ + i * NK + k] = g[i][k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NKC; ++k) {
          double v = 0.0;
#pragma unroll
          for (int r = 0; r < NR; ++r)
            v += Gt[r * GSZ + offx[k]] * Gt[(NR + r) * GSZ + offy[k]] * Gt[(2 * NR + r) * GSZ + offz[k]];
          acc[k] += v;
        }
        __syncthreads();
      }
    }
    // cartesian scales and the value cutoff, applied by the last write of the
    // HRR (whose output index is v's [a][b][c][d])
    auto final_value = [&](int it, double x) -> double {
      const int d = it % ND, c = (it / ND) % NC, b = (it / NCD) % NB, a = it / (NB * NCD);
      const double v = x * scale_rt(LA, a) * scale_rt(LB, b) * scale_rt(LC, c) * scale_rt(LD, d);
      return fabs(v) < ba.value_cut ? 0.0 : v;
    };
#pragma unroll
    for (int k = 0; k < NKC; ++k) {
      const int c = t + NT * k;
      if (c < NE * NF) buf0[c] = LB == 0 && LD == 0 ? final_value(c, acc[k]) : acc[k];
    }
    __syncthreads();

    // bra HRR: level j out[a in range(LA, LAB-j)][b in cart(j)][f] from level j-1
    double *in = buf0, *out = buf1;
#pragma unroll
    for (int j = 1; j <= LB; ++j) {
      const int *tj = tab + S::tab_off(j - 1);
      const int items = S::bra_rows(j) * NF;
      for (int it = t; it < items; it += NT) {
        const int row = it / NF, f = it - row * NF;
        const int e = tj[row], dir = e >> 28;
        const double abv = dir == 0 ? bra.ABx : (dir == 1 ? bra.ABy : bra.ABz);
        const double x = in[(e & 16383) * NF + f] + abv * in[((e >> 14) & 16383) * NF + f];
        out[it] = j == LB && LD == 0 ? final_value(it, x) : x;
      }
      __syncthreads();
      double *tmp = in;
      in = out;
      out = tmp;
    }
    // ket HRR on the inner index: out[ab][c in range(LC, LCD-j)][d in cart(j)]
#pragma unroll
    for (int j = 1; j <= LD; ++j) {
      const int *tj = tab + S::tab_off(LB + j - 1);
      const int nin = S::ket_rows(j - 1), nout = S::ket_rows(j);
      for (int it = t; it < NAB * nout; it += NT) {
        const int ab = it / nout, cdi = it - ab * nout;
        const int e = tj[cdi], dir = e >> 28;
        const double cdv = dir == 0 ? ket.ABx : (dir == 1 ? ket.ABy : ket.ABz);
        const double x = in[ab * nin + (e & 16383)] + cdv * in[ab * nin + ((e >> 14) & 16383)];
        out[it] = j == LD ? final_value(it, x) : x;
      }
      __syncthreads();
      double *tmp = in;
      in = out;
      out = tmp;
    }

    double s = 1.0;
    if (bra.sameAB) s *= 0.5;
    if (ket.sameAB) s *= 0.5;
    if (cd.diag && ib == ik) s *= 0.5;
    const double *v = in;
    // six contracted blocks, one output element per lane iteration; the sums
    // run over compile-time strides from the output's base
    constexpr int n1 = NAB, n2 = n1 + NCD, n3 = n2 + NAC, n4 = n3 + NBD, n5 = n4 + NAD, n6 = n5 + NBC;
    for (int it = t; it < n6; it += NT) {
      double acc2 = 0.0, fac;
      uint32_t fi, fj;
      if (it < n1) {  // F[ab] += 4s sum_cd v D[cd]
        const int a = it / NB, b = it % NB;
        const double *vb = v + it * NCD;
#pragma unroll 4
        for (int k = 0; k < NCD; ++k) acc2 += vb[k] * Dcd[k];
        fi = bra.bfA + gpos_rt(LA, a), fj = bra.bfB + gpos_rt(LB, b), fac = 4.0 * s;
      } else if (it < n2) {  // F[cd] += 4s sum_ab v D[ab]
        const int k0 = it - n1, c = k0 / ND, d = k0 % ND;
        const double *vb = v + k0;
#pragma unroll 4
        for (int k = 0; k < NAB; ++k) acc2 += vb[k * NCD] * Dab[k];
        fi = ket.bfA + gpos_rt(LC, c), fj = ket.bfB + gpos_rt(LD, d), fac = 4.0 * s;
      } else if (it < n3) {  // F[ac] -= s sum_bd v D[bd]
        const int k0 = it - n2, a = k0 / NC, c = k0 % NC;
        const double *vb = v + a * NB * NCD + c * ND;
#pragma unroll 4
        for (int b = 0; b < NB; ++b)
#pragma unroll 4
          for (int d = 0; d < ND; ++d) acc2 += vb[b * NCD + d] * Dbd[b * ND + d];
        fi = bra.bfA + gpos_rt(LA, a), fj = ket.bfA + gpos_rt(LC, c), fac = -s * ks;
      } else if (it < n4) {  // F[bd] -= s sum_ac v D[ac]
        const int k0 = it - n3, b = k0 / ND, d = k0 % ND;
        const double *vb = v + b * NCD + d;
#pragma unroll 4
        for (int a = 0; a < NA; ++a)
#pragma unroll 4
          for (int c = 0; c < NC; ++c) acc2 += vb[a * NB * NCD + c * ND] * Dac[a * NC + c];
        fi = bra.bfB + gpos_rt(LB, b), fj = ket.bfB + gpos_rt(LD, d), fac = -s * ks;
      } else if (it < n5) {  // F[ad] -= s sum_bc v D[bc]
        const int k0 = it - n4, a = k0 / ND, d = k0 % ND;
        const double *vb = v + a * NB * NCD + d;
#pragma unroll 4
        for (int b = 0; b < NB; ++b)
#pragma unroll 4
          for (int c = 0; c < NC; ++c) acc2 += vb[b * NCD + c * ND] * Dbc[b * NC + c];
        fi = bra.bfA + gpos_rt(LA, a), fj = ket.bfB + gpos_rt(LD, d), fac = -s * ks;
      } else {  // F[bc] -= s sum_ad v D[ad]
        const int k0 = it - n5, b = k0 / NC, c = k0 % NC;
        const double *vb = v + b * NCD + c * ND;
#pragma unroll 4
        for (int a = 0; a < NA; ++a)
#pragma unroll 4
          for (int d = 0; d < ND; ++d) acc2 += vb[a * NB * NCD + d] * Dad[a * ND + d];
        fi = bra.bfB + gpos_rt(LB, b), fj = ket.bfA + gpos_rt(LC, c), fac = -s * ks;
      }
      if (acc2 != 0.0) fock_add(F, fi, fj, fac * acc2);
    }
    __syncthreads();  // buffers and density blocks are reused by the next quartet
  }
}

// =================================================== group-per-quartet path
// G lanes own one quartet; a one-wave workgroup holds 64/G quartets side by
// side, each with its own LDS slot: 2-D Rys tables of the current primitive
// quartet, two HRR buffers, the six density blocks.  Loop bounds are uniform
// over the wave (primitive loop to the wave's longest quartet, others
// predicated), so the single-wave barriers are cheap and always matched.
// components of angular momenta l0..L (runtime L >= l0 - 1)
__device__ __forceinline__ int nrange_rt(int l0, int L) {
  return ((L + 1) * (L + 2) * (L + 3) - l0 * (l0 + 1) * (l0 + 2)) / 6;
}

// The HRR levels in a second form, for levels whose columns (bra) or rows
// (ket) are fewer than the lanes: lanes over (component, column) pairs.  A
// lane decodes its component once and finds the positions of its three
// neighbours one unit higher; the loop over the level's new cartesian
// components stays compile-time.  Same arithmetic as the one-lane-per-column
// form below; grp_hrr_cost picks the form with fewer issued instructions.
template <int LA, int LAB, int J, int NF, int G, int OS>
__device__ __forceinline__ void grp_bra_level_af(const double *__restrict__ in, double *__restrict__ out, int gl,
                                                 double abx, double aby, double abz) {
  constexpr Range<J, J> rb{};
  constexpr int na = nrange(LA, LAB - J), nbm = ncart(J - 1), nb = ncart(J);
  for (int t = gl; t < na * NF; t += G) {
    const int a = t / NF, f = t - a * NF;
    int L, ax, ay, az;
    range_decode(LA, a, L, ax, ay, az);
    // positions of a + e_x, a + e_y, a + e_z in range(LA, LAB - J + 1)
    const int base = nrange_rt(LA, L), s1 = ay + az + 1;
    const int px = base + cidx(ax, ay, az), py = base + (s1 * (s1 + 1)) / 2 + az, pz = py + 1;
    const double *i0 = in + a * nbm * NF + f;
    const double *ix = in + px * nbm * NF + f, *iy = in + py * nbm * NF + f, *iz = in + pz * nbm * NF + f;
    double *o = out + a * nb * OS + f;
#pragma unroll
    for (int b = 0; b < nb; ++b) {
      const int bx = rb.lx[b], by = rb.ly[b], bz = rb.lz[b];
      const int dir = bx ? 0 : (by ? 1 : 2);
      const int bmi = cidx(bx - (dir == 0), by - (dir == 1), bz - (dir == 2));
      const double *ip = dir == 0 ? ix : (dir == 1 ? iy : iz);
      const double ab = dir == 0 ? abx : (dir == 1 ? aby : abz);
      o[b * OS] = ip[bmi * NF] + ab * i0[bmi * NF];
    }
  }
}
template <int LC, int LCD, int J, int NAB, int G, int KS>
__device__ __forceinline__ void grp_ket_level_rc(const double *__restrict__ in, double *__restrict__ out, int gl,
                                                 double cdx, double cdy, double cdz) {
  constexpr Range<J, J> rd{};
  constexpr int nc = nrange(LC, LCD - J), ndm = ncart(J - 1), nd = ncart(J);
  for (int t = gl; t < NAB * nc; t += G) {
    const int ab = t / nc, c = t - ab * nc;
    int L, cx, cy, cz;
    range_decode(LC, c, L, cx, cy, cz);
    const int base = nrange_rt(LC, L), s1 = cy + cz + 1;
    const int px = base + cidx(cx, cy, cz), py = base + (s1 * (s1 + 1)) / 2 + cz, pz = py + 1;
    const double *ip = in + ab * KS;
    const double *i0 = ip + c * ndm, *ix = ip + px * ndm, *iy = ip + py * ndm, *iz = ip + pz * ndm;
    double *op = out + ab * KS + c * nd;
#pragma unroll
    for (int d = 0; d < nd; ++d) {
      const int dx = rd.lx[d], dy = rd.ly[d], dz = rd.lz[d];
      const int dir = dx ? 0 : (dy ? 1 : 2);
      const int dmi = cidx(dx - (dir == 0), dy - (dir == 1), dz - (dir == 2));
      const double *i1 = dir == 0 ? ix : (dir == 1 ? iy : iz);
      const double cdv = dir == 0 ? cdx : (dir == 1 ? cdy : cdz);
      op[d] = i1[dmi] + cdv * i0[dmi];
    }
  }
}
// Whether the pair form issues fewer instructions (estimate: 4 per output,
// 20 per decoded component) for a level of `lines` columns (bra) or rows
// (ket), m components per line and n new cartesian components each, on g
// lanes: one lane per line does m n outputs, the pair form n per position
__host__ __device__ constexpr bool grp_pair_form(int lines, int m, int n, int g) {
  return ((lines * m + g - 1) / g) * (n * 4 + 20) < ((lines + g - 1) / g) * (m * n * 4);
}

template <int LA, int LB, int LC, int LD>
struct GroupSize {
  static constexpr int NEF = nrange(LA, LA + LB) * nrange(LC, LC + LD);
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  static constexpr int raw = NEF / 8 <= 16 ? 16 : (NEF / 8 <= 32 ? 32 : 64);
  static constexpr int G = (3 * NR > raw) ? 64 : raw;
};

__host__ __device__ constexpr int odd_at_least(int n) { return n | 1; }
// longest row of the ket-phase arrays: NF at level 0, then each level's output
__host__ __device__ constexpr int ket_row_max(int LC, int LD) {
  int m = nrange(LC, LC + LD);
  for (int j = 1; j <= LD; ++j) {
    const int s = nrange(LC, LC + LD - j) * ncart(j);
    m = s > m ? s : m;
  }
  return m;
}
__host__ __device__ constexpr int grp_buffer(int LA, int LB, int LC, int LD, int KS) {
  const int NF = nrange(LC, LC + LD);
  int m = ncart(LA) * ncart(LB) * KS;
  if (LB > 0) {
    const int s0 = nrange(LA, LA + LB) * NF;
    m = s0 > m ? s0 : m;
    for (int j = 1; j < LB; ++j) {
      const int s = nrange(LA, LA + LB - j) * ncart(j) * NF;
      m = s > m ? s : m;
    }
  }
  return m;
}

// bra HRR level J for the columns f of this lane: out[a in range(LA, LAB-J)][b in
// cart(J)][f] (row stride OS per (a, b)) from level J-1 (row stride NF)
template <int LA, int LAB, int J, int NF, int G, int OS>
__device__ __forceinline__ void grp_bra_level(const double *__restrict__ in, double *__restrict__ out, int gl,
                                              double abx, double aby, double abz) {
  constexpr Range<LA, LAB - J> ra{};
  constexpr Range<J, J> rb{};
  constexpr int nbm = ncart(J - 1), nb = ncart(J);
  for (int f = gl; f < NF; f += G) {
#pragma unroll
    for (int a = 0; a < ra.n; ++a)
#pragma unroll
      for (int b = 0; b < nb; ++b) {
        const int bx = rb.lx[b], by = rb.ly[b], bz = rb.lz[b];
        const int dir = bx ? 0 : (by ? 1 : 2);
        const int bmi = cidx(bx - (dir == 0), by - (dir == 1), bz - (dir == 2));
        const int ax = ra.lx[a] + (dir == 0), ay = ra.ly[a] + (dir == 1), az = ra.lz[a] + (dir == 2);
        const int api = nrange(LA, ra.l[a]) + cidx(ax, ay, az);
        const double ab = dir == 0 ? abx : (dir == 1 ? aby : abz);
        out[(a * nb + b) * OS + f] = in[(api * nbm + bmi) * NF + f] + ab * in[(a * nbm + bmi) * NF + f];
      }
  }
}

// ket HRR level J for the rows ab of this lane (row stride KS in and out)
template <int LC, int LCD, int J, int NAB, int G, int KS>
__device__ __forceinline__ void grp_ket_level(const double *__restrict__ in, double *__restrict__ out, int gl,
                                              double cdx, double cdy, double cdz) {
  constexpr Range<LC, LCD - J> rc{};
  constexpr Range<J, J> rd{};
  constexpr int ndm = ncart(J - 1), nd = ncart(J);
  for (int ab = gl; ab < NAB; ab += G) {
    const double *ip = in + ab * KS;
    double *op = out + ab * KS;
#pragma unroll
    for (int c = 0; c < rc.n; ++c)
#pragma unroll
      for (int d = 0; d < nd; ++d) {
        const int dx = rd.lx[d], dy = rd.ly[d], dz = rd.lz[d];
        const int dir = dx ? 0 : (dy ? 1 : 2);
        const int dmi = cidx(dx - (dir == 0), dy - (dir == 1), dz - (dir == 2));
        const int cx = rc.lx[c] + (dir == 0), cy = rc.ly[c] + (dir == 1), cz = rc.lz[c] + (dir == 2);
        const int cpi = nrange(LC, rc.l[c]) + cidx(cx, cy, cz);
        const double cdv = dir == 0 ? cdx : (dir == 1 ? cdy : cdz);
        op[c * nd + d] = ip[cpi * ndm + dmi] + cdv * ip[c * ndm + dmi];
      }
  }
}

// LDS bytes of a one-wave group workgroup that still allow 3 of them per SIMD
constexpr int kGroupLdsTarget = 160 * 1024 / 12;

// Root split P of the row/column lane map of the primitive loop (lanes per
// quartet = nl * P): the smallest split whose NQ = 64 / (nl P) slots of
// `slot` doubles fit kGroupLdsTarget, else the largest split that fits a
// wave; the 3 NR table lanes must fit the group.  0 when no split fits.
__host__ __device__ constexpr int grp_parts(int nl, int nr, int slot) {
  int best = 0;
  for (int p = 1; p <= nr && nl * p <= 64; ++p) {
    if (nl * p < 3 * nr) continue;
    best = p;
    if ((64 / (nl * p)) * slot * 8 <= kGroupLdsTarget) return p;
  }
  return best;
}

// LDS cycles of one wave access to the distinct doubles at addresses
// a[0..n) (32 banks of 4 bytes, double d on banks 2d and 2d+1 mod 32): the
// largest number of addresses that share a bank pair
constexpr int lds_cycles(const int *a, int n) {
  int cnt[16] = {};
  for (int i = 0; i < n; ++i) ++cnt[a[i] % 16];
  int m = 0;
  for (int k = 0; k < 16; ++k) m = cnt[k] > m ? cnt[k] : m;
  return m;
}

// Slot of a group (unpadded, doubles): the union of table and HRR buffers
// (uni), the density blocks when they do not fit the free HRR buffer (dbx),
// and room for as many rule records of the primitive loop as one pass of the
// wave makes, but never more than the slot with separate density blocks
// (db) holds
__host__ __device__ constexpr int grp_slot(int uni, int gs, int kt, int nq, int nr, int db, int dbx) {
  const int rtw = 64 / (nq * nr) / kt * kt, need = kt * gs + rtw * (2 * nr + 14);
  const int lo = uni + dbx, hi = uni + db;
  const int want = need < hi ? need : hi;
  return want > lo ? want : lo;
}

// LDS cycles of one wave-wide 64-bit access, lane l at address a[l] (in
// doubles, all distinct; -1: lane inactive): the LDS serves 16 lanes of such
// an access per clock, so each 16-lane group costs the largest number of
// its addresses on one bank pair
constexpr int lds_cycles_b64(const int *a, int n) {
  int total = 0;
  for (int g0 = 0; g0 < n; g0 += 16) {
    int cnt[16] = {};
    for (int i = g0; i < g0 + 16 && i < n; ++i)
      if (a[i] >= 0) ++cnt[a[i] % 16];
    int m = 0;
    for (int k = 0; k < 16; ++k) m = cnt[k] > m ? cnt[k] : m;
    total += m;
  }
  return total;
}

// Padding of the 2-D table stride (gsz0 + gp doubles) and of the slot
// (+ sp doubles), chosen to minimise the LDS cycles of the primitive loop:
// the accumulation reads (one instruction reads the same table element
// offset of every slot, root part and row of the lanes) and the table writes
// (one element of every (slot, quartet, root, direction) table), weighted by
// how often each runs per KT primitive quartets.  Candidates that lose a
// table pass's quartets, leave 64 KB, or push a slot set that fitted three
// waves per SIMD out of that budget are skipped.  Returns gp * 16 + sp
// (0 for the lane-per-quartet classes, which have no slots, and for strided
// maps of more than 128 components; strided maps pad the slot only).
template <int LA, int LB, int LC, int LD>
constexpr int grp_pad(int map, int p, int nq, int g, int nr, int gsz0, int bufp, int db, int dbx) {
  constexpr int LAB = LA + LB, LCD = LC + LD, NI = LAB + 1, NK = LCD + 1;
  constexpr Range<LA, LAB> re{};
  constexpr Range<LC, LCD> rf{};
  constexpr int NE = re.n, NF = rf.n;
  if (NE * NF <= 100 && hrr_max_buffer(LA, LB, LC, LD) <= 100) return 0;
  if (map == 0 && NE * NF > 128) return 0;  // (keeps the search within the compiler's step budget)
  int best = 0, best_cost = -1, kt0 = 0;
  bool fit3 = false;
  // row/column maps: the distinct (slot, root part, exponent) triples each
  // 16-lane group of an accumulation read touches, per direction (lanes
  // with equal triples read the same address); independent of the padding
  int tq[3][4][16] = {}, tp[3][4][16] = {}, tx[3][4][16] = {}, tn[3][4] = {};
  if (map == 1 || map == 2)
    for (int dim = 0; dim < 3; ++dim)
      for (int l = 0; l < 64; ++l) {
        const int q = l / g, gl = l - q * g, row = gl / p, pp = gl - row * p, grp = l / 16;
        if (q >= nq) continue;
        int x = 0;
        if (map == 1 && row < NE) x = dim == 0 ? re.lx[row] : (dim == 1 ? re.ly[row] : re.lz[row]);
        if (map == 2 && row < NF) x = dim == 0 ? rf.lx[row] : (dim == 1 ? rf.ly[row] : rf.lz[row]);
        bool dup = false;
        for (int j = 0; j < tn[dim][grp]; ++j)
          dup = dup || (tq[dim][grp][j] == q && tp[dim][grp][j] == pp && tx[dim][grp][j] == x);
        if (!dup) {
          const int j = tn[dim][grp]++;
          tq[dim][grp][j] = q;
          tp[dim][grp][j] = pp;
          tx[dim][grp][j] = x;
        }
      }
  for (int gp = 0; gp < (map == 0 ? 1 : 16); ++gp)
    for (int sp = 0; sp < 16; ++sp) {
      const int gsz = gsz0 + gp, gs = 3 * nr * gsz;
      const int uni = 2 * gs > 2 * bufp ? 2 * gs : 2 * bufp;
      int kt = g / (3 * nr) < uni / gs ? g / (3 * nr) : uni / gs;
      kt = kt < 1 ? 1 : kt;
      const int slot = grp_slot(uni, gs, kt, nq, nr, db, dbx) + sp;
      if (gp == 0 && sp == 0) {
        kt0 = kt;
        fit3 = nq * slot * 8 <= kGroupLdsTarget;
      }
      if (kt != kt0 || nq * slot * 8 > 64 * 1024 || (fit3 && nq * slot * 8 > kGroupLdsTarget)) continue;
      int addr[64] = {};
      int cost = 0;
      // accumulation reads
      if (map == 1 || map == 2) {
        // lane q g + row p + part reads row (column) `row` of its part's
        // table: offset part gsz + (its exponent) xs; per 16-lane group, the
        // most distinct addresses on one bank pair
        const int xs = map == 1 ? NK : 1, nj = map == 1 ? NK : NI;
        for (int dim = 0; dim < 3; ++dim) {
          int cyc = 0;
          for (int grp = 0; grp < 4; ++grp) {
            int cnt[16] = {};
            for (int j = 0; j < tn[dim][grp]; ++j)
              ++cnt[(tq[dim][grp][j] * slot + (dim * nr + tp[dim][grp][j]) * gsz + tx[dim][grp][j] * xs) % 16];
            int m = 0;
            for (int k = 0; k < 16; ++k) m = cnt[k] > m ? cnt[k] : m;
            cyc += m;
          }
          cost += kt * nj * ((nr + p - 1) / p) * cyc;
        }
      } else {
        for (int k = 0; k * g < NE * NF; ++k)
          for (int dim = 0; dim < 3; ++dim) {
            bool seen[64] = {};  // table offsets ec * NK + fc < NI * NK <= 64
            for (int l = 0; l < g; ++l) {
              const int c = l + g * k;
              if (c >= NE * NF) break;
              const int e = c / NF, f = c % NF;
              const int ec = dim == 0 ? re.lx[e] : (dim == 1 ? re.ly[e] : re.lz[e]);
              const int fc = dim == 0 ? rf.lx[f] : (dim == 1 ? rf.ly[f] : rf.lz[f]);
              seen[ec * NK + fc] = true;
            }
            int n = 0;
            for (int q = 0; q < nq; ++q)
              for (int o = 0; o < NI * NK && o < 64; ++o)
                if (seen[o] && n < 64) addr[n++] = q * slot + o;
            cost += kt * nr * lds_cycles(addr, n);
          }
      }
      // table writes: lane q g + tt 3 nr + task writes table (task % 3, task / 3) of buffer tt
      if (map == 1 || map == 2) {
        for (int l = 0; l < 64; ++l) {
          const int q = l / g, gl = l - q * g, tt = gl / (3 * nr), task = gl - tt * (3 * nr);
          addr[l] = q < nq && tt < kt ? q * slot + tt * gs + ((task % 3) * nr + task / 3) * gsz : -1;
        }
        cost += NI * NK * lds_cycles_b64(addr, 64);
      } else {
        int n = 0;
        for (int q = 0; q < nq; ++q)
          for (int m = 0; m < kt * 3 * nr && n < 64; ++m) addr[n++] = q * slot + m * gsz;
        cost += NI * NK * lds_cycles(addr, n);
      }
      if (best_cost < 0 || cost < best_cost) {
        best_cost = cost;
        best = gp * 16 + sp;
      }
    }
  return best;
}

// classes whose group form runs the HRR in registers (Grp::REG): one
// column of (e0|f0) and one row (ab| of the ket phase per lane
__host__ __device__ constexpr bool grp_reg_hrr(int LA, int LB, int LC, int LD) {
  // (dd|dd), and the (df| classes with a ket HRR or an f ket shell but
  // (df|ff) (46-row columns; rows of at most 46 columns and two ket levels).
  // Measured on the C-N chain against the level-by-level form: (df|df) 2.41
  // -> 1.11 ms, (df|dd) 2.36 -> 1.32, (df|dp) 1.35 -> 1.03, (df|fs) 0.93 ->
  // 0.84, (df|pp) 0.46 -> 0.41; (df|ds) 1.09 -> 1.34 and (df|ps) 0.72 ->
  // 1.29 stay level by level
  return (LA == 2 && LB == 2 && LC == 2 && LD == 2) ||
         (LA == 3 && LB == 2 && LC >= 1 && LD <= 2 && LC >= LD && (LD > 0 || LC == 3));
}

template <int LA, int LB, int LC, int LD>
struct Grp {
  static constexpr int LAB = LA + LB, LCD = LC + LD;
  static constexpr int NR = (LAB + LCD) / 2 + 1;
  static constexpr int NI = LAB + 1, NK = LCD + 1;
  static constexpr int NE = nrange(LA, LAB), NF = nrange(LC, LCD);
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NABCD = NA * NB * NC * ND;
  static constexpr int BUF = hrr_max_buffer(LA, LB, LC, LD);
  static constexpr int NAB = NA * NB, NCD = NC * ND, NAC = NA * NC, NBD = NB * ND, NAD = NA * ND, NBC = NB * NC;
  static constexpr int DB = NAB + NCD + NAC + NBD + NAD + NBC;
  // row stride of the ket-phase arrays [ab][...]: the longest ket-phase row,
  // made odd so that lanes on consecutive rows hit distinct LDS banks
  static constexpr int KS = odd_at_least(ket_row_max(LC, LD));
  // register form of the HRR (REG): the bra HRR runs one column f of
  // (e0|f0) per lane and the ket HRR one row ab per lane, both in registers;
  // LDS holds (e0|f0) as [e][f], then (ab|f0) as [ab][f] (odd row stride TS),
  // then the digestion's partial sums ([ab][cd] for J(CD), stride JS, and
  // the four exchange blocks per row ab).  For the classes whose columns and
  // rows fit a lane ((dd|dd)).
  static constexpr bool REG = grp_reg_hrr(LA, LB, LC, LD);
  static constexpr int TS = odd_at_least(NF), JS = odd_at_least(NCD);
  static constexpr int REGA = NE * NF + NAB * TS, REGB = NAB * JS + NAB * 2 * (NC + ND);
  static constexpr int CBUF = hrr_max_buffer(LA, LB, 0, 0), RBUF = hrr_max_buffer(0, 0, LC, LD);
  // buffers of the HRR: bra phase [e][b][f] (stride NF), ket phase [ab][KS]
  // (REG: half of the larger of the two layouts above)
  static constexpr int BUFP = REG ? ((REGA > REGB ? REGA : REGB) + 1) / 2 : grp_buffer(LA, LB, LC, LD, KS);
  // slot before padding: table buffers (at least two), dead once the
  // primitive loop ends, aliased by the HRR buffers; the density blocks of
  // the digestion go into the HRR buffer the last level left free when they
  // fit there (DBX = 0), else after the buffers (DBX doubles; always in REG)
  static constexpr int GSZ0 = NI * NK;
  static constexpr int UNI0 = 2 * 3 * NR * GSZ0 > 2 * BUFP ? 2 * 3 * NR * GSZ0 : 2 * BUFP;
  static constexpr int DBX = REG || DB > BUFP ? DB : 0;
  static constexpr int SLOT0 = UNI0 + DBX;
  // lane map of the primitive loop.  MAP 1: a lane per row e of (e0|f0)
  // accumulating all its columns f; MAP 2: a lane per column f accumulating
  // all rows e; in both the roots are split over P lanes (partial sums added
  // at the end), and a lane reads one row (column) of each 2-D table per root
  // at compile-time offsets.  MAP 0 (rows or columns longer than a wave): the
  // NE NF components strided over G lanes, three table reads per product.
  // The map is taken only where it keeps at least as many quartets per wave
  // as the strided one (the table lanes' work per primitive quartet is the
  // same for any map, so fewer quartets per wave cost more than the saved
  // reads: e.g. (ps|dd), rows of 3).
  static constexpr int NL = NE >= NF ? NE : NF;
  static constexpr int P0 = NL <= 64 ? grp_parts(NL, NR, SLOT0) : 0;
  static constexpr int P = P0 && 64 / (NL * P0) >= 64 / GroupSize<LA, LB, LC, LD>::G ? P0 : 0;
  static constexpr int MAP = P == 0 ? 0 : (NE >= NF ? 1 : 2);
  static constexpr int G = MAP ? NL * P : GroupSize<LA, LB, LC, LD>::G;
  static constexpr int NQ = 64 / G;
  static constexpr int NKC = (NE * NF + G - 1) / G;
  // table stride (one 2-D table of NI x NK, padded) and slot, padded
  // against LDS bank conflicts
  static constexpr int PAD = grp_pad<LA, LB, LC, LD>(MAP, P, NQ, G, NR, GSZ0, BUFP, DB, DBX);
  static constexpr int GSZ = GSZ0 + PAD / 16;
  static constexpr int GS = 3 * NR * GSZ;
  static constexpr int UNI = 2 * GS > 2 * BUFP ? 2 * GS : 2 * BUFP;
  // primitive quartets whose 2-D tables one pass of the table lanes makes
  // (3 NR lanes each), held in the HRR space of the slot
  static constexpr int KT0 = G / (3 * NR) < UNI / GS ? G / (3 * NR) : UNI / GS;
  static constexpr int KT = KT0 < 1 ? 1 : KT0;
  static constexpr int SLOT = grp_slot(UNI, GS, KT, NQ, NR, DB, DBX) + PAD % 16;
  static constexpr bool fits = SLOT * NQ * 8 <= 64 * 1024;
  // rule records of the primitive loop, one per primitive quartet:
  // u[NR], w[NR] (times 2 pi^5/2 K_P K_Q / sqrt(p+q)), 1/(p+q), q/(p+q),
  // p/(p+q), 1/2p, 1/2q, P-Q, P-A, Q-C.  They live in the part of the slot
  // the tables leave free (HRR and density space, written only after the
  // loop); RT of them per slot (a multiple of KT) are made in one pass of the
  // wave, one lane per (slot, primitive quartet, root).
  static constexpr int REC = 2 * NR + 14;
  static constexpr int RT0 = (SLOT - KT * GS) / REC;
  static constexpr int RT1 = RT0 < 64 / (NQ * NR) ? RT0 : 64 / (NQ * NR);
  static constexpr int RT = RT1 / KT * KT;
};

// The group kernels run one wave per workgroup.  A wave's LDS operations
// complete in program order, so between the LDS phases of such a kernel a
// compiler-level fence at wavefront scope is enough: no s_barrier and no
// drain of the wave's outstanding memory operations.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lanes per output of a digestion block of `no` outputs on a slot of g
// lanes, splitting a summed index of extent x: the largest power of two
// that keeps no x split <= g and the split <= x
__host__ __device__ constexpr int split_lanes(int no, int x, int g) {
  int s = 1;
  while (2 * s * no <= g && 2 * s <= x) s *= 2;
  return s;
}

// Sum of the partial sums of SP adjacent lanes (SP a power of two) into the
// first of them; the other lanes' results are meaningless
template <int SP>
__device__ __forceinline__ double part_sum(double t) {
#pragma unroll
  for (int d = SP / 2; d >= 1; d /= 2) t += __shfl_down(t, d, 64);
  return t;
}

__device__ __forceinline__ int gpos_of(int L, int k) {
  int x, y, z;
  cart_decode(L, k, x, y, z);
  return gamess_pos(x, y, z);
}

// Persistent loop over a contiguous chunk of the class's quartets, NQ per
// one-wave workgroup and round: rule records, 2-D tables and accumulation
// per batch of primitive quartets, the HRR, then digestion by atomic adds to
// F (the bra pair's Coulomb block summed in registers while consecutive
// rounds keep the bra pair; the sums of blocks with fewer outputs than lanes
// split over the idle lanes).  Keeping the exchange blocks on chip -- per row
// tile in LDS rows, or per run of kets sharing a shell -- cut this kernel's
// atomics 5-9x on C60 but not its time: its atomics overlap its arithmetic,
// and the LDS rows cost occupancy.  Classes of s, p and d shells up to total
// angular momentum 7 fit 168 VGPRs, i.e. 3 waves per SIMD, with at most 40
// bytes of spill per lane ((dp|dp), (dd|pp), (dd|dp)); measured faster than
// 2 waves without spill.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void grp_run(const ClassDev &cd, const BuildArgs &ba, int64_t c0, int64_t c1,
                                        double *lds, int *s_npq, int *s_nk, int *s_boff, int *s_koff) {
  using S = Grp<LA, LB, LC, LD>;
  constexpr int G = S::G, NQ = S::NQ, NR = S::NR, NI = S::NI, NK = S::NK, NE = S::NE, NF = S::NF, NKC = S::NKC;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND, NABCD = S::NABCD;
  constexpr int GSZ = S::GSZ;  // stride of the 2-D tables (padded against bank conflicts)
  constexpr int RT = S::RT, REC = S::REC, KT = S::KT;
  static_assert(RT >= KT && KT * S::GS <= S::UNI, "no room for the tables and rule records");
  static_assert(!S::REG || (LB > 0 && 2 * S::BUFP <= S::UNI), "register HRR: layout");

  // lanes past the NQ groups (G need not divide 64) idle: their gl lies past
  // every loop bound and their slots are never valid
  const int lane = threadIdx.x;
  const bool idle = lane >= NQ * G;
  const int slot = idle ? NQ - 1 : lane / G;
  const int gl = idle ? (1 << 20) : lane - slot * G;
  double *Gt = lds + slot * S::SLOT;
  double *buf0 = Gt, *buf1 = buf0 + S::BUFP;

  // MAP 0: table offsets of the lane's components
  int offx[NKC], offy[NKC], offz[NKC];
#pragma unroll
  for (int k = 0; k < NKC; ++k) {
    const int c = gl + G * k;
    if (S::MAP == 0 && c < NE * NF) {
      int Le, ex, ey, ez, Lf, fx, fy, fz;
      range_decode(LA, c / NF, Le, ex, ey, ez);
      range_decode(LC, c % NF, Lf, fx, fy, fz);
      offx[k] = ex * NK + fx;
      offy[k] = ey * NK + fy;
      offz[k] = ez * NK + fz;
    } else {
      offx[k] = offy[k] = offz[k] = 0;
    }
  }
  // MAP 1/2: the lane's row e (column f) and root part
  constexpr int P = S::P, NRP = P ? (NR + P - 1) / P : 1;
  const int part = P ? gl % P : 0;
  int rx = 0, ry = 0, rz = 0;
  if (S::MAP != 0 && gl < G) {
    int Lr;
    range_decode(S::MAP == 1 ? LA : LC, gl / P, Lr, rx, ry, rz);
  }
  // table element of the lane's row (MAP 1: [rx][j]) or column (MAP 2: [j][rx])
  const int rowx = S::MAP == 1 ? rx * NK : rx, rowy = S::MAP == 1 ? ry * NK : ry, rowz = S::MAP == 1 ? rz * NK : rz;

  // the slot's J(AB) block for the run of quartets sharing bra pair jab_pair
  constexpr int NJL = (S::NAB + G - 1) / G;
  double jab[NJL];
#pragma unroll
  for (int k = 0; k < NJL; ++k) jab[k] = 0.0;
  int32_t jab_pair = -1, jab_a = 0, jab_b = 0;
  auto flush_jab = [&]() {
    if (jab_pair < 0) return;
#pragma unroll
    for (int k = 0; k < NJL; ++k) {
      const int i = gl + k * G;
      if (i < S::NAB && jab[k] != 0.0)
        fock_add(ba.F, jab_a + gpos_rt(LA, i / S::NB), jab_b + gpos_rt(LB, i % S::NB), jab[k]);
      jab[k] = 0.0;
    }
  };

  // one round of NQ quartets (one per slot): integrals, HRR, digestion
  auto round = [&](bool valid, int32_t ib, int32_t ik) {
    PairRec bra{}, ket{};
    if (valid) {
      bra = ba.pairs[ib];
      ket = ba.pairs[ik];
    }
    const int nk = valid ? ket.nprim : 0;
    const int npq = valid ? bra.nprim * ket.nprim : 0;
    if (gl == 0) {
      s_npq[slot] = npq;
      s_nk[slot] = nk;
      s_boff[slot] = bra.prim_off;
      s_koff[slot] = ket.prim_off;
    }
    wave_sync();
    int maxpq = 0;
#pragma unroll
    for (int s = 0; s < NQ; ++s) maxpq = s_npq[s] > maxpq ? s_npq[s] : maxpq;

    constexpr int NACC = S::MAP == 0 ? NKC : (S::MAP == 1 ? NF : NE);
    double acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; ++k) acc[k] = 0.0;

    // rule records of primitive quartets t0 .. t0+RT-1 of every slot, one
    // lane per (slot, quartet, root): the root and weight, and (root 0) the
    // prefactor, reciprocals and displacements the recursion needs
    auto rules = [&](int t0) {
      const int sa = lane / (RT * NR), rem = lane - sa * (RT * NR), tt = rem / NR, r = rem - tt * NR;
      const int t = t0 + tt;
      if (sa < NQ && t < s_npq[sa]) {
        const int nka = s_nk[sa];
        // t / nka in FP32: both are small integers, the quotient's fraction
        // stays >= 0.5/nka away from an integer, far above the rounding error
        const int ip = (int)(((float)t + 0.5f) * __builtin_amdgcn_rcpf((float)nka)), iq = t - ip * nka;
        const PrimRec P = ba.prims[s_boff[sa] + ip];
        const PrimRec Q = ba.prims[s_koff[sa] + iq];
        const double rs = nr_rsqrt(P.p + Q.p), ipq = rs * rs;
        const double pqx = P.Px - Q.Px, pqy = P.Py - Q.Py, pqz = P.Pz - Q.Pz;
        const double X = P.p * Q.p * ipq * (pqx * pqx + pqy * pqy + pqz * pqz);
        const double pref = kTwoPi52 * P.K * Q.K * rs;
        double *rec = lds + sa * S::SLOT + KT * S::GS + tt * REC;
        double ur, wr;
        rys_root<NR>(X, cd.xswitch, r, ur, wr);
        rec[r] = ur;
        rec[NR + r] = wr * pref;
        if (r == 0) {
        rec[2 * NR] = ipq;
        rec[2 * NR + 1] = Q.p * ipq;
        rec[2 * NR + 2] = P.p * ipq;
        rec[2 * NR + 3] = 0.5 * nr_rcp(P.p);
        rec[2 * NR + 4] = 0.5 * nr_rcp(Q.p);
        rec[2 * NR + 5] = pqx;
        rec[2 * NR + 6] = pqy;
        rec[2 * NR + 7] = pqz;
        rec[2 * NR + 8] = P.PAx;
        rec[2 * NR + 9] = P.PAy;
        rec[2 * NR + 10] = P.PAz;
        rec[2 * NR + 11] = Q.PAx;
        rec[2 * NR + 12] = Q.PAy;
        rec[2 * NR + 13] = Q.PAz;
        }
      }
    };
    // 2-D Rys tables of primitive quartets t0 .. t0+KT-1 (3 NR lanes each:
    // one root and one direction per lane) from their rule records, quartet
    // t0+tt into table buffer tt
    auto tables = [&](int t0) {
      const int tt = gl / (3 * NR), task = gl - tt * (3 * NR), t = t0 + tt;
      if (tt < KT && t < npq) {
        const int r = task / 3, dim = task - 3 * r;
        const double *rec = Gt + KT * S::GS + (t % RT) * REC;
        const double u = rec[r], ipq = rec[2 * NR], a = rec[2 * NR + 1], c = rec[2 * NR + 2];
        const double hp = rec[2 * NR + 3], hq = rec[2 * NR + 4];
        const double pqd = rec[2 * NR + 5 + dim], pa = rec[2 * NR + 8 + dim], qc = rec[2 * NR + 11 + dim];
        const double fa = a * u, fc = c * u;
        const double b00 = 0.5 * u * ipq, b10 = hp - hp * fa, b01 = hq - hq * fc;
        const double g00 = dim == 2 ? rec[NR + r] : 1.0;
        double g[NI][NK];
        rys_2d<NI, NK>(g, pa - fa * pqd, qc + fc * pqd, b10, b01, b00, g00);
        double *dst = Gt + tt * S::GS + (dim * NR + r) * GSZ;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
          for (int k = 0; k < NK; ++k) dst[i * NK + k] = g[i][k];
      }
    };
    // KT primitive quartets per pass: (rule records when a batch is due),
    // tables, then every lane accumulates them
    for (int t0 = 0; t0 < maxpq; t0 += KT) {
      if (t0 % RT == 0) {  // the last batch's readers are past a fence
        rules(t0);
        wave_sync();
      }
      tables(t0);
      wave_sync();
#pragma unroll
      for (int tt = 0; tt < KT; ++tt) {
        const int t = t0 + tt;
        const double *cur = Gt + tt * S::GS;
        if constexpr (S::MAP == 0) {
          if (t < npq) {
#pragma unroll
            for (int k = 0; k < NKC; ++k) {
              double v = 0.0;
#pragma unroll
              for (int r = 0; r < NR; ++r)
                v += cur[r * GSZ + offx[k]] * cur[(NR + r) * GSZ + offy[k]] * cur[(2 * NR + r) * GSZ + offz[k]];
              acc[k] += v;
            }
          }
        } else {
          if (t < npq && gl < G) {
            // MAP 1: loop over the columns f; MAP 2: over the rows e
            constexpr int L0 = S::MAP == 1 ? LC : LA, L1 = S::MAP == 1 ? S::LCD : S::LAB;
            constexpr int NJ = S::MAP == 1 ? NK : NI, SJ = S::MAP == 1 ? 1 : NK;
            constexpr Range<L0, L1> rl{};
            const double *bx = cur + part * GSZ + rowx, *by = cur + (NR + part) * GSZ + rowy,
                         *bz = cur + (2 * NR + part) * GSZ + rowz;
#pragma unroll
            for (int rr = 0; rr < NRP; ++rr) {
              if (NR % P == 0 || part + rr * P < NR) {
                double X[NJ], Y[NJ], Z[NJ];
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                  X[j] = bx[rr * P * GSZ + j * SJ];
                  Y[j] = by[rr * P * GSZ + j * SJ];
                  Z[j] = bz[rr * P * GSZ + j * SJ];
                }
#pragma unroll
                for (int k = 0; k < NACC; ++k) acc[k] += X[rl.lx[k]] * Y[rl.ly[k]] * Z[rl.lz[k]];
              }
            }
          }
        }
      }
      wave_sync();
    }
    // (e0|f0) to LDS: rows e of stride NF for the bra HRR, or directly the
    // ket-phase rows [a][KS] when there is no bra HRR
    constexpr int ACCS = LB > 0 ? NF : S::KS;
    if constexpr (S::MAP == 0) {
#pragma unroll
      for (int k = 0; k < NKC; ++k) {
        const int c = gl + G * k;
        if (c < NE * NF) {
          const int e = c / NF;
          buf0[e * ACCS + (c - e * NF)] = acc[k];
        }
      }
    } else {
      // the P partial sums of a row (column) sit on adjacent lanes
#pragma unroll
      for (int k = 0; k < NACC; ++k) {
        double v = acc[k];
#pragma unroll
        for (int o = 1; o < P; ++o) v += __shfl(acc[k], lane + o, 64);
        acc[k] = v;
      }
      if (gl < G && part == 0) {
        const int rl = gl / P;
#pragma unroll
        for (int k = 0; k < NACC; ++k) {
          if (S::MAP == 1)
            buf0[rl * ACCS + k] = acc[k];
          else
            buf0[k * ACCS + rl] = acc[k];
        }
      }
    }
    wave_sync();

    if constexpr (S::REG) {
      constexpr int NAB = S::NAB, NCD = S::NCD, TS = S::TS, JS = S::JS;
      // the six density blocks, after the buffers
      double *DAB = Gt + S::UNI, *DCD = DAB + NAB, *DAC = DCD + NCD, *DBD = DAC + S::NAC, *DAD = DBD + S::NBD,
             *DBC = DAD + S::NAD;
      if (valid) {
        const double *Dm = ba.D, *Dk = ba.DK;  // Coulomb and exchange densities
        for (int i = gl; i < NAB; i += G) DAB[i] = Dm[tri(bra.bfA + gpos_rt(LA, i / NB), bra.bfB + gpos_rt(LB, i % NB))];
        for (int i = gl; i < NCD; i += G) DCD[i] = Dm[tri(ket.bfA + gpos_rt(LC, i / ND), ket.bfB + gpos_rt(LD, i % ND))];
        for (int i = gl; i < S::NAC; i += G) DAC[i] = Dk[tri(bra.bfA + gpos_rt(LA, i / NC), ket.bfA + gpos_rt(LC, i % NC))];
        for (int i = gl; i < S::NBD; i += G) DBD[i] = Dk[tri(bra.bfB + gpos_rt(LB, i / ND), ket.bfB + gpos_rt(LD, i % ND))];
        for (int i = gl; i < S::NAD; i += G) DAD[i] = Dk[tri(bra.bfA + gpos_rt(LA, i / ND), ket.bfB + gpos_rt(LD, i % ND))];
        for (int i = gl; i < S::NBC; i += G) DBC[i] = Dk[tri(bra.bfB + gpos_rt(LB, i / NC), ket.bfA + gpos_rt(LC, i % NC))];
      }
      // bra HRR: one column f per lane, in registers; (ab|f) to [ab][f]
      double *bufT = Gt + NE * NF;
      for (int f = gl; f < NF; f += G) {
        double c0[S::CBUF], c1[S::CBUF];
#pragma unroll
        for (int e = 0; e < NE; ++e) c0[e] = buf0[e * NF + f];
        hrr_bra<LA, LB, 1>(c0, c1, bra.ABx, bra.ABy, bra.ABz);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab) bufT[ab * TS + f] = (LB % 2) ? c1[ab] : c0[ab];
      }
      wave_sync();
      // ket HRR: rows ab = gl, gl + G, ... of the lane, in registers, then
      // the cartesian scales and the value cutoff
      constexpr int NJ = (NAB + G - 1) / G;
      double v[NJ][NCD];
#pragma unroll
      for (int k = 0; k < NJ; ++k) {
        const int r = gl + k * G;
        if (valid && r < NAB) {
          double r0[S::RBUF], r1[S::RBUF];
#pragma unroll
          for (int f = 0; f < NF; ++f) r0[f] = bufT[r * TS + f];
          hrr_ket<LC, LD, 1>(r0, r1, ket.ABx, ket.ABy, ket.ABz);
          constexpr Range<LC, LC> rc{};
          constexpr Range<LD, LD> rd{};
          const double sab = scale_rt(LA, r / NB) * scale_rt(LB, r % NB);
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
              const double x = ((LD % 2) ? r1[c * ND + d] : r0[c * ND + d]) * sab *
                               (cart_scale(rc.lx[c], rc.ly[c], rc.lz[c]) * cart_scale(rd.lx[d], rd.ly[d], rd.lz[d]));
              v[k][c * ND + d] = fabs(x) < ba.value_cut ? 0.0 : x;
            }
        }
      }
      wave_sync();  // [ab][f] read: its space takes the partial sums
      // digestion: each lane's row ab contributes to the six blocks; the
      // sums over the other rows go through LDS
      double s = 1.0;
      if (bra.sameAB) s *= 0.5;
      if (ket.sameAB) s *= 0.5;
      if (cd.diag && ib == ik) s *= 0.5;
      double *JC = Gt, *KAC = JC + NAB * JS, *KBD = KAC + NAB * NC, *KAD = KBD + NAB * ND, *KBC = KAD + NAB * ND;
      if (valid && jab_pair != ib) {
        flush_jab();
        jab_pair = ib;
        jab_a = bra.bfA;
        jab_b = bra.bfB;
      }
#pragma unroll
      for (int k = 0; k < NJ; ++k) {
        const int r = gl + k * G;
        if (!(valid && r < NAB)) continue;
        const double *w = v[k];
        const int a = r / NB, b = r - a * NB;
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < NCD; ++i) t += w[i] * DCD[i];
        jab[k] += 4.0 * s * t;  // J(AB): the row's own element, in registers over the bra pair's run
        const double dab = DAB[r];
#pragma unroll
        for (int i = 0; i < NCD; ++i) JC[r * JS + i] = w[i] * dab;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double kac = 0.0, kbc = 0.0;
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            kac += w[c * ND + d] * DBD[b * ND + d];
            kbc += w[c * ND + d] * DAD[a * ND + d];
          }
          KAC[(a * NB + b) * NC + c] = kac;
          KBC[(b * NA + a) * NC + c] = kbc;
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          double kbd = 0.0, kad = 0.0;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            kbd += w[c * ND + d] * DAC[a * NC + c];
            kad += w[c * ND + d] * DBC[b * NC + c];
          }
          KBD[(b * NA + a) * ND + d] = kbd;
          KAD[(a * NB + b) * ND + d] = kad;
        }
      }
      wave_sync();
      if (valid) {
        double *F = ba.F;
        const double sk = s * ba.k_scale;
        auto put = [&](int64_t fi, int64_t fj, double x) {
          if (x != 0.0) fock_add(F, fi, fj, x);
        };
        for (int i = gl; i < NCD; i += G) {
          double t = 0.0;
#pragma unroll
          for (int ab = 0; ab < NAB; ++ab) t += JC[ab * JS + i];
          put(ket.bfA + gpos_rt(LC, i / ND), ket.bfB + gpos_rt(LD, i % ND), 4.0 * s * t);
        }
        for (int i = gl; i < NA * NC; i += G) {
          const int a = i / NC, c = i - a * NC;
          double t = 0.0;
#pragma unroll
          for (int b = 0; b < NB; ++b) t += KAC[(a * NB + b) * NC + c];
          put(bra.bfA + gpos_rt(LA, a), ket.bfA + gpos_rt(LC, c), -sk * t);
        }
        for (int i = gl; i < NB * ND; i += G) {
          const int b = i / ND, d = i - b * ND;
          double t = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a) t += KBD[(b * NA + a) * ND + d];
          put(bra.bfB + gpos_rt(LB, b), ket.bfB + gpos_rt(LD, d), -sk * t);
        }
        for (int i = gl; i < NA * ND; i += G) {
          const int a = i / ND, d = i - a * ND;
          double t = 0.0;
#pragma unroll
          for (int b = 0; b < NB; ++b) t += KAD[(a * NB + b) * ND + d];
          put(bra.bfA + gpos_rt(LA, a), ket.bfB + gpos_rt(LD, d), -sk * t);
        }
        for (int i = gl; i < NB * NC; i += G) {
          const int b = i / NC, c = i - b * NC;
          double t = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a) t += KBC[(b * NA + a) * NC + c];
          put(bra.bfB + gpos_rt(LB, b), ket.bfA + gpos_rt(LC, c), -sk * t);
        }
      }
    } else {
    // bra HRR, lanes over the columns f; the last level writes rows [ab] of
    // stride KS for the ket HRR, whose lanes run over the rows ab.  Component
    // loops are compile-time, so every LDS address is a lane base plus a
    // constant and the loads of one level issue back to back.
    double *in = buf0, *out = buf1;
    if constexpr (LB >= 1) {
      if constexpr (grp_pair_form(NF, nrange(LA, S::LAB - 1), ncart(1), G))
        grp_bra_level_af<LA, S::LAB, 1, NF, G, (LB == 1 ? S::KS : NF)>(in, out, gl, bra.ABx, bra.ABy, bra.ABz);
      else
        grp_bra_level<LA, S::LAB, 1, NF, G, (LB == 1 ? S::KS : NF)>(in, out, gl, bra.ABx, bra.ABy, bra.ABz);
      wave_sync();
      double *t = in; in = out; out = t;
    }
    if constexpr (LB >= 2) {
      if constexpr (grp_pair_form(NF, nrange(LA, S::LAB - 2), ncart(2), G))
        grp_bra_level_af<LA, S::LAB, 2, NF, G, (LB == 2 ? S::KS : NF)>(in, out, gl, bra.ABx, bra.ABy, bra.ABz);
      else
        grp_bra_level<LA, S::LAB, 2, NF, G, (LB == 2 ? S::KS : NF)>(in, out, gl, bra.ABx, bra.ABy, bra.ABz);
      wave_sync();
      double *t = in; in = out; out = t;
    }
    if constexpr (LB >= 3) {
      if constexpr (grp_pair_form(NF, nrange(LA, S::LAB - 3), ncart(3), G))
        grp_bra_level_af<LA, S::LAB, 3, NF, G, S::KS>(in, out, gl, bra.ABx, bra.ABy, bra.ABz);
      else
        grp_bra_level<LA, S::LAB, 3, NF, G, S::KS>(in, out, gl, bra.ABx, bra.ABy, bra.ABz);
      wave_sync();
      double *t = in; in = out; out = t;
    }
    if constexpr (LD >= 1) {
      if constexpr (grp_pair_form(S::NAB, nrange(LC, S::LCD - 1), ncart(1), G))
        grp_ket_level_rc<LC, S::LCD, 1, S::NAB, G, S::KS>(in, out, gl, ket.ABx, ket.ABy, ket.ABz);
      else
        grp_ket_level<LC, S::LCD, 1, S::NAB, G, S::KS>(in, out, gl, ket.ABx, ket.ABy, ket.ABz);
      wave_sync();
      double *t = in; in = out; out = t;
    }
    if constexpr (LD >= 2) {
      if constexpr (grp_pair_form(S::NAB, nrange(LC, S::LCD - 2), ncart(2), G))
        grp_ket_level_rc<LC, S::LCD, 2, S::NAB, G, S::KS>(in, out, gl, ket.ABx, ket.ABy, ket.ABz);
      else
        grp_ket_level<LC, S::LCD, 2, S::NAB, G, S::KS>(in, out, gl, ket.ABx, ket.ABy, ket.ABz);
      wave_sync();
      double *t = in; in = out; out = t;
    }
    if constexpr (LD >= 3) {
      if constexpr (grp_pair_form(S::NAB, nrange(LC, S::LCD - 3), ncart(3), G))
        grp_ket_level_rc<LC, S::LCD, 3, S::NAB, G, S::KS>(in, out, gl, ket.ABx, ket.ABy, ket.ABz);
      else
        grp_ket_level<LC, S::LCD, 3, S::NAB, G, S::KS>(in, out, gl, ket.ABx, ket.ABy, ket.ABz);
      wave_sync();
      double *t = in; in = out; out = t;
    }
    constexpr int KS = S::KS, NCD = NC * ND;
    // the six density blocks: in the free HRR buffer (at least BUFP doubles
    // from `out`) or after the buffers
    double *DAB = S::DBX ? Gt + S::UNI : out, *DCD = DAB + S::NAB, *DAC = DCD + S::NCD, *DBD = DAC + S::NAC,
           *DAD = DBD + S::NBD, *DBC = DAD + S::NAD;
    // scale + value cutoff in place (rows ab of stride KS); gather the six density blocks
    if (valid) {
      constexpr Range<LC, LC> rc{};
      constexpr Range<LD, LD> rd{};
      for (int ab = gl; ab < S::NAB; ab += G) {
        const double sab = scale_rt(LA, ab / NB) * scale_rt(LB, ab % NB);
        double *row = in + ab * KS;
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < ND; ++d) {
            const double x = row[c * ND + d] * sab * (cart_scale(rc.lx[c], rc.ly[c], rc.lz[c]) *
                                                      cart_scale(rd.lx[d], rd.ly[d], rd.lz[d]));
            row[c * ND + d] = fabs(x) < ba.value_cut ? 0.0 : x;
          }
      }
      const double *Dm = ba.D, *Dk = ba.DK;  // Coulomb and exchange densities
      for (int i = gl; i < S::NAB; i += G) DAB[i] = Dm[tri(bra.bfA + gpos_rt(LA, i / NB), bra.bfB + gpos_rt(LB, i % NB))];
      for (int i = gl; i < NCD; i += G) DCD[i] = Dm[tri(ket.bfA + gpos_rt(LC, i / ND), ket.bfB + gpos_rt(LD, i % ND))];
      for (int i = gl; i < S::NAC; i += G) DAC[i] = Dk[tri(bra.bfA + gpos_rt(LA, i / NC), ket.bfA + gpos_rt(LC, i % NC))];
      for (int i = gl; i < S::NBD; i += G) DBD[i] = Dk[tri(bra.bfB + gpos_rt(LB, i / ND), ket.bfB + gpos_rt(LD, i % ND))];
      for (int i = gl; i < S::NAD; i += G) DAD[i] = Dk[tri(bra.bfA + gpos_rt(LA, i / ND), ket.bfB + gpos_rt(LD, i % ND))];
      for (int i = gl; i < S::NBC; i += G) DBC[i] = Dk[tri(bra.bfB + gpos_rt(LB, i / NC), ket.bfA + gpos_rt(LC, i % NC))];
    }
    wave_sync();
    if (valid) {
      double s = 1.0;
      if (bra.sameAB) s *= 0.5;
      if (ket.sameAB) s *= 0.5;
      if (cd.diag && ib == ik) s *= 0.5;
      const double *__restrict__ v = in;
      double *F = ba.F;
      auto put = [&](int64_t fi, int64_t fj, double x) {
        if (x != 0.0) fock_add(F, fi, fj, x);
      };
      // J(AB): F[ab] += 4s sum_cd v D[cd], summed in registers while the
      // slot's bra pair stays the same
      if (jab_pair != ib) {
        flush_jab();
        jab_pair = ib;
        jab_a = bra.bfA;
        jab_b = bra.bfB;
      }
#pragma unroll
      for (int k = 0; k < NJL; ++k) {
        const int i = gl + k * G;
        if (i < S::NAB) {
          double t = 0.0;
#pragma unroll
          for (int e = 0; e < NCD; ++e) t += v[i * KS + e] * DCD[e];
          jab[k] += 4.0 * s * t;
        }
      }
      // The other five blocks: an output's sum runs over two component
      // indices; when the slot's lanes outnumber the outputs, the outer one is
      // split over SP adjacent lanes (part p takes p, p + SP, ...) whose
      // partial sums meet in part 0.  Offsets stay compile-time past a
      // per-lane base.
      // J(CD): F[cd] += 4s sum_ab v D[ab], rows ab split
      {
        constexpr int SP = split_lanes(NCD, S::NAB, G);
        for (int i0 = gl; i0 < NCD * SP; i0 += G) {
          const int i = i0 / SP, p = i0 - i * SP, c = i / ND, d = i - c * ND;
          const double *vb = v + p * KS + i, *db = DAB + p;
          double t = 0.0;
#pragma unroll
          for (int k = 0; k * SP < S::NAB; ++k)
            if (S::NAB % SP == 0 || p + k * SP < S::NAB) t += vb[k * SP * KS] * db[k * SP];
          t = part_sum<SP>(t);
          if (p == 0) put(ket.bfA + gpos_rt(LC, c), ket.bfB + gpos_rt(LD, d), 4.0 * s * t);
        }
      }
      // K(AC): F[ac] -= s sum_bd v D[bd], b split
      {
        constexpr int SP = split_lanes(S::NAC, NB, G);
        for (int i0 = gl; i0 < S::NAC * SP; i0 += G) {
          const int i = i0 / SP, p = i0 - i * SP, a = i / NC, c = i - a * NC;
          const double *vb = v + (a * NB + p) * KS + c * ND, *db = DBD + p * ND;
          double t = 0.0;
#pragma unroll
          for (int k = 0; k * SP < NB; ++k)
            if (NB % SP == 0 || p + k * SP < NB)
#pragma unroll
              for (int d = 0; d < ND; ++d) t += vb[k * SP * KS + d] * db[k * SP * ND + d];
          t = part_sum<SP>(t);
          if (p == 0) put(bra.bfA + gpos_rt(LA, a), ket.bfA + gpos_rt(LC, c), -s * ba.k_scale * t);
        }
      }
      // K(BD): F[bd] -= s sum_ac v D[ac], a split
      {
        constexpr int SP = split_lanes(S::NBD, NA, G);
        for (int i0 = gl; i0 < S::NBD * SP; i0 += G) {
          const int i = i0 / SP, p = i0 - i * SP, b = i / ND, d = i - b * ND;
          const double *vb = v + (p * NB + b) * KS + d, *db = DAC + p * NC;
          double t = 0.0;
#pragma unroll
          for (int k = 0; k * SP < NA; ++k)
            if (NA % SP == 0 || p + k * SP < NA)
#pragma unroll
              for (int c = 0; c < NC; ++c) t += vb[k * SP * NB * KS + c * ND] * db[k * SP * NC + c];
          t = part_sum<SP>(t);
          if (p == 0) put(bra.bfB + gpos_rt(LB, b), ket.bfB + gpos_rt(LD, d), -s * ba.k_scale * t);
        }
      }
      // K(AD): F[ad] -= s sum_bc v D[bc], b split
      {
        constexpr int SP = split_lanes(S::NAD, NB, G);
        for (int i0 = gl; i0 < S::NAD * SP; i0 += G) {
          const int i = i0 / SP, p = i0 - i * SP, a = i / ND, d = i - a * ND;
          const double *vb = v + (a * NB + p) * KS + d, *db = DBC + p * NC;
          double t = 0.0;
#pragma unroll
          for (int k = 0; k * SP < NB; ++k)
            if (NB % SP == 0 || p + k * SP < NB)
#pragma unroll
              for (int c = 0; c < NC; ++c) t += vb[k * SP * KS + c * ND] * db[k * SP * NC + c];
          t = part_sum<SP>(t);
          if (p == 0) put(bra.bfA + gpos_rt(LA, a), ket.bfB + gpos_rt(LD, d), -s * ba.k_scale * t);
        }
      }
      // K(BC): F[bc] -= s sum_ad v D[ad], a split
      {
        constexpr int SP = split_lanes(S::NBC, NA, G);
        for (int i0 = gl; i0 < S::NBC * SP; i0 += G) {
          const int i = i0 / SP, p = i0 - i * SP, b = i / NC, c = i - b * NC;
          const double *vb = v + (p * NB + b) * KS + c * ND, *db = DAD + p * ND;
          double t = 0.0;
#pragma unroll
          for (int k = 0; k * SP < NA; ++k)
            if (NA % SP == 0 || p + k * SP < NA)
#pragma unroll
              for (int d = 0; d < ND; ++d) t += vb[k * SP * NB * KS + d] * db[k * SP * ND + d];
          t = part_sum<SP>(t);
          if (p == 0) put(bra.bfB + gpos_rt(LB, b), ket.bfA + gpos_rt(LC, c), -s * ba.k_scale * t);
        }
      }
    }
    }
    wave_sync();  // slots are reused by the next round
  };

  // the quartets [c0, c1) of the class's share, NQ per round: consecutive
  // rounds stay in the same or the next segment, so each lane searches the
  // segments once and then steps forward, and they share the bra pair (its
  // records and density blocks stay in cache)
  int seg = -1;
  for (int64_t base = c0; base < c1; base += NQ) {
    const bool valid = !idle && base + slot < c1;
    int32_t ib = 0, ik = 0;
    if (valid) {
      const int64_t q = cd.q_begin + base + slot;
      if (seg < 0) seg = find_segment(cd, q, 0, cd.nseg);
      while (seg + 1 < cd.nseg && cd.seg_start[seg + 1] <= q) ++seg;
      quartet_of_segment(cd, seg, q, ib, ik);
    }
    round(valid, ib, ik);
  }
  flush_jab();
}

template <int LA, int LB, int LC, int LD>
__attribute__((amdgpu_waves_per_eu(grp_reg_hrr(LA, LB, LC, LD) && LA + LB + LC + LD == 8               ? 2
                                   : LA <= 2 && LB <= 2 && LC <= 2 && LD <= 2 && LA + LB + LC + LD <= 7 ? 3
                                                                                                        : 1,
                                   8)))
__global__ __launch_bounds__(64) void eri_group_kernel(ClassDev cd, BuildArgs ba) {
  using S = Grp<LA, LB, LC, LD>;
  constexpr int NQ = S::NQ;
  __shared__ double lds[NQ * S::SLOT];
  __shared__ int s_npq[NQ], s_nk[NQ], s_boff[NQ], s_koff[NQ];
  // a contiguous chunk of the class's quartets per workgroup (a multiple of
  // NQ).  Workgroups resident at the same time take chunks spread over the
  // whole quartet range (chunk = workgroup x chunk_stride(grid) mod grid, a
  // permutation), so their atomics meet fewer common Fock rows.
  const int64_t nq = cd.q_end - cd.q_begin;
  const int64_t per = ((nq + gridDim.x - 1) / gridDim.x + NQ - 1) / NQ * NQ;
  const int64_t chunk = (int64_t)blockIdx.x * cd.chunk_stride % gridDim.x;
  const int64_t c0 = chunk * per, c1 = c0 + per < nq ? c0 + per : nq;
  grp_run<LA, LB, LC, LD>(cd, ba, c0, c1, lds, s_npq, s_nk, s_boff, s_koff);
}

// ----------------------------------------------------- segment lengths
#if !defined(LIBERI_CLASS_GROUP) || LIBERI_CLASS_GROUP == -1 || LIBERI_CLASS_GROUP == -2
__global__ void segment_lengths_kernel(const double *bra_x, const int32_t *bra_group, const int32_t *bra_pos,
                                       int32_t nbra, const double *ket_x, const int32_t *gbeg, const int32_t *gend,
                                       int32_t ngroup, int32_t diag, int32_t cmp, double cut, int64_t *seg_len) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)nbra * ngroup) return;
  const int i = (int)(id / ngroup), g = (int)(id % ngroup);
  const double xi = bra_x[i];
  int lo = gbeg[g], hi = gend[g];
  int limit = hi;
  if (diag) {
    const int gi = bra_group[i];
    if (g > gi) {
      seg_len[id] = 0;
      return;
    }
    if (g == gi) limit = bra_pos[i] + 1;  // kets up to and including the bra itself
  }
  // first position in [lo, hi) whose product fails the test (x descending)
  int a = lo, b = hi;
  if (cmp != CMP_NONE) {
    while (a < b) {
      const int m = (a + b) >> 1;
      const double tst = xi * ket_x[m];
      const bool pass = cmp == CMP_GT ? (tst > cut) : !(tst < cut);
      if (pass)
        a = m + 1;
      else
        b = m;
    }
  } else {
    a = hi;
  }
  const int end = a < limit ? a : limit;
  seg_len[id] = end > lo ? end - lo : 0;
}

// ------------------------------------------------------ Schwarz integrals
// Schwarz value of a shell pair: sqrt(max over components |(ab|ab)|) of the
// normalised cartesian components, the quantity the reference takes as
// schwrz_int (src/interface/liberi_interface.F90:128).  One lane per pair;
// the 2-D Rys integrals G(i, k) of each root (i on A, k on C = A) are moved
// to the diagonal (a b | a b) in one step with the binomial form of the
// transfer relation, I(a,b,a,b) = sum_t sum_u C(b,t) C(b,u) AB^(2b-t-u)
// G(a+t, a+u), since the ket pair is the bra pair (CD = AB).
__host__ __device__ constexpr int binom(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

template <int LA, int LB, int NI>
__device__ __forceinline__ void diag_transfer(const double (&g)[NI][NI], double ab, double (&h)[LA + 1][LB + 1]) {
  double abp[2 * LB + 1];
  abp[0] = 1.0;
#pragma unroll
  for (int k = 1; k <= 2 * LB; ++k) abp[k] = abp[k - 1] * ab;
#pragma unroll
  for (int a = 0; a <= LA; ++a)
#pragma unroll
    for (int b = 0; b <= LB; ++b) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t <= b; ++t)
#pragma unroll
        for (int u = 0; u <= b; ++u) s += (double)(binom(b, t) * binom(b, u)) * abp[2 * b - t - u] * g[a + t][a + u];
      h[a][b] = s;
    }
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void schwarz_kernel(const PairRec *pairs, const PrimRec *prims, const int32_t *list,
                                                     int32_t n, double xswitch, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int L = LA + LB, NR = L + 1, NI = L + 1;
  constexpr int NA = ncart(LA), NB = ncart(LB);
  constexpr Range<LA, LA> ra{};
  constexpr Range<LB, LB> rb{};
  const PairRec pr = pairs[list[i]];
  double acc[NA * NB];
#pragma unroll
  for (int k = 0; k < NA * NB; ++k) acc[k] = 0.0;
  for (int ip = 0; ip < pr.nprim; ++ip) {
    const PrimRec P = prims[pr.prim_off + ip];
    for (int iq = 0; iq < pr.nprim; ++iq) {
      const PrimRec Q = prims[pr.prim_off + iq];
      const double pq = P.p + Q.p, ipq = 1.0 / pq;
      const double pqx = P.Px - Q.Px, pqy = P.Py - Q.Py, pqz = P.Pz - Q.Pz;
      const double X = P.p * Q.p * ipq * (pqx * pqx + pqy * pqy + pqz * pqz);
      const double pref = kTwoPi52 * P.K * Q.K * sqrt(ipq);
      double u[NR], w[NR];
      rys_rule<NR>(X, xswitch, u, w);
      const double hp = 0.5 / P.p, hq = 0.5 / Q.p;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double t2 = u[r];
        const double b00 = 0.5 * t2 * ipq;
        const double b10 = hp - hp * Q.p * t2 * ipq;
        const double b01 = hq - hq * P.p * t2 * ipq;
        const double fa = Q.p * ipq * t2, fc = P.p * ipq * t2;
        double h[3][LA + 1][LB + 1];
        {
          double g[NI][NI];
          rys_2d<NI, NI>(g, P.PAx - fa * pqx, Q.PAx + fc * pqx, b10, b01, b00, 1.0);
          diag_transfer<LA, LB, NI>(g, pr.ABx, h[0]);
          rys_2d<NI, NI>(g, P.PAy - fa * pqy, Q.PAy + fc * pqy, b10, b01, b00, 1.0);
          diag_transfer<LA, LB, NI>(g, pr.ABy, h[1]);
          rys_2d<NI, NI>(g, P.PAz - fa * pqz, Q.PAz + fc * pqz, b10, b01, b00, w[r] * pref);
          diag_transfer<LA, LB, NI>(g, pr.ABz, h[2]);
        }
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b)
            acc[a * NB + b] += h[0][ra.lx[a]][rb.lx[b]] * h[1][ra.ly[a]][rb.ly[b]] * h[2][ra.lz[a]][rb.lz[b]];
      }
    }
  }
  double mx = 0.0;
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double sc = cart_scale(ra.lx[a], ra.ly[a], ra.lz[a]) * cart_scale(rb.lx[b], rb.ly[b], rb.lz[b]);
      const double v = fabs(acc[a * NB + b]) * sc * sc;
      mx = v > mx ? v : mx;
    }
  out[i] = sqrt(mx);
}

template <int LA, int LB>
static int launch_schwarz_t(const PairRec *pairs, const PrimRec *prims, const int32_t *list, int32_t n, double *out,
                            hipStream_t st) {
  const int tpb = 64;
  hipLaunchKernelGGL((schwarz_kernel<LA, LB>), dim3((unsigned)((n + tpb - 1) / tpb)), dim3(tpb), 0, st, pairs,
                     prims, list, n, rys_xmax_host[LA + LB + 1], out);
  return hipGetLastError();
}

int launch_schwarz(int la, int lb, const PairRec *pairs, const PrimRec *prims, const int32_t *list, int32_t n,
                   double *out, void *stream) {
  if (n <= 0) return hipSuccess;
  hipStream_t st = (hipStream_t)stream;
  switch (la * 4 + lb) {
    case 0: return launch_schwarz_t<0, 0>(pairs, prims, list, n, out, st);
    case 4: return launch_schwarz_t<1, 0>(pairs, prims, list, n, out, st);
    case 5: return launch_schwarz_t<1, 1>(pairs, prims, list, n, out, st);
    case 8: return launch_schwarz_t<2, 0>(pairs, prims, list, n, out, st);
    case 9: return launch_schwarz_t<2, 1>(pairs, prims, list, n, out, st);
    case 10: return launch_schwarz_t<2, 2>(pairs, prims, list, n, out, st);
    case 12: return launch_schwarz_t<3, 0>(pairs, prims, list, n, out, st);
    case 13: return launch_schwarz_t<3, 1>(pairs, prims, list, n, out, st);
    case 14: return launch_schwarz_t<3, 2>(pairs, prims, list, n, out, st);
    case 15: return launch_schwarz_t<3, 3>(pairs, prims, list, n, out, st);
    default: return hipErrorInvalidValue;
  }
}

// F = 0 as a kernel (not a memset node) so that a build can be captured in a graph
__global__ void fill_zero_kernel(double *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// max |p| over the array: block maximum, then one atomic max per block on the
// bit pattern (order-preserving for non-negative doubles)
__global__ void abs_max_kernel(const double *p, int64_t n, double *out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(p[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax((unsigned long long *)out, (unsigned long long)__double_as_longlong(red[0]));
}

int launch_abs_max(const double *p, int64_t n, double *out, void *stream) {
  if (n <= 0) return hipSuccess;
  const int64_t need = (n + 255) / 256;
  const unsigned g = (unsigned)(need < 1024 ? need : 1024);
  hipLaunchKernelGGL(abs_max_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, p, n, out);
  return hipGetLastError();
}

__global__ void add_kernel(const double *a, const double *b, double *out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] + b[i];
}

int launch_add(const double *a, const double *b, double *out, int64_t n, void *stream) {
  if (n <= 0) return hipSuccess;
  const int64_t need = (n + 255) / 256;
  const unsigned g = (unsigned)(need < 4096 ? need : 4096);
  hipLaunchKernelGGL(add_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, a, b, out, n);
  return hipGetLastError();
}

int launch_fill_zero(double *p, int64_t n, void *stream) {
  if (n <= 0) return hipSuccess;
  const int tpb = 256;
  const int64_t need = (n + tpb - 1) / tpb;
  const unsigned g = (unsigned)(need < 4096 ? need : 4096);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(g), dim3(tpb), 0, (hipStream_t)stream, p, n);
  return hipGetLastError();
}

int launch_segment_lengths(const double *bra_x, const int32_t *bra_group, const int32_t *bra_pos, int32_t nbra,
                           const double *ket_x, const int32_t *gbeg, const int32_t *gend, int32_t ngroup,
                           int32_t diag, int32_t cmp, double cut, int64_t *seg_len, void *stream) {
  const int64_t n = (int64_t)nbra * ngroup;
  if (n == 0) return hipSuccess;
  const int tpb = 256;
  hipLaunchKernelGGL(segment_lengths_kernel, dim3((unsigned)((n + tpb - 1) / tpb)), dim3(tpb), 0,
                     (hipStream_t)stream, bra_x, bra_group, bra_pos, nbra, ket_x, gbeg, gend, ngroup, diag, cmp,
                     cut, seg_len);
  return hipGetLastError();
}
#endif

// ----------------------------------------------------------- dispatch
template <int LA, int LB, int LC, int LD>
constexpr bool use_small() {
  return nrange(LA, LA + LB) * nrange(LC, LC + LD) <= 100 && hrr_max_buffer(LA, LB, LC, LD) <= 100;
}

// LDS bytes of the row-tiled kernel of the class in orientation TR, or -1
// when the class has no such kernel
// (dynamic) LDS bytes of the row accumulators of the class in orientation TR
template <int LA, int LB, int LC, int LD, bool TR>
int64_t row_acc_bytes(const int32_t *ncol) {
  using R = RowShape<LA, LB, LC, LD, TR>;
  return (int64_t)sizeof(double) * (R::NFIX + R::NROW * row_columns<LA, LB, LC, LD, TR>(ncol));
}

// LDS bytes of the row-tiled form of the class in orientation TR (the lane
// kernel, or the wave-split kernel of a group class), or -1 when the class
// has none or its LDS exceeds what that kernel may take
template <int LA, int LB, int LC, int LD, bool TR>
int64_t row_lds(const int32_t *ncol, double xswitch) {
  using R = RowShape<LA, LB, LC, LD, TR>;
  if constexpr (!R::exists) {
    return -1;
  } else if constexpr (SplitShape<LA, LB, LC, LD>::exists) {
    const int64_t b = split_lds<LA, LB, LC, LD, TR>(ncol, xswitch);
    return b <= kSplitLdsBytes ? b : -1;
  } else if constexpr (use_small<LA, LB, LC, LD>()) {
    const int64_t b = row_acc_bytes<LA, LB, LC, LD, TR>(ncol) +
                      (int64_t)sizeof(double) * rys_table_used(Small<LA, LB, LC, LD>::NR, xswitch);
    return b <= kRowLdsBytes ? b : -1;
  } else {
    return -1;
  }
}

// group classes with an f shell also get the cooperative form when its
// buffers fit the LDS; which one runs is timed per class
template <int LA, int LB, int LC, int LD>
constexpr bool has_coop_alt() {
  return !use_small<LA, LB, LC, LD>() && Grp<LA, LB, LC, LD>::fits && Coop<LA, LB, LC, LD>::IN_LDS &&
         (LA == 3 || LB == 3 || LC == 3 || LD == 3);
}

template <int LA, int LB, int LC, int LD, bool TR>
static int launch_row(const ClassDev &cd, const BuildArgs &ba, hipStream_t st) {
  const int64_t lds = row_lds<LA, LB, LC, LD, TR>(ba.ncol, cd.xswitch);
  if (lds < 0 || cd.t_end <= cd.t_begin) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(cd.t_end - cd.t_begin));
  if constexpr (!RowShape<LA, LB, LC, LD, TR>::exists) {
    return hipErrorInvalidValue;
  } else if constexpr (use_small<LA, LB, LC, LD>() && !SplitShape<LA, LB, LC, LD>::exists) {
    hipLaunchKernelGGL((eri_row_kernel<LA, LB, LC, LD, TR>), grid, dim3(row_threads<LA, LB, LC, LD>()), (unsigned)lds,
                       st, cd, ba);
  } else if constexpr (SplitShape<LA, LB, LC, LD>::exists) {
    // (its dynamic LDS limit was raised by prepare_class at setup)
    hipLaunchKernelGGL((eri_split_kernel<LA, LB, LC, LD, TR>), grid, dim3(SplitShape<LA, LB, LC, LD>::NT),
                       (unsigned)lds, st, cd, ba);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
static int launch_one(const ClassDev &cd, const BuildArgs &ba, hipStream_t st, int nwg) {
  const int64_t n = cd.q_end - cd.q_begin;
  if (n <= 0) return hipSuccess;
  if constexpr (use_small<LA, LB, LC, LD>()) {
    const int tpb = 256;
    if (cd.row_mode) {
      return cd.transposed ? launch_row<LA, LB, LC, LD, true>(cd, ba, st) : launch_row<LA, LB, LC, LD, false>(cd, ba, st);
    } else {
      hipLaunchKernelGGL((eri_small_kernel<LA, LB, LC, LD>), dim3((unsigned)((n + tpb - 1) / tpb)), dim3(tpb), 0,
                         st, cd, ba);
    }
  } else if constexpr (Grp<LA, LB, LC, LD>::fits) {
    if (cd.row_mode) {
      if constexpr (SplitShape<LA, LB, LC, LD>::exists)
        return cd.transposed ? launch_row<LA, LB, LC, LD, true>(cd, ba, st)
                             : launch_row<LA, LB, LC, LD, false>(cd, ba, st);
      else
        return hipErrorInvalidValue;
    }
    if (cd.coop) {
      if constexpr (has_coop_alt<LA, LB, LC, LD>()) {
        using S = Coop<LA, LB, LC, LD>;
        const int64_t g = n < nwg ? n : nwg;
        hipLaunchKernelGGL((eri_coop_kernel<LA, LB, LC, LD>), dim3((unsigned)g), dim3(S::NT), 0, st, cd, ba);
        return hipGetLastError();
      } else {
        return hipErrorInvalidValue;
      }
    }
    constexpr int NQ = Grp<LA, LB, LC, LD>::NQ;
    const int64_t need = (n + NQ - 1) / NQ;
    const int64_t g = need < kGroupWorkgroups ? need : kGroupWorkgroups;
    ClassDev cg = cd;
    cg.chunk_stride = chunk_stride(g);
    hipLaunchKernelGGL((eri_group_kernel<LA, LB, LC, LD>), dim3((unsigned)g), dim3(64), 0, st, cg, ba);
  } else {
    using S = Coop<LA, LB, LC, LD>;
    const int64_t g = n < nwg ? n : nwg;
    hipLaunchKernelGGL((eri_coop_kernel<LA, LB, LC, LD>), dim3((unsigned)g), dim3(S::NT), 0, st, cd, ba);
  }
  return hipGetLastError();
}

// Kernel attributes a class's launches need, set outside any stream capture
// (setup): the wave-split kernels take more than 64 KB of dynamic LDS.
template <int LA, int LB, int LC, int LD>
static int prepare_one() {
  if constexpr (SplitShape<LA, LB, LC, LD>::exists) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&eri_split_kernel<LA, LB, LC, LD, false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSplitLdsBytes);
    if (e == hipSuccess && RowShape<LA, LB, LC, LD, true>::exists)
      e = hipFuncSetAttribute(reinterpret_cast<const void *>(&eri_split_kernel<LA, LB, LC, LD, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSplitLdsBytes);
    return e;
  }
  return hipSuccess;
}

template <int LA, int LB, int LC, int LD>
static int64_t scratch_one() {
  if constexpr (use_small<LA, LB, LC, LD>() || Grp<LA, LB, LC, LD>::fits)
    return 0;
  else
    return Coop<LA, LB, LC, LD>::scratch();
}

// class c -> angular momenta (A >= B within a pair: the higher-L shell first)
#define LIBERI_CLASS_LIST(X)                                                                              \
  X(0, 0, 0, 0, 0) X(1, 0, 0, 1, 0) X(2, 0, 0, 1, 1) X(3, 1, 0, 1, 1) X(4, 1, 0, 1, 0) X(5, 1, 1, 1, 1)   \
  X(6, 0, 0, 2, 0) X(7, 0, 0, 2, 1) X(8, 1, 0, 2, 0) X(9, 1, 0, 2, 1) X(10, 0, 0, 2, 2)                   \
  X(11, 1, 0, 2, 2) X(12, 1, 1, 2, 0) X(13, 2, 0, 2, 0) X(14, 2, 2, 1, 1) X(15, 2, 1, 2, 1)               \
  X(16, 2, 1, 2, 0) X(17, 2, 1, 1, 1) X(18, 2, 2, 2, 0) X(19, 2, 2, 2, 1) X(20, 2, 2, 2, 2)               \
  X(21, 3, 3, 3, 3) X(22, 0, 0, 3, 0) X(23, 0, 0, 3, 1) X(24, 0, 0, 3, 2) X(25, 1, 0, 3, 0)               \
  X(26, 3, 3, 3, 0) X(27, 3, 0, 1, 1) X(28, 3, 0, 2, 0) X(29, 2, 1, 3, 0) X(30, 2, 2, 3, 0)               \
  X(31, 3, 0, 3, 0) X(32, 3, 1, 1, 0) X(33, 3, 1, 1, 1) X(34, 3, 1, 2, 0) X(35, 3, 1, 2, 1)               \
  X(36, 2, 2, 3, 1) X(37, 3, 1, 3, 0) X(38, 3, 1, 3, 1) X(39, 3, 2, 1, 0) X(40, 3, 3, 3, 1)               \
  X(41, 3, 2, 1, 1) X(42, 3, 2, 2, 0) X(43, 3, 2, 2, 1) X(44, 3, 2, 2, 2) X(45, 3, 2, 3, 0)               \
  X(46, 3, 2, 3, 1) X(47, 3, 2, 3, 2) X(48, 3, 3, 0, 0) X(49, 3, 3, 1, 0) X(50, 3, 3, 1, 1)               \
  X(51, 3, 3, 2, 0) X(52, 3, 3, 2, 1) X(53, 3, 3, 2, 2) X(54, 3, 3, 3, 2)

// The 55 instantiations are spread over LIBERI_CLASS_GROUPS translation units
// (class id modulo the group count) so they compile in parallel; the unit
// built with LIBERI_CLASS_GROUP=-1 holds the dispatcher and the common
// kernels, the one built with -3 the fused group launch; without the macro
// one unit holds everything.
#define LIBERI_CLASS_LIST_G0(X) X(0, 0, 0, 0, 0) X(8, 1, 0, 2, 0) X(16, 2, 1, 2, 0) X(24, 0, 0, 3, 2) X(32, 3, 1, 1, 0) X(40, 3, 3, 3, 1) X(48, 3, 3, 0, 0)
#define LIBERI_CLASS_LIST_G1(X) X(1, 0, 0, 1, 0) X(9, 1, 0, 2, 1) X(17, 2, 1, 1, 1) X(25, 1, 0, 3, 0) X(33, 3, 1, 1, 1) X(41, 3, 2, 1, 1) X(49, 3, 3, 1, 0)
#define LIBERI_CLASS_LIST_G2(X) X(2, 0, 0, 1, 1) X(10, 0, 0, 2, 2) X(18, 2, 2, 2, 0) X(26, 3, 3, 3, 0) X(34, 3, 1, 2, 0) X(42, 3, 2, 2, 0) X(50, 3, 3, 1, 1)
#define LIBERI_CLASS_LIST_G3(X) X(3, 1, 0, 1, 1) X(11, 1, 0, 2, 2) X(19, 2, 2, 2, 1) X(27, 3, 0, 1, 1) X(35, 3, 1, 2, 1) X(43, 3, 2, 2, 1) X(51, 3, 3, 2, 0)
#define LIBERI_CLASS_LIST_G4(X) X(4, 1, 0, 1, 0) X(12, 1, 1, 2, 0) X(20, 2, 2, 2, 2) X(28, 3, 0, 2, 0) X(36, 2, 2, 3, 1) X(44, 3, 2, 2, 2) X(52, 3, 3, 2, 1)
#define LIBERI_CLASS_LIST_G5(X) X(5, 1, 1, 1, 1) X(13, 2, 0, 2, 0) X(21, 3, 3, 3, 3) X(29, 2, 1, 3, 0) X(37, 3, 1, 3, 0) X(45, 3, 2, 3, 0) X(53, 3, 3, 2, 2)
#define LIBERI_CLASS_LIST_G6(X) X(6, 0, 0, 2, 0) X(14, 2, 2, 1, 1) X(22, 0, 0, 3, 0) X(30, 2, 2, 3, 0) X(38, 3, 1, 3, 1) X(46, 3, 2, 3, 1) X(54, 3, 3, 3, 2)
#define LIBERI_CLASS_LIST_G7(X) X(7, 0, 0, 2, 1) X(15, 2, 1, 2, 1) X(23, 0, 0, 3, 1) X(31, 3, 0, 3, 0) X(39, 3, 2, 1, 0) X(47, 3, 2, 3, 2)

#define DECL_CLASS(id, a, b, cc, d)                                                       \
  int launch_class_##id(const ClassDev &cd, const BuildArgs &ba, hipStream_t st, int nwg); \
  int64_t scratch_class_##id();                                                            \
  int prepare_class_##id();
LIBERI_CLASS_LIST(DECL_CLASS)
#undef DECL_CLASS

#define DEF_CLASS(id, a, b, cc, d)                                                         \
  int launch_class_##id(const ClassDev &cd, const BuildArgs &ba, hipStream_t st, int nwg) { \
    return launch_one<a, b, cc, d>(cd, ba, st, nwg);                                        \
  }                                                                                         \
  int64_t scratch_class_##id() { return scratch_one<a, b, cc, d>(); }                      \
  int prepare_class_##id() { return prepare_one<a, b, cc, d>(); }

#ifndef LIBERI_CLASS_GROUP
#define LIBERI_CLASS_GROUP -2
#endif
#if LIBERI_CLASS_GROUP == 0 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G0(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 1 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G1(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 2 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G2(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 3 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G3(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 4 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G4(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 5 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G5(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 6 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G6(DEF_CLASS)
#endif
#if LIBERI_CLASS_GROUP == 7 || LIBERI_CLASS_GROUP == -2
LIBERI_CLASS_LIST_G7(DEF_CLASS)
#endif
#undef DEF_CLASS

// ------------------------------------------------- fused group launch
// The s, p, d classes of the group path that run three waves per SIMD (all
// but (dd|dd)) in one persistent launch fed from a device work queue: items
// (class, range of quartets) in an array, taken by the workgroups one at a
// time from an atomic counter, so the classes of a build share the grid and
// no class waits for another's tail.  Each item runs the class's group
// kernel body on one wave; the LDS is the largest slot set among the classes.
#define LIBERI_FUSED_LIST(X) \
  X(11, 1, 0, 2, 2) X(14, 2, 2, 1, 1) X(15, 2, 1, 2, 1) X(16, 2, 1, 2, 0) X(17, 2, 1, 1, 1) X(18, 2, 2, 2, 0) X(19, 2, 2, 2, 1)

#if LIBERI_CLASS_GROUP == -3 || LIBERI_CLASS_GROUP == -2
constexpr int fused_lds_doubles() {
  int m = 0;
#define X(id, a, b, cc, d) m = Grp<a, b, cc, d>::NQ * Grp<a, b, cc, d>::SLOT > m ? Grp<a, b, cc, d>::NQ * Grp<a, b, cc, d>::SLOT : m;
  LIBERI_FUSED_LIST(X)
#undef X
  return m;
}
constexpr int fused_nq() {
  int m = 0;
#define X(id, a, b, cc, d) m = Grp<a, b, cc, d>::NQ > m ? Grp<a, b, cc, d>::NQ : m;
  LIBERI_FUSED_LIST(X)
#undef X
  return m;
}
#define X(id, a, b, cc, d)                                                                                    \
  static_assert(!use_small<a, b, cc, d>() && Grp<a, b, cc, d>::fits && a + b + cc + d <= 7 && a <= 2 && b <= 2 && \
                    cc <= 2 && d <= 2,                                                                         \
                "fused classes are group classes at three waves per SIMD");
LIBERI_FUSED_LIST(X)
#undef X

// one item's class body as a call of its own, so that each class keeps its
// own register allocation instead of one allocation over all of them
template <int LA, int LB, int LC, int LD>
__device__ __noinline__ void grp_run_item(const ClassDev &cd, const BuildArgs &ba, int64_t c0, int64_t c1,
                                          double *lds, int *s_npq, int *s_nk, int *s_boff, int *s_koff) {
  grp_run<LA, LB, LC, LD>(cd, ba, c0, c1, lds, s_npq, s_nk, s_boff, s_koff);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void eri_group_fused_kernel(
    const ClassDev *cds, BuildArgs ba, const FusedItem *items, int64_t nitems, unsigned long long *head) {
  __shared__ double lds[fused_lds_doubles()];
  __shared__ int s_npq[fused_nq()], s_nk[fused_nq()], s_boff[fused_nq()], s_koff[fused_nq()];
  for (;;) {
    unsigned long long it = 0;
    if (threadIdx.x == 0) it = atomicAdd(head, 1ull);
    it = __shfl(it, 0, 64);
    if (it >= (unsigned long long)nitems) break;
    const FusedItem w = items[it];
    switch (w.cls) {
#define X(id, a, b, cc, d)                                                                 \
  case id:                                                                                 \
    grp_run_item<a, b, cc, d>(cds[id], ba, w.c0, w.c1, lds, s_npq, s_nk, s_boff, s_koff); \
    break;
      LIBERI_FUSED_LIST(X)
#undef X
      default:
        break;
    }
    wave_sync();  // the next item reuses the slots
  }
}

int launch_group_fused(const ClassDev *cds, const BuildArgs &ba, const FusedItem *items, int64_t nitems,
                       unsigned long long *head, int nwg, void *stream) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(eri_group_fused_kernel, dim3((unsigned)nwg), dim3(64), 0, (hipStream_t)stream, cds, ba, items,
                     nitems, head);
  return hipGetLastError();
}

int class_fused_nq(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return Grp<a, b, cc, d>::NQ;
    LIBERI_FUSED_LIST(X)
#undef X
    default:
      return 0;
  }
}
#endif

#if LIBERI_CLASS_GROUP == -1 || LIBERI_CLASS_GROUP == -2
int launch_class(int c, const ClassDev &cd, const BuildArgs &ba, void *stream, int nwg) {
  hipStream_t st = (hipStream_t)stream;
  if (!(cd.xswitch <= class_rys_xmax(c))) return hipErrorInvalidValue;
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return launch_class_##id(cd, ba, st, nwg);
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

double class_rys_xmax(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return rys_xmax_host[(a + b + cc + d) / 2 + 1];
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return 0.0;
  }
}

int class_nroots(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return (a + b + cc + d) / 2 + 1;
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return 0;
  }
}

int class_kernel_kind(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return use_small<a, b, cc, d>() ? 0 : (Grp<a, b, cc, d>::fits ? 1 : 2);
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return -1;
  }
}

int class_has_coop(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return has_coop_alt<a, b, cc, d>() ? 1 : 0;
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return 0;
  }
}

int64_t class_row_lds(int c, int transposed, const int32_t *ncol, double xswitch) {
  switch (c) {
#define X(id, a, b, cc, d)                                                                      \
  case id:                                                                                      \
    return transposed ? row_lds<a, b, cc, d, true>(ncol, xswitch) : row_lds<a, b, cc, d, false>(ncol, xswitch);
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return -1;
  }
}

int class_has_split(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return SplitShape<a, b, cc, d>::exists && !use_small<a, b, cc, d>() ? 1 : 0;
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return 0;
  }
}

int prepare_class(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return prepare_class_##id();
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

int64_t class_scratch_doubles(int c) {
  switch (c) {
#define X(id, a, b, cc, d) \
  case id:                 \
    return scratch_class_##id();
    LIBERI_CLASS_LIST(X)
#undef X
    default:
      return 0;
  }
}
#endif

}  // namespace liberi
