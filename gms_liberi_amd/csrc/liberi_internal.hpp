// Internal data model of the MI355X Fock build (host + device).
//
// The reference keeps, per pair class, arrays of primitive-pair data behind
// CSR offsets (shell_pair_t, src/liberi_types.f90:80-92) and rebuilds the
// screened pair lists on the host for every kernel call
// (e.g. src/rhf/rot_axis/int0001.F90:65-84).  Here every shell pair of every
// class lives once in HBM as a fixed-size record, every primitive pair as a
// 64-byte record, and each of the 55 classes owns its screened, sorted pair
// lists plus a segment table that enumerates the surviving quartets without
// materialising them.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace liberi {

constexpr int kNumClasses = 55;
constexpr int kNumPairClasses = 10;
constexpr double kRefCutoff = 5.00000000000000018e-11;  // src/liberi_parameters.F90:38
constexpr double kTwoPi52 = 34.98683665524971863;        // 2 pi^(5/2), :20

// pair classes in the order of shell_pair_container_t (src/liberi_types.f90:94-109)
enum PairClass { P_SS, P_SP, P_PP, P_SD, P_PD, P_DD, P_SF, P_PF, P_DF, P_FF };
constexpr int kPairL[kNumPairClasses][2] = {{0, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2},
                                            {2, 2}, {0, 3}, {1, 3}, {2, 3}, {3, 3}};

enum QuartetTest { CMP_GT, CMP_GE, CMP_NONE };
enum PreScreen { PS_CROSS, PS_SELF };

struct ClassDef {
  const char *name;
  int bra, ket;  // pair classes; bra == ket marks a diagonal class (kl <= ij)
  int cmp, ps;
};

// Dispatch order of src/liberi_driver.F90:69-158.  Quartet test and pair
// pre-screen per kernel file: the hand-written rotated-axis kernels normalise
// each list's pre-screen by the OTHER list's largest Schwarz value
// (int0001.F90:67,76), the generated Rys/ERIC kernels by their own
// (int3300_rysgen.F90, int0030_ericgen.F90); int0000/int0101/int1111 skip when
// x*x < cut (int0000.F90:107), int2222gen and int3333 have no quartet test
// (int2222_rysgen.F90:428-431), every other kernel keeps x*x > cut.
extern const ClassDef kClasses[kNumClasses];

// One shell pair, A = the shell of higher angular momentum (A first for equal L).
struct PairRec {
  int32_t prim_off;  // first primitive pair record
  int32_t nprim;
  int32_t bfA, bfB;  // first basis function (0-based) of A and B
  double ABx, ABy, ABz;
  int32_t sameAB;    // A and B are one shell
  int32_t shA;       // shell index of A (0-based)
  int32_t cA, cB;    // compact column of A's / B's first function among all
                     // functions of that angular momentum (shell rank * ncart)
  double x;          // Schwarz value of the pair
};

// One primitive pair (a in A, b in B): p = a + b, P, P - A, and
// K = c_a c_b exp(-ab/p |A-B|^2) / p.
// largest basis: n(n+1) stays below 2^32, so packed indices fit 32 bits
constexpr int32_t kMaxBasis = 65535;

struct alignas(64) PrimRec {
  double p, Px, Py, Pz, PAx, PAy, PAz, K;
};

// Device view of one class's quartet enumeration.  Kept pair lists are sorted
// by (primitive count, Schwarz value descending); a segment is one kept bra
// row against one primitive-count group of kets, and its surviving kets are a
// prefix of that group (the quartet test is monotone in the ket's Schwarz
// value).  Quartet q of the class lies in segment s with
// seg_start[s] <= q < seg_start[s+1], ket = ket_list[seg_ket0[s] + q - seg_start[s]].
struct ClassDev {
  const int32_t *bra_list;  // pair indices (into the PairRec array)
  const int32_t *ket_list;
  const int64_t *seg_start;  // nseg + 1
  const int32_t *seg_bra;    // bra pair index of the segment
  const int32_t *seg_ket0;   // position in ket_list of the segment's first ket
  int32_t nseg;
  int32_t diag;
  int64_t q_begin, q_end;    // this rank's share of the class's quartets
  double xswitch;            // Rys argument from which the asymptotic rule is used
  // Row tiles: runs of at most the class's tile length (kTileLens) of one row.  Tile t
  // covers [tile_q[t], tile_q[t+1]) and starts in segment tile_seg[t]; this
  // rank's tiles are [t_begin, t_end).  row_mode selects the row-tiled kernel.
  const int64_t *tile_q;     // ntile + 1
  const int32_t *tile_seg;   // ntile
  int32_t t_begin, t_end;
  int32_t row_mode;
  // 0: rows run over the class's bra pairs, segments over its kets (as
  // above); 1: the roles are swapped -- seg_bra holds ket pairs and ket_list
  // bra pairs -- so that a row tile fixes the ket pair (non-diagonal classes)
  int32_t transposed;
  // group kernels: chunk_stride(grid) of the launch (set at launch)
  int64_t chunk_stride;
  // 1: the mixed form may evaluate this class's quartets in FP32 (chosen per
  // class at the first build under a given switch tau)
  int32_t fp32;
  // 1: a group class with a cooperative form (class_has_coop) runs that form
  int32_t coop;
};

struct BuildArgs {
  const PairRec *pairs;
  const PrimRec *prims;
  const double *D;   // density of the Coulomb blocks
  const double *DK;  // density of the exchange blocks (= D for RHF)
  double k_scale;    // exchange weight (1 for RHF: F = J[D] - K[D]/2; 2 for a UHF spin: J[Da+Db] - K[Ds])
  double *F;
  double value_cut;
  double *scratch;       // per-workgroup scratch for the cooperative kernels
  int64_t scratch_per_wg;  // doubles
  int32_t nbf;           // basis functions
  // compact columns of the row-tiled accumulators: the functions of angular
  // momentum L in shell order, col_bf[L][k] = basis function of column k
  const int32_t *col_bf[4];
  int32_t ncol[4];
  // mixed precision: quartets with x_bra * x_ket * dmax[0] < fp32_tau in FP32
  double fp32_tau;
  const double *dmax;        // max |D| of this build (device)
  uint64_t *fp32_count;      // quartets evaluated in FP32 (device counter)
};

// Launch the kernel of class c over [q_begin, q_end).  Returns hipError_t.
int launch_class(int c, const ClassDev &cd, const BuildArgs &ba, void *stream, int num_wg_coop);
// 1 when group class c has the wave-split row-tiled form (its row form)
int class_has_split(int c);
// kernel attributes the launches of class c need (the wave-split kernels'
// dynamic LDS above 64 KB), on the current device; called at setup, outside
// any stream capture.  Returns hipError_t.
int prepare_class(int c);
// doubles of scratch the cooperative kernel of class c needs per workgroup (0 if none)
int64_t class_scratch_doubles(int c);
// Rys roots of class c and the end of its tabulated range
int class_nroots(int c);
double class_rys_xmax(int c);
// LDS bytes of the row-tiled kernel of class c with rows over the bra pairs
// (transposed = 0) or the ket pairs (1), given the compact column counts per
// angular momentum and the Rys switch (the kernel stages the table below it);
// -1 when the class has no such kernel
int64_t class_row_lds(int c, int transposed, const int32_t *ncol, double xswitch);
// kernel shape of class c: 0 lane per quartet, 1 workgroup group per
// quartet, 2 cooperative workgroup per quartet
int class_kernel_kind(int c);
// 1 when group class c also has the cooperative workgroup-per-quartet form
// with its buffers in LDS (classes with an f shell), selected by ClassDev::coop
int class_has_coop(int c);

// Fused group launch: one work item = quartets [c0, c1) of class cls (relative
// to the class's share, c0 a multiple of its quartets per round)
struct FusedItem {
  int32_t cls, pad;
  int64_t c0, c1;
};
// quartets per round of class c in the fused group launch (0: not fused)
int class_fused_nq(int c);
// the items[0 .. nitems) of the build, taken through the counter *head
// (zero before the launch); cds: ClassDev of every class by class index
int launch_group_fused(const ClassDev *cds, const BuildArgs &ba, const FusedItem *items, int64_t nitems,
                       unsigned long long *head, int nwg, void *stream);
// longest row tile, quartets
// row-tile lengths (quartets of one row per workgroup) the autotuner times
// per row-tiled class; kDefaultTile where it does not time
constexpr int kNumTileLens = 4;
constexpr int64_t kTileLens[kNumTileLens] = {512, 1024, 4096, 16384};
constexpr int kDefaultTile = 2;
// LDS the row-tiled kernel may take per workgroup
constexpr int64_t kRowLdsBytes = 64 * 1024;
// workgroups the cooperative kernels use
constexpr int kCoopWorkgroups = 2048;
// one-wave workgroups the group-per-quartet kernels use at most (persistent loop)
constexpr int kGroupWorkgroups = 65536;

// Order in which the workgroups of a persistent group-kernel launch take the
// contiguous chunks of a class's quartet range: workgroup b takes chunk
// b * s mod grid.  The stride s is coprime with the grid, so the map is a
// permutation of 0 .. grid-1 for every grid size, and close to grid divided by
// the golden ratio, so that workgroups resident at the same time (consecutive
// b) work on distant parts of the range.
constexpr int64_t chunk_stride(int64_t grid) {
  if (grid <= 2) return 1;
  int64_t s = (int64_t)((double)grid * 0.6180339887498949) | 1;
  for (;; ++s) {  // ends at the latest at s = grid - 1, coprime with grid
    int64_t a = grid, b = s;
    while (b != 0) {
      const int64_t t = a % b;
      a = b;
      b = t;
    }
    if (a == 1) return s;
  }
}

// Schwarz values sqrt(max |(ab|ab)|) of the n shell pairs list[0..n) whose
// shells have angular momenta la >= lb (A the la shell); out[i] for list[i].
int launch_schwarz(int la, int lb, const PairRec *pairs, const PrimRec *prims, const int32_t *list, int32_t n,
                   double *out, void *stream);

// Schwarz integrals of a whole basis, host side (api.cpp): every shell pair
// i >= j with its primitive pairs (A = the shell of higher angular momentum),
// the pairs grouped by (la, lb) in `list` (group la*4+lb is
// list[off[k] .. off[k+1])), and per pair its packed index i(i+1)/2 + j.
struct SchwarzStage {
  std::vector<PairRec> pairs;
  std::vector<PrimRec> prims;
  std::vector<int64_t> where;
  std::vector<int32_t> list;
  size_t off[17];
};
// checked basis -> stage (LIBERI_OK or a status with the message set)
int schwarz_stage(const char *who, int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                  const int32_t *ang_mom, const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                  const int32_t *atom_loc, const double *exponents, const double *cs, const double *cp,
                  const double *cd, const double *cf, const double *coords, SchwarzStage &s);
// a gfx950 device is current (LIBERI_OK or LIBERI_EDEVICE with the message set)
int schwarz_check_device(const char *who);
// records and list uploaded into the given device buffers, one kernel per
// (la, lb) group writing out[list position]; returns a hipError_t
int schwarz_enqueue(const SchwarzStage &s, PairRec *dp, PrimRec *dq, int32_t *dl, double *dout, void *stream);
// out (list order) -> schwrz_int (packed by shell pair)
void schwarz_scatter(const SchwarzStage &s, const double *out, double *schwrz_int);
// records the message of liberi_c_last_error and returns code
int set_last_error(int code, const std::string &msg);

// p[0..n) = 0 on the stream (a kernel, so it may sit in a captured graph)
int launch_fill_zero(double *p, int64_t n, void *stream);
// out = a + b over n elements
int launch_add(const double *a, const double *b, double *out, int64_t n, void *stream);
// out[0] = max |p[0..n)| (out must be zero on entry)
int launch_abs_max(const double *p, int64_t n, double *out, void *stream);

// Segment lengths on the device: one thread per (kept bra row, ket group).
int launch_segment_lengths(const double *bra_x, const int32_t *bra_group, const int32_t *bra_pos_in_group,
                           int32_t nbra, const double *ket_x, const int32_t *group_begin,
                           const int32_t *group_end, int32_t ngroup, int32_t diag, int32_t cmp,
                           double cut, int64_t *seg_len, void *stream);

}  // namespace liberi
