/*
 * CPU oracle for the Schwarz-screened RHF Fock build (test infrastructure).
 *
 * This library is the checker the HIP product path is compared against.  It
 * restates, in plain C, the reference's algorithm for `compute_integrals`
 * (src/liberi_driver.F90:52-167 of gms_libERI):
 *   - shell classification and pair lists (src/liberi_shell_pair.F90:108-346),
 *   - the per-class pair pre-screen and quartet Schwarz test of every one of
 *     the 55 dispatched kernels (e.g. src/rhf/rot_axis/int0000.F90:57-107,
 *     src/rhf/rot_axis/int0001.F90:65-121, src/rhf/rys/int3300_rysgen.F90),
 *   - the value cutoff |v| < 5e-11 and the DIRFCK digestion into a packed
 *     lower-triangular Fock matrix (src/rhf/rot_axis/int0000.F90:241-311).
 * The contracted ERIs themselves come from an independent textbook engine
 * (McMurchie-Davidson Hermite expansion with a series/recursion Boys
 * function), so that agreement with the HIP path (Rys quadrature with
 * tabulated roots) checks two unrelated evaluations against each other.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg load this
 * library.  The product never does.
 */
#ifndef LIBERI_ORACLE_H
#define LIBERI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Basis exactly as liberi_setup receives it (src/interface/liberi_interface.F90:72-129):
 * 1-based sh_loc / atom_num / atom_loc, ang_mom 1..4 = s..f, coefficient
 * arrays of length mxgtot*nsh, coords of length 3*natoms (bohr), schwrz of
 * length nsh*(nsh+1)/2 (may be NULL for oracle_schwarz). */
typedef struct {
  int32_t nsh, natoms, num_bas, mxgtot;
  const int32_t *ang_mom, *contr_num, *sh_loc, *atom_num, *atom_loc;
  const double *exponents, *coef_s, *coef_p, *coef_d, *coef_f;
  const double *coords, *schwrz;
} oracle_basis;

#define ORACLE_NCLASS 55

int oracle_nclass(void);
/* kernel name of class c in dispatch order of src/liberi_driver.F90:69-158 */
const char *oracle_class_name(int c);
/* class c's bra and ket pair classes (0..9, shell_pair_container_t order),
 * quartet test (0: keep if >, 1: skip if <, 2: none) and pre-screen
 * normalisation (0: the other list's maximum, 1: its own); -1 if c is out of range */
int oracle_class_def(int c, int def[4]);

/* Per-class surviving contracted shell-quartet counts (pure screening). */
int oracle_screen_counts(const oracle_basis *b, double cutoff, int64_t counts[ORACLE_NCLASS]);

/* Full Fock build: F (packed, n(n+1)/2) is overwritten.  counts may be NULL.
 * sample_stride > 1 evaluates only every sample_stride-th surviving quartet
 * of each class (used to time a bounded sample); 1 = the full build.
 * nthreads <= 0 uses the OpenMP default.  Returns number of quartets
 * evaluated, or -1 on error. */
int64_t oracle_fock(const oracle_basis *b, const double *D, double *F, double cutoff,
                    int64_t counts[ORACLE_NCLASS], int64_t sample_stride, int nthreads);

/* Same, restricted to the classes with enable[c] != 0 (enable may be NULL). */
int64_t oracle_fock_classes(const oracle_basis *b, const double *D, double *F, double cutoff,
                            int64_t counts[ORACLE_NCLASS], int64_t sample_stride, int nthreads,
                            const int32_t *enable);

/* Same; ref_switch != 0 takes the reference's large-argument Boys values
 * F_m = (2m-1)!!/(2T)^m sqrt(pi/4T) past each class's switch point (T > 25
 * for its rotated-axis and ERIC kernels, X >= 37..65 by root count for its
 * Rys kernels) instead of the exact ones. */
int64_t oracle_fock_classes_sw(const oracle_basis *b, const double *D, double *F, double cutoff,
                               int64_t counts[ORACLE_NCLASS], int64_t sample_stride, int nthreads,
                               const int32_t *enable, int ref_switch);

/* Schwarz integrals sqrt(max_{a in i, b in j} |(ab|ab)|), packed by shell pair
 * (index i(i-1)/2 + j, 1-based, i >= j), with the kernels' component scaling. */
int oracle_schwarz(const oracle_basis *b, double *schwrz_out);

/* One normalized cartesian block (ab|cd) of shells i,j,k,l (0-based), in
 * GAMESS component order, out[((a*nb + b)*nc + c)*nd + d]. */
int oracle_shell_quartet(const oracle_basis *b, int i, int j, int k, int l, double *out);

/* Boys function F_m(T), used by the engine; exported for tests. */
double oracle_boys(int m, double T);

#ifdef __cplusplus
}
#endif
#endif
