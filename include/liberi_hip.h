/*
 * liberi_hip -- MI355X-native Schwarz-screened RHF Fock build.
 *
 * C ABI of the drop-in replacement for gms_libERI's handle API
 * (src/interface/liberi_interface.F90) and its GAMESS direct-link wrapper
 * (src/interface/gms_liberi_wrapper.F90).  Plain pointers and sizes only.
 * Every entry point returns LIBERI_OK (0) or a LIBERI_E* status and records a
 * message readable with liberi_c_last_error() (thread-local).
 *
 * Array conventions are the reference's: 1-based sh_loc / atom_num /
 * atom_loc, ang_mom 1..4 = s..f, coordinates in bohr, density and Fock packed
 * lower-triangular with idx(i,j) = i(i-1)/2 + j (1-based, i >= j), density
 * not doubled off the diagonal, Fock overwritten with S o (J[D] - K[D]/2)
 * (1 on the diagonal of S, 2 off it).
 */
#ifndef LIBERI_HIP_H
#define LIBERI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIBERI_OK 0
#define LIBERI_EARG 1    /* bad argument (NULL, negative size, inconsistent basis) */
#define LIBERI_ESTATE 2  /* call out of order, e.g. fock build before setup */
#define LIBERI_ESIZE 3   /* matrix size n != num_bas(num_bas+1)/2, or a basis above a fixed staging size */
#define LIBERI_EDEVICE 4 /* HIP runtime failure or no usable gfx950 device */
#define LIBERI_ENOMEM 5  /* device or host allocation failed */

#define LIBERI_NCLASS 55 /* integral classes, dispatch order of src/liberi_driver.F90:69-158 */

typedef struct liberi_ctx liberi_ctx;

/* Replaces liberi_create (src/interface/liberi_interface.F90:46-55). */
int liberi_c_create(liberi_ctx **ctx);

/* Replaces liberi_destroy (src/interface/liberi_interface.F90:57-68): cleans
 * up if set up, frees the handle, sets *ctx = NULL. */
int liberi_c_destroy(liberi_ctx **ctx);

/* Replaces liberi_setup (src/interface/liberi_interface.F90:72-200): deep-copies
 * the basis, builds the shell-pair structure-of-arrays and the screened
 * quartet lists on the device.  Calling it on a set-up handle cleans up first
 * (:133-136).  Arrays: ang_mom..end_bas have nsh entries; exponents and
 * contr_coef_* mxgtot*nsh; coords 3*natoms; schwrz_int nsh(nsh+1)/2 (packed by
 * shell pair like the density).  my_rank/num_procs select this rank's share
 * of every class, as the reference's rank split does. */
int liberi_c_setup(liberi_ctx *ctx, int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                   const int32_t *ang_mom, const int32_t *contr_num, const int32_t *sh_loc,
                   const int32_t *atom_num, const int32_t *atom_loc, const int32_t *start_bas,
                   const int32_t *end_bas, const double *exponents, const double *contr_coef_s,
                   const double *contr_coef_p, const double *contr_coef_d, const double *contr_coef_f,
                   const double *coords, const double *schwrz_int, int32_t my_rank, int32_t num_procs);

/* Schwarz integrals on the device: schwrz_int[i(i+1)/2 + j] (0-based shells
 * i >= j) = sqrt(max over components |(ab|ab)|), the normalised-component
 * convention of the schwrz_int argument of liberi_setup
 * (src/interface/liberi_interface.F90:128), which the reference takes from its
 * caller (GAMESS) and never computes.  Same basis arguments as
 * liberi_c_setup; needs no handle.  Blocks until done. */
int liberi_c_schwarz(int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot, const int32_t *ang_mom,
                     const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                     const int32_t *atom_loc, const double *exponents, const double *contr_coef_s,
                     const double *contr_coef_p, const double *contr_coef_d, const double *contr_coef_f,
                     const double *coords, double *schwrz_int);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, serving loops that forbid hipMalloc
 * and hipFree (hipFree synchronises the whole device, stalling the builds of
 * other requests in flight): liberi_c_schwarz with the same arguments and
 * results, which allocates no device memory and creates no stream.  Its
 * records, pair list and results are staged in one library-owned device array
 * of LIBERI_SCHWARZ_SCRATCH_BYTES per GPU, held until the process ends; a
 * basis whose staging needs more (about 64 bytes per primitive pair plus 80
 * per shell pair: some 900k primitive pairs) fails with LIBERI_ESIZE before
 * any device work.  Runs on `stream` (a hipStream_t; NULL = the default
 * stream) and returns when schwrz_int is written.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU race, callers serialise
 * them. */
#define LIBERI_SCHWARZ_SCRATCH_BYTES (64LL << 20)
int liberi_c_schwarz_noalloc(int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot, const int32_t *ang_mom,
                             const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                             const int32_t *atom_loc, const double *exponents, const double *contr_coef_s,
                             const double *contr_coef_p, const double *contr_coef_d, const double *contr_coef_f,
                             const double *coords, double *schwrz_int, void *stream);

/* Replaces liberi_fock_build (src/interface/liberi_interface.F90:202-227) and
 * through it compute_integrals (src/liberi_driver.F90:52-167).  Host arrays of
 * n = num_bas(num_bas+1)/2 doubles; fock is overwritten.  Blocks until done. */
int liberi_c_fock_build(liberi_ctx *ctx, const double *density, double *fock, int64_t n);

/* Same build on device-resident arrays (hipMalloc'd, n doubles each) on the
 * given hipStream_t (NULL = the handle's stream).  Asynchronous: returns after
 * enqueueing.  No counterpart in the reference (which maps D/F per call,
 * src/liberi_driver.F90:66,163); for callers that keep D and F in HBM. */
int liberi_c_fock_build_device(liberi_ctx *ctx, const double *d_density, double *d_fock, int64_t n,
                               void *stream);

/* Unrestricted (UHF) Fock matrices, the FOCK_B quantity the reference's MDI
 * plugin registers but never computes (src/interface/liberi_mdi_api.F90:222-224,
 * 404-409): Fa = S o (J[Da+Db] - K[Da]), Fb = S o (J[Da+Db] - K[Db]), with
 * Da, Db the packed spin densities (same packing as the RHF density; Da = Db =
 * D/2 gives Fa = Fb = the RHF F of D).  Two passes of the build (the integrals
 * are evaluated twice).  Host arrays, blocking; the _device form takes HBM
 * arrays and a hipStream_t and is asynchronous. */
int liberi_c_fock_build_uhf(liberi_ctx *ctx, const double *density_a, const double *density_b, double *fock_a,
                            double *fock_b, int64_t n);
int liberi_c_fock_build_uhf_device(liberi_ctx *ctx, const double *d_density_a, const double *d_density_b,
                                   double *d_fock_a, double *d_fock_b, int64_t n, void *stream);

/* Replaces liberi_cleanup (src/interface/liberi_interface.F90:229-260):
 * releases the basis and all device data; the handle can be set up again. */
int liberi_c_cleanup(liberi_ctx *ctx);

/* Message of the last failing call on this thread ("" if none). */
const char *liberi_c_last_error(void);

/* Runtime knobs (compile-time parameters in the reference,
 * src/liberi_parameters.F90:38).  Take effect at the next setup unless noted. */
#define LIBERI_OPT_SCHWARZ_CUTOFF 1 /* quartet screening cutoff, default 5.00000000000000018e-11 */
#define LIBERI_OPT_VALUE_CUTOFF 2   /* |integral| skip threshold, default 5.00000000000000018e-11 */
#define LIBERI_OPT_FP32_TAU 3       /* mixed precision: quartets of the lane-per-quartet classes whose
                                       Schwarz bound x_ab x_cd times max|D| is below tau are evaluated
                                       in FP32 (Rys roots/weights and all Fock accumulation stay FP64),
                                       in those classes where the first build under this tau measures
                                       FP32 faster (liberi_stats.class_fp32); 0 = plain FP64 (default).
                                       Read at each build. */
#define LIBERI_OPT_EXACT_LARGE_T 4  /* 0 (default): large-argument forms from the reference's switch
                                       points (T > 25 for its Boys-based classes, X >= 37..65 by root
                                       count for its Rys classes); 1: only where exact to double.
                                       Read at setup. */
#define LIBERI_OPT_CLASS_TIMING 5   /* 1: time every class launch with HIP events (stats.class_ms) */
#define LIBERI_OPT_ONLY_CLASS 6     /* >= 0: build only this class (diagnostics); -1 (default): all */
#define LIBERI_OPT_PRIM_CUTOFF 7    /* drop primitive pairs whose share of any integral they can
                                       enter (with the pairs that pass the quartet test with
                                       theirs) is bounded below this (default 1e-18; 0 keeps every
                                       primitive pair) */
#define LIBERI_OPT_ROW_TILES 8      /* digestion: row tiles accumulate the exchange blocks on chip
                                       per tile of quartets sharing one shell pair; direct digestion
                                       adds every element of every quartet to F atomically.  1
                                       (default): per class, see LIBERI_OPT_AUTOTUNE; 0: direct
                                       everywhere; 2: row tiles wherever they fit.  Read at each
                                       build. */
#define LIBERI_OPT_AUTOTUNE 9       /* 1 (default): the first build after setup times direct and
                                       row-tiled digestion per class and keeps the faster; 0: fixed
                                       choice (row tiles for the lane-per-quartet classes and the
                                       wave-split row tiles for the d classes that have them).
                                       Setting it makes the next build decide again. */
#define LIBERI_OPT_GRAPH 10         /* 1 (default): batched form -- the build's class launches run on
                                       four side streams (largest first onto the least-loaded
                                       stream), captured once into a HIP graph that later builds on
                                       the same D and F replay; 0: one stream, launched eagerly.
                                       Per-class timing and ONLY_CLASS builds run eagerly.  Read at
                                       each build; any option change drops the captured graph. */
#define LIBERI_OPT_RANK_SPLIT 12     /* rank shares of every class (my_rank/num_procs of setup): 0 (default)
                                        contiguous shares of equal estimated cost (primitive quartets plus
                                        a per-quartet term); 1 equal quartet counts, the reference's split
                                        (e.g. src/rhf/rot_axis/int0001.F90:100-104).  Either way the ranks'
                                        partial Fock matrices sum to the full one.  Read at setup. */
#define LIBERI_OPT_FUSED 13          /* 0 (default): one launch per class; 1: the seven s, p, d classes
                                        of the group path with the smallest LDS slots run in one
                                        persistent launch fed from a device work queue (slower on the
                                        workloads measured).  Per-class timing and ONLY_CLASS builds
                                        launch per class.  Read at each build. */
#define LIBERI_OPT_TILE_LEN 14       /* row-tile length of the row-tiled classes, in quartets of one row
                                        per workgroup: 0 (default) timed per class at the first build
                                        (LIBERI_OPT_AUTOTUNE; 4096 where untimed), or 512, 1024, 4096, 16384
                                        for every class.  Read at each build. */
#define LIBERI_OPT_ROW_ORIENT 15     /* orientation of the row tiles of a non-diagonal class: 0 (default)
                                        rows over the pair whose Coulomb block is larger, among the
                                        orientations whose accumulators fit; 1 rows over the bra pairs,
                                        2 over the ket pairs, where that orientation fits (else as 0).
                                        Read at setup. */
#define LIBERI_OPT_MAX_DEVICE_BYTES 11 /* HBM budget of the handle in bytes (0, the default: none).  An
                                          allocation that would take the handle's device data past it
                                          fails with LIBERI_ENOMEM; a setup that fails that way (or any
                                          other way) leaves the handle released, ready for another
                                          setup.  For callers that admit work by HBM footprint. */
int liberi_c_set_option(liberi_ctx *ctx, int32_t option, double value);
int liberi_c_get_option(liberi_ctx *ctx, int32_t option, double *value);

/* Statistics of the handle (no counterpart in the reference, which computes
 * the surviving quartet counts implicitly and reports none). */
typedef struct {
  int64_t quartets[LIBERI_NCLASS];      /* surviving contracted shell quartets per class (this rank) */
  int64_t prim_quartets[LIBERI_NCLASS]; /* primitive quartets of those */
  int64_t total_quartets;
  int64_t fp32_quartets;                /* evaluated in FP32 in the last build */
  double last_build_ms;                 /* device time of the last build */
  int64_t device_bytes;                 /* HBM held by the handle */
  double class_ms[LIBERI_NCLASS];       /* per-class device time of the last build, when
                                           LIBERI_OPT_CLASS_TIMING is on (else 0) */
  int64_t prim_pairs;                   /* primitive pairs kept on the device */
  int64_t prim_pairs_dropped;           /* primitive pairs removed by LIBERI_OPT_PRIM_CUTOFF */
  int32_t class_variant[LIBERI_NCLASS]; /* kernel form used per class in the last build: 0 its own
                                           launch with direct atomics, 1 row tiles over bra pairs,
                                           2 row tiles over ket pairs, 3 the cooperative workgroup-per-
                                           quartet form (group classes with an f shell, timed against
                                           the group form at the first build; the only form of
                                           (ff|dd), (ff|fd), (ff|ff)), 4 inside the fused group
                                           launch (LIBERI_OPT_FUSED); -1 class inactive or not yet built */
  int32_t class_fp32[LIBERI_NCLASS];    /* 1: the mixed form (LIBERI_OPT_FP32_TAU > 0) evaluates this
                                           class's qualifying quartets in FP32, chosen per class from
                                           measured time at the first build under that tau; 0: FP64 */
  int32_t class_tile[LIBERI_NCLASS];    /* row-tile length (quartets of one row per workgroup: 512,
                                           1024, 4096 or 16384) of a row-tiled class, timed per class at the
                                           first build after setup (LIBERI_OPT_AUTOTUNE) or 4096;
                                           0 where the class is not row-tiled */
} liberi_stats;
int liberi_c_get_stats(liberi_ctx *ctx, liberi_stats *out);

/* Name of class c (reference kernel name, e.g. "int0000"). */
const char *liberi_c_class_name(int32_t c);

/* GAMESS direct-link entry points with one process-wide handle: replace
 * liberi_init / liberi_fock / liberi_finalize of
 * src/interface/gms_liberi_wrapper.F90:29-146.  Note the argument order of
 * liberi_c_init (natoms, num_bas, nsh, mxgtot, ...) as in :29.  liberi_c_fock
 * returns LIBERI_ESTATE before init and LIBERI_ESIZE for a wrong n (:117-127). */
int liberi_c_init(int32_t natoms, int32_t num_bas, int32_t nsh, int32_t mxgtot, const int32_t *ang_mom,
                  const int32_t *contr_num, const int32_t *sh_loc, const int32_t *atom_num,
                  const int32_t *atom_loc, const int32_t *start_bas, const int32_t *end_bas,
                  const double *exponents, const double *contr_coef_s, const double *contr_coef_p,
                  const double *contr_coef_d, const double *contr_coef_f, const double *coords,
                  const double *schwrz_int, int32_t my_rank, int32_t num_procs);
int liberi_c_fock(const double *density, double *fock, int64_t n);
int liberi_c_finalize(void);

#ifdef __cplusplus
}
#endif
#endif
