// Schwarz integrals without any device allocation: liberi_c_schwarz_noalloc.
//
// liberi_c_schwarz allocates four device buffers (pair records, primitive-pair
// records, the pair list, the results) and a stream per call, and frees them
// again; hipFree synchronises the whole device, which stalls every other
// build in flight, and neither call is allowed under HIP graph capture.  This
// variant stages the same data in one library-owned device array, `scratch`,
// and runs on the caller's stream.
//
// The array lives as long as the process, once per GPU the library runs on.
// Host code never passes its name to the runtime or to a kernel: its device
// address on each GPU comes from hipGetSymbolAddress, called once per device
// with that device current, and only that address is used.  Calls that
// overlap on one GPU share the array and race; callers serialise them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/liberi_hip.h"
#include "liberi_internal.hpp"

using namespace liberi;

static __device__ __align__(256) unsigned char scratch[LIBERI_SCHWARZ_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device (nullptr: not looked up yet)

// address of this device's copy of `scratch` (looked up on first use)
int scratch_address(const char *who, void **ptr) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_last_error(LIBERI_EDEVICE, std::string(who) + ": no usable current device");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_last_error(LIBERI_EDEVICE, std::string(who) + ": hipGetSymbolAddress(scratch): " +
                                                hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return LIBERI_OK;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" int liberi_c_schwarz_noalloc(int32_t nsh, int32_t natoms, int32_t num_bas, int32_t mxgtot,
                                        const int32_t *ang_mom, const int32_t *contr_num, const int32_t *sh_loc,
                                        const int32_t *atom_num, const int32_t *atom_loc, const double *exponents,
                                        const double *cs, const double *cp, const double *cd, const double *cf,
                                        const double *coords, double *schwrz_int, void *stream) {
  const char *who = "liberi_c_schwarz_noalloc";
  if (!schwrz_int) return set_last_error(LIBERI_EARG, std::string(who) + ": NULL output");
  SchwarzStage s;
  if (int rc = schwarz_stage(who, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc,
                             exponents, cs, cp, cd, cf, coords, s))
    return rc;
  // layout of the staging area: pair records | primitive records | list | results
  const size_t o_pairs = 0;
  const size_t o_prims = o_pairs + align256(s.pairs.size() * sizeof(PairRec));
  const size_t o_list = o_prims + align256(s.prims.size() * sizeof(PrimRec));
  const size_t o_out = o_list + align256(s.list.size() * sizeof(int32_t));
  const size_t need = o_out + s.list.size() * sizeof(double);
  if (need > (size_t)LIBERI_SCHWARZ_SCRATCH_BYTES)
    return set_last_error(LIBERI_ESIZE, std::string(who) + ": the basis needs " + std::to_string(need) +
                                            " bytes of staging (" + std::to_string(s.pairs.size()) +
                                            " shell pairs, " + std::to_string(s.prims.size()) +
                                            " primitive pairs), above the fixed " +
                                            std::to_string((long long)LIBERI_SCHWARZ_SCRATCH_BYTES) +
                                            " bytes; use liberi_c_schwarz");
  if (int rc = schwarz_check_device(who)) return rc;
  void *base = nullptr;
  if (int rc = scratch_address(who, &base)) return rc;
  unsigned char *b = (unsigned char *)base;
  hipStream_t st = (hipStream_t)stream;
  std::vector<double> out(s.list.size());
  hipError_t e = (hipError_t)schwarz_enqueue(s, (PairRec *)(b + o_pairs), (PrimRec *)(b + o_prims),
                                             (int32_t *)(b + o_list), (double *)(b + o_out), st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data(), b + o_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return set_last_error(LIBERI_EDEVICE, std::string(who) + ": " + hipGetErrorString(e));
  schwarz_scatter(s, out.data(), schwrz_int);
  return LIBERI_OK;
}
