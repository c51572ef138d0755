// chkdiv over caller-owned device arrays with no allocation: the partial
// sums and maxima of the reduction live in a fixed device array of this
// library instead of memory a context allocates, so the call works where
// hipMalloc cannot be called (inside HIP graph capture, in serving loops,
// under harnesses that forbid allocation).
#include <array>
#include <mutex>

#include "../../include/snac_amd.h"
#include "common.hpp"

namespace snac {

void set_last_error(const std::string &msg);

namespace {

// SNAC_NOALLOC_MAX_PARTS partial sums followed by as many partial maxima
// (one pair per 64x4 tile of a plane, per plane), 4 MiB in all.  Each GPU
// holds its own copy for the life of the process.
constexpr long kParts = SNAC_NOALLOC_MAX_PARTS;
static __device__ __align__(256) double scratch[2 * kParts];

// device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (the array's name is only a host-side stand-in)
std::mutex g_mu;
std::array<double *, 64> g_ptr{};

double *scratch_on(int device) {
  if (device < 0 || device >= (int)g_ptr.size()) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    SNAC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null address for the reduction scratch");
    g_ptr[device] = (double *)p;
  }
  return g_ptr[device];
}

// restores the caller's current device on every exit
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int d) {
    SNAC_HIP(hipGetDevice(&prev));
    if (d != prev) SNAC_HIP(hipSetDevice(d));
  }
  ~DeviceScope() {
    int now = -1;
    if (hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace
}  // namespace snac

extern "C" int snac_chkdiv_noalloc(int device, int nx, int ny, int nz, long sx, long sy, const double *u,
                                   const double *v, const double *w, const double *dxf, const double *dyf,
                                   const double *dzf, double vol_all, double *out, void *stream) {
  using namespace snac;
  try {
    set_last_error("");
    if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("chkdiv_noalloc: block sizes must be >= 1");
    if (sx < nx + 2 || sy < ny + 2)
      throw std::runtime_error("chkdiv_noalloc: pitches must be sx >= nx+2 and sy >= ny+2");
    if (!u || !v || !w || !dxf || !dyf || !dzf || !out) throw std::runtime_error("chkdiv_noalloc: null pointer");
    if (!(vol_all > 0.0)) throw std::runtime_error("chkdiv_noalloc: vol_all must be > 0");
    const Geo g{nx, ny, nz, sx, sy};
    const long np = reduce_nparts(g);
    if (np > kParts)
      throw std::runtime_error("chkdiv_noalloc: the block needs " + std::to_string(np) +
                               " partial sums, more than the " + std::to_string(kParts) +
                               " the fixed scratch holds ((nx/64 rounded up) * (ny/4 rounded up) * nz)");
    DeviceScope ds(device);
    double *ptr = scratch_on(device);
    Metrics m{};
    m.dxf = dxf;
    m.dyf = dyf;
    m.dzf = dzf;
    hipStream_t st = (hipStream_t)stream;
    // every partial the final reduction reads is written first: nothing to clear
    launch_chkdiv_part(st, g, m, u, v, w, vol_all, ptr, ptr + kParts);
    launch_reduce_final(st, ptr, ptr + kParts, (int)np, out);
    return 0;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  } catch (...) {
    set_last_error("unknown error");
    return -1;
  }
}
