// noalloc.cpp -- qcuHopNoAlloc: qcuHop for callers that cannot allocate device memory
// (code under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// qcuHop stages the caller's interface spinor in the internal layout of the requested
// precision, runs the hop into a second internal field and converts that back; its two
// staging fields are library allocations that grow on demand.  Here the two fields
// live in one file-scope device array, `scratch`, of kNoAllocScratchBytes, which every
// device the library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol): the
// library's init call looks up the device address once per device with
// hipGetSymbolAddress, and the launches take that address.
//
// The kernels of qcuHop run unchanged, so the result is the same bit for bit.  Nothing
// in the array has to start at zero: the conversion writes every site of the staged
// input and the hop every site of its output before either is read.  Calls that
// overlap on one device race on the array.
#include <cstring>

#include "qcu_host.h"

namespace qcu {

constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;  // two staging fields of 16 MiB
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

static void* g_scratch_dev[kNoAllocMaxDevices];  // device address of `scratch`, per device

// with `device` current (called once per device from the library's init)
void noalloc_setup(int device) {
  QREQ(device >= 0 && device < kNoAllocMaxDevices, "device index out of range for the no-alloc scratch table");
  void* p = nullptr;
  QHIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  QREQ(p != nullptr, "hipGetSymbolAddress returned no address for the no-alloc scratch");
  g_scratch_dev[device] = p;
}

static void* scratch_ptr() {
  const int d = ctx().device;
  QREQ(d >= 0 && d < kNoAllocMaxDevices, "device index out of range for the no-alloc scratch table");
  void* p = g_scratch_dev[d];
  if (!p) throw Error(QCU_ERROR_STATE, "qcuHopNoAlloc: scratch address not set up (library not initialised)");
  return p;
}

}  // namespace qcu

using namespace qcu;

extern "C" int qcuHopNoAlloc(qcuGauge h, void* out, const void* in, int parity, int dagger, int t_boundary,
                             int recon, int precision) {
  int rc = QCU_SUCCESS;
  try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    ensure_init();
    QREQ(h && out && in, "qcuHopNoAlloc: NULL argument");
    QREQ(parity == 0 || parity == 1, "parity must be 0 or 1");
    QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    QREQ(t_boundary == 1 || t_boundary == -1, "t_boundary must be +1 or -1");
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    const long vh = g.geo.vh;
    const size_t fb = spinor_bytes(precision, vh);
    const size_t off = (fb + 255) & ~(size_t)255;  // the output field starts on a 256-byte boundary
    if (off + fb > kNoAllocScratchBytes) {
      char m[256];
      snprintf(m, sizeof m,
               "qcuHopNoAlloc: a half spinor of %ld sites in %s needs %zu B of staging, above the %zu B "
               "maximum of the static scratch; use qcuHop",
               vh, prec_name(precision), off + fb, kNoAllocScratchBytes);
      throw Error(QCU_ERROR_INVALID_VALUE, m);
    }
    // the packed links are made (allocated) on first use by the other entry points;
    // this one only uses them
    if (!g.packed[precision][recon == 12 ? 1 : 0].p)
      throw Error(QCU_ERROR_STATE,
                  "qcuHopNoAlloc: no packed links of this precision and recon on this gauge yet (one qcuHop, "
                  "qcuHopBatch or solve with them packs them)");
    char* base = static_cast<char*>(scratch_ptr());
    void* a = base;
    void* b = base + off;
    launch_spinor_from_aos(precision, in, 0, vec(a), vh, 1, stream());
    hop(g, precision, recon, EPI_PLAIN, parity, dagger ? 1 : 0, (double)t_boundary, vec(a), vec(b), vec(nullptr),
        nullptr, 1.0, 1, nullptr, stream());
    launch_spinor_to_aos(precision, vec(b), out, 0, vh, 1, stream());
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    rc = e.code;
  } catch (const std::exception& e) {
    set_last_error(QCU_ERROR_STATE, e.what());
    rc = QCU_ERROR_STATE;
  }
  return rc;
}
