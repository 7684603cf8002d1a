// uno_eigh_packed_noalloc: the packed symmetric eigensolver of uno_eigh_packed for callers
// that cannot allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc).  Its device memory is one file-scope array, `scratch`,
// of a fixed size (matrices up to kNoAllocMaxN functions); each GPU the library runs on
// holds its own copy of the array until the process ends.
//
// Host code never uses the array's name as an address: on a device's first call the
// device-side address is taken with hipGetSymbolAddress (the device current) and kept per
// device; kernels receive that pointer.  Every call on one GPU shares that GPU's array and
// finds the previous call's contents in it: the solver's kernels write every element they
// read first (the reflector rows are zero-filled by the solver itself), so nothing needs
// clearing between calls.  Calls that overlap on one GPU (several host threads) race on the
// array; like the rest of the library, the entry point is not re-entrant.
#include <hip/hip_runtime.h>
#include <cstring>
#include "../host/engine.hpp"
#include "kernels.hpp"

namespace {

constexpr int kNoAllocMaxN = 256;
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 32;   // doubles: 256 bytes
constexpr size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
// dense A (destroyed), dense Z, the packed input, then the solver's work space
// (7 n^2 + small vectors, see eigh_ws_doubles)
constexpr size_t kMaxSq = (size_t)kNoAllocMaxN * kNoAllocMaxN;
constexpr size_t kScratchDoubles = 2 * kMaxSq + round_up(kMaxSq / 2 + kNoAllocMaxN) + 7 * kMaxSq + 16 * 1024;

static __device__ __align__(256) double scratch[kScratchDoubles];   // 5.1 MiB per GPU

double* g_scratch[kMaxDevices] = {};

}  // namespace

extern "C" int uno_eigh_packed_noalloc(int n, double* a, double* z, double* evals) {
    using namespace uno;
    if (n <= 0 || !a || !z || !evals) { set_error("bad argument"); return UNO_EINVAL; }
    if (n > kNoAllocMaxN) {
        set_error("uno_eigh_packed_noalloc: n = " + std::to_string(n) + " exceeds its fixed work space (n <= " +
                  std::to_string(kNoAllocMaxN) + "); use uno_eigh_packed");
        return UNO_EINVAL;
    }
    int dev = 0;
    UNO_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return UNO_EDEVICE; }
    if (!g_scratch[dev]) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return UNO_EDEVICE;
        }
        g_scratch[dev] = p;
    }
    double* base = g_scratch[dev];
    const size_t nn = (size_t)n * n, n2 = (size_t)n * (n + 1) / 2;
    double* A = base;
    double* Z = A + round_up(nn);
    double* pk = Z + round_up(nn);
    double* ws = pk + round_up(n2);
    if ((size_t)(ws - base) + eigh_ws_doubles(n) > kScratchDoubles) {
        set_error("uno_eigh_packed_noalloc: work space layout exceeds the fixed array");
        return UNO_EINVAL;
    }
    hipStream_t st = nullptr;   // the null stream: no stream is created either
    double hev[kNoAllocMaxN];
    int hcl[kNoAllocMaxN];
    UNO_HIP(hipMemcpyAsync(pk, a, n2 * sizeof(double), hipMemcpyHostToDevice, st));
    UNO_HIP(k_unpack(n, pk, A, st));
    const int rc = eigh_in_workspace(n, A, Z, evals, ws, hev, hcl, st);
    if (rc) return rc;
    UNO_HIP(hipMemcpyAsync(z, Z, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    UNO_HIP(hipStreamSynchronize(st));
    std::memset(a, 0, n2 * sizeof(double));   // input destroyed like rsp
    return UNO_OK;
}
