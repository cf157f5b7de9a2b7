// The allocation-free variant of the dense LDL^T factor-and-solve entry
// point, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its scratch (the 64-column panel buffer, the D^-1/2 vector, the pivot
// error word) and its launch descriptors live in one file-scope device
// array, `scratch`, of kScratchBytes per device.  The array's device address
// is looked up once per device (scratchInit, called by amls_hip_create with
// the device current) and kept per device; every later use goes through
// that address.  Calls on one device share the array, so calls that overlap
// on one device race; each call sets the error word itself, and every other
// scratch element is written before it is read.
#include <climits>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"

namespace amls {

namespace {
constexpr size_t kScratchBytes = size_t(8) << 20;   // 8 MiB per device
std::mutex g_scratchMu;
std::vector<void*> g_scratchPtr;                     // device address per device id
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

size_t noallocScratchBytes() { return kScratchBytes; }

// With `device` current: records the device address of `scratch` on it.
hipError_t scratchInit(int device) {
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (device < 0) return hipErrorInvalidDevice;
    if (static_cast<int>(g_scratchPtr.size()) <= device) g_scratchPtr.resize(device + 1, nullptr);
    if (g_scratchPtr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchPtr[device] = ptr;
    return hipSuccess;
}

namespace {
char* scratchOf(int device) {
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (device < 0 || device >= static_cast<int>(g_scratchPtr.size()) || !g_scratchPtr[device])
        throw StatusError(AMLS_E_STATE, "scratch of this device was not set up");
    return static_cast<char*>(g_scratchPtr[device]);
}
inline size_t up256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

// A (n x n, device, lower used) := its LDL^T factor (L strict lower, D on the
// diagonal); B (n x nrhs, device) := A^-1 B.  The same launches as
// amls_hip_ldlt_solve, with scratch and descriptors in `scratch`.
void ldltSolveNoalloc(Ctx& c, int64_t n, double* A, int64_t nrhs, double* B) {
    if (n <= 0 || nrhs < 0) throw StatusError(AMLS_E_ARG, "ldlt_solve_noalloc: bad size");
    if (n > INT32_MAX || nrhs > INT32_MAX)
        throw StatusError(AMLS_E_ARG, "ldlt_solve_noalloc: n and nrhs must fit in 32 bits");
    // panel scratch, D^-1/2 and the error word first, descriptors after them
    const size_t oW = 0, oq = up256(size_t(n) * 64 * 8), oe = oq + up256(size_t(n) * 8),
                 od = oe + 256;
    if (od > kScratchBytes)
        throw StatusError(AMLS_E_ARG, "ldlt_solve_noalloc: n = " + std::to_string(n) +
                                          " needs more than the " + std::to_string(kScratchBytes) +
                                          " bytes of scratch");
    if (!A || (nrhs > 0 && !B)) throw StatusError(AMLS_E_ARG, "null matrix argument");
    c.requireGpu();
    char* base = scratchOf(c.device);
    double* dW = reinterpret_cast<double*>(base + oW);
    double* dq = reinterpret_cast<double*>(base + oq);
    int* derr = reinterpret_cast<int*>(base + oe);
    Blob blob;
    blob.useExternal(base + od, kScratchBytes - od);
    Plan plan(&blob, nullptr);   // no event timing: events are allocations too
    addLdltSolve(plan, A, n, B, nrhs, dW, dq, derr);
    if (od + blob.size() > kScratchBytes)
        throw StatusError(AMLS_E_ARG, "ldlt_solve_noalloc: n = " + std::to_string(n) +
                                          ", nrhs = " + std::to_string(nrhs) + " needs " +
                                          std::to_string(od + blob.size()) +
                                          " bytes of scratch, the maximum is " +
                                          std::to_string(kScratchBytes));
    plan.run(c.stream);
    int e = 0;
    AMLS_HIP_CHECK(hipMemcpyAsync(&e, derr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    AMLS_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (e != INT_MAX) throw StatusError(AMLS_E_NOT_SPD, "LDL^T: non-positive or non-finite pivot");
}

}  // namespace amls
