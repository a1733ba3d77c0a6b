// qsg_nearest_facet_probe_noalloc: the nearest-facet probe for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  It runs the same kernel as
// qsg_nearest_facet_probe (k_nearest_probe, the tracking kernels' search) but
// stages the rays and results through one fixed device array, `scratch`,
// instead of a per-call allocation.
//
// The array holds kNoAllocMaxRays rays (88 B each: position, direction and
// nudged position 3 x 8 B each, distance 8 B, cell and facet 4 B each), in
// six sections at fixed offsets.  Every GPU the library runs on keeps its own
// copy until the process ends.  Its device address is looked up once per
// device (hipGetSymbolAddress with that device current) and only that address
// is handed to copies and kernels.  Calls on one device are serialised by a
// host lock held until the call's stream has drained, so calls from several
// host threads do not overlap on the array; nothing in it needs clearing,
// because each call writes every entry it reads back.
#include <hip/hip_runtime.h>

#include <mutex>

#include "qs_device.h"
#include "qsg.h"
#include "qsg_internal.h"

using qsgi::hipCheck;
using qsgi::QsgError;

namespace {
constexpr uint64_t kNoAllocMaxRays = QSG_PROBE_NOALLOC_MAX_RAYS;
constexpr size_t kVec3Bytes = kNoAllocMaxRays * 3 * sizeof(double);
constexpr size_t kOffXyz = 0;
constexpr size_t kOffDir = kOffXyz + kVec3Bytes;
constexpr size_t kOffOut = kOffDir + kVec3Bytes;
constexpr size_t kOffDist = kOffOut + kVec3Bytes;
constexpr size_t kOffCell = kOffDist + kNoAllocMaxRays * sizeof(double);
constexpr size_t kOffFacet = kOffCell + kNoAllocMaxRays * sizeof(int32_t);
constexpr size_t kScratchBytes = kOffFacet + kNoAllocMaxRays * sizeof(int32_t);
static_assert(kOffCell % 256 == 0 && kOffFacet % 256 == 0, "sections stay 256-byte aligned");
}  // namespace

// File-scope device array: in host code this name is only a handle for
// hipGetSymbolAddress, never an address to pass to a copy or a kernel.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex[kMaxDevices];
unsigned char* g_scratchPtr[kMaxDevices] = {};  // device address of `scratch`, per device

// The calling thread's current device must be `device`; call with
// g_scratchMutex[device] held.
unsigned char* scratchAddress(int device) {
    if (!g_scratchPtr[device]) {
        void* ptr = nullptr;
        hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratchPtr[device] = static_cast<unsigned char*>(ptr);
    }
    return g_scratchPtr[device];
}
}  // namespace

extern "C" int qsg_nearest_facet_probe_noalloc(qsg_problem p, uint64_t n, const double* xyz, const double* dir,
                                               const int32_t* cell, double* distance, int32_t* facet,
                                               double* xyz_out) {
    QSG_GUARD({
        if (!p || !xyz || !dir || !cell || !distance || !facet || !xyz_out)
            throw QsgError(QSG_ERR_ARG, "NULL argument");
        if (n > kNoAllocMaxRays)
            throw QsgError(QSG_ERR_ARG, "qsg_nearest_facet_probe_noalloc takes at most " +
                                            std::to_string(kNoAllocMaxRays) + " rays per call");
        for (uint64_t q = 0; q < n; q++)
            if (cell[q] < 0 || cell[q] >= p->tables.nCells) throw QsgError(QSG_ERR_ARG, "cell id outside the mesh");
        if (n == 0) return QSG_OK;
        if (p->device < 0 || p->device >= kMaxDevices) throw QsgError(QSG_ERR_ARG, "device index out of range");
        std::lock_guard<std::mutex> lock(g_scratchMutex[p->device]);
        hipCheck(hipSetDevice(p->device), "hipSetDevice");
        unsigned char* base = scratchAddress(p->device);
        double* dXyz = reinterpret_cast<double*>(base + kOffXyz);
        double* dDir = reinterpret_cast<double*>(base + kOffDir);
        double* dOut = reinterpret_cast<double*>(base + kOffOut);
        double* dDist = reinterpret_cast<double*>(base + kOffDist);
        int32_t* dCell = reinterpret_cast<int32_t*>(base + kOffCell);
        int32_t* dFacet = reinterpret_cast<int32_t*>(base + kOffFacet);
        const size_t v3 = (size_t)n * 3 * sizeof(double);
        try {
            hipCheck(hipMemcpyAsync(dXyz, xyz, v3, hipMemcpyHostToDevice, p->stream), "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(dDir, dir, v3, hipMemcpyHostToDevice, p->stream), "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(dCell, cell, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, p->stream),
                     "hipMemcpyAsync");
            qs::launchNearestProbe(p->ctx(), p->dp, dXyz, dDir, dCell, dDist, dFacet, dOut, n);
            hipCheck(hipGetLastError(), "k_nearest_probe launch");
            hipCheck(hipMemcpyAsync(distance, dDist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, p->stream),
                     "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(facet, dFacet, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, p->stream),
                     "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(xyz_out, dOut, v3, hipMemcpyDeviceToHost, p->stream), "hipMemcpyAsync");
            hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        } catch (...) {
            (void)hipStreamSynchronize(p->stream);
            throw;
        }
    })
}
