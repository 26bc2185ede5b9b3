// The device-resident batch without allocation: inspire_respond_batch_device
// with its device scratch (digit addresses, transformed rows, the packing
// sums and b sums, transformed keys) in one array of this library's code
// object instead of context workspaces, for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops and harnesses
// that forbid hipMalloc).
//
// The array is a file-scope __device__ variable: every GPU the library runs
// on holds its own copy, at its own address, until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): the device address is looked up once per device,
// with that device current, when a context is created there, and that
// address is what the kernels receive.  Calls on one GPU share the array;
// this entry point holds a per-device lock for the whole call (it returns
// after the batch has finished), so calls from several threads or lanes on
// one GPU run one after another rather than racing.  Every part of the
// scratch a call reads is written earlier in the same call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "abi.hpp"
#include "context.hpp"
#include "pack.hpp"
#include "respond.hpp"

namespace inspire {
namespace {

// 24 MiB: 64 InsPIRe^2 queries at d = 2048 with two moduli and 16 columns
// (rows 96 KiB, step-1 sums 256 KiB and b sums 8 KiB per query), or 256
// InsPIRe^0 queries (rows only).
constexpr size_t kScratchBytes = (size_t)24 << 20;
constexpr uint32_t kMaxQueries = 1024;  // host staging of the digit addresses lives on the stack
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct DeviceScratch {
    std::once_flag once;
    hipError_t err = hipErrorNotInitialized;
    uint8_t* ptr = nullptr;  // device address of `scratch` on this device
    std::mutex call_mu;      // one call at a time on this device
};
DeviceScratch g_scratch[kMaxDevices];

}  // namespace

void noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;  // refused by the entry point
    DeviceScratch& D = g_scratch[device];
    std::call_once(D.once, [&] {
        void* p = nullptr;
        D.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (D.err == hipSuccess) D.ptr = static_cast<uint8_t*>(p);
        else (void)hipGetLastError();
    });
}

}  // namespace inspire

using namespace inspire;

extern "C" int inspire_respond_batch_device_noalloc(inspire_ctx* ctx, const inspire_device_batch* batch) {
    INSPIRE_TRY({
        if (!ctx || !batch) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        LaneRead lr(ctx);
        const int dev = ctx->device;
        if (dev < 0 || dev >= kMaxDevices)
            throw Error{INSPIRE_ERR_UNSUPPORTED, "the no-allocation scratch covers devices 0.." +
                                                     std::to_string(kMaxDevices - 1)};
        DeviceScratch& D = g_scratch[dev];
        if (!D.ptr)
            throw Error{INSPIRE_ERR_DEVICE, std::string("no-allocation scratch unavailable on device ") +
                                                std::to_string(dev) + ": " + hipGetErrorString(D.err)};
        if (batch->count == 0) return INSPIRE_OK;
        if (batch->count > kMaxQueries)
            throw Error{INSPIRE_ERR_CAPACITY, "the no-allocation entry point takes at most " +
                                                  std::to_string(kMaxQueries) + " queries per call"};
        const DeviceBatchPlan P = device_batch_plan(ctx, *batch);
        if (P.tree)
            throw Error{INSPIRE_ERR_UNSUPPORTED,
                        "tree packing is not on the no-allocation entry point (its node state lives in context "
                        "workspaces)"};
        if (P.packed && !P.tables)
            throw Error{INSPIRE_ERR_UNSUPPORTED,
                        "the InspiRING tables of " + std::to_string(P.cols) +
                            " columns are not built yet (an allocating respond call on this context builds them)"};
        if (P.bytes > kScratchBytes) {
            const size_t per = (P.bytes + batch->count - 1) / batch->count;
            throw Error{INSPIRE_ERR_CAPACITY, "this batch needs " + std::to_string(P.bytes) +
                                                  " bytes of scratch; the no-allocation entry point holds " +
                                                  std::to_string(kScratchBytes) + " (about " +
                                                  std::to_string(kScratchBytes / per) + " queries of this shape)"};
        }
        uint64_t host_meta[kMaxQueries + 2];
        std::lock_guard<std::mutex> dl(D.call_mu);
        ProfSuspend ps(ctx);
        respond_batch_device(ctx, *batch, D.ptr, host_meta);
        return INSPIRE_OK;
    })
}
