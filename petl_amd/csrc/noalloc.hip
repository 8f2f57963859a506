// backproject_gpu for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops): the fast back projector's workspace is a fixed device array instead of a buffer the
// library allocates.  The kernel tables are built beforehand by prepare_gpu() (or any earlier projector
// call on the same geometry); the entry point itself allocates nothing.
#include <map>
#include <mutex>
#include "petl_common.h"
#include "petl_model.h"
#include "../../include/petl_c_interface.h"

namespace petl {
namespace {

// the largest back projector workspace the entry point takes: one planogram's u0-gathered columns,
// N_v1 x N_u1 x (in-plane voxels) floats (66 MB for two 64 x 64-crystal planograms and a 64^3 volume)
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;

// Every device the library runs on holds one copy of this array until the process ends.  Host code uses
// the per-device address hipGetSymbolAddress returns, never the array's name.
__device__ __align__(256) float scratch[NOALLOC_SCRATCH_BYTES / sizeof(float)];

std::mutex scratch_mu;
std::map<int, float*> scratch_ptr;  // device -> address of its copy of `scratch`

float* scratch_for_current_device()
{
    int dev = -1;
    PETL_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(scratch_mu);
    auto it = scratch_ptr.find(dev);
    if (it != scratch_ptr.end())
        return it->second;
    float* p = nullptr;
    PETL_HIP_CHECK(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)));
    if (!p)
        throw hip_error("backproject_gpu_noalloc: no device address for the scratch array");
    scratch_ptr[dev] = p;
    return p;
}

template <typename Fn>
bool guarded_noalloc(Fn&& fn)
{
    try {
        return fn();
    } catch (const std::exception& e) {
        fprintf(stderr, "PETL backproject_gpu_noalloc failed: %s\n", e.what());
    }
    return false;
}

}  // namespace
}  // namespace petl

using petl::Model;

extern "C" {

bool prepare_gpu(void)
{
    Model* m = petl::current_model();
    if (!m->allDefined())
        return false;
    return petl::guarded_noalloc([&] {
        m->prepare();
        return true;
    });
}

bool backproject_gpu_noalloc(float** d_g, float* d_f, void* stream, bool sync)
{
    Model* m = petl::current_model();
    if (!d_g || !d_f || !m->allDefined())
        return false;
    return petl::guarded_noalloc([&] {
        if (!m->prepared())
            throw petl::hip_error("the kernel tables of this geometry are not built: call prepare_gpu() (or any "
                                  "projector call) first");
        if (m->dev.h2_need > petl::NOALLOC_SCRATCH_BYTES)
            throw petl::hip_error("this geometry's back projector workspace (" + std::to_string(m->dev.h2_need) +
                                  " bytes) exceeds the fixed scratch array (" +
                                  std::to_string(petl::NOALLOC_SCRATCH_BYTES) + " bytes)");
        // set-up on the first call on each device, then the array's address on that device
        float* ws = petl::scratch_for_current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        m->back(d_g, d_f, m->back_mode, s, 0, 1, ws);
        if (sync)
            PETL_HIP_CHECK(hipStreamSynchronize(s));
        return true;
    });
}

}  // extern "C"
