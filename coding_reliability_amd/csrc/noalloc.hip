// Allocation-free exhaustive AN pair histogram: the pair kernel of
// cr_an_pairs_tuned with its job descriptor, work-queue head and histogram
// copies in a file-scope device array instead of context buffers, enqueued on
// the caller's stream with no host synchronisation, so that callers that
// cannot allocate (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc) can run it.
#include <mutex>

#include "cr_device.h"
#include "cr_internal.h"

namespace cr {
namespace {

// Scratch layout (bytes): the PairJob at 0, the work-queue head at 256, the
// kHistCopies histogram copies (hist_copy_stride(65) = 80 entries apart) at
// 512.  One array per GPU, alive until the process ends.
constexpr size_t kJobOffset = 0, kQueueOffset = 256, kCopiesOffset = 512;
constexpr uint32_t kMaxBins = 65;  // distances 0..64: every codeword of up to 64 bits
constexpr size_t kScratchBytes = 32768;
static_assert(kCopiesOffset + (size_t)dev::kHistCopies * 80 * sizeof(uint64_t) <= kScratchBytes, "scratch size");
static_assert(sizeof(PairJob) <= kQueueOffset, "job slot");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
int g_cus[kMaxDevices] = {};

// Writes the job descriptor (a kernel argument) into the scratch and clears
// the queue head and the histogram copies: every call starts from zero,
// whatever the previous call on this GPU left there.
__global__ __launch_bounds__(256) void noalloc_setup(PairJob job, unsigned char *base, uint32_t entries) {
    const uint32_t t = threadIdx.x;
    if (t < sizeof(PairJob) / sizeof(uint32_t))
        reinterpret_cast<uint32_t *>(base + kJobOffset)[t] = reinterpret_cast<const uint32_t *>(&job)[t];
    if (t == 0) *reinterpret_cast<unsigned long long *>(base + kQueueOffset) = 0;
    unsigned long long *copies = reinterpret_cast<unsigned long long *>(base + kCopiesOffset);
    for (uint32_t e = t; e < entries; e += blockDim.x) copies[e] = 0;
}

// raw[b] = sum of the copies' bin b for b < need, 0 for need <= b < nbins.
__global__ __launch_bounds__(128) void noalloc_reduce(const unsigned char *base, uint32_t need, uint32_t stride,
                                                      uint64_t *raw, uint32_t nbins) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    const unsigned long long *copies = reinterpret_cast<const unsigned long long *>(base + kCopiesOffset);
    uint64_t s = 0;
    if (b < need)
        for (uint32_t c = 0; c < dev::kHistCopies; ++c) s += copies[(uint64_t)c * stride + b];
    raw[b] = s;
}

}  // namespace

// The scratch address of `device` (made current by the caller): looked up
// once per device with hipGetSymbolAddress, each GPU holding its own copy of
// the array at its own address.  Host code uses only this address, never the
// array's name.
cr_status noalloc_setup_device(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(CR_ERR_NO_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return CR_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(CR_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                    hipGetErrorString(e));
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess || cus <= 0) return fail(CR_ERR_DEVICE, "CU count query failed on device %d", device);
    g_cus[device] = cus;
    g_ptr[device] = p;
    return CR_OK;
}

}  // namespace cr

using namespace cr;

extern "C" cr_status cr_an_pairs_noalloc(int device, uint64_t A, uint64_t num_words, uint64_t *raw_device,
                                         uint32_t nbins, void *stream) {
    clear_error();
    if (!raw_device) return fail(CR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (A == 0 || num_words == 0) return fail(CR_ERR_INVALID_ARGUMENT, "A and num_words must be >= 1");
    if ((unsigned __int128)(num_words - 1) * A >> 64)
        return fail(CR_ERR_UNSUPPORTED, "codewords (num_words-1)*A exceed 64 bits");
    const uint32_t need = pair_max_distance(A, num_words) + 1;
    if (nbins < need) return fail(CR_ERR_INVALID_ARGUMENT, "nbins=%u but distances reach %u", nbins, need - 1);
    if (need > kMaxBins)
        return fail(CR_ERR_UNSUPPORTED, "%u histogram bins exceed the %u the static scratch holds", need, kMaxBins);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(CR_ERR_NO_DEVICE, "device %d not in [0, %d)", device, ndev);
    CR_HIP(hipSetDevice(device));
    CR_TRY(noalloc_setup_device(device));
    unsigned char *base = static_cast<unsigned char *>(g_ptr[device]);
    hipStream_t s = static_cast<hipStream_t>(stream);

    // the same geometry as cr_an_pairs_tuned with automatic tuning; the
    // histogram always as kHistCopies copies (summed by noalloc_reduce)
    PairLaunchConfig cfg;
    cfg.lds_rows = need + 1;
    const uint32_t per_cu = pair_waves_per_cu(8, cfg.lds_rows);
    uint32_t R = 0, Q = 0;
    auto_pair_tuning(num_words, per_cu * (uint32_t)g_cus[device], R, Q);
    cfg.R = R;
    const long double pairs = (long double)num_words * (long double)(num_words - 1) / 2;
    cfg.allow_static = pairs < kStaticPairs;
    cfg.zero_counter = false;

    PairJob job{};
    job.A = A;
    job.N = num_words;
    job.mode = pair_mode(A, num_words);
    job.nbins = need;
    pair_job_geometry(job, R, Q);
    job.item_begin = 0;
    job.out_copies = dev::kHistCopies;
    job.out = reinterpret_cast<unsigned long long *>(base + kCopiesOffset);
    const uint32_t cstride = (uint32_t)dev::hist_copy_stride(need);
    const uint32_t entries = cstride * dev::kHistCopies;

    hipLaunchKernelGGL(noalloc_setup, dim3(1), dim3(256), 0, s, job, base, entries);
    CR_HIP(hipGetLastError());
    CR_TRY(launch_pairs(reinterpret_cast<const PairJob *>(base + kJobOffset), 1, job.items,
                        reinterpret_cast<unsigned long long *>(base + kQueueOffset), cfg, g_cus[device], s, nullptr,
                        nullptr));
    hipLaunchKernelGGL(noalloc_reduce, dim3((nbins + 127) / 128), dim3(128), 0, s, base, need, cstride, raw_device,
                       nbins);
    CR_HIP(hipGetLastError());
    return CR_OK;
}
