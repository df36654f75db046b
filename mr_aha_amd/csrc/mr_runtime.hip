// Runtime pieces of libmr_aha_c: backend identity, HIP error hook and the
// per-device cache of FFT twiddle tables.

#include "mr_hip.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

void mr_backend_clear_error() { (void)hipGetLastError(); }

// "cuda" is the value that makes the reference's bindings treat every array
// argument as a device pointer (python/mr_aha/_lib.py:700-715); this library
// follows that device-pointer ABI on HIP devices.
extern "C" MR_EXPORT const char *mr_backend(void) { return "cuda"; }

namespace {
std::mutex g_tw_mu;
std::map<std::tuple<int, int, int>, void *> g_tw;  // (device, n, bytes per real) -> table
}  // namespace

template <typename R> int mr_fft_get(int n, mr_fft_len<R> *out)
{
    using V = typename cvec<R>::T;
    MR_REQUIRE(n >= 1, "FFT length %d must be >= 1", n);
    const size_t lds = ((size_t)n + 2 * (size_t)mr_fft_ns(n)) * sizeof(V);
    if (lds > 64 * 1024)
        MR_FAIL(MR_ERR_UNSUPPORTED, "FFT length %d exceeds the single-pencil LDS limit", n);
    out->n = n;
    out->nst = mr_fft_factor(n, out->radix);
    if (out->nst < 0) MR_FAIL(MR_ERR_UNSUPPORTED, "FFT length %d has too many factors", n);
    int dev = 0;
    MR_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_tuple(dev, n, (int)sizeof(R));
    auto it = g_tw.find(key);
    if (it == g_tw.end()) {
        std::vector<V> h(n);
        for (int q = 0; q < n; ++q) {
            // exact quadrant values, long double elsewhere
            const long double a = -2.0L * 3.14159265358979323846264338327950288L * q / n;
            h[q].x = (R)cosl(a);
            h[q].y = (R)sinl(a);
            if ((4LL * q) % n == 0) {
                const int quad = (int)((4LL * q) / n);
                const R cs[4] = {1, 0, -1, 0}, sn[4] = {0, -1, 0, 1};
                h[q].x = cs[quad];
                h[q].y = sn[quad];
            }
        }
        // stage-major copy: entry r * Ns + k - 1 of stage (R, Ns) is
        // w_n^(k r n / (Ns R)); the stages tile [0, n - 1)
        std::vector<V> hs(n, h[0]);
        int Ns = 1;
        for (int s = 0; s < out->nst; ++s) {
            const int rad = out->radix[s];
            const int stride = n / (Ns * rad);
            for (int r = 1; r < rad; ++r)
                for (int k = 0; k < Ns; ++k) hs[r * Ns + k - 1] = h[(long long)k * r * stride];
            Ns *= rad;
        }
        // Uploaded on a private non-blocking stream: a legacy-stream copy
        // would be ordered after every blocking stream, including one that
        // another thread is capturing into a graph.
        void *d = nullptr;
        hipStream_t up = nullptr;
        MR_HIP_CHECK(hipMalloc(&d, 2 * n * sizeof(V)));
        MR_HIP_CHECK(hipStreamCreateWithFlags(&up, hipStreamNonBlocking));
        hipError_t e = hipMemcpyAsync(d, h.data(), n * sizeof(V), hipMemcpyHostToDevice, up);
        if (e == hipSuccess)
            e = hipMemcpyAsync(reinterpret_cast<V *>(d) + n, hs.data(), n * sizeof(V),
                               hipMemcpyHostToDevice, up);
        if (e == hipSuccess) e = hipStreamSynchronize(up);
        (void)hipStreamDestroy(up);
        if (e != hipSuccess) {
            (void)hipFree(d);
            MR_FAIL(MR_ERR_CUDA, "uploading the twiddles of length %d failed: %s", n,
                    hipGetErrorString(e));
        }
        it = g_tw.emplace(key, d).first;
    }
    out->tw = reinterpret_cast<const V *>(it->second);
    out->tws = out->tw + n;
    mr_fft_fill_divs(out);
    return MR_SUCCESS;
}

template int mr_fft_get<float>(int, mr_fft_len<float> *);
template int mr_fft_get<double>(int, mr_fft_len<double> *);

void mr_fft_release_all()
{
    std::lock_guard<std::mutex> lk(g_tw_mu);
    for (auto &kv : g_tw) (void)hipFree(kv.second);
    g_tw.clear();
}

// ---------------------------------------------------------------------------
// Streaming copy: the attainable HBM bandwidth that the passes are measured
// against (read + write of `bytes`). 16-byte vector loads and stores with the
// non-temporal hint, eight per work-item in flight, a grid of a few
// workgroups per CU striding over the buffer.
// ---------------------------------------------------------------------------
namespace {
typedef unsigned int mr_cp_u4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) mr_k_stream_copy(const mr_cp_u4 *__restrict__ src,
                                                         mr_cp_u4 *__restrict__ dst, long long n)
{
    constexpr int U = 8;
    const long long stride = (long long)gridDim.x * 256 * U;
    for (long long b = (long long)blockIdx.x * 256 * U + threadIdx.x; b < n; b += stride) {
        mr_cp_u4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b + (long long)u * 256 < n) v[u] = __builtin_nontemporal_load(src + b + (long long)u * 256);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b + (long long)u * 256 < n) __builtin_nontemporal_store(v[u], dst + b + (long long)u * 256);
    }
}
}  // namespace

extern "C" MR_EXPORT int mr_stream_copy(const void *d_src, void *d_dst, int64_t bytes,
                                        void *stream)
{
    mr_clear_error();
    MR_REQUIRE(d_src != nullptr && d_dst != nullptr, "mr_stream_copy: NULL buffer");
    MR_REQUIRE(bytes >= 0 && bytes % 16 == 0, "mr_stream_copy: %lld bytes is not a multiple of 16",
               (long long)bytes);
    MR_REQUIRE(((uintptr_t)d_src & 15) == 0 && ((uintptr_t)d_dst & 15) == 0,
               "mr_stream_copy: buffers must be 16-byte aligned");
    if (bytes == 0) return MR_SUCCESS;
    int dev = 0, ncu = 0;
    MR_HIP_CHECK(hipGetDevice(&dev));
    MR_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long long n = bytes / 16;
    long long blocks = (n + 256 * 8 - 1) / (256 * 8);
    const long long cap = (long long)ncu * 8;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(mr_k_stream_copy, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, (const mr_cp_u4 *)d_src, (mr_cp_u4 *)d_dst, n);
    MR_LAUNCH_CHECK("mr_k_stream_copy");
    return MR_SUCCESS;
}
