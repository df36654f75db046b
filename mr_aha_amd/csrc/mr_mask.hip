// mr_calc_mask on gfx950: nearest upsampled grid point of every sample,
// (2r+1)^d box dilation with periodic wrap, ascending compaction.
//
// Same integer result as src/cuda/mr_calc_mask.cu:63-181, 329-341: the grid
// index is rintf(x * (float)(M / 2pi)) + M/2 mod M computed in float32, the
// 2D mask of a stack is replicated over its Mz planes.

#include "mr_hip.hpp"

#include <hipcub/hipcub.hpp>

__global__ void mr_k_mask_mark(unsigned char *marks, const float *traj, long long ns, int ndim,
                               int m0, int m1, int m2, float s0, float s1, float s2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const int m[3] = {m0, m1, m2};
    const float sc[3] = {s0, s1, s2};
    long long idx = 0, stride = 1;
    for (int a = 0; a < ndim; ++a) {
        const float prod = __fmul_rn(traj[t + ns * a], sc[a]);
        long long i = (long long)rintf(prod) + m[a] / 2;
        i %= m[a];
        if (i < 0) i += m[a];
        idx += i * stride;
        stride *= m[a];
    }
    marks[idx] = 1;
}

__global__ void mr_k_mask_dilate(unsigned char *out, const unsigned char *in, long long n, int m,
                                 long long stride, int r)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)((t / stride) % m);
    const long long base = t - (long long)i * stride;
    unsigned char v = in[t];
    for (int d = -r; d <= r && !v; ++d) {
        int j = (i + d) % m;
        if (j < 0) j += m;
        v = in[base + (long long)j * stride];
    }
    out[t] = v;
}

__global__ void mr_k_mask_replicate(int64_t *mask, long long nplane, long long plane, int Mz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nplane * (long long)(Mz - 1)) return;
    const long long i = t % nplane, kz = t / nplane + 1;
    mask[i + nplane * kz] = mask[i] + plane * kz;
}

static int mr_calc_mask_impl(const mr_config *cfg, const float *d_traj, int64_t Ntrajs,
                             int64_t Nkx, int64_t *d_mask, int64_t *num_out)
{
    hipStream_t s = (hipStream_t)cfg->stream;
    const bool three = cfg->traj_type == MR_TRAJ_3D;
    const int m[3] = {cfg->phi2_dim[0], cfg->phi2_dim[1], three ? cfg->phi2_dim[2] : 1};
    const long long n = (long long)m[0] * m[1] * m[2];
    const long long ns = (long long)Ntrajs * Nkx;
    const double tau = 6.283185307179586;
    unsigned char *a = nullptr, *b = nullptr;
    int64_t *d_cnt = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = MR_SUCCESS;
    hipError_t e = hipMalloc((void **)&a, n);
    if (e == hipSuccess) e = hipMalloc((void **)&b, n);
    if (e == hipSuccess) e = hipMalloc((void **)&d_cnt, sizeof(int64_t));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        rc = MR_ERR_ALLOC;
        mr_set_error("failed to allocate the %lld-element mask grid", n);
    }
    if (rc == MR_SUCCESS) {
        e = hipMemsetAsync(a, 0, n, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(mr_k_mask_mark, dim3((unsigned)mr_ceil_div(ns, 256)), dim3(256), 0,
                               s, a, d_traj, ns, three ? 3 : 2, m[0], m[1], m[2],
                               (float)(m[0] / tau), (float)(m[1] / tau), (float)(m[2] / tau));
            e = hipGetLastError();
        }
        long long stride = 1;
        for (int ax = 0; ax < 3 && e == hipSuccess; ++ax) {
            if (m[ax] > 1 && cfg->mask_radius > 0) {
                hipLaunchKernelGGL(mr_k_mask_dilate, dim3((unsigned)mr_ceil_div(n, 256)),
                                   dim3(256), 0, s, b, a, n, m[ax], stride, cfg->mask_radius);
                e = hipGetLastError();
                unsigned char *t = a; a = b; b = t;
            }
            stride *= m[ax];
        }
        hipcub::CountingInputIterator<int64_t> first(0);
        int64_t *dst = d_mask;
        int64_t *scratch_out = nullptr;
        if (e == hipSuccess && !dst) {
            e = hipMalloc((void **)&scratch_out, n * sizeof(int64_t));
            dst = scratch_out;
        }
        if (e == hipSuccess)
            e = hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, first, a, dst, d_cnt, (int64_t)n, s);
        if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1);
        if (e == hipSuccess)
            e = hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, first, a, dst, d_cnt, (int64_t)n, s);
        int64_t kept = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&kept, d_cnt, sizeof(int64_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (scratch_out) (void)hipFree(scratch_out);
        if (e == hipSuccess && cfg->traj_type == MR_TRAJ_STACK) {
            const int Mz = cfg->phi2_dim[2];
            if (d_mask && Mz > 1 && kept > 0) {
                const long long total = kept * (long long)(Mz - 1);
                hipLaunchKernelGGL(mr_k_mask_replicate, dim3((unsigned)mr_ceil_div(total, 256)),
                                   dim3(256), 0, s, d_mask, (long long)kept,
                                   (long long)m[0] * m[1], Mz);
                e = hipGetLastError();
                if (e == hipSuccess) e = hipStreamSynchronize(s);
            }
            kept *= Mz;
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            rc = MR_ERR_CUDA;
            mr_set_error("mr_calc_mask failed: %s", hipGetErrorString(e));
        } else {
            *num_out = kept;
            if (cfg->verbose)
                printf("[mr_aha] calc_mask: kept %lld of %lld upsampled locations (%.1f%%) within "
                       "radius %d of a sample\n",
                       (long long)kept, (long long)mr_pad_size(cfg),
                       100.0 * (double)kept / (double)mr_pad_size(cfg), cfg->mask_radius);
        }
    }
    if (tmp) (void)hipFree(tmp);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (d_cnt) (void)hipFree(d_cnt);
    return rc;
}

extern "C" MR_EXPORT int mr_calc_mask(const mr_config *cfg, const float *d_traj, int64_t Ntrajs,
                                      int64_t Nkx, int64_t *d_mask, int64_t *num_masked_out)
{
    mr_clear_error();
    MR_REQUIRE(cfg != nullptr, "mr_calc_mask: cfg is NULL");
    MR_REQUIRE(d_traj != nullptr, "mr_calc_mask: d_traj is NULL");
    MR_REQUIRE(num_masked_out != nullptr, "mr_calc_mask: num_masked_out is NULL");
    MR_REQUIRE(Ntrajs > 0 && Nkx > 0, "mr_calc_mask: Ntrajs = %lld and Nkx = %lld must be > 0",
               (long long)Ntrajs, (long long)Nkx);
    *num_masked_out = 0;
    // The mask depends only on the geometry: validate an unmasked copy so a
    // configuration still waiting for its num_masked is accepted.
    mr_config c = *cfg;
    c.is_masked = 0;
    MR_TRY(mr_config_validate(&c));
    int prev = -1;
    int rc = mr_push_device(cfg, &prev);
    if (rc == MR_SUCCESS) rc = mr_calc_mask_impl(&c, d_traj, Ntrajs, Nkx, d_mask, num_masked_out);
    mr_pop_device(prev);
    return rc;
}
