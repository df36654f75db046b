// Toeplitz kernel build on gfx950 (mr_calc_toep, mr_calc_toep_from_psf).
//
//   phi2_toep[kappa, t] = (1/prod M_tr) sum_delta psf_t[delta] e^{+2 pi i kappa.delta/M}
//   psf_t[delta]        = sum_m c_{t,m} e^{-i delta.x_m}
// (reference: src/cuda/mr_calc_toep.cu:10-24, 171-200, 498-643). Transformed
// axes: x, y for 2D and stack (every kz plane its own component, no shift
// and no scaling along z), x, y, z for 3D.
//
// MI355X pipeline per chunk of components q = z + Nkz * t:
//   1. gather-spread (FP64) of the readout-constant strengths onto the
//      2x fine grid (mr_grid_spread: no atomics, one strength multiply per
//      readout run of each bin),
//   2. forward FFT of the fine grids (FP64, sign -1),
//   3. deconvolution by the ES kernel's transform and crop to the M modes,
//      written in natural order,
//   4. inverse FFT over the M grid (FP64, sign +1),
//   5. natural -> centered shift, 1/prod M, FP32 store (or mask gather).

#include "mr_grid.hpp"

#include <chrono>
#include <cstring>

// c[q * Ntrajs + s] = packed element t of the Gram block of readout s and kz
// plane z (q = z + Nkz t); phi2 is [Ntrajs, Nkz, L, L] column-major
// (src/cuda/mr_kernels.cu:307-333 packing).
__global__ void mr_k_pack_strengths(const float2 *phi2, long long Ntrajs, int Nkz, int L,
                                    int upper, int T, double2 *c)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)T * Nkz * Ntrajs;
    if (id >= total) return;
    const long long s = id % Ntrajs;
    const long long q = id / Ntrajs;
    const int z = (int)(q % Nkz);
    const int t = (int)(q / Nkz);
    // packed offset t -> (i, j): walk columns
    int i = 0, j = 0, acc = 0;
    for (j = 0; j < L; ++j) {
        const int len = upper ? j + 1 : L - j;
        if (t < acc + len) {
            i = upper ? (t - acc) : j + (t - acc);
            break;
        }
        acc += len;
    }
    const float2 v = phi2[s + Ntrajs * (z + (long long)Nkz * (i + (long long)L * j))];
    c[id] = make_double2((double)v.x, (double)v.y);
}

// nat[j + Ptr*q] = fine[(delta mod nf) + nfine*q] * dc0[jx] dc1[jy] dc2[jz]
__global__ void mr_k_deconv(const double2 *fine, mr_grid g, long long Ptr, int Qc,
                            const double *dc0, const double *dc1, const double *dc2, double2 *nat)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Ptr * Qc) return;
    const long long q = id / Ptr;
    const long long j = id - q * Ptr;
    const int jx = (int)(j % g.M[0]);
    const int jy = (int)((j / g.M[0]) % g.M[1]);
    const int jz = (int)(j / ((long long)g.M[0] * g.M[1]));
    int f[3];
    const int jj[3] = {jx, jy, jz};
    for (int a = 0; a < 3; ++a) {
        const int M = g.M[a];
        int dlt = (jj[a] < M - M / 2) ? jj[a] : jj[a] - M;
        if (dlt < 0) dlt += g.nf[a];
        f[a] = dlt;
    }
    const long long fl = f[0] + (long long)g.nf[0] * (f[1] + (long long)g.nf[1] * f[2]);
    const double s = dc0[jx] * dc1[jy] * dc2[jz];
    const double2 v = fine[fl + g.nfine * q];
    nat[id] = make_double2(v.x * s, v.y * s);
}

// centered psf [Ptr, Qc] (float) -> natural-order double
__global__ void mr_k_psf_to_nat(const float2 *psf, int m0, int m1, int m2, int s0, int s1, int s2,
                                long long Ptr, int Qc, double2 *nat)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Ptr * Qc) return;
    const long long q = id / Ptr;
    const long long j = id - q * Ptr;
    int k0 = (int)(j % m0) + s0; if (k0 >= m0) k0 -= m0;
    int k1 = (int)((j / m0) % m1) + s1; if (k1 >= m1) k1 -= m1;
    int k2 = (int)(j / ((long long)m0 * m1)) + s2; if (k2 >= m2) k2 -= m2;
    const float2 v = psf[k0 + (long long)m0 * (k1 + (long long)m1 * k2) + Ptr * q];
    nat[id] = make_double2((double)v.x, (double)v.y);
}

// Natural -> centered, scale, FP32 store of components q0 .. q0+Qc-1.
// Unmasked: out[k + Ptr*q]. Masked: out[row + nm*t] for row = inv[k + Ptr*z].
__global__ void mr_k_toep_store(const double2 *nat, int m0, int m1, int m2, int s0, int s1, int s2,
                                long long Ptr, int Qc, long long q0, int Nkz, double scale,
                                const int *inv, long long nm, float2 *out)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Ptr * Qc) return;
    const long long qq = id / Ptr;
    const long long k = id - qq * Ptr;
    int j0 = (int)(k % m0) - s0; if (j0 < 0) j0 += m0;
    int j1 = (int)((k / m0) % m1) - s1; if (j1 < 0) j1 += m1;
    int j2 = (int)(k / ((long long)m0 * m1)) - s2; if (j2 < 0) j2 += m2;
    const double2 v = nat[j0 + (long long)m0 * (j1 + (long long)m1 * j2) + Ptr * qq];
    const float2 o = make_float2((float)(v.x * scale), (float)(v.y * scale));
    const long long q = q0 + qq;
    if (!inv) {
        out[k + Ptr * q] = o;
        return;
    }
    const long long z = q % Nkz, t = q / Nkz;
    const int row = inv[k + Ptr * z];
    if (row >= 0) out[row + nm * t] = o;
}

// inv[centered linear index] = mask row, -1 elsewhere
__global__ void mr_k_mask_rows(int *inv, long long P, const int64_t *mask, long long nm)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm) return;
    const long long k = mask[t];
    if (k >= 0 && k < P) inv[k] = (int)t;
}
__global__ void mr_k_fill_i32(int *p, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

namespace {

struct toep_geom {
    int d;        // transformed axes
    int M[3];     // transformed grid (1 beyond d)
    int shift[3]; // centering shift per transformed axis
    long long Ptr;
    int Nkz;
    double scale;
};

toep_geom mr_toep_geom(const mr_config *cfg)
{
    toep_geom g;
    const bool three = cfg->traj_type == MR_TRAJ_3D;
    g.d = three ? 3 : 2;
    g.M[0] = cfg->phi2_dim[0];
    g.M[1] = cfg->phi2_dim[1];
    g.M[2] = three ? cfg->phi2_dim[2] : 1;
    for (int a = 0; a < 3; ++a) g.shift[a] = g.M[a] / 2;
    g.Ptr = (long long)g.M[0] * g.M[1] * g.M[2];
    g.Nkz = (cfg->traj_type == MR_TRAJ_STACK) ? cfg->phi2_dim[2] : 1;
    g.scale = 1.0 / (double)g.Ptr;
    return g;
}

// Inverse FFT over the transformed M grid of Qc natural-order volumes.
int mr_toep_ifft(const toep_geom &tg, double2 *nat, int Qc, hipStream_t s)
{
    mr_fft_len<double> f[3];
    for (int a = 0; a < tg.d; ++a) MR_TRY(mr_fft_get<double>(tg.M[a], &f[a]));
    const long long m0 = tg.M[0], m1 = tg.M[1], m2 = tg.M[2];
    MR_TRY((mr_fft_pass<double, true>(f[0], mr_ld_axis<double2, true>{nat, 1, (int)m0},
                                      mr_st_axis<double2, true>{nat, 1, (int)m0},
                                      m1 * m2 * Qc, 0, 4, s, "toep ifft x")));
    MR_TRY((mr_fft_pass<double, true>(f[1], mr_ld_axis<double2, false>{nat, m0, (int)m1},
                                      mr_st_axis<double2, false>{nat, m0, (int)m1},
                                      m0 * m2 * Qc, 0, 16, s, "toep ifft y")));
    if (tg.d == 3)
        MR_TRY((mr_fft_pass<double, true>(f[2], mr_ld_axis<double2, false>{nat, m0 * m1, (int)m2},
                                          mr_st_axis<double2, false>{nat, m0 * m1, (int)m2},
                                          m0 * m1 * Qc, 0, 16, s, "toep ifft z")));
    return MR_SUCCESS;
}

struct toep_bufs {
    double2 *c = nullptr, *fine = nullptr, *nat = nullptr;
    double *dc[3] = {nullptr, nullptr, nullptr};
    int *inv = nullptr;
    ~toep_bufs()
    {
        if (c) (void)hipFree(c);
        if (fine) (void)hipFree(fine);
        if (nat) (void)hipFree(nat);
        for (double *p : dc) if (p) (void)hipFree(p);
        if (inv) (void)hipFree(inv);
    }
};

int mr_build_mask_rows(const mr_config *cfg, const int64_t *mask, hipStream_t s, int **inv)
{
    const long long P = mr_pad_size(cfg), nm = cfg->num_masked;
    MR_HIP_CHECK(hipMalloc((void **)inv, P * sizeof(int)));
    hipLaunchKernelGGL(mr_k_fill_i32, dim3((unsigned)mr_ceil_div(P, 256)), dim3(256), 0, s, *inv,
                       P, -1);
    MR_LAUNCH_CHECK("mr_k_fill_i32");
    hipLaunchKernelGGL(mr_k_mask_rows, dim3((unsigned)mr_ceil_div(nm, 256)), dim3(256), 0, s,
                       *inv, P, mask, nm);
    MR_LAUNCH_CHECK("mr_k_mask_rows");
    return MR_SUCCESS;
}

}  // namespace

int mr_calc_toep_impl(const mr_config *cfg, const mr_complex *d_phi2, const float *d_traj,
                      int64_t Ntrajs, int64_t Nkx, const int64_t *d_mask, mr_complex *d_out)
{
    hipStream_t s = (hipStream_t)cfg->stream;
    const toep_geom tg = mr_toep_geom(cfg);
    const int T = (int)mr_num_tri(cfg->L);
    const long long Q = (long long)T * tg.Nkz;
    const int ndim = (cfg->traj_type == MR_TRAJ_3D) ? 3 : 2;

    mr_grid g;
    mr_grid_setup(&g, tg.d, tg.M, cfg->eps);
    toep_bufs b;
    mr_grid_samples smp;
    MR_TRY(mr_grid_samples_build(g, d_traj, Ntrajs, Nkx, ndim, s, &smp));
    struct samples_guard {
        mr_grid_samples *p;
        ~samples_guard() { mr_grid_samples_free(p); }
    } guard{&smp};

    MR_HIP_CHECK(hipMalloc((void **)&b.c, Q * Ntrajs * sizeof(double2)));
    hipLaunchKernelGGL(mr_k_pack_strengths, dim3((unsigned)mr_ceil_div(Q * Ntrajs, 256)),
                       dim3(256), 0, s, reinterpret_cast<const float2 *>(d_phi2), (long long)Ntrajs,
                       tg.Nkz, cfg->L, cfg->triangulation == MR_TRI_UPPER ? 1 : 0, T, b.c);
    MR_LAUNCH_CHECK("mr_k_pack_strengths");
    for (int a = 0; a < 3; ++a) {
        std::vector<double> dc = mr_grid_deconv(g, a);
        MR_HIP_CHECK(hipMalloc((void **)&b.dc[a], dc.size() * sizeof(double)));
        // stream-ordered, not on the legacy stream (see mr_fft_get)
        MR_HIP_CHECK(hipMemcpyAsync(b.dc[a], dc.data(), dc.size() * sizeof(double),
                                    hipMemcpyHostToDevice, s));
        MR_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (cfg->is_masked) MR_TRY(mr_build_mask_rows(cfg, d_mask, s, &b.inv));

    // Chunk of components: fine grids + natural-order grids within budget.
    size_t budget = 0;
    MR_TRY(mr_mem_budget(cfg, &budget));
    const size_t per_q = ((size_t)g.nfine + (size_t)tg.Ptr) * sizeof(double2);
    long long Qc = (long long)(budget / 2 / per_q);
    const long long cap = mr_env_ll("MR_AHA_TOEP_MAX_CHUNK", 0);
    if (cap > 0 && Qc > cap) Qc = cap;
    if (Qc > Q) Qc = Q;
    if (Qc > 65535LL * 16) Qc = 65535LL * 16;
    if (Qc < 1) Qc = 1;
    hipError_t e = hipMalloc((void **)&b.fine, (size_t)g.nfine * Qc * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc((void **)&b.nat, (size_t)tg.Ptr * Qc * sizeof(double2));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        MR_FAIL(MR_ERR_ALLOC,
                "failed to allocate the calc_toep chunk buffers (chunk = %lld); lower "
                "mr_config.mem_fraction or the upsampled grid size",
                Qc);
    }
    for (long long q0 = 0; q0 < Q; q0 += Qc) {
        const int qn = (int)((Q - q0 < Qc) ? Q - q0 : Qc);
        MR_TRY(mr_grid_spread(g, smp, b.c + q0 * Ntrajs, Ntrajs, qn, b.fine, s));
        MR_TRY(mr_grid_fft(g, b.fine, nullptr, qn, false, s));
        MR_GRID_CHECK(mr_ceil_div(tg.Ptr * qn, 256), 256, "mr_k_deconv");
        hipLaunchKernelGGL(mr_k_deconv, dim3((unsigned)mr_ceil_div(tg.Ptr * qn, 256)), dim3(256),
                           0, s, b.fine, g, tg.Ptr, qn, b.dc[0], b.dc[1], b.dc[2], b.nat);
        MR_LAUNCH_CHECK("mr_k_deconv");
        MR_TRY(mr_toep_ifft(tg, b.nat, qn, s));
        hipLaunchKernelGGL(mr_k_toep_store, dim3((unsigned)mr_ceil_div(tg.Ptr * qn, 256)),
                           dim3(256), 0, s, b.nat, tg.M[0], tg.M[1], tg.M[2], tg.shift[0],
                           tg.shift[1], tg.shift[2], tg.Ptr, qn, q0, tg.Nkz, tg.scale, b.inv,
                           (long long)cfg->num_masked, reinterpret_cast<float2 *>(d_out));
        MR_LAUNCH_CHECK("mr_k_toep_store");
        if (cfg->verbose)
            printf("[mr_aha] calc_toep: components %lld-%lld of %lld issued\n", q0, q0 + qn - 1, Q);
    }
    MR_HIP_CHECK(hipStreamSynchronize(s));
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_calc_toep(const mr_config *cfg, const mr_complex *d_phi2,
                                      const float *d_traj, int64_t Ntrajs, int64_t Nkx,
                                      const int64_t *d_mask, mr_complex *d_phi2_toep)
{
    mr_clear_error();
    MR_REQUIRE(cfg != nullptr, "mr_calc_toep: cfg is NULL");
    MR_REQUIRE(d_phi2 != nullptr, "mr_calc_toep: d_phi2 is NULL");
    MR_REQUIRE(d_traj != nullptr, "mr_calc_toep: d_traj is NULL");
    MR_REQUIRE(d_phi2_toep != nullptr, "mr_calc_toep: d_phi2_toep is NULL");
    MR_REQUIRE(Ntrajs > 0, "mr_calc_toep: Ntrajs = %lld must be > 0", (long long)Ntrajs);
    MR_REQUIRE(Nkx > 0, "mr_calc_toep: Nkx = %lld must be > 0", (long long)Nkx);
    if (cfg->is_masked)
        MR_REQUIRE(d_mask != nullptr, "mr_calc_toep: d_mask is required when is_masked is set");
    MR_TRY(mr_config_validate(cfg));
    int prev = -1;
    int rc = mr_push_device(cfg, &prev);
    if (rc == MR_SUCCESS)
        rc = mr_calc_toep_impl(cfg, d_phi2, d_traj, Ntrajs, Nkx, d_mask, d_phi2_toep);
    mr_pop_device(prev);
    return rc;
}

static int mr_calc_toep_from_psf_impl(const mr_config *cfg, const mr_complex *d_psf, int64_t t0,
                                      int64_t n_comp, const int64_t *d_mask, mr_complex *d_out)
{
    hipStream_t s = (hipStream_t)cfg->stream;
    const toep_geom tg = mr_toep_geom(cfg);
    const long long Qc = n_comp * tg.Nkz;
    toep_bufs b;
    if (cfg->is_masked) MR_TRY(mr_build_mask_rows(cfg, d_mask, s, &b.inv));
    if (hipMalloc((void **)&b.nat, (size_t)tg.Ptr * Qc * sizeof(double2)) != hipSuccess) {
        (void)hipGetLastError();
        MR_FAIL(MR_ERR_ALLOC, "failed to allocate %lld scratch elements for mr_calc_toep_from_psf",
                tg.Ptr * Qc);
    }
    MR_GRID_CHECK(mr_ceil_div(tg.Ptr * Qc, 256), 256, "mr_k_psf_to_nat");
    hipLaunchKernelGGL(mr_k_psf_to_nat, dim3((unsigned)mr_ceil_div(tg.Ptr * Qc, 256)), dim3(256),
                       0, s, reinterpret_cast<const float2 *>(d_psf), tg.M[0], tg.M[1], tg.M[2],
                       tg.shift[0], tg.shift[1], tg.shift[2], tg.Ptr, (int)Qc, b.nat);
    MR_LAUNCH_CHECK("mr_k_psf_to_nat");
    MR_TRY(mr_toep_ifft(tg, b.nat, (int)Qc, s));
    hipLaunchKernelGGL(mr_k_toep_store, dim3((unsigned)mr_ceil_div(tg.Ptr * Qc, 256)), dim3(256),
                       0, s, b.nat, tg.M[0], tg.M[1], tg.M[2], tg.shift[0], tg.shift[1],
                       tg.shift[2], tg.Ptr, (int)Qc, (long long)t0 * tg.Nkz, tg.Nkz, tg.scale,
                       b.inv, (long long)cfg->num_masked, reinterpret_cast<float2 *>(d_out));
    MR_LAUNCH_CHECK("mr_k_toep_store");
    MR_HIP_CHECK(hipStreamSynchronize(s));
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_calc_toep_from_psf(const mr_config *cfg, const mr_complex *d_psf,
                                               int64_t t0, int64_t n_comp, const int64_t *d_mask,
                                               mr_complex *d_phi2_toep)
{
    mr_clear_error();
    MR_REQUIRE(cfg != nullptr, "mr_calc_toep_from_psf: cfg is NULL");
    MR_REQUIRE(d_psf != nullptr, "mr_calc_toep_from_psf: d_psf is NULL");
    MR_REQUIRE(d_phi2_toep != nullptr, "mr_calc_toep_from_psf: d_phi2_toep is NULL");
    MR_TRY(mr_config_validate(cfg));
    const int64_t ntri = mr_num_tri(cfg->L);
    MR_REQUIRE(t0 >= 0 && n_comp >= 1 && t0 + n_comp <= ntri,
               "mr_calc_toep_from_psf: components [%lld, %lld) outside [0, %lld)", (long long)t0,
               (long long)(t0 + n_comp), (long long)ntri);
    if (cfg->is_masked)
        MR_REQUIRE(d_mask != nullptr,
                   "mr_calc_toep_from_psf: d_mask is required when is_masked is set");
    int prev = -1;
    int rc = mr_push_device(cfg, &prev);
    if (rc == MR_SUCCESS) rc = mr_calc_toep_from_psf_impl(cfg, d_psf, t0, n_comp, d_mask, d_phi2_toep);
    mr_pop_device(prev);
    return rc;
}
