// method = MR_OP_NUFFT engine: per coil a type-2 NUFFT (sign -1), the
// per-readout Gram multiply k'_l = sum_j k_j B(j, l), and a type-1 NUFFT
// (sign +1) back (reference: src/cuda/mr_nufft_op.cu:613-700,
// python/mr_aha/nufft.py:473-552). For a stack the image is modulated by
// zmod(z) = exp(2 pi i floor(Nz/2) z / Nz), transformed along z with the
// unnormalised DFT pair, and kz plane K uses the Gram blocks of plane K.
//
// MI355X realisation: FP64 ES gridding of mr_grid.hip (the same sorted
// sample table serves the type-2 interpolation and the type-1 gather
// spreading), FP64 fine-grid FFTs of the project FFT engine, all L (x Nkz)
// transforms of a coil batched in one launch per stage.

#include "mr_grid.hpp"

#include <new>
#include <vector>

struct mr_nufft_eng {
    mr_config cfg;
    hipStream_t s;
    mr_grid g;
    mr_grid_samples smp;
    int d;          // NUFFT dimensions
    int Nkz;        // kz planes (stack) or 1
    int Q;          // transforms per coil: L * Nkz
    long long Nimg; // image voxels
    long long Ntrajs, Nkx, Mnu;
    float2 *phi2;   // Gram blocks [Ntrajs, Nkz, L, L]
    double *dc[3];
    double2 *img;   // [Nimg, L] (z fastest after x, y): image-side work
    double2 *acc;   // [Nimg, L] coil sum
    double2 *fine;  // [Q, nfine]
    double2 *k1;    // [Q, Mnu]
    double2 *k2;    // [Q, Mnu]
    mr_fft_len<double> fz;
};

static void mr_nufft_free(mr_nufft_eng *e)
{
    mr_grid_samples_free(&e->smp);
    for (double *p : e->dc) if (p) (void)hipFree(p);
    void *bufs[] = {e->phi2, e->img, e->acc, e->fine, e->k1, e->k2};
    for (void *p : bufs) if (p) (void)hipFree(p);
    delete e;
}

void mr_nufft_eng_destroy(mr_nufft_eng *e)
{
    if (e) mr_nufft_free(e);
}

int mr_nufft_eng_create(const mr_config *cfg, const mr_complex *d_phi2, const float *d_traj,
                        int64_t Ntrajs, int64_t Nkx, double upsampfac, mr_nufft_eng **out)
{
    *out = nullptr;
    if (upsampfac != 0.0 && upsampfac != 2.0)
        MR_FAIL(MR_ERR_UNSUPPORTED, "mr_op_create: nufft_upsampfac = %g; this build grids with 2.0",
                upsampfac);
    mr_nufft_eng *e = new (std::nothrow) mr_nufft_eng();
    if (!e) MR_FAIL(MR_ERR_ALLOC, "failed to allocate the NUFFT engine");
    e->cfg = *cfg;
    e->s = (hipStream_t)cfg->stream;
    const bool three = cfg->traj_type == MR_TRAJ_3D;
    e->d = three ? 3 : 2;
    e->Nkz = (cfg->traj_type == MR_TRAJ_STACK) ? cfg->n_dim[2] : 1;
    e->Q = cfg->L * e->Nkz;
    e->Nimg = mr_image_size(cfg);
    e->Ntrajs = Ntrajs;
    e->Nkx = Nkx;
    e->Mnu = Ntrajs * Nkx;
    int modes[3] = {cfg->n_dim[0], cfg->n_dim[1], three ? cfg->n_dim[2] : 1};
    mr_grid_setup(&e->g, e->d, modes, cfg->eps);
    int rc = mr_grid_samples_build(e->g, d_traj, Ntrajs, Nkx, e->d, e->s, &e->smp);
    if (rc != MR_SUCCESS) { mr_nufft_free(e); return rc; }
    const size_t phi2_bytes = (size_t)Ntrajs * e->Nkz * cfg->L * cfg->L * sizeof(float2);
    const size_t img_bytes = (size_t)e->Nimg * cfg->L * sizeof(double2);
    hipError_t he = hipMalloc((void **)&e->phi2, phi2_bytes);
    if (he == hipSuccess) he = hipMalloc((void **)&e->img, img_bytes);
    if (he == hipSuccess) he = hipMalloc((void **)&e->acc, img_bytes);
    if (he == hipSuccess) he = hipMalloc((void **)&e->fine, (size_t)e->Q * e->g.nfine * sizeof(double2));
    if (he == hipSuccess) he = hipMalloc((void **)&e->k1, (size_t)e->Q * e->Mnu * sizeof(double2));
    if (he == hipSuccess) he = hipMalloc((void **)&e->k2, (size_t)e->Q * e->Mnu * sizeof(double2));
    for (int a = 0; a < 3 && he == hipSuccess; ++a) {
        std::vector<double> dc = mr_grid_deconv(e->g, a);
        he = hipMalloc((void **)&e->dc[a], dc.size() * sizeof(double));
        // stream-ordered (never the legacy stream, which could tie this call
        // to another thread's graph capture); synchronised below
        if (he == hipSuccess)
            he = hipMemcpyAsync(e->dc[a], dc.data(), dc.size() * sizeof(double),
                                hipMemcpyHostToDevice, e->s);
        if (he == hipSuccess) he = hipStreamSynchronize(e->s);
    }
    if (he == hipSuccess)
        he = hipMemcpyAsync(e->phi2, d_phi2, phi2_bytes, hipMemcpyDeviceToDevice, e->s);
    if (he == hipSuccess) he = hipStreamSynchronize(e->s);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        mr_nufft_free(e);
        MR_FAIL(MR_ERR_ALLOC, "failed to allocate the NUFFT engine buffers: %s", hipGetErrorString(he));
    }
    if (e->Nkz > 1) {
        rc = mr_fft_get<double>(e->Nkz, &e->fz);
        if (rc != MR_SUCCESS) { mr_nufft_free(e); return rc; }
    }
    *out = e;
    return MR_SUCCESS;
}

// img[i + Nimg*l] = S_c[i] x_l[i] zmod(z)
__global__ void mr_k_nu_img(const float2 *x, const float2 *S, long long Nimg, long long NxNy, int L,
                            int Nz, int stack, double2 *img)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Nimg * L) return;
    const long long i = id % Nimg;
    const float2 xv = x[id], sv = S[i];
    double re = (double)sv.x * xv.x - (double)sv.y * xv.y;
    double im = (double)sv.x * xv.y + (double)sv.y * xv.x;
    if (stack) {
        const long long z = i / NxNy;
        double sn, cs;
        sincospi(2.0 * (double)((Nz / 2) * z % Nz) / Nz, &sn, &cs);
        const double r2 = re * cs - im * sn, i2 = re * sn + im * cs;
        re = r2;
        im = i2;
    }
    img[id] = make_double2(re, im);
}

// Type-2 pre-correction: fine[q][g] = img(r) dc(r) at the fine point of
// centered image coordinate r, zero elsewhere. The 2D planes of a stack are
// transforms q = K + Nkz * l of img viewed as [NxNy, Nz * L].
__global__ void mr_k_nu_pad(const double2 *img, mr_grid g, long long Nplane, int Q, const double *dc0,
                            const double *dc1, const double *dc2, double2 *fine)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.nfine * Q) return;
    const long long q = id / g.nfine;
    const long long gl = id - q * g.nfine;
    int gg[3];
    gg[0] = (int)(gl % g.nf[0]);
    gg[1] = (int)((gl / g.nf[0]) % g.nf[1]);
    gg[2] = (int)(gl / ((long long)g.nf[0] * g.nf[1]));
    long long src = 0, mul = 1;
    double w = 1.0;
    for (int a = 0; a < 3; ++a) {
        const int N = g.M[a], nf = g.nf[a];
        const int r = (gg[a] <= nf / 2) ? gg[a] : gg[a] - nf;  // signed fine frequency
        if (r < -(N / 2) || r > N - N / 2 - 1) {
            fine[id] = make_double2(0.0, 0.0);
            return;
        }
        const int iimg = r + N / 2;       // image index
        int jn = r; if (jn < 0) jn += N;  // natural mode index
        const double *dc = (a == 0) ? dc0 : (a == 1) ? dc1 : dc2;
        w *= dc[jn];
        src += mul * iimg;
        mul *= N;
    }
    const double2 v = img[src + Nplane * q];
    fine[id] = make_double2(v.x * w, v.y * w);
}

// Type-1 post-correction: img(r) = fine[q][f(r)] dc(r).
__global__ void mr_k_nu_crop(const double2 *fine, mr_grid g, long long Nplane, int Q, const double *dc0,
                             const double *dc1, const double *dc2, double2 *img)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Nplane * Q) return;
    const long long q = id / Nplane;
    const long long i = id - q * Nplane;
    int ii[3];
    ii[0] = (int)(i % g.M[0]);
    ii[1] = (int)((i / g.M[0]) % g.M[1]);
    ii[2] = (int)(i / ((long long)g.M[0] * g.M[1]));
    long long fl = 0, mul = 1;
    double w = 1.0;
    for (int a = 0; a < 3; ++a) {
        const int N = g.M[a], nf = g.nf[a];
        const int r = ii[a] - N / 2;
        int jn = r; if (jn < 0) jn += N;
        int f = r; if (f < 0) f += nf;
        const double *dc = (a == 0) ? dc0 : (a == 1) ? dc1 : dc2;
        w *= dc[jn];
        fl += mul * f;
        mul *= nf;
    }
    const double2 v = fine[fl + g.nfine * q];
    img[id] = make_double2(v.x * w, v.y * w);
}

// k2[(K + Nkz l), m] = sum_j k1[(K + Nkz j), m] B_{s(m),K}(j, l)
__global__ void mr_k_nu_gram(const double2 *k1, const float2 *phi2, long long Mnu, long long Ntrajs,
                             int Nkz, int L, double2 *k2)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Mnu * Nkz) return;
    const long long m = id % Mnu;
    const int K = (int)(id / Mnu);
    const long long s = m % Ntrajs;
    const float2 *B = phi2 + s + Ntrajs * (long long)K;  // entry (j,l) at + Ntrajs*Nkz*(j + L*l)
    const long long bs = Ntrajs * Nkz;
    for (int l = 0; l < L; ++l) {
        double re = 0.0, im = 0.0;
        for (int j = 0; j < L; ++j) {
            const double2 v = k1[m + Mnu * (K + (long long)Nkz * j)];
            const float2 b = B[bs * (j + (long long)L * l)];
            re += v.x * b.x - v.y * b.y;
            im += v.x * b.y + v.y * b.x;
        }
        k2[m + Mnu * (K + (long long)Nkz * l)] = make_double2(re, im);
    }
}

// acc (+)= conj(S_c zmod) img ; final pass writes float out
__global__ void mr_k_nu_acc(const double2 *img, const float2 *S, long long Nimg, long long NxNy, int L,
                            int Nz, int stack, int first, double2 *acc)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Nimg * L) return;
    const long long i = id % Nimg;
    const float2 sv = S[i];
    double sr = sv.x, si = sv.y;
    if (stack) {
        const long long z = i / NxNy;
        double sn, cs;
        sincospi(2.0 * (double)((Nz / 2) * z % Nz) / Nz, &sn, &cs);
        const double r2 = sr * cs - si * sn, i2 = sr * sn + si * cs;
        sr = r2;
        si = i2;
    }
    const double2 v = img[id];
    const double re = sr * v.x + si * v.y, im = sr * v.y - si * v.x;
    const double2 a = first ? make_double2(0.0, 0.0) : acc[id];
    acc[id] = make_double2(a.x + re, a.y + im);
}

__global__ void mr_k_nu_out(const double2 *acc, long long n, float2 *out)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = make_float2((float)acc[id].x, (float)acc[id].y);
}

static int mr_nu_zfft(mr_nufft_eng *e, bool inverse)
{
    const long long NxNy = (long long)e->cfg.n_dim[0] * e->cfg.n_dim[1];
    const int Nz = e->Nkz;
    mr_ld_axis<double2, false> ld{e->img, NxNy, Nz};
    mr_st_axis<double2, false> st{e->img, NxNy, Nz};
    if (inverse)
        return mr_fft_pass<double, true>(e->fz, ld, st, NxNy * e->cfg.L, 0, 16, e->s, "nufft z");
    return mr_fft_pass<double, false>(e->fz, ld, st, NxNy * e->cfg.L, 0, 16, e->s, "nufft z");
}

int mr_nufft_eng_execute(mr_nufft_eng *e, const mr_complex *d_x, const mr_complex *d_SEs,
                         mr_complex *d_out)
{
    const mr_config &c = e->cfg;
    const int L = c.L;
    const bool stack = e->Nkz > 1 || c.traj_type == MR_TRAJ_STACK;
    const long long NxNy = (long long)c.n_dim[0] * c.n_dim[1];
    const long long Nplane = stack ? NxNy : e->Nimg;
    const auto *x = reinterpret_cast<const float2 *>(d_x);
    const auto *S = reinterpret_cast<const float2 *>(d_SEs);
    const unsigned bimg = (unsigned)mr_ceil_div(e->Nimg * L, 256);
    for (int coil = 0; coil < c.n_coils; ++coil) {
        const float2 *Sc = S + e->Nimg * coil;
        hipLaunchKernelGGL(mr_k_nu_img, dim3(bimg), dim3(256), 0, e->s, x, Sc, e->Nimg, NxNy, L,
                           c.n_dim[2], stack ? 1 : 0, e->img);
        MR_LAUNCH_CHECK("mr_k_nu_img");
        if (e->Nkz > 1) MR_TRY(mr_nu_zfft(e, false));
        hipLaunchKernelGGL(mr_k_nu_pad, dim3((unsigned)mr_ceil_div(e->g.nfine * e->Q, 256)), dim3(256),
                           0, e->s, e->img, e->g, Nplane, e->Q, e->dc[0], e->dc[1], e->dc[2], e->fine);
        MR_LAUNCH_CHECK("mr_k_nu_pad");
        MR_TRY(mr_grid_fft(e->g, e->fine, nullptr, e->Q, false, e->s));
        MR_TRY(mr_grid_interp(e->g, e->smp, e->fine, e->Q, e->k1, e->s));
        hipLaunchKernelGGL(mr_k_nu_gram, dim3((unsigned)mr_ceil_div(e->Mnu * e->Nkz, 256)), dim3(256),
                           0, e->s, e->k1, e->phi2, e->Mnu, e->Ntrajs, e->Nkz, L, e->k2);
        MR_LAUNCH_CHECK("mr_k_nu_gram");
        MR_TRY(mr_grid_spread_samples(e->g, e->smp, e->k2, e->Q, e->fine, e->s));
        MR_TRY(mr_grid_fft(e->g, e->fine, nullptr, e->Q, true, e->s));
        hipLaunchKernelGGL(mr_k_nu_crop, dim3((unsigned)mr_ceil_div(Nplane * e->Q, 256)), dim3(256), 0,
                           e->s, e->fine, e->g, Nplane, e->Q, e->dc[0], e->dc[1], e->dc[2], e->img);
        MR_LAUNCH_CHECK("mr_k_nu_crop");
        if (e->Nkz > 1) MR_TRY(mr_nu_zfft(e, true));
        hipLaunchKernelGGL(mr_k_nu_acc, dim3(bimg), dim3(256), 0, e->s, e->img, Sc, e->Nimg, NxNy, L,
                           c.n_dim[2], stack ? 1 : 0, coil == 0 ? 1 : 0, e->acc);
        MR_LAUNCH_CHECK("mr_k_nu_acc");
    }
    hipLaunchKernelGGL(mr_k_nu_out, dim3(bimg), dim3(256), 0, e->s, e->acc, e->Nimg * L,
                       reinterpret_cast<float2 *>(d_out));
    MR_LAUNCH_CHECK("mr_k_nu_out");
    return MR_SUCCESS;
}
