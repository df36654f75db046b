// ES-kernel gridding on gfx950 (see mr_grid.hpp for the conventions).

#include "mr_grid.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>

void mr_grid_setup(mr_grid *g, int d, const int *M, double eps)
{
    g->d = d;
    int w = (int)std::ceil(std::log10(1.0 / eps)) + 1;
    if (w < 2) w = 2;
    if (w > MR_GRID_MAXW) w = MR_GRID_MAXW;
    g->w = w;
    g->beta = 2.30 * w;
    g->nfine = 1;
    g->nbins = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < d) {
            const int bs = (d == 2) ? 8 : 4;
            int need = 2 * M[a];
            if (need < 2 * w) need = 2 * w;
            const int nf = bs * mr_fft_next_smooth((need + bs - 1) / bs);
            g->M[a] = M[a];
            g->nf[a] = nf;
            g->bs[a] = bs;
            g->nb[a] = nf / bs;
        } else {
            g->M[a] = 1;
            g->nf[a] = 1;
            g->bs[a] = 1;
            g->nb[a] = 1;
        }
        g->nfine *= g->nf[a];
        g->nbins *= g->nb[a];
    }
}

// Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_n.
static void mr_gauss_legendre(int n, std::vector<double> &x, std::vector<double> &wq)
{
    x.resize(n);
    wq.resize(n);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        wq[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

std::vector<double> mr_grid_deconv(const mr_grid &g, int a)
{
    const int M = g.M[a];
    std::vector<double> out(M, 1.0);
    if (a >= g.d) return out;
    std::vector<double> zq, wq;
    mr_gauss_legendre(2 * g.w + 60, zq, wq);
    const int nf = g.nf[a];
    const int half = M / 2;
    for (int j = 0; j < M; ++j) {
        const int delta = (j < M - half) ? j : j - M;
        double s = 0.0;
        for (size_t i = 0; i < zq.size(); ++i) {
            const double phi = std::exp(g.beta * (std::sqrt(1.0 - zq[i] * zq[i]) - 1.0));
            s += wq[i] * phi * std::cos(delta * M_PI * g.w * zq[i] / nf);
        }
        out[j] = 1.0 / (0.5 * g.w * s);
    }
    return out;
}

__device__ __forceinline__ double mr_es(double z, double beta)
{
    const double t = 1.0 - z * z;
    return (t > 0.0) ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

__device__ __forceinline__ int mr_first_point(float x, int nf, int w, double *u_out)
{
    const double u = (double)x * (double)nf * (0.5 / M_PI);
    *u_out = u;
    return (int)ceil(u - 0.5 * w);
}

__global__ void mr_k_grid_keys(const float *traj, long long n, long long Ntrajs, long long Nkx,
                               mr_grid g, unsigned long long *keys, int *vals)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    long long bin = 0, mul = 1;
    for (int a = 0; a < g.d; ++a) {
        double u;
        int i0 = mr_first_point(traj[m + n * a], g.nf[a], g.w, &u);
        i0 %= g.nf[a];
        if (i0 < 0) i0 += g.nf[a];
        bin += mul * (i0 / g.bs[a]);
        mul *= g.nb[a];
    }
    const long long s = m % Ntrajs, r = m / Ntrajs;
    keys[m] = (unsigned long long)bin * (unsigned long long)n + (unsigned long long)(s * Nkx + r);
    vals[m] = (int)m;
}

__global__ void mr_k_grid_fill(const float *traj, long long n, long long Ntrajs, mr_grid g,
                               const unsigned long long *keys, const int *perm, int *i0o, int *rd,
                               double *wt, int *bin_lo, int *bin_hi)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long m = perm[i];
    for (int a = 0; a < g.d; ++a) {
        double u;
        const int i0 = mr_first_point(traj[m + n * a], g.nf[a], g.w, &u);
        int im = i0 % g.nf[a];
        if (im < 0) im += g.nf[a];
        i0o[i * 3 + a] = im;
        for (int k = 0; k < g.w; ++k)
            wt[(i * g.d + a) * g.w + k] = mr_es(2.0 * ((double)(i0 + k) - u) / g.w, g.beta);
    }
    for (int a = g.d; a < 3; ++a) i0o[i * 3 + a] = 0;
    rd[i] = (int)(m % Ntrajs);
    const unsigned long long b = keys[i] / (unsigned long long)n;
    if (i == 0 || keys[i - 1] / (unsigned long long)n != b) bin_lo[b] = (int)i;
    if (i == n - 1 || keys[i + 1] / (unsigned long long)n != b) bin_hi[b] = (int)(i + 1);
}

void mr_grid_samples_free(mr_grid_samples *s)
{
    if (s->i0) (void)hipFree(s->i0);
    if (s->rd) (void)hipFree(s->rd);
    if (s->wt) (void)hipFree(s->wt);
    if (s->bin_lo) (void)hipFree(s->bin_lo);
    if (s->bin_hi) (void)hipFree(s->bin_hi);
    if (s->perm) (void)hipFree(s->perm);
    *s = mr_grid_samples();
}

int mr_grid_samples_build(const mr_grid &g, const float *traj, long long Ntrajs, long long Nkx,
                          int ndim, hipStream_t stream, mr_grid_samples *out)
{
    (void)ndim;
    const long long n = Ntrajs * Nkx;
    MR_REQUIRE(n > 0 && n < 0x7fffffffLL, "gridding: %lld samples outside [1, 2^31)", n);
    MR_REQUIRE(g.nbins < 0x7fffffffLL, "gridding: too many bins (%lld)", g.nbins);
    mr_grid_samples s;
    s.n = n;
    unsigned long long *keys = nullptr, *keys2 = nullptr;
    int *vals = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipSuccess;
#define MR_GA(ptr, bytes) if (e == hipSuccess) e = hipMalloc((void **)&(ptr), (bytes))
    MR_GA(keys, n * sizeof(unsigned long long));
    MR_GA(keys2, n * sizeof(unsigned long long));
    MR_GA(vals, n * sizeof(int));
    MR_GA(s.perm, n * sizeof(int));
    MR_GA(s.i0, n * 3 * sizeof(int));
    MR_GA(s.rd, n * sizeof(int));
    MR_GA(s.wt, (size_t)n * g.d * g.w * sizeof(double));
    MR_GA(s.bin_lo, g.nbins * sizeof(int));
    MR_GA(s.bin_hi, g.nbins * sizeof(int));
#undef MR_GA
    int rc = MR_SUCCESS;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        mr_set_error("gridding: failed to allocate the sample tables for %lld samples", n);
        rc = MR_ERR_ALLOC;
    }
    unsigned long long maxkey = (unsigned long long)g.nbins * (unsigned long long)n;
    int end_bit = 1;
    while (end_bit < 64 && (maxkey >> end_bit) != 0) ++end_bit;
    if (rc == MR_SUCCESS) {
        hipLaunchKernelGGL(mr_k_grid_keys, dim3((unsigned)mr_ceil_div(n, 256)), dim3(256), 0,
                           stream, traj, n, Ntrajs, Nkx, g, keys, vals);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys2, vals, s.perm,
                                                  (int)n, 0, end_bit, stream);
        if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1);
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, vals, s.perm,
                                                  (int)n, 0, end_bit, stream);
        if (e == hipSuccess) e = hipMemsetAsync(s.bin_lo, 0, g.nbins * sizeof(int), stream);
        if (e == hipSuccess) e = hipMemsetAsync(s.bin_hi, 0, g.nbins * sizeof(int), stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(mr_k_grid_fill, dim3((unsigned)mr_ceil_div(n, 256)), dim3(256), 0,
                               stream, traj, n, Ntrajs, g, keys2, s.perm, s.i0, s.rd, s.wt,
                               s.bin_lo, s.bin_hi);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            mr_set_error("gridding: sample sort failed: %s", hipGetErrorString(e));
            rc = MR_ERR_CUFINUFFT;
        }
    }
    if (tmp) (void)hipFree(tmp);
    if (keys) (void)hipFree(keys);
    if (keys2) (void)hipFree(keys2);
    if (vals) (void)hipFree(vals);
    if (rc != MR_SUCCESS) {
        mr_grid_samples_free(&s);
        return rc;
    }
    *out = s;
    return MR_SUCCESS;
}

// ---------------------------------------------------------------------------
// Type-1 gather spreading. One 64-lane workgroup per bin-sized tile of the
// fine grid (8x8 in 2D, 4x4x4 in 3D); lane = fine point. Samples of the
// candidate bins are staged in LDS 64 at a time and every lane adds its
// kernel weight. With readout-constant strengths (PER_SAMPLE false) the
// weights of consecutive samples of one readout are summed first and the
// CB strengths are applied once per readout run.
// ---------------------------------------------------------------------------

template <int D, int CB, bool PER_SAMPLE>
__global__ void __launch_bounds__(64)
mr_k_spread(mr_grid g, const int *__restrict__ i0s, const int *__restrict__ rds,
            const double *__restrict__ wts, const int *__restrict__ perm,
            const int *__restrict__ bin_lo, const int *__restrict__ bin_hi,
            const double2 *__restrict__ str, long long sstride, int Q, double2 *__restrict__ fine)
{
    __shared__ int sh_i0[64][3];
    __shared__ int sh_key[64];
    __shared__ double sh_w[64][D][MR_GRID_MAXW];
    const int lane = threadIdx.x;
    long long tile = blockIdx.x;
    int t[3], p[3];
    t[0] = (int)(tile % g.nb[0]);
    t[1] = (int)((tile / g.nb[0]) % g.nb[1]);
    t[2] = (int)(tile / ((long long)g.nb[0] * g.nb[1]));
    if (D == 2) { p[0] = lane & 7; p[1] = lane >> 3; p[2] = 0; }
    else { p[0] = lane & 3; p[1] = (lane >> 2) & 3; p[2] = lane >> 4; }
    int gp[3];
    for (int a = 0; a < 3; ++a) gp[a] = t[a] * g.bs[a] + p[a];
    const int q0 = blockIdx.y * CB;
    const int nq = (Q - q0 < CB) ? Q - q0 : CB;
    const int w = g.w;

    double2 acc[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) acc[q] = make_double2(0.0, 0.0);
    double wsum = 0.0;
    int cur = -1;

    int nbr[3];
    for (int a = 0; a < 3; ++a) nbr[a] = (a < D) ? (w - 1 + g.bs[a] - 1) / g.bs[a] : 0;
    for (int k2 = nbr[2]; k2 >= 0; --k2)
    for (int k1 = nbr[1]; k1 >= 0; --k1)
    for (int k0 = nbr[0]; k0 >= 0; --k0) {
        const int b0 = (t[0] - k0 + g.nb[0] * 4) % g.nb[0];
        const int b1 = (t[1] - k1 + g.nb[1] * 4) % g.nb[1];
        const int b2 = (t[2] - k2 + g.nb[2] * 4) % g.nb[2];
        const long long B = b0 + (long long)g.nb[0] * (b1 + (long long)g.nb[1] * b2);
        const int lo = bin_lo[B], hi = bin_hi[B];
        for (int base = lo; base < hi; base += 64) {
            const int cnt = (hi - base < 64) ? hi - base : 64;
            __syncthreads();
            if (lane < cnt) {
                const long long i = base + lane;
                for (int a = 0; a < 3; ++a) sh_i0[lane][a] = i0s[i * 3 + a];
                sh_key[lane] = PER_SAMPLE ? perm[i] : rds[i];
                for (int a = 0; a < D; ++a)
                    for (int k = 0; k < w; ++k) sh_w[lane][a][k] = wts[(i * D + a) * w + k];
            }
            __syncthreads();
            for (int j = 0; j < cnt; ++j) {
                double wv = 1.0;
                bool in = true;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    int o = gp[a] - sh_i0[j][a];
                    if (o < 0) o += g.nf[a];
                    in = in && (o < w);
                    wv *= in ? sh_w[j][a][in ? o : 0] : 0.0;
                }
                if (!in) wv = 0.0;
                const int key = sh_key[j];
                if (PER_SAMPLE) {
#pragma unroll
                    for (int q = 0; q < CB; ++q)
                        if (q < nq) {
                            const double2 c = str[(long long)(q0 + q) * sstride + key];
                            acc[q].x += wv * c.x;
                            acc[q].y += wv * c.y;
                        }
                } else {
                    if (key != cur) {
                        if (cur >= 0) {
#pragma unroll
                            for (int q = 0; q < CB; ++q)
                                if (q < nq) {
                                    const double2 c = str[(long long)(q0 + q) * sstride + cur];
                                    acc[q].x += wsum * c.x;
                                    acc[q].y += wsum * c.y;
                                }
                        }
                        cur = key;
                        wsum = 0.0;
                    }
                    wsum += wv;
                }
            }
        }
    }
    if (!PER_SAMPLE && cur >= 0) {
#pragma unroll
        for (int q = 0; q < CB; ++q)
            if (q < nq) {
                const double2 c = str[(long long)(q0 + q) * sstride + cur];
                acc[q].x += wsum * c.x;
                acc[q].y += wsum * c.y;
            }
    }
    if (gp[0] < g.nf[0] && gp[1] < g.nf[1] && gp[2] < g.nf[2]) {
        const long long gl = gp[0] + (long long)g.nf[0] * (gp[1] + (long long)g.nf[1] * gp[2]);
#pragma unroll
        for (int q = 0; q < CB; ++q)
            if (q < nq) fine[(long long)(q0 + q) * g.nfine + gl] = acc[q];
    }
}

template <bool PER_SAMPLE>
static int mr_spread_launch(const mr_grid &g, const mr_grid_samples &s, const double2 *str,
                            long long sstride, int Q, double2 *fine, hipStream_t stream)
{
    constexpr int CB = 16;
    const dim3 grid((unsigned)g.nbins, (unsigned)((Q + CB - 1) / CB));
    if (g.d == 2)
        hipLaunchKernelGGL((mr_k_spread<2, CB, PER_SAMPLE>), grid, dim3(64), 0, stream, g, s.i0,
                           s.rd, s.wt, s.perm, s.bin_lo, s.bin_hi, str, sstride, Q, fine);
    else
        hipLaunchKernelGGL((mr_k_spread<3, CB, PER_SAMPLE>), grid, dim3(64), 0, stream, g, s.i0,
                           s.rd, s.wt, s.perm, s.bin_lo, s.bin_hi, str, sstride, Q, fine);
    MR_LAUNCH_CHECK("mr_k_spread");
    return MR_SUCCESS;
}

int mr_grid_spread(const mr_grid &g, const mr_grid_samples &s, const double2 *str,
                   long long Ntrajs, int Q, double2 *fine, hipStream_t stream)
{
    return mr_spread_launch<false>(g, s, str, Ntrajs, Q, fine, stream);
}

int mr_grid_spread_samples(const mr_grid &g, const mr_grid_samples &s, const double2 *str, int Q,
                           double2 *fine, hipStream_t stream)
{
    return mr_spread_launch<true>(g, s, str, s.n, Q, fine, stream);
}

// Type-2 interpolation, one thread per (sorted sample, set q).
template <int D>
__global__ void mr_k_interp(mr_grid g, const int *__restrict__ i0s, const double *__restrict__ wts,
                            const int *__restrict__ perm, long long n, const double2 *__restrict__ fine,
                            int Q, double2 *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n || q >= Q) return;
    const int w = g.w;
    const double2 *f = fine + (long long)q * g.nfine;
    const double *wx = wts + (i * D + 0) * w;
    const double *wy = wts + (i * D + 1) * w;
    const double *wz = (D == 3) ? wts + (i * D + 2) * w : nullptr;
    const int ix = i0s[i * 3], iy = i0s[i * 3 + 1], iz = i0s[i * 3 + 2];
    double ar = 0.0, ai = 0.0;
    const int nz = (D == 3) ? w : 1;
    for (int kz = 0; kz < nz; ++kz) {
        int gz = 0;
        double fz = 1.0;
        if (D == 3) {
            gz = iz + kz;
            if (gz >= g.nf[2]) gz -= g.nf[2];
            fz = wz[kz];
        }
        for (int ky = 0; ky < w; ++ky) {
            int gy = iy + ky;
            if (gy >= g.nf[1]) gy -= g.nf[1];
            const double fy = fz * wy[ky];
            const double2 *row = f + (long long)g.nf[0] * (gy + (long long)g.nf[1] * gz);
            for (int kx = 0; kx < w; ++kx) {
                int gx = ix + kx;
                if (gx >= g.nf[0]) gx -= g.nf[0];
                const double wv = fy * wx[kx];
                const double2 v = row[gx];
                ar += wv * v.x;
                ai += wv * v.y;
            }
        }
    }
    out[(long long)q * n + perm[i]] = make_double2(ar, ai);
}

int mr_grid_interp(const mr_grid &g, const mr_grid_samples &s, const double2 *fine, int Q,
                   double2 *out, hipStream_t stream)
{
    const dim3 grid((unsigned)mr_ceil_div(s.n, 256), (unsigned)Q);
    if (g.d == 2)
        hipLaunchKernelGGL(mr_k_interp<2>, grid, dim3(256), 0, stream, g, s.i0, s.wt, s.perm, s.n,
                           fine, Q, out);
    else
        hipLaunchKernelGGL(mr_k_interp<3>, grid, dim3(256), 0, stream, g, s.i0, s.wt, s.perm, s.n,
                           fine, Q, out);
    MR_LAUNCH_CHECK("mr_k_interp");
    return MR_SUCCESS;
}

int mr_grid_fft(const mr_grid &g, double2 *fine, double2 *, int Q, bool inverse,
                hipStream_t stream)
{
    mr_fft_len<double> f[3];
    for (int a = 0; a < g.d; ++a) MR_TRY(mr_fft_get<double>(g.nf[a], &f[a]));
    const long long nx = g.nf[0], ny = g.nf[1], nz = g.nf[2];
    if (!inverse) {
        MR_TRY((mr_fft_pass<double, false>(f[0], mr_ld_axis<double2, true>{fine, 1, (int)nx},
                                           mr_st_axis<double2, true>{fine, 1, (int)nx},
                                           ny * nz * Q, 0, 4, stream, "grid x")));
        MR_TRY((mr_fft_pass<double, false>(f[1], mr_ld_axis<double2, false>{fine, nx, (int)ny},
                                           mr_st_axis<double2, false>{fine, nx, (int)ny},
                                           nx * nz * Q, 0, 16, stream, "grid y")));
        if (g.d == 3)
            MR_TRY((mr_fft_pass<double, false>(
                f[2], mr_ld_axis<double2, false>{fine, nx * ny, (int)nz},
                mr_st_axis<double2, false>{fine, nx * ny, (int)nz}, nx * ny * Q, 0, 16, stream,
                "grid z")));
    } else {
        MR_TRY((mr_fft_pass<double, true>(f[0], mr_ld_axis<double2, true>{fine, 1, (int)nx},
                                          mr_st_axis<double2, true>{fine, 1, (int)nx},
                                          ny * nz * Q, 0, 4, stream, "grid x")));
        MR_TRY((mr_fft_pass<double, true>(f[1], mr_ld_axis<double2, false>{fine, nx, (int)ny},
                                          mr_st_axis<double2, false>{fine, nx, (int)ny},
                                          nx * nz * Q, 0, 16, stream, "grid y")));
        if (g.d == 3)
            MR_TRY((mr_fft_pass<double, true>(
                f[2], mr_ld_axis<double2, false>{fine, nx * ny, (int)nz},
                mr_st_axis<double2, false>{fine, nx * ny, (int)nz}, nx * ny * Q, 0, 16, stream,
                "grid z")));
    }
    return MR_SUCCESS;
}
