// CPU NUFFT spreading and interpolation for the CPU baseline of the
// benchmark: the FINUFFT-style A^H A (method="nufft") of the reference, whose
// GPU form is src/cuda/mr_nufft_op.cu:613-700 and whose Python form is
// python/mr_aha/nufft.py:473-552. TEST INFRASTRUCTURE (oracle side): only
// bench.py's CPU-baseline leg and tests/ load it, never the product.
//
// Kernel: exponential of semicircle (Barnett, Magland & af Klinteberg, SIAM J.
// Sci. Comput. 41(5), 2019), phi(z) = exp(beta (sqrt(1 - (2 z / w)^2) - 1))
// on |z| <= w/2 in fine-grid units, w = ceil(log10(1/eps)) + 1, beta =
// 2.30 w, upsampling 2. The FFTs and the deconvolution by the kernel's
// Fourier transform are done by the Python harness (oracle/cpu_nufft.py).
//
// Layouts: a fine grid is [nf2][nf1][nf0][B] complex (B vectors interleaved
// fastest, so that one sample's kernel weights serve every vector of the
// batch); strengths and values at the samples are [M][B] complex; the
// sample coordinates are radians in [-pi, pi), one array per axis (axis 0
// fastest on the grid). OpenMP: interpolation splits the samples; spreading
// splits the slowest grid axis into bands of at least w planes, and a band
// takes the samples whose support starts in it or in the band before it, so
// no two threads write one grid point.

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

template <typename R> struct es_kernel {
    int w;
    R beta, half, inv_half;
    // w weights of the support that starts at grid index i0 = ceil(u - w/2)
    inline int weights(R u, R *wt) const
    {
        const R s = std::ceil(u - half);
        for (int k = 0; k < w; ++k) {
            const R z = (s + (R)k - u) * inv_half;
            const R t = (R)1 - z * z;
            wt[k] = t > (R)0 ? std::exp(beta * (std::sqrt(t) - (R)1)) : (R)0;
        }
        return (int)s;
    }
};

template <typename R> es_kernel<R> make_kernel(int w, double beta)
{
    es_kernel<R> k;
    k.w = w;
    k.beta = (R)beta;
    k.half = (R)w / (R)2;
    k.inv_half = (R)2 / (R)w;
    return k;
}

inline int64_t wrap(int64_t i, int64_t n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

// grid position of radians x on a grid of nf points over [0, 2 pi)
template <typename R> inline R grid_pos(R x, int64_t nf)
{
    return x * (R)nf / (R)(2.0 * M_PI);
}

constexpr int WMAX = 16;

template <typename R>
void interp(int ndim, const int64_t *nf, int64_t B, const std::complex<R> *grid, int64_t M,
            const float *const *pts, int w, double beta, std::complex<R> *out)
{
    const es_kernel<R> ker = make_kernel<R>(w, beta);
#pragma omp parallel
    {
        R wt[3][WMAX];
        std::vector<std::complex<R>> acc(B);
#pragma omp for schedule(static)
        for (int64_t m = 0; m < M; ++m) {
            int64_t i0[3] = {0, 0, 0};
            for (int a = 0; a < ndim; ++a)
                i0[a] = ker.weights(grid_pos<R>((R)pts[a][m], nf[a]), wt[a]);
            for (int64_t b = 0; b < B; ++b) acc[b] = 0;
            const int kz = ndim == 3 ? w : 1;
            for (int c = 0; c < kz; ++c) {
                const int64_t iz = ndim == 3 ? wrap(i0[2] + c, nf[2]) : 0;
                const R wz = ndim == 3 ? wt[2][c] : (R)1;
                for (int bb = 0; bb < w; ++bb) {
                    const int64_t iy = wrap(i0[1] + bb, nf[1]);
                    const R wyz = wz * wt[1][bb];
                    const std::complex<R> *row = grid + ((iz * nf[1] + iy) * nf[0]) * B;
                    for (int aa = 0; aa < w; ++aa) {
                        const int64_t ix = wrap(i0[0] + aa, nf[0]);
                        const R wv = wyz * wt[0][aa];
                        const std::complex<R> *g = row + ix * B;
                        for (int64_t b = 0; b < B; ++b) acc[b] += wv * g[b];
                    }
                }
            }
            std::complex<R> *o = out + m * B;
            for (int64_t b = 0; b < B; ++b) o[b] = acc[b];
        }
    }
}

template <typename R>
void spread(int ndim, const int64_t *nf, int64_t B, std::complex<R> *grid, int64_t M,
            const float *const *pts, int w, double beta, const std::complex<R> *str)
{
    const es_kernel<R> ker = make_kernel<R>(w, beta);
    const int sa = ndim - 1;                 // slowest axis
    const int64_t ns = nf[sa];
    const int nth = omp_get_max_threads();
    int64_t bs = ns / (4 * (int64_t)nth);
    if (bs < w) bs = w;
    int64_t nb = ns / bs;
    if (nb < 1) nb = 1;
    auto band_of = [&](int64_t i) { return i / bs < nb ? i / bs : nb - 1; };
    // bucket the samples by the band their support starts in
    std::vector<int64_t> start(M), cnt(nb + 1, 0), order(M);
    for (int64_t m = 0; m < M; ++m) {
        const R u = grid_pos<R>((R)pts[sa][m], ns);
        start[m] = wrap((int64_t)std::ceil(u - ker.half), ns);
        ++cnt[band_of(start[m]) + 1];
    }
    for (int64_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    {
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t m = 0; m < M; ++m) order[fill[band_of(start[m])]++] = m;
    }
    const int64_t plane = (ndim == 3 ? nf[1] * nf[0] : nf[0]) * B;
#pragma omp parallel
    {
        R wt[3][WMAX];
#pragma omp for schedule(dynamic, 1)
        for (int64_t band = 0; band < nb; ++band) {
            const int64_t lo = band * bs, hi = band == nb - 1 ? ns : lo + bs;
            for (int pass = 0; pass < 2; ++pass) {
                const int64_t src = pass == 0 ? band : (band + nb - 1) % nb;
                if (pass == 1 && src == band) break;
                for (int64_t q = cnt[src]; q < cnt[src + 1]; ++q) {
                    const int64_t m = order[q];
                    int64_t i0[3] = {0, 0, 0};
                    for (int a = 0; a < ndim; ++a)
                        i0[a] = ker.weights(grid_pos<R>((R)pts[a][m], nf[a]), wt[a]);
                    const std::complex<R> *s = str + m * B;
                    for (int c = 0; c < w; ++c) {
                        const int64_t is = wrap(i0[sa] + c, ns);
                        if (is < lo || is >= hi) continue;
                        const R ws = wt[sa][c];
                        std::complex<R> *pl = grid + is * plane;
                        if (ndim == 3) {
                            for (int bb = 0; bb < w; ++bb) {
                                const int64_t iy = wrap(i0[1] + bb, nf[1]);
                                const R wy = ws * wt[1][bb];
                                std::complex<R> *row = pl + iy * nf[0] * B;
                                for (int aa = 0; aa < w; ++aa) {
                                    const R wv = wy * wt[0][aa];
                                    std::complex<R> *g = row + wrap(i0[0] + aa, nf[0]) * B;
                                    for (int64_t b = 0; b < B; ++b) g[b] += wv * s[b];
                                }
                            }
                        } else {
                            for (int aa = 0; aa < w; ++aa) {
                                const R wv = ws * wt[0][aa];
                                std::complex<R> *g = pl + wrap(i0[0] + aa, nf[0]) * B;
                                for (int64_t b = 0; b < B; ++b) g[b] += wv * s[b];
                            }
                        }
                    }
                }
            }
        }
    }
}

}  // namespace

extern "C" {

// single precision (complex64 data)
int mr_cpu_interp_f(int ndim, const int64_t *nf, int64_t B, const void *grid, int64_t M,
                    const float *px, const float *py, const float *pz, int w, double beta,
                    void *out)
{
    if (ndim < 2 || ndim > 3 || w < 2 || w > WMAX) return -1;
    const float *pts[3] = {px, py, pz};
    interp<float>(ndim, nf, B, static_cast<const std::complex<float> *>(grid), M, pts, w, beta,
                  static_cast<std::complex<float> *>(out));
    return 0;
}

int mr_cpu_spread_f(int ndim, const int64_t *nf, int64_t B, void *grid, int64_t M,
                    const float *px, const float *py, const float *pz, int w, double beta,
                    const void *str)
{
    if (ndim < 2 || ndim > 3 || w < 2 || w > WMAX) return -1;
    const float *pts[3] = {px, py, pz};
    spread<float>(ndim, nf, B, static_cast<std::complex<float> *>(grid), M, pts, w, beta,
                  static_cast<const std::complex<float> *>(str));
    return 0;
}

// double precision (complex128 data; the FP64 cross-check at small eps)
int mr_cpu_interp_d(int ndim, const int64_t *nf, int64_t B, const void *grid, int64_t M,
                    const float *px, const float *py, const float *pz, int w, double beta,
                    void *out)
{
    if (ndim < 2 || ndim > 3 || w < 2 || w > WMAX) return -1;
    const float *pts[3] = {px, py, pz};
    interp<double>(ndim, nf, B, static_cast<const std::complex<double> *>(grid), M, pts, w, beta,
                   static_cast<std::complex<double> *>(out));
    return 0;
}

int mr_cpu_spread_d(int ndim, const int64_t *nf, int64_t B, void *grid, int64_t M,
                    const float *px, const float *py, const float *pz, int w, double beta,
                    const void *str)
{
    if (ndim < 2 || ndim > 3 || w < 2 || w > WMAX) return -1;
    const float *pts[3] = {px, py, pz};
    spread<double>(ndim, nf, B, static_cast<std::complex<double> *>(grid), M, pts, w, beta,
                   static_cast<const std::complex<double> *>(str));
    return 0;
}

int mr_cpu_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
