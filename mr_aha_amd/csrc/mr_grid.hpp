// Non-uniform gridding shared by the Toeplitz kernel build (type-1, sign -1)
// and the NUFFT operator (type-1 / type-2): exponential-of-semicircle (ES)
// kernel of Barnett, Magland and af Klinteberg, "A parallel non-uniform fast
// Fourier transform library based on an exponential of semicircle kernel",
// SIAM J. Sci. Comput. 41(5) (2019): width w = ceil(log10(1/eps)) + 1,
// phi(z) = exp(beta (sqrt(1 - z^2) - 1)), beta = 2.30 w, upsampling 2.
//
// Sample m at angle x (radians) sits at fine-grid coordinate u = x nf/(2 pi);
// it touches the w points i0 .. i0+w-1 with i0 = ceil(u - w/2) (mod nf) and
// weight phi(2 (i - u) / w). The fine grid is cut into bins of bs points per
// axis (bs divides nf); a sample belongs to the bin of its first point i0.
#pragma once

#include "mr_hip.hpp"

#include <vector>

#define MR_GRID_MAXW 16

struct mr_grid {
    int d;           // transformed dimensions (2 or 3)
    int nf[3];       // fine grid, 1 beyond d
    int M[3];        // modes per axis, 1 beyond d
    int bs[3];       // bin size per axis
    int nb[3];       // bins per axis
    int w;           // kernel width
    double beta;
    long long nfine; // prod nf
    long long nbins; // prod nb
};

// Fill the geometry for `d` transformed axes with mode counts M[0..d-1].
void mr_grid_setup(mr_grid *g, int d, const int *M, double eps);

// Deconvolution factors 1/phi_hat(delta) for the modes of axis a in
// natural order (index j <-> delta = j or j - M). Host vector.
std::vector<double> mr_grid_deconv(const mr_grid &g, int a);

// Per-sample geometry, sorted by (bin, readout, readout position).
struct mr_grid_samples {
    long long n = 0;        // samples
    int *i0 = nullptr;      // [n * 3] first fine point per axis (mod nf), sorted order
    int *rd = nullptr;      // [n] readout of each sorted sample
    double *wt = nullptr;   // [n * d * w] weights per axis, sorted order
    int *bin_lo = nullptr;  // [nbins] first sorted sample of each bin
    int *bin_hi = nullptr;  // [nbins] one past the last
    int *perm = nullptr;    // [n] original sample index of each sorted sample
};

// Build the sorted sample table from a [Ntrajs, Nkx, ndim] trajectory
// (first `g.d` coordinates used) on `stream`. Synchronous.
int mr_grid_samples_build(const mr_grid &g, const float *traj, long long Ntrajs, long long Nkx,
                          int ndim, hipStream_t stream, mr_grid_samples *out);
void mr_grid_samples_free(mr_grid_samples *s);

// Type-1 spreading of Q strength sets (strength of sample m, set q:
// str[q * Ntrajs + readout(m)], i.e. constant along a readout) onto Q fine
// grids fine[q * nfine + g], complex double. Gather formulation: every fine
// point sums the samples of the bins that can reach it; no atomics.
int mr_grid_spread(const mr_grid &g, const mr_grid_samples &s, const double2 *str,
                   long long Ntrajs, int Q, double2 *fine, hipStream_t stream);

// Same with per-sample strengths str[q * nsamp + m] (m = original index).
int mr_grid_spread_samples(const mr_grid &g, const mr_grid_samples &s, const double2 *str,
                           int Q, double2 *fine, hipStream_t stream);

// Type-2 interpolation: out[q * nsamp + m] = sum_points fine[q][p] weight.
int mr_grid_interp(const mr_grid &g, const mr_grid_samples &s, const double2 *fine, int Q,
                   double2 *out, hipStream_t stream);

// Forward or inverse FFT of Q fine grids in place (double precision).
int mr_grid_fft(const mr_grid &g, double2 *fine, double2 *scratch, int Q, bool inverse,
                hipStream_t stream);
