// Toeplitz A^H A apply on MI355X: plan life cycle and the execute pipeline.
//
// Per coil group (G coils, V = L*G rank volumes), the operator
//   AhAx = sum_c S_c^H P^H F^-1 T F P S_c x           (mr_aha.h of the reference, :580-603)
// runs as fused axis passes of the project FFT engine and one block pass:
//   F1  x-FFT of every image row, loading S_c * x_l zero-padded in natural
//       order (pad and sensitivity fused into the load)    -> A1 [Mx, Ny, Nz, V]
//   F2  y-FFT, zero-padding the Ny rows on load            -> A2 [Mx, My, Nz, V]
//   F3  z-FFT (3D / stack), zero-padding the Nz planes    -> K  [Mx, My, Mz, V]
//   BD  per-location Hermitian block multiply v <- v B(k), in place on K
//   I3  inverse z-FFT, keeping the Nz image planes        -> A2'
//   I2  inverse y-FFT, keeping the Ny image rows          -> A1'
//   I1  inverse x-FFT, crop, conj(S_c) and the coil sum fused into the store
// Only the image region is ever loaded or stored outside k-space: the zero
// rows of the padding are synthesised in LDS and the discarded rows of the
// inverse transforms are never written.
//
// Reference behaviour followed: stage order and coil grouping of
// src/cuda/mr_aha.cu:694-790; index maps of src/cuda/mr_kernels.cu:98-176;
// the block walk of src/cuda/mr_kernels_bd.cu:74-139 (each stored element
// read once and applied to both of its mirror positions).

#include "mr_hip.hpp"
#include "mr_aha/mr_aha_ext.h"
#include "mr_xbd.hpp"
#include "mr_fft_ct.hpp"

#include <cstdlib>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

// ---------------------------------------------------------------------------
// Fused pass functors
// ---------------------------------------------------------------------------

// F1 load: pencil p = v + V * yz (the V rank volumes of one image row are
// neighbouring pencils, so the row of x_l and S_c is reused from cache),
// v = l + L * c; natural element e.
struct mr_ld_pad_sx {
    static constexpr bool kElemFast = true;
    struct PB {
        long long xo, so;
    };
    const float2 *x, *S;
    long long N, NyNz, V;
    int Nx, Mx, cnx, L, coil0;
    __device__ int nsub() const { return 1; }
    __device__ PB pencil(long long p, int) const
    {
        const long long yz = p / V, v = p - yz * V;
        const int l = (int)(v % L), c = (int)(v / L);
        return PB{(long long)Nx * yz + N * l, (long long)Nx * yz + N * (coil0 + c)};
    }
    __device__ float2 get(const PB &b, int e) const
    {
        int i = e + cnx;
        if (i >= Mx) i -= Mx;
        return (i < Nx) ? cmul(S[b.so + i], x[b.xo + i]) : make_float2(0.f, 0.f);
    }
};

// Row store for F1's pencil order p = v + V * yz -> row yz + NyNz * v.
struct mr_st_rows_vfast {
    static constexpr bool kElemFast = true;
    using PB = long long;
    float2 *out;
    long long NyNz, V;
    int Mx;
    __device__ PB pencil(long long p, int) const
    {
        const long long yz = p / V, v = p - yz * V;
        return (long long)Mx * (yz + NyNz * v);
    }
    __device__ void put(const PB &b, int, int e, int, float2 v, float2 *) const { out[b + e] = v; }
    __device__ void flush(long long, int, float2 *) const {}
};

// I1 load: pencil p = l + L * yz, round `sub` = coil of the group.
struct mr_ld_row_coil {
    static constexpr bool kElemFast = true;
    using PB = long long;
    const float2 *in;
    long long NyNz;
    int Mx, L, ncoil;
    __device__ int nsub() const { return ncoil; }
    __device__ PB pencil(long long p, int sub) const
    {
        const long long yz = p / L, l = p - yz * L;
        return (long long)Mx * (yz + NyNz * (l + (long long)L * sub));
    }
    __device__ float2 get(const PB &b, int e) const { return in[b + e]; }
};

// I1 store: crop, conj(S_c) and coil sum in LDS; flush writes AhAx.
struct mr_st_crop_conj_sum {
    static constexpr bool kElemFast = true;
    using PB = long long;  // offset of the coil's map row
    float2 *out;
    const float2 *S;
    long long N, NyNz;
    int Nx, Mx, cnx, coil0, accumulate, L;
    __device__ PB pencil(long long p, int sub) const
    {
        const long long yz = p / L;
        return (long long)Nx * yz + N * (coil0 + sub);
    }
    __device__ void put(const PB &b, int pl, int e, int sub, float2 v, float2 *acc) const
    {
        int i = e + cnx;
        if (i >= Mx) i -= Mx;
        if (i >= Nx) return;
        const float2 t = cmulc(S[b + i], v);
        float2 *a = acc + (size_t)pl * Nx + i;
        *a = (sub == 0) ? t : cadd(*a, t);
    }
    __device__ void flush(long long p0, int tp, float2 *acc) const
    {
        for (int pl = 0; pl < tp; ++pl) {
            const long long p = p0 + pl;
            const long long yz = p / L, l = p - yz * L;
            float2 *o = out + (long long)Nx * yz + N * l;
            const float2 *a = acc + (size_t)pl * Nx;
            for (int i = threadIdx.x; i < Nx; i += blockDim.x)
                o[i] = accumulate ? cadd(o[i], a[i]) : a[i];
        }
    }
};

// Outer-axis passes. The slowest spatial axis (z in 3D and stack-of-stars,
// y in 2D) is transformed on the image side, before the inner axes, with the
// sensitivity product fused into its load, and its inverse runs last with
// crop, conj(S) and the coil sum fused into its store. "plane" is the stride
// of that axis (Nx*Ny, or Nx in 2D); Nz its image length and Mz its transform
// length (Mz = Nz for stack-of-stars kz, which is never padded).
// Pencils are enumerated p = lt + TXZ * (v + V * tile) with the in-plane
// location i = tile * TXZ + lt (TXZ a power of two dividing plane): a tile of
// TXZ neighbouring locations is revisited for every volume v by consecutive
// workgroups, which reuse its rows of x and S from cache.
struct mr_zmap {
    long long V;
    int lgTXZ;
    __device__ void split(long long p, long long &i, long long &v) const
    {
        const long long lt = p & ((1LL << lgTXZ) - 1), rest = p >> lgTXZ;
        const long long tile = rest / V;
        v = rest - tile * V;
        i = (tile << lgTXZ) + lt;
    }
};

// Outer forward load: S_c x_l along the axis in natural order, zero-padded
// (image z = (e + cnz) mod Mz when < Nz).
struct mr_ld_sx_z {
    static constexpr bool kElemFast = false;
    struct PB {
        long long xo, so;
    };
    const float2 *x, *S;
    mr_zmap zm;
    long long NxNy, N;
    int Nz, Mz, cnz, L, coil0;
    __device__ int nsub() const { return 1; }
    __device__ PB pencil(long long p, int) const
    {
        long long i, v;
        zm.split(p, i, v);
        const int l = (int)(v % L), c = (int)(v / L);
        return PB{i + N * l, i + N * (coil0 + c)};
    }
    __device__ float2 get(const PB &b, int e) const
    {
        int z = e + cnz;
        if (z >= Mz) z -= Mz;
        if (z >= Nz) return make_float2(0.f, 0.f);
        const long long o = NxNy * z;
        return cmul(S[b.so + o], x[b.xo + o]);
    }
};

// Outer forward store: [plane, Mz (natural k), V].
struct mr_st_z {
    static constexpr bool kElemFast = false;
    using PB = long long;
    float2 *out;
    mr_zmap zm;
    long long NxNy;
    int Mz;
    __device__ PB pencil(long long p, int) const
    {
        long long i, v;
        zm.split(p, i, v);
        return i + NxNy * Mz * v;
    }
    __device__ void put(const PB &b, int, int e, int, float2 val, float2 *) const
    {
        out[b + NxNy * e] = val;
    }
    __device__ void flush(long long, int, float2 *) const {}
};

// Outer inverse load: pencil (i, l), round `sub` = coil; [plane, Mz, V] input.
struct mr_ld_z_coil {
    static constexpr bool kElemFast = false;
    using PB = long long;
    const float2 *in;
    mr_zmap zm;  // V = L here
    long long NxNy;
    int Mz, L, ncoil;
    __device__ int nsub() const { return ncoil; }
    __device__ PB pencil(long long p, int sub) const
    {
        long long i, l;
        zm.split(p, i, l);
        return i + NxNy * Mz * (l + (long long)L * sub);
    }
    __device__ float2 get(const PB &b, int e) const { return in[b + NxNy * e]; }
};

// Outer inverse store: image z = (e + cnz) mod Mz (elements with z >= Nz
// are cropped), conj(S_c), coil sum in LDS; flush writes AhAx[i, z, l].
struct mr_st_z_conj_sum {
    static constexpr bool kElemFast = false;
    using PB = long long;
    float2 *out;
    const float2 *S;
    mr_zmap zm;  // V = L
    long long NxNy, N;
    int Nz, Mz, cnz, coil0, accumulate;
    __device__ PB pencil(long long p, int sub) const
    {
        long long i, l;
        zm.split(p, i, l);
        return i + N * (coil0 + sub);
    }
    __device__ void put(const PB &b, int pl, int e, int sub, float2 v, float2 *acc) const
    {
        int z = e + cnz;
        if (z >= Mz) z -= Mz;
        if (z >= Nz) return;
        const float2 t = cmulc(S[b + NxNy * z], v);
        float2 *a = acc + (size_t)pl * Nz + z;
        *a = (sub == 0) ? t : cadd(*a, t);
    }
    __device__ void flush(long long p0, int tp, float2 *acc) const
    {
        for (int q = threadIdx.x; q < tp * Nz; q += blockDim.x) {
            const int z = q / tp, pl = q - z * tp;
            long long i, l;
            zm.split(p0 + pl, i, l);
            float2 *o = out + i + NxNy * z + N * l;
            const float2 a = acc[(size_t)pl * Nz + z];
            *o = accumulate ? cadd(*o, a) : a;
        }
    }
};

// ---------------------------------------------------------------------------
// Block-diagonal multiply
// ---------------------------------------------------------------------------

// y_c = sum_r v_r A_rc over NC vectors, walking the stored triangle in packed
// order once: packed element (r, c) adds A_rc v_r to y_c and, off the
// diagonal, conj(A_rc) v_c to y_r.
template <int L, int NC>
__device__ __forceinline__ void mr_bd_walk(const float2 (&v)[NC][L], float2 (&y)[NC][L],
                                           const float2 *bp, long long stride, int upper)
{
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int i = 0; i < L; ++i) y[k][i] = make_float2(0.f, 0.f);
#define MR_BD_EL(r, c)                                                         \
    do {                                                                       \
        const float2 a = *bp;                                                  \
        bp += stride;                                                          \
        _Pragma("unroll") for (int k = 0; k < NC; ++k) {                       \
            y[k][c] = cmac(y[k][c], a, v[k][r]);                         \
            if ((r) != (c)) y[k][r] = cmacc(y[k][r], a, v[k][c]);        \
        }                                                                      \
    } while (0)
    if (upper) {
#pragma unroll
        for (int c = 0; c < L; ++c) {
            asm volatile("" ::: "memory");
#pragma unroll
            for (int r = 0; r <= c; ++r) MR_BD_EL(r, c);
        }
    } else {
#pragma unroll
        for (int c = 0; c < L; ++c) {
            asm volatile("" ::: "memory");
#pragma unroll
            for (int r = c; r < L; ++r) MR_BD_EL(r, c);
        }
    }
#undef MR_BD_EL
}

// In place over K [P, L, ncoil] (natural order). Unmasked: kernel [P, T]
// centered. Masked: inv[loc] is the kernel row of natural location loc or -1
// (a location outside the mask multiplies by zero).
//
// One workgroup per tile of TP consecutive locations: the tile's T packed
// kernel entries are staged once in LDS (coalesced: consecutive natural
// locations are consecutive centered locations except at the one wrap point
// of the x axis) and every coil of the group is multiplied from LDS, so the
// kernel streams from HBM once per apply. Thread (p, slot) handles location
// p for coils slot, slot + nslot, ...
template <int L, int TP, bool MASKED>
__global__ void __launch_bounds__(256)
mr_k_bd(float2 *__restrict__ K, const float2 *__restrict__ ker, const int *__restrict__ inv,
        long long kstride, mr_dims d, int ncoil)
{
    constexpr int T = L * (L + 1) / 2;
    __shared__ float2 kt[T * TP];
    __shared__ long long kb[TP];
    const long long P = d.pad_size;
    const long long loc0 = mr_xcd_block(blockIdx.x, gridDim.x) * TP;
    if (threadIdx.x < TP) {
        const long long loc = loc0 + threadIdx.x;
        long long b = -1;
        if (loc < P) b = MASKED ? (long long)inv[loc] : mr_nat_to_cent_lin(loc, d);
        kb[threadIdx.x] = b;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < T * TP; f += blockDim.x) {
        const int t = f / TP, p = f - t * TP;
        const long long b = kb[p];
        kt[f] = (b >= 0) ? ker[b + kstride * t] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int p = threadIdx.x % TP;
    const int slot = threadIdx.x / TP;
    const int nslot = blockDim.x / TP;
    const long long loc = loc0 + p;
    if (loc >= P) return;
    for (int c = slot; c < ncoil; c += nslot) {
        float2 *g = K + loc + P * (long long)L * c;
        float2 v[1][L], y[1][L];
#pragma unroll
        for (int l = 0; l < L; ++l) v[0][l] = g[P * l];
        mr_bd_walk<L, 1>(v, y, kt + p, TP, d.tri_upper);
#pragma unroll
        for (int l = 0; l < L; ++l) g[P * l] = y[0][l];
    }
}

// inv[natural location] = mask row, -1 elsewhere (mask holds centered
// linear indices; src/cuda/mr_kernels.cu:259-281).
__global__ void mr_k_fill_int(int *p, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void mr_k_mask_inverse(int *inv, const int64_t *mask, long long nm, mr_dims d)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm) return;
    long long k = mask[t];
    if (k < 0 || k >= d.pad_size) return;
    const int k0 = (int)(k % d.m[0]);
    const long long q = k / d.m[0];
    const int k1 = (int)(q % d.m[1]);
    const int k2 = (int)(q / d.m[1]);
    int j0 = k0 - d.c_m[0]; if (j0 < 0) j0 += d.m[0];
    int j1 = k1 - d.c_m[1]; if (j1 < 0) j1 += d.m[1];
    int j2 = k2 - d.c_m[2]; if (j2 < 0) j2 += d.m[2];
    inv[j0 + (long long)d.m[0] * (j1 + (long long)d.m[1] * j2)] = (int)t;
}

typedef void (*mr_bd_fn)(float2 *, const float2 *, const int *, long long, mr_dims, int);

template <int TP, bool MASKED> static mr_bd_fn mr_bd_pick_tp(int L)
{
    switch (L) {
#define MR_BD_CASE(n) case n: return mr_k_bd<n, TP, MASKED>;
        MR_BD_CASE(1) MR_BD_CASE(2) MR_BD_CASE(3) MR_BD_CASE(4)
        MR_BD_CASE(5) MR_BD_CASE(6) MR_BD_CASE(7) MR_BD_CASE(8)
        MR_BD_CASE(9) MR_BD_CASE(10) MR_BD_CASE(11) MR_BD_CASE(12)
        MR_BD_CASE(13) MR_BD_CASE(14) MR_BD_CASE(15) MR_BD_CASE(16)
#undef MR_BD_CASE
    default: return nullptr;
    }
}

template <bool MASKED> static mr_bd_fn mr_bd_pick(int L, int TP)
{
    if (TP == 16) return mr_bd_pick_tp<16, MASKED>(L);
    if (TP == 64) return mr_bd_pick_tp<64, MASKED>(L);
    return mr_bd_pick_tp<32, MASKED>(L);
}

// ---------------------------------------------------------------------------
// Fused x pass (x-last pipeline)
// ---------------------------------------------------------------------------

// When the outer axes are transformed on the image side first, the x axis is
// the last one: for one k-space row r = (ky, kz) (natural order) the forward
// x-FFT, the block multiply of the row's Mx locations and the inverse x-FFT
// need nothing outside the row. One workgroup takes the row for a group of up
// to G coils of Y [Nx, R, V] (R = My*Mz rows, V = L * coils):
//   load   the L*G pencils of Nx image columns, zero-padded to Mx in natural
//          order (pad map of src/cuda/mr_kernels.cu:111-128)
//   FFT    forward, radix stages in LDS
//   block  v <- v B(k) at each of the Mx locations; the row's T*Mx kernel
//          entries are read coalesced (centered kx runs contiguously except at
//          the one wrap point) and shared, through L2, by the G coils and by
//          the coil groups of the row, which are consecutive workgroups of
//          one XCD (src/cuda/mr_kernels_bd.cu:74-139 for the walk)
//   IFFT   inverse
//   store  the Nx image columns back in place (crop of mr_kernels.cu:158-165)
// Padded k-space never reaches HBM: per row and coil the pass moves 2*L*Nx
// values plus the kernel row. Twiddles are read from the global table (L1
// resident) so LDS holds exactly G*L*Mx values and two 80 KB workgroups can
// share a CU.

template <bool INV>
__device__ __forceinline__ void mr_xbd_fft(float2 *A, const mr_fft_len<float> &f, int tp)
{
    int Ns = 1;
    for (int s = 0; s < f.nst; ++s) {
        const int RR = f.radix[s];
        const mr_fdiv &md = f.mdiv[s], &nd = f.nsdiv[s];
        switch (RR) {
        case 2: mr_stage_reg<2, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        case 3: mr_stage_reg<3, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        case 4: mr_stage_reg<4, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        case 5: mr_stage_reg<5, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        case 6: mr_stage_reg<6, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        case 7: mr_stage_reg<7, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        case 10: mr_stage_reg<10, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        default: mr_stage_reg<8, INV, float2, MR_XBD_THREADS, MR_XBD_EMAX>(A, f.tws, f.n, f.n, tp, Ns, md, nd); break;
        }
        Ns *= RR;
    }
}

template <int L, bool MASKED>
__global__ void __launch_bounds__(MR_XBD_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
mr_k_xbd(float2 *__restrict__ Y, const float2 *__restrict__ ker, const int *__restrict__ inv,
         long long kstride, mr_dims d, mr_fft_len<float> fx, mr_fdiv nxdiv, long long R, int G,
         int ncoil, int ncg)
{
    constexpr int NW = MR_XBD_THREADS / 64;
    extern __shared__ __align__(16) unsigned char mr_smem[];
    float2 *A = reinterpret_cast<float2 *>(mr_smem);
    const int Nx = d.n[0], Mx = d.m[0], cnx = d.c_n[0];
    const long long q = mr_xcd_block(blockIdx.x, gridDim.x);
    const long long r = q / ncg;
    const int c0 = (int)(q - r * ncg) * G;
    const int gn = (ncoil - c0 < G) ? ncoil - c0 : G;
    const int tp = gn * L;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // pencil pl = g*L + l of the workgroup is volume v = L*c0 + pl; every
    // wavefront loads and stores whole pencils, so a pencil's base is uniform
    const float2 *Yr = Y + (long long)Nx * r + (long long)Nx * R * L * c0;
    const long long pstride = (long long)Nx * R;
    const unsigned rowbytes = (unsigned)Nx * 8u;

    for (int pl = wave; pl < tp; pl += NW) {
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, rowbytes);
        float2 *Ap = A + pl * Mx;
        for (int e = lane; e < Mx; e += 64) {
            int s = e + cnx;
            if (s >= Mx) s -= Mx;
            Ap[e] = mr_bload(rs, (unsigned)s * 8u);   // s >= Nx reads zero: the padding
        }
    }
    __syncthreads();
    mr_xbd_fft<false>(A, fx, tp);

    const int My = d.m[1];
    const int jy = (int)(r % My), jz = (int)(r / My);
    int ky = jy + d.c_m[1]; if (ky >= My) ky -= My;
    int kz = jz + d.c_m[2]; if (kz >= d.m[2]) kz -= d.m[2];
    const long long rowc = (long long)Mx * (ky + (long long)My * kz);
    for (int w = threadIdx.x; w < gn * Mx; w += MR_XBD_THREADS) {
        const int g = (int)mr_fdiv_q(fx.ndiv, (unsigned)w), e = w - g * Mx;
        long long b;
        if (MASKED) {
            b = inv[e + (long long)Mx * r];
        } else {
            int kx = e + d.c_m[0];
            if (kx >= Mx) kx -= Mx;
            b = rowc + kx;
        }
        float2 *a = A + (size_t)g * L * Mx + e;
        if (b < 0) {
#pragma unroll
            for (int l = 0; l < L; ++l) a[l * Mx] = make_float2(0.f, 0.f);
        } else if (d.tri_upper) {
            mr_bd_walk_g<L, true>(a, Mx, ker, kstride, (unsigned)(b * 8));
        } else {
            mr_bd_walk_g<L, false>(a, Mx, ker, kstride, (unsigned)(b * 8));
        }
    }
    __syncthreads();
    mr_xbd_fft<true>(A, fx, tp);

    for (int pl = wave; pl < tp; pl += NW) {
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, rowbytes);
        const float2 *Ap = A + pl * Mx;
        for (int s = lane; s < Nx; s += 64) {
            int e = s - cnx;
            if (e < 0) e += Mx;
            mr_bstore(rs, (unsigned)s * 8u, Ap[e]);
        }
    }
}

typedef void (*mr_xbd_fn)(float2 *, const float2 *, const int *, long long, mr_dims,
                          mr_fft_len<float>, mr_fdiv, long long, int, int, int);

template <bool MASKED> static mr_xbd_fn mr_xbd_pick(int L)
{
    switch (L) {
#define MR_XBD_CASE(n) case n: return mr_k_xbd<n, MASKED>;
        MR_XBD_CASE(1) MR_XBD_CASE(2) MR_XBD_CASE(3) MR_XBD_CASE(4)
        MR_XBD_CASE(5) MR_XBD_CASE(6) MR_XBD_CASE(7) MR_XBD_CASE(8)
        MR_XBD_CASE(9) MR_XBD_CASE(10) MR_XBD_CASE(11) MR_XBD_CASE(12)
        MR_XBD_CASE(13) MR_XBD_CASE(14) MR_XBD_CASE(15) MR_XBD_CASE(16)
#undef MR_XBD_CASE
    default: return nullptr;
    }
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------

// Optional per-stage timing (MR_AHA_HIP_PROFILE=1, read at every execute):
// an event is recorded on the plan's stream after every stage launch, so the
// interval between consecutive events is that stage's kernel time.
enum { MR_ST_F1, MR_ST_F2, MR_ST_F3, MR_ST_BD, MR_ST_I3, MR_ST_I2, MR_ST_I1, MR_ST_OTHER,
       MR_N_ST };

struct mr_prof {
    std::vector<hipEvent_t> ev;   // ev[0] start of an execute, then one per stage
    std::vector<int> stage;       // stage of ev[i+1] relative to ev[i]; -1 = new execute
    double ms[MR_N_ST] = {0};
    int calls[MR_N_ST] = {0};
};

struct mr_aha_plan_s {
    mr_config cfg;
    mr_dims dims;
    hipStream_t stream;
    int group;
    float2 *bufA;
    float2 *bufB;
    int *d_inv;
    mr_fft_len<float> fx, fy, fz;
    mr_prof prof;
    bool prof_on;
    int xbd_g;          // coils per workgroup of the fused x pass; 0 = separate passes
    size_t xbd_lds;     // its dynamic LDS bytes
    mr_fdiv nxdiv;
    mr_xbd_ct_fn xbd_ct;  // kernel specialised for (Mx, L), or null: mr_k_xbd
    int xbd_nt;           // its workgroup size
    int tile[4];          // pencils per workgroup of Fo, Fy, Iy, Io (0: default rule)
    int no_wide;          // 1: no two-coil 1024-thread fused workgroups
    mr_xbd_ct_fn xbd_mf;  // xbd_ct with the block step on the matrix cores, or null
    mr_xbd_ct_fn xbd_pd;  // persistent form of xbd_ct (one workgroup per CU), or null
    int no_mf;            // 1: keep the FMA block walks (built-in rule; the autotuner may pick xbd_mf)
    bool tuning;          // mr_aha_plan_autotune is timing stages
    int kf;               // kernel format used by execute (MR_KF_*)
    const unsigned *kq;   // 16-bit entries and block scales (kf != FP32)
    const float *ksc;
};

static int mr_prof_mark(mr_aha_plan p, int stage)
{
    if (!p->prof_on) return MR_SUCCESS;
    hipEvent_t e;
    MR_HIP_CHECK(hipEventCreate(&e));
    p->prof.ev.push_back(e);
    p->prof.stage.push_back(stage);
    MR_HIP_CHECK(hipEventRecord(e, p->stream));
    return MR_SUCCESS;
}

// Fold the recorded intervals into the per-stage totals (synchronises).
static int mr_prof_collect(mr_aha_plan p)
{
    mr_prof &f = p->prof;
    if (f.ev.empty()) return MR_SUCCESS;
    MR_HIP_CHECK(hipEventSynchronize(f.ev.back()));
    for (size_t i = 1; i < f.ev.size(); ++i) {
        const int st = f.stage[i];
        if (st < 0) continue;
        float t = 0.f;
        MR_HIP_CHECK(hipEventElapsedTime(&t, f.ev[i - 1], f.ev[i]));
        f.ms[st] += t;
        f.calls[st] += 1;
    }
    for (hipEvent_t e : f.ev) (void)hipEventDestroy(e);
    f.ev.clear();
    f.stage.clear();
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_aha_plan_profile(mr_aha_plan p, double *ms, int *calls, int n)
{
    mr_clear_error();
    MR_REQUIRE(p != nullptr, "mr_aha_plan_profile: plan is NULL");
    MR_TRY(mr_prof_collect(p));
    for (int i = 0; i < n && i < MR_N_ST; ++i) {
        if (ms) ms[i] = p->prof.ms[i];
        if (calls) calls[i] = p->prof.calls[i];
        p->prof.ms[i] = 0.0;
        p->prof.calls[i] = 0;
    }
    return MR_SUCCESS;
}

int mr_push_device(const mr_config *cfg, int *prev)
{
    *prev = -1;
    if (cfg->gpu_device_id < 0) return MR_SUCCESS;
    int cur = 0;
    MR_HIP_CHECK(hipGetDevice(&cur));
    if (cur != cfg->gpu_device_id) {
        int count = 0;
        MR_HIP_CHECK(hipGetDeviceCount(&count));
        MR_REQUIRE(cfg->gpu_device_id < count, "device %d is not visible (%d visible)",
                   cfg->gpu_device_id, count);
        MR_HIP_CHECK(hipSetDevice(cfg->gpu_device_id));
        *prev = cur;
    }
    return MR_SUCCESS;
}

void mr_pop_device(int prev)
{
    if (prev >= 0) (void)hipSetDevice(prev);
}

int mr_mem_budget(const mr_config *cfg, size_t *budget)
{
    size_t fr = 0, tot = 0;
    MR_HIP_CHECK(hipMemGetInfo(&fr, &tot));
    double f = cfg->mem_fraction;
    if (!(f > 0.0)) f = 0.5;
    if (f > 0.95) f = 0.95;
    *budget = (size_t)(f * (double)fr);
    return MR_SUCCESS;
}

static void mr_plan_free_buffers(mr_aha_plan p)
{
    for (hipEvent_t e : p->prof.ev) (void)hipEventDestroy(e);
    p->prof.ev.clear();
    p->prof.stage.clear();
    if (p->bufA) (void)hipFree(p->bufA);
    if (p->bufB) (void)hipFree(p->bufB);
    if (p->d_inv) (void)hipFree(p->d_inv);
    p->bufA = p->bufB = nullptr;
    p->d_inv = nullptr;
}

// Decide whether this plan runs the fused x pass: rank up to 16, an x length
// made of the unrolled radices, and one coil's padded row within the LDS of a
// workgroup. (Mx, L) pairs with a compile-time kernel use it. Default G: the
// most coils that keep a workgroup within 80 KB of LDS, so two workgroups
// share a CU. MR_AHA_HIP_XBD=0 selects the separate passes,
// MR_AHA_HIP_XBD_CT=0 the run-time-length kernel; MR_AHA_HIP_XBD_G sets G.
static int mr_xbd_setup(mr_aha_plan p)
{
    p->xbd_g = 0;
    p->xbd_lds = 0;
    p->xbd_ct = nullptr;
    p->xbd_mf = nullptr;
    p->xbd_pd = nullptr;
    p->xbd_nt = MR_XBD_THREADS;
    const mr_config &cfg = p->cfg;
    const int L = cfg.L, Mx = p->dims.m[0];
    if (mr_env_ll("MR_AHA_HIP_XBD", 1) == 0 || L > 16) return MR_SUCCESS;
    for (int i = 0; i < p->fx.nst; ++i)
        if (!mr_fft_unrolled(p->fx.radix[i])) return MR_SUCCESS;
    const long long per_elems = (long long)L * Mx;
    const size_t per = (size_t)per_elems * sizeof(float2);
    if (per_elems > MR_XBD_EMAX) return MR_SUCCESS;
    // kernel entries are addressed with 32-bit byte offsets: one column, plus
    // up to L - 1 columns of scalar offset
    const long long kcol = cfg.is_masked ? (long long)cfg.num_masked : p->dims.pad_size;
    if (kcol * 8 * L > 0xFFFFFF00LL) return MR_SUCCESS;
    int dev = 0, maxlds = 0;
    MR_HIP_CHECK(hipGetDevice(&dev));
    MR_HIP_CHECK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (per > (size_t)maxlds) return MR_SUCCESS;
    long long gmax = MR_XBD_EMAX / per_elems;
    if (gmax > (long long)(maxlds / per)) gmax = (long long)(maxlds / per);
    const void *fn = nullptr;
    int ct_gmax = 0;
    mr_xbd_ct_fn ct = nullptr;
    if (mr_env_ll("MR_AHA_HIP_XBD_CT", 1) != 0)
        ct = mr_xbd_ct_pick(Mx, L, cfg.is_masked != 0, p->kf, &ct_gmax);
    if (ct) {
        fn = (const void *)ct;
        if (gmax > ct_gmax) gmax = ct_gmax;
    } else if (p->kf != MR_KF_FP32) {
        return MR_SUCCESS;
    } else {
        mr_xbd_fn gf = cfg.is_masked ? mr_xbd_pick<true>(L) : mr_xbd_pick<false>(L);
        if (!gf) return MR_SUCCESS;
        fn = (const void *)gf;
    }
    if (gmax > cfg.n_coils) gmax = cfg.n_coils;
    long long g = (long long)(81920 / per);
    g = mr_env_ll("MR_AHA_HIP_XBD_G", g);
    if (g > gmax) g = gmax;
    if (g < 1) g = 1;
    int nt = MR_XBD_THREADS;
    // Wide form (one coil per 80 KB): 1024 threads walk two coils per
    // workgroup and load each kernel entry once for both
    if (ct && g == 1 && cfg.n_coils >= 2 && 2 * per <= (size_t)maxlds && !p->no_wide &&
        mr_env_ll("MR_AHA_HIP_XBD_WIDE", 1) != 0) {
        mr_xbd_ct_fn wide = mr_xbd_ct_pick_wide(Mx, L, cfg.is_masked != 0, p->kf);
        if (wide) {
            ct = wide;
            fn = (const void *)wide;
            g = 2;
            nt = 1024;
        }
    }
    const size_t lds = per * (size_t)g;
    if (lds > 65536 &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess) {
        (void)hipGetLastError();
        return MR_SUCCESS;
    }
    // Persistent form of the wide workgroups: the image fills the centred
    // half of the x grid (Mx = 2 Nx, centre Nx/2) and its rows are whole
    // 16-byte units, so the next row's image can stream into the LDS slots
    // the current row's result does not occupy
    mr_xbd_ct_fn pd = nullptr;
    if (nt == 1024 && 2 * p->dims.n[0] == Mx && 4 * p->dims.c_n[0] == Mx &&
        p->dims.n[0] % 4 == 0 && mr_env_ll("MR_AHA_HIP_XBD_PD", 1) != 0) {
        pd = mr_xbd_ct_pick_persistent(Mx, L, cfg.is_masked != 0, p->kf);
        if (pd && hipFuncSetAttribute((const void *)pd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds) != hipSuccess) {
            (void)hipGetLastError();
            pd = nullptr;
        }
    }
    // The same kernel with its block step on the matrix cores: FP32 kernel,
    // unmasked, L a multiple of 4 and an even kx centering shift (adjacent
    // locations then have adjacent kernel entries). It runs when
    // MR_AHA_HIP_MFMA=1 was set at plan creation or the autotuner measured it
    // faster; the built-in rule keeps the FMA walks (measured faster at
    // 320x320x56, L=16: the matrix-core step loads every off-diagonal entry
    // twice, once per output it feeds).
    mr_xbd_ct_fn mf = nullptr;
    if (ct && !cfg.is_masked && p->kf == MR_KF_FP32 && L % 4 == 0 && (Mx / 2) % 2 == 0) {
        int mgmax = 0;
        mf = (nt > MR_XBD_THREADS) ? mr_xbd_ct_pick_wide(Mx, L, false, p->kf, true)
                                   : mr_xbd_ct_pick(Mx, L, false, p->kf, &mgmax, true);
        if (mf && lds > 65536 &&
            hipFuncSetAttribute((const void *)mf, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess) {
            (void)hipGetLastError();
            mf = nullptr;
        }
    }
    p->xbd_g = (int)g;
    p->xbd_lds = lds;
    p->xbd_ct = ct;
    p->xbd_mf = mf;
    p->xbd_pd = pd;
    p->xbd_nt = nt;
    p->nxdiv = mr_fdiv_make((unsigned)p->dims.n[0]);
    return MR_SUCCESS;
}

static int mr_plan_create_impl(const mr_config *cfg, mr_aha_plan *out)
{
    MR_TRY(mr_config_validate(cfg));
    mr_aha_plan p = new (std::nothrow) mr_aha_plan_s();
    if (!p) MR_FAIL(MR_ERR_ALLOC, "failed to allocate the plan struct");
    p->cfg = *cfg;
    p->stream = (hipStream_t)cfg->stream;
    // matrix-core block step of the fused pass: opt-in (or the autotuner's pick)
    p->no_mf = mr_env_ll("MR_AHA_HIP_MFMA", 0) == 1 ? 0 : 1;
    mr_dims_from_config(cfg, &p->dims);
    int rc = mr_fft_get<float>(cfg->phi2_dim[0], &p->fx);
    if (rc == MR_SUCCESS) rc = mr_fft_get<float>(cfg->phi2_dim[1], &p->fy);
    if (rc == MR_SUCCESS) rc = mr_fft_get<float>(cfg->phi2_dim[2], &p->fz);
    if (rc == MR_SUCCESS) rc = mr_xbd_setup(p);
    if (rc != MR_SUCCESS) { delete p; return rc; }

    // Work buffers per coil. Fused x pass: A [Nx, Ny, Mz, L] (outer axis
    // transformed; unused in 2D) and B [Nx, My, Mz, L] (the fused pass works in
    // place). Separate passes: two ping-pong grids [Mx, My, Mz, L].
    const mr_dims &d = p->dims;
    size_t a_per, b_per;
    if (p->xbd_g > 0) {
        a_per = (cfg->n_spatial == 2) ? 0 : (size_t)d.n[0] * d.n[1] * d.m[2] * cfg->L;
        b_per = (size_t)d.n[0] * d.m[1] * d.m[2] * cfg->L;
    } else {
        a_per = b_per = (size_t)d.pad_size * cfg->L;
    }
    // Coil group: as many coils as fit in mem_fraction of the free memory
    // (MR_AHA_HIP_COIL_GROUP caps it).
    size_t budget = 0;
    rc = mr_mem_budget(cfg, &budget);
    if (rc != MR_SUCCESS) { delete p; return rc; }
    const long long want = mr_env_ll("MR_AHA_HIP_COIL_GROUP", cfg->n_coils);
    int group = (int)((want < 1) ? 1 : (want > cfg->n_coils ? cfg->n_coils : want));
    const size_t per_coil = (a_per + b_per) * sizeof(float2);
    const size_t inv_bytes = cfg->is_masked ? (size_t)p->dims.pad_size * sizeof(int) : 0;
    while (group > 1 && per_coil * group + inv_bytes > budget) --group;
    for (;; --group) {
        const size_t abytes = a_per * group * sizeof(float2);
        const size_t bbytes = b_per * group * sizeof(float2);
        hipError_t e = hipSuccess;
        if (abytes) e = hipMalloc((void **)&p->bufA, abytes);
        if (e == hipSuccess) e = hipMalloc((void **)&p->bufB, bbytes);
        if (e == hipSuccess && inv_bytes) e = hipMalloc((void **)&p->d_inv, inv_bytes);
        if (e == hipSuccess) break;
        (void)hipGetLastError();
        mr_plan_free_buffers(p);
        if (group == 1) {
            delete p;
            MR_FAIL(MR_ERR_ALLOC, "failed to allocate %zu bytes of work buffers for one coil: %s",
                    abytes + bbytes + inv_bytes, hipGetErrorString(e));
        }
    }
    p->group = group;
    mr_clear_error();
    if (cfg->verbose)
        printf("[mr_aha] plan: %d coil(s) per group, grid [%d %d %d], L = %d\n", group,
               cfg->phi2_dim[0], cfg->phi2_dim[1], cfg->phi2_dim[2], cfg->L);
    *out = p;
    return MR_SUCCESS;
}

int mr_aha_plan_group(mr_aha_plan p) { return p->group; }

// Switch the kernel format the plan's executions read: MR_KF_FP32 (the
// d_phi2_toep argument) or a 16-bit format held at q, sc (mr_kf_quantize).
// A 16-bit format needs the compiled fused x pass for this shape.
int mr_aha_plan_set_kernel_format(mr_aha_plan p, int kf, const unsigned *q, const float *sc)
{
    const int old = p->kf;
    p->kf = kf;
    p->kq = q;
    p->ksc = sc;
    const int rc = mr_xbd_setup(p);
    if (rc != MR_SUCCESS || (kf != MR_KF_FP32 && !p->xbd_ct)) {
        p->kf = old;
        p->kq = nullptr;
        p->ksc = nullptr;
        (void)mr_xbd_setup(p);
        if (rc != MR_SUCCESS) return rc;
        MR_FAIL(MR_ERR_UNSUPPORTED,
                "16-bit kernel: no compiled fused x pass for x grid %d, L = %d%s (available: "
                "x grid 256, 320, 384, 512 or 640 with L 4, 8 or 16, unmasked)",
                p->dims.m[0], p->cfg.L, p->cfg.is_masked ? ", masked" : "");
    }
    return MR_SUCCESS;
}
extern "C" MR_EXPORT int mr_aha_plan_xbd_group(mr_aha_plan p)
{
    mr_clear_error();
    MR_REQUIRE(p != nullptr, "mr_aha_plan_xbd_group: plan is NULL");
    return p->xbd_g;
}
extern "C" MR_EXPORT int mr_aha_plan_xbd_form(mr_aha_plan p)
{
    mr_clear_error();
    MR_REQUIRE(p != nullptr, "mr_aha_plan_xbd_form: plan is NULL");
    if (p->xbd_g == 0) return 0;
    if (p->xbd_ct && p->xbd_mf && !p->no_mf) return 3;
    if (p->xbd_ct && p->xbd_pd) return 2;
    return 1;
}

extern "C" MR_EXPORT int mr_aha_plan_create(const mr_config *cfg, mr_aha_plan *plan_out)
{
    mr_clear_error();
    MR_REQUIRE(plan_out != nullptr, "mr_aha_plan_create: plan_out is NULL");
    *plan_out = nullptr;
    MR_REQUIRE(cfg != nullptr, "mr_aha_plan_create: cfg is NULL");
    int prev = -1;
    int rc = mr_push_device(cfg, &prev);
    if (rc == MR_SUCCESS) rc = mr_plan_create_impl(cfg, plan_out);
    mr_pop_device(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_aha_plan_destroy(mr_aha_plan plan)
{
    mr_clear_error();
    if (!plan) return MR_SUCCESS;
    int prev = -1;
    (void)mr_push_device(&plan->cfg, &prev);
    mr_plan_free_buffers(plan);
    mr_pop_device(prev);
    delete plan;
    return MR_SUCCESS;
}

int mr_bdt_launch(const mr_dims &d, float2 *K, const float2 *ker, const int *inv,
                  long long kstride, int nc, bool masked, hipStream_t s);

static int mr_blockdiag(mr_aha_plan p, float2 *K, int nc, const float2 *ker)
{
    const mr_dims d = p->dims;
    const bool masked = p->cfg.is_masked != 0;
    const long long kstride = masked ? (long long)p->cfg.num_masked : d.pad_size;
    // Tile width (locations) and coil slots per workgroup; MR_AHA_HIP_BD_TP
    // (16, 32, 64) and MR_AHA_HIP_BD_SLOTS override them (read per call).
    // Default slots: the fewest that still give every slot the same number of
    // coils, ceil(nc / ceil(nc / smax)) with smax = 256 / TP, so no slot idles
    // through a last round (12 coils at TP 32: 6 slots of 2 coils each).
    int TP = (int)mr_env_ll("MR_AHA_HIP_BD_TP", 32);
    if (TP != 16 && TP != 64) TP = 32;
    const int smax = 256 / TP;
    int slots = (int)mr_ceil_div(nc, mr_ceil_div(nc, smax));
    slots = (int)mr_env_ll("MR_AHA_HIP_BD_SLOTS", slots);
    if (slots < 1) slots = 1;
    if (slots > nc) slots = nc;
    while (slots > 1 && TP * slots > 256) --slots;
    mr_bd_fn fn = masked ? mr_bd_pick<true>(d.L, TP) : mr_bd_pick<false>(d.L, TP);
    if (fn) {
        const long long blocks = mr_ceil_div(d.pad_size, TP);
        MR_GRID_CHECK(blocks, TP * slots, "mr_k_bd");
        hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(TP * slots), 0, p->stream, K, ker,
                           p->d_inv, kstride, d, nc);
        MR_LAUNCH_CHECK("mr_k_bd");
        return MR_SUCCESS;
    }
    // ranks without a compiled kernel: the tiled kernel (LDS panels)
    return mr_bdt_launch(d, K, ker, p->d_inv, kstride, nc, masked, p->stream);
}

// ---------------------------------------------------------------------------
// Double-buffered outer (z) passes of the x-last route (stack of stars: a
// compiled z length with Nz = Mz), built like mr_fft_axis_dma. A tile is 16
// consecutive xy positions x 8 ranks = 128 pencils of [element][pencil]; its
// source rows (16 xy, one 128-byte line each) stream straight into LDS with
// global_load_lds_dwordx4 while the previous tile is transformed and
// stored. The first radix stage reads the loaded rows and writes the work
// buffer, so the next tile's rows can land as soon as it is done.
// ---------------------------------------------------------------------------

constexpr int MR_ZD_LB = 8;              // ranks per tile
constexpr int MR_ZD_PT = 16 * MR_ZD_LB;  // pencils per tile
// LDS stride of one rank's N rows of 16 xy, padded by one row so that the
// four ranks a wavefront reads at once start on alternate bank halves
template <int N> constexpr int mr_zd_xs() { return N * 16 + 16; }

// Forward: element e of pencil (rank l, xy) is S_c x_l at image z = (e + cnz)
// mod N (the centering of mr_ld_sx_z). Per tile (xy tile q, rank block lb)
// the coils in turn: the tile's x rows are loaded once and serve every coil;
// coil c's S_c rows stream in while coil c-1 is transformed and stored (two
// LDS buffers), and the next tile's x rows while the last coil is. Output
// [plane, N, V] (v = l + L c) as mr_st_z.
template <int N>
__global__ void __launch_bounds__(1024)
mr_k_fo_dma(const float2 *tws, const float2 *__restrict__ x, const float2 *__restrict__ S,
            float2 *__restrict__ out, long long NxNy, long long Nimg, int L, int nc, int coil0,
            int cnz, long long ntiles)
{
    constexpr int NT = 1024, NW = NT / 64, PT = MR_ZD_PT, XS = mr_zd_xs<N>(), NG = N / 8;
    static_assert(N % 8 == 0, "rows load in groups of 8");
    __shared__ __align__(16) float2 Bx[MR_ZD_LB * XS];   // [rank][z][xy]
    __shared__ __align__(16) float2 Bs0[N * 16];         // [z][xy], even coils
    __shared__ __align__(16) float2 Bs1[N * 16];         // odd coils
    __shared__ __align__(16) float2 A[N * PT];           // [e][rank * 16 + xy]
    float2 w[MR_EP_NST][MR_EP_KMAX];
    mr_ep_twload<N, 0, NT, PT>(w, mr_tw_buf{mr_rsrc(tws, (unsigned)N * 8u)});
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nlb = L / MR_ZD_LB;
    const int col = (lane & 7) * 2, row = lane >> 3;
    // the x rows of tile t (8 ranks x N rows), and/or the S_c rows
    auto fill = [&](long long t, int c, float2 *bs, bool with_x) {
        const int lb = (int)(t % nlb);
        const long long xy0 = (t / nlb) * 16 + col;
        const int g0 = with_x ? 0 : MR_ZD_LB * NG;
        for (int gi = g0 + wave; gi < (MR_ZD_LB + 1) * NG; gi += NW) {
            const int rk = gi / NG, z0 = (gi - rk * NG) * 8;   // rk == 8: the S_c rows
            if (rk < MR_ZD_LB)
                __builtin_amdgcn_global_load_lds(
                    x + xy0 + NxNy * (z0 + row) + Nimg * (lb * MR_ZD_LB + rk),
                    (__attribute__((address_space(3))) void *)(Bx + rk * XS + z0 * 16), 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    S + xy0 + NxNy * (z0 + row) + Nimg * (coil0 + c),
                    (__attribute__((address_space(3))) void *)(bs + z0 * 16), 16, 0, 0);
        }
    };
    const int p = threadIdx.x & (PT - 1);
    auto coil = [&](long long t, int c, const float2 *bcur, float2 *bnext) {
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's row loads
        __syncthreads();                 // everyone's
        mr_ep_stage<N, 0, false, NT, PT, true>(A, w[0], [&](int pp, int e) {
            int z = e + cnz;
            if (z >= N) z -= N;
            const int xy = pp & 15;
            return cmul(bcur[z * 16 + xy], Bx[(pp >> 4) * XS + z * 16 + xy]);
        });
        if (c + 1 < nc) fill(t, c + 1, bnext, false);
        else if (t + gridDim.x < ntiles) fill(t + gridDim.x, 0, bnext, true);
        mr_ep_fft<N, 1, false, NT, PT>(A, w);
        const long long v = (t % nlb) * MR_ZD_LB + (p >> 4) + (long long)L * c;
        float2 *o = out + (t / nlb) * 16 + (p & 15) + NxNy * (long long)N * v;
        for (int e = threadIdx.x / PT; e < N; e += NT / PT) o[NxNy * e] = A[e * PT + p];
        mr_lds_barrier();                // A and bcur are read out before they are refilled
    };
    long long t = blockIdx.x;
    if (t >= ntiles) return;
    fill(t, 0, Bs0, true);
    bool odd = false;   // the current coil's S rows are in Bs1
    for (; t < ntiles; t += gridDim.x) {
        for (int c = 0; c < nc; ++c) {
            if (!odd) coil(t, c, Bs0, Bs1);
            else coil(t, c, Bs1, Bs0);
            odd = !odd;
        }
    }
}

// Inverse: per tile (xy tile q, rank block lb) the coils in turn; coil c's
// rows [plane, N, V] (v = l + L c) and its S_c rows stream in while coil c-1
// is transformed. The output element e is image z = (e + cnz) mod N; each
// work-item keeps the coil sum of conj(S_c) times its elements in registers
// (the order of mr_st_z_conj_sum) and writes AhAx[xy, z, l] (added to it when
// `accumulate`) after the last coil. The S rows of consecutive coils
// alternate between two LDS buffers.
template <int N>
__global__ void __launch_bounds__(1024)
mr_k_io_dma(const float2 *tws, const float2 *__restrict__ in, const float2 *__restrict__ S,
            float2 *__restrict__ out, long long NxNy, long long Nimg, int L, int nc, int coil0,
            int cnz, int accumulate, long long ntiles)
{
    constexpr int NT = 1024, NW = NT / 64, PT = MR_ZD_PT, XS = mr_zd_xs<N>(), NG = N / 8;
    constexpr int KE = N * PT / NT;   // elements per work-item
    static_assert(N % 8 == 0 && (N * PT) % NT == 0, "tile shape");
    __shared__ __align__(16) float2 Bx[MR_ZD_LB * XS];   // [rank][e][xy]
    __shared__ __align__(16) float2 Bs0[N * 16];         // [z][xy], even coils
    __shared__ __align__(16) float2 Bs1[N * 16];         // odd coils
    __shared__ __align__(16) float2 W[N * PT];           // [e][rank * 16 + xy]
    float2 w[MR_EP_NST][MR_EP_KMAX];
    mr_ep_twload<N, 0, NT, PT>(w, mr_tw_buf{mr_rsrc(tws, (unsigned)N * 8u)});
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nlb = L / MR_ZD_LB;
    const int p = threadIdx.x & (PT - 1), xy = p & 15, rl = p >> 4;
    const int e0 = threadIdx.x >> 7;   // elements e0 + 8k, k < KE
    auto fill = [&](long long t, int c, float2 *bs) {
        const int lb = (int)(t % nlb);
        const long long xy0 = (t / nlb) * 16 + (lane & 7) * 2;
        const int row = lane >> 3;
        for (int gi = wave; gi < (MR_ZD_LB + 1) * NG; gi += NW) {
            const int rk = gi / NG, z0 = (gi - rk * NG) * 8;   // rk == 8: the S_c rows
            if (rk < MR_ZD_LB)
                __builtin_amdgcn_global_load_lds(
                    in + xy0 + NxNy * (z0 + row) +
                        NxNy * (long long)N * (lb * MR_ZD_LB + rk + (long long)L * c),
                    (__attribute__((address_space(3))) void *)(Bx + rk * XS + z0 * 16), 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds(
                    S + xy0 + NxNy * (z0 + row) + Nimg * (coil0 + c),
                    (__attribute__((address_space(3))) void *)(bs + z0 * 16), 16, 0, 0);
        }
    };
    // one coil of tile t: first stage from the loaded rows, then the next
    // (tile, coil)'s rows start loading, then the other stages and the coil
    // sum with this coil's S rows
    float2 acc[KE];
    auto coil = [&](long long t, int c, const float2 *bcur, float2 *bnext) {
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's row loads
        __syncthreads();                 // everyone's
        mr_ep_stage<N, 0, true, NT, PT, true>(W, w[0], [&](int pp, int e) {
            return Bx[(pp >> 4) * XS + e * 16 + (pp & 15)];
        });
        long long tn = t;
        int cn = c + 1;
        if (cn == nc) {
            cn = 0;
            tn = t + gridDim.x;
        }
        if (tn < ntiles) fill(tn, cn, bnext);
        mr_ep_fft<N, 1, true, NT, PT>(W, w);
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            int z = e0 + 8 * k + cnz;
            if (z >= N) z -= N;
            const float2 v = cmulc(bcur[z * 16 + xy], W[(e0 + 8 * k) * PT + p]);
            acc[k] = (c == 0) ? v : cadd(acc[k], v);
        }
        mr_lds_barrier();                // W and bcur are read out before they are refilled
    };
    long long t = blockIdx.x;
    if (t >= ntiles) return;
    fill(t, 0, Bs0);
    bool odd = false;   // the current coil's S rows are in Bs1
    for (; t < ntiles; t += gridDim.x) {
        for (int c = 0; c < nc; ++c) {
            if (!odd) coil(t, c, Bs0, Bs1);
            else coil(t, c, Bs1, Bs0);
            odd = !odd;
        }
        const long long q = t / nlb;
        const int l = (int)(t % nlb) * MR_ZD_LB + rl;
        float2 *o = out + q * 16 + xy + Nimg * l;
        if (accumulate) {   // a later coil group: add to the earlier groups' sum
#pragma unroll
            for (int k = 0; k < KE; ++k) {
                int z = e0 + 8 * k + cnz;
                if (z >= N) z -= N;
                o[NxNy * z] = cadd(o[NxNy * z], acc[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < KE; ++k) {
                int z = e0 + 8 * k + cnz;
                if (z >= N) z -= N;
                o[NxNy * z] = acc[k];
            }
        }
    }
}

// Run Fo / Io on the double-buffered kernels when the shape allows (z length
// 56 = Mz = Nz, whole 16-position xy tiles, ranks in blocks of 8, 16-byte
// aligned arrays); false: use the one-tile passes.
static bool mr_zd_ok(const mr_dims &d, int L, long long plane, const void *a, const void *b,
                     const void *c)
{
    return d.m[2] == 56 && d.n[2] == 56 && plane % 16 == 0 && L % MR_ZD_LB == 0 &&
           d.image_size % 2 == 0 && (((unsigned long long)a | (unsigned long long)b |
                                      (unsigned long long)c) & 15) == 0 &&
           mr_env_ll("MR_AHA_HIP_ZDMA", 1) != 0;
}

static long long mr_zd_grid(long long ntiles)
{
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess || ncu <= 0)
            ncu = 256;
    }
    return ntiles < ncu ? ntiles : ncu;
}

// x-last pipeline for one coil group (coils coil0 .. coil0 + nc - 1):
//   Fo   outer axis (z; y in 2D): S_c x_l, zero-pad, FFT      -> A [Nx, Ny, Mz, V]
//   Fy   y-FFT, zero-padding the Ny rows (3D and stack)      -> B [Nx, My, Mz, V]
//   XBD  fused x-FFT, block multiply, inverse x-FFT, in place on B
//   Iy   inverse y-FFT, keeping the Ny image rows            -> A
//   Io   inverse outer FFT, crop, conj(S_c), coil sum         -> out
// In 2D the outer axis is y itself: Fo writes B [Nx, My, V] and Io reads it.
static int mr_execute_xlast(mr_aha_plan p, const float2 *x, const float2 *S, const float2 *ker,
                            float2 *out, int coil0, int nc)
{
    const mr_config &cfg = p->cfg;
    const mr_dims d = p->dims;
    hipStream_t s = p->stream;
    const int L = cfg.L;
    const int Nx = d.n[0], Ny = d.n[1], Nz = d.n[2];
    const int My = d.m[1], Mz = d.m[2];
    const long long N = d.image_size;
    const long long V = (long long)L * nc;
    const bool two = cfg.n_spatial == 2;
    const long long plane = two ? (long long)Nx : (long long)Nx * Ny;
    const int No = two ? Ny : Nz, Mo = two ? My : Mz, cno = two ? d.c_n[1] : d.c_n[2];
    const mr_fft_len<float> &fo = two ? p->fy : p->fz;
    int lgT = 0;
    while (lgT < 6 && (plane % (2LL << lgT)) == 0) ++lgT;
    float2 *A = two ? p->bufB : p->bufA, *B = p->bufB;
    {
        mr_zmap zm{V, lgT};
        mr_ld_sx_z ld{x, S, zm, plane, N, No, Mo, cno, L, coil0};
        mr_st_z st{A, zm, plane, Mo};
        // pencils per workgroup: the autotuned width, else 2^lgT up to 64;
        // MR_AHA_HIP_FO_TX overrides both, and likewise for the other passes
        // below
        const int txo = (int)mr_env_ll("MR_AHA_HIP_FO_TX", p->tile[0] > 0 ? p->tile[0] : 1 << lgT);
        if (!two && mr_zd_ok(d, L, plane, x, S, A)) {
            const long long nt = (plane / 16) * (L / MR_ZD_LB);
            hipLaunchKernelGGL(mr_k_fo_dma<56>, dim3((unsigned)mr_zd_grid(nt)), dim3(1024), 0, s,
                               fo.tws, x, S, A, plane, N, L, nc, coil0, cno, nt);
            MR_LAUNCH_CHECK("mr_k_fo_dma");
        } else {
            MR_TRY((mr_fft_pass_ct<false>(fo, ld, st, plane * V, 0, txo, s, "Fo")));
        }
        MR_TRY(mr_prof_mark(p, two ? MR_ST_F2 : MR_ST_F3));
    }
    if (!two) {
        mr_ld_pad<float2, false> ld{A, Nx, My, Ny, d.c_n[1]};
        mr_st_axis<float2, false> st{B, Nx, My};
        const int txy = (int)mr_env_ll("MR_AHA_HIP_FY_TX", p->tile[1] > 0 ? p->tile[1] : 16);
        MR_TRY((mr_fft_pass_ct<false>(p->fy, ld, st, (long long)Nx * Mz * V, 0, txy, s, "Fy")));
        MR_TRY(mr_prof_mark(p, MR_ST_F2));
    }
    {
        const bool masked = cfg.is_masked != 0;
        const long long kstride = masked ? (long long)cfg.num_masked : d.pad_size;
        const long long R = (long long)My * Mz;
        const int G = p->xbd_g;
        const int ncg = (nc + G - 1) / G;
        const long long blocks = R * ncg;
        if (blocks > 0x7fffffffLL)
            MR_FAIL(MR_ERR_UNSUPPORTED, "fused x pass: %lld workgroups exceed the launch limit",
                    blocks);
        MR_GRID_CHECK(blocks, p->xbd_ct ? p->xbd_nt : MR_XBD_THREADS, "fused x pass");
        if (p->xbd_ct) {
            // the matrix-core block step reads kernel entries two locations
            // per 16-byte load
            const bool mf = p->xbd_mf && !p->no_mf && ((unsigned long long)ker & 15) == 0;
            // the persistent form walks all items with one workgroup per CU
            const bool pd = !mf && p->xbd_pd && ((unsigned long long)B & 15) == 0;
            const mr_xbd_ct_fn fn = mf ? p->xbd_mf : pd ? p->xbd_pd : p->xbd_ct;
            const long long grid = pd ? mr_zd_grid(blocks) : blocks;
            hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3((unsigned)p->xbd_nt), p->xbd_lds,
                               s, B, ker, p->d_inv, kstride, d, p->fx.tws, R, G, nc, ncg, p->kq,
                               p->ksc);
            MR_LAUNCH_CHECK("mr_k_xbd_ct");
        } else {
            mr_xbd_fn fn = masked ? mr_xbd_pick<true>(L) : mr_xbd_pick<false>(L);
            hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(MR_XBD_THREADS), p->xbd_lds, s, B,
                               ker, p->d_inv, kstride, d, p->fx, p->nxdiv, R, G, nc, ncg);
            MR_LAUNCH_CHECK("mr_k_xbd");
        }
        MR_TRY(mr_prof_mark(p, MR_ST_BD));
    }
    if (!two) {
        mr_ld_axis<float2, false> ld{B, Nx, My};
        mr_st_crop<float2, false> st{A, Nx, My, Ny, d.c_n[1]};
        const int txy = (int)mr_env_ll("MR_AHA_HIP_IY_TX", p->tile[2] > 0 ? p->tile[2] : 16);
        MR_TRY((mr_fft_pass_ct<true>(p->fy, ld, st, (long long)Nx * Mz * V, 0, txy, s, "Iy")));
        MR_TRY(mr_prof_mark(p, MR_ST_I2));
    }
    {
        mr_zmap zm{L, lgT};
        mr_ld_z_coil ld{A, zm, plane, Mo, L, nc};
        mr_st_z_conj_sum st{out, S, zm, plane, N, No, Mo, cno, coil0, coil0 > 0 ? 1 : 0};
        // 32 pencils: the coil-sum accumulator makes 64 cost two workgroups per CU
        const int txi = (int)mr_env_ll("MR_AHA_HIP_IO_TX",
                                       p->tile[3] > 0 ? p->tile[3] : 1 << (lgT < 5 ? lgT : 5));
        if (!two && mr_zd_ok(d, L, plane, A, S, out)) {
            const long long nt = (plane / 16) * (L / MR_ZD_LB);
            hipLaunchKernelGGL(mr_k_io_dma<56>, dim3((unsigned)mr_zd_grid(nt)), dim3(1024), 0, s,
                               fo.tws, A, S, out, plane, N, L, nc, coil0, cno, coil0 > 0 ? 1 : 0,
                               nt);
            MR_LAUNCH_CHECK("mr_k_io_dma");
        } else {
            MR_TRY((mr_fft_pass_ct<true>(fo, ld, st, plane * L, (size_t)No, txi, s, "Io")));
        }
        MR_TRY(mr_prof_mark(p, two ? MR_ST_I2 : MR_ST_I3));
    }
    return MR_SUCCESS;
}

static int mr_execute_impl(mr_aha_plan p, const float2 *x, const float2 *S, const float2 *ker,
                           const int64_t *mask, float2 *out)
{
    const mr_config &cfg = p->cfg;
    const mr_dims d = p->dims;
    hipStream_t s = p->stream;
    const int L = cfg.L;
    const int Nx = d.n[0], Ny = d.n[1], Nz = d.n[2];
    const int Mx = d.m[0], My = d.m[1], Mz = d.m[2];
    const long long N = d.image_size;
    const long long NyNz = (long long)Ny * Nz;
    const bool three = cfg.n_spatial == 3;
    p->prof_on = p->tuning || mr_env_ll("MR_AHA_HIP_PROFILE", 0) != 0;
    MR_TRY(mr_prof_mark(p, -1));

    if (cfg.is_masked) {
        const long long P = d.pad_size, nm = cfg.num_masked;
        hipLaunchKernelGGL(mr_k_fill_int, dim3((unsigned)mr_ceil_div(P, 256)), dim3(256), 0, s,
                           p->d_inv, P, -1);
        MR_LAUNCH_CHECK("mr_k_fill_int");
        hipLaunchKernelGGL(mr_k_mask_inverse, dim3((unsigned)mr_ceil_div(nm, 256)), dim3(256), 0,
                           s, p->d_inv, mask, nm, d);
        MR_LAUNCH_CHECK("mr_k_mask_inverse");
        MR_TRY(mr_prof_mark(p, MR_ST_OTHER));
    }

    const bool zfirst = cfg.traj_type == MR_TRAJ_STACK && Mz == Nz;
    const long long NxNy = (long long)Nx * Ny;
    int lgTXZ = 0;
    while (lgTXZ < 6 && (NxNy % (2LL << lgTXZ)) == 0) ++lgTXZ;

    for (int coil0 = 0; coil0 < cfg.n_coils; coil0 += p->group) {
        const int nc = (cfg.n_coils - coil0 < p->group) ? cfg.n_coils - coil0 : p->group;
        if (p->xbd_g > 0) {
            MR_TRY(mr_execute_xlast(p, x, S, ker, out, coil0, nc));
            continue;
        }
        const long long V = (long long)L * nc;
        float2 *A = p->bufA, *B = p->bufB;

        if (zfirst) {
            // Z1: S_c x_l, z-FFT on the image -> A [Nx, Ny, Nz(kz), V]
            mr_zmap zm{V, lgTXZ};
            mr_ld_sx_z ld{x, S, zm, NxNy, N, Nz, Nz, d.c_n[2], L, coil0};
            mr_st_z st{A, zm, NxNy, Nz};
            MR_TRY((mr_fft_pass_ct<false>(p->fz, ld, st, NxNy * V, 0, 1 << lgTXZ, s, "Z1")));
            MR_TRY(mr_prof_mark(p, MR_ST_F3));
            // F1: zero-padded x-FFT -> B [Mx, Ny, Nz, V]
            mr_ld_pad<float2, true> l1{A, 1, Mx, Nx, d.c_n[0]};
            mr_st_axis<float2, true> s1{B, 1, Mx};
            MR_TRY((mr_fft_pass_ct<false>(p->fx, l1, s1, NyNz * V, 0, 8, s, "F1")));
            MR_TRY(mr_prof_mark(p, MR_ST_F1));
            std::swap(A, B);  // x-transformed rows now in A
        } else {
            // F1: pad * S, x-FFT -> A [Mx, Ny, Nz, V]
            mr_ld_pad_sx ld{x, S, N, NyNz, V, Nx, Mx, d.c_n[0], L, coil0};
            mr_st_rows_vfast st{A, NyNz, V, Mx};
            MR_TRY((mr_fft_pass_ct<false>(p->fx, ld, st, NyNz * V, 0, 8, s, "F1")));
            MR_TRY(mr_prof_mark(p, MR_ST_F1));
        }
        // F2: y-FFT -> B [Mx, My, Nz, V]
        {
            mr_ld_pad<float2, false> ld{A, Mx, My, Ny, d.c_n[1]};
            mr_st_axis<float2, false> st{B, Mx, My};
            MR_TRY((mr_fft_pass_ct<false>(p->fy, ld, st, (long long)Mx * Nz * V, 0, 16, s, "F2")));
            MR_TRY(mr_prof_mark(p, MR_ST_F2));
        }
        float2 *K = B, *free_buf = A;
        if (three && !zfirst) {
            // F3: z-FFT -> A [Mx, My, Mz, V]
            mr_ld_pad<float2, false> ld{B, (long long)Mx * My, Mz, Nz, d.c_n[2]};
            mr_st_axis<float2, false> st{A, (long long)Mx * My, Mz};
            MR_TRY((mr_fft_pass_ct<false>(p->fz, ld, st, (long long)Mx * My * V, 0, 64, s, "F3")));
            MR_TRY(mr_prof_mark(p, MR_ST_F3));
            K = A;
            free_buf = B;
        }
        MR_TRY(mr_blockdiag(p, K, nc, ker));
        MR_TRY(mr_prof_mark(p, MR_ST_BD));
        float2 *src = K;
        if (three && !zfirst) {
            // I3: inverse z-FFT, keep Nz planes -> free_buf [Mx, My, Nz, V]
            mr_ld_axis<float2, false> ld{K, (long long)Mx * My, Mz};
            mr_st_crop<float2, false> st{free_buf, (long long)Mx * My, Mz, Nz, d.c_n[2]};
            MR_TRY((mr_fft_pass_ct<true>(p->fz, ld, st, (long long)Mx * My * V, 0, 64, s, "I3")));
            MR_TRY(mr_prof_mark(p, MR_ST_I3));
            src = free_buf;
            free_buf = K;
        }
        // I2: inverse y-FFT, keep Ny rows -> [Mx, Ny, Nz, V]
        float2 *dst = (src == A) ? B : A;
        {
            mr_ld_axis<float2, false> ld{src, Mx, My};
            mr_st_crop<float2, false> st{dst, Mx, My, Ny, d.c_n[1]};
            MR_TRY((mr_fft_pass_ct<true>(p->fy, ld, st, (long long)Mx * Nz * V, 0, 16, s, "I2")));
            MR_TRY(mr_prof_mark(p, MR_ST_I2));
        }
        if (zfirst) {
            // I1: inverse x-FFT, crop -> src [Nx, Ny, Nz, V]
            mr_ld_axis<float2, true> l1{dst, 1, Mx};
            mr_st_crop<float2, true> s1{src, 1, Mx, Nx, d.c_n[0]};
            MR_TRY((mr_fft_pass_ct<true>(p->fx, l1, s1, NyNz * V, 0, 8, s, "I1")));
            MR_TRY(mr_prof_mark(p, MR_ST_I1));
            // Z2: inverse z-FFT, conj(S_c), coil sum -> out
            mr_zmap zm{L, lgTXZ};
            mr_ld_z_coil ld{src, zm, NxNy, Nz, L, nc};
            mr_st_z_conj_sum st{out, S, zm, NxNy, N, Nz, Nz, d.c_n[2], coil0, coil0 > 0 ? 1 : 0};
            MR_TRY((mr_fft_pass_ct<true>(p->fz, ld, st, NxNy * L, (size_t)Nz, 1 << lgTXZ, s, "Z2")));
            MR_TRY(mr_prof_mark(p, MR_ST_I3));
        } else {
            // I1: inverse x-FFT, crop, conj(S) and coil sum -> out
            mr_ld_row_coil ld{dst, NyNz, Mx, L, nc};
            mr_st_crop_conj_sum st{out, S, N, NyNz, Nx, Mx, d.c_n[0], coil0, coil0 > 0 ? 1 : 0, L};
            MR_TRY((mr_fft_pass_ct<true>(p->fx, ld, st, NyNz * L, (size_t)Nx, 8, s, "I1")));
            MR_TRY(mr_prof_mark(p, MR_ST_I1));
        }
    }
    if (cfg.verbose) {
        MR_HIP_CHECK(hipStreamSynchronize(s));
        printf("[mr_aha] AhA applied over %d coil(s) in groups of %d, grid [%d %d %d] -> "
               "[%d %d %d], L = %d%s\n",
               cfg.n_coils, p->group, Nx, Ny, Nz, Mx, My, Mz, L, cfg.is_masked ? ", masked" : "");
    }
    return MR_SUCCESS;
}

int mr_aha_plan_execute_raw(mr_aha_plan plan, const mr_complex *d_x, const mr_complex *d_SEs,
                            const mr_complex *d_phi2_toep, const int64_t *d_mask,
                            mr_complex *d_AhAx)
{
    return mr_execute_impl(plan, reinterpret_cast<const float2 *>(d_x),
                           reinterpret_cast<const float2 *>(d_SEs),
                           reinterpret_cast<const float2 *>(d_phi2_toep), d_mask,
                           reinterpret_cast<float2 *>(d_AhAx));
}

extern "C" MR_EXPORT int mr_aha_plan_execute(mr_aha_plan plan, const mr_complex *d_x,
                                             const mr_complex *d_SEs,
                                             const mr_complex *d_phi2_toep, const int64_t *d_mask,
                                             mr_complex *d_AhAx)
{
    mr_clear_error();
    MR_REQUIRE(plan != nullptr, "mr_aha_plan_execute: plan is NULL");
    MR_REQUIRE(d_x != nullptr, "mr_aha_plan_execute: d_x is NULL");
    MR_REQUIRE(d_SEs != nullptr, "mr_aha_plan_execute: d_SEs is NULL");
    MR_REQUIRE(d_phi2_toep != nullptr, "mr_aha_plan_execute: d_phi2_toep is NULL");
    MR_REQUIRE(d_AhAx != nullptr, "mr_aha_plan_execute: d_AhAx is NULL");
    MR_REQUIRE((const void *)d_AhAx != (const void *)d_x,
               "mr_aha_plan_execute: d_AhAx must not alias d_x");
    if (plan->cfg.is_masked)
        MR_REQUIRE(d_mask != nullptr,
                   "mr_aha_plan_execute: d_mask is required when is_masked is set");
    int prev = -1;
    int rc = mr_push_device(&plan->cfg, &prev);
    if (rc == MR_SUCCESS)
        rc = mr_aha_plan_execute_raw(plan, d_x, d_SEs, d_phi2_toep, d_mask, d_AhAx);
    mr_pop_device(prev);
    return rc;
}

// Autotuner of the fused route: for each axis pass (Fo, Fy, Iy, Io; 2D: Fo,
// Io) the tile widths 1..64 pencils are timed in turn on the caller's data
// (`reps` applies each, HIP events per stage) and the fastest is kept; then
// the fused x pass's two-coil wide form against one coil per workgroup.
// Every candidate computes the same arithmetic per pencil, so results do not
// depend on the choice. The per-stage profile totals are cleared.
static int mr_autotune_impl(mr_aha_plan p, const mr_complex *x, const mr_complex *S,
                            const mr_complex *ker, const int64_t *mask, mr_complex *out, int reps)
{
    if (p->xbd_g == 0) return MR_SUCCESS;  // separate passes: fixed tiling
    const bool two = p->cfg.n_spatial == 2;
    auto zero_prof = [&]() -> int {
        MR_TRY(mr_prof_collect(p));
        for (int i = 0; i < MR_N_ST; ++i) {
            p->prof.ms[i] = 0.0;
            p->prof.calls[i] = 0;
        }
        return MR_SUCCESS;
    };
    // mean ms of one stage over `reps` applies
    auto stage_ms = [&](int stage, double *ms) -> int {
        MR_TRY(zero_prof());
        for (int r = 0; r < reps; ++r) MR_TRY(mr_aha_plan_execute_raw(p, x, S, ker, mask, out));
        MR_TRY(mr_prof_collect(p));
        *ms = p->prof.calls[stage] > 0 ? p->prof.ms[stage] / p->prof.calls[stage] : 0.0;
        return MR_SUCCESS;
    };
    p->tuning = true;
    int rc = mr_aha_plan_execute_raw(p, x, S, ker, mask, out);  // warm-up
    const int idx[4] = {0, 1, 2, 3};
    const int stg[4] = {two ? MR_ST_F2 : MR_ST_F3, MR_ST_F2, MR_ST_I2, two ? MR_ST_I2 : MR_ST_I3};
    for (int k = 0; k < 4 && rc == MR_SUCCESS; ++k) {
        const int t = idx[k];
        if (two && (t == 1 || t == 2)) continue;  // 2D: no separate y passes
        int best = p->tile[t];
        double best_ms = 1e30;
        for (int cand = 1; cand <= 64 && rc == MR_SUCCESS; cand *= 2) {
            p->tile[t] = cand;
            double ms = 0.0;
            rc = stage_ms(stg[t], &ms);
            if (rc == MR_SUCCESS && ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
        p->tile[t] = best;
    }
    if (rc == MR_SUCCESS && p->xbd_nt > MR_XBD_THREADS) {
        double wide_ms = 0.0, narrow_ms = 0.0;
        rc = stage_ms(MR_ST_BD, &wide_ms);
        if (rc == MR_SUCCESS) {
            p->no_wide = 1;
            rc = mr_xbd_setup(p);
            if (rc == MR_SUCCESS) rc = stage_ms(MR_ST_BD, &narrow_ms);
            if (rc == MR_SUCCESS && narrow_ms >= wide_ms) {
                p->no_wide = 0;
                rc = mr_xbd_setup(p);
            }
        }
    }
    // block step on the matrix cores against the FMA walks
    if (rc == MR_SUCCESS && p->xbd_mf) {
        double mf_ms = 0.0, fma_ms = 0.0;
        p->no_mf = 0;
        rc = stage_ms(MR_ST_BD, &mf_ms);
        if (rc == MR_SUCCESS) {
            p->no_mf = 1;
            rc = stage_ms(MR_ST_BD, &fma_ms);
            if (rc == MR_SUCCESS && mf_ms < fma_ms) p->no_mf = 0;
        }
    }
    p->tuning = false;
    const int rc2 = zero_prof();
    return rc != MR_SUCCESS ? rc : rc2;
}

extern "C" MR_EXPORT int mr_aha_plan_autotune(mr_aha_plan plan, const mr_complex *d_x,
                                              const mr_complex *d_SEs,
                                              const mr_complex *d_phi2_toep,
                                              const int64_t *d_mask, mr_complex *d_AhAx, int reps)
{
    mr_clear_error();
    MR_REQUIRE(plan != nullptr, "mr_aha_plan_autotune: plan is NULL");
    MR_REQUIRE(d_x != nullptr && d_SEs != nullptr && d_phi2_toep != nullptr && d_AhAx != nullptr,
               "mr_aha_plan_autotune: NULL array");
    MR_REQUIRE((const void *)d_AhAx != (const void *)d_x,
               "mr_aha_plan_autotune: d_AhAx must not alias d_x");
    MR_REQUIRE(reps >= 1 && reps <= 100, "mr_aha_plan_autotune: reps %d is not in 1..100", reps);
    if (plan->cfg.is_masked)
        MR_REQUIRE(d_mask != nullptr,
                   "mr_aha_plan_autotune: d_mask is required when is_masked is set");
    int prev = -1;
    int rc = mr_push_device(&plan->cfg, &prev);
    if (rc == MR_SUCCESS) rc = mr_autotune_impl(plan, d_x, d_SEs, d_phi2_toep, d_mask, d_AhAx, reps);
    mr_pop_device(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_aha_plan_tiles(mr_aha_plan plan, int *tiles, int n)
{
    mr_clear_error();
    MR_REQUIRE(plan != nullptr, "mr_aha_plan_tiles: plan is NULL");
    MR_REQUIRE(tiles != nullptr || n == 0, "mr_aha_plan_tiles: tiles is NULL");
    const int v[6] = {plan->tile[0], plan->tile[1], plan->tile[2], plan->tile[3],
                      plan->xbd_nt > MR_XBD_THREADS ? 1 : 0,
                      (plan->xbd_mf && !plan->no_mf) ? 1 : 0};
    for (int i = 0; i < n && i < 6; ++i) tiles[i] = v[i];
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_aha_toeplitz(const mr_config *cfg, const mr_complex *d_x,
                                         const mr_complex *d_SEs, const mr_complex *d_phi2_toep,
                                         const int64_t *d_mask, mr_complex *d_AhAx)
{
    mr_aha_plan plan = nullptr;
    int rc = mr_aha_plan_create(cfg, &plan);
    if (rc != MR_SUCCESS) return rc;
    rc = mr_aha_plan_execute(plan, d_x, d_SEs, d_phi2_toep, d_mask, d_AhAx);
    char msg[512];
    snprintf(msg, sizeof(msg), "%s", mr_last_error());
    const int rd = mr_aha_plan_destroy(plan);
    if (rc != MR_SUCCESS) {
        mr_set_error("%s", msg);
        return rc;
    }
    return rd;
}
