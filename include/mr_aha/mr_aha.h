/*
 * mr_aha.h — C ABI of the MI355X (gfx950) Toeplitz A^H A library.
 *
 * Drop-in for the single-GPU `method = "toeplitz"` path of the reference
 * library: every type, field, enumerator value and function signature below
 * has the same layout and meaning as the reference's public header, so a
 * binding written against it (ctypes, MEX, C) links against
 * `libmr_aha_c.so` unchanged. Each declaration names the reference
 * declaration it stands in for (include/mr_aha/mr_aha.h:<line>).
 *
 * Conventions (identical to the reference):
 *  - Arrays are column-major; a row-major caller reverses the dimension
 *    order, the memory is the same ([Nx, Ny, Nz, L] == numpy (L, Nz, Ny, Nx)).
 *  - Image and k-space arrays are centered: index floor(M/2) is the origin.
 *  - A Gram block is Phi2 = Phi_T Phi_T^H for a rank x time basis Phi; every
 *    block multiply treats the rank vector v of one location as a row vector
 *    multiplied from the right, v <- v Phi2.
 *  - Every array argument is a device pointer (HIP device memory on the
 *    calling device); the caller allocates inputs and outputs.
 *  - Entry points return 0 on success or a negative mr_status, and record a
 *    message retrievable with mr_last_error() on the calling thread.
 *
 * MI355X notes: `stream` carries a hipStream_t; status MR_ERR_CUDA (-4)
 * reports a HIP runtime failure and MR_ERR_CUFFT (-5) a failure of the
 * library's own FFT engine (the library calls no vendor FFT). mr_backend()
 * returns "cuda", the string that makes the reference's bindings pass device
 * pointers ("device-pointer ABI").
 */
#ifndef MR_AHA_MR_AHA_H
#define MR_AHA_MR_AHA_H

#include <stddef.h>
#include <stdint.h>

#ifdef _WIN32
#  ifdef MR_AHA_BUILD_SHARED
#    define MR_EXPORT __declspec(dllexport)
#  else
#    define MR_EXPORT
#  endif
#else
#  define MR_EXPORT __attribute__((visibility("default")))
#endif

#if defined(_MSC_VER)
#  define MR_ALIGN8 __declspec(align(8))
#else
#  define MR_ALIGN8 __attribute__((aligned(8)))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Interleaved single-precision complex, 8-byte aligned (numpy complex64,
 * std::complex<float>, HIP float2). Reference: mr_aha.h:179-182. */
typedef struct MR_ALIGN8 mr_complex_s {
    float x; /* real      */
    float y; /* imaginary */
} mr_complex;

/* Default neighbourhood radius of mr_calc_mask. Reference: mr_aha.h:205. */
#define MR_DEFAULT_MASK_RADIUS 2

/* Status codes. Reference: mr_aha.h:225-237. */
typedef enum mr_status {
    MR_SUCCESS            =  0,
    MR_ERR_INVALID_ARG    = -1,
    MR_ERR_INVALID_CONFIG = -2,
    MR_ERR_ALLOC          = -3,
    MR_ERR_CUDA           = -4,  /* HIP runtime failure on MI355X          */
    MR_ERR_CUFFT          = -5,  /* failure of the library's FFT engine     */
    MR_ERR_CUFINUFFT      = -6,  /* failure of the library's gridding       */
    MR_ERR_UNSUPPORTED    = -7,
    MR_ERR_INTERNAL       = -8,
    MR_ERR_METAL          = -9,  /* never returned by this build            */
    MR_ERR_VKFFT          = -10  /* never returned by this build            */
} mr_status;

/* Static text of a status. Reference: mr_aha.h:246. */
MR_EXPORT const char *mr_status_string(int status);
/* Thread-local message of the last failure. Reference: mr_aha.h:265. */
MR_EXPORT const char *mr_last_error(void);
/* "cuda": arrays are device pointers. Reference: mr_aha.h:276. */
MR_EXPORT const char *mr_backend(void);

/* Trajectory family. Reference: mr_aha.h:291-297. */
typedef enum mr_traj_type {
    MR_TRAJ_2D    = 0,
    MR_TRAJ_3D    = 1,
    MR_TRAJ_STACK = 2
} mr_traj_type;

/* Stored triangle of each packed Hermitian block (column-major packing).
 * lower: (i, j), i >= j at t = j*L - j*(j-1)/2 + (i - j)
 * upper: (i, j), i <= j at t = j*(j+1)/2 + i
 * Reference: mr_aha.h:314-317. */
typedef enum mr_triangulation {
    MR_TRI_LOWER = 0,
    MR_TRI_UPPER = 1
} mr_triangulation;

/* Operator description. Field order and types fix the ABI: 104 bytes on
 * LP64 with a 4-byte hole before num_masked. Reference: mr_aha.h:332-414. */
typedef struct mr_config {
    mr_traj_type     traj_type;
    int              n_spatial;     /* 2 for 2D, 3 for 3D and stack        */
    int              n_dim[3];      /* image size [Nx, Ny, Nz]              */
    int              phi2_dim[3];   /* upsampled grid [Mx, My, Mz]          */
    int              L;             /* rank                                  */
    mr_triangulation triangulation;
    int              is_masked;
    int64_t          num_masked;
    int              n_coils;
    double           eps;           /* gridding tolerance, (0, 1)           */
    int              gpu_device_id; /* -1: caller's current device          */
    void            *stream;        /* hipStream_t or NULL                  */
    double           mem_fraction;  /* share of free memory for work buffers*/
    int              verbose;
    int              mask_radius;
} mr_config;

/* Reference: mr_aha.h:433. */
MR_EXPORT int mr_config_defaults(mr_config *cfg, mr_traj_type traj_type);
/* Reference: mr_aha.h:457. */
MR_EXPORT int mr_config_validate(const mr_config *cfg);

/* Element counts (not bytes). Reference: mr_aha.h:474-518. */
MR_EXPORT int64_t mr_num_tri(int L);
MR_EXPORT int64_t mr_image_size(const mr_config *cfg);
MR_EXPORT int64_t mr_pad_size(const mr_config *cfg);
MR_EXPORT int64_t mr_phi2_toep_size(const mr_config *cfg);
MR_EXPORT int64_t mr_x_size(const mr_config *cfg);
MR_EXPORT int64_t mr_se_size(const mr_config *cfg);

/* Reusable single-GPU plan: work buffers and FFT plans. Reference:
 * mr_aha.h:535. */
typedef struct mr_aha_plan_s *mr_aha_plan;

/* Reference: mr_aha.h:578. */
MR_EXPORT int mr_aha_plan_create(const mr_config *cfg, mr_aha_plan *plan_out);
/* AhAx = sum_c S_c^H P^H F^-1 T F P S_c x, asynchronous on cfg->stream.
 * Reference: mr_aha.h:635-640. */
MR_EXPORT int mr_aha_plan_execute(mr_aha_plan         plan,
                                  const mr_complex   *d_x,
                                  const mr_complex   *d_SEs,
                                  const mr_complex   *d_phi2_toep,
                                  const int64_t      *d_mask,
                                  mr_complex         *d_AhAx);
/* Reference: mr_aha.h:653. */
MR_EXPORT int mr_aha_plan_destroy(mr_aha_plan plan);
/* Create + execute + destroy. Reference: mr_aha.h:670-675. */
MR_EXPORT int mr_aha_toeplitz(const mr_config  *cfg,
                              const mr_complex *d_x,
                              const mr_complex *d_SEs,
                              const mr_complex *d_phi2_toep,
                              const int64_t    *d_mask,
                              mr_complex       *d_AhAx);

/* Operator object. Reference: mr_aha.h:746-898. */
#define MR_OP_MAX_DEVICES 16
#define MR_OP_ALL_DEVICES (-1)
#define MR_OP_SHARE_PRIMARY 1

typedef enum mr_op_method {
    MR_OP_TOEPLITZ = 0,
    MR_OP_NUFFT    = 1
} mr_op_method;

typedef enum mr_toep_dist {
    MR_TOEP_DIST_AUTO      = 0,
    MR_TOEP_DIST_SPLIT     = 1,
    MR_TOEP_DIST_REPLICATE = 2,
    MR_TOEP_DIST_BROADCAST = 3
} mr_toep_dist;

/* 56 bytes on LP64. Reference: mr_aha.h:814-842. */
typedef struct mr_op_inputs {
    const mr_complex *d_SEs;
    const mr_complex *d_phi2;
    const float      *d_traj;
    int64_t           Ntrajs;
    int64_t           Nkx;
    const int64_t    *d_mask;
    const mr_complex *d_phi2_toep;
} mr_op_inputs;

/* 48 bytes on LP64. Reference: mr_aha.h:849-874. */
typedef struct mr_op_options {
    mr_op_method method;
    const int   *devices;
    int          n_devices;
    mr_toep_dist toep_dist;
    int          flags;
    double       nufft_upsampfac;
    int          nufft_gpu_method;
} mr_op_options;

typedef struct mr_op_s *mr_op;

#define MR_OP_TOEP_SUPPLIED  0
#define MR_OP_TOEP_BROADCAST 1
#define MR_OP_TOEP_SPLIT     2
#define MR_OP_TOEP_REPLICATE 3
#define MR_OP_TOEP_NONE      4

/* 320 bytes on LP64. Reference: mr_aha.h:905-952. */
typedef struct mr_op_info_t {
    int    method;
    int    n_devices;
    int    devices[MR_OP_MAX_DEVICES];
    int    coil_begin[MR_OP_MAX_DEVICES];
    int    coil_count[MR_OP_MAX_DEVICES];
    int    group[MR_OP_MAX_DEVICES];
    int    n_split_coils;
    int    peer_access;
    double peer_bandwidth_gbs;
    int    toep_source;
    int    toep_devices;
    double toep_ms;
    double toep_component_ms;
    double toep_gather_gbs;
    double create_ms;
} mr_op_info_t;

/* Reference: mr_aha.h:967. */
MR_EXPORT int mr_op_options_defaults(mr_op_options *opt);
/* Synchronous; one device (the primary). Reference: mr_aha.h:1004-1007. */
MR_EXPORT int mr_op_create(const mr_config     *cfg,
                           const mr_op_inputs  *in,
                           const mr_op_options *opt,
                           mr_op               *op_out);
/* Reference: mr_aha.h:1026. */
MR_EXPORT int mr_op_execute(mr_op op, const mr_complex *d_x, mr_complex *d_AhAx);
/* Reference: mr_aha.h:1038. */
MR_EXPORT int mr_op_info(mr_op op, mr_op_info_t *info);
/* Reference: mr_aha.h:1059. */
MR_EXPORT int mr_op_kernel(mr_op op, int k, const mr_complex **d_phi2_toep);
/* Reference: mr_aha.h:1070. */
MR_EXPORT int mr_op_destroy(mr_op op);
/* Reference: mr_aha.h:1086. */
MR_EXPORT int mr_op_abandon(mr_op op);

/* Toeplitz kernel from Gram blocks and trajectory. Reference:
 * mr_aha.h:1150-1156. */
MR_EXPORT int mr_calc_toep(const mr_config  *cfg,
                           const mr_complex *d_phi2,
                           const float      *d_traj,
                           int64_t           Ntrajs,
                           int64_t           Nkx,
                           const int64_t    *d_mask,
                           mr_complex       *d_phi2_toep);
/* Finish components [t0, t0+n_comp) from point-spread functions.
 * Reference: mr_aha.h:1199-1204. */
MR_EXPORT int mr_calc_toep_from_psf(const mr_config  *cfg,
                                    const mr_complex *d_psf,
                                    int64_t           t0,
                                    int64_t           n_comp,
                                    const int64_t    *d_mask,
                                    mr_complex       *d_phi2_toep);
/* Neighbourhood mask from the trajectory. Reference: mr_aha.h:1254-1259. */
MR_EXPORT int mr_calc_mask(const mr_config *cfg,
                           const float     *d_traj,
                           int64_t          Ntrajs,
                           int64_t          Nkx,
                           int64_t         *d_mask,
                           int64_t         *num_masked_out);

/* Packed library version major*10000 + minor*100 + patch.
 * Reference: src/common/mr_shared_stub.cpp:23. */
MR_EXPORT int mr_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MR_AHA_MR_AHA_H */
