// The A^H A operator object (mr_op_*) on one MI355X.
//
// Follows the single-device ("direct") path of src/cuda/mr_op.cu:
// validation (:951-990), candidate devices (src/cuda/mr_toep_mgpu.cu:
// 189-232), kernel computed in place or supplied (:873-945), maps aliased
// with MR_OP_SHARE_PRIMARY (:745-754), engine on the caller's stream,
// report (:1260-1281). Requests for further devices are validated and then
// served by the primary device alone: the operator is a sum over coils and
// the result is the same on one device (the multi-device split is not part
// of this library).

#include "mr_hip.hpp"
#include "mr_aha/mr_aha_ext.h"

#include <chrono>
#include <cstring>
#include <new>

int mr_calc_toep_impl(const mr_config *cfg, const mr_complex *d_phi2, const float *d_traj,
                      int64_t Ntrajs, int64_t Nkx, const int64_t *d_mask, mr_complex *d_out);
int mr_aha_plan_execute_raw(mr_aha_plan plan, const mr_complex *d_x, const mr_complex *d_SEs,
                            const mr_complex *d_phi2_toep, const int64_t *d_mask,
                            mr_complex *d_AhAx);
int mr_aha_plan_group(mr_aha_plan p);
int mr_aha_plan_set_kernel_format(mr_aha_plan p, int kf, const unsigned *q, const float *sc);
int mr_kf_quantize(int kf, const float2 *ker, long long kcol, int L, int upper, unsigned *q,
                   float *sc, hipStream_t stream);

struct mr_nufft_eng;
int mr_nufft_eng_create(const mr_config *cfg, const mr_complex *d_phi2, const float *d_traj,
                        int64_t Ntrajs, int64_t Nkx, double upsampfac, mr_nufft_eng **out);
int mr_nufft_eng_execute(mr_nufft_eng *e, const mr_complex *d_x, const mr_complex *d_SEs,
                         mr_complex *d_out);
void mr_nufft_eng_destroy(mr_nufft_eng *e);

struct mr_op_s {
    mr_config cfg;
    mr_op_options opt;
    int primary;
    hipStream_t stream;
    mr_aha_plan plan;
    mr_nufft_eng *nu;
    const mr_complex *d_SEs;
    const int64_t *d_mask;
    const mr_complex *d_toep;
    mr_complex *own_SEs;
    int64_t *own_mask;
    mr_complex *own_toep;
    unsigned *own_kq;   // 16-bit kernel entries (mr_op_set_kernel_format)
    float *own_ksc;     // and their block scales
    int kf;
    mr_op_info_t info;
};

static void mr_op_release(mr_op p)
{
    if (p->plan) mr_aha_plan_destroy(p->plan);
    if (p->nu) mr_nufft_eng_destroy(p->nu);
    if (p->own_SEs) (void)hipFree(p->own_SEs);
    if (p->own_mask) (void)hipFree(p->own_mask);
    if (p->own_toep) (void)hipFree(p->own_toep);
    if (p->own_kq) (void)hipFree(p->own_kq);
    if (p->own_ksc) (void)hipFree(p->own_ksc);
    (void)hipGetLastError();
    delete p;
}

// MR_AHA_TOEP_SPLIT overrides opt->toep_dist (src/cuda/mr_toep_mgpu.cu:156-185).
static int mr_op_resolve_dist(int requested)
{
    int m = requested;
    const char *e = std::getenv("MR_AHA_TOEP_SPLIT");
    if (e && e[0]) {
        if (!strcmp(e, "auto")) m = MR_TOEP_DIST_AUTO;
        else if (!strcmp(e, "split")) m = MR_TOEP_DIST_SPLIT;
        else if (!strcmp(e, "replicate")) m = MR_TOEP_DIST_REPLICATE;
        else if (!strcmp(e, "broadcast") || !strcmp(e, "off")) m = MR_TOEP_DIST_BROADCAST;
        else
            MR_FAIL(MR_ERR_INVALID_ARG,
                    "MR_AHA_TOEP_SPLIT='%s' is not one of auto, split, replicate, broadcast (off)",
                    e);
    }
    if (m < MR_TOEP_DIST_AUTO || m > MR_TOEP_DIST_BROADCAST)
        MR_FAIL(MR_ERR_INVALID_ARG, "toep_dist = %d is not a valid mr_toep_dist", m);
    return MR_SUCCESS;
}

static int mr_op_candidates(const mr_config *cfg, const int *devices, int n_devices, int *primary)
{
    const char *who = "mr_op_create";
    const int all = (devices == nullptr && n_devices == MR_OP_ALL_DEVICES);
    MR_REQUIRE(all || (n_devices >= 0 && n_devices <= MR_OP_MAX_DEVICES),
               "%s: n_devices must be in [0, %d] or MR_OP_ALL_DEVICES (with devices NULL)", who,
               MR_OP_MAX_DEVICES);
    MR_REQUIRE(n_devices <= 0 || devices != nullptr, "%s: devices is NULL but n_devices = %d",
               who, n_devices);
    int visible = 0;
    MR_HIP_CHECK(hipGetDeviceCount(&visible));
    int pr = cfg->gpu_device_id;
    if (pr < 0) MR_HIP_CHECK(hipGetDevice(&pr));
    MR_REQUIRE(pr < visible, "%s: primary device %d is not visible", who, pr);
    if (!all)
        for (int k = 0; k < n_devices; ++k)
            MR_REQUIRE(devices[k] >= 0 && devices[k] < visible,
                       "%s: device %d is not visible (%d visible)", who, devices[k], visible);
    *primary = pr;
    return MR_SUCCESS;
}

static int mr_op_check(const mr_config *cfg, const mr_op_inputs *in, const mr_op_options *o)
{
    MR_TRY(mr_config_validate(cfg));
    MR_REQUIRE(o->method == MR_OP_TOEPLITZ || o->method == MR_OP_NUFFT,
               "mr_op_create: method = %d is neither MR_OP_TOEPLITZ nor MR_OP_NUFFT",
               (int)o->method);
    MR_REQUIRE(in->d_SEs != nullptr, "mr_op_create: d_SEs is NULL");
    const int need_gram = (o->method == MR_OP_NUFFT || in->d_phi2_toep == nullptr);
    if (need_gram) {
        MR_REQUIRE(in->d_phi2 != nullptr, "mr_op_create: d_phi2 is NULL");
        MR_REQUIRE(in->d_traj != nullptr, "mr_op_create: d_traj is NULL");
        MR_REQUIRE(in->Ntrajs > 0 && in->Nkx > 0,
                   "mr_op_create: Ntrajs = %lld and Nkx = %lld must be > 0",
                   (long long)in->Ntrajs, (long long)in->Nkx);
    }
    if (o->method == MR_OP_NUFFT) {
        MR_REQUIRE(!cfg->is_masked,
                   "mr_op_create: the NUFFT method has no kernel to mask; set is_masked = 0 (or "
                   "use the Toeplitz method)");
        MR_REQUIRE(in->d_phi2_toep == nullptr,
                   "mr_op_create: d_phi2_toep is only used by the Toeplitz method");
        MR_REQUIRE(o->nufft_upsampfac == 0.0 || o->nufft_upsampfac > 1.0,
                   "mr_op_create: nufft_upsampfac = %g must be > 1 (or 0 for the default 2.0)",
                   o->nufft_upsampfac);
    } else if (cfg->is_masked) {
        MR_REQUIRE(in->d_mask != nullptr, "mr_op_create: d_mask is required when is_masked is set");
    }
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_op_options_defaults(mr_op_options *opt)
{
    mr_clear_error();
    MR_REQUIRE(opt != nullptr, "mr_op_options_defaults: opt is NULL");
    std::memset(opt, 0, sizeof(*opt));
    opt->method = MR_OP_TOEPLITZ;
    opt->toep_dist = MR_TOEP_DIST_AUTO;
    opt->flags = MR_OP_SHARE_PRIMARY;
    return MR_SUCCESS;
}

static int mr_op_create_impl(const mr_config *cfg, const mr_op_inputs *in,
                             const mr_op_options *o, mr_op *op_out)
{
    const auto t_create = std::chrono::steady_clock::now();
    MR_TRY(mr_op_check(cfg, in, o));
    const bool toeplitz = o->method == MR_OP_TOEPLITZ;
    const bool supplied = toeplitz && in->d_phi2_toep != nullptr;
    if (toeplitz && !supplied) MR_TRY(mr_op_resolve_dist((int)o->toep_dist));
    int primary = 0;
    MR_TRY(mr_op_candidates(cfg, o->devices, o->n_devices, &primary));

    mr_op p = new (std::nothrow) mr_op_s();
    if (!p) MR_FAIL(MR_ERR_ALLOC, "failed to allocate the operator struct");
    p->cfg = *cfg;
    p->cfg.gpu_device_id = primary;
    p->opt = *o;
    p->opt.devices = nullptr;
    p->primary = primary;
    p->stream = (hipStream_t)cfg->stream;
    if (hipSetDevice(primary) != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess) {
        mr_set_error("mr_op_create: HIP error on the primary device %d: %s", primary,
                     hipGetErrorString(hipGetLastError()));
        mr_op_release(p);
        return MR_ERR_CUDA;
    }
    const size_t c8 = sizeof(mr_complex);
    const size_t se_bytes = (size_t)mr_se_size(cfg) * c8;
    const size_t mask_bytes = (toeplitz && cfg->is_masked) ? (size_t)cfg->num_masked * sizeof(int64_t) : 0;
    const size_t toep_bytes = toeplitz ? (size_t)mr_phi2_toep_size(cfg) * c8 : 0;
    const bool share = (o->flags & MR_OP_SHARE_PRIMARY) != 0;
    int rc = MR_SUCCESS;
    double toep_ms = 0.0;

    // Constants: alias or copy (device to device on the caller's stream).
    if (share) {
        p->d_SEs = in->d_SEs;
        p->d_mask = mask_bytes ? in->d_mask : nullptr;
    } else {
        hipError_t e = hipMalloc((void **)&p->own_SEs, se_bytes);
        if (e == hipSuccess && mask_bytes) e = hipMalloc((void **)&p->own_mask, mask_bytes);
        if (e == hipSuccess)
            e = hipMemcpyAsync(p->own_SEs, in->d_SEs, se_bytes, hipMemcpyDeviceToDevice, p->stream);
        if (e == hipSuccess && mask_bytes)
            e = hipMemcpyAsync(p->own_mask, in->d_mask, mask_bytes, hipMemcpyDeviceToDevice, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            mr_op_release(p);
            MR_FAIL(MR_ERR_ALLOC, "failed to copy the maps%s: %s", mask_bytes ? " and the mask" : "",
                    hipGetErrorString(e));
        }
        p->d_SEs = p->own_SEs;
        p->d_mask = p->own_mask;
    }

    if (toeplitz) {
        if (supplied) {
            if (share) {
                p->d_toep = in->d_phi2_toep;
            } else {
                hipError_t e = hipMalloc((void **)&p->own_toep, toep_bytes);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(p->own_toep, in->d_phi2_toep, toep_bytes,
                                       hipMemcpyDeviceToDevice, p->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    mr_op_release(p);
                    MR_FAIL(MR_ERR_ALLOC, "failed to copy the %.2f GB kernel: %s",
                            toep_bytes / 1e9, hipGetErrorString(e));
                }
                p->d_toep = p->own_toep;
            }
        } else {
            const auto t0 = std::chrono::steady_clock::now();
            if (hipMalloc((void **)&p->own_toep, toep_bytes) != hipSuccess) {
                (void)hipGetLastError();
                mr_op_release(p);
                MR_FAIL(MR_ERR_ALLOC, "device %d: failed to allocate the %.2f GB copy of phi2_toep",
                        primary, toep_bytes / 1e9);
            }
            rc = mr_calc_toep_impl(&p->cfg, in->d_phi2, in->d_traj, in->Ntrajs, in->Nkx,
                                   p->d_mask ? p->d_mask : in->d_mask, p->own_toep);
            if (rc != MR_SUCCESS) {
                mr_op_release(p);
                return rc;
            }
            p->d_toep = p->own_toep;
            toep_ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        }
        mr_config pc = p->cfg;
        pc.verbose = 0;
        rc = mr_aha_plan_create(&pc, &p->plan);
    } else {
        rc = mr_nufft_eng_create(&p->cfg, in->d_phi2, in->d_traj, in->Ntrajs, in->Nkx,
                                 o->nufft_upsampfac, &p->nu);
    }
    if (rc != MR_SUCCESS) {
        mr_op_release(p);
        return rc;
    }

    mr_op_info_t *nf = &p->info;
    std::memset(nf, 0, sizeof(*nf));
    nf->method = (int)o->method;
    nf->n_devices = 1;
    nf->devices[0] = primary;
    nf->coil_begin[0] = 0;
    nf->coil_count[0] = cfg->n_coils;
    nf->group[0] = toeplitz ? mr_aha_plan_group(p->plan) : 1;
    nf->toep_source = toeplitz ? (supplied ? MR_OP_TOEP_SUPPLIED : MR_OP_TOEP_BROADCAST)
                               : MR_OP_TOEP_NONE;
    nf->toep_devices = (toeplitz && !supplied) ? 1 : 0;
    nf->toep_ms = toep_ms;
    nf->create_ms = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - t_create).count();
    if (cfg->verbose)
        printf("[mr_aha] %s operator: %d coil(s) on device %d, created in %.0f ms\n",
               toeplitz ? "Toeplitz" : "NUFFT", cfg->n_coils, primary, nf->create_ms);
    *op_out = p;
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_op_create(const mr_config *cfg, const mr_op_inputs *in,
                                      const mr_op_options *opt, mr_op *op_out)
{
    mr_clear_error();
    MR_REQUIRE(op_out != nullptr, "mr_op_create: op_out is NULL");
    *op_out = nullptr;
    MR_REQUIRE(cfg != nullptr, "mr_op_create: cfg is NULL");
    MR_REQUIRE(in != nullptr, "mr_op_create: in is NULL");
    mr_op_options o;
    if (opt == nullptr) mr_op_options_defaults(&o);
    else o = *opt;
    int prev = 0;
    MR_HIP_CHECK(hipGetDevice(&prev));
    const int rc = mr_op_create_impl(cfg, in, &o, op_out);
    (void)hipSetDevice(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_op_execute(mr_op op, const mr_complex *d_x, mr_complex *d_AhAx)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_execute: op is NULL");
    MR_REQUIRE(d_x != nullptr, "mr_op_execute: d_x is NULL");
    MR_REQUIRE(d_AhAx != nullptr, "mr_op_execute: d_AhAx is NULL");
    MR_REQUIRE((const void *)d_AhAx != (const void *)d_x,
               "mr_op_execute: d_AhAx must not alias d_x");
    int prev = 0;
    MR_HIP_CHECK(hipGetDevice(&prev));
    int rc = MR_SUCCESS;
    if (prev != op->primary && hipSetDevice(op->primary) != hipSuccess)
        MR_FAIL(MR_ERR_CUDA, "mr_op_execute: cannot select device %d", op->primary);
    if (op->plan)
        rc = mr_aha_plan_execute_raw(op->plan, d_x, op->d_SEs, op->d_toep, op->d_mask, d_AhAx);
    else
        rc = mr_nufft_eng_execute(op->nu, d_x, op->d_SEs, d_AhAx);
    if (prev != op->primary) (void)hipSetDevice(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_op_info(mr_op op, mr_op_info_t *info)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_info: op is NULL");
    MR_REQUIRE(info != nullptr, "mr_op_info: info is NULL");
    *info = op->info;
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_op_kernel(mr_op op, int k, const mr_complex **d_phi2_toep)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_kernel: op is NULL");
    MR_REQUIRE(d_phi2_toep != nullptr, "mr_op_kernel: d_phi2_toep is NULL");
    if (op->opt.method != MR_OP_TOEPLITZ)
        MR_FAIL(MR_ERR_UNSUPPORTED, "mr_op_kernel: a NUFFT operator has no kernel");
    MR_REQUIRE(k >= 0 && k < 1, "mr_op_kernel: device index %d outside [0, %d)", k, 1);
    *d_phi2_toep = op->d_toep;
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_op_destroy(mr_op op)
{
    mr_clear_error();
    if (!op) return MR_SUCCESS;
    int prev = 0;
    const hipError_t ce = hipGetDevice(&prev);
    (void)hipSetDevice(op->primary);
    (void)hipStreamSynchronize(op->stream);
    (void)hipGetLastError();
    mr_op_release(op);
    if (ce == hipSuccess) (void)hipSetDevice(prev);
    return MR_SUCCESS;
}

// After a device reset the operator's device memory no longer exists; only
// the host side is released.
extern "C" MR_EXPORT int mr_op_abandon(mr_op op)
{
    mr_clear_error();
    if (!op) return MR_SUCCESS;
    if (op->plan) {
        // The plan's buffers died with the context; release the host struct
        // through the normal path, whose hipFree calls fail harmlessly.
        mr_aha_plan_destroy(op->plan);
    }
    if (op->nu) mr_nufft_eng_destroy(op->nu);
    (void)hipGetLastError();
    delete op;
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_op_profile(mr_op op, double *ms, int *calls, int n)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_profile: op is NULL");
    if (!op->plan) MR_FAIL(MR_ERR_UNSUPPORTED, "mr_op_profile: only Toeplitz operators are profiled");
    return mr_aha_plan_profile(op->plan, ms, calls, n);
}

// Stream, image-vector length (N * L) and device of an operator.
int mr_op_stream_len(mr_op op, hipStream_t *s, long long *n, int *device)
{
    MR_REQUIRE(op != nullptr, "operator is NULL");
    *s = op->stream;
    *n = (long long)mr_x_size(&op->cfg);
    *device = op->primary;
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_op_set_kernel_format(mr_op op, int format)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_set_kernel_format: op is NULL");
    MR_REQUIRE(format == MR_KERNEL_FP32 || format == MR_KERNEL_BF16 || format == MR_KERNEL_FP16,
               "mr_op_set_kernel_format: format %d is not MR_KERNEL_FP32, _BF16 or _FP16", format);
    if (!op->plan)
        MR_FAIL(MR_ERR_UNSUPPORTED, "mr_op_set_kernel_format: a NUFFT operator has no kernel");
    int prev = 0;
    MR_HIP_CHECK(hipGetDevice(&prev));
    if (prev != op->primary) MR_HIP_CHECK(hipSetDevice(op->primary));
    int rc = MR_SUCCESS;
    unsigned *q = nullptr;
    float *sc = nullptr;
    do {
        if (format != MR_KERNEL_FP32) {
            const long long kcol = op->cfg.is_masked ? (long long)op->cfg.num_masked
                                                     : (long long)mr_pad_size(&op->cfg);
            const int L = op->cfg.L;
            const long long T = (long long)L * (L + 1) / 2;
            const long long NB = (long long)((L + 3) / 4) * ((L + 3) / 4 + 1) / 2;
            if (hipMalloc((void **)&q, (size_t)(T * kcol) * 4) != hipSuccess ||
                hipMalloc((void **)&sc, (size_t)(NB * kcol) * 4) != hipSuccess) {
                (void)hipGetLastError();
                mr_set_error("mr_op_set_kernel_format: failed to allocate the 16-bit kernel "
                             "(%lld bytes)", (T + NB) * kcol * 4);
                rc = MR_ERR_ALLOC;
                break;
            }
            rc = mr_kf_quantize(format, reinterpret_cast<const float2 *>(op->d_toep), kcol, L,
                                op->cfg.triangulation == MR_TRI_UPPER ? 1 : 0, q, sc, op->stream);
            if (rc != MR_SUCCESS) break;
        }
        rc = mr_aha_plan_set_kernel_format(op->plan, format, q, sc);
        if (rc != MR_SUCCESS) break;
        // earlier applies may still read the previous buffers
        if (hipStreamSynchronize(op->stream) != hipSuccess) {
            mr_set_error("mr_op_set_kernel_format: stream synchronisation failed");
            rc = MR_ERR_CUDA;
            break;
        }
        if (op->own_kq) (void)hipFree(op->own_kq);
        if (op->own_ksc) (void)hipFree(op->own_ksc);
        op->own_kq = q;
        op->own_ksc = sc;
        q = nullptr;
        sc = nullptr;
        op->kf = format;
    } while (false);
    if (q) (void)hipFree(q);
    if (sc) (void)hipFree(sc);
    if (prev != op->primary) (void)hipSetDevice(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_op_autotune(mr_op op, const mr_complex *d_x, mr_complex *d_AhAx,
                                        int reps)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_autotune: op is NULL");
    if (!op->plan) return MR_SUCCESS;  // NUFFT operators have nothing to tune
    int prev = 0;
    MR_HIP_CHECK(hipGetDevice(&prev));
    if (prev != op->primary) MR_HIP_CHECK(hipSetDevice(op->primary));
    const int rc = mr_aha_plan_autotune(op->plan, d_x, op->d_SEs, op->d_toep, op->d_mask, d_AhAx,
                                        reps);
    if (prev != op->primary) (void)hipSetDevice(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_op_tiles(mr_op op, int *tiles, int n)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_tiles: op is NULL");
    if (!op->plan) {
        for (int i = 0; i < n && tiles; ++i) tiles[i] = 0;
        return MR_SUCCESS;
    }
    return mr_aha_plan_tiles(op->plan, tiles, n);
}

extern "C" MR_EXPORT int mr_op_xbd_group(mr_op op)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_xbd_group: op is NULL");
    if (!op->plan) return 0;
    return mr_aha_plan_xbd_group(op->plan);
}

extern "C" MR_EXPORT int mr_op_xbd_form(mr_op op)
{
    mr_clear_error();
    MR_REQUIRE(op != nullptr, "mr_op_xbd_form: op is NULL");
    if (!op->plan) return 0;
    return mr_aha_plan_xbd_form(op->plan);
}
