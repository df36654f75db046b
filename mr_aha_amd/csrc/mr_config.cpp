// Host-only part of libmr_aha_c: error reporting, configuration defaults
// and validation, derived sizes, version.
//
// Behaviour follows src/common/mr_config.cpp of the reference: the same
// defaults (:74-104), the same validation rules in the same order with the
// same messages (:106-173; bindings and tests match on the text), the same
// size helpers (:179-215).

#include "mr_internal.hpp"

#include <cstdarg>
#include <cstdlib>
#include <cstring>

namespace {
constexpr int kErrLen = 512;
thread_local char g_err[kErrLen] = {0};
}  // namespace

void mr_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, kErrLen, fmt, ap);
    va_end(ap);
}

void mr_clear_error()
{
    g_err[0] = '\0';
    mr_backend_clear_error();
}

long long mr_env_ll(const char *name, long long def)
{
    const char *e = std::getenv(name);
    return (e && e[0]) ? std::atoll(e) : def;
}

extern "C" MR_EXPORT const char *mr_last_error(void) { return g_err; }

extern "C" MR_EXPORT const char *mr_status_string(int status)
{
    switch (status) {
    case MR_SUCCESS:            return "success";
    case MR_ERR_INVALID_ARG:    return "invalid argument";
    case MR_ERR_INVALID_CONFIG: return "invalid configuration";
    case MR_ERR_ALLOC:          return "GPU memory allocation failed";
    case MR_ERR_CUDA:           return "HIP runtime error";
    case MR_ERR_CUFFT:          return "FFT engine error";
    case MR_ERR_CUFINUFFT:      return "gridding error";
    case MR_ERR_UNSUPPORTED:    return "unsupported configuration";
    case MR_ERR_INTERNAL:       return "internal error";
    case MR_ERR_METAL:          return "Metal error";
    case MR_ERR_VKFFT:          return "VkFFT error";
    default:                    return "unknown status";
    }
}

extern "C" MR_EXPORT int mr_version(void) { return 1 * 10000 + 0 * 100 + 0; }

extern "C" MR_EXPORT int mr_config_defaults(mr_config *cfg, mr_traj_type traj_type)
{
    mr_clear_error();
    MR_REQUIRE(cfg != nullptr, "mr_config_defaults: cfg is NULL");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->traj_type = traj_type;
    cfg->n_spatial = (traj_type == MR_TRAJ_2D) ? 2 : 3;
    for (int d = 0; d < 3; ++d) {
        cfg->n_dim[d] = 1;
        cfg->phi2_dim[d] = 1;
    }
    cfg->L = 1;
    cfg->triangulation = MR_TRI_LOWER;
    cfg->n_coils = 1;
    cfg->eps = 1e-6;
    cfg->gpu_device_id = 0;
    cfg->stream = nullptr;
    cfg->mem_fraction = 0.5;
    cfg->mask_radius = MR_DEFAULT_MASK_RADIUS;
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int mr_config_validate(const mr_config *cfg)
{
    mr_clear_error();
    MR_REQUIRE(cfg != nullptr, "mr_config_validate: cfg is NULL");
    MR_REQUIRE_CFG(cfg->traj_type == MR_TRAJ_2D || cfg->traj_type == MR_TRAJ_3D ||
                       cfg->traj_type == MR_TRAJ_STACK,
                   "traj_type must be one of MR_TRAJ_2D, MR_TRAJ_3D, MR_TRAJ_STACK");
    const int want = (cfg->traj_type == MR_TRAJ_2D) ? 2 : 3;
    MR_REQUIRE_CFG(cfg->n_spatial == want,
                   "n_spatial is %d but trajectory type requires %d", cfg->n_spatial, want);
    for (int d = 0; d < 3; ++d) {
        MR_REQUIRE_CFG(cfg->n_dim[d] >= 1, "n_dim[%d] = %d must be >= 1", d, cfg->n_dim[d]);
        MR_REQUIRE_CFG(cfg->phi2_dim[d] >= 1, "phi2_dim[%d] = %d must be >= 1", d,
                       cfg->phi2_dim[d]);
        if (d >= cfg->n_spatial)
            MR_REQUIRE_CFG(cfg->n_dim[d] == 1 && cfg->phi2_dim[d] == 1,
                           "n_dim[%d] and phi2_dim[%d] must be 1 for a %dD trajectory", d, d,
                           cfg->n_spatial);
        MR_REQUIRE_CFG(cfg->phi2_dim[d] >= cfg->n_dim[d],
                       "phi2_dim[%d] = %d must not be smaller than n_dim[%d] = %d", d,
                       cfg->phi2_dim[d], d, cfg->n_dim[d]);
    }
    if (cfg->traj_type == MR_TRAJ_STACK)
        MR_REQUIRE_CFG(cfg->phi2_dim[2] == cfg->n_dim[2],
                       "for MR_TRAJ_STACK, phi2_dim[2] (Mz = %d) must equal n_dim[2] (Nz = %d)",
                       cfg->phi2_dim[2], cfg->n_dim[2]);
    MR_REQUIRE_CFG(cfg->L >= 1, "L = %d must be >= 1", cfg->L);
    MR_REQUIRE_CFG(cfg->n_coils >= 1, "n_coils = %d must be >= 1", cfg->n_coils);
    MR_REQUIRE_CFG(cfg->triangulation == MR_TRI_LOWER || cfg->triangulation == MR_TRI_UPPER,
                   "triangulation must be MR_TRI_LOWER or MR_TRI_UPPER");
    MR_REQUIRE_CFG(cfg->eps > 0.0 && cfg->eps < 1.0, "eps = %g must lie in (0, 1)", cfg->eps);
    MR_REQUIRE_CFG(cfg->mask_radius >= 0, "mask_radius = %d must be >= 0", cfg->mask_radius);
    if (cfg->is_masked) {
        const int64_t pad = mr_pad_size(cfg);
        MR_REQUIRE_CFG(cfg->num_masked > 0, "num_masked = %lld must be > 0 when is_masked is set",
                       (long long)cfg->num_masked);
        MR_REQUIRE_CFG(cfg->num_masked <= pad,
                       "num_masked = %lld exceeds the upsampled grid size %lld",
                       (long long)cfg->num_masked, (long long)pad);
    }
    return MR_SUCCESS;
}

extern "C" MR_EXPORT int64_t mr_num_tri(int L)
{
    return (L <= 0) ? 0 : (int64_t)L * (L + 1) / 2;
}

extern "C" MR_EXPORT int64_t mr_image_size(const mr_config *cfg)
{
    return cfg ? (int64_t)cfg->n_dim[0] * cfg->n_dim[1] * cfg->n_dim[2] : 0;
}

extern "C" MR_EXPORT int64_t mr_pad_size(const mr_config *cfg)
{
    return cfg ? (int64_t)cfg->phi2_dim[0] * cfg->phi2_dim[1] * cfg->phi2_dim[2] : 0;
}

extern "C" MR_EXPORT int64_t mr_phi2_toep_size(const mr_config *cfg)
{
    if (!cfg) return 0;
    return (cfg->is_masked ? cfg->num_masked : mr_pad_size(cfg)) * mr_num_tri(cfg->L);
}

extern "C" MR_EXPORT int64_t mr_x_size(const mr_config *cfg)
{
    return cfg ? mr_image_size(cfg) * cfg->L : 0;
}

extern "C" MR_EXPORT int64_t mr_se_size(const mr_config *cfg)
{
    return cfg ? mr_image_size(cfg) * cfg->n_coils : 0;
}

void mr_dims_from_config(const mr_config *cfg, mr_dims *d)
{
    std::memset(d, 0, sizeof(*d));
    for (int i = 0; i < 3; ++i) {
        d->n[i] = cfg->n_dim[i];
        d->m[i] = cfg->phi2_dim[i];
        d->c_n[i] = cfg->n_dim[i] / 2;
        d->c_m[i] = cfg->phi2_dim[i] / 2;
    }
    d->L = cfg->L;
    d->tri_upper = (cfg->triangulation == MR_TRI_UPPER) ? 1 : 0;
    d->image_size = (long long)mr_image_size(cfg);
    d->pad_size = (long long)mr_pad_size(cfg);
}
