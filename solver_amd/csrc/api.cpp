// C ABI entry points of libpeec_hip: context lifetime, error reporting and
// the partial-element row calls.  The FFT operator, preconditioner and GMRES
// entry points live in operator.hip and gmres.cpp.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "context.h"

static thread_local char g_err[1024] = "";

void peec_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *peec_last_error(void) { return g_err; }

extern "C" int peec_ctx_create(int device, peec_ctx **out)
{
    if (!out) {
        peec_set_error("peec_ctx_create: null output pointer");
        return PEEC_ERR_ARG;
    }
    int ndev = 0;
    PEEC_HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        peec_set_error("peec_ctx_create: device %d out of range (%d devices)", device, ndev);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(device));
    peec_ctx *c = new peec_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        peec_set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return PEEC_ERR_HIP;
    }
    *out = c;
    return PEEC_OK;
}

extern "C" int peec_ctx_destroy(peec_ctx *ctx)
{
    if (!ctx) return PEEC_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_all();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PEEC_OK;
}

extern "C" int peec_set_stop_flag(peec_ctx *ctx, volatile const int32_t *flag)
{
    if (!ctx) {
        peec_set_error("peec_set_stop_flag: null context");
        return PEEC_ERR_ARG;
    }
    ctx->stop_flag = flag;
    return PEEC_OK;
}

bool peec_ctx::stop_requested() const { return stop_flag && *stop_flag != 0; }

void *peec_ctx::scratch(size_t bytes, int slot)
{
    if (slot < 0 || slot >= kScratchSlots) return nullptr;
    if (scratch_bytes[slot] >= bytes && scratch_ptr[slot]) return scratch_ptr[slot];
    if (scratch_ptr[slot]) (void)hipFree(scratch_ptr[slot]);
    scratch_ptr[slot] = nullptr;
    scratch_bytes[slot] = 0;
    if (hipMalloc(&scratch_ptr[slot], bytes) != hipSuccess) return nullptr;
    scratch_bytes[slot] = bytes;
    return scratch_ptr[slot];
}

int peec_ctx::prof_begin(int kind, double bytes)
{
    if (!prof_on) return -1;
    ProfRec r;
    r.kind = kind;
    r.bytes = bytes;
    for (hipEvent_t *e : {&r.a, &r.b}) {
        if (!event_pool.empty()) {
            *e = event_pool.back();
            event_pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            return -1;
        }
    }
    (void)hipEventRecord(r.a, stream);
    prof.push_back(r);
    return (int)prof.size() - 1;
}

void peec_ctx::prof_end(int idx)
{
    if (idx < 0) return;
    (void)hipEventRecord(prof[(size_t)idx].b, stream);
}

extern "C" int peec_profile(peec_ctx *ctx, int enable, double *ms, double *bytes, int64_t *count, int nkinds)
{
    if (!ctx) {
        peec_set_error("peec_profile: null context");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nkinds; k++) {
        if (ms) ms[k] = 0.0;
        if (bytes) bytes[k] = 0.0;
        if (count) count[k] = 0;
    }
    for (auto &r : ctx->prof) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess && r.kind < nkinds) {
            if (ms) ms[r.kind] += t;
            if (bytes) bytes[r.kind] += r.bytes;
            if (count) count[r.kind] += 1;
        }
        ctx->event_pool.push_back(r.a);
        ctx->event_pool.push_back(r.b);
    }
    ctx->prof.clear();
    ctx->prof_on = enable != 0;
    return PEEC_OK;
}

void peec_ctx::release_all()
{
    for (auto &r : prof) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    prof.clear();
    for (auto e : event_pool) (void)hipEventDestroy(e);
    event_pool.clear();
    for (int i = 0; i < kScratchSlots; i++) {
        if (scratch_ptr[i]) (void)hipFree(scratch_ptr[i]);
        scratch_ptr[i] = nullptr;
        scratch_bytes[i] = 0;
    }
    release_kernels();
    release_operator();
}

static int copy_centers(peec_ctx *ctx, const double *centers, int64_t N, double **d)
{
    *d = (double *)ctx->scratch((size_t)N * 3 * sizeof(double), 0);
    if (!*d) {
        peec_set_error("out of device memory for %lld centres", (long long)N);
        return PEEC_ERR_NOMEM;
    }
    PEEC_HIP_TRY(hipMemcpyAsync(*d, centers, (size_t)N * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    return PEEC_OK;
}

extern "C" int peec_lp_row(peec_ctx *ctx, const double *centri_m, const double *centri_n, int64_t N, double sx,
                           double sy, double sz, double sx2, double sy2, double sz2, int dc, double *out)
{
    if (!ctx || !centri_m || (N > 0 && (!centri_n || !out)) || N < 0 || dc < 1 || dc > 3) {
        peec_set_error("peec_lp_row: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (ctx->stop_requested()) return PEEC_STOPPED;
    if (N == 0) return PEEC_OK;
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    double *dc_ = nullptr;
    int rc = copy_centers(ctx, centri_n, N, &dc_);
    if (rc) return rc;
    double *dout = (double *)ctx->scratch((size_t)N * sizeof(double), 1);
    if (!dout) {
        peec_set_error("peec_lp_row: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    rc = peec_lp_row_device(ctx->stream, centri_m, dc_, N, sx, sy, sz, sx2, sy2, sz2, dc, 1.0, dout);
    if (rc) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(out, dout, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->stop_requested() ? PEEC_STOPPED : PEEC_OK;
}

extern "C" int peec_p_row(peec_ctx *ctx, const double *centers, const double *centers2, int64_t N, double sx,
                          double sy, double sz, int caser, int caser2, double *out)
{
    if (!ctx || !centers || (N > 0 && (!centers2 || !out)) || N < 0 || caser < 1 || caser > 6 || caser2 < 1 ||
        caser2 > 6) {
        peec_set_error("peec_p_row: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (ctx->stop_requested()) return PEEC_STOPPED;
    if (N == 0) return PEEC_OK;
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    double *dc_ = nullptr;
    int rc = copy_centers(ctx, centers2, N, &dc_);
    if (rc) return rc;
    double *dout = (double *)ctx->scratch((size_t)N * sizeof(double), 1);
    if (!dout) {
        peec_set_error("peec_p_row: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    rc = peec_p_row_device(ctx->stream, centers, dc_, N, sx, sy, sz, caser, caser2, 1.0, dout);
    if (rc) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(out, dout, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->stop_requested() ? PEEC_STOPPED : PEEC_OK;
}
