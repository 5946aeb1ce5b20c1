/*
 * capi.hip - the extern "C" entry points of include/bcmg.h.
 *
 * Every entry point catches all C++ exceptions and returns a status code;
 * the message is kept per thread for bcmg_last_error().
 */
#include "bcmg_internal.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

/* The second iterate of bcmg_relax_noalloc: one array per GPU, alive until
   the process ends. Host code never uses this name as an address: each
   device's copy is reached through the pointer hipGetSymbolAddress returns
   for it (scratch_address below). */
static __device__ __align__(256) double scratch[BCMG_RELAX_NOALLOC_MAX_ROWS];

struct bcmg_ctx_s {
    bcmg::Ctx c;
    double *scratch = nullptr; /* this device's copy of `scratch` */
};

struct bcmg_csr_s {
    bcmg_ctx ctx = nullptr;
    bcmg::DevCsr A;
};

struct bcmg_boot_s {
    bcmg_ctx ctx = nullptr;
    bcmg_csr A = nullptr;
    bcmg::Boot B;
};

struct bcmg_server_s {
    bcmg_ctx ctx = nullptr;
    bcmg::Server *S = nullptr;
};

namespace bcmg {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

} // namespace bcmg

using namespace bcmg;

#define API_BEGIN try {
#define API_END                                                  \
    }                                                            \
    catch (const bcmg::Error &e) {                               \
        set_last_error(e.what());                                \
        return e.code;                                           \
    }                                                            \
    catch (const std::bad_alloc &) {                             \
        set_last_error("host allocation failed");                \
        return BCMG_ENOMEM;                                      \
    }                                                            \
    catch (const std::exception &e) {                            \
        set_last_error(e.what());                                \
        return BCMG_EINVAL;                                      \
    }                                                            \
    return BCMG_OK;

#define REQUIRE(cond, msg)                                       \
    do {                                                         \
        if (!(cond)) throw bcmg::Error(BCMG_EINVAL, msg);        \
    } while (0)

namespace {

/* device of the context current, its pool serving the call's allocations */
struct CtxScope {
    bcmg::PoolScope pool;
    explicit CtxScope(bcmg_ctx ctx) : pool(&ctx->c.pool) { BCMG_HIP(hipSetDevice(ctx->c.device)); }
};

void check_csr_host(int n, int nnz, const int *row, const int *col)
{
    REQUIRE(n > 0 && nnz >= n && row && col, "CSR: bad sizes or null pointers");
    REQUIRE(row[0] == 0 && row[n] == nnz, "CSR: row pointer must start at 0 and end at nnz");
    for (int i = 0; i < n; i++) {
        REQUIRE(row[i + 1] >= row[i], "CSR: row pointer not monotone");
        bool diag = false;
        for (int j = row[i]; j < row[i + 1]; j++) {
            REQUIRE(col[j] >= 0 && col[j] < n, "CSR: column index out of range");
            REQUIRE(j == row[i] || col[j] > col[j - 1], "CSR: columns must be strictly ascending");
            diag = diag || col[j] == i;
        }
        REQUIRE(diag, "CSR: every row needs a diagonal entry");
    }
}

/* from_file: the reference's parser (src/AMG/AMG.cu:480-526) accepts cycle
   types 0..2; a parameter block may also ask for the K-cycle (3), which the
   reference's cycle implements (src/SOLVER/GAMG_cycle.cu:86-99) */
int params_check(const bcmg_params &p, bool from_file = false)
{
    if (p.solver_type < 0 || p.solver_type > 2) return BCMG_ESETTINGS;
    if (p.aggrtype != 0 && p.aggrtype != 1) return BCMG_ESETTINGS;
    if (p.cycle_type < 0 || p.cycle_type > (from_file ? 2 : 3)) return BCMG_ESETTINGS;
    if (p.coarse_solver != 0 && p.coarse_solver != 4) return BCMG_ESETTINGS;
    if (p.relax_type != 0 && p.relax_type != 4) return BCMG_ESETTINGS;
    if (p.max_hrc < 1 || p.aggrsweeps < 1 || p.max_levels < 1 || p.itnlim < 1)
        return BCMG_EINVAL;
    if (p.prerelax_sweeps < 0 || p.postrelax_sweeps < 0 || p.relaxnumber_coarse < 0)
        return BCMG_EINVAL;
    return BCMG_OK;
}

const Level &get_level(bcmg_boot B, int k, int l)
{
    REQUIRE(B, "null composite");
    REQUIRE(k >= 0 && k < (int)B->B.H.size(), "hierarchy index out of range");
    REQUIRE(l >= 0 && l < B->B.H[k]->num_levels(), "level index out of range");
    return B->B.H[k]->L[l];
}

/* Device address of `scratch` on the current device, resolved once per
   device and kept: every GPU holds its own copy at its own address. */
double *scratch_address(int device)
{
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static double *addr[MAX_DEVICES] = {};
    REQUIRE(device >= 0 && device < MAX_DEVICES, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        BCMG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        REQUIRE(p != nullptr, "no device address for the relaxation scratch");
        addr[device] = static_cast<double *>(p);
    }
    return addr[device];
}

void d2h(bcmg_ctx ctx, void *dst, const void *src, size_t bytes)
{
    if (!bytes || !dst) return;
    BCMG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
}

} // namespace

extern "C" {

const char *bcmg_last_error(void) { return g_last_error.c_str(); }

/* src/AMG/AMG.cu:480-526 with the line parser of src/AMG/utility.cu:7-25:
   one value per line, text after '%' ignored. */
int bcmg_params_from_file(const char *path, bcmg_params *out)
{
    if (!path || !out) return BCMG_EINVAL;
    FILE *fp = std::fopen(path, "r");
    if (!fp) {
        set_last_error("Setting file not found!");
        return BCMG_ESETTINGS;
    }
    char buf[1025];
    bool ok = true;
    auto token = [&]() -> const char * {
        if (!std::fgets(buf, 1024, fp)) {
            ok = false;
            return nullptr;
        }
        char *save = nullptr;
        const char *t = strtok_r(buf, "%", &save);
        if (!t) ok = false;
        return t;
    };
    auto gi = [&]() -> int {
        int v = 0;
        const char *t = token();
        if (!t || std::sscanf(t, "%d", &v) != 1) ok = false;
        return v;
    };
    auto gd = [&]() -> double {
        double v = 0;
        const char *t = token();
        if (!t || std::sscanf(t, "%lf", &v) != 1) ok = false;
        return v;
    };
    bcmg_params p;
    std::memset(&p, 0, sizeof(p));
    p.solver_type = gi();
    p.max_hrc = gi();
    p.conv_ratio = gd();
    p.matchtype = 3;
    p.aggrsweeps = gi();
    p.aggrtype = gi();
    p.max_levels = gi();
    p.cycle_type = gi();
    p.coarse_solver = gi();
    p.relax_type = gi();
    p.relaxnumber_coarse = gi();
    p.prerelax_sweeps = gi();
    p.postrelax_sweeps = gi();
    p.itnlim = gi();
    p.rtol = gd();
    std::fclose(fp);
    if (!ok) {
        set_last_error("settings file: expected 14 lines, one value each");
        return BCMG_ESETTINGS;
    }
    const int rc = params_check(p, true);
    if (rc == BCMG_ESETTINGS) {
        set_last_error("settings file: value outside the allowed set");
        return rc;
    }
    *out = p;
    return BCMG_OK;
}

namespace {

int ctx_create(int device, hipStream_t ext, bcmg_ctx *out)
{
    if (!out) return BCMG_EINVAL;
    *out = nullptr;
    API_BEGIN
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        throw Error(BCMG_ENODEV, "no HIP device with that index");
    hipDeviceProp_t prop;
    BCMG_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw Error(BCMG_ENODEV, std::string("library is built for gfx950, device is ") +
                                     prop.gcnArchName);
    auto *ctx = new bcmg_ctx_s();
    try {
        ctx->c.device = device;
        BCMG_HIP(hipSetDevice(device));
        ctx->scratch = scratch_address(device);
        if (ext) {
            int sdev = -1;
            BCMG_HIP(hipStreamGetDevice(ext, &sdev));
            if (sdev != device) throw Error(BCMG_EINVAL, "the stream belongs to another device");
            ctx->c.stream = ext;
            ctx->c.own_stream = false;
        } else {
            BCMG_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
        }
        BCMG_HIP(hipHostMalloc((void **)&ctx->c.h_scalars, 64 * sizeof(double), hipHostMallocDefault));
        BCMG_HIP(hipHostMalloc((void **)&ctx->c.h_ints, 256 * sizeof(int), hipHostMallocDefault));
        ctx->c.ensure_reductions(1);
        BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    } catch (...) {
        delete ctx;
        throw;
    }
    *out = ctx;
    API_END
}

} // namespace

int bcmg_ctx_create(int device, bcmg_ctx *out) { return ctx_create(device, nullptr, out); }

int bcmg_ctx_create_on_stream(int device, void *stream, bcmg_ctx *out)
{
    if (!stream) {
        set_last_error("a created stream is required (not the null stream)");
        return BCMG_EINVAL;
    }
    return ctx_create(device, static_cast<hipStream_t>(stream), out);
}

int bcmg_ctx_destroy(bcmg_ctx ctx)
{
    if (!ctx) return BCMG_OK;
    API_BEGIN
    CtxScope scope_(ctx);
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
    API_END
}

int bcmg_sync(bcmg_ctx ctx)
{
    if (!ctx) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    API_END
}

int bcmg_dmalloc(bcmg_ctx ctx, size_t bytes, void **ptr)
{
    if (!ctx || !ptr) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(BCMG_ENOMEM, "device allocation failed");
    }
    API_END
}

int bcmg_dfree(bcmg_ctx ctx, void *ptr)
{
    if (!ctx) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    if (ptr) BCMG_HIP(hipFree(ptr));
    API_END
}

int bcmg_memcpy_h2d(bcmg_ctx ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    if (bytes) {
        BCMG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
        BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    }
    API_END
}

int bcmg_memcpy_d2h(bcmg_ctx ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    d2h(ctx, dst, src, bytes);
    API_END
}

int bcmg_memcpy_d2d(bcmg_ctx ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    if (bytes) {
        BCMG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
        BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    }
    API_END
}

int bcmg_csr_upload(bcmg_ctx ctx, int n, int nnz, const int *row, const int *col,
                    const double *val, bcmg_csr *out)
{
    if (!ctx || !out || !val) return BCMG_EINVAL;
    *out = nullptr;
    API_BEGIN
    check_csr_host(n, nnz, row, col);
    CtxScope scope_(ctx);
    auto *h = new bcmg_csr_s();
    try {
        h->ctx = ctx;
        DevCsr &A = h->A;
        A.n = A.m = n;
        A.nnz = nnz;
        A.row.alloc(n + 1);
        A.col.alloc(nnz);
        A.val.alloc(nnz);
        hipStream_t s = ctx->c.stream;
        BCMG_HIP(hipMemcpyAsync(A.row.get(), row, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
        BCMG_HIP(hipMemcpyAsync(A.col.get(), col, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
        BCMG_HIP(hipMemcpyAsync(A.val.get(), val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
        BCMG_HIP(hipStreamSynchronize(s));
        analyze_rows(ctx->c, A);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    API_END
}

int bcmg_csr_download(bcmg_csr A, int *row, int *col, double *val)
{
    if (!A) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(A->ctx);
    d2h(A->ctx, row, A->A.row.get(), sizeof(int) * (A->A.n + 1));
    d2h(A->ctx, col, A->A.col.get(), sizeof(int) * A->A.nnz);
    d2h(A->ctx, val, A->A.val.get(), sizeof(double) * A->A.nnz);
    API_END
}

int bcmg_csr_info(bcmg_csr A, int *n, int *nnz)
{
    if (!A) return BCMG_EINVAL;
    if (n) *n = A->A.n;
    if (nnz) *nnz = A->A.nnz;
    return BCMG_OK;
}

int bcmg_csr_column_bytes(bcmg_csr A, int *bytes)
{
    if (!A || !bytes) return BCMG_EINVAL;
    *bytes = A->A.use_pattern() ? 0 : A->A.use_narrow() ? 2 : 4;
    return BCMG_OK;
}

int bcmg_csr_set_columns(bcmg_csr A, int bytes)
{
    if (!A) return BCMG_EINVAL;
    if (bytes == 4) {
        A->A.col_mode = 0;
    } else if (bytes == 2) {
        if (!A->A.narrow) {
            set_last_error("bcmg_csr_set_columns: no 16-bit column copy for this matrix");
            return BCMG_EUNSUP;
        }
        A->A.col_mode = 1;
    } else if (bytes == 0) {
        if (A->A.pk == 0) {
            set_last_error("bcmg_csr_set_columns: no stencil pattern for this matrix");
            return BCMG_EUNSUP;
        }
        A->A.col_mode = 2;
    } else {
        return BCMG_EINVAL;
    }
    return BCMG_OK;
}

int bcmg_csr_free(bcmg_csr A)
{
    if (!A) return BCMG_OK;
    API_BEGIN
    CtxScope scope_(A->ctx);
    BCMG_HIP(hipStreamSynchronize(A->ctx->c.stream));
    delete A;
    API_END
}

int bcmg_setup(bcmg_ctx ctx, bcmg_csr A, const bcmg_params *p, bcmg_boot *out)
{
    if (!ctx || !A || !p || !out) return BCMG_EINVAL;
    *out = nullptr;
    const int rc = params_check(*p);
    if (rc) {
        set_last_error("invalid settings");
        return rc;
    }
    API_BEGIN
    CtxScope scope_(ctx);
    auto *h = new bcmg_boot_s();
    try {
        h->ctx = ctx;
        h->A = A;
        build_boot(ctx->c, A->A, *p, h->B);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    API_END
}

int bcmg_set_precision(bcmg_boot B, int precision)
{
    if (!B) return BCMG_EINVAL;
    if (precision != BCMG_PREC_FP64 && precision != BCMG_PREC_FP32 &&
        precision != BCMG_PREC_BF16) {
        set_last_error("precision must be BCMG_PREC_FP64, BCMG_PREC_FP32 or BCMG_PREC_BF16");
        return BCMG_EINVAL;
    }
    API_BEGIN
    CtxScope scope_(B->ctx);
    set_precision(B->ctx->c, B->B, precision);
    API_END
}

int bcmg_boot_free(bcmg_boot B)
{
    if (!B) return BCMG_OK;
    API_BEGIN
    CtxScope scope_(B->ctx);
    BCMG_HIP(hipStreamSynchronize(B->ctx->c.stream));
    delete B;
    API_END
}

int bcmg_fcg(bcmg_ctx ctx, bcmg_boot B, const double *b, double *x, int *iters, double *resid)
{
    if (!ctx || !B || !b || !x) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    fcg_solve(ctx->c, B->B, b, x, iters, resid, nullptr);
    API_END
}

int bcmg_fcg_multi(bcmg_ctx ctx, bcmg_boot B, int k, const double *const *b_dev,
                   double *const *x_dev, int *iters, double *resid, int *status)
{
    if (!ctx || !B || k < 0 || (k > 0 && (!b_dev || !x_dev || !iters || !resid || !status)))
        return BCMG_EINVAL;
    for (int j = 0; j < k; j++)
        if (!b_dev[j] || !x_dev[j]) {
            set_last_error("bcmg_fcg_multi: null column pointer");
            return BCMG_EINVAL;
        }
    API_BEGIN
    CtxScope scope_(ctx);
    fcg_multi(ctx->c, B->B, k, b_dev, x_dev, iters, resid, status);
    API_END
}

int bcmg_server_create(bcmg_ctx ctx, size_t hbm_budget_bytes, int slots_per_operator,
                       bcmg_server *out)
{
    if (!ctx || !out) return BCMG_EINVAL;
    *out = nullptr;
    API_BEGIN
    CtxScope scope_(ctx);
    auto *h = new bcmg_server_s();
    try {
        h->ctx = ctx;
        h->S = server_create(ctx->c, hbm_budget_bytes, slots_per_operator);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    API_END
}

int bcmg_server_destroy(bcmg_server S)
{
    if (!S) return BCMG_OK;
    API_BEGIN
    CtxScope scope_(S->ctx);
    server_destroy(S->S);
    delete S;
    API_END
}

int bcmg_server_add_operator(bcmg_server S, bcmg_boot B, int *op_id)
{
    if (!S || !B || !op_id) return BCMG_EINVAL;
    if (B->ctx != S->ctx) {
        set_last_error("the composite belongs to another context");
        return BCMG_EINVAL;
    }
    API_BEGIN
    CtxScope scope_(S->ctx);
    *op_id = server_add(*S->S, B->B);
    API_END
}

int bcmg_server_submit(bcmg_server S, int op_id, const double *b_dev, double *x_dev,
                       long long *req_id)
{
    if (!S || !b_dev || !x_dev || !req_id) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(S->ctx);
    *req_id = server_submit(*S->S, op_id, b_dev, x_dev);
    API_END
}

int bcmg_server_step(bcmg_server S, int *completed)
{
    if (!S) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(S->ctx);
    const int c = server_step(*S->S);
    if (completed) *completed = c;
    API_END
}

int bcmg_server_result(bcmg_server S, long long req_id, int *done, int *iters, double *resid,
                       int *status, double *latency_ms)
{
    if (!S) return BCMG_EINVAL;
    if (!server_result(*S->S, req_id, done, iters, resid, status, latency_ms)) {
        set_last_error("unknown (or already returned) request id");
        return BCMG_EINVAL;
    }
    return BCMG_OK;
}

int bcmg_server_completed(bcmg_server S, long long *ids, int cap, int *count)
{
    if (!S || !count || cap < 0 || (cap > 0 && !ids)) return BCMG_EINVAL;
    *count = server_completed(*S->S, ids, cap);
    return BCMG_OK;
}

int bcmg_server_info(bcmg_server S, int *queued, int *running, size_t *bytes_in_use, int *lanes)
{
    if (!S) return BCMG_EINVAL;
    server_info(*S->S, queued, running, bytes_in_use, lanes);
    return BCMG_OK;
}

int bcmg_solve(int n, int nnz, const int *row, const int *col, const double *val,
               const bcmg_params *p, const double *rhs, double *x, int *iters, double *resid,
               double *setup_ms, double *solve_ms)
{
    if (!p || !rhs || !x) return BCMG_EINVAL;
    bcmg_ctx ctx = nullptr;
    bcmg_csr A = nullptr;
    bcmg_boot B = nullptr;
    double *db = nullptr, *dx = nullptr;
    int rc = bcmg_ctx_create(0, &ctx);
    if (rc) return rc;
    auto cleanup = [&]() {
        if (db) bcmg_dfree(ctx, db);
        if (dx) bcmg_dfree(ctx, dx);
        bcmg_boot_free(B);
        bcmg_csr_free(A);
        bcmg_ctx_destroy(ctx);
    };
    rc = bcmg_csr_upload(ctx, n, nnz, row, col, val, &A);
    if (!rc) rc = bcmg_dmalloc(ctx, sizeof(double) * n, (void **)&db);
    if (!rc) rc = bcmg_dmalloc(ctx, sizeof(double) * n, (void **)&dx);
    if (!rc) rc = bcmg_memcpy_h2d(ctx, db, rhs, sizeof(double) * n);
    auto t0 = std::chrono::steady_clock::now();
    if (!rc) rc = bcmg_setup(ctx, A, p, &B);
    auto t1 = std::chrono::steady_clock::now();
    if (!rc) rc = bcmg_fcg(ctx, B, db, dx, iters, resid);
    auto t2 = std::chrono::steady_clock::now();
    if (!rc) rc = bcmg_memcpy_d2h(ctx, x, dx, sizeof(double) * n);
    if (setup_ms) *setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (solve_ms) *solve_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    const std::string msg = g_last_error;
    cleanup();
    if (rc) g_last_error = msg;
    return rc;
}

int bcmg_boot_info(bcmg_boot B, int *n_hrc, double *estimated_ratio)
{
    if (!B) return BCMG_EINVAL;
    if (n_hrc) *n_hrc = (int)B->B.H.size();
    if (estimated_ratio) *estimated_ratio = B->B.estimated_ratio;
    return BCMG_OK;
}

int bcmg_hier_info(bcmg_boot B, int k, int *levels, double *op_cmplx, double *op_wcmplx,
                   double *avg_cratio, double *ratio_after)
{
    if (!B || k < 0 || k >= (int)B->B.H.size()) return BCMG_EINVAL;
    const Hierarchy &H = *B->B.H[k];
    if (levels) *levels = H.num_levels();
    if (op_cmplx) *op_cmplx = H.op_cmplx;
    if (op_wcmplx) *op_wcmplx = H.op_wcmplx;
    if (avg_cratio) *avg_cratio = H.avg_cratio;
    if (ratio_after)
        *ratio_after = k < (int)B->B.ratio_after.size() ? B->B.ratio_after[k] : B->B.estimated_ratio;
    return BCMG_OK;
}

int bcmg_time_setup(bcmg_ctx ctx, int enable)
{
    if (!ctx) return BCMG_EINVAL;
    ctx->c.time_setup = enable != 0;
    return BCMG_OK;
}

int bcmg_setup_timing(bcmg_boot B, int k, bcmg_setup_times *out)
{
    if (!B || !out || k < -1 || k >= (int)B->B.H.size()) return BCMG_EINVAL;
    double t[PH_COUNT] = {0};
    for (int h = 0; h < (int)B->B.t_phase.size(); h++) {
        if (k >= 0 && h != k) continue;
        const std::vector<double> &v = B->B.t_phase[h];
        for (int i = 0; i < PH_COUNT && i < (int)v.size(); i++) t[i] += v[i];
    }
    out->weights_ms = t[PH_WEIGHTS];
    out->matching_ms = t[PH_MATCHING];
    out->aggregation_ms = t[PH_AGGREGATION];
    out->galerkin_ms = t[PH_GALERKIN];
    out->smoother_ms = t[PH_SMOOTHER];
    out->inner_ms = t[PH_INNER];
    out->hierarchy_ms = t[PH_HIER];
    out->total_ms = k < 0 ? B->B.t_setup_ms : t[PH_HIER];
    return BCMG_OK;
}

int bcmg_fcg_restarts(bcmg_boot B, int *restarts)
{
    if (!B || !restarts) return BCMG_EINVAL;
    *restarts = B->B.restarts;
    return BCMG_OK;
}

int bcmg_hier_levels(bcmg_boot B, int k)
{
    if (!B || k < 0 || k >= (int)B->B.H.size()) return BCMG_EINVAL;
    return B->B.H[k]->num_levels();
}

int bcmg_level_size(bcmg_boot B, int k, int l, int *n, int *nnz)
{
    API_BEGIN
    const Level &lv = get_level(B, k, l);
    if (n) *n = lv.mat().n;
    if (nnz) *nnz = lv.mat().nnz;
    API_END
}

int bcmg_level_A(bcmg_boot B, int k, int l, int *row, int *col, double *val)
{
    API_BEGIN
    const Level &lv = get_level(B, k, l);
    CtxScope scope_(B->ctx);
    const DevCsr &A = lv.mat();
    d2h(B->ctx, row, A.row.get(), sizeof(int) * (A.n + 1));
    d2h(B->ctx, col, A.col.get(), sizeof(int) * A.nnz);
    d2h(B->ctx, val, A.val.get(), sizeof(double) * A.nnz);
    API_END
}

int bcmg_level_P(bcmg_boot B, int k, int l, int *agg, double *p)
{
    API_BEGIN
    const Level &lv = get_level(B, k, l);
    REQUIRE(l + 1 < B->B.H[k]->num_levels(), "the coarsest level has no prolongator");
    CtxScope scope_(B->ctx);
    d2h(B->ctx, agg, lv.agg.get(), sizeof(int) * lv.mat().n);
    d2h(B->ctx, p, lv.p.get(), sizeof(double) * lv.mat().n);
    API_END
}

int bcmg_level_Pinfo(bcmg_boot B, int k, int l, int *nnz, int *smoothed, double *omega)
{
    API_BEGIN
    const Level &lv = get_level(B, k, l);
    REQUIRE(l + 1 < B->B.H[k]->num_levels(), "the coarsest level has no prolongator");
    if (nnz) *nnz = lv.smoothed ? lv.Pneg.nnz : lv.mat().n;
    if (smoothed) *smoothed = lv.smoothed ? 1 : 0;
    if (omega) *omega = lv.omega;
    API_END
}

int bcmg_level_Pcsr(bcmg_boot B, int k, int l, int *row, int *col, double *val)
{
    API_BEGIN
    const Level &lv = get_level(B, k, l);
    REQUIRE(l + 1 < B->B.H[k]->num_levels(), "the coarsest level has no prolongator");
    CtxScope scope_(B->ctx);
    const int n = lv.mat().n;
    if (lv.smoothed) {
        const DevCsr &P = lv.Pneg;
        d2h(B->ctx, row, P.row.get(), sizeof(int) * (n + 1));
        d2h(B->ctx, col, P.col.get(), sizeof(int) * P.nnz);
        if (val) {
            d2h(B->ctx, val, P.val.get(), sizeof(double) * P.nnz);
            for (int e = 0; e < P.nnz; e++) val[e] = -val[e];
        }
    } else {
        if (row)
            for (int i = 0; i <= n; i++) row[i] = i;
        d2h(B->ctx, col, lv.agg.get(), sizeof(int) * n);
        d2h(B->ctx, val, lv.p.get(), sizeof(double) * n);
    }
    API_END
}

int bcmg_level_DM(bcmg_boot B, int k, int l, double *D, double *M, int *hasD, int *hasM)
{
    API_BEGIN
    const Level &lv = get_level(B, k, l);
    CtxScope scope_(B->ctx);
    if (hasD) *hasD = lv.hasD;
    if (hasM) *hasM = lv.hasM;
    if (lv.hasD) d2h(B->ctx, D, lv.D.get(), sizeof(double) * lv.mat().n);
    if (lv.hasM) d2h(B->ctx, M, lv.M.get(), sizeof(double) * lv.mat().n);
    API_END
}

int bcmg_spmv(bcmg_ctx ctx, bcmg_csr A, const double *x, double *y)
{
    if (!ctx || !A || !x || !y) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    launch_rows(ctx->c.stream, A->A, OP_SPMV, x, nullptr, nullptr, y);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    API_END
}

int bcmg_residual(bcmg_ctx ctx, bcmg_csr A, const double *x, const double *f, double *y)
{
    if (!ctx || !A || !x || !f || !y) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    launch_rows(ctx->c.stream, A->A, OP_RESID, x, f, nullptr, y);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    API_END
}

int bcmg_relax(bcmg_ctx ctx, bcmg_csr A, const double *d, const double *f, double *u,
               double *u2, int sweeps)
{
    if (!ctx || !A || !d || !f || !u || !u2 || sweeps < 0) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    double *x = u, *y = u2;
    for (int k = 0; k < sweeps; k++) {
        launch_rows(ctx->c.stream, A->A, OP_JACOBI, x, f, d, y);
        std::swap(x, y);
    }
    if (x != u) launch_copy(ctx->c.stream, A->A.n, x, u);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    API_END
}

int bcmg_relax_noalloc(bcmg_ctx ctx, bcmg_csr A, const double *d, const double *f, double *u,
                       int sweeps)
{
    if (!ctx || !A || !d || !f || !u || sweeps < 0) return BCMG_EINVAL;
    if (A->A.n > BCMG_RELAX_NOALLOC_MAX_ROWS) {
        set_last_error("bcmg_relax_noalloc: " + std::to_string(A->A.n) +
                       " rows exceed the static scratch of " +
                       std::to_string(BCMG_RELAX_NOALLOC_MAX_ROWS));
        return BCMG_EINVAL;
    }
    API_BEGIN
    /* no pool scope: nothing here allocates; the kernels are bcmg_relax's */
    BCMG_HIP(hipSetDevice(ctx->c.device));
    double *x = u, *y = ctx->scratch;
    for (int k = 0; k < sweeps; k++) {
        launch_rows(ctx->c.stream, A->A, OP_JACOBI, x, f, d, y);
        std::swap(x, y);
    }
    if (x != u) launch_copy(ctx->c.stream, A->A.n, x, u);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    API_END
}

int bcmg_weights(bcmg_ctx ctx, bcmg_csr A, const double *w, double *W)
{
    if (!ctx || !A || !w || !W) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    match_weights(ctx->c, A->A, w, W);
    API_END
}

int bcmg_matching(bcmg_ctx ctx, bcmg_csr A, const double *W, int *M_raw, int *M_sym,
                  int *n_unsym)
{
    if (!ctx || !A || !W || !M_raw || !M_sym) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    const int nu = match_suitor(ctx->c, A->A, W, M_raw, M_sym, true);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    if (n_unsym) *n_unsym = nu;
    API_END
}

int bcmg_aggregate(bcmg_ctx ctx, int n, const int *M_sym, const double *w, int *agg, double *p,
                   int *n_coarse)
{
    if (!ctx || n <= 0 || !M_sym || !w || !agg || !p || !n_coarse) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    DBuf<int> members(2 * (size_t)n);
    *n_coarse = aggregate_pairs(ctx->c, n, M_sym, w, agg, p, members.get(), true);
    API_END
}

int bcmg_galerkin(bcmg_ctx ctx, bcmg_csr A, const int *agg, const double *p, int n_coarse,
                  bcmg_csr *Ac)
{
    if (!ctx || !A || !agg || !p || n_coarse <= 0 || !Ac) return BCMG_EINVAL;
    *Ac = nullptr;
    API_BEGIN
    CtxScope scope_(ctx);
    DBuf<int> members(2 * (size_t)n_coarse);
    members_from_agg(ctx->c, A->A.n, agg, n_coarse, members.get());
    auto *h = new bcmg_csr_s();
    try {
        h->ctx = ctx;
        galerkin_pairwise(ctx->c, A->A, agg, p, members.get(), n_coarse, h->A);
        BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    } catch (...) {
        delete h;
        throw;
    }
    *Ac = h;
    API_END
}

int bcmg_spgemm(bcmg_ctx ctx, int nl, int ml, int mr, const int *lrow, const int *lcol,
                const double *lval, const int *rrow, const int *rcol, const double *rval,
                long long slots_per_pass, bcmg_csr *C)
{
    if (!ctx || nl < 0 || ml < 0 || mr < 0 || !lrow || !rrow || !C) return BCMG_EINVAL;
    *C = nullptr;
    API_BEGIN
    CtxScope scope_(ctx);
    /* the operands' indices are checked on the host before any kernel reads
       through them */
    std::vector<int> hl((size_t)nl + 1), hr((size_t)ml + 1);
    d2h(ctx, hl.data(), lrow, sizeof(int) * hl.size());
    d2h(ctx, hr.data(), rrow, sizeof(int) * hr.size());
    auto check_rows = [](const std::vector<int> &rp) {
        REQUIRE(rp[0] == 0, "CSR: row pointers must start at 0");
        for (size_t i = 1; i < rp.size(); i++)
            REQUIRE(rp[i] >= rp[i - 1], "CSR: row pointers must not decrease");
    };
    check_rows(hl);
    check_rows(hr);
    const int lnnz = hl[nl], rnnz = hr[ml];
    REQUIRE(lnnz == 0 || (lcol && lval), "null left operand arrays");
    REQUIRE(rnnz == 0 || (rcol && rval), "null right operand arrays");
    std::vector<int> cl(lnnz), cr(rnnz);
    d2h(ctx, cl.data(), lcol, sizeof(int) * lnnz);
    d2h(ctx, cr.data(), rcol, sizeof(int) * rnnz);
    for (int c : cl) REQUIRE(c >= 0 && c < ml, "left operand: column index out of range");
    for (int c : cr) REQUIRE(c >= 0 && c < mr, "right operand: column index out of range");
    bool right_unique = true;
    for (int i = 0; i < ml && right_unique; i++)
        for (int e = hr[i] + 1; e < hr[i + 1]; e++)
            if (cr[e] <= cr[e - 1]) {
                right_unique = false;
                break;
            }
    const CsrRef L{nl, ml, lnnz, lrow, lcol, lval};
    const CsrRef R{ml, mr, rnnz, rrow, rcol, rval};
    auto *h = new bcmg_csr_s();
    try {
        h->ctx = ctx;
        spgemm(ctx->c, L, R, h->A, slots_per_pass, right_unique);
        resolve_tiles(ctx->c, {&h->A});
        make_narrow(ctx->c, {&h->A});
    } catch (...) {
        delete h;
        throw;
    }
    *C = h;
    API_END
}

int bcmg_precond_apply(bcmg_ctx ctx, bcmg_boot B, const double *r, double *x)
{
    if (!ctx || !B || !r || !x) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    precond_apply(ctx->c, B->B, r, x, false);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    API_END
}

int bcmg_use_graphs(bcmg_ctx ctx, int enable)
{
    if (!ctx) return BCMG_EINVAL;
    ctx->c.use_graphs = enable != 0;
    return BCMG_OK;
}

int bcmg_time_smoother(bcmg_ctx ctx, int enable)
{
    if (!ctx) return BCMG_EINVAL;
    ctx->c.time_smoother = enable != 0;
    ctx->c.ev_used = 0;
    ctx->c.ev_launches.clear();
    return BCMG_OK;
}

int bcmg_smoother_timing(bcmg_ctx ctx, double *total_ms, int *launches)
{
    if (!ctx) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    BCMG_HIP(hipStreamSynchronize(ctx->c.stream));
    double tot = 0.0;
    int count = 0;
    for (size_t i = 0; i + 1 < ctx->c.ev_used; i += 2) {
        float ms = 0.f;
        BCMG_HIP(hipEventElapsedTime(&ms, ctx->c.ev_pool[i], ctx->c.ev_pool[i + 1]));
        tot += ms;
        count += ctx->c.ev_launches[i / 2];
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = count;
    ctx->c.ev_used = 0;
    ctx->c.ev_launches.clear();
    API_END
}

int bcmg_time_spmv(bcmg_ctx ctx, bcmg_csr A, const double *x, double *y, int reps,
                   double *avg_ms)
{
    if (!ctx || !A || !x || !y || reps < 1 || !avg_ms) return BCMG_EINVAL;
    API_BEGIN
    CtxScope scope_(ctx);
    hipEvent_t e0, e1;
    BCMG_HIP(hipEventCreate(&e0));
    BCMG_HIP(hipEventCreate(&e1));
    launch_rows(ctx->c.stream, A->A, OP_SPMV, x, nullptr, nullptr, y);
    BCMG_HIP(hipEventRecord(e0, ctx->c.stream));
    for (int k = 0; k < reps; k++) launch_rows(ctx->c.stream, A->A, OP_SPMV, x, nullptr, nullptr, y);
    BCMG_HIP(hipEventRecord(e1, ctx->c.stream));
    BCMG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    BCMG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    API_END
}

} // extern "C"
