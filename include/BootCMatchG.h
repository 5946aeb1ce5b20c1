/*
 * BootCMatchG.h - C++ facade with the names, types and error behaviour of
 * BootCMatchG's public surface (src/BootCMatchG.h and what it includes),
 * implemented over the C-ABI of bcmg.h (libbcmg.so, gfx950 kernels).
 *
 * Code written against the reference's solve_example (src/solve_example.cu)
 * compiles unchanged:
 *     CSR *A_host = readMatrixFromFile(path, 0, false);
 *     vector<vtype> *rhs = Vector::init<vtype>(A_host->n, true, true);
 *     Vector::fillWithValue(rhs, 1.);
 *     params p = AMG::Params::initFromFile(settings_path);
 *     vector<vtype> *x = solve(A_host, p, rhs);
 *     Vector::printOnFile(x, "x.txt");
 *     Vector::free(x); CSRm::free(A_host);
 *
 * Like the reference, the facade keeps one process-wide device context
 * (src/utility/handles.cu), prints the reference's #META lines and timings,
 * and terminates the process where the reference does (bad settings file:
 * exit(-1), src/AMG/AMG.cu:8-18,485-488; zero initial residual: exit(1),
 * src/SOLVER/FCG.cu:459-462). Define BCMG_FACADE_THROW before including this
 * header to get a std::runtime_error instead.
 */
#ifndef BCMG_BOOTCMATCHG_H
#define BCMG_BOOTCMATCHG_H

#include "bcmg.h"

#include <algorithm>
#include <cassert>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

/* src/utility/setting.h:20-25 */
#define vtype double
#define itype int
#define stype int

typedef bcmg_params params;

/* src/matrix/vector.cu:12-17 */
template <typename T>
struct vector {
    int n;
    bool on_the_device;
    T *val;
};

/* src/matrix/CSR.h:13-28 (descr is kept as an opaque pointer) */
typedef struct {
    stype nnz;
    stype n;
    stype m;
    bool on_the_device;
    bool is_symmetric;
    vtype *val;
    itype *col;
    itype *row;
    void *descr;
} CSR;

namespace bcmg_facade {

inline void fail(int code, const char *what)
{
#ifdef BCMG_FACADE_THROW
    throw std::runtime_error(std::string(what) + ": " + bcmg_last_error());
#else
    std::cout << what << ": " << bcmg_last_error() << "\n";
    std::exit(code);
#endif
}

inline void check(int rc, const char *what)
{
    if (rc != BCMG_OK) fail(rc == BCMG_EZERORHS ? 1 : -1, what);
}

/* the process-wide context (the reference's handles, src/utility/handles.cu) */
inline bcmg_ctx ctx()
{
    static bcmg_ctx c = nullptr;
    if (!c) check(bcmg_ctx_create(0, &c), "device context");
    return c;
}

/* #META <name> int|float <value> (src/utility/eval.cu:17-26) */
inline void meta(const char *name, double v, int is_float)
{
    if (is_float)
        std::printf("#META %s float %le\n", name, v);
    else
        std::printf("#META %s int %d\n", name, (int)v);
}

} // namespace bcmg_facade

namespace AMG {
namespace Params {
/* src/AMG/AMG.cu:480-526 */
inline params initFromFile(const char *path)
{
    params p;
    const int rc = bcmg_params_from_file(path, &p);
    if (rc != BCMG_OK) {
        std::cout << (rc == BCMG_ESETTINGS && std::string(bcmg_last_error()).find("not found") !=
                                                   std::string::npos
                          ? "Setting file not found!"
                          : bcmg_last_error())
                  << "\n";
#ifdef BCMG_FACADE_THROW
        throw std::runtime_error(bcmg_last_error());
#else
        std::exit(-1);
#endif
    }
    return p;
}
} // namespace Params
} // namespace AMG

namespace Vector {

template <typename T>
vector<T> *init(int n, bool allocate_mem, bool on_the_device)
{
    auto *v = new vector<T>();
    v->n = n;
    v->on_the_device = on_the_device;
    v->val = nullptr;
    if (allocate_mem) {
        if (on_the_device)
            bcmg_facade::check(bcmg_dmalloc(bcmg_facade::ctx(), sizeof(T) * (size_t)n,
                                            (void **)&v->val),
                               "device allocation");
        else
            v->val = static_cast<T *>(std::malloc(sizeof(T) * (size_t)n));
    }
    return v;
}

template <typename T>
vector<T> *copyToHost(vector<T> *v_d)
{
    vector<T> *h = init<T>(v_d->n, true, false);
    bcmg_facade::check(bcmg_memcpy_d2h(bcmg_facade::ctx(), h->val, v_d->val, sizeof(T) * v_d->n),
                       "copy to host");
    return h;
}

template <typename T>
vector<T> *copyToDevice(vector<T> *v)
{
    vector<T> *d = init<T>(v->n, true, true);
    bcmg_facade::check(bcmg_memcpy_h2d(bcmg_facade::ctx(), d->val, v->val, sizeof(T) * v->n),
                       "copy to device");
    return d;
}

template <typename T>
void fillWithValue(vector<T> *v, T value)
{
    if (!v->on_the_device) {
        std::fill(v->val, v->val + v->n, value);
        return;
    }
    std::vector<T> tmp(v->n, value);
    bcmg_facade::check(bcmg_memcpy_h2d(bcmg_facade::ctx(), v->val, tmp.data(), sizeof(T) * v->n),
                       "fill");
}

template <typename T>
void free(vector<T> *v)
{
    if (!v) return;
    if (v->on_the_device)
        bcmg_dfree(bcmg_facade::ctx(), v->val);
    else
        std::free(v->val);
    delete v;
}

/* src/matrix/vector.cu:444-466: one value per line, "%.14e" */
template <typename T>
void printOnFile(vector<T> *v_, const char *path)
{
    FILE *file = std::fopen(path, "w");
    assert(file);
    vector<T> *v = v_->on_the_device ? copyToHost(v_) : v_;
    for (int i = 0; i < v->n; i++) std::fprintf(file, "%.14e\n", v->val[i]);
    if (v_->on_the_device) free(v);
    std::fclose(file);
}

} // namespace Vector

namespace CSRm {

/* src/matrix/CSR.cu:21-70 */
inline CSR *init(stype n, stype m, stype nnz, bool allocate_mem, bool on_the_device,
                 bool is_symmetric)
{
    assert(n > 0 && m > 0 && nnz >= 0);
    CSR *A = new CSR();
    A->nnz = nnz;
    A->n = n;
    A->m = m;
    A->on_the_device = on_the_device;
    A->is_symmetric = is_symmetric;
    A->val = nullptr;
    A->col = nullptr;
    A->row = nullptr;
    A->descr = nullptr;
    if (allocate_mem) {
        if (on_the_device) {
            bcmg_ctx c = bcmg_facade::ctx();
            bcmg_facade::check(bcmg_dmalloc(c, sizeof(vtype) * nnz, (void **)&A->val), "alloc");
            bcmg_facade::check(bcmg_dmalloc(c, sizeof(itype) * nnz, (void **)&A->col), "alloc");
            bcmg_facade::check(bcmg_dmalloc(c, sizeof(itype) * (n + 1), (void **)&A->row), "alloc");
        } else {
            A->val = static_cast<vtype *>(std::malloc(sizeof(vtype) * nnz));
            A->col = static_cast<itype *>(std::malloc(sizeof(itype) * nnz));
            A->row = static_cast<itype *>(std::malloc(sizeof(itype) * (n + 1)));
        }
    }
    return A;
}

inline void free(CSR *A)
{
    if (!A) return;
    if (A->on_the_device) {
        bcmg_ctx c = bcmg_facade::ctx();
        bcmg_dfree(c, A->val);
        bcmg_dfree(c, A->col);
        bcmg_dfree(c, A->row);
    } else {
        std::free(A->val);
        std::free(A->col);
        std::free(A->row);
    }
    delete A;
}

/* src/matrix/CSR.cu:194-217 */
inline CSR *copyToDevice(CSR *A)
{
    assert(!A->on_the_device);
    CSR *d = init(A->n, A->m, A->nnz, true, true, A->is_symmetric);
    bcmg_ctx c = bcmg_facade::ctx();
    bcmg_facade::check(bcmg_memcpy_h2d(c, d->val, A->val, sizeof(vtype) * A->nnz), "copy");
    bcmg_facade::check(bcmg_memcpy_h2d(c, d->col, A->col, sizeof(itype) * A->nnz), "copy");
    bcmg_facade::check(bcmg_memcpy_h2d(c, d->row, A->row, sizeof(itype) * (A->n + 1)), "copy");
    return d;
}

/* src/matrix/CSR.cu:245-269 */
inline CSR *copyToHost(CSR *A_d)
{
    assert(A_d->on_the_device);
    CSR *h = init(A_d->n, A_d->m, A_d->nnz, true, false, A_d->is_symmetric);
    bcmg_ctx c = bcmg_facade::ctx();
    bcmg_facade::check(bcmg_memcpy_d2h(c, h->val, A_d->val, sizeof(vtype) * A_d->nnz), "copy");
    bcmg_facade::check(bcmg_memcpy_d2h(c, h->col, A_d->col, sizeof(itype) * A_d->nnz), "copy");
    bcmg_facade::check(bcmg_memcpy_d2h(c, h->row, A_d->row, sizeof(itype) * (A_d->n + 1)), "copy");
    return h;
}

} // namespace CSRm

namespace bcmg_facade {

/* one coordinate entry of a matrix file (0-based) */
struct Coo {
    int r, c;
    double v;
};

/* COO entries to a host CSR of n rows and m columns: columns ascending within
   a row, duplicates summed (the reference's readers keep file order and
   duplicates, src/matrix/matrixIO.cu:145-170, 237-261, which its kernels then
   read as unsorted rows; the library requires sorted, unique columns) */
inline CSR *csr_from_coo(std::vector<Coo> &e, long long n, long long m, bool symmetric);

inline void read_fail(FILE *fp, const char *what)
{
    if (fp) std::fclose(fp);
    fail(-1, what);
}

/* MTX2 (src/matrix/matrixIO.cu:194-289): "rows nnz", then the nnz 1-based row
   indices, the nnz 1-based column indices and the nnz values, whitespace
   separated; square with n = rows */
inline CSR *read_mtx2(FILE *fp)
{
    long long n = 0, L = 0;
    if (std::fscanf(fp, "%lld %lld", &n, &L) != 2 || n <= 0 || L < 0 || n > 0x7fffffff ||
        L > 0x7fffffff)
        read_fail(fp, "readMatrixFromFile: bad MTX2 header");
    std::vector<Coo> e((size_t)L);
    for (auto &x : e) {
        long long i = 0;
        if (std::fscanf(fp, "%lld", &i) != 1 || i < 1 || i > n)
            read_fail(fp, "readMatrixFromFile: bad MTX2 row index");
        x.r = (int)(i - 1);
    }
    for (auto &x : e) {
        long long j = 0;
        if (std::fscanf(fp, "%lld", &j) != 1 || j < 1 || j > n)
            read_fail(fp, "readMatrixFromFile: bad MTX2 column index");
        x.c = (int)(j - 1);
    }
    for (auto &x : e)
        if (std::fscanf(fp, "%lf", &x.v) != 1) read_fail(fp, "readMatrixFromFile: short MTX2 file");
    std::fclose(fp);
    return csr_from_coo(e, n, n, false);
}

} // namespace bcmg_facade

/* Read a matrix file into a host CSR with sorted columns
   (src/matrix/matrixIO.cu:4-30). m_type 0: MatrixMarket "coordinate
   real|integer general|symmetric" (:33-192; a symmetric file is expanded to
   both triangles); m_type 1: MTX2 (:194-289). Duplicates are summed. */
inline CSR *readMatrixFromFile(const char *path, int m_type, bool loadOnDevice = true)
{
    if (m_type != 0 && m_type != 1) {
        std::cout << "You need to specify an input matrix type with the argument -F/--inputype\n";
        std::exit(1);
    }
    FILE *fp = std::fopen(path, "r");
    if (!fp) {
        std::cout << "file not found: " << path << "\n";
        std::exit(-1);
    }
    if (m_type == 1) {
        CSR *A = bcmg_facade::read_mtx2(fp);
        if (loadOnDevice) {
            CSR *d = CSRm::copyToDevice(A);
            CSRm::free(A);
            return d;
        }
        return A;
    }
    char line[1025];
    if (!std::fgets(line, 1024, fp)) bcmg_facade::fail(-1, "readMatrixFromFile: empty file");
    std::string banner(line);
    for (auto &ch : banner) ch = (char)std::tolower(ch);
    const bool symmetric = banner.find("symmetric") != std::string::npos;
    if (banner.find("coordinate") == std::string::npos)
        bcmg_facade::fail(-1, "readMatrixFromFile: only coordinate format");
    do {
        if (!std::fgets(line, 1024, fp)) bcmg_facade::fail(-1, "readMatrixFromFile: no size line");
    } while (line[0] == '%');
    long long M = 0, N = 0, L = 0;
    if (std::sscanf(line, "%lld %lld %lld", &M, &N, &L) != 3)
        bcmg_facade::fail(-1, "readMatrixFromFile: bad size line");
    if (M <= 0 || N <= 0 || L < 0 || M > 0x7fffffff || N > 0x7fffffff || L > 0x3fffffff)
        bcmg_facade::read_fail(fp, "readMatrixFromFile: bad size line");
    std::vector<bcmg_facade::Coo> e;
    e.reserve((size_t)(symmetric ? 2 * L : L));
    for (long long k = 0; k < L; k++) {
        long long i, j;
        double v = 1.0;
        if (!std::fgets(line, 1024, fp) || std::sscanf(line, "%lld %lld %lf", &i, &j, &v) < 2)
            bcmg_facade::read_fail(fp, "readMatrixFromFile: short file");
        if (i < 1 || i > M || j < 1 || j > N)
            bcmg_facade::read_fail(fp, "readMatrixFromFile: index out of range");
        e.push_back({(int)(i - 1), (int)(j - 1), v});
        if (symmetric && i != j) e.push_back({(int)(j - 1), (int)(i - 1), v});
    }
    std::fclose(fp);
    CSR *A = bcmg_facade::csr_from_coo(e, M, N, symmetric);
    if (loadOnDevice) {
        CSR *d = CSRm::copyToDevice(A);
        CSRm::free(A);
        return d;
    }
    return A;
}

inline CSR *bcmg_facade::csr_from_coo(std::vector<Coo> &e, long long n, long long m,
                                      bool symmetric)
{
    std::stable_sort(e.begin(), e.end(), [](const Coo &a, const Coo &b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    size_t u = 0;
    for (size_t k = 0; k < e.size(); k++) {
        if (u > 0 && e[u - 1].r == e[k].r && e[u - 1].c == e[k].c)
            e[u - 1].v += e[k].v;
        else
            e[u++] = e[k];
    }
    e.resize(u);
    CSR *A = CSRm::init((stype)n, (stype)m, (stype)u, true, false, symmetric);
    std::fill(A->row, A->row + n + 1, 0);
    for (size_t k = 0; k < u; k++) {
        A->row[e[k].r + 1]++;
        A->col[k] = e[k].c;
        A->val[k] = e[k].v;
    }
    for (long long i = 0; i < n; i++) A->row[i + 1] += A->row[i];
    return A;
}

namespace bcmg_facade {

/* What the reference prints while it builds the composite, per hierarchy
   (AMG::Hierarchy::printInfo, src/AMG/AMG.cu:241-251; the #META lines and
   TOTAL_MUL_TIME of src/AMG/matchingAggregation.cu:408-421; the estimated
   ratio, src/BOOT/bootstrap.cu:76) and after it (the hierarchy count and
   each coarsest size, src/BOOT/bootstrap.cu:81-86). TOTAL_MUL_TIME is the
   event-timed Galerkin phase (0 unless setup timing was on); a hierarchy
   that shares the first one's storage was not rebuilt and shows 0. */
inline void print_setup(bcmg_boot B)
{
    int nh = 0;
    double ratio = 0;
    check(bcmg_boot_info(B, &nh, &ratio), "composite info");
    std::vector<int> coarsest(nh, 0);
    for (int k = 0; k < nh; k++) {
        int levels = 0;
        double cx = 0, cw = 0, cr = 0, ra = 0;
        check(bcmg_hier_info(B, k, &levels, &cx, &cw, &cr, &ra), "hierarchy info");
        for (int l = 0; l < levels; l++) {
            int n = 0, nnz = 0;
            check(bcmg_level_size(B, k, l, &n, &nnz), "level size");
            const float avg_nnz = (float)nnz / (float)n;
            std::cout << "A" << l << " n: " << n << " nnz: " << nnz << " avg_nnz: " << avg_nnz
                      << "\n";
            coarsest[k] = n;
        }
        std::cout << "\nCurrent cmplx for V-cycle: " << cx;
        std::cout << "\nCurrent cmplx for W-cycle: " << cw;
        std::cout << "\nAverage Coarsening Ratio: " << cr << "\n";
        meta("agg;level_number", levels - 1, 0);
        meta("agg;avg_coarse_ratio", cr, 1);
        meta("agg;OpCmplx", cx, 1);
        bcmg_setup_times t;
        check(bcmg_setup_timing(B, k, &t), "setup timing");
        std::cout << "TOTAL_MUL_TIME: " << (float)t.galerkin_ms << "\n\n";
        std::cout << "estimated_ratio " << ra << "\n";
    }
    std::cout << "Number of hierarchy: " << nh << "\n";
    for (int k = 0; k < nh; k++) std::printf("%d] %d\n", k, coarsest[k]);
}

/* the #META lines and timings of src/SOLVER/solver.cu:106-118 */
inline void print_solve(bcmg_boot B, int num_iter, double residual, float boot_time,
                        float solve_time)
{
    int n_hrc = 0;
    double ratio = 0.0;
    check(bcmg_boot_info(B, &n_hrc, &ratio), "composite info");
    meta("agg;hierarchy_levels_num", n_hrc, 0);
    meta("agg;final_estimated_ratio", ratio, 1);
    meta("sol;num_iteration", num_iter, 0);
    meta("sol;residual", residual, 1);
    meta("time;bootstrap_time", boot_time, 1);
    meta("time;solve_time", solve_time, 1);
    std::cout << "bootstrap_time: " << (boot_time / 1000) << " sec \n";
    std::cout << "solve_time: " << (solve_time / 1000) << " sec \n";
    std::cout << "time_for_iteration: " << (solve_time / num_iter / 1000) << " sec \n";
}

} // namespace bcmg_facade

/* src/SOLVER/solver.cu:20-129: A_host is a host CSR, rhs a device vector;
   returns a new device vector x (FCG from x0 = 0), caller frees it. */
inline vector<vtype> *solve(CSR *A_host, const params p, vector<vtype> *rhs)
{
    using namespace bcmg_facade;
    bcmg_ctx c = ctx();
    assert(!A_host->on_the_device && rhs->on_the_device);
    bcmg_csr A = nullptr;
    bcmg_boot B = nullptr;
    check(bcmg_csr_upload(c, A_host->n, A_host->nnz, A_host->row, A_host->col, A_host->val, &A),
          "copy matrix to device");
    std::cout << "BUILDING....:\n\n";
    check(bcmg_time_setup(c, 1), "setup timing");
    auto t0 = std::chrono::steady_clock::now();
    check(bcmg_setup(c, A, &p, &B), "AMG setup");
    auto t1 = std::chrono::steady_clock::now();
    print_setup(B);
    vector<vtype> *Sol = Vector::init<vtype>(A_host->n, true, true);
    std::cout << "SOLVING....:\n\n";
    int num_iter = 0;
    double residual = 0.0;
    const int rc = bcmg_fcg(c, B, rhs->val, Sol->val, &num_iter, &residual);
    auto t2 = std::chrono::steady_clock::now();
    if (rc != BCMG_OK) {
        bcmg_boot_free(B);
        bcmg_csr_free(A);
        check(rc, "FCG");
    }
    const float boot_time = std::chrono::duration<float, std::milli>(t1 - t0).count();
    const float solve_time = std::chrono::duration<float, std::milli>(t2 - t1).count();
    print_solve(B, num_iter, residual, boot_time, solve_time);
    bcmg_boot_free(B);
    bcmg_csr_free(A);
    return Sol;
}

#endif /* BCMG_BOOTCMATCHG_H */
