/*
 * bcmg_cli.cpp - command-line solver over the library (bootcmatchg_amd/bcmg_solve).
 *
 *   bcmg_solve [options] MATRIX SETTINGS
 *
 * With no options it does what the reference's driver does
 * (src/solve_example.cu:8-38): MatrixMarket in, b = 1, one setup + FCG solve
 * with the 14-value settings file, the solution written to x.txt as "%.14e"
 * lines, the #META lines and timings on stdout, the same exit codes. The
 * options add what the library offers beyond that:
 *
 *   -f, --format mtx|mtx2      matrix file format (readMatrixFromFile m_type 0|1)
 *   -b, --rhs ones|rand:SEED|FILE
 *                              right-hand side: all ones (default); uniform[-1,1]
 *                              from SplitMix64 stream SEED (the generator of the
 *                              Python package, bootcmatchg_amd/operators.py); or
 *                              one value per line from FILE
 *   -o, --out FILE             solution file (default x.txt; "-" writes none)
 *   -p, --precision fp64|fp32|bf16
 *                              preconditioner precision (FCG itself stays FP64)
 *   -r, --repeat K             K solves on the one built composite (x0 = 0 each
 *                              time); the last solution is written
 *   -t, --phases               print the event-timed setup phases
 *   -h, --help
 *
 * Exit codes: 0 done; 1 bad command line, zero right-hand side or missing
 * matrix-type argument (as the reference); 255 bad settings or matrix file
 * (the reference's exit(-1)); 2 a solve that did not converge within itnlim
 * or broke down.
 */
#include "BootCMatchG.h"

#include <cinttypes>
#include <fstream>

namespace {

void usage(const char *prog)
{
    std::printf("usage: %s [-f mtx|mtx2] [-b ones|rand:SEED|FILE] [-o FILE] "
                "[-p fp64|fp32|bf16] [-r K] [-t] MATRIX SETTINGS\n",
                prog);
}

/* SplitMix64 draw k (k = 0, 1, ...) of stream seed as uniform[-1, 1] */
double uniform_pm1(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
}

struct Options {
    int m_type = 0;
    std::string rhs = "ones";
    std::string out = "x.txt";
    int precision = BCMG_PREC_FP64;
    int repeat = 1;
    bool phases = false;
    const char *matrix = nullptr, *settings = nullptr;
};

bool parse(int argc, char **argv, Options &o)
{
    std::vector<const char *> pos;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto value = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::printf("%s needs a value\n", name);
                return nullptr;
            }
            return argv[++i];
        };
        const char *v = nullptr;
        if (a == "-h" || a == "--help") {
            usage(argv[0]);
            std::exit(0);
        } else if (a == "-f" || a == "--format") {
            if (!(v = value("--format"))) return false;
            if (std::string(v) == "mtx")
                o.m_type = 0;
            else if (std::string(v) == "mtx2")
                o.m_type = 1;
            else
                return false;
        } else if (a == "-b" || a == "--rhs") {
            if (!(v = value("--rhs"))) return false;
            o.rhs = v;
        } else if (a == "-o" || a == "--out") {
            if (!(v = value("--out"))) return false;
            o.out = v;
        } else if (a == "-p" || a == "--precision") {
            if (!(v = value("--precision"))) return false;
            const std::string s = v;
            if (s == "fp64")
                o.precision = BCMG_PREC_FP64;
            else if (s == "fp32")
                o.precision = BCMG_PREC_FP32;
            else if (s == "bf16")
                o.precision = BCMG_PREC_BF16;
            else
                return false;
        } else if (a == "-r" || a == "--repeat") {
            if (!(v = value("--repeat"))) return false;
            o.repeat = std::atoi(v);
            if (o.repeat < 1) return false;
        } else if (a == "-t" || a == "--phases") {
            o.phases = true;
        } else if (!a.empty() && a[0] == '-' && a != "-") {
            std::printf("unknown option %s\n", a.c_str());
            return false;
        } else {
            pos.push_back(argv[i]);
        }
    }
    if (pos.size() != 2) return false;
    o.matrix = pos[0];
    o.settings = pos[1];
    return true;
}

std::vector<double> make_rhs(const std::string &spec, int n)
{
    std::vector<double> b((size_t)n, 1.0);
    if (spec == "ones") return b;
    if (spec.rfind("rand:", 0) == 0) {
        const uint64_t seed = std::strtoull(spec.c_str() + 5, nullptr, 10);
        for (int i = 0; i < n; i++) b[i] = uniform_pm1(seed, (uint64_t)i);
        return b;
    }
    std::ifstream f(spec);
    if (!f) {
        std::cout << "file not found: " << spec << "\n";
        std::exit(-1);
    }
    for (int i = 0; i < n; i++)
        if (!(f >> b[i])) {
            std::cout << "right-hand side file " << spec << ": fewer than " << n << " values\n";
            std::exit(-1);
        }
    return b;
}

} // namespace

int main(int argc, char **argv)
{
    using namespace bcmg_facade;
    Options o;
    if (!parse(argc, argv, o)) {
        usage(argv[0]);
        return 1;
    }
    std::cout << "READING INPUT MATRIX....:\n\n";
    CSR *A_host = readMatrixFromFile(o.matrix, o.m_type, false);
    const params p = AMG::Params::initFromFile(o.settings);
    const int n = A_host->n;
    const std::vector<double> b = make_rhs(o.rhs, n);

    bcmg_ctx c = ctx();
    bcmg_csr A = nullptr;
    bcmg_boot B = nullptr;
    check(bcmg_csr_upload(c, n, A_host->nnz, A_host->row, A_host->col, A_host->val, &A),
          "copy matrix to device");
    double *db = nullptr, *dx = nullptr;
    check(bcmg_dmalloc(c, sizeof(double) * n, (void **)&db), "device allocation");
    check(bcmg_dmalloc(c, sizeof(double) * n, (void **)&dx), "device allocation");
    check(bcmg_memcpy_h2d(c, db, b.data(), sizeof(double) * n), "copy rhs to device");

    std::cout << "BUILDING....:\n\n";
    check(bcmg_time_setup(c, 1), "setup timing");
    auto t0 = std::chrono::steady_clock::now();
    check(bcmg_setup(c, A, &p, &B), "AMG setup");
    if (o.precision != BCMG_PREC_FP64) check(bcmg_set_precision(B, o.precision), "precision");
    auto t1 = std::chrono::steady_clock::now();
    print_setup(B);
    if (o.phases) {
        bcmg_setup_times t;
        check(bcmg_setup_timing(B, -1, &t), "setup timing");
        std::printf("setup phases (ms): weights %.3f matching %.3f aggregation %.3f "
                    "galerkin %.3f smoother %.3f inner %.3f total %.3f\n",
                    t.weights_ms, t.matching_ms, t.aggregation_ms, t.galerkin_ms, t.smoother_ms,
                    t.inner_ms, t.total_ms);
    }

    std::cout << "SOLVING....:\n\n";
    int status = 0, num_iter = 0;
    double residual = 0.0;
    float solve_time = 0.f;
    for (int r = 0; r < o.repeat; r++) {
        auto s0 = std::chrono::steady_clock::now();
        const int rc = bcmg_fcg(c, B, db, dx, &num_iter, &residual);
        auto s1 = std::chrono::steady_clock::now();
        if (rc == BCMG_EZERORHS || rc == BCMG_EBREAKDOWN) {
            std::cout << "FCG: " << bcmg_last_error() << "\n";
            return rc == BCMG_EZERORHS ? 1 : 2;
        }
        check(rc, "FCG");
        solve_time = std::chrono::duration<float, std::milli>(s1 - s0).count();
        if (o.repeat > 1) std::printf("solve %d: %.3f ms, %d iterations\n", r, solve_time, num_iter);
    }
    const float boot_time = std::chrono::duration<float, std::milli>(t1 - t0).count();
    print_solve(B, num_iter, residual, boot_time, solve_time);
    if (num_iter >= p.itnlim) {
        std::cout << "not converged within itnlim = " << p.itnlim << "\n";
        status = 2;
    }

    if (o.out != "-") {
        vector<vtype> x{n, true, dx};
        Vector::printOnFile(&x, o.out.c_str());
    }
    bcmg_dfree(c, db);
    bcmg_dfree(c, dx);
    bcmg_boot_free(B);
    bcmg_csr_free(A);
    CSRm::free(A_host);
    return status;
}
