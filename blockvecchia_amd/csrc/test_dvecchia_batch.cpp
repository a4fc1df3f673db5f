/*
 * test_dvecchia_batch -- the Block Vecchia maximum-likelihood driver on libbv.
 *
 * Same command line, defaults, console lines and summary CSV as the
 * reference driver (src/test_Xvecchia_batch.cpp with the option parser of
 * include/vecchia_helper_data.h:256-739 and the writers of
 * include/utils.h:27-191), with the likelihood evaluated by libbv on the GPU
 * (bv_create / llh_Xvecchia_batch) and the outer loop by bv_optim.h in place
 * of NLopt's BOBYQA.
 *
 *   --perf      one evaluation at the true theta with y = 0 on generated
 *               locations (GenerateXYLoc: jittered grid, Morton order)
 *   otherwise   locations / observations from CSV (--xy_path, --obs_path),
 *               maximisation from --kernel_init within [lower, upper]
 *
 * Only the default block ordering (random) is built; --permutation with
 * another method exits with an error.
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "bv.h"
#include "bv_optim.h"

struct Opts {
    int kernel = 1, num_params = 3, num_loc = 2000, time_flag = 0, t_slots = 0;
    double sigma = 1.0, beta = 0.1, nu = 0.5, nugget = 0.0, dist_scale = 1.0;
    double sigma_init = 0.01, beta_init = 0.01, nu_init = 0.01, nugget_init = 0.01;
    double nu_time = 0.5, beta_time = 2.2, sep = 0.5, aux = 0.0;
    double nu_time_init = 0.01, beta_time_init = 0.01, sep_init = 0.01;
    int knn = 1, randomordering = 1, mortonordering = 0, kdtreeordering = 0;
    int hilbertordering = 0, mmdordering = 0;
    int seed = 0, perf = 0, vecchia = 1, vecchia_cs = 100, vecchia_bc = 200;
    double tol = 1e-5;
    int maxiter = 1000;
    double lower_bound = 0.01, upper_bound = 3.0;
    int omp_numthreads = 40, device = 0, earth = 0;
    std::string xy_path, obs_path;
};

static void display_help() {
    std::cout << "Usage: test_dvecchia_batch [options]\n"
              << "Options:\n"
              << "  --help                this message\n"
              << "  --ikernel             true kernel parameters sigma^2:range:smooth[:nugget] (? = unknown)\n"
              << "  --kernel              kernel name: univariate_matern_stationary_no_nugget,\n"
              << "                        univariate_powexp_stationary_no_nugget,\n"
              << "                        univariate_powexp_nugget_stationary,\n"
              << "                        univariate_matern_spacetime_stationary\n"
              << "  --kernel_init         starting parameters of the maximisation, same form\n"
              << "  --vecchia_bc          [int] number of blocks (k-means clusters)\n"
              << "  --vecchia_cs          [int] conditioning set size\n"
              << "  --num_loc             [int] number of locations\n"
              << "  --dist_scale          [double] distance scale\n"
              << "  --knn                 nearest-neighbour conditioning sets (default)\n"
              << "  --perf                one evaluation at the true parameters with obs = 0\n"
              << "  --seed                [int] seed of the generated locations and k-means\n"
              << "  --3D                  3-D locations (x, y, z)\n"
              << "  --earth               great-circle distance (x = latitude, y = longitude)\n"
              << "  --xy_path             [string] locations CSV\n"
              << "  --obs_path            [string] observations CSV\n"
              << "  --tol                 [int] x tolerance of the maximisation, 5 -> 1e-5\n"
              << "  --maxiter             [int] maximum number of likelihood evaluations\n"
              << "  --lower_bound         [double] lower bound of every parameter\n"
              << "  --upper_bound         [double] upper bound of every parameter (< 100)\n"
              << "  --t_slots             [int] time slots (space-time), num_loc % t_slots == 0\n"
              << "  --omp_threads         [int] host threads\n"
              << "  --permutation         [string] block ordering (random)\n"
              << std::endl;
}

/* a:b:c[:d[:e:f]] with "?" for unknown; returns the number of fields */
static int parse_colon(const char *arg, double v[6], bool known[6]) {
    char s[6][10];
    const int info = sscanf(arg, "%9[^:]:%9[^:]:%9[^:]:%9[^:]:%9[^:]:%9[^:]", s[0], s[1], s[2],
                            s[3], s[4], s[5]);
    for (int i = 0; i < 6; ++i) {
        known[i] = false;
        if (i < info && strcmp(s[i], "?") != 0) {
            v[i] = atof(s[i]);
            known[i] = true;
        }
    }
    return info;
}

/* the option parser of include/vecchia_helper_data.h:256-739: the same flags,
   defaults, checks and exit codes */
static void parse_opts(int argc, char **argv, Opts &o) {
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        const bool has = i + 1 < argc;
        if (!strcmp(a, "--xy_path") && has) {
            o.xy_path = argv[++i];
            std::cout << "xy_path: " << o.xy_path << std::endl;
        } else if (!strcmp(a, "--obs_path") && has) {
            o.obs_path = argv[++i];
            std::cout << "obs_path: " << o.obs_path << std::endl;
        } else if (!strcmp(a, "--num_loc") && has) {
            int v = 0;
            const int info = sscanf(argv[++i], "%d", &v);
            if (info == 1 && v > 0) o.num_loc = v;
            else {
                fprintf(stderr, "error: --num_loc %s is invalid; ensure num_loc > 0.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--dist_scale") && has) {
            double v = 0;
            if (sscanf(argv[++i], "%lf", &v) == 1 && v > 0) o.dist_scale = v;
            else {
                fprintf(stderr, "error: --dist_scale %s is invalid; ensure dist_scale > 0.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--vecchia_cs") && has) {
            int v = 0;
            if (sscanf(argv[++i], "%d", &v) == 1 && v >= 0) {
                o.vecchia_cs = v;
                o.vecchia = 1;
            } else {
                fprintf(stderr, "error: --vecchia_cs %s is invalid; ensure 0 <= vecchia_cs <= N.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--vecchia_bc") && has) {
            int v = 0;
            if (sscanf(argv[++i], "%d", &v) == 1 && v > 0) o.vecchia_bc = v;
            else {
                fprintf(stderr, "error: --vecchia_bc %s is invalid; ensure 0 < vecchia_bc <= N.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--perf") && has) {
            o.perf = 1;
            o.maxiter = 1;
        } else if (!strcmp(a, "--knn") && has) {
            o.knn = 1;
        } else if (!strcmp(a, "--3D") && has) {
            o.time_flag = 1;
            o.t_slots = 0;
        } else if (!strcmp(a, "--earth") && has) {
            o.earth = 1;
        } else if (!strcmp(a, "--permutation") && has) {
            const std::string r = argv[++i];
            if (r == "random") o.randomordering = 1;
            else {
                o.randomordering = 0;
                if (r == "morton") o.mortonordering = 1;
                if (r == "kdtree") o.kdtreeordering = 1;
                if (r == "hilbert") o.hilbertordering = 1;
                if (r == "mmd") o.mmdordering = 1;
            }
            const int sum = o.randomordering + o.mortonordering + o.kdtreeordering +
                            o.hilbertordering + o.mmdordering;
            if (sum > 1) {
                std::cout << "Please only contain one permutation methods, you are containing "
                          << sum << std::endl;
                exit(0);
            }
        } else if (!strcmp(a, "--kernel") && has) {
            const char *k = argv[++i];
            if (!strcmp(k, "univariate_matern_stationary_no_nugget")) {
                fprintf(stderr, "You are using the Matern Kernel (sigma^2, range, smooth)!\n");
                o.kernel = 1; o.num_params = 3;
            } else if (!strcmp(k, "univariate_powexp_stationary_no_nugget")) {
                fprintf(stderr, "You are using the Power exponential Kernel (sigma^2, range, smooth)!\n");
                o.kernel = 2; o.num_params = 3;
            } else if (!strcmp(k, "univariate_powexp_nugget_stationary")) {
                fprintf(stderr, "You are using the Power exponential Kernel with nugget (sigma^2, range, smooth, nugget)!\n");
                o.kernel = 3; o.num_params = 4;
            } else if (!strcmp(k, "univariate_matern_spacetime_stationary")) {
                fprintf(stderr, "You are using the space-time Kernel (sigma^2, range_spatial, smooth_spatial, range_time, smooth_time, seperability, aux)!\n");
                o.kernel = 4; o.num_params = 7; o.time_flag = 1;
            } else {
                fprintf(stderr, "Unsupported kernel type: %s\n", k);
                exit(1);
            }
        } else if (!strcmp(a, "--t_slots") && has) {
            int v = 0;
            if (sscanf(argv[++i], "%d", &v) == 1 && v >= 0) o.t_slots = v;
        } else if (!strcmp(a, "--ikernel") && has) {
            double v[6];
            bool k[6];
            const int info = parse_colon(argv[++i], v, k);
            if (info < 3 || info > 6) {
                printf("Other kernels have been developing on the way!");
                exit(0);
            }
            if (k[0]) o.sigma = v[0];
            if (k[1]) o.beta = v[1];
            if (k[2]) o.nu = v[2];
            if (k[3]) o.nugget = v[3];
            if (info == 6) {
                if (k[3]) o.beta_time = v[3];
                if (k[4]) o.nu_time = v[4];
                if (k[5]) o.sep = v[5];
            }
        } else if (!strcmp(a, "--kernel_init") && has) {
            double v[6];
            bool k[6];
            const int info = parse_colon(argv[++i], v, k);
            if (k[0]) o.sigma_init = v[0];
            if (k[1]) o.beta_init = v[1];
            if (k[2]) o.nu_init = v[2];
            if (k[3]) o.nugget_init = v[3];
            if (info == 6) {
                if (k[3]) o.beta_time_init = v[3];
                if (k[4]) o.nu_time_init = v[4];
                if (k[5]) o.sep_init = v[5];
            }
        } else if (!strcmp(a, "--seed") && has) {
            int v = 0;
            sscanf(argv[++i], "%d", &v);
            o.seed = v;
        } else if (!strcmp(a, "--omp_threads") && has) {
            o.omp_numthreads = atoi(argv[++i]);
        } else if (!strcmp(a, "--maxiter") && has) {
            int v = 0;
            if (sscanf(argv[++i], "%d", &v) == 1 && v > 0) o.maxiter = v;
            else {
                fprintf(stderr, "error: --maxiter %s is invalid; ensure maxiter > 0 and be integer.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--tol") && has) {
            int v = 0;
            if (sscanf(argv[++i], "%d", &v) == 1 && v > 0) o.tol = pow(10, -v);
            else {
                fprintf(stderr, "error: --tol %s is invalid; ensure tol > 0.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--lower_bound") && has) {
            double v = 0;
            if (sscanf(argv[++i], "%lf", &v) == 1 && v > 0) o.lower_bound = v;
            else {
                fprintf(stderr, "error: --lower_bound %s is invalid; ensure lower_bound > 0.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "--upper_bound") && has) {
            double v = 0;
            if (sscanf(argv[++i], "%lf", &v) == 1 && v < 100) o.upper_bound = v;
            else {
                fprintf(stderr, "error: --upper_bound %s is invalid; ensure upper_bound < 100.\n", argv[i]);
                exit(1);
            }
        } else if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
            display_help();
            exit(0);
        } else {
            fprintf(stderr, "error: unrecognized option %s\n", a);
            exit(1);
        }
    }
}

/* glibc rand() scaled to [lo, hi), as uniform_distribution (src/ckernel.cpp:36-41) */
static double uniform_draw(double lo, double hi) {
    const double r = (double)rand() / (double)(1.0 + RAND_MAX);
    return r * (hi - lo) + lo;
}

static uint32_t spread_bits(uint32_t x) {
    x &= 0x0000ffff;
    x = (x ^ (x << 8)) & 0x00ff00ff;
    x = (x ^ (x << 4)) & 0x0f0f0f0f;
    x = (x ^ (x << 2)) & 0x33333333;
    x = (x ^ (x << 1)) & 0x55555555;
    return x;
}

static uint32_t gather_bits(uint32_t x) {
    x &= 0x55555555;
    x = (x ^ (x >> 1)) & 0x33333333;
    x = (x ^ (x >> 2)) & 0x0f0f0f0f;
    x = (x ^ (x >> 4)) & 0x00ff00ff;
    x = (x ^ (x >> 8)) & 0x0000ffff;
    return x;
}

static int cmp_u32(const void *a, const void *b) {
    const uint32_t u = *(const uint32_t *)a, v = *(const uint32_t *)b;
    return u < v ? -1 : (u > v ? 1 : 0);
}

/* Jittered grid on [0,1]^2 (cell centres +- 0.4 cell, glibc rand after
   srand(seed)), then quantised to 16 bits and sorted along the Morton curve:
   the locations of GenerateXYLoc (src/ckernel.cpp:43-76,
   include/ckernel.h:43-123). */
static void generate_xy(int n, int seed, std::vector<double> &x, std::vector<double> &y) {
    srand(seed);
    x.assign(n, 0.0);
    y.assign(n, 0.0);
    const int s = (int)ceil(sqrt((double)n));
    int idx = 0;
    for (int i = 0; i < s && idx < n; ++i)
        for (int j = 0; j < s && idx < n; ++j) {
            x[idx] = ((i + 1) - 0.5 + uniform_draw(-0.4, 0.4)) / s;
            y[idx] = ((j + 1) - 0.5 + uniform_draw(-0.4, 0.4)) / s;
            ++idx;
        }
    std::vector<uint32_t> z(n);
    for (int i = 0; i < n; ++i) {
        const uint16_t qx = (uint16_t)(x[i] * (double)UINT16_MAX + .5);
        const uint16_t qy = (uint16_t)(y[i] * (double)UINT16_MAX + .5);
        z[i] = (spread_bits(qy) << 1) + spread_bits(qx);
    }
    qsort(z.data(), n, sizeof(uint32_t), cmp_u32);
    for (int i = 0; i < n; ++i) {
        x[i] = (double)(uint16_t)gather_bits(z[i]) / (double)UINT16_MAX;
        y[i] = (double)(uint16_t)gather_bits(z[i] >> 1) / (double)UINT16_MAX;
    }
}

/* GenerateXYLoc_ST (src/ckernel.cpp:78-117): n spatial sites repeated over
   t_slots time slots, z = (slot + 1) / t_slots, no Morton sort. */
static void generate_xy_st(int n, int t_slots, int seed, std::vector<double> &x,
                           std::vector<double> &y, std::vector<double> &z) {
    srand(seed);
    const size_t N = (size_t)n * t_slots;
    x.assign(N, 0.0); y.assign(N, 0.0); z.assign(N, 0.0);
    const int s = (int)ceil(sqrt((double)n));
    int idx = 0;
    for (int i = 0; i < s && idx < n; ++i)
        for (int j = 0; j < s && idx < n; ++j) {
            x[idx] = ((i + 1) - 0.5 + uniform_draw(-0.4, 0.4)) / s;
            y[idx] = ((j + 1) - 0.5 + uniform_draw(-0.4, 0.4)) / s;
            z[idx] = 1.0 / t_slots;
            ++idx;
        }
    for (int j = 1; j < t_slots; ++j)
        for (int i = 0; i < n; ++i) {
            x[i + (size_t)j * n] = x[i];
            y[i + (size_t)j * n] = y[i];
            z[i + (size_t)j * n] = (double)(j + 1) / t_slots;
        }
}

/* loadXYcsv / loadObscsv (include/utils.h:27-87): "x,y[,z]" per line, no
   header; one value per line.  Reads at most n rows. */
static bool load_xy(const std::string &path, int n, bool tflag, std::vector<double> &x,
                    std::vector<double> &y, std::vector<double> &z) {
    std::ifstream f(path);
    if (!f) {
        std::cerr << "Error: File " << path << " does not exist\n";
        exit(0);
    }
    x.assign(n, 0.0); y.assign(n, 0.0); z.assign(tflag ? n : 0, 0.0);
    std::string line;
    int i = 0;
    while (i < n && std::getline(f, line)) {
        std::stringstream ss(line);
        std::string tok;
        std::getline(ss, tok, ',');
        x[i] = std::stod(tok);
        std::getline(ss, tok, ',');
        y[i] = std::stod(tok);
        if (tflag) {
            std::getline(ss, tok, ',');
            z[i] = std::stod(tok);
        }
        ++i;
    }
    return i == n;
}

static bool load_obs(const std::string &path, int n, std::vector<double> &obs) {
    std::ifstream f(path);
    if (!f) {
        std::cerr << "Error: File " << path << " does not exist\n";
        exit(0);
    }
    obs.assign(n, 0.0);
    std::string line;
    int i = 0;
    while (i < n && std::getline(f, line)) {
        std::istringstream ss(line);
        double v;
        if (ss >> v) obs[i++] = v;
    }
    return i == n;
}

/* saveLogFileSum (include/utils.h:133-191): console summary and
   ./log/locs_<n>_cs_<cs>_bc_<bc>_seed_<seed>_kernel_<s>:<b>:<nu>_random */
static void save_log_sum(int iterations, const double *theta, double max_llh, double whole_time,
                         const Opts &o) {
    if (mkdir("./log", 0777) == 0 || errno == EEXIST)
        printf("Directory exists or was created successfully.\n");
    else
        printf("Failed to create the directory.\n");
    std::string path = "./log/locs_" + std::to_string(o.num_loc) + "_cs_" +
                       std::to_string(o.vecchia_cs) + "_bc_" + std::to_string(o.vecchia_bc) +
                       "_seed_" + std::to_string(o.seed) + "_kernel_" + std::to_string(o.sigma) +
                       ":" + std::to_string(o.beta) + ":" + std::to_string(o.nu);
    if (o.mortonordering) path += "_morton";
    else if (o.randomordering) path += "_random";
    else if (o.kdtreeordering) path += "_kdtree";
    else if (o.hilbertordering) path += "_hilbert";
    else if (o.mmdordering) path += "_mmd";
    printf("Total Number of Iterations = %d \n", iterations);
    printf("Total Optimization Time = %lf secs \n", whole_time);
    if (o.kernel == 1 || o.kernel == 2) {
        printf("Model Parameters (Variance, range, smoothness): (%.8f, %.8f, %.8f) -> Loglik: %.18f \n",
               theta[0], theta[1], theta[2], max_llh);
        std::ofstream out(path);
        out << "Iterations, variance, range, smoothness, log-likelihood, time " << std::endl;
        out << iterations << ", " << theta[0] << ", " << theta[1] << ", " << theta[2] << ", "
            << std::setprecision(std::numeric_limits<double>::max_digits10) << max_llh << ", "
            << whole_time << std::endl;
    } else if (o.kernel == 3) {
        printf("Model Parameters (Variance, range, smoothness, nugget): (%.8f, %.8f, %.8f, %.8f) -> Loglik: %.18f \n",
               theta[0], theta[1], theta[2], theta[3], max_llh);
        std::ofstream out(path);
        out << "Iterations, variance, range, smoothness, nugget, log-likelihood, time " << std::endl;
        out << iterations << ", " << theta[0] << ", " << theta[1] << ", " << theta[2] << ", "
            << theta[3] << ", " << std::setprecision(std::numeric_limits<double>::max_digits10)
            << max_llh << ", " << whole_time << std::endl;
    }
}

int main(int argc, char **argv) {
    Opts o;
    parse_opts(argc, argv, o);
    if (!o.randomordering) {
        fprintf(stderr, "error: only the random block ordering is built in this driver\n");
        return 1;
    }
    if (o.time_flag && o.t_slots != 0 && o.num_loc % o.t_slots != 0) {
        fprintf(stderr, "Your number of locations cannot be divided by t_slot!");
        return -1;
    }
    const int np = o.num_params;
    std::vector<double> x, y, z, obs;
    std::vector<double> theta0(np, 0.0);
    int metric = 0;
    if (o.perf) {
        if (o.time_flag) {
            if (o.t_slots <= 0) {
                fprintf(stderr, "error: --perf with 3-D locations needs --t_slots > 0\n");
                return 1;
            }
            generate_xy_st(o.num_loc / o.t_slots, o.t_slots, o.seed, x, y, z);
        } else {
            generate_xy(o.num_loc, o.seed, x, y);
        }
        obs.assign(o.num_loc, 0.0);
        const double t[7] = {o.sigma, o.beta, o.nu, o.kernel == 3 ? o.nugget : o.beta_time,
                             o.nu_time, o.sep, o.aux};
        for (int i = 0; i < np; ++i) theta0[i] = t[i];
    } else {
        std::string xy = o.xy_path, zp = o.obs_path;
        if (xy.empty()) {
            fprintf(stderr, "You are using the default the path, please see in file ./src/test_Xvecchia_batch.cpp!\n");
            xy = "./simu_ds/20ks_" + std::to_string(o.beta) + "_" + std::to_string(o.nu) +
                 "/LOC_20000_univariate_matern_stationary_" + std::to_string(o.seed);
            zp = "./simu_ds/20ks_" + std::to_string(o.beta) + "_" + std::to_string(o.nu) +
                 "/Z1_20000_univariate_matern_stationary_" + std::to_string(o.seed);
        }
        if (o.earth) metric = 1;
        if (!load_xy(xy, o.num_loc, o.time_flag, x, y, z) || !load_obs(zp, o.num_loc, obs)) {
            fprintf(stderr, "error: fewer than --num_loc %d rows in the input files\n", o.num_loc);
            return 1;
        }
        const double t[7] = {o.sigma_init, o.beta_init, o.nu_init,
                             o.kernel == 3 ? o.nugget_init : o.beta_time_init, o.nu_time_init,
                             o.sep_init, 0.0};
        for (int i = 0; i < np; ++i) theta0[i] = t[i];
    }
    const int n = (int)x.size();
    o.time_flag ? fprintf(stderr, "You were using the 3D ordering. \n")
                : fprintf(stderr, "You were using the 2D ordering. \n");

    /* partition and conditioning sets (host) */
    bv_blocks blk;
    memset(&blk, 0, sizeof(blk));
    int rc = bv_partition(n, x.data(), y.data(), o.time_flag ? z.data() : nullptr, obs.data(),
                          o.vecchia_bc, o.vecchia_cs, o.seed, 50, o.time_flag, metric,
                          o.omp_numthreads, &blk);
    if (rc != BV_OK) {
        fprintf(stderr, "error: partition failed (%d)\n", rc);
        return 1;
    }
    bv_problem p;
    memset(&p, 0, sizeof(p));
    p.n = blk.n;
    p.batch_count = blk.batch_count;
    p.cs = blk.cs;
    p.batch_num = blk.batch_num;
    p.x = blk.x; p.y = blk.y; p.z = blk.z; p.obs = blk.obs;
    p.nbr = blk.cs > 0 ? blk.nbr : nullptr;
    p.kernel = o.kernel;
    p.distance_metric = metric;
    p.time_flag = o.time_flag;
    p.dist_scale = o.dist_scale;
    p.device = o.device;
    p.verbose = 1;
    bv_ctx *ctx = nullptr;
    rc = bv_create(&p, &ctx);
    if (rc != BV_OK) {
        fprintf(stderr, "error: bv_create failed (%d)\n", rc);
        bv_blocks_free(&blk);
        return 1;
    }

    /* bounds of src/test_Xvecchia_batch.cpp:706-720 */
    std::vector<double> lb(np, o.lower_bound), ub(np, o.upper_bound);
    if (o.kernel == 4) {
        ub[4] = 1.0;
        ub[5] = 1.0;
        lb[6] = 0.0;
        ub[6] = 0.0;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const bvopt::Result r = bvopt::maximize(
        [&](const std::vector<double> &th) {
            return llh_Xvecchia_batch((unsigned)np, th.data(), nullptr, ctx);
        },
        theta0, lb, ub, o.tol, o.maxiter);
    if (o.kernel == 1 || o.kernel == 2)
        printf("Found maximum at f(%g, %g, %g) = %0.10g\n", r.x[0], r.x[1], r.x[2], r.f);
    else if (o.kernel == 3)
        printf("Found maximum at f(%g, %g, %g, %g) = %0.10g\n", r.x[0], r.x[1], r.x[2], r.x[3], r.f);
    else
        printf("Found maximum at f(%g, %g, %g, %g, %g, %g) = %0.10g\n", r.x[0], r.x[1], r.x[2],
               r.x[3], r.x[4], r.x[5], r.f);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    const double whole = (double)(t1.tv_sec - t0.tv_sec) + (t1.tv_nsec - t0.tv_nsec) / 1e9;
    save_log_sum(bv_iterations(ctx), r.x.data(), r.f, whole, o);
    fprintf(stderr, "Done! \n");
    bv_destroy(ctx);
    bv_blocks_free(&blk);
    return 0;
}
