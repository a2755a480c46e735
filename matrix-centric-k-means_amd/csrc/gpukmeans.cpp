// gpukmeans.cpp -- command line with the reference's flags, input formats and
// report lines (src/main.cu:15-131, src/include/utils.hpp:92-172,
// src/include/input_parser.hpp:15-206), running on the gfx950 engine.
//
//   gpukmeans -d D -n N -k K -m MAXITER -b random -f linear -l 2
//             [-i in.csv|in.svm] [-o out.csv] [-r RUNS] [-s SEED] [-t TOL] [-c 0|1]
// Extensions (not in the reference):
//   -p/--precision fp32|bf16x3|bf16|fp8   distance GEMM form (default fp32)
//   --label-col first|last|none       CSV class-label column (default first, as
//                                     the reference reader; its shipped data
//                                     files put the label last)
//   --cpu-semantics                   reference CPU k-means semantics (seeded
//                                     distinct rows, shift test) instead of the
//                                     reference GPU engine's
//   --no-stage-times                  omit the [PERFORMANCE] stage-time lines
//                                     (printed by default, as the reference's
//                                     harness expects them)
//   --batch-runs                      run the -r repetitions together as one
//                                     batch (kmc_batch_*: one launch per stage
//                                     for all of them, each chunk of iterations
//                                     replayed as a HIP graph); needs
//                                     --cpu-semantics, -b random and the linear
//                                     kernel
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "kmeans.hpp"

namespace {

const int EXIT_ARGS = 1, EXIT_DEV = 2, EXIT_INFILE = 4;
const char* BOLDBLUE = "\033[1m\033[34m";
const char* RESET = "\033[0m";

struct Flag {
  const char* shortname;
  const char* longname;
  bool takes_value;
};
// Long names as src/include/utils.hpp:147-159.
const Flag kFlags[] = {
    {"h", "help", false},        {"d", "n-dimensions", true}, {"n", "n-samples", true},
    {"k", "n-clusters", true},   {"m", "maxiter", true},      {"o", "out-file", true},
    {"i", "in-file", true},      {"r", "runs", true},         {"s", "seed", true},
    {"t", "tolerance", true},    {"c", "check", true},        {"b", "init", true},
    {"f", "kernel", true},       {"l", "level", true},        {"p", "precision", true},
    {nullptr, "label-col", true}, {nullptr, "cpu-semantics", false},
    {nullptr, "no-stage-times", false}, {nullptr, "batch-runs", false}};

void usage() {
  std::cout
      << "gpukmeans is an implementation of the K-means algorithm that uses a GPU\n"
         "Usage:\n  gpukmeans [OPTION...]\n\n"
         "  -h, --help               Print usage\n"
         "  -d, --n-dimensions arg   Number of dimensions of a point\n"
         "  -n, --n-samples arg      Number of points\n"
         "  -k, --n-clusters arg     Number of clusters\n"
         "  -m, --maxiter arg        Maximum number of iterations\n"
         "  -o, --out-file arg       Output filename\n"
         "  -i, --in-file arg        Input filename\n"
         "  -r, --runs arg           Number of k-means runs (default: 1)\n"
         "  -s, --seed arg           Seed for centroids generator\n"
         "  -t, --tolerance arg      Tolerance to declare convergence (default: FLT_EPSILON)\n"
         "  -c, --check arg          Whether or not to check convergence (default: 0)\n"
         "  -b, --init arg           Method to use to initialize centroids\n"
         "  -f, --kernel arg         kernel function\n"
         "  -l, --level arg          Optimization level\n"
         "  -p, --precision arg      fp32 | bf16x3 | bf16 | fp8 (default: fp32)\n"
         "      --label-col arg      first | last | none (default: first)\n"
         "      --cpu-semantics      reference CPU k-means semantics\n"
         "      --no-stage-times     omit the [PERFORMANCE] stage-time lines\n"
         "      --batch-runs         run the -r repetitions as one batch\n";
}

[[noreturn]] void missing(const char* arg) {
  std::cerr << "Invalid or missing argument: " << arg << std::endl;
  std::exit(EXIT_ARGS);
}

const Flag* find_flag(const std::string& s, bool is_long) {
  for (const Flag& f : kFlags) {
    if (is_long && s == f.longname) return &f;
    if (!is_long && f.shortname && s == f.shortname) return &f;
  }
  return nullptr;
}

std::map<std::string, std::string> parse(int argc, char** argv) {
  std::map<std::string, std::string> out;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    std::string name, value;
    bool has_value = false;
    const Flag* f = nullptr;
    if (a.rfind("--", 0) == 0) {
      std::string body = a.substr(2);
      const size_t eq = body.find('=');
      if (eq != std::string::npos) {
        value = body.substr(eq + 1);
        body = body.substr(0, eq);
        has_value = true;
      }
      f = find_flag(body, true);
    } else if (a.size() >= 2 && a[0] == '-') {
      f = find_flag(a.substr(1, 1), false);
      if (a.size() > 2) {
        value = a.substr(2);
        has_value = true;
      }
    }
    if (!f) {
      std::cerr << "Invalid or missing argument: " << a << std::endl;
      std::exit(EXIT_ARGS);
    }
    name = f->longname;
    if (f->takes_value && !has_value) {
      if (i + 1 >= argc) missing(f->longname);
      value = argv[++i];
    }
    out[name] = value;
  }
  return out;
}

long get_int(const std::map<std::string, std::string>& a, const char* name, const long* def) {
  auto it = a.find(name);
  if (it == a.end()) {
    if (def) return *def;
    missing(name);
  }
  char* end = nullptr;
  const long v = std::strtol(it->second.c_str(), &end, 10);
  if (!end || *end != '\0' || it->second.empty()) missing(name);
  return v;
}

// Random points as src/include/input_parser.hpp:50-68: mt19937(seed), values
// uniform in [-1e5, 1e5), drawn row by row.
std::vector<Point<float>*> random_points(int seed, int d, size_t n) {
  std::mt19937 eng(seed);
  std::uniform_real_distribution<float> distr(-1e5, 1e5);
  std::vector<Point<float>*> pts(n);
  std::vector<float> row(d);
  for (size_t i = 0; i < n; ++i) {
    for (int j = 0; j < d; ++j) row[j] = distr(eng);
    pts[i] = new Point<float>(row.data(), d);
  }
  return pts;
}

// CSV reader (src/include/input_parser.hpp:103-137): first line is a header;
// the label column is skipped (first column in the reference).
bool read_csv(std::istream& in, int d, size_t n, const std::string& label_col,
              std::vector<Point<float>*>& pts) {
  std::string line;
  std::getline(in, line);  // header
  std::vector<float> row(d);
  size_t i = 0;
  while (i < n && std::getline(in, line)) {
    std::vector<std::string> tok;
    std::stringstream ss(line);
    std::string t;
    while (std::getline(ss, t, ',')) tok.push_back(t);
    size_t off = label_col == "first" ? 1 : 0;
    std::fill(row.begin(), row.end(), 0.0f);
    for (int j = 0; j < d && off + j < tok.size(); ++j) row[j] = (float)std::atof(tok[off + j].c_str());
    pts[i++] = new Point<float>(row.data(), d);
  }
  return i == n;
}

// libsvm reader (src/include/input_parser.hpp:76-100): "idx:value" pairs,
// 1-based indices, missing features are 0.
bool read_svm(std::istream& in, int d, size_t n, std::vector<Point<float>*>& pts) {
  std::string line;
  std::vector<float> row(d);
  size_t i = 0;
  while (i < n && std::getline(in, line)) {
    std::fill(row.begin(), row.end(), 0.0f);
    std::stringstream ss(line);
    std::string tok;
    while (std::getline(ss, tok, ' ')) {
      const size_t c = tok.find(':');
      if (c == std::string::npos) continue;
      const int idx = std::atoi(tok.substr(0, c).c_str()) - 1;
      if (idx >= 0 && idx < d) row[idx] = (float)std::atof(tok.substr(c + 1).c_str());
    }
    pts[i++] = new Point<float>(row.data(), d);
  }
  return i == n;
}

// CSV writer (src/include/input_parser.hpp:167-183).
void write_csv(std::ostream& o, const std::vector<Point<float>*>& pts, int d) {
  o << "cluster,";
  for (int j = 0; j < d; ++j) o << "d" << j << (j != d - 1 ? "," : "");
  o << std::endl;
  for (const Point<float>* p : pts) {
    o << p->getCluster() << ",";
    for (int j = 0; j < d; ++j) o << std::setprecision(8) << p->get(j) << (j != d - 1 ? "," : "");
    o << std::endl;
  }
}

// The -r repetitions as one batch (the reference's loop is src/main.cu:87-115:
// one Kmeans object and one run per repetition, the first a warm-up). Every
// repetition is one problem of a kmc_batch with the same points and the same
// seeded rows, so all of them run in one launch per stage; the batch is run
// once untimed (the warm-up), then timed, and each repetition is charged an
// equal share of the timed run. The batch follows the reference CPU k-means'
// rules (shift test, empty cluster -> zero vector) with exact fixed-point
// centroid sums rounded once. Prints the sequential loop's report lines.
int batch_runs(std::vector<Point<float>*>& pts, int d, uint32_t k, float tol, int seed,
               uint32_t runs, uint64_t maxiter, bool check_converged, const std::string& prec,
               const kmc::DeviceInfo& dev) {
  const size_t n = pts.size();
  std::vector<float> X(n * (size_t)d);
  for (size_t i = 0; i < n; ++i)
    for (int j = 0; j < d; ++j) X[i * d + j] = pts[i]->get(j);
  std::vector<float> C0;
  try {
    C0 = kmc::seeded_rows(pts.data(), X, n, (uint32_t)d, k, &seed);
  } catch (const kmc::Error& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  double sum_x2 = 0.0;  // the score is the inertia less sum ||x||^2
  for (float v : X) sum_x2 += (double)v * (double)v;
  std::vector<kmc_problem> probs(runs);
  for (uint32_t r = 0; r < runs; ++r) probs[r] = {n, (uint32_t)d, k, X.data(), (size_t)d, C0.data()};
  kmc_batch_config cfg{};
  cfg.tol = tol;
  cfg.precision = prec == "bf16x3" ? KMC_PREC_BF16X3 : prec == "bf16" ? KMC_PREC_BF16
                  : prec == "fp8"  ? KMC_PREC_FP8 : KMC_PREC_FP32;
  cfg.device = dev.ordinal;
  kmc_batch* b = nullptr;
  const auto i0 = std::chrono::high_resolution_clock::now();
  kmc_status st = kmc_batch_create(probs.data(), runs, &cfg, &b);
  const auto i1 = std::chrono::high_resolution_clock::now();
  if (st != KMC_OK) {
    std::cerr << "kmc_batch_create: " << kmc_status_string(st) << std::endl;
    return st == KMC_ERR_NO_DEVICE ? EXIT_DEV : 1;
  }
  const uint32_t flags = (check_converged ? KMC_RUN_CHECK_CONVERGED : 0u) | KMC_RUN_GRAPH;
  std::vector<kmc_batch_result> res(runs);
  st = kmc_batch_run(b, maxiter, flags, res.data(), nullptr);  // warm-up
  const auto t0 = std::chrono::high_resolution_clock::now();
  if (st == KMC_OK) st = kmc_batch_run(b, maxiter, flags, res.data(), nullptr);
  const auto t1 = std::chrono::high_resolution_clock::now();
  std::vector<int32_t> lab(n);
  if (st == KMC_OK) st = kmc_batch_get_labels(b, 0, lab.data());
  if (st != KMC_OK) {
    const char* m = kmc_batch_last_error(b);
    std::cerr << "kmc_batch_run: " << (m ? m : kmc_status_string(st)) << std::endl;
    kmc_batch_destroy(b);
    return 1;
  }
  kmc_batch_destroy(b);
  const double per_run = std::chrono::duration<double>(t1 - t0).count() / runs;
  const double init_per_run = std::chrono::duration<double>(i1 - i0).count() / runs;
  printf("%s", BOLDBLUE);
  double score = 0;
  for (uint32_t r = 0; r < runs; ++r) {
    if (res[r].converged_at < maxiter)
      printf("K-means converged at iteration %lu - ", (unsigned long)res[r].converged_at);
    else
      printf("K-means did NOT converge - ");
    printf("Time: %lf\n", per_run);
    const float s = (float)(res[r].inertia - sum_x2);
    printf("Objective score: %lf\n", s);
    score += s;
  }
  printf("GPU_Kmeans: %lfs (%u runs)\n", per_run, runs);
  printf("Init_time: %lfs \n", init_per_run);
  printf("Score: %lf\n", score / runs);
  printf("%s", RESET);
  for (size_t i = 0; i < n; ++i) pts[i]->setCluster((unsigned int)lab[i]);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const auto args = parse(argc, argv);
  if (args.count("help")) {
    usage();
    return 0;
  }
  const long one = 1, zero = 0;
  const int d = (int)get_int(args, "n-dimensions", nullptr);
  const size_t n = (size_t)get_int(args, "n-samples", nullptr);
  const uint32_t k = (uint32_t)get_int(args, "n-clusters", nullptr);
  const uint64_t maxiter = (uint64_t)get_int(args, "maxiter", nullptr);
  const std::string out_file = args.count("out-file") ? args.at("out-file") : "None";
  float tol = FLT_EPSILON;
  if (args.count("tolerance")) tol = std::strtof(args.at("tolerance").c_str(), nullptr);
  const uint32_t runs = (uint32_t)get_int(args, "runs", &one);
  const bool check_converged = get_int(args, "check", &zero) != 0;
  if (!args.count("init")) missing("init");
  if (!args.count("kernel")) missing("kernel");
  const std::string init = args.at("init"), kernel = args.at("kernel");
  const int level = (int)get_int(args, "level", nullptr);
  int seed = args.count("seed") ? (int)get_int(args, "seed", nullptr) : std::rand();
  const std::string prec = args.count("precision") ? args.at("precision") : "fp32";
  const std::string label_col = args.count("label-col") ? args.at("label-col") : "first";
  if (d <= 0 || n == 0 || k == 0) missing("n-dimensions/n-samples/n-clusters");

  std::vector<Point<float>*> pts(n, nullptr);
  if (args.count("in-file")) {
    const std::string f = args.at("in-file");
    std::ifstream in(f);
    bool ok = false;
    if (in) {
      if (f.find(".csv") != std::string::npos) {
        std::cout << "READING CSV" << std::endl;
        ok = read_csv(in, d, n, label_col, pts);
      } else if (f.find(".svm") != std::string::npos) {
        ok = read_svm(in, d, n, pts);
      }
    }
    if (!ok) {
      std::cerr << "Invalid input file" << std::endl;
      return EXIT_INFILE;
    }
  } else {
    pts = random_points(seed, d, n);
  }

  kmc::DeviceInfo dev;
  if (!kmc::query_device(0, &dev)) {
    std::cerr << "There are no available device(s) that support HIP" << std::endl;
    return EXIT_DEV;
  }

  Kmeans::InitMethod im;
  if (init == "random") im = Kmeans::InitMethod::random;
  else if (init == "plusplus") im = Kmeans::InitMethod::plus_plus;
  else {
    printf("Invalid init method: %s\n", init.c_str());
    return 1;
  }
  Kmeans::Kernel kf;
  if (kernel == "linear") kf = Kmeans::Kernel::linear;
  else if (kernel == "polynomial") kf = Kmeans::Kernel::polynomial;
  else if (kernel == "sigmoid") kf = Kmeans::Kernel::sigmoid;
  else if (kernel == "rbf") kf = Kmeans::Kernel::rbf;
  else {
    printf("Invalid kernel: %s\n", kernel.c_str());
    return 1;
  }
  if (level == 1 || level == 3 || level == 4) {
    // the reference's init_kernel_mtx refuses these (src/kernels/kernels.cuh:670-713)
    std::cerr << "ERROR, level " << level << " should never be used." << std::endl;
    return 1;
  }
  Kmeans::Options opts;
  opts.reference_gpu = !args.count("cpu-semantics");
  // the [PERFORMANCE] stage lines that experiments/driver.py:108-124 parses
  opts.report_stage_times = !args.count("no-stage-times");
  if (prec == "fp32") opts.precision = Kmeans::Precision::fp32;
  else if (prec == "bf16x3") opts.precision = Kmeans::Precision::bf16x3;
  else if (prec == "fp8") opts.precision = Kmeans::Precision::fp8;
  else if (prec == "bf16") opts.precision = Kmeans::Precision::bf16;
  else missing("precision");

  if (args.count("batch-runs")) {
    if (kf != Kmeans::Kernel::linear || opts.reference_gpu || im != Kmeans::InitMethod::random) {
      std::cerr << "--batch-runs needs the linear kernel, --cpu-semantics and -b random"
                << std::endl;
      return EXIT_ARGS;
    }
    const int st = batch_runs(pts, d, k, tol, seed, runs, maxiter, check_converged, prec, dev);
    if (st != 0) return st;
    if (out_file != "None") {
      std::ofstream fout(out_file);
      write_csv(fout, pts, d);
    }
    for (Point<float>* p : pts) delete p;
    return 0;
  }

  printf("%s", BOLDBLUE);
  double tot_time = 0, init_time = 0, score = 0;
  try {
    for (uint32_t r = 0; r < runs; ++r) {
      const auto i0 = std::chrono::high_resolution_clock::now();
      Kmeans km(n, (uint32_t)d, k, tol, &seed, pts.data(), &dev, im, kf, level, opts);
      const auto i1 = std::chrono::high_resolution_clock::now();
      if (r > 0) init_time += std::chrono::duration<double>(i1 - i0).count();
      const auto t0 = std::chrono::high_resolution_clock::now();
      const uint64_t conv = km.run(maxiter, check_converged);
      const auto t1 = std::chrono::high_resolution_clock::now();
      const double dur = std::chrono::duration<double>(t1 - t0).count();
      if (r > 0) tot_time += dur;  // the first run is a warm-up (src/main.cu:104)
      if (conv < maxiter) printf("K-means converged at iteration %lu - ", (unsigned long)conv);
      else printf("K-means did NOT converge - ");
      printf("Time: %lf\n", dur);
      printf("Objective score: %lf\n", km.get_score());
      size_t used = 0;
      if (kmc_memory_footprint(dev.ordinal, &used) == KMC_OK)
        std::cout << "MEMORY FOOTPRINT: " << (size_t)(used / 1e6) << " MB" << std::endl;
      score += km.get_score();
    }
  } catch (const kmc::Error& e) {
    printf("%s", RESET);
    std::cerr << e.what() << std::endl;
    return e.status == KMC_ERR_NO_DEVICE ? EXIT_DEV : 1;
  }
  // as the reference, the averages divide by runs-1 (src/main.cu:118-120)
  printf("GPU_Kmeans: %lfs (%u runs)\n", tot_time / (runs - 1), runs);
  printf("Init_time: %lfs \n", init_time / (runs - 1));
  printf("Score: %lf\n", score / runs);
  printf("%s", RESET);

  if (out_file != "None") {
    std::ofstream fout(out_file);
    write_csv(fout, pts, d);
  }
  for (Point<float>* p : pts) delete p;
  return 0;
}
