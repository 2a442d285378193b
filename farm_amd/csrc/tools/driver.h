// Shared command-line handling of the drop-in executables.
//
// Flags follow the reference's get_parameters (src/geometry/query_context.cpp:
// 227-272): -s/--source, -t/--target, -n/--threads, -v/--vpr,
// -b/--batch_size, -m/--merge_threshold, --NLow, -u/--unorll_size, and the
// switches -r/--rasterize, -g/--gpu, -a/--approximation, -h/--hierachy.  The
// reference's -w and -d never reach its option table (query_context.cpp:
// 245-247 end the chain with a comma); here they are registered:
// -w/--category_bits b selects 2^b fill categories for the approximate modes
// (exact results do not depend on it) and -d/--within_distance sets the
// within distance in km (default 10, the reference's fixed value).
#pragma once
#include <sys/time.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/farm_capi.h"
#include "../../../include/farm_host.h"

struct DriverArgs {
    std::string source, target;
    int threads = (int)std::thread::hardware_concurrency();
    int vpr = 10;
    unsigned long long batch = 0;
    float merge_threshold = 0.9f;
    int nlow = 1;
    int unroll = 16;
    int category_bits = 4;
    int within_km = 10;
    bool rasterize = false, gpu = false, approx = false, hierarchy = false;
    int device = 0;
    int tile = 0;
    int chunk = 0;                // target segments per chunk box (--chunk)
    std::string tune;             // tuning table file (--tune)
    bool plain = false;
    bool gpu_candidates = false;  // build the candidate pairs on the device instead of the host R-tree
    std::string model;            // trained MLP for the learned -a mode (--model)
    std::string wkt;              // intersection: write the polygons of A n B here (--wkt)
};

inline double now_ms() {
    struct timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec * 1000.0 + t.tv_usec / 1000.0;
}

inline void logf(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
}

inline void usage(const char *prog) {
    printf("%s usage:\n"
           "  -s, --source PATH       source polygons (.dt)\n"
           "  -t, --target PATH       target polygons (.dt)\n"
           "  -n, --threads N         host threads\n"
           "  -v, --vpr N             vertices per raster pixel (default 10)\n"
           "  -b, --batch_size N      candidate pairs per GPU batch (0 = all; a batch whose pixel\n"
           "                          pairs exceed the per-call limit is split in halves)\n"
           "  -m, --merge_threshold X accepted for compatibility (exact results do not depend on it)\n"
           "  -w, --category_bits B   intersect: 2^B fill categories (B in 2..8, default 4) for -a and the\n"
           "                          learned features; exact results do not depend on it\n"
           "  -d, --within_distance D within distance in km (default 10)\n"
           "  --NLow N                coarsest hierarchy layer size (default 1)\n"
           "  -u, --unorll_size N     accepted for compatibility\n"
           "  -r, -g, -a, -h          rasterize / gpu / approximation / hierarchy\n"
           "  --plain                 one-pair-per-workgroup kernel instead of the batched one\n"
           "  --tile E                source edges per lane in the batched kernel (1, 2, 4 or 8)\n"
           "  --chunk C               target segments per chunk box in the batched kernel (8, 16 or 32)\n"
           "  --tune FILE             intersect: the refinement configuration of the set's shape bucket\n"
           "                          from FILE; when FILE has none, autotune on the first batch and\n"
           "                          add it to FILE (--tile / --chunk override it)\n"
           "  --model FILE            with -a: the learned mode with this trained MLP (farm_amd/learned.py\n"
           "                          writes it) instead of the reference heuristic\n"
           "  --gpu-candidates        build the candidate pairs on the device (same pairs as the R-tree)\n"
           "  --wkt FILE              intersection: write A n B of every intersecting pair as WKT lines\n"
           "                          <source id> <target id> MULTIPOLYGON (...), tab-separated\n"
           "  --device N              HIP device\n",
           prog);
}

inline bool parse_args(int argc, char **argv, DriverArgs &a) {
    for (int i = 1; i < argc; i++) {
        std::string k = argv[i], v;
        size_t eq = k.find('=');
        bool has_eq = k.rfind("--", 0) == 0 && eq != std::string::npos;
        if (has_eq) {
            v = k.substr(eq + 1);
            k = k.substr(0, eq);
        }
        auto val = [&](void) -> const char * {
            if (has_eq) return v.c_str();
            if (i + 1 >= argc) {
                printf("Undefined Option!\n");
                exit(0);
            }
            return argv[++i];
        };
        if (k == "--help") {
            usage(argv[0]);
            exit(0);
        } else if (k == "-r" || k == "--rasterize") a.rasterize = true;
        else if (k == "-g" || k == "--gpu") a.gpu = true;
        else if (k == "-a" || k == "--approximation") a.approx = true;
        else if (k == "-h" || k == "--hierachy") a.hierarchy = true;
        else if (k == "--plain") a.plain = true;
        else if (k == "-s" || k == "--source") a.source = val();
        else if (k == "-t" || k == "--target") a.target = val();
        else if (k == "-n" || k == "--threads") a.threads = atoi(val());
        else if (k == "-v" || k == "--vpr") a.vpr = atoi(val());
        else if (k == "-b" || k == "--batch_size") a.batch = strtoull(val(), nullptr, 10);
        else if (k == "-m" || k == "--merge_threshold") a.merge_threshold = (float)atof(val());
        else if (k == "--NLow") a.nlow = atoi(val());
        else if (k == "-u" || k == "--unorll_size") a.unroll = atoi(val());
        else if (k == "-w" || k == "--category_bits") a.category_bits = atoi(val());
        else if (k == "-d" || k == "--within_distance") a.within_km = atoi(val());
        else if (k == "--device") a.device = atoi(val());
        else if (k == "--tile") a.tile = atoi(val());
        else if (k == "--chunk") a.chunk = atoi(val());
        else if (k == "--tune") a.tune = val();
        else if (k == "--gpu-candidates") a.gpu_candidates = true;
        else if (k == "--model") a.model = val();
        else if (k == "--wkt") a.wkt = val();
        else {
            // boost::program_options rejection of the reference (query_context.cpp:250-255)
            printf("Undefined Option!\n");
            exit(0);
        }
    }
    if (a.threads < 1) a.threads = 1;
    if (a.category_bits < 2 || a.category_bits > 8) {
        fprintf(stderr, "-w must be in 2..8 (2^w fill categories in a byte)\n");
        return false;
    }
    if (a.vpr < 1) {
        fprintf(stderr, "-v must be positive\n");
        return false;
    }
    return true;
}

inline farm_hostset *load_or_die(const std::string &path) {
    double t = now_ms();
    farm_hostset *s = farm_hostset_load_dt(path.c_str());
    if (!s) {
        logf("%s does not exist or is not a .dt file", path.c_str());
        exit(1);
    }
    logf("loaded %lld polygons from %s takes %.3f ms", (long long)farm_hostset_size(s), path.c_str(),
         now_ms() - t);
    return s;
}

#define DIE_IF(h, rc, what)                                                                 \
    do {                                                                                    \
        int rc_ = (rc);                                                                     \
        if (rc_ != FARM_OK) {                                                               \
            logf("%s failed (%d): %s", what, rc_, (h) ? farm_last_error(h) : "");           \
            exit(1);                                                                        \
        }                                                                                   \
    } while (0)

// --gpu-candidates: the pairs of farm_candidates(src, tgt, km, skip_self),
// built on the device after the polygon upload (target ids shifted by
// b_base) and kept there as the pair list; returns their count.
inline uint64_t device_candidates(farm_handle h, const farm_hostset *src, const farm_hostset *tgt, double km,
                                  int skip_self, uint32_t b_base) {
    const int64_t ns = farm_hostset_size(src), nt = farm_hostset_size(tgt);
    std::vector<double> ma(4 * (size_t)ns), mb(4 * (size_t)nt);
    DIE_IF(h, farm_hostset_mbbs(src, ma.data()), "source boxes");
    DIE_IF(h, farm_hostset_mbbs(tgt, mb.data()), "target boxes");
    const double t = now_ms();
    uint64_t n = 0;
    DIE_IF(h, farm_generate_pairs(h, ma.data(), ns, mb.data(), nt, km, skip_self, 0, b_base, &n), "generate pairs");
    logf("device candidate pairs: %llu takes %.3f ms", (unsigned long long)n, now_ms() - t);
    return n;
}

// A trained MLP for the learned -a mode, as farm_amd/learned.py save_model
// writes it: the 8 bytes "FARMMLP1", u32 parameter count, u32 feature count,
// f32 accept and reject logits, the parameters, then the feature
// normalisation (means, then reciprocal deviations), all little-endian.
struct MlpModel {
    std::vector<float> params, norm;
    float accept = 0, reject = 0;
};

inline bool load_model(const std::string &path, MlpModel &m) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    uint32_t np = 0, nf = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, "FARMMLP1", 8) && fread(&np, 4, 1, f) == 1 &&
              fread(&nf, 4, 1, f) == 1 && np == FARM_MLP_NPARAMS && nf == FARM_NFEAT &&
              fread(&m.accept, 4, 1, f) == 1 && fread(&m.reject, 4, 1, f) == 1;
    if (ok) {
        m.params.resize(np);
        m.norm.resize(2 * nf);
        ok = fread(m.params.data(), 4, np, f) == np && fread(m.norm.data(), 4, 2 * nf, f) == 2 * nf;
    }
    fclose(f);
    return ok;
}

// farm_intersect over pairs [b, e), split in halves while a window's pixel
// pairs exceed the per-call limit (FARM_ECAPACITY) or the device memory
// (FARM_ENOMEM); adds the counters.
inline void intersect_split(farm_handle h, uint64_t b, uint64_t e, const farm_opts *o, uint64_t &found,
                            uint64_t &tests) {
    farm_stats st;
    const int rc = farm_intersect(h, b, e, o, nullptr, &st);
    if ((rc == FARM_ECAPACITY || rc == FARM_ENOMEM) && e - b > 1) {
        const uint64_t m = b + (e - b) / 2;
        intersect_split(h, b, m, o, found, tests);
        intersect_split(h, m, e, o, found, tests);
        return;
    }
    DIE_IF(h, rc, "intersect");
    found += st.found;
    tests += st.segment_tests;
}

