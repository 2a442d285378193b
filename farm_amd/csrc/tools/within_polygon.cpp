// within_polygon: drop-in for the reference's within_polygon /
// within_polygon_cpu drivers (src/query/within_polygon.cpp:26-116,
// within_polygon_cpu.cpp:20-146): a self-join of the -s set; candidate pairs
// are every (i, j), i != j, whose MBBs overlap after j's MBB is expanded by
// the within distance (both orders kept; built on the device with
// --gpu-candidates); the -h hierarchy is built on the
// host and the within decision runs on the GPU in -b batches.  Prints
// count-found like query_context::print_stats (query_context.cpp:173).
#include <vector>

#include "driver.h"

int main(int argc, char **argv) {
    DriverArgs a;
    if (!parse_args(argc, argv, a)) return 1;
    if (a.source.empty()) {
        usage(argv[0]);
        return 1;
    }
    if (!a.hierarchy) {
        // Ideal::within reads layers[0] (Ideal.cpp:2483), which only -h builds
        logf("within_polygon needs -h (hierarchy)");
        return 1;
    }
    if (!a.target.empty() && a.target != a.source)
        logf("-t is ignored: within_polygon joins the -s set with itself");
    farm_hostset *src = load_or_die(a.source);

    double t = now_ms();
    uint32_t *pa = nullptr, *pb = nullptr;
    int64_t npairs = 0;
    if (!a.gpu_candidates) {
        npairs = farm_candidates(src, src, (double)a.within_km, 1, a.threads, &pa, &pb);
        if (npairs < 0) {
            logf("candidate generation failed");
            return 1;
        }
        logf("rtree query finished: %lld candidate pairs takes %.3f ms", (long long)npairs, now_ms() - t);
    }

    t = now_ms();
    if (farm_hostset_rasterize_h(src, a.vpr, a.nlow, a.threads) != FARM_OK) {
        logf("rasterization failed");
        return 1;
    }
    const int64_t bad = farm_hostset_raster_failures(src);
    if (bad >= 0) {
        logf("raster builder rejects polygon %lld", (long long)bad);
        return 1;
    }
    logf("preprocess finished takes %.3f ms", now_ms() - t);

    farm_handle h = nullptr;
    if (farm_open(a.device, &h) != FARM_OK) {
        logf("cannot open HIP device %d", a.device);
        return 1;
    }
    t = now_ms();
    farm_polyset ps;
    DIE_IF(h, farm_hostset_view(src, &ps), "view");
    DIE_IF(h, farm_upload_polygons(h, &ps), "upload polygons");
    if (a.gpu_candidates) npairs = (int64_t)device_candidates(h, src, src, (double)a.within_km, 1, 0);
    else DIE_IF(h, farm_upload_pairs(h, pa, pb, (uint64_t)npairs), "upload pairs");
    logf("preprocess for gpu finished takes %.3f ms", now_ms() - t);

    const uint64_t batch = a.batch ? a.batch : (uint64_t)npairs;
    uint64_t found = 0, dists = 0;
    t = now_ms();
    for (uint64_t b = 0; b < (uint64_t)npairs; b += batch) {
        const uint64_t e = std::min<uint64_t>(b + batch, (uint64_t)npairs);
        farm_stats st;
        DIE_IF(h, farm_within(h, b, e, (double)a.within_km, nullptr, nullptr, &st), "within");
        found += st.found;
        dists += st.segment_tests;
    }
    const double qms = now_ms() - t;
    logf("count-found:\t%llu", (unsigned long long)found);
    logf("distance-evaluations:\t%llu", (unsigned long long)dists);
    logf("query finished takes %.3f ms (%.1f pairs/s)", qms, npairs / (qms / 1000.0));
    farm_close(h);
    farm_free(pa);
    farm_free(pb);
    farm_hostset_free(src);
    return 0;
}
