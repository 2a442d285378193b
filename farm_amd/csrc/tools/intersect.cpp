// intersect: drop-in for the reference's intersect / intersect_cpu drivers
// (src/query/intersect.cpp:39-97, intersect_cpu.cpp:65-126): load two .dt
// sets, generate candidate pairs with the host R-tree (or on the device with
// --gpu-candidates), rasterize on the host, then run the exact (or -a
// approximate) intersect on the GPU in -b batches.
// Prints count-found like query_context::print_stats (query_context.cpp:173).
#include <vector>

#include "driver.h"

int main(int argc, char **argv) {
    DriverArgs a;
    if (!parse_args(argc, argv, a)) return 1;
    if (a.source.empty() || a.target.empty()) {
        usage(argv[0]);
        return 1;
    }
    farm_hostset *src = load_or_die(a.source);
    farm_hostset *tgt = load_or_die(a.target);
    const int64_t ns = farm_hostset_size(src);

    double t = now_ms();
    uint32_t *pa = nullptr, *pb = nullptr;
    int64_t npairs = 0;
    if (!a.gpu_candidates) {
        npairs = farm_candidates(src, tgt, 0.0, 0, a.threads, &pa, &pb);
        if (npairs < 0) {
            logf("candidate generation failed");
            return 1;
        }
        logf("rtree query finished: %lld candidate pairs takes %.3f ms", (long long)npairs, now_ms() - t);
        for (int64_t i = 0; i < npairs; i++) pb[i] += (uint32_t)ns;  // sources then targets in one set
    }

    t = now_ms();
    farm_hostset *all = farm_hostset_concat(src, tgt);
    if (farm_hostset_rasterize(all, a.vpr, a.threads) != FARM_OK) {
        logf("rasterization failed");
        return 1;
    }
    const int64_t bad = farm_hostset_raster_failures(all);
    if (bad >= 0) {
        // the reference raster builder asserts (or never terminates) here
        logf("raster builder rejects polygon %lld (%s %lld)", (long long)bad, bad < ns ? "source" : "target",
             (long long)(bad < ns ? bad : bad - ns));
        return 1;
    }
    // -w b: fill classes at 2^b categories for -a (exact results do not depend on it)
    if (farm_hostset_set_categories(all, a.category_bits) != FARM_OK) {
        logf("-w %d: bad category bits", a.category_bits);
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
    DIE_IF(h, farm_hostset_view(all, &ps), "view");
    DIE_IF(h, farm_upload_polygons(h, &ps), "upload polygons");
    if (a.gpu_candidates) npairs = (int64_t)device_candidates(h, src, tgt, 0.0, 0, (uint32_t)ns);
    else DIE_IF(h, farm_upload_pairs(h, pa, pb, (uint64_t)npairs), "upload pairs");
    logf("preprocess for gpu finished takes %.3f ms", now_ms() - t);

    farm_opts o;
    memset(&o, 0, sizeof o);
    o.mode = a.approx ? FARM_MODE_APPROX_HEURISTIC : (a.plain ? FARM_MODE_EXACT_PLAIN : FARM_MODE_EXACT_BATCHED);
    o.tile_a = a.tile;
    o.tile_b = a.chunk;
    o.merge_threshold = a.merge_threshold;
    if (!a.tune.empty() && o.mode == FARM_MODE_EXACT_BATCHED && npairs > 0) {
        // the set's shape bucket from the file, else tuned on the first batch
        FILE *f = fopen(a.tune.c_str(), "r");
        if (f) {
            fclose(f);
            DIE_IF(h, farm_tuning_load(h, a.tune.c_str()), "load tuning");
        }
        farm_tuning tn;
        DIE_IF(h, farm_get_tuning(h, &tn), "tuning");
        if (tn.ms <= 0) {
            const uint64_t w = std::min<uint64_t>((uint64_t)npairs, a.batch ? a.batch : 3000000);
            DIE_IF(h, farm_autotune_config(h, 0, w, &tn), "autotune");
            DIE_IF(h, farm_tuning_save(h, a.tune.c_str()), "save tuning");
            logf("autotuned on %llu pairs: tile %d, chunk %d, %d workgroups per CU (%.3f ms)",
                 (unsigned long long)w, tn.tile_a, tn.chunk, tn.blocks_per_cu, tn.ms);
        } else {
            logf("tuning of shape bucket %d from %s: tile %d, chunk %d, %d workgroups per CU", tn.shape_bucket,
                 a.tune.c_str(), tn.tile_a, tn.chunk, tn.blocks_per_cu);
        }
    }
    if (!a.model.empty()) {
        if (!a.approx) {
            logf("--model needs -a");
            return 1;
        }
        MlpModel m;
        if (!load_model(a.model, m)) {
            logf("%s is not a trained MLP file", a.model.c_str());
            return 1;
        }
        DIE_IF(h, farm_mlp_set_params(h, m.params.data(), m.params.size(), m.norm.data()), "load MLP");
        o.mode = FARM_MODE_APPROX_LEARNED;
        o.approx_threshold = m.accept;
        o.approx_reject = m.reject;
        logf("learned -a: accept logit > %g, reject < %g, exact refinement between", m.accept, m.reject);
    }
    // -b 0 is one batch of all pairs, as the reference; any batch whose
    // BORDER x BORDER pixel pairs exceed the per-call limit is split in halves
    const uint64_t batch = a.batch ? a.batch : std::max<uint64_t>(1, (uint64_t)npairs);
    uint64_t found = 0, tests = 0;
    t = now_ms();
    for (uint64_t b = 0; b < (uint64_t)npairs; b += batch)
        intersect_split(h, b, std::min<uint64_t>(b + batch, (uint64_t)npairs), &o, found, tests);
    const double qms = now_ms() - t;
    logf("count-found:\t%llu", (unsigned long long)found);
    logf("segment-tests:\t%llu", (unsigned long long)tests);
    logf("query finished takes %.3f ms (%.1f pairs/s)", qms, npairs / (qms / 1000.0));
    farm_close(h);
    farm_free(pa);
    farm_free(pb);
    farm_hostset_free(all);
    farm_hostset_free(src);
    farm_hostset_free(tgt);
    return 0;
}
