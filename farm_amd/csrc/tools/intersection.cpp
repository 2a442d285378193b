// intersection: drop-in for the reference's intersection / intersection_cpu
// drivers (src/query/intersection.cpp:40-97, intersection_cpu.cpp:67-124):
// load two .dt sets, candidate pairs from the host R-tree (on the device with
// --gpu-candidates), host rasters, then
// the intersection query on the GPU in -b batches: crossing points, boundary
// pieces and their statuses (where the reference stops) and the area of each
// pair's intersection.  Prints count-found (pairs with a positive area) and
// the summed area.  -a is accepted; the areas are exact either way.
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
        for (int64_t i = 0; i < npairs; i++) pb[i] += (uint32_t)ns;
    }

    t = now_ms();
    farm_hostset *all = farm_hostset_concat(src, tgt);
    if (farm_hostset_rasterize(all, a.vpr, a.threads) != FARM_OK) {
        logf("rasterization failed");
        return 1;
    }
    const int64_t bad = farm_hostset_raster_failures(all);
    if (bad >= 0) {
        logf("raster builder rejects polygon %lld (%s %lld)", (long long)bad, bad < ns ? "source" : "target",
             (long long)(bad < ns ? bad : bad - ns));
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

    const uint64_t batch = a.batch ? a.batch : (uint64_t)npairs;
    uint64_t found = 0, points = 0, pieces = 0;
    double total_area = 0.0;
    std::vector<int32_t> counts;
    std::vector<double> area;
    FILE *wkt = nullptr;
    if (!a.wkt.empty() && !(wkt = fopen(a.wkt.c_str(), "w"))) {
        logf("cannot write %s", a.wkt.c_str());
        return 1;
    }
    std::vector<farm_piece> pcs(1 << 16);
    std::vector<uint32_t> wa, wb;
    t = now_ms();
    for (uint64_t b = 0; b < (uint64_t)npairs; b += batch) {
        const uint64_t e = std::min<uint64_t>(b + batch, (uint64_t)npairs);
        counts.resize(5 * (e - b));
        area.resize(e - b);
        farm_stats st;
        if (!wkt) {
            DIE_IF(h, farm_intersection_area(h, b, e, counts.data(), area.data(), &st), "intersection");
        } else {
            // the boundary pieces of A n B, then their rings as WKT
            uint64_t np = 0;
            int rc = farm_intersection_pieces(h, b, e, counts.data(), area.data(), pcs.data(), pcs.size(), &np, &st);
            if (rc == FARM_ECAPACITY && np > pcs.size()) {
                pcs.resize(np);
                rc = farm_intersection_pieces(h, b, e, counts.data(), area.data(), pcs.data(), pcs.size(), &np, &st);
            }
            DIE_IF(h, rc, "intersection pieces");
            wa.resize(e - b);
            wb.resize(e - b);
            if (a.gpu_candidates) DIE_IF(h, farm_download_pairs(h, b, e, wa.data(), wb.data()), "download pairs");
            else {
                std::copy(pa + b, pa + e, wa.begin());
                std::copy(pb + b, pb + e, wb.begin());
            }
            uint64_t len = 0;
            char *text = farm_pieces_wkt(all, wa.data(), wb.data(), e - b, pcs.data(), np, &len);
            if (!text) {
                logf("could not assemble the polygons of A n B");
                return 1;
            }
            // "<k>\t<wkt>" per window index k -> "<source id>\t<target id>\t<wkt>"
            for (char *line = text; *line;) {
                char *nl = strchr(line, '\n');
                if (nl) *nl = 0;
                char *tab = strchr(line, '\t');
                if (tab) {
                    const uint64_t k = strtoull(line, nullptr, 10);
                    fprintf(wkt, "%u\t%u\t%s\n", wa[k], wb[k] - (uint32_t)ns, tab + 1);
                }
                if (!nl) break;
                line = nl + 1;
            }
            farm_free(text);
        }
        found += st.found;
        for (uint64_t k = 0; k < e - b; k++) {
            if (counts[5 * k] > 0) points += (uint64_t)counts[5 * k];
            pieces += (uint64_t)counts[5 * k + 1];
            total_area += area[k];
        }
    }
    const double qms = now_ms() - t;
    logf("count-found:\t%llu", (unsigned long long)found);
    logf("crossing-points:\t%llu", (unsigned long long)points);
    logf("boundary-pieces:\t%llu", (unsigned long long)pieces);
    logf("total-area:\t%.12g", total_area);
    logf("query finished takes %.3f ms (%.1f pairs/s)", qms, npairs / (qms / 1000.0));
    if (wkt) fclose(wkt);
    farm_close(h);
    farm_free(pa);
    farm_free(pb);
    farm_hostset_free(all);
    farm_hostset_free(src);
    farm_hostset_free(tgt);
    return 0;
}
