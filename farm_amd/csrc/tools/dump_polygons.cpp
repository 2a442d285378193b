// dump_polygons in.wkt out.dt: drop-in for the reference converter
// (src/datagen/dump_polygons.cpp:4-15): parse one POLYGON WKT per line
// (load_polygon_wkt, storage.cpp:45-94) and write the .dt format
// (dump_polygons_to_file, storage.cpp:158-193).
#include "driver.h"

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s in.wkt out.dt\n", argv[0]);
        return 1;
    }
    farm_hostset *s = farm_hostset_load_wkt(argv[1]);
    if (!s) {
        logf("cannot parse %s", argv[1]);
        return 1;
    }
    int rc = farm_hostset_save_dt(s, argv[2]);
    if (rc != FARM_OK) {
        logf("cannot write %s", argv[2]);
        return 1;
    }
    logf("dumped %lld polygons to %s", (long long)farm_hostset_size(s), argv[2]);
    farm_hostset_free(s);
    return 0;
}
