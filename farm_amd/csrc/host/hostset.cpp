// Host polygon storage (.dt / WKT), raster packing into the device layout,
// and candidate-pair generation with a host R-tree.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <thread>
#include <vector>

#include <unistd.h>

#include "../../../include/farm_host.h"
#include "raster.h"

struct farm_hostset {
    // stored rings (as in the file; no appended vertex)
    std::vector<uint64_t> ring_off{0};
    std::vector<float> xy;
    // holes, kept for .dt round trips only
    std::vector<uint64_t> hole_first{0};  // per polygon: range into hole_off
    std::vector<uint64_t> hole_off{0};
    std::vector<float> hole_xy;
    std::vector<double> mbb;  // 4 per polygon
    // device layout (after rasterize)
    bool rastered = false;
    int64_t raster_fail = -1;
    std::vector<uint8_t> raster_ok;  // per polygon
    std::vector<farm_raster_hdr> hdr;
    std::vector<uint8_t> status;
    std::vector<double> areas;  // per pixel, indexed like status
    std::vector<uint32_t> pixoff, seqs, voff;
    std::vector<double> vnodes;
    std::vector<float> vx, vy;
    // -w b: fill classes at cc = 2^b categories (empty: the status bytes, 16)
    std::vector<uint8_t> fill;
    int cc = 16;
    // hierarchy (-h)
    bool hier = false;
    std::vector<farm_layer> layers;
    std::vector<uint8_t> lstatus;
};

namespace {

// VertexSequence::getMBR (vertex_sequence.cpp:152-167) over the decoded ring
// (stored vertices plus the repeated first one, which changes nothing).
void ring_mbb(const float *xy, uint64_t n, double *m) {
    double lx = 100000.0, ly = 100000.0, hx = -100000.0, hy = -100000.0;
    for (uint64_t k = 0; k < n; k++) {
        lx = std::min(lx, (double)xy[2 * k]);
        hx = std::max(hx, (double)xy[2 * k]);
        ly = std::min(ly, (double)xy[2 * k + 1]);
        hy = std::max(hy, (double)xy[2 * k + 1]);
    }
    const float bias = 1e-9f;
    m[0] = lx - bias;
    m[1] = ly - bias;
    m[2] = hx + bias;
    m[3] = hy + bias;
}

void push_polygon(farm_hostset *s, const float *xy, uint64_t nv) {
    s->xy.insert(s->xy.end(), xy, xy + 2 * nv);
    s->ring_off.push_back(s->xy.size() / 2);
    s->hole_first.push_back(s->hole_off.size() - 1);
    double m[4];
    ring_mbb(xy, nv, m);
    s->mbb.insert(s->mbb.end(), m, m + 4);
}

int64_t npoly(const farm_hostset *s) { return (int64_t)s->ring_off.size() - 1; }

// ---------------------------------------------------------------- WKT ----
// read_float (util.h:129-144): skip to the next number character, take the
// run of [-.0-9e], convert with stof (strtof rounding).
bool is_num(char c) { return c == '-' || c == '.' || (c <= '9' && c >= '0') || c == 'e'; }
float read_float(const char *w, size_t &off, size_t len, bool &ok) {
    while (off < len && !is_num(w[off])) off++;
    char tmp[100];
    int k = 0;
    while (off < len && is_num(w[off]) && k < 99) tmp[k++] = w[off++];
    tmp[k] = 0;
    if (k == 0) {
        ok = false;
        return 0.f;
    }
    return strtof(tmp, nullptr);
}
void skip_sp(const char *w, size_t &off, size_t len) {
    while (off < len && (w[off] == ' ' || w[off] == '\t' || w[off] == '\n' || w[off] == '\r')) off++;
}
// read_vertices (storage.cpp:637-684): vertex count = commas before ')' + 1
bool read_ring(const char *w, size_t &off, size_t len, std::vector<float> &out) {
    skip_sp(w, off, len);
    if (off >= len || w[off++] != '(') return false;
    size_t c = off;
    int nv = 0;
    while (c < len && w[c++] != ')')
        if (c < len && w[c] == ',') nv++;
    nv++;
    out.clear();
    bool ok = true;
    for (int i = 0; i < nv; i++) {
        out.push_back(read_float(w, off, len, ok));
        out.push_back(read_float(w, off, len, ok));
    }
    skip_sp(w, off, len);
    if (off >= len || w[off++] != ')') return false;
    return ok;
}

}  // namespace

namespace farm {

// A set from closed rings already in the stored layout (every ring has at
// least 3 stored vertices), taking the arrays over; MBBs in parallel.
farm_hostset *hostset_adopt(std::vector<uint64_t> &&ring_off, std::vector<float> &&xy, int nthreads) {
    farm_hostset *s = new farm_hostset();
    const int64_t n = (int64_t)ring_off.size() - 1;
    s->ring_off = std::move(ring_off);
    s->xy = std::move(xy);
    s->hole_first.assign((size_t)n + 1, 0);
    s->mbb.resize(4 * (size_t)n);
    const int nth = std::max(1, nthreads);
    auto work = [&](int t) {
        for (int64_t i = t; i < n; i += nth)
            ring_mbb(s->xy.data() + 2 * s->ring_off[i], s->ring_off[i + 1] - s->ring_off[i], &s->mbb[4 * i]);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; t++) th.emplace_back(work, t);
    work(0);
    for (auto &q : th) q.join();
    return s;
}

}  // namespace farm

extern "C" {

farm_hostset *farm_hostset_from_rings(int64_t n, const uint64_t *ring_off, const float *xy) {
    farm_hostset *s = new farm_hostset();
    for (int64_t i = 0; i < n; i++) {
        uint64_t b = ring_off[i], e = ring_off[i + 1];
        if (e < b) {
            delete s;
            return nullptr;
        }
        if (e - b + 1 < 3) continue;  // storage.cpp:341
        push_polygon(s, xy + 2 * b, e - b);
    }
    return s;
}

// load_binary_file (storage.cpp:361-449): the meta table at the end gives
// the first record; records are decoded back to back (load_unit
// storage.cpp:320-360, MyPolygon::decode MyPolygon.cpp:747-759) and rings with
// fewer than 3 stored vertices are dropped (storage.cpp:341).  Ids follow
// file order.  The file is read with parallel preads, one pass over the
// record headers finds every kept ring and hole, and the vertex copies and
// boxes then run in parallel over ranges of polygons.
farm_hostset *farm_hostset_load_dt(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return nullptr;
    fseek(f, 0, SEEK_END);
    const long fsz_l = ftell(f);
    if (fsz_l < 8) {
        fclose(f);
        return nullptr;
    }
    const uint64_t fsz = (uint64_t)fsz_l;
    std::unique_ptr<char[]> storage(new (std::nothrow) char[fsz]);  // no zero fill: pread writes all of it
    if (!storage) {
        fclose(f);
        return nullptr;
    }
    struct {
        char *p;
        char *data() const { return p; }
    } buf{storage.get()};
    const int fd = fileno(f);
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    {
        const uint64_t piece = (fsz + nth - 1) / nth;
        std::atomic<bool> ok(true);
        std::vector<std::thread> th;
        for (int t = 0; t < nth; t++)
            th.emplace_back([&, t]() {
                uint64_t o = t * piece, e = std::min(fsz, o + piece);
                while (o < e) {
                    const ssize_t r = pread(fd, buf.data() + o, e - o, (off_t)o);
                    if (r <= 0) {
                        ok = false;
                        return;
                    }
                    o += (uint64_t)r;
                }
            });
        for (auto &t : th) t.join();
        fclose(f);
        if (!ok) return nullptr;
    }
    uint64_t n;
    memcpy(&n, buf.data() + fsz - 8, 8);
    const uint64_t meta_bytes = 48;  // PolygonMeta (MyPolygon.h:29-35)
    if (n == 0 || n > (fsz - 8) / meta_bytes) return nullptr;
    const char *meta = buf.data() + fsz - 8 - n * meta_bytes;
    uint64_t first_off, last_off;
    uint32_t last_size;
    memcpy(&first_off, meta + 8, 8);
    memcpy(&last_off, meta + (n - 1) * meta_bytes + 8, 8);
    memcpy(&last_size, meta + (n - 1) * meta_bytes, 4);
    const uint64_t data_end = last_off + last_size;
    if (data_end > fsz - 8 - n * meta_bytes || first_off > data_end) return nullptr;
    // one pass over the record headers: the kept rings and their holes
    struct Rec {
        uint64_t ring, nv, hole0;  // ring offset, stored vertices, first entry in holes
    };
    std::vector<Rec> recs;
    std::vector<std::pair<uint64_t, uint64_t>> holes;  // (offset, vertices)
    uint64_t off = first_off;
    auto rd64 = [&](uint64_t &v) -> bool {
        if (off + 8 > data_end) return false;
        memcpy(&v, buf.data() + off, 8);
        off += 8;
        return true;
    };
    while (off < data_end) {
        uint64_t nh, nv;
        if (!rd64(nh) || !rd64(nv) || nv == 0 || nv > (data_end - off) / 8) return nullptr;
        const uint64_t ring = off;
        off += nv * 8;
        const size_t h0 = holes.size();
        for (uint64_t h = 0; h < nh; h++) {
            uint64_t hn;
            if (!rd64(hn) || hn > (data_end - off) / 8) return nullptr;
            holes.push_back({off, hn});
            off += hn * 8;
        }
        if (nv + 1 < 3) {
            holes.resize(h0);
            continue;
        }
        recs.push_back({ring, nv, (uint64_t)h0});
    }
    // sizes from prefix sums, then the copies in parallel
    farm_hostset *s = new farm_hostset();
    const size_t np = recs.size();
    s->ring_off.resize(np + 1);
    s->hole_first.resize(np + 1);
    s->hole_off.resize(holes.size() + 1);
    s->ring_off[0] = 0;
    s->hole_off[0] = 0;
    for (size_t i = 0; i < np; i++) s->ring_off[i + 1] = s->ring_off[i] + recs[i].nv;
    for (size_t k = 0; k < holes.size(); k++) s->hole_off[k + 1] = s->hole_off[k] + holes[k].second;
    // hole_first[i + 1]: the hole_off entry that ends polygon i's holes
    s->hole_first[0] = 0;
    for (size_t i = 0; i < np; i++)
        s->hole_first[i + 1] = (i + 1 < np ? recs[i + 1].hole0 : holes.size());
    s->xy.resize(2 * s->ring_off[np]);
    s->hole_xy.resize(2 * s->hole_off[holes.size()]);
    s->mbb.resize(4 * np);
    {
        std::atomic<size_t> next(0);
        const size_t chunk = 1024;
        std::vector<std::thread> th;
        for (int t = 0; t < nth; t++)
            th.emplace_back([&]() {
                for (;;) {
                    const size_t c0 = next.fetch_add(chunk);
                    if (c0 >= np) break;
                    for (size_t i = c0; i < std::min(np, c0 + chunk); i++) {
                        const float *ring = reinterpret_cast<const float *>(buf.data() + recs[i].ring);
                        memcpy(&s->xy[2 * s->ring_off[i]], ring, 8 * recs[i].nv);
                        ring_mbb(ring, recs[i].nv, &s->mbb[4 * i]);
                        const uint64_t hend = i + 1 < np ? recs[i + 1].hole0 : holes.size();
                        for (uint64_t k = recs[i].hole0; k < hend; k++)
                            memcpy(&s->hole_xy[2 * s->hole_off[k]], buf.data() + holes[k].first, 8 * holes[k].second);
                    }
                }
            });
        for (auto &t : th) t.join();
    }
    return s;
}

farm_hostset *farm_hostset_load_wkt(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return nullptr;
    farm_hostset *s = new farm_hostset();
    std::string line;
    std::vector<float> ring;
    std::vector<std::vector<float>> holes;
    char chunk[1 << 16];
    auto handle = [&](const std::string &l) -> bool {
        if (l.empty()) return true;
        size_t off = l.find('(');
        if (off == std::string::npos) return false;
        const char *w = l.c_str();
        size_t len = l.size();
        skip_sp(w, off, len);
        if (w[off++] != '(') return false;
        if (!read_ring(w, off, len, ring)) return false;
        skip_sp(w, off, len);
        holes.clear();
        while (off < len && w[off] == ',') {
            off++;
            std::vector<float> h;
            if (!read_ring(w, off, len, h)) return false;
            holes.push_back(std::move(h));
            skip_sp(w, off, len);
        }
        // the WKT loader keeps every polygon; the .dt loader drops tiny rings
        s->xy.insert(s->xy.end(), ring.begin(), ring.end());
        s->ring_off.push_back(s->xy.size() / 2);
        double m[4];
        ring_mbb(ring.data(), ring.size() / 2, m);
        s->mbb.insert(s->mbb.end(), m, m + 4);
        for (auto &h : holes) {
            s->hole_xy.insert(s->hole_xy.end(), h.begin(), h.end());
            s->hole_off.push_back(s->hole_xy.size() / 2);
        }
        s->hole_first.push_back(s->hole_off.size() - 1);
        return true;
    };
    bool ok = true;
    while (ok) {
        size_t got = fread(chunk, 1, sizeof chunk, f);
        for (size_t i = 0; i < got && ok; i++) {
            if (chunk[i] == '\n') {
                ok = handle(line);
                line.clear();
            } else {
                line.push_back(chunk[i]);
            }
        }
        if (got < sizeof chunk) break;
    }
    if (ok && !line.empty()) ok = handle(line);
    fclose(f);
    if (!ok) {
        delete s;
        return nullptr;
    }
    return s;
}

// dump_polygons_to_file (storage.cpp:158-193); layout of MyPolygon::encode
// (MyPolygon.cpp:736-745) and PolygonMeta (MyPolygon.h:29-35).
int farm_hostset_save_dt(const farm_hostset *s, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) return FARM_EINVAL;
    const int64_t n = npoly(s);
    std::vector<char> meta((size_t)n * 48);
    uint64_t cur = 0;
    for (int64_t i = 0; i < n; i++) {
        const uint64_t nv = s->ring_off[i + 1] - s->ring_off[i];
        const uint64_t h0 = s->hole_first[i], h1 = s->hole_first[i + 1];
        const uint64_t nh = h1 - h0;
        uint64_t size = 8 + 8 + 8 * nv;
        fwrite(&nh, 8, 1, f);
        fwrite(&nv, 8, 1, f);
        fwrite(s->xy.data() + 2 * s->ring_off[i], 8, nv, f);
        for (uint64_t h = h0; h < h1; h++) {
            uint64_t hn = s->hole_off[h + 1] - s->hole_off[h];
            fwrite(&hn, 8, 1, f);
            fwrite(s->hole_xy.data() + 2 * s->hole_off[h], 8, hn, f);
            size += 8 + 8 * hn;
        }
        char *m = meta.data() + i * 48;
        uint32_t sz32 = (uint32_t)size, nv32 = (uint32_t)nv;
        memcpy(m, &sz32, 4);
        memcpy(m + 4, &nv32, 4);
        memcpy(m + 8, &cur, 8);
        memcpy(m + 16, &s->mbb[4 * i], 32);
        cur += size;
    }
    fwrite(meta.data(), 1, meta.size(), f);
    uint64_t nn = (uint64_t)n;
    fwrite(&nn, 8, 1, f);
    int rc = ferror(f) ? FARM_EINVAL : FARM_OK;
    fclose(f);
    return rc;
}

farm_hostset *farm_hostset_concat(const farm_hostset *a, const farm_hostset *b) {
    farm_hostset *s = new farm_hostset();
    for (const farm_hostset *src : {a, b}) {
        const uint64_t vbase = s->xy.size() / 2, hbase = s->hole_off.size() - 1, hxy = s->hole_xy.size() / 2;
        s->xy.insert(s->xy.end(), src->xy.begin(), src->xy.end());
        for (size_t i = 1; i < src->ring_off.size(); i++) s->ring_off.push_back(src->ring_off[i] + vbase);
        s->hole_xy.insert(s->hole_xy.end(), src->hole_xy.begin(), src->hole_xy.end());
        for (size_t i = 1; i < src->hole_off.size(); i++) s->hole_off.push_back(src->hole_off[i] + hxy);
        for (size_t i = 1; i < src->hole_first.size(); i++) s->hole_first.push_back(src->hole_first[i] + hbase);
        s->mbb.insert(s->mbb.end(), src->mbb.begin(), src->mbb.end());
    }
    return s;
}

void farm_hostset_free(farm_hostset *s) { delete s; }
int64_t farm_hostset_size(const farm_hostset *s) { return npoly(s); }

int farm_hostset_ring(const farm_hostset *s, int64_t i, const float **xy, int64_t *nv) {
    if (i < 0 || i >= npoly(s)) return FARM_EINVAL;
    *xy = s->xy.data() + 2 * s->ring_off[i];
    *nv = (int64_t)(s->ring_off[i + 1] - s->ring_off[i]);
    return FARM_OK;
}

int farm_hostset_mbbs(const farm_hostset *s, double *out) {
    memcpy(out, s->mbb.data(), s->mbb.size() * sizeof(double));
    return FARM_OK;
}

int farm_hostset_closed_rings(const farm_hostset *s, uint64_t *vert_off, float *x, float *y, int nthreads) {
    if (!s || !vert_off) return FARM_EINVAL;
    const int64_t n = npoly(s);
    uint64_t t = 0;
    for (int64_t i = 0; i < n; i++) {
        vert_off[i] = t;
        t += s->ring_off[i + 1] - s->ring_off[i] + 1;
    }
    vert_off[n] = t;
    if (!x || !y) return FARM_OK;
    const int nth = std::max(1, nthreads);
    auto work = [&](int k) {
        for (int64_t i = k; i < n; i += nth) {
            const uint64_t r0 = s->ring_off[i], r1 = s->ring_off[i + 1], o = vert_off[i];
            for (uint64_t v = r0; v < r1; v++) {
                x[o + v - r0] = s->xy[2 * v];
                y[o + v - r0] = s->xy[2 * v + 1];
            }
            x[o + r1 - r0] = s->xy[2 * r0];  // the decoder's repeated first vertex
            y[o + r1 - r0] = s->xy[2 * r0 + 1];
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nth; k++) th.emplace_back(work, k);
    work(0);
    for (auto &q : th) q.join();
    return FARM_OK;
}

namespace {

int rasterize_all(farm_hostset *s, int vpr, bool hier, int nlow, int nthreads) {
    if (vpr <= 0 || (hier && nlow < 1)) return FARM_EINVAL;
    const int64_t n = npoly(s);
    if (nthreads < 1) nthreads = 1;
    std::vector<farm::RasterOut> outs(n);
    std::vector<std::vector<farm::LayerOut>> lays(hier ? n : 0);
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        farm::RasterScratch *sc = farm::raster_scratch_new();
        std::vector<float> ring;
        for (;;) {
            int64_t b = next.fetch_add(8);
            if (b >= n) break;
            for (int64_t i = b; i < std::min(n, b + 8); i++) {
                const uint64_t r0 = s->ring_off[i], r1 = s->ring_off[i + 1];
                ring.assign(s->xy.begin() + 2 * r0, s->xy.begin() + 2 * r1);
                ring.push_back(ring[0]);  // VertexSequence::decode appends p[0]
                ring.push_back(ring[1]);
                if (hier)
                    farm::build_raster_hier(ring.data(), (int)(r1 - r0 + 1), &s->mbb[4 * i], vpr, nlow, outs[i],
                                            lays[i], sc);
                else
                    farm::build_raster(ring.data(), (int)(r1 - r0 + 1), &s->mbb[4 * i], vpr, outs[i], sc);
            }
        }
        farm::raster_scratch_free(sc);
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();

    // pack: prefix sums, then parallel copy
    s->hdr.assign(n, farm_raster_hdr());
    uint64_t np = 0, npo = 0, nsq = 0, nvt = 0, nvl = 0, nvn = 0, nly = 0, nls = 0;
    s->raster_fail = -1;
    s->layers.clear();
    s->raster_ok.assign(n, 1);
    for (int64_t i = 0; i < n; i++) {
        const farm::RasterOut &o = outs[i];
        if (!o.ok) {
            s->raster_ok[i] = 0;
            if (s->raster_fail < 0) s->raster_fail = i;
        }
        farm_raster_hdr &h = s->hdr[i];
        memset(&h, 0, sizeof h);
        h.mbr[0] = o.g.low[0];
        h.mbr[1] = o.g.low[1];
        h.mbr[2] = o.g.high[0];
        h.mbr[3] = o.g.high[1];
        h.step[0] = o.g.step[0];
        h.step[1] = o.g.step[1];
        h.dimx = o.g.dimx;
        h.dimy = o.g.dimy;
        h.pix_start = np;
        h.pixoff_start = npo;
        h.seq_start = nsq;
        h.vert_start = nvt;
        h.nverts = (uint32_t)(s->ring_off[i + 1] - s->ring_off[i] + 1);
        h.nseqs = (uint32_t)(o.seqs.size() / 2);
        h.vline_start = nvl;
        h.vnode_start = nvn;
        h.flags = o.ok ? 0 : FARM_HDR_REJECTED;
        if (hier) {
            const auto &L = lays[i];
            h.layer_start = nly;
            h.nlayers = (int32_t)L.size();
            for (size_t k = 0; k + 1 < L.size(); k++) {
                farm_layer l;
                l.low[0] = L[k].g.low[0];
                l.low[1] = L[k].g.low[1];
                l.step = L[k].g.step[0];
                l.dimx = L[k].g.dimx;
                l.dimy = L[k].g.dimy;
                l.status_start = nls;
                s->layers.push_back(l);
                nls += L[k].status.size();
                nly++;
            }
        }
        np += o.status.size();
        npo += o.pixoff.size();
        nsq += o.seqs.size() / 2;
        nvt += h.nverts;
        nvl += o.voff.size();
        nvn += o.vnodes.size();
    }
    s->status.assign(np, 0);
    s->areas.assign(np, 0.0);
    s->pixoff.assign(npo, 0);
    s->seqs.assign(2 * nsq, 0);
    s->vx.assign(nvt, 0.f);
    s->vy.assign(nvt, 0.f);
    s->voff.assign(nvl, 0);
    s->vnodes.assign(nvn, 0.0);
    s->lstatus.assign(nls, 0);
    next = 0;
    auto pack = [&]() {
        for (;;) {
            int64_t b = next.fetch_add(64);
            if (b >= n) break;
            for (int64_t i = b; i < std::min(n, b + 64); i++) {
                farm::RasterOut &o = outs[i];
                const farm_raster_hdr &h = s->hdr[i];
                std::copy(o.status.begin(), o.status.end(), s->status.begin() + h.pix_start);
                std::copy(o.areas.begin(), o.areas.end(), s->areas.begin() + h.pix_start);
                std::copy(o.pixoff.begin(), o.pixoff.end(), s->pixoff.begin() + h.pixoff_start);
                std::copy(o.seqs.begin(), o.seqs.end(), s->seqs.begin() + 2 * h.seq_start);
                std::copy(o.voff.begin(), o.voff.end(), s->voff.begin() + h.vline_start);
                std::copy(o.vnodes.begin(), o.vnodes.end(), s->vnodes.begin() + h.vnode_start);
                const uint64_t r0 = s->ring_off[i];
                const uint32_t nvs = h.nverts - 1;
                for (uint32_t k = 0; k < nvs; k++) {
                    s->vx[h.vert_start + k] = s->xy[2 * (r0 + k)];
                    s->vy[h.vert_start + k] = s->xy[2 * (r0 + k) + 1];
                }
                s->vx[h.vert_start + nvs] = s->xy[2 * r0];
                s->vy[h.vert_start + nvs] = s->xy[2 * r0 + 1];
                if (hier) {
                    for (int k = 0; k + 1 < h.nlayers; k++) {
                        const auto &st = lays[i][k].status;
                        std::copy(st.begin(), st.end(),
                                  s->lstatus.begin() + s->layers[h.layer_start + k].status_start);
                    }
                    std::vector<farm::LayerOut>().swap(lays[i]);
                }
                farm::RasterOut().status.swap(o.status);  // release early
                std::vector<double>().swap(o.areas);
                std::vector<uint32_t>().swap(o.seqs);
            }
        }
    };
    th.clear();
    for (int t = 0; t < nthreads; t++) th.emplace_back(pack);
    for (auto &t : th) t.join();
    s->rastered = true;
    s->hier = hier;
    return FARM_OK;
}

}  // namespace

int farm_hostset_rasterize(farm_hostset *s, int vpr, int nthreads) { return rasterize_all(s, vpr, false, 1, nthreads); }

int farm_hostset_rasterize_h(farm_hostset *s, int vpr, int nlow, int nthreads) {
    return rasterize_all(s, vpr, true, nlow, nthreads);
}

int farm_hostset_view(const farm_hostset *s, farm_polyset *o) {
    if (!s->rastered) return FARM_ESTATE;
    o->npolys = npoly(s);
    o->hdr = s->hdr.data();
    o->status = s->status.data();
    o->nstatus = s->status.size();
    o->pixoff = s->pixoff.data();
    o->npixoff = s->pixoff.size();
    o->seqs = s->seqs.data();
    o->nseqs = s->seqs.size() / 2;
    o->x = s->vx.data();
    o->y = s->vy.data();
    o->nverts = s->vx.size();
    o->voff = s->voff.data();
    o->nvoff = s->voff.size();
    o->vnodes = s->vnodes.data();
    o->nvnodes = s->vnodes.size();
    // layers and per-pixel areas are uploaded for within queries only
    o->layers = s->hier ? s->layers.data() : nullptr;
    o->nlayers = s->hier ? s->layers.size() : 0;
    o->lstatus = s->hier ? s->lstatus.data() : nullptr;
    o->nlstatus = s->hier ? s->lstatus.size() : 0;
    o->areas = s->hier ? s->areas.data() : nullptr;
    o->fill = s->fill.empty() ? nullptr : s->fill.data();
    o->category_count = s->cc;
    o->reserved = 0;
    return FARM_OK;
}

int farm_hostset_set_categories(farm_hostset *s, int category_bits) {
    if (!s->rastered) return FARM_ESTATE;
    if (category_bits < 2 || category_bits > 8) return FARM_EINVAL;
    const int cc = 1 << category_bits;
    s->cc = cc;
    s->fill.clear();
    if (cc == FARM_CATEGORY_COUNT) return FARM_OK;
    s->fill.assign(s->status.size(), 0);
    for (const farm_raster_hdr &h : s->hdr) {
        const double pix_area = h.step[0] * h.step[1];
        const uint64_t np = (uint64_t)h.dimx * (uint64_t)h.dimy;
        for (uint64_t p = 0; p < np; p++) {
            const uint8_t st = s->status[h.pix_start + p];
            uint8_t f = 0;
            if (st == FARM_CATEGORY_COUNT - 1) {
                f = (uint8_t)(cc - 1);
            } else if (st != 0) {
                // encodePixelArea (Ideal.cpp:64-88) with cc categories, then the
                // BORDER clamp of calculate_fullness (Ideal.cpp:905-912)
                const double ratio = s->areas[h.pix_start + p] / pix_area;
                int c;
                if (std::abs(ratio - 1.0) < 1e-9) c = cc - 1;
                else if (std::abs(ratio) < 1e-9) c = 0;
                else {
                    c = (int)std::ceil(ratio * (cc - 2));
                    if (c >= cc) c = cc - 1;
                }
                f = (uint8_t)std::max(1, std::min(c, cc - 2));
            }
            s->fill[h.pix_start + p] = f;
        }
    }
    return FARM_OK;
}

int farm_hostset_layer(const farm_hostset *s, int64_t i, int k, farm_layer *out, const uint8_t **status) {
    if (!s->rastered || !s->hier || i < 0 || i >= npoly(s)) return FARM_EINVAL;
    const farm_raster_hdr &h = s->hdr[i];
    if (k < 0 || k >= h.nlayers) return FARM_EINVAL;
    if (k == h.nlayers - 1) {
        out->low[0] = h.mbr[0];
        out->low[1] = h.mbr[1];
        out->step = h.step[0];
        out->dimx = h.dimx;
        out->dimy = h.dimy;
        out->status_start = h.pix_start;
        *status = s->status.data() + h.pix_start;
    } else {
        *out = s->layers[h.layer_start + k];
        *status = s->lstatus.data() + out->status_start;
    }
    return FARM_OK;
}

int farm_hostset_areas(const farm_hostset *s, int64_t i, const double **areas) {
    if (!s->rastered || i < 0 || i >= npoly(s)) return FARM_EINVAL;
    *areas = s->areas.data() + s->hdr[i].pix_start;
    return FARM_OK;
}

int64_t farm_hostset_raster_failures(const farm_hostset *s) { return s->raster_fail; }

int farm_hostset_raster_ok(const farm_hostset *s, uint8_t *out) {
    if (!s->rastered) return FARM_ESTATE;
    memcpy(out, s->raster_ok.data(), s->raster_ok.size());
    return FARM_OK;
}

void farm_free(void *p) { free(p); }

}  // extern "C"

// ------------------------------------------------------------ R-tree ----
// Sort-Tile-Recursive bulk-loaded R-tree over double boxes.  Any correct
// spatial index yields the same candidate set: every pair whose MBBs overlap
// inclusively (RTree::Overlap, RTree.h:1567-1580).
namespace {

struct RNode {
    double b[4];
    int32_t first, count;  // children (internal) or entries (leaf)
    int32_t leaf;
};

struct RTree {
    static const int M = 16;
    std::vector<RNode> nodes;
    std::vector<int32_t> items;
    int root = -1;

    void build(const double *boxes, int64_t n) {
        nodes.clear();
        items.resize(n);
        for (int64_t i = 0; i < n; i++) items[i] = (int32_t)i;
        if (n == 0) return;
        auto cx = [&](int32_t i) { return boxes[4 * i] + boxes[4 * i + 2]; };
        auto cy = [&](int32_t i) { return boxes[4 * i + 1] + boxes[4 * i + 3]; };
        const int64_t nleaf = (n + M - 1) / M;
        const int64_t slices = (int64_t)std::ceil(std::sqrt((double)nleaf));
        std::sort(items.begin(), items.end(), [&](int32_t a, int32_t b) { return cx(a) < cx(b); });
        const int64_t per_slice = slices * M;
        for (int64_t s0 = 0; s0 < n; s0 += per_slice) {
            int64_t s1 = std::min(n, s0 + per_slice);
            std::sort(items.begin() + s0, items.begin() + s1, [&](int32_t a, int32_t b) { return cy(a) < cy(b); });
        }
        std::vector<int32_t> level;
        for (int64_t s0 = 0; s0 < n; s0 += M) {
            RNode nd;
            nd.first = (int32_t)s0;
            nd.count = (int32_t)std::min<int64_t>(M, n - s0);
            nd.leaf = 1;
            nd.b[0] = nd.b[1] = 1e300;
            nd.b[2] = nd.b[3] = -1e300;
            for (int k = 0; k < nd.count; k++) {
                const double *bb = boxes + 4 * items[s0 + k];
                nd.b[0] = std::min(nd.b[0], bb[0]);
                nd.b[1] = std::min(nd.b[1], bb[1]);
                nd.b[2] = std::max(nd.b[2], bb[2]);
                nd.b[3] = std::max(nd.b[3], bb[3]);
            }
            level.push_back((int32_t)nodes.size());
            nodes.push_back(nd);
        }
        while (level.size() > 1) {
            std::vector<int32_t> up;
            // children of one parent must be contiguous in `nodes`
            std::vector<RNode> moved;
            for (size_t s0 = 0; s0 < level.size(); s0 += M) {
                RNode nd;
                nd.leaf = 0;
                nd.count = (int32_t)std::min<size_t>(M, level.size() - s0);
                nd.first = level[s0];
                nd.b[0] = nd.b[1] = 1e300;
                nd.b[2] = nd.b[3] = -1e300;
                for (int k = 0; k < nd.count; k++) {
                    const RNode &c = nodes[level[s0 + k]];
                    nd.b[0] = std::min(nd.b[0], c.b[0]);
                    nd.b[1] = std::min(nd.b[1], c.b[1]);
                    nd.b[2] = std::max(nd.b[2], c.b[2]);
                    nd.b[3] = std::max(nd.b[3], c.b[3]);
                }
                moved.push_back(nd);
            }
            for (auto &nd : moved) {
                up.push_back((int32_t)nodes.size());
                nodes.push_back(nd);
            }
            level.swap(up);
        }
        root = level[0];
    }

    static bool overlap(const double *a, const double *b) {
        return !(a[0] > b[2] || b[0] > a[2] || a[1] > b[3] || b[1] > a[3]);
    }

    template <class F>
    void search(const double *q, const double *boxes, F &&emit) const {
        if (root < 0) return;
        int32_t stack[128];
        int sp = 0;
        stack[sp++] = root;
        while (sp) {
            const RNode &nd = nodes[stack[--sp]];
            if (!overlap(nd.b, q)) continue;
            if (nd.leaf) {
                for (int k = 0; k < nd.count; k++) {
                    int32_t it = items[nd.first + k];
                    if (overlap(boxes + 4 * it, q)) emit(it);
                }
            } else {
                for (int k = 0; k < nd.count; k++) stack[sp++] = nd.first + k;
            }
        }
    }
};

const double kDegPerKmLat = 360.0 / 40076.0;

// degree_per_kilometer_longitude (util.h:73-96): the 900-entry table equals
// 360 / (sin((90 - k/10) * PI / 180) * 40076), PI = 3.14159265, printed with
// six decimals; it is rebuilt from that formula.
struct Dpk {
    double v[900];
    Dpk() {
        const double PI = 3.14159265;
        char buf[64];
        for (int k = 0; k < 900; k++) {
            double f = 360.0 / (std::sin((90 - k / 10.0) * PI / 180) * 40076);
            snprintf(buf, sizeof buf, "%.6f", f);
            v[k] = strtod(buf, nullptr);
        }
    }
    double lon(double lat) const {
        double a = std::abs(lat);
        if (a == 90) a = 89.9;
        return v[(int)(a * 10)];
    }
};
const Dpk kDpk;

}  // namespace

// Rings of A n B from the IN pieces of each pair (the chaining of the
// reference's rebuild_polygons, query_intersection.cu:393-502): start at an
// unused piece, follow a -> ring vertices -> b, continue with the unused
// piece that starts at b, until the chain returns to its start.
extern "C" char *farm_pieces_wkt(const farm_hostset *s, const uint32_t *pa, const uint32_t *pb, uint64_t npairs,
                                 const farm_piece *pieces, uint64_t npieces, uint64_t *len) {
    if (!s || !s->rastered || (npieces && !pieces) || (npairs && (!pa || !pb))) return nullptr;
    std::vector<uint64_t> order(npieces);
    for (uint64_t i = 0; i < npieces; i++) {
        if (pieces[i].pair >= npairs || pieces[i].ring < 0 || pieces[i].ring > 1) return nullptr;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint64_t x, uint64_t y) { return pieces[x].pair < pieces[y].pair; });
    auto key = [](float x, float y) {
        uint32_t ux, uy;
        memcpy(&ux, &x, 4);
        memcpy(&uy, &y, 4);
        return ((uint64_t)ux << 32) | uy;
    };
    std::string out;
    char num[64];
    auto put = [&](float x, float y, bool first) {
        snprintf(num, sizeof num, "%s%.9g %.9g", first ? "" : ", ", (double)x, (double)y);
        out += num;
    };
    for (uint64_t lo = 0; lo < npieces;) {
        uint64_t hi = lo;
        const uint32_t k = pieces[order[lo]].pair;
        while (hi < npieces && pieces[order[hi]].pair == k) hi++;
        const uint32_t ids[2] = {pa[k], pb[k]};
        if ((int64_t)ids[0] >= npoly(s) || (int64_t)ids[1] >= npoly(s)) return nullptr;
        snprintf(num, sizeof num, "%u\tMULTIPOLYGON (", k);
        out += num;
        bool first_poly = true;
        // pieces follow their ring's direction; when the two rings turn
        // opposite ways the target's pieces are walked backwards, so that
        // every piece ends where the next one starts
        auto orient = [&](uint32_t id) {
            const farm_raster_hdr &h = s->hdr[id];
            double a = 0;
            const double ox = s->vx[h.vert_start], oy = s->vy[h.vert_start];
            for (uint32_t v = 0; v + 1 < h.nverts; v++)
                a += ((double)s->vx[h.vert_start + v] - ox) * ((double)s->vy[h.vert_start + v + 1] - oy) -
                     ((double)s->vx[h.vert_start + v + 1] - ox) * ((double)s->vy[h.vert_start + v] - oy);
            return a < 0 ? -1 : 1;
        };
        const bool flip_b = orient(ids[0]) != orient(ids[1]);
        auto rev = [&](const farm_piece &p) { return flip_b && p.ring == 1; };
        std::unordered_map<uint64_t, std::vector<uint64_t>> starts;
        std::vector<char> used(hi - lo, 0);
        for (uint64_t i = lo; i < hi; i++) {
            const farm_piece &p = pieces[order[i]];
            if (p.nvert >= 0) starts[rev(p) ? key(p.bx, p.by) : key(p.ax, p.ay)].push_back(i - lo);
        }
        for (uint64_t i = lo; i < hi; i++) {
            if (used[i - lo]) continue;
            const farm_piece &p0 = pieces[order[i]];
            const farm_raster_hdr &h0 = s->hdr[ids[p0.ring]];
            out += first_poly ? "((" : ", ((";
            first_poly = false;
            if (p0.nvert < 0) {
                // a whole ring (containment)
                for (uint32_t v = 0; v < h0.nverts; v++) put(s->vx[h0.vert_start + v], s->vy[h0.vert_start + v], v == 0);
                used[i - lo] = 1;
                out += "))";
                continue;
            }
            bool first_pt = true;
            uint64_t cur = i - lo;
            const float sx0 = rev(p0) ? p0.bx : p0.ax, sy0 = rev(p0) ? p0.by : p0.ay;
            for (;;) {
                used[cur] = 1;
                const farm_piece &p = pieces[order[lo + cur]];
                const farm_raster_hdr &h = s->hdr[ids[p.ring]];
                const int m = (int)h.nverts - 1;
                const bool r = rev(p);
                put(r ? p.bx : p.ax, r ? p.by : p.ay, first_pt);
                first_pt = false;
                for (int t = 0; t < p.nvert; t++) {
                    const int v = (p.first + (r ? p.nvert - 1 - t : t)) % m;
                    put(s->vx[h.vert_start + v], s->vy[h.vert_start + v], false);
                }
                // the next piece starts where this one ends
                const float ex = r ? p.ax : p.bx, ey = r ? p.ay : p.by;
                uint64_t next = UINT64_MAX;
                auto it = starts.find(key(ex, ey));
                if (it != starts.end())
                    for (uint64_t c : it->second)
                        if (!used[c]) {
                            next = c;
                            break;
                        }
                if (next == UINT64_MAX && ex == sx0 && ey == sy0) {
                    put(ex, ey, false);  // back at the ring's start
                    break;
                }
                if (next == UINT64_MAX) {
                    // Two crossing points a few ulps apart (one crossing
                    // computed from two edge pairs) with a BORDER piece
                    // between them: bridge to the nearest unused start
                    // within 64 ulps, else close the ring here.
                    const double tol = std::ldexp(std::max({1.0, (double)std::fabs(ex), (double)std::fabs(ey)}), -17);
                    double best = tol;
                    for (uint64_t c = 0; c < hi - lo; c++) {
                        if (used[c]) continue;
                        const farm_piece &q = pieces[order[lo + c]];
                        if (q.nvert < 0) continue;
                        const float qx = rev(q) ? q.bx : q.ax, qy = rev(q) ? q.by : q.ay;
                        const double dd = std::max(std::fabs((double)qx - ex), std::fabs((double)qy - ey));
                        if (dd <= best) {
                            best = dd;
                            next = c;
                        }
                    }
                    put(ex, ey, false);
                    if (next == UINT64_MAX) break;
                }
                cur = next;
            }
            put(sx0, sy0, false);  // close the ring at its start
            out += "))";
        }
        out += ")\n";
        lo = hi;
    }
    char *r = (char *)malloc(out.size() + 1);
    if (!r) return nullptr;
    memcpy(r, out.c_str(), out.size() + 1);
    if (len) *len = out.size();
    return r;
}

extern "C" int farm_km_table(double *out) {
    if (!out) return 1;
    for (int k = 0; k < 900; k++) out[k] = kDpk.v[k];
    return 0;
}

namespace farm {
const double *km_lon_table() { return kDpk.v; }

// box::expand(d, geography) (Box.cpp:200-214): the query box of a within
// join grows by d km, in longitude by the table entry of its low latitude
void expand_box_km(double q[4], double km) {
    const double sx = km * kDpk.lon(q[1]);
    const double sy = km * kDegPerKmLat;
    q[0] -= sx;
    q[2] += sx;
    q[1] -= sy;
    q[3] += sy;
}
}  // namespace farm

extern "C" int64_t farm_candidates(const farm_hostset *a, const farm_hostset *b, double expand_km, int skip_self,
                                   int nthreads, uint32_t **pa, uint32_t **pb) {
    const int64_t na = npoly(a), nb = npoly(b);
    RTree rt;
    rt.build(a->mbb.data(), na);
    if (nthreads < 1) nthreads = 1;
    const int64_t chunk = 256;
    const int64_t nchunks = (nb + chunk - 1) / chunk;
    std::vector<std::vector<uint32_t>> res_a(nchunks), res_b(nchunks);
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        std::vector<uint32_t> hits;
        for (;;) {
            int64_t c = next.fetch_add(1);
            if (c >= nchunks) break;
            for (int64_t j = c * chunk; j < std::min(nb, (c + 1) * chunk); j++) {
                double q[4];
                memcpy(q, &b->mbb[4 * j], sizeof q);
                if (expand_km > 0) farm::expand_box_km(q, expand_km);
                hits.clear();
                rt.search(q, a->mbb.data(), [&](int32_t i) {
                    if (skip_self && i == j) return;
                    hits.push_back((uint32_t)i);
                });
                std::sort(hits.begin(), hits.end());
                for (uint32_t i : hits) {
                    res_a[c].push_back(i);
                    res_b[c].push_back((uint32_t)j);
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    int64_t total = 0;
    for (auto &v : res_a) total += (int64_t)v.size();
    *pa = (uint32_t *)malloc(sizeof(uint32_t) * std::max<int64_t>(total, 1));
    *pb = (uint32_t *)malloc(sizeof(uint32_t) * std::max<int64_t>(total, 1));
    if (!*pa || !*pb) return -1;
    int64_t o = 0;
    for (int64_t c = 0; c < nchunks; c++) {
        std::copy(res_a[c].begin(), res_a[c].end(), *pa + o);
        std::copy(res_b[c].begin(), res_b[c].end(), *pb + o);
        o += (int64_t)res_a[c].size();
    }
    return total;
}
