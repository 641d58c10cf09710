// runvx.cpp -- command-line runner for GDF graphs on the MI355X engine (the reference's
// utilities/runvx: options runvx.cpp:63-174, GDF README.md:14-200), for the vision-node path.
//
//   runvx [options] <file.gdf> [argument(s)]
//   runvx [options] node <kernelName> [argument(s)]
//
// Arguments are object descriptions ("image:1920,1080,U008:READ,in.raw"), reachable from the GDF as
// $1, $2, ...  Options:
//   -frames:<n>               run the graph n times (default 1)
//   -affinity:CPU|GPU[<n>]    context affinity (this engine runs every kernel on the GPU; CPU
//                             requests fall back with a warning)
//   -dump-profile             print per-graph and per-node timing
//   -discard-compare-errors   keep running after a compare mismatch
//   -v                        echo the GDF and the schedule
// Object I/O (utilities/runvx/vxImage.cpp:560-640, 1185-1265):
//   READ,<file>       raw planes / items; a file holding several frames is read frame after frame
//                     (a single-frame file feeds every frame)
//   WRITE,<file>      after every frame
//   COMPARE,<file>[,rect{<sx>;<sy>;<ex>;<ey>}][,err{<min>;<max>}][,checksum|checksum-save-instead-of-test]
//                     after every frame: raw frames read one after another from the file (a file of
//                     exactly one frame serves every frame) (exact, or
//                     pixel differences img - ref within [min, max] in the image's pixel type), or MD5
//                     checksums of the region's rows (one 32-hex-digit line per frame; the -save- form
//                     writes them instead).  A '%' in the file name makes it a per-frame printf pattern
//                     of (frame, width, height).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <VX/vx.h>
#include <vx_ext_amd.h>

#include "md5.hpp"

namespace {

struct CompareSpec {
    bool rect = false;
    vx_rectangle_t region{0, 0, 0, 0};
    float errMin = 0.f, errMax = 0.f;
    bool checksum = false, checksumSave = false;
};

struct IoItem {
    vx_reference ref;
    std::string name, op, file;
    CompareSpec cmp;
    FILE *fp = nullptr;          // COMPARE / READ stream kept open across frames
    std::vector<uint8_t> once;   // READ: a single-frame file
};

void VX_CALLBACK registry(void *obj, vx_reference ref, const char *name, const char *params)
{
    auto *list = (std::vector<IoItem> *)obj;
    std::string p = params;
    const size_t c = p.find(',');
    IoItem it;
    it.ref = ref;
    it.name = name;
    it.op = c == std::string::npos ? p : p.substr(0, c);
    it.file = c == std::string::npos ? "" : p.substr(c + 1);
    for (char &ch : it.op) ch = (char)toupper((unsigned char)ch);
    if (it.op == "COMPARE") {
        // options after the file name: ,rect{..} ,err{..} ,checksum ,checksum-save-instead-of-test
        size_t k = it.file.find(',');
        std::string opts = k == std::string::npos ? "" : it.file.substr(k);
        if (k != std::string::npos) it.file.resize(k);
        while (!opts.empty() && opts[0] == ',') {
            size_t e = opts.find(',', 1);
            std::string o = opts.substr(1, e == std::string::npos ? std::string::npos : e - 1);
            opts = e == std::string::npos ? "" : opts.substr(e);
            std::string lo = o;
            for (char &ch : lo) ch = (char)tolower((unsigned char)ch);
            if (!lo.compare(0, 4, "rect")) {
                int v[4] = {0, 0, 0, 0};
                if (sscanf(o.c_str() + 4, "{%d;%d;%d;%d}", &v[0], &v[1], &v[2], &v[3]) == 4) {
                    it.cmp.rect = true;
                    it.cmp.region = {(vx_uint32)v[0], (vx_uint32)v[1], (vx_uint32)v[2], (vx_uint32)v[3]};
                }
            } else if (!lo.compare(0, 3, "err")) {
                sscanf(o.c_str() + 3, "{%f;%f}", &it.cmp.errMin, &it.cmp.errMax);
            } else if (lo == "checksum") {
                it.cmp.checksum = true;
            } else if (lo == "checksum-save-instead-of-test") {
                it.cmp.checksumSave = true;
            } else {
                fprintf(stderr, "runvx: invalid image compare option: %s\n", o.c_str());
            }
        }
    }
    list->push_back(it);
}

bool read_file(const std::string &fn, std::vector<uint8_t> &data)
{
    FILE *fp = fopen(fn.c_str(), "rb");
    if (!fp) return false;
    fseek(fp, 0, SEEK_END);
    data.resize((size_t)ftell(fp));
    fseek(fp, 0, SEEK_SET);
    const bool ok = fread(data.data(), 1, data.size(), fp) == data.size();
    fclose(fp);
    return ok;
}

// the object's bytes in file order: image planes row by row without padding, array items, scalar value
vx_status object_bytes(vx_reference ref, std::vector<uint8_t> &out, bool read)
{
    vx_enum type = 0;
    vxQueryReference(ref, VX_REFERENCE_TYPE, &type, sizeof(type));
    if (type == VX_TYPE_IMAGE) {
        vx_image img = (vx_image)ref;
        vx_size planes = 0;
        vx_uint32 w = 0, h = 0;
        vxQueryImage(img, VX_IMAGE_PLANES, &planes, sizeof(planes));
        vxQueryImage(img, VX_IMAGE_WIDTH, &w, sizeof(w));
        vxQueryImage(img, VX_IMAGE_HEIGHT, &h, sizeof(h));
        size_t off = 0;
        std::vector<uint8_t> result;
        for (vx_uint32 p = 0; p < planes; p++) {
            vx_rectangle_t rect{0, 0, w, h};
            vx_map_id id;
            vx_imagepatch_addressing_t addr;
            void *ptr = nullptr;
            vx_status s = vxMapImagePatch(img, &rect, p, &id, &addr, &ptr, read ? VX_READ_ONLY : VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST, VX_NOGAP_X);
            if (s) return s;
            const size_t bytes = (size_t)addr.dim_x * addr.stride_x * addr.dim_y;
            if (read) result.insert(result.end(), (uint8_t *)ptr, (uint8_t *)ptr + bytes);
            else {
                if (off + bytes > out.size()) { vxUnmapImagePatch(img, id); return VX_ERROR_INVALID_DIMENSION; }
                memcpy(ptr, out.data() + off, bytes);
            }
            off += bytes;
            vxUnmapImagePatch(img, id);
        }
        if (read) out = result;
        return VX_SUCCESS;
    }
    if (type == VX_TYPE_ARRAY) {
        vx_array arr = (vx_array)ref;
        vx_size n = 0, isz = 0, cap = 0;
        vxQueryArray(arr, VX_ARRAY_NUMITEMS, &n, sizeof(n));
        vxQueryArray(arr, VX_ARRAY_ITEMSIZE, &isz, sizeof(isz));
        vxQueryArray(arr, VX_ARRAY_CAPACITY, &cap, sizeof(cap));
        if (read) {
            out.assign(n * isz, 0);
            return n ? vxCopyArrayRange(arr, 0, n, isz, out.data(), VX_READ_ONLY, VX_MEMORY_TYPE_HOST) : VX_SUCCESS;
        }
        vxTruncateArray(arr, 0);
        const vx_size items = std::min(out.size() / isz, cap);
        return items ? vxAddArrayItems(arr, items, out.data(), isz) : VX_SUCCESS;
    }
    if (type == VX_TYPE_SCALAR) {
        uint8_t v[8] = {0};
        if (read) {
            vx_status s = vxCopyScalar((vx_scalar)ref, v, VX_READ_ONLY, VX_MEMORY_TYPE_HOST);
            vx_enum t;
            vxQueryScalar((vx_scalar)ref, VX_SCALAR_TYPE, &t, sizeof(t));
            out.assign(v, v + (t == VX_TYPE_SIZE || t == VX_TYPE_UINT64 || t == VX_TYPE_INT64 || t == VX_TYPE_FLOAT64 ? 8 : 4));
            return s;
        }
        memcpy(v, out.data(), std::min<size_t>(out.size(), 8));
        return vxCopyScalar((vx_scalar)ref, v, VX_WRITE_ONLY, VX_MEMORY_TYPE_HOST);
    }
    return VX_ERROR_NOT_SUPPORTED;
}

struct PlaneGeom {
    uint32_t w, h, bpp;
};

// packed planes of an image (rows without padding) and each plane's geometry
vx_status image_planes(vx_image img, std::vector<uint8_t> &data, std::vector<PlaneGeom> &geom)
{
    vx_size planes = 0;
    vx_uint32 w = 0, h = 0;
    vxQueryImage(img, VX_IMAGE_PLANES, &planes, sizeof(planes));
    vxQueryImage(img, VX_IMAGE_WIDTH, &w, sizeof(w));
    vxQueryImage(img, VX_IMAGE_HEIGHT, &h, sizeof(h));
    data.clear();
    geom.clear();
    for (vx_uint32 p = 0; p < planes; p++) {
        vx_rectangle_t rect{0, 0, w, h};
        vx_map_id id;
        vx_imagepatch_addressing_t addr;
        void *ptr = nullptr;
        vx_status s = vxMapImagePatch(img, &rect, p, &id, &addr, &ptr, VX_READ_ONLY, VX_MEMORY_TYPE_HOST, VX_NOGAP_X);
        if (s) return s;
        const size_t bytes = (size_t)addr.dim_x * addr.stride_x * addr.dim_y;
        data.insert(data.end(), (uint8_t *)ptr, (uint8_t *)ptr + bytes);
        geom.push_back({addr.dim_x, addr.dim_y, (uint32_t)addr.stride_x});
        vxUnmapImagePatch(img, id);
    }
    return VX_SUCCESS;
}

std::string frame_file(const std::string &pattern, int frame, vx_reference ref)
{
    if (pattern.find('%') == std::string::npos) return pattern;
    vx_uint32 w = 0, h = 0;
    vx_enum type = 0;
    vxQueryReference(ref, VX_REFERENCE_TYPE, &type, sizeof(type));
    if (type == VX_TYPE_IMAGE) {
        vxQueryImage((vx_image)ref, VX_IMAGE_WIDTH, &w, sizeof(w));
        vxQueryImage((vx_image)ref, VX_IMAGE_HEIGHT, &h, sizeof(h));
    }
    char buf[4096];
    snprintf(buf, sizeof(buf), pattern.c_str(), frame, w, h);
    return buf;
}

// pixels of the region outside [errMin, errMax] (img - ref in the pixel type: S16 / U16 / S32 /
// U32 / F32 planes compare as such, everything else byte by byte)
size_t count_mismatches(vx_df_image fmt, const std::vector<uint8_t> &mine, const uint8_t *ref, const std::vector<PlaneGeom> &geom,
                        uint32_t w, uint32_t h, const CompareSpec &c)
{
    size_t bad = 0, off = 0;
    for (const PlaneGeom &g : geom) {
        const uint32_t sx = c.rect ? c.region.start_x * g.w / w : 0, sy = c.rect ? c.region.start_y * g.h / h : 0;
        const uint32_t ex = c.rect ? std::min(g.w, c.region.end_x * g.w / w) : g.w;
        const uint32_t ey = c.rect ? std::min(g.h, c.region.end_y * g.h / h) : g.h;
        for (uint32_t y = sy; y < ey; y++)
            for (uint32_t x = sx; x < ex; x++) {
                const size_t at = off + ((size_t)y * g.w + x) * g.bpp;
                double d;
                if (fmt == VX_DF_IMAGE_S16) d = (double)*(const int16_t *)&mine[at] - (double)*(const int16_t *)(ref + at);
                else if (fmt == VX_DF_IMAGE_U16) d = (double)*(const uint16_t *)&mine[at] - (double)*(const uint16_t *)(ref + at);
                else if (fmt == VX_DF_IMAGE_S32) d = (double)*(const int32_t *)&mine[at] - (double)*(const int32_t *)(ref + at);
                else if (fmt == VX_DF_IMAGE_U32) d = (double)*(const uint32_t *)&mine[at] - (double)*(const uint32_t *)(ref + at);
                else {
                    for (uint32_t b = 0; b < g.bpp; b++) {
                        const double e = (double)mine[at + b] - (double)ref[at + b];
                        if (e < c.errMin || e > c.errMax) { bad++; break; }
                    }
                    continue;
                }
                if (d < c.errMin || d > c.errMax) bad++;
            }
        off += (size_t)g.w * g.h * g.bpp;
    }
    return bad;
}

std::string region_checksum(const std::vector<uint8_t> &mine, const std::vector<PlaneGeom> &geom, uint32_t w, uint32_t h,
                            const CompareSpec &c)
{
    mvxe::Md5 md5;
    size_t off = 0;
    for (const PlaneGeom &g : geom) {
        const uint32_t sx = c.rect ? c.region.start_x * g.w / w : 0, sy = c.rect ? c.region.start_y * g.h / h : 0;
        const uint32_t ex = c.rect ? std::min(g.w, c.region.end_x * g.w / w) : g.w;
        const uint32_t ey = c.rect ? std::min(g.h, c.region.end_y * g.h / h) : g.h;
        for (uint32_t y = sy; y < ey; y++) md5.update(&mine[off + ((size_t)y * g.w + sx) * g.bpp], (size_t)(ex - sx) * g.bpp);
        off += (size_t)g.w * g.h * g.bpp;
    }
    return md5.hex();
}

// one frame's COMPARE of an item; returns -1 on an I/O failure, else the number of mismatches
long compare_frame(IoItem &it, int frame, bool verbose)
{
    std::vector<uint8_t> mine;
    vx_enum type = 0;
    vxQueryReference(it.ref, VX_REFERENCE_TYPE, &type, sizeof(type));
    const bool perFrame = it.file.find('%') != std::string::npos;
    const std::string fn = frame_file(it.file, frame, it.ref);
    if (!it.fp && it.once.empty()) {
        it.fp = fopen(fn.c_str(), it.cmp.checksumSave ? "w" : (it.cmp.checksum ? "r" : "rb"));
        if (!it.fp) return -1;
    }
    // raw reference data: a file holding exactly one frame serves every frame
    auto next_ref = [&](size_t bytes, std::vector<uint8_t> &ref) -> bool {
        if (!it.once.empty()) { ref = it.once; return ref.size() == bytes; }
        if (!perFrame && frame == 0) {
            fseek(it.fp, 0, SEEK_END);
            const long total = ftell(it.fp);
            fseek(it.fp, 0, SEEK_SET);
            if (total == (long)bytes) {
                it.once.resize(bytes);
                const bool ok = fread(it.once.data(), 1, bytes, it.fp) == bytes;
                fclose(it.fp);
                it.fp = nullptr;
                ref = it.once;
                return ok;
            }
        }
        ref.resize(bytes);
        return fread(ref.data(), 1, bytes, it.fp) == bytes;
    };
    long bad = 0;
    if (type != VX_TYPE_IMAGE) {
        // arrays and scalars: exact bytes, successive records of the same size
        if (object_bytes(it.ref, mine, true)) return -1;
        std::vector<uint8_t> ref;
        if (!next_ref(mine.size(), ref)) return -1;
        for (size_t k = 0; k < mine.size(); k++) bad += mine[k] != ref[k];
        if (bad) printf("runvx: COMPARE %s: %ld byte(s) differ from frame %d of %s\n", it.name.c_str(), bad, frame, fn.c_str());
    } else {
        vx_image img = (vx_image)it.ref;
        std::vector<PlaneGeom> geom;
        vx_uint32 w = 0, h = 0;
        vx_df_image fmt = 0;
        vxQueryImage(img, VX_IMAGE_WIDTH, &w, sizeof(w));
        vxQueryImage(img, VX_IMAGE_HEIGHT, &h, sizeof(h));
        vxQueryImage(img, VX_IMAGE_FORMAT, &fmt, sizeof(fmt));
        if (image_planes(img, mine, geom)) return -1;
        if (it.cmp.checksumSave) {
            fprintf(it.fp, "%s\n", region_checksum(mine, geom, w, h, it.cmp).c_str());
        } else if (it.cmp.checksum) {
            char want[64] = {0};
            if (fscanf(it.fp, "%63s", want) != 1) return -1;
            const std::string got = region_checksum(mine, geom, w, h, it.cmp);
            if (got != want) {
                printf("runvx: COMPARE %s: checksum %s instead of %s (frame %d of %s)\n", it.name.c_str(), got.c_str(), want, frame, fn.c_str());
                bad = 1;
            }
        } else {
            std::vector<uint8_t> ref;
            if (!next_ref(mine.size(), ref)) return -1;
            bad = (long)count_mismatches(fmt, mine, ref.data(), geom, w, h, it.cmp);
            if (bad) printf("runvx: COMPARE %s: %ld pixel(s) differ from frame %d of %s\n", it.name.c_str(), bad, frame, fn.c_str());
        }
    }
    if (!bad && verbose) printf("runvx: COMPARE %s: match (frame %d)\n", it.name.c_str(), frame);
    if (perFrame && it.fp) { fclose(it.fp); it.fp = nullptr; }
    return bad;
}

// one frame's READ of an item: the next frame of a multi-frame file, or the single frame again
bool read_frame(IoItem &it, int frame)
{
    if (!it.once.empty()) return object_bytes(it.ref, it.once, false) == VX_SUCCESS;
    const std::string fn = frame_file(it.file, frame, it.ref);
    std::vector<uint8_t> cur;
    if (object_bytes(it.ref, cur, true)) return false;   // the frame size
    const size_t frameBytes = cur.size();
    if (frame == 0 || fn != it.file) {
        std::vector<uint8_t> all;
        if (!read_file(fn, all)) return false;
        if (fn == it.file && all.size() == frameBytes) {
            it.once = all;
            return object_bytes(it.ref, it.once, false) == VX_SUCCESS;
        }
        if (fn != it.file) return object_bytes(it.ref, all, false) == VX_SUCCESS;
        it.fp = fopen(fn.c_str(), "rb");
        if (!it.fp) return false;
    }
    std::vector<uint8_t> data(frameBytes);
    if (!it.fp || fread(data.data(), 1, frameBytes, it.fp) != frameBytes) return false;
    return object_bytes(it.ref, data, false) == VX_SUCCESS;
}

int usage()
{
    fprintf(stderr, "usage: runvx [-frames:<n>] [-affinity:CPU|GPU[<n>]] [-dump-profile] [-discard-compare-errors] [-v] "
                    "<file.gdf>|node <kernel> [argument(s)]\n");
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    int frames = 1;
    bool profile = false, discardCompare = false, verbose = false;
    AgoTargetAffinityInfo aff{};
    int a = 1;
    for (; a < argc && argv[a][0] == '-'; a++) {
        if (!strncmp(argv[a], "-frames:", 8)) frames = atoi(argv[a] + 8);
        else if (!strncmp(argv[a], "-affinity:", 10)) {
            aff.device_type = !strncmp(argv[a] + 10, "CPU", 3) ? AGO_TARGET_AFFINITY_CPU : AGO_TARGET_AFFINITY_GPU;
            aff.device_info = (vx_uint32)atoi(argv[a] + 13);
        } else if (!strcmp(argv[a], "-dump-profile")) profile = true;
        else if (!strcmp(argv[a], "-discard-compare-errors")) discardCompare = true;
        else if (!strcmp(argv[a], "-v")) verbose = true;
        else if (!strcmp(argv[a], "-h")) return usage();
        else { fprintf(stderr, "runvx: unknown option %s\n", argv[a]); return usage(); }
    }
    if (a >= argc) return usage();
    std::string text;
    if (!strcmp(argv[a], "node")) {
        if (a + 1 >= argc) return usage();
        text = std::string("node ") + argv[a + 1];
        for (int k = a + 2; k < argc; k++) text += " $" + std::to_string(k - a - 1);
        text += "\n";
        a += 2;
    } else {
        std::vector<uint8_t> raw;
        if (!read_file(argv[a], raw)) { fprintf(stderr, "runvx: unable to read %s\n", argv[a]); return 1; }
        text.assign(raw.begin(), raw.end());
        a += 1;
    }
    // arguments become data objects $1..$n: declare them in front of the graph text
    std::string pre;
    for (int k = a; k < argc; k++) pre += "data __runvx_arg" + std::to_string(k - a + 1) + " = " + argv[k] + "\n";
    for (int k = argc - a; k >= 1; k--) {
        const std::string from = "$" + std::to_string(k), to = "__runvx_arg" + std::to_string(k);
        for (size_t p = 0; (p = text.find(from, p)) != std::string::npos;) {
            const size_t e = p + from.size();
            if (e < text.size() && isdigit((unsigned char)text[e])) { p = e; continue; }
            text.replace(p, from.size(), to);
            p += to.size();
        }
    }
    text = pre + text;

    vx_context ctx = vxCreateContext();
    if (aff.device_type) vxSetContextAttribute(ctx, VX_CONTEXT_ATTRIBUTE_AMD_AFFINITY, &aff, sizeof(aff));
    vx_graph graph = vxCreateGraph(ctx);
    std::vector<IoItem> io;
    AgoGraphImportInfo info{};
    info.text = (vx_char *)text.c_str();
    info.dumpToConsole = verbose ? 1 : 0;
    info.data_registry_callback_f = registry;
    info.data_registry_callback_obj = &io;
    int rc = 0;
    vx_status s = vxSetGraphAttribute(graph, VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT, &info, sizeof(info));
    if (s) { fprintf(stderr, "runvx: GDF import failed (%d)\n", s); rc = 1; }
    auto t0 = std::chrono::steady_clock::now();
    if (!rc && (s = vxVerifyGraph(graph))) { fprintf(stderr, "runvx: vxVerifyGraph failed (%d)\n", s); rc = 1; }
    auto t1 = std::chrono::steady_clock::now();
    if (!rc) printf("runvx: graph verified in %.3f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
    double total = 0, best = 1e30;
    int done = 0;
    for (int f = 0; f < frames && !rc; f++) {
        bool eos = false;
        for (IoItem &it : io) {
            if (it.op != "READ") continue;
            if (!read_frame(it, f)) {
                if (f == 0) { fprintf(stderr, "runvx: READ %s from %s failed\n", it.name.c_str(), it.file.c_str()); rc = 1; }
                eos = true;   // end of a multi-frame input: stop, as the reference does
            }
        }
        if (rc || eos) break;
        auto a0 = std::chrono::steady_clock::now();
        if ((s = vxProcessGraph(graph))) { fprintf(stderr, "runvx: vxProcessGraph failed (%d) at frame %d\n", s, f); rc = 1; break; }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a0).count();
        total += ms;
        best = std::min(best, ms);
        for (IoItem &it : io) {
            if (it.op == "WRITE") {
                std::vector<uint8_t> data;
                FILE *fp = nullptr;
                if (object_bytes(it.ref, data, true) || !(fp = fopen(it.file.c_str(), "wb")) ||
                    fwrite(data.data(), 1, data.size(), fp) != data.size()) {
                    fprintf(stderr, "runvx: WRITE %s to %s failed\n", it.name.c_str(), it.file.c_str());
                    rc = 1;
                }
                if (fp) fclose(fp);
            } else if (it.op == "COMPARE") {
                const long bad = compare_frame(it, f, verbose);
                if (bad < 0) {
                    fprintf(stderr, "runvx: COMPARE %s with %s failed\n", it.name.c_str(), it.file.c_str());
                    rc = 1;
                } else if (bad && !discardCompare) {
                    rc = 2;
                }
            }
        }
        done++;
    }
    for (IoItem &it : io)
        if (it.fp) fclose(it.fp);
    if (!rc) printf("runvx: %d frame(s): avg %.3f ms, min %.3f ms per frame\n", done, done ? total / done : 0.0, done ? best : 0.0);
    if (profile && !rc) {
        vx_perf_t perf;
        vxQueryGraph(graph, VX_GRAPH_PERFORMANCE, &perf, sizeof(perf));
        printf("runvx: profile: graph num %llu avg %.3f ms min %.3f ms max %.3f ms\n", (unsigned long long)perf.num, perf.avg * 1e-6,
               perf.min * 1e-6, perf.max * 1e-6);
    }
    vxReleaseGraph(&graph);
    vxReleaseContext(&ctx);
    return rc;
}
