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
// Object I/O: READ,<file> (raw planes / items, loaded before the first frame), WRITE,<file>
// (after every frame), COMPARE,<file> (after every frame, exact bytes).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <VX/vx.h>
#include <vx_ext_amd.h>

namespace {

struct IoItem {
    vx_reference ref;
    std::string name, op, file;
};

void VX_CALLBACK registry(void *obj, vx_reference ref, const char *name, const char *params)
{
    auto *list = (std::vector<IoItem> *)obj;
    std::string p = params;
    const size_t c = p.find(',');
    IoItem it{ref, name, c == std::string::npos ? p : p.substr(0, c), c == std::string::npos ? "" : p.substr(c + 1)};
    for (char &ch : it.op) ch = (char)toupper((unsigned char)ch);
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
    for (IoItem &it : io) {
        if (rc || it.op != "READ") continue;
        std::vector<uint8_t> data;
        if (!read_file(it.file, data) || object_bytes(it.ref, data, false)) {
            fprintf(stderr, "runvx: READ %s from %s failed\n", it.name.c_str(), it.file.c_str());
            rc = 1;
        }
    }
    double total = 0, best = 1e30;
    for (int f = 0; f < frames && !rc; f++) {
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
                std::vector<uint8_t> mine, ref;
                if (object_bytes(it.ref, mine, true) || !read_file(it.file, ref)) {
                    fprintf(stderr, "runvx: COMPARE %s with %s failed\n", it.name.c_str(), it.file.c_str());
                    rc = 1;
                    continue;
                }
                size_t diff = mine.size() == ref.size() ? 0 : std::max(mine.size(), ref.size());
                for (size_t k = 0; k < std::min(mine.size(), ref.size()); k++) diff += mine[k] != ref[k];
                if (diff) {
                    printf("runvx: COMPARE %s: %zu bytes differ from %s\n", it.name.c_str(), diff, it.file.c_str());
                    if (!discardCompare) rc = 2;
                } else if (verbose) {
                    printf("runvx: COMPARE %s: match\n", it.name.c_str());
                }
            }
        }
    }
    if (!rc) printf("runvx: %d frame(s): avg %.3f ms, min %.3f ms per frame\n", frames, total / frames, best);
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
