// vx_gdf.cpp -- GDF (graph description format) text import: vxSetGraphAttribute(
// VX_GRAPH_ATTRIBUTE_AMD_IMPORT_FROM_TEXT).
//
// Grammar of the reference's reader (agoReadGraphFromStringInternal, ago_interface.cpp:456-1068) and
// runvx's object descriptions (utilities/runvx/README.md:14-200, vxParameter.cpp:212-330), for the
// commands and objects of the vision-node path:
//   data <name> = <object>:<config>[:<io>]      image, virtual-image / image-virtual, uniform-image /
//                                                image-uniform, pyramid, array, scalar, threshold,
//                                                matrix, remap (each also *-virtual / virtual-*)
//   node <kernel-name> <arg>... [attr:BORDER_MODE:UNDEFINED|REPLICATE|CONSTANT,<v>]
//                              [attr:AFFINITY:CPU|GPU[<n>][,<group>]]
//   arguments: data names, name[<level>] for pyramid levels, $<n> for AgoGraphImportInfo::ref[n-1],
//              null, !<ENUM> for an enum scalar.
//   '#' starts a comment, a trailing '\' continues a line.
// ":INIT,<values>" on scalars, thresholds and matrices is applied here; any other third field
// (":READ,<file>", ":WRITE,<file>", ":COMPARE,<file>") goes to data_registry_callback_f, which is how
// runvx attaches file I/O (ago_interface.cpp:616-632).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <sstream>

#include "engine.hpp"

using namespace mvxe;

namespace {

struct EnumName { const char *name; vx_enum value; };

// names accepted for enums and types in GDF text (the reference's table is ago_util.cpp:33-147)
const EnumName kNames[] = {
    {"CHANNEL_0", VX_CHANNEL_0}, {"CHANNEL_1", VX_CHANNEL_1}, {"CHANNEL_2", VX_CHANNEL_2}, {"CHANNEL_3", VX_CHANNEL_3},
    {"CHANNEL_R", VX_CHANNEL_R}, {"CHANNEL_G", VX_CHANNEL_G}, {"CHANNEL_B", VX_CHANNEL_B}, {"CHANNEL_A", VX_CHANNEL_A},
    {"CHANNEL_Y", VX_CHANNEL_Y}, {"CHANNEL_U", VX_CHANNEL_U}, {"CHANNEL_V", VX_CHANNEL_V},
    {"NEAREST_NEIGHBOR", VX_INTERPOLATION_NEAREST_NEIGHBOR}, {"BILINEAR", VX_INTERPOLATION_BILINEAR}, {"AREA", VX_INTERPOLATION_AREA},
    {"BINARY", VX_THRESHOLD_TYPE_BINARY}, {"RANGE", VX_THRESHOLD_TYPE_RANGE},
    {"NORM_L1", VX_NORM_L1}, {"NORM_L2", VX_NORM_L2},
    {"BORDER_MODE_UNDEFINED", VX_BORDER_UNDEFINED}, {"BORDER_MODE_REPLICATE", VX_BORDER_REPLICATE},
    {"BORDER_MODE_CONSTANT", VX_BORDER_CONSTANT},
    {"CRITERIA_ITERATIONS", VX_TERM_CRITERIA_ITERATIONS}, {"CRITERIA_EPSILON", VX_TERM_CRITERIA_EPSILON},
    {"CRITERIA_BOTH", VX_TERM_CRITERIA_BOTH},
    {"RECTANGLE", VX_TYPE_RECTANGLE}, {"KEYPOINT", VX_TYPE_KEYPOINT}, {"COORDINATES2D", VX_TYPE_COORDINATES2D},
    {"COORDINATES3D", VX_TYPE_COORDINATES3D}, {"COORDINATES2DF", VX_TYPE_COORDINATES2DF}, {"DF_IMAGE", VX_TYPE_DF_IMAGE},
    {"ENUM", VX_TYPE_ENUM}, {"UINT64", VX_TYPE_UINT64}, {"INT64", VX_TYPE_INT64}, {"UINT32", VX_TYPE_UINT32},
    {"INT32", VX_TYPE_INT32}, {"UINT16", VX_TYPE_UINT16}, {"INT16", VX_TYPE_INT16}, {"UINT8", VX_TYPE_UINT8},
    {"INT8", VX_TYPE_INT8}, {"CHAR", VX_TYPE_CHAR}, {"FLOAT32", VX_TYPE_FLOAT32}, {"FLOAT64", VX_TYPE_FLOAT64},
    {"SIZE", VX_TYPE_SIZE}, {"BOOL", VX_TYPE_BOOL},
};

std::string upper(std::string s)
{
    for (char &c : s) c = (char)toupper((unsigned char)c);
    return s;
}

std::vector<std::string> split(const std::string &s, char d)
{
    std::vector<std::string> out;
    std::string cur;
    int depth = 0;  // commas inside {...} do not split
    for (char c : s) {
        if (c == '{') depth++;
        if (c == '}') depth--;
        if (c == d && depth == 0) { out.push_back(cur); cur.clear(); }
        else cur += c;
    }
    out.push_back(cur);
    return out;
}

bool to_u32(const std::string &s, uint32_t &v)
{
    char *end = nullptr;
    unsigned long x = strtoul(s.c_str(), &end, 0);
    if (s.empty() || *end) return false;
    v = (uint32_t)x;
    return true;
}

}  // namespace

namespace mvxe {

const char *type_name(vx_enum type)
{
    switch (type) {
    case VX_TYPE_IMAGE: return "image";
    case VX_TYPE_SCALAR: return "scalar";
    case VX_TYPE_THRESHOLD: return "threshold";
    case VX_TYPE_ARRAY: return "array";
    case VX_TYPE_PYRAMID: return "pyramid";
    case VX_TYPE_MATRIX: return "matrix";
    case VX_TYPE_REMAP: return "remap";
    case VX_TYPE_GRAPH: return "graph";
    case VX_TYPE_NODE: return "node";
    default: return "object";
    }
}

vx_enum enum_from_name(const char *name, bool *ok)
{
    std::string n = upper(name);
    if (!n.compare(0, 3, "VX_")) n = n.substr(3);
    for (const EnumName &e : kNames)
        if (n == e.name) { if (ok) *ok = true; return e.value; }
    if (ok) *ok = false;
    return 0;
}

vx_df_image format_from_name(const char *name)
{
    std::string n = upper(name);
    if (n == "U8" || n == "UINT8") n = "U008";
    if (n == "S16" || n == "INT16") n = "S016";
    if (n == "U16" || n == "UINT16") n = "U016";
    if (n == "U32" || n == "UINT32") n = "U032";
    if (n == "S32" || n == "INT32") n = "S032";
    if (n == "RGB") n = "RGB2";
    if (n == "RGBX") n = "RGBA";
    if (n.size() != 4) return 0;
    return VX_DF_IMAGE(n[0], n[1], n[2], n[3]);
}

}  // namespace mvxe

namespace {

struct Importer {
    _vx_graph *graph;
    _vx_context *ctx;
    const AgoGraphImportInfo *info;
    std::map<std::string, _vx_reference *> names;
    std::vector<_vx_reference *> created;  // virtual objects and inline scalars: released after import
    int lineno = 0;

    vx_status error(vx_status s, const std::string &msg, const std::string &line)
    {
        log(graph, s, "ERROR: GDF line %d: %s\n>>>> %s\n", lineno, msg.c_str(), line.c_str());
        return s;
    }

    _vx_reference *lookup(const std::string &arg)
    {
        if (arg.empty()) return nullptr;
        if (arg[0] == '$') {
            const int idx = atoi(arg.c_str() + 1) - 1;
            return (idx >= 0 && idx < (int)info->num_ref) ? info->ref[idx] : nullptr;
        }
        std::string name = arg;
        int level = -1;
        const size_t lb = arg.find('[');
        if (lb != std::string::npos && arg.back() == ']') {
            name = arg.substr(0, lb);
            level = atoi(arg.c_str() + lb + 1);
        }
        _vx_reference *r = nullptr;
        auto it = names.find(name);
        if (it != names.end()) r = it->second;
        else {
            std::lock_guard<std::recursive_mutex> g(ctx->lock);
            for (_vx_reference *x : ctx->refs)
                if (x->name == name && x->type != VX_TYPE_NODE) { r = x; break; }
        }
        if (r && level >= 0) {
            if (!is_ref(r, VX_TYPE_PYRAMID) || level >= (int)((_vx_pyramid *)r)->images.size()) return nullptr;
            r = ((_vx_pyramid *)r)->images[level];
        }
        return r;
    }

    _vx_scalar *scalar_from(vx_enum type, const std::string &value)
    {
        _vx_scalar *s = make_scalar(ctx, type, nullptr);
        if (type == VX_TYPE_FLOAT32) s->value.f32 = (float)atof(value.c_str());
        else if (type == VX_TYPE_FLOAT64) s->value.f64 = atof(value.c_str());
        else if (type == VX_TYPE_ENUM) {
            bool ok = false;
            s->value.e = enum_from_name(value.c_str(), &ok);
            if (!ok) s->value.e = (vx_enum)strtol(value.c_str(), nullptr, 0);
        } else if (type == VX_TYPE_DF_IMAGE) s->value.df = format_from_name(value.c_str());
        else if (type == VX_TYPE_BOOL) s->value.b = (upper(value) == "TRUE" || atoi(value.c_str())) ? vx_true_e : vx_false_e;
        else if (type == VX_TYPE_INT64 || type == VX_TYPE_UINT64 || type == VX_TYPE_SIZE) s->value.u64 = strtoull(value.c_str(), nullptr, 0);
        else {
            const long long v = strtoll(value.c_str(), nullptr, 0);
            switch (s->item_size()) {
            case 1: s->value.u8 = (uint8_t)v; break;
            case 2: { uint16_t x = (uint16_t)v; memcpy(&s->value, &x, 2); break; }
            default: s->value.i32 = (int32_t)v; break;
            }
        }
        return s;
    }

    // INIT on thresholds: "{lower;upper}" / "lower,upper" / "value"
    vx_status init_threshold(_vx_threshold *t, const std::vector<std::string> &io)
    {
        std::vector<int> v;
        for (size_t i = 1; i < io.size(); i++) {
            std::string x = io[i];
            for (char &c : x) if (c == '{' || c == '}' || c == ';') c = ' ';
            std::istringstream is(x);
            int k;
            while (is >> k) v.push_back(k);
        }
        if (t->threshType == VX_THRESHOLD_TYPE_RANGE && v.size() >= 2) {
            memset(&t->lower, 0, sizeof(t->lower));
            memset(&t->upper, 0, sizeof(t->upper));
            t->lower.S32 = v[0];
            t->upper.S32 = v[1];
        } else if (t->threshType == VX_THRESHOLD_TYPE_BINARY && v.size() >= 1) {
            memset(&t->value, 0, sizeof(t->value));
            t->value.S32 = v[0];
        } else {
            return VX_ERROR_INVALID_PARAMETERS;
        }
        return VX_SUCCESS;
    }

    vx_status init_matrix(_vx_matrix *m, const std::vector<std::string> &io)
    {
        std::string all;
        for (size_t i = 1; i < io.size(); i++) all += io[i] + " ";
        for (char &c : all) if (c == '{' || c == '}' || c == ';' || c == ',') c = ' ';
        std::istringstream is(all);
        const size_t n = m->columns * m->rows;
        for (size_t i = 0; i < n; i++) {
            double v;
            if (!(is >> v)) return VX_ERROR_INVALID_PARAMETERS;
            if (m->dataType == VX_TYPE_FLOAT32) ((float *)m->data.data())[i] = (float)v;
            else if (m->dataType == VX_TYPE_INT32) ((int32_t *)m->data.data())[i] = (int32_t)v;
            else m->data[i] = (uint8_t)v;
        }
        return VX_SUCCESS;
    }

    _vx_reference *create(const std::string &desc, std::string &ioOut, std::string &err)
    {
        std::vector<std::string> parts = split(desc, ':');
        std::string kind = parts[0];
        std::string cfg = parts.size() > 1 ? parts[1] : "";
        ioOut.clear();
        for (size_t i = 2; i < parts.size(); i++) ioOut += (i > 2 ? ":" : "") + parts[i];
        bool isVirtual = false;
        for (const char *v : {"-virtual", "virtual-"}) {
            const size_t p = kind.find(v);
            if (p != std::string::npos) { isVirtual = true; kind.erase(p, strlen(v)); }
        }
        std::vector<std::string> a = split(cfg, ',');
        _vx_reference *r = nullptr;
        if (kind == "image" || kind == "uniform-image" || kind == "image-uniform") {
            uint32_t w = 0, h = 0;
            if (a.size() < 3 || !to_u32(a[0], w) || !to_u32(a[1], h)) { err = "image:<width>,<height>,<format>"; return nullptr; }
            const vx_df_image f = format_from_name(a[2].c_str());
            if (!f) { err = "image format"; return nullptr; }
            if (kind != "image") {
                vx_pixel_value_t v{};
                const unsigned long long x = a.size() > 3 ? strtoull(a[3].c_str(), nullptr, 0) : 0;
                memcpy(&v, &x, sizeof(x));
                r = vxCreateUniformImage(ctx, w, h, f, &v);
            } else {
                r = isVirtual ? vxCreateVirtualImage(graph, w, h, f) : vxCreateImage(ctx, w, h, f);
            }
        } else if (kind == "pyramid") {
            uint32_t n = 0, w = 0, h = 0;
            if (a.size() < 5 || !to_u32(a[0], n) || !to_u32(a[2], w) || !to_u32(a[3], h)) {
                err = "pyramid:<levels>,half|orb|<scale>,<width>,<height>,<format>";
                return nullptr;
            }
            const std::string sc = upper(a[1]);
            const float scale = sc == "HALF" ? VX_SCALE_PYRAMID_HALF : sc == "ORB" ? VX_SCALE_PYRAMID_ORB : (float)atof(a[1].c_str());
            const vx_df_image f = format_from_name(a[4].c_str());
            r = isVirtual ? vxCreateVirtualPyramid(graph, n, scale, w, h, f) : vxCreatePyramid(ctx, n, scale, w, h, f);
        } else if (kind == "array") {
            bool ok = false;
            const vx_enum t = a.size() >= 2 ? enum_from_name(a[0].c_str(), &ok) : 0;
            uint32_t cap = 0;
            if (!ok || !to_u32(a[1], cap)) { err = "array:<type>,<capacity>"; return nullptr; }
            r = isVirtual ? vxCreateVirtualArray(graph, t, cap) : vxCreateArray(ctx, t, cap);
        } else if (kind == "scalar") {
            bool ok = false;
            const vx_enum t = a.size() >= 1 ? enum_from_name(a[0].c_str(), &ok) : 0;
            if (!ok) { err = "scalar:<type>,<value>"; return nullptr; }
            std::string value = cfg.size() > a[0].size() ? cfg.substr(a[0].size() + 1) : "0";
            r = scalar_from(t, value);
            if (isVirtual) { r->isVirtual = true; r->scope = graph; }
        } else if (kind == "threshold") {
            bool ok = false;
            const vx_enum tt = a.size() >= 2 ? enum_from_name(a[0].c_str(), &ok) : 0;
            if (!ok) { err = "threshold:BINARY|RANGE,<input-format>[,<output-format>]"; return nullptr; }
            vx_df_image in = format_from_name(a[1].c_str());
            const vx_df_image out = a.size() > 2 ? format_from_name(a[2].c_str()) : VX_DF_IMAGE_U8;
            r = vxCreateThresholdForImage(ctx, tt, in, out ? out : VX_DF_IMAGE_U8);
        } else if (kind == "matrix") {
            bool ok = false;
            const vx_enum t = a.size() >= 3 ? enum_from_name(a[0].c_str(), &ok) : 0;
            uint32_t cols = 0, rows = 0;
            if (!ok || !to_u32(a[1], cols) || !to_u32(a[2], rows)) { err = "matrix:<type>,<columns>,<rows>"; return nullptr; }
            r = vxCreateMatrix(ctx, t, cols, rows);
        } else if (kind == "remap") {
            uint32_t v[4];
            if (a.size() < 4 || !to_u32(a[0], v[0]) || !to_u32(a[1], v[1]) || !to_u32(a[2], v[2]) || !to_u32(a[3], v[3])) {
                err = "remap:<src-width>,<src-height>,<dst-width>,<dst-height>";
                return nullptr;
            }
            r = vxCreateRemap(ctx, v[0], v[1], v[2], v[3]);
        } else {
            err = "object type '" + kind + "' is not supported";
            return nullptr;
        }
        if (!r) { err = "cannot create " + kind; return nullptr; }
        if (isVirtual) created.push_back(r);
        // INIT is part of the description; every other I/O request goes to the registry callback
        std::vector<std::string> io = split(ioOut, ',');
        if (!ioOut.empty() && upper(io[0]) == "INIT") {
            vx_status s = VX_SUCCESS;
            if (r->type == VX_TYPE_THRESHOLD) s = init_threshold((_vx_threshold *)r, io);
            else if (r->type == VX_TYPE_MATRIX) s = init_matrix((_vx_matrix *)r, io);
            else if (r->type == VX_TYPE_SCALAR && io.size() > 1) {
                _vx_scalar *t = scalar_from(((_vx_scalar *)r)->dataType, io[1]);
                ((_vx_scalar *)r)->value = t->value;
                release_ref(t, true);
            } else s = VX_ERROR_NOT_SUPPORTED;
            if (s) { err = "INIT values"; return nullptr; }
            ioOut.clear();
        }
        return r;
    }

    // the context's kernels: the built-in table and user kernels (vxAddUserKernel, modules loaded by
    // "import"), by full name or by the part after the last '.'
    const KernelDef *find_kernel(const std::string &name)
    {
        std::lock_guard<std::recursive_mutex> lk(graph->context->lock);
        for (_vx_kernel *k : graph->context->kernels)
            if (name == k->def->name) return k->def;
        for (_vx_kernel *k : graph->context->kernels) {
            const char *p = strrchr(k->def->name, '.');
            if (p && name == p + 1) return k->def;
        }
        return nullptr;
    }

    vx_status node_line(const std::vector<std::string> &arg, const std::string &line)
    {
        const KernelDef *def = find_kernel(arg[1]);
        if (!def) return error(VX_ERROR_INVALID_PARAMETERS, "kernel not supported: " + arg[1], line);
        _vx_node *n = create_node(graph, def, {});
        if (!n) return error(VX_FAILURE, "cannot create node", line);
        // the graph holds the node; the importer's own reference ends with this line
        struct Drop { _vx_node *n; ~Drop() { release_ref(n, true); } } drop{n};
        size_t p = 0;
        for (size_t i = 2; i < arg.size(); i++) {
            const std::string &x = arg[i];
            if (!x.compare(0, 5, "attr:")) {
                if (!x.compare(5, 12, "BORDER_MODE:")) {
                    const std::string m = x.substr(17);
                    memset(&n->border, 0, sizeof(n->border));
                    if (m == "UNDEFINED") n->border.mode = VX_BORDER_UNDEFINED;
                    else if (m == "REPLICATE") n->border.mode = VX_BORDER_REPLICATE;
                    else if (!m.compare(0, 9, "CONSTANT,")) {
                        n->border.mode = VX_BORDER_CONSTANT;
                        n->border.constant_value.U32 = (uint32_t)strtoul(m.c_str() + 9, nullptr, 0);
                    } else return error(VX_ERROR_INVALID_PARAMETERS, "border mode", line);
                } else if (!x.compare(5, 9, "AFFINITY:")) {
                    const std::string d = x.substr(14);
                    n->affinity.device_type = !d.compare(0, 3, "CPU") ? AGO_TARGET_AFFINITY_CPU
                                            : !d.compare(0, 3, "GPU") ? AGO_TARGET_AFFINITY_GPU : 0;
                    if (!n->affinity.device_type) return error(VX_ERROR_INVALID_PARAMETERS, "affinity", line);
                    n->affinity.device_info = (uint32_t)atoi(d.c_str() + 3);
                    const size_t comma = d.find(',');
                    n->affinity.group = comma == std::string::npos ? 0 : (uint32_t)atoi(d.c_str() + comma + 1);
                } else return error(VX_ERROR_INVALID_PARAMETERS, "attribute", line);
                continue;
            }
            if (p >= n->params.size()) return error(VX_ERROR_INVALID_PARAMETERS, "too many arguments", line);
            _vx_reference *r = nullptr;
            if (x == "null") r = nullptr;
            else if (x[0] == '!') {
                bool ok = false;
                const vx_enum e = enum_from_name(x.c_str() + 1, &ok);
                if (!ok) return error(VX_ERROR_INVALID_PARAMETERS, "unknown enum " + x, line);
                r = make_scalar(ctx, VX_TYPE_ENUM, &e);
                created.push_back(r);
            } else if (!(r = lookup(x))) {
                return error(VX_ERROR_INVALID_PARAMETERS, "undefined data " + x, line);
            }
            if (r) {
                r->intCount++;
                n->params[p] = r;
            }
            p++;
        }
        return VX_SUCCESS;
    }

    vx_status run(const std::string &text)
    {
        std::istringstream in(text);
        std::string raw, line;
        while (std::getline(in, raw)) {
            lineno++;
            while (!raw.empty() && (raw.back() == '\r' || raw.back() == '\n')) raw.pop_back();
            if (!raw.empty() && raw.back() == '\\') {
                line += raw.substr(0, raw.size() - 1) + " ";
                continue;
            }
            line += raw;
            std::string l = line;
            line.clear();
            const size_t hash = l.find('#');
            if (hash != std::string::npos) l = l.substr(0, hash);
            std::istringstream ws(l);
            std::vector<std::string> arg;
            std::string t;
            while (ws >> t) arg.push_back(t);
            if (arg.empty()) continue;
            if (info->dumpToConsole) log(graph, VX_SUCCESS, "%s\n", l.c_str());
            if (arg[0] == "data" && arg.size() == 4 && arg[2] == "=") {
                std::string io, err;
                _vx_reference *r = create(arg[3], io, err);
                if (!r) return error(VX_FAILURE, "data: " + err, l);
                r->name = arg[1];
                names[arg[1]] = r;
                if (r->type == VX_TYPE_PYRAMID) {
                    _vx_pyramid *py = (_vx_pyramid *)r;
                    for (size_t k = 0; k < py->images.size(); k++) py->images[k]->name = arg[1] + "[" + std::to_string(k) + "]";
                }
                if (!io.empty() && info->data_registry_callback_f) {
                    r->extCount++;
                    info->data_registry_callback_f(info->data_registry_callback_obj, r, arg[1].c_str(), io.c_str());
                }
            } else if (arg[0] == "node" && arg.size() >= 2) {
                vx_status s = node_line(arg, l);
                if (s) return s;
            } else if (arg[0] == "import" && arg.size() == 2) {
                vx_status s = vxLoadKernels(ctx, arg[1].c_str());
                if (s) return error(s, "import " + arg[1], l);
            } else if (arg[0] == "affinity" && arg.size() == 2) {
                const std::string d = upper(arg[1]);
                graph->affinity.device_type = !d.compare(0, 3, "CPU") ? AGO_TARGET_AFFINITY_CPU : AGO_TARGET_AFFINITY_GPU;
                graph->affinity.device_info = (uint32_t)atoi(d.c_str() + 3);
            } else {
                return error(VX_ERROR_NOT_SUPPORTED, "unsupported command '" + arg[0] + "'", l);
            }
        }
        return VX_SUCCESS;
    }
};

}  // namespace

namespace mvxe {

vx_status import_graph_text(_vx_graph *graph, const AgoGraphImportInfo *info)
{
    if (!info || !info->text) return VX_ERROR_INVALID_PARAMETERS;
    std::string text = info->text;
    if (!text.compare(0, 5, "file ")) {
        std::string fn = text.substr(5);
        while (!fn.empty() && isspace((unsigned char)fn.back())) fn.pop_back();
        FILE *fp = fopen(fn.c_str(), "rb");
        if (!fp) {
            log(graph, VX_FAILURE, "ERROR: GDF import: unable to open %s\n", fn.c_str());
            return VX_FAILURE;
        }
        text.clear();
        char buf[4096];
        size_t k;
        while ((k = fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, k);
        fclose(fp);
    }
    Importer im{graph, graph->context, info, {}, {}, 0};
    vx_status s = im.run(text);
    // nodes hold what they use; inline scalars and virtual objects need no other owner
    for (_vx_reference *r : im.created) release_ref(r, true);
    invalidate_graph(graph);
    return s;
}

}  // namespace mvxe
