// Field snapshots as .npy files (NumPy format 1.0 / 2.0 / 3.0, float32 little
// endian, C order, shape (height, width)): the format of the reference's
// state dumps (LifeCore/IO/NpyWriter.cpp:42-171, readNPYFloat32 /
// writeNPYFloat32), applied to a world's field.  The writer pads the
// preamble to a multiple of 64 bytes as NumPy does; the reader takes any
// version-1-to-3 header whose dtype, order and shape match the world.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.h"

using namespace lenia;

namespace {

// value text of `key` in the header dict ("'key': value"), up to the next
// top-level ',' or '}' (tuples kept whole)
bool dict_value(const std::string &d, const char *key, std::string &out) {
    const std::string k1 = std::string("'") + key + "'", k2 = std::string("\"") + key + "\"";
    size_t p = d.find(k1);
    if (p == std::string::npos) p = d.find(k2);
    if (p == std::string::npos) return false;
    p = d.find(':', p);
    if (p == std::string::npos) return false;
    ++p;
    while (p < d.size() && d[p] == ' ') ++p;
    int depth = 0;
    size_t e = p;
    for (; e < d.size(); ++e) {
        char c = d[e];
        if (c == '(') ++depth;
        else if (c == ')') --depth;
        else if ((c == ',' || c == '}') && depth == 0) break;
    }
    out = d.substr(p, e - p);
    while (!out.empty() && out.back() == ' ') out.pop_back();
    return true;
}

bool parse_shape(const std::string &t, std::vector<uint64_t> &shape) {
    shape.clear();
    if (t.size() < 2 || t.front() != '(' || t.back() != ')') return false;
    const std::string inner = t.substr(1, t.size() - 2);
    size_t p = 0;
    while (p < inner.size()) {
        while (p < inner.size() && (inner[p] == ' ' || inner[p] == ',')) ++p;
        if (p >= inner.size()) break;
        uint64_t v = 0;
        size_t q = p;
        while (q < inner.size() && inner[q] >= '0' && inner[q] <= '9') {
            v = v * 10 + (uint64_t)(inner[q] - '0');
            if (v > (1ull << 40)) return false;
            ++q;
        }
        if (q == p) return false;
        shape.push_back(v);
        p = q;
    }
    return true;
}

}  // namespace

extern "C" int lenia_world_save_npy(lenia_world *w, const char *path) {
    if (!w || !path) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_save_npy: NULL argument");
    const uint32_t W = w->desc.width, H = w->desc.height;
    std::vector<float> buf((size_t)W * H);
    int rc = lenia_world_read_state(w, buf.data(), buf.size());
    if (rc) return rc;
    std::string dict = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(H) + ", " +
                       std::to_string(W) + "), }";
    const size_t fixed = 6 + 2 + 2;
    const size_t pad = (64 - (fixed + dict.size() + 1) % 64) % 64;
    dict.append(pad, ' ');
    dict.push_back('\n');
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(LENIA_ERR_INVALID_ARG, "cannot open %s for writing", path);
    const unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(dict.size() & 0xFF),
                                   (unsigned char)(dict.size() >> 8)};
    bool ok = std::fwrite(pre, 1, 10, f) == 10 && std::fwrite(dict.data(), 1, dict.size(), f) == dict.size() &&
              std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(LENIA_ERR_INVALID_ARG, "short write on %s", path);
    return LENIA_OK;
}

extern "C" int lenia_world_load_npy(lenia_world *w, const char *path) {
    if (!w || !path) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_load_npy: NULL argument");
    const uint32_t W = w->desc.width, H = w->desc.height;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(LENIA_ERR_INVALID_ARG, "cannot open %s for reading", path);
    unsigned char pre[8];
    if (std::fread(pre, 1, 8, f) != 8 || std::memcmp(pre, "\x93NUMPY", 6) != 0 || pre[6] < 1 || pre[6] > 3) {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: not an .npy file of format version 1-3", path);
    }
    size_t hlen = 0;
    unsigned char lb[4] = {0, 0, 0, 0};
    const size_t nlb = pre[6] == 1 ? 2 : 4;
    if (std::fread(lb, 1, nlb, f) != nlb) {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: truncated header", path);
    }
    hlen = (size_t)lb[0] | ((size_t)lb[1] << 8) | ((size_t)lb[2] << 16) | ((size_t)lb[3] << 24);
    if (hlen == 0 || hlen > (1u << 20)) {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: bad header length", path);
    }
    std::string dict(hlen, '\0');
    if (std::fread(&dict[0], 1, hlen, f) != hlen) {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: truncated header", path);
    }
    std::string descr, order, shapeText;
    std::vector<uint64_t> shape;
    const bool parsed = dict_value(dict, "descr", descr) && dict_value(dict, "fortran_order", order) &&
                        dict_value(dict, "shape", shapeText) && parse_shape(shapeText, shape);
    if (!parsed) {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: cannot parse the header", path);
    }
    if (descr != "'<f4'" && descr != "\"<f4\"") {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: dtype %s, expected little-endian float32 ('<f4')", path,
                    descr.c_str());
    }
    if (order != "False") {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: Fortran-ordered arrays are not supported", path);
    }
    if (shape.size() != 2 || shape[0] != H || shape[1] != W) {
        std::fclose(f);
        return fail(LENIA_ERR_INVALID_ARG, "%s: shape %s does not match the world (%u, %u)", path,
                    shapeText.c_str(), H, W);
    }
    std::vector<float> buf((size_t)W * H);
    const size_t got = std::fread(buf.data(), sizeof(float), buf.size(), f);
    std::fclose(f);
    if (got != buf.size())
        return fail(LENIA_ERR_INVALID_ARG, "%s: truncated data (%zu of %zu values)", path, got, buf.size());
    return lenia_world_write_state(w, buf.data(), buf.size());
}
