// glTF 2.0 reader for polyline (hair) and triangle assets: the part of assimp 6.0.1's glTF2
// importer that ModelLoader::LoadFromFile relies on (source/resources/model/model_loader.cpp:
// 139-312; importer flags Triangulate | FlipUVs | GenNormals | GenBoundingBoxes, :278),
// restated from the glTF 2.0 specification, and Node::GetWorldMatrix
// (source/resources/model/model.cpp:6-18). Host code only.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "asset.hpp"

namespace {

thread_local std::string g_asset_error;

int asset_fail(int code, const std::string& msg) {
    g_asset_error = msg;
    return code;
}

// ---- JSON (RFC 8259) --------------------------------------------------------------------

struct Json {
    enum Kind : uint8_t { kNull, kBool, kNum, kStr, kArr, kObj };
    Kind kind = kNull;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> items;        // array elements, or object values
    std::vector<std::string> keys;  // object keys, parallel to items

    const Json* get(const char* k) const {
        if (kind != kObj) return nullptr;
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == k) return &items[i];
        return nullptr;
    }
    const Json* at(size_t i) const { return kind == kArr && i < items.size() ? &items[i] : nullptr; }
    size_t size() const { return kind == kArr ? items.size() : 0; }
};

class JsonParser {
public:
    JsonParser(const char* b, const char* e) : b_(b), p_(b), e_(e) {}

    bool parse(Json& out, std::string& err) {
        ws();
        if (!value(out, 0)) {
            err = "JSON: " + err_ + " at byte " + std::to_string(p_ - b_);
            return false;
        }
        ws();
        if (p_ != e_) {
            err = "JSON: trailing characters at byte " + std::to_string(p_ - b_);
            return false;
        }
        return true;
    }

private:
    static constexpr int kMaxDepth = 256;

    bool fail(const char* m) {
        if (err_.empty()) err_ = m;
        return false;
    }
    void ws() {
        while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    bool lit(const char* s) {
        size_t n = std::strlen(s);
        if (size_t(e_ - p_) < n || std::memcmp(p_, s, n) != 0) return fail("invalid literal");
        p_ += n;
        return true;
    }

    bool value(Json& v, int depth) {
        if (depth > kMaxDepth) return fail("nesting too deep");
        if (p_ >= e_) return fail("unexpected end");
        switch (*p_) {
            case '{': return object(v, depth);
            case '[': return array(v, depth);
            case '"': v.kind = Json::kStr; return string(v.str);
            case 't': v.kind = Json::kBool; v.b = true; return lit("true");
            case 'f': v.kind = Json::kBool; v.b = false; return lit("false");
            case 'n': v.kind = Json::kNull; return lit("null");
            default: v.kind = Json::kNum; return number(v.num);
        }
    }

    bool object(Json& v, int depth) {
        v.kind = Json::kObj;
        ++p_;
        ws();
        if (p_ < e_ && *p_ == '}') {
            ++p_;
            return true;
        }
        for (;;) {
            ws();
            if (p_ >= e_ || *p_ != '"') return fail("expected a key");
            v.keys.emplace_back();
            if (!string(v.keys.back())) return false;
            ws();
            if (p_ >= e_ || *p_ != ':') return fail("expected ':'");
            ++p_;
            ws();
            v.items.emplace_back();
            if (!value(v.items.back(), depth + 1)) return false;
            ws();
            if (p_ < e_ && *p_ == ',') {
                ++p_;
                continue;
            }
            if (p_ < e_ && *p_ == '}') {
                ++p_;
                return true;
            }
            return fail("expected ',' or '}'");
        }
    }

    bool array(Json& v, int depth) {
        v.kind = Json::kArr;
        ++p_;
        ws();
        if (p_ < e_ && *p_ == ']') {
            ++p_;
            return true;
        }
        for (;;) {
            ws();
            v.items.emplace_back();
            if (!value(v.items.back(), depth + 1)) return false;
            ws();
            if (p_ < e_ && *p_ == ',') {
                ++p_;
                continue;
            }
            if (p_ < e_ && *p_ == ']') {
                ++p_;
                return true;
            }
            return fail("expected ',' or ']'");
        }
    }

    static void utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s += char(cp);
        } else if (cp < 0x800) {
            s += char(0xC0 | (cp >> 6));
            s += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += char(0xE0 | (cp >> 12));
            s += char(0x80 | ((cp >> 6) & 0x3F));
            s += char(0x80 | (cp & 0x3F));
        } else {
            s += char(0xF0 | (cp >> 18));
            s += char(0x80 | ((cp >> 12) & 0x3F));
            s += char(0x80 | ((cp >> 6) & 0x3F));
            s += char(0x80 | (cp & 0x3F));
        }
    }

    bool hex4(uint32_t& cp) {
        if (e_ - p_ < 4) return fail("short \\u escape");
        cp = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p_++;
            cp <<= 4;
            if (c >= '0' && c <= '9') cp |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') cp |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') cp |= uint32_t(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }

    bool string(std::string& s) {
        ++p_;  // opening quote
        while (p_ < e_) {
            unsigned char c = static_cast<unsigned char>(*p_++);
            if (c == '"') return true;
            if (c < 0x20) return fail("control character in string");
            if (c != '\\') {
                s += char(c);
                continue;
            }
            if (p_ >= e_) break;
            char esc = *p_++;
            switch (esc) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {  // high surrogate: a low one must follow
                        uint32_t lo;
                        if (e_ - p_ < 2 || p_[0] != '\\' || p_[1] != 'u') return fail("lone surrogate");
                        p_ += 2;
                        if (!hex4(lo)) return false;
                        if (lo < 0xDC00 || lo >= 0xE000) return fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        return fail("lone surrogate");
                    }
                    utf8(s, cp);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        return fail("unterminated string");
    }

    bool number(double& out) {
        const char* s = p_;
        if (p_ < e_ && *p_ == '-') ++p_;
        if (p_ >= e_) return fail("bad number");
        if (*p_ == '0') {
            ++p_;
        } else if (*p_ >= '1' && *p_ <= '9') {
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        } else {
            return fail("unexpected character");
        }
        if (p_ < e_ && *p_ == '.') {
            ++p_;
            if (p_ >= e_ || *p_ < '0' || *p_ > '9') return fail("bad fraction");
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
            ++p_;
            if (p_ < e_ && (*p_ == '+' || *p_ == '-')) ++p_;
            if (p_ >= e_ || *p_ < '0' || *p_ > '9') return fail("bad exponent");
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        // the grammar above is the JSON one, which strtod reads the same way in the "C" locale
        std::string tok(s, p_);
        errno = 0;
        out = std::strtod(tok.c_str(), nullptr);
        return true;
    }

    const char* b_;
    const char* p_;
    const char* e_;
    std::string err_;
};

// ---- buffers ------------------------------------------------------------------------------

bool base64(const char* s, size_t n, std::vector<uint8_t>& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    while (n && s[n - 1] == '=') --n;
    out.clear();
    out.reserve(n * 3 / 4);
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = 0; i < n; ++i) {
        int v = val(s[i]);
        if (v < 0) return false;
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(uint8_t((acc >> bits) & 0xFF));
        }
    }
    return bits < 6;  // a single leftover symbol cannot end a valid encoding
}

std::string percent_decode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size()) {
            char h[3] = {s[i + 1], s[i + 2], 0};
            char* end = nullptr;
            long v = std::strtol(h, &end, 16);
            if (end == h + 2) {
                o += char(v);
                i += 2;
                continue;
            }
        }
        o += s[i];
    }
    return o;
}

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    std::streamsize n = f.tellg();
    if (n < 0) return false;
    f.seekg(0);
    out.resize(size_t(n));
    return n == 0 || bool(f.read(reinterpret_cast<char*>(out.data()), n));
}

uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

// ---- glTF ----------------------------------------------------------------------------------

enum : uint32_t {
    kByte = 5120,
    kUByte = 5121,
    kShort = 5122,
    kUShort = 5123,
    kUInt = 5125,
    kFloat = 5126,
};

uint32_t component_size(uint32_t t) {
    switch (t) {
        case kByte:
        case kUByte: return 1;
        case kShort:
        case kUShort: return 2;
        case kUInt:
        case kFloat: return 4;
        default: return 0;
    }
}

int type_components(const std::string& t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    return 0;
}

struct View {
    const uint8_t* base = nullptr;  // null: every element is zero
    size_t stride = 0;
    uint32_t count = 0;
    int comps = 0;
    uint32_t ctype = 0;
    bool normalized = false;

    // component c of element i as float; normalized integers map as the glTF specification's
    // "Animations" / "Meshes" sections state (unsigned: c / max, signed: max(c / max, -1))
    float f(uint32_t i, int c) const {
        if (!base) return 0.0f;
        const uint8_t* p = base + size_t(i) * stride + size_t(c) * component_size(ctype);
        switch (ctype) {
            case kFloat: {
                float v;
                std::memcpy(&v, p, 4);
                return v;
            }
            case kUByte: return normalized ? float(p[0]) / 255.0f : float(p[0]);
            case kByte: {
                float v = float(int8_t(p[0]));
                return normalized ? std::fmax(v / 127.0f, -1.0f) : v;
            }
            case kUShort: {
                uint16_t u;
                std::memcpy(&u, p, 2);
                return normalized ? float(u) / 65535.0f : float(u);
            }
            case kShort: {
                int16_t s;
                std::memcpy(&s, p, 2);
                return normalized ? std::fmax(float(s) / 32767.0f, -1.0f) : float(s);
            }
            case kUInt: return float(rd32(p));
            default: return 0.0f;
        }
    }
    uint32_t u(uint32_t i) const {
        if (!base) return 0;
        const uint8_t* p = base + size_t(i) * stride;
        switch (ctype) {
            case kUByte: return p[0];
            case kUShort: {
                uint16_t v;
                std::memcpy(&v, p, 2);
                return v;
            }
            case kUInt: return rd32(p);
            default: return 0;
        }
    }
};

class Loader {
public:
    Loader(const std::string& base_dir, bool has_dir) : dir_(base_dir), has_dir_(has_dir) {}

    int load(const uint8_t* data, size_t size, vkhrt_asset& out) {
        const char* json_b = reinterpret_cast<const char*>(data);
        const char* json_e = json_b + size;
        if (size >= 12 && rd32(data) == 0x46546C67u) {  // "glTF": binary container
            if (rd32(data + 4) != 2) return err(VKHRT_ERROR_UNSUPPORTED, "GLB container version is not 2");
            const uint64_t total = rd32(data + 8);
            if (total > size) return err(VKHRT_ERROR_INVALID_ARGUMENT, "GLB length exceeds the data");
            size_t off = 12;
            bool have_json = false;
            while (off + 8 <= total) {
                const uint32_t len = rd32(data + off), type = rd32(data + off + 4);
                if (off + 8 + uint64_t(len) > total) return err(VKHRT_ERROR_INVALID_ARGUMENT, "GLB chunk overruns the file");
                const uint8_t* body = data + off + 8;
                if (!have_json) {
                    if (type != 0x4E4F534Au) return err(VKHRT_ERROR_INVALID_ARGUMENT, "GLB does not start with a JSON chunk");
                    json_b = reinterpret_cast<const char*>(body);
                    json_e = json_b + len;
                    have_json = true;
                } else if (type == 0x004E4942u && !bin_) {
                    bin_ = body;
                    bin_size_ = len;
                }
                off += 8 + size_t(len);
            }
            if (!have_json) return err(VKHRT_ERROR_INVALID_ARGUMENT, "GLB without a JSON chunk");
        }
        // a UTF-8 byte order mark may precede the JSON text
        if (json_e - json_b >= 3 && std::memcmp(json_b, "\xEF\xBB\xBF", 3) == 0) json_b += 3;
        std::string perr;
        JsonParser jp(json_b, json_e);
        if (!jp.parse(root_, perr)) return err(VKHRT_ERROR_INVALID_ARGUMENT, perr);
        if (root_.kind != Json::kObj) return err(VKHRT_ERROR_INVALID_ARGUMENT, "the glTF root is not an object");
        const Json* asset = root_.get("asset");
        const Json* version = asset ? asset->get("version") : nullptr;
        if (!version || version->kind != Json::kStr) return err(VKHRT_ERROR_INVALID_ARGUMENT, "missing asset.version");
        if (version->str.empty() || version->str[0] != '2' || (version->str.size() > 1 && version->str[1] != '.'))
            return err(VKHRT_ERROR_UNSUPPORTED, "glTF version " + version->str + " (2.x is read)");
        int rc;
        if ((rc = buffers()) != VKHRT_SUCCESS) return rc;
        if ((rc = materials(out)) != VKHRT_SUCCESS) return rc;
        if ((rc = meshes(out)) != VKHRT_SUCCESS) return rc;
        if ((rc = nodes(out)) != VKHRT_SUCCESS) return rc;
        for (uint32_t n = 0; n < out.nodes.size(); ++n)
            for (uint32_t m : out.nodes[n].meshes) out.instances.push_back({n, m});
        return VKHRT_SUCCESS;
    }

private:
    int err(int code, const std::string& m) { return asset_fail(code, m); }

    const Json* arr(const Json& o, const char* k) const {
        const Json* a = o.get(k);
        return a && a->kind == Json::kArr ? a : nullptr;
    }
    // non-negative integer property (required unless a default is given)
    bool uint_prop(const Json& o, const char* k, uint64_t& v, bool required, uint64_t dflt = 0) const {
        const Json* j = o.get(k);
        if (!j) {
            v = dflt;
            return !required;
        }
        if (j->kind != Json::kNum || !(j->num >= 0.0) || j->num != std::floor(j->num) || j->num > 1.8e19) return false;
        v = uint64_t(j->num);
        return true;
    }
    bool floats(const Json& o, const char* k, float* v, size_t n) const {
        const Json* j = o.get(k);
        if (!j) return true;
        if (j->kind != Json::kArr || j->items.size() != n) return false;
        for (size_t i = 0; i < n; ++i) {
            if (j->items[i].kind != Json::kNum) return false;
            v[i] = float(j->items[i].num);
        }
        return true;
    }

    int buffers() {
        const Json* bs = arr(root_, "buffers");
        for (size_t i = 0; bs && i < bs->items.size(); ++i) {
            const Json& b = bs->items[i];
            uint64_t len;
            if (!uint_prop(b, "byteLength", len, true))
                return err(VKHRT_ERROR_INVALID_ARGUMENT, "buffer " + std::to_string(i) + ": bad byteLength");
            bufs_.emplace_back();
            Buf& B = bufs_.back();
            const Json* uri = b.get("uri");
            if (!uri) {
                if (i != 0 || !bin_)
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "buffer " + std::to_string(i) + " has no uri and no GLB binary chunk");
                B.data = bin_;
                B.size = bin_size_;
            } else if (uri->kind != Json::kStr) {
                return err(VKHRT_ERROR_INVALID_ARGUMENT, "buffer " + std::to_string(i) + ": uri is not a string");
            } else if (uri->str.compare(0, 5, "data:") == 0) {
                size_t comma = uri->str.find(',');
                if (comma == std::string::npos || uri->str.rfind(";base64", comma) == std::string::npos)
                    return err(VKHRT_ERROR_UNSUPPORTED, "buffer " + std::to_string(i) + ": data URI is not base64");
                if (!base64(uri->str.data() + comma + 1, uri->str.size() - comma - 1, B.own))
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "buffer " + std::to_string(i) + ": bad base64");
                B.data = B.own.data();
                B.size = B.own.size();
            } else {
                if (!has_dir_)
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "buffer " + std::to_string(i) + " is an external file and no directory was given");
                const std::string path = dir_.empty() ? percent_decode(uri->str) : dir_ + "/" + percent_decode(uri->str);
                if (!read_file(path, B.own)) return err(VKHRT_ERROR_INVALID_ARGUMENT, "cannot read buffer file " + path);
                B.data = B.own.data();
                B.size = B.own.size();
            }
            if (B.size < len)
                return err(VKHRT_ERROR_INVALID_ARGUMENT, "buffer " + std::to_string(i) + " holds " + std::to_string(B.size) +
                                                             " bytes, byteLength says " + std::to_string(len));
            B.size = size_t(len);
        }
        return VKHRT_SUCCESS;
    }

    // accessor `index` checked against its buffer view, as a strided view
    int accessor(uint64_t index, View& v, const std::string& what) {
        const Json* as = arr(root_, "accessors");
        const Json* a = as ? as->at(size_t(index)) : nullptr;
        if (!a) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": accessor " + std::to_string(index) + " does not exist");
        uint64_t ctype, count, aoff;
        const Json* type = a->get("type");
        if (!uint_prop(*a, "componentType", ctype, true) || !component_size(uint32_t(ctype)) ||
            !uint_prop(*a, "count", count, true) || count > 0xFFFFFFFFull || !uint_prop(*a, "byteOffset", aoff, false) ||
            !type || type->kind != Json::kStr || !type_components(type->str))
            return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": malformed accessor " + std::to_string(index));
        if (a->get("sparse")) return err(VKHRT_ERROR_UNSUPPORTED, what + ": sparse accessors are not read");
        const Json* norm = a->get("normalized");
        v = View{};
        v.ctype = uint32_t(ctype);
        v.count = uint32_t(count);
        v.comps = type_components(type->str);
        v.normalized = norm && norm->kind == Json::kBool && norm->b;
        const uint32_t elem = uint32_t(v.comps) * component_size(v.ctype);
        uint64_t bv;
        if (!a->get("bufferView")) return VKHRT_SUCCESS;  // all zeros
        if (!uint_prop(*a, "bufferView", bv, true))
            return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bad bufferView of accessor " + std::to_string(index));
        const Json* vs = arr(root_, "bufferViews");
        const Json* view = vs ? vs->at(size_t(bv)) : nullptr;
        uint64_t buf, voff, vlen, stride;
        if (!view || !uint_prop(*view, "buffer", buf, true) || !uint_prop(*view, "byteOffset", voff, false) ||
            !uint_prop(*view, "byteLength", vlen, true) || !uint_prop(*view, "byteStride", stride, false, elem))
            return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": malformed bufferView " + std::to_string(bv));
        if (buf >= bufs_.size()) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bufferView " + std::to_string(bv) + " names a missing buffer");
        if (stride < elem) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": byteStride below the element size");
        const Buf& B = bufs_[size_t(buf)];
        if (voff + vlen > B.size) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bufferView " + std::to_string(bv) + " overruns its buffer");
        if (count && aoff + stride * (count - 1) + elem > vlen)
            return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": accessor " + std::to_string(index) + " overruns its bufferView");
        v.base = B.data + voff + aoff;
        v.stride = size_t(stride);
        return VKHRT_SUCCESS;
    }

    int materials(vkhrt_asset& out) {
        const Json* ms = arr(root_, "materials");
        for (size_t i = 0; ms && i < ms->items.size(); ++i) {
            const Json& m = ms->items[i];
            vkhrt_asset_material mat{};
            for (int k = 0; k < 4; ++k) mat.base_color[k] = 1.0f;
            mat.metallic = mat.roughness = 1.0f;
            bool ok = floats(m, "emissiveFactor", mat.emissive, 3);
            if (const Json* pbr = m.get("pbrMetallicRoughness")) {
                ok = ok && floats(*pbr, "baseColorFactor", mat.base_color, 4);
                const Json* mf = pbr->get("metallicFactor");
                const Json* rf = pbr->get("roughnessFactor");
                if (mf && mf->kind == Json::kNum) mat.metallic = float(mf->num);
                if (rf && rf->kind == Json::kNum) mat.roughness = float(rf->num);
            }
            if (!ok) return err(VKHRT_ERROR_INVALID_ARGUMENT, "material " + std::to_string(i) + ": malformed factors");
            out.materials.push_back(mat);
        }
        return VKHRT_SUCCESS;
    }

    int meshes(vkhrt_asset& out) {
        const Json* ms = arr(root_, "meshes");
        for (size_t mi = 0; ms && mi < ms->items.size(); ++mi) {
            const Json& m = ms->items[mi];
            const Json* prims = arr(m, "primitives");
            if (!prims) return err(VKHRT_ERROR_INVALID_ARGUMENT, "mesh " + std::to_string(mi) + " has no primitives");
            first_prim_.push_back(uint32_t(out.meshes.size()));
            const Json* mname = m.get("name");
            for (size_t pi = 0; pi < prims->items.size(); ++pi) {
                const std::string what = "mesh " + std::to_string(mi) + " primitive " + std::to_string(pi);
                out.meshes.emplace_back();
                vk::AssetMesh& M = out.meshes.back();
                if (mname && mname->kind == Json::kStr) M.name = mname->str;
                int rc = primitive(prims->items[pi], what, out, M);
                if (rc != VKHRT_SUCCESS) return rc;
            }
        }
        first_prim_.push_back(uint32_t(out.meshes.size()));
        return VKHRT_SUCCESS;
    }

    int primitive(const Json& p, const std::string& what, const vkhrt_asset& out, vk::AssetMesh& M) {
        uint64_t mode, mat;
        if (!uint_prop(p, "mode", mode, false, 4) || mode > 6) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bad mode");
        if (mode == 0) return err(VKHRT_ERROR_UNSUPPORTED, what + ": POINTS primitives are not read");
        M.primitive = mode <= 3 ? VKHRT_PRIMITIVE_LINES : VKHRT_PRIMITIVE_TRIANGLES;
        if (p.get("material")) {
            if (!uint_prop(p, "material", mat, true) || mat >= out.materials.size())
                return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": material index out of range");
            M.material = int32_t(mat);
        }
        const Json* attr = p.get("attributes");
        uint64_t ia;
        if (!attr || attr->kind != Json::kObj || !uint_prop(*attr, "POSITION", ia, true))
            return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": no POSITION attribute");
        View pos, nrm, uv, idx;
        int rc = accessor(ia, pos, what + " POSITION");
        if (rc != VKHRT_SUCCESS) return rc;
        if (pos.comps != 3 || pos.ctype != kFloat) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": POSITION is not VEC3 float");
        const uint32_t nv = pos.count;
        if (uint64_t(nv) * 8 > (1ull << 31)) return err(VKHRT_ERROR_UNSUPPORTED, what + ": too many vertices");
        bool has_n = false, has_uv = false;
        if (attr->get("NORMAL")) {
            if (!uint_prop(*attr, "NORMAL", ia, true)) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bad NORMAL");
            if ((rc = accessor(ia, nrm, what + " NORMAL")) != VKHRT_SUCCESS) return rc;
            if (nrm.comps != 3 || nrm.ctype != kFloat || nrm.count != nv)
                return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": NORMAL is not VEC3 float per vertex");
            has_n = true;
        }
        if (attr->get("TEXCOORD_0")) {
            if (!uint_prop(*attr, "TEXCOORD_0", ia, true)) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bad TEXCOORD_0");
            if ((rc = accessor(ia, uv, what + " TEXCOORD_0")) != VKHRT_SUCCESS) return rc;
            const bool ok_type = uv.ctype == kFloat || ((uv.ctype == kUByte || uv.ctype == kUShort) && uv.normalized);
            if (uv.comps != 2 || !ok_type || uv.count != nv)
                return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": TEXCOORD_0 is not VEC2 per vertex");
            has_uv = true;
        }
        // the vertex sequence the primitive's mode walks
        std::vector<uint32_t> seq;
        if (p.get("indices")) {
            if (!uint_prop(p, "indices", ia, true)) return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": bad indices");
            if ((rc = accessor(ia, idx, what + " indices")) != VKHRT_SUCCESS) return rc;
            if (idx.comps != 1 || !(idx.ctype == kUByte || idx.ctype == kUShort || idx.ctype == kUInt))
                return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": indices are not unsigned scalars");
            seq.resize(idx.count);
            for (uint32_t i = 0; i < idx.count; ++i) {
                seq[i] = idx.u(i);
                if (seq[i] >= nv)
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, what + ": index " + std::to_string(seq[i]) +
                                                                 " is not below the vertex count " + std::to_string(nv));
            }
        } else {
            seq.resize(nv);
            for (uint32_t i = 0; i < nv; ++i) seq[i] = i;
        }
        // primitives by mode (glTF 2.0 specification, "Meshes: Primitive topology")
        const size_t n = seq.size();
        std::vector<uint32_t>& I = M.indices;
        switch (mode) {
            case 1:  // LINES; an odd trailing index is dropped
                I.assign(seq.begin(), seq.begin() + (n / 2) * 2);
                break;
            case 2:  // LINE_LOOP
                if (n >= 2)
                    for (size_t i = 0; i < n; ++i) I.insert(I.end(), {seq[i], seq[(i + 1) % n]});
                break;
            case 3:  // LINE_STRIP
                for (size_t i = 0; i + 1 < n; ++i) I.insert(I.end(), {seq[i], seq[i + 1]});
                break;
            case 4:  // TRIANGLES; trailing indices short of a triangle are dropped
                I.assign(seq.begin(), seq.begin() + (n / 3) * 3);
                break;
            case 5:  // TRIANGLE_STRIP: {i, i + 1 + i % 2, i + 2 - i % 2}
                for (size_t i = 0; i + 2 < n; ++i) I.insert(I.end(), {seq[i], seq[i + 1 + i % 2], seq[i + 2 - i % 2]});
                break;
            case 6:  // TRIANGLE_FAN: {i + 1, i + 2, 0}
                for (size_t i = 0; i + 2 < n; ++i) I.insert(I.end(), {seq[i + 1], seq[i + 2], seq[0]});
                break;
        }
        M.vertices.assign(size_t(nv) * 8, 0.0f);
        for (uint32_t i = 0; i < nv; ++i) {
            float* v = &M.vertices[size_t(i) * 8];
            for (int c = 0; c < 3; ++c) v[c] = pos.f(i, c);
            if (has_n)
                for (int c = 0; c < 3; ++c) v[3 + c] = nrm.f(i, c);
            if (has_uv)
                for (int c = 0; c < 2; ++c) v[6 + c] = uv.f(i, c);
        }
        if (!has_n && M.primitive == VKHRT_PRIMITIVE_TRIANGLES) face_normals(M);
        // aiProcess_GenBoundingBoxes: the box of every vertex
        for (int c = 0; c < 3; ++c) {
            M.lo[c] = nv ? INFINITY : 0.0f;
            M.hi[c] = nv ? -INFINITY : 0.0f;
        }
        for (uint32_t i = 0; i < nv; ++i)
            for (int c = 0; c < 3; ++c) {
                M.lo[c] = std::fmin(M.lo[c], M.vertices[size_t(i) * 8 + c]);
                M.hi[c] = std::fmax(M.hi[c], M.vertices[size_t(i) * 8 + c]);
            }
        return VKHRT_SUCCESS;
    }

    // aiProcess_GenNormals on a mesh without normals: n = normalize((v1 - v0) x (v2 - v0)) per
    // face (left unnormalized when its length is 0), stored at each of the face's vertices, so a
    // vertex shared by several faces keeps the last one's.
    static void face_normals(vk::AssetMesh& M) {
        float* V = M.vertices.data();
        for (size_t f = 0; f + 2 < M.indices.size(); f += 3) {
            const float* a = V + size_t(M.indices[f]) * 8;
            const float* b = V + size_t(M.indices[f + 1]) * 8;
            const float* c = V + size_t(M.indices[f + 2]) * 8;
            const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (len > 0.0f)
                for (float& x : n) x /= len;
            for (int k = 0; k < 3; ++k) std::memcpy(V + size_t(M.indices[f + k]) * 8 + 3, n, sizeof(n));
        }
    }

    // node transform: `matrix` (column-major) or T * R * S with R from the unit quaternion
    // (x, y, z, w); composing with translation and scale matrices only moves or scales
    // entries, so the float result is [R * diag(s) | t]
    bool local_matrix(const Json& n, float m[16]) const {
        for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        if (n.get("matrix")) return floats(n, "matrix", m, 16);
        float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
        if (!floats(n, "translation", t, 3) || !floats(n, "rotation", q, 4) || !floats(n, "scale", s, 3)) return false;
        const float x = q[0], y = q[1], z = q[2], w = q[3];
        const float R[3][3] = {
            {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - z * w), 2.0f * (x * z + y * w)},
            {2.0f * (x * y + z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - x * w)},
            {2.0f * (x * z - y * w), 2.0f * (y * z + x * w), 1.0f - 2.0f * (x * x + y * y)},
        };
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) m[c * 4 + r] = R[r][c] * s[c];
        for (int r = 0; r < 3; ++r) m[12 + r] = t[r];
        return true;
    }

    int nodes(vkhrt_asset& out) {
        const Json* ns = arr(root_, "nodes");
        const size_t nn = ns ? ns->items.size() : 0;
        // children lists, each node owned by at most one parent
        std::vector<std::vector<uint32_t>> kids(nn);
        std::vector<int> parent(nn, -1);
        for (size_t i = 0; i < nn; ++i) {
            const Json* ch = arr(ns->items[i], "children");
            for (size_t k = 0; ch && k < ch->items.size(); ++k) {
                const Json& c = ch->items[k];
                if (c.kind != Json::kNum || !(c.num >= 0) || c.num != std::floor(c.num) || c.num >= double(nn))
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "node " + std::to_string(i) + ": bad child index");
                const uint32_t ci = uint32_t(c.num);
                if (parent[ci] >= 0 || ci == i)
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "node " + std::to_string(ci) + " has more than one parent");
                parent[ci] = int(i);
                kids[i].push_back(ci);
            }
        }
        // scene roots: the default scene's (`scene`, else scene 0), or every parentless node
        std::vector<uint32_t> roots;
        const Json* scenes = arr(root_, "scenes");
        if (scenes && !scenes->items.empty()) {
            uint64_t si;
            if (!uint_prop(root_, "scene", si, false, 0) || si >= scenes->items.size())
                return err(VKHRT_ERROR_INVALID_ARGUMENT, "bad default scene");
            const Json* sn = arr(scenes->items[size_t(si)], "nodes");
            for (size_t k = 0; sn && k < sn->items.size(); ++k) {
                const Json& c = sn->items[k];
                if (c.kind != Json::kNum || !(c.num >= 0) || c.num != std::floor(c.num) || c.num >= double(nn))
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "scene: bad node index");
                if (parent[size_t(c.num)] >= 0) return err(VKHRT_ERROR_INVALID_ARGUMENT, "scene: a root node has a parent");
                roots.push_back(uint32_t(c.num));
            }
        } else {
            for (size_t i = 0; i < nn; ++i)
                if (parent[i] < 0) roots.push_back(uint32_t(i));
        }
        // pre-order walk (ProcessNode, model_loader.cpp:208-246) with an explicit stack
        std::vector<uint8_t> seen(nn, 0);
        struct Item {
            uint32_t gltf;
            int32_t parent;
        };
        std::vector<Item> stack;
        if (roots.size() != 1) {
            vk::AssetNode R;
            R.name = "ROOT";
            for (int i = 0; i < 16; ++i) R.local[i] = (i % 5 == 0) ? 1.0f : 0.0f;
            out.nodes.push_back(R);
            for (size_t k = roots.size(); k-- > 0;) stack.push_back({roots[k], 0});
        } else {
            stack.push_back({roots[0], -1});
        }
        while (!stack.empty()) {
            const Item it = stack.back();
            stack.pop_back();
            if (seen[it.gltf]) return err(VKHRT_ERROR_INVALID_ARGUMENT, "node " + std::to_string(it.gltf) + " is reached twice");
            seen[it.gltf] = 1;
            const Json& n = ns->items[it.gltf];
            vk::AssetNode N;
            N.parent = it.parent;
            if (const Json* nm = n.get("name"))
                if (nm->kind == Json::kStr) N.name = nm->str;
            if (!local_matrix(n, N.local)) return err(VKHRT_ERROR_INVALID_ARGUMENT, "node " + std::to_string(it.gltf) + ": malformed transform");
            if (n.get("mesh")) {
                uint64_t mi;
                if (!uint_prop(n, "mesh", mi, true) || mi + 1 >= first_prim_.size())
                    return err(VKHRT_ERROR_INVALID_ARGUMENT, "node " + std::to_string(it.gltf) + ": mesh index out of range");
                for (uint32_t m = first_prim_[size_t(mi)]; m < first_prim_[size_t(mi) + 1]; ++m) N.meshes.push_back(m);
            }
            const int32_t self = int32_t(out.nodes.size());
            out.nodes.push_back(std::move(N));
            for (size_t k = kids[it.gltf].size(); k-- > 0;) stack.push_back({kids[it.gltf][k], self});
        }
        for (uint32_t i = 0; i < out.nodes.size(); ++i) vk::world_matrix(out.nodes, i, out.nodes[i].world);
        return VKHRT_SUCCESS;
    }

    struct Buf {
        const uint8_t* data = nullptr;
        size_t size = 0;
        std::vector<uint8_t> own;
    };

    std::string dir_;
    bool has_dir_;
    Json root_;
    const uint8_t* bin_ = nullptr;
    size_t bin_size_ = 0;
    std::vector<Buf> bufs_;
    std::vector<uint32_t> first_prim_;  // glTF mesh -> first of its primitives' meshes
};

int load_into(const uint8_t* data, size_t size, const std::string& dir, bool has_dir, vkhrt_asset** out) {
    std::unique_ptr<vkhrt_asset> a(new (std::nothrow) vkhrt_asset());
    if (!a) return asset_fail(VKHRT_ERROR_OUT_OF_MEMORY, "asset");
    try {
        Loader L(dir, has_dir);
        int rc = L.load(data, size, *a);
        if (rc != VKHRT_SUCCESS) return rc;
    } catch (const std::bad_alloc&) {
        return asset_fail(VKHRT_ERROR_OUT_OF_MEMORY, "out of host memory while loading the asset");
    }
    *out = a.release();
    return VKHRT_SUCCESS;
}

}  // namespace

namespace vk {

// glm's mat4 product with the float sums taken in column order, ((a0 + a1) + a2) + a3
static void mat4_mul(const float A[16], const float B[16], float C[16]) {
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r) {
            float s = A[0 * 4 + r] * B[j * 4 + 0];
            s = s + A[1 * 4 + r] * B[j * 4 + 1];
            s = s + A[2 * 4 + r] * B[j * 4 + 2];
            s = s + A[3 * 4 + r] * B[j * 4 + 3];
            C[j * 4 + r] = s;
        }
}

void world_matrix(const std::vector<AssetNode>& nodes, uint32_t i, float out[16]) {
    float m[16], t[16];
    std::memcpy(m, nodes[i].local, sizeof(m));
    for (int32_t p = nodes[i].parent; p >= 0; p = nodes[size_t(p)].parent) {
        mat4_mul(nodes[size_t(p)].local, m, t);
        std::memcpy(m, t, sizeof(m));
    }
    std::memcpy(out, m, sizeof(m));
}

}  // namespace vk

extern "C" {

int vkhrt_asset_load(const char* path, vkhrt_asset** out) {
    if (!path || !out) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::vector<uint8_t> data;
    if (!read_file(path, data)) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    std::string p(path);
    const size_t slash = p.find_last_of('/');
    return load_into(data.data(), data.size(), slash == std::string::npos ? std::string() : p.substr(0, slash), true, out);
}

int vkhrt_asset_load_memory(const void* data, size_t size, const char* base_dir, vkhrt_asset** out) {
    if (!out || (size && !data)) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    return load_into(static_cast<const uint8_t*>(data), size, base_dir ? base_dir : "", base_dir != nullptr, out);
}

void vkhrt_asset_destroy(vkhrt_asset* a) { delete a; }

const char* vkhrt_asset_last_error(void) { return g_asset_error.c_str(); }

int vkhrt_asset_info_get(const vkhrt_asset* a, vkhrt_asset_info* out) {
    if (!a || !out) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->mesh_count = uint32_t(a->meshes.size());
    out->node_count = uint32_t(a->nodes.size());
    out->material_count = uint32_t(a->materials.size());
    out->instance_count = uint32_t(a->instances.size());
    for (const vk::AssetMesh& m : a->meshes)
        (m.primitive == VKHRT_PRIMITIVE_LINES ? out->line_mesh_count : out->triangle_mesh_count)++;
    return VKHRT_SUCCESS;
}

int vkhrt_asset_mesh_get(const vkhrt_asset* a, uint32_t i, vkhrt_asset_mesh* out) {
    if (!a || !out) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (i >= a->meshes.size()) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "mesh index out of range");
    const vk::AssetMesh& m = a->meshes[i];
    out->primitive = m.primitive;
    out->material = m.material;
    out->vertex_count = uint32_t(m.vertices.size() / 8);
    out->index_count = uint32_t(m.indices.size());
    out->vertices = m.vertices.data();
    out->indices = m.indices.data();
    std::memcpy(out->bounds_min, m.lo, sizeof(m.lo));
    std::memcpy(out->bounds_max, m.hi, sizeof(m.hi));
    out->name = m.name.c_str();
    return VKHRT_SUCCESS;
}

int vkhrt_asset_node_get(const vkhrt_asset* a, uint32_t i, vkhrt_asset_node* out) {
    if (!a || !out) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (i >= a->nodes.size()) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "node index out of range");
    const vk::AssetNode& n = a->nodes[i];
    out->parent = n.parent;
    out->mesh_count = uint32_t(n.meshes.size());
    out->meshes = n.meshes.data();
    std::memcpy(out->local, n.local, sizeof(n.local));
    std::memcpy(out->world, n.world, sizeof(n.world));
    out->name = n.name.c_str();
    return VKHRT_SUCCESS;
}

int vkhrt_asset_material_get(const vkhrt_asset* a, uint32_t i, vkhrt_asset_material* out) {
    if (!a || !out) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (i >= a->materials.size()) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "material index out of range");
    *out = a->materials[i];
    return VKHRT_SUCCESS;
}

int vkhrt_asset_instance_get(const vkhrt_asset* a, uint32_t i, vkhrt_asset_instance* out) {
    if (!a || !out) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (i >= a->instances.size()) return asset_fail(VKHRT_ERROR_INVALID_ARGUMENT, "instance index out of range");
    const vk::AssetInstance& I = a->instances[i];
    out->node = I.node;
    out->mesh = I.mesh;
    const float* W = a->nodes[I.node].world;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) out->transform[r * 4 + c] = W[c * 4 + r];
    return VKHRT_SUCCESS;
}

}  // extern "C"
