// gltf.cpp -- see gltf.hpp.
#include "gltf.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "sol_math.hpp"

namespace sol {
namespace gltf {

// ---- JSON -------------------------------------------------------------------
const Json *Json::get(const std::string &key) const
{
    if (t != Obj) return nullptr;
    for (const auto &kv : o)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
double Json::num(const std::string &key, double def) const
{
    const Json *v = get(key);
    return (v && v->t == Num) ? v->n : def;
}
int64_t Json::integer(const std::string &key, int64_t def) const
{
    const Json *v = get(key);
    return (v && v->t == Num) ? (int64_t)v->n : def;
}
std::string Json::str(const std::string &key, const std::string &def) const
{
    const Json *v = get(key);
    return (v && v->t == Str) ? v->s : def;
}

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &text) : s(text) {}
    [[noreturn]] void fail(const char *what) const
    {
        throw std::runtime_error(std::string("glTF JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
    }
    bool eat(char c)
    {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    std::string string()
    {
        if (!eat('"')) fail("expected string");
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                char e = s[i++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (i + 4 > s.size()) fail("bad \\u");
                    unsigned cp = (unsigned)std::stoul(s.substr(i, 4), nullptr, 16);
                    i += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) {
                        out += (char)(0xC0 | (cp >> 6));
                        out += (char)(0x80 | (cp & 0x3F));
                    } else {
                        out += (char)(0xE0 | (cp >> 12));
                        out += (char)(0x80 | ((cp >> 6) & 0x3F));
                        out += (char)(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    Json value()
    {
        ws();
        if (i >= s.size()) fail("unexpected end");
        Json v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.t = Json::Obj;
            if (eat('}')) return v;
            do {
                std::string k = string();
                if (!eat(':')) fail("expected ':'");
                v.o.emplace_back(std::move(k), value());
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            ++i;
            v.t = Json::Arr;
            if (eat(']')) return v;
            do {
                v.a.push_back(value());
            } while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            v.t = Json::Str;
            v.s = string();
        } else if (s.compare(i, 4, "true") == 0) {
            i += 4;
            v.t = Json::Bool;
            v.b = true;
        } else if (s.compare(i, 5, "false") == 0) {
            i += 5;
            v.t = Json::Bool;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char *begin = s.c_str() + i;
            char *end = nullptr;
            v.n = std::strtod(begin, &end);
            if (end == begin) fail("bad value");
            i += (size_t)(end - begin);
            v.t = Json::Num;
        }
        return v;
    }
};

std::vector<uint8_t> base64_decode(const std::string &in)
{
    static int8_t T[256];
    static bool init = false;
    if (!init) {
        for (int k = 0; k < 256; ++k) T[k] = -1;
        const char *a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int k = 0; k < 64; ++k) T[(uint8_t)a[k]] = (int8_t)k;
        init = true;
    }
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (char ch : in) {
        const int8_t d = T[(uint8_t)ch];
        if (d < 0) continue;  // '=' padding and whitespace
        acc = (acc << 6) | (uint32_t)d;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)((acc >> bits) & 0xFF));
        }
    }
    return out;
}

std::vector<uint8_t> read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::string dirname_of(const std::string &path)
{
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

int comp_count(const std::string &type)
{
    if (type == "SCALAR") return 1;
    if (type == "VEC2") return 2;
    if (type == "VEC3") return 3;
    if (type == "VEC4") return 4;
    if (type == "MAT4") return 16;
    throw std::runtime_error("unsupported accessor type " + type);
}

int comp_size(int component)
{
    switch (component) {
    case 5120:
    case 5121: return 1;
    case 5122:
    case 5123: return 2;
    case 5125:
    case 5126: return 4;
    default: throw std::runtime_error("unsupported component type " + std::to_string(component));
    }
}

}  // namespace

Json parse_json(const std::string &text)
{
    Parser p(text);
    Json v = p.value();
    return v;
}

std::vector<float> Document::read_floats(int64_t ai) const
{
    if (ai < 0 || (size_t)ai >= accessors.size()) return {};
    const Accessor &A = accessors[(size_t)ai];
    std::vector<float> out(A.count * (uint64_t)A.ncomp, 0.0f);
    if (A.view < 0) return out;
    const View &V = views[(size_t)A.view];
    const std::vector<uint8_t> &B = buffers[(size_t)V.buffer];
    const uint64_t csz = (uint64_t)comp_size(A.component);
    const uint64_t stride = V.stride ? V.stride : csz * (uint64_t)A.ncomp;
    for (uint64_t e = 0; e < A.count; ++e)
        for (int k = 0; k < A.ncomp; ++k) {
            const uint64_t at = V.offset + A.offset + e * stride + (uint64_t)k * csz;
            if (at + csz > B.size()) throw std::runtime_error("accessor out of buffer range");
            const uint8_t *p = &B[at];
            float v = 0.0f;
            switch (A.component) {
            case 5126: std::memcpy(&v, p, 4); break;
            case 5121: v = A.normalized ? (float)p[0] / 255.0f : (float)p[0]; break;
            case 5123: {
                uint16_t u;
                std::memcpy(&u, p, 2);
                v = A.normalized ? (float)u / 65535.0f : (float)u;
                break;
            }
            case 5125: {
                uint32_t u;
                std::memcpy(&u, p, 4);
                v = (float)u;
                break;
            }
            case 5120: v = A.normalized ? std::fmax((float)(int8_t)p[0] / 127.0f, -1.0f) : (float)(int8_t)p[0]; break;
            case 5122: {
                int16_t q;
                std::memcpy(&q, p, 2);
                v = A.normalized ? std::fmax((float)q / 32767.0f, -1.0f) : (float)q;
                break;
            }
            }
            out[e * (uint64_t)A.ncomp + (uint64_t)k] = v;
        }
    return out;
}

std::vector<uint32_t> Document::read_indices(int64_t ai) const
{
    const Accessor &A = accessors.at((size_t)ai);
    std::vector<uint32_t> out(A.count, 0u);
    if (A.view < 0) return out;
    const View &V = views[(size_t)A.view];
    const std::vector<uint8_t> &B = buffers[(size_t)V.buffer];
    const uint64_t csz = (uint64_t)comp_size(A.component);
    const uint64_t stride = V.stride ? V.stride : csz;
    for (uint64_t e = 0; e < A.count; ++e) {
        const uint64_t at = V.offset + A.offset + e * stride;
        if (at + csz > B.size()) throw std::runtime_error("index accessor out of buffer range");
        uint32_t u = 0;
        if (csz == 1) u = B[at];
        else if (csz == 2) {
            uint16_t h;
            std::memcpy(&h, &B[at], 2);
            u = h;
        } else {
            std::memcpy(&u, &B[at], 4);
        }
        out[e] = u;
    }
    return out;
}

Document load(const std::string &path)
{
    std::vector<uint8_t> raw = read_file(path);
    std::string json;
    std::vector<uint8_t> glb_bin;
    bool is_glb = raw.size() >= 12 && std::memcmp(raw.data(), "glTF", 4) == 0;
    if (is_glb) {
        size_t off = 12;
        while (off + 8 <= raw.size()) {
            uint32_t len, type;
            std::memcpy(&len, &raw[off], 4);
            std::memcpy(&type, &raw[off + 4], 4);
            off += 8;
            if (off + len > raw.size()) throw std::runtime_error("truncated GLB chunk");
            if (type == 0x4E4F534Au) json.assign((const char *)&raw[off], len);
            else if (type == 0x004E4942u) glb_bin.assign(raw.begin() + (long)off, raw.begin() + (long)(off + len));
            off += len;
        }
    } else {
        json.assign(raw.begin(), raw.end());
    }
    const Json root = parse_json(json);
    Document d;
    const std::string dir = dirname_of(path);
    if (const Json *bufs = root.get("buffers")) {
        for (const Json &b : bufs->a) {
            const std::string uri = b.str("uri");
            if (uri.empty()) {
                d.buffers.push_back(glb_bin);
            } else if (uri.rfind("data:", 0) == 0) {
                const size_t comma = uri.find(',');
                if (comma == std::string::npos) throw std::runtime_error("bad data URI");
                d.buffers.push_back(base64_decode(uri.substr(comma + 1)));
            } else {
                d.buffers.push_back(read_file(dir + uri));
            }
        }
    }
    if (const Json *vs = root.get("bufferViews")) {
        for (const Json &v : vs->a) {
            View V;
            V.buffer = v.integer("buffer", 0);
            V.offset = (uint64_t)v.integer("byteOffset", 0);
            V.length = (uint64_t)v.integer("byteLength", 0);
            V.stride = (uint64_t)v.integer("byteStride", 0);
            if (V.buffer < 0 || (size_t)V.buffer >= d.buffers.size()) throw std::runtime_error("bad bufferView");
            d.views.push_back(V);
        }
    }
    if (const Json *as = root.get("accessors")) {
        for (const Json &a : as->a) {
            Accessor A;
            A.view = a.integer("bufferView", -1);
            A.offset = (uint64_t)a.integer("byteOffset", 0);
            A.component = (int)a.integer("componentType", 5126);
            A.count = (uint64_t)a.integer("count", 0);
            A.ncomp = comp_count(a.str("type", "SCALAR"));
            const Json *nz = a.get("normalized");
            A.normalized = nz && nz->t == Json::Bool && nz->b;
            if (A.view >= (int64_t)d.views.size()) throw std::runtime_error("bad accessor view");
            d.accessors.push_back(A);
        }
    }
    if (const Json *ms = root.get("meshes")) {
        for (const Json &m : ms->a) {
            std::vector<Primitive> prims;
            if (const Json *ps = m.get("primitives")) {
                for (const Json &p : ps->a) {
                    Primitive P;
                    if (const Json *at = p.get("attributes")) {
                        P.position = at->integer("POSITION", -1);
                        P.normal = at->integer("NORMAL", -1);
                        P.texcoord0 = at->integer("TEXCOORD_0", -1);
                        P.color0 = at->integer("COLOR_0", -1);
                    }
                    P.indices = p.integer("indices", -1);
                    P.material = p.integer("material", -1);
                    prims.push_back(P);
                }
            }
            d.meshes.push_back(prims);
            d.mesh_names.push_back(m.str("name"));
        }
    }
    if (const Json *mats = root.get("materials")) {
        for (const Json &m : mats->a) {
            Material M;
            if (const Json *pbr = m.get("pbrMetallicRoughness")) {
                if (const Json *bc = pbr->get("baseColorFactor"))
                    for (size_t k = 0; k < 4 && k < bc->a.size(); ++k) M.base_color[k] = (float)bc->a[k].n;
                M.metallic = (float)pbr->num("metallicFactor", 1.0);
                M.roughness = (float)pbr->num("roughnessFactor", 1.0);
            }
            if (const Json *e = m.get("emissiveFactor"))
                for (size_t k = 0; k < 3 && k < e->a.size(); ++k) M.emissive[k] = (float)e->a[k].n;
            d.materials.push_back(M);
        }
    }
    if (const Json *ns = root.get("nodes")) {
        for (const Json &n : ns->a) {
            Node N;
            N.mesh = n.integer("mesh", -1);
            N.camera = n.integer("camera", -1);
            if (const Json *ch = n.get("children"))
                for (const Json &c : ch->a) N.children.push_back((int64_t)c.n);
            Mat4 local;
            if (const Json *mx = n.get("matrix")) {
                float a[16];
                for (int k = 0; k < 16; ++k) a[k] = (size_t)k < mx->a.size() ? (float)mx->a[(size_t)k].n : 0.0f;
                local = Mat4::from_cols_array(a);
            } else {
                Vec3 t{0, 0, 0}, s{1, 1, 1};
                float q[4] = {0, 0, 0, 1};
                if (const Json *tr = n.get("translation")) t = {(float)tr->a[0].n, (float)tr->a[1].n, (float)tr->a[2].n};
                if (const Json *sc = n.get("scale")) s = {(float)sc->a[0].n, (float)sc->a[1].n, (float)sc->a[2].n};
                if (const Json *r = n.get("rotation"))
                    for (int k = 0; k < 4; ++k) q[k] = (float)r->a[(size_t)k].n;
                local = Mat4::from_translation(t) * Mat4::from_quat(q[0], q[1], q[2], q[3]) * Mat4::from_scale(s);
            }
            std::memcpy(N.matrix, local.data(), sizeof(N.matrix));
            d.nodes.push_back(N);
        }
    }
    if (const Json *cs = root.get("cameras")) {
        for (const Json &c : cs->a) {
            Camera C;
            if (c.str("type") == "perspective") {
                if (const Json *p = c.get("perspective")) {
                    C.perspective = true;
                    C.yfov = (float)p->num("yfov", 0.0);
                    C.znear = (float)p->num("znear", 0.01);
                    C.has_zfar = p->get("zfar") != nullptr;
                    C.zfar = (float)p->num("zfar", 100.0);
                }
            }
            d.cameras.push_back(C);
        }
    }
    return d;
}

}  // namespace gltf
}  // namespace sol
