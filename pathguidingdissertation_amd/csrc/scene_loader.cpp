// scene_loader.cpp — the reference's scene files on the host: scene.json (camera,
// instances, environment map), .gem static meshes, PNG and Radiance HDR textures;
// and the image files the renderer writes (RayTracer::savePNG / saveHDR).
//
// Behaviour follows SceneLoader.h:104-269 (loadInstance / loadScene) with the
// semantics of the reference's GEMLoader.h (JSON values and properties:
// GEMLoader.h:39-147, 524-894; mesh files: GEMLoader.h:268-458) and of the image
// decoding Texture::load relies on (Imaging.h:47-80). The reference links
// stb_image v2.30 for PNG/HDR decoding; that third-party algorithm is restated here
// from the PNG (RFC 2083) and Radiance RGBE formats, with stb's conventions: the
// file's own channel count, 8-bit samples, RGBE mantissa * 2^(e - 136).
// Every arithmetic step that shapes triangles (instance transform, normal matrix,
// normalisation, Triangle::init area) is FP32 in the reference's operation order.
#include <zlib.h>

#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "host_prep.hpp"
#include "host_scene.hpp"

namespace pg {
namespace {

// ------------------------------------------------------------------ JSON (GEMLoader.h:524-804)
// The reference's parser: no string escapes (a string ends at the next quote),
// numbers converted with stof, objects keyed in sorted order, any separator.
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    float num = 0.f;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    // GEMJson::asStr: std::to_string of numbers and booleans, "" for the rest
    std::string as_str() const {
        if (kind == BOOL) return b ? "1" : "0";
        if (kind == NUM) return std::to_string(num);
        if (kind == STR) return str;
        return "";
    }
};

struct JParser {
    const std::string& s;
    size_t pos = 0;
    std::string err;
    explicit JParser(const std::string& t) : s(t) {}
    char peek() const { return pos < s.size() ? s[pos] : 0; }
    char get() { return pos < s.size() ? s[pos++] : (pos++, 0); }
    void ws() { while (pos < s.size() && isspace((unsigned char)s[pos])) ++pos; }
    JVal value(int depth) {
        JVal v;
        if (depth > 64) { err = "JSON nested too deeply"; return v; }
        ws();
        char c = peek();
        if (c == 'n') { pos += 4; return v; }
        if (c == 't' || c == 'f') { v.kind = JVal::BOOL; v.b = c == 't'; pos += (c == 't') ? 4 : 5; return v; }
        if (c == '-' || isdigit((unsigned char)c)) return number();
        if (c == '"') return string();
        if (c == '[') {
            get(); ws();
            v.kind = JVal::ARR;
            if (peek() == ']') { get(); return v; }
            while (err.empty()) {
                if (pos >= s.size()) { err = "unterminated JSON array"; break; }
                v.arr.push_back(value(depth + 1));
                ws();
                if (get() == ']') break;
                ws();
            }
            return v;
        }
        if (c == '{') {
            get(); ws();
            v.kind = JVal::OBJ;
            if (peek() == '}') { get(); return v; }
            while (err.empty()) {
                if (pos >= s.size()) { err = "unterminated JSON object"; break; }
                ws();
                JVal k = string();
                ws();
                get();  // ':'
                v.obj[k.str] = value(depth + 1);
                ws();
                if (get() == '}') break;
                ws();
            }
            return v;
        }
        return v;
    }
    JVal number() {
        size_t start = pos;
        if (peek() == '-') get();
        if (peek() == '0') get();
        else while (isdigit((unsigned char)peek())) get();
        if (peek() == '.') { get(); while (isdigit((unsigned char)peek())) get(); }
        if (peek() == 'e' || peek() == 'E') {
            get();
            if (peek() == '+' || peek() == '-') get();
            while (isdigit((unsigned char)peek())) get();
        }
        JVal v;
        v.kind = JVal::NUM;
        std::string tok = s.substr(start, pos - start);
        char* end = nullptr;
        errno = 0;
        v.num = strtof(tok.c_str(), &end);
        if (end == tok.c_str() || errno == ERANGE) err = "bad JSON number '" + tok + "'";
        return v;
    }
    JVal string() {
        JVal v;
        v.kind = JVal::STR;
        if (get() != '"') { err = "expected a JSON string"; return v; }
        while (true) {
            if (pos >= s.size()) { err = "unterminated JSON string"; break; }
            char c = get();
            if (c == '"') break;
            v.str.push_back(c);
        }
        return v;
    }
};

// GEMProperty::getValue / getValuesAsVector3 (GEMLoader.h:58-145): std::stof / std::stoi
// of the string, the default when it is empty or does not convert.
bool parse_float(const std::string& v, float& out) {
    const char* p = v.c_str();
    char* end = nullptr;
    errno = 0;
    float f = strtof(p, &end);
    if (end == p || errno == ERANGE) return false;
    out = f;
    return true;
}
float prop_float(const std::map<std::string, std::string>& m, const char* k, float dflt) {
    auto it = m.find(k);
    if (it == m.end() || it->second.empty()) return dflt;
    float f;
    return parse_float(it->second, f) ? f : dflt;
}
int prop_int(const std::map<std::string, std::string>& m, const char* k, int dflt) {
    auto it = m.find(k);
    if (it == m.end() || it->second.empty()) return dflt;
    const char* p = it->second.c_str();
    char* end = nullptr;
    errno = 0;
    long v = strtol(p, &end, 10);
    if (end == p || errno == ERANGE || v > INT32_MAX || v < INT32_MIN) return dflt;
    return (int)v;
}
std::string prop_str(const std::map<std::string, std::string>& m, const char* k) {
    auto it = m.find(k);
    return it == m.end() ? std::string() : it->second;
}
void prop_vec3(const std::map<std::string, std::string>& m, const char* k, float* out) {
    std::vector<float> vals;
    std::string v = prop_str(m, k);
    std::stringstream ss(v);
    std::string word;
    while (std::getline(ss, word, ' ')) {
        float f = 0.f;
        if (word.empty() || !parse_float(word, f)) f = 0.f;
        vals.push_back(f);
    }
    while (vals.size() < 3) vals.push_back(0.f);
    out[0] = vals[0]; out[1] = vals[1]; out[2] = vals[2];
}

bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// ------------------------------------------------------------------ .gem (GEMLoader.h:268-458)
constexpr uint32_t kGemMagic = 4058972161u;
struct GemVertex { float p[3], n[3], t[3], u, v; };  // GEMStaticVertex, 44 bytes
static_assert(sizeof(GemVertex) == 44, "GEM static vertex");
struct GemMesh { std::vector<GemVertex> verts; std::vector<uint32_t> idx; };

struct Reader {
    const std::string& b;
    size_t pos = 0;
    bool bad = false;
    explicit Reader(const std::string& s) : b(s) {}
    bool take(void* dst, size_t n) {
        if (bad || pos + n > b.size()) { bad = true; return false; }
        std::memcpy(dst, b.data() + pos, n);
        pos += n;
        return true;
    }
    uint32_t u32() { uint32_t v = 0; take(&v, 4); return v; }
    std::string str() {
        int32_t l = 0;
        take(&l, 4);
        if (l < 0 || bad) { bad = true; return std::string(); }
        std::string s(l, '\0');
        if (l) take(&s[0], (size_t)l);
        return std::string(s.c_str());  // loadString stops at an embedded NUL
    }
};

std::string load_gem(const std::string& path, std::vector<GemMesh>& meshes) {
    std::string buf;
    if (!read_file(path, buf)) return "cannot read mesh file " + path;
    Reader r(buf);
    if (r.u32() != kGemMagic) return path + " is not a GE Model File";
    uint32_t animated = r.u32();
    if (animated) return path + ": animated GEM meshes carry no static vertices (SceneLoader.h:208-226 reads only those)";
    uint32_t n = r.u32();
    for (uint32_t i = 0; i < n && !r.bad; ++i) {
        GemMesh m;
        uint32_t np = r.u32();
        for (uint32_t k = 0; k < np && !r.bad; ++k) { r.str(); r.str(); }  // material properties (unused)
        uint32_t nv = r.u32();
        if (r.bad || (uint64_t)nv * sizeof(GemVertex) > buf.size()) return path + ": truncated mesh";
        m.verts.resize(nv);
        if (nv) r.take(m.verts.data(), (size_t)nv * sizeof(GemVertex));
        uint32_t ni = r.u32();
        if (r.bad || (uint64_t)ni * 4 > buf.size()) return path + ": truncated mesh";
        m.idx.resize(ni);
        if (ni) r.take(m.idx.data(), (size_t)ni * 4);
        meshes.push_back(std::move(m));
    }
    if (r.bad) return path + ": truncated mesh";
    return std::string();
}

inline HV3 cross(const HV3& a, const HV3& b) {                                 // MyMath.h:134-140
    return HV3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
inline float length(const HV3& a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }  // MyMath.h:125
inline HV3 normalized(const HV3& a) { float l = 1.f / length(a); return HV3(a.x * l, a.y * l, a.z * l); }

// ------------------------------------------------------------------ textures (Imaging.h:22-80)
struct TexCache {
    std::map<std::string, int32_t> index;
};

// Texture::load: ".hdr" in the name selects the float path; otherwise 8-bit PNG or
// JPEG samples times 1/255. A file that does not decode becomes the 1x1 white default.
int32_t add_texture(pg_host_scene* s, TexCache& cache, const std::string& path) {
    auto it = cache.index.find(path);
    if (it != cache.index.end()) return it->second;
    Image img;
    bool hdr = false;
    std::string e = read_image_file(path, img, hdr);
    int32_t id = (int32_t)s->tex_w.size();
    s->tex_offset.push_back(s->texels.size() / 3);
    s->tex_path.push_back(path);
    if (!e.empty() || img.w <= 0 || img.h <= 0) {
        s->warnings.push_back("texture " + path + ": " + (e.empty() ? "empty" : e) + " (white default)");
        s->tex_w.push_back(1); s->tex_h.push_back(1);
        s->texels.insert(s->texels.end(), {1.f, 1.f, 1.f});
    } else {
        size_t n = (size_t)img.w * img.h, ch = (size_t)img.channels;
        s->tex_w.push_back(img.w); s->tex_h.push_back(img.h);
        size_t base = s->texels.size();
        s->texels.resize(base + 3 * n);
        float* o = s->texels.data() + base;
        const float scale = 1.f / 255.f;                                       // Imaging.h:20
        // texel i reads samples i*ch .. i*ch+2 as the reference does (for 1- and
        // 2-channel files that spills into the next pixel; past the end reads 0)
        size_t total = n * ch;
        for (size_t i = 0; i < n; ++i)
            for (size_t c = 0; c < 3; ++c) {
                size_t k = i * ch + c;
                if (hdr) o[3 * i + c] = k < total ? img.f32[k] : 0.f;
                else o[3 * i + c] = k < total ? (float)img.u8[k] * scale : 0.f;
            }
    }
    cache.index[path] = id;
    return id;
}

}  // namespace

// ------------------------------------------------------------------ PNG (RFC 2083)
namespace {
inline uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
inline int paeth(int a, int b, int c) {
    int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}
}  // namespace

std::string read_png(const std::string& path, Image& out) {
    std::string buf;
    if (!read_file(path, buf)) return "cannot read";
    const uint8_t* d = (const uint8_t*)buf.data();
    size_t n = buf.size();
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) return "not a PNG file";
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    while (pos + 12 <= n) {
        uint32_t len = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        if (len > n - pos - 12) return "truncated chunk";
        const uint8_t* body = d + pos + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return "bad IHDR";
            w = be32(body); h = be32(body + 4); depth = body[8]; ctype = body[9]; interlace = body[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24)) return "bad image size";
    if ((int64_t)w * h > kMaxImagePixels) return "PNG image above 2^28 pixels";
    if (depth != 8) return "only 8-bit PNG samples are supported";
    if (interlace) return "interlaced PNG is not supported";
    int comps;
    switch (ctype) {
    case 0: comps = 1; break;
    case 2: comps = 3; break;
    case 3: comps = 1; break;
    case 4: comps = 2; break;
    case 6: comps = 4; break;
    default: return "unknown PNG colour type";
    }
    size_t stride = (size_t)w * comps;
    size_t raw_n = (stride + 1) * h;
    std::vector<uint8_t> raw(raw_n);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return "zlib init failed";
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    int zr = inflate(&zs, Z_FINISH);
    size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || got != raw_n) return "corrupt image data";
    std::vector<uint8_t> px(stride * h);
    int bpp = comps;
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* src = raw.data() + y * (stride + 1);
        uint8_t ft = src[0];
        ++src;
        uint8_t* dst = px.data() + y * stride;
        const uint8_t* prv = y ? px.data() + (y - 1) * stride : nullptr;
        for (size_t x = 0; x < stride; ++x) {
            int a = x >= (size_t)bpp ? dst[x - bpp] : 0;
            int b = prv ? prv[x] : 0;
            int c = (prv && x >= (size_t)bpp) ? prv[x - bpp] : 0;
            int v = src[x];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: return "bad PNG filter";
            }
            dst[x] = (uint8_t)v;
        }
    }
    out.w = (int32_t)w; out.h = (int32_t)h;
    size_t npx = (size_t)w * h;
    if (ctype == 3) {
        // palette: 3 channels, 4 when tRNS holds alphas
        if (plte.size() < 3) return "missing palette";
        int oc = trns.empty() ? 3 : 4;
        out.channels = oc;
        out.u8.resize(npx * oc);
        size_t np = plte.size() / 3;
        for (size_t i = 0; i < npx; ++i) {
            size_t k = px[i];
            if (k >= np) return "palette index out of range";
            for (int c = 0; c < 3; ++c) out.u8[i * oc + c] = plte[3 * k + c];
            if (oc == 4) out.u8[i * oc + 3] = k < trns.size() ? trns[k] : 255;
        }
    } else if (!trns.empty() && (ctype == 0 || ctype == 2)) {
        // a tRNS colour key adds an alpha channel
        int oc = comps + 1;
        out.channels = oc;
        out.u8.resize(npx * oc);
        for (size_t i = 0; i < npx; ++i) {
            bool key = true;
            for (int c = 0; c < comps; ++c) {
                uint8_t v = px[i * comps + c];
                out.u8[i * oc + c] = v;
                size_t tk = 2 * (size_t)c + 1;
                if (tk >= trns.size() || trns[tk] != v) key = false;
            }
            out.u8[i * oc + comps] = key ? 0 : 255;
        }
    } else {
        out.channels = comps;
        out.u8 = std::move(px);
    }
    return std::string();
}

std::string read_image_file(const std::string& path, Image& out, bool& hdr) {
    hdr = path.find(".hdr") != std::string::npos;
    if (hdr) return read_hdr(path, out);
    uint8_t sig[8] = {0};
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return "cannot read";
    size_t got = std::fread(sig, 1, 8, f);
    std::fclose(f);
    if (got >= 8 && sig[0] == 137 && sig[1] == 'P' && sig[2] == 'N' && sig[3] == 'G') return read_png(path, out);
    if (got >= 3 && sig[0] == 0xFF && sig[1] == 0xD8 && sig[2] == 0xFF) return read_jpeg(path, out);
    return "unsupported image format (PNG, JPEG and Radiance HDR are decoded)";
}

// ------------------------------------------------------------------ Radiance HDR
std::string read_hdr(const std::string& path, Image& out) {
    std::string buf;
    if (!read_file(path, buf)) return "cannot read";
    size_t pos = 0;
    auto line = [&](std::string& l) {
        l.clear();
        while (pos < buf.size() && buf[pos] != '\n') l.push_back(buf[pos++]);
        if (pos < buf.size()) ++pos;
    };
    std::string l;
    line(l);
    if (l != "#?RADIANCE" && l != "#?RGBE") return "not a Radiance file";
    bool fmt = false;
    while (true) {
        if (pos >= buf.size()) return "truncated header";
        line(l);
        if (l.empty()) break;
        if (l == "FORMAT=32-bit_rle_rgbe") fmt = true;
    }
    if (!fmt) return "unsupported HDR format";
    line(l);
    int H = 0, W = 0;
    if (std::sscanf(l.c_str(), "-Y %d +X %d", &H, &W) != 2 || H <= 0 || W <= 0) return "unsupported HDR orientation";
    if (W > (1 << 24) || H > (1 << 24) || (int64_t)W * H > kMaxImagePixels) return "HDR image above 2^28 pixels";
    out.w = W; out.h = H; out.channels = 3;
    out.f32.assign((size_t)W * H * 3, 0.f);
    const uint8_t* d = (const uint8_t*)buf.data();
    size_t n = buf.size();
    auto convert = [&](size_t px, const uint8_t* rgbe) {
        float* o = out.f32.data() + 3 * px;
        if (rgbe[3] != 0) {
            float f = ldexpf(1.f, (int)rgbe[3] - (128 + 8));
            o[0] = rgbe[0] * f; o[1] = rgbe[1] * f; o[2] = rgbe[2] * f;
        } else {
            o[0] = o[1] = o[2] = 0.f;
        }
    };
    auto flat = [&](size_t first_px) -> std::string {
        for (size_t i = first_px; i < (size_t)W * H; ++i) {
            if (pos + 4 > n) return "truncated pixels";
            convert(i, d + pos);
            pos += 4;
        }
        return std::string();
    };
    if (W < 8 || W >= 32768) return flat(0);
    std::vector<uint8_t> sl((size_t)W * 4);
    for (int y = 0; y < H; ++y) {
        if (pos + 4 > n) return "truncated pixels";
        const uint8_t* q = d + pos;
        int len = ((int)q[2] << 8) | q[3];
        if (q[0] != 2 || q[1] != 2 || (q[2] & 0x80)) {
            if (y != 0) return "mixed HDR scanline encodings";
            return flat(0);  // not run-length encoded: the whole image is flat
        }
        if (len != W) return "bad HDR scanline length";
        pos += 4;
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < W) {
                if (pos >= n) return "truncated pixels";
                int cnt = d[pos++];
                if (cnt > 128) {
                    cnt -= 128;
                    if (pos >= n || x + cnt > W) return "bad HDR run";
                    uint8_t v = d[pos++];
                    for (int k = 0; k < cnt; ++k) sl[(size_t)(x++) * 4 + c] = v;
                } else {
                    if (cnt == 0 || pos + cnt > n || x + cnt > W) return "bad HDR run";
                    for (int k = 0; k < cnt; ++k) sl[(size_t)(x++) * 4 + c] = d[pos++];
                }
            }
        }
        for (int x = 0; x < W; ++x) convert((size_t)y * W + x, &sl[(size_t)x * 4]);
    }
    return std::string();
}

// ------------------------------------------------------------------ loadScene
namespace {
std::string load_scene(const std::string& dir, uint32_t flags, pg_host_scene* s) {
    std::string text;
    if (!read_file(dir + "/scene.json", text)) return "cannot read " + dir + "/scene.json";
    JParser jp(text);
    JVal root = jp.value(0);
    if (!jp.err.empty()) return "scene.json: " + jp.err;
    if (root.kind != JVal::OBJ) return "scene.json: top level is not an object";
    // GEMScene::load (GEMLoader.h:857-878): non-array entries are scene properties;
    // every array holds instances
    std::map<std::string, std::string> props;
    std::vector<const JVal*> instances;
    for (const auto& kv : root.obj) {
        if (kv.second.kind != JVal::ARR) {
            if (!props.count(kv.first)) props[kv.first] = kv.second.as_str();
        } else {
            for (const JVal& v : kv.second.arr) instances.push_back(&v);
        }
    }
    // camera (SceneLoader.h:222-244)
    int width = prop_int(props, "width", 1920), height = prop_int(props, "height", 1080);
    float fov = prop_float(props, "fov", 45.0f);
    if (width < 1 || height < 1) return "scene.json: width and height must be >= 1";
    prop_vec3(props, "from", s->cam.from);
    prop_vec3(props, "to", s->cam.to);
    prop_vec3(props, "up", s->cam.up);
    s->cam.fov_deg = fov;
    s->cam.width = (float)width;
    s->cam.height = (float)height;
    if (prop_int(props, "flipX", 0) == 1) return "scene.json: flipX cameras are not supported";
    TexCache cache;
    for (const JVal* inst : instances) {
        // GEMScene::parseInstance (GEMLoader.h:822-847)
        std::string mesh_file;
        float world[16];
        for (int i = 0; i < 16; ++i) world[i] = 0.f;
        std::map<std::string, std::string> mp;
        if (inst->kind == JVal::OBJ) {
            for (const auto& kv : inst->obj) {
                bool core = false;
                if (kv.first == "filename") { mesh_file = kv.second.as_str(); core = true; }
                if (kv.first == "world") {
                    if (kv.second.arr.size() < 16) return "scene.json: instance world matrix needs 16 numbers";
                    for (int i = 0; i < 16; ++i) world[i] = kv.second.arr[i].num;
                    core = true;
                }
                if (!core) mp[kv.first] = kv.second.as_str();
            }
        }
        // material (SceneLoader.h:112-181); one per instance, in instance order
        std::string bsdf = prop_str(mp, "bsdf");
        pg_material m;
        std::memset(&m, 0, sizeof(m));
        bool known = true;
        if (bsdf == "diffuse") m.type = PG_BSDF_DIFFUSE;
        else if (bsdf == "orennayar") { m.type = PG_BSDF_ORENNAYAR; m.roughness = prop_float(mp, "alpha", 1.0f); }
        else if (bsdf == "glass") {
            m.type = PG_BSDF_GLASS;
            m.int_ior = prop_float(mp, "intIOR", 1.33f); m.ext_ior = prop_float(mp, "extIOR", 1.0f);
        } else if (bsdf == "mirror") m.type = PG_BSDF_MIRROR;
        else if (bsdf == "plastic") {
            m.type = PG_BSDF_PLASTIC;
            m.int_ior = prop_float(mp, "intIOR", 1.33f); m.ext_ior = prop_float(mp, "extIOR", 1.0f);
            m.roughness = prop_float(mp, "roughness", 1.0f);
        } else if (bsdf == "dielectric") {
            m.int_ior = prop_float(mp, "intIOR", 1.33f); m.ext_ior = prop_float(mp, "extIOR", 1.0f);
            float r = prop_float(mp, "roughness", 1.0f);
            if (r < 0.001f) m.type = PG_BSDF_GLASS;
            else { m.type = PG_BSDF_DIELECTRIC; m.roughness = r; }
        } else if (bsdf == "conductor") {
            m.type = PG_BSDF_CONDUCTOR;
            prop_vec3(mp, "eta", m.eta);
            prop_vec3(mp, "k", m.k);
            m.roughness = prop_float(mp, "roughness", 1.0f);
        } else {
            known = false;
        }
        if (!known) {  // "Error in loading": the instance is skipped (SceneLoader.h:182-186)
            s->warnings.push_back("instance " + mesh_file + ": unknown bsdf '" + bsdf + "' skipped");
            continue;
        }
        std::vector<GemMesh> meshes;
        std::string e = load_gem(dir + "/" + mesh_file, meshes);
        if (!e.empty()) {
            if ((flags & PG_LOAD_SKIP_MISSING) != 0) {
                s->warnings.push_back(e + " (instance skipped)");
                continue;
            }
            return e;
        }
        if (!prop_str(mp, "emission").empty()) prop_vec3(mp, "emission", m.emission);
        if (prop_int(mp, "coatingThickness", 0) > 0) m.type += PG_BSDF_LAYERED;  // LayeredBSDF (SceneLoader.h:173-181)
        int32_t tex = add_texture(s, cache, dir + "/" + prop_str(mp, "reflectance"));
        const float* t0 = s->texels.data() + 3 * s->tex_offset[tex];
        m.albedo[0] = t0[0]; m.albedo[1] = t0[1]; m.albedo[2] = t0[2];
        s->mats.push_back(m);
        s->mat_tex.push_back(tex);
        uint32_t mat_index = (uint32_t)s->mats.size() - 1;
        // geometry (SceneLoader.h:188-219): points by the instance matrix, normals by
        // its inverse transpose then normalised; indices of later meshes are offset by
        // the number of indices read so far, as the reference does
        M4 T;
        std::memcpy(T.m, world, sizeof(world));
        M4 N = T.inverse().transpose();
        std::vector<HV3> vp, vn;
        std::vector<float> vu, vv;
        std::vector<uint32_t> idx;
        for (const GemMesh& gm : meshes) {
            for (const GemVertex& gv : gm.verts) {
                vp.push_back(T.mul_point(HV3(gv.p[0], gv.p[1], gv.p[2])));
                vn.push_back(normalized(N.mul_vec(HV3(gv.n[0], gv.n[1], gv.n[2]))));
                vu.push_back(gv.u); vv.push_back(gv.v);
            }
            uint32_t offset = (uint32_t)idx.size();
            for (uint32_t k : gm.idx) idx.push_back(offset + k);
        }
        for (size_t i = 0; i + 2 < idx.size(); i += 3) {
            uint32_t a = idx[i], b = idx[i + 1], c = idx[i + 2];
            if (a >= vp.size() || b >= vp.size() || c >= vp.size()) return mesh_file + ": vertex index out of range";
            HV3 e1 = vp[c] - vp[b], e2 = vp[a] - vp[c];                          // Triangle::init, Geometry.h:86-93
            float area = length(cross(e1, e2)) * 0.5f;
            if (!(area > 0.f)) continue;
            const uint32_t v3[3] = {a, b, c};
            for (int k = 0; k < 3; ++k) {
                const HV3& p = vp[v3[k]];
                const HV3& q = vn[v3[k]];
                s->pos.insert(s->pos.end(), {p.x, p.y, p.z});
                s->nrm.insert(s->nrm.end(), {q.x, q.y, q.z});
                s->uv.insert(s->uv.end(), {vu[v3[k]], vv[v3[k]]});
            }
            s->mat.push_back(mat_index);
        }
    }
    // environment map (SceneLoader.h:252-258)
    std::string envf = prop_str(props, "envmap");
    if (!envf.empty()) {
        TexCache env_cache;
        pg_host_scene tmp;
        add_texture(&tmp, env_cache, dir + "/" + envf);
        s->env_w = tmp.tex_w[0]; s->env_h = tmp.tex_h[0];
        s->env = std::move(tmp.texels);
        for (auto& w : tmp.warnings) s->warnings.push_back("envmap: " + w);
    }
    if (s->mat.empty()) return "scene has no triangles";
    return std::string();
}
}  // namespace
}  // namespace pg

static std::string g_load_error;

static pg_status write_png(const char* path, const uint8_t* rgb, int32_t width, int32_t height);
static pg_status write_hdr(const char* path, const float* rgb, int32_t width, int32_t height);

// host allocation failures of the writers become status codes, never exceptions
template <class F>
static pg_status guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        g_load_error = "out of host memory";
        return PG_ERR_NOMEM;
    } catch (const std::exception& x) {
        g_load_error = x.what();
        return PG_ERR_INVALID;
    }
}

extern "C" {


pg_status pg_scene_load(const char* scene_dir, uint32_t flags, pg_host_scene** out) {
    if (!scene_dir || !out) return PG_ERR_INVALID;
    *out = nullptr;
    pg_host_scene* s = new pg_host_scene();
    std::string e;
    try {
        e = pg::load_scene(scene_dir, flags, s);
    } catch (const std::bad_alloc&) {
        delete s;
        g_load_error = "out of host memory";
        return PG_ERR_NOMEM;
    } catch (const std::exception& x) {
        delete s;
        g_load_error = std::string("scene load failed: ") + x.what();
        return PG_ERR_INVALID;
    }
    if (!e.empty()) {
        delete s;
        g_load_error = e;
        return PG_ERR_INVALID;
    }
    g_load_error.clear();
    *out = s;
    return PG_OK;
}

const char* pg_scene_load_error(void) { return g_load_error.c_str(); }

pg_status pg_scene_texture_sizes(const pg_host_scene* s, uint32_t* n_textures, uint64_t* n_texels,
                                 uint32_t* n_warnings) {
    if (!s) return PG_ERR_INVALID;
    if (n_textures) *n_textures = (uint32_t)s->tex_w.size();
    if (n_texels) *n_texels = s->texels.size() / 3;
    if (n_warnings) *n_warnings = (uint32_t)s->warnings.size();
    return PG_OK;
}

pg_status pg_scene_copy_textures(const pg_host_scene* s, float* uv, int32_t* material_texture, float* texels,
                                 uint64_t* texel_offset, int32_t* tex_width, int32_t* tex_height) {
    if (!s) return PG_ERR_INVALID;
    bool has = !s->mat_tex.empty();
    size_t nt = s->mat.size();
    if (uv) {
        if (has) std::memcpy(uv, s->uv.data(), s->uv.size() * sizeof(float));
        else std::memset(uv, 0, nt * 6 * sizeof(float));
    }
    if (material_texture) {
        for (size_t i = 0; i < s->mats.size(); ++i) material_texture[i] = has ? s->mat_tex[i] : -1;
    }
    if (texels && !s->texels.empty()) std::memcpy(texels, s->texels.data(), s->texels.size() * sizeof(float));
    for (size_t i = 0; i < s->tex_w.size(); ++i) {
        if (texel_offset) texel_offset[i] = s->tex_offset[i];
        if (tex_width) tex_width[i] = s->tex_w[i];
        if (tex_height) tex_height[i] = s->tex_h[i];
    }
    return PG_OK;
}

const char* pg_scene_warning(const pg_host_scene* s, uint32_t i) {
    if (!s || i >= s->warnings.size()) return nullptr;
    return s->warnings[i].c_str();
}

pg_status pg_image_read(const char* path, int32_t* width, int32_t* height, int32_t* channels, float* rgb) {
    if (!path || !width || !height) return PG_ERR_INVALID;
    pg::Image img;
    std::string p(path);
    bool hdr = false;
    std::string e;
    try {
        e = pg::read_image_file(p, img, hdr);
    } catch (const std::bad_alloc&) {
        g_load_error = p + ": out of host memory";
        return PG_ERR_NOMEM;
    } catch (const std::exception& x) {
        g_load_error = p + ": " + x.what();
        return PG_ERR_INVALID;
    }
    if (!e.empty()) { g_load_error = p + ": " + e; return PG_ERR_INVALID; }
    *width = img.w; *height = img.h;
    if (channels) *channels = img.channels;
    if (rgb) {
        size_t n = (size_t)img.w * img.h * img.channels;
        for (size_t i = 0; i < n; ++i) rgb[i] = hdr ? img.f32[i] : (float)img.u8[i];
    }
    return PG_OK;
}

// RayTracer::savePNG (Renderer.h:751): 8-bit RGB, one zlib stream of filter-0 rows.
pg_status pg_image_write_png(const char* path, const uint8_t* rgb, int32_t width, int32_t height) {
    return guarded([&] { return write_png(path, rgb, width, height); });
}

pg_status pg_image_write_hdr(const char* path, const float* rgb, int32_t width, int32_t height) {
    return guarded([&] { return write_hdr(path, rgb, width, height); });
}

}  // extern "C"

static pg_status write_png(const char* path, const uint8_t* rgb, int32_t width, int32_t height) {
    if (!path || !rgb || width < 1 || height < 1) return PG_ERR_INVALID;
    size_t stride = (size_t)width * 3;
    std::vector<uint8_t> raw((stride + 1) * height);
    for (int32_t y = 0; y < height; ++y) {
        raw[y * (stride + 1)] = 0;
        std::memcpy(&raw[y * (stride + 1) + 1], rgb + y * stride, stride);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return PG_ERR_NOMEM;
    std::string out("\x89PNG\r\n\x1a\n", 8);
    auto chunk = [&](const char* type, const uint8_t* d, uint32_t n) {
        uint8_t len[4] = {(uint8_t)(n >> 24), (uint8_t)(n >> 16), (uint8_t)(n >> 8), (uint8_t)n};
        out.append((const char*)len, 4);
        std::string body(type, 4);
        if (n) body.append((const char*)d, n);
        out += body;
        uint32_t crc = (uint32_t)crc32(0L, (const Bytef*)body.data(), (uInt)body.size());
        uint8_t c[4] = {(uint8_t)(crc >> 24), (uint8_t)(crc >> 16), (uint8_t)(crc >> 8), (uint8_t)crc};
        out.append((const char*)c, 4);
    };
    uint8_t ihdr[13] = {(uint8_t)(width >> 24), (uint8_t)(width >> 16), (uint8_t)(width >> 8), (uint8_t)width,
                        (uint8_t)(height >> 24), (uint8_t)(height >> 16), (uint8_t)(height >> 8), (uint8_t)height,
                        8, 2, 0, 0, 0};
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), (uint32_t)zn);
    chunk("IEND", nullptr, 0);
    std::ofstream f(path, std::ios::binary);
    if (!f) return PG_ERR_INVALID;
    f.write(out.data(), (std::streamsize)out.size());
    return f ? PG_OK : PG_ERR_INVALID;
}

// Film::save (Imaging.h:263-272) writes the mean image as Radiance RGBE: the shared
// exponent of the largest component (frexp), mantissas scaled by 256 / max and
// truncated; scanlines run-length encoded per component when 8 <= width < 32768.
static pg_status write_hdr(const char* path, const float* rgb, int32_t width, int32_t height) {
    if (!path || !rgb || width < 1 || height < 1) return PG_ERR_INVALID;
    std::ofstream f(path, std::ios::binary);
    if (!f) return PG_ERR_INVALID;
    char head[128];
    int hn = std::snprintf(head, sizeof(head), "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", height, width);
    f.write(head, hn);
    std::vector<uint8_t> sl((size_t)width * 4);
    std::vector<uint8_t> out;
    for (int32_t y = 0; y < height; ++y) {
        for (int32_t x = 0; x < width; ++x) {
            const float* p = rgb + 3 * ((size_t)y * width + x);
            float mx = p[0] > p[1] ? p[0] : p[1];
            mx = mx > p[2] ? mx : p[2];
            uint8_t* e = &sl[(size_t)x * 4];
            if (!(mx >= 1e-32f)) { e[0] = e[1] = e[2] = e[3] = 0; continue; }
            int ex = 0;
            float norm = (float)std::frexp(mx, &ex) * 256.0f / mx;
            for (int c = 0; c < 3; ++c) e[c] = (uint8_t)(p[c] * norm);
            e[3] = (uint8_t)(ex + 128);
        }
        out.clear();
        if (width >= 8 && width < 32768) {
            uint8_t lead[4] = {2, 2, (uint8_t)(width >> 8), (uint8_t)(width & 255)};
            out.insert(out.end(), lead, lead + 4);
            for (int c = 0; c < 4; ++c)
                for (int32_t x = 0; x < width;) {
                    int32_t n = width - x < 128 ? width - x : 128;
                    out.push_back((uint8_t)n);
                    for (int32_t k = 0; k < n; ++k) out.push_back(sl[(size_t)(x + k) * 4 + c]);
                    x += n;
                }
        } else {
            out = sl;
        }
        f.write((const char*)out.data(), (std::streamsize)out.size());
    }
    return f ? PG_OK : PG_ERR_INVALID;
}
