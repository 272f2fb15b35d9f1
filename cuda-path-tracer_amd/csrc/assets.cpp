// Asset I/O (host only).  The reference reads assets through assimp
// (AssetImporter.cpp:12-25) and images through SOIL, neither part of this
// build; this is a small parser of the OBJ/MTL subset a path tracer needs:
// v (with an optional vertex colour), vn, vt, f (v, v/vt, v//vn, v/vt/vn,
// negative indices), usemtl, mtllib; newmtl, Kd, Ke, Ns, map_Kd, map_bump /
// bump; a TGA reader for the textures, and a PNG writer for images.
#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pt_assets.h"

namespace {

thread_local std::string g_error;

struct Vec {
    float x, y, z;
};

struct Material {
    Vec kd{1.0f, 1.0f, 1.0f};  // AssetImporter.cpp:98 default diffuse
    Vec ke{0.0f, 0.0f, 0.0f};  // :100 default emission
    float ns = 0.0f;           // :101 default shininess
    std::string map_kd;        // diffuse texture (aiTextureType_DIFFUSE, :92)
    std::string map_bump;      // normal map (map_bump / bump -> aiTextureType_HEIGHT, :93)
};

struct Corner {
    int v, t, n;  // 0-based; t, n = -1 when absent
};

struct Face {
    std::vector<Corner> c;
    int mat;
};

struct ParseError {
    std::string msg;
};

std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

Vec read_vec(std::istringstream& ls, const std::string& what) {
    Vec v;
    if (!(ls >> v.x >> v.y >> v.z)) throw ParseError{"bad " + what};
    return v;
}

// OBJ index (1-based, negative = relative to the end) -> 0-based.
int resolve(long i, size_t n, const std::string& what) {
    long k = i > 0 ? i - 1 : static_cast<long>(n) + i;
    if (i == 0 || k < 0 || k >= static_cast<long>(n)) throw ParseError{"index out of range in " + what};
    return static_cast<int>(k);
}

// Texture statement: options (-bm 0.5, -o u v w, ...) may precede the file
// name, which is the last token; Windows separators become '/'.
std::string texture_path(std::istringstream& ls, const std::string& dir) {
    std::string tok, last;
    while (ls >> tok) last = tok;
    for (char& ch : last)
        if (ch == '\\') ch = '/';
    return last.empty() ? last : dir + last;
}

void load_mtl(const std::string& path, std::map<std::string, Material>& mats) {
    std::ifstream f(path);
    if (!f) throw ParseError{"cannot open material library " + path};
    std::string line;
    Material* cur = nullptr;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            std::string name;
            ls >> name;
            cur = &mats[name];
            *cur = Material{};
        } else if (cur && key == "Kd") {
            cur->kd = read_vec(ls, "Kd in " + path);
        } else if (cur && key == "Ke") {
            cur->ke = read_vec(ls, "Ke in " + path);
        } else if (cur && key == "Ns") {
            if (!(ls >> cur->ns)) throw ParseError{"bad Ns in " + path};
        } else if (cur && key == "map_Kd") {
            cur->map_kd = texture_path(ls, dir_of(path));
        } else if (cur && (key == "map_bump" || key == "bump") && cur->map_bump.empty()) {
            cur->map_bump = texture_path(ls, dir_of(path));
        }
    }
}

Vec add(Vec a, Vec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec sub(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec scale(Vec a, float s) { return {a.x * s, a.y * s, a.z * s}; }
Vec cross(Vec a, Vec b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// glm::normalize: v * (1 / sqrt(dot(v, v)))
Vec normalize(Vec v) {
    const float inv = 1.0f / std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z);
    return {v.x * inv, v.y * inv, v.z * inv};
}

void put3(float* dst, Vec v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

struct Obj {
    std::vector<Vec> pos, nrm, col;
    std::vector<bool> has_col;
    std::vector<std::array<float, 2>> uv;
    std::vector<Face> faces;
    std::vector<Material> used;  // materials in first-use order; faces index them
    size_t ntris = 0;
};

Obj parse_obj(const char* path) {
    std::ifstream f(path);
    if (!f) throw ParseError{std::string("cannot open ") + path + ": " + std::strerror(errno)};
    Obj o;
    std::map<std::string, Material> mats;
    std::map<std::string, int> used_index;
    int cur = -1;
    std::string line;
    size_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        const std::string where = std::string(path) + ":" + std::to_string(lineno);
        if (key == "v") {
            o.pos.push_back(read_vec(ls, "vertex at " + where));
            Vec c;
            const bool col = static_cast<bool>(ls >> c.x >> c.y >> c.z);  // optional vertex colour
            o.col.push_back(col ? c : Vec{0.0f, 0.0f, 0.0f});
            o.has_col.push_back(col);
        } else if (key == "vn") {
            o.nrm.push_back(read_vec(ls, "normal at " + where));
        } else if (key == "vt") {
            std::array<float, 2> t{0.0f, 0.0f};
            if (!(ls >> t[0])) throw ParseError{"bad texture coordinate at " + where};
            if (!(ls >> t[1])) t[1] = 0.0f;
            o.uv.push_back(t);
        } else if (key == "mtllib") {
            std::string name;
            while (ls >> name) load_mtl(dir_of(path) + name, mats);
        } else if (key == "usemtl") {
            std::string name;
            ls >> name;
            auto it = used_index.find(name);
            if (it == used_index.end()) {
                auto m = mats.find(name);
                o.used.push_back(m == mats.end() ? Material{} : m->second);
                it = used_index.emplace(name, static_cast<int>(o.used.size()) - 1).first;
            }
            cur = it->second;
        } else if (key == "f") {
            Face fc;
            fc.mat = cur;
            std::string tok;
            while (ls >> tok) {
                Corner c{-1, -1, -1};
                const size_t s1 = tok.find('/');
                c.v = resolve(std::strtol(tok.substr(0, s1).c_str(), nullptr, 10), o.pos.size(), where);
                if (s1 != std::string::npos) {
                    const size_t s2 = tok.find('/', s1 + 1);
                    const std::string ts = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
                    if (!ts.empty()) c.t = resolve(std::strtol(ts.c_str(), nullptr, 10), o.uv.size(), where);
                    if (s2 != std::string::npos && s2 + 1 < tok.size())
                        c.n = resolve(std::strtol(tok.substr(s2 + 1).c_str(), nullptr, 10), o.nrm.size(), where);
                }
                fc.c.push_back(c);
            }
            if (fc.c.size() < 3) throw ParseError{"face with fewer than 3 corners at " + where};
            o.faces.push_back(fc);
        }
    }
    for (const Face& fc : o.faces) o.ntris += fc.c.size() - 2;
    if (o.ntris >= (1u << 30)) throw ParseError{"too many triangles"};
    return o;
}

// Triangles, infos and (if `detail`) surface detail of a parsed file, in file
// order with ids 0..  tex_index maps a texture path to its index.
void emit(const Obj& o, const pt_obj_options& opt, pt_triangle* tris, pt_triangle_info* infos,
          pt_triangle_detail* detail, const std::map<std::string, uint32_t>* tex_index) {
    const float zs = opt.left_handed ? -1.0f : 1.0f;
    auto conv = [&](Vec v) { return Vec{v.x, v.y, v.z * zs}; };
    uint32_t id = 0;
    for (const Face& fc : o.faces) {
        const Material m = fc.mat >= 0 ? o.used[static_cast<size_t>(fc.mat)] : Material{};
        for (size_t k = 1; k + 1 < fc.c.size(); ++k) {
            Corner c[3] = {fc.c[0], fc.c[k], fc.c[k + 1]};
            if (opt.left_handed) std::swap(c[0], c[2]);  // reversed winding
            Vec p[3];
            for (int j = 0; j < 3; ++j) p[j] = conv(o.pos[static_cast<size_t>(c[j].v)]);
            pt_triangle& t = tris[id];
            put3(t.v0, p[0]);
            put3(t.v1, p[1]);
            put3(t.v2, p[2]);
            t.id = id;
            const bool have_n = c[0].n >= 0 && c[1].n >= 0 && c[2].n >= 0;
            const Vec face_n = normalize(cross(sub(p[1], p[0]), sub(p[2], p[0])));
            Vec n[3];
            for (int j = 0; j < 3; ++j) n[j] = have_n ? conv(o.nrm[static_cast<size_t>(c[j].n)]) : face_n;
            if (infos) {
                pt_triangle_info& in = infos[id];
                std::memset(&in, 0, sizeof in);
                float* dst[3] = {in.n0, in.n1, in.n2};
                for (int j = 0; j < 3; ++j) put3(dst[j], n[j]);
                put3(in.albedo, opt.normalize_albedo ? normalize(m.kd) : m.kd);
                put3(in.emission, m.ke);
                in.specularity = m.ns;
            }
            if (detail) {
                // PathTracingApp.cpp:59-93: vertex colour from the file or the
                // material colour, uv from the file (else 0), tangent frame.
                pt_triangle_detail& d = detail[id];
                std::memset(&d, 0, sizeof d);
                float uv[3][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
                for (int j = 0; j < 3; ++j) {
                    const size_t v = static_cast<size_t>(c[j].v);
                    put3(d.color[j], o.has_col[v] ? o.col[v] : m.kd);
                    if (c[j].t >= 0) {
                        uv[j][0] = o.uv[static_cast<size_t>(c[j].t)][0];
                        uv[j][1] = o.uv[static_cast<size_t>(c[j].t)][1];
                    }
                    d.uv[j][0] = uv[j][0];
                    d.uv[j][1] = uv[j][1];
                }
                // Per-face tangent frame from the uv derivatives: T along +u,
                // B along +v (e1 = T du1 + B dv1, e2 = T du2 + B dv2); with
                // degenerate uvs T = normalize(e1) and B = cross(N, T).
                const Vec e1 = sub(p[1], p[0]), e2 = sub(p[2], p[0]);
                const float du1 = uv[1][0] - uv[0][0], dv1 = uv[1][1] - uv[0][1];
                const float du2 = uv[2][0] - uv[0][0], dv2 = uv[2][1] - uv[0][1];
                const float det = du1 * dv2 - du2 * dv1;
                Vec tg, bt;
                if (std::fabs(det) > 1e-20f) {
                    const float r = 1.0f / det;
                    tg = scale(sub(scale(e1, dv2), scale(e2, dv1)), r);
                    bt = scale(sub(scale(e2, du1), scale(e1, du2)), r);
                } else {
                    tg = normalize(e1);
                    bt = cross(face_n, tg);
                }
                for (int j = 0; j < 3; ++j) {
                    put3(d.tangent[j], tg);
                    put3(d.bitangent[j], bt);
                }
                auto tex = [&](const std::string& path) {
                    if (path.empty() || !tex_index) return PT_NO_TEXTURE;
                    return tex_index->at(path);
                };
                d.albedo_tex = tex(m.map_kd);
                d.normal_tex = tex(m.map_bump);
            }
            ++id;
        }
    }
}

struct Image {
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> rgba;  // row 0 = bottom
};

// Truevision TGA: uncompressed (2, 3) and run-length (10, 11) true-colour or
// grey images of 8, 24 or 32 bits; rows stored bottom-up unless the
// descriptor's bit 5 says top-down, right-to-left if bit 4.  BGR(A) -> RGBA,
// grey g -> (g, g, g, 255), no alpha -> 255.
Image load_tga(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ParseError{"cannot open texture " + path};
    std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (b.size() < 18) throw ParseError{"truncated TGA header in " + path};
    const uint32_t idlen = b[0], cmtype = b[1], itype = b[2];
    const uint32_t cmlen = b[5] | (b[6] << 8), cmbits = b[7];
    Image im;
    im.w = b[12] | (b[13] << 8);
    im.h = b[14] | (b[15] << 8);
    const uint32_t bpp = b[16], desc = b[17];
    const bool grey = itype == 3 || itype == 11, rle = itype == 10 || itype == 11;
    if (!(itype == 2 || itype == 3 || itype == 10 || itype == 11))
        throw ParseError{"unsupported TGA image type " + std::to_string(itype) + " in " + path};
    if (grey ? bpp != 8 : (bpp != 24 && bpp != 32))
        throw ParseError{"unsupported TGA pixel depth " + std::to_string(bpp) + " in " + path};
    if (im.w == 0 || im.h == 0 || im.w > 16384 || im.h > 16384) throw ParseError{"bad TGA size in " + path};
    size_t at = 18 + idlen + (cmtype ? static_cast<size_t>(cmlen) * ((cmbits + 7) / 8) : 0);
    const uint32_t bytes = bpp / 8;
    const size_t npx = static_cast<size_t>(im.w) * im.h;
    std::vector<uint8_t> px(npx * bytes);
    if (!rle) {
        if (b.size() < at + px.size()) throw ParseError{"truncated TGA pixels in " + path};
        std::memcpy(px.data(), b.data() + at, px.size());
    } else {
        size_t k = 0;
        while (k < npx) {
            if (at >= b.size()) throw ParseError{"truncated TGA run in " + path};
            const uint32_t hd = b[at++], cnt = (hd & 0x7Fu) + 1u;
            if (k + cnt > npx) throw ParseError{"TGA run past the image in " + path};
            if (hd & 0x80u) {
                if (at + bytes > b.size()) throw ParseError{"truncated TGA run in " + path};
                for (uint32_t r = 0; r < cnt; ++r, ++k) std::memcpy(&px[k * bytes], &b[at], bytes);
                at += bytes;
            } else {
                if (at + static_cast<size_t>(cnt) * bytes > b.size()) throw ParseError{"truncated TGA run in " + path};
                std::memcpy(&px[k * bytes], &b[at], static_cast<size_t>(cnt) * bytes);
                at += static_cast<size_t>(cnt) * bytes;
                k += cnt;
            }
        }
    }
    im.rgba.resize(npx * 4);
    const bool top_down = (desc & 0x20u) != 0, right_left = (desc & 0x10u) != 0;
    for (uint32_t y = 0; y < im.h; ++y) {
        const uint32_t dy = top_down ? im.h - 1 - y : y;  // stored row y -> bottom-up row dy
        for (uint32_t x = 0; x < im.w; ++x) {
            const uint32_t dx = right_left ? im.w - 1 - x : x;
            const uint8_t* s = &px[(static_cast<size_t>(y) * im.w + x) * bytes];
            uint8_t* d = &im.rgba[(static_cast<size_t>(dy) * im.w + dx) * 4];
            if (grey) {
                d[0] = d[1] = d[2] = s[0];
                d[3] = 255;
            } else {
                d[0] = s[2];
                d[1] = s[1];
                d[2] = s[0];
                d[3] = bytes == 4 ? s[3] : 255;
            }
        }
    }
    return im;
}

}  // namespace

struct pt_asset {
    std::vector<pt_triangle> tris;
    std::vector<pt_triangle_info> infos;
    std::vector<pt_triangle_detail> detail;
    std::vector<Image> images;
};

extern "C" {

int pt_asset_error(char* buf, size_t len) {
    if (!buf || len == 0) return PT_E_INVALID_ARG;
    std::snprintf(buf, len, "%s", g_error.c_str());
    return PT_OK;
}

int pt_load_obj(const char* path, const pt_obj_options* options, pt_triangle* tris, pt_triangle_info* infos,
                uint32_t* count) {
    g_error.clear();
    if (!path || !count) {
        g_error = "NULL argument";
        return PT_E_INVALID_ARG;
    }
    const pt_obj_options opt = options ? *options : pt_obj_options{1u, 1u};
    try {
        const Obj o = parse_obj(path);
        if (!tris) {
            *count = static_cast<uint32_t>(o.ntris);
            return PT_OK;
        }
        if (*count < o.ntris) throw ParseError{"output arrays hold fewer triangles than the file has"};
        emit(o, opt, tris, infos, nullptr, nullptr);
        *count = static_cast<uint32_t>(o.ntris);
        return PT_OK;
    } catch (const ParseError& e) {
        g_error = e.msg;
    } catch (const std::exception& e) {
        g_error = e.what();
    }
    return PT_E_INVALID_ARG;
}

int pt_asset_load_obj(const char* path, const pt_obj_options* options, pt_asset** out) {
    g_error.clear();
    if (!path || !out) {
        g_error = "NULL argument";
        return PT_E_INVALID_ARG;
    }
    *out = nullptr;
    const pt_obj_options opt = options ? *options : pt_obj_options{1u, 1u};
    try {
        const Obj o = parse_obj(path);
        auto a = std::make_unique<pt_asset>();
        std::map<std::string, uint32_t> tex_index;
        for (const Material& m : o.used) {
            for (const std::string* p : {&m.map_kd, &m.map_bump}) {
                if (p->empty() || tex_index.count(*p)) continue;
                std::string ext = p->size() >= 4 ? p->substr(p->size() - 4) : std::string();
                for (char& ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
                if (ext != ".tga") throw ParseError{"unsupported texture format (TGA only): " + *p};
                tex_index.emplace(*p, static_cast<uint32_t>(a->images.size()));
                a->images.push_back(load_tga(*p));
            }
        }
        a->tris.resize(o.ntris);
        a->infos.resize(o.ntris);
        a->detail.resize(o.ntris);
        emit(o, opt, a->tris.data(), a->infos.data(), a->detail.data(), &tex_index);
        *out = a.release();
        return PT_OK;
    } catch (const ParseError& e) {
        g_error = e.msg;
    } catch (const std::exception& e) {
        g_error = e.what();
    }
    return PT_E_INVALID_ARG;
}

int pt_asset_info(const pt_asset* a, uint32_t* ntris, uint32_t* ntextures) {
    if (!a) return PT_E_INVALID_ARG;
    if (ntris) *ntris = static_cast<uint32_t>(a->tris.size());
    if (ntextures) *ntextures = static_cast<uint32_t>(a->images.size());
    return PT_OK;
}

int pt_asset_copy(const pt_asset* a, pt_triangle* tris, pt_triangle_info* infos, pt_triangle_detail* detail) {
    if (!a) return PT_E_INVALID_ARG;
    const size_t n = a->tris.size();
    if (tris && n) std::memcpy(tris, a->tris.data(), n * sizeof(pt_triangle));
    if (infos && n) std::memcpy(infos, a->infos.data(), n * sizeof(pt_triangle_info));
    if (detail && n) std::memcpy(detail, a->detail.data(), n * sizeof(pt_triangle_detail));
    return PT_OK;
}

int pt_asset_textures(const pt_asset* a, pt_texture* textures) {
    if (!a || (!textures && !a->images.empty())) return PT_E_INVALID_ARG;
    for (size_t k = 0; k < a->images.size(); ++k)
        textures[k] = pt_texture{a->images[k].w, a->images[k].h, a->images[k].rgba.data()};
    return PT_OK;
}

int pt_asset_free(pt_asset* a) {
    delete a;
    return PT_OK;
}

int pt_load_tga(const char* path, uint32_t* w, uint32_t* h, uint8_t* rgba) {
    g_error.clear();
    if (!path || !w || !h) {
        g_error = "NULL argument";
        return PT_E_INVALID_ARG;
    }
    try {
        const Image im = load_tga(path);
        if (rgba) {
            if (*w != im.w || *h != im.h) throw ParseError{"output size does not match the image"};
            std::memcpy(rgba, im.rgba.data(), im.rgba.size());
        }
        *w = im.w;
        *h = im.h;
        return PT_OK;
    } catch (const ParseError& e) {
        g_error = e.msg;
    } catch (const std::exception& e) {
        g_error = e.what();
    }
    return PT_E_INVALID_ARG;
}

namespace {

// PNG (ISO/IEC 15948) with a zlib stream (RFC 1950) of stored deflate blocks
// (RFC 1951, BTYPE 00): no compression, no dependency.
uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int b = 0; b < 8; ++b) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        ready = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back(static_cast<uint8_t>(x >> s));
}

void put_chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
    put_be32(out, static_cast<uint32_t>(data.size()));
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = crc32_update(0xFFFFFFFFu, out.data() + at, 4 + data.size()) ^ 0xFFFFFFFFu;
    put_be32(out, crc);
}

uint8_t quantize(float v, bool srgb) {
    if (!(v > 0.0f)) return 0;  // also NaN
    if (v >= 1.0f) return 255;
    double x = v;
    if (srgb) x = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
    return static_cast<uint8_t>(std::floor(x * 255.0 + 0.5));
}

}  // namespace

int pt_write_png(const char* path, const float* rgba, uint32_t w, uint32_t h, uint32_t flags) {
    try {
        if (!path || !rgba || w == 0 || h == 0) throw ParseError{"pt_write_png: NULL argument or empty image"};
        if (static_cast<uint64_t>(w) * 3 + 1 > 0x7FFFFFFFull || static_cast<uint64_t>(h) > 0x7FFFFFFFull)
            throw ParseError{"pt_write_png: image too large"};
        const bool srgb = (flags & PT_PNG_SRGB) != 0, top_down = (flags & PT_PNG_TOP_DOWN) != 0;
        // Filtered scanlines: filter type 0, then RGB bytes; PNG rows run top to bottom.
        const size_t row = static_cast<size_t>(w) * 3 + 1;
        std::vector<uint8_t> raw(row * h);
        for (uint32_t y = 0; y < h; ++y) {
            const uint32_t src = top_down ? y : h - 1 - y;
            uint8_t* d = raw.data() + row * y;
            d[0] = 0;
            const float* s = rgba + static_cast<size_t>(src) * w * 4;
            for (uint32_t x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) d[1 + 3 * x + c] = quantize(s[4 * x + c], srgb);
        }
        std::vector<uint8_t> z;
        z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
        z.push_back(0x78);
        z.push_back(0x01);
        size_t at = 0;
        do {
            const size_t n = std::min<size_t>(65535, raw.size() - at);
            z.push_back(at + n == raw.size() ? 1 : 0);
            z.push_back(static_cast<uint8_t>(n));
            z.push_back(static_cast<uint8_t>(n >> 8));
            z.push_back(static_cast<uint8_t>(~n));
            z.push_back(static_cast<uint8_t>(~n >> 8));
            z.insert(z.end(), raw.begin() + static_cast<std::ptrdiff_t>(at),
                     raw.begin() + static_cast<std::ptrdiff_t>(at + n));
            at += n;
        } while (at < raw.size());
        uint32_t a = 1, b = 0;  // Adler-32
        for (uint8_t v : raw) {
            a = (a + v) % 65521u;
            b = (b + a) % 65521u;
        }
        put_be32(z, (b << 16) | a);
        std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
        std::vector<uint8_t> ihdr;
        put_be32(ihdr, w);
        put_be32(ihdr, h);
        ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, filter 0, no interlace
        put_chunk(out, "IHDR", ihdr);
        put_chunk(out, "IDAT", z);
        put_chunk(out, "IEND", {});
        std::FILE* f = std::fopen(path, "wb");
        if (!f) throw ParseError{std::string("cannot open ") + path + " for writing"};
        const size_t put = std::fwrite(out.data(), 1, out.size(), f);
        const int closed = std::fclose(f);
        if (put != out.size() || closed != 0) throw ParseError{std::string("cannot write ") + path};
        return PT_OK;
    } catch (const ParseError& e) {
        g_error = e.msg;
    } catch (const std::exception& e) {
        g_error = e.what();
    }
    return PT_E_INVALID_ARG;
}

}  // extern "C"
