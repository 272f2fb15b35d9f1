// Wavefront OBJ/MTL import (host only).  The reference reads assets through
// assimp (AssetImporter.cpp:12-25), which is not part of this build; this is
// a small parser of the OBJ/MTL subset a path tracer needs: v, vn, f (v,
// v/vt, v//vn, v/vt/vn, negative indices), usemtl, mtllib; newmtl, Kd, Ke, Ns.
#include <algorithm>
#include <cerrno>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
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
};

struct Corner {
    int v, n;  // 0-based, n = -1 when absent
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
        }
    }
}

Vec sub(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
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

}  // namespace

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
        std::ifstream f(path);
        if (!f) throw ParseError{std::string("cannot open ") + path + ": " + std::strerror(errno)};
        std::vector<Vec> pos, nrm;
        std::vector<Face> faces;
        std::map<std::string, Material> mats;
        std::vector<Material> used;  // material of each face by index
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
                pos.push_back(read_vec(ls, "vertex at " + where));
            } else if (key == "vn") {
                nrm.push_back(read_vec(ls, "normal at " + where));
            } else if (key == "mtllib") {
                std::string name;
                while (ls >> name) load_mtl(dir_of(path) + name, mats);
            } else if (key == "usemtl") {
                std::string name;
                ls >> name;
                auto it = used_index.find(name);
                if (it == used_index.end()) {
                    auto m = mats.find(name);
                    used.push_back(m == mats.end() ? Material{} : m->second);
                    it = used_index.emplace(name, static_cast<int>(used.size()) - 1).first;
                }
                cur = it->second;
            } else if (key == "f") {
                Face fc;
                fc.mat = cur;
                std::string tok;
                while (ls >> tok) {
                    Corner c{-1, -1};
                    const size_t s1 = tok.find('/');
                    c.v = resolve(std::strtol(tok.substr(0, s1).c_str(), nullptr, 10), pos.size(), where);
                    if (s1 != std::string::npos) {
                        const size_t s2 = tok.find('/', s1 + 1);
                        if (s2 != std::string::npos && s2 + 1 < tok.size())
                            c.n = resolve(std::strtol(tok.substr(s2 + 1).c_str(), nullptr, 10), nrm.size(), where);
                    }
                    fc.c.push_back(c);
                }
                if (fc.c.size() < 3) throw ParseError{"face with fewer than 3 corners at " + where};
                faces.push_back(fc);
            }
        }
        size_t n = 0;
        for (const Face& fc : faces) n += fc.c.size() - 2;
        if (n >= (1u << 30)) throw ParseError{"too many triangles"};
        if (!tris) {
            *count = static_cast<uint32_t>(n);
            return PT_OK;
        }
        if (*count < n) throw ParseError{"output arrays hold fewer triangles than the file has"};
        const float zs = opt.left_handed ? -1.0f : 1.0f;
        auto conv = [&](Vec v) { return Vec{v.x, v.y, v.z * zs}; };
        uint32_t id = 0;
        for (const Face& fc : faces) {
            const Material m = fc.mat >= 0 ? used[static_cast<size_t>(fc.mat)] : Material{};
            for (size_t k = 1; k + 1 < fc.c.size(); ++k) {
                Corner c[3] = {fc.c[0], fc.c[k], fc.c[k + 1]};
                if (opt.left_handed) std::swap(c[0], c[2]);  // reversed winding
                Vec p[3];
                for (int j = 0; j < 3; ++j) p[j] = conv(pos[static_cast<size_t>(c[j].v)]);
                pt_triangle& t = tris[id];
                put3(t.v0, p[0]);
                put3(t.v1, p[1]);
                put3(t.v2, p[2]);
                t.id = id;
                if (infos) {
                    pt_triangle_info& in = infos[id];
                    std::memset(&in, 0, sizeof in);
                    const bool have_n = c[0].n >= 0 && c[1].n >= 0 && c[2].n >= 0;
                    const Vec face_n = normalize(cross(sub(p[1], p[0]), sub(p[2], p[0])));
                    float* dst[3] = {in.n0, in.n1, in.n2};
                    for (int j = 0; j < 3; ++j)
                        put3(dst[j], have_n ? conv(nrm[static_cast<size_t>(c[j].n)]) : face_n);
                    put3(in.albedo, opt.normalize_albedo ? normalize(m.kd) : m.kd);
                    put3(in.emission, m.ke);
                    in.specularity = m.ns;
                }
                ++id;
            }
        }
        *count = id;
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
