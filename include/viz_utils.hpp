// viz_utils.hpp -- the reference driver's output files.
// Replaces include/viz_utils.hpp:19-132 with the same functions and meaning:
//   dump_vorticity_png          8-bit grey image of sum_i |w_i| W(|p - x_i|)
//                               at the pixel centres ((i+1/2)/res, (j+1/2)/res),
//                               pixel = clamp(128 + int(value * 127 / max|value|),
//                               0, 255), row j = y index; periodic: the minimum
//                               image; r^2 < 1e-12 skipped (viz_utils.hpp:26-82)
//   dump_velocity_svg_centered  one arrow per particle, centred on it along its
//                               velocity, length vel_scale |v|, ends wrapped when
//                               periodic, y flipped (viz_utils.hpp:84-132)
// The field is evaluated on the GPU (vpmc_query_view over an all-particle cell
// list whose cells are at least h_w wide) instead of the reference's host loop
// over every particle per pixel; the PNG is written without an image library
// (zlib stream of stored blocks).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_vector_alias.hpp"
#include "spatial_hash.hpp"
#include "vortex_particle_mc.hpp"
#include "vpmc.h"

inline float length(float2 v) { return std::sqrt(v.x * v.x + v.y * v.y); }

namespace vpmc_viz {

// CRC-32 (ISO 3309 polynomial, reflected 0xEDB88320) as PNG chunks use it
inline uint32_t crc32(const unsigned char *p, size_t n, uint32_t c = 0xFFFFFFFFu)
{
    static const std::vector<uint32_t> table = [] {
        std::vector<uint32_t> t(256);
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t v = k;
            for (int b = 0; b < 8; ++b) v = (v & 1u) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            t[k] = v;
        }
        return t;
    }();
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c;
}

inline void put_be32(std::vector<unsigned char> &o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}

inline void put_chunk(std::vector<unsigned char> &o, const char *tag,
                      const std::vector<unsigned char> &data)
{
    put_be32(o, (uint32_t)data.size());
    std::vector<unsigned char> td(tag, tag + 4);
    td.insert(td.end(), data.begin(), data.end());
    o.insert(o.end(), td.begin(), td.end());
    put_be32(o, crc32(td.data(), td.size()) ^ 0xFFFFFFFFu);
}

// 8-bit grey PNG of img (h rows of w bytes, row 0 first): filter 0 rows in a
// zlib stream of stored deflate blocks (RFC 1950/1951)
inline std::vector<unsigned char> encode_png(const std::vector<unsigned char> &img, int w, int h)
{
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (w + 1));
    for (int r = 0; r < h; ++r) {
        raw.push_back(0);
        raw.insert(raw.end(), img.begin() + (size_t)r * w, img.begin() + (size_t)(r + 1) * w);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char)(len & 0xFF));
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)(~len & 0xFF));
        z.push_back((unsigned char)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;   // Adler-32
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<unsigned char> o = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 0, 0, 0, 0});   // depth 8, grey, deflate, filter 0, no interlace
    put_chunk(o, "IHDR", ihdr);
    put_chunk(o, "IDAT", z);
    put_chunk(o, "IEND", {});
    return o;
}

}  // namespace vpmc_viz

inline void dump_vorticity_png(const Simulation &sim, const std::string &file, int res = 512,
                               bool periodic = false)
{
    const int N = sim.particles_.N_cur;
    const float h_w = sim.P_.h_w, Lx = sim.P_.Lx, Ly = sim.P_.Ly;
    if (res <= 0 || N <= 0) throw std::invalid_argument("dump_vorticity_png: res, N_cur");
    // all-particle cell list, cells >= h_w wide so that 3 x 3 cells cover the support
    const int rx = (int)std::floor(Lx / h_w), ry = (int)std::floor(Ly / h_w);
    if (rx < 3 || ry < 3) throw std::invalid_argument("dump_vorticity_png: h_w > L / 3");
    SpatialHash H;
    H.resize_cells(rx, ry);
    H.resize_particles(N);
    std::vector<int> iota(N);
    for (int i = 0; i < N; ++i) iota[i] = i;
    dvec<int> sub(iota);
    FillCells_cuda(rx, ry, Lx, Ly, N, raw_ptr(sub), sim.particles_.d_p(), raw_ptr(H.cell_num),
                   raw_ptr(H.cell_acc), raw_ptr(H.vp_cell));
    CountingSort_cuda(N, raw_ptr(H.cell_num), raw_ptr(H.cell_acc), raw_ptr(H.vp_cell),
                      raw_ptr(H.vp_sort));
    const vpmc_view v{(const float *)sim.particles_.d_p(), sim.particles_.d_omega(), N,
                      raw_ptr(H.vp_sort), raw_ptr(H.cell_acc), rx, ry, Lx, Ly, N, raw_ptr(sub)};
    std::vector<float2> px((size_t)res * res);   // [j][i]: row j = y
    for (int j = 0; j < res; ++j)
        for (int i = 0; i < res; ++i)
            px[(size_t)j * res + i] = make_float2((i + 0.5f) / res, (j + 0.5f) / res);
    dvec<float2> dpx(px);
    dvec<float> val(px.size());
    const int st = vpmc_query_view(&v, (int)px.size(), (const float *)raw_ptr(dpx), 0, 0, 1, h_w,
                                   periodic ? 1 : 0, raw_ptr(val), nullptr);
    if (st != VPMC_OK)
        throw std::runtime_error(std::string("dump_vorticity_png: ") + vpmc_status_string(st));
    vpmc_detail::hip_check(hipDeviceSynchronize(), "dump_vorticity_png");
    const std::vector<float> w = val.to_host();
    float max_abs = 0.f;
    for (float x : w) max_abs = std::max(max_abs, std::fabs(x));
    const float inv = max_abs > 0 ? 127.f / max_abs : 1.f;
    std::vector<unsigned char> img(w.size());
    for (size_t k = 0; k < w.size(); ++k)
        img[k] = (unsigned char)std::clamp(128 + int(w[k] * inv), 0, 255);
    const std::vector<unsigned char> png = vpmc_viz::encode_png(img, res, res);
    std::ofstream f(file, std::ios::binary);
    if (!f) throw std::runtime_error("dump_vorticity_png: cannot open " + file);
    f.write(reinterpret_cast<const char *>(png.data()), (std::streamsize)png.size());
}

inline void dump_velocity_svg_centered(const Simulation &sim, const std::string &file,
                                       float vel_scale = 0.2f, bool periodic = false)
{
    const std::filesystem::path parent = std::filesystem::path(file).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
    const int N = sim.particles_.N_cur;
    const std::vector<float2> pos = sim.particles_.pos.to_host();
    const std::vector<float2> vel = sim.particles_.vel.to_host();
    std::ofstream o(file);
    if (!o.is_open()) {
        std::cerr << "x open " << file << " failed\n";
        return;
    }
    o << "<?xml version=\"1.0\" standalone=\"no\"?>\n"
      << "<svg xmlns=\"http://www.w3.org/2000/svg\" width=\"600\" height=\"600\" "
         "viewBox=\"0 0 1 1\">\n"
      << "<style>.arrow{stroke:#ff4d00;stroke-width:0.0015;marker-end:url(#head);}</style>\n"
      << "<defs><marker id=\"head\" orient=\"auto\" markerWidth=\"4\" markerHeight=\"4\" "
         "refX=\"0.1\" refY=\"2\"><path d=\"M0,0 V4 L4,2 Z\" fill=\"#ff4d00\"/></marker></defs>\n";
    for (int i = 0; i < N && (size_t)i < vel.size(); ++i) {
        const float len = length(vel[i]);
        if (len < 1e-8f) continue;
        const float ux = vel[i].x / len, uy = vel[i].y / len, half = 0.5f * vel_scale * len;
        float ax = pos[i].x - half * ux, ay = pos[i].y - half * uy;
        float bx = pos[i].x + half * ux, by = pos[i].y + half * uy;
        if (periodic) {
            ax -= std::floor(ax);
            ay -= std::floor(ay);
            bx -= std::floor(bx);
            by -= std::floor(by);
        }
        o << "<line class=\"arrow\" x1=\"" << ax << "\" y1=\"" << 1 - ay << "\" x2=\"" << bx
          << "\" y2=\"" << 1 - by << "\"/>\n";
    }
    o << "</svg>\n";
}
