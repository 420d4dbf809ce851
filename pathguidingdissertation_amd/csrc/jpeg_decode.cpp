// jpeg_decode.cpp — JPEG decoding for the scene loader's textures (Texture::load reads
// JPEG files through stb_image, Imaging.h:62). Written from ITU-T T.81 (JPEG) and JFIF:
// baseline and extended sequential Huffman (SOF0/SOF1) and progressive Huffman (SOF2:
// spectral selection and successive approximation, EOB runs), restart intervals, any
// sampling factors up to 4, grayscale and YCbCr. Dequantised blocks go through a
// separable FP64 IDCT; subsampled chroma is upsampled with the triangle ("fancy")
// filter for 2x horizontal/vertical factors and by replication otherwise; YCbCr to RGB
// uses the JFIF matrix. Output samples are 8-bit, one or three channels per pixel (the
// channel count stb_image reports for these files). Arithmetic coding, lossless and
// hierarchical modes, 12-bit samples and CMYK are rejected with a reason.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_scene.hpp"

namespace pg {
namespace {

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    // canonical code tables (T.81 Annex C): per length the first code, the index of its
    // first value and the largest code; 9-bit fast lookup
    uint8_t vals[256] = {0};
    int32_t maxcode[18] = {0}, valptr[17] = {0}, mincode[17] = {0};
    uint16_t fast[512];  // (length << 8) | value, 0 = slow path
    bool ok = false;
    bool build(const uint8_t* counts, const uint8_t* v, int nv) {
        std::memcpy(vals, v, (size_t)nv);
        std::memset(fast, 0, sizeof(fast));
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            valptr[len] = k;
            mincode[len] = code;
            code += counts[len - 1];
            k += counts[len - 1];
            if (counts[len - 1]) maxcode[len] = code - 1;
            else maxcode[len] = -1;
            if (code > (1 << len)) return false;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        // fast table for codes of up to 9 bits
        code = 0;
        k = 0;
        for (int len = 1; len <= 9; ++len) {
            for (int i = 0; i < counts[len - 1]; ++i, ++k, ++code) {
                int shift = 9 - len;
                for (int f = 0; f < (1 << shift); ++f) fast[(code << shift) | f] = (uint16_t)((len << 8) | vals[k]);
            }
            code <<= 1;
        }
        ok = true;
        return true;
    }
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0;
    int bw = 0, bh = 0;          // blocks per line / column of the padded component
    std::vector<int16_t> coef;   // bw * bh * 64, natural order
    int dc_pred = 0;
    int td = 0, ta = 0;
};

struct Decoder {
    const uint8_t* d;
    size_t n, pos = 0;
    std::string err;
    uint16_t qt[4][64];
    Huff hdc[4], hac[4];
    std::vector<Comp> comps;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, adobe_rgb = false, have_adobe = false;
    int adobe_transform = -1;
    int restart = 0;
    int eobrun = 0;
    // bit reader
    uint32_t bitbuf = 0;
    int bitcnt = 0;
    bool marker_hit = false;

    Decoder(const uint8_t* p, size_t len) : d(p), n(len) { std::memset(qt, 0, sizeof(qt)); }

    int u8() { return pos < n ? d[pos++] : (pos++, 0); }
    int u16() { int a = u8(); return (a << 8) | u8(); }

    void reset_bits() { bitbuf = 0; bitcnt = 0; marker_hit = false; }
    void fill() {
        while (bitcnt <= 24) {
            int b = 0;
            if (!marker_hit && pos < n) {
                b = d[pos];
                if (b == 0xFF) {
                    int nx = pos + 1 < n ? d[pos + 1] : 0;
                    if (nx == 0x00) {
                        pos += 2;
                    } else {
                        marker_hit = true;  // leave the marker for the scan loop; feed zeros
                        b = 0;
                    }
                } else {
                    ++pos;
                }
            }
            bitbuf |= (uint32_t)b << (24 - bitcnt);
            bitcnt += 8;
        }
    }
    int bits(int k) {
        if (k == 0) return 0;
        if (bitcnt < k) fill();
        int v = (int)(bitbuf >> (32 - k));
        bitbuf <<= k;
        bitcnt -= k;
        return v;
    }
    int bit() { return bits(1); }
    int decode(const Huff& h) {
        if (bitcnt < 16) fill();
        int look = (int)(bitbuf >> (32 - 9));
        uint16_t f = h.fast[look];
        if (f) {
            int len = f >> 8;
            bitbuf <<= len;
            bitcnt -= len;
            return f & 0xFF;
        }
        int code = 0;
        for (int len = 1; len <= 16; ++len) {
            code = (code << 1) | bit();
            if (h.maxcode[len] >= 0 && code <= h.maxcode[len]) {
                int idx = h.valptr[len] + code - h.mincode[len];
                return (idx >= 0 && idx < 256) ? h.vals[idx] : 0;
            }
        }
        err = "bad Huffman code";
        return 0;
    }
    static int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }
    int receive_extend(int s) { return extend(bits(s), s); }
    // DC difference of a category from the stream: 8-bit samples allow categories
    // 0..11 (T.81 F.1.2.1.1); anything larger is a corrupt table
    int receive_dc(int t) {
        if (t > 11) {
            if (err.empty()) err = "bad DC category";
            return 0;
        }
        return receive_extend(t);
    }
    // the DC predictor stays inside the range a coefficient can hold
    static int clamp_dc(int64_t v) { return (int)std::min<int64_t>(std::max<int64_t>(v, -65536), 65535); }

    bool read_sof(int marker) {
        int len = u16();
        size_t end = pos + (size_t)len - 2;
        int p = u8();
        if (p != 8) { err = "only 8-bit JPEG samples are supported"; return false; }
        height = u16();
        width = u16();
        int nc = u8();
        if (width <= 0 || height <= 0) { err = "bad JPEG size (DNL not supported)"; return false; }
        if ((int64_t)width * height > kMaxImagePixels) { err = "JPEG above 2^28 pixels"; return false; }
        if (nc != 1 && nc != 3) { err = "only 1- or 3-component JPEGs are supported"; return false; }
        if (!comps.empty()) { err = "more than one frame header"; return false; }
        hmax = vmax = 1;
        comps.assign(nc, Comp());
        for (int i = 0; i < nc; ++i) {
            comps[i].id = u8();
            int hv = u8();
            comps[i].h = hv >> 4;
            comps[i].v = hv & 15;
            comps[i].tq = u8() & 3;
            if (comps[i].h < 1 || comps[i].h > 4 || comps[i].v < 1 || comps[i].v > 4) {
                err = "bad sampling factor";
                return false;
            }
            hmax = std::max(hmax, comps[i].h);
            vmax = std::max(vmax, comps[i].v);
        }
        // upsampling replicates each sample hmax / h times: factors must divide the maxima
        for (const Comp& c : comps)
            if (hmax % c.h != 0 || vmax % c.v != 0) {
                err = "sampling factors that do not divide the largest ones are not supported";
                return false;
            }
        progressive = marker == 0xC2;
        mcux = (width + 8 * hmax - 1) / (8 * hmax);
        mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (Comp& c : comps) {
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.coef.assign((size_t)c.bw * c.bh * 64, 0);
        }
        pos = end;
        return true;
    }

    bool read_dqt() {
        int len = u16();
        size_t end = pos + (size_t)len - 2;
        while (pos < end) {
            int pq = u8();
            int t = pq & 3, prec = pq >> 4;
            for (int k = 0; k < 64; ++k) qt[t][kZigzag[k]] = (uint16_t)(prec ? u16() : u8());
        }
        pos = end;
        return true;
    }

    bool read_dht() {
        int len = u16();
        size_t end = pos + (size_t)len - 2;
        while (pos < end) {
            int tc = u8();
            uint8_t counts[16];
            int total = 0;
            for (int i = 0; i < 16; ++i) { counts[i] = (uint8_t)u8(); total += counts[i]; }
            if (total > 256) { err = "bad Huffman table"; return false; }
            uint8_t v[256];
            for (int i = 0; i < total; ++i) v[i] = (uint8_t)u8();
            Huff& h = (tc >> 4) ? hac[tc & 3] : hdc[tc & 3];
            if (!h.build(counts, v, total)) { err = "bad Huffman table"; return false; }
        }
        pos = end;
        return true;
    }

    int16_t* block(Comp& c, int bx, int by) { return &c.coef[((size_t)by * c.bw + bx) * 64]; }

    // one block of a sequential scan (T.81 F.2.2)
    void decode_seq(Comp& c, int16_t* b) {
        int t = decode(hdc[c.td]);
        int diff = receive_dc(t);
        c.dc_pred = clamp_dc((int64_t)c.dc_pred + diff);
        b[0] = (int16_t)c.dc_pred;
        for (int k = 1; k < 64;) {
            int rs = decode(hac[c.ta]);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) break;
            b[kZigzag[k]] = (int16_t)receive_extend(s);
            ++k;
        }
    }
    // progressive scans (T.81 G.1.2)
    void decode_dc_prog(Comp& c, int16_t* b, int ah, int al) {
        if (ah == 0) {
            int t = decode(hdc[c.td]);
            c.dc_pred = clamp_dc((int64_t)c.dc_pred + receive_dc(t));
            b[0] = (int16_t)((int64_t)c.dc_pred * (int64_t(1) << al));
        } else if (bit()) {
            b[0] = (int16_t)(b[0] | (1 << al));
        }
    }
    void decode_ac_first(Comp& c, int16_t* b, int ss, int se, int al) {
        if (eobrun > 0) { --eobrun; return; }
        for (int k = ss; k <= se;) {
            int rs = decode(hac[c.ta]);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {
                    eobrun = (1 << r) - 1;
                    if (r) eobrun += bits(r);
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) break;
            b[kZigzag[k]] = (int16_t)(receive_extend(s) * (1 << al));
            ++k;
        }
    }
    void decode_ac_refine(Comp& c, int16_t* b, int ss, int se, int al) {
        const int p1 = 1 << al, m1 = -1 * (1 << al);
        int k = ss;
        if (eobrun <= 0) {
            for (; k <= se;) {
                int rs = decode(hac[c.ta]);
                int r = rs >> 4, s = rs & 15;
                int val = 0;
                if (s == 0) {
                    if (r < 15) {
                        eobrun = 1 << r;
                        if (r) eobrun += bits(r);
                        break;  // the remaining coefficients are refined below
                    }
                } else {
                    val = bit() ? p1 : m1;
                }
                while (k <= se) {
                    int16_t& cf = b[kZigzag[k]];
                    if (cf != 0) {
                        if (bit() && (cf & p1) == 0) cf = (int16_t)(cf >= 0 ? cf + p1 : cf + m1);
                    } else {
                        if (r == 0) {
                            if (val) cf = (int16_t)val;
                            ++k;
                            break;
                        }
                        --r;
                    }
                    ++k;
                }
            }
        }
        if (eobrun > 0) {
            for (; k <= se; ++k) {
                int16_t& cf = b[kZigzag[k]];
                if (cf != 0 && bit() && (cf & p1) == 0) cf = (int16_t)(cf >= 0 ? cf + p1 : cf + m1);
            }
            --eobrun;
        }
    }

    bool read_scan() {
        int len = u16();
        size_t end = pos + (size_t)len - 2;
        int ns = u8();
        std::vector<int> idx;
        for (int i = 0; i < ns; ++i) {
            int cid = u8(), t = u8();
            int k = -1;
            for (size_t j = 0; j < comps.size(); ++j)
                if (comps[j].id == cid) k = (int)j;
            if (k < 0) { err = "scan names an unknown component"; return false; }
            comps[k].td = t >> 4;
            comps[k].ta = t & 15;
            if (comps[k].td > 3 || comps[k].ta > 3) { err = "bad table selector"; return false; }
            idx.push_back(k);
        }
        int ss = u8(), se = u8(), a = u8();
        int ah = a >> 4, al = a & 15;
        pos = end;
        if (!progressive) { ss = 0; se = 63; }
        if (ss > se || se > 63) { err = "bad spectral selection"; return false; }
        for (int k : idx) {
            if (ss == 0 && (progressive ? ah == 0 : true) && !hdc[comps[k].td].ok) { err = "missing DC table"; return false; }
            if (se > 0 && (!progressive || ss > 0) && !hac[comps[k].ta].ok) { err = "missing AC table"; return false; }
        }
        reset_bits();
        for (Comp& c : comps) c.dc_pred = 0;
        eobrun = 0;
        int restarts_left = restart;
        auto block_op = [&](Comp& c, int16_t* b) {
            if (!progressive) decode_seq(c, b);
            else if (ss == 0) decode_dc_prog(c, b, ah, al);
            else if (ah == 0) decode_ac_first(c, b, ss, se, al);
            else decode_ac_refine(c, b, ss, se, al);
        };
        auto handle_restart = [&]() -> bool {
            if (restart == 0) return true;
            if (--restarts_left > 0) return true;
            restarts_left = restart;
            // expect an RSTn marker: drop buffered bits, find the marker
            reset_bits();
            while (pos + 1 < n && !(d[pos] == 0xFF && d[pos + 1] >= 0xD0 && d[pos + 1] <= 0xD7)) {
                if (d[pos] == 0xFF && d[pos + 1] != 0 && d[pos + 1] != 0xFF) return false;  // another marker: stop
                ++pos;
            }
            if (pos + 1 < n) pos += 2;
            for (Comp& c : comps) c.dc_pred = 0;
            eobrun = 0;
            return true;
        };
        if (idx.size() == 1) {
            // non-interleaved: the component's own block grid, clipped to its size
            Comp& c = comps[idx[0]];
            int cw = (width * c.h + hmax - 1) / hmax, chh = (height * c.v + vmax - 1) / vmax;
            int nbx = (cw + 7) / 8, nby = (chh + 7) / 8;
            for (int by = 0; by < nby && err.empty(); ++by)
                for (int bx = 0; bx < nbx && err.empty(); ++bx) {
                    block_op(c, block(c, bx, by));
                    if (!handle_restart()) return true;
                }
        } else {
            for (int my = 0; my < mcuy && err.empty(); ++my)
                for (int mx = 0; mx < mcux && err.empty(); ++mx) {
                    for (int k : idx) {
                        Comp& c = comps[k];
                        for (int y = 0; y < c.v; ++y)
                            for (int x = 0; x < c.h; ++x) block_op(c, block(c, mx * c.h + x, my * c.v + y));
                    }
                    if (!handle_restart()) return true;
                }
        }
        // continue after the entropy-coded segment: the next marker
        while (pos + 1 < n && !(d[pos] == 0xFF && d[pos + 1] != 0x00 && !(d[pos + 1] >= 0xD0 && d[pos + 1] <= 0xD7)))
            ++pos;
        return err.empty();
    }

    bool parse() {
        if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) { err = "not a JPEG file"; return false; }
        pos = 2;
        bool frame = false;
        while (pos < n) {
            if (d[pos] != 0xFF) { ++pos; continue; }
            int m = d[pos + 1];
            pos += 2;
            if (m == 0xFF) { --pos; continue; }
            if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) continue;
            if (m == 0xD9) break;
            if (m == 0xC0 || m == 0xC1 || m == 0xC2) {
                if (!read_sof(m)) return false;
                frame = true;
            } else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                err = "unsupported JPEG process (lossless, hierarchical or arithmetic)";
                return false;
            } else if (m == 0xC4) {
                if (!read_dht()) return false;
            } else if (m == 0xDB) {
                if (!read_dqt()) return false;
            } else if (m == 0xDD) {
                u16();
                restart = u16();
            } else if (m == 0xDA) {
                if (!frame) { err = "scan before frame"; return false; }
                if (!read_scan()) return false;
            } else if (m == 0xEE) {  // Adobe APP14: colour transform flag
                int len = u16();
                size_t end = pos + (size_t)len - 2;
                if (len >= 14 && pos + 12 <= n && std::memcmp(d + pos, "Adobe", 5) == 0) {
                    have_adobe = true;
                    adobe_transform = d[pos + 11];
                }
                pos = end;
            } else {
                int len = u16();
                pos += (size_t)len - 2;
            }
        }
        if (!frame) { err = "no JPEG frame"; return false; }
        return true;
    }
};

// FP64 8x8 inverse DCT (T.81 A.3.3) of one dequantised block, level shift and clamp.
void idct_block(const int16_t* coef, const uint16_t* q, uint8_t* out, int stride) {
    static double cosv[8][8];
    static bool init = false;
    if (!init) {
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                cosv[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0);
        init = true;
    }
    double f[64], tmp[64];
    for (int i = 0; i < 64; ++i) f[i] = (double)coef[i] * (double)q[i];
    for (int y = 0; y < 8; ++y)  // rows: v fixed, over u
        for (int x = 0; x < 8; ++x) {
            double s = 0.0;
            for (int u = 0; u < 8; ++u) s += cosv[x][u] * f[y * 8 + u];
            tmp[y * 8 + x] = s;
        }
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) {
            double s = 0.0;
            for (int v = 0; v < 8; ++v) s += cosv[y][v] * tmp[v * 8 + x];
            double r = std::floor(s * 0.25 + 128.5);
            out[y * stride + x] = (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
        }
}

inline uint8_t clamp8(double v) {
    double r = std::floor(v + 0.5);
    return (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
}

}  // namespace

std::string read_jpeg(const std::string& path, Image& out) {
    std::string buf;
    {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) return "cannot read";
        std::fseek(f, 0, SEEK_END);
        long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (sz <= 0) { std::fclose(f); return "empty file"; }
        buf.resize((size_t)sz);
        size_t got = std::fread(&buf[0], 1, (size_t)sz, f);
        std::fclose(f);
        if (got != (size_t)sz) return "cannot read";
    }
    Decoder dec((const uint8_t*)buf.data(), buf.size());
    if (!dec.parse()) return dec.err;
    const int W = dec.width, H = dec.height, nc = (int)dec.comps.size();
    // component planes at their own resolution
    std::vector<std::vector<uint8_t>> planes(nc);
    std::vector<int> pw(nc), ph(nc);
    for (int ci = 0; ci < nc; ++ci) {
        Comp& c = dec.comps[ci];
        pw[ci] = c.bw * 8;
        ph[ci] = c.bh * 8;
        planes[ci].assign((size_t)pw[ci] * ph[ci], 0);
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx)
                idct_block(dec.block(c, bx, by), dec.qt[c.tq], &planes[ci][(size_t)by * 8 * pw[ci] + bx * 8], pw[ci]);
    }
    out.w = W;
    out.h = H;
    out.channels = nc == 1 ? 1 : 3;
    out.u8.assign((size_t)W * H * out.channels, 0);
    // sample of component ci at full-resolution pixel (x, y): triangle filter for 2x
    // factors (each output sample 3/4 nearest + 1/4 next source sample per axis),
    // replication for other ratios
    auto sample = [&](int ci, int x, int y) -> double {
        const Comp& c = dec.comps[ci];
        int fx = dec.hmax / c.h, fy = dec.vmax / c.v;
        const std::vector<uint8_t>& p = planes[ci];
        int w = pw[ci];
        int cw = (W * c.h + dec.hmax - 1) / dec.hmax, chh = (H * c.v + dec.vmax - 1) / dec.vmax;
        auto at = [&](int sx, int sy) -> double {
            sx = sx < 0 ? 0 : (sx >= cw ? cw - 1 : sx);
            sy = sy < 0 ? 0 : (sy >= chh ? chh - 1 : sy);
            return p[(size_t)sy * w + sx];
        };
        if (fx == 1 && fy == 1) return at(x, y);
        if ((fx == 2 || fx == 1) && (fy == 2 || fy == 1)) {
            // integer triangle filter rounded to 8 bits, as libjpeg's fancy upsampling:
            // 2x1: (3 near + far + 2) / 4; 2x2: (9 near + 3 + 3 + far + 8) / 16
            int sx = x / fx, sy = y / fy;
            int nx = fx == 2 ? (x % 2 == 0 ? sx - 1 : sx + 1) : sx;
            int ny = fy == 2 ? (y % 2 == 0 ? sy - 1 : sy + 1) : sy;
            if (fx == 2 && fy == 2)
                return (double)(((int)(9 * at(sx, sy) + 3 * at(nx, sy) + 3 * at(sx, ny) + at(nx, ny)) + 8) >> 4);
            if (fx == 2) return (double)(((int)(3 * at(sx, sy) + at(nx, sy)) + 2) >> 2);
            return (double)(((int)(3 * at(sx, sy) + at(sx, ny)) + 2) >> 2);
        }
        return at(x / fx, y / fy);
    };
    // JFIF YCbCr -> RGB, unless an Adobe marker says the components are RGB
    bool ycc = nc == 3 && !(dec.have_adobe && dec.adobe_transform == 0);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint8_t* o = &out.u8[((size_t)y * W + x) * out.channels];
            if (nc == 1) {
                o[0] = clamp8(sample(0, x, y));
                continue;
            }
            double Y = sample(0, x, y), cb = sample(1, x, y), cr = sample(2, x, y);
            if (ycc) {
                o[0] = clamp8(Y + 1.402 * (cr - 128.0));
                o[1] = clamp8(Y - 0.344136 * (cb - 128.0) - 0.714136 * (cr - 128.0));
                o[2] = clamp8(Y + 1.772 * (cb - 128.0));
            } else {
                o[0] = clamp8(Y); o[1] = clamp8(cb); o[2] = clamp8(cr);
            }
        }
    return std::string();
}

}  // namespace pg
