// mc_tiff.cpp -- TIFF directories, strip/tile decoding and raw EER strips.
//
// The reference reads TIFF movies and EER files through libtiff:
// CLoadTiffHeader (TiffUtil/CLoadTiffHeader.cpp:18-147) takes the size,
// directory count, sample mode and strip/tile geometry of the first
// directory; CLoadTiffImage (TiffUtil/CLoadTiffImage.cpp:50-160) decodes a
// directory strip by strip (or tile by tile) with TIFFReadEncodedStrip and
// stores the rows bottom-up (TIFF row r -> image row H-1-r); CLoadEerFrames
// (EerUtil/CLoadEerFrames.cpp:72-95) concatenates a directory's raw strips
// (TIFFReadRawStrip).  This file reads the same structures directly from the
// file: classic TIFF and BigTIFF in either byte order; uncompressed, LZW
// (TIFF flavour: MSB-first codes, early width change), Deflate (zlib) and
// PackBits strips or tiles; horizontal-differencing predictor; 8-, 16-bit
// integer and 32-bit float single-sample images.
#include "mc_tiff.h"

#include <zlib.h>

#include <cstring>

namespace {

bool rd(FILE* f, uint64_t off, void* p, size_t n) {
    if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
    return fread(p, 1, n, f) == n;
}

uint16_t u16(const unsigned char* p, bool be) {
    return be ? (uint16_t)((p[0] << 8) | p[1]) : (uint16_t)(p[0] | (p[1] << 8));
}
uint32_t u32(const unsigned char* p, bool be) {
    return be ? ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]
              : (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
uint64_t u64(const unsigned char* p, bool be) {
    return be ? ((uint64_t)u32(p, true) << 32) | u32(p + 4, true)
              : (uint64_t)u32(p, false) | ((uint64_t)u32(p + 4, false) << 32);
}

int type_size(int type) {
    switch (type) {
        case 1: case 2: case 6: case 7: return 1;   // BYTE ASCII SBYTE UNDEFINED
        case 3: case 8: return 2;                   // SHORT SSHORT
        case 4: case 9: case 11: case 13: return 4; // LONG SLONG FLOAT IFD
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;  // RATIONAL DOUBLE LONG8 ...
        default: return 0;
    }
}

// Integer values of one directory entry (SHORT, LONG, LONG8 and the signed
// forms), or the RATIONAL / DOUBLE / FLOAT value as a double in `real`.
bool entry_values(TiffFile& t, int type, uint64_t count, const unsigned char* val, int inl,
                  uint64_t max_count, std::vector<uint64_t>& out, double* real) {
    const int ts = type_size(type);
    if (ts == 0 || count == 0 || count > max_count) return false;
    const uint64_t bytes = (uint64_t)ts * count;
    std::vector<unsigned char> buf((size_t)bytes);
    if (bytes <= (uint64_t)inl) {
        memcpy(buf.data(), val, (size_t)bytes);
    } else {
        const uint64_t off = inl == 8 ? u64(val, t.be) : u32(val, t.be);
        if (!rd(t.f, off, buf.data(), (size_t)bytes)) return false;
    }
    out.resize((size_t)count);
    for (uint64_t i = 0; i < count; i++) {
        const unsigned char* p = buf.data() + i * ts;
        switch (type) {
            case 1: case 6: case 7: out[i] = p[0]; break;
            case 3: case 8: out[i] = u16(p, t.be); break;
            case 4: case 9: case 13: out[i] = u32(p, t.be); break;
            case 16: case 17: case 18: out[i] = u64(p, t.be); break;
            case 5: {
                const double num = u32(p, t.be), den = u32(p + 4, t.be);
                if (real && i == 0) *real = den != 0 ? num / den : 0.0;
                out[i] = 0;
                break;
            }
            case 11: {
                uint32_t b = u32(p, t.be);
                float v;
                memcpy(&v, &b, 4);
                if (real && i == 0) *real = v;
                out[i] = 0;
                break;
            }
            case 12: {
                uint64_t b = u64(p, t.be);
                double v;
                memcpy(&v, &b, 8);
                if (real && i == 0) *real = v;
                out[i] = 0;
                break;
            }
            default: return false;
        }
    }
    return true;
}

// TIFF LZW: codes of 9..12 bits, most significant bit first; 256 clears the
// table, 257 ends the strip; the code width grows once the next free entry
// reaches 2^width - 1 (the "early change" of TIFF 6.0).
bool lzw_decode(const unsigned char* in, size_t n, unsigned char* out, size_t cap) {
    static thread_local std::vector<uint16_t> prefix(4096);
    static thread_local std::vector<unsigned char> suffix(4096), first(4096);
    static thread_local std::vector<uint16_t> len(4096);
    std::vector<unsigned char> stack(4096);
    for (int i = 0; i < 256; i++) {
        prefix[i] = 0xffff;
        suffix[i] = first[i] = (unsigned char)i;
        len[i] = 1;
    }
    size_t o = 0;
    uint64_t bitpos = 0;
    const uint64_t nbits = (uint64_t)n * 8;
    int width = 9, next = 258, old = -1;
    auto get = [&](int w) -> int {
        if (bitpos + (uint64_t)w > nbits) return 257;
        int v = 0;
        for (int i = 0; i < w; i++) {
            const uint64_t b = bitpos + i;
            v = (v << 1) | ((in[b >> 3] >> (7 - (b & 7))) & 1);
        }
        bitpos += w;
        return v;
    };
    auto emit = [&](int code) -> bool {
        int l = len[code];
        if (o + (size_t)l > cap) l = (int)(cap - o);  // a last strip may decode past the image
        int c = code, k = len[code];
        while (k > 0) {
            stack[--k] = suffix[c];
            c = prefix[c];
        }
        memcpy(out + o, stack.data(), (size_t)l);
        o += (size_t)l;
        return true;
    };
    while (o < cap) {
        int code = get(width);
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            code = get(width);
            if (code == 257) break;
            if (code > 255) return false;
            emit(code);
            old = code;
            continue;
        }
        if (old < 0) return false;
        if (code < next) {
            emit(code);
            if (next < 4096) {
                prefix[next] = (uint16_t)old;
                suffix[next] = first[code];
                first[next] = first[old];
                len[next] = (uint16_t)(len[old] + 1);
                next++;
            }
        } else if (code == next && next < 4096) {
            prefix[next] = (uint16_t)old;
            suffix[next] = first[old];
            first[next] = first[old];
            len[next] = (uint16_t)(len[old] + 1);
            next++;
            emit(code);
        } else {
            return false;
        }
        old = code;
        if (next >= (1 << width) - 1 && width < 12) width++;
    }
    if (o < cap) memset(out + o, 0, cap - o);
    return true;
}

bool packbits_decode(const unsigned char* in, size_t n, unsigned char* out, size_t cap) {
    size_t i = 0, o = 0;
    while (i < n && o < cap) {
        const int c = (signed char)in[i++];
        if (c >= 0) {
            const size_t k = std::min<size_t>((size_t)c + 1, std::min(n - i, cap - o));
            memcpy(out + o, in + i, k);
            i += (size_t)c + 1;
            o += k;
        } else if (c != -128) {
            if (i >= n) break;
            const size_t k = std::min<size_t>((size_t)(1 - c), cap - o);
            memset(out + o, in[i++], k);
            o += k;
        }
    }
    if (o < cap) memset(out + o, 0, cap - o);
    return true;
}

// One strip or tile: raw bytes -> `cap` decoded bytes in file byte order.
bool decode_block(TiffFile& t, const TiffDir& d, size_t i, unsigned char* out, size_t cap,
                  std::vector<unsigned char>& raw) {
    if (i >= d.off.size() || i >= d.cnt.size()) return false;
    raw.resize((size_t)d.cnt[i]);
    if (!raw.empty() && !rd(t.f, d.off[i], raw.data(), raw.size())) return false;
    switch (d.comp) {
        case 1:
            memcpy(out, raw.data(), std::min(cap, raw.size()));
            if (raw.size() < cap) memset(out + raw.size(), 0, cap - raw.size());
            return true;
        case 5: return lzw_decode(raw.data(), raw.size(), out, cap);
        case 8:
        case 32946: {
            uLongf got = (uLongf)cap;
            const int rc = uncompress(out, &got, raw.data(), (uLong)raw.size());
            if (rc != Z_OK && rc != Z_BUF_ERROR) return false;
            if (got < cap) memset(out + got, 0, cap - got);
            return true;
        }
        case 32773: return packbits_decode(raw.data(), raw.size(), out, cap);
        default: return false;
    }
}

// Byte order and horizontal differencing of decoded rows (w samples each).
void fix_rows(const TiffDir& d, bool be, unsigned char* p, int rows, int w) {
    const int bs = d.bits / 8;
    for (int r = 0; r < rows; r++) {
        unsigned char* row = p + (size_t)r * w * bs;
        if (bs > 1 && be) {  // samples to the (little-endian) host order
            for (int x = 0; x < w; x++) {
                unsigned char* s = row + (size_t)x * bs;
                for (int k = 0; k < bs / 2; k++) std::swap(s[k], s[bs - 1 - k]);
            }
        }
        if (d.pred == 2) {
            if (bs == 1) {
                for (int x = 1; x < w; x++) row[x] = (unsigned char)(row[x] + row[x - 1]);
            } else if (bs == 2) {
                uint16_t* s = reinterpret_cast<uint16_t*>(row);
                for (int x = 1; x < w; x++) s[x] = (uint16_t)(s[x] + s[x - 1]);
            } else if (bs == 4) {
                uint32_t* s = reinterpret_cast<uint32_t*>(row);
                for (int x = 1; x < w; x++) s[x] = s[x] + s[x - 1];
            }
        }
    }
}

}  // namespace

bool tiff_open(const char* path, TiffFile& t) {
    t.f = fopen(path, "rb");
    if (!t.f) {
        t.err = std::string("cannot open ") + path;
        return false;
    }
    if (fseeko(t.f, 0, SEEK_END) != 0) {
        t.err = "cannot size file";
        return false;
    }
    t.size = (uint64_t)ftello(t.f);
    unsigned char h[16];
    if (!rd(t.f, 0, h, 8)) {
        t.err = "short file";
        return false;
    }
    if (h[0] == 'I' && h[1] == 'I') t.be = false;
    else if (h[0] == 'M' && h[1] == 'M') t.be = true;
    else {
        t.err = "not a TIFF file";
        return false;
    }
    const int magic = u16(h + 2, t.be);
    const bool big = magic == 43;
    if (magic != 42 && !big) {
        t.err = "not a TIFF file";
        return false;
    }
    uint64_t ifd = 0;
    if (big) {
        if (!rd(t.f, 0, h, 16) || u16(h + 4, t.be) != 8) {
            t.err = "bad BigTIFF header";
            return false;
        }
        ifd = u64(h + 8, t.be);
    } else {
        ifd = u32(h + 4, t.be);
    }
    const int es = big ? 20 : 12, inl = big ? 8 : 4;
    std::vector<uint64_t> seen;
    while (ifd != 0) {
        for (uint64_t s : seen)
            if (s == ifd) {
                t.err = "directory loop";
                return false;
            }
        seen.push_back(ifd);
        unsigned char cb[8];
        if (!rd(t.f, ifd, cb, big ? 8 : 2)) {
            t.err = "truncated directory";
            return false;
        }
        const uint64_t ne = big ? u64(cb, t.be) : u16(cb, t.be);
        if (ne == 0 || ne > 4096) {
            t.err = "bad directory";
            return false;
        }
        std::vector<unsigned char> ents((size_t)(ne * es + inl));
        if (!rd(t.f, ifd + (big ? 8 : 2), ents.data(), ents.size())) {
            t.err = "truncated directory";
            return false;
        }
        TiffDir d;
        std::vector<uint64_t> v;
        for (uint64_t e = 0; e < ne; e++) {
            const unsigned char* p = ents.data() + e * es;
            const int tag = u16(p, t.be), type = u16(p + 2, t.be);
            const uint64_t count = big ? u64(p + 4, t.be) : u32(p + 4, t.be);
            const unsigned char* val = p + (big ? 12 : 8);
            double real = 0;
            const bool arr = tag == 273 || tag == 279 || tag == 324 || tag == 325;
            if (!arr && tag != 256 && tag != 257 && tag != 258 && tag != 259 && tag != 277 &&
                tag != 278 && tag != 282 && tag != 283 && tag != 317 && tag != 322 && tag != 323 &&
                tag != 339)
                continue;
            if (!entry_values(t, type, count, val, inl, arr ? (1u << 24) : 16, v, &real)) {
                t.err = "bad directory entry " + std::to_string(tag);
                return false;
            }
            switch (tag) {
                case 256: d.w = (int)v[0]; break;
                case 257: d.h = (int)v[0]; break;
                case 258: d.bits = (int)v[0]; break;
                case 259: d.comp = (int)v[0]; break;
                case 273: case 324: d.off = v; break;
                case 277: d.spp = (int)v[0]; break;
                case 278: d.rps = (int)std::min<uint64_t>(v[0], 1u << 30); break;
                case 279: case 325: d.cnt = v; break;
                case 282: d.has_xres = true; d.xres = real; break;
                case 283: d.has_yres = true; d.yres = real; break;
                case 317: d.pred = (int)v[0]; break;
                case 322: d.tw = (int)v[0]; break;
                case 323: d.th = (int)v[0]; break;
                case 339: d.fmt = (int)v[0]; break;
            }
        }
        // RowsPerStrip defaults to 2^32 - 1 (one strip); a strip never
        // holds more rows than the image has
        if (d.rps <= 0 || d.rps > d.h) d.rps = d.h;
        // every block inside the file; image and tile sizes bounded
        for (size_t i = 0; i < d.off.size() && i < d.cnt.size(); i++)
            if (d.off[i] > t.size || d.cnt[i] > t.size - d.off[i]) {
                t.err = "strip or tile beyond the end of the file";
                return false;
            }
        if (d.w < 0 || d.h < 0 || (uint64_t)d.w * (uint64_t)d.h > (1ull << 34) ||
            (uint64_t)d.tw * (uint64_t)d.th > (1ull << 28) || d.bits < 0 || d.bits > 64) {
            t.err = "image or tile size out of range";
            return false;
        }
        t.dirs.push_back(std::move(d));
        ifd = big ? u64(ents.data() + ne * es, t.be) : u32(ents.data() + ne * es, t.be);
    }
    if (t.dirs.empty()) {
        t.err = "no image directory";
        return false;
    }
    return true;
}

int tiff_mode(const TiffDir& d) {
    // TiffUtil/CLoadTiffHeader.cpp:111-135
    if (d.spp != 1) return -1;
    if (d.bits == 8) return 0;
    if (d.bits == 16) return d.fmt == 2 ? 1 : 6;
    if (d.fmt == 3 && d.bits == 32) return 2;
    return -1;
}

bool tiff_decode(TiffFile& t, const TiffDir& d, unsigned char* out) {
    const int mode = tiff_mode(d);
    if (mode < 0 || d.w < 1 || d.h < 1) return false;
    if (d.pred != 1 && d.pred != 2) return false;
    const bool tiled = d.tw > 0 && d.th > 0;
    const size_t nblk = tiled ? (size_t)((d.w + d.tw - 1) / d.tw) * ((d.h + d.th - 1) / d.th)
                              : (size_t)((d.h + d.rps - 1) / d.rps);
    if (d.off.size() < nblk || d.cnt.size() < nblk) return false;
    const int bs = d.bits / 8;
    const size_t row = (size_t)d.w * bs;
    std::vector<unsigned char> raw, blk;
    if (d.tw > 0 && d.th > 0) {
        // TiffUtil/CLoadTiffImage.cpp:118-160: tiles placed bottom-up
        const int ntx = (d.w + d.tw - 1) / d.tw, nty = (d.h + d.th - 1) / d.th;
        blk.resize((size_t)d.tw * d.th * bs);
        for (int ty = 0; ty < nty; ty++)
            for (int tx = 0; tx < ntx; tx++) {
                if (!decode_block(t, d, (size_t)ty * ntx + tx, blk.data(), blk.size(), raw)) return false;
                fix_rows(d, t.be, blk.data(), d.th, d.tw);
                const int x0 = tx * d.tw, nx = std::min(d.tw, d.w - x0);
                for (int r = 0; r < d.th; r++) {
                    const int y = ty * d.th + r;
                    if (y >= d.h) break;
                    memcpy(out + (size_t)(d.h - 1 - y) * row + (size_t)x0 * bs,
                           blk.data() + (size_t)r * d.tw * bs, (size_t)nx * bs);
                }
            }
        return true;
    }
    // TiffUtil/CLoadTiffImage.cpp:72-116: strips placed bottom-up
    const int ns = (d.h + d.rps - 1) / d.rps;
    blk.resize((size_t)d.rps * row);
    for (int s = 0; s < ns; s++) {
        const int rows = std::min(d.rps, d.h - s * d.rps);
        if (!decode_block(t, d, (size_t)s, blk.data(), (size_t)rows * row, raw)) return false;
        fix_rows(d, t.be, blk.data(), rows, d.w);
        for (int r = 0; r < rows; r++) {
            const int y = s * d.rps + r;
            memcpy(out + (size_t)(d.h - 1 - y) * row, blk.data() + (size_t)r * row, row);
        }
    }
    return true;
}

bool tiff_raw(TiffFile& t, const TiffDir& d, unsigned char* out, uint64_t* bytes) {
    uint64_t tot = 0;
    for (size_t i = 0; i < d.off.size() && i < d.cnt.size(); i++) {
        if (out && d.cnt[i] && !rd(t.f, d.off[i], out + tot, (size_t)d.cnt[i])) return false;
        tot += d.cnt[i];
    }
    *bytes = tot;
    return true;
}
