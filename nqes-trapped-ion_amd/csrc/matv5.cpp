// matv5.cpp -- MAT v5 reader/writer (see matv5.hpp).
//
// Layout (MATLAB "MAT-File Format", level 5): a 128-byte header (116 bytes of text,
// 8 bytes subsystem offset, version 0x0100, endian indicator "IM"), then data
// elements, each an 8-byte tag (type, byte count) plus data padded to 8 bytes, or a
// "small" element with byte count and type packed in the first 4 bytes and up to 4
// data bytes in the next 4.  A numeric variable is an miMATRIX element holding array
// flags (miUINT32), dimensions (miINT32), the name (miINT8), the real part and, when
// the complex flag (0x0800) is set, the imaginary part.
#include "matv5.hpp"

#include <zlib.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <ctime>
#include <stdexcept>

namespace matv5 {

namespace {

enum {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6,
    miSINGLE = 7, miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15
};

struct Reader {
    const uint8_t *p;
    size_t n, off = 0;
    std::string file;
    [[noreturn]] void fail(const std::string &why) const
    {
        throw std::runtime_error(file + ": malformed MAT v5 file (" + why + ")");
    }
    uint32_t u32(size_t at) const
    {
        if (at + 4 > n) fail("truncated");
        uint32_t v;
        memcpy(&v, p + at, 4);
        return v;
    }
};

// one element at r.off: returns type, data pointer and byte count; advances r.off
bool next_element(Reader &r, uint32_t &type, const uint8_t *&data, uint32_t &nbytes)
{
    if (r.off + 8 > r.n) return false;
    uint32_t w0 = r.u32(r.off);
    if (w0 >> 16) {  // small data element
        type = w0 & 0xffff;
        nbytes = w0 >> 16;
        if (nbytes > 4) r.fail("small element longer than 4 bytes");
        data = r.p + r.off + 4;
        r.off += 8;
        return true;
    }
    type = w0;
    nbytes = r.u32(r.off + 4);
    if (r.off + 8 + (size_t)nbytes > r.n) r.fail("element runs past the end");
    data = r.p + r.off + 8;
    size_t adv = 8 + (size_t)nbytes;
    if (type != miCOMPRESSED) adv = (adv + 7) & ~(size_t)7;
    r.off += adv;
    return true;
}

void to_double(uint32_t type, const uint8_t *d, uint32_t nbytes, std::vector<double> &out, const Reader &r)
{
    size_t esz;
    switch (type) {
    case miINT8: case miUINT8: esz = 1; break;
    case miINT16: case miUINT16: esz = 2; break;
    case miINT32: case miUINT32: case miSINGLE: esz = 4; break;
    case miDOUBLE: case miINT64: case miUINT64: esz = 8; break;
    default: r.fail("unsupported numeric type " + std::to_string(type));
    }
    size_t cnt = nbytes / esz;
    out.resize(cnt);
    for (size_t i = 0; i < cnt; i++) {
        const uint8_t *q = d + i * esz;
        switch (type) {
        case miINT8: out[i] = (int8_t)q[0]; break;
        case miUINT8: out[i] = q[0]; break;
        case miINT16: { int16_t v; memcpy(&v, q, 2); out[i] = v; break; }
        case miUINT16: { uint16_t v; memcpy(&v, q, 2); out[i] = v; break; }
        case miINT32: { int32_t v; memcpy(&v, q, 4); out[i] = v; break; }
        case miUINT32: { uint32_t v; memcpy(&v, q, 4); out[i] = v; break; }
        case miSINGLE: { float v; memcpy(&v, q, 4); out[i] = v; break; }
        case miDOUBLE: { double v; memcpy(&v, q, 8); out[i] = v; break; }
        case miINT64: { int64_t v; memcpy(&v, q, 8); out[i] = (double)v; break; }
        case miUINT64: { uint64_t v; memcpy(&v, q, 8); out[i] = (double)v; break; }
        }
    }
}

void parse_matrix(const uint8_t *d, uint32_t nbytes, const std::string &file, std::vector<Var> &out)
{
    Reader r{d, nbytes, 0, file};
    uint32_t type, nb;
    const uint8_t *el;
    if (!next_element(r, type, el, nb) || type != miUINT32 || nb < 8) r.fail("array flags");
    uint32_t flags;
    memcpy(&flags, el, 4);
    Var v;
    v.mclass = flags & 0xff;
    v.complex_ = (flags & 0x0800) != 0;
    if (v.mclass < 6 || v.mclass > 15) return;  // cell, struct, char, sparse...: skipped
    if (!next_element(r, type, el, nb) || type != miINT32) r.fail("dimensions");
    for (uint32_t i = 0; i < nb / 4; i++) {
        int32_t dd;
        memcpy(&dd, el + 4 * i, 4);
        v.dims.push_back(dd);
    }
    if (!next_element(r, type, el, nb) || (type != miINT8 && type != miUINT8)) r.fail("array name");
    v.name.assign((const char *)el, nb);
    if (!next_element(r, type, el, nb)) r.fail("real part");
    to_double(type, el, nb, v.re, r);
    if (v.complex_) {
        if (!next_element(r, type, el, nb)) r.fail("imaginary part");
        to_double(type, el, nb, v.im, r);
    }
    if ((int64_t)v.re.size() != v.numel() || (v.complex_ && (int64_t)v.im.size() != v.numel()))
        r.fail("data size of " + v.name + " does not match its dimensions");
    out.push_back(std::move(v));
}

void parse_elements(const uint8_t *p, size_t n, const std::string &file, std::vector<Var> &out)
{
    Reader r{p, n, 0, file};
    uint32_t type, nb;
    const uint8_t *el;
    while (next_element(r, type, el, nb)) {
        if (type == miMATRIX) {
            parse_matrix(el, nb, file, out);
        } else if (type == miCOMPRESSED) {
            std::vector<uint8_t> buf(1 << 16);
            z_stream zs;
            memset(&zs, 0, sizeof(zs));
            if (inflateInit(&zs) != Z_OK) r.fail("zlib init");
            zs.next_in = const_cast<Bytef *>(el);
            zs.avail_in = nb;
            size_t have = 0;
            int ret;
            do {
                if (have == buf.size()) buf.resize(buf.size() * 2);
                zs.next_out = buf.data() + have;
                zs.avail_out = (uInt)(buf.size() - have);
                ret = inflate(&zs, Z_NO_FLUSH);
                have = buf.size() - zs.avail_out;
                if (ret != Z_OK && ret != Z_STREAM_END) {
                    inflateEnd(&zs);
                    r.fail("zlib inflate");
                }
            } while (ret != Z_STREAM_END);
            inflateEnd(&zs);
            parse_elements(buf.data(), have, file, out);
        }
    }
}

void put(std::vector<uint8_t> &b, const void *src, size_t n)
{
    const uint8_t *s = (const uint8_t *)src;
    b.insert(b.end(), s, s + n);
}

void put_u32(std::vector<uint8_t> &b, uint32_t v) { put(b, &v, 4); }

void pad8(std::vector<uint8_t> &b)
{
    while (b.size() % 8) b.push_back(0);
}

} // namespace

std::vector<Var> read_file(const std::string &path)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error(path + ": cannot open");
    std::vector<uint8_t> data;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) data.insert(data.end(), chunk, chunk + got);
    fclose(f);
    if (data.size() < 128) throw std::runtime_error(path + ": not a MAT v5 file (too short)");
    if (data[126] != 'I' || data[127] != 'M')
        throw std::runtime_error(path + ": not a little-endian MAT v5 file");
    std::vector<Var> out;
    parse_elements(data.data() + 128, data.size() - 128, path, out);
    return out;
}

const Var *find(const std::vector<Var> &vars, const std::string &name)
{
    for (const auto &v : vars)
        if (v.name == name) return &v;
    return nullptr;
}

void write_file(const std::string &path, const std::vector<Var> &vars)
{
    std::vector<uint8_t> b(128, ' ');
    char hdr[256];
    time_t now = time(nullptr);
    char tbuf[64];
    strftime(tbuf, sizeof(tbuf), "%a %b %d %H:%M:%S %Y", localtime(&now));
    snprintf(hdr, sizeof(hdr), "MATLAB 5.0 MAT-file, Platform: x86_64-pc-linux-gnu, Created by: nqes-mi355x on %s",
             tbuf);
    memcpy(b.data(), hdr, std::min<size_t>(strlen(hdr), 116));
    memset(b.data() + 116, 0, 8);
    b[124] = 0x00;
    b[125] = 0x01;
    b[126] = 'I';
    b[127] = 'M';
    for (const auto &v : vars) {
        std::vector<uint8_t> m;
        put_u32(m, miUINT32);
        put_u32(m, 8);
        put_u32(m, 6u | (v.complex_ ? 0x0800u : 0u));  // mxDOUBLE_CLASS
        put_u32(m, 0);
        put_u32(m, miINT32);
        put_u32(m, (uint32_t)(4 * v.dims.size()));
        for (auto d : v.dims) {
            int32_t dd = (int32_t)d;
            put(m, &dd, 4);
        }
        pad8(m);
        put_u32(m, miINT8);
        put_u32(m, (uint32_t)v.name.size());
        put(m, v.name.data(), v.name.size());
        pad8(m);
        put_u32(m, miDOUBLE);
        put_u32(m, (uint32_t)(8 * v.re.size()));
        put(m, v.re.data(), 8 * v.re.size());
        pad8(m);
        if (v.complex_) {
            put_u32(m, miDOUBLE);
            put_u32(m, (uint32_t)(8 * v.im.size()));
            put(m, v.im.data(), 8 * v.im.size());
            pad8(m);
        }
        put_u32(b, miMATRIX);
        put_u32(b, (uint32_t)m.size());
        put(b, m.data(), m.size());
    }
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error(path + ": cannot create");
    bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error(path + ": write failed");
}

} // namespace matv5
