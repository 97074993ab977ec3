// io.cpp -- readers and writers for the reference's input files: OSKAR
// binary (.vis) and the W-kernel FITS cubes (declared in mi355_io.h).
//
// Host C++ only.  The formats are the reference's (utils/binary, cfitsio as
// read_kernel.c uses it); the code is this library's own: one index of
// chunks per binary file, CRC-32C by a byte-wise table, FITS headers parsed
// card by card and data converted from big-endian on the fly.
#include "gg_status.hpp"
#include "mi355_io.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>

using ggs::fail;
using ggs::ok;

namespace {

#define IO_TRY(expr)                                                         \
    do { int s_ = (expr); if (s_ != MI355_OK) return s_; } while (0)

// ---------------------------------------------------------------------------
// CRC-32C (Castagnoli; reflected polynomial 0x82F63B78, initial value and
// final XOR 0xFFFFFFFF), the check code of OSKAR binary chunks.
// ---------------------------------------------------------------------------
struct Crc32c
{
    uint32_t t[256];
    Crc32c()
    {
        for (uint32_t i = 0; i < 256; ++i)
        {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
            t[i] = c;
        }
    }
    // Continues a finished code `crc` over more bytes and finishes it again.
    // A code equal to the initial value is taken as "nothing yet", the
    // convention of the OSKAR library's update routine (oskar_crc.c:130),
    // which the stored codes follow.
    uint32_t update(uint32_t crc, const void* data, size_t n) const
    {
        if (crc != 0xFFFFFFFFu) crc ^= 0xFFFFFFFFu;
        const unsigned char* p = static_cast<const unsigned char*>(data);
        for (size_t i = 0; i < n; ++i) crc = (crc >> 8) ^ t[(crc ^ p[i]) & 0xFFu];
        return crc ^ 0xFFFFFFFFu;
    }
    uint32_t compute(const void* data, size_t n) const { return update(0xFFFFFFFFu, data, n); }
};

const Crc32c& crc32c()
{
    static const Crc32c c;
    return c;
}

uint32_t le32(const unsigned char* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
uint64_t le64(const unsigned char* p)
{
    return (uint64_t)le32(p) | ((uint64_t)le32(p + 4) << 32);
}
void put_le32(unsigned char* p, uint32_t v)
{
    for (int i = 0; i < 4; ++i) p[i] = (unsigned char)(v >> (8 * i));
}
void put_le64(unsigned char* p, uint64_t v)
{
    for (int i = 0; i < 8; ++i) p[i] = (unsigned char)(v >> (8 * i));
}

// Size of one payload element for a data type code (0: unknown type).
size_t element_bytes(int type)
{
    size_t b = 0;
    if (type & MI355_BIN_CHAR) b = 1;
    else if (type & MI355_BIN_INT) b = 4;
    else if (type & MI355_BIN_SINGLE) b = 4;
    else if (type & MI355_BIN_DOUBLE) b = 8;
    else return 0;
    if (type & MI355_BIN_COMPLEX) b *= 2;
    if (type & MI355_BIN_MATRIX) b *= 4;
    return b;
}

struct Chunk
{
    bool extended = false;
    int data_type = 0, id_group = 0, id_tag = 0, user_index = 0;
    std::string name_group, name_tag;
    long long payload_offset = 0;
    size_t payload_size = 0;
    bool has_crc = false;
    uint32_t crc = 0, crc_header = 0;
};

} // namespace

struct mi355_bin
{
    FILE* f = nullptr;
    char mode = 'r';
    int version = 0;
    int search_start = 0;
    std::vector<Chunk> chunks;
};

namespace {

// Header: "OSKARBIN\0", format version, then (version 1 only) endianness
// and type sizes; 64 bytes in all (private_binary.h:43-54).
int read_header(mi355_bin* h, const char* path)
{
    unsigned char hd[64];
    if (fread(hd, 1, 64, h->f) != 64)
        return fail(MI355_ERR_FORMAT, "%s: shorter than an OSKAR binary header", path);
    if (memcmp(hd, "OSKARBIN", 8) != 0)
        return fail(MI355_ERR_FORMAT, "%s: not an OSKAR binary file", path);
    h->version = hd[9];
    if (h->version > 2)
        return fail(MI355_ERR_FORMAT, "%s: unknown OSKAR binary format version %d", path, h->version);
    if (h->version == 1)
    {
        // written by a machine of this byte order and these type sizes?
        if (hd[10] != 0 || hd[12] != sizeof(int) || hd[14] != sizeof(float) ||
                hd[15] != sizeof(double))
            return fail(MI355_ERR_FORMAT, "%s: version-1 file from an incompatible machine", path);
    }
    return MI355_OK;
}

// Tag: 'T', 0x40 + version, 'G', element size, flags, data type, group id
// (or name length), tag id (or name length), user index (LE32), block size
// (LE64) = bytes to the next tag (private_binary.h:70-128).
int index_chunks(mi355_bin* h, const char* path)
{
    if (fseeko(h->f, 0, SEEK_END) != 0) return fail(MI355_ERR_IO, "%s: seek failed", path);
    const long long file_size = ftello(h->f);
    if (fseeko(h->f, 64, SEEK_SET) != 0) return fail(MI355_ERR_IO, "%s: seek failed", path);
    for (;;)
    {
        unsigned char tag[20];
        const size_t got = fread(tag, 1, 20, h->f);
        if (got == 0) break;
        const size_t n = h->chunks.size();
        if (got != 20) return fail(MI355_ERR_FORMAT, "%s: truncated tag of chunk %zu", path, n);
        const int flags = tag[4];
        if (tag[0] != 'T' || tag[2] != 'G' || (flags & 0x1F) != 0)
            return fail(MI355_ERR_FORMAT, "%s: chunk %zu: not a tag", path, n);
        const int fv = tag[1] - 0x40;
        if (fv < 1 || fv > 2)
            return fail(MI355_ERR_FORMAT, "%s: chunk %zu: unknown format version %d", path, n, fv);
        Chunk c;
        c.data_type = tag[5];
        if (fv > 1)
        {
            // element size must match this machine's type (create.c:156-192)
            size_t el = tag[3];
            if (c.data_type & MI355_BIN_MATRIX) el /= 4;
            if (c.data_type & MI355_BIN_COMPLEX) el /= 2;
            size_t want = 0;
            if (c.data_type & MI355_BIN_CHAR) want = 1;
            else if (c.data_type & MI355_BIN_INT) want = sizeof(int);
            else if (c.data_type & MI355_BIN_SINGLE) want = sizeof(float);
            else if (c.data_type & MI355_BIN_DOUBLE) want = sizeof(double);
            if (!want || el != want)
                return fail(MI355_ERR_FORMAT, "%s: chunk %zu: unsupported data type 0x%02x / "
                        "element size %d", path, n, c.data_type, tag[3]);
        }
        c.id_group = tag[6];
        c.id_tag = tag[7];
        c.user_index = (int)le32(tag + 8);
        const uint64_t block = le64(tag + 12);
        c.has_crc = (flags & 0x40) != 0;
        uint64_t payload = block - (c.has_crc ? 4 : 0);
        uint32_t crc = crc32c().compute(tag, 20);
        if (flags & 0x80)
        {
            c.extended = true;
            const size_t lg = tag[6], lt = tag[7];
            std::vector<char> names(lg + lt);
            if (lg + lt > payload || fread(names.data(), 1, lg + lt, h->f) != lg + lt)
                return fail(MI355_ERR_FORMAT, "%s: chunk %zu: truncated tag names", path, n);
            payload -= lg + lt;
            c.name_group.assign(names.data(), strnlen(names.data(), lg));
            c.name_tag.assign(names.data() + lg, strnlen(names.data() + lg, lt));
            crc = crc32c().update(crc, names.data(), lg + lt);
        }
        c.payload_offset = ftello(h->f);
        c.payload_size = (size_t)payload;
        c.crc_header = crc;
        if (c.payload_offset + (long long)payload + (c.has_crc ? 4 : 0) > file_size)
            return fail(MI355_ERR_FORMAT, "%s: chunk %zu runs past the end of the file", path, n);
        if (fseeko(h->f, (off_t)payload, SEEK_CUR) != 0)
            return fail(MI355_ERR_IO, "%s: seek failed", path);
        if (c.has_crc)
        {
            unsigned char cb[4];
            if (fread(cb, 1, 4, h->f) != 4)
                return fail(MI355_ERR_FORMAT, "%s: chunk %zu: missing CRC", path, n);
            c.crc = le32(cb);
        }
        h->chunks.push_back(std::move(c));
    }
    return MI355_OK;
}

int query(const mi355_bin* h, unsigned char data_type, unsigned char id_group,
        unsigned char id_tag, int user_index, int* index)
{
    for (int i = std::max(0, h->search_start); i < (int)h->chunks.size(); ++i)
    {
        const Chunk& c = h->chunks[i];
        if (!c.extended && (c.data_type == (int)data_type || !data_type) &&
                c.id_group == (int)id_group && c.id_tag == (int)id_tag &&
                c.user_index == user_index)
        {
            *index = i;
            return MI355_OK;
        }
    }
    return fail(MI355_ERR_FORMAT, "tag (type 0x%02x, group %d, tag %d, index %d) not found "
            "at or after chunk %d", data_type, id_group, id_tag, user_index, h->search_start);
}

int read_chunk(mi355_bin* h, int index, size_t data_size, void* data)
{
    const Chunk& c = h->chunks[index];
    if (c.payload_size == 0) return MI355_OK;
    if (!data || data_size < c.payload_size)
        return fail(MI355_ERR_INVALID_ARG, "chunk %d holds %zu bytes, buffer has %zu", index,
                c.payload_size, data_size);
    if (fseeko(h->f, (off_t)c.payload_offset, SEEK_SET) != 0)
        return fail(MI355_ERR_IO, "seek to chunk %d failed", index);
    if (fread(data, 1, c.payload_size, h->f) != c.payload_size)
        return fail(MI355_ERR_IO, "read of chunk %d failed", index);
    if (c.has_crc && crc32c().update(c.crc_header, data, c.payload_size) != c.crc)
        return fail(MI355_ERR_FORMAT, "chunk %d: CRC-32C mismatch", index);
    return MI355_OK;
}

// ---------------------------------------------------------------------------
// FITS (NOST 100-2.0 / FITS 4.0): 2880-byte records, 80-character cards,
// big-endian data.
// ---------------------------------------------------------------------------
constexpr size_t kRecord = 2880;

struct Hdu
{
    std::map<std::string, std::string> keys;    // keyword -> raw value field
    long long data_offset = 0, data_bytes = 0;
    bool primary = false;
};

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) ++a;
    while (b > a && isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string upper(std::string s)
{
    for (char& ch : s) ch = (char)toupper((unsigned char)ch);
    return s;
}

// The value of a card (columns 11-80), without its comment.
std::string card_value(const char* card)
{
    std::string v(card + 10, 70);
    bool in_str = false;
    for (size_t i = 0; i < v.size(); ++i)
    {
        if (v[i] == '\'')
        {
            if (in_str && i + 1 < v.size() && v[i + 1] == '\'') { ++i; continue; }
            in_str = !in_str;
        }
        else if (v[i] == '/' && !in_str)
        {
            v.resize(i);
            break;
        }
    }
    return trim(v);
}

bool parse_string(const std::string& raw, std::string* out)
{
    if (raw.size() < 2 || raw[0] != '\'') return false;
    std::string s;
    for (size_t i = 1; i < raw.size(); ++i)
    {
        if (raw[i] == '\'')
        {
            if (i + 1 < raw.size() && raw[i + 1] == '\'') { s += '\''; ++i; continue; }
            break;
        }
        s += raw[i];
    }
    while (!s.empty() && s.back() == ' ') s.pop_back();   // trailing blanks are not significant
    *out = s;
    return true;
}

bool parse_double(const std::string& raw, double* out)
{
    std::string s = raw;
    for (char& ch : s) if (ch == 'D' || ch == 'd') ch = 'E';   // Fortran exponent
    if (s.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const double v = strtod(s.c_str(), &end);
    if (end == s.c_str() || trim(end).size() != 0) return false;
    *out = v;
    return true;
}

bool parse_ll(const std::string& raw, long long* out)
{
    if (raw.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const long long v = strtoll(raw.c_str(), &end, 10);
    if (end != raw.c_str() && trim(end).empty() && errno == 0) { *out = v; return true; }
    double d;
    if (!parse_double(raw, &d) || !(d > -9.3e18 && d < 9.3e18)) return false;
    *out = (long long)d;       // a real value read as an integer truncates
    return true;
}

int hdu_int(const Hdu& h, const char* key, long long* out, const char* path)
{
    auto it = h.keys.find(key);
    if (it == h.keys.end() || !parse_ll(it->second, out))
        return fail(MI355_ERR_FORMAT, "%s: missing or bad '%s' FITS header keyword", path, key);
    return MI355_OK;
}

int read_hdus(FILE* f, const char* path, std::vector<Hdu>* hdus)
{
    if (fseeko(f, 0, SEEK_END) != 0) return fail(MI355_ERR_IO, "%s: seek failed", path);
    const long long size = ftello(f);
    long long pos = 0;
    while (pos + (long long)kRecord <= size)
    {
        Hdu h;
        h.primary = hdus->empty();
        bool end = false;
        char rec[kRecord];
        while (!end)
        {
            if (fseeko(f, pos, SEEK_SET) != 0 || fread(rec, 1, kRecord, f) != kRecord)
                return fail(MI355_ERR_FORMAT, "%s: truncated FITS header", path);
            pos += kRecord;
            for (size_t c = 0; c < kRecord; c += 80)
            {
                const char* card = rec + c;
                const std::string key = trim(std::string(card, 8));
                if (key == "END") { end = true; break; }
                if (key.empty() || card[8] != '=' || card[9] != ' ') continue;   // commentary
                if (!h.keys.count(key)) h.keys[key] = card_value(card);
            }
        }
        if (h.primary && !h.keys.count("SIMPLE"))
            return fail(MI355_ERR_FORMAT, "%s: not a FITS file", path);
        long long bitpix = 0, naxis = 0;
        IO_TRY(hdu_int(h, "BITPIX", &bitpix, path));
        IO_TRY(hdu_int(h, "NAXIS", &naxis, path));
        long long n = naxis > 0 ? 1 : 0;
        for (long long i = 1; i <= naxis; ++i)
        {
            long long d = 0;
            const std::string k = "NAXIS" + std::to_string(i);
            IO_TRY(hdu_int(h, k.c_str(), &d, path));
            n *= d;
        }
        long long pcount = 0, gcount = 1;
        if (h.keys.count("PCOUNT")) IO_TRY(hdu_int(h, "PCOUNT", &pcount, path));
        if (h.keys.count("GCOUNT")) IO_TRY(hdu_int(h, "GCOUNT", &gcount, path));
        h.data_offset = pos;
        h.data_bytes = (std::llabs(bitpix) / 8) * gcount * (pcount + n);
        pos += (h.data_bytes + (long long)kRecord - 1) / (long long)kRecord * (long long)kRecord;
        hdus->push_back(std::move(h));
    }
    if (hdus->empty()) return fail(MI355_ERR_FORMAT, "%s: not a FITS file", path);
    return MI355_OK;
}

// Big-endian element i of a FITS array of the given BITPIX, as double.
double fits_element(const unsigned char* p, long long bitpix)
{
    unsigned char b[8];
    const int w = (int)(std::llabs(bitpix) / 8);
    for (int k = 0; k < w; ++k) b[k] = p[w - 1 - k];
    switch (bitpix)
    {
    case 8: return (double)p[0];
    case 16: { int16_t v; memcpy(&v, b, 2); return v; }
    case 32: { int32_t v; memcpy(&v, b, 4); return v; }
    case 64: { int64_t v; memcpy(&v, b, 8); return (double)v; }
    case -32: { float v; memcpy(&v, b, 4); return v; }
    default: { double v; memcpy(&v, b, 8); return v; }
    }
}

std::string fits_card(const std::string& key, const std::string& value, const char* comment = "")
{
    char card[81];
    snprintf(card, sizeof(card), "%-8.8s= %-70.70s", key.c_str(),
            (value + (comment[0] ? std::string(" / ") + comment : std::string())).c_str());
    return std::string(card, 80);
}
std::string fits_num(long long v)
{
    char b[32];
    snprintf(b, sizeof(b), "%20lld", v);
    return b;
}
std::string fits_real(double v)
{
    char b[40];
    snprintf(b, sizeof(b), "%.17G", v);
    std::string s = b;
    if (s.find_first_of(".EN") == std::string::npos) s += ".";   // a real, not an integer
    if (s.size() < 20) s = std::string(20 - s.size(), ' ') + s;
    return s;
}
std::string fits_str(const std::string& v)
{
    std::string s = v;
    while (s.size() < 8) s += ' ';
    return "'" + s + "'";
}

int write_record_padded(FILE* f, const std::string& bytes, char pad, const char* path)
{
    std::string b = bytes;
    if (b.size() % kRecord) b.append(kRecord - b.size() % kRecord, pad);
    if (fwrite(b.data(), 1, b.size(), f) != b.size())
        return fail(MI355_ERR_IO, "%s: write failed", path);
    return MI355_OK;
}

template <typename T>
std::string big_endian(const T* v, size_t n)
{
    std::string out(n * sizeof(T), '\0');
    for (size_t i = 0; i < n; ++i)
    {
        unsigned char b[sizeof(T)];
        memcpy(b, &v[i], sizeof(T));
        for (size_t k = 0; k < sizeof(T); ++k) out[i * sizeof(T) + k] = (char)b[sizeof(T) - 1 - k];
    }
    return out;
}

int write_image_hdu(FILE* f, const char* path, int precision, const long long* dims, int naxis,
        const void* data, const std::vector<std::string>& extra, bool extend)
{
    std::string hdr;
    hdr += fits_card("SIMPLE", fits_num(0).replace(19, 1, "T"));
    hdr += fits_card("BITPIX", fits_num(precision == MI355_BIN_DOUBLE ? -64 : -32));
    hdr += fits_card("NAXIS", fits_num(naxis));
    size_t n = 1;
    for (int i = 0; i < naxis; ++i)
    {
        hdr += fits_card("NAXIS" + std::to_string(i + 1), fits_num(dims[i]));
        n *= (size_t)dims[i];
    }
    if (extend) hdr += fits_card("EXTEND", fits_num(0).replace(19, 1, "T"));
    for (const std::string& c : extra) hdr += c;
    hdr += std::string("END") + std::string(77, ' ');
    IO_TRY(write_record_padded(f, hdr, ' ', path));
    const std::string body = precision == MI355_BIN_DOUBLE
            ? big_endian(static_cast<const double*>(data), n)
            : big_endian(static_cast<const float*>(data), n);
    return write_record_padded(f, body, '\0', path);
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

MI355_EXPORT int mi355_bin_open(const char* path, char mode, mi355_bin** handle)
{
    if (!path || !handle) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    *handle = nullptr;
    if (mode != 'r' && mode != 'w') return fail(MI355_ERR_INVALID_ARG, "mode must be 'r' or 'w'");
    FILE* f = fopen(path, mode == 'r' ? "rb" : "wb");
    if (!f) return fail(MI355_ERR_IO, "cannot open '%s': %s", path, strerror(errno));
    mi355_bin* h = new mi355_bin();
    h->f = f;
    h->mode = mode;
    int st = MI355_OK;
    if (mode == 'r')
    {
        st = read_header(h, path);
        if (st == MI355_OK) st = index_chunks(h, path);
    }
    else
    {
        unsigned char hd[64] = {0};
        memcpy(hd, "OSKARBIN", 8);
        hd[9] = 2;
        h->version = 2;
        if (fwrite(hd, 1, 64, f) != 64) st = fail(MI355_ERR_IO, "%s: write failed", path);
    }
    if (st != MI355_OK)
    {
        fclose(f);
        delete h;
        return st;
    }
    *handle = h;
    return ok();
}

MI355_EXPORT int mi355_bin_close(mi355_bin* h)
{
    if (!h) return ok();
    const int bad = fclose(h->f);
    const char mode = h->mode;
    delete h;
    if (bad && mode == 'w') return fail(MI355_ERR_IO, "close failed: %s", strerror(errno));
    return ok();
}

MI355_EXPORT int mi355_bin_num_tags(const mi355_bin* h)
{
    return h ? (int)h->chunks.size() : 0;
}

MI355_EXPORT int mi355_bin_set_query_search_start(mi355_bin* h, int start)
{
    if (!h) return fail(MI355_ERR_INVALID_ARG, "NULL handle");
    if (start < 0 || start >= (int)h->chunks.size())
        return fail(MI355_ERR_INVALID_ARG, "search start %d outside the %zu chunks", start,
                h->chunks.size());
    h->search_start = start;
    return ok();
}

MI355_EXPORT int mi355_bin_query(const mi355_bin* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index, size_t* payload_size,
        int* chunk_index)
{
    if (!h) return fail(MI355_ERR_INVALID_ARG, "NULL handle");
    int i = -1;
    IO_TRY(query(h, data_type, id_group, id_tag, user_index, &i));
    if (payload_size) *payload_size = h->chunks[i].payload_size;
    if (chunk_index) *chunk_index = i;
    return ok();
}

MI355_EXPORT int mi355_bin_read(mi355_bin* h, unsigned char data_type, unsigned char id_group,
        unsigned char id_tag, int user_index, size_t data_size, void* data)
{
    if (!h) return fail(MI355_ERR_INVALID_ARG, "NULL handle");
    if (h->mode != 'r') return fail(MI355_ERR_INVALID_ARG, "file not open for reading");
    int i = -1;
    IO_TRY(query(h, data_type, id_group, id_tag, user_index, &i));
    IO_TRY(read_chunk(h, i, data_size, data));
    return ok();
}

MI355_EXPORT int mi355_bin_write(mi355_bin* h, unsigned char data_type, unsigned char id_group,
        unsigned char id_tag, int user_index, size_t data_size, const void* data)
{
    if (!h) return fail(MI355_ERR_INVALID_ARG, "NULL handle");
    if (h->mode != 'w') return fail(MI355_ERR_INVALID_ARG, "file not open for writing");
    const size_t el = element_bytes(data_type);
    if (!el || el > 255) return fail(MI355_ERR_INVALID_ARG, "unknown data type 0x%02x", data_type);
    if (data_size && !data) return fail(MI355_ERR_INVALID_ARG, "NULL data");
    unsigned char tag[20] = {0};
    tag[0] = 'T'; tag[1] = 0x40 + 2; tag[2] = 'G'; tag[3] = (unsigned char)el;
    tag[4] = 0x40;                       // CRC-32C present
    tag[5] = data_type; tag[6] = id_group; tag[7] = id_tag;
    put_le32(tag + 8, (uint32_t)user_index);
    put_le64(tag + 12, (uint64_t)data_size + 4);
    uint32_t crc = crc32c().compute(tag, 20);
    crc = crc32c().update(crc, data, data_size);
    unsigned char cb[4];
    put_le32(cb, crc);
    if (fwrite(tag, 1, 20, h->f) != 20 || (data_size && fwrite(data, 1, data_size, h->f) != data_size) ||
            fwrite(cb, 1, 4, h->f) != 4)
        return fail(MI355_ERR_IO, "write failed");
    return ok();
}

MI355_EXPORT int mi355_vis_read_header(mi355_bin* h, mi355_vis_header* hd)
{
    if (!h || !hd) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    memset(hd, 0, sizeof(*hd));
    const int start = h->search_start;
    h->search_start = 0;
    struct { int tag; int* out; } ints[] = {
        {MI355_VIS_HEADER_NUM_TAGS_PER_BLOCK, &hd->num_tags_per_block},
        {MI355_VIS_HEADER_AMP_TYPE, &hd->amp_type},
        {MI355_VIS_HEADER_COORD_PRECISION, &hd->coord_precision},
        {MI355_VIS_HEADER_MAX_TIMES_PER_BLOCK, &hd->max_times_per_block},
        {MI355_VIS_HEADER_NUM_TIMES_TOTAL, &hd->num_times_total},
        {MI355_VIS_HEADER_NUM_CHANNELS_TOTAL, &hd->num_channels_total},
        {MI355_VIS_HEADER_NUM_STATIONS, &hd->num_stations}};
    int st = MI355_OK;
    for (auto& e : ints)
    {
        st = mi355_bin_read(h, MI355_BIN_INT, MI355_VIS_GROUP_HEADER, (unsigned char)e.tag, 0,
                sizeof(int), e.out);
        if (st != MI355_OK) break;
    }
    if (st == MI355_OK)
        st = mi355_bin_read(h, MI355_BIN_DOUBLE, MI355_VIS_GROUP_HEADER,
                MI355_VIS_HEADER_FREQ_START_HZ, 0, sizeof(double), &hd->freq_start_hz);
    h->search_start = start;
    if (st != MI355_OK) return st;
    if (hd->max_times_per_block < 1 || hd->num_times_total < 0 || hd->num_stations < 0)
        return fail(MI355_ERR_FORMAT, "visibility header holds invalid sizes");
    hd->num_baselines = hd->num_stations * (hd->num_stations - 1) / 2;
    hd->num_blocks = (hd->num_times_total + hd->max_times_per_block - 1) / hd->max_times_per_block;
    return ok();
}

MI355_EXPORT int mi355_vis_read_block(mi355_bin* h, const mi355_vis_header* hd, int block,
        size_t capacity, size_t* block_size, void* uu, void* vv, void* ww, void* vis)
{
    if (!h || !hd || !block_size || !uu || !vv || !ww)
        return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    if (block < 0 || block >= hd->num_blocks)
        return fail(MI355_ERR_INVALID_ARG, "block %d outside [0, %d)", block, hd->num_blocks);
    if (hd->coord_precision != MI355_BIN_SINGLE && hd->coord_precision != MI355_BIN_DOUBLE)
        return fail(MI355_ERR_FORMAT, "coordinate precision 0x%02x", hd->coord_precision);
    IO_TRY(mi355_bin_set_query_search_start(h, block * hd->num_tags_per_block));
    int dims[6];
    IO_TRY(mi355_bin_read(h, MI355_BIN_INT, MI355_VIS_GROUP_BLOCK,
            MI355_VIS_BLOCK_DIM_START_AND_SIZE, block, sizeof(dims), dims));
    const size_t n = (size_t)std::max(dims[2], 0) * (size_t)hd->num_baselines;
    if (n > capacity)
        return fail(MI355_ERR_INVALID_ARG, "block %d has %zu points, arrays hold %zu", block, n,
                capacity);
    const size_t ce = hd->coord_precision == MI355_BIN_DOUBLE ? 8 : 4;
    void* c[3] = {uu, vv, ww};
    const unsigned char tags[3] = {MI355_VIS_BLOCK_BASELINE_UU, MI355_VIS_BLOCK_BASELINE_VV,
                                   MI355_VIS_BLOCK_BASELINE_WW};
    for (int k = 0; k < 3; ++k)
        IO_TRY(mi355_bin_read(h, (unsigned char)hd->coord_precision, MI355_VIS_GROUP_BLOCK,
                tags[k], block, ce * n, c[k]));
    // metres -> wavelengths (read_vis.c:49-73)
    const double scale = hd->freq_start_hz / 299792458.0;
    if (hd->coord_precision == MI355_BIN_DOUBLE)
    {
        for (int k = 0; k < 3; ++k)
        {
            double* p = static_cast<double*>(c[k]);
            for (size_t i = 0; i < n; ++i) p[i] *= scale;
        }
    }
    else
    {
        const float s = (float)scale;
        for (int k = 0; k < 3; ++k)
        {
            float* p = static_cast<float*>(c[k]);
            for (size_t i = 0; i < n; ++i) p[i] *= s;
        }
    }
    if (vis)
    {
        const size_t ae = element_bytes(hd->amp_type);
        if (!ae) return fail(MI355_ERR_FORMAT, "amplitude type 0x%02x", hd->amp_type);
        IO_TRY(mi355_bin_read(h, (unsigned char)hd->amp_type, MI355_VIS_GROUP_BLOCK,
                MI355_VIS_BLOCK_CROSS_CORRELATIONS, block, ae * n, vis));
    }
    *block_size = n;
    return ok();
}

MI355_EXPORT void mi355_io_free(void* p) { free(p); }

MI355_EXPORT int mi355_fits_read_kernel(const char* path, int precision, mi355_kernel_info* info,
        void** data, int** support)
{
    if (!path || !info || !data) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    if (precision != MI355_BIN_SINGLE && precision != MI355_BIN_DOUBLE)
        return fail(MI355_ERR_INVALID_ARG, "precision must be MI355_BIN_SINGLE or _DOUBLE");
    *data = nullptr;
    if (support) *support = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(MI355_ERR_IO, "cannot open '%s': %s", path, strerror(errno));
    std::vector<Hdu> hdus;
    int st = read_hdus(f, path, &hdus);
    void* out = nullptr;
    int* sup = nullptr;
    do
    {
        if (st != MI355_OK) break;
        const Hdu& p = hdus[0];
        long long naxis = 0, bitpix = 0, d[3] = {0, 0, 0};
        hdu_int(p, "NAXIS", &naxis, path);
        if (naxis != 3) { st = fail(MI355_ERR_FORMAT, "%s: need a 3D kernel FITS cube", path); break; }
        hdu_int(p, "BITPIX", &bitpix, path);
        if ((st = hdu_int(p, "NAXIS1", &d[0], path)) || (st = hdu_int(p, "NAXIS2", &d[1], path)) ||
                (st = hdu_int(p, "NAXIS3", &d[2], path)))
            break;
        info->conv_size_half = (int)d[0];
        info->num_w_planes = (int)d[2];
        long long ov = 0, gs = 0;
        double cell = 0, ws = 0;
        auto real_key = [&](const char* k, double* v) {
            auto it = p.keys.find(k);
            if (it == p.keys.end() || !parse_double(it->second, v))
                return fail(MI355_ERR_FORMAT, "%s: missing '%s' FITS header keyword", path, k);
            return (int)MI355_OK;
        };
        if ((st = hdu_int(p, "OVERSAMP", &ov, path)) || (st = hdu_int(p, "GRIDSIZE", &gs, path)) ||
                (st = real_key("CELLSIZE", &cell)) || (st = real_key("W_SCALE", &ws)))
            break;
        info->oversample = (int)ov;
        info->grid_size = (int)gs;
        info->cellsize_rad = cell;
        info->w_scale = ws;
        double bscale = 1.0, bzero = 0.0;
        if (p.keys.count("BSCALE")) parse_double(p.keys.at("BSCALE"), &bscale);
        if (p.keys.count("BZERO")) parse_double(p.keys.at("BZERO"), &bzero);
        const size_t n = (size_t)(d[0] * d[1] * d[2]);
        const size_t w = (size_t)(std::llabs(bitpix) / 8);
        if (!(bitpix == 8 || bitpix == 16 || bitpix == 32 || bitpix == 64 || bitpix == -32 ||
                bitpix == -64) || (long long)(n * w) > p.data_bytes)
        { st = fail(MI355_ERR_FORMAT, "%s: bad primary data (BITPIX %lld)", path, bitpix); break; }
        std::vector<unsigned char> raw(n * w);
        if (fseeko(f, p.data_offset, SEEK_SET) != 0 || fread(raw.data(), 1, raw.size(), f) != raw.size())
        { st = fail(MI355_ERR_IO, "%s: read of the kernel data failed", path); break; }
        const bool scaled = bscale != 1.0 || bzero != 0.0;
        out = calloc(n ? n : 1, precision == MI355_BIN_DOUBLE ? sizeof(double) : sizeof(float));
        if (!out) { st = fail(MI355_ERR_OUT_OF_MEMORY, "host allocation failed"); break; }
        for (size_t i = 0; i < n; ++i)
        {
            const unsigned char* q = raw.data() + i * w;
            if (precision == MI355_BIN_SINGLE)
            {
                float v;
                if (bitpix == -32 && !scaled)
                {
                    unsigned char b[4] = {q[3], q[2], q[1], q[0]};
                    memcpy(&v, b, 4);           // bit for bit (NaNs included)
                }
                else
                {
                    const double x = fits_element(q, bitpix);   // (-0 stays -0 unscaled)
                    v = (float)(scaled ? x * bscale + bzero : x);
                }
                static_cast<float*>(out)[i] = v;
            }
            else
                static_cast<double*>(out)[i] = scaled ? fits_element(q, bitpix) * bscale + bzero
                                                      : fits_element(q, bitpix);
        }
        if (support)
        {
            const Hdu* t = nullptr;
            for (size_t k = 1; k < hdus.size() && !t; ++k)
            {
                std::string xt, name;
                auto ix = hdus[k].keys.find("XTENSION");
                if (ix == hdus[k].keys.end() || !parse_string(ix->second, &xt) ||
                        upper(xt) != "BINTABLE")
                    continue;
                for (const char* nk : {"EXTNAME", "HDUNAME"})
                {
                    auto in = hdus[k].keys.find(nk);
                    if (in != hdus[k].keys.end() && parse_string(in->second, &name) &&
                            upper(name) == "W_KERNELS")
                        t = &hdus[k];
                }
            }
            if (!t) { st = fail(MI355_ERR_FORMAT, "%s: missing 'W_KERNELS' FITS binary table", path); break; }
            long long row = 0, rows = 0;
            std::string tform;
            if ((st = hdu_int(*t, "NAXIS1", &row, path)) || (st = hdu_int(*t, "NAXIS2", &rows, path)))
                break;
            auto itf = t->keys.find("TFORM1");
            if (itf == t->keys.end() || !parse_string(itf->second, &tform))
            { st = fail(MI355_ERR_FORMAT, "%s: W_KERNELS has no TFORM1", path); break; }
            tform = upper(trim(tform));
            size_t pos = 0;
            long long rep = 1;
            while (pos < tform.size() && isdigit((unsigned char)tform[pos])) ++pos;
            if (pos) rep = atoll(tform.substr(0, pos).c_str());
            const char code = pos < tform.size() ? tform[pos] : '?';
            int bp = 0;
            switch (code)
            {
            case 'B': bp = 8; break;
            case 'I': bp = 16; break;
            case 'J': bp = 32; break;
            case 'K': bp = 64; break;
            case 'E': bp = -32; break;
            case 'D': bp = -64; break;
            default: break;
            }
            const int num = info->num_w_planes;
            if (!bp || rep < 1 || rep * rows < num)
            { st = fail(MI355_ERR_FORMAT, "%s: unable to read kernel support sizes (TFORM1 '%s')",
                    path, tform.c_str()); break; }
            // the column's elements must lie inside a row (element i is read
            // at row i / rep, byte (i % rep) * width)
            if (row < 1 || rows < 1 || rep > row / (std::abs(bp) / 8) || row > (1ll << 40) / rows)
            { st = fail(MI355_ERR_FORMAT, "%s: W_KERNELS column (TFORM1 '%s') does not fit a "
                    "row of %lld bytes", path, tform.c_str(), row); break; }
            double tscal = 1.0, tzero = 0.0;
            if (t->keys.count("TSCAL1")) parse_double(t->keys.at("TSCAL1"), &tscal);
            if (t->keys.count("TZERO1")) parse_double(t->keys.at("TZERO1"), &tzero);
            std::vector<unsigned char> tab((size_t)(row * rows));
            if (fseeko(f, t->data_offset, SEEK_SET) != 0 || fread(tab.data(), 1, tab.size(), f) != tab.size())
            { st = fail(MI355_ERR_IO, "%s: read of W_KERNELS failed", path); break; }
            sup = static_cast<int*>(calloc(num ? num : 1, sizeof(int)));
            if (!sup) { st = fail(MI355_ERR_OUT_OF_MEMORY, "host allocation failed"); break; }
            const int ew = std::abs(bp) / 8;
            for (int i = 0; i < num; ++i)
            {
                const long long r = i / rep, e = i % rep;
                const double v = fits_element(tab.data() + r * row + e * ew, bp) * tscal + tzero;
                if (!(v > -2147483649.0 && v < 2147483648.0))
                { st = fail(MI355_ERR_FORMAT, "%s: support value out of int range", path); break; }
                sup[i] = (int)v;
            }
        }
    } while (false);
    fclose(f);
    if (st != MI355_OK)
    {
        free(out);
        free(sup);
        return st;
    }
    *data = out;
    if (support) *support = sup;
    return ok();
}

MI355_EXPORT int mi355_fits_write_kernel(const char* path, int precision,
        const mi355_kernel_info* info, const void* data, const int* support)
{
    if (!path || !info || !data) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    if (precision != MI355_BIN_SINGLE && precision != MI355_BIN_DOUBLE)
        return fail(MI355_ERR_INVALID_ARG, "precision must be MI355_BIN_SINGLE or _DOUBLE");
    if (info->conv_size_half < 1 || info->num_w_planes < 1)
        return fail(MI355_ERR_INVALID_ARG, "empty kernel cube");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(MI355_ERR_IO, "cannot open '%s': %s", path, strerror(errno));
    const long long dims[3] = {info->conv_size_half, info->conv_size_half, info->num_w_planes};
    std::vector<std::string> keys = {
        fits_card("OVERSAMP", fits_num(info->oversample), "oversampling of the kernels"),
        fits_card("GRIDSIZE", fits_num(info->grid_size), "grid size in cells"),
        fits_card("CELLSIZE", fits_real(info->cellsize_rad), "cell size [rad]"),
        fits_card("W_SCALE", fits_real(info->w_scale), "w scale")};
    int st = write_image_hdu(f, path, precision, dims, 3, data, keys, support != nullptr);
    if (st == MI355_OK && support)
    {
        std::string hdr;
        hdr += fits_card("XTENSION", fits_str("BINTABLE"));
        hdr += fits_card("BITPIX", fits_num(8));
        hdr += fits_card("NAXIS", fits_num(2));
        hdr += fits_card("NAXIS1", fits_num(4));
        hdr += fits_card("NAXIS2", fits_num(info->num_w_planes));
        hdr += fits_card("PCOUNT", fits_num(0));
        hdr += fits_card("GCOUNT", fits_num(1));
        hdr += fits_card("TFIELDS", fits_num(1));
        hdr += fits_card("TTYPE1", fits_str("SUPPORT"));
        hdr += fits_card("TFORM1", fits_str("1J"));
        hdr += fits_card("EXTNAME", fits_str("W_KERNELS"));
        hdr += std::string("END") + std::string(77, ' ');
        st = write_record_padded(f, hdr, ' ', path);
        if (st == MI355_OK)
        {
            std::vector<int32_t> v(support, support + info->num_w_planes);
            st = write_record_padded(f, big_endian(v.data(), v.size()), '\0', path);
        }
    }
    if (fclose(f) != 0 && st == MI355_OK) st = fail(MI355_ERR_IO, "%s: close failed", path);
    return st == MI355_OK ? ok() : st;
}

MI355_EXPORT int mi355_fits_write_image(const char* path, int precision, int width, int height,
        const void* data)
{
    if (!path || !data || width < 1 || height < 1) return fail(MI355_ERR_INVALID_ARG, "bad argument");
    if (precision != MI355_BIN_SINGLE && precision != MI355_BIN_DOUBLE)
        return fail(MI355_ERR_INVALID_ARG, "precision must be MI355_BIN_SINGLE or _DOUBLE");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(MI355_ERR_IO, "cannot open '%s': %s", path, strerror(errno));
    const long long dims[2] = {width, height};
    int st = write_image_hdu(f, path, precision, dims, 2, data, {}, false);
    if (fclose(f) != 0 && st == MI355_OK) st = fail(MI355_ERR_IO, "%s: close failed", path);
    return st == MI355_OK ? ok() : st;
}

} // extern "C"
