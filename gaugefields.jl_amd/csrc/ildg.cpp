// ildg.cpp -- ILDG/LIME gauge-configuration files, host side of libgfhip.
//
// LIME: a file is a sequence of records, each a 144-byte big-endian header (32-bit magic
// 0x456789ab, 16-bit version, 16-bit flags with message-begin 0x8000 / message-end 0x4000,
// 64-bit payload length, 128-byte NUL-padded type string) followed by the payload padded
// with zeros to a multiple of 8 bytes.
//
// ILDG payload ("ildg-binary-data"): for t, z, y, x (x fastest) and mu = 1..4, the 3x3 link
// matrix row by row (column fastest), each element a big-endian (re, im) pair of 32- or 64-bit
// IEEE floats.  This is the order the reference writes and reads (_save_binarydata,
// ildg_format.jl:697-746; load_binarydata!, ildg_format.jl:862-887).  The "ildg-format" XML
// record carries the extents and precision (ildg_format_xml, ildg_format.jl:34-48); a file
// without one is read at the caller's extents, its precision inferred from the payload size.
//
// Host arrays use the link layout of the rest of the library: double[4][V][9][2], element
// (i, j) at complex offset i + 3 j, site s = x + NX (y + NY (z + NZ t)).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gfhip.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return GF_EINVAL;
}

constexpr uint32_t kLimeMagic = 0x456789abu;
constexpr int kHeader = 144;
constexpr size_t kChunkSites = 4096;  // sites per read/write call

uint64_t load_be(const unsigned char* p, int nbytes) {
    uint64_t v = 0;
    for (int i = 0; i < nbytes; i++) v = (v << 8) | p[i];
    return v;
}
void store_be(unsigned char* p, uint64_t v, int nbytes) {
    for (int i = nbytes - 1; i >= 0; i--) {
        p[i] = (unsigned char)(v & 0xff);
        v >>= 8;
    }
}

double decode_real(const unsigned char* p, int w) {
    if (w == 8) {
        const uint64_t b = load_be(p, 8);
        double d;
        std::memcpy(&d, &b, 8);
        return d;
    }
    const uint32_t b = (uint32_t)load_be(p, 4);
    float f;
    std::memcpy(&f, &b, 4);
    return (double)f;
}
void encode_real(unsigned char* p, double v, int w) {
    if (w == 8) {
        uint64_t b;
        std::memcpy(&b, &v, 8);
        store_be(p, b, 8);
    } else {
        const float f = (float)v;
        uint32_t b;
        std::memcpy(&b, &f, 4);
        store_be(p, b, 4);
    }
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

struct Record {
    std::string type;
    uint64_t offset;  // file offset of the payload
    uint64_t length;
};

int scan(FILE* f, std::vector<Record>& recs) {
    unsigned char h[kHeader];
    uint64_t pos = 0;
    for (;;) {
        if (fseeko(f, (off_t)pos, SEEK_SET) != 0) return fail("seek failed");
        const size_t n = std::fread(h, 1, kHeader, f);
        if (n == 0) break;
        if (n != (size_t)kHeader) return fail("truncated LIME record header");
        if ((uint32_t)load_be(h, 4) != kLimeMagic) return fail("not a LIME file (bad record magic)");
        Record r;
        r.length = load_be(h + 8, 8);
        r.offset = pos + kHeader;
        char t[129];
        std::memcpy(t, h + 16, 128);
        t[128] = 0;
        r.type = t;
        recs.push_back(r);
        pos = r.offset + ((r.length + 7) / 8) * 8;
    }
    if (recs.empty()) return fail("empty LIME file");
    return GF_OK;
}

bool xml_int(const std::string& x, const char* tag, long long* v) {
    const std::string open = std::string("<") + tag + ">";
    const size_t a = x.find(open);
    if (a == std::string::npos) return false;
    char* end = nullptr;
    const long long r = std::strtoll(x.c_str() + a + open.size(), &end, 10);
    if (end == x.c_str() + a + open.size()) return false;
    *v = r;
    return true;
}

// Header facts of one file: the binary record and what the ildg-format record says.
struct Info {
    Record bin{};
    bool have_bin = false;
    long long dims[4] = {0, 0, 0, 0};  // 0: no ildg-format extents
    long long precision = 0;           // 0: not stated
};

int read_info(FILE* f, Info& info) {
    std::vector<Record> recs;
    int rc = scan(f, recs);
    if (rc != GF_OK) return rc;
    for (const Record& r : recs) {
        if (r.type == "ildg-format" && r.length < (1u << 20)) {
            std::string xml(r.length, '\0');
            if (fseeko(f, (off_t)r.offset, SEEK_SET) != 0 || std::fread(&xml[0], 1, r.length, f) != r.length)
                return fail("truncated ildg-format record");
            long long l[4];
            const char* tags[4] = {"lx", "ly", "lz", "lt"};
            bool all = true;
            for (int m = 0; m < 4; m++) all = all && xml_int(xml, tags[m], &l[m]);
            if (all)
                for (int m = 0; m < 4; m++) info.dims[m] = l[m];
            long long p;
            if (xml_int(xml, "precision", &p)) info.precision = p;
        }
        if (r.type == "ildg-binary-data" && !info.have_bin) {
            info.bin = r;
            info.have_bin = true;
        }
    }
    if (!info.have_bin) return fail("no ildg-binary-data record");
    return GF_OK;
}

int volume(const int* dims, long long* V) {
    *V = 1;
    for (int m = 0; m < 4; m++) {
        if (dims[m] <= 0) return fail("lattice extents must be positive");
        *V *= dims[m];
    }
    return GF_OK;
}

}  // namespace

extern "C" {

const char* gf_ildg_last_error(void) { return g_err.c_str(); }

int gf_ildg_info(const char* path, int* dims, int* precision_bits, long long* payload_bytes) {
    if (!path) return fail("null argument");
    File file;
    file.f = std::fopen(path, "rb");
    if (!file.f) return fail(std::string("cannot open ") + path);
    Info info;
    const int rc = read_info(file.f, info);
    if (rc != GF_OK) return rc;
    if (dims)
        for (int m = 0; m < 4; m++) dims[m] = (int)info.dims[m];
    if (precision_bits) *precision_bits = (int)info.precision;
    if (payload_bytes) *payload_bytes = (long long)info.bin.length;
    return GF_OK;
}

int gf_ildg_read(const char* path, const int* dims, double* ref, int* precision_bits) {
    if (!path || !dims || !ref) return fail("null argument");
    long long V;
    int rc = volume(dims, &V);
    if (rc != GF_OK) return rc;
    File file;
    file.f = std::fopen(path, "rb");
    if (!file.f) return fail(std::string("cannot open ") + path);
    Info info;
    rc = read_info(file.f, info);
    if (rc != GF_OK) return rc;
    if (info.dims[0] != 0)
        for (int m = 0; m < 4; m++)
            if (info.dims[m] != dims[m]) return fail("ildg-format extents differ from the requested extents");
    const uint64_t n_real = (uint64_t)V * 4 * 9 * 2;
    long long prec = info.precision;
    if (prec == 0) {
        prec = info.bin.length == n_real * 8 ? 64 : (info.bin.length == n_real * 4 ? 32 : 0);
        if (prec == 0)
            return fail("ildg-binary-data length does not match an SU(3) configuration of these extents");
    }
    if (prec != 32 && prec != 64) return fail("ILDG precision must be 32 or 64");
    if (info.bin.length != n_real * (uint64_t)(prec / 8))
        return fail("ildg-binary-data length does not match an SU(3) configuration of these extents");
    const int w = (int)prec / 8;
    const size_t site_bytes = (size_t)4 * 9 * 2 * w;
    std::vector<unsigned char> buf(site_bytes * kChunkSites);
    if (fseeko(file.f, (off_t)info.bin.offset, SEEK_SET) != 0) return fail("seek failed");
    for (long long s0 = 0; s0 < V; s0 += (long long)kChunkSites) {
        const size_t ns = (size_t)((V - s0) < (long long)kChunkSites ? (V - s0) : (long long)kChunkSites);
        if (std::fread(buf.data(), 1, ns * site_bytes, file.f) != ns * site_bytes)
            return fail("truncated ildg-binary-data");
        const unsigned char* p = buf.data();
        for (size_t k = 0; k < ns; k++) {
            const size_t s = (size_t)s0 + k;
            for (int mu = 0; mu < 4; mu++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double* o = ref + 2 * (((size_t)mu * (size_t)V + s) * 9 + i + 3 * j);
                        o[0] = decode_real(p, w);
                        o[1] = decode_real(p + w, w);
                        p += 2 * w;
                    }
        }
    }
    if (precision_bits) *precision_bits = (int)prec;
    return GF_OK;
}

int gf_ildg_write(const char* path, const int* dims, const double* ref, int precision_bits) {
    if (!path || !dims || !ref) return fail("null argument");
    if (precision_bits != 32 && precision_bits != 64) return fail("ILDG precision must be 32 or 64");
    long long V;
    int rc = volume(dims, &V);
    if (rc != GF_OK) return rc;
    File file;
    file.f = std::fopen(path, "wb");
    if (!file.f) return fail(std::string("cannot create ") + path);
    FILE* f = file.f;
    char xml[512];
    std::snprintf(xml, sizeof xml,
                  "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat xmlns=\"http://www.lqcd.org/ildg\">\n"
                  "  <version>1.2</version>\n  <field>su3gauge</field>\n  <precision>%d</precision>\n"
                  "  <lx>%d</lx>\n  <ly>%d</ly>\n  <lz>%d</lz>\n  <lt>%d</lt>\n</ildgFormat>\n",
                  precision_bits, dims[0], dims[1], dims[2], dims[3]);
    static const unsigned char zeros[8] = {0};
    auto record = [&](const char* type, uint64_t len, bool mb, bool me) {
        unsigned char h[kHeader];
        std::memset(h, 0, sizeof h);
        store_be(h, kLimeMagic, 4);
        store_be(h + 4, 1, 2);  // LIME version 1
        store_be(h + 6, (mb ? 0x8000u : 0u) | (me ? 0x4000u : 0u), 2);
        store_be(h + 8, len, 8);
        std::strncpy((char*)h + 16, type, 127);
        return std::fwrite(h, 1, kHeader, f) == (size_t)kHeader;
    };
    auto pad = [&](uint64_t len) { return len % 8 == 0 || std::fwrite(zeros, 1, 8 - len % 8, f) == 8 - len % 8; };
    const uint64_t xl = std::strlen(xml);
    bool ok = record("ildg-format", xl, true, false) && std::fwrite(xml, 1, xl, f) == xl && pad(xl);
    const int w = precision_bits / 8;
    const size_t site_bytes = (size_t)4 * 9 * 2 * w;
    const uint64_t blen = (uint64_t)V * site_bytes;
    ok = ok && record("ildg-binary-data", blen, false, true);
    std::vector<unsigned char> buf(site_bytes * kChunkSites);
    for (long long s0 = 0; s0 < V && ok; s0 += (long long)kChunkSites) {
        const size_t ns = (size_t)((V - s0) < (long long)kChunkSites ? (V - s0) : (long long)kChunkSites);
        unsigned char* p = buf.data();
        for (size_t k = 0; k < ns; k++) {
            const size_t s = (size_t)s0 + k;
            for (int mu = 0; mu < 4; mu++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        const double* o = ref + 2 * (((size_t)mu * (size_t)V + s) * 9 + i + 3 * j);
                        encode_real(p, o[0], w);
                        encode_real(p + w, o[1], w);
                        p += 2 * w;
                    }
        }
        ok = std::fwrite(buf.data(), 1, ns * site_bytes, f) == ns * site_bytes;
    }
    ok = ok && pad(blen);
    file.f = nullptr;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? GF_OK : fail(std::string("write failed: ") + path);
}

}  // extern "C"
