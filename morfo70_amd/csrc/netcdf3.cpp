// netcdf3.cpp -- NetCDF classic-format reader and writer (see netcdf3.h).
#include "netcdf3.h"

#include <cstdio>
#include <cstring>
#include <memory>

namespace m70nc {
namespace {

constexpr uint32_t TAG_DIMENSION = 0x0A, TAG_VARIABLE = 0x0B, TAG_ATTRIBUTE = 0x0C;

int type_size(int t)
{
    switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: return 0;
    }
}

uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

// ------------------------------------------------------------------ reading
struct Reader {
    FILE *fp = nullptr;
    bool ok = true;
    ~Reader() { if (fp) std::fclose(fp); }
    bool bytes(void *p, size_t n)
    {
        if (!ok || std::fread(p, 1, n, fp) != n) ok = false;
        return ok;
    }
    uint32_t u32()
    {
        unsigned char b[4] = {};
        bytes(b, 4);
        return (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
    }
    uint64_t u64()
    {
        const uint64_t hi = u32();
        return hi << 32 | u32();
    }
    std::string name()
    {
        const uint32_t n = u32();
        if (!ok || n > (1u << 20)) {
            ok = false;
            return {};
        }
        std::string s(n, '\0');
        bytes(&s[0], n);
        char pad[4];
        bytes(pad, pad4(n) - n);
        return s;
    }
    bool atts(std::vector<Attr> &out)
    {
        const uint32_t tag = u32(), n = u32();
        if (tag == 0 && n == 0) return ok;
        if (tag != TAG_ATTRIBUTE) return ok = false;
        for (uint32_t q = 0; q < n && ok; ++q) {
            Attr a;
            a.name = name();
            a.type = (int)u32();
            const uint32_t cnt = u32();
            const int sz = type_size(a.type);
            if (!sz || cnt > (1u << 24)) return ok = false;
            std::vector<unsigned char> raw(pad4((uint64_t)cnt * sz));
            bytes(raw.data(), raw.size());
            if (a.type == NC_CHAR) {
                a.text.assign((const char *)raw.data(), cnt);
            } else {
                for (uint32_t k = 0; k < cnt; ++k) a.values.push_back(decode(raw.data() + (size_t)k * sz, a.type));
            }
            out.push_back(a);
        }
        return ok;
    }
    static double decode(const unsigned char *b, int type)
    {
        switch (type) {
        case NC_BYTE: return (double)(int8_t)b[0];
        case NC_SHORT: return (double)(int16_t)((uint16_t)b[0] << 8 | b[1]);
        case NC_INT: return (double)(int32_t)((uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3]);
        case NC_FLOAT: {
            const uint32_t u = (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
            float f;
            std::memcpy(&f, &u, 4);
            return (double)f;
        }
        case NC_DOUBLE: {
            uint64_t u = 0;
            for (int k = 0; k < 8; ++k) u = u << 8 | b[k];
            double d;
            std::memcpy(&d, &u, 8);
            return d;
        }
        default: return 0.0;
        }
    }
};

// ------------------------------------------------------------------ writing
struct Writer {
    std::vector<unsigned char> buf;
    void u32(uint32_t v)
    {
        for (int s = 24; s >= 0; s -= 8) buf.push_back((unsigned char)(v >> s));
    }
    void u64(uint64_t v)
    {
        u32((uint32_t)(v >> 32));
        u32((uint32_t)v);
    }
    void name(const std::string &s)
    {
        u32((uint32_t)s.size());
        buf.insert(buf.end(), s.begin(), s.end());
        buf.resize(pad4(buf.size()), 0);
    }
    void f64(double d)
    {
        uint64_t u;
        std::memcpy(&u, &d, 8);
        u64(u);
    }
};

}  // namespace

bool read_header(const std::string &path, File &f, std::string &err)
{
    Reader r;
    r.fp = std::fopen(path.c_str(), "rb");
    if (!r.fp) {
        err = "cannot open " + path;
        return false;
    }
    f = File{};
    f.path = path;
    char magic[4];
    if (!r.bytes(magic, 4) || std::memcmp(magic, "CDF", 3) != 0 || (magic[3] != 1 && magic[3] != 2)) {
        err = path + ": not a NetCDF classic (CDF-1/CDF-2) file";
        return false;
    }
    f.version = magic[3];
    const uint32_t numrecs = r.u32();
    uint32_t tag = r.u32(), n = r.u32();
    if (tag == TAG_DIMENSION) {
        for (uint32_t q = 0; q < n && r.ok; ++q) {
            std::string nm = r.name();
            const uint32_t len = r.u32();
            f.dims.emplace_back(nm, len);
        }
    } else if (tag != 0 || n != 0) {
        r.ok = false;
    }
    r.atts(f.gatts);
    tag = r.u32();
    n = r.u32();
    if (tag == TAG_VARIABLE) {
        for (uint32_t q = 0; q < n && r.ok; ++q) {
            Var v;
            v.name = r.name();
            const uint32_t nd = r.u32();
            if (nd > 64) {
                r.ok = false;
                break;
            }
            for (uint32_t k = 0; k < nd; ++k) v.dims.push_back((int)r.u32());
            r.atts(v.atts);
            v.type = (int)r.u32();
            v.vsize = r.u32();
            v.begin = (f.version == 2) ? r.u64() : r.u32();
            f.vars.push_back(v);
        }
    } else if (tag != 0 || n != 0) {
        r.ok = false;
    }
    if (!r.ok) {
        err = path + ": malformed header";
        return false;
    }
    for (const Var &v : f.vars)
        for (int d : v.dims)
            if (d < 0 || d >= (int)f.dims.size() || (f.dims[d].second == 0 && numrecs != 0)) {
                err = path + ": record variables are not supported (" + v.name + ")";
                return false;
            }
    return true;
}

bool read_var(const File &f, const std::string &name, std::vector<double> &out, std::string &err)
{
    const Var *v = nullptr;
    for (const Var &q : f.vars)
        if (q.name == name) v = &q;
    if (!v) {
        err = f.path + ": no variable " + name;
        return false;
    }
    const int sz = type_size(v->type);
    if (!sz || v->type == NC_CHAR) {
        err = f.path + ": variable " + name + " is not numeric";
        return false;
    }
    uint64_t n = 1;
    for (int d : v->dims) n *= f.dims[d].second;
    Reader r;
    r.fp = std::fopen(f.path.c_str(), "rb");
    if (!r.fp) {
        err = "cannot open " + f.path;
        return false;
    }
#if defined(_WIN32)
    const int sk = _fseeki64(r.fp, (long long)v->begin, SEEK_SET);
#else
    const int sk = fseeko(r.fp, (off_t)v->begin, SEEK_SET);
#endif
    std::vector<unsigned char> raw((size_t)(n * sz));
    if (sk != 0 || !r.bytes(raw.data(), raw.size())) {
        err = f.path + ": short data for " + name;
        return false;
    }
    out.resize((size_t)n);
    for (uint64_t k = 0; k < n; ++k) out[(size_t)k] = Reader::decode(raw.data() + (size_t)k * sz, v->type);
    return true;
}

bool read_att(const File &f, const std::string &name, double &out, std::string &err)
{
    for (const Attr &a : f.gatts)
        if (a.name == name && !a.values.empty()) {
            out = a.values[0];
            return true;
        }
    err = f.path + ": no numeric global attribute " + name;
    return false;
}

bool write_file(const std::string &path, const std::vector<std::pair<std::string, uint64_t>> &dims,
                const std::vector<std::pair<std::string, double>> &gatts, const std::vector<OutVar> &vars,
                bool force_64bit, std::string &err)
{
    std::vector<uint64_t> count(vars.size(), 1);
    for (size_t q = 0; q < vars.size(); ++q)
        for (int d : vars[q].dims) {
            if (d < 0 || d >= (int)dims.size()) {
                err = "variable " + vars[q].name + ": bad dimension id";
                return false;
            }
            count[q] *= dims[d].second;
        }
    // the header, twice: first to measure it, then with the offsets
    auto header = [&](int version, uint64_t data0) {
        Writer w;
        const char magic[4] = {'C', 'D', 'F', (char)version};
        w.buf.insert(w.buf.end(), magic, magic + 4);
        w.u32(0);   // numrecs
        w.u32(TAG_DIMENSION);
        w.u32((uint32_t)dims.size());
        for (auto &d : dims) {
            w.name(d.first);
            w.u32((uint32_t)d.second);
        }
        if (gatts.empty()) {
            w.u32(0);
            w.u32(0);
        } else {
            w.u32(TAG_ATTRIBUTE);
            w.u32((uint32_t)gatts.size());
            for (auto &a : gatts) {
                w.name(a.first);
                w.u32(NC_DOUBLE);
                w.u32(1);
                w.f64(a.second);
            }
        }
        w.u32(TAG_VARIABLE);
        w.u32((uint32_t)vars.size());
        uint64_t off = data0;
        for (size_t q = 0; q < vars.size(); ++q) {
            w.name(vars[q].name);
            w.u32((uint32_t)vars[q].dims.size());
            for (int d : vars[q].dims) w.u32((uint32_t)d);
            w.u32(0);   // no variable attributes
            w.u32(0);
            w.u32(NC_DOUBLE);
            const uint64_t vsize = count[q] * 8;
            w.u32(vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)vsize);
            if (version == 2) w.u64(off);
            else w.u32((uint32_t)off);
            off += vsize;
        }
        return std::make_pair(w.buf, off);
    };
    uint64_t total_data = 0;
    for (uint64_t c : count) total_data += c * 8;
    int version = 1;
    uint64_t hsize = header(2, 0).first.size();   // an upper bound for either version
    if (force_64bit || hsize + total_data > 0x7FFFFFFFull) version = 2;
    hsize = header(version, 0).first.size();
    auto hdr = header(version, hsize);
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        err = "cannot create " + path;
        return false;
    }
    bool ok = std::fwrite(hdr.first.data(), 1, hdr.first.size(), fp) == hdr.first.size();
    std::vector<unsigned char> chunk;
    for (size_t q = 0; q < vars.size() && ok; ++q) {
        const double *p = vars[q].data;
        const uint64_t n = count[q];
        const uint64_t step = 1 << 16;
        for (uint64_t k0 = 0; k0 < n && ok; k0 += step) {
            const uint64_t k1 = (k0 + step < n) ? k0 + step : n;
            chunk.resize((size_t)(k1 - k0) * 8);
            for (uint64_t k = k0; k < k1; ++k) {
                uint64_t u;
                std::memcpy(&u, p + k, 8);
                for (int b = 0; b < 8; ++b) chunk[(size_t)(k - k0) * 8 + b] = (unsigned char)(u >> (56 - 8 * b));
            }
            ok = std::fwrite(chunk.data(), 1, chunk.size(), fp) == chunk.size();
        }
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) err = "write failed: " + path;
    return ok;
}

}  // namespace m70nc
