// mr_netcdf3.cpp -- NetCDF-3 classic / 64-bit-offset reader (see mr_netcdf3.h).
#include "mr_netcdf3.h"

#include <cstdio>
#include <cstring>

namespace mr {

namespace {

struct Cursor {
    const std::vector<uint8_t> &b;
    size_t pos = 0;
    bool ok = true;
    explicit Cursor(const std::vector<uint8_t> &bytes) : b(bytes) {}
    bool need(size_t n)
    {
        if (!ok || pos + n > b.size()) { ok = false; return false; }
        return true;
    }
    uint32_t u32()
    {
        if (!need(4)) return 0;
        uint32_t v = (uint32_t)b[pos] << 24 | (uint32_t)b[pos + 1] << 16 |
                     (uint32_t)b[pos + 2] << 8 | (uint32_t)b[pos + 3];
        pos += 4;
        return v;
    }
    uint64_t u64()
    {
        uint64_t hi = u32();
        uint64_t lo = u32();
        return hi << 32 | lo;
    }
    std::string name()
    {
        uint32_t n = u32();
        if (!need(n)) return {};
        std::string s(reinterpret_cast<const char *>(&b[pos]), n);
        pos += (n + 3u) & ~3u;
        if (pos > b.size()) ok = false;
        return s;
    }
    void skip(size_t n)
    {
        if (need(n)) pos += n;
    }
};

size_t type_size(NcType t)
{
    switch (t) {
    case NcType::Byte: case NcType::Char: return 1;
    case NcType::Short: return 2;
    case NcType::Int: case NcType::Float: return 4;
    case NcType::Double: return 8;
    }
    return 0;
}

bool valid_type(uint32_t t) { return t >= 1 && t <= 6; }

void skip_atts(Cursor &c)
{
    uint32_t tag = c.u32();
    uint32_t n = c.u32();
    if (tag == 0 && n == 0) return;
    if (tag != 0x0Cu) { c.ok = false; return; }
    for (uint32_t i = 0; i < n && c.ok; i++) {
        c.name();
        uint32_t t = c.u32();
        uint32_t ne = c.u32();
        if (!valid_type(t)) { c.ok = false; return; }
        size_t sz = type_size((NcType)t) * ne;
        c.skip((sz + 3) & ~size_t(3));
    }
}

template <class T>
T be_load(const uint8_t *p)
{
    uint8_t tmp[sizeof(T)];
    for (size_t i = 0; i < sizeof(T); i++) tmp[i] = p[sizeof(T) - 1 - i];
    T v;
    std::memcpy(&v, tmp, sizeof(T));
    return v;
}

const NcVar *find_var(const NcFile &f, const std::string &name)
{
    for (const auto &v : f.vars)
        if (v.name == name) return &v;
    return nullptr;
}

// Offsets are formed in 128 bits, so no header value (a 64-bit begin near
// 2^64, huge dimensions or record counts) can wrap them past the checks.
typedef unsigned __int128 u128;

// Whether every element of variable v lies inside the file, checked before
// anything is allocated for it: the last element of a fixed-size variable,
// and of the last record's slab of a record variable.
bool var_fits(const NcFile &f, const NcVar &v, uint64_t &n_out)
{
    const u128 ts = type_size(v.type), size = f.bytes.size();
    const u128 per = v.n_per_record;
    if (v.is_record) {
        const u128 n = per * (u128)f.numrecs;
        if (n > (u128)UINT64_MAX) return false;
        n_out = (uint64_t)n;
        if (n == 0) return true;
        // a slab longer than the record would overlap the next record's data
        if (f.numrecs > 1 && per * ts > (u128)f.recsize) return false;
        const u128 last = (u128)v.begin + (u128)(f.numrecs - 1) * (u128)f.recsize + per * ts;
        return last <= size;
    }
    n_out = v.n_per_record;
    if (per == 0) return true;
    return (u128)v.begin + per * ts <= size;
}

// Element e of variable v (file order across records) as a raw pointer;
// var_fits has checked that every element lies inside the file.
const uint8_t *elem_ptr(const NcFile &f, const NcVar &v, uint64_t e)
{
    const u128 ts = type_size(v.type);
    u128 off;
    if (v.is_record) {
        const uint64_t rec = e / v.n_per_record, in = e % v.n_per_record;
        off = (u128)v.begin + (u128)rec * f.recsize + (u128)in * ts;
    } else {
        off = (u128)v.begin + (u128)e * ts;
    }
    if (off + ts > (u128)f.bytes.size()) return nullptr;
    return &f.bytes[(size_t)off];
}

template <class Out>
int read_as(const NcFile &f, const std::string &name, std::vector<Out> &out, std::string &msg)
{
    const NcVar *v = find_var(f, name);
    if (!v) {
        msg = "variable '" + name + "' not found";
        return -3;
    }
    uint64_t n = 0;
    if (!var_fits(f, *v, n)) {
        msg = "variable '" + name + "' extends past the end of the file";
        return -3;
    }
    out.resize((size_t)n);
    for (uint64_t e = 0; e < n; e++) {
        const uint8_t *p = elem_ptr(f, *v, e);
        if (!p) {
            msg = "variable '" + name + "' extends past the end of the file";
            return -3;
        }
        switch (v->type) {
        case NcType::Byte: out[e] = (Out)(int8_t)p[0]; break;      // DataType::I8
        case NcType::Char: out[e] = (Out)(uint8_t)p[0]; break;     // DataType::U8
        case NcType::Short: out[e] = (Out)be_load<int16_t>(p); break;
        case NcType::Int: out[e] = (Out)be_load<int32_t>(p); break;
        case NcType::Float: out[e] = (Out)be_load<float>(p); break;
        case NcType::Double: out[e] = (Out)be_load<double>(p); break;
        }
    }
    return 0;
}

}  // namespace

int nc3_open(const std::string &path, NcFile &f, std::string &msg)
{
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        msg = "cannot open '" + path + "'";
        return -2;
    }
    std::fseek(fp, 0, SEEK_END);
    long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz < 0) { std::fclose(fp); msg = "cannot size '" + path + "'"; return -2; }
    f.bytes.resize((size_t)sz);
    if (sz > 0 && std::fread(f.bytes.data(), 1, (size_t)sz, fp) != (size_t)sz) {
        std::fclose(fp);
        msg = "cannot read '" + path + "'";
        return -2;
    }
    std::fclose(fp);

    Cursor c(f.bytes);
    if (f.bytes.size() < 4 || f.bytes[0] != 'C' || f.bytes[1] != 'D' || f.bytes[2] != 'F') {
        msg = "'" + path + "' is not a NetCDF-3 file";
        return -3;
    }
    f.version = f.bytes[3];
    if (f.version != 1 && f.version != 2) {
        msg = "unsupported NetCDF format version " + std::to_string(f.version);
        return -3;
    }
    c.pos = 4;
    uint32_t numrecs = c.u32();
    bool streaming = numrecs == 0xFFFFFFFFu;
    f.numrecs = streaming ? 0 : numrecs;

    // dim_list
    uint32_t tag = c.u32(), n = c.u32();
    if (!(tag == 0 && n == 0)) {
        if (tag != 0x0Au) { msg = "bad dimension list"; return -3; }
        for (uint32_t i = 0; i < n && c.ok; i++) {
            std::string nm = c.name();
            uint64_t len = c.u32();
            f.dims.emplace_back(nm, len);
        }
    }
    skip_atts(c);   // global attributes
    // var_list
    tag = c.u32(); n = c.u32();
    if (!(tag == 0 && n == 0)) {
        if (tag != 0x0Bu) { msg = "bad variable list"; return -3; }
        for (uint32_t i = 0; i < n && c.ok; i++) {
            NcVar v;
            v.name = c.name();
            uint32_t nd = c.u32();
            for (uint32_t d = 0; d < nd && c.ok; d++) v.dimids.push_back((int32_t)c.u32());
            skip_atts(c);
            uint32_t t = c.u32();
            if (!valid_type(t)) { msg = "bad variable type"; return -3; }
            v.type = (NcType)t;
            v.vsize = c.u32();
            v.begin = f.version == 1 ? (uint64_t)c.u32() : c.u64();
            for (size_t d = 0; d < v.dimids.size(); d++) {
                int32_t id = v.dimids[d];
                if (id < 0 || (size_t)id >= f.dims.size()) { msg = "bad dimension id"; return -3; }
                uint64_t len = f.dims[(size_t)id].second;
                if (d == 0 && len == 0) { v.is_record = true; continue; }
                if (len != 0 && v.n_per_record > UINT64_MAX / len) {
                    msg = "variable '" + v.name + "' has more elements than 2^64";
                    return -3;
                }
                v.n_per_record *= len;
            }
            f.vars.push_back(v);
        }
    }
    if (!c.ok) { msg = "truncated NetCDF header"; return -3; }

    // record size: sum of the record variables' vsize; a lone record variable
    // is not padded between records (format specification, "vsize").
    size_t n_rec_vars = 0;
    for (const auto &v : f.vars) {
        if (v.is_record) { n_rec_vars++; f.recsize += v.vsize; }
    }
    if (n_rec_vars == 1) {
        for (const auto &v : f.vars) {
            if (!v.is_record) continue;
            if (v.n_per_record > UINT64_MAX / type_size(v.type)) {
                msg = "record variable '" + v.name + "' is too large";
                return -3;
            }
            f.recsize = v.n_per_record * type_size(v.type);
        }
    }
    if (streaming && f.recsize > 0) {
        uint64_t first = UINT64_MAX;
        for (const auto &v : f.vars)
            if (v.is_record && v.begin < first) first = v.begin;
        if (first != UINT64_MAX && f.bytes.size() > first)
            f.numrecs = (f.bytes.size() - first) / f.recsize;
    }
    return 0;
}

int nc3_read_f32(const NcFile &f, const std::string &name, std::vector<float> &out,
                 std::string &msg)
{
    return read_as(f, name, out, msg);
}

int nc3_read_f64(const NcFile &f, const std::string &name, std::vector<double> &out,
                 std::string &msg)
{
    return read_as(f, name, out, msg);
}

}  // namespace mr
