// rp_ncin.cpp -- reader of netCDF classic files (CDF-1, CDF-2, CDF-5) for the
// ROMS grid and history input, declared in include/rp_ncin.h.  Host code.
//
// The header is parsed once at open (dimensions, global and per-variable
// attributes, variables with their type, shape and file offset); a read
// seeks to the variable's slice and converts the big-endian values to
// double.  Record variables are interleaved record by record; the record
// size is the sum of their per-record sizes, each rounded up to 4 bytes,
// except that a file with a single record variable stores its records
// unpadded.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rp_abi.h"
#include "../../include/rp_ncin.h"

namespace {

enum : int32_t {
    T_BYTE = 1, T_CHAR = 2, T_SHORT = 3, T_INT = 4, T_FLOAT = 5, T_DOUBLE = 6,
    T_UBYTE = 7, T_USHORT = 8, T_UINT = 9, T_INT64 = 10, T_UINT64 = 11
};

int type_size(int32_t t)
{
    switch (t) {
    case T_BYTE: case T_CHAR: case T_UBYTE: return 1;
    case T_SHORT: case T_USHORT: return 2;
    case T_INT: case T_FLOAT: case T_UINT: return 4;
    case T_DOUBLE: case T_INT64: case T_UINT64: return 8;
    default: return 0;
    }
}

struct Att {
    std::string name;
    int32_t type = 0;
    std::vector<unsigned char> raw;   // big-endian values, unpadded
    int64_t n = 0;
};

struct Dim {
    std::string name;
    int64_t len = 0;   // 0: the record dimension
};

struct Var {
    std::string name;
    std::vector<int64_t> dimids;
    std::vector<Att> atts;
    int32_t type = 0;
    uint64_t begin = 0;
    bool rec = false;
    int64_t per_rec = 1;   // values per record (record variable) or in all (fixed)
};

struct Reader {
    FILE* f = nullptr;
    int version = 1;
    int64_t numrecs = 0;
    uint64_t recsize = 0;
    std::vector<Dim> dims;
    std::vector<Att> gatts;
    std::vector<Var> vars;
    std::string err;
};

std::string g_open_err;

// cursor over the header bytes
struct Cur {
    const std::vector<unsigned char>& b;
    size_t p = 0;
    bool bad = false;
    explicit Cur(const std::vector<unsigned char>& buf) : b(buf) {}
    uint64_t be(int n)
    {
        if (p + (size_t)n > b.size()) {
            bad = true;
            return 0;
        }
        uint64_t v = 0;
        for (int i = 0; i < n; i++) v = (v << 8) | b[p + i];
        p += (size_t)n;
        return v;
    }
    void skip(uint64_t n)
    {
        if (p + n > b.size()) bad = true;
        else p += (size_t)n;
    }
};

bool read_atts(Cur& c, int nn, std::vector<Att>& out)
{
    const uint32_t tag = (uint32_t)c.be(4);
    const uint64_t n = c.be(nn);
    if (c.bad) return false;
    if (tag == 0 && n == 0) return true;
    if (tag != 0x0C || n > (1u << 20)) return false;
    out.resize(n);
    for (auto& a : out) {
        const uint64_t ln = c.be(nn);
        if (c.bad || ln > (1u << 16) || c.p + ln > c.b.size()) return false;
        a.name.assign((const char*)&c.b[c.p], ln);
        c.skip((ln + 3) & ~3ull);
        a.type = (int32_t)c.be(4);
        a.n = (int64_t)c.be(nn);
        const int ts = type_size(a.type);
        if (c.bad || ts == 0 || a.n < 0 || a.n > (1LL << 31)) return false;
        const uint64_t bytes = (uint64_t)a.n * ts;
        if (c.p + bytes > c.b.size()) return false;
        a.raw.assign(c.b.begin() + (long)c.p, c.b.begin() + (long)(c.p + bytes));
        c.skip((bytes + 3) & ~3ull);
        if (c.bad) return false;
    }
    return true;
}

int parse(Reader* r, const std::vector<unsigned char>& buf, uint64_t file_size)
{
    Cur c(buf);
    const int nn = r->version == 5 ? 8 : 4;   // NON_NEG width
    c.skip(4);
    const uint64_t nr = c.be(nn);
    // dimensions
    {
        const uint32_t tag = (uint32_t)c.be(4);
        const uint64_t n = c.be(nn);
        if (c.bad) return -1;
        if (!(tag == 0 && n == 0)) {
            if (tag != 0x0A || n > 65536) return -1;
            r->dims.resize(n);
            for (auto& d : r->dims) {
                const uint64_t ln = c.be(nn);
                if (c.bad || ln > 65536 || c.p + ln > buf.size()) return -1;
                d.name.assign((const char*)&buf[c.p], ln);
                c.skip((ln + 3) & ~3ull);
                d.len = (int64_t)c.be(nn);
            }
        }
    }
    if (!read_atts(c, nn, r->gatts)) return -1;
    // variables
    {
        const uint32_t tag = (uint32_t)c.be(4);
        const uint64_t n = c.be(nn);
        if (c.bad) return -1;
        if (!(tag == 0 && n == 0)) {
            if (tag != 0x0B || n > 65536) return -1;
            r->vars.resize(n);
            for (auto& v : r->vars) {
                const uint64_t ln = c.be(nn);
                if (c.bad || ln > 65536 || c.p + ln > buf.size()) return -1;
                v.name.assign((const char*)&buf[c.p], ln);
                c.skip((ln + 3) & ~3ull);
                const uint64_t nd = c.be(nn);
                if (c.bad || nd > 64) return -1;
                v.dimids.resize(nd);
                for (auto& id : v.dimids) {
                    id = (int64_t)c.be(nn);
                    if (id < 0 || id >= (int64_t)r->dims.size()) return -1;
                }
                if (!read_atts(c, nn, v.atts)) return -1;
                v.type = (int32_t)c.be(4);
                (void)c.be(nn);                              // vsize: recomputed below
                v.begin = c.be(r->version == 1 ? 4 : 8);
                if (c.bad || type_size(v.type) == 0) return -1;
                v.per_rec = 1;
                for (size_t k = 0; k < v.dimids.size(); k++) {
                    const Dim& d = r->dims[(size_t)v.dimids[k]];
                    if (d.len == 0) {
                        if (k != 0) return -1;               // only the first may be unlimited
                        v.rec = true;
                    } else {
                        v.per_rec *= d.len;
                    }
                }
            }
        }
    }
    // record size and count
    int nrec_vars = 0;
    uint64_t rs = 0, first_rec = 0;
    for (const Var& v : r->vars)
        if (v.rec) {
            const uint64_t bytes = (uint64_t)v.per_rec * type_size(v.type);
            rs += (bytes + 3) & ~3ull;
            if (nrec_vars == 0 || v.begin < first_rec) first_rec = v.begin;
            nrec_vars++;
        }
    if (nrec_vars == 1)
        for (const Var& v : r->vars)
            if (v.rec) rs = (uint64_t)v.per_rec * type_size(v.type);
    r->recsize = rs;
    const uint64_t streaming = r->version == 5 ? ~0ull : 0xFFFFFFFFull;
    if (nr == streaming)
        r->numrecs = (rs > 0 && file_size > first_rec) ? (int64_t)((file_size - first_rec) / rs) : 0;
    else
        r->numrecs = (int64_t)nr;
    return 0;
}

int fail(Reader* r, int code, const std::string& m)
{
    r->err = m;
    return code;
}

const Var* find_var(const Reader* r, const char* name)
{
    if (!name) return nullptr;
    for (const Var& v : r->vars)
        if (v.name == name) return &v;
    return nullptr;
}

double value_at(const unsigned char* p, int32_t t)
{
    uint64_t u = 0;
    const int n = type_size(t);
    for (int i = 0; i < n; i++) u = (u << 8) | p[i];
    switch (t) {
    case T_BYTE: return (double)(int8_t)u;
    case T_CHAR: case T_UBYTE: return (double)(uint8_t)u;
    case T_SHORT: return (double)(int16_t)u;
    case T_USHORT: return (double)(uint16_t)u;
    case T_INT: return (double)(int32_t)u;
    case T_UINT: return (double)(uint32_t)u;
    case T_INT64: return (double)(int64_t)u;
    case T_UINT64: return (double)u;
    case T_FLOAT: {
        const uint32_t w = (uint32_t)u;
        float x;
        std::memcpy(&x, &w, 4);
        return (double)x;
    }
    default: {
        double x;
        std::memcpy(&x, &u, 8);
        return x;
    }
    }
}

}  // namespace

extern "C" {

int rp_ncin_open(void** out, const char* path)
{
    if (!out || !path) return RP_E_ARG;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        g_open_err = std::string("cannot open ") + path + ": " + std::strerror(errno);
        return RP_E_IO;
    }
    std::fseek(f, 0, SEEK_END);
    const long long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    unsigned char magic[8] = {0};
    const size_t got = std::fread(magic, 1, 8, f);
    if (got >= 4 && magic[0] == 0x89 && magic[1] == 'H' && magic[2] == 'D' && magic[3] == 'F') {
        std::fclose(f);
        g_open_err = std::string(path) + ": netCDF-4 (HDF5) file; only the classic formats are read";
        return RP_E_IO;
    }
    if (got < 4 || magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F' ||
        (magic[3] != 1 && magic[3] != 2 && magic[3] != 5)) {
        std::fclose(f);
        g_open_err = std::string(path) + ": not a netCDF classic file";
        return RP_E_IO;
    }
    Reader* r = new Reader;
    r->f = f;
    r->version = magic[3];
    // the header ends before the first variable's data; read it in growing
    // chunks until it parses (headers beyond 256 MiB are taken as malformed)
    std::vector<unsigned char> buf;
    int rc = -1;
    for (long long want = 1 << 16; rc != 0 && want <= (1LL << 28); want *= 4) {
        const long long n = want < size ? want : size;
        buf.resize((size_t)n);
        std::fseek(f, 0, SEEK_SET);
        if (std::fread(buf.data(), 1, (size_t)n, f) != (size_t)n) break;
        Reader tmp;
        tmp.version = r->version;
        rc = parse(&tmp, buf, (uint64_t)size);
        if (rc == 0) {
            tmp.f = f;
            *r = std::move(tmp);
        }
        if (n == size) break;
    }
    if (rc != 0) {
        std::fclose(f);
        delete r;
        g_open_err = std::string(path) + ": malformed netCDF header";
        return RP_E_IO;
    }
    *out = r;
    return RP_OK;
}

int rp_ncin_dim(void* h, const char* name, int64_t* len)
{
    Reader* r = (Reader*)h;
    if (!r || !name || !len) return RP_E_ARG;
    for (const Dim& d : r->dims)
        if (d.name == name) {
            *len = d.len == 0 ? r->numrecs : d.len;
            return RP_OK;
        }
    return fail(r, RP_E_ARG, std::string("no dimension ") + name);
}

int rp_ncin_var(void* h, const char* name, int32_t* ndims, int64_t shape[8], int32_t* is_record)
{
    Reader* r = (Reader*)h;
    if (!r || !ndims || !shape || !is_record) return RP_E_ARG;
    const Var* v = find_var(r, name);
    if (!v) return fail(r, RP_E_ARG, std::string("no variable ") + (name ? name : "(null)"));
    if (v->dimids.size() > 8) return fail(r, RP_E_RANGE, "variable of rank above 8");
    *ndims = (int32_t)v->dimids.size();
    for (size_t k = 0; k < v->dimids.size(); k++) {
        const Dim& d = r->dims[(size_t)v->dimids[k]];
        shape[k] = d.len == 0 ? r->numrecs : d.len;
    }
    *is_record = v->rec ? 1 : 0;
    return RP_OK;
}

int rp_ncin_get(void* h, const char* name, int64_t rec, double* out, int64_t count)
{
    Reader* r = (Reader*)h;
    if (!r || (!out && count > 0)) return RP_E_ARG;
    const Var* v = find_var(r, name);
    if (!v) return fail(r, RP_E_ARG, std::string("no variable ") + (name ? name : "(null)"));
    if (count != v->per_rec)
        return fail(r, RP_E_ARG, std::string(name) + ": count " + std::to_string(count) + " differs from " +
                                     std::to_string(v->per_rec) + " values");
    uint64_t off = v->begin;
    if (v->rec) {
        if (rec < 0 || rec >= r->numrecs)
            return fail(r, RP_E_RANGE, std::string(name) + ": record " + std::to_string(rec) + " of " +
                                           std::to_string(r->numrecs));
        off += (uint64_t)rec * r->recsize;
    }
    const int ts = type_size(v->type);
    std::vector<unsigned char> b((size_t)count * ts);
    if (std::fseek(r->f, (long)off, SEEK_SET) != 0 || std::fread(b.data(), 1, b.size(), r->f) != b.size())
        return fail(r, RP_E_IO, std::string(name) + ": short read");
    for (int64_t i = 0; i < count; i++) out[i] = value_at(&b[(size_t)i * ts], v->type);
    return RP_OK;
}

int rp_ncin_att_text(void* h, const char* var, const char* att, char* buf, int64_t size, int64_t* len)
{
    Reader* r = (Reader*)h;
    if (!r || !att || !buf || size < 1) return RP_E_ARG;
    const std::vector<Att>* atts = &r->gatts;
    if (var) {
        const Var* v = find_var(r, var);
        if (!v) return fail(r, RP_E_ARG, std::string("no variable ") + var);
        atts = &v->atts;
    }
    for (const Att& a : *atts)
        if (a.name == att) {
            if (a.type != T_CHAR) return fail(r, RP_E_ARG, std::string(att) + " is not text");
            const int64_t n = (int64_t)a.raw.size() < size - 1 ? (int64_t)a.raw.size() : size - 1;
            std::memcpy(buf, a.raw.data(), (size_t)n);
            buf[n] = 0;
            if (len) *len = (int64_t)a.raw.size();
            return RP_OK;
        }
    return fail(r, RP_E_ARG, std::string("no attribute ") + att);
}

const char* rp_ncin_error(void* h)
{
    return h ? ((Reader*)h)->err.c_str() : g_open_err.c_str();
}

int rp_ncin_close(void* h)
{
    Reader* r = (Reader*)h;
    if (!r) return RP_E_ARG;
    if (r->f) std::fclose(r->f);
    delete r;
    return RP_OK;
}

}  // extern "C"
