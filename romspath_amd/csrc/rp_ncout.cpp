// rp_ncout.cpp -- the trajectory file of createNetCDF / writeNetCDF
// (ROMSPath.f90:1140-2157) as netCDF classic with 64-bit offsets (CDF-2).
//
// Format (netCDF classic file format specification, version 2): a header of
// magic, record count, dimension list, global attribute list and variable
// list, all big-endian and padded to 4 bytes; then the fixed-size variables
// at their `begin` offsets; then the records, each holding one slab of every
// record variable in variable-list order.  The header is written once; each
// record is appended and the record count at byte 4 is updated after it.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/rp_abi.h"
#include "../../include/rp_ncout.h"

namespace {

enum { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum : uint32_t { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };
enum { DIM_NUMPAR = 0, DIM_TIME = 1 };

struct Att {
    std::string name;
    int type;
    std::string text;            // NC_CHAR
    std::vector<double> num;     // NC_INT, NC_FLOAT, NC_DOUBLE
};

struct Var {
    std::string name;
    int type;
    bool rec;                    // dims (time, numpar) or (time); else (numpar)
    bool per_particle;
    std::vector<Att> atts;
    uint64_t vsize = 0, begin = 0;
};

struct Be {
    std::vector<unsigned char> b;
    void u32(uint32_t v)
    {
        for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void u64(uint64_t v)
    {
        u32((uint32_t)(v >> 32));
        u32((uint32_t)v);
    }
    void pad()
    {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string& s)
    {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
    void att(const Att& a)
    {
        name(a.name);
        u32((uint32_t)a.type);
        if (a.type == NC_CHAR) {
            u32((uint32_t)a.text.size());
            b.insert(b.end(), a.text.begin(), a.text.end());
        } else {
            u32((uint32_t)a.num.size());
            for (double v : a.num) {
                if (a.type == NC_INT) {
                    u32((uint32_t)(int32_t)v);
                } else if (a.type == NC_FLOAT) {
                    const float f = (float)v;
                    uint32_t u;
                    std::memcpy(&u, &f, 4);
                    u32(u);
                } else {
                    uint64_t u;
                    std::memcpy(&u, &v, 8);
                    u64(u);
                }
            }
        }
        pad();
    }
    void atts(const std::vector<Att>& list)
    {
        if (list.empty()) {
            u32(0);
            u32(0);
            return;
        }
        u32(TAG_ATT);
        u32((uint32_t)list.size());
        for (const Att& a : list) att(a);
    }
};

Att text(const char* n, const std::string& v) { return Att{n, NC_CHAR, v, {}}; }
Att num(const char* n, int type, double v) { return Att{n, type, std::string(), {v}}; }

std::vector<Att> desc(const char* long_name, const char* units, const char* field)
{
    std::vector<Att> a;
    a.push_back(text("long_name", long_name));
    if (units) a.push_back(text("units", units));
    if (field) a.push_back(text("field", field));
    return a;
}

struct Writer {
    FILE* f = nullptr;
    int64_t numpar = 0;
    int flags = 0;
    int ngrid = 1;
    std::vector<Var> vars;
    uint64_t rec_begin = 0, recsize = 0;
    int64_t nrec = 0;
    std::string err;
    std::vector<unsigned char> buf;
};

thread_local std::string g_create_err;

int wfail(Writer* w, int code, const std::string& msg)
{
    w->err = msg;
    return code;
}

// values rounded to the variable's type, big-endian, into w->buf
void encode(Writer* w, int type, const double* v, const int32_t* iv, int64_t n)
{
    const size_t sz = type == NC_DOUBLE ? 8 : 4;
    w->buf.resize((size_t)n * sz);
    unsigned char* o = w->buf.data();
    for (int64_t i = 0; i < n; i++) {
        const double x = iv ? (double)iv[i] : v[i];
        if (type == NC_DOUBLE) {
            uint64_t u;
            std::memcpy(&u, &x, 8);
            for (int s = 56; s >= 0; s -= 8) *o++ = (unsigned char)(u >> s);
        } else {
            const float f = (float)x;
            uint32_t u;
            std::memcpy(&u, &f, 4);
            for (int s = 24; s >= 0; s -= 8) *o++ = (unsigned char)(u >> s);
        }
    }
}

int put_at(Writer* w, uint64_t off, const void* p, size_t n)
{
    if (fseeko(w->f, (off_t)off, SEEK_SET) != 0 || fwrite(p, 1, n, w->f) != n)
        return wfail(w, RP_E_IO, "write failed at offset " + std::to_string(off));
    return RP_OK;
}

}  // namespace

extern "C" {

int rp_nc_create(void** nc, const char* path, int64_t numpar, int32_t ngrid, int32_t flags,
                 const rp_nc_meta* meta, const double* dob)
{
    if (!nc || !path || !meta || !dob || numpar < 1 || ngrid < 1 || ngrid > RP_MAXGRID) {
        g_create_err = "rp_nc_create: invalid argument";
        return RP_E_ARG;
    }
    *nc = nullptr;
    // classic-format slabs are sized by 32-bit vsize fields
    if ((uint64_t)numpar * 8 > 0xFFFFFFFCull) {
        g_create_err = "rp_nc_create: too many particles for one netCDF record slab";
        return RP_E_RANGE;
    }
    Writer* w = new Writer;
    w->numpar = numpar;
    w->flags = flags;
    w->ngrid = ngrid;
    auto add = [&](const char* name, int type, bool rec, bool pp, std::vector<Att> atts) {
        Var v;
        v.name = name;
        v.type = type;
        v.rec = rec;
        v.per_particle = pp;
        v.atts = std::move(atts);
        w->vars.push_back(std::move(v));
    };
    const std::string tu = meta->time_units ? meta->time_units : "seconds";
    {
        std::vector<Att> a = desc("Model time", nullptr, nullptr);
        a.push_back(text("units", tu));
        a.push_back(text("field", "model_time, scalar, series"));
        add("model_time", NC_DOUBLE, true, false, a);
    }
    add("dob", NC_FLOAT, false, true,
        desc("Date of Birth of particles in seconds from model start", "seconds", "age, scalar, series"));
    add("age", NC_DOUBLE, true, true, desc("age of particles", "seconds", "age, scalar, series"));
    add("lon", NC_FLOAT, true, true, desc("longitude of particles", "decimal degrees E", "lon, scalar, series"));
    add("lat", NC_FLOAT, true, true, desc("latitude of particles", "decimal degrees N", "lat, scalar, series"));
    add("zp", NC_FLOAT, true, true,
        desc("vertical position of the particle in ROMS z-coordinates, zp=0 is mean sea level, "
             "positive is up.", "meters", "zp, scalar, series"));
    add("status", NC_FLOAT, true, true,
        desc("identification number for particle behavior or status",
             "nondimensional, see ROMSPath User Guide", "status, scalar, series"));
    if (ngrid > 1) add("GID", NC_FLOAT, true, true, desc("Grid ID", "none", nullptr));
    if (flags & RP_NC_COLLISIONS) {
        add("hitBottom", NC_FLOAT, true, true,
            desc("# of times Particle Collided with Bottom", "Number of Collisions",
                 "hitBottom, scalar, series"));
        add("hitLand", NC_FLOAT, true, true,
            desc("# of times Particle Collided with Land", "Number of Collisions",
                 "hitLand, scalar, series"));
    }
    if (flags & RP_NC_SALTTEMP) {
        add("salinity", NC_FLOAT, true, true,
            desc("Salinity at the particle's location", nullptr, "salinity, scalar, series"));
        add("temperature", NC_FLOAT, true, true,
            desc("Temperature at the particle's location", "degree Celsius",
                 "temperature, scalar, series"));
    }
    if (flags & RP_NC_BOTTOM) {
        add("HOB", NC_FLOAT, true, true,
            desc("Particle Minimum height above bottom since last output time step", "m",
                 "HOB, scalar, series"));
        add("bustr", NC_FLOAT, true, true, desc("U bottom stress at minimum HOB", "newton meter-2", nullptr));
        add("bvstr", NC_FLOAT, true, true, desc("V bottom stress at minimum HOB", "newton meter-2", nullptr));
    }
    if (flags & RP_NC_WATERDEPTH)
        add("WaterDepth", NC_FLOAT, true, true,
            desc("Local Water Depth at Particle Location", "m", "WaterDepth, scalar, series"));
    if (flags & RP_NC_ZETA)
        add("zeta", NC_FLOAT, true, true,
            desc("free surface displacement at particle locations", "m", "zeta, scalar, series"));
    if (flags & RP_NC_BATH)
        add("h", NC_FLOAT, true, true, desc("bathymetry at particle locations", "m", "H, scalar, series"));

    std::vector<Att> g;
    auto str = [](const char* s) { return std::string(s ? s : ""); };
    g.push_back(text("title", "Particle tracking output using ROMS data."));
    g.push_back(text("author", "Elias Hunter (hunter@marine.rutgers.edu)"));
    g.push_back(num("adv_time_step", NC_INT, meta->idt));
    g.push_back(num("days_run", NC_FLOAT, meta->days));
    g.push_back(num("Behavior", NC_INT, meta->behavior));
    if (meta->open_boundary) g.push_back(text("OpenOceanBoundary", "TRUE"));
    g.push_back(num("swimfast", NC_DOUBLE, meta->swimfast));
    if (meta->vturb_on) {
        g.push_back(text("Vertical_Turb", "True"));
        g.push_back(num("seed", NC_INT, meta->seed));
        g.push_back(num("turb_time_step", NC_DOUBLE, meta->deltat));
        g.push_back(num("spline_error_cutoff", NC_DOUBLE, meta->serr));
        g.push_back(num("spline_smooth_param", NC_DOUBLE, meta->smth));
        g.push_back(num("subset_scale", NC_DOUBLE, meta->sub));
        g.push_back(num("AKs_bakcground", NC_DOUBLE, meta->aksback));
    } else {
        g.push_back(text("Vertical_Turb", "False"));
    }
    if (meta->hturb_on) {
        g.push_back(text("Horiz_Turb", "True"));
        if (!meta->vturb_on) g.push_back(num("seed", NC_INT, meta->seed));
    } else {
        g.push_back(text("Horiz_Turb", "False"));
    }
    if (meta->run_name) g.push_back(text("run_name", meta->run_name));
    g.push_back(text("executable", str(meta->executable)));
    g.push_back(text("output_loc", str(meta->output_loc)));
    g.push_back(text("administrator", str(meta->administrator)));
    g.push_back(text("institution", str(meta->institution)));
    {
        // date_and_time -> "(I4'/'I0.2'/'I0.2' 'I0.2':'I0.2':'I0.2)"
        char when[32] = "";
        const time_t now = time(nullptr);
        struct tm lt;
        if (localtime_r(&now, &lt))
            std::snprintf(when, sizeof when, "%4d/%02d/%02d %02d:%02d:%02d", lt.tm_year + 1900,
                          lt.tm_mon + 1, lt.tm_mday, lt.tm_hour, lt.tm_min, lt.tm_sec);
        g.push_back(text("Start_datetime", when));
    }
    g.push_back(text("SaltTempMean", meta->salt_temp_mean ? "TRUE" : "FALSE"));
    g.push_back(num("TempOffset", NC_DOUBLE, meta->temp_offset));
    g.push_back(text("source", "romspath_amd (MI355X particle step), netCDF classic 64-bit offset"));

    // slab sizes, then the header with placeholder offsets to learn its length
    for (Var& v : w->vars) {
        const uint64_t sz = v.type == NC_DOUBLE ? 8 : 4;
        v.vsize = v.per_particle ? sz * (uint64_t)numpar : sz;
    }
    auto header = [&]() {
        Be h;
        h.b.insert(h.b.end(), {'C', 'D', 'F', 2});
        h.u32(0);   // numrecs
        h.u32(TAG_DIM);
        h.u32(2);
        h.name("numpar");
        h.u32((uint32_t)numpar);
        h.name("time");
        h.u32(0);   // the record dimension
        h.atts(g);
        h.u32(TAG_VAR);
        h.u32((uint32_t)w->vars.size());
        for (const Var& v : w->vars) {
            h.name(v.name);
            if (v.rec && v.per_particle) {
                h.u32(2);
                h.u32(DIM_TIME);
                h.u32(DIM_NUMPAR);
            } else {
                h.u32(1);
                h.u32(v.rec ? DIM_TIME : DIM_NUMPAR);
            }
            h.atts(v.atts);
            h.u32((uint32_t)v.type);
            h.u32((uint32_t)v.vsize);
            h.u64(v.begin);
        }
        return h;
    };
    uint64_t off = header().b.size();
    for (Var& v : w->vars)
        if (!v.rec) {
            v.begin = off;
            off += v.vsize;
        }
    w->rec_begin = off;
    for (Var& v : w->vars)
        if (v.rec) {
            v.begin = off;
            off += v.vsize;
            w->recsize += v.vsize;
        }
    const Be h = header();
    w->f = std::fopen(path, "wb+");
    if (!w->f) {
        g_create_err = std::string("rp_nc_create: cannot create ") + path;
        delete w;
        return RP_E_IO;
    }
    int rc = put_at(w, 0, h.b.data(), h.b.size());
    if (!rc) {
        encode(w, NC_FLOAT, dob, nullptr, numpar);
        for (const Var& v : w->vars)
            if (v.name == "dob") rc = put_at(w, v.begin, w->buf.data(), w->buf.size());
    }
    if (rc || std::fflush(w->f) != 0) {
        g_create_err = "rp_nc_create: " + (w->err.empty() ? std::string("write failed") : w->err);
        std::fclose(w->f);
        delete w;
        return RP_E_IO;
    }
    *nc = w;
    return RP_OK;
}

int rp_nc_write(void* nc, double model_time, const double* age, const double* lon,
                const double* lat, const double* zp, const double* status, const int32_t* gid,
                const double* wd, const double* zeta, const double* bath, const double* salt,
                const double* temp, const double* hob, const double* bustr, const double* bvstr)
{
    Writer* w = (Writer*)nc;
    if (!w) return RP_E_ARG;
    if (!age || !lon || !lat || !zp || !status) return wfail(w, RP_E_ARG, "null particle column");
    if ((w->ngrid > 1 && !gid) || ((w->flags & RP_NC_WATERDEPTH) && !wd) ||
        ((w->flags & RP_NC_ZETA) && !zeta) || ((w->flags & RP_NC_BATH) && !bath) ||
        ((w->flags & RP_NC_SALTTEMP) && (!salt || !temp)) ||
        ((w->flags & RP_NC_BOTTOM) && (!hob || !bustr || !bvstr)))
        return wfail(w, RP_E_ARG, "a column the file defines is missing");
    if (w->nrec >= 0x7FFFFFFF) return wfail(w, RP_E_RANGE, "record count limit");
    const uint64_t base = w->rec_begin + (uint64_t)w->nrec * w->recsize;
    int rc = RP_OK;
    for (const Var& v : w->vars) {
        if (!v.rec) continue;
        const double* src = nullptr;
        const int32_t* isrc = nullptr;
        if (v.name == "model_time") src = &model_time;
        else if (v.name == "age") src = age;
        else if (v.name == "lon") src = lon;
        else if (v.name == "lat") src = lat;
        else if (v.name == "zp") src = zp;
        else if (v.name == "status") src = status;
        else if (v.name == "GID") isrc = gid;
        else if (v.name == "WaterDepth") src = wd;
        else if (v.name == "zeta") src = zeta;
        else if (v.name == "h") src = bath;
        else if (v.name == "salinity") src = salt;
        else if (v.name == "temperature") src = temp;
        else if (v.name == "HOB") src = hob;
        else if (v.name == "bustr") src = bustr;
        else if (v.name == "bvstr") src = bvstr;
        if (!src && !isrc) {
            // hitBottom / hitLand: the reference's counters stay at their
            // initial 0 (the increment, advection_module.f90:173, is commented out)
            w->buf.assign((size_t)w->numpar * 4, 0);
        } else {
            encode(w, v.type, src, isrc, v.per_particle ? w->numpar : 1);
        }
        if ((rc = put_at(w, base + (v.begin - w->rec_begin), w->buf.data(), w->buf.size()))) return rc;
    }
    w->nrec++;
    const unsigned char nr[4] = {(unsigned char)(w->nrec >> 24), (unsigned char)(w->nrec >> 16),
                                 (unsigned char)(w->nrec >> 8), (unsigned char)w->nrec};
    if ((rc = put_at(w, 4, nr, 4))) return rc;
    if (std::fflush(w->f) != 0) return wfail(w, RP_E_IO, "flush failed");
    return RP_OK;
}

int64_t rp_nc_records(void* nc) { return nc ? ((Writer*)nc)->nrec : -1; }

const char* rp_nc_error(void* nc)
{
    return nc ? ((Writer*)nc)->err.c_str() : g_create_err.c_str();
}

int rp_nc_close(void* nc)
{
    Writer* w = (Writer*)nc;
    if (!w) return RP_E_ARG;
    const int bad = std::fclose(w->f) != 0;
    delete w;
    return bad ? RP_E_IO : RP_OK;
}

}  // extern "C"
