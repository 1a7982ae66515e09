// rp_nc_check.cpp -- host-only driver of the netCDF reader and writer, built
// with AddressSanitizer and UndefinedBehaviorSanitizer (romspath_amd/build.py,
// build_nc_check) for the tests that feed the reader damaged files.
//
//   rp_nc_check read <file> [var ...]   open, query and read every named
//                                       variable (all records); failures
//                                       must come back as codes
//   rp_nc_check write <file>            write a small trajectory file, then
//                                       read every variable back
//
// Exit status 0 when every call returned (a code or success) and the
// round trip matched; a sanitizer report aborts with its own status.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rp_abi.h"
#include "../../include/rp_ncin.h"
#include "../../include/rp_ncout.h"

namespace {

int read_file(const char* path, int nvar, char** vars)
{
    void* f = nullptr;
    const int rc = rp_ncin_open(&f, path);
    if (rc != 0) {
        std::printf("open %s: %d %s\n", path, rc, rp_ncin_error(nullptr));
        return 0;
    }
    for (const char* d : {"xi_rho", "eta_rho", "s_rho", "ocean_time", "numpar", "time", "nope"}) {
        int64_t len = -1;
        (void)rp_ncin_dim(f, d, &len);
    }
    char buf[64];
    int64_t alen = 0;
    (void)rp_ncin_att_text(f, nullptr, "title", buf, sizeof buf, &alen);
    (void)rp_ncin_att_text(f, "ocean_time", "units", buf, sizeof buf, &alen);
    for (int k = 0; k < nvar; k++) {
        int32_t nd = 0, isrec = 0;
        int64_t shape[8] = {0};
        if (rp_ncin_var(f, vars[k], &nd, shape, &isrec) != 0) continue;
        if (nd < 0 || nd > 8) {
            std::printf("var %s: rank %d out of range\n", vars[k], nd);
            return 1;
        }
        int64_t count = 1;
        bool sane = true;
        for (int i = isrec ? 1 : 0; i < nd; i++) {
            if (shape[i] < 0 || shape[i] > (int64_t)1 << 24) sane = false;
            else count *= shape[i];
        }
        if (!sane || count > (int64_t)1 << 24) {
            // a damaged header may announce huge shapes: a read of the whole
            // announced slice must still fail with a code, not fault
            std::vector<double> small(16);
            (void)rp_ncin_get(f, vars[k], 0, small.data(), (int64_t)small.size());
            continue;
        }
        std::vector<double> out((size_t)count + 1);
        const int64_t nrec = isrec ? shape[0] : 1;
        for (int64_t r = 0; r < nrec && r < 64; r++)
            (void)rp_ncin_get(f, vars[k], r, out.data(), count);
        (void)rp_ncin_get(f, vars[k], nrec + 3, out.data(), count);   // past the last record
        (void)rp_ncin_get(f, vars[k], 0, out.data(), count + 1);      // wrong count
    }
    return rp_ncin_close(f) == 0 ? 0 : 1;
}

int write_roundtrip(const char* path)
{
    const int64_t n = 37;
    rp_nc_meta m;
    std::memset(&m, 0, sizeof m);
    m.time_units = "seconds since 2000-01-01 00:00:00";
    m.run_name = "sanitizer round trip";
    m.idt = 120;
    m.days = 1.0;
    m.seed = 3;
    std::vector<double> dob(n), age(n), lon(n), lat(n), z(n), st(n), wd(n), ze(n), h(n);
    std::vector<int32_t> gid(n);
    for (int64_t i = 0; i < n; i++) {
        dob[i] = 60.0 * (double)i;
        gid[i] = 1 + (int32_t)(i % 2);
    }
    void* nc = nullptr;
    int rc = rp_nc_create(&nc, path, n, 2, 7, &m, dob.data());
    if (rc != 0) {
        std::printf("create: %d %s\n", rc, rp_nc_error(nullptr));
        return 1;
    }
    for (int rec = 0; rec < 3; rec++) {
        for (int64_t i = 0; i < n; i++) {
            age[i] = 120.0 * rec;
            lon[i] = -70.0 + 0.01 * (double)i;
            lat[i] = 40.0 + 0.02 * (double)rec;
            z[i] = -1.0 - (double)i;
            st[i] = 1.0;
            wd[i] = 30.0;
            ze[i] = 0.1;
            h[i] = 29.9;
        }
        rc = rp_nc_write(nc, 3600.0 * rec, age.data(), lon.data(), lat.data(), z.data(), st.data(),
                         gid.data(), wd.data(), ze.data(), h.data(), nullptr, nullptr, nullptr,
                         nullptr, nullptr);
        if (rc != 0) return 1;
    }
    if (rp_nc_records(nc) != 3 || rp_nc_close(nc) != 0) return 1;
    void* f = nullptr;
    if (rp_ncin_open(&f, path) != 0) return 1;
    std::vector<double> got(n);
    if (rp_ncin_get(f, "zp", 2, got.data(), n) != 0) return 1;
    for (int64_t i = 0; i < n; i++)
        if (got[i] != (double)(float)(-1.0 - (double)i)) return 1;
    char units[128];
    int64_t len = 0;
    if (rp_ncin_att_text(f, "model_time", "units", units, sizeof units, &len) != 0) return 1;
    if (std::strcmp(units, m.time_units) != 0) return 1;
    return rp_ncin_close(f) == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc >= 3 && std::strcmp(argv[1], "read") == 0) return read_file(argv[2], argc - 3, argv + 3);
    if (argc == 3 && std::strcmp(argv[1], "write") == 0) return write_roundtrip(argv[2]);
    std::fprintf(stderr, "usage: rp_nc_check read <file> [var ...] | write <file>\n");
    return 2;
}
