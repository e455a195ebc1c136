// netcdf_io.cpp -- dump / restart files in the reference's layout, written
// and read without libnetcdf.
//
// Format: netCDF classic (CDF-1, 32-bit offsets, big-endian), as created by
// the reference's CreateNetCDF / WriteNetCDF (src/netcdf.F90:39-580), or its
// 64-bit-offset variant (CDF-2) when a variable begins past 2 GiB, with
// itype = 1: dimensions nx, ny, nz, one; global attributes version (int),
// history (text), gamma, sigma, theta, ck, cm, f, tayl, psi, zeta, xmax, ymax
// (double); variables time(one) double with units="UNKNOWN",
// iteration(one) int, timestep(one) double, then R, T, U, V, W, Bx, By, Bz
// over (nz, ny, nx) as double (iprec 8) or float (iprec 4).  The reader
// accepts any CDF-1 or CDF-2 file holding those variables (double or float)
// and any attribute/dimension order.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhd.h"

namespace {

enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C };

const char* kFields[8] = {"R", "T", "U", "V", "W", "Bx", "By", "Bz"};
const char* kAttrs[11] = {"gamma", "sigma", "theta", "ck", "cm", "f", "tayl", "psi", "zeta", "xmax", "ymax"};

struct Out {
    std::vector<unsigned char> b;
    void u32(uint32_t v)
    {
        for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void i32(int32_t v) { u32((uint32_t)v); }
    void f64(double d)
    {
        uint64_t v;
        memcpy(&v, &d, 8);
        for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
    }
    void f32(float f)
    {
        uint32_t v;
        memcpy(&v, &f, 4);
        u32(v);
    }
    void pad4()
    {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string& s)
    {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad4();
    }
};

size_t type_size(int t)
{
    switch (t) {
    case NC_BYTE:
    case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT:
    case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    }
    return 0;
}

struct In {
    const unsigned char* p;
    size_t n, o = 0;
    bool ok = true;
    uint32_t u32()
    {
        if (o + 4 > n) { ok = false; return 0; }
        uint32_t v = ((uint32_t)p[o] << 24) | ((uint32_t)p[o + 1] << 16) | ((uint32_t)p[o + 2] << 8) | p[o + 3];
        o += 4;
        return v;
    }
    std::string name()
    {
        uint32_t len = u32();
        if (!ok || o + len > n) { ok = false; return ""; }
        std::string s((const char*)p + o, len);
        o += (len + 3) & ~3u;
        return s;
    }
    void skip_values(int type, uint32_t count)
    {
        size_t bytes = (type_size(type) * count + 3) & ~(size_t)3;
        if (o + bytes > n) { ok = false; return; }
        o += bytes;
    }
};

// fn(lo, hi) over [0, n) in contiguous chunks on up to 16 host threads
// (the byte-order conversions of the field arrays)
template <typename F>
void par_for(size_t n, F fn)
{
    unsigned t = std::thread::hardware_concurrency();
    t = t < 1 ? 1 : (t > 16 ? 16 : t);
    if (n < ((size_t)1 << 20)) t = 1;
    std::vector<std::thread> th;
    for (unsigned i = 1; i < t; ++i) {
        try {
            th.emplace_back(fn, n * i / t, n * (i + 1) / t);
        } catch (...) {
            fn(n * i / t, n * (i + 1) / t);   // no thread to be had: this chunk here
        }
    }
    fn((size_t)0, n / t);
    for (auto& x : th) x.join();
}

double be_f64(const unsigned char* q)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v = (v << 8) | q[i];
    double d;
    memcpy(&d, &v, 8);
    return d;
}

// The header of a dump: dimensions, global attributes and the variables with
// their begin offsets.  CDF-1 (32-bit signed offsets) when every variable
// begins below 2^31, as the reference's nf_create(NF_CLOBBER) writes;
// otherwise CDF-2 (64-bit offsets), which libnetcdf and scipy read as well.
// Returns 0, or MHD_E_ARG when a variable exceeds the format's 4 GiB limit.
int build_header(const mhd_dump_t* meta, int iprec, Out& h, int* version_out)
{
    const int nx = meta->nx, ny = meta->ny, nz = meta->nz;
    const size_t n3 = (size_t)nx * ny * nz;
    const int ftype = iprec == 8 ? NC_DOUBLE : NC_FLOAT;
    const size_t fsize = (size_t)iprec * n3;
    if (fsize > 0xfffffffcULL) return MHD_E_ARG;   // fixed-size variables of CDF-1/2
    for (int version = 1; version <= 2; ++version) {
        h.b.clear();
        h.b.insert(h.b.end(), {'C', 'D', 'F', (unsigned char)version});
        h.u32(0);  // numrecs (no record dimension)
        // dimensions
        h.u32(NC_DIMENSION);
        h.u32(4);
        const char* dn[4] = {"nx", "ny", "nz", "one"};
        const int dl[4] = {nx, ny, nz, 1};
        for (int i = 0; i < 4; ++i) {
            h.name(dn[i]);
            h.u32((uint32_t)dl[i]);
        }
        // global attributes
        const double* pv = &meta->gamma;
        h.u32(NC_ATTRIBUTE);
        h.u32(13);
        h.name("version");
        h.u32(NC_INT);
        h.u32(1);
        h.i32(meta->version);
        std::string hist(meta->history, strnlen(meta->history, sizeof(meta->history)));
        h.name("history");
        h.u32(NC_CHAR);
        h.u32((uint32_t)hist.size());
        h.b.insert(h.b.end(), hist.begin(), hist.end());
        h.pad4();
        for (int i = 0; i < 11; ++i) {
            h.name(kAttrs[i]);
            h.u32(NC_DOUBLE);
            h.u32(1);
            h.f64(pv[i]);
        }
        // variables: sizes and offsets follow the header, so lay the header
        // out with placeholder offsets first and patch them
        struct V { const char* name; int ndim; int dims[3]; int type; size_t vsize; bool units; };
        std::vector<V> vars = {{"time", 1, {3, 0, 0}, NC_DOUBLE, 8, true},
                               {"iteration", 1, {3, 0, 0}, NC_INT, 4, false},
                               {"timestep", 1, {3, 0, 0}, NC_DOUBLE, 8, false}};
        for (int f = 0; f < 8; ++f)
            vars.push_back({kFields[f], 3, {2, 1, 0}, ftype, (fsize + 3) & ~(size_t)3, false});
        h.u32(NC_VARIABLE);
        h.u32((uint32_t)vars.size());
        std::vector<size_t> begin_pos;
        for (const V& v : vars) {
            h.name(v.name);
            h.u32((uint32_t)v.ndim);
            for (int d = 0; d < v.ndim; ++d) h.u32((uint32_t)v.dims[d]);
            if (v.units) {
                h.u32(NC_ATTRIBUTE);
                h.u32(1);
                h.name("units");
                h.u32(NC_CHAR);
                h.u32(7);
                h.b.insert(h.b.end(), {'U', 'N', 'K', 'N', 'O', 'W', 'N'});
                h.pad4();
            } else {
                h.u32(0);
                h.u32(0);
            }
            h.u32((uint32_t)v.type);
            h.u32((uint32_t)v.vsize);
            begin_pos.push_back(h.b.size());
            h.u32(0);
            if (version == 2) h.u32(0);
        }
        size_t off = h.b.size();
        bool fits = true;
        for (size_t i = 0; i < vars.size(); ++i) {
            if (version == 1) {
                if (off > 0x7fffffffULL) { fits = false; break; }   // signed 32-bit offsets
                for (int s = 0; s < 4; ++s) h.b[begin_pos[i] + s] = (unsigned char)(off >> (24 - 8 * s));
            } else {
                for (int s = 0; s < 8; ++s) h.b[begin_pos[i] + s] = (unsigned char)((uint64_t)off >> (56 - 8 * s));
            }
            off += vars[i].vsize;
        }
        if (fits) {
            *version_out = version;
            return 0;
        }
    }
    return MHD_E_ARG;
}

}  // namespace

extern "C" {

long mhd_dump_header(const mhd_dump_t* meta, int iprec, unsigned char* buf, long cap, int* version)
{
    if (!meta || (iprec != 4 && iprec != 8) || meta->nx < 1 || meta->ny < 1 || meta->nz < 1) return MHD_E_ARG;
    Out h;
    int v = 0;
    const int e = build_header(meta, iprec, h, &v);
    if (e) return e;
    if (version) *version = v;
    if (buf && cap >= (long)h.b.size()) memcpy(buf, h.b.data(), h.b.size());
    return (long)h.b.size();
}

int mhd_write_dump(const char* path, const mhd_dump_t* meta, const double* const fields[8], int iprec)
{
    if (!path || !meta || !fields || (iprec != 4 && iprec != 8)) return MHD_E_ARG;
    const int nx = meta->nx, ny = meta->ny, nz = meta->nz;
    if (nx < 1 || ny < 1 || nz < 1) return MHD_E_ARG;
    const size_t n3 = (size_t)nx * ny * nz;
    for (int f = 0; f < 8; ++f)
        if (!fields[f]) return MHD_E_ARG;
    Out h;
    int version = 0;
    {
        const int e = build_header(meta, iprec, h, &version);
        if (e) return e;
    }
    const size_t fsize = (size_t)iprec * n3;
    // fields converted to big-endian on the host threads, one buffer per
    // field (a mapped output file measured slower on the GPU hosts)
    FILE* fp = fopen(path, "wb");
    if (!fp) return -13;
    bool ok = fwrite(h.b.data(), 1, h.b.size(), fp) == h.b.size();
    Out d;
    d.f64(meta->time);
    d.i32(meta->iteration);
    d.f64(meta->timestep);
    ok = ok && fwrite(d.b.data(), 1, d.b.size(), fp) == d.b.size();
    const size_t padded = (fsize + 3) & ~(size_t)3;
    std::vector<unsigned char> buf(padded);
    for (int f = 0; f < 8 && ok; ++f) {
        const double* src = fields[f];
        unsigned char* bp = buf.data();
        if (iprec == 8) {
            par_for(n3, [=](size_t lo, size_t hi) {
                for (size_t i = lo; i < hi; ++i) {
                    uint64_t v;
                    memcpy(&v, &src[i], 8);
                    v = __builtin_bswap64(v);
                    memcpy(bp + 8 * i, &v, 8);
                }
            });
        } else {
            par_for(n3, [=](size_t lo, size_t hi) {
                for (size_t i = lo; i < hi; ++i) {
                    const float x = (float)src[i];
                    uint32_t v;
                    memcpy(&v, &x, 4);
                    v = __builtin_bswap32(v);
                    memcpy(bp + 4 * i, &v, 4);
                }
            });
        }
        for (size_t i = fsize; i < padded; ++i) bp[i] = 0;
        ok = fwrite(bp, 1, padded, fp) == padded;
    }
    ok = (fclose(fp) == 0) && ok;
    return ok ? MHD_OK : -14;
}

// the file mapped read-only: a metadata query touches only the header pages
struct Mapped {
    const unsigned char* p = nullptr;
    size_t n = 0;
    ~Mapped()
    {
        if (p) munmap((void*)p, n);
    }
    const unsigned char* data() const { return p; }
    size_t size() const { return n; }
};

static int map_file(const char* path, Mapped& m)
{
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return -13;
    struct stat st;
    if (fstat(fd, &st) != 0) { close(fd); return -13; }
    if (st.st_size < 8) { close(fd); return -15; }
    void* p = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (p == MAP_FAILED) return -13;
    m.p = (const unsigned char*)p;
    m.n = (size_t)st.st_size;
    return 0;
}

// fields may be NULL to query meta only (dimensions, attributes, scalars)
int mhd_read_dump(const char* path, mhd_dump_t* meta, double* const fields[8])
{
    if (!path || !meta) return MHD_E_ARG;
    Mapped data;
    int e = map_file(path, data);
    if (e) return e;
    In in{data.data(), data.size()};
    if (memcmp(data.data(), "CDF", 3) != 0 || (data.data()[3] != 1 && data.data()[3] != 2)) return -15;
    const bool off64 = data.data()[3] == 2;   // CDF-2: 64-bit begin offsets
    in.o = 4;
    in.u32();  // numrecs
    std::vector<std::string> dname;
    std::vector<uint32_t> dlen;
    uint32_t tag = in.u32(), cnt = in.u32();
    if (tag == NC_DIMENSION)
        for (uint32_t i = 0; i < cnt && in.ok; ++i) {
            dname.push_back(in.name());
            dlen.push_back(in.u32());
        }
    memset(meta, 0, sizeof(*meta));
    for (size_t i = 0; i < dname.size(); ++i) {
        if (dname[i] == "nx") meta->nx = (int)dlen[i];
        if (dname[i] == "ny") meta->ny = (int)dlen[i];
        if (dname[i] == "nz") meta->nz = (int)dlen[i];
    }
    double* pv = &meta->gamma;
    tag = in.u32();
    cnt = in.u32();
    if (tag == NC_ATTRIBUTE)
        for (uint32_t i = 0; i < cnt && in.ok; ++i) {
            std::string nm = in.name();
            int type = (int)in.u32();
            uint32_t count = in.u32();
            const unsigned char* q = data.data() + in.o;
            if (in.o + 8 > data.size()) return -15;   // the value reads below stay inside the file
            if (nm == "version" && type == NC_INT && count == 1)
                meta->version = (int)(((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3]);
            if (nm == "history" && type == NC_CHAR) {
                size_t L = count < sizeof(meta->history) - 1 ? count : sizeof(meta->history) - 1;
                if (in.o + L > data.size()) return -15;
                memcpy(meta->history, q, L);
                meta->history[L] = 0;
            }
            for (int a = 0; a < 11; ++a)
                if (nm == kAttrs[a] && type == NC_DOUBLE && count >= 1) pv[a] = be_f64(q);
            in.skip_values(type, count);
        }
    tag = in.u32();
    cnt = in.u32();
    if (!in.ok || tag != NC_VARIABLE) return -15;
    const size_t n3 = (size_t)meta->nx * meta->ny * meta->nz;
    int found = 0;
    for (uint32_t i = 0; i < cnt && in.ok; ++i) {
        std::string nm = in.name();
        uint32_t nd = in.u32();
        for (uint32_t d = 0; d < nd; ++d) in.u32();
        uint32_t at = in.u32(), ac = in.u32();
        if (at == NC_ATTRIBUTE)
            for (uint32_t a = 0; a < ac && in.ok; ++a) {
                in.name();
                int type = (int)in.u32();
                uint32_t count = in.u32();
                in.skip_values(type, count);
            }
        int type = (int)in.u32();
        in.u32();  // vsize
        size_t begin = in.u32();
        if (off64) begin = (begin << 32) | in.u32();
        if (!in.ok || begin + 8 > data.size()) return -15;
        const unsigned char* q = data.data() + begin;
        if (nm == "time" && type == NC_DOUBLE) meta->time = be_f64(q);
        if (nm == "timestep" && type == NC_DOUBLE) meta->timestep = be_f64(q);
        if (nm == "iteration" && type == NC_INT) meta->iteration = (int)(((uint32_t)q[0] << 24) | (q[1] << 16) | (q[2] << 8) | q[3]);
        for (int f = 0; f < 8; ++f) {
            if (nm != kFields[f]) continue;
            const size_t need = n3 * type_size(type);
            if ((type != NC_DOUBLE && type != NC_FLOAT) || begin + need > data.size()) return -15;
            ++found;
            if (fields && fields[f]) {
                double* dst = fields[f];
                if (type == NC_DOUBLE)
                    par_for(n3, [=](size_t lo, size_t hi) {
                        for (size_t k = lo; k < hi; ++k) {
                            uint64_t v;
                            memcpy(&v, q + 8 * k, 8);
                            v = __builtin_bswap64(v);
                            memcpy(&dst[k], &v, 8);
                        }
                    });
                else
                    par_for(n3, [=](size_t lo, size_t hi) {
                        for (size_t k = lo; k < hi; ++k) {
                            uint32_t v;
                            memcpy(&v, q + 4 * k, 4);
                            v = __builtin_bswap32(v);
                            float x;
                            memcpy(&x, &v, 4);
                            dst[k] = (double)x;
                        }
                    });
            }
        }
    }
    if (!in.ok) return -15;
    return found == 8 ? MHD_OK : -15;
}

}  // extern "C"
