/*
 * HDF5 snapshot and geometry files, host code only (no GPU): the on-disk side of
 * K7/node/save_fields.py:10-48 -- one file per saved step, one gzip-compressed dataset
 * per field component, named by the component ('ex' ... 'hz') -- and of the geometry
 * file read at K7/main.py:70-80: root attributes nx, ny, nz (int64 scalars) and
 * coeff_use (a fixed-length NULLPAD ASCII string, one byte for ''), as
 * K7/test/create_sample_h5.py:3-8 writes them through h5py.
 *
 * Datasets are chunked with the deflate filter at the caller's level (h5py's
 * compression='gzip' is level 4). The chunks are compressed on `nthreads` host threads
 * with zlib's compress2 -- the zlib stream the HDF5 deflate filter decodes -- and written
 * in chunk order by the calling thread with H5Dwrite_chunk, so any HDF5 reader inflates
 * them through its own filter pipeline. A bounded window of compressed chunks is held
 * at once. Chunks are at most 1 MiB: the full extent of the fast axes, the slowest axes
 * halved until the chunk fits.
 *
 * Built by fdtd_accelerate_amd/build.py with g++ against the HDF5 1.10 C library and
 * zlib, into libmifdtd_h5.so (declared in include/mifdtd_h5.h).
 */
#include "mifdtd_h5.h"

#include <hdf5.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// Owns an HDF5 identifier and closes it with the matching H5?close.
struct Hid {
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid()
    {
        if (id >= 0) close(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    bool ok() const { return id >= 0; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

constexpr size_t kChunkBytes = size_t(1) << 20;

// Chunk shape: the dataset's extent, slowest axes halved (rounding up) until the chunk
// holds at most kChunkBytes.
void pick_chunk(int rank, const hsize_t *dims, size_t esize, hsize_t *chunk)
{
    for (int r = 0; r < rank; ++r) chunk[r] = dims[r];
    for (;;) {
        size_t bytes = esize;
        for (int r = 0; r < rank; ++r) bytes *= chunk[r];
        if (bytes <= kChunkBytes) return;
        int a = 0;
        while (a < rank && chunk[a] == 1) ++a;
        if (a == rank) return;
        chunk[a] = (chunk[a] + 1) / 2;
    }
}

struct Grid3 {   // dataset and chunk extents padded to three axes (leading 1s)
    hsize_t d[3], c[3], n[3];
};

Grid3 grid3(int rank, const hsize_t *dims, const hsize_t *chunk)
{
    Grid3 g;
    for (int r = 0; r < 3; ++r) {
        const int s = r - (3 - rank);
        g.d[r] = s >= 0 ? dims[s] : 1;
        g.c[r] = s >= 0 ? chunk[s] : 1;
        g.n[r] = (g.d[r] + g.c[r] - 1) / g.c[r];
    }
    return g;
}

// Copy chunk `idx` (row-major over the chunk grid) of the C-order array into `raw`
// (chunk-shaped, zero where the chunk passes the dataset's edge); origin in elements.
void gather_chunk(const Grid3 &g, size_t esize, const unsigned char *src, size_t idx, unsigned char *raw,
                  hsize_t origin[3])
{
    const hsize_t c2 = idx % g.n[2], c1 = (idx / g.n[2]) % g.n[1], c0 = idx / (g.n[2] * g.n[1]);
    origin[0] = c0 * g.c[0];
    origin[1] = c1 * g.c[1];
    origin[2] = c2 * g.c[2];
    const hsize_t e0 = std::min(g.c[0], g.d[0] - origin[0]);
    const hsize_t e1 = std::min(g.c[1], g.d[1] - origin[1]);
    const hsize_t e2 = std::min(g.c[2], g.d[2] - origin[2]);
    if (e0 < g.c[0] || e1 < g.c[1] || e2 < g.c[2]) std::memset(raw, 0, g.c[0] * g.c[1] * g.c[2] * esize);
    for (hsize_t a0 = 0; a0 < e0; ++a0)
        for (hsize_t a1 = 0; a1 < e1; ++a1) {
            const size_t s = (((origin[0] + a0) * g.d[1] + origin[1] + a1) * g.d[2] + origin[2]) * esize;
            const size_t t = ((a0 * g.c[1] + a1) * g.c[2]) * esize;
            std::memcpy(raw + t, src + s, e2 * esize);
        }
}

int write_dataset(hid_t file, const char *name, const void *data, int f64, int rank, const hsize_t *dims,
                  int level, int nthreads)
{
    const size_t esize = f64 ? 8 : 4;
    hsize_t chunk[3];
    pick_chunk(rank, dims, esize, chunk);
    Hid dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    if (!dcpl.ok() || H5Pset_chunk(dcpl.id, rank, chunk) < 0 || H5Pset_deflate(dcpl.id, (unsigned)level) < 0 ||
        H5Pset_fill_time(dcpl.id, H5D_FILL_TIME_NEVER) < 0)
        return fail(MIFDTD_H5_EIO, "%s: dataset creation properties", name);
    Hid space(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Hid dset(H5Dcreate2(file, name, f64 ? H5T_IEEE_F64LE : H5T_IEEE_F32LE, space.id, H5P_DEFAULT, dcpl.id,
                        H5P_DEFAULT),
             H5Dclose);
    if (!space.ok() || !dset.ok()) return fail(MIFDTD_H5_EIO, "%s: cannot create the dataset", name);

    const Grid3 g = grid3(rank, dims, chunk);
    const size_t total = g.n[0] * g.n[1] * g.n[2];
    const size_t raw_bytes = g.c[0] * g.c[1] * g.c[2] * esize;
    const size_t window = 4 * (size_t)nthreads;
    const unsigned char *src = static_cast<const unsigned char *>(data);

    std::vector<std::vector<unsigned char>> out(total);
    std::vector<char> ready(total, 0);
    std::mutex m;
    std::condition_variable cv_ready, cv_space;
    size_t next = 0, written = 0;
    bool abort = false;
    int zerr = Z_OK;

    auto work = [&]() {
        std::vector<unsigned char> raw(raw_bytes);
        for (;;) {
            size_t idx;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_space.wait(lk, [&] { return abort || next >= total || next < written + window; });
                if (abort || next >= total) return;
                idx = next++;
            }
            hsize_t origin[3];
            gather_chunk(g, esize, src, idx, raw.data(), origin);
            uLongf len = compressBound(raw_bytes);
            std::vector<unsigned char> z(len);
            const int zr = compress2(z.data(), &len, raw.data(), raw_bytes, level);
            {
                std::lock_guard<std::mutex> lk(m);
                if (zr != Z_OK) {
                    abort = true;
                    zerr = zr;
                } else {
                    z.resize(len);
                    out[idx] = std::move(z);
                    ready[idx] = 1;
                }
            }
            cv_ready.notify_all();
            if (zr != Z_OK) cv_space.notify_all();
        }
    };
    auto worker = [&]() {  // an allocation failure stops the write instead of the process
        try {
            work();
        } catch (...) {
            {
                std::lock_guard<std::mutex> lk(m);
                abort = true;
                zerr = Z_MEM_ERROR;
            }
            cv_ready.notify_all();
            cv_space.notify_all();
        }
    };
    std::vector<std::thread> pool;
    try {
        for (int t = 0; t < nthreads; ++t) pool.emplace_back(worker);
    } catch (...) {  // thread creation failed: stop the ones started
        {
            std::lock_guard<std::mutex> lk(m);
            abort = true;
        }
        cv_space.notify_all();
        for (auto &t : pool) t.join();
        return fail(MIFDTD_H5_EIO, "%s: cannot start %d compression threads", name, nthreads);
    }

    int rc = 0;
    for (size_t i = 0; i < total; ++i) {
        std::vector<unsigned char> z;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_ready.wait(lk, [&] { return abort || ready[i]; });
            if (abort) break;
            z = std::move(out[i]);
        }
        hsize_t origin3[3], offset[3];
        const hsize_t c2 = i % g.n[2], c1 = (i / g.n[2]) % g.n[1], c0 = i / (g.n[2] * g.n[1]);
        origin3[0] = c0 * g.c[0];
        origin3[1] = c1 * g.c[1];
        origin3[2] = c2 * g.c[2];
        for (int r = 0; r < rank; ++r) offset[r] = origin3[r + 3 - rank];
        if (H5Dwrite_chunk(dset.id, H5P_DEFAULT, 0, offset, z.size(), z.data()) < 0) {
            rc = fail(MIFDTD_H5_EIO, "%s: writing chunk %zu failed", name, i);
            std::lock_guard<std::mutex> lk(m);
            abort = true;
        } else {
            std::lock_guard<std::mutex> lk(m);
            written = i + 1;
        }
        cv_space.notify_all();
        if (rc) break;
    }
    {
        std::lock_guard<std::mutex> lk(m);
        if (written < total) abort = true;
    }
    cv_space.notify_all();
    for (auto &t : pool) t.join();
    if (rc) return rc;
    if (zerr != Z_OK) return fail(MIFDTD_H5_EIO, "%s: zlib compress2 failed (%d)", name, zerr);
    if (written != total) return fail(MIFDTD_H5_EIO, "%s: %zu of %zu chunks written", name, written, total);
    return 0;
}

int check_dims(int rank, const long long *dims, hsize_t *hd)
{
    if (rank < 1 || rank > 3) return fail(MIFDTD_H5_EINVAL, "rank %d is not 1, 2 or 3", rank);
    for (int r = 0; r < rank; ++r) {
        if (dims[r] < 1) return fail(MIFDTD_H5_EINVAL, "dims[%d] = %lld < 1", r, dims[r]);
        hd[r] = (hsize_t)dims[r];
    }
    return 0;
}

int write_i64_attr(hid_t loc, const char *name, long long v)
{
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(loc, name, H5T_STD_I64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!sp.ok() || !a.ok() || H5Awrite(a.id, H5T_NATIVE_LLONG, &v) < 0)
        return fail(MIFDTD_H5_EIO, "cannot write attribute %s", name);
    return 0;
}

int read_i64_attr(hid_t loc, const char *name, long long *v)
{
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(MIFDTD_H5_EFORMAT, "attribute '%s' is missing", name);
    Hid t(H5Aget_type(a.id), H5Tclose);
    if (H5Tget_class(t.id) != H5T_INTEGER) return fail(MIFDTD_H5_EFORMAT, "attribute '%s' is not an integer", name);
    if (H5Aread(a.id, H5T_NATIVE_LLONG, v) < 0) return fail(MIFDTD_H5_EIO, "cannot read attribute '%s'", name);
    return 0;
}

// Runs an entry point's body; no C++ exception crosses the C interface.
template <class F>
int guarded(const char *what, F &&body)
{
    try {
        return body();
    } catch (const std::exception &e) {
        return fail(MIFDTD_H5_EIO, "%s: %s", what, e.what());
    } catch (...) {
        return fail(MIFDTD_H5_EIO, "%s: unknown C++ exception", what);
    }
}

}  // namespace

extern "C" {

const char *mifdtd_h5_last_error(void) { return g_err; }

int mifdtd_h5_write_fields(const char *path, int ndset, const char *const *names, const void *const *data, int f64,
                           int rank, const long long *dims, int gzip_level, int nthreads)
{
    quiet();
    if (!path || !names || !data || !dims) return fail(MIFDTD_H5_EINVAL, "null argument");
    if (ndset < 1) return fail(MIFDTD_H5_EINVAL, "ndset %d < 1", ndset);
    if (gzip_level < 0 || gzip_level > 9) return fail(MIFDTD_H5_EINVAL, "gzip level %d outside 0..9", gzip_level);
    if (nthreads < 1 || nthreads > 256) return fail(MIFDTD_H5_EINVAL, "nthreads %d outside 1..256", nthreads);
    hsize_t hd[3];
    if (int rc = check_dims(rank, dims, hd)) return rc;
    for (int n = 0; n < ndset; ++n)
        if (!names[n] || !data[n]) return fail(MIFDTD_H5_EINVAL, "dataset %d: null name or data", n);
    return guarded(path, [&]() -> int {
        Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        if (!file.ok()) return fail(MIFDTD_H5_EIO, "cannot create %s", path);
        for (int n = 0; n < ndset; ++n)
            if (int rc = write_dataset(file.id, names[n], data[n], f64, rank, hd, gzip_level, nthreads)) return rc;
        return 0;
    });
}

int mifdtd_h5_write_geometry(const char *path, long long nx, long long ny, long long nz, const char *coeff_use)
{
    quiet();
    if (!path || !coeff_use) return fail(MIFDTD_H5_EINVAL, "null argument");
    return guarded(path, [&]() -> int {
        Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        if (!file.ok()) return fail(MIFDTD_H5_EIO, "cannot create %s", path);
        if (int rc = write_i64_attr(file.id, "nx", nx)) return rc;
        if (int rc = write_i64_attr(file.id, "ny", ny)) return rc;
        if (int rc = write_i64_attr(file.id, "nz", nz)) return rc;
        const size_t len = std::strlen(coeff_use);
        std::vector<char> buf(std::max<size_t>(len, 1), 0);
        std::memcpy(buf.data(), coeff_use, len);
        Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
        Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
        if (!t.ok() || H5Tset_size(t.id, buf.size()) < 0 || H5Tset_strpad(t.id, H5T_STR_NULLPAD) < 0)
            return fail(MIFDTD_H5_EIO, "coeff_use string type");
        Hid a(H5Acreate2(file.id, "coeff_use", t.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        if (!a.ok() || H5Awrite(a.id, t.id, buf.data()) < 0) return fail(MIFDTD_H5_EIO, "cannot write coeff_use");
        return 0;
    });
}

int mifdtd_h5_read_geometry(const char *path, long long *nx, long long *ny, long long *nz, char *coeff_use, int cap)
{
    quiet();
    if (!path || !nx || !ny || !nz || !coeff_use || cap < 1) return fail(MIFDTD_H5_EINVAL, "null argument");
    return guarded(path, [&]() -> int {
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!file.ok()) return fail(MIFDTD_H5_EIO, "cannot open %s as HDF5", path);
        if (int rc = read_i64_attr(file.id, "nx", nx)) return rc;
        if (int rc = read_i64_attr(file.id, "ny", ny)) return rc;
        if (int rc = read_i64_attr(file.id, "nz", nz)) return rc;
        Hid a(H5Aopen(file.id, "coeff_use", H5P_DEFAULT), H5Aclose);
        if (!a.ok()) return fail(MIFDTD_H5_EFORMAT, "attribute 'coeff_use' is missing");
        Hid ft(H5Aget_type(a.id), H5Tclose);
        if (H5Tget_class(ft.id) != H5T_STRING) return fail(MIFDTD_H5_EFORMAT, "attribute 'coeff_use' is not a string");
        std::string s;
        if (H5Tis_variable_str(ft.id) > 0) {
            Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
            H5Tset_size(mt.id, H5T_VARIABLE);
            char *p = nullptr;
            if (H5Aread(a.id, mt.id, &p) < 0) return fail(MIFDTD_H5_EIO, "cannot read coeff_use");
            s = p ? p : "";
            H5free_memory(p);
        } else {
            const size_t n = H5Tget_size(ft.id);
            std::vector<char> buf(n + 1, 0);
            Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
            H5Tset_size(mt.id, n);
            H5Tset_strpad(mt.id, H5T_STR_NULLPAD);
            if (H5Aread(a.id, mt.id, buf.data()) < 0) return fail(MIFDTD_H5_EIO, "cannot read coeff_use");
            s.assign(buf.data(), std::strlen(buf.data()));
            while (!s.empty() && s.back() == ' ') s.pop_back();   // a SPACEPAD string
        }
        if ((int)s.size() >= cap) return fail(MIFDTD_H5_EINVAL, "coeff_use '%s' longer than %d", s.c_str(), cap - 1);
        std::memcpy(coeff_use, s.c_str(), s.size() + 1);
        return 0;
    });
}

int mifdtd_h5_dataset_info(const char *path, const char *name, int *rank, long long *dims, int *f64,
                           int *gzip_level)
{
    quiet();
    if (!path || !name || !rank || !dims || !f64 || !gzip_level) return fail(MIFDTD_H5_EINVAL, "null argument");
    return guarded(path, [&]() -> int {
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!file.ok()) return fail(MIFDTD_H5_EIO, "cannot open %s as HDF5", path);
        Hid d(H5Dopen2(file.id, name, H5P_DEFAULT), H5Dclose);
        if (!d.ok()) return fail(MIFDTD_H5_EFORMAT, "no dataset '%s' in %s", name, path);
        Hid sp(H5Dget_space(d.id), H5Sclose);
        const int r = H5Sget_simple_extent_ndims(sp.id);
        if (r < 1 || r > MIFDTD_H5_MAX_RANK) return fail(MIFDTD_H5_EFORMAT, "%s: rank %d", name, r);
        hsize_t hd[MIFDTD_H5_MAX_RANK];
        H5Sget_simple_extent_dims(sp.id, hd, nullptr);
        *rank = r;
        for (int i = 0; i < r; ++i) dims[i] = (long long)hd[i];
        Hid t(H5Dget_type(d.id), H5Tclose);
        if (H5Tget_class(t.id) != H5T_FLOAT) return fail(MIFDTD_H5_EFORMAT, "%s is not a floating-point dataset", name);
        *f64 = H5Tget_size(t.id) == 8;
        *gzip_level = -1;
        Hid pl(H5Dget_create_plist(d.id), H5Pclose);
        const int nf = H5Pget_nfilters(pl.id);
        for (int i = 0; i < nf; ++i) {
            unsigned flags = 0, cd[8] = {0}, fc = 0;
            size_t ncd = 8;
            char fname[64];
            if (H5Pget_filter2(pl.id, (unsigned)i, &flags, &ncd, cd, sizeof(fname), fname, &fc) == H5Z_FILTER_DEFLATE)
                *gzip_level = ncd > 0 ? (int)cd[0] : 0;
        }
        return 0;
    });
}

int mifdtd_h5_read_dataset(const char *path, const char *name, void *out, long long nbytes, int f64)
{
    quiet();
    if (!path || !name || !out) return fail(MIFDTD_H5_EINVAL, "null argument");
    return guarded(path, [&]() -> int {
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (!file.ok()) return fail(MIFDTD_H5_EIO, "cannot open %s as HDF5", path);
        Hid d(H5Dopen2(file.id, name, H5P_DEFAULT), H5Dclose);
        if (!d.ok()) return fail(MIFDTD_H5_EFORMAT, "no dataset '%s' in %s", name, path);
        Hid t(H5Dget_type(d.id), H5Tclose);
        if (!t.ok() || H5Tget_class(t.id) != H5T_FLOAT)
            return fail(MIFDTD_H5_EFORMAT, "%s is not a floating-point dataset", name);
        Hid sp(H5Dget_space(d.id), H5Sclose);
        const long long n = (long long)H5Sget_simple_extent_npoints(sp.id);
        if (n * (f64 ? 8 : 4) != nbytes)
            return fail(MIFDTD_H5_EINVAL, "%s holds %lld elements, not %lld bytes of %s", name, n, nbytes,
                        f64 ? "float64" : "float32");
        if (H5Dread(d.id, f64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
            return fail(MIFDTD_H5_EIO, "cannot read %s", name);
        return 0;
    });
}

}  // extern "C"
