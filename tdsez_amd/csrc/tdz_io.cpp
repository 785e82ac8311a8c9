// HDF5 restart files and observables traces (C ABI of include/tdz_io.h).
//
// Checkpoint layout after TDSEZCheckpointer (src/checkpoint.cpp:101-216): root
// attributes version/step/time/time_step/dofs/propagator (64-bit integers as PETSc
// built with 64-bit indices writes them, doubles for times) and complex vectors as
// [n][2] double datasets carrying the attribute complex = 1, the shape PETSc's HDF5
// viewer gives a complex Vec. Trace layout after the observables writer
// (src/manager.cpp:51-201): extendable datasets of real columns, chunked by rows.
#include "tdz_io.h"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;
constexpr long long kVersion = 1;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

// RAII for HDF5 identifiers
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

hid_t chk(hid_t id, const char *what)
{
    if (id < 0) throw Error(std::string("HDF5: ") + what);
    return id;
}

void chk(herr_t e, const char *what, int)
{
    if (e < 0) throw Error(std::string("HDF5: ") + what);
}

void write_attr(hid_t loc, const char *name, hid_t type, const void *v)
{
    Hid sp(chk(H5Screate(H5S_SCALAR), "scalar space"), H5Sclose);
    Hid a(chk(H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), name), H5Aclose);
    chk(H5Awrite(a, type, v), name, 0);
}

bool has_attr(hid_t loc, const char *name) { return H5Aexists(loc, name) > 0; }

void read_attr(hid_t loc, const char *name, hid_t type, void *v)
{
    if (!has_attr(loc, name)) throw Error(std::string("checkpoint: attribute ") + name + " missing");
    Hid a(chk(H5Aopen(loc, name, H5P_DEFAULT), name), H5Aclose);
    chk(H5Aread(a, type, v), name, 0);
}

void write_complex(hid_t file, const char *name, long long n, const double *v)
{
    const hsize_t dims[2] = {(hsize_t)n, 2};
    Hid sp(chk(H5Screate_simple(2, dims, nullptr), "space"), H5Sclose);
    Hid ds(chk(H5Dcreate2(file, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), name), H5Dclose);
    chk(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v), name, 0);
    const long long one = 1;
    write_attr(ds, "complex", H5T_NATIVE_LLONG, &one);
}

bool read_complex(hid_t file, const char *name, long long n, double *v)
{
    if (H5Lexists(file, name, H5P_DEFAULT) <= 0) return false;
    Hid ds(chk(H5Dopen2(file, name, H5P_DEFAULT), name), H5Dclose);
    Hid sp(chk(H5Dget_space(ds), "space"), H5Sclose);
    hsize_t dims[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(sp) != 2) throw Error(std::string("checkpoint: ") + name + " is not [n][2]");
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    if ((long long)dims[0] != n || dims[1] != 2)
        throw Error(std::string("checkpoint: ") + name + " has the wrong shape");
    chk(H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v), name, 0);
    return true;
}

struct Trace {
    hid_t file = -1;
    int dim = 3;
};

// append rows x cols doubles to an extendable dataset, creating it on first use
void append_rows(hid_t file, const char *name, int cols, long long rows, const double *buf)
{
    if (rows <= 0 || !buf) return;
    hid_t ds;
    hsize_t cur[2] = {0, (hsize_t)cols};
    if (H5Lexists(file, name, H5P_DEFAULT) > 0) {
        ds = chk(H5Dopen2(file, name, H5P_DEFAULT), name);
        Hid sp(chk(H5Dget_space(ds), "space"), H5Sclose);
        H5Sget_simple_extent_dims(sp, cur, nullptr);
        if ((int)cur[1] != cols) {
            H5Dclose(ds);
            throw Error(std::string("trace: dataset ") + name + " has another column count");
        }
    } else {
        const hsize_t maxd[2] = {H5S_UNLIMITED, (hsize_t)cols}, chunk[2] = {256, (hsize_t)cols};
        Hid dcpl(chk(H5Pcreate(H5P_DATASET_CREATE), "dcpl"), H5Pclose);
        chk(H5Pset_chunk(dcpl, 2, chunk), "chunk", 0);
        Hid sp(chk(H5Screate_simple(2, cur, maxd), "space"), H5Sclose);
        ds = chk(H5Dcreate2(file, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT), name);
    }
    Hid dset(ds, H5Dclose);
    const hsize_t nd[2] = {cur[0] + (hsize_t)rows, (hsize_t)cols};
    chk(H5Dset_extent(dset, nd), "extend", 0);
    Hid fsp(chk(H5Dget_space(dset), "space"), H5Sclose);
    const hsize_t start[2] = {cur[0], 0}, count[2] = {(hsize_t)rows, (hsize_t)cols};
    chk(H5Sselect_hyperslab(fsp, H5S_SELECT_SET, start, nullptr, count, nullptr), "select", 0);
    Hid msp(chk(H5Screate_simple(2, count, nullptr), "memspace"), H5Sclose);
    chk(H5Dwrite(dset, H5T_NATIVE_DOUBLE, msp, fsp, H5P_DEFAULT, buf), name, 0);
}

}  // namespace

extern "C" {

const char *tdz_io_last_error(void) { return g_err.c_str(); }

int tdz_io_checkpoint_write(const char *path, long long step, double time, double dt, int propagator,
                            long long dofs, const double *psi, int nslots, int order, const int *guess_count,
                            const double *const *guess)
{
    return guarded([&] {
        if (!path || !psi || dofs <= 0) throw Error("checkpoint: bad arguments");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        const std::string tmp = std::string(path) + ".tmp";
        {
            Hid f(chk(H5Fcreate(tmp.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), tmp.c_str()), H5Fclose);
            const long long prop = propagator;
            write_attr(f, "version", H5T_NATIVE_LLONG, &kVersion);
            write_attr(f, "step", H5T_NATIVE_LLONG, &step);
            write_attr(f, "time", H5T_NATIVE_DOUBLE, &time);
            write_attr(f, "time_step", H5T_NATIVE_DOUBLE, &dt);
            write_attr(f, "dofs", H5T_NATIVE_LLONG, &dofs);
            write_attr(f, "propagator", H5T_NATIVE_LLONG, &prop);
            write_complex(f, "psi", dofs, psi);
            if (propagator != 0 && guess && guess_count)   // CN keeps no guess history
                for (int s = 0; s < nslots; ++s) {
                    char name[64];
                    const long long cnt = guess_count[s];
                    std::snprintf(name, sizeof name, "guess_count_%d", s);
                    write_attr(f, name, H5T_NATIVE_LLONG, &cnt);
                    for (int j = 0; j < guess_count[s] && j < order; ++j) {
                        std::snprintf(name, sizeof name, "guess_%d_%d", s, j);
                        write_complex(f, name, dofs, guess[(size_t)s * order + j]);
                    }
                }
            chk(H5Fflush(f, H5F_SCOPE_GLOBAL), "flush", 0);
        }
        if (std::rename(tmp.c_str(), path) != 0)
            throw Error("checkpoint: could not rename " + tmp + " to " + path);
    });
}

int tdz_io_checkpoint_read(const char *path, long long dofs, double dt, int propagator, long long *step,
                           double *time, double *psi, int nslots, int order, int *guess_count, double *const *guess)
{
    return guarded([&] {
        if (!path || !psi) throw Error("checkpoint: bad arguments");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        FILE *probe = std::fopen(path, "rb");
        if (!probe) throw Error(std::string("Restart = 1: no checkpoint file ") + path);
        std::fclose(probe);
        Hid f(chk(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), path), H5Fclose);
        long long version = 0, st = 0, n = -1, prop = -1;
        double t = 0, dtf = 0;
        read_attr(f, "version", H5T_NATIVE_LLONG, &version);
        read_attr(f, "step", H5T_NATIVE_LLONG, &st);
        read_attr(f, "time", H5T_NATIVE_DOUBLE, &t);
        read_attr(f, "time_step", H5T_NATIVE_DOUBLE, &dtf);
        read_attr(f, "dofs", H5T_NATIVE_LLONG, &n);
        read_attr(f, "propagator", H5T_NATIVE_LLONG, &prop);
        char msg[512];
        if (version != kVersion) {
            std::snprintf(msg, sizeof msg, "checkpoint %s: version %lld, expected %lld", path, version, kVersion);
            throw Error(msg);
        }
        if (n != dofs) {
            std::snprintf(msg, sizeof msg, "checkpoint %s has %lld DOFs, this grid %lld: restart with the same grid",
                          path, n, dofs);
            throw Error(msg);
        }
        if (std::fabs(dtf - dt) > 1e-12 * std::fabs(dt)) {
            std::snprintf(msg, sizeof msg, "checkpoint %s was written with TimeStep = %g, this input has %g", path,
                          dtf, dt);
            throw Error(msg);
        }
        if (prop != propagator) {
            std::snprintf(msg, sizeof msg, "checkpoint %s was written by another Propagator", path);
            throw Error(msg);
        }
        if (!read_complex(f, "psi", dofs, psi)) throw Error(std::string("checkpoint ") + path + ": no psi");
        if (step) *step = st;
        if (time) *time = t;
        if (guess && guess_count)
            for (int s = 0; s < nslots; ++s) {
                char name[64];
                std::snprintf(name, sizeof name, "guess_count_%d", s);
                long long cnt = 0;
                if (has_attr(f, name)) read_attr(f, name, H5T_NATIVE_LLONG, &cnt);
                if (cnt > order) cnt = order;
                int got = 0;
                for (int j = 0; j < cnt; ++j) {
                    std::snprintf(name, sizeof name, "guess_%d_%d", s, j);
                    if (!read_complex(f, name, dofs, guess[(size_t)s * order + j])) break;
                    ++got;
                }
                guess_count[s] = got;
            }
    });
}

int tdz_io_trace_open(const char *path, int dim, int append, void **handle)
{
    return guarded([&] {
        if (!path || !handle || dim < 1 || dim > 3) throw Error("trace: bad arguments");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        hid_t f = -1;
        FILE *probe = append ? std::fopen(path, "rb") : nullptr;
        if (probe) {
            std::fclose(probe);
            f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
        } else {
            f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        }
        chk(f, path);
        Trace *t = new Trace;
        t->file = f;
        t->dim = dim;
        *handle = t;
    });
}

int tdz_io_trace_append(void *handle, long long rows, const double *dipoles, const double *energies,
                        const double *autocorr)
{
    return guarded([&] {
        Trace *t = static_cast<Trace *>(handle);
        if (!t || t->file < 0) throw Error("trace: not open");
        append_rows(t->file, "dipoles", 1 + 3 * t->dim, rows, dipoles);
        append_rows(t->file, "energies", 7, rows, energies);
        append_rows(t->file, "autocorrelation", 3, rows, autocorr);
        chk(H5Fflush(t->file, H5F_SCOPE_GLOBAL), "flush", 0);
    });
}

int tdz_io_trace_close(void *handle)
{
    return guarded([&] {
        Trace *t = static_cast<Trace *>(handle);
        if (!t) return;
        if (t->file >= 0) H5Fclose(t->file);
        delete t;
    });
}

int tdz_io_trace_read(const char *path, const char *dataset, long long *rows, int *cols, double *out)
{
    return guarded([&] {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Hid f(chk(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), path), H5Fclose);
        if (H5Lexists(f, dataset, H5P_DEFAULT) <= 0) throw Error(std::string("trace: no dataset ") + dataset);
        Hid ds(chk(H5Dopen2(f, dataset, H5P_DEFAULT), dataset), H5Dclose);
        Hid sp(chk(H5Dget_space(ds), "space"), H5Sclose);
        hsize_t dims[2] = {0, 0};
        H5Sget_simple_extent_dims(sp, dims, nullptr);
        if (rows) *rows = (long long)dims[0];
        if (cols) *cols = (int)dims[1];
        if (out && dims[0] > 0) chk(H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out), dataset, 0);
    });
}

}  // extern "C"
