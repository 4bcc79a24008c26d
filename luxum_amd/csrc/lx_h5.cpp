// lx_h5.cpp -- HDF5 output and input in the reference's layout.
//
// Replaces fileio (lib/io/fileio.cc:55-247) for one process: a file per
// output, one FP64 dataset per snapshot named "t=<t>" (the reference formats
// t with an ostream, fileio.cc:175), shaped by the global extents of the
// field (staggered axes have n-1 entries), C order with k fastest.  The
// reference's 3D solvers write |E|^2 at equal indices into an Ez-shaped
// array (maxwell_three_dim_pml.cc:264-272, zero outside i<Nx-1, j<Ny-1,
// k<Nz-1); the 2D solvers write Ey.
//
// libhdf5 (serial, 1.10 C API) is loaded with dlopen on first use: the path
// comes from LX_HDF5_LIB or the usual sonames; a missing library is an
// LX_EIO error, not a load-time dependency of libluxum.
//
// lx_writer: write() gathers the field on the compute stream into a device
// staging buffer (the step loop waits only for that kernel); a writer thread
// copies it to a pinned host slot on the DMA engine (lx_dma.hpp) once the
// gather is done and puts it into the file, so the time loop keeps running
// while files are written.  Without the HSA runtime's copy the slot is
// copied on the side stream instead.
#include <hip/hip_runtime.h>
#include <dlfcn.h>

#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <thread>

#include "../../include/luxum.h"
#include "lx_sim.hpp"
#include "lx_util.hpp"

using lx::fail;

namespace {

// ---- the HDF5 1.10 C API subset, resolved at run time ----------------------
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0;

struct H5 {
    herr_t (*open)(void);
    hid_t (*Fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*Fopen)(const char *, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*Sclose)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*Dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*Dopen2)(hid_t, const char *, hid_t);
    hid_t (*Dget_space)(hid_t);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*Dclose)(hid_t);
    int (*Lexists)(hid_t, const char *, hid_t);
    herr_t (*Eset_auto2)(hid_t, void *, void *);
    hid_t *native_double;
    bool ok = false;
};

std::mutex g_h5_mu;   // serialises every call into the library

H5 *h5api() {
    static H5 api;
    static bool tried = false;
    if (tried) return api.ok ? &api : nullptr;
    tried = true;
    const char *env = std::getenv("LX_HDF5_LIB");
    const char *cands[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    void *h = nullptr;
    for (const char *c : cands)
        if (c && *c && (h = dlopen(c, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) return nullptr;
    bool good = true;
    auto sym = [&](const char *n) {
        void *p = dlsym(h, n);
        if (!p) good = false;
        return p;
    };
    api.open = (herr_t(*)(void))sym("H5open");
    api.Fcreate = (hid_t(*)(const char *, unsigned, hid_t, hid_t))sym("H5Fcreate");
    api.Fopen = (hid_t(*)(const char *, unsigned, hid_t))sym("H5Fopen");
    api.Fclose = (herr_t(*)(hid_t))sym("H5Fclose");
    api.Screate_simple = (hid_t(*)(int, const hsize_t *, const hsize_t *))sym("H5Screate_simple");
    api.Sclose = (herr_t(*)(hid_t))sym("H5Sclose");
    api.Sget_simple_extent_dims = (int (*)(hid_t, hsize_t *, hsize_t *))sym("H5Sget_simple_extent_dims");
    api.Dcreate2 = (hid_t(*)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t))sym("H5Dcreate2");
    api.Dopen2 = (hid_t(*)(hid_t, const char *, hid_t))sym("H5Dopen2");
    api.Dget_space = (hid_t(*)(hid_t))sym("H5Dget_space");
    api.Dwrite = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *))sym("H5Dwrite");
    api.Dread = (herr_t(*)(hid_t, hid_t, hid_t, hid_t, hid_t, void *))sym("H5Dread");
    api.Dclose = (herr_t(*)(hid_t))sym("H5Dclose");
    api.Lexists = (int (*)(hid_t, const char *, hid_t))sym("H5Lexists");
    api.Eset_auto2 = (herr_t(*)(hid_t, void *, void *))sym("H5Eset_auto2");
    api.native_double = (hid_t *)sym("H5T_NATIVE_DOUBLE_g");
    if (!good || api.open() < 0) return nullptr;
    api.Eset_auto2(0, nullptr, nullptr);   // errors come back as status codes, not stderr
    api.ok = true;
    return &api;
}

lx_status need_h5(H5 **out) {
    *out = h5api();
    if (!*out) return fail(LX_EIO, "libhdf5 could not be loaded (set LX_HDF5_LIB to libhdf5.so of the 1.10 C API)");
    return LX_OK;
}

std::string dataset_name(double t) {
    std::ostringstream o;   // the reference writes  temp << "t=" << t
    o << "t=" << t;
    return o.str();
}

// write one dataset into an open file (caller holds g_h5_mu)
lx_status put(H5 *a, hid_t f, const std::string &name, const double *data, const hsize_t dims[3]) {
    const hid_t sp = a->Screate_simple(3, dims, nullptr);
    if (sp < 0) return fail(LX_EIO, "H5Screate_simple failed");
    if (a->Lexists(f, name.c_str(), kDefault) > 0) {
        a->Sclose(sp);
        return fail(LX_EIO, "dataset %s already exists", name.c_str());
    }
    const hid_t ds = a->Dcreate2(f, name.c_str(), *a->native_double, sp, kDefault, kDefault, kDefault);
    if (ds < 0) {
        a->Sclose(sp);
        return fail(LX_EIO, "H5Dcreate2(%s) failed", name.c_str());
    }
    const herr_t e = a->Dwrite(ds, *a->native_double, 0, 0, kDefault, data);
    a->Dclose(ds);
    a->Sclose(sp);
    if (e < 0) return fail(LX_EIO, "H5Dwrite(%s) failed", name.c_str());
    return LX_OK;
}

}  // namespace

// ---- the asynchronous writer -------------------------------------------------
struct lx_writer {
    lx_sim *sim = nullptr;
    int what = LX_OUT_FIELD, comp = 0;
    int ext[3] = {0, 0, 0};
    size_t n = 0;
    hid_t file = -1;
    double *dstage[2] = {nullptr, nullptr};    // device staging per slot
    double *host[2] = {nullptr, nullptr};      // pinned host slot
    hipEvent_t copied[2] = {nullptr, nullptr}; // D2H of the slot done (side-stream copies)
    hipEvent_t gathered[2] = {nullptr, nullptr};   // gather into dstage done (side-stream copies)
    lx::DmaSlot dma[2];
    bool use_dma = false;
    double slot_t[2] = {0, 0};
    bool busy[2] = {false, false};
    int next = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> queue;
    bool stop = false;
    lx_status err = LX_OK;
    std::string errmsg;
    int64_t written = 0;
};

static void writer_loop(lx_writer *w) {
    (void)hipSetDevice(w->sim->ctx->device);
    H5 *a = h5api();
    for (;;) {
        int slot;
        {
            std::unique_lock<std::mutex> lk(w->mu);
            w->cv.wait(lk, [&] { return w->stop || !w->queue.empty(); });
            if (w->queue.empty()) return;
            slot = w->queue.front();
            w->queue.pop_front();
        }
        lx_status st = LX_OK;
        std::string msg;
        if (w->use_dma) {
            if ((st = lx::dma_copy_marked(&w->dma[slot], w->host[slot], w->dstage[slot], sizeof(double) * w->n)))
                msg = lx_last_error();
        } else if (hipEventSynchronize(w->copied[slot]) != hipSuccess) {
            st = LX_EDEVICE;
            msg = "snapshot copy failed";
        }
        if (!st) {
            const hsize_t dims[3] = {(hsize_t)w->ext[0], (hsize_t)w->ext[1], (hsize_t)w->ext[2]};
            std::lock_guard<std::mutex> lk(g_h5_mu);
            st = put(a, w->file, dataset_name(w->slot_t[slot]), w->host[slot], dims);
            if (st) msg = lx_last_error();
        }
        std::lock_guard<std::mutex> lk(w->mu);
        if (st && !w->err) {
            w->err = st;
            w->errmsg = msg;
        }
        if (!st) w->written++;
        w->busy[slot] = false;
        w->cv.notify_all();
    }
}

extern "C" {

lx_status lx_h5_write_dataset(const char *path, const char *name, const double *data, const int64_t dims[3],
                             int32_t append) {
    if (!path || !name || !data || !dims) return fail(LX_EINVAL, "lx_h5_write: null pointer");
    H5 *a;
    lx_status st = need_h5(&a);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_h5_mu);
    const hid_t f = append ? a->Fopen(path, kAccRdwr, kDefault) : a->Fcreate(path, kAccTrunc, kDefault, kDefault);
    if (f < 0) return fail(LX_EIO, "cannot %s %s", append ? "open" : "create", path);
    const hsize_t d[3] = {(hsize_t)dims[0], (hsize_t)dims[1], (hsize_t)dims[2]};
    st = put(a, f, name, data, d);
    a->Fclose(f);
    return st;
}

lx_status lx_h5_write(const char *path, double t, const double *data, const int64_t dims[3], int32_t append) {
    return lx_h5_write_dataset(path, dataset_name(t).c_str(), data, dims, append);
}

lx_status lx_h5_read(const char *path, const char *dataset, double *dst, size_t n, int64_t dims[3]) {
    if (!path || !dataset) return fail(LX_EINVAL, "lx_h5_read: null pointer");
    H5 *a;
    lx_status st = need_h5(&a);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_h5_mu);
    const hid_t f = a->Fopen(path, kAccRdonly, kDefault);
    if (f < 0) return fail(LX_EIO, "cannot open %s", path);
    const hid_t ds = a->Dopen2(f, dataset, kDefault);
    if (ds < 0) {
        a->Fclose(f);
        return fail(LX_EIO, "no dataset %s in %s", dataset, path);
    }
    const hid_t sp = a->Dget_space(ds);
    hsize_t d[3] = {1, 1, 1};
    const int rank = a->Sget_simple_extent_dims(sp, d, nullptr);
    if (rank < 1 || rank > 3) st = fail(LX_EIO, "dataset %s has rank %d", dataset, rank);
    size_t m = 1;
    for (int q = 0; q < 3; ++q) {
        if (q >= rank) d[q] = 1;
        m *= (size_t)d[q];
        if (dims) dims[q] = (int64_t)d[q];
    }
    if (!st && dst) {
        if (n != m) st = fail(LX_EINVAL, "buffer holds %zu values, dataset %zu", n, m);
        else if (a->Dread(ds, *a->native_double, 0, 0, kDefault, dst) < 0) st = fail(LX_EIO, "H5Dread(%s) failed", dataset);
    }
    a->Sclose(sp);
    a->Dclose(ds);
    a->Fclose(f);
    return st;
}

lx_status lx_h5_dataset_name(double t, char *buf, size_t n) {
    if (!buf || n == 0) return fail(LX_EINVAL, "lx_h5_dataset_name: no buffer");
    const std::string s = dataset_name(t);
    if (s.size() + 1 > n) return fail(LX_EINVAL, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return LX_OK;
}

lx_status lx_writer_open(lx_sim *s, const char *path, int32_t what, int32_t comp, lx_writer **out) {
    if (!s || !path || !out) return fail(LX_EINVAL, "lx_writer_open: null pointer");
    int ext[3];
    if (what == LX_OUT_E2) {
        if (s->planar) return fail(LX_EINVAL, "|E|^2 output is the 3D solvers' (2D writes Ey)");
        lx::comp_extent(s->N, 0, LX_EZ, ext);
    } else if (what == LX_OUT_FIELD) {
        if (comp < 0 || comp > 5 || !lx::comp_extent(s->N, s->planar, comp, ext))
            return fail(LX_EINVAL, "bad component %d", comp);
    } else {
        return fail(LX_EINVAL, "unknown output kind %d", what);
    }
    H5 *a;
    lx_status st = need_h5(&a);
    if (st) return st;
    lx_writer *w = new (std::nothrow) lx_writer;
    if (!w) return fail(LX_ENOMEM, "out of host memory");
    w->sim = s;
    w->what = what;
    w->comp = comp;
    for (int q = 0; q < 3; ++q) w->ext[q] = ext[q];
    w->n = (size_t)ext[0] * ext[1] * ext[2];
    {
        std::lock_guard<std::mutex> lk(g_h5_mu);
        w->file = a->Fcreate(path, kAccTrunc, kDefault, kDefault);
    }
    if (w->file < 0) {
        delete w;
        return fail(LX_EIO, "cannot create %s", path);
    }
    bool good = hipSetDevice(s->ctx->device) == hipSuccess;
    for (int q = 0; q < 2 && good; ++q) {
        good = hipMalloc(&w->dstage[q], sizeof(double) * w->n) == hipSuccess &&
               hipHostMalloc(&w->host[q], sizeof(double) * w->n, hipHostMallocDefault) == hipSuccess &&
               hipEventCreateWithFlags(&w->copied[q], hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&w->gathered[q], hipEventDisableTiming) == hipSuccess;
    }
    if (good) w->use_dma = lx::dma_eligible(w->host[0], w->dstage[0], sizeof(double) * w->n) &&
                           lx::dma_eligible(w->host[1], w->dstage[1], sizeof(double) * w->n);
    if (!good) {
        (void)hipGetLastError();
        w->th = std::thread();
        lx_writer_close(w);
        return fail(LX_ENOMEM, "snapshot buffers of %zu values could not be allocated", w->n);
    }
    w->th = std::thread(writer_loop, w);
    *out = w;
    return LX_OK;
}

lx_status lx_writer_write(lx_writer *w, double t) {
    if (!w) return fail(LX_EINVAL, "null writer");
    lx_sim *s = w->sim;
    int slot;
    {
        std::unique_lock<std::mutex> lk(w->mu);
        if (w->err) return fail(w->err, "%s", w->errmsg.c_str());
        slot = w->next;
        w->cv.wait(lk, [&] { return !w->busy[slot]; });   // at most two snapshots in flight
        w->busy[slot] = true;
        w->next ^= 1;
    }
    lx_status st = hipSetDevice(s->ctx->device) == hipSuccess ? LX_OK : fail(LX_EDEVICE, "hipSetDevice failed");
    if (!st && !w->use_dma) st = lx::ensure_side_stream(s);
    if (!st) lx::member_enter(s);
    // gather on the compute stream (ordered with the steps); the writer
    // thread copies the slot out once the gather is done, or the side stream does
    if (!st) st = lx::gather_output(s, w->what, w->comp, w->dstage[slot], w->ext);
    if (!st && w->use_dma) st = lx::dma_mark(&w->dma[slot], s->cs);
    if (!st && !w->use_dma &&
        (hipEventRecord(w->gathered[slot], s->cs) != hipSuccess || hipStreamWaitEvent(s->ss, w->gathered[slot], 0) != hipSuccess ||
         hipMemcpyAsync(w->host[slot], w->dstage[slot], sizeof(double) * w->n, hipMemcpyDeviceToHost, s->ss) != hipSuccess ||
         hipEventRecord(w->copied[slot], s->ss) != hipSuccess))
        st = fail(LX_EDEVICE, "snapshot copy could not be enqueued");
    std::lock_guard<std::mutex> lk(w->mu);
    if (st) {
        w->busy[slot] = false;
        w->cv.notify_all();
        return st;
    }
    w->slot_t[slot] = t;
    w->queue.push_back(slot);
    w->cv.notify_all();
    return LX_OK;
}

lx_status lx_writer_close(lx_writer *w) {
    if (!w) return LX_OK;
    {
        std::lock_guard<std::mutex> lk(w->mu);
        w->stop = true;
        w->cv.notify_all();
    }
    if (w->th.joinable()) w->th.join();
    lx_status st = w->err;
    std::string msg = w->errmsg;
    if (w->file >= 0) {
        H5 *a = h5api();
        std::lock_guard<std::mutex> lk(g_h5_mu);
        if (a && a->Fclose(w->file) < 0 && !st) {
            st = LX_EIO;
            msg = "H5Fclose failed";
        }
    }
    for (int q = 0; q < 2; ++q) {
        lx::dma_release(&w->dma[q]);
        if (w->copied[q]) (void)hipEventSynchronize(w->copied[q]);
        if (w->copied[q]) (void)hipEventDestroy(w->copied[q]);
        if (w->gathered[q]) (void)hipEventSynchronize(w->gathered[q]);
        if (w->gathered[q]) (void)hipEventDestroy(w->gathered[q]);
        if (w->dstage[q]) (void)hipFree(w->dstage[q]);
        if (w->host[q]) (void)hipHostFree(w->host[q]);
    }
    delete w;
    if (st) return fail(st, "%s", msg.c_str());
    return LX_OK;
}

int64_t lx_writer_count(lx_writer *w) {
    if (!w) return 0;
    std::lock_guard<std::mutex> lk(w->mu);
    return w->written;
}

}  // extern "C"
