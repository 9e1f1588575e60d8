// Host-side ownership of device memory and graphs for the engines: RAII
// buffers over gchi_malloc/gchi_free, the one-time upload of a graph into an
// HBM-resident CSC (gchi_csc_build), and status checking that turns a library
// error into the reference's fatal log.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "gchi.h"
#include "graphchi_types.hpp"
#include "logger/logger.hpp"

namespace graphchi {
namespace device {

inline void check(int st, const char *what) {
    if (st != GCHI_OK) {
        char buf[1024];
        gchi_last_error(buf, sizeof(buf));
        logstream(LOG_FATAL) << what << ": " << gchi_strerror(st) << ": " << buf << std::endl;
    }
}

inline void open() {
    static bool opened = false;
    if (!opened) {
        check(gchi_device_open(0), "gchi_device_open");
        opened = true;
    }
}

template <typename T>
class buffer {
  public:
    buffer() {}
    explicit buffer(size_t n) { resize(n); }
    ~buffer() { release(); }
    buffer(const buffer &) = delete;
    buffer &operator=(const buffer &) = delete;
    buffer(buffer &&o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    buffer &operator=(buffer &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void resize(size_t n) {
        release();
        n_ = n;
        void *p = nullptr;
        check(gchi_malloc(&p, (n ? n : 1) * sizeof(T)), "gchi_malloc");
        p_ = static_cast<T *>(p);
    }
    void release() {
        if (p_) gchi_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void upload(const T *h, size_t n) { check(gchi_memcpy(p_, h, n * sizeof(T), nullptr), "upload"); }
    void download(T *h, size_t n) const { check(gchi_memcpy(h, p_, n * sizeof(T), nullptr), "download"); }
    std::vector<T> to_host(size_t n) const {
        std::vector<T> v(n);
        if (n) download(v.data(), n);
        return v;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// Pinned host memory (gchi_host_alloc): what the one-time uploads read from.
template <typename T>
class pinned {
  public:
    pinned() {}
    explicit pinned(size_t n) { resize(n); }
    ~pinned() { release(); }
    pinned(const pinned &) = delete;
    pinned &operator=(const pinned &) = delete;
    void resize(size_t n) {
        release();
        void *p = nullptr;
        check(gchi_host_alloc(&p, (n ? n : 1) * sizeof(T)), "gchi_host_alloc");
        p_ = static_cast<T *>(p);
        n_ = n;
    }
    void release() {
        if (p_) gchi_host_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

class workspace : public buffer<uint8_t> {
  public:
    explicit workspace(size_t n) : buffer<uint8_t>(n) {}
};

// A graph resident in HBM as CSC + degrees.
struct graph {
    uint32_t n = 0;
    uint64_t e = 0;
    buffer<uint32_t> src, dst, in_off, in_src, outdeg, indeg;
    gchi_dgraph dg{};

    graph(const std::vector<vid_t> &h_src, const std::vector<vid_t> &h_dst, uint32_t nvertices, bool keep_edges) {
        open();
        n = nvertices;
        e = h_src.size();
        src.resize(e);
        dst.resize(e);
        if (e) {
            src.upload(h_src.data(), e);
            dst.upload(h_dst.data(), e);
        }
        build(keep_edges);
    }
    // From an edge list already in HBM (the device shard decode).
    graph(buffer<uint32_t> &&d_src, buffer<uint32_t> &&d_dst, uint64_t nedges, uint32_t nvertices, bool keep_edges)
        : n(nvertices), e(nedges), src(std::move(d_src)), dst(std::move(d_dst)) {
        open();
        build(keep_edges);
    }

  private:
    void build(bool keep_edges) {
        in_off.resize((size_t)n + 1);
        in_src.resize(e);
        outdeg.resize(n);
        indeg.resize(n);
        size_t ws = 0;
        check(gchi_csc_build_workspace(e, n, &ws), "csc workspace");
        workspace w(ws);
        check(gchi_csc_build(src.get(), dst.get(), e, n, in_off.get(), in_src.get(), outdeg.get(), indeg.get(),
                             w.get(), ws, nullptr),
              "gchi_csc_build");
        check(gchi_work_status(w.get(), nullptr), "gchi_csc_build status");  // waits for the stream
        dg = gchi_dgraph{n, 0, e, in_off.get(), in_src.get(), outdeg.get(), indeg.get()};
        if (!keep_edges) {
            src.release();
            dst.release();
        }
    }
};

}  // namespace device
}  // namespace graphchi
