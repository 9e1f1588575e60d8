// Reads the adjacency stream of a shard back into (src, dst) pairs: the format
// finish_shard writes (src/preprocessing/sharder.hpp:773-819) and
// memory_shard::load_vertices decodes (src/shards/memoryshard.hpp:448-467):
//   count byte (0xff -> uint32 count follows), then count uint32 dst ids for
//   the current src; a 0 byte followed by k means 1 + k vertices without
//   out-edges in this shard.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "api/chifilenames.hpp"
#include "engine/device.hpp"
#include "logger/logger.hpp"

namespace graphchi {

inline std::vector<uint8_t> read_file_bytes(const std::string &path) {
    std::vector<uint8_t> buf;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) logstream(LOG_FATAL) << "could not open " << path << std::endl;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize(n > 0 ? (size_t)n : 0);
    if (n > 0 && fread(buf.data(), 1, (size_t)n, f) != (size_t)n) {
        fclose(f);
        logstream(LOG_FATAL) << "short read from " << path << std::endl;
    }
    fclose(f);
    return buf;
}

// Appends the shard's edges; returns false on a malformed stream.
inline bool decode_adjacency(const std::vector<uint8_t> &adj, std::vector<vid_t> &src, std::vector<vid_t> &dst) {
    size_t p = 0, n = adj.size();
    uint64_t vid = 0;
    while (p < n) {
        uint8_t ns = adj[p++];
        if (ns == 0) {
            if (p >= n) return false;
            vid += 1 + adj[p++];
            continue;
        }
        uint32_t cnt = ns;
        if (ns == 0xff) {
            if (p + 4 > n) return false;
            memcpy(&cnt, &adj[p], 4);
            p += 4;
        }
        if (p + 4ull * cnt > n) return false;
        for (uint32_t k = 0; k < cnt; k++) {
            uint32_t d;
            memcpy(&d, &adj[p], 4);
            p += 4;
            src.push_back((vid_t)vid);
            dst.push_back(d);
        }
        vid++;
    }
    return true;
}

// All edges of a sharded graph, shard by shard.
inline void load_shard_edges(const std::string &base, int nshards, std::vector<vid_t> &src, std::vector<vid_t> &dst) {
    std::vector<vid_t> ends = load_interval_ends(base, nshards);
    for (size_t p = 0; p < ends.size(); p++) {
        std::vector<uint8_t> adj = read_file_bytes(filename_shard_adj(base, (int)p, nshards));
        if (!decode_adjacency(adj, src, dst))
            logstream(LOG_FATAL) << "malformed shard " << filename_shard_adj(base, (int)p, nshards) << std::endl;
    }
}

inline uint64_t file_size_or_die(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) logstream(LOG_FATAL) << "could not open " << path << std::endl;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fclose(f);
    return n > 0 ? (uint64_t)n : 0;
}

// Reads the first n bytes of a file into pinned host memory (the upload source).
inline void read_file_pinned(const std::string &path, uint64_t n, device::pinned<uint8_t> &buf) {
    if (buf.size() < n) buf.resize(n);
    if (!n) return;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) logstream(LOG_FATAL) << "could not open " << path << std::endl;
    const size_t got = fread(buf.get(), 1, n, f);
    fclose(f);
    if (got != n) logstream(LOG_FATAL) << "short read from " << path << std::endl;
}

// The device load path: every shard's .adj bytes are read into pinned host
// memory, uploaded once and decoded in HBM, all shards in one call
// (gchi_shard_decode_many), and the edges stay there for the device build.
// Same edge sequence as load_shard_edges.  Returns the edge count.
inline uint64_t load_shard_edges_device(const std::string &base, int nshards, device::buffer<uint32_t> &src,
                                        device::buffer<uint32_t> &dst, std::vector<uint64_t> *per_shard = nullptr) {
    device::open();
    std::vector<vid_t> ends = load_interval_ends(base, nshards);
    std::vector<device::buffer<uint8_t>> bytes(ends.size());
    std::vector<uint64_t> nbytes(ends.size()), count(ends.size());
    uint64_t largest = 0;
    for (size_t p = 0; p < ends.size(); p++) {
        nbytes[p] = file_size_or_die(filename_shard_adj(base, (int)p, nshards));
        largest = std::max(largest, nbytes[p]);
    }
    device::pinned<uint8_t> staging(largest);
    for (size_t p = 0; p < ends.size(); p++) {
        read_file_pinned(filename_shard_adj(base, (int)p, nshards), nbytes[p], staging);
        bytes[p].resize(nbytes[p]);
        if (nbytes[p]) bytes[p].upload(staging.get(), nbytes[p]);
    }
    std::vector<const uint8_t *> ptrs(ends.size());
    for (size_t p = 0; p < ends.size(); p++) ptrs[p] = bytes[p].get();
    size_t ws = 0;
    device::check(gchi_shard_decode_many_workspace(nbytes.data(), (uint32_t)ends.size(), &ws),
                  "shard decode workspace");
    device::workspace w(ws);
    uint64_t cap = 0;  // every edge takes 4 bytes of its stream
    for (uint64_t b : nbytes) cap += b / 4;
    src.resize(cap);
    dst.resize(cap);
    device::check(gchi_shard_decode_many(ptrs.data(), nbytes.data(), (uint32_t)ends.size(), src.get(), dst.get(), cap,
                                         count.data(), nullptr, w.get(), ws, nullptr),
                  ("decoding the shards of " + base).c_str());
    uint64_t total = 0;
    for (uint64_t c : count) total += c;
    if (per_shard) *per_shard = count;
    return total;
}

}  // namespace graphchi
