// On-disk names of a sharded graph (src/api/chifilenames.hpp:65-153, 200-275,
// 400-421): the files the device sharder writes and the engine reads.
#pragma once
#include <sys/stat.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "graphchi_types.hpp"
#include "logger/logger.hpp"

namespace graphchi {

template <typename VertexDataType>
std::string filename_vertex_data(std::string base) {
    std::stringstream ss;
    ss << base << "." << sizeof(VertexDataType) << "B.vout";
    return ss.str();
}
inline std::string filename_degree_data(std::string base) { return base + "_degs.bin"; }
inline std::string filename_intervals(std::string base, int nshards) {
    std::stringstream ss;
    ss << base << "." << nshards << ".intervals";
    return ss.str();
}
inline std::string filename_shard_adj(std::string base, int p, int nshards) {
    std::stringstream ss;
    ss << base << ".edata_azv." << p << "_" << nshards << ".adj";
    return ss.str();
}
template <typename EdgeDataType>
std::string filename_shard_edata(std::string base, int p, int nshards) {
    std::stringstream ss;
    ss << base << ".edata.";
#ifndef GRAPHCHI_DISABLE_COMPRESSION
    ss << ".Z.";
#endif
    ss << "e" << sizeof(EdgeDataType) << "B." << p << "_" << nshards;
    return ss.str();
}
inline std::string dirname_shard_edata_block(std::string edata, size_t blocksize) {
    std::stringstream ss;
    ss << edata << "_blockdir_" << blocksize;
    return ss.str();
}

inline bool file_exists(const std::string &f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}

inline size_t get_num_vertices(std::string base) {
    std::ifstream in((base + ".numvertices").c_str());
    size_t n = 0;
    if (!(in >> n)) logstream(LOG_FATAL) << "could not read " << base << ".numvertices" << std::endl;
    return n;
}

inline std::vector<vid_t> load_interval_ends(std::string base, int nshards) {
    std::ifstream in(filename_intervals(base, nshards).c_str());
    std::vector<vid_t> ends;
    unsigned long x;
    while (in >> x) ends.push_back((vid_t)x);
    return ends;
}

// Number of shards of an already sharded graph, 0 when none are found: the
// adjacency files, the intervals and the degree file must all exist.
inline int find_shards(std::string base, std::string shard_string = "auto") {
    int lo = 1, hi = 2400;
    if (shard_string != "auto" && shard_string != "0") lo = hi = atoi(shard_string.c_str());
    if (!file_exists(filename_degree_data(base))) return 0;
    for (int n = lo; n <= hi; n++) {
        if (!file_exists(filename_intervals(base, n))) continue;
        std::vector<vid_t> ends = load_interval_ends(base, n);
        bool ok = !ends.empty();
        for (size_t p = 0; ok && p < ends.size(); p++) ok = file_exists(filename_shard_adj(base, (int)p, n));
        if (ok) return n;
    }
    return 0;
}

}  // namespace graphchi
