// Graph input conversion (src/preprocessing/conversions.hpp:210-288, 611-790):
// edge lists ("<from> <to> [value]" with tab, comma or space delimiters, '#'
// and '%' comments) are parsed on the device and handed to the device sharder,
// unless the graph is already sharded.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "api/chifilenames.hpp"
#include "engine/device.hpp"
#include "logger/logger.hpp"
#include "preprocessing/sharder.hpp"
#include "util/cmdopts.hpp"

namespace graphchi {

// convert_edgelist's loop (conversions.hpp:210-288) on the device: the file is
// read into pinned host memory, uploaded once and parsed by gchi_parse_edgelist
// (the same fields, comments, atoi values and self-loop filter); the pairs stay
// in HBM for the sharder.  A line without two fields stops with the fatal log,
// as the reference's assertion does.
template <typename EdgeDataType, typename FinalEdgeDataType>
void convert_edgelist(std::string inputfile, sharder<EdgeDataType, FinalEdgeDataType> &sh) {
    device::open();
    FILE *f = fopen(inputfile.c_str(), "rb");
    if (!f) logstream(LOG_FATAL) << "Could not load: " << inputfile << std::endl;
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    const uint64_t n = len > 0 ? (uint64_t)len : 0;
    device::pinned<uint8_t> text(n);
    if (n && fread(text.get(), 1, n, f) != n) logstream(LOG_FATAL) << "short read from " << inputfile << std::endl;
    fclose(f);
    device::buffer<uint8_t> d_text(n);
    if (n) d_text.upload(text.get(), n);
    size_t wb = 0;
    device::check(gchi_parse_edgelist_workspace(n, &wb), "gchi_parse_edgelist_workspace");
    device::workspace w(wb);
    uint64_t ne = 0;
    uint32_t maxid = 0;
    device::check(gchi_parse_edgelist(d_text.get(), n, nullptr, nullptr, 0, &ne, &maxid, w.get(), wb, nullptr),
                  "Input file is not in right format (gchi_parse_edgelist)");
    device::buffer<uint32_t> src(ne), dst(ne);
    if (ne)
        device::check(gchi_parse_edgelist(d_text.get(), n, src.get(), dst.get(), ne, &ne, &maxid, w.get(), wb, nullptr),
                      "Input file is not in right format (gchi_parse_edgelist)");
    sh.add_device_edges(std::move(src), std::move(dst), ne, maxid);
}

// Binary pairs of uint32 (from, to), back to back.
template <typename EdgeDataType, typename FinalEdgeDataType>
void convert_binedgelist(std::string inputfile, sharder<EdgeDataType, FinalEdgeDataType> &sh) {
    FILE *f = fopen(inputfile.c_str(), "rb");
    if (!f) logstream(LOG_FATAL) << "Could not load: " << inputfile << std::endl;
    vid_t pair[2];
    while (fread(pair, sizeof(vid_t), 2, f) == 2) sh.preprocessing_add_edge(pair[0], pair[1]);
    fclose(f);
}

template <typename EdgeDataType, typename FinalEdgeDataType>
int convert(std::string basefilename, std::string nshards_string, bool no_values = false) {
    sharder<EdgeDataType, FinalEdgeDataType> sh(basefilename);
    if (no_values) sh.set_no_edgevalues();
    std::string type = get_option_string("filetype", "edgelist");
    sh.start_preprocessing();
    if (type == "edgelist") convert_edgelist(basefilename, sh);
    else if (type == "binedgelist") convert_binedgelist(basefilename, sh);
    else logstream(LOG_FATAL) << "filetype must be edgelist or binedgelist" << std::endl;
    sh.end_preprocessing();
    return sh.execute_sharding(nshards_string);
}

template <typename EdgeDataType, typename FinalEdgeType>
int convert_if_notexists(std::string basefilename, std::string nshards_string, bool &didexist) {
    int n = find_shards(basefilename, nshards_string);
    didexist = n > 0;
    if (n) {
        logstream(LOG_INFO) << "Found preprocessed files for " << basefilename << ", num shards=" << n << std::endl;
        return n;
    }
    return convert<EdgeDataType, FinalEdgeType>(basefilename, nshards_string);
}

template <typename EdgeDataType, typename FinalEdgeType>
int convert_if_notexists(std::string basefilename, std::string nshards_string) {
    bool b;
    return convert_if_notexists<EdgeDataType, FinalEdgeType>(basefilename, nshards_string, b);
}

template <typename EdgeDataType>
int convert_if_notexists(std::string basefilename, std::string nshards_string) {
    bool b;
    return convert_if_notexists<EdgeDataType, EdgeDataType>(basefilename, nshards_string, b);
}

template <typename EdgeDataType>
int convert_if_notexists_novalues(std::string basefilename, std::string nshards_string) {
    // default-valued edge data of sizeof(EdgeDataType) bytes, as convert<dummyC<ET>, ET>
    // (conversions.hpp:705-726)
    int n = find_shards(basefilename, nshards_string);
    if (n) return n;
    return convert<dummy, EdgeDataType>(basefilename, nshards_string);
}

}  // namespace graphchi
