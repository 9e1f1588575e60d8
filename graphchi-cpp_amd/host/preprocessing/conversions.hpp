// Graph input conversion (src/preprocessing/conversions.hpp:210-288, 611-790):
// edge lists ("<from> <to> [value]" with tab, comma or space delimiters, '#'
// and '%' comments) are read and handed to the device sharder, unless the
// graph is already sharded.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "api/chifilenames.hpp"
#include "logger/logger.hpp"
#include "preprocessing/sharder.hpp"
#include "util/cmdopts.hpp"

namespace graphchi {

template <typename EdgeDataType, typename FinalEdgeDataType>
void convert_edgelist(std::string inputfile, sharder<EdgeDataType, FinalEdgeDataType> &sh) {
    FILE *f = fopen(inputfile.c_str(), "r");
    if (!f) logstream(LOG_FATAL) << "Could not load: " << inputfile << std::endl;
    char line[1024];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '%') continue;
        char *p = line, *end = nullptr;
        while (*p == ' ' || *p == '\t' || *p == ',') p++;
        if (*p == '\n' || *p == '\r' || *p == 0) continue;
        unsigned long a = strtoul(p, &end, 10);
        if (end == p) logstream(LOG_FATAL) << "Input file is not in right format: " << line << std::endl;
        p = end;
        while (*p == ' ' || *p == '\t' || *p == ',') p++;
        unsigned long b = strtoul(p, &end, 10);
        if (end == p) logstream(LOG_FATAL) << "Input file is not in right format: " << line << std::endl;
        sh.preprocessing_add_edge((vid_t)a, (vid_t)b);
    }
    fclose(f);
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
