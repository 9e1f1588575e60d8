// The sharder's interface (src/preprocessing/sharder.hpp:332-548):
// start_preprocessing / preprocessing_add_edge / end_preprocessing /
// execute_sharding.  Edges are collected on the host (self-loops dropped,
// max id tracked, :427-446), uploaded once, and sharded on the device
// (gchi_shard_build: interval cut, per-shard (src,dst) order, degrees), which
// also writes the reference's file layout (gchi_shard_write_files).
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include "engine/device.hpp"
#include "graphchi_types.hpp"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "util/cmdopts.hpp"

namespace graphchi {

template <typename EdgeDataType, typename FinalEdgeDataType = EdgeDataType>
class sharder {
  public:
    explicit sharder(std::string basefilename) : base_(basefilename), m_("sharder") {}

    void set_no_edgevalues() { no_values_ = true; }
    void start_preprocessing() {
        m_.start_time("preprocessing");
        src_.clear();
        dst_.clear();
        max_id_ = 0;
    }
    void preprocessing_add_edge(vid_t from, vid_t to, EdgeDataType = EdgeDataType(), bool = false) {
        if (from == to) return;
        src_.push_back(from);
        dst_.push_back(to);
        max_id_ = std::max(max_id_, std::max(from, to));
    }
    void end_preprocessing() { m_.stop_time("preprocessing"); }

    int execute_sharding(std::string nshards_string) {
        m_.start_time("execute_sharding");
        device::open();
        int nshards = 0;
        if (nshards_string.find("auto") != std::string::npos || nshards_string == "0") {
            device::check(gchi_shard_auto_count(src_.size(), (uint32_t)sizeof(FinalEdgeDataType),
                                                get_option_int("membudget_mb", 1024), &nshards),
                          "gchi_shard_auto_count");
        } else {
            nshards = atoi(nshards_string.c_str());
        }
        if (nshards <= 0) logstream(LOG_FATAL) << "number of shards must be positive" << std::endl;
        const uint64_t e = src_.size();
        device::buffer<uint32_t> s(e), d(e);
        if (e) {
            s.upload(src_.data(), e);
            d.upload(dst_.data(), e);
        }
        gchi_shards *sh = nullptr;
        device::check(gchi_shard_build(s.get(), d.get(), e, max_id_, nshards, &sh, nullptr), "gchi_shard_build");
        // edge-data blocks zlib-compressed unless the application opted out, as
        // the reference's sharder (src/util/ioutil.hpp:144-199)
#ifndef GRAPHCHI_DISABLE_COMPRESSION
        const int wflags = GCHI_SHARD_ZLIB;
#else
        const int wflags = 0;
#endif
        int st = gchi_shard_write_files_ex(sh, base_.c_str(), no_values_ ? 0 : (uint32_t)sizeof(FinalEdgeDataType),
                                           wflags, nullptr);
        int created = 0;
        gchi_shard_info(sh, &created, nullptr, nullptr);
        gchi_shard_destroy(sh);
        device::check(st, "gchi_shard_write_files");
        m_.stop_time("execute_sharding");
        logstream(LOG_INFO) << "Created " << created << " shards (" << nshards << " requested) for " << e
                            << " edges" << std::endl;
        return nshards;
    }

  private:
    std::string base_;
    metrics m_;
    bool no_values_ = false;
    std::vector<vid_t> src_, dst_;
    vid_t max_id_ = 0;
};

}  // namespace graphchi
