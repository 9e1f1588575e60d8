// The sharder's interface (src/preprocessing/sharder.hpp:332-548):
// start_preprocessing / preprocessing_add_edge / end_preprocessing /
// execute_sharding.  Edges added one by one are collected on the host
// (self-loops dropped, max id tracked, :427-446) and uploaded once; edge lists
// parsed on the device arrive already in HBM (add_device_edges).  Sharding runs
// on the device (gchi_shard_build: interval cut, per-shard (src,dst) order,
// degrees), which also writes the reference's file layout.
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
        dev_src_.release();
        dev_dst_.release();
        dev_e_ = 0;
        max_id_ = 0;
    }
    void preprocessing_add_edge(vid_t from, vid_t to, EdgeDataType = EdgeDataType(), bool = false) {
        if (from == to) return;
        src_.push_back(from);
        dst_.push_back(to);
        max_id_ = std::max(max_id_, std::max(from, to));
    }
    // Edges already in HBM, self-loops dropped (the device edge-list parse).
    void add_device_edges(device::buffer<uint32_t> &&src, device::buffer<uint32_t> &&dst, uint64_t nedges,
                          vid_t max_id) {
        if (!src_.empty() || dev_e_) logstream(LOG_FATAL) << "sharder: edges from two inputs" << std::endl;
        dev_src_ = std::move(src);
        dev_dst_ = std::move(dst);
        dev_e_ = nedges;
        max_id_ = std::max(max_id_, max_id);
    }
    void end_preprocessing() { m_.stop_time("preprocessing"); }

    int execute_sharding(std::string nshards_string) {
        m_.start_time("execute_sharding");
        device::open();
        int nshards = 0;
        if (nshards_string.find("auto") != std::string::npos || nshards_string == "0") {
            device::check(gchi_shard_auto_count(src_.size() + dev_e_, (uint32_t)sizeof(FinalEdgeDataType),
                                                get_option_int("membudget_mb", 1024), &nshards),
                          "gchi_shard_auto_count");
        } else {
            nshards = atoi(nshards_string.c_str());
        }
        if (nshards <= 0) logstream(LOG_FATAL) << "number of shards must be positive" << std::endl;
        const uint64_t e = dev_e_ ? dev_e_ : src_.size();
        device::buffer<uint32_t> s, d;
        if (dev_e_) {
            s = std::move(dev_src_);
            d = std::move(dev_dst_);
        } else {
            s.resize(e);
            d.resize(e);
            if (e) {
                s.upload(src_.data(), e);
                d.upload(dst_.data(), e);
            }
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
    device::buffer<uint32_t> dev_src_, dev_dst_;
    uint64_t dev_e_ = 0;
    vid_t max_id_ = 0;
};

}  // namespace graphchi
