// order_by_degree (src/preprocessing/util/orderbydegree.hpp:58-143): ranks the
// vertices of a sharded graph by (in+out degree, id), writes the translation
// table <base>.vertexmap (vertexmap[old] = new id), relabels every edge and
// shards the result as <base>_degord with the same number of shards.  The
// shards are decoded, ranked and relabelled on the device (gchi_shard_decode_many,
// gchi_degree_order, gchi_relabel_edges), then sharded by the device sharder.
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "api/chifilenames.hpp"
#include "engine/device.hpp"
#include "metrics/metrics.hpp"
#include "shards/shard_reader.hpp"

namespace graphchi {

template <typename EdgeDataType>
int order_by_degree(std::string &base_filename, int nshards, metrics &m) {
    metrics_entry me = m.start_time();
    device::open();
    const size_t n = get_num_vertices(base_filename);
    device::buffer<uint32_t> s, d;
    const uint64_t e = load_shard_edges_device(base_filename, nshards, s, d);
    device::buffer<uint32_t> s2(e), d2(e), vmap(n);
    size_t ws = 0;
    device::check(gchi_degree_order_workspace(e, (uint32_t)n, &ws), "degree order workspace");
    {
        device::workspace w(ws);
        device::check(gchi_degree_order(s.get(), d.get(), e, (uint32_t)n, vmap.get(), w.get(), ws, nullptr),
                      "gchi_degree_order");
    }
    std::vector<uint32_t> h_vmap = vmap.to_host(n);
    FILE *f = fopen((base_filename + ".vertexmap").c_str(), "wb");
    if (!f || (n && fwrite(h_vmap.data(), 4, n, f) != n)) logstream(LOG_FATAL) << "could not write vertexmap" << std::endl;
    fclose(f);
    device::check(gchi_relabel_edges(s.get(), d.get(), e, (uint32_t)n, vmap.get(), s2.get(), d2.get(), nullptr),
                  "gchi_relabel_edges");
    // the highest rank belongs to a vertex with edges whenever there are edges
    const uint32_t max_id = e ? (uint32_t)(n - 1) : 0;
    gchi_shards *sh = nullptr;
    device::check(gchi_shard_build(s2.get(), d2.get(), e, max_id, nshards, &sh, nullptr), "gchi_shard_build");
    std::string out = base_filename + "_degord";
    int st = gchi_shard_write_files(sh, out.c_str(), (uint32_t)sizeof(EdgeDataType), nullptr);
    gchi_shard_destroy(sh);
    device::check(st, "gchi_shard_write_files");
    m.stop_time(me, "order_by_degree");
    return nshards;
}

}  // namespace graphchi
