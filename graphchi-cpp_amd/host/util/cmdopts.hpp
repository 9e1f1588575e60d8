// Command-line and configuration options (src/util/cmdopts.hpp:112-249,
// src/util/configfile.hpp:83-119): "key value" pairs or --key=value on the
// command line override conf/graphchi.local.cnf or conf/graphchi.cnf under
// $GRAPHCHI_ROOT (or the working directory).  A missing file is not an error:
// the reference's shipped settings are the defaults.
#pragma once
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>

#include "logger/logger.hpp"

namespace graphchi {

inline std::map<std::string, std::string> &option_table() {
    static std::map<std::string, std::string> t;
    return t;
}

inline std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

inline void load_config_file(const std::string &path) {
    std::ifstream in(path.c_str());
    std::string line;
    while (std::getline(in, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
        if (!k.empty() && !option_table().count(k)) option_table()[k] = v;
    }
}

inline void graphchi_init(int argc, const char **argv) {
    auto &t = option_table();
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a.rfind("--", 0) == 0) {
            size_t eq = a.find('=');
            if (eq != std::string::npos) t[a.substr(2, eq - 2)] = a.substr(eq + 1);
            else if (i + 1 < argc) t[a.substr(2)] = argv[++i];
        } else if (i + 1 < argc) {
            t[a] = argv[++i];
        }
    }
    const char *root = getenv("GRAPHCHI_ROOT");
    std::string dir = root ? std::string(root) : std::string(".");
    load_config_file(dir + "/conf/graphchi.local.cnf");
    load_config_file(dir + "/conf/graphchi.cnf");
    // Without a configuration file the values the reference ships in
    // conf/graphchi.cnf apply (it refuses to start without that file,
    // src/util/configfile.hpp:83-92): membudget_mb sets the automatic shard
    // count (sharder.hpp:565-575), so it must be 800 here too.
    static const char *const defaults[][2] = {{"loadthreads", "4"},      {"niothreads", "2"},
                                              {"membudget_mb", "800"},   {"cachesize_mb", "0"},
                                              {"io.blocksize", "1048576"}, {"mmap", "0"}};
    for (const auto &kv : defaults)
        if (!t.count(kv[0])) t[kv[0]] = kv[1];
}

inline void set_conf(std::string key, std::string value) { option_table()[key] = value; }

inline std::string get_option_string(const char *name, std::string def) {
    auto it = option_table().find(name);
    return it == option_table().end() ? def : it->second;
}
inline std::string get_option_string(const char *name) {
    auto it = option_table().find(name);
    if (it == option_table().end()) logstream(LOG_FATAL) << "missing required option: " << name << std::endl;
    return it->second;
}
inline std::string get_option_string_interactive(const char *name, std::string /*options*/) {
    return get_option_string(name);
}
inline int get_option_int(const char *name, int def) {
    auto it = option_table().find(name);
    return it == option_table().end() ? def : atoi(it->second.c_str());
}
inline int get_option_int(const char *name) { return atoi(get_option_string(name).c_str()); }
inline long get_option_long(const char *name, long def) {
    auto it = option_table().find(name);
    return it == option_table().end() ? def : atol(it->second.c_str());
}
inline float get_option_float(const char *name, float def) {
    auto it = option_table().find(name);
    return it == option_table().end() ? def : (float)atof(it->second.c_str());
}

}  // namespace graphchi
