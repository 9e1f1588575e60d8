// Component-size report for a label file (src/util/labelanalysis.hpp:66-189):
// counts vertices per label, writes "<label>,<size>" lines, largest first, to
// <base>.components and prints the largest components.
#pragma once
#include <algorithm>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "api/vertex_aggregator.hpp"

namespace graphchi {

template <typename LabelType>
void analyze_labels(std::string basefilename, int printtop = 20) {
    std::vector<LabelType> lab = read_vertex_values<LabelType>(basefilename);
    std::map<LabelType, size_t> counts;
    for (auto &l : lab) counts[l]++;
    std::vector<std::pair<size_t, LabelType>> bysize;
    for (auto &kv : counts) bysize.push_back({kv.second, kv.first});
    std::sort(bysize.begin(), bysize.end(), [](const std::pair<size_t, LabelType> &a,
                                               const std::pair<size_t, LabelType> &b) {
        return a.first > b.first || (a.first == b.first && a.second < b.second);
    });
    FILE *f = fopen((basefilename + ".components").c_str(), "w");
    if (f) {
        for (auto &p : bysize) fprintf(f, "%llu,%llu\n", (unsigned long long)p.second, (unsigned long long)p.first);
        fclose(f);
    }
    printf("Number of components: %zu\n", bysize.size());
    for (int i = 0; i < printtop && i < (int)bysize.size(); i++)
        printf("%d. label: %llu, size: %llu\n", i + 1, (unsigned long long)bysize[i].second,
               (unsigned long long)bysize[i].first);
}

}  // namespace graphchi
