// Command-line driver with the reference's flags and pipeline (main.cpp:41-97,
// flags.cpp:3-101): Input, RB_Graph, Projection, Init_entry, Search; appends
// the reference's out.txt lines (run header, n_aabbs, time search, recall@10).
//
//   rte_search --data_path base.fvecs --query_path q.fvecs --gt_path gt.ivecs
//              --graph_path g.nsw --data_name cache_dir [--n_candidates 128]
//              [--topk 0] [--max_iter 100] [--search_width 1] [--ALGO 1]
//              [--point_ratio 0.000128] [--expand_ratio 0.2] [--dim 0]
//              [--mode 0] [--device 0] [--ids_out ids.ivecs] [--reorder 0]
//   rte_search --check_file data.{txt,fvecs,fbin,hvecs,bvecs,hh}
//              (reads one data file as the pipeline would and prints
//               "n d sum", no GPU work)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "graph.h"

int main(int argc, char** argv) {
  std::map<std::string, std::string> f = {
      {"data_name", "sift1M"},   {"data_path", ""},        {"query_path", ""},
      {"gt_path", ""},           {"graph_path", ""},       {"n_candidates", "128"},
      {"topk", "0"},             {"max_iter", "100"},      {"search_width", "1"},
      {"ALGO", "1"},             {"point_ratio", "0.000128"}, {"expand_ratio", "0.2"},
      {"max_hits", "1"},         {"n_subspaces", "1"},     {"buffer_size", "200000"},
      {"dim", "0"},              {"mode", "0"},            {"device", "0"},
      {"ids_out", ""},           {"out_file", "out.txt"},  {"reorder", "0"},
      {"check_file", ""}};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) {
      fprintf(stderr, "unexpected argument %s\n", argv[i]);
      return 2;
    }
    a = a.substr(2);
    std::string v;
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      v = a.substr(eq + 1);
      a = a.substr(0, eq);
    } else if (i + 1 < argc) {
      v = argv[++i];
    }
    if (!f.count(a)) {
      fprintf(stderr, "unknown flag --%s\n", a.c_str());
      return 2;
    }
    f[a] = v;
  }
  try {
    if (!f["check_file"].empty()) {
      int n = 0, d = 0;
      std::vector<float> v;
      rte_host::read_data(f["check_file"], n, d, v);
      double sum = 0.0;
      for (float x : v) sum += x;
      printf("%d %d %.17g\n", n, d, sum);
      return 0;
    }
    {
      std::ofstream out(f["out_file"], std::ios::app);  // main.cpp:60-73
      out << "\n---------- n_candidates = " << f["n_candidates"] << "\tpoint_ratio = "
          << f["point_ratio"] << "\tmax_iter = " << f["max_iter"] << "\tgraph_path = "
          << f["graph_path"] << "\tDIM = " << f["dim"] << "\tREORDER = " << f["reorder"]
          << " ----------\n\n";
    }
    rte_host::Graph g(atoi(f["n_subspaces"].c_str()), atoi(f["buffer_size"].c_str()),
                      atoi(f["n_candidates"].c_str()), atoi(f["max_hits"].c_str()),
                      atof(f["expand_ratio"].c_str()), atof(f["point_ratio"].c_str()),
                      f["data_name"], f["data_path"], f["query_path"], f["gt_path"],
                      f["graph_path"], atoi(f["ALGO"].c_str()), atoi(f["search_width"].c_str()),
                      atoi(f["topk"].c_str()), atoi(f["max_iter"].c_str()),
                      atoi(f["device"].c_str()), atoi(f["mode"].c_str()));
    g.set_out_file(f["out_file"]);
    if (atoi(f["dim"].c_str()) > 0) g.set_search_dim(atoi(f["dim"].c_str()));
    g.set_reorder(atoi(f["reorder"].c_str()) != 0);
    g.Input();
    g.RB_Graph();
    g.Projection();
    g.Init_entry();
    g.Search();
    if (!f["ids_out"].empty()) {  // ivecs rows of the topk ids
      std::ofstream o(f["ids_out"], std::ios::binary);
      const int k = g.topk();
      for (int q = 0; q < g.nq(); ++q) {
        o.write((const char*)&k, 4);
        o.write((const char*)&g.results()[(size_t)q * k], 4LL * k);
      }
    }
    g.CleanUp();
  } catch (const std::exception& e) {
    fprintf(stderr, "rte_search: %s\n", e.what());
    return 1;
  }
  return 0;
}
