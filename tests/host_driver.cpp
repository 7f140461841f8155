// Test driver for the host-side C++ (bfs-gpu-sycl_amd/host/*.hpp): compiled
// with g++ by tests/test_host.py and run on the CPU.
//   meta FILE                      -> meta records as JSON
//   load ROOT NAME P               -> loaded graph summary + concatenated CSC
//                                     written to ROOT/loaded.bin
//   write ROOT NAME P IN.bin       -> writes dataset from IN.bin, prints breaks
//   breaks P COUNTS.bin            -> nnz-split breaks for per-row counts
//   json FILE                      -> parse and re-dump (indent 4)
//   validate IN.bin ROOTV LEV.bin  -> certificate check result
//   args ...                       -> CommandLineParser result
//   stats T1 T2 ...                -> GPURun's timing record of these run
//                                     times (summary table, then the JSON
//                                     record on the last line)
// IN.bin layout: u64 n, u64 nnz, u64 col_ptr[n+1], u32 row_idx[nnz].
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "bfsx_json.hpp"
#include "functions.hpp"
#include "graph_load.hpp"
#include "run_bfs_gpu.hpp"
#include "validate.hpp"

using bfsx::Json;

static void read_csc(const std::string &path, std::vector<uint64_t> &cp, std::vector<uint32_t> &ri) {
  std::ifstream f(path, std::ios::binary);
  uint64_t n = 0, nnz = 0;
  f.read((char *)&n, 8);
  f.read((char *)&nnz, 8);
  cp.resize(n + 1);
  ri.resize(nnz);
  f.read((char *)cp.data(), (std::streamsize)(8 * (n + 1)));
  f.read((char *)ri.data(), (std::streamsize)(4 * nnz));
  if (!f) throw std::runtime_error("short csc file");
}

static void write_csc(const std::string &path, const std::vector<uint64_t> &cp,
                      const std::vector<uint32_t> &ri) {
  std::ofstream f(path, std::ios::binary);
  const uint64_t n = cp.size() - 1, nnz = ri.size();
  f.write((const char *)&n, 8);
  f.write((const char *)&nnz, 8);
  f.write((const char *)cp.data(), (std::streamsize)(8 * cp.size()));
  f.write((const char *)ri.data(), (std::streamsize)(4 * ri.size()));
}

int main(int argc, char **argv) {
  try {
    const std::string cmd = argc > 1 ? argv[1] : "";
    if (cmd == "meta") {
      std::vector<uint32_t> m32;
      std::vector<uint64_t> m64;
      bfsx::readFromMM_meta(argv[2], m32, m64);
      Json out = Json::array();
      for (size_t k = 0; k < m64.size(); k++) {
        Json r = Json::object();
        r["rows"] = m32[3 * k];
        r["cols"] = m32[3 * k + 1];
        r["nnz"] = m64[k];
        r["startRow"] = m32[3 * k + 2];
        out.push_back(r);
      }
      std::cout << out.dump() << std::endl;
    } else if (cmd == "load") {
      const std::string root = argv[2], name = argv[3];
      const uint32_t P = (uint32_t)std::stoul(argv[4]);
      bfsx::CSRGraph g = bfsx::loadMatrix(P, name, root);
      std::vector<uint64_t> cp;
      std::vector<uint32_t> ri;
      if (g.isMultiDimensional) bfsx::concat_partitions(g.indptrMulti, g.indsMulti, cp, ri);
      else if (g.wideOffsets) { cp = g.indptr64; ri = g.inds; }
      else { cp.assign(g.indptr.begin(), g.indptr.end()); ri = g.inds; }
      write_csc(root + "/loaded.bin", cp, ri);
      Json out = Json::object();
      uint64_t rows = 0;
      for (size_t i = 0; i < g.meta32.size(); i += 3) rows += g.meta32[i];
      out["rows"] = rows;
      out["cols"] = g.meta32[1];
      out["multi"] = g.isMultiDimensional;
      out["wide"] = g.wideOffsets;
      out["nnz"] = (uint64_t)ri.size();
      std::cout << out.dump() << std::endl;
    } else if (cmd == "write") {
      std::vector<uint64_t> cp;
      std::vector<uint32_t> ri;
      read_csc(argv[5], cp, ri);
      auto br = bfsx::write_dataset(argv[2], argv[3], (uint32_t)std::stoul(argv[4]), cp, ri);
      std::cout << Json(br).dump() << std::endl;
    } else if (cmd == "breaks") {
      std::ifstream f(argv[3], std::ios::binary | std::ios::ate);
      const size_t n = (size_t)f.tellg() / 8;
      f.seekg(0);
      std::vector<uint64_t> counts(n);
      f.read((char *)counts.data(), (std::streamsize)(8 * n));
      std::cout << Json(bfsx::nnz_split_breaks(counts, (uint32_t)std::stoul(argv[2]))).dump()
                << std::endl;
    } else if (cmd == "json") {
      std::ifstream f(argv[2]);
      Json j;
      f >> j;
      std::cout << j.dump(4);
    } else if (cmd == "validate") {
      std::vector<uint64_t> cp;
      std::vector<uint32_t> ri;
      read_csc(argv[2], cp, ri);
      const uint32_t root = (uint32_t)std::stoul(argv[3]);
      std::ifstream f(argv[4], std::ios::binary);
      std::vector<int> lev(cp.size() - 1);
      f.read((char *)lev.data(), (std::streamsize)(4 * lev.size()));
      const uint32_t n_rows = argc > 5 ? (uint32_t)std::stoul(argv[5]) : (uint32_t)lev.size();
      const std::string why = bfsx::validate_levels(n_rows, cp.data(), ri.data(), root, lev);
      Json out = Json::object();
      out["ok"] = why.empty();
      out["why"] = why;
      out["edges"] = bfsx::traversed_edges(n_rows, cp.data(), lev);
      std::cout << out.dump() << std::endl;
    } else if (cmd == "args") {
      bfsx::CommandLineParser p;
      p.addArgument("num_runs", "1");
      p.addArgument("dataset", "default");
      p.addArgument("root", "1");
      p.addArgument("output", "data.json");
      p.parseArguments(argc - 1, argv + 1);
      Json out = Json::object();
      for (const char *k : {"num_runs", "dataset", "root", "output"}) out[k] = p.getArgument(k);
      std::cout << out.dump() << std::endl;
    } else if (cmd == "stats") {
      std::vector<double> t;
      for (int i = 2; i < argc; i++) t.push_back(std::stod(argv[i]));
      Json out = Json::object();
      bfsx::timing_record(t, out);
      std::cout << "\n" << out.dump() << std::endl;
    } else {
      std::cerr << "unknown command" << std::endl;
      return 2;
    }
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "[ERROR] " << e.what() << std::endl;
    return 1;
  }
}
