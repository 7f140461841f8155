// Binary CSC dataset loader with the reference's on-disk layout and entry
// points (src/graph_load.hpp:7-131):
//
//   $PWD/dataset/<name>-csc-<P>/<name>-csc-meta.bin        P records of
//       <u32 rows, u32 cols, u64 nnz, u32 startRow> (20 bytes, packed;
//       written by scripts/generator.py:182-185)
//   $PWD/dataset/<name>-csc-<P>/<name>-csc-<k>-indptr.bin  cols+1 offsets
//   $PWD/dataset/<name>-csc-<P>/<name>-csc-<k>-inds.bin    nnz destinations
//
// Offsets are uint32, or uint64 when the writer kept scipy's int64 array
// (generator.py:125-127 narrows only when nnz fits); the width is inferred
// from the file size.  Unlike the reference, every inconsistency (missing,
// empty, truncated or mis-sized file) raises std::runtime_error with the
// file name instead of reading past the data.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace bfsx {

struct CSRGraph {
  std::vector<uint32_t> meta32;  // rows, cols, startRow per partition
  std::vector<uint64_t> meta64;  // nnz per partition
  std::vector<uint32_t> indptr;  // one partition, 4-byte offsets
  std::vector<uint32_t> inds;
  std::vector<std::vector<uint32_t>> indptrMulti;  // P > 1
  std::vector<std::vector<uint32_t>> indsMulti;
  bool isMultiDimensional = false;
  std::vector<uint64_t> indptr64;  // one partition stored with 8-byte offsets
  bool wideOffsets = false;
};

inline uint64_t file_size_or_throw(const std::string &path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("File not found: " + path);
  const std::streamoff n = f.tellg();
  if (n <= 0) throw std::runtime_error("File size is 0: " + path);
  return (uint64_t)n;
}

// Appends the whole file as raw little-endian words of type T.
template <class T>
inline void read_words(const std::string &path, std::vector<T> &buffer) {
  const uint64_t bytes = file_size_or_throw(path);
  if (bytes % sizeof(T)) throw std::runtime_error("File size is not a multiple of the word size: " + path);
  std::ifstream f(path, std::ios::binary);
  const size_t old = buffer.size();
  buffer.resize(old + bytes / sizeof(T));
  if (!f.read(reinterpret_cast<char *>(buffer.data() + old), (std::streamsize)bytes))
    throw std::runtime_error("Error reading file " + path);
}

// src/graph_load.hpp:23-48 (raw uint32 words)
inline void readFromMM(const char *filename, std::vector<uint32_t> &buffer) {
  read_words<uint32_t>(filename, buffer);
}

// src/graph_load.hpp:49-92 (20-byte meta records)
inline void readFromMM_meta(const char *filename, std::vector<uint32_t> &buffer32,
                            std::vector<uint64_t> &buffer64) {
  const std::string path(filename);
  const uint64_t bytes = file_size_or_throw(path);
  if (bytes % 20) throw std::runtime_error("Meta file holds a partial record: " + path);
  std::ifstream f(path, std::ios::binary);
  for (uint64_t r = 0; r < bytes / 20; r++) {
    unsigned char rec[20];
    if (!f.read(reinterpret_cast<char *>(rec), 20)) throw std::runtime_error("Error reading file " + path);
    uint32_t rows, cols, start;
    uint64_t nnz;
    std::memcpy(&rows, rec, 4);
    std::memcpy(&cols, rec + 4, 4);
    std::memcpy(&nnz, rec + 8, 8);
    std::memcpy(&start, rec + 16, 4);
    buffer32.push_back(rows);
    buffer32.push_back(cols);
    buffer64.push_back(nnz);
    buffer32.push_back(start);
  }
}

inline std::string dataset_prefix(uint32_t partitionCount, const std::string &datasetName,
                                  const std::string &root) {
  return root + "/dataset/" + datasetName + "-csc-" + std::to_string(partitionCount) + "/" +
         datasetName + "-csc-";
}

// src/graph_load.hpp:93-131; `root` defaults to $PWD as in the reference.
inline CSRGraph loadMatrix(uint32_t partitionCount, std::string datasetName,
                           std::string root = "") {
  if (partitionCount == 0) throw std::runtime_error("partition count must be >= 1");
  if (root.empty()) {
    const char *pwd = std::getenv("PWD");
    root = pwd ? pwd : ".";
  }
  CSRGraph g;
  const std::string pre = dataset_prefix(partitionCount, datasetName, root);
  readFromMM_meta((pre + "meta.bin").c_str(), g.meta32, g.meta64);
  if (g.meta64.size() != partitionCount)
    throw std::runtime_error("Meta file holds " + std::to_string(g.meta64.size()) +
                             " records for " + std::to_string(partitionCount) + " partitions");
  const uint64_t cols = g.meta32[1];
  for (uint32_t k = 0; k < partitionCount; k++) {
    if (g.meta32[3 * k + 1] != cols) throw std::runtime_error("partitions disagree on the column count");
    const std::string ip = pre + std::to_string(k) + "-indptr.bin";
    const std::string ix = pre + std::to_string(k) + "-inds.bin";
    const uint64_t ib = file_size_or_throw(ip);
    const uint64_t nnz = g.meta64[k];
    const bool wide = ib == (cols + 1) * 8 && ib != (cols + 1) * 4;
    if (ib != (cols + 1) * 4 && !wide)
      throw std::runtime_error("indptr size does not match cols+1 offsets: " + ip);
    if (nnz > 0 && file_size_or_throw(ix) != nnz * 4)
      throw std::runtime_error("inds size does not match nnz: " + ix);
    if (partitionCount == 1) {
      if (wide) {
        read_words<uint64_t>(ip, g.indptr64);
        g.wideOffsets = true;
      } else {
        read_words<uint32_t>(ip, g.indptr);
      }
      if (nnz > 0) read_words<uint32_t>(ix, g.inds);
    } else {
      if (wide) throw std::runtime_error("8-byte offsets are supported for one partition only");
      g.isMultiDimensional = true;
      g.indptrMulti.emplace_back();
      g.indsMulti.emplace_back();
      read_words<uint32_t>(ip, g.indptrMulti.back());
      if (nnz > 0) read_words<uint32_t>(ix, g.indsMulti.back());
    }
  }
  return g;
}

// One CSC holding every partition: partition k stores all columns but only
// its destination range, with global row indices (generator.py:135), so the
// lists of column u are concatenated in partition order.  64-bit offsets.
inline void concat_partitions(const std::vector<std::vector<uint32_t>> &indptrs,
                              const std::vector<std::vector<uint32_t>> &indss,
                              std::vector<uint64_t> &col_ptr, std::vector<uint32_t> &row_idx) {
  const size_t P = indptrs.size();
  if (P == 0) throw std::runtime_error("no partitions");
  const size_t n1 = indptrs[0].size();
  for (const auto &ip : indptrs)
    if (ip.size() != n1) throw std::runtime_error("partitions disagree on the column count");
  col_ptr.assign(n1, 0);
  uint64_t total = 0;
  for (size_t u = 0; u + 1 < n1; u++) {
    col_ptr[u] = total;
    for (size_t k = 0; k < P; k++) total += indptrs[k][u + 1] - indptrs[k][u];
  }
  col_ptr[n1 - 1] = total;
  row_idx.resize(total);
  uint64_t o = 0;
  for (size_t u = 0; u + 1 < n1; u++)
    for (size_t k = 0; k < P; k++)
      for (uint64_t e = indptrs[k][u]; e < indptrs[k][u + 1]; e++) {
        if (e >= indss[k].size()) throw std::runtime_error("indptr points past inds");
        row_idx[o++] = indss[k][e];
      }
}

// ---------------------------------------------------------------------------
// writer: the layout of scripts/generator.py:115-195
// ---------------------------------------------------------------------------
// round8 of generator.py:11-12: `int(a) + 4 & ~7` (nearest multiple of 8).
inline uint64_t round8(uint64_t a) { return (a + 4) & ~7ull; }

// Destination-range breaks of nnzSplit (generator.py:22-40): cumulative
// entries per destination row, ideal breaks at k * total / P (numpy arange
// semantics), first row reaching each break, then round8, last = V.
inline std::vector<uint64_t> nnz_split_breaks(const std::vector<uint64_t> &row_counts,
                                              uint32_t P) {
  const uint64_t V = row_counts.size();
  std::vector<uint64_t> cum(V);
  uint64_t t = 0;
  for (uint64_t i = 0; i < V; i++) cum[i] = (t += row_counts[i]);
  const double total = (double)t, step = total / P;
  const uint64_t n_breaks = t ? (uint64_t)std::ceil(total / step) : 1;
  std::vector<uint64_t> br;
  for (uint64_t k = 0; k < n_breaks; k++) {
    const double target = (double)k * step;
    uint64_t lo = 0, hi = V;  // first i with cum[i] >= target
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if ((double)cum[mid] >= target) hi = mid;
      else lo = mid + 1;
    }
    br.push_back(round8(lo));
  }
  br.push_back(round8(V));
  return br;
}

template <class W>
inline void write_words(const std::string &path, const W *data, size_t n) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) throw std::runtime_error("cannot write " + path);
  f.write(reinterpret_cast<const char *>(data), (std::streamsize)(n * sizeof(W)));
  if (!f) throw std::runtime_error("error writing " + path);
}

// Writes a CSC graph (col_ptr over sources, sorted destinations) as dataset
// `name` with P destination-range partitions split by nnz; returns the
// partition start rows (P + 1 entries).  Offsets are written as uint32 when
// the partition's nnz fits, else uint64 (generator.py:125-127).
inline std::vector<uint64_t> write_dataset(const std::string &root, const std::string &name,
                                           uint32_t P, const std::vector<uint64_t> &col_ptr,
                                           const std::vector<uint32_t> &row_idx) {
  const uint64_t V = col_ptr.size() - 1;
  std::vector<uint64_t> row_counts(V, 0);
  for (uint32_t v : row_idx) row_counts[v]++;
  std::vector<uint64_t> br = nnz_split_breaks(row_counts, P);
  const std::string dir = root + "/dataset/" + name + "-csc-" + std::to_string(P);
  std::filesystem::create_directories(dir);
  const std::string pre = dir + "/" + name + "-csc-";
  std::ofstream meta(pre + "meta.bin", std::ios::binary | std::ios::trunc);
  if (!meta) throw std::runtime_error("cannot write " + pre + "meta.bin");
  for (size_t k = 0; k + 1 < br.size(); k++) {
    const uint64_t lo = std::min<uint64_t>(br[k], V), hi = std::min<uint64_t>(br[k + 1], V);
    std::vector<uint64_t> ip(V + 1, 0);
    std::vector<uint32_t> ix;
    for (uint64_t u = 0; u < V; u++) {
      for (uint64_t e = col_ptr[u]; e < col_ptr[u + 1]; e++)
        if (row_idx[e] >= lo && row_idx[e] < hi) ix.push_back(row_idx[e]);
      ip[u + 1] = ix.size();
    }
    const std::string ks = std::to_string(k);
    if (ix.size() <= 0xFFFFFFFFull) {
      std::vector<uint32_t> ip32(ip.begin(), ip.end());
      write_words(pre + ks + "-indptr.bin", ip32.data(), ip32.size());
    } else {
      write_words(pre + ks + "-indptr.bin", ip.data(), ip.size());
    }
    write_words(pre + ks + "-inds.bin", ix.data(), ix.size());
    const uint32_t rows = (uint32_t)(hi - lo), cols = (uint32_t)V, start = (uint32_t)lo;
    const uint64_t nnz = ix.size();
    meta.write(reinterpret_cast<const char *>(&rows), 4);
    meta.write(reinterpret_cast<const char *>(&cols), 4);
    meta.write(reinterpret_cast<const char *>(&nnz), 8);
    meta.write(reinterpret_cast<const char *>(&start), 4);
  }
  return br;
}

}  // namespace bfsx
