// What the reference's src/host.cpp takes from its SYCL headers, supplied by
// the MI355X engine instead.  host.cpp includes this file in place of
// src/run_bfs_gpu.hpp (which pulls in sycl/sycl.hpp, src/functions.hpp and,
// through include/unrolled_loop.hpp, SYCL pipe types), after including the
// reference's own include/json.hpp:
//
//   GPURun                  src/run_bfs_gpu.hpp:534-542 (C ABI underneath)
//   CommandLineParser       src/functions.hpp:7-61
//   areAllRowsEqualToLast   src/functions.hpp:131-163
//   removeAnomalies, removeAnomaliesnofilter, calculateMean,
//   calculateStandardDeviation, printRow, formatDouble
//                           src/functions.hpp:63-129
//   DeviceInfo, printDeviceInfo
//                           src/functions.hpp:165-186 (run_bfs_cpu.hpp takes
//                           a vector of DeviceInfo)
//
// Include it by a path that cannot resolve into the reference's src/
// directory (for example `#include <bfsx_dropin.hpp>` with -I on this
// directory): a quoted "run_bfs_gpu.hpp" in src/host.cpp would find the
// SYCL header next to it first.
#pragma once

#include <algorithm>
#include <cassert>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <string>
#include <vector>

#include "functions.hpp"
#include "run_bfs_gpu.hpp"

using bfsx::areAllRowsEqualToLast;
using bfsx::calculateMean;
using bfsx::calculateStandardDeviation;
using bfsx::CommandLineParser;
using bfsx::formatDouble;
using bfsx::GPURun;
using bfsx::printRow;
using bfsx::removeAnomalies;
using bfsx::removeAnomaliesnofilter;

// Per-device edge statistics the reference's CPU BFS fills in
// (src/run_bfs_cpu.hpp) and host.cpp prints with VERBOSE=1.
struct DeviceInfo {
  int gpuID;
  long long edgesProcessed;
  double percentage;
};

inline void printDeviceInfo(const std::vector<DeviceInfo> &devices) {
  std::cout << std::left << std::setw(10) << "GPU ID" << std::setw(30) << "Edges Processed (*)"
            << std::setw(15) << "Percentage" << std::endl;
  std::cout << std::string(55, '-') << std::endl;
  for (const DeviceInfo &d : devices)
    std::cout << std::left << std::setw(10) << d.gpuID << std::setw(30) << d.edgesProcessed
              << std::setw(15) << std::fixed << std::setprecision(2) << d.percentage << " %"
              << std::endl;
  std::cout << std::string(55, '-') << std::endl;
}
