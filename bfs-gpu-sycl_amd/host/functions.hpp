// CLI and statistics helpers with the reference's names and behaviour
// (src/functions.hpp:7-163):
//   CommandLineParser   `--name=value` or `--name value`; only registered
//                       names are taken, unknown ones are ignored
//                       (:15-40); printArguments lists them sorted (:51-56)
//   removeAnomalies     keep values within threshold * stddev of the mean
//                       (:83-94)
//   removeAnomaliesnofilter  drop the first 10 % of runs (:96-108)
//   areAllRowsEqualToLast    every run's levels equal the last run's
//                            (:131-163)
#pragma once

#include <cmath>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace bfsx {

class CommandLineParser {
 public:
  void addArgument(const std::string &name, const std::string &defaultValue) {
    args_[name] = defaultValue;
  }

  void parseArguments(int argc, char *argv[]) {
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      if (a.rfind("--", 0) != 0) continue;
      const size_t eq = a.find('=');
      if (eq != std::string::npos) {
        set(a.substr(2, eq - 2), a.substr(eq + 1));
      } else if (i + 1 < argc && argv[i + 1][0] != '-') {
        set(a.substr(2), argv[++i]);
      } else {
        std::cerr << "Warning: No value provided for argument " << a << std::endl;
      }
    }
  }

  std::string getArgument(const std::string &name) const {
    auto it = args_.find(name);
    return it == args_.end() ? std::string() : it->second;
  }

  void printArguments() const {
    for (const auto &kv : args_)
      std::cout << std::setw(20) << std::left << "- " + kv.first << std::setw(20) << kv.second
                << std::endl;
  }

 private:
  void set(const std::string &k, const std::string &v) {
    auto it = args_.find(k);
    if (it != args_.end()) it->second = v;
  }
  std::map<std::string, std::string> args_;
};

inline double calculateMean(const std::vector<double> &data) {
  double s = 0;
  for (double x : data) s += x;
  return data.empty() ? 0.0 : s / data.size();
}

inline double calculateStandardDeviation(const std::vector<double> &data, double mean) {
  double s = 0;
  for (double x : data) s += (x - mean) * (x - mean);
  return data.empty() ? 0.0 : std::sqrt(s / data.size());
}

inline std::vector<double> removeAnomalies(const std::vector<double> &data, double threshold) {
  const double mean = calculateMean(data);
  const double sd = calculateStandardDeviation(data, mean);
  std::vector<double> out;
  for (double x : data)
    if (std::abs(x - mean) <= threshold * sd) out.push_back(x);
  return out;
}

inline std::vector<double> removeAnomaliesnofilter(const std::vector<double> &data) {
  const size_t skip = (size_t)(data.size() * 0.1);
  return std::vector<double>(data.begin() + (ptrdiff_t)skip, data.end());
}

inline std::string formatDouble(double value, int precision = 2) {
  std::ostringstream o;
  o << std::fixed << std::setprecision(precision) << value;
  return o.str();
}

inline void printRow(const std::string &label, const std::string &value) {
  std::cout << "- " << std::left << std::setw(24) << label << " " << std::setw(24) << value
            << "\n";
}

inline bool areAllRowsEqualToLast(const std::vector<std::vector<int>> &runs, int size) {
  if (runs.size() < 2) return true;
  const auto &last = runs.back();
  for (size_t i = 0; i + 1 < runs.size(); ++i) {
    if (runs[i] == last) continue;
    std::cout << "\n distanceGPU[" << i << "] does not match the last distance \n";
    for (size_t j = 0; j < last.size(); ++j)
      if (runs[i][j] != last[j]) {
        std::cout << "Mismatch at index " << j << ": " << runs[i][j] << " vs " << last[j]
                  << std::endl;
        break;
      }
    (void)size;
    return false;
  }
  return true;
}

}  // namespace bfsx
