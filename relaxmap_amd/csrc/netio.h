// Host-side network files: the reference's .net / .txt readers and the
// .tree / .clu writers.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace rmx {

struct NetFile {
  int32_t n = 0;
  std::vector<int64_t> ptr;      // out-CSR after dedup, self-loop and w <= 0 removal
  std::vector<int32_t> idx;
  std::vector<double> w;
  std::vector<double> node_w;
  std::vector<std::string> names;
  int32_t n_self = 0, n_double = 0;
  int64_t n_lines = 0;
};

// 0 on success; -1 malformed (the reference exits), -2 unreadable.
// .net and .txt follow the reference's parsers; .rmxcsr is the binary CSR
// written by write_binary.
int read_network(const std::string &path, NetFile &net, std::string &err);
int write_binary(const std::string &path, const NetFile &net, std::string &err);
void set_parse_threads(unsigned t);   // 0: all hardware threads
int write_tree(const std::string &path, double codelength_nats, const std::vector<uint32_t> &mod,
               const std::vector<double> &p, const std::vector<std::string> &names, std::string &err);
int write_clu(const std::string &path, const std::vector<uint32_t> &mod, std::string &err);

}  // namespace rmx
