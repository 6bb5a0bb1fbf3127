#include "command_file.h"

#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>

namespace bbts {

namespace {

class reader {
 public:
  reader(const uint8_t *d, size_t n) : _d(d), _n(n) {}
  template <class T>
  T val(const char *what) {
    if (_n - _p < sizeof(T)) fail(what);
    T v;
    std::memcpy(&v, _d + _p, sizeof(T));
    _p += sizeof(T);
    return v;
  }
  std::string str(const char *what) {
    uint64_t n = val<uint64_t>(what);
    if (n > _n - _p) fail(what);
    std::string s(reinterpret_cast<const char *>(_d + _p), (size_t)n);
    _p += (size_t)n;
    return s;
  }
  std::vector<std::string> strs(const char *what) {
    uint64_t n = val<uint64_t>(what);
    if (n > (_n - _p) / 8) fail(what);  // every string takes at least its 8-byte length
    std::vector<std::string> out((size_t)n);
    for (auto &s : out) s = str(what);
    return out;
  }
  // std::tuple<tid_t, node_id_t> as libstdc++ stores it: node at +0, tid at +4
  void tids(std::vector<int32_t> &tid, std::vector<int32_t> &node, const char *what) {
    uint64_t n = val<uint64_t>(what);
    if (n > (_n - _p) / 8) fail(what);
    tid.resize((size_t)n);
    node.resize((size_t)n);
    for (size_t i = 0; i < (size_t)n; ++i) {
      node[i] = val<int32_t>(what);
      tid[i] = val<int32_t>(what);
    }
  }
  std::vector<uint32_t> words(const char *what) {
    uint64_t n = val<uint64_t>(what);
    if (n > (_n - _p) / 4) fail(what);
    std::vector<uint32_t> out((size_t)n);
    for (auto &w : out) w = val<uint32_t>(what);
    return out;
  }
  bool done() const { return _p == _n; }
  [[noreturn]] void fail(const char *what) const {
    throw std::runtime_error("command file: truncated or malformed " + std::string(what) + " at byte " +
                             std::to_string(_p));
  }

 private:
  const uint8_t *_d;
  size_t _n, _p = 0;
};

}  // namespace

std::vector<file_command_t> parse_bbts_binary(const uint8_t *data, size_t len) {
  reader r(data, len);
  const uint64_t n = r.val<uint64_t>("command count");
  if (n > len / 53) r.fail("command count");  // the smallest command takes 53 bytes
  std::vector<file_command_t> out((size_t)n);
  for (auto &c : out) {
    c.type = r.val<int32_t>("command type");
    if (c.type < 0 || c.type > 3) r.fail("command type");
    c.ud_name = r.str("ud name");
    c.is_gpu = r.val<uint8_t>("is_gpu flag") != 0;
    c.input_types = r.strs("input types");
    c.output_types = r.strs("output types");
    r.tids(c.inputs, c.input_nodes, "inputs");
    r.tids(c.outputs, c.output_nodes, "outputs");
    c.params = r.words("parameters");
  }
  if (!r.done()) r.fail("trailing bytes");
  return out;
}

std::vector<file_command_t> parse_bbts_text(const char *data, size_t len) {
  std::istringstream in(std::string(data, len));
  auto bad = [&](const std::string &what) {
    throw std::runtime_error("command file: malformed " + what + " near offset " +
                             std::to_string((long long)in.tellg()));
  };
  auto count = [&](const char *what) {
    long long v = -1;
    if (!(in >> v) || v < 0 || v > (long long)len) bad(what);
    return (size_t)v;
  };
  struct spec {
    std::string name;
    std::vector<std::string> in, out;
  };
  std::map<int32_t, spec> specs;
  const size_t nfun = count("function count");
  for (size_t f = 0; f < nfun; ++f) {
    int32_t id;
    spec s;
    if (!(in >> id >> s.name)) bad("function spec");
    s.in.resize(count("input type count"));
    for (auto &t : s.in)
      if (!(in >> t)) bad("input type");
    s.out.resize(count("output type count"));
    for (auto &t : s.out)
      if (!(in >> t)) bad("output type");
    specs[id] = std::move(s);
  }
  const size_t ncmd = count("command count");
  std::vector<file_command_t> out(ncmd);
  for (auto &c : out) {
    int32_t ud;
    std::string type;
    if (!(in >> ud >> type)) bad("command");
    if (type == "APPLY") c.type = 0;
    else if (type == "REDUCE") c.type = 1;
    else if (type == "DELETE") c.type = 3;
    else bad("command type '" + type + "'");
    c.inputs.resize(count("input count"));
    for (auto &t : c.inputs)
      if (!(in >> t)) bad("input tid");
    c.outputs.resize(count("output count"));
    for (auto &t : c.outputs)
      if (!(in >> t)) bad("output tid");
    c.input_nodes.assign(c.inputs.size(), 0);
    c.output_nodes.assign(c.outputs.size(), 0);
    const size_t np = count("parameter count");
    c.params.resize(np);
    for (auto &w : c.params) {
      std::string kind;
      if (!(in >> kind)) bad("parameter");
      if (kind == "int") {
        int32_t v;
        if (!(in >> v)) bad("int parameter");
        std::memcpy(&w, &v, 4);
      } else if (kind == "float") {
        float v;
        if (!(in >> v)) bad("float parameter");
        std::memcpy(&w, &v, 4);
      } else if (kind == "uint") {
        uint32_t v;
        if (!(in >> v)) bad("uint parameter");
        w = v;
      } else {
        bad("parameter kind '" + kind + "'");
      }
    }
    if (c.type != 3) {
      auto it = specs.find(ud);
      if (it == specs.end()) bad("function id " + std::to_string(ud));
      c.ud_name = it->second.name;
      c.input_types = it->second.in;
      c.output_types = it->second.out;
      c.gpu_unspecified = true;
    }
  }
  std::string rest;
  if (in >> rest) bad("trailing token '" + rest + "'");
  return out;
}

}  // namespace bbts
