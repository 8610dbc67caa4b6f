// psv_config.cpp -- INI reading, validation and host-side setup arrays.
// Pure host C++; compiled with -ffp-contract=off so that every expression is
// rounded operation by operation as in the reference build.
#include "psv_config.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

namespace psv {

namespace {

// Moment-tensor angle conversion uses this truncated constant
// (src/fdModel.cpp:14, used at :238-245).
constexpr double kTruncatedPi = 3.14159265;

std::string lower(std::string s) {
  for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

std::string rstrip(const std::string &s) {
  size_t e = s.size();
  while (e > 0 && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(0, e);
}

std::string lstrip(const std::string &s) {
  size_t b = 0;
  while (b < s.size() && std::isspace((unsigned char)s[b])) ++b;
  return s.substr(b);
}

// Position of the first character of `s` (from `from`) that is in `chars`,
// or of a ';' that directly follows whitespace (an inline comment).
size_t find_chars_or_comment(const std::string &s, size_t from,
                             const char *chars) {
  bool was_space = false;
  size_t i = from;
  for (; i < s.size(); ++i) {
    char ch = s[i];
    if (chars && std::strchr(chars, ch)) break;
    if (was_space && ch == ';') break;
    was_space = std::isspace((unsigned char)ch) != 0;
  }
  return i;
}

// INI contents keyed by "section=name" (lower case).  Repeated keys and
// continuation lines are joined with '\n', as the inih reader does.
std::map<std::string, std::string> read_ini(const std::string &path) {
  std::ifstream in(path);
  if (!in.good()) throw ValueError("Can't load .ini file.");
  std::map<std::string, std::string> values;
  std::string line, section, prev_name;
  bool first = true;
  while (std::getline(in, line)) {
    if (first && line.size() >= 3 && (unsigned char)line[0] == 0xEF &&
        (unsigned char)line[1] == 0xBB && (unsigned char)line[2] == 0xBF)
      line = line.substr(3);
    first = false;
    std::string r = rstrip(line);
    std::string start = lstrip(r);
    bool indented = start.size() < r.size();
    if (start.empty()) continue;
    if (start[0] == ';' || start[0] == '#') continue;
    auto store = [&](const std::string &name, const std::string &value) {
      std::string key = lower(section) + "=" + lower(name);
      auto it = values.find(key);
      if (it != values.end() && !it->second.empty())
        it->second += "\n" + value;
      else
        values[key] = value;
    };
    if (!prev_name.empty() && indented) {
      // continuation of the previous value
      size_t e = find_chars_or_comment(start, 0, nullptr);
      store(prev_name, rstrip(start.substr(0, e)));
      continue;
    }
    if (start[0] == '[') {
      size_t e = find_chars_or_comment(start, 1, "]");
      if (e < start.size() && start[e] == ']') {
        section = start.substr(1, e - 1);
        prev_name.clear();
      }
      continue;  // a malformed section line is skipped, as inih does
    }
    size_t e = find_chars_or_comment(start, 0, "=:");
    if (e < start.size() && (start[e] == '=' || start[e] == ':')) {
      std::string name = rstrip(start.substr(0, e));
      std::string rest = start.substr(e + 1);
      size_t c = find_chars_or_comment(rest, 0, nullptr);
      std::string value = rstrip(lstrip(rest.substr(0, c)));
      store(name, value);
      prev_name = name;
    }
  }
  return values;
}

struct Ini {
  std::map<std::string, std::string> v;
  const std::string &get(const char *sec, const char *name) const {
    auto it = v.find(lower(std::string(sec)) + "=" + lower(std::string(name)));
    if (it == v.end())
      throw ValueError(std::string("Missing configuration value [") + sec +
                       "] " + name);
    return it->second;
  }
  // strtol with base 0: the longest valid prefix ("8000;   int" -> 8000).
  int get_int(const char *sec, const char *name) const {
    const std::string &s = get(sec, name);
    char *end = nullptr;
    long n = std::strtol(s.c_str(), &end, 0);
    if (end == s.c_str())
      throw ValueError(std::string("Not an integer: [") + sec + "] " + name);
    return (int)n;
  }
  double get_real(const char *sec, const char *name) const {
    const std::string &s = get(sec, name);
    char *end = nullptr;
    double d = std::strtod(s.c_str(), &end);
    if (end == s.c_str())
      throw ValueError(std::string("Not a number: [") + sec + "] " + name);
    return d;
  }
};

std::string remove_spaces(const std::string &s) {
  std::string out;
  out.reserve(s.size());
  for (char ch : s)
    if (!std::isspace((unsigned char)ch)) out.push_back(ch);
  return out;
}

// Comma-separated tokens, each converted with atof semantics.
std::vector<double> split_atof(const std::string &s) {
  std::vector<double> out;
  size_t begin = 0;
  while (true) {
    size_t comma = s.find(',', begin);
    std::string tok = s.substr(begin, comma == std::string::npos
                                          ? std::string::npos
                                          : comma - begin);
    out.push_back(std::atof(tok.c_str()));
    if (comma == std::string::npos) break;
    begin = comma + 1;
  }
  return out;
}

std::vector<int> to_ints(const std::vector<double> &d) {
  std::vector<int> out;
  out.reserve(d.size());
  for (double x : d) out.push_back((int)x);
  return out;
}

}  // namespace

// "{a, b, c} ; comment": drop whitespace, keep what precedes the first '}',
// drop the first character, split on ','.
std::vector<double> parse_list(const std::string &text) {
  std::string s = remove_spaces(text);
  size_t close = s.find('}');
  if (close == std::string::npos)
    throw ValueError("List value has no closing '}': " + text);
  s = s.substr(0, close);
  s = s.empty() ? s : s.substr(1);
  return split_atof(s);
}

// "{{a, b}, {c}};": drop whitespace and the two leading characters, split
// the groups on "},{", end the last group at "}};" (or at the end).
std::vector<std::vector<int>> parse_nested_int_list(const std::string &text) {
  std::string s = remove_spaces(text);
  s = s.size() >= 2 ? s.substr(2) : std::string();
  std::vector<std::vector<int>> out;
  size_t pos;
  while ((pos = s.find("},{")) != std::string::npos) {
    out.push_back(to_ints(split_atof(s.substr(0, pos))));
    s = s.substr(pos + 3);
  }
  pos = s.find("}};");
  out.push_back(to_ints(split_atof(s.substr(0, pos))));
  return out;
}

Config parse_ini_file(const std::string &path) {
  Ini r{read_ini(path)};
  Config c;
  c.nt = r.get_int("domain", "nt");
  c.nx_inner = r.get_int("domain", "nx_inner");
  c.nz_inner = r.get_int("domain", "nz_inner");
  c.nx_inner_boundary = r.get_int("domain", "nx_inner_boundary");
  c.nz_inner_boundary = r.get_int("domain", "nz_inner_boundary");
  c.dx = r.get_real("domain", "dx");
  c.dz = r.get_real("domain", "dz");
  c.dt = r.get_real("domain", "dt");
  c.np_boundary = r.get_int("boundary", "np_boundary");
  c.np_factor = r.get_real("boundary", "np_factor");
  c.scalar_rho = r.get_real("medium", "scalar_rho");
  c.scalar_vp = r.get_real("medium", "scalar_vp");
  c.scalar_vs = r.get_real("medium", "scalar_vs");
  c.npx = r.get_int("basis", "npx");
  c.npz = r.get_int("basis", "npz");
  c.peak_frequency = r.get_real("sources", "peak_frequency");
  c.source_timeshift = r.get_real("sources", "source_timeshift");
  c.delay_cycles_per_shot = r.get_real("sources", "delay_cycles_per_shot");
  c.n_sources = r.get_int("sources", "n_sources");
  c.n_shots = r.get_int("sources", "n_shots");
  c.ix_sources = to_ints(parse_list(r.get("sources", "ix_sources")));
  c.iz_sources = to_ints(parse_list(r.get("sources", "iz_sources")));
  c.moment_angles = parse_list(r.get("sources", "moment_angles"));
  c.which_source_to_fire_in_which_shot =
      parse_nested_int_list(r.get("sources", "which_source_to_fire_in_which_shot"));
  c.nr = r.get_int("receivers", "nr");
  c.ix_receivers = to_ints(parse_list(r.get("receivers", "ix_receivers")));
  c.iz_receivers = to_ints(parse_list(r.get("receivers", "iz_receivers")));
  c.snapshot_interval = r.get_int("inversion", "snapshot_interval");
  c.observed_data_folder = r.get("output", "observed_data_folder");
  c.stf_folder = r.get("output", "stf_folder");
  return c;
}

Derived derive(const Config &c) {
  Derived d;
  if (c.nt < 1 || c.nx_inner < 1 || c.nz_inner < 1 || c.np_boundary < 0)
    throw ValueError("nt, nx_inner and nz_inner must be positive and "
                     "np_boundary non-negative");
  if (c.n_sources < 0 || c.n_shots < 0 || c.nr < 0)
    throw ValueError("n_sources, n_shots and nr must be non-negative");
  // src/fdModel.cpp:398-407
  d.nx = c.nx_inner + c.np_boundary * 2;
  d.nz = c.nz_inner + c.np_boundary * 2;
  d.nx_free_parameters = c.nx_inner - c.nx_inner_boundary * 2;
  d.nz_free_parameters = c.nz_inner - c.nz_inner_boundary * 2;
  if (c.npx < 1 || c.npz < 1)
    throw ValueError("basis npx and npz must be positive");
  if (d.nx_free_parameters < 1 || d.nz_free_parameters < 1)
    throw ValueError("inner boundary leaves no free parameters");
  if (d.nx_free_parameters % c.npx != 0 || d.nz_free_parameters % c.npz != 0)
    throw ValueError("free-parameter region is not divisible by the basis "
                     "(npx, npz)");
  d.free_parameters = 3 * d.nx_free_parameters * d.nz_free_parameters /
                      (c.npx * c.npz);
  // src/fdModel.cpp:414-457
  if ((int)c.ix_sources.size() != c.n_sources ||
      (int)c.iz_sources.size() != c.n_sources ||
      (int)c.moment_angles.size() != c.n_sources)
    throw ValueError(
        "Dimension mismatch between n_sources and sources.ix_sources, "
        "sources.iz_sources or sources.moment_angles");
  if ((int)c.which_source_to_fire_in_which_shot.size() != c.n_shots)
    throw ValueError(
        "Mismatch between n_shots and sources.which_source_to_fire_in_which_shot");
  int total = 0;
  for (const auto &shot : c.which_source_to_fire_in_which_shot)
    total += (int)shot.size();
  if (total != c.n_sources)
    throw ValueError(
        "Mismatch between n_sources and "
        "sources.which_source_to_fire_in_which_shot.");
  if ((int)c.ix_receivers.size() != c.nr || (int)c.iz_receivers.size() != c.nr)
    throw ValueError(
        "Mismatch between nr and receivers.ix_receivers or receivers.iz_receivers");
  // src/fdModel.cpp:460-467
  if (c.snapshot_interval < 1)
    throw ValueError("snapshot_interval must be positive");
  d.alpha = 1.0 / c.peak_frequency;
  d.snapshots = c.nt / c.snapshot_interval;

  // src/fdModel.cpp:197-212
  for (int s = 0; s < c.n_sources; ++s) {
    d.ix_src.push_back(c.ix_sources[s] + c.np_boundary);
    d.iz_src.push_back(c.iz_sources[s] + c.np_boundary);
  }
  for (int r = 0; r < c.nr; ++r) {
    d.ix_rec.push_back(c.ix_receivers[r] + c.np_boundary);
    d.iz_rec.push_back(c.iz_receivers[r] + c.np_boundary);
  }
  // Guards the reference leaves undefined (out-of-bounds writes).
  for (const auto &shot : c.which_source_to_fire_in_which_shot)
    for (int s : shot) {
      if (s < 0 || s >= c.n_sources)
        throw ValueError("which_source_to_fire_in_which_shot names a source "
                         "index outside [0, n_sources)");
      if (d.ix_src[s] < 1 || d.ix_src[s] > d.nx - 2 || d.iz_src[s] < 1 ||
          d.iz_src[s] > d.nz - 2)
        throw ValueError("source " + std::to_string(s) +
                         " injects outside the grid");
    }
  for (int r = 0; r < c.nr; ++r)
    if (d.ix_rec[r] < 0 || d.ix_rec[r] >= d.nx || d.iz_rec[r] < 0 ||
        d.iz_rec[r] >= d.nz)
      throw ValueError("receiver " + std::to_string(r) + " lies outside the grid");

  // Ricker wavelet per source; the delay counts the source's position in
  // its shot list (src/fdModel.cpp:214-231).
  d.stf.assign((size_t)c.n_sources * c.nt, 0.0);
  for (int i_shot = 0; i_shot < c.n_shots; ++i_shot) {
    const auto &shot = c.which_source_to_fire_in_which_shot[i_shot];
    for (int i_source = 0; i_source < (int)shot.size(); ++i_source) {
      for (unsigned int it = 0; it < (unsigned int)c.nt; ++it) {
        double t = it * c.dt;
        double f = 1.0 / d.alpha;
        double shifted = t - 1.4 / f - c.delay_cycles_per_shot * i_source / f;
        double a = M_PI * f * shifted;
        d.stf[(size_t)shot[i_source] * c.nt + it] =
            (1 - 2 * (a * a)) * std::exp(-(a * a));
      }
    }
  }
  // src/fdModel.cpp:233-246
  d.moment.assign((size_t)c.n_sources * 4, 0.0);
  for (int s = 0; s < c.n_sources; ++s) {
    double arg = c.moment_angles[s] * kTruncatedPi / 180.0;
    d.moment[s * 4 + 0] = std::cos(arg) * 1e15;
    d.moment[s * 4 + 1] = -std::sin(arg) * 1e15;
    d.moment[s * 4 + 2] = -std::sin(arg) * 1e15;
    d.moment[s * 4 + 3] = -std::cos(arg) * 1e15;
  }
  // A point's taper depends only on how many of the np_boundary nested
  // rectangles contain it (src/fdModel.cpp:263-298).
  d.taper_lut.resize(c.np_boundary + 1);
  for (int count = 0; count <= c.np_boundary; ++count) {
    double x = c.np_factor * (c.np_boundary - (double)count);
    d.taper_lut[count] = std::exp(-(x * x));
  }
  return d;
}

}  // namespace psv
