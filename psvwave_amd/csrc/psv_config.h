// psv_config.h -- configuration, validation and host-side setup arrays for
// the MI355X fdModel.  Pure host C++ (no HIP).
//
// Mirrors the configuration surface of psvWave's fdModel: the INI schema read
// by fdModel::parse_configuration_file (src/fdModel.cpp:470-523), the derived
// sizes and checks of fdModel::parse_parameters (src/fdModel.cpp:390-468) and
// the setup products of fdModel::initialize_arrays (src/fdModel.cpp:193-299).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace psv {

// Raised for the conditions where the reference throws std::invalid_argument
// or std::length_error; the C-ABI maps it to a "value error" status.
struct ValueError : public std::runtime_error {
  explicit ValueError(const std::string &m) : std::runtime_error(m) {}
};

struct Config {
  // [domain]
  int nt = 0, nx_inner = 0, nz_inner = 0, nx_inner_boundary = 0,
      nz_inner_boundary = 0;
  double dx = 0, dz = 0, dt = 0;
  // [boundary]
  int np_boundary = 0;
  double np_factor = 0;
  // [medium]
  double scalar_rho = 0, scalar_vp = 0, scalar_vs = 0;
  // [basis]
  int npx = 1, npz = 1;
  // [sources]
  double peak_frequency = 0, source_timeshift = 0, delay_cycles_per_shot = 0;
  int n_sources = 0, n_shots = 0;
  std::vector<int> ix_sources, iz_sources;  // inner-grid indices as given
  std::vector<double> moment_angles;
  std::vector<std::vector<int>> which_source_to_fire_in_which_shot;
  // [receivers]
  int nr = 0;
  std::vector<int> ix_receivers, iz_receivers;  // inner-grid indices as given
  // [inversion]
  int snapshot_interval = 1;
  // [output]
  std::string observed_data_folder = ".", stf_folder = ".";
};

// Sizes and setup arrays derived from a Config.
struct Derived {
  int nx = 0, nz = 0;
  int nx_free_parameters = 0, nz_free_parameters = 0, free_parameters = 0;
  int snapshots = 0;
  double alpha = 0;
  // source / receiver grid indices with np_boundary added
  // (src/fdModel.cpp:197-212)
  std::vector<int> ix_src, iz_src, ix_rec, iz_rec;
  std::vector<double> stf;     // [n_sources][nt], Ricker (src/fdModel.cpp:214-231)
  std::vector<double> moment;  // [n_sources][2][2] (src/fdModel.cpp:233-246)
  // taper value by nesting count 0..np_boundary (src/fdModel.cpp:263-298)
  std::vector<double> taper_lut;
};

// Reads an INI file the way the reference's inih reader does (keys are
// case-insensitive, ';' after whitespace starts an inline comment, integers
// and reals are parsed as the longest valid prefix).
Config parse_ini_file(const std::string &path);

// parse_string_to_vector (src/fdModel.cpp:1574-1599) and
// parse_string_to_nested_int_vector (src/fdModel.cpp:1601-1659).
std::vector<double> parse_list(const std::string &text);
std::vector<std::vector<int>> parse_nested_int_list(const std::string &text);

// Validates and derives; throws ValueError for inconsistent configurations.
Derived derive(const Config &c);

}  // namespace psv
