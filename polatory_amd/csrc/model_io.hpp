// polatory's binary model and interpolant files.
//
// The format is the reference's native-endian stream of its fields
// (common/io.hpp:15-186): trivially copyable values as their bytes; strings and
// vectors as a size_t count followed by the elements; Eigen matrices as their
// dynamic extents (Eigen::Index) followed by the coefficients in storage
// order (polatory's Mat<M, N> is row-major unless N == 1).
//
//   Rbf          short name (string), parameters (vector<double>),
//                anisotropy (Mat3, 9 doubles row-major)        rbf/rbf_io.hpp:9-35
//   Model        rbfs (vector<Rbf>), poly_degree (int), nugget  model.hpp:304-320
//   Bbox         min (1x3), max (1x3)                           geometry/bbox3d.hpp:82-98
//   Interpolant  model, fitted (bool), centers (N x 3), grad_centers (M x 3),
//                bbox, weights (vector)                         interpolant.hpp:345-367
#pragma once

#include <iosfwd>
#include <memory>
#include <string>

#include "interpolant.hpp"

namespace pam {

void write_model(std::ostream& os, const Model& m);
Model read_model(std::istream& is);
void save_model(const std::string& filename, const Model& m);
Model load_model(const std::string& filename);

// Interpolants with gradient data (grad_centers) are refused on load: this
// library fits and evaluates values only.
void save_interpolant(const std::string& filename, const Interpolant& ip);
std::unique_ptr<Interpolant> load_interpolant(const std::string& filename);

}  // namespace pam
