// INRIA-layout PLY scene loader (host C++), the reference's `ops.loadPly`
// (pybind.cpp:12-120, called from example.py:21-30).
//
// Vertex record: 62 little-endian floats -- position[3], normal[3],
// f_dc[3], f_rest[45] (channel-major: 15 R, 15 G, 15 B), opacity (logit),
// scale[3] (log), rotation[4] (w, x, y, z).  Output tensors (CPU, FP32):
// position [P,3]; SH [P,48] as 16 coefficients x RGB; opacity = sigmoid;
// cov3D [P,6] = upper triangle of Sigma = (S R)^T (S R) with the quaternion
// normalised first.  All products are evaluated in the same order as the
// reference's column-major 3x3 matrix library so the FP32 results match.

#include <torch/extension.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

namespace ps_ply {

struct Mat3 {
  float m[3][3];  // m[col][row]
};

static Mat3 mul(const Mat3& a, const Mat3& b) {
  Mat3 r;
  for (int c = 0; c < 3; ++c)
    for (int w = 0; w < 3; ++w)
      r.m[c][w] = a.m[0][w] * b.m[c][0] + a.m[1][w] * b.m[c][1] + a.m[2][w] * b.m[c][2];
  return r;
}

static Mat3 transpose(const Mat3& a) {
  Mat3 r;
  for (int c = 0; c < 3; ++c)
    for (int w = 0; w < 3; ++w) r.m[c][w] = a.m[w][c];
  return r;
}

// computeCov3D (pybind.cpp:22-56).
void cov3d_from(const float scale[3], float mod, const float q[4], float out[6]) {
  Mat3 S;
  std::memset(&S, 0, sizeof(S));
  S.m[0][0] = mod * scale[0];
  S.m[1][1] = mod * scale[1];
  S.m[2][2] = mod * scale[2];
  const float r = q[0], x = q[1], y = q[2], z = q[3];
  const float v[9] = {1.f - 2.f * (y * y + z * z), 2.f * (x * y - r * z), 2.f * (x * z + r * y),
                      2.f * (x * y + r * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - r * x),
                      2.f * (x * z - r * y), 2.f * (y * z + r * x), 1.f - 2.f * (x * x + y * y)};
  Mat3 R;
  for (int c = 0; c < 3; ++c)
    for (int w = 0; w < 3; ++w) R.m[c][w] = v[c * 3 + w];
  const Mat3 M = mul(S, R);
  const Mat3 Sig = mul(transpose(M), M);
  out[0] = Sig.m[0][0];
  out[1] = Sig.m[0][1];
  out[2] = Sig.m[0][2];
  out[3] = Sig.m[1][1];
  out[4] = Sig.m[1][2];
  out[5] = Sig.m[2][2];
}

// 4-vector normalisation: v * (1 / sqrt((x*x + y*y) + (z*z + w*w))).
static void normalize4(float q[4]) {
  const float t0 = q[0] * q[0], t1 = q[1] * q[1], t2 = q[2] * q[2], t3 = q[3] * q[3];
  const float inv = 1.0f / std::sqrt((t0 + t1) + (t2 + t3));
  for (int i = 0; i < 4; ++i) q[i] = q[i] * inv;
}

std::tuple<int64_t, at::Tensor, at::Tensor, at::Tensor, at::Tensor> load_ply(
    const std::string& name) {
  std::ifstream file(name, std::ios::binary);
  if (!file.is_open()) throw std::runtime_error("Failed to open file: " + name);
  int64_t n = 0;
  int n_float_props = 0;
  bool end_header = false, binary_le = false, in_vertex = false;
  std::string line;
  while (std::getline(file, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.rfind("format", 0) == 0) binary_le = line.find("binary_little_endian") != std::string::npos;
    if (line.rfind("element", 0) == 0) {
      in_vertex = line.rfind("element vertex", 0) == 0;
      if (in_vertex) {
        n = std::stoll(line.substr(15));
        if (n <= 0) throw std::runtime_error("Vertex number is not positive");
      }
    }
    if (in_vertex && line.rfind("property", 0) == 0) {
      if (line.rfind("property float ", 0) != 0)
        throw std::runtime_error("vertex properties must be float: " + line);
      ++n_float_props;
    }
    if (line == "end_header") {
      end_header = true;
      break;
    }
  }
  if (!end_header) throw std::runtime_error("Cannot find end of header");
  if (!binary_le) throw std::runtime_error("only binary_little_endian PLY files are supported");
  if (n_float_props != 62)
    throw std::runtime_error("expected 62 float properties per vertex (INRIA layout), found " +
                             std::to_string(n_float_props));
  auto opt = at::TensorOptions().dtype(at::kFloat);
  at::Tensor pos = at::empty({n, 3}, opt), shs = at::empty({n, 48}, opt),
             opac = at::empty({n}, opt), cov = at::empty({n, 6}, opt);
  float* P = pos.data_ptr<float>();
  float* SH = shs.data_ptr<float>();
  float* O = opac.data_ptr<float>();
  float* C = cov.data_ptr<float>();
  std::vector<float> rec(62);
  for (int64_t i = 0; i < n; ++i) {
    file.read(reinterpret_cast<char*>(rec.data()), 62 * sizeof(float));
    if (!file) throw std::runtime_error("PLY file truncated at vertex " + std::to_string(i));
    const float* sh = rec.data() + 6;  // f_dc[3] then f_rest[45]
    for (int k = 0; k < 3; ++k) P[3 * i + k] = rec[k];
    for (int c = 0; c < 3; ++c) SH[48 * i + c] = sh[c];
    for (int j = 1; j < 16; ++j) {
      SH[48 * i + 3 * j + 0] = sh[3 + (j - 1)];
      SH[48 * i + 3 * j + 1] = sh[3 + (j - 1) + 15];
      SH[48 * i + 3 * j + 2] = sh[3 + (j - 1) + 30];
    }
    O[i] = 1.0f / (1.0f + std::exp(-rec[54]));
    const float scale[3] = {std::exp(rec[55]), std::exp(rec[56]), std::exp(rec[57])};
    float q[4] = {rec[58], rec[59], rec[60], rec[61]};
    normalize4(q);
    cov3d_from(scale, 1.0f, q, C + 6 * i);
  }
  return std::make_tuple(n, pos, shs, opac, cov);
}

}  // namespace ps_ply
