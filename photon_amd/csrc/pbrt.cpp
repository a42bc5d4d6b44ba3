// pbrt-v4 scene files behind phx_pbrt_load: the reference's reader
// (parse_pbrt_file, pbrt_parser.cpp:233-637, with its PLY reader
// ply_reader.cpp:64-251 and image reader image_loader.cpp:19-104) and its
// conversion to a renderable scene (convert_pbrt_scene,
// pbrt_to_photon.cpp:58-503).  Host code: the result is a phx_scene_desc
// over arrays this module owns, plus the camera and the file's film, sampler
// and integrator settings.
//
// The reader accepts what the reference's does and ignores what it ignores:
// unknown directives and their arguments are skipped token by token
// (Translate among them), anonymous Material blocks are skipped, only named
// materials bind, "spectrum" parameters take a named spectrum only inside
// brackets, LookAt sets the camera without touching the transform, and
// ConcatTransform pre-multiplies the current transform.  Image files: PFM,
// binary PNM, Radiance HDR and PNG (8 and 16 bit, not interlaced) are read,
// low-dynamic-range ones converted to linear with the 2.2 power curve of the
// reference's loader; a texture in another format, like one that fails to
// load there, is left out with a warning.
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "context.h"

using namespace phx_detail;

namespace {

// ---- small float algebra in the reference's operation order -----------------
struct F3 {
  float x, y, z;
};
F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 scale(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(F3 a) { return std::sqrt(dot(a, a)); }
F3 normalize(F3 a) {
  const float l = length(a);
  return l > 0.f ? scale(a, 1.f / l) : F3{0.f, 0.f, 0.f};
}

// The parser's 16-float current transform; mul16(a, b)[r*4+c] = sum_k
// a[r*4+k] b[k*4+c] (pbrt_parser.cpp:331-340).
using M16 = std::array<float, 16>;
M16 identity16() {
  M16 m{};
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.f : 0.f;
  return m;
}
M16 mul16(const M16 &a, const M16 &b) {
  M16 r{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      float s = 0.f;
      for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
      r[i * 4 + j] = s;
    }
  return r;
}

// A 4x4 matrix read from the 16 floats as columns (mat4_from_pbrt_column_major,
// pbrt_to_photon.cpp:25-32) and the Mat4 operations the conversion uses
// (math_mat4.h:79-145).
struct Mat4 {
  float m[4][4];
  static Mat4 from_columns(const M16 &cm) {
    Mat4 r{};
    for (int col = 0; col < 4; ++col)
      for (int row = 0; row < 4; ++row) r.m[row][col] = cm[col * 4 + row];
    return r;
  }
  F3 point(F3 p) const {
    const float x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    const float y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    const float z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    const float w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (w != 0.f && w != 1.f) {
      const float iw = 1.f / w;
      return {x * iw, y * iw, z * iw};
    }
    return {x, y, z};
  }
  F3 direction(F3 d) const {
    return {m[0][0] * d.x + m[0][1] * d.y + m[0][2] * d.z, m[1][0] * d.x + m[1][1] * d.y + m[1][2] * d.z,
            m[2][0] * d.x + m[2][1] * d.y + m[2][2] * d.z};
  }
  // inverse of the affine part (a singular 3x3 gives the zero matrix)
  Mat4 inverse() const {
    Mat4 r{};
    const float a00 = m[0][0], a01 = m[0][1], a02 = m[0][2];
    const float a10 = m[1][0], a11 = m[1][1], a12 = m[1][2];
    const float a20 = m[2][0], a21 = m[2][1], a22 = m[2][2];
    const float det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
    if (det == 0.f) return r;
    r.m[3][3] = 1.f;
    const float id = 1.f / det;
    r.m[0][0] = (a11 * a22 - a12 * a21) * id;
    r.m[0][1] = (a02 * a21 - a01 * a22) * id;
    r.m[0][2] = (a01 * a12 - a02 * a11) * id;
    r.m[1][0] = (a12 * a20 - a10 * a22) * id;
    r.m[1][1] = (a00 * a22 - a02 * a20) * id;
    r.m[1][2] = (a02 * a10 - a00 * a12) * id;
    r.m[2][0] = (a10 * a21 - a11 * a20) * id;
    r.m[2][1] = (a01 * a20 - a00 * a21) * id;
    r.m[2][2] = (a00 * a11 - a01 * a10) * id;
    const float tx = m[0][3], ty = m[1][3], tz = m[2][3];
    r.m[0][3] = -(r.m[0][0] * tx + r.m[0][1] * ty + r.m[0][2] * tz);
    r.m[1][3] = -(r.m[1][0] * tx + r.m[1][1] * ty + r.m[1][2] * tz);
    r.m[2][3] = -(r.m[2][0] * tx + r.m[2][1] * ty + r.m[2][2] * tz);
    return r;
  }
};

// ---- the file's contents -----------------------------------------------------
struct Texture {
  std::string value_type, class_type, filename;
  std::vector<float> rgb;  // linear RGB, rows from the top
  int width = 0, height = 0;
};
struct NamedMat {
  std::string type;
  F3 reflectance{0.5f, 0.5f, 0.5f}, eta{0.f, 0.f, 0.f}, k{0.f, 0.f, 0.f};
  float roughness = 0.f, uroughness = -1.f, vroughness = -1.f, eta_scalar = 1.5f;
  std::string reflectance_texture;
};
struct Mesh {
  std::vector<float> positions, normals, uvs;
  std::vector<int> indices;
  std::string material, alpha_texture;
  F3 emission{0.f, 0.f, 0.f};
  bool emissive = false;
  M16 transform = identity16();
};
struct Camera {
  float fov = 45.f, lensradius = 0.f, focaldistance = 0.f;
  M16 transform = identity16();
  F3 from{0.f, 0.f, 0.f}, at{0.f, 0.f, 1.f}, up{0.f, 1.f, 0.f};
  bool lookat = false;
};
struct File {
  int width = 512, height = 512, spp = 64, max_depth = 8;
  Camera camera;
  std::map<std::string, Texture> textures;  // by name (the conversion's order)
  std::map<std::string, NamedMat> materials;
  std::vector<Mesh> meshes;
  std::map<std::string, std::vector<Mesh>> objects;
  std::vector<std::pair<std::string, M16>> instances;
  std::string env_file;
  M16 env_transform = identity16();
  float env_scale = 1.f;
  bool env = false;
  uint64_t max_instances_per_object = 500, max_total_triangles = 500000000;
};

// ---- text: Include / Import lines spliced in (pbrt_parser.cpp:201-231) --------
std::string dir_of(const std::string &path) {
  const size_t s = path.find_last_of("/\\");
  return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}
std::string read_spliced(const std::string &path, const std::string &root, int depth) {
  if (depth > 10) return std::string();
  std::ifstream f(path);
  if (!f.is_open()) {
    std::fprintf(stderr, "  Warning: cannot open %s\n", path.c_str());
    return std::string();
  }
  std::string out, line;
  while (std::getline(f, line)) {
    const size_t b = line.find_first_not_of(" \t");
    const std::string t = b == std::string::npos ? line : line.substr(b);
    if (t.size() > 5 && (t.compare(0, 7, "Include") == 0 || t.compare(0, 6, "Import") == 0)) {
      const size_t q1 = t.find('"'), q2 = t.rfind('"');
      if (q1 != std::string::npos && q2 > q1) {
        out += read_spliced(root + t.substr(q1 + 1, q2 - q1 - 1), root, depth + 1);
        continue;
      }
    }
    out += line;
    out += '\n';
  }
  return out;
}

// ---- tokens (pbrt_parser.cpp:19-94) ---------------------------------------------
enum class Tk { Str, Num, Open, Close, Word, End };
struct Token {
  Tk kind = Tk::End;
  std::string text;
  float num = 0.f;
};
class Lexer {
 public:
  explicit Lexer(const std::string &s) : p_(s.data()), e_(s.data() + s.size()) {}
  Token next() {
    for (;;) {
      while (p_ < e_) {
        if (std::isspace(static_cast<unsigned char>(*p_))) {
          ++p_;
        } else if (*p_ == '#') {
          while (p_ < e_ && *p_ != '\n') ++p_;
        } else {
          break;
        }
      }
      if (p_ >= e_) return Token{};
      const char c = *p_;
      if (c == '[' || c == ']') {
        ++p_;
        return Token{c == '[' ? Tk::Open : Tk::Close, std::string(1, c), 0.f};
      }
      if (c == '"') {
        const char *s = ++p_;
        while (p_ < e_ && *p_ != '"') ++p_;
        Token t{Tk::Str, std::string(s, p_), 0.f};
        if (p_ < e_) ++p_;
        return t;
      }
      if (c == '-' || c == '+' || c == '.' || std::isdigit(static_cast<unsigned char>(c))) {
        const char *s = p_;
        if (c == '-' || c == '+') ++p_;
        while (p_ < e_ && (std::isdigit(static_cast<unsigned char>(*p_)) || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
                           *p_ == '-' || *p_ == '+'))
          ++p_;
        const std::string txt(s, p_);
        return Token{Tk::Num, txt, std::strtof(txt.c_str(), nullptr)};
      }
      if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
        const char *s = p_;
        while (p_ < e_ && (std::isalnum(static_cast<unsigned char>(*p_)) || *p_ == '_')) ++p_;
        return Token{Tk::Word, std::string(s, p_), 0.f};
      }
      ++p_;  // any other character is passed over
    }
  }
  Token peek() {
    const char *keep = p_;
    Token t = next();
    p_ = keep;
    return t;
  }

 private:
  const char *p_, *e_;
};

// One "type name" parameter and its values (parse_param, pbrt_parser.cpp:104-160).
struct Param {
  std::string type, name;
  std::vector<float> floats;
  std::vector<int> ints;
  std::vector<std::string> strings;
  bool integral() const { return type == "integer" || type == "bool"; }
};
Param read_param(Lexer &lx) {
  Param pv;
  const Token t = lx.next();
  if (t.kind != Tk::Str) return pv;
  const size_t sp = t.text.find(' ');
  if (sp == std::string::npos) {
    pv.name = t.text;
  } else {
    pv.type = t.text.substr(0, sp);
    pv.name = t.text.substr(sp + 1);
  }
  const Token v = lx.peek();
  if (v.kind == Tk::Str && (pv.type == "string" || pv.type == "texture")) {
    lx.next();
    pv.strings.push_back(v.text);
    return pv;
  }
  if (v.kind == Tk::Num && (pv.type == "float" || pv.integral())) {
    lx.next();
    if (pv.integral()) pv.ints.push_back(int(v.num));
    else pv.floats.push_back(v.num);
    return pv;
  }
  if (v.kind != Tk::Open) return pv;
  lx.next();
  for (;;) {
    const Token w = lx.peek();
    if (w.kind == Tk::Close) {
      lx.next();
      break;
    }
    if (w.kind == Tk::End) break;
    lx.next();
    if (w.kind == Tk::Num) {
      if (pv.integral()) pv.ints.push_back(int(w.num));
      else pv.floats.push_back(w.num);
    } else if (w.kind == Tk::Str) {
      pv.strings.push_back(w.text);
    }
  }
  return pv;
}
template <class F>
void each_param(Lexer &lx, F f) {
  while (lx.peek().kind == Tk::Str) f(read_param(lx));
}
void skip_params(Lexer &lx) {
  each_param(lx, [](const Param &) {});
}

// sRGB stand-ins for measured metal spectra (pbrt_parser.cpp:170-183).
F3 named_spectrum(const std::string &n) {
  static const std::map<std::string, F3> table = {
      {"metal-Ag-eta", {0.155f, 0.116f, 0.138f}}, {"metal-Ag-k", {4.828f, 3.122f, 2.146f}},
      {"metal-Au-eta", {0.143f, 0.374f, 1.442f}}, {"metal-Au-k", {3.983f, 2.387f, 1.603f}},
      {"metal-Cu-eta", {0.200f, 0.924f, 1.102f}}, {"metal-Cu-k", {3.912f, 2.452f, 2.142f}},
      {"metal-Al-eta", {1.345f, 0.965f, 0.612f}}, {"metal-Al-k", {7.474f, 6.400f, 5.275f}},
      {"metal-Fe-eta", {2.912f, 2.950f, 2.584f}}, {"metal-Fe-k", {3.070f, 2.930f, 2.771f}}};
  const auto it = table.find(n);
  return it == table.end() ? F3{0.f, 0.f, 0.f} : it->second;
}

// ---- PLY (ply_reader.cpp:64-251) -----------------------------------------------
struct PlyProp {
  std::string name, type, count_type, value_type;
  bool list = false;
};
struct PlyElem {
  std::string name;
  int count = 0;
  std::vector<PlyProp> props;
};
int ply_type_size(const std::string &t) {
  if (t == "double" || t == "float64" || t == "int64" || t == "uint64") return 8;
  if (t == "uchar" || t == "uint8" || t == "char" || t == "int8") return 1;
  if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
  return 4;
}
template <class T>
T read_le(std::istream &in) {
  T v{};
  in.read(reinterpret_cast<char *>(&v), sizeof v);
  return v;
}
int read_ply_int(std::istream &in, const std::string &t) {
  if (t == "uchar" || t == "uint8") return int(read_le<uint8_t>(in));
  if (t == "short" || t == "int16") return int(read_le<int16_t>(in));
  if (t == "ushort" || t == "uint16") return int(read_le<uint16_t>(in));
  if (t == "uint" || t == "uint32") return int(read_le<uint32_t>(in));
  return int(read_le<int32_t>(in));  // int, int32 and anything else
}
bool read_ply(const std::string &path, Mesh &mesh) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) {
    std::fprintf(stderr, "PLY: cannot open %s\n", path.c_str());
    return false;
  }
  std::string line;
  std::getline(in, line);
  if (line.find("ply") == std::string::npos) return false;
  bool ascii = false, big = false;
  std::vector<PlyElem> elems;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '\r') continue;
    if (line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw == "end_header") break;
    if (kw == "format") {
      std::string f;
      ls >> f;
      ascii = f == "ascii";
      big = f == "binary_big_endian";
    } else if (kw == "element") {
      PlyElem e;
      ls >> e.name >> e.count;
      elems.push_back(e);
    } else if (kw == "property" && !elems.empty()) {
      PlyProp p;
      std::string w;
      ls >> w;
      if (w == "list") {
        p.list = true;
        ls >> p.count_type >> p.value_type >> p.name;
      } else {
        p.type = w;
        ls >> p.name;
      }
      elems.back().props.push_back(p);
    }
  }
  if (big) {
    std::fprintf(stderr, "PLY: big-endian not supported: %s\n", path.c_str());
    return false;
  }
  const PlyElem *ve = nullptr, *fe = nullptr;
  for (const PlyElem &e : elems) {
    if (e.name == "vertex") ve = &e;
    if (e.name == "face") fe = &e;
  }
  if (!ve || !fe) return false;
  int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, iu = -1, iv = -1;
  for (int i = 0; i < int(ve->props.size()); ++i) {
    const std::string &n = ve->props[i].name;
    if (n == "x") ix = i;
    else if (n == "y") iy = i;
    else if (n == "z") iz = i;
    else if (n == "nx") inx = i;
    else if (n == "ny") iny = i;
    else if (n == "nz") inz = i;
    else if (n == "u" || n == "s" || n == "texture_u") iu = i;
    else if (n == "v" || n == "t" || n == "texture_v") iv = i;
  }
  const bool has_n = inx >= 0 && iny >= 0 && inz >= 0, has_uv = iu >= 0 && iv >= 0;
  const size_t nv = size_t(std::max(ve->count, 0));
  mesh.positions.assign(nv * 3, 0.f);
  if (has_n) mesh.normals.assign(nv * 3, 0.f);
  if (has_uv) mesh.uvs.assign(nv * 2, 0.f);
  auto fan = [&](const std::vector<int> &f) {
    for (size_t j = 1; j + 1 < f.size(); ++j) {
      mesh.indices.push_back(f[0]);
      mesh.indices.push_back(f[j]);
      mesh.indices.push_back(f[j + 1]);
    }
  };
  constexpr int kMaxProps = 16;  // values kept per vertex record
  for (size_t v = 0; v < nv; ++v) {
    float vals[kMaxProps] = {};
    for (int p = 0; p < int(ve->props.size()); ++p) {
      float x = 0.f;
      if (ascii) {
        in >> x;
      } else if (ve->props[p].type == "float" || ve->props[p].type == "float32") {
        x = read_le<float>(in);
      } else {
        in.seekg(ply_type_size(ve->props[p].type), std::ios::cur);
      }
      if (p < kMaxProps) vals[p] = x;
    }
    if (ix >= 0 && ix < kMaxProps) mesh.positions[v * 3] = vals[ix];
    if (iy >= 0 && iy < kMaxProps) mesh.positions[v * 3 + 1] = vals[iy];
    if (iz >= 0 && iz < kMaxProps) mesh.positions[v * 3 + 2] = vals[iz];
    if (has_n && std::max({inx, iny, inz}) < kMaxProps) {
      mesh.normals[v * 3] = vals[inx];
      mesh.normals[v * 3 + 1] = vals[iny];
      mesh.normals[v * 3 + 2] = vals[inz];
    }
    if (has_uv && std::max(iu, iv) < kMaxProps) {
      mesh.uvs[v * 2] = vals[iu];
      mesh.uvs[v * 2 + 1] = vals[iv];
    }
  }
  for (int f = 0; f < fe->count; ++f) {
    if (ascii) {
      int n = 0;
      in >> n;
      std::vector<int> idx(size_t(std::max(n, 0)));
      for (int &k : idx) in >> k;
      fan(idx);
      continue;
    }
    for (const PlyProp &p : fe->props) {
      if (p.list && (p.name == "vertex_indices" || p.name == "vertex_index")) {
        const int n = read_ply_int(in, p.count_type);
        std::vector<int> idx(size_t(std::max(n, 0)));
        for (int &k : idx) k = read_ply_int(in, p.value_type);
        fan(idx);
      } else if (p.list) {
        const int n = read_ply_int(in, p.count_type);
        in.seekg(std::streamoff(std::max(n, 0)) * ply_type_size(p.value_type), std::ios::cur);
      } else {
        in.seekg(ply_type_size(p.type), std::ios::cur);
      }
    }
  }
  return !mesh.positions.empty() && !mesh.indices.empty();
}

// ---- images (image_loader.cpp:19-104) -------------------------------------------
// Low-dynamic-range channel to linear as stb_image's float loader does it.
float ldr_linear(float c) { return std::pow(c, 2.2f); }

bool read_pfm(const std::string &path, Texture &t) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return false;
  std::string magic, dims, sc;
  std::getline(in, magic);
  if (magic != "PF" && magic != "Pf") return false;
  const int ch = magic == "PF" ? 3 : 1;
  std::getline(in, dims);
  int w = 0, h = 0;
  std::sscanf(dims.c_str(), "%d %d", &w, &h);
  if (w <= 0 || h <= 0) return false;
  std::getline(in, sc);
  float s = 0.f;
  std::sscanf(sc.c_str(), "%f", &s);
  float mag = std::fabs(s);
  if (mag == 0.f) mag = 1.f;
  t.width = w;
  t.height = h;
  t.rgb.assign(size_t(w) * h * 3, 0.f);
  std::vector<float> row(size_t(w) * ch);
  for (int y = h - 1; y >= 0; --y) {  // rows are stored bottom first
    in.read(reinterpret_cast<char *>(row.data()), std::streamsize(row.size() * sizeof(float)));
    for (int x = 0; x < w; ++x)
      for (int c = 0; c < 3; ++c) t.rgb[(size_t(y) * w + x) * 3 + c] = row[size_t(x) * ch + (ch == 3 ? c : 0)] * mag;
  }
  return true;
}

// 8-bit-per-channel rows to linear RGB (grey and grey+alpha widen; alpha drops).
void ldr_rows(const std::vector<uint8_t> &px, int w, int h, int comps, Texture &t) {
  t.width = w;
  t.height = h;
  t.rgb.assign(size_t(w) * h * 3, 0.f);
  for (size_t i = 0; i < size_t(w) * h; ++i)
    for (int c = 0; c < 3; ++c) {
      const int src = comps >= 3 ? c : 0;
      t.rgb[i * 3 + c] = ldr_linear(float(px[i * comps + src]) / 255.f);
    }
}

bool read_pnm(const std::string &path, Texture &t) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return false;
  std::string magic;
  in >> magic;
  if (magic != "P5" && magic != "P6") return false;
  auto number = [&]() {
    int v = -1;
    for (;;) {
      in >> std::ws;
      if (in.peek() == '#') {
        std::string c;
        std::getline(in, c);
        continue;
      }
      in >> v;
      return v;
    }
  };
  const int w = number(), h = number(), maxv = number();
  if (w <= 0 || h <= 0 || maxv <= 0 || maxv > 255) return false;
  in.get();  // the single whitespace before the samples
  const int comps = magic == "P6" ? 3 : 1;
  std::vector<uint8_t> px(size_t(w) * h * comps);
  in.read(reinterpret_cast<char *>(px.data()), std::streamsize(px.size()));
  if (!in) return false;
  ldr_rows(px, w, h, comps, t);
  return true;
}

// Radiance RGBE: flat and new-style run-length scanlines; value = byte * 2^(e - 136).
bool read_hdr(const std::string &path, Texture &t) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return false;
  std::string line;
  std::getline(in, line);
  if (line.rfind("#?", 0) != 0) return false;
  while (std::getline(in, line) && !line.empty()) {
  }
  std::getline(in, line);
  char ya[3] = {0}, xa[3] = {0};
  int h = 0, w = 0;
  if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::strcmp(ya, "-Y") != 0 ||
      std::strcmp(xa, "+X") != 0 || w <= 0 || h <= 0)
    return false;
  t.width = w;
  t.height = h;
  t.rgb.assign(size_t(w) * h * 3, 0.f);
  std::vector<uint8_t> scan(size_t(w) * 4);
  for (int y = 0; y < h; ++y) {
    uint8_t head[4];
    in.read(reinterpret_cast<char *>(head), 4);
    if (!in) return false;
    if (w >= 8 && w < 32768 && head[0] == 2 && head[1] == 2 && !(head[2] & 0x80) && ((head[2] << 8) | head[3]) == w) {
      for (int c = 0; c < 4; ++c)
        for (int x = 0; x < w;) {
          int n = in.get();
          if (n < 0) return false;
          if (n > 128) {
            n -= 128;
            const int v = in.get();
            for (int k = 0; k < n && x < w; ++k) scan[size_t(x++) * 4 + c] = uint8_t(v);
          } else {
            for (int k = 0; k < n && x < w; ++k) scan[size_t(x++) * 4 + c] = uint8_t(in.get());
          }
        }
    } else {
      std::memcpy(scan.data(), head, 4);
      in.read(reinterpret_cast<char *>(scan.data() + 4), std::streamsize(scan.size() - 4));
    }
    for (int x = 0; x < w; ++x) {
      const uint8_t *q = &scan[size_t(x) * 4];
      const float f = q[3] ? std::ldexp(1.f, int(q[3]) - 136) : 0.f;
      for (int c = 0; c < 3; ++c) t.rgb[(size_t(y) * w + x) * 3 + c] = float(q[c]) * f;
    }
  }
  return true;
}

// PNG: 8- and 16-bit grey, grey+alpha, RGB, RGBA and palette images, not
// interlaced (16-bit samples keep their high byte).
bool read_png(const std::string &path, Texture &t) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return false;
  uint8_t sig[8];
  in.read(reinterpret_cast<char *>(sig), 8);
  static const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (!in || std::memcmp(sig, kSig, 8) != 0) return false;
  auto be32 = [](const uint8_t *b) { return (uint32_t(b[0]) << 24) | (uint32_t(b[1]) << 16) | (uint32_t(b[2]) << 8) | b[3]; };
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte;
  for (;;) {
    uint8_t hd[8];
    in.read(reinterpret_cast<char *>(hd), 8);
    if (!in) return false;
    const uint32_t len = be32(hd);
    const std::string type(reinterpret_cast<char *>(hd + 4), 4);
    std::vector<uint8_t> data(len);
    in.read(reinterpret_cast<char *>(data.data()), len);
    in.seekg(4, std::ios::cur);  // CRC
    if (!in) return false;
    if (type == "IHDR" && len >= 13) {
      w = be32(&data[0]);
      h = be32(&data[4]);
      depth = data[8];
      ctype = data[9];
      interlace = data[12];
    } else if (type == "PLTE") {
      plte = data;
    } else if (type == "IDAT") {
      idat.insert(idat.end(), data.begin(), data.end());
    } else if (type == "IEND") {
      break;
    }
  }
  const int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
  if (!w || !h || !chans || interlace || (depth != 8 && depth != 16) || (ctype == 3 && depth != 8)) return false;
  const size_t bpp = size_t(chans) * (depth / 8), stride = size_t(w) * bpp;
  std::vector<uint8_t> raw((stride + 1) * h);
  uLongf raw_len = uLongf(raw.size());
  if (uncompress(raw.data(), &raw_len, idat.data(), uLong(idat.size())) != Z_OK || raw_len != raw.size()) return false;
  std::vector<uint8_t> img(stride * h), prev(stride, 0);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t filter = raw[y * (stride + 1)];
    const uint8_t *src = &raw[y * (stride + 1) + 1];
    uint8_t *row = &img[y * stride];
    for (size_t i = 0; i < stride; ++i) {
      const int a = i >= bpp ? row[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
      int pred = 0;
      if (filter == 1) pred = a;
      else if (filter == 2) pred = b;
      else if (filter == 3) pred = (a + b) / 2;
      else if (filter == 4) {
        const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
        pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
      }
      row[i] = uint8_t(src[i] + pred);
    }
    std::memcpy(prev.data(), row, stride);
  }
  // to 8-bit samples (high byte of 16-bit ones), palette expanded to RGB
  const int comps = ctype == 3 ? 3 : chans;
  std::vector<uint8_t> px(size_t(w) * h * comps);
  for (size_t i = 0; i < size_t(w) * h; ++i) {
    if (ctype == 3) {
      const size_t k = size_t(img[i]) * 3;
      for (int c = 0; c < 3; ++c) px[i * 3 + c] = k + 2 < plte.size() ? plte[k + c] : 0;
    } else {
      for (int c = 0; c < chans; ++c) px[i * chans + c] = img[i * bpp + size_t(c) * (depth / 8)];
    }
  }
  ldr_rows(px, int(w), int(h), comps, t);
  return true;
}

bool read_image(const std::string &path, Texture &t) {
  std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : std::string();
  std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return char(std::tolower(c)); });
  bool ok = false;
  if (ext == ".pfm") ok = read_pfm(path, t);
  else if (ext == ".hdr") ok = read_hdr(path, t);
  else if (ext == ".png") ok = read_png(path, t);
  else if (ext == ".ppm" || ext == ".pgm" || ext == ".pnm") ok = read_pnm(path, t);
  if (!ok) std::fprintf(stderr, "Warning: cannot load texture: %s\n", path.c_str());
  return ok;
}

// ---- the parser (parse_pbrt_file, pbrt_parser.cpp:233-637) ----------------------
struct GraphicsState {
  std::string material;
  M16 ctm = identity16();
  F3 area_L{0.f, 0.f, 0.f};
  bool area = false;
};

void tessellate_sphere(float radius, Mesh &m) {
  const int slices = 32, stacks = 16;
  for (int j = 0; j <= stacks; ++j) {
    const float theta = float(j) / float(stacks) * 3.14159265f;
    const float st = std::sin(theta), ct = std::cos(theta);
    for (int i = 0; i <= slices; ++i) {
      const float phi = float(i) / float(slices) * 6.28318530f;
      const float cp = std::cos(phi), sp = std::sin(phi);
      m.positions.insert(m.positions.end(), {st * cp * radius, ct * radius, st * sp * radius});
      m.normals.insert(m.normals.end(), {st * cp, ct, st * sp});
    }
  }
  for (int j = 0; j < stacks; ++j)
    for (int i = 0; i < slices; ++i) {
      const int a = j * (slices + 1) + i, b = a + 1, c = a + slices + 1, d = c + 1;
      m.indices.insert(m.indices.end(), {a, c, b, b, c, d});
    }
}

File parse_file(const std::string &path, bool use_instancing) {
  const std::string root = dir_of(path);
  const std::string text = read_spliced(path, root, 0);
  Lexer lx(text);
  File f;
  std::vector<GraphicsState> stack;
  GraphicsState gs;
  bool in_world = false;
  std::string open_object;
  std::map<std::string, uint64_t> per_object;
  uint64_t tris = 0;
  auto add_mesh = [&](Mesh &&m) {
    if (!open_object.empty()) {
      f.objects[open_object].push_back(std::move(m));
    } else {
      tris += m.indices.size() / 3;
      f.meshes.push_back(std::move(m));
    }
  };
  auto new_mesh = [&]() {
    Mesh m;
    m.material = gs.material;
    m.transform = gs.ctm;
    if (gs.area) {
      m.emissive = true;
      m.emission = gs.area_L;
    }
    return m;
  };
  for (;;) {
    const Token t = lx.next();
    if (t.kind == Tk::End) break;
    if (t.kind != Tk::Word) continue;
    const std::string &d = t.text;
    if (d == "Integrator") {
      lx.next();
      each_param(lx, [&](const Param &p) {
        if (p.name == "maxdepth" && !p.ints.empty()) f.max_depth = p.ints[0];
      });
    } else if (d == "Sampler") {
      lx.next();
      each_param(lx, [&](const Param &p) {
        if (p.name == "pixelsamples" && !p.ints.empty()) f.spp = p.ints[0];
      });
    } else if (d == "Film") {
      lx.next();
      each_param(lx, [&](const Param &p) {
        if (p.name == "xresolution" && !p.ints.empty()) f.width = p.ints[0];
        else if (p.name == "yresolution" && !p.ints.empty()) f.height = p.ints[0];
      });
    } else if (d == "PixelFilter") {
      lx.next();
      skip_params(lx);
    } else if (d == "Camera") {
      lx.next();
      f.camera.transform = gs.ctm;
      each_param(lx, [&](const Param &p) {
        if (p.floats.empty()) return;
        if (p.name == "fov") f.camera.fov = p.floats[0];
        else if (p.name == "lensradius") f.camera.lensradius = p.floats[0];
        else if (p.name == "focaldistance") f.camera.focaldistance = p.floats[0];
      });
    } else if (d == "Transform") {
      if (lx.next().kind == Tk::Open) {
        for (int i = 0; i < 16; ++i) gs.ctm[i] = lx.next().num;
        lx.next();
      }
    } else if (d == "LookAt") {
      float v[9];
      for (float &x : v) x = lx.next().num;
      f.camera.from = {v[0], v[1], v[2]};
      f.camera.at = {v[3], v[4], v[5]};
      f.camera.up = {v[6], v[7], v[8]};
      f.camera.lookat = true;
    } else if (d == "Scale") {
      const float sx = lx.next().num, sy = lx.next().num, sz = lx.next().num;
      const M16 s = {sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, sz, 0, 0, 0, 0, 1};
      gs.ctm = mul16(gs.ctm, s);
    } else if (d == "Rotate") {
      const float angle = lx.next().num;
      float ax = lx.next().num, ay = lx.next().num, az = lx.next().num;
      const float len = std::sqrt(ax * ax + ay * ay + az * az);
      if (len > 0.f) {
        ax /= len;
        ay /= len;
        az /= len;
      }
      const float rad = angle * 3.14159265f / 180.f;
      const float c = std::cos(rad), s = std::sin(rad), u = 1.f - c;
      const M16 r = {u * ax * ax + c,      u * ax * ay - s * az, u * ax * az + s * ay, 0,
                     u * ax * ay + s * az, u * ay * ay + c,      u * ay * az - s * ax, 0,
                     u * ax * az - s * ay, u * ay * az + s * ax, u * az * az + c,      0,
                     0,                    0,                    0,                    1};
      gs.ctm = mul16(gs.ctm, r);
    } else if (d == "ConcatTransform") {
      if (lx.next().kind == Tk::Open) {
        M16 m{};
        for (int i = 0; i < 16; ++i) m[i] = lx.next().num;
        lx.next();
        gs.ctm = mul16(m, gs.ctm);
      }
    } else if (d == "Identity") {
      gs.ctm = identity16();
    } else if (d == "WorldBegin") {
      in_world = true;
      gs.ctm = identity16();
    } else if (d == "AttributeBegin") {
      stack.push_back(gs);
    } else if (d == "AttributeEnd") {
      if (!stack.empty()) {
        gs = stack.back();
        stack.pop_back();
      }
    } else if (d == "MakeNamedMaterial") {
      const std::string name = lx.next().text;
      NamedMat m;
      each_param(lx, [&](const Param &p) {
        if (p.name == "type" && !p.strings.empty()) {
          m.type = p.strings[0];
        } else if (p.name == "reflectance" && p.type == "texture" && !p.strings.empty()) {
          m.reflectance_texture = p.strings[0];
        } else if (p.name == "reflectance" && p.floats.size() >= 3) {
          m.reflectance = {p.floats[0], p.floats[1], p.floats[2]};
        } else if (p.name == "roughness" && !p.floats.empty()) {
          m.roughness = p.floats[0];
        } else if (p.name == "eta") {
          if (!p.strings.empty()) m.eta = named_spectrum(p.strings[0]);
          else if (p.floats.size() >= 3) m.eta = {p.floats[0], p.floats[1], p.floats[2]};
          else if (p.floats.size() == 1) m.eta_scalar = p.floats[0];
        } else if (p.name == "k") {
          if (!p.strings.empty()) m.k = named_spectrum(p.strings[0]);
          else if (p.floats.size() >= 3) m.k = {p.floats[0], p.floats[1], p.floats[2]};
        } else if (p.name == "uroughness" && !p.floats.empty()) {
          m.uroughness = p.floats[0];
        } else if (p.name == "vroughness" && !p.floats.empty()) {
          m.vroughness = p.floats[0];
        }
      });
      f.materials[name] = m;
    } else if (d == "NamedMaterial") {
      gs.material = lx.next().text;
    } else if (d == "AreaLightSource") {
      lx.next();
      gs.area = true;
      each_param(lx, [&](const Param &p) {
        if (p.name == "L" && p.floats.size() >= 3) gs.area_L = {p.floats[0], p.floats[1], p.floats[2]};
      });
    } else if (d == "Shape") {
      const std::string type = lx.next().text;
      if (type == "trianglemesh") {
        Mesh m = new_mesh();
        each_param(lx, [&](const Param &p) {
          if ((p.name == "P" || p.name == "p") && p.type == "point3") m.positions = p.floats;
          else if (p.name == "indices" && p.type == "integer") m.indices = p.ints;
          else if (p.name == "N" && p.type == "normal") m.normals = p.floats;
          else if (p.name == "uv" && p.type == "point2") m.uvs = p.floats;
        });
        if (!m.positions.empty() && !m.indices.empty()) add_mesh(std::move(m));
      } else if (type == "plymesh") {
        Mesh m = new_mesh();
        std::string file;
        each_param(lx, [&](const Param &p) {
          if (p.name == "filename" && !p.strings.empty()) file = p.strings[0];
          else if (p.name == "alpha" && p.type == "texture" && !p.strings.empty()) m.alpha_texture = p.strings[0];
        });
        if (!file.empty()) {
          if (read_ply(root + file, m)) add_mesh(std::move(m));
          else std::fprintf(stderr, "Warning: failed to load PLY: %s\n", (root + file).c_str());
        }
      } else if (type == "sphere") {
        Mesh m = new_mesh();
        float radius = 1.f;
        each_param(lx, [&](const Param &p) {
          if (p.name == "radius" && !p.floats.empty()) radius = p.floats[0];
        });
        tessellate_sphere(radius, m);
        add_mesh(std::move(m));
      } else {
        skip_params(lx);
      }
    } else if (d == "Texture") {
      const std::string name = lx.next().text;
      Texture tx;
      tx.value_type = lx.next().text;
      tx.class_type = lx.next().text;
      each_param(lx, [&](const Param &p) {
        if (p.name == "filename" && !p.strings.empty()) tx.filename = p.strings[0];
      });
      f.textures[name] = tx;
    } else if (d == "LightSource") {
      const std::string type = lx.next().text;
      if (type == "infinite") {
        f.env = true;
        f.env_transform = gs.ctm;
        each_param(lx, [&](const Param &p) {
          if (p.name == "filename" && !p.strings.empty()) f.env_file = p.strings[0];
          else if ((p.name == "scale" || p.name == "L") && p.floats.size() >= 3)
            f.env_scale = (p.floats[0] + p.floats[1] + p.floats[2]) / 3.f;
        });
      } else {
        skip_params(lx);
      }
    } else if (d == "ObjectBegin") {
      open_object = lx.next().text;
    } else if (d == "ObjectEnd") {
      open_object.clear();
    } else if (d == "ObjectInstance") {
      const std::string name = lx.next().text;
      const auto it = f.objects.find(name);
      if (it == f.objects.end()) continue;
      f.instances.emplace_back(name, gs.ctm);
      if (use_instancing || tris >= f.max_total_triangles) continue;
      uint64_t &placed = per_object[name];
      if (placed >= f.max_instances_per_object) continue;
      uint64_t obj_tris = 0;
      for (const Mesh &m : it->second) obj_tris += m.indices.size() / 3;
      if (tris + obj_tris > f.max_total_triangles) continue;
      for (const Mesh &m : it->second) {
        Mesh c = m;
        c.transform = gs.ctm;
        f.meshes.push_back(std::move(c));
      }
      tris += obj_tris;
      ++placed;
    } else if (d == "Material" || d == "MakeNamedMedium") {
      lx.next();
      skip_params(lx);
    } else if (d == "ReverseOrientation" || d == "TransformBegin" || d == "TransformEnd" || d == "MediumInterface") {
      skip_params(lx);
    }
  }
  (void)in_world;
  return f;
}

}  // namespace

// ---- the converted scene ---------------------------------------------------------
struct phx_pbrt_scene {
  std::vector<float> positions, normals, texcoords, texture_pixels, env_pixels;
  std::vector<uint32_t> indices, material_ids, emissive_ids;
  std::vector<float> emissive_areas;
  std::vector<phx_material> materials;
  std::vector<phx_light> lights;
  std::vector<phx_texture_info> texture_infos;
  uint32_t env_w = 0, env_h = 0;
  float env_w2t[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  phx_camera camera{};
  uint32_t width = 512, height = 512, spp = 64, max_depth = 8;
  uint32_t objects = 0, instances = 0, meshes = 0;
};

namespace {

// convert_pbrt_scene (pbrt_to_photon.cpp:58-503) without the instanced
// geometry lists (this renderer draws the flattened meshes).
void convert(File &f, const std::string &root, phx_pbrt_scene &out) {
  for (auto &kv : f.textures)
    if (kv.second.class_type == "imagemap" && !kv.second.filename.empty() && kv.second.rgb.empty())
      read_image(root + kv.second.filename, kv.second);
  std::map<std::string, int32_t> tex_id;
  std::vector<const Texture *> texs;
  for (const auto &kv : f.textures)
    if (!kv.second.rgb.empty() && kv.second.width > 0 && kv.second.height > 0) {
      tex_id[kv.first] = int32_t(texs.size());
      texs.push_back(&kv.second);
    }
  std::map<std::string, uint32_t> mat_id;
  auto blended_roughness = [](const NamedMat &p) {
    return p.uroughness >= 0.f ? (p.uroughness + (p.vroughness >= 0.f ? p.vroughness : p.uroughness)) * 0.5f
                               : p.roughness;
  };
  for (const auto &kv : f.materials) {
    const NamedMat &p = kv.second;
    mat_id[kv.first] = uint32_t(out.materials.size());
    phx_material m = default_material();
    if (!p.reflectance_texture.empty()) {
      const auto it = tex_id.find(p.reflectance_texture);
      if (it != tex_id.end()) m.base_color_tex = it->second;
    }
    auto set_base = [&](F3 c) {
      m.base_color[0] = c.x;
      m.base_color[1] = c.y;
      m.base_color[2] = c.z;
    };
    if (p.type == "diffuse") {
      set_base(p.reflectance);
      m.roughness = 1.f;
      m.metallic = 0.f;
    } else if (p.type == "conductor") {
      auto f0 = [](float eta, float k) {
        const float num = (eta - 1.f) * (eta - 1.f) + k * k, den = (eta + 1.f) * (eta + 1.f) + k * k;
        return den > 0.f ? num / den : 0.f;
      };
      set_base({f0(p.eta.x, p.k.x), f0(p.eta.y, p.k.y), f0(p.eta.z, p.k.z)});
      m.metallic = 1.f;
      m.roughness = blended_roughness(p);
    } else if (p.type == "dielectric") {
      set_base({1.f, 1.f, 1.f});
      m.ior = p.eta_scalar;
      m.transmission = 1.f;
      m.roughness = blended_roughness(p);
    } else if (p.type == "coateddiffuse") {
      set_base(p.reflectance);
      m.metallic = 0.f;
      const float r = blended_roughness(p);
      m.roughness = r;
      m.clearcoat = 1.f;
      m.clearcoat_roughness = r * 0.5f;
    } else if (p.type == "diffusetransmission") {
      set_base(p.reflectance);
      m.roughness = 1.f;
      m.metallic = 0.f;
      m.transmission = 0.5f;
    } else {
      set_base(p.reflectance);
      m.roughness = p.roughness;
    }
    out.materials.push_back(m);
  }
  if (out.materials.empty()) {
    phx_material m = default_material();
    m.base_color[0] = m.base_color[1] = m.base_color[2] = 0.5f;
    m.roughness = 1.f;
    out.materials.push_back(m);
    mat_id["__default__"] = 0;
  }
  // every triangle gets three vertices of its own
  for (const Mesh &mesh : f.meshes) {
    const Mat4 xf = Mat4::from_columns(mesh.transform);
    uint32_t mid = 0;
    const auto mit = mat_id.find(mesh.material);
    if (mit != mat_id.end()) mid = mit->second;
    const size_t nverts = mesh.positions.size() / 3;
    const bool has_uv = mesh.uvs.size() >= nverts * 2, has_n = mesh.normals.size() >= mesh.positions.size();
    if (!mesh.alpha_texture.empty()) {
      const auto ait = tex_id.find(mesh.alpha_texture);
      if (ait != tex_id.end()) {
        phx_material am = out.materials[mid];
        am.alpha_tex = ait->second;
        mid = uint32_t(out.materials.size());
        out.materials.push_back(am);
      }
    }
    if (mesh.emissive) {
      phx_material em = default_material();
      em.base_color[0] = em.base_color[1] = em.base_color[2] = 0.f;
      em.emission[0] = mesh.emission.x;
      em.emission[1] = mesh.emission.y;
      em.emission[2] = mesh.emission.z;
      em.emission_strength = 1.f;
      em.roughness = 1.f;
      mid = uint32_t(out.materials.size());
      out.materials.push_back(em);
    }
    auto vertex = [&](int i) { return F3{mesh.positions[size_t(i) * 3], mesh.positions[size_t(i) * 3 + 1],
                                        mesh.positions[size_t(i) * 3 + 2]}; };
    auto normal_in = [&](int i) { return F3{mesh.normals[size_t(i) * 3], mesh.normals[size_t(i) * 3 + 1],
                                           mesh.normals[size_t(i) * 3 + 2]}; };
    for (size_t t = 0; t + 2 < mesh.indices.size(); t += 3) {
      const int i0 = mesh.indices[t], i1 = mesh.indices[t + 1], i2 = mesh.indices[t + 2];
      // a triangle naming a vertex the mesh does not have is left out (the
      // reference reads past its arrays there)
      if (std::min({i0, i1, i2}) < 0 || size_t(std::max({i0, i1, i2})) >= nverts) continue;
      const F3 p[3] = {xf.point(vertex(i0)), xf.point(vertex(i1)), xf.point(vertex(i2))};
      const uint32_t base = uint32_t(out.positions.size() / 3);
      for (int k = 0; k < 3; ++k) {
        out.positions.insert(out.positions.end(), {p[k].x, p[k].y, p[k].z});
        out.indices.push_back(base + uint32_t(k));
      }
      if (has_n) {
        const int ii[3] = {i0, i1, i2};
        for (int k = 0; k < 3; ++k) {
          const F3 n = normalize(xf.direction(normal_in(ii[k])));
          out.normals.insert(out.normals.end(), {n.x, n.y, n.z});
        }
      } else {
        const F3 n = normalize(cross(sub(p[1], p[0]), sub(p[2], p[0])));
        for (int k = 0; k < 3; ++k) out.normals.insert(out.normals.end(), {n.x, n.y, n.z});
      }
      if (has_uv) {
        const int ii[3] = {i0, i1, i2};
        for (int k = 0; k < 3; ++k)
          out.texcoords.insert(out.texcoords.end(), {mesh.uvs[size_t(ii[k]) * 2], mesh.uvs[size_t(ii[k]) * 2 + 1]});
      } else {
        out.texcoords.insert(out.texcoords.end(), {0.f, 0.f, 0.f, 0.f, 0.f, 0.f});
      }
      const uint32_t tri = uint32_t(out.material_ids.size());
      out.material_ids.push_back(mid);
      if (mesh.emissive) {
        out.emissive_ids.push_back(tri);
        out.emissive_areas.push_back(0.5f * length(cross(sub(p[1], p[0]), sub(p[2], p[0]))));
      }
    }
  }
  // textures, in name order
  uint32_t off = 0;
  for (const Texture *t : texs) {
    phx_texture_info info{};
    info.offset = off;
    info.width = uint32_t(t->width);
    info.height = uint32_t(t->height);
    out.texture_infos.push_back(info);
    out.texture_pixels.insert(out.texture_pixels.end(), t->rgb.begin(), t->rgb.end());
    off += uint32_t(t->width) * uint32_t(t->height);
  }
  // an Area light record per emissive mesh (pbrt_to_photon.cpp:351-393); the
  // light tables leave such records out, the emissive triangles stand for them
  if (!out.emissive_ids.empty()) {
    float total_area = 0.f;
    for (float a : out.emissive_areas) total_area += a;
    for (const Mesh &mesh : f.meshes) {
      if (!mesh.emissive) continue;
      const Mat4 xf = Mat4::from_columns(mesh.transform);
      auto vtx = [&](int i) {
        return xf.point({mesh.positions[size_t(i) * 3], mesh.positions[size_t(i) * 3 + 1],
                         mesh.positions[size_t(i) * 3 + 2]});
      };
      const size_t nv = mesh.positions.size() / 3;
      auto ok = [&](size_t k) { return k < mesh.indices.size() && mesh.indices[k] >= 0 && size_t(mesh.indices[k]) < nv; };
      if (!ok(0) || !ok(1)) continue;
      const F3 p0 = vtx(mesh.indices[0]), p1 = vtx(mesh.indices[1]);
      const F3 p3 = mesh.indices.size() >= 6 && ok(5) ? vtx(mesh.indices[5]) : p0;
      phx_light l{};  // pt::Light{} (light.h:19-35), then the area record
      l.spot_angle = 0.5f;
      l.spot_falloff = 0.1f;
      l.type = PHX_LIGHT_AREA;
      const F3 e1 = sub(p1, p0), e2 = sub(p3, p0), cr = cross(e1, e2);
      const float cl = length(cr);
      const F3 dir = cl > 0.f ? scale(cr, 1.f / cl) : F3{0.f, -1.f, 0.f};
      const float v[] = {p0.x, p0.y, p0.z, dir.x, dir.y, dir.z, mesh.emission.x, mesh.emission.y, mesh.emission.z};
      std::memcpy(l.position, v, 12);
      std::memcpy(l.direction, v + 3, 12);
      std::memcpy(l.color, v + 6, 12);
      l.intensity = 1.f;
      l.mesh_prim_begin = out.emissive_ids[0];
      l.mesh_prim_count = uint32_t(mesh.indices.size() / 3);
      l.area = total_area;
      l.edge1[0] = e1.x;
      l.edge1[1] = e1.y;
      l.edge1[2] = e1.z;
      l.edge2[0] = e2.x;
      l.edge2[1] = e2.y;
      l.edge2[2] = e2.z;
      out.lights.push_back(l);
    }
  }
  // the environment: pixels times the light's scale, rotated by the inverse
  // of its transform (pbrt_to_photon.cpp:400-437)
  if (f.env && !f.env_file.empty()) {
    Texture img;
    if (read_image(root + f.env_file, img)) {
      out.env_w = uint32_t(img.width);
      out.env_h = uint32_t(img.height);
      out.env_pixels.resize(img.rgb.size());
      for (size_t i = 0; i < img.rgb.size(); ++i) out.env_pixels[i] = img.rgb[i] * f.env_scale;
      const Mat4 inv = Mat4::from_columns(f.env_transform).inverse();
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out.env_w2t[r * 3 + c] = inv.m[r][c];
    }
  }
  // the camera (pbrt_to_photon.cpp:439-461)
  F3 pos, dir, up;
  if (f.camera.lookat) {
    pos = f.camera.from;
    dir = normalize(sub(f.camera.at, pos));
    up = normalize(f.camera.up);
  } else {
    const Mat4 c2w = Mat4::from_columns(f.camera.transform).inverse();
    pos = c2w.point({0.f, 0.f, 0.f});
    dir = normalize(c2w.direction({0.f, 0.f, 1.f}));
    up = normalize(c2w.direction({0.f, 1.f, 0.f}));
  }
  const float aspect = float(f.width) / float(f.height);
  const F3 at = add(pos, dir);
  const float from3[3] = {pos.x, pos.y, pos.z}, at3[3] = {at.x, at.y, at.z}, up3[3] = {up.x, up.y, up.z};
  phx_camera_perspective(&out.camera, from3, at3, up3, f.camera.fov, aspect, 0.f, 1.f);
  if (f.camera.lensradius > 0.f) {
    out.camera.lens_radius = f.camera.lensradius;
    out.camera.focus_dist = f.camera.focaldistance > 0.f ? f.camera.focaldistance : length(dir);
  }
  out.width = uint32_t(std::max(f.width, 1));
  out.height = uint32_t(std::max(f.height, 1));
  out.spp = uint32_t(std::max(f.spp, 1));
  out.max_depth = uint32_t(std::max(f.max_depth, 1));
  out.meshes = uint32_t(f.meshes.size());
  out.objects = uint32_t(f.objects.size());
  out.instances = uint32_t(f.instances.size());
  // no triangles: one degenerate triangle, which no ray hits, stands in for
  // the empty mesh (every camera ray misses, as with the reference's empty BVH)
  if (out.indices.empty()) {
    out.positions.assign(9, 0.f);
    out.normals.assign({0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f});
    out.texcoords.assign(6, 0.f);
    out.indices = {0u, 1u, 2u};
    out.material_ids = {0u};
  }
}

}  // namespace

extern "C" {

int phx_pbrt_load(const char *path, int use_instancing, phx_pbrt_scene **out) {
  if (!path || !out) return PHX_ERR_INVALID;
  *out = nullptr;
  std::unique_ptr<phx_pbrt_scene> s(new phx_pbrt_scene());
  try {
    File f = parse_file(path, use_instancing != 0);
    convert(f, dir_of(path), *s);
  } catch (const std::bad_alloc &) {
    return PHX_ERR_ALLOC;
  }
  *out = s.release();
  return PHX_OK;
}

int phx_pbrt_scene_desc(const phx_pbrt_scene *s, phx_scene_desc *d, phx_camera *camera, phx_pbrt_info *info) {
  if (!s) return PHX_ERR_INVALID;
  if (d) {
    *d = phx_scene_desc{};
    d->positions = s->positions.data();
    d->vertex_count = uint32_t(s->positions.size() / 3);
    d->indices = s->indices.data();
    d->triangle_count = uint32_t(s->indices.size() / 3);
    d->normals = s->normals.data();
    d->texcoords = s->texcoords.data();
    d->material_ids = s->material_ids.data();
    d->materials = s->materials.data();
    d->material_count = uint32_t(s->materials.size());
    d->lights = s->lights.empty() ? nullptr : s->lights.data();
    d->light_count = uint32_t(s->lights.size());
    d->emissive_prim_ids = s->emissive_ids.empty() ? nullptr : s->emissive_ids.data();
    d->emissive_prim_areas = s->emissive_areas.empty() ? nullptr : s->emissive_areas.data();
    d->emissive_count = uint32_t(s->emissive_ids.size());
    d->texture_pixels = s->texture_pixels.empty() ? nullptr : s->texture_pixels.data();
    d->texture_pixel_count = uint32_t(s->texture_pixels.size() / 3);
    d->texture_infos = s->texture_infos.empty() ? nullptr : s->texture_infos.data();
    d->texture_count = uint32_t(s->texture_infos.size());
    for (int k = 0; k < PHX_ATTRIBUTE_MESH_COUNT; ++k) d->attribute_offset[k] = -1;
    d->env_width = s->env_w;
    d->env_height = s->env_h;
    d->env_pixels = s->env_pixels.empty() ? nullptr : s->env_pixels.data();
    std::memcpy(d->env_world_to_tex, s->env_w2t, sizeof s->env_w2t);
    d->env_visible = 1;
  }
  if (camera) *camera = s->camera;
  if (info) {
    info->width = s->width;
    info->height = s->height;
    info->spp = s->spp;
    info->max_depth = s->max_depth;
    info->meshes = s->meshes;
    info->objects = s->objects;
    info->instances = s->instances;
    info->textures = uint32_t(s->texture_infos.size());
  }
  return PHX_OK;
}

void phx_pbrt_free(phx_pbrt_scene *s) { delete s; }

}  // extern "C"
